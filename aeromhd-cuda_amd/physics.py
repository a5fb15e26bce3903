"""Drop-in for the reference's physics.py (physics.py:284-429).

The numerical operators run as gfx950 HIP kernels of libaeromhd_hip.so:
``cons_to_prim``, ``prim_to_cons``, ``muscl_x``, ``muscl_y``, ``hlld_flux_x``,
``hlld_flux_y`` keep the reference's names, arguments and return values.
Arrays are torch tensors on the ROCm device (the role CuPy arrays play in the
reference); NumPy arrays are accepted too and NumPy arrays are returned for
them.  ``xp`` is a small float64 array namespace on the device for callers
that build initial conditions with ``xp.zeros`` / ``xp.where`` / ``xp.sin``
the way the reference's drivers do (main.py:51-59).

``mhd_rankine_hugoniot`` is the host-side jump solver of physics.py:385-429
(run once per initial condition).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import amhd_lib as _L
from config import (  # noqa: F401  (re-exported like the reference)
    NVAR, FLOOR_RHO, FLOOR_PR,
    RHO, MX, MY, MZ, BX, BY, BZ, EN, PSI, RHOC,
    iRHO, iVX, iVY, iVZ, iBX, iBY, iBZ, iPR, iPSI, iCLR,
)


# ------------------------------------------------------------------ arrays
def _torch():
    return _L.require_gpu()


def _to_device(a):
    """-> (contiguous float64 device tensor, came_from_numpy)."""
    torch = _torch()
    if isinstance(a, torch.Tensor):
        t = a
        if t.device.type != "cuda":
            t = t.to("cuda")
        if t.dtype != torch.float64:
            t = t.to(torch.float64)
        return t.contiguous(), False
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda"), True


def _from_device(t, host: bool):
    return t.cpu().numpy() if host else t


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _stream():
    return _L.stream_handle(_torch())


class _DeviceArrays:
    """float64 array functions on the ROCm device (CuPy-like subset)."""

    pi = np.pi
    float64 = np.float64
    newaxis = None

    @staticmethod
    def _t():
        return _torch()

    def _arr(self, a):
        torch = self._t()
        if isinstance(a, torch.Tensor):
            return a
        return torch.as_tensor(np.asarray(a, dtype=np.float64), device="cuda")

    def asarray(self, a, dtype=None):
        return self._arr(a)

    ascontiguousarray = asarray

    def zeros(self, shape, dtype=None):
        torch = self._t()
        return torch.zeros(shape, dtype=torch.float64, device="cuda")

    def ones(self, shape, dtype=None):
        torch = self._t()
        return torch.ones(shape, dtype=torch.float64, device="cuda")

    def empty(self, shape, dtype=None):
        torch = self._t()
        return torch.empty(shape, dtype=torch.float64, device="cuda")

    def zeros_like(self, a):
        return self._t().zeros_like(self._arr(a))

    def empty_like(self, a):
        return self._t().empty_like(self._arr(a))

    def arange(self, *args):
        return self._arr(np.arange(*args, dtype=np.float64))

    def linspace(self, *args, **kw):
        return self._arr(np.linspace(*args, **kw))

    def meshgrid(self, a, b, indexing="xy"):
        return self._t().meshgrid(self._arr(a), self._arr(b), indexing=indexing)

    def where(self, cond, a, b):
        torch = self._t()
        cond = self._arr(cond) if not isinstance(cond, torch.Tensor) else cond
        a = a if isinstance(a, torch.Tensor) else torch.tensor(float(a), dtype=torch.float64, device="cuda")
        b = b if isinstance(b, torch.Tensor) else torch.tensor(float(b), dtype=torch.float64, device="cuda")
        return torch.where(cond, a, b)

    def maximum(self, a, b):
        torch = self._t()
        a, b = self._arr(a), self._arr(b)
        return torch.maximum(a, b)

    def minimum(self, a, b):
        torch = self._t()
        return torch.minimum(self._arr(a), self._arr(b))

    def clip(self, a, lo, hi):
        return self._arr(a).clamp(lo, hi)

    def max(self, a):
        return self._arr(a).max()

    def min(self, a):
        return self._arr(a).min()

    def sum(self, a):
        return self._arr(a).sum()

    def mean(self, a):
        return self._arr(a).mean()

    def __getattr__(self, name):
        # elementwise math (sin, cos, tanh, exp, sqrt, abs, ...)
        torch = self._t()
        fn = getattr(torch, name)

        def wrapped(a, *args, **kw):
            return fn(self._arr(a), *args, **kw)

        return wrapped


xp = _DeviceArrays()
cp = xp


# ------------------------------------------------------------- operators
def cons_to_prim(U, gamma):
    """Conservative -> primitive (physics.py:284-301), HIP kernel."""
    Ud, host = _to_device(U)
    W = _torch().empty_like(Ud)
    n = Ud.numel() // NVAR
    _L.call("amhd_cons_to_prim", _ptr(Ud), _ptr(W), n, float(gamma), _stream())
    return _from_device(W, host)


def prim_to_cons(W, gamma):
    """Primitive -> conservative (physics.py:304-320), HIP kernel."""
    Wd, host = _to_device(W)
    U = _torch().empty_like(Wd)
    n = Wd.numel() // NVAR
    _L.call("amhd_prim_to_cons", _ptr(Wd), _ptr(U), n, float(gamma), _stream())
    return _from_device(U, host)


def _muscl(W, direction: int):
    Wd, host = _to_device(W)
    if Wd.dim() != 3 or Wd.shape[0] != NVAR:
        raise ValueError(f"muscl: W must have shape (NVAR, nxt, nyt), got {tuple(Wd.shape)}")
    nxt, nyt = int(Wd.shape[1]), int(Wd.shape[2])
    shape = (NVAR, nxt - 3, nyt) if direction == 0 else (NVAR, nxt, nyt - 3)
    torch = _torch()
    WL = torch.empty(shape, dtype=torch.float64, device="cuda")
    WR = torch.empty(shape, dtype=torch.float64, device="cuda")
    _L.call("amhd_muscl", _ptr(Wd), nxt, nyt, direction, _ptr(WL), _ptr(WR), _stream())
    return _from_device(WL, host), _from_device(WR, host)


def muscl_x(W):
    """MUSCL / van Leer reconstruction in x (physics.py:323-335)."""
    return _muscl(W, 0)


def muscl_y(W):
    """MUSCL / van Leer reconstruction in y (physics.py:338-350)."""
    return _muscl(W, 1)


def _hlld(WL, WR, gamma, ch, direction: int, smax_arr=None):
    WLd, host = _to_device(WL)
    WRd, _ = _to_device(WR)
    if WLd.shape != WRd.shape or WLd.shape[0] != NVAR:
        raise ValueError("hlld: WL and WR must have the same shape (NVAR, ...)")
    n = WLd.numel() // NVAR
    F = _torch().empty_like(WLd)
    sa = None
    if smax_arr is not None:
        if int(np.prod(np.shape(smax_arr))) != n:
            raise ValueError(f"hlld: smax_arr must hold one value per face ({n})")
        sa = _torch().empty(n, dtype=WLd.dtype, device=WLd.device)
    smax = C.c_double(0.0)
    _L.call("amhd_hlld_flux", _ptr(WLd), _ptr(WRd), _ptr(F), _ptr(sa) if sa is not None else None, C.byref(smax), n,
            direction, float(gamma), float(ch), FLOOR_RHO, FLOOR_PR, _L.AMHD_F64, _stream())
    if smax_arr is not None:
        if isinstance(smax_arr, np.ndarray):
            smax_arr[...] = sa.cpu().numpy().reshape(smax_arr.shape)
        else:
            smax_arr.copy_(sa.reshape(smax_arr.shape))
    return _from_device(F, host), float(smax.value)


def hlld_flux_x(WL, WR, gamma, ch, smax_arr=None):
    """HLLD flux through x faces, one thread per face (physics.py:353-370).

    Returns ``(F, smax)`` with ``smax = max(|SL|, |SR|)`` over the faces;
    ``smax_arr`` (optional, one value per face) receives each face's
    ``max(|SL|, |SR|)`` as the reference kernel's ``smax_arr``
    (physics.py:108-109)."""
    return _hlld(WL, WR, gamma, ch, 0, smax_arr)


def hlld_flux_y(WL, WR, gamma, ch, smax_arr=None):
    """HLLD flux through y faces by the (vx,vy),(Bx,By) rotation
    (physics.py:373-379), done inside the kernel."""
    return _hlld(WL, WR, gamma, ch, 1, smax_arr)


# ---------------------------------------------------- Rankine-Hugoniot (host)
def mhd_rankine_hugoniot(gamma, M, rho1, p1, By1):
    """Post-shock state of a perpendicular MHD (or hydro) shock moving into
    (rho1, p1, By1) at Mach M (physics.py:385-429).

    Hydro: closed-form jump.  MHD: Newton iteration on the energy residual
    for the compression ratio r with a central-difference derivative, r kept
    in [1.001, (g+1)/(g-1) - 0.001].  Returns (rho2, p2, vx2, By2, vs)."""
    g = gamma
    M2 = M * M
    if abs(By1) < 1e-14:
        r = (g + 1) * M2 / ((g - 1) * M2 + 2)
        cs1 = np.sqrt(g * p1 / rho1)
        vs = M * cs1
        p2 = p1 * (2 * g * M2 - (g - 1)) / (g + 1)
        vx2 = vs * (1 - 1.0 / r)
        return rho1 * r, p2, vx2, 0.0, vs

    cf1 = np.sqrt(g * p1 / rho1 + By1 ** 2 / rho1)
    vs = M * cf1
    r_hi = (g + 1) / (g - 1)

    def p_behind(r):
        return p1 + rho1 * vs ** 2 * (1 - 1.0 / r) + 0.5 * By1 ** 2 * (1 - r ** 2)

    h_up = 0.5 * vs ** 2 + g / (g - 1) * p1 / rho1 + By1 ** 2 / rho1

    def residual(r):
        h_dn = 0.5 * vs ** 2 / r ** 2 + g / (g - 1) * p_behind(r) / (rho1 * r) + r * By1 ** 2 / rho1
        return h_up - h_dn

    r = min((g + 1) * M2 / ((g - 1) * M2 + 2), r_hi - 0.01)
    for _ in range(100):
        h = 1e-6 * r
        slope = (residual(r + h) - residual(r - h)) / (2 * h)
        if abs(slope) < 1e-30:
            break
        r_next = min(max(r - residual(r) / slope, 1.001), r_hi - 0.001)
        converged = abs(r_next - r) < 1e-12
        r = r_next
        if converged:
            break
    return rho1 * r, p_behind(r), vs * (1 - 1.0 / r), r * By1, vs
