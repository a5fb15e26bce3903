"""CPU oracle of the reference's MHD update -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper over liboracle.so, the plain-C restatement in oracle.c of
physics.py:43-379 and solver.py:158-333 on the reference's ghosted layout
(float64, -ffp-contract=off).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it, and only as what results are compared
with / the timed CPU baseline: the product (aeromhd-cuda_amd/) never
imports it.

What pins each oracle function independently of its own text (tests under
tests/ named in brackets):

* orc_hlld_flux -- bit-for-bit against the reference's own kernel source
  (physics.py:43-275) compiled by hipcc into oracle/_ref and run on the GPU
  [tests/golden/hlld_ref_vectors.npz, test_oracle_pins.py::test_hlld_matches_reference_kernel];
  supersonic branches equal the physical flux, F(W, W) = F(W), the GLM
  2x2 face solution, the exact moving contact, x<->y rotation and mirror
  symmetry [test_oracle_pins.py].
* orc_muscl_x / _y -- piecewise-constant data reproduces the cell values,
  extrema get zero slope, linear data gives the face midpoint, TVD bounds,
  x/y transpose symmetry [test_oracle_pins.py].
* orc_cons_to_prim / orc_prim_to_cons -- round trip, closed-form pressure.
* orc_compute_dt -- Brio-Wu at t = 0: cf_max = sqrt(14.1), ch = 1.5 cf_max,
  dt = 0.4 dx / ch [test_oracle_pins.py].
* orc_apply_bc -- ghost images by construction on index grids.
* orc_step_ssprk3 / orc_run -- exact conservation with periodic boundaries,
  uniform state preserved, x<->y transpose symmetry, and the reference's
  published V&V: Brio-Wu checks (main.py:88-97), Alfven L1(By) errors
  (PDF p.3: ~3.1e-8, 8.7e-9, 2.2e-9, 5.5e-10) and order >= 1.5 (published
  1.96), contact p and vx constant [test_oracle_vv.py].  The stage weights
  by the scheme's temporal order: on a checkerboard state (every slope 0,
  so the semi-discrete system is a smooth ODE) the error falls 6-12x per
  halving of dt, and the same step composed with wrong weights fails
  [test_oracle_pins_step.py::test_ssprk3_*].  The HIP path reproduces the
  reference's published RMI results through this step: peak integral
  mixing widths 0.0968 / 0.0879 / 0.0579 (PDF p.5) and max|divB| peaks
  ~13 / ~27 (PDF p.4) [test_gpu_workloads.py], and equals the oracle for
  the first 150 steps of each case.
* orc_compute_rhs's Powell source (solver.py:242-266) -- R(on) - R(off) =
  -divB (0, B, v, v.B, 0, 0) on fields with Bx linear in x and By linear
  in y, whose divergence is known [test_oracle_pins_step.py::test_powell_*].
* orc_enforce_bounds -- case by case: C in [0, 1] kept, C > 1 -> rho,
  C < 0 -> 0, rho below the floor -> the floor
  [test_oracle_pins_step.py::test_scalar_bounds_case_by_case].
* orc_boundary_flux -- the textbook ideal-MHD energy flux summed over two
  non-uniform boundary columns (every other column holding unrelated
  values), and the energy balance of an RMI run: the raw drift is
  large and the drift corrected by the accumulated flux is ~0, while the
  opposite sign breaks it (PDF p.1) [test_oracle_pins_step.py::test_boundary_*].
  The sum order of the reference (CuPy's reduction) is not published, so
  the flux is compared at 1e-13 relative, not bit for bit.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
NV = 10


class OrcParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("gamma", C.c_double), ("dx", C.c_double),
                ("dy", C.c_double), ("bcx_periodic", C.c_int), ("bcy_periodic", C.c_int),
                ("powell", C.c_int), ("glm_alpha", C.c_double)]


def build(force: bool = False) -> str:
    """gcc -O3 -fopenmp -ffp-contract=off: IEEE per-operation FP64 (x86-64
    baseline ISA, no FMA)."""
    src = os.path.join(HERE, "oracle.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        tmp = LIB + ".tmp"
        subprocess.run(["gcc", "-O3", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                        "-std=c11", "-o", tmp, src, "-lm"], check=True)
        os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp, vp, i, l, d = C.POINTER(C.c_double), C.c_void_p, C.c_int, C.c_long, C.c_double
        pp = C.POINTER(OrcParams)
        sig = {
            "orc_set_threads": (None, [i]), "orc_get_threads": (i, []),
            "orc_cons_to_prim": (None, [vp, vp, l, d]), "orc_prim_to_cons": (None, [vp, vp, l, d]),
            "orc_muscl_x": (None, [vp, i, i, vp, vp]), "orc_muscl_y": (None, [vp, i, i, vp, vp]),
            "orc_hlld_flux": (None, [vp, vp, vp, dp, l, i, d, d, d, d]),
            "orc_apply_bc": (None, [vp, i, i, i, i]), "orc_enforce_bounds": (None, [vp, l]),
            "orc_compute_dt": (d, [vp, l, d, d, d, d, dp]), "orc_boundary_flux": (d, [vp, pp]),
            "orc_compute_rhs": (None, [vp, vp, pp, d]), "orc_step_ssprk3": (d, [vp, pp, d, d, d]),
            "orc_run": (i, [vp, pp, d, d, i, dp, dp, dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def params(cfg) -> OrcParams:
    """orc_params of a Config (BC kinds, Powell gate solver.py:304)."""
    return OrcParams(int(cfg.nx), int(cfg.ny), float(cfg.gamma), float(cfg.dx), float(cfg.dy),
                     1 if cfg.get_bc_x() == "periodic" else 0, 1 if cfg.bc_y_type == "periodic" else 0,
                     1 if (cfg.powell_source and cfg.B_transverse > 0) else 0, float(cfg.glm_alpha))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def threads() -> int:
    return int(lib().orc_get_threads())


def cons_to_prim(U, gamma):
    U = _f64(U)
    W = np.empty_like(U)
    lib().orc_cons_to_prim(_p(U), _p(W), U.size // NV, float(gamma))
    return W


def prim_to_cons(W, gamma):
    W = _f64(W)
    U = np.empty_like(W)
    lib().orc_prim_to_cons(_p(W), _p(U), W.size // NV, float(gamma))
    return U


def muscl_x(W):
    W = _f64(W)
    nxt, nyt = W.shape[1], W.shape[2]
    WL = np.empty((NV, nxt - 3, nyt))
    WR = np.empty_like(WL)
    lib().orc_muscl_x(_p(W), nxt, nyt, _p(WL), _p(WR))
    return WL, WR


def muscl_y(W):
    W = _f64(W)
    nxt, nyt = W.shape[1], W.shape[2]
    WL = np.empty((NV, nxt, nyt - 3))
    WR = np.empty_like(WL)
    lib().orc_muscl_y(_p(W), nxt, nyt, _p(WL), _p(WR))
    return WL, WR


def hlld_flux(WL, WR, gamma, ch, direction=0, floor_rho=1e-12, floor_pr=1e-12):
    WL, WR = _f64(WL), _f64(WR)
    n = WL.size // NV
    F = np.empty_like(WL)
    smax = C.c_double(0.0)
    lib().orc_hlld_flux(_p(WL), _p(WR), _p(F), C.byref(smax), n, int(direction), float(gamma), float(ch),
                        float(floor_rho), float(floor_pr))
    return F, smax.value


def apply_bc(U, cfg):
    U = _f64(U).copy()
    p = params(cfg)
    lib().orc_apply_bc(_p(U), p.nx, p.ny, p.bcx_periodic, p.bcy_periodic)
    return U


def enforce_bounds(U):
    U = _f64(U).copy()
    lib().orc_enforce_bounds(_p(U), U.size // NV)
    return U


def compute_dt(W, cfg):
    """-> (dt, ch) as solver.py:206-223 over the whole array W."""
    W = _f64(W)
    ch = C.c_double()
    dt = lib().orc_compute_dt(_p(W), W.size // NV, float(cfg.gamma), float(cfg.cfl), float(cfg.dx),
                              float(cfg.dy), C.byref(ch))
    return dt, ch.value


def boundary_flux(U, cfg):
    U = _f64(U)
    p = params(cfg)
    return lib().orc_boundary_flux(_p(U), C.byref(p))


def compute_rhs(U, cfg, ch):
    """-> (R, U with BC applied) as solver.py:268-308."""
    U = _f64(U).copy()
    R = np.empty_like(U)
    p = params(cfg)
    lib().orc_compute_rhs(_p(U), _p(R), C.byref(p), float(ch))
    return R, U


def psi_decay(cfg, ch, dt):
    """solver.py:331-333 on the host (numpy exp, as the reference)."""
    return float(np.exp(-cfg.glm_alpha * ch * dt / min(cfg.dx, cfg.dy))) if ch > 0 else 1.0


def step(U, cfg, dt, ch, decay=None):
    """One step_ssprk3 (solver.py:310-333) -> (U_new, boundary flux)."""
    U = _f64(U).copy()
    p = params(cfg)
    if decay is None:
        decay = psi_decay(cfg, ch, dt)
    bf = lib().orc_step_ssprk3(_p(U), C.byref(p), float(dt), float(ch), float(decay))
    return U, bf


def run_loop(U, cfg, n_steps, t=0.0):
    """The time loop of main.py:66-73 / solver.py:475-482 in Python over
    the oracle (host numpy exp for the decay, as the reference).  Returns
    (U, t, steps, last_dt, cumulative boundary energy)."""
    U = _f64(U).copy()
    steps, last, cum = 0, 0.0, 0.0
    while t < cfg.t_end and steps < n_steps:
        dt, ch = compute_dt(cons_to_prim(U, cfg.gamma), cfg)
        dt = min(dt, cfg.t_end - t)
        if dt <= 1e-16:
            break
        U, bf = step(U, cfg, dt, ch)
        cum += bf * dt
        t += dt
        steps += 1
        last = dt
    return U, t, steps, last, cum
