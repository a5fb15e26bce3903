"""CPU oracle of the learned-closure convolutions -- TEST INFRASTRUCTURE ONLY.

NumPy restatement of the three 3x3 convolution passes that
aeromhd-cuda_amd/csrc/amhd_cnn.hip runs on BF16 matrix cores:

* conv_fwd:         y[p, co] = act(b[co] + sum_{tap, ci} x[p + off(tap), ci] W[co, tap*Cin + ci])
* conv_bwd_data:    dx[p, ci] = mask * sum_{tap, co} dy[p - off(tap), co] W[co, tap*Cin + ci]
* conv_bwd_weight:  dW[co, tap*Cin + ci] = sum_p dy[p, co] x[p + off(tap), ci];  db = sum_p dy

on NHWC arrays (n, nx, ny, C), zero padding in x and periodic in y.  With
``bf16=True`` the operands are rounded to BF16 exactly where the kernels
round them (activations and weights are BF16 tensors; FP32 gradients are
rounded when loaded into the MFMA); sums are in float64.

and the inference glue of the same file: the network input (features),
the donor-cell rhoC flux (mix_fluxes), the limited source
(limiter_theta for positivity, limiter_theta_mix for the bounds of the
concentration, limited_source) and the limited anti-diffusion of the
concentration across x faces (sharpen_x).

Pins (tests/test_closure_oracle.py): the forward equals
torch.nn.functional.conv2d on the CPU (float64) with the same padding; the
backward passes equal torch autograd of that conv2d; the adjoint identity
<conv(x), y> = <x, conv^T(y)> holds; a finite-difference check of the whole
closure loss (with the donor-cell mixing); limiter_theta meets its
definition face by face (pressure recomputed independently) and gives
theta = 1 for a vanishing correction; every cell of U0 + dt *
limited_source keeps the rho / p bound for random face weights below
theta and keeps C within its neighbourhood's bounds (the convexity
arguments), where the unlimited correction breaks both; limited_source
with unit thetas equals source_from_flux and conserves every field;
sharpen_x reproduces a four-cell profile solved by hand (unsaturated and
saturated), keeps C within its bounds where the unlimited pass leaves them,
conserves rhoC row by row, leaves every other plane and a y-uniform
profile's uniformity alone, and is the identity at kappa = 0.
"""
from __future__ import annotations

import numpy as np

TAPS = [(di, dj) for di in (-1, 0, 1) for dj in (-1, 0, 1)]   # tap = (di+1)*3 + (dj+1)


def bf16_round(a) -> np.ndarray:
    """float32 -> nearest-even BF16 -> float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, (u >> 16) << 16, r).astype(np.uint32)
    return r.view(np.float32)


def kpad(cin: int) -> int:
    return (9 * cin + 31) // 32 * 32


def shift(x: np.ndarray, di: int, dj: int) -> np.ndarray:
    """out[n, i, j] = x[n, i + di, (j + dj) mod ny], zero where i + di is
    outside [0, nx)."""
    n, nx, ny, c = x.shape
    out = np.zeros_like(x)
    src = np.roll(x, -dj, axis=2)
    lo, hi = max(0, -di), min(nx, nx - di)
    out[:, lo:hi] = src[:, lo + di:hi + di]
    return out


def conv_fwd(x, W, b, relu: bool, bf16: bool = True):
    x = np.asarray(x, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    if bf16:
        x = bf16_round(x).astype(np.float64)
        W = bf16_round(W).astype(np.float64)
    cin = x.shape[3]
    y = np.zeros(x.shape[:3] + (W.shape[0],))
    for t, (di, dj) in enumerate(TAPS):
        y += shift(x, di, dj) @ W[:, t * cin:(t + 1) * cin].T
    y += np.asarray(b, dtype=np.float64)
    return np.maximum(y, 0.0) if relu else y


def conv_bwd_data(dy, W, cin: int, mask=None, bf16: bool = True):
    dy = np.asarray(dy, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    if bf16:
        dy = bf16_round(dy).astype(np.float64)
        W = bf16_round(W).astype(np.float64)
    dx = np.zeros(dy.shape[:3] + (cin,))
    for t, (di, dj) in enumerate(TAPS):
        dx += shift(dy, -di, -dj) @ W[:, t * cin:(t + 1) * cin]
    if mask is not None:
        dx = np.where(np.asarray(mask) > 0, dx, 0.0)
    return dx


def conv_bwd_weight(x, dy, bf16: bool = True):
    x = np.asarray(x, dtype=np.float64)
    dy64 = np.asarray(dy, dtype=np.float64)
    dyr = bf16_round(dy64).astype(np.float64) if bf16 else dy64
    if bf16:
        x = bf16_round(x).astype(np.float64)
    cin, cout = x.shape[3], dy64.shape[3]
    dW = np.zeros((cout, kpad(cin)))
    a = dyr.reshape(-1, cout)
    for t, (di, dj) in enumerate(TAPS):
        dW[:, t * cin:(t + 1) * cin] = a.T @ shift(x, di, dj).reshape(-1, cin)
    return dW, dy64.reshape(-1, cout).sum(axis=0)


# ------------------------------------------------------------------ network
LAYERS = [(16, 32, True), (32, 32, True), (32, 32, True), (32, 32, False)]
FEAT = [0, 1, 2, 3, 4, 5, 6, 7, 9]


def net_forward(x, Ws, bs, bf16: bool = True):
    """Layer inputs and the FP32-accumulated output of the 4-layer net;
    hidden activations are stored in BF16 (as the kernels store them)."""
    acts = [np.asarray(x, dtype=np.float64)]
    h = acts[0]
    for li, (cin, cout, relu) in enumerate(LAYERS):
        y = conv_fwd(h, Ws[li], bs[li], relu, bf16=bf16)
        if li < len(LAYERS) - 1:
            h = bf16_round(y).astype(np.float64) if bf16 else y
            acts.append(h)
        else:
            h = y
    return h, acts


def net_backward(acts, dout, Ws, bf16: bool = True):
    gW, gb = [None] * len(LAYERS), [None] * len(LAYERS)
    dy = np.asarray(dout, dtype=np.float64)
    for li in range(len(LAYERS) - 1, -1, -1):
        cin, cout, _ = LAYERS[li]
        gW[li], gb[li] = conv_bwd_weight(acts[li], dy, bf16=bf16)
        if li > 0:
            dy = conv_bwd_data(dy, Ws[li], cin, mask=acts[li], bf16=bf16)
    return gW, gb


RC = 8   # position of rhoC among the network fields (FEAT)


def mix_fluxes(fx, fy, C):
    """Face fluxes of the 9 fields with the rhoC channel rebuilt: donor-cell
    transport of the mass flux plus the network's mixing flux,
    F_rhoC = C_donor * F_rho + G.  fx, fy: (n, nx, ny, 9) at faces i-1/2,
    j-1/2; C (n, nx, ny); the donor of face i-1/2 is cell i-1 when
    F_rho > 0, else cell i (likewise in y).  Returns fx, fy, Cx, Cy."""
    Cx = np.where(fx[..., 0] > 0, np.roll(C, 1, axis=1), C)
    Cy = np.where(fy[..., 0] > 0, np.roll(C, 1, axis=2), C)
    fx = fx.copy()
    fy = fy.copy()
    fx[..., RC] = fx[..., RC] + Cx * fx[..., 0]
    fy[..., RC] = fy[..., RC] + Cy * fy[..., 0]
    return fx, fy, Cx, Cy


def divergence(fx, fy, dx, dy):
    """-(d_x fx + d_y fy) with fx at i-1/2 (zero on both x boundaries) and
    fy at j-1/2 (periodic)."""
    fx = fx.copy()
    fx[:, 0] = 0.0
    fxe = np.concatenate([fx, np.zeros_like(fx[:, :1])], axis=1)
    return -(fxe[:, 1:] - fxe[:, :-1]) / dx - (np.roll(fy, -1, axis=2) - fy) / dy


def source_from_flux(dF, dx, dy, C):
    """S = -div of the face corrections dF (n, nx, ny, 18), already in
    physical units: channels 0..8 at x faces i-1/2, 9..17 at y faces j-1/2,
    the rhoC channels being the mixing flux G (mix_fluxes).  C: the
    concentration of the state the correction is added to."""
    n9 = len(FEAT)
    fx, fy, _, _ = mix_fluxes(dF[..., :n9], dF[..., n9:2 * n9], C)
    return divergence(fx, fy, dx, dy)


def concentration(U):
    """C = rhoC / rho of states U (n, 10, nx, ny)."""
    return U[:, 9] / U[:, 0]


def loss(out, R, fscale, rstd, dx, dy, C):
    """Normalised MSE of S(out * fscale) against R (n, 9, nx, ny)."""
    S = source_from_flux(out[..., :18] * np.concatenate([fscale, fscale]), dx, dy, C)
    e = (S - np.transpose(R, (0, 2, 3, 1))) / rstd
    return float(np.mean(e * e))


def loss_grad_out(out, R, fscale, rstd, dx, dy, C):
    """d loss / d out (adjoints of the divergence and of the donor-cell
    mixing, by hand; the donor choice is piecewise constant)."""
    n9 = len(FEAT)
    fs = np.concatenate([fscale, fscale])
    sc = out[..., :18] * fs
    fx, fy, Cx, Cy = mix_fluxes(sc[..., :n9], sc[..., n9:2 * n9], C)
    S = divergence(fx, fy, dx, dy)
    g = 2.0 * (S - np.transpose(R, (0, 2, 3, 1))) / rstd ** 2 / S.size
    gx = np.zeros_like(g)                    # d/d fx_i: S_i has +fx_i/dx, S_{i-1} has -fx_i/dx
    gx[:, 1:] = (g[:, 1:] - g[:, :-1]) / dx
    gy = (g - np.roll(g, 1, axis=2)) / dy
    gx[..., 0] += Cx * gx[..., RC]
    gy[..., 0] += Cy * gy[..., RC]
    d = np.zeros_like(out)
    d[..., :n9] = gx * fscale
    d[..., n9:2 * n9] = gy * fscale
    return d


# --------------------------------------------------------- inference glue
THETAS = [1.0, 0.5, 0.25, 0.125, 0.0625, 0.03125]


def features(U, fields, mean, std):
    """BF16 network input (n, nx, ny, 16) of states U (n, planes, nx, ny):
    channel f = (float32(U[fields[f]]) - mean[f]) / std[f] in float32."""
    U = np.asarray(U)
    n, _, nx, ny = U.shape
    out = np.zeros((n, nx, ny, 16), dtype=np.float32)
    for f, pl in enumerate(fields):
        v = (U[:, pl].astype(np.float32) - np.float32(mean[f])) / np.float32(std[f])
        out[..., f] = v.astype(np.float32)
    return bf16_round(out)


def _pressure(U, gamma):
    ke = 0.5 * (U[:, 1] ** 2 + U[:, 2] ** 2 + U[:, 3] ** 2) / U[:, 0]
    me = 0.5 * (U[:, 4] ** 2 + U[:, 5] ** 2 + U[:, 6] ** 2)
    return (gamma - 1.0) * (U[:, 7] - ke - me)


def _faces(f, dt, dx, dy, fy=None):
    """The increments the four faces of each cell bring alone, dt-scaled:
    (4, n, nx, ny, ...) for i-1/2, i+1/2, j-1/2, j+1/2; f / fy at faces
    i-1/2 / j-1/2 (x faces on the domain boundary bring nothing)."""
    xl = f.copy()
    xl[:, 0] = 0.0
    xr = np.zeros_like(f)
    xr[:, :-1] = f[:, 1:]
    return np.stack([xl * dt / dx, -xr * dt / dx, fy * dt / dy, -np.roll(fy, -1, axis=2) * dt / dy])


def admissible(U0, W, gamma, keep):
    """rho > 0, rho >= keep*rho0 and p >= keep*p0 (the bound is rho0 / p0
    itself where those are not positive); U0, W (n, 10, nx, ny)."""
    rho0, p0 = U0[:, 0], _pressure(U0, gamma)
    rmin = np.where(rho0 > 0, keep * rho0, rho0)
    pmin = np.where(p0 > 0, keep * p0, p0)
    return (W[:, 0] > 0) & (W[:, 0] >= rmin) & (_pressure(W, gamma) >= pmin)


def _largest(ok_at):
    """Per cell, the largest t of THETAS with ok_at(t) (else 0)."""
    out = None
    done = None
    for t in THETAS:
        ok = ok_at(t)
        if out is None:
            out = np.zeros(ok.shape)
            done = np.zeros(ok.shape, dtype=bool)
        ok = ok & ~done
        out[ok] = t
        done |= ok
    return out


def limiter_theta(U0, dF, fscale, dt, dx, dy, gamma, keep):
    """Per-cell limiter of the non-mixing corrections: the smallest over
    the cell's four faces of the largest t of THETAS (else 0) for which
    U0 + 4 t D_face alone is admissible.  U0 (n, 10, nx, ny), dF (n, nx,
    ny, 32) -> theta (n, nx, ny)."""
    U0 = np.asarray(U0, dtype=np.float64)
    n9 = len(FEAT)
    dF = np.asarray(dF, dtype=np.float64)
    theta = np.ones(U0[:, 0].shape)
    for Df in _faces(dF[..., :n9] * fscale, dt, dx, dy, dF[..., n9:2 * n9] * fscale):
        D = np.zeros_like(U0)
        D[:, FEAT] = np.transpose(Df, (0, 3, 1, 2))
        D[:, 9] = 0.0                       # the mixing channel has its own limiter
        theta = np.minimum(theta, _largest(lambda t: admissible(U0, U0 + 4.0 * t * D, gamma, keep)))
    return theta


def face_weights(theta):
    """min over each face's two cells: wx at i-1/2 (0 at i = 0), wy at j-1/2."""
    wx = np.zeros_like(theta)
    wx[:, 1:] = np.minimum(theta[:, 1:], theta[:, :-1])
    wy = np.minimum(theta, np.roll(theta, 1, axis=2))
    return wx, wy


def local_bounds(a):
    """min and max over each cell and its face neighbours (no neighbour
    beyond the x boundaries, y periodic)."""
    lo, hi = a.copy(), a.copy()
    for nb in (np.roll(a, 1, axis=2), np.roll(a, -1, axis=2)):
        lo, hi = np.minimum(lo, nb), np.maximum(hi, nb)
    lo[:, 1:] = np.minimum(lo[:, 1:], a[:, :-1])
    hi[:, 1:] = np.maximum(hi[:, 1:], a[:, :-1])
    lo[:, :-1] = np.minimum(lo[:, :-1], a[:, 1:])
    hi[:, :-1] = np.maximum(hi[:, :-1], a[:, 1:])
    return lo, hi


def limiter_theta_mix(U0, dF, theta, fscale, dt, dx, dy, Ub=None):
    """Per-cell limiter of the mixing flux G: with the mass flux limited by
    theta's face weights and its donor-cell rhoC transport applied, the
    smallest over the faces of the largest t of THETAS (else 0) for which
    the face's G alone, times 4 t, keeps C = rhoC/rho within the min and max
    of C over the cell and its face neighbours in U0 and, if given, in Ub
    as well (the state before the step: Zalesak's bounds)."""
    U0 = np.asarray(U0, dtype=np.float64)
    n9 = len(FEAT)
    dF = np.asarray(dF, dtype=np.float64)
    C = concentration(U0)
    wx, wy = face_weights(theta)
    fx = dF[..., :n9] * fscale
    fy = dF[..., n9:2 * n9] * fscale
    mx = np.zeros_like(fx)
    my = np.zeros_like(fy)
    mx[..., 0] = wx * fx[..., 0]
    my[..., 0] = wy * fy[..., 0]
    mx, my, _, _ = mix_fluxes(mx, my, C)     # rho and donor rhoC parts only
    Sm = divergence(mx, my, dx, dy)
    rho1 = U0[:, 0] + dt * Sm[..., 0]
    rc1 = U0[:, 9] + dt * Sm[..., RC]
    clo, chi = local_bounds(C)
    if Ub is not None:
        blo, bhi = local_bounds(concentration(np.asarray(Ub, dtype=np.float64)))
        clo, chi = np.minimum(clo, blo), np.maximum(chi, bhi)
    th = np.ones(theta.shape)
    for G in _faces(fx[..., RC], dt, dx, dy, fy[..., RC]):
        th = np.minimum(th, _largest(lambda t: (rc1 + 4.0 * t * G >= clo * rho1) & (rc1 + 4.0 * t * G <= chi * rho1)))
    return th


def limited_source(dF, theta, theta_mix, fscale, dx, dy, C):
    """-div of the face corrections, the non-mixing channels scaled by
    theta's face weights and the mixing flux G by theta_mix's; the rhoC
    flux is C_donor * (weighted F_rho) + weighted G.  Returns (n, nx, ny, 9)."""
    n9 = len(FEAT)
    dF = np.asarray(dF, dtype=np.float64)
    wx, wy = face_weights(theta)
    gx, gy = face_weights(theta_mix)
    fx = dF[..., :n9] * fscale * wx[..., None]
    fy = dF[..., n9:2 * n9] * fscale * wy[..., None]
    fx[..., RC] = dF[..., RC] * fscale[RC] * gx
    fy[..., RC] = dF[..., n9 + RC] * fscale[RC] * gy
    fx, fy, _, _ = mix_fluxes(fx, fy, C)
    return divergence(fx, fy, dx, dy)


def sharpen_x(U, Ub, kappa):
    """Limited anti-diffusion of C across the x faces, one pass of
    flux-corrected transport (Zalesak 1979, J. Comput. Phys. 31, 335) on
    the states U (n, 10, nx, ny).  Face i+1/2 (i < nx-1) carries the
    anti-diffusive rhoC flux A = kappa * (rho_i + rho_{i+1}) / 2 *
    (C_{i+1} - C_i), positive towards i+1 (from the lower concentration to
    the higher); x-boundary faces carry nothing.  Each cell's inflows
    P+ and outflows P- (sums of the positive and negative parts of its two
    face fluxes) are scaled by R+ = min(1, Q+ / P+) and R- = min(1, Q- / P-)
    with Q+ = rho (C_hi - C), Q- = rho (C - C_lo), C_lo / C_hi the min / max
    of C over the cell and its face neighbours in U and, if given, in Ub
    (the state before the step); face i+1/2 takes min(R+ of the receiving
    cell, R- of the giving one).  Every cell keeps C within [C_lo, C_hi],
    rho and every other plane are untouched, and each row conserves rhoC.
    Returns the new states."""
    U = np.array(U, dtype=np.float64)
    rho, rc = U[:, 0], U[:, 9]
    C = concentration(U)
    A = np.zeros(rho.shape)                                          # A[:, i]: face i+1/2
    A[:, :-1] = kappa * 0.5 * (rho[:, :-1] + rho[:, 1:]) * (C[:, 1:] - C[:, :-1])
    Al = np.zeros_like(A)                                            # face i-1/2
    Al[:, 1:] = A[:, :-1]
    gin = np.maximum(Al, 0.0) + np.maximum(-A, 0.0)
    gout = np.maximum(A, 0.0) + np.maximum(-Al, 0.0)
    lo, hi = local_bounds(C)
    if Ub is not None:
        blo, bhi = local_bounds(concentration(np.asarray(Ub, dtype=np.float64)))
        lo, hi = np.minimum(lo, blo), np.maximum(hi, bhi)
    qp = np.maximum(rho * hi - rc, 0.0)
    qm = np.maximum(rc - rho * lo, 0.0)
    rp = np.where(gin > 0, np.minimum(1.0, qp / np.where(gin > 0, gin, 1.0)), 0.0)
    rm = np.where(gout > 0, np.minimum(1.0, qm / np.where(gout > 0, gout, 1.0)), 0.0)
    L = np.zeros_like(A)
    a = A[:, :-1]
    L[:, :-1] = np.where(a > 0, np.minimum(rp[:, 1:], rm[:, :-1]), np.minimum(rp[:, :-1], rm[:, 1:])) * a
    Ll = np.zeros_like(L)
    Ll[:, 1:] = L[:, :-1]
    U[:, 9] = rc + (Ll - L)
    return U
