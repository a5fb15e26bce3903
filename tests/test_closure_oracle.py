"""Pins of the closure convolution oracle against an independent library
(torch.nn.functional.conv2d and autograd on the CPU, float64) and against
the adjoint identity."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

import closure_oracle as co


def torch_conv(x, W, b, cin, cout):
    """NHWC x; W[co, tap*cin+ci] -> conv2d with zero pad in x, circular in y."""
    xt = torch.as_tensor(x).permute(0, 3, 1, 2)                      # N C H W
    xt = F.pad(xt, (1, 1, 0, 0), mode="circular")                     # y (W axis) periodic
    xt = F.pad(xt, (0, 0, 1, 1))                                      # x (H axis) zeros
    w = torch.as_tensor(W[:, :9 * cin]).reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
    return F.conv2d(xt, w, torch.as_tensor(b)).permute(0, 2, 3, 1)


@pytest.mark.parametrize("cin,cout", [(16, 32), (32, 32), (32, 16)])
def test_fwd_and_backward_match_torch(cin, cout):
    rng = np.random.default_rng(cin + cout)
    x = rng.standard_normal((2, 7, 6, cin))
    W = np.zeros((cout, co.kpad(cin)))
    W[:, :9 * cin] = rng.standard_normal((cout, 9 * cin)) * 0.2
    b = rng.standard_normal(cout)
    y = co.conv_fwd(x, W, b, relu=False, bf16=False)
    xt = torch.tensor(x, requires_grad=True)
    Wt = torch.tensor(W, requires_grad=True)
    bt = torch.tensor(b, requires_grad=True)
    yt = torch_conv(xt, Wt, bt, cin, cout)
    np.testing.assert_allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    g = rng.standard_normal(y.shape)
    yt.backward(torch.tensor(g))
    np.testing.assert_allclose(co.conv_bwd_data(g, W, cin, bf16=False), xt.grad.numpy(), rtol=1e-12, atol=1e-12)
    dW, db = co.conv_bwd_weight(x, g, bf16=False)
    np.testing.assert_allclose(dW, Wt.grad.numpy(), rtol=1e-12, atol=1e-11)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_adjoint_identity():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 5, 9, 16))
    W = rng.standard_normal((32, co.kpad(16)))
    y = rng.standard_normal((1, 5, 9, 32))
    lhs = np.sum(co.conv_fwd(x, W, np.zeros(32), relu=False, bf16=False) * y)
    rhs = np.sum(x * co.conv_bwd_data(y, W, 16, bf16=False))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.0e38, 1e-40, np.inf], dtype=np.float32)
    r = co.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125   # ties to even, then up
    assert np.isinf(r[5]) and np.all(np.isfinite(r[:5]))
    t = torch.tensor(v).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(r, t)


def test_network_gradient_finite_differences():
    """The oracle's network backward (float64, no rounding) against central
    differences of the closure loss."""
    rng = np.random.default_rng(11)
    n, nx, ny = 1, 6, 5
    x = np.zeros((n, nx, ny, 16))
    x[..., :9] = rng.standard_normal((n, nx, ny, 9))
    Ws, bs = [], []
    for cin, cout, _ in co.LAYERS:
        W = np.zeros((cout, co.kpad(cin)))
        W[:, :9 * cin] = rng.standard_normal((cout, 9 * cin)) * 0.15
        Ws.append(W)
        bs.append(rng.standard_normal(cout) * 0.1)
    R = rng.standard_normal((n, 9, nx, ny))
    Cc = rng.uniform(0, 1, (n, nx, ny))
    fscale, rstd, dx, dy = rng.uniform(0.5, 2, 9), rng.uniform(0.5, 2, 9), 0.3, 0.2
    out, acts = co.net_forward(x, Ws, bs, bf16=False)
    gW, gb = co.net_backward(acts, co.loss_grad_out(out, R, fscale, rstd, dx, dy, Cc), Ws, bf16=False)

    def L():
        return co.loss(co.net_forward(x, Ws, bs, bf16=False)[0], R, fscale, rstd, dx, dy, Cc)

    h = 1e-6
    for li, (cin, cout, _) in enumerate(co.LAYERS):
        for (a, b) in ((0, 0), (cout - 1, 9 * cin - 1), (3, 40)):
            old = Ws[li][a, b]
            Ws[li][a, b] = old + h
            lp = L()
            Ws[li][a, b] = old - h
            lm = L()
            Ws[li][a, b] = old
            assert gW[li][a, b] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-9), (li, a, b)
        old = bs[li][1]
        bs[li][1] = old + h
        lp = L()
        bs[li][1] = old - h
        lm = L()
        bs[li][1] = old
        assert gb[li][1] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-9)


def _state(rng, n, nx, ny, gamma):
    """Random conserved states with rho in [0.5, 2] and p in [0.2, 1]."""
    U = np.zeros((n, 10, nx, ny))
    rho = rng.uniform(0.5, 2.0, (n, nx, ny))
    v = rng.uniform(-1, 1, (3, n, nx, ny))
    B = rng.uniform(-1, 1, (3, n, nx, ny))
    p = rng.uniform(0.2, 1.0, (n, nx, ny))
    U[:, 0] = rho
    U[:, 1:4] = np.moveaxis(rho * v, 0, 1)
    U[:, 4:7] = np.moveaxis(B, 0, 1)
    U[:, 7] = p / (gamma - 1) + 0.5 * rho * (v ** 2).sum(0) + 0.5 * (B ** 2).sum(0)
    U[:, 9] = rho * rng.uniform(0, 1, (n, nx, ny))
    return U


def _admissible(U0, D, t, gamma, keep):
    U = U0 + t * D
    p = lambda A: (gamma - 1) * (A[7] - 0.5 * (A[1] ** 2 + A[2] ** 2 + A[3] ** 2) / A[0]
                                 - 0.5 * (A[4] ** 2 + A[5] ** 2 + A[6] ** 2))
    return U[0] > 0 and U[0] >= keep * U0[0] and p(U) >= keep * p(U0)


def test_limiter_theta_definition():
    """theta = min over faces of the largest grid t with U0 + 4 t D_face
    admissible (checked cell by cell with an independent pressure), and the
    next grid value up fails for some face; theta = 1 for a vanishing
    correction."""
    rng = np.random.default_rng(21)
    gamma, dx, dy, dt, keep = 5.0 / 3.0, 0.1, 0.08, 0.01, 0.5
    U0 = _state(rng, 1, 8, 6, gamma)
    dF = rng.standard_normal((1, 8, 6, 32)) * 0.7
    fs = rng.uniform(0.5, 2.0, 9)
    th = co.limiter_theta(U0, dF, fs, dt, dx, dy, gamma, keep)
    n9 = 9
    grid = co.THETAS + [0.0]
    n_lim = 0
    for i in range(8):
        for j in range(6):
            faces = []
            fx = dF[0, :, :, :n9] * fs
            fy = dF[0, :, :, n9:18] * fs
            faces.append(fx[i, j] * dt / dx if i > 0 else np.zeros(9))
            faces.append(-fx[i + 1, j] * dt / dx if i + 1 < 8 else np.zeros(9))
            faces.append(fy[i, j] * dt / dy)
            faces.append(-fy[i, (j + 1) % 6] * dt / dy)
            t = th[0, i, j]
            assert t in grid
            for Df in faces:
                D = np.zeros(10)
                D[co.FEAT] = Df
                assert t == 0.0 or _admissible(U0[0, :, i, j], 4 * t * D, 1.0, gamma, keep)
            if t < 1.0:
                n_lim += 1
                up = grid[grid.index(t) - 1]
                fails = []
                for Df in faces:
                    D = np.zeros(10)
                    D[co.FEAT] = Df
                    fails.append(not _admissible(U0[0, :, i, j], 4 * up * D, 1.0, gamma, keep))
                assert any(fails)
    assert 0 < n_lim < 48
    assert np.all(co.limiter_theta(U0, dF * 1e-9, fs, dt, dx, dy, gamma, keep) == 1.0)


def test_limited_correction_keeps_the_bound():
    """Every cell of U0 + dt * S stays admissible when S is built from face
    weights in [0, min(theta_L, theta_R)] -- the limiter's guarantee --
    while the unlimited correction breaks the bound somewhere."""
    rng = np.random.default_rng(23)
    gamma, dx, dy, dt, keep = 5.0 / 3.0, 0.1, 0.08, 0.02, 0.5
    U0 = _state(rng, 2, 9, 7, gamma)
    C = co.concentration(U0)
    dF = rng.standard_normal((2, 9, 7, 32)) * 4.0
    fs = rng.uniform(0.5, 2.0, 9)
    th = co.limiter_theta(U0, dF, fs, dt, dx, dy, gamma, keep)
    assert np.any(th < 1)
    S0 = co.source_from_flux(dF[..., :18] * np.concatenate([fs, fs]), dx, dy, C)
    W0 = U0.copy()
    W0[:, co.FEAT] += dt * np.transpose(S0, (0, 3, 1, 2))
    assert not np.all(co.admissible(U0, W0, gamma, keep))
    for trial in range(3):
        # random face weights below the face limiter: scale dF face by face
        w = rng.uniform(0, 1, (2, 9, 7, 32)).astype(np.float64)
        S = co.limited_source(dF * w, th, np.ones_like(th), fs, dx, dy, C)
        W = U0.copy()
        W[:, co.FEAT] += dt * np.transpose(S, (0, 3, 1, 2))
        assert np.all(co.admissible(U0, W, gamma, keep)), trial


def test_mixing_limiter_keeps_concentration_bounds():
    """C of U0 + dt * limited_source stays within the min / max of C over
    each cell and its face neighbours in U0 (a discrete maximum principle);
    without the mixing limiter it does not."""
    rng = np.random.default_rng(24)
    gamma, dx, dy, dt, keep = 5.0 / 3.0, 0.1, 0.08, 0.02, 0.5
    U0 = _state(rng, 2, 10, 8, gamma)
    U0[:, 9] = U0[:, 0] * (rng.uniform(0, 1, (2, 10, 8)) > 0.5)     # pure cells, C in {0, 1}
    U0[:, 9, 3:5] = U0[:, 0, 3:5] * 0.4                              # and a mixed band
    C = co.concentration(U0)
    dF = rng.standard_normal((2, 10, 8, 32)) * 1.5
    fs = rng.uniform(0.5, 2.0, 9)
    th = co.limiter_theta(U0, dF, fs, dt, dx, dy, gamma, keep)
    tm = co.limiter_theta_mix(U0, dF, th, fs, dt, dx, dy)
    assert np.any(tm < 1) and np.any(tm > 0)
    lo, hi = co.local_bounds(C)

    def c_after(theta_mix):
        S = co.limited_source(dF, th, theta_mix, fs, dx, dy, C)
        W = U0.copy()
        W[:, co.FEAT] += dt * np.transpose(S, (0, 3, 1, 2))
        return W[:, 9] / W[:, 0]

    c1 = c_after(tm)
    assert np.all(c1 >= lo - 1e-12) and np.all(c1 <= hi + 1e-12)
    # with the bounds widened by a second state (the state before the step)
    Ub = U0.copy()
    Ub[:, 9] = U0[:, 0] * rng.uniform(0, 1, (2, 10, 8))
    tb = co.limiter_theta_mix(U0, dF, th, fs, dt, dx, dy, Ub)
    assert np.all(tb >= tm) and np.any(tb > tm)
    blo, bhi = co.local_bounds(co.concentration(Ub))
    cb = c_after(tb)
    assert np.all(cb >= np.minimum(lo, blo) - 1e-12) and np.all(cb <= np.maximum(hi, bhi) + 1e-12)
    c0 = c_after(np.ones_like(tm))
    assert np.any(c0 < lo - 1e-6) or np.any(c0 > hi + 1e-6)
    # the donor-cell part alone (no mixing flux) keeps the bounds for any theta
    c2 = c_after(np.zeros_like(tm))
    assert np.all(c2 >= lo - 1e-12) and np.all(c2 <= hi + 1e-12)


def test_limited_source_reduces_and_conserves():
    rng = np.random.default_rng(22)
    dF = rng.standard_normal((2, 7, 5, 32))
    fs = rng.uniform(0.5, 2.0, 9)
    C = rng.uniform(0, 1, (2, 7, 5))
    one = np.ones((2, 7, 5))
    S1 = co.limited_source(dF, one, one, fs, 0.3, 0.2, C)
    np.testing.assert_allclose(S1, co.source_from_flux(dF[..., :18] * np.concatenate([fs, fs]), 0.3, 0.2, C),
                               rtol=1e-13, atol=1e-13)
    th = rng.choice(co.THETAS + [0.0], (2, 7, 5))
    tm = rng.choice(co.THETAS + [0.0], (2, 7, 5))
    S = co.limited_source(dF, th, tm, fs, 0.3, 0.2, C)
    np.testing.assert_allclose(S.sum(axis=(1, 2)), 0.0, atol=1e-12)
    zero = np.zeros((2, 7, 5))
    assert np.all(co.limited_source(dF, zero, zero, fs, 0.3, 0.2, C) == 0.0)


def _profile(c_row, ny=3, rho=None):
    """One instance whose every y row holds the x profile c_row."""
    nx = len(c_row)
    U = np.zeros((1, 10, nx, ny))
    U[0, 0] = 1.0 if rho is None else np.asarray(rho, dtype=np.float64)[:, None]
    U[0, 9] = U[0, 0] * np.asarray(c_row, dtype=np.float64)[:, None]
    return U


@pytest.mark.parametrize("kappa,expect", [(0.1, [1.0, 0.8, 0.2, 0.0]), (1.0, [1.0, 1.0, 0.0, 0.0])])
def test_sharpen_four_cell_profile(kappa, expect):
    """rho = 1, C = (1, 0.75, 0.25, 0) along x.  The outer faces move
    nothing (their flux would push C of the end cells, already at their
    neighbourhood's max and min, beyond it); the middle face's flux
    kappa * (0.25 - 0.75) = -kappa / 2 moves rhoC from cell 2 to cell 1.
    Cell 1 may gain Q+ = 0.25 and its inflow is kappa / 2; cell 2 may lose
    Q- = 0.25 and its outflow is kappa / 2.  kappa = 0.1: both ratios are
    above 1, the whole 0.05 moves: (1, 0.8, 0.2, 0).  kappa = 1: both are
    0.5, a quarter moves: the sharp step (1, 1, 0, 0)."""
    U = _profile([1.0, 0.75, 0.25, 0.0])
    out = co.sharpen_x(U, None, kappa)
    np.testing.assert_allclose(out[0, 9] / out[0, 0], np.repeat(np.asarray(expect)[:, None], 3, axis=1),
                               rtol=0, atol=1e-15)
    # the mirrored profile sharpens the mirrored way
    outm = co.sharpen_x(_profile([0.0, 0.25, 0.75, 1.0]), None, kappa)
    np.testing.assert_allclose(outm[0, 9, :, 0], np.asarray(expect)[::-1], rtol=0, atol=1e-15)


def test_sharpen_bounds_conservation_and_identity():
    """Random states: C stays within the min / max over each cell and its
    face neighbours (of the state, and of a second state when given), where
    the unlimited fluxes (every ratio 1) leave them; rhoC is conserved in
    every row (x faces only); every other plane is untouched; kappa = 0 is
    the identity; a y-uniform state stays y-uniform; the pass narrows a
    smeared interface (sum of C (1 - C) falls)."""
    rng = np.random.default_rng(41)
    U = _state(rng, 2, 12, 7, 5.0 / 3.0)
    Ub = _state(rng, 2, 12, 7, 5.0 / 3.0)
    for kappa in (0.05, 0.5, 5.0):
        for b in (None, Ub):
            out = co.sharpen_x(U, b, kappa)
            C1 = out[:, 9] / out[:, 0]
            lo, hi = co.local_bounds(co.concentration(U))
            if b is not None:
                blo, bhi = co.local_bounds(co.concentration(b))
                lo, hi = np.minimum(lo, blo), np.maximum(hi, bhi)
            assert np.all(C1 >= lo - 1e-13) and np.all(C1 <= hi + 1e-13)
            np.testing.assert_allclose(out[:, 9].sum(axis=1), U[:, 9].sum(axis=1), rtol=1e-13)
            np.testing.assert_array_equal(np.delete(out, 9, axis=1), np.delete(U, 9, axis=1))
            assert np.any(out[:, 9] != U[:, 9])
    # the unlimited fluxes break the bounds at kappa = 5
    C = co.concentration(U)
    A = 5.0 * 0.5 * (U[:, 0, :-1] + U[:, 0, 1:]) * (C[:, 1:] - C[:, :-1])
    rc = U[:, 9].copy()
    rc[:, :-1] -= A
    rc[:, 1:] += A
    lo, hi = co.local_bounds(C)
    assert np.any(rc / U[:, 0] > hi + 1e-6) or np.any(rc / U[:, 0] < lo - 1e-6)
    np.testing.assert_array_equal(co.sharpen_x(U, Ub, 0.0), U)
    x = np.arange(16)
    Up = _profile(0.5 * (1 - np.tanh((x - 7.5) / 2.5)), ny=5, rho=1 + x / 16)
    out = co.sharpen_x(Up, None, 0.2)
    assert np.all(out[0, 9] == out[0, 9, :, :1])
    c0, c1 = co.concentration(Up)[0, :, 0], co.concentration(out)[0, :, 0]
    assert (c1 * (1 - c1)).sum() < (c0 * (1 - c0)).sum()


def test_donor_cell_flux():
    """F_rhoC = C_donor F_rho + G with the donor upstream of the mass flux."""
    fx = np.zeros((1, 3, 2, 9))
    fy = np.zeros((1, 3, 2, 9))
    C = np.array([[[0.1, 0.2], [0.3, 0.4], [0.5, 0.6]]])
    fx[0, 1, 0, 0], fx[0, 2, 1, 0] = 2.0, -3.0       # face 1-1/2 from cell 0; face 2-1/2 from cell 2
    fx[0, 1, 0, 8] = 0.25
    fy[0, 1, 0, 0], fy[0, 2, 1, 0] = 1.0, -1.0       # face j=0-1/2 from cell j=1 (periodic); face 1-1/2 from j=1
    ox, oy, _, _ = co.mix_fluxes(fx, fy, C)
    assert ox[0, 1, 0, 8] == 0.25 + 0.1 * 2.0
    assert ox[0, 2, 1, 8] == 0.6 * -3.0
    assert oy[0, 1, 0, 8] == 0.4 * 1.0
    assert oy[0, 2, 1, 8] == 0.6 * -1.0
