"""Pins of the CPU oracle against facts that do not come from its own text:
textbook fluxes, the HLLD solver's exact contact, the GLM closed form, the
reference kernel's own outputs (golden vectors), closed-form CFL values,
conservation and symmetries.  Each check is chosen so that a plausible
restatement mistake (dropped term, wrong sign, swapped index) fails it."""
import os

import numpy as np
import pytest

from helpers import (NV, physical_flux_x, physical_flux_y, prim_to_cons_np, random_prim, rel_linf,
                     smooth_state, transpose_state)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ HLLD
def test_hlld_supersonic_equals_physical_flux(orc):
    """SL >= 0 picks F(W_L); SR <= 0 picks F(W_R) (physics.py:177-198)."""
    rng = np.random.default_rng(1)
    n = 2000
    WL = random_prim(rng, n, vmax=0.5)
    WL[4] = rng.uniform(-1, 1, n)
    WR = random_prim(rng, n, vmax=0.5)
    WR[4] = WL[4]
    gamma = 5.0 / 3.0
    for sign in (+1.0, -1.0):
        a, b = WL.copy(), WR.copy()
        a[1] += sign * 25.0
        b[1] += sign * 25.0
        F, smax = orc.hlld_flux(a, b, gamma, 0.0)
        up = a if sign > 0 else b
        ref = physical_flux_x(up, gamma, up[4])
        np.testing.assert_allclose(F, ref, rtol=1e-13, atol=1e-12)
        assert smax > 20.0


def test_hlld_y_supersonic_equals_physical_y_flux(orc):
    rng = np.random.default_rng(2)
    n = 1000
    W = random_prim(rng, n, vmax=0.5)
    W[2] += 30.0
    F, _ = orc.hlld_flux(W, W, 1.4, 0.0, direction=1)
    np.testing.assert_allclose(F, physical_flux_y(W, 1.4, W[5]), rtol=1e-13, atol=1e-12)


def test_hlld_consistency_uniform_state(orc):
    """F(W, W) = F(W) with the GLM fluxes psi, ch^2 Bx."""
    rng = np.random.default_rng(3)
    W = random_prim(rng, 4000)
    ch = 7.5
    F, _ = orc.hlld_flux(W, W, 5.0 / 3.0, ch)
    ref = physical_flux_x(W, 5.0 / 3.0, W[4], psi_flux=W[8], bx_flux=ch * ch * W[4])
    np.testing.assert_allclose(F, ref, rtol=1e-10, atol=1e-10)


def test_hlld_glm_face_solution(orc):
    """Upwind solution of the 2x2 GLM system (physics.py:76-82):
    psi* = (psiL+psiR)/2 - ch (BxR-BxL)/2, Bx* = (BxL+BxR)/2 - (psiR-psiL)/(2 ch)."""
    rng = np.random.default_rng(4)
    WL, WR = random_prim(rng, 3000), random_prim(rng, 3000)
    ch = 3.25
    F, _ = orc.hlld_flux(WL, WR, 5.0 / 3.0, ch)
    psi_f = 0.5 * (WL[8] + WR[8]) - 0.5 * ch * (WR[4] - WL[4])
    bx_f = 0.5 * (WL[4] + WR[4]) - (WR[8] - WL[8]) / (2 * ch)
    np.testing.assert_allclose(F[4], psi_f, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(F[8], ch * ch * bx_f, rtol=1e-13, atol=1e-13)


def test_hlld_exact_moving_contact(orc):
    """An isolated contact (p, v, B continuous; rho, C jump) is resolved
    exactly: the flux is the upwind physical flux."""
    rng = np.random.default_rng(5)
    n = 2000
    WL = random_prim(rng, n)
    WR = WL.copy()
    WR[0] = rng.uniform(0.1, 4.0, n)
    WR[9] = rng.uniform(0, 1, n)
    WL[8] = WR[8] = 0.0
    ch = 10.0
    F, _ = orc.hlld_flux(WL, WR, 5.0 / 3.0, ch)
    up = np.where(WL[1] >= 0, WL, WR)
    ref = physical_flux_x(up, 5.0 / 3.0, WL[4], psi_flux=0.0, bx_flux=ch * ch * WL[4])
    np.testing.assert_allclose(F, ref, rtol=1e-9, atol=1e-9)


def test_hlld_rotation_matches_y_physics(orc):
    """A y face of W equals the x face of the rotated state, rotated back."""
    rng = np.random.default_rng(6)
    WL, WR = random_prim(rng, 1500), random_prim(rng, 1500)
    perm = [0, 2, 1, 3, 5, 4, 6, 7, 8, 9]
    Fy, sy = orc.hlld_flux(WL, WR, 1.4, 2.0, direction=1)
    Fx, sx = orc.hlld_flux(WL[perm], WR[perm], 1.4, 2.0, direction=0)
    np.testing.assert_array_equal(Fy, Fx[perm])
    assert sx == sy


def test_hlld_matches_reference_kernel(orc):
    """Bit-for-bit against the reference's own HLLD kernel source
    (physics.py:43-275), compiled by hipcc with contraction off and run on
    the GPU to produce tests/golden/hlld_ref_vectors.npz
    (tests/golden/make_hlld_vectors.py)."""
    path = os.path.join(GOLDEN, "hlld_ref_vectors.npz")
    if not os.path.exists(path):
        pytest.fail("golden vectors missing: run tests/golden/make_hlld_vectors.py on a GPU")
    z = np.load(path, allow_pickle=False)
    for case in sorted({k.split("__")[0] for k in z.files}):
        F, smax = orc.hlld_flux(z[f"{case}__WL"], z[f"{case}__WR"], float(z[f"{case}__gamma"]),
                                float(z[f"{case}__ch"]))
        np.testing.assert_array_equal(F, z[f"{case}__F"], err_msg=case)
        assert smax == float(np.max(z[f"{case}__smax"])), case


# ----------------------------------------------------------------- MUSCL
def test_muscl_piecewise_constant_and_extrema(orc):
    W = np.ones((NV, 12, 9))
    W[:, 5:, :] = 2.0
    WL, WR = orc.muscl_x(W)
    # faces away from the jump reproduce the cell value exactly
    np.testing.assert_array_equal(WL[:, :2], 1.0)
    np.testing.assert_array_equal(WR[:, 6:], 2.0)
    # an isolated extremum gets zero slope
    W2 = np.ones((NV, 8, 5))
    W2[:, 4] = 3.0
    WL2, WR2 = orc.muscl_x(W2)
    np.testing.assert_array_equal(WL2[:, 3], 3.0)   # face 3 left state = cell 4 (the peak)
    np.testing.assert_array_equal(WR2[:, 2], 3.0)   # face 2 right state = cell 4


def test_muscl_linear_data_gives_midpoints(orc):
    nxt, nyt = 16, 7
    i = np.arange(nxt)[:, None] * np.ones(nyt)[None, :]
    W = np.stack([3.0 + 0.25 * (v + 1) * i for v in range(NV)])
    WL, WR = orc.muscl_x(W)
    for v in range(NV):
        mid = 3.0 + 0.25 * (v + 1) * (np.arange(nxt - 3) + 1.5)
        np.testing.assert_allclose(WL[v, :, 0], mid, rtol=1e-14)
        np.testing.assert_allclose(WR[v, :, 0], mid, rtol=1e-14)


def test_muscl_tvd_bounds_and_floors(orc):
    rng = np.random.default_rng(7)
    W = np.cumsum(rng.uniform(0, 1, (NV, 40, 6)), axis=1)
    WL, WR = orc.muscl_x(W)
    lo, hi = W[:, 1:-2], W[:, 2:-1]
    assert np.all(WL >= lo - 1e-12) and np.all(WL <= hi + 1e-12)
    assert np.all(WR >= lo - 1e-12) and np.all(WR <= hi + 1e-12)
    Wn = -np.ones((NV, 8, 8))
    WLn, WRn = orc.muscl_y(Wn)
    assert np.all(WLn[0] == 1e-12) and np.all(WRn[7] == 1e-12)
    assert np.all(WLn[1] == -1.0)


def test_muscl_y_is_transposed_x(orc):
    rng = np.random.default_rng(8)
    W = rng.uniform(0.5, 2, (NV, 11, 13))
    WLy, WRy = orc.muscl_y(W)
    WLx, WRx = orc.muscl_x(np.ascontiguousarray(np.transpose(W, (0, 2, 1))))
    np.testing.assert_array_equal(WLy, np.transpose(WLx, (0, 2, 1)))
    np.testing.assert_array_equal(WRy, np.transpose(WRx, (0, 2, 1)))


# --------------------------------------------------- conversions and CFL
def test_cons_prim_round_trip_and_pressure(orc):
    rng = np.random.default_rng(9)
    W = random_prim(rng, 5000)
    U = orc.prim_to_cons(W, 1.4)
    np.testing.assert_allclose(U, prim_to_cons_np(W, 1.4), rtol=1e-14, atol=1e-15)
    W2 = orc.cons_to_prim(U, 1.4)
    np.testing.assert_allclose(W2, W, rtol=1e-11, atol=1e-12)


def test_compute_dt_brio_wu_closed_form(orc):
    """Brio-Wu at t = 0: cf_max = sqrt(gamma p/rho + B^2/rho) on the right
    state = sqrt(14.1), ch = 1.5 cf_max, dt = cfl dx / ch (v = 0)."""
    from amhd_ics import brio_wu

    cfg, W = brio_wu(800, cfl=0.4)
    dt, ch = orc.compute_dt(W, cfg)
    cf = np.sqrt(14.1)
    assert ch == pytest.approx(1.5 * cf, rel=1e-15)
    assert dt == pytest.approx(0.4 * cfg.dx / (1.5 * cf), rel=1e-14)
    assert dt == pytest.approx(8.877e-5, rel=1e-3)


def test_apply_bc_images(orc):
    from config import Config

    for bcx, bcy in (("periodic", "periodic"), ("extrapolation", "periodic"), ("extrapolation", "outflow")):
        cfg = Config(nx=6, ny=5, bc_x_type=bcx, bc_y_type=bcy)
        U = np.arange(NV * 10 * 9, dtype=np.float64).reshape(NV, 10, 9) + 1.0
        G = orc.apply_bc(U, cfg)
        np.testing.assert_array_equal(G[:, 2:-2, 2:-2], U[:, 2:-2, 2:-2])
        for gi in range(10):
            for gj in range(9):
                i, j = gi - 2, gj - 2
                si = (i % 6) if bcx == "periodic" else min(max(i, 0), 5)
                sj = (j % 5) if bcy == "periodic" else min(max(j, 0), 4)
                want = U[:, si + 2, sj + 2].copy()
                if bcx != "periodic" and (i < 0 or i >= 6):
                    want[8] = 0.0
                np.testing.assert_array_equal(G[:, gi, gj], want)


# --------------------------------------------------------------- the step
def _periodic_cfg(nx, ny, powell=False):
    from config import Config

    return Config(nx=nx, ny=ny, x_min=0, x_max=1, y_min=0, y_max=ny / nx, gamma=5 / 3, cfl=0.3,
                  bc_x_type="periodic", bc_y_type="periodic", powell_source=powell,
                  B_transverse=0.3 if powell else 0.0)


def test_step_conserves_with_periodic_boundaries(orc):
    cfg = _periodic_cfg(32, 24)
    U = prim_to_cons_np(smooth_state(32, 24, seed=3), cfg.gamma)
    s0 = U[:, 2:-2, 2:-2].sum(axis=(1, 2))
    Uc = U
    for _ in range(5):
        dt, ch = orc.compute_dt(orc.cons_to_prim(Uc, cfg.gamma), cfg)
        Uc, _ = orc.step(Uc, cfg, dt, ch)
    s1 = Uc[:, 2:-2, 2:-2].sum(axis=(1, 2))
    for v in (0, 1, 2, 3, 4, 5, 6, 7, 9):
        assert abs(s1[v] - s0[v]) <= 1e-12 * max(1.0, np.abs(U[v]).sum()), v


def test_step_preserves_uniform_state(orc):
    cfg = _periodic_cfg(16, 16)
    W = np.zeros((NV, 20, 20))
    W[0], W[1], W[2], W[4], W[5], W[7], W[9] = 1.3, 0.4, -0.7, 0.6, -0.2, 0.9, 0.25
    U = prim_to_cons_np(W, cfg.gamma)
    dt, ch = orc.compute_dt(W, cfg)
    U1, _ = orc.step(U, cfg, dt, ch)
    assert np.max(np.abs(rel_linf(U1[[0, 1, 2, 3, 4, 5, 6, 7, 9]], U[[0, 1, 2, 3, 4, 5, 6, 7, 9]]))) < 1e-14


@pytest.mark.parametrize("powell", [False, True])
def test_step_transpose_symmetry(orc, powell):
    """Step(T U) == T Step(U) bit for bit on a square periodic grid (x and y
    sweeps enter R symmetrically when dx == dy)."""
    cfg = _periodic_cfg(20, 20, powell=powell)
    U = prim_to_cons_np(smooth_state(20, 20, seed=11), cfg.gamma)
    dt, ch = orc.compute_dt(orc.cons_to_prim(U, cfg.gamma), cfg)
    A, _ = orc.step(U, cfg, dt, ch)
    B, _ = orc.step(transpose_state(U), cfg, dt, ch)
    np.testing.assert_array_equal(transpose_state(A), B)


def test_boundary_flux_uniform_flow(orc):
    """Uniform flow: both boundary columns carry the same flux, so the net
    boundary energy flux is zero (the closed form of a single column is
    pinned in test_oracle_pins_step.py)."""
    from config import Config

    cfg = Config(nx=10, ny=6, x_min=0, x_max=1, y_min=0, y_max=0.6, bc_x_type="extrapolation")
    W = np.zeros((NV, 14, 10))
    W[0], W[1], W[4], W[5], W[7] = 1.0, 0.5, 0.3, 0.2, 1.0
    U = prim_to_cons_np(W, cfg.gamma)
    assert orc.boundary_flux(U, cfg) == pytest.approx(0.0, abs=1e-14)
