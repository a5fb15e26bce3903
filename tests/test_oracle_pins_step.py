"""Pins of the oracle's step-level functions against facts that do not come
from its own text: the Powell source (solver.py:242-266) on fields whose
divergence is known in closed form, the boundary energy flux
(solver.py:225-240) against the textbook energy flux and against the energy
balance of an RMI run, the scalar bounds (solver.py:199-204) case by case,
the SSP-RK3 stage weights (solver.py:319-327) through the scheme's temporal
order, and the host diagnostics (solver.py:335-444) on profiles whose
mixing width, mixedness, mode amplitude and enstrophy are known.

Where a pin could pass for a wrong implementation, the test also runs a
planted mistake (composed in the test from oracle parts, on the CPU) and
shows that the pin rejects it."""
import numpy as np
import pytest

from helpers import NV, physical_flux_x, prim_to_cons_np, smooth_state


# ------------------------------------------------------------------ Powell
def _linear_b_state(nx, ny, cfg, b_slope, d_slope):
    """Ghosted primitive state with Bx = 0.4 + b (x - x_min), By = -0.3 +
    d (y - y_min), everything else uniform and nonzero.  Its cell-centred
    divergence is b + d exactly in every cell whose stencil is interior."""
    x = cfg.x_min + (np.arange(nx + 4) - 2 + 0.5) * cfg.dx
    y = cfg.y_min + (np.arange(ny + 4) - 2 + 0.5) * cfg.dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    W = np.zeros((NV, nx + 4, ny + 4))
    W[0], W[1], W[2], W[3] = 1.3, 0.35, -0.25, 0.15
    W[4] = 0.4 + b_slope * (X - cfg.x_min)
    W[5] = -0.3 + d_slope * (Y - cfg.y_min)
    W[6], W[7], W[8], W[9] = 0.2, 0.9, 0.0, 0.4
    return W


def _powell_cfg(nx, ny, on):
    from config import Config

    return Config(nx=nx, ny=ny, x_min=0.0, x_max=1.0, y_min=0.0, y_max=ny / nx, gamma=5 / 3,
                  bc_x_type="extrapolation", bc_y_type="extrapolation", powell_source=on,
                  B_transverse=0.3 if on else 0.0)


def expected_powell(W, divb):
    """S = -divB (0, Bx, By, Bz, vx, vy, vz, v.B, 0, 0): the 8-wave source of
    Powell et al. (1999), solver.py:256-264."""
    S = np.zeros_like(W)
    S[1], S[2], S[3] = -divb * W[4], -divb * W[5], -divb * W[6]
    S[4], S[5], S[6] = -divb * W[1], -divb * W[2], -divb * W[3]
    S[7] = -divb * (W[1] * W[4] + W[2] * W[5] + W[3] * W[6])
    return S


@pytest.mark.parametrize("b_slope,d_slope", [(0.8, 0.0), (0.0, -1.1), (0.6, 0.9)])
def test_powell_source_closed_form(orc, b_slope, d_slope):
    """R(Powell on) - R(Powell off) = S(divB = b + d) in every cell whose
    divergence stencil stays inside the interior.  A sign, a factor 1/2, a
    swapped component (v for B) or a missing y term fails."""
    nx, ny = 16, 12
    on, off = _powell_cfg(nx, ny, True), _powell_cfg(nx, ny, False)
    W = _linear_b_state(nx, ny, on, b_slope, d_slope)
    U = prim_to_cons_np(W, on.gamma)
    ch = 3.0
    R_on, _ = orc.compute_rhs(U, on, ch)
    R_off, _ = orc.compute_rhs(U, off, ch)
    S = expected_powell(W, b_slope + d_slope)
    inner = (slice(3, nx + 1), slice(3, ny + 1))   # interior cells 1..n-2
    scale = np.max(np.abs(S[:, inner[0], inner[1]])) + 1e-300
    for v in range(NV):
        got = (R_on - R_off)[v][inner]
        want = S[v][inner]
        assert np.max(np.abs(got - want)) <= 1e-12 * scale, v
    for v in (0, 8, 9):
        assert np.array_equal((R_on - R_off)[v][inner], np.zeros_like(S[v][inner])), v


def test_powell_gate_needs_transverse_field(orc):
    """solver.py:304: the source is added only if powell_source and
    B_transverse > 0."""
    import oracle

    cfg = _powell_cfg(8, 8, True)
    assert oracle.params(cfg).powell == 1
    cfg.B_transverse = 0.0
    assert oracle.params(cfg).powell == 0


# ------------------------------------------------------------ boundary flux
def _column_state(ny, seed, gamma):
    rng = np.random.default_rng(seed)
    W = np.zeros((NV, ny))
    W[0] = rng.uniform(0.5, 2.0, ny)
    W[1:4] = rng.uniform(-1.0, 1.0, (3, ny))
    W[4:7] = rng.uniform(-0.8, 0.8, (3, ny))
    W[7] = rng.uniform(0.3, 2.0, ny)
    W[8] = rng.uniform(-0.1, 0.1, ny)
    W[9] = rng.uniform(0.0, 1.0, ny)
    return W


def test_boundary_flux_textbook_energy_flux(orc):
    """Net boundary energy flux = sum_j F_E(first interior column) dy -
    sum_j F_E(last interior column) dy, with F_E = (E + p + B^2/2) vx -
    Bx (v.B) the textbook ideal-MHD energy flux.  The two columns carry
    different non-uniform states and every other column (ghosts included)
    holds unrelated values, so a wrong column, a wrong sign, a missing dy or
    a dropped magnetic term fails."""
    from config import Config

    nx, ny = 7, 9
    cfg = Config(nx=nx, ny=ny, x_min=0.0, x_max=0.7, y_min=0.0, y_max=0.45, gamma=1.4,
                 bc_x_type="extrapolation")
    rng = np.random.default_rng(5)
    W = np.zeros((NV, nx + 4, ny + 4))
    for i in range(nx + 4):
        W[:, i, :] = _column_state(ny + 4, 100 + i, cfg.gamma)
    U = prim_to_cons_np(W, cfg.gamma)
    first, last = W[:, 2, 2:-2], W[:, nx + 1, 2:-2]
    fl = physical_flux_x(first, cfg.gamma, first[4])[7]
    fr = physical_flux_x(last, cfg.gamma, last[4])[7]
    want = float(np.sum(fl) * cfg.dy - np.sum(fr) * cfg.dy)
    got = orc.boundary_flux(U, cfg)
    assert abs(want) > 0.1
    assert got == pytest.approx(want, rel=1e-13, abs=1e-14)
    # periodic x: no physical boundary, no flux (solver.py:229)
    cfg.bc_x_type = "periodic"
    assert orc.boundary_flux(U, cfg) == 0.0
    del rng


def test_boundary_flux_closes_the_energy_balance(orc):
    """The reference's published energy check (solver.py:507-511, results
    PDF p.1): on an RMI run the raw energy drift (E(t) - E0) / E0 is large,
    because the post-shock gas flows in through the x boundary, and the
    drift corrected by the accumulated boundary flux is close to 0."""
    from amhd_ics import rmi_uniform_bx, cons_from_prim_host

    cfg, W = rmi_uniform_bx(96, 32)
    cfg.t_end = 1e30
    U0 = cons_from_prim_host(W, cfg.gamma)
    e = lambda U: float(np.sum(U[7, 2:-2, 2:-2]) * cfg.dx * cfg.dy)
    e0 = e(U0)
    U, t, steps, _, cum = orc.run_loop(U0, cfg, 150)
    raw = (e(U) - e0) / e0
    corrected = (e(U) - e0 - cum) / e0
    assert steps == 150
    assert raw > 0.02                       # the inflow raised E by > 2 %
    assert abs(corrected) < 2e-3 * raw      # the boundary flux accounts for it
    # with the sign of the flux reversed the balance fails
    assert abs((e(U) - e0 + cum) / e0) > 1.5 * raw


# ------------------------------------------------------------ scalar bounds
def test_scalar_bounds_case_by_case(orc):
    """solver.py:199-204: rhoC = max(rho, 1e-12) * clip(rhoC / max(rho,
    1e-12), 0, 1).  C in [0, 1] keeps rhoC to within one rounding of the
    quotient; C > 1 gives rho (exactly); C < 0 gives 0; rho below the floor
    uses the floor, not rho; rho itself is never changed."""
    n = 8
    U = np.zeros((NV, n))
    U[0] = [1.0, 3.0, 0.7, 2.0, 5.0, 1e-15, -1.0, 1e-15]
    U[9] = [0.25, 1.0, 0.7, 2.5, -0.3, 3e-12, 0.5, 4e-16]
    U[1:9] = np.arange(8)[None, :] + 1.0
    out = orc.enforce_bounds(U)
    np.testing.assert_array_equal(out[:9], U[:9])
    rc = out[9]
    for k in (0, 1, 2):                       # 0 <= C <= 1
        assert abs(rc[k] - U[9, k]) <= 2.0 ** -52 * abs(U[9, k]) * 2
    assert rc[3] == 2.0                       # C = 1.25 -> rho
    assert rc[4] == 0.0 and not np.signbit(rc[4])   # C < 0 -> 0
    assert rc[5] == 1e-12                     # rho < floor, C > 1 -> floor
    assert rc[6] == 1e-12                     # negative rho -> floor
    assert abs(rc[7] - 4e-16) <= 1e-30        # rho < floor, C = 4e-4 of the floor


# ---------------------------------------------------------------- SSP-RK3
def _fixed_dt_run(orc, U, cfg, ch, T, n, stepper=None):
    dt = T / n
    for _ in range(n):
        if stepper is None:
            U, _ = orc.step(U, cfg, dt, ch, decay=1.0)
        else:
            U = stepper(U, dt)
    return U


def _order_cfg():
    from config import Config

    return Config(nx=16, ny=4, x_min=0.0, x_max=1.0, y_min=0.0, y_max=0.25, gamma=5 / 3,
                  bc_x_type="periodic", bc_y_type="periodic", powell_source=False, B_transverse=0.0,
                  glm_alpha=0.0)


def _checkerboard(cfg):
    """Cells alternating in x between two states A and B.  Every MUSCL
    stencil then has dm dp < 0, so each slope is 0 and stays 0 (the pattern
    keeps its period), the HLLD fan is pinned to [-ch, ch] and the contact
    side never changes: the semi-discrete system is the smooth ODE
    dA/dt = (F(B, A) - F(A, B)) / dx = -dB/dt, on which a third-order
    integrator shows its order cleanly (the limiter's switching on smooth
    profiles would blur it)."""
    A = np.array([1.0, 0.8, 0.1, -0.05, 0.6, 0.3, 0.1, 1.0, 0.0, 0.3])
    B = np.array([1.4, 0.7, -0.1, 0.05, 0.6, 0.2, -0.1, 1.3, 0.0, 0.7])
    W = np.zeros((NV, cfg.nx + 4, cfg.ny + 4))
    for i in range(cfg.nx + 4):
        W[:, i, :] = (A if i % 2 == 0 else B)[:, None]
    return prim_to_cons_np(W, cfg.gamma)


def _temporal_errors(orc, stepper=None):
    cfg = _order_cfg()
    U0 = _checkerboard(cfg)
    dt0, ch = orc.compute_dt(orc.cons_to_prim(U0, cfg.gamma), cfg)
    T = 8 * dt0
    ref = _fixed_dt_run(orc, U0, cfg, ch, T, 256)
    errs = []
    for n in (8, 16, 32):
        U = _fixed_dt_run(orc, U0, cfg, ch, T, n, stepper)
        errs.append(float(np.max(np.abs(U[:, 2:-2, 2:-2] - ref[:, 2:-2, 2:-2]))))
    return errs, (U0, cfg, ch)


def test_ssprk3_is_third_order_in_time(orc):
    """Shu-Osher SSP-RK3 (solver.py:319-327) is third order: on a fixed grid
    the error against a 32x smaller step falls about 8x per halving of dt
    (>= 6 required).  Any other set of stage weights loses an order or
    consistency."""
    errs, _ = _temporal_errors(orc)
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 6.0 <= r1 <= 12.0 and 6.0 <= r2 <= 12.0, (errs, r1, r2)


def _composed_step(orc, cfg, ch, weights):
    """One SSP-RK3-shaped step composed in the test from the oracle's
    compute_rhs, bounds and BC, with the given stage weights
    ((a2, b2), (a3, b3)): U1 = U0 + dt L(U0), U2 = a2 U0 + b2 (U1 + dt L(U1)),
    U3 = a3 U0 + b3 (U2 + dt L(U2))."""
    (a2, b2), (a3, b3) = weights

    def stage(U):
        R, Ub = orc.compute_rhs(U, cfg, ch)
        return R, Ub

    def fix(U):
        return orc.apply_bc(orc.enforce_bounds(U), cfg)

    def step(U0, dt):
        R, U0b = stage(U0)
        U1 = fix(U0b + dt * R)
        R, U1b = stage(U1)
        U2 = fix(a2 * U0b + b2 * (U1b + dt * R))
        R, U2b = stage(U2)
        return fix(a3 * U0b + b3 * (U2b + dt * R))

    return step


def test_ssprk3_pin_rejects_wrong_stage_weights(orc):
    """The order pin has teeth: the same step composed from oracle parts
    with the Shu-Osher weights (3/4, 1/4), (1/3, 2/3) equals the oracle's
    step bit for bit, and with the stage-2 weights of a midpoint rule
    (1/2, 1/2) or the stage-3 weights swapped it falls below order 3."""
    errs_ok, (U0, cfg, ch) = _temporal_errors(orc)
    ref_err = errs_ok[2]
    dt = 1e-3
    good = _composed_step(orc, cfg, ch, ((0.75, 0.25), (1.0 / 3.0, 2.0 / 3.0)))
    np.testing.assert_array_equal(good(U0, dt), orc.step(U0, cfg, dt, ch, decay=1.0)[0])
    for w in (((0.5, 0.5), (1.0 / 3.0, 2.0 / 3.0)), ((0.75, 0.25), (2.0 / 3.0, 1.0 / 3.0))):
        errs, _ = _temporal_errors(orc, stepper=_composed_step(orc, cfg, ch, w))
        r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
        assert min(r1, r2) < 5.0 or errs[2] > 1e3 * ref_err, (w, errs)


# ------------------------------------------------------- host diagnostics
def _diag_cfg(nx=400, ny=64):
    from config import Config

    return Config(nx=nx, ny=ny, x_min=0.0, x_max=4.0, y_min=0.0, y_max=1.0, gamma=5 / 3,
                  perturbation_mode=3)


def _state_from_c(cfg, C, rho=None, vx=None, vy=None):
    W = np.zeros((NV, cfg.nx + 4, cfg.ny + 4))
    W[0] = 1.0 if rho is None else rho
    if vx is not None:
        W[1] = vx
    if vy is not None:
        W[2] = vy
    W[7] = 1.0
    W[9] = C
    return prim_to_cons_np(W, cfg.gamma)


def _centres(cfg):
    x = cfg.x_min + (np.arange(cfg.nx + 4) - 2 + 0.5) * cfg.dx
    y = cfg.y_min + (np.arange(cfg.ny + 4) - 2 + 0.5) * cfg.dy
    return np.meshgrid(x, y, indexing="ij")


def test_diag_mixing_width_of_a_tanh_profile():
    """C = (1 + tanh((x - x0) / d)) / 2 gives int C (1 - C) dx = d / 2, a
    threshold width (0.01 < C < 0.99) of 2 atanh(0.98) d to within a cell,
    and mixedness 1 (no fluctuation about the row mean)."""
    import amhd_diag

    cfg = _diag_cfg()
    X, _ = _centres(cfg)
    for d in (0.05, 0.12):
        U = _state_from_c(cfg, 0.5 * (1.0 + np.tanh((X - 2.0) / d)))
        assert amhd_diag.mixing_width(U, cfg) == pytest.approx(d / 2, rel=1e-4)
        g = amhd_diag.diagnostics(U, cfg)
        assert g["mw_integral"] == pytest.approx(d / 2, rel=1e-4)
        assert abs(g["mw_thresh"] - 2 * np.arctanh(0.98) * d) <= 2 * cfg.dx
        assert g["mixedness"] == pytest.approx(1.0, abs=1e-12)


def test_diag_segregated_interface_has_zero_mixedness():
    """A sharp interface x_if(y) = x0 + a sin(2 pi m y / Ly) with C in {0, 1}:
    the y-variance of C equals C_bar (1 - C_bar) in every row, so the
    mixedness is 0; the interface's mode-m amplitude is a."""
    import amhd_diag

    cfg = _diag_cfg(nx=800, ny=128)
    X, Y = _centres(cfg)
    a = 0.2
    xif = 2.0 + a * np.sin(2 * np.pi * cfg.perturbation_mode * Y / (cfg.y_max - cfg.y_min))
    U = _state_from_c(cfg, (X > xif).astype(float))
    g = amhd_diag.diagnostics(U, cfg)
    assert g["mixedness"] == pytest.approx(0.0, abs=1e-12)
    assert g["pert_amp"] == pytest.approx(a, rel=0.02)
    amps = g["mode_amps"].copy()
    amps[cfg.perturbation_mode] = 0.0
    assert np.max(amps) < 0.05 * a


def test_diag_enstrophy_and_stagnation_pressure():
    """Solid-body rotation v = W (-(y - yc), x - xc) has vorticity 2W in
    every cell, so the enstrophy mean(rho w^2) is rho (2 W)^2; the
    stagnation pressure max(p + rho v^2 / 2 + B^2 / 2) is reached at the
    corner farthest from the centre."""
    import amhd_diag

    cfg = _diag_cfg(nx=64, ny=48)
    X, Y = _centres(cfg)
    om, rho = 0.7, 1.6
    vx, vy = -om * (Y - 0.5), om * (X - 2.0)
    U = _state_from_c(cfg, 0.5 + 0 * X, rho=rho, vx=vx, vy=vy)
    g = amhd_diag.diagnostics(U, cfg)
    assert g["enstrophy"] == pytest.approx(rho * (2 * om) ** 2, rel=1e-12)
    r2 = (X[2:-2, 2:-2] - 2.0) ** 2 + (Y[2:-2, 2:-2] - 0.5) ** 2
    assert g["stag"] == pytest.approx(1.0 + 0.5 * rho * om * om * np.max(r2), rel=1e-12)


def test_diag_divb_matches_the_linear_field():
    """divB of the diagnostics (solver.py:419-428) on Bx = b x, By = d y is
    b + d away from the edge cells; the total energy is sum(E) dx dy."""
    import amhd_diag

    cfg = _diag_cfg(nx=40, ny=30)
    X, Y = _centres(cfg)
    W = np.zeros((NV, cfg.nx + 4, cfg.ny + 4))
    W[0], W[7] = 1.0, 1.0
    W[4], W[5] = 0.3 * X, -0.8 * Y
    U = prim_to_cons_np(W, cfg.gamma)
    g = amhd_diag.diagnostics(U, cfg)
    assert g["divB_max"] == pytest.approx(0.5, rel=1e-10)
    assert g["energy_total"] == pytest.approx(float(np.sum(U[7, 2:-2, 2:-2])) * cfg.dx * cfg.dy, rel=1e-14)


def _host_reductions(U, cfg, ng=2):
    """The per-row / per-column reductions amhd_flow_diagnostics,
    amhd_concentration_profile and amhd_global_sums return, formed on the
    host from the same definitions."""
    import amhd_diag

    W = amhd_diag._c2p(U, cfg.gamma)
    sl = (slice(ng, -ng), slice(ng, -ng))
    rho, p, vx, vy = W[0][sl], W[7][sl], W[1][sl], W[2][sl]
    bx, by, bz = W[4][sl], W[5][sl], W[6][sl]
    cc = np.clip(W[9][sl], 0.0, 1.0)
    cbar = cc.mean(axis=1)
    cvar = np.mean((cc - cbar[:, None]) ** 2, axis=1)
    omega = np.gradient(vy, cfg.dx, axis=0) - np.gradient(vx, cfg.dy, axis=1)
    row_ens = np.sum(rho * omega ** 2, axis=1)
    row_stag = np.max(p + 0.5 * rho * (vx ** 2 + vy ** 2) + 0.5 * (bx ** 2 + by ** 2 + bz ** 2), axis=1)
    x = cfg.x_min + (np.arange(cfg.nx) + 0.5) * cfg.dx
    pos = amhd_diag.interface_positions(cc, x, cfg.dx)
    g = amhd_diag.diagnostics(U, cfg)
    return cbar, cvar, row_ens, row_stag, pos, g["mass_total"], g["energy_total"], g["divB_max"], g["divB_L2"]


def test_diag_from_reductions_equals_diagnostics():
    """The device path of compute_diagnostics (row and column reductions,
    then amhd_diag.from_reductions on the host) gives what
    amhd_diag.diagnostics gives on the whole field: a perturbed interface in
    a rotating flow, so the interface modes, the local enstrophy window and
    the stagnation maximum are all non-trivial."""
    import amhd_diag

    cfg = _diag_cfg(nx=200, ny=64)
    X, Y = _centres(cfg)
    xif = 2.0 + 0.1 * np.sin(2 * np.pi * cfg.perturbation_mode * Y / (cfg.y_max - cfg.y_min))
    C = 0.5 * (1.0 + np.tanh((X - xif) / 0.04))
    U = _state_from_c(cfg, C, rho=1.0 + 2.0 * C, vx=-0.4 * (Y - 0.5), vy=0.3 * np.sin(X))
    g = amhd_diag.diagnostics(U, cfg)
    d = amhd_diag.from_reductions(cfg, *_host_reductions(U, cfg))
    assert set(d) == set(g)
    assert g["pert_amp"] > 0.05 and g["enstrophy_local"] != g["enstrophy"]
    for k in g:
        if k == "mode_amps":
            assert np.allclose(d[k], g[k], rtol=1e-12, atol=1e-14)
        else:
            assert d[k] == pytest.approx(g[k], rel=1e-12), k
