"""x-slab decomposition on one GPU: slabs with halo exchange before every
RK stage and a dt from the maxima over all slabs give the undivided
solver's state bit for bit."""
import numpy as np
import pytest

from helpers import prim_to_cons_np, smooth_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def undivided(cfg, U, n_steps):
    """The host-driven loop of solver.py:475-482 on one context."""
    from solver import MHDSolver

    s = MHDSolver(cfg)
    s.U = U
    for _ in range(n_steps):
        dt = min(s.compute_dt(), cfg.t_end - s.t)
        s.dt = dt
        s.step_ssprk3()
        s.t += dt
        s.step += 1
    return s.ctx.download_numpy()[:, 2:-2, 2:-2], s.t, s._cumulative_boundary_energy


@pytest.mark.parametrize("case,precision,n_slabs", [("sieve", "fp64", 4), ("sieve", "fp32", 3),
                                                    ("periodic", "fp64", 2)])
def test_slabs_equal_undivided(torch_gpu, case, precision, n_slabs):
    from amhd_ics import cons_from_prim_host, magnetic_sieve
    from amhd_slab import SlabSolver
    from config import Config

    if case == "sieve":
        cfg, W = magnetic_sieve(nx=256, ny=128, seed=5, precision=precision, max_steps=10 ** 6)
        U = cons_from_prim_host(W, cfg.gamma)
    else:
        cfg = Config(nx=96, ny=64, x_min=0, x_max=1.5, y_min=0, y_max=1, gamma=5 / 3, bc_x_type="periodic",
                     bc_y_type="periodic", B_transverse=0.3, precision=precision, t_end=1e9, max_steps=10 ** 6)
        U = prim_to_cons_np(smooth_state(96, 64, seed=8, amp=0.3), cfg.gamma)
    say(f"slabs {case} {precision} n={n_slabs}")
    ref, t_ref, bf_ref = undivided(cfg, U, 25)
    sl = SlabSolver(cfg, n_slabs)
    sl.upload(U)
    assert sl.advance(25) == 25
    got = sl.download()
    np.testing.assert_array_equal(got, ref)
    assert sl.t == t_ref
    assert sl.boundary_energy() == pytest.approx(bf_ref, rel=1e-12, abs=1e-14)


def test_slab_halo_changes_the_answer(torch_gpu):
    """The halo rows are read: when the exchange before stage 2 is skipped
    (the halos still hold the stage-1 input), the slabs depart from the
    undivided run within the stencil reach of their shared edge only."""
    from amhd_ics import cons_from_prim_host, magnetic_sieve
    from amhd_slab import SlabSolver

    cfg, W = magnetic_sieve(nx=128, ny=64, seed=6, precision="fp64", max_steps=10 ** 6)
    U = cons_from_prim_host(W, cfg.gamma)
    ref, _, _ = undivided(cfg, U, 1)
    sl = SlabSolver(cfg, 5)          # slab edges at rows 25 (the shock), 51, 76, 102
    sl.upload(U)
    dt, ch = sl.compute_dt()
    exchange = sl._exchange
    sl._exchange = lambda stage: None if stage == 2 else exchange(stage)
    sl.step_ssprk3(dt, ch)
    got = sl.download()
    bad = np.nonzero(np.any(got != ref, axis=(0, 2)))[0]
    edges = [i0 for i0, _ in sl.bounds[1:]]
    assert len(bad) > 0, "skipping an exchange changed nothing"
    assert all(any(b - 4 <= r <= b + 3 for b in edges) for r in bad), (bad, edges)
