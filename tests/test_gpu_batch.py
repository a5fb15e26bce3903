"""Batched persistent form: every instance of a batch must equal its own
single-instance device loop bit for bit (same dtype; the result does not
depend on the tiling), stop at its own t_end / max_steps, and the tuner
must pick one of its candidates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def cases(precision):
    from amhd_ics import magnetic_sieve, rmi_uniform_bx, cons_from_prim_host

    out = []
    for nx, ny, mach, atw, beta, mode in ((256, 128, 10.0, 0.5, 10.0, 4), (192, 100, 3.0, 0.3, 2.0, 2),
                                          (320, 160, 12.0, 0.8, 90.0, 7), (128, 64, 6.0, 0.2, 1.2, 1)):
        cfg, W = rmi_uniform_bx(nx, ny, precision=precision, mach=mach, atwood=atw, beta=beta, mode=mode)
        out.append((cfg, cons_from_prim_host(W, cfg.gamma)))
    cfg, W = magnetic_sieve(nx=160, ny=96, seed=5, precision=precision)   # Powell on
    out.append((cfg, cons_from_prim_host(W, cfg.gamma)))
    for cfg, _ in out:
        cfg.t_end, cfg.max_steps = 1e9, 10 ** 6
    return out


def single(cfg, U, n):
    from solver import MHDSolver, to_numpy

    s = MHDSolver(cfg)
    s.U = U
    s.advance(n)
    return to_numpy(s.U), s.t, s.step


@pytest.mark.parametrize("precision,steps", [("fp32", 25), ("fp64", 12)])
def test_batch_equals_single_instances(torch_gpu, precision, steps):
    from amhd_batch import BatchSolver

    cs = cases(precision)
    b = BatchSolver(cs)
    say("batch", precision, b.info())
    took = b.run(steps)
    assert list(took) == [steps] * len(cs)
    for k, (cfg, U) in enumerate(cs):
        Us, ts, ss = single(cfg, U, steps)
        np.testing.assert_array_equal(b.download(k), Us, err_msg=f"instance {k}")
        assert b.t[k] == ts and b.steps[k] == ss


def test_batch_tiling_invariance_and_graph(torch_gpu):
    from amhd_batch import BatchSolver

    cs = cases("fp32")
    res = []
    for rows, graph in ((0, True), (3, True), (17, False)):
        b = BatchSolver(cs)
        b.set_rows(rows)
        b.run(9, use_graph=graph)
        res.append([b.download(k) for k in range(len(cs))])
        say("rows", rows, b.info())
    for other in res[1:]:
        for a, c in zip(res[0], other):
            np.testing.assert_array_equal(a, c)


def test_batch_per_instance_stopping(torch_gpu):
    from amhd_batch import BatchSolver

    cs = cases("fp32")[:3]
    cs[0][0].max_steps = 4
    cs[1][0].max_steps = 11
    b = BatchSolver(cs)
    b.run(7)
    assert list(b.steps) == [4, 7, 7]
    b.run(7)
    assert list(b.steps) == [4, 11, 14]
    Us, ts, ss = single(cs[1][0], cs[1][1], 11)
    np.testing.assert_array_equal(b.download(1), Us)


def test_tuners_pick_a_candidate(torch_gpu, tmp_path, monkeypatch):
    import amhd_tune
    from amhd_batch import BatchSolver
    from amhd_ics import rmi_uniform_bx
    from physics import prim_to_cons
    from solver import MHDSolver

    monkeypatch.setenv("AMHD_TUNE_CACHE", str(tmp_path / "tune.json"))
    cfg, W = rmi_uniform_bx(256, 128, precision="fp32")
    s = MHDSolver(cfg)
    s.U = prim_to_cons(W, cfg.gamma)
    before = s.U.clone()
    best = amhd_tune.tune_single(s, candidates=(0, 4, 16), steps=3)
    assert best in (0, 4, 16)
    np.testing.assert_array_equal(s.U.cpu().numpy(), before.cpu().numpy())
    assert amhd_tune.tune_single(s, candidates=(0, 4, 16)) == best      # cached
    b = BatchSolver(cases("fp32")[:2])
    assert amhd_tune.tune_batch(b, candidates=(0, 5), steps=2) in (0, 5)
