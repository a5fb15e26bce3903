"""Learned closure on the GPU: MFMA lane layout, every convolution pass
against the CPU oracle, whole-network gradients against the oracle chain,
and (with the committed trained weights) the gain over the plain coarse
solver as measured against the fine-grid solver."""
import ctypes as C
import os

import numpy as np
import pytest

import closure_oracle as co

pytestmark = pytest.mark.gpu
PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "aeromhd-cuda_amd")


def _p(t):
    return C.c_void_p(t.data_ptr())


def test_mfma_lane_layout(torch_gpu):
    import amhd_lib

    torch = torch_gpu
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 32)).astype(np.float32)
    B = rng.standard_normal((32, 16)).astype(np.float32)
    At, Bt = torch.tensor(A, device="cuda"), torch.tensor(B, device="cuda")
    D = torch.empty(16, 16, device="cuda")
    amhd_lib.call("amhd_mfma_probe", _p(At), _p(Bt), _p(D), amhd_lib.stream_handle(torch))
    ref = co.bf16_round(A).astype(np.float64) @ co.bf16_round(B).astype(np.float64)
    np.testing.assert_allclose(D.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("cin,cout,relu,last", [(16, 32, True, False), (32, 32, True, False), (32, 32, False, True)])
def test_conv_passes_vs_oracle(torch_gpu, cin, cout, relu, last):
    import amhd_lib

    torch = torch_gpu
    rng = np.random.default_rng(cin + 3 * cout)
    n, nx, ny = 2, 9, 7
    x = co.bf16_round(np.abs(rng.standard_normal((n, nx, ny, cin))) * (rng.random((n, nx, ny, cin)) > 0.3))
    W = np.zeros((cout, co.kpad(cin)), dtype=np.float32)
    W[:, :9 * cin] = rng.standard_normal((cout, 9 * cin)) * 0.2
    b = rng.standard_normal(cout).astype(np.float32)
    xt = torch.tensor(x, device="cuda").to(torch.bfloat16)
    Wt = torch.tensor(W, device="cuda").to(torch.bfloat16)
    bt = torch.tensor(b, device="cuda")
    y = torch.empty(n, nx, ny, cout, device="cuda", dtype=torch.float32 if last else torch.bfloat16)
    s = amhd_lib.stream_handle(torch)
    amhd_lib.call("amhd_conv3x3_fwd", _p(xt), n, nx, ny, cin, _p(Wt), _p(bt), cout, int(relu), _p(y), int(last), s)
    ref = co.conv_fwd(x, W, b, relu)
    got = y.float().cpu().numpy()
    if last:
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5)
    else:   # stored in BF16: at most one BF16 ulp from the rounded oracle
        assert np.all(np.abs(got - co.bf16_round(ref)) <= 2.0 ** -7 * np.abs(ref) + 1e-6)
    # backward data with the ReLU mask of the layer input, and weights
    dy = rng.standard_normal((n, nx, ny, cout)).astype(np.float32)
    dyt = torch.tensor(dy, device="cuda")
    WT = W[:, :9 * cin].reshape(cout, 9, cin).transpose(2, 1, 0).reshape(cin, 9 * cout)
    WTp = np.zeros((cin, co.kpad(cout)), dtype=np.float32)
    WTp[:, :9 * cout] = WT
    WTt = torch.tensor(WTp, device="cuda").to(torch.bfloat16)
    dx = torch.empty(n, nx, ny, cin, device="cuda")
    amhd_lib.call("amhd_conv3x3_bwd_data", _p(dyt), n, nx, ny, cout, _p(WTt), cin, _p(xt), _p(dx), s)
    np.testing.assert_allclose(dx.cpu().numpy(), co.conv_bwd_data(dy, W, cin, mask=x), rtol=1e-4, atol=1e-4)
    dW = torch.empty(cout, co.kpad(cin), device="cuda")
    db = torch.empty(cout, device="cuda")
    ws = torch.empty(5 * cout * co.kpad(cin), device="cuda")
    amhd_lib.call("amhd_conv3x3_bwd_weight", _p(xt), _p(dyt), n, nx, ny, cin, cout, _p(dW), _p(db), _p(ws), 5, s)
    rW, rb = co.conv_bwd_weight(x, dy)
    np.testing.assert_allclose(dW.cpu().numpy(), rW, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(db.cpu().numpy(), rb, rtol=1e-5, atol=1e-4)


def test_network_gradients_vs_oracle(torch_gpu):
    """The network through the HIP kernels vs the CPU oracle chain, one piece
    at a time: the forward (hidden activations are stored in BF16, so a
    value can land one BF16 ulp away: norm-wise 1e-2), the loss gradient at
    the GPU's own output (1e-4), and every weight / bias gradient from the
    GPU's own activations and output gradient (norm-wise 1e-3)."""
    torch = torch_gpu
    import amhd_closure as cl

    rng = np.random.default_rng(5)
    n, nx, ny = 2, 16, 12
    model = cl.FluxCNN(seed=3)
    X = torch.tensor(rng.standard_normal((n, 9, nx, ny)), device="cuda", dtype=torch.float32)
    R = torch.tensor(rng.standard_normal((n, 9, nx, ny)), device="cuda", dtype=torch.float32)
    Cp = torch.tensor(rng.uniform(0, 1, (n, nx, ny)), device="cuda", dtype=torch.float32)
    Cph = Cp.cpu().numpy().astype(np.float64)
    model.stats = {"in_mean": np.zeros(9, np.float32), "in_std": np.ones(9, np.float32),
                   "r_std": rng.uniform(0.5, 2, 9).astype(np.float32),
                   "f_scale": rng.uniform(0.5, 2, 9).astype(np.float32)}
    st = model.stats
    dx, dy = 0.25, 0.2
    loss, dout, acts = cl.loss_and_grad(model, X, R, Cp, st, dx, dy)
    gW, gb = model.backward(acts, dout, n_split=3)
    Ws = [W.cpu().numpy() for W in model.W]
    bs = [b.cpu().numpy() for b in model.b]
    gacts = [a.float().cpu().numpy().astype(np.float64) for a in acts]
    out_gpu = model.forward(acts[0]).cpu().numpy().astype(np.float64)
    out_ref, _ = co.net_forward(gacts[0], Ws, bs)
    assert np.linalg.norm(out_gpu - out_ref) <= 1e-2 * np.linalg.norm(out_ref)
    Rn = R.cpu().numpy().astype(np.float64)
    assert loss == pytest.approx(co.loss(out_gpu, Rn, st["f_scale"], st["r_std"], dx, dy, Cph), rel=1e-5)
    dref = co.loss_grad_out(out_gpu, Rn, st["f_scale"].astype(np.float64), st["r_std"].astype(np.float64), dx, dy,
                            Cph)
    np.testing.assert_allclose(dout.cpu().numpy(), dref, rtol=1e-4, atol=1e-9)
    rW, rb = co.net_backward(gacts, dout.cpu().numpy().astype(np.float64), Ws)
    for li in range(len(co.LAYERS)):
        g, r = gW[li].cpu().numpy(), rW[li] * model.wmask[li].cpu().numpy()
        assert np.linalg.norm(g - r) <= 1e-3 * np.linalg.norm(r), li
        assert np.linalg.norm(gb[li].cpu().numpy() - rb[li]) <= 1e-3 * np.linalg.norm(rb[li]) + 1e-9, li


@pytest.mark.parametrize("dtype", [0, 1])
def test_closure_glue_vs_oracle(torch_gpu, dtype):
    """Network input and limited source of n states in place (FP64 and
    FP32 states): the features bit for bit; theta equal to the oracle's in
    all but at most one cell (double sums in another order can land on the
    other side of a bound); the source equal to the oracle's from the GPU's
    theta (FP32: 1e-6 of max|S|, FP64: 1e-12); without the limiter, the
    plain divergence; every field conserved; in place, U + dt*S with rho
    and p of every cell at or above half their values."""
    torch = torch_gpu
    import amhd_closure as cl
    import amhd_lib as L
    from test_closure_oracle import _state

    lib = L.load()
    rng = np.random.default_rng(30 + dtype)
    gamma, dx, dy, dt, keep = 5.0 / 3.0, 0.1, 0.08, 0.01, 0.5
    n, nx, ny = 3, 12, 10
    tdt = torch.float64 if dtype == 0 else torch.float32
    Ud = torch.tensor(_state(rng, n, nx, ny, gamma), dtype=tdt, device="cuda").contiguous()
    Uh = Ud.cpu().numpy().astype(np.float64)
    esz = Ud.element_size()
    ptrs = [Ud.data_ptr() + m * 10 * nx * ny * esz for m in range(n)]
    stats = {"in_mean": rng.standard_normal(9).astype(np.float32),
             "in_std": rng.uniform(0.5, 2, 9).astype(np.float32)}
    x = torch.empty(n, nx, ny, 16, device="cuda", dtype=torch.bfloat16)
    cl.features(ptrs, dtype, cl.FEAT, nx, ny, stats, x)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(x.float().cpu().numpy(), co.features(Uh, cl.FEAT, stats["in_mean"], stats["in_std"]))

    dF = torch.tensor(rng.standard_normal((n, nx, ny, 32)) * 0.3, dtype=torch.float32, device="cuda")
    fs = rng.uniform(0.5, 2, 9).astype(np.float32)
    theta = torch.empty(2 * n * nx * ny, device="cuda")
    src = torch.empty(n, 10, nx, ny, device="cuda", dtype=tdt)
    sp = [src.data_ptr() + m * 10 * nx * ny * esz for m in range(n)]
    V = C.c_void_p * n

    def run(th, out=sp, acc=0, bnd=None):
        L.check(lib.amhd_closure_source(_p(dF), V(*ptrs), V(*out), dtype, cl._int9(cl.FEAT), n, nx, ny, cl._f9(fs),
                                        dx, dy, cl._dvec([dt] * n), gamma, keep, _p(th) if th is not None else None,
                                        bnd, acc, L.stream_handle(torch)), "amhd_closure_source")
        torch.cuda.synchronize()
        return src.cpu().numpy().astype(np.float64)

    S = run(theta)
    # the guard quantities of U + dt S (nothing written), FP64 against NumPy
    out = np.empty((n, 4))
    work = torch.empty(32 * 64 * 4, device="cuda", dtype=torch.float64)
    L.check(lib.amhd_closure_check(V(*ptrs), V(*sp), cl._dvec([dt] * n), dtype, n, nx, ny, gamma, dx, dy, _p(work),
                                   out.ctypes.data_as(C.POINTER(C.c_double)), L.stream_handle(torch)),
            "amhd_closure_check")
    Wc = Uh + dt * S
    pc = (gamma - 1) * (Wc[:, 7] - 0.5 * (Wc[:, 1] ** 2 + Wc[:, 2] ** 2 + Wc[:, 3] ** 2) / Wc[:, 0]
                        - 0.5 * (Wc[:, 4] ** 2 + Wc[:, 5] ** 2 + Wc[:, 6] ** 2))
    bx, by = Wc[:, 4], Wc[:, 5]
    div = ((0.5 * (bx[:, 1:-1, 1:-1] + bx[:, 2:, 1:-1]) - 0.5 * (bx[:, :-2, 1:-1] + bx[:, 1:-1, 1:-1])) / dx
           + (0.5 * (by[:, 1:-1, 1:-1] + by[:, 1:-1, 2:]) - 0.5 * (by[:, 1:-1, :-2] + by[:, 1:-1, 1:-1])) / dy)
    np.testing.assert_allclose(out[:, 0], Wc[:, 0].min(axis=(1, 2)), rtol=1e-12)
    np.testing.assert_allclose(out[:, 1], pc.min(axis=(1, 2)), rtol=1e-12)
    np.testing.assert_allclose(out[:, 2], np.abs(div).max(axis=(1, 2)), rtol=1e-12)
    assert np.all(out[:, 3] == 0)
    dFh = dF.cpu().numpy().astype(np.float64)
    fs64 = fs.astype(np.float64)
    Ch = co.concentration(Uh)
    th_ref = co.limiter_theta(Uh, dFh, fs64, dt, dx, dy, gamma, keep)
    th_all = theta.cpu().numpy().astype(np.float64)
    th_gpu = th_all[:n * nx * ny].reshape(n, nx, ny)
    tm_gpu = th_all[n * nx * ny:].reshape(n, nx, ny)
    assert np.sum(th_gpu != th_ref) <= 1
    assert 0 < np.sum(th_gpu < 1) < th_gpu.size
    tm_ref = co.limiter_theta_mix(Uh, dFh, th_gpu, fs64, dt, dx, dy)
    assert np.sum(tm_gpu != tm_ref) <= 1
    assert 0 < np.sum(tm_gpu < 1) < tm_gpu.size
    tol = 1e-6 if dtype == 1 else 1e-12
    ref = co.limited_source(dFh, th_gpu, tm_gpu, fs64, dx, dy, Ch)
    scale = np.abs(ref).max()
    np.testing.assert_allclose(np.transpose(S[:, cl.FEAT], (0, 2, 3, 1)), ref, rtol=0, atol=tol * scale)
    assert np.all(S[:, 8] == 0)
    assert np.all(np.abs(S[:, cl.FEAT].sum(axis=(2, 3))) <= 1e-5 * np.abs(S).sum(axis=(2, 3))[:, cl.FEAT] + 1e-9)
    S1 = run(None)
    ref1 = co.source_from_flux(dFh[..., :18] * np.concatenate([fs64, fs64]), dx, dy, Ch)
    np.testing.assert_allclose(np.transpose(S1[:, cl.FEAT], (0, 2, 3, 1)), ref1, rtol=0,
                               atol=tol * np.abs(ref1).max())
    # concentration bounds widened by a second state (the state before a step)
    Ub = torch.tensor(_state(rng, n, nx, ny, gamma), dtype=tdt, device="cuda").contiguous()
    bptrs = [Ub.data_ptr() + m * 10 * nx * ny * esz for m in range(n)]
    run(theta, bnd=V(*bptrs))
    tb_gpu = theta.cpu().numpy().astype(np.float64)[n * nx * ny:].reshape(n, nx, ny)
    tb_ref = co.limiter_theta_mix(Uh, dFh, th_gpu, fs64, dt, dx, dy, Ub.cpu().numpy().astype(np.float64))
    assert np.sum(tb_gpu != tb_ref) <= 1
    assert np.all(tb_gpu >= tm_gpu) and np.any(tb_gpu > tm_gpu)
    # in place on the states: U += dt * S with the limiter taken on U itself
    run(theta, out=ptrs, acc=1)
    Unew = Ud.cpu().numpy().astype(np.float64)
    exp = Uh.copy()
    exp[:, cl.FEAT] += dt * np.transpose(ref, (0, 3, 1, 2))
    np.testing.assert_allclose(Unew, exp, rtol=2 * tol, atol=tol * np.abs(Uh).max())
    assert np.all(Unew[:, 8] == Uh[:, 8])
    # the limiters' guarantees on the result: rho and p keep half their
    # values, C stays within its neighbourhood's bounds
    p0, p1 = co._pressure(Uh, gamma), co._pressure(Unew, gamma)
    assert np.all(Unew[:, 0] >= keep * Uh[:, 0] * (1 - 10 * tol))
    assert np.all(p1 >= keep * p0 - 10 * tol * np.abs(Uh[:, 7]))
    lo, hi = co.local_bounds(Ch)
    c1 = Unew[:, 9] / Unew[:, 0]
    assert np.all(c1 >= lo - 10 * tol) and np.all(c1 <= hi + 10 * tol)


@pytest.mark.parametrize("dtype", [0, 1])
def test_closure_sharpen_vs_oracle(torch_gpu, dtype):
    """amhd_closure_sharpen on 35 random states (two launches of up to 32
    instances), with and without the second state's bounds, at a kappa that
    leaves some faces unlimited and one that saturates them: equal to the
    oracle (FP64: 1e-12 of max|rhoC|, FP32 states: 1e-6), C within the
    bounds, rhoC conserved row by row, every other plane bit for bit as
    before, kappa = 0 a no-op."""
    torch = torch_gpu
    import amhd_lib as L
    from test_closure_oracle import _state

    lib = L.load()
    rng = np.random.default_rng(50 + dtype)
    gamma = 5.0 / 3.0
    n, nx, ny = 35, 12, 10
    tdt = torch.float64 if dtype == 0 else torch.float32
    U0 = _state(rng, n, nx, ny, gamma)
    U0[:, 9] = U0[:, 0] * np.clip(rng.uniform(-0.3, 1.3, (n, nx, ny)), 0, 1)   # pure and mixed cells
    Ub = torch.tensor(_state(rng, n, nx, ny, gamma), dtype=tdt, device="cuda").contiguous()
    Ubh = Ub.cpu().numpy().astype(np.float64)
    esz = 8 if dtype == 0 else 4
    work = torch.empty(3 * 32 * nx * ny, device="cuda", dtype=torch.float64)
    V = C.c_void_p * n
    tol = 1e-12 if dtype == 0 else 1e-6
    for kappa in (0.05, 2.0):
        for with_b in (False, True):
            Ud = torch.tensor(U0, dtype=tdt, device="cuda").contiguous()
            Uh = Ud.cpu().numpy().astype(np.float64)
            ptrs = V(*[Ud.data_ptr() + m * 10 * nx * ny * esz for m in range(n)])
            bptrs = V(*[Ub.data_ptr() + m * 10 * nx * ny * esz for m in range(n)]) if with_b else None
            print(f"sharpen dtype {dtype} kappa {kappa} bounds {with_b}", flush=True)
            L.check(lib.amhd_closure_sharpen(ptrs, bptrs, dtype, n, nx, ny, kappa, _p(work), L.stream_handle(torch)),
                    "amhd_closure_sharpen")
            torch.cuda.synchronize()
            out = Ud.cpu().numpy().astype(np.float64)
            ref = co.sharpen_x(Uh, Ubh if with_b else None, kappa)
            scale = np.abs(Uh[:, 9]).max()
            np.testing.assert_allclose(out[:, 9], ref[:, 9], rtol=0, atol=tol * scale)
            assert np.any(np.abs(out[:, 9] - Uh[:, 9]) > 100 * tol * scale)
            np.testing.assert_array_equal(np.delete(out, 9, axis=1), np.delete(Uh, 9, axis=1))
            np.testing.assert_allclose(out[:, 9].sum(axis=1), Uh[:, 9].sum(axis=1), rtol=0,
                                       atol=10 * tol * np.abs(Uh[:, 9]).sum(axis=1).max())
            lo, hi = co.local_bounds(co.concentration(Uh))
            if with_b:
                blo, bhi = co.local_bounds(co.concentration(Ubh))
                lo, hi = np.minimum(lo, blo), np.maximum(hi, bhi)
            c1 = out[:, 9] / out[:, 0]
            assert np.all(c1 >= lo - 10 * tol) and np.all(c1 <= hi + 10 * tol)
    Ud = torch.tensor(U0, dtype=tdt, device="cuda").contiguous()
    before = Ud.cpu().numpy()
    ptrs = V(*[Ud.data_ptr() + m * 10 * nx * ny * esz for m in range(n)])
    L.check(lib.amhd_closure_sharpen(ptrs, None, dtype, n, nx, ny, 0.0, None, L.stream_handle(torch)),
            "amhd_closure_sharpen")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(Ud.cpu().numpy(), before)


def test_closure_gain_over_plain_coarse(torch_gpu):
    """The shipped network (closure_weights.npz, which applies the mixing
    flux alone, with its calibrated gain, then the sharpening of C at its
    calibrated coefficient) against the plain coarse solver on two held-out
    seeds, measured against the fine 2048x1024 solver restricted to the
    coarse grid: the mixing-width error within 5% and under a fifth of the
    plain solver's, the concentration error below the plain solver's, the
    density (which neither the mixing flux nor the sharpening changes) as
    the plain solver's, no fallbacks; the sharpening alone (no network, its
    own calibrated coefficient) within 5% as well."""
    import amhd_closure as cl

    path = os.path.join(PKG, "closure_weights.npz")
    assert os.path.exists(path), "closure_weights.npz (trained weights) missing"
    model = cl.FluxCNN.load(path)
    assert cl.applied_fields(model.stats) == [cl.RC]
    rows = cl.evaluate([210, 217], model, log=print, variants=("closure", "sharpen"))
    for r in rows:
        assert r["mw_err_closure"] < 0.2 * r["mw_err_plain"], r
        assert r["mw_err_closure"] < 0.05, r
        assert r["l1c_closure"] < r["l1c_plain"], r
        assert r["mw_err_sharpen"] < 0.05, r
        assert r["l1rho_closure"] == pytest.approx(r["l1rho_plain"], rel=1e-6), r
        assert r["fallbacks_closure"] == 0, r


def _coarse_sieve(seed):
    from amhd_ics import cons_from_prim_host, magnetic_sieve

    cfg, W = magnetic_sieve(nx=256, ny=128, precision="fp32", seed=seed, max_steps=10 ** 6)
    cfg.t_end = 1e30
    return cfg, cons_from_prim_host(W, cfg.gamma)


def test_sharpening_step_is_the_plain_step_then_the_oracle(torch_gpu):
    """One host-driven step of the sharpening alone (ClosureSolver without a
    network, kappa > 0) equals the plain step of the same dt from the same
    state followed by the oracle's sharpen_x on the post-step state with the
    pre-step state's bounds (FP32 states: 1e-6 of max|rhoC|); only rhoC
    differs from the plain step, and it does differ."""
    import amhd_closure as cl

    cfg, U = _coarse_sieve(216)
    kappa = 0.3
    post = {}
    for tag, kw in (("plain", {}), ("sharp", {"kappa": kappa})):
        cs = cl.ClosureSolver(cfg, None, **kw)
        cs.s.U = U.copy()
        cs.begin()
        pre = cs.s.ctx.download_numpy()
        dt = 0.5 * cs.s.compute_dt()
        print(f"sharpening step {tag}: dt {dt:.6e}", flush=True)
        assert cs.step(dt) is True
        post[tag] = (pre, dt, cs.s.ctx.download_numpy())
    (pre_p, dt_p, post_p), (pre_s, dt_s, post_s) = post["plain"], post["sharp"]
    np.testing.assert_array_equal(pre_p, pre_s)
    assert dt_p == dt_s
    inner = (slice(None), slice(2, -2), slice(2, -2))
    ref = co.sharpen_x(post_p[inner][None].astype(np.float64), pre_p[inner][None].astype(np.float64), kappa)[0]
    scale = np.abs(post_p[9][2:-2, 2:-2]).max()
    np.testing.assert_allclose(post_s[9][2:-2, 2:-2], ref[9], rtol=0, atol=1e-6 * scale)
    assert np.any(np.abs(post_s[9][2:-2, 2:-2] - post_p[9][2:-2, 2:-2]) > 1e-4 * scale)
    np.testing.assert_array_equal(np.delete(post_s[inner], 9, axis=0), np.delete(post_p[inner], 9, axis=0))


def test_closure_fallback_is_the_plain_step(torch_gpu):
    """A guard that rejects every correction (divB floor -1e30) leaves every
    step the plain one: the run equals the plain coarse solver's bit for
    bit, step for step."""
    import amhd_closure as cl

    model = cl.FluxCNN.load(os.path.join(PKG, "closure_weights.npz"))
    cfg, U = _coarse_sieve(211)
    runs = {}
    for tag, m, floor in (("plain", None, 1.0), ("rejected", model, -1e30)):
        cs = cl.ClosureSolver(cfg, m, divb_floor=floor)
        cs.s.U = U.copy()
        cs.run(1e30, max_steps=40)
        runs[tag] = (cs.s.ctx.download_numpy(), cs.s.t, cs.s.step, cs.fallbacks)
    assert runs["rejected"][3] == runs["rejected"][2] == 40
    np.testing.assert_array_equal(runs["rejected"][0], runs["plain"][0])
    assert runs["rejected"][1] == runs["plain"][1]


def test_closure_guard_catches_bad_corrections(torch_gpu):
    """Unlimited corrections of every field at 100x the calibrated scale
    drive the pressure negative; the guard rejects those steps and the
    state stays finite with positive density and pressure."""
    import amhd_closure as cl

    model = cl.FluxCNN.load(os.path.join(PKG, "closure_weights.npz"))
    cfg, U = _coarse_sieve(212)
    cs = cl.ClosureSolver(cfg, model, keep=None, fields=list(range(cl.N_IN)), gain=100.0)
    cs.s.U = U
    cs.run(1e30, max_steps=30)
    rho_min, p_min, divb, nonfinite = cs.check()
    print(f"fallbacks {cs.fallbacks}/{cs.s.step}, rho_min {rho_min:.3e}, p_min {p_min:.3e}", flush=True)
    assert cs.fallbacks > 0
    assert nonfinite == 0 and rho_min > 0 and p_min > 0


@pytest.mark.parametrize("form", ["closure", "sharpen"])
def test_batched_closure_equals_single_instances(torch_gpu, form):
    """Three held-out seeds in one BatchClosure (one batched step, one
    network pass, one batched correction and sharpening per step; or, for
    the sharpening alone, the batched step and one batched sharpening) end
    bit for bit where ClosureSolver takes each alone, with the same step
    counts."""
    import amhd_closure as cl

    model = cl.FluxCNN.load(os.path.join(PKG, "closure_weights.npz"))
    kw = {}
    if form == "sharpen":
        kw = {"kappa": float(model.stats["kappa_alone"])}
        model = None
    cases = []
    for seed, t_end in ((213, 0.004), (214, 0.0035), (215, 0.003)):
        cfg, U = _coarse_sieve(seed)
        cfg.t_end = t_end
        cases.append((cfg, U))
    singles = []
    for cfg, U in cases:
        cs = cl.ClosureSolver(cfg, model, **kw)
        cs.s.U = U.copy()
        cs.run(cfg.t_end)
        singles.append((cs.s.ctx.download_numpy(), cs.s.step, cs.s.t))
    bc = cl.BatchClosure(cases, model, **kw)
    bc.run()
    for k, (U1, steps, t) in enumerate(singles):
        print(f"instance {k}: steps {bc.b.steps[k]} / {steps}, fallbacks {bc.fallbacks[k]}", flush=True)
        assert bc.b.steps[k] == steps
        assert bc.b.t[k] == t
        np.testing.assert_array_equal(bc.b.download(k), U1)
