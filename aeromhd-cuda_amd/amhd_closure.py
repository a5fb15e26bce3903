"""Learned flux-correction closure on an 8x-coarsened grid (no counterpart
in the reference).

A 4-layer 3x3 CNN (32 channels; BF16 weights, FP32 accumulation, run by
the BF16-MFMA kernels of libaeromhd_hip.so) maps the coarse conserved
state -- 9 fields: rho, m, B, E, rhoC -- to corrections of the x- and
y-face fluxes of those 9 fields.  Their discrete divergence enters the
coarse SSP-RK3 step as a constant source S = -div(dF) in every stage
(amhd_set_source), which adds dt*S per step and, being a flux difference,
conserves every corrected quantity exactly.  Face corrections on the two
x boundaries are held at zero; y is periodic.

Training data come from this solver: fine 2048x1024 magnetic-sieve runs
(FP32) restricted to the coarse grid by 8x8 block means of the conserved
variables.  For a pair of restricted states (U_a at t, U_b at t + dt,
dt = 8 fine steps) the target is the source the plain coarse step misses:
r = (U_b - RK3_coarse(U_a, dt)) / dt.  The loss is the normalised mean
squared error between S(CNN(U_a)) and r.

At run time every corrected step is checked (density and pressure
positive, no non-finite values, max|divB| not above 1.5x its pre-step
value plus a floor); a failing step is discarded (amhd_rewind) and redone
by the plain coarse solver.

The fine-grid solver judges the result: integral mixing width and L1
density error of the coarse run against the restricted fine run.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

import amhd_lib as _L
from config import Config

FEAT = [0, 1, 2, 3, 4, 5, 6, 7, 9]        # conserved fields seen by the network (psi excluded)
N_IN, N_OUT = len(FEAT), 2 * len(FEAT)
LAYERS = [(16, 32, True), (32, 32, True), (32, 32, True), (32, 32, False)]   # (Cin, Cout, relu), padded widths
FACTOR = 8


def kpad(c: int) -> int:
    return (9 * c + 31) // 32 * 32


def _p(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


# ------------------------------------------------------------------ network
class FluxCNN:
    """Parameters (FP32 masters on the device), BF16 copies and the
    forward / backward passes through the HIP MFMA kernels."""

    def __init__(self, seed: int = 0):
        torch = _L.require_gpu()
        self.lib = _L.load()
        g = torch.Generator().manual_seed(seed)
        self.W, self.b, self.wmask = [], [], []
        for li, (cin, cout, _) in enumerate(LAYERS):
            real_in = N_IN if li == 0 else cin
            real_out = N_OUT if li == len(LAYERS) - 1 else cout
            W = torch.zeros(cout, kpad(cin))
            std = math.sqrt(2.0 / (9 * real_in)) * (0.1 if li == len(LAYERS) - 1 else 1.0)
            w = torch.randn(real_out, 9, real_in, generator=g) * std
            Wv = W[:, :9 * cin].view(cout, 9, cin)
            Wv[:real_out, :, :real_in] = w
            m = torch.zeros_like(W)
            m[:, :9 * cin].view(cout, 9, cin)[:real_out, :, :real_in] = 1.0
            self.W.append(W.cuda())
            self.b.append(torch.zeros(cout).cuda())
            self.wmask.append(m.cuda())
        self.stats = None   # normalisation, set by fit_stats / load
        self.refresh()

    # BF16 weights (W[co][k]) and transposed weights (WT[ci][tap*Cout+co])
    def refresh(self) -> None:
        torch = _L.require_gpu()
        self.Wb, self.WTb = [], []
        for (cin, cout, _), W in zip(LAYERS, self.W):
            self.Wb.append(W.to(torch.bfloat16).contiguous())
            wt = W[:, :9 * cin].reshape(cout, 9, cin).permute(2, 1, 0).reshape(cin, 9 * cout)
            WT = torch.zeros(cin, kpad(cout), device=W.device)
            WT[:, :9 * cout] = wt
            self.WTb.append(WT.to(torch.bfloat16).contiguous())

    def parameters(self):
        return self.W + self.b

    def forward(self, x, keep: bool = False):
        """x: (n, nx, ny, 16) BF16.  Returns the (n, nx, ny, 32) FP32 output
        (channels 0..17 used) and, with keep, the layer inputs."""
        torch = _L.require_gpu()
        n, nx, ny, _ = x.shape
        acts = [x]
        stream = _L.stream_handle(torch)
        h = x
        for li, (cin, cout, relu) in enumerate(LAYERS):
            last = li == len(LAYERS) - 1
            y = torch.empty(n, nx, ny, cout, device=x.device, dtype=torch.float32 if last else torch.bfloat16)
            _L.check(self.lib.amhd_conv3x3_fwd(_p(h), n, nx, ny, cin, _p(self.Wb[li]), _p(self.b[li]), cout,
                                               1 if relu else 0, _p(y), 1 if last else 0, stream), "amhd_conv3x3_fwd")
            h = y
            if keep and not last:
                acts.append(y)
        return (h, acts) if keep else h

    def backward(self, acts, dout, n_split: int = 64):
        """Gradients of every W and b for d(loss)/d(output) = dout (FP32,
        (n, nx, ny, 32)); the padded entries stay zero."""
        torch = _L.require_gpu()
        n, nx, ny, _ = dout.shape
        stream = _L.stream_handle(torch)
        gW = [None] * len(LAYERS)
        gb = [None] * len(LAYERS)
        dy = dout.contiguous()
        for li in range(len(LAYERS) - 1, -1, -1):
            cin, cout, _ = LAYERS[li]
            dW = torch.empty(cout, kpad(cin), device=dout.device)
            db = torch.empty(cout, device=dout.device)
            ws = torch.empty(n_split * cout * kpad(cin), device=dout.device)
            _L.check(self.lib.amhd_conv3x3_bwd_weight(_p(acts[li]), _p(dy), n, nx, ny, cin, cout, _p(dW), _p(db),
                                                      _p(ws), n_split, stream), "amhd_conv3x3_bwd_weight")
            gW[li] = dW * self.wmask[li]
            gb[li] = db
            if li > 0:
                dx = torch.empty(n, nx, ny, cin, device=dout.device)
                _L.check(self.lib.amhd_conv3x3_bwd_data(_p(dy), n, nx, ny, cout, _p(self.WTb[li]), cin,
                                                        _p(acts[li]), _p(dx), stream), "amhd_conv3x3_bwd_data")
                dy = dx
        return gW, gb

    # ------------------------------------------------------------ state
    def state_dict(self) -> dict:
        d = {f"W{i}": W.cpu().numpy() for i, W in enumerate(self.W)}
        d.update({f"b{i}": b.cpu().numpy() for i, b in enumerate(self.b)})
        for k, v in (self.stats or {}).items():
            d[f"stats_{k}"] = np.asarray(v)
        return d

    def save(self, path: str) -> None:
        np.savez_compressed(path, **self.state_dict())

    @classmethod
    def load(cls, path: str) -> "FluxCNN":
        torch = _L.require_gpu()
        z = np.load(path, allow_pickle=False)
        m = cls()
        for i in range(len(LAYERS)):
            m.W[i] = torch.as_tensor(z[f"W{i}"]).cuda()
            m.b[i] = torch.as_tensor(z[f"b{i}"]).cuda()
        m.stats = {k[6:]: z[k] for k in z.files if k.startswith("stats_")}
        m.refresh()
        return m


# ------------------------------------------------------- features / source
def features(Uc, stats):
    """(n, 10, nx, ny) coarse conserved (device) -> normalised NHWC BF16
    (n, nx, ny, 16) with the 9 FEAT fields in channels 0..8."""
    torch = _L.require_gpu()
    mean = torch.as_tensor(stats["in_mean"], device=Uc.device, dtype=torch.float32).view(1, -1, 1, 1)
    std = torch.as_tensor(stats["in_std"], device=Uc.device, dtype=torch.float32).view(1, -1, 1, 1)
    x = (Uc[:, FEAT].float() - mean) / std
    out = torch.zeros(Uc.shape[0], Uc.shape[2], Uc.shape[3], 16, device=Uc.device, dtype=torch.bfloat16)
    out[..., :N_IN] = x.permute(0, 2, 3, 1).to(torch.bfloat16)
    return out


def source_from_flux(dF, dx: float, dy: float):
    """S = -div(dF) for face corrections dF (n, nx, ny, 18): channels 0..8
    at the x face i-1/2 of cell i (zero on both x boundaries), 9..17 at the
    y face j-1/2 (periodic).  Returns (n, nx, ny, 9)."""
    torch = _L.require_gpu()
    fx = dF[..., :N_IN]
    fx = torch.cat([torch.zeros_like(fx[:, :1]), fx[:, 1:], torch.zeros_like(fx[:, :1])], dim=1)
    fy = dF[..., N_IN:N_OUT]
    sx = -(fx[:, 1:] - fx[:, :-1]) / dx
    sy = -(torch.roll(fy, -1, dims=2) - fy) / dy
    return sx + sy


def restrict(Ug, factor: int = FACTOR):
    """Ghosted (10, nx+4, ny+4) -> interior block means (10, nx/f, ny/f) (float64)."""
    U = Ug[:, 2:-2, 2:-2]
    nv, nx, ny = U.shape
    return U.reshape(nv, nx // factor, factor, ny // factor, factor).mean(dim=(2, 4))


def ghosted(Ui):
    """Interior (10, nx, ny) -> ghosted (10, nx+4, ny+4) with edge copies
    (the solver ignores uploaded ghosts; they only keep the array shape)."""
    torch = _L.require_gpu()
    return torch.nn.functional.pad(Ui.unsqueeze(0), (2, 2, 2, 2), mode="replicate")[0]


def coarse_config(fine_cfg, factor: int = FACTOR) -> Config:
    c = Config(**{k: getattr(fine_cfg, k) for k in fine_cfg.__dataclass_fields__})
    c.nx, c.ny = fine_cfg.nx // factor, fine_cfg.ny // factor
    c.precision = "fp32"
    return c


# -------------------------------------------------------------- fine data
def fine_run(seed: int, nx: int = 2048, ny: int = 1024, steps: int = 2000, on_sample=None, sample_steps=()):
    """A fine magnetic-sieve run (FP32) for `seed`; calls on_sample(solver)
    when the step count reaches each entry of sample_steps."""
    from amhd_ics import magnetic_sieve
    from physics import prim_to_cons
    from solver import MHDSolver

    cfg, W = magnetic_sieve(nx, ny, precision="fp32", seed=seed, max_steps=steps)
    cfg.t_end = 1e30
    s = MHDSolver(cfg)
    s.U = prim_to_cons(W, cfg.gamma)
    s._sync_state()          # samplers read the context directly, also at step 0
    for st in sorted(sample_steps):
        if st > s.step:
            s.advance(st - s.step)
        if on_sample:
            on_sample(s)
    if s.step < steps:
        s.advance(steps - s.step)
    return s


def generate_pairs(seeds, pairs_per_run: int = 16, span: int = 8, steps: int = 2000, nx: int = 2048,
                   ny: int = 1024, log=None):
    """Training pairs (U_a[FEAT], r[FEAT]) on the coarse grid from fine runs."""
    torch = _L.require_gpu()
    from solver import MHDSolver

    starts = [int(v) for v in np.linspace(0, steps - span, pairs_per_run)]
    X, R = [], []
    coarse = None
    for si, seed in enumerate(seeds):
        t0 = time.time()
        state = {}

        def sample(s):
            nonlocal coarse
            if coarse is None:
                coarse = MHDSolver(coarse_config(s.cfg))
            Ua = restrict(s.ctx.download())
            ta = s.t
            s.advance(span)
            Ub = restrict(s.ctx.download())
            dt = s.t - ta
            coarse.U = ghosted(Ua)
            coarse.t, coarse.step = 0.0, 0
            coarse.compute_dt()
            coarse.dt = dt
            coarse.step_ssprk3()
            Uc = coarse.ctx.download()[:, 2:-2, 2:-2]
            X.append(Ua[FEAT].float())
            R.append(((Ub - Uc) / dt)[FEAT].float())
            state["n"] = state.get("n", 0) + 1

        fine_run(seed, nx, ny, steps, sample, starts)
        if log:
            log(f"[closure] data seed {seed}: {state.get('n', 0)} pairs, {time.time() - t0:.1f} s")
    return torch.stack(X), torch.stack(R)


def fit_stats(X, R, dx: float) -> dict:
    in_mean = X.mean(dim=(0, 2, 3)).cpu().numpy()
    in_std = X.std(dim=(0, 2, 3)).clamp_min(1e-6).cpu().numpy()
    r_std = R.std(dim=(0, 2, 3)).clamp_min(1e-12).cpu().numpy()
    return {"in_mean": in_mean, "in_std": in_std, "r_std": r_std, "f_scale": r_std * dx}


# ---------------------------------------------------------------- training
def loss_and_grad(model, X, R, stats, dx, dy):
    """Normalised MSE of S(CNN(X)) against R and d(loss)/d(output)."""
    torch = _L.require_gpu()
    Ug = torch.zeros(X.shape[0], 10, X.shape[2], X.shape[3], device=X.device)
    Ug[:, FEAT] = X
    x = features(Ug, stats)
    out, acts = model.forward(x, keep=True)
    o = out.detach().requires_grad_(True)
    fscale = torch.as_tensor(np.concatenate([stats["f_scale"], stats["f_scale"]]), device=X.device,
                             dtype=torch.float32)
    S = source_from_flux(o[..., :N_OUT] * fscale, dx, dy)
    rstd = torch.as_tensor(stats["r_std"], device=X.device, dtype=torch.float32)
    err = (S - R.permute(0, 2, 3, 1)) / rstd
    loss = (err * err).mean()
    loss.backward()
    return float(loss.detach()), o.grad, acts


def loss_only(model, X, R, stats, dx, dy) -> float:
    """The loss of loss_and_grad without the gradient."""
    torch = _L.require_gpu()
    Ug = torch.zeros(X.shape[0], 10, X.shape[2], X.shape[3], device=X.device)
    Ug[:, FEAT] = X
    out = model.forward(features(Ug, stats))
    fscale = torch.as_tensor(np.concatenate([stats["f_scale"], stats["f_scale"]]), device=X.device,
                             dtype=torch.float32)
    S = source_from_flux(out[..., :N_OUT] * fscale, dx, dy)
    rstd = torch.as_tensor(stats["r_std"], device=X.device, dtype=torch.float32)
    err = (S - R.permute(0, 2, 3, 1)) / rstd
    return float((err * err).mean())


class Adam:
    def __init__(self, params, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        torch = _L.require_gpu()
        self.p, self.lr, self.b1, self.b2, self.eps = params, lr, b1, b2, eps
        self.m = [torch.zeros_like(p) for p in params]
        self.v = [torch.zeros_like(p) for p in params]
        self.t = 0

    def step(self, grads):
        self.t += 1
        for p, g, m, v in zip(self.p, grads, self.m, self.v):
            m.mul_(self.b1).add_(g, alpha=1 - self.b1)
            v.mul_(self.b2).addcmul_(g, g, value=1 - self.b2)
            mh = m / (1 - self.b1 ** self.t)
            vh = v / (1 - self.b2 ** self.t)
            p.sub_(self.lr * mh / (vh.sqrt() + self.eps))


def train(X, R, dx, dy, iters: int = 2000, batch: int = 16, lr: float = 1e-3, seed: int = 0, log=None,
          holdout=None):
    torch = _L.require_gpu()
    model = FluxCNN(seed)
    model.stats = fit_stats(X, R, dx)
    opt = Adam(model.parameters(), lr=lr)
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    for it in range(iters):
        idx = torch.randint(0, X.shape[0], (batch,), generator=g)
        loss, dout, acts = loss_and_grad(model, X[idx], R[idx], model.stats, dx, dy)
        gW, gb = model.backward(acts, dout)
        opt.step(gW + gb)
        for W, m in zip(model.W, model.wmask):
            W.mul_(m)
        model.refresh()
        if it % 100 == 0 or it == iters - 1:
            rec = {"iter": it, "loss": loss}
            if holdout is not None:
                rec["holdout"] = eval_loss(model, holdout[0], holdout[1], dx, dy)
            hist.append(rec)
            if log:
                log(f"[closure] iter {it} {rec}")
    return model, hist


def eval_loss(model, X, R, dx, dy, batch: int = 16) -> float:
    tot, n = 0.0, 0
    for k in range(0, X.shape[0], batch):
        l = loss_only(model, X[k:k + batch], R[k:k + batch], model.stats, dx, dy)
        tot += l * min(batch, X.shape[0] - k)
        n += min(batch, X.shape[0] - k)
    return tot / max(n, 1)


def baseline_loss(R, stats) -> float:
    """Loss of the zero correction (the plain coarse solver)."""
    torch = _L.require_gpu()
    rstd = torch.as_tensor(stats["r_std"], device=R.device, dtype=torch.float32).view(1, -1, 1, 1)
    return float(((R / rstd) ** 2).mean())


# ------------------------------------------------------------ coarse solver
class ClosureSolver:
    """Coarse MHDSolver with the learned flux correction and the plain-step
    fallback."""

    def __init__(self, cfg, model: FluxCNN | None, divb_floor: float = 1.0):
        from solver import MHDSolver

        self.s = MHDSolver(cfg)
        self.model = model
        self.divb_floor = divb_floor
        self.fallbacks = 0
        self.src = None

    def check(self):
        out = (C.c_double * 4)()
        _L.check(self.s.ctx.lib.amhd_state_check(self.s.ctx.h, out), "amhd_state_check")
        return tuple(out)

    def set_source(self):
        torch = _L.require_gpu()
        cfg = self.s.cfg
        U = self.s.ctx.download()[:, 2:-2, 2:-2].unsqueeze(0)
        with torch.no_grad():
            out = self.model.forward(features(U, self.model.stats))
            fs = torch.as_tensor(np.concatenate([self.model.stats["f_scale"]] * 2), device=U.device,
                                 dtype=torch.float32)
            S = source_from_flux(out[..., :N_OUT] * fs, cfg.dx, cfg.dy)[0]
        src = torch.zeros(10, cfg.nx, cfg.ny, device=U.device, dtype=torch.float32)
        src[FEAT] = S.permute(2, 0, 1)
        self.src = src.contiguous()
        torch.cuda.current_stream().synchronize()
        _L.check(self.s.ctx.lib.amhd_set_source(self.s.ctx.h, _p(self.src)), "amhd_set_source")

    def clear_source(self):
        _L.check(self.s.ctx.lib.amhd_set_source(self.s.ctx.h, None), "amhd_set_source")

    def run(self, t_end: float, max_steps: int = 10 ** 7):
        s = self.s
        s._sync_state()
        div0 = self.check()[2]
        while s.t < t_end and s.step < max_steps:
            dt = min(s.compute_dt(), t_end - s.t)
            if dt <= 1e-16:
                break
            s.dt = dt
            if self.model is not None:
                self.set_source()
                s.step_ssprk3()
                rho_min, p_min, divb, nonfinite = self.check()
                ok = rho_min > 0 and p_min > 0 and nonfinite == 0 and divb <= 1.5 * div0 + self.divb_floor
                self.clear_source()
                if not ok:
                    _L.check(s.ctx.lib.amhd_rewind(s.ctx.h), "amhd_rewind")
                    s.step_ssprk3()
                    self.fallbacks += 1
                div0 = self.check()[2]
            else:
                s.step_ssprk3()
            s.t += dt
            s.step += 1
        return s


# ------------------------------------------------------------------- judge
def judge(seeds, model, steps: int = 2000, nx: int = 2048, ny: int = 1024, log=None):
    """Fine run (the judge) vs plain coarse and closure-corrected coarse runs
    from the restricted initial condition to the fine run's final time."""
    torch = _L.require_gpu()
    from amhd_diag import mixing_width

    rows = []
    for seed in seeds:
        fine = {}

        def grab0(s):
            fine["U0"] = restrict(s.ctx.download())

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = fine_run(seed, nx, ny, steps, grab0, [0])
        torch.cuda.synchronize()
        fine_wall = time.perf_counter() - t0
        Uf = s.ctx.download()
        t_end = s.t
        mw_f = mixing_width(Uf.cpu().numpy(), s.cfg)
        rho_f = restrict(Uf)[0]
        ccfg = coarse_config(s.cfg)
        rec = {"seed": seed, "t_end": t_end, "mw_fine": mw_f, "fine_wall_s": fine_wall}
        for tag, m in (("plain", None), ("closure", model)):
            cs = ClosureSolver(ccfg, m)
            cs.s.U = ghosted(fine["U0"])
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            cs.run(t_end)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t1
            Uc = cs.s.ctx.download()
            rho_c = Uc[0, 2:-2, 2:-2]
            rec[f"mw_{tag}"] = mixing_width(Uc.cpu().numpy(), ccfg)
            rec[f"l1rho_{tag}"] = float((rho_c - rho_f).abs().mean() / rho_f.abs().mean())
            rec[f"steps_{tag}"] = cs.s.step
            rec[f"wall_{tag}_s"] = wall
            rec[f"fallbacks_{tag}"] = cs.fallbacks
        for tag in ("plain", "closure"):
            rec[f"mw_err_{tag}"] = abs(rec[f"mw_{tag}"] - mw_f) / max(abs(mw_f), 1e-30)
        rows.append(rec)
        if log:
            log(f"[closure] judge seed {seed}: " + ", ".join(f"{k}={v:.4g}" if isinstance(v, float) else f"{k}={v}"
                                                             for k, v in rec.items()))
    return rows
