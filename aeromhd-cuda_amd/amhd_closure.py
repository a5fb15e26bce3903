"""Learned flux-correction closure on an 8x-coarsened grid (no counterpart
in the reference).

A 4-layer 3x3 CNN (32 channels; BF16 weights, FP32 accumulation, run by
the BF16-MFMA kernels of libaeromhd_hip.so) maps the coarse conserved
state -- 9 fields: rho, m, B, E, rhoC -- to corrections of the x- and
y-face fluxes of those 9 fields.  Their discrete divergence S = -div(dF)
is added to the coarse step as dt*S and, being a flux difference, conserves
every corrected quantity exactly.  Face corrections on the two
x boundaries are held at zero; y is periodic.

Training data come from this solver: fine 2048x1024 magnetic-sieve runs
(FP32) restricted to the coarse grid by 8x8 block means of the conserved
variables.  For a pair of restricted states (U_a at t, U_b at t + dt,
dt = 8 fine steps) the target is the source the plain coarse step misses:
r = (U_b - RK3_coarse(U_a, dt)) / dt.  The loss is the normalised mean
squared error between S(CNN(U_a)) and r.  The rhoC face flux is not
predicted freely: it is C_donor * F_rho (the concentration upwind of the
mass-flux correction, taken from the state the correction is added to)
plus the network's mixing flux G, so that a mass correction carries its
concentration and C cannot leak into pure fluid.  A model may apply only
some of the corrections (stats["fields"]; the shipped one applies the
mixing flux G alone): it is trained and run with the others held at zero.
After the first fit, rounds of data aggregation add pairs taken along the
corrected solver's own trajectories (rollout_pairs).

At run time a corrected step is the plain coarse SSP-RK3 step followed by
U += dt*S(U_pre) (the form of the target), applied in place by
amhd_closure_source through a positivity limiter: per cell the largest
theta in {1, 1/2, .., 1/32, 0} for which the post-step state plus
theta*dt*S keeps half of its density and pressure, and each face takes the
smaller theta of its two cells, so the limited correction is still a flux
difference.  The corrected state is checked before it is written
(amhd_closure_check: density and pressure positive, no non-finite values,
max|divB| not above 1.5x its pre-step value plus a floor); a correction
that fails is not applied, and the plain coarse step stands.  After an
accepted correction, amhd_closure_sharpen moves rhoC across the x faces
against the gradient of C (limited anti-diffusion, flux-corrected
transport with the pre- and post-step states' bounds) with the coefficient
stats["kappa"], calibrated on training seeds after the gain: the plain
coarse step smears the interface across x faces faster than the limited
mixing flux can undo.  The same pass without the network (_Sharpener,
stats["kappa_alone"]) is evaluated beside it.  BatchClosure does all of
this for many instances in one batched step.

The fine-grid solver is the reference: integral mixing width and L1
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

    def forward(self, x, keep: bool = False, stream=None, bufs=None):
        """x: (n, nx, ny, 16) BF16.  Returns the (n, nx, ny, 32) FP32 output
        (channels 0..17 used) and, with keep, the layer inputs.  `bufs`
        (from alloc_outputs) and `stream` let a caller run the layers on
        its own stream into memory it owns."""
        torch = _L.require_gpu()
        n, nx, ny, _ = x.shape
        acts = [x]
        if stream is None:
            stream = _L.stream_handle(torch)
        if bufs is None:
            bufs = self.alloc_outputs(n, nx, ny, x.device)
        h = x
        for li, (cin, cout, relu) in enumerate(LAYERS):
            last = li == len(LAYERS) - 1
            y = bufs[li]
            _L.check(self.lib.amhd_conv3x3_fwd(_p(h), n, nx, ny, cin, _p(self.Wb[li]), _p(self.b[li]), cout,
                                               1 if relu else 0, _p(y), 1 if last else 0, stream), "amhd_conv3x3_fwd")
            h = y
            if keep and not last:
                acts.append(y)
        return (h, acts) if keep else h

    @staticmethod
    def alloc_outputs(n, nx, ny, device="cuda"):
        torch = _L.require_gpu()
        return [torch.empty(n, nx, ny, cout, device=device,
                            dtype=torch.float32 if li == len(LAYERS) - 1 else torch.bfloat16)
                for li, (_, cout, _) in enumerate(LAYERS)]

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
PAIR_FIELDS = list(range(N_IN))   # planes of the network fields in a (n, 9, nx, ny) pair tensor


def _int9(v):
    return (C.c_int * N_IN)(*[int(a) for a in v])


def _f9(v):
    return (C.c_float * N_IN)(*[float(a) for a in v])


def features(states, dtype: int, fields, nx: int, ny: int, stats, out, stream=None):
    """Normalised NHWC BF16 network input (n, nx, ny, 16) of n device arrays
    (`states`: addresses of 10- or 9-plane arrays of nx*ny cells, `fields`:
    the plane of each network field), by amhd_closure_features."""
    torch = _L.require_gpu()
    n = len(states)
    ptrs = (C.c_void_p * n)(*states)
    _L.check(_L.load().amhd_closure_features(ptrs, dtype, _int9(fields), n, nx, ny, _f9(stats["in_mean"]),
                                             _f9(stats["in_std"]), _p(out),
                                             stream if stream is not None else _L.stream_handle(torch)),
             "amhd_closure_features")
    return out


def pair_features(X, stats):
    """Network input of training states X (n, 9, nx, ny) FP32."""
    torch = _L.require_gpu()
    X = X.contiguous()
    n, _, nx, ny = X.shape
    out = torch.empty(n, nx, ny, 16, device=X.device, dtype=torch.bfloat16)
    step = N_IN * nx * ny * X.element_size()
    return features([X.data_ptr() + m * step for m in range(n)], 1, PAIR_FIELDS, nx, ny, stats, out)


RC = N_IN - 1        # position of rhoC among the network fields


def source_from_flux(dF, dx: float, dy: float, C):
    """S = -div of the face corrections dF (n, nx, ny, 18) in physical
    units: channels 0..8 at the x face i-1/2 of cell i (zero on both x
    boundaries), 9..17 at the y face j-1/2 (periodic).  The rhoC channels
    carry the mixing flux G; the rhoC face flux is C_donor * F_rho + G with
    the donor upwind of F_rho and C (n, nx, ny) the concentration of the
    state the correction is added to.  Returns (n, nx, ny, 9)."""
    torch = _L.require_gpu()
    fx = dF[..., :N_IN]
    fy = dF[..., N_IN:N_OUT]
    Cx = torch.where(fx[..., 0] > 0, torch.roll(C, 1, dims=1), C)
    Cy = torch.where(fy[..., 0] > 0, torch.roll(C, 1, dims=2), C)
    fx = torch.cat([fx[..., :RC], (fx[..., RC] + Cx * fx[..., 0]).unsqueeze(-1)], dim=-1)
    fy = torch.cat([fy[..., :RC], (fy[..., RC] + Cy * fy[..., 0]).unsqueeze(-1)], dim=-1)
    fx = torch.cat([torch.zeros_like(fx[:, :1]), fx[:, 1:], torch.zeros_like(fx[:, :1])], dim=1)
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
    """Training pairs on the coarse grid from fine runs: (X, R, Cp) =
    (U_a[FEAT], r[FEAT], C of RK3_coarse(U_a)), stacked on the device."""
    torch = _L.require_gpu()
    from solver import MHDSolver

    starts = [int(v) for v in np.linspace(0, steps - span, pairs_per_run)]
    X, R, CP = [], [], []
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
            CP.append((Uc[9] / Uc[0]).float())
            state["n"] = state.get("n", 0) + 1

        fine_run(seed, nx, ny, steps, sample, starts)
        if log:
            log(f"[closure] data seed {seed}: {state.get('n', 0)} pairs, {time.time() - t0:.1f} s")
    return torch.stack(X), torch.stack(R), torch.stack(CP)


def rollout_pairs(model, seeds, starts: int = 10, horizon: int = 25, span: int = 8, steps: int = 2000,
                  nx: int = 2048, ny: int = 1024, keep: float = 0.5, pullback: float = 0.1, log=None):
    """Pairs along the closure's own coarse trajectories (data aggregation):
    the fine run is restricted every `span` fine steps (F_k at t_k); from
    `starts` points F_k0 the corrected coarse solver steps with
    dt_k = t_{k+1} - t_k for `horizon` steps, and step k gives the pair
    (V_k, r_k, C of RK3(V_k)) with the tendency-matching target
    r_k = [(F_{k+1} - F_k) - (RK3(V_k) - V_k) + pullback * (F_k - V_k)] / dt_k:
    the fine solution's change minus the coarse step's own change at the
    visited state, plus a fraction of the drift (all of it, pullback = 1,
    asks one step to undo the accumulated error, which V_k alone cannot
    tell; at V_k = F_k every pullback gives the first-fit target)."""
    torch = _L.require_gpu()
    X, R, CP = [], [], []
    for seed in seeds:
        t0 = time.time()
        F, T = [], []

        def grab(s):
            F.append(restrict(s.ctx.download()).float())
            T.append(s.t)

        fs = fine_run(seed, nx, ny, steps, grab, range(0, steps + 1, span))
        cs = ClosureSolver(coarse_config(fs.cfg), model, keep=keep)
        n0 = len(X)
        for k0 in np.linspace(0, len(F) - horizon - 1, starts).astype(int):
            cs.s.U = ghosted(F[k0].double())
            cs.s.t = T[k0]
            cs.begin()
            for k in range(k0, k0 + horizon):
                dt = T[k + 1] - T[k]
                if dt > cs.s.compute_dt():       # the coarse CFL step is shorter: stop here
                    break
                V = cs.s.ctx.download()[:, 2:-2, 2:-2]
                post = {}

                def grab_plain():
                    post["U"] = cs.s.ctx.download()[:, 2:-2, 2:-2]

                cs.step(dt, on_plain=grab_plain)
                Uc = post["U"]
                X.append(V[FEAT].float())
                Fk, Fk1 = F[k].double(), F[k + 1].double()
                R.append(((Fk1 - Fk - (Uc - V) + pullback * (Fk - V)) / dt)[FEAT].float())
                CP.append((Uc[9] / Uc[0]).float())
        del F
        if log:
            log(f"[closure] rollout seed {seed}: {len(X) - n0} pairs, {cs.fallbacks} fallbacks, "
                f"{time.time() - t0:.1f} s")
    return torch.stack(X), torch.stack(R), torch.stack(CP)


def join(a, b):
    """Concatenate two pair sets (X, R, Cp)."""
    torch = _L.require_gpu()
    return tuple(torch.cat([x, y]) for x, y in zip(a, b))


def fit_stats(X, R, dx: float) -> dict:
    in_mean = X.mean(dim=(0, 2, 3)).cpu().numpy()
    in_std = X.std(dim=(0, 2, 3)).clamp_min(1e-6).cpu().numpy()
    r_std = R.std(dim=(0, 2, 3)).clamp_min(1e-12).cpu().numpy()
    return {"in_mean": in_mean, "in_std": in_std, "r_std": r_std, "f_scale": r_std * dx}


# ---------------------------------------------------------------- training
def applied_fields(stats) -> list:
    """The network fields whose corrections a model applies (all by default)."""
    f = stats.get("fields") if stats is not None else None
    return list(range(N_IN)) if f is None else [int(v) for v in np.asarray(f).ravel()]


def _scales(stats, device):
    """Per-channel output scale with the fields a model does not apply zeroed,
    and the normalising r_std."""
    torch = _L.require_gpu()
    mask = np.zeros(N_IN, dtype=np.float32)
    mask[applied_fields(stats)] = 1.0
    fs = np.asarray(stats["f_scale"], dtype=np.float32) * mask
    fscale = torch.as_tensor(np.concatenate([fs, fs]), device=device, dtype=torch.float32)
    rstd = torch.as_tensor(stats["r_std"], device=device, dtype=torch.float32)
    return fscale, rstd


def loss_and_grad(model, X, R, Cp, stats, dx, dy):
    """Normalised MSE of S(CNN(X)) against R over the applied fields and
    d(loss)/d(output)."""
    torch = _L.require_gpu()
    out, acts = model.forward(pair_features(X, stats), keep=True)
    o = out.detach().requires_grad_(True)
    fscale, rstd = _scales(stats, X.device)
    S = source_from_flux(o[..., :N_OUT] * fscale, dx, dy, Cp)
    err = ((S - R.permute(0, 2, 3, 1)) / rstd)[..., applied_fields(stats)]
    loss = (err * err).mean()
    loss.backward()
    return float(loss.detach()), o.grad, acts


def loss_only(model, X, R, Cp, stats, dx, dy) -> float:
    """The loss of loss_and_grad without the gradient."""
    out = model.forward(pair_features(X, stats))
    fscale, rstd = _scales(stats, X.device)
    S = source_from_flux(out[..., :N_OUT] * fscale, dx, dy, Cp)
    err = ((S - R.permute(0, 2, 3, 1)) / rstd)[..., applied_fields(stats)]
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


def train(data, dx, dy, iters: int = 2000, batch: int = 16, lr: float = 1e-3, seed: int = 0, log=None,
          holdout=None, model=None, fields=None):
    """Adam on the pairs data = (X, R, Cp) (generate_pairs); a given model
    is trained further with its own normalisation.  fields: the network
    fields whose corrections the model applies (None: all)."""
    torch = _L.require_gpu()
    X, R, Cp = data
    if model is None:
        model = FluxCNN(seed)
        model.stats = fit_stats(X, R, dx)
        if fields is not None:
            model.stats["fields"] = np.asarray(sorted(fields), dtype=np.int32)
    opt = Adam(model.parameters(), lr=lr)
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    for it in range(iters):
        opt.lr = lr * 0.5 * (1.0 + math.cos(math.pi * it / iters))     # cosine decay
        idx = torch.randint(0, X.shape[0], (batch,), generator=g)
        loss, dout, acts = loss_and_grad(model, X[idx], R[idx], Cp[idx], model.stats, dx, dy)
        gW, gb = model.backward(acts, dout)
        opt.step(gW + gb)
        for W, m in zip(model.W, model.wmask):
            W.mul_(m)
        model.refresh()
        if it % 100 == 0 or it == iters - 1:
            rec = {"iter": it, "loss": loss}
            if holdout is not None:
                rec["holdout"] = eval_loss(model, holdout, dx, dy)
            hist.append(rec)
            if log:
                log(f"[closure] iter {it} {rec}")
    return model, hist


def eval_loss(model, data, dx, dy, batch: int = 16) -> float:
    X, R, Cp = data
    tot, n = 0.0, 0
    for k in range(0, X.shape[0], batch):
        l = loss_only(model, X[k:k + batch], R[k:k + batch], Cp[k:k + batch], model.stats, dx, dy)
        tot += l * min(batch, X.shape[0] - k)
        n += min(batch, X.shape[0] - k)
    return tot / max(n, 1)


def baseline_loss(R, stats) -> float:
    """Loss of the zero correction (the plain coarse solver)."""
    torch = _L.require_gpu()
    rstd = torch.as_tensor(stats["r_std"], device=R.device, dtype=torch.float32).view(1, -1, 1, 1)
    return float(((R / rstd) ** 2)[:, applied_fields(stats)].mean())


# ------------------------------------------------------------ coarse solver
class _Corrector:
    """The learned correction of n coarse states of one grid: network pass
    on the pre-step states, limited source against the post-step states,
    the guard evaluated on U + dt S before anything is written, and
    U += dt S for the instances that pass (the others keep their plain
    step: the fallback), then, on those, the limited anti-diffusion of C
    across x faces (amhd_closure_sharpen) with the coefficient kappa.  All
    launches on `stream`."""

    def __init__(self, model: FluxCNN, n: int, cfg, dtype: int, stream, keep, divb_floor, fields, gain,
                 kappa=None):
        torch = _L.require_gpu()
        self.model, self.n, self.cfg, self.dtype, self.stream = model, n, cfg, dtype, stream
        self.keep, self.divb_floor = keep, divb_floor
        if fields is None:
            fields = applied_fields(model.stats)
        self.mask = np.zeros(N_IN, dtype=np.float32)
        self.mask[list(fields)] = 1.0
        if gain is None:
            gain = float(np.asarray(model.stats.get("gain", 1.0)))
        self.mask *= np.float32(gain)
        if kappa is None:
            kappa = float(np.asarray(model.stats.get("kappa", 0.0)))
        if not (kappa >= 0.0 and np.isfinite(kappa)):
            raise ValueError(f"kappa must be finite and >= 0, got {kappa}")
        self.kappa = float(kappa)
        nx, ny = cfg.nx, cfg.ny
        tdt = torch.float32 if dtype == _L.AMHD_F32 else torch.float64
        self.x = torch.empty(n, nx, ny, 16, device="cuda", dtype=torch.bfloat16)
        self.outs = FluxCNN.alloc_outputs(n, nx, ny)
        self.theta = torch.ones(2 * n * nx * ny, device="cuda", dtype=torch.float32)   # theta, theta_mix
        self.src = torch.zeros(n, 10, nx, ny, device="cuda", dtype=tdt)
        self.work = torch.empty(32 * 64 * 4, device="cuda", dtype=torch.float64)
        self.sharp_work = (torch.empty(3 * min(n, 32) * nx * ny, device="cuda", dtype=torch.float64)
                           if self.kappa > 0 else None)
        step = 10 * nx * ny * self.src.element_size()
        self.src_ptrs = [self.src.data_ptr() + m * step for m in range(n)]
        torch.cuda.synchronize()

    def _plan(self):
        """The ctypes arguments that stay the same from step to step (buffers,
        weights, scales, stream), built once and again only when the model's
        BF16 weights or its statistics are replaced: a step then makes only
        the calls themselves."""
        m = self.model
        if getattr(self, "_plan_wb", None) is m.Wb and self._plan_stats is m.stats:
            return
        cfg, n = self.cfg, self.n
        conv, h = [], self.x
        for li, (cin, cout, relu) in enumerate(LAYERS):
            y = self.outs[li]
            conv.append((_p(h), n, cfg.nx, cfg.ny, cin, _p(m.Wb[li]), _p(m.b[li]), cout, 1 if relu else 0, _p(y),
                         1 if li == len(LAYERS) - 1 else 0, self.stream))
            h = y
        self._conv_args = conv
        self._feat_args = (_int9(FEAT), n, cfg.nx, cfg.ny, _f9(m.stats["in_mean"]), _f9(m.stats["in_std"]),
                           _p(self.x), self.stream)
        self._fields = _int9(FEAT)
        self._fscale = _f9(m.stats["f_scale"] * self.mask)
        self._srcs = (C.c_void_p * n)(*self.src_ptrs)
        self._plan_wb, self._plan_stats = m.Wb, m.stats

    def predict(self, pre):
        """Features and the four convolutions of the network pass, on the
        stream; the same launches as features() and FluxCNN.forward."""
        if len(pre) != self.n:
            raise ValueError(f"predict: {len(pre)} states for a corrector of {self.n}")
        self._plan()
        lib = self.model.lib
        _L.check(lib.amhd_closure_features((C.c_void_p * len(pre))(*pre), self.dtype, *self._feat_args),
                 "amhd_closure_features")
        for a in self._conv_args:
            _L.check(lib.amhd_conv3x3_fwd(*a), "amhd_conv3x3_fwd")

    def check(self, ptrs, dt) -> np.ndarray:
        """(n, 4): min rho, min p, max|divB|, non-finite cells of U + dt S."""
        self._plan()
        cfg, n = self.cfg, len(ptrs)
        out = np.empty((n, 4))
        srcs = self._srcs if n == self.n else (C.c_void_p * n)(*self.src_ptrs[:n])
        _L.check(self.model.lib.amhd_closure_check((C.c_void_p * n)(*ptrs), srcs, _dvec(dt), self.dtype, n, cfg.nx,
                                                   cfg.ny, float(cfg.gamma), float(cfg.dx), float(cfg.dy),
                                                   _p(self.work), out.ctypes.data_as(C.POINTER(C.c_double)),
                                                   self.stream), "amhd_closure_check")
        return out

    def correct(self, pre, post, dt, div0):
        """Source against the post-step states (bounds also from the pre-step
        ones), guard, apply, sharpen C on the accepted states (bounds again
        from both states); returns (accepted mask, |divB| of the states now
        current)."""
        self._plan()
        cfg, n = self.cfg, self.n
        lib = self.model.lib
        V = C.c_void_p * n
        lim = self.keep is not None
        _L.check(lib.amhd_closure_source(_p(self.outs[-1]), V(*post), self._srcs, self.dtype, self._fields, n,
                                         cfg.nx, cfg.ny, self._fscale,
                                         float(cfg.dx), float(cfg.dy), _dvec(dt), float(cfg.gamma),
                                         float(self.keep) if lim else 0.0, _p(self.theta) if lim else None,
                                         V(*pre) if lim else None, 0, self.stream), "amhd_closure_source")
        r = self.check(post, dt)
        ok = ((r[:, 0] > 0) & (r[:, 1] > 0) & (r[:, 3] == 0) & (r[:, 2] <= 1.5 * np.asarray(div0) + self.divb_floor)
              & (np.asarray(dt) > 0))
        idx = np.nonzero(ok)[0]
        if len(idx):
            k = len(idx)
            W = C.c_void_p * k
            _L.check(lib.amhd_closure_apply(W(*[post[i] for i in idx]), W(*[self.src_ptrs[i] for i in idx]),
                                            _dvec([dt[i] for i in idx]), self.dtype, k, cfg.nx, cfg.ny, self.stream),
                     "amhd_closure_apply")
            if self.kappa > 0:
                _L.check(lib.amhd_closure_sharpen(W(*[post[i] for i in idx]), W(*[pre[i] for i in idx]), self.dtype,
                                                  k, cfg.nx, cfg.ny, self.kappa, _p(self.sharp_work), self.stream),
                         "amhd_closure_sharpen")
        div_now = r[:, 2].copy()
        if not ok.all():
            div_now[~ok] = self.check(post, np.zeros(n))[~ok, 2]
        return ok, div_now


class _Sharpener:
    """The limited anti-diffusion of C across x faces without the network:
    after each plain step, amhd_closure_sharpen on the post-step states with
    the pre-step states' bounds.  The interface of _Corrector; there is no
    guard to take (only rhoC moves, and C stays within its bounds)."""

    def __init__(self, n: int, cfg, dtype: int, stream, kappa: float):
        torch = _L.require_gpu()
        if not (kappa > 0.0 and np.isfinite(kappa)):
            raise ValueError(f"the sharpening needs a finite kappa > 0, got {kappa}")
        self.lib = _L.load()
        self.n, self.cfg, self.dtype, self.stream, self.kappa = n, cfg, dtype, stream, float(kappa)
        self.work = torch.empty(3 * min(n, 32) * cfg.nx * cfg.ny, device="cuda", dtype=torch.float64)

    def predict(self, pre):
        pass

    def check(self, ptrs, dt) -> np.ndarray:
        return np.zeros((len(ptrs), 4))

    def correct(self, pre, post, dt, div0):
        idx = [i for i in range(len(post)) if dt[i] > 0]
        if idx:
            V = C.c_void_p * len(idx)
            _L.check(self.lib.amhd_closure_sharpen(V(*[post[i] for i in idx]), V(*[pre[i] for i in idx]), self.dtype,
                                                   len(idx), self.cfg.nx, self.cfg.ny, self.kappa, _p(self.work),
                                                   self.stream), "amhd_closure_sharpen")
        return np.ones(len(post), dtype=bool), np.asarray(div0, dtype=np.float64)


def _dvec(v):
    v = [float(x) for x in v]
    return (C.c_double * len(v))(*v)


def check_kappa(kappa) -> None:
    """The sharpening coefficient: None (the model's) or a finite number >= 0."""
    if kappa is not None and not (isinstance(kappa, (int, float, np.floating)) and math.isfinite(kappa)
                                  and kappa >= 0):
        raise ValueError(f"kappa must be None or a finite number >= 0, got {kappa!r}")


def check_closure_boundaries(cfg) -> None:
    """The closure's kernels (features, face corrections, limiters, guard)
    wrap y periodically and carry no correction through the x boundaries,
    the boundaries of the RMI runs it is trained on; other boundary
    conditions would wrap or drop correction fluxes and break conservation."""
    if cfg.bc_y_type != "periodic" or cfg.get_bc_x() == "periodic":
        raise ValueError("the learned closure needs periodic y and non-periodic x boundaries "
                         f"(got bc_x {cfg.get_bc_x()!r}, bc_y {cfg.bc_y_type!r})")


class ClosureSolver:
    """Coarse MHDSolver with the learned flux correction, its limiters and
    the plain-step fallback.  Every launch of a corrected step runs on the
    context's stream and reads the state in place."""

    def __init__(self, cfg, model: FluxCNN | None, divb_floor: float = 1.0, keep: float | None = 0.5,
                 fields=None, gain: float | None = None, kappa: float | None = None):
        """fields: the network fields (positions in FEAT) whose corrections
        are applied, None for the model's own (all unless it was trained
        on fewer); the others keep the plain solver's fluxes.  gain: a
        factor on the corrections, None for the model's calibrated one
        (stats["gain"], 1 if absent).  keep: the positivity limiter's bound
        (None: no limiters).  kappa: the coefficient of the limited
        anti-diffusion of C across x faces after each accepted correction,
        None for the model's calibrated one (stats["kappa"], 0 -- none --
        if absent).  model None: the plain solver, followed by that
        anti-diffusion alone when kappa > 0."""
        from solver import MHDSolver

        check_closure_boundaries(cfg)
        check_kappa(kappa)
        self.s = MHDSolver(cfg)
        self.model = model
        self.fallbacks = 0
        self._opts = (keep, divb_floor, fields, gain, kappa)
        self.cor = None

    def check(self):
        out = (C.c_double * 4)()
        _L.check(self.s.ctx.lib.amhd_state_check(self.s.ctx.h, out), "amhd_state_check")
        return tuple(out)

    def _state_ptr(self) -> int:
        ctx = self.s.ctx
        state = C.c_void_p()
        _L.check(ctx.lib.amhd_state_ptr(ctx.h, C.byref(state)), "amhd_state_ptr")
        return state.value

    def begin(self):
        """Upload an assigned state and take the guard's reference |divB|."""
        self.s._sync_state()
        kappa = self._opts[4]
        if self.cor is None and (self.model is not None or (kappa is not None and kappa > 0)):
            ctx = self.s.ctx
            st = C.c_void_p()
            _L.check(ctx.lib.amhd_get_stream(ctx.h, C.byref(st)), "amhd_get_stream")
            self.cor = (_Corrector(self.model, 1, self.s.cfg, ctx.dtype, st, *self._opts) if self.model is not None
                        else _Sharpener(1, self.s.cfg, ctx.dtype, st, kappa))
        self.div0 = self.check()[2]

    def step(self, dt: float | None = None, on_plain=None) -> bool:
        """One coarse step -- the plain SSP-RK3 step followed by
        U += dt * S(U_pre), the form the network is trained on,
        r = (U_b - RK3(U_a)) / dt -- guarded: a failing correction is not
        applied and the plain step stands.  dt None: the device loop's CFL
        step (bounded by cfg.t_end); else that dt with the host-driven step.
        on_plain() runs between the plain step and the correction.  Returns
        whether the correction was applied, or None when no step ran."""
        s = self.s
        pre = self._state_ptr() if self.cor is not None else None
        if self.cor is not None:
            self.cor.predict([pre])
        if dt is None:
            if s.advance(1, use_graph=False) == 0:   # one step: a 3-step graph would replay two no-ops
                return None
            dt = s.dt
        else:
            s.compute_dt()                 # sets cfg.glm_ch for the step
            s.dt = dt
            s.step_ssprk3()
            s.t += dt
            s.step += 1
        if self.cor is None:
            return True
        post = self._state_ptr()
        if on_plain is not None:
            on_plain()
        ok, div_now = self.cor.correct([pre], [post], [dt], [self.div0])
        self.div0 = float(div_now[0])
        if not ok[0]:
            self.fallbacks += 1
        return bool(ok[0])

    def run(self, t_end: float, max_steps: int = 10 ** 7):
        """Steps of the device loop (CFL dt) until t_end or max_steps."""
        s = self.s
        s.cfg.t_end = t_end
        s.cfg.max_steps = max_steps
        self.begin()
        while s.t < t_end and s.step < max_steps:
            if self.step() is None:
                break
        return s


class BatchClosure:
    """The corrected coarse solver over n instances of one grid size: one
    batched SSP-RK3 step (amhd_batch_run), one network pass over all
    instances and one batched limited correction per step; each instance
    keeps its own dt, t_end, guard and fallback.  An instance steps exactly
    as ClosureSolver.run would step it alone."""

    def __init__(self, cases, model: FluxCNN | None, divb_floor: float = 1.0, keep: float | None = 0.5,
                 fields=None, gain: float | None = None, kappa: float | None = None):
        check_kappa(kappa)
        torch = _L.require_gpu()
        from amhd_batch import BatchSolver

        cfgs = [c for c, _ in cases]
        if len({(c.nx, c.ny) for c in cfgs}) != 1:
            raise ValueError("BatchClosure: every instance must have the same grid")
        for c in cfgs:
            check_closure_boundaries(c)
        self.b = BatchSolver(cases, precision=cfgs[0].precision)
        self.model = model
        self.n = len(cases)
        self.fallbacks = np.zeros(self.n, dtype=np.int64)
        dtype = _L.AMHD_F32 if self.b.precision in ("fp32", "float32") else _L.AMHD_F64
        # the corrector's launches go to the stream current at construction;
        # the end of each step waits for that stream itself
        self.tstream = torch.cuda.current_stream()
        self.stream = C.c_void_p(self.tstream.cuda_stream)
        if model is not None:
            self.cor = _Corrector(model, self.n, cfgs[0], dtype, self.stream, keep, divb_floor, fields, gain, kappa)
        elif kappa is not None and kappa > 0:
            self.cor = _Sharpener(self.n, cfgs[0], dtype, self.stream, kappa)
        else:
            self.cor = None

    def _ptrs(self):
        n = self.n
        cur, prev = (C.c_void_p * n)(), (C.c_void_p * n)()
        _L.check(self.b.lib.amhd_batch_state_ptrs(self.b.h, cur, prev), "amhd_batch_state_ptrs")
        return [c for c in cur], [p for p in prev]

    def run(self):
        """Every instance to its own cfg.t_end / cfg.max_steps."""
        torch = _L.require_gpu()
        b = self.b
        if self.cor is None:
            while b.run(1).sum() > 0:
                pass
            return
        cur, _ = self._ptrs()
        div0 = self.cor.check(cur, np.zeros(self.n))[:, 2]
        while True:
            pre, _ = self._ptrs()
            # The network pass (this stream) and the batched step (the
            # batch's stream) overlap: both only read the pre-step states,
            # and the step writes the other two of each instance's three
            # state buffers.  amhd_batch_run returns after its step has
            # finished; the correction below queues behind the network pass.
            self.cor.predict(pre)
            taken = b.run(1)
            if taken.sum() == 0:
                break
            dts = (C.c_double * self.n)()
            _L.check(b.lib.amhd_batch_last_dt(b.h, dts), "amhd_batch_last_dt")
            dt = np.where(taken > 0, np.array(dts[:]), 0.0)
            post, _ = self._ptrs()
            ok, div0 = self.cor.correct(pre, post, dt, div0)
            self.fallbacks += ((~ok) & (taken > 0)).astype(np.int64)
            # the next step reads the corrected states on the batch's stream:
            # wait for the corrector's stream (whichever stream is current now)
            self.tstream.synchronize()


# ------------------------------------------------------- evaluation
VARIANTS = {
    "closure": dict(keep=0.5),                       # the model's own fields, gain and kappa
    "network": dict(keep=0.5, kappa=0.0, gain="alone"),   # ... without the sharpening of C, at the
                                                          # network's own gain (stats["gain_alone"])
    "nolimit": dict(keep=None, kappa=0.0),           # ... and without the limiters
    # the sharpening alone (no network pass) at its own calibrated
    # coefficient (stats["kappa_alone"])
    "sharpen": dict(network=False, kappa="alone"),
    # ablations of the network itself (gain 1)
    "all": dict(keep=0.5, fields=list(range(N_IN)), gain=1.0, kappa=0.0),
    "mix": dict(keep=0.5, fields=[RC], gain=1.0, kappa=0.0),
    "mass": dict(keep=0.5, fields=[0, RC], gain=1.0, kappa=0.0),
    "hydro": dict(keep=0.5, fields=[0, 1, 2, 3, 7, RC], gain=1.0, kappa=0.0),
}


def _resolve(model, kw):
    """(model or None, ClosureSolver keywords) of a variant: kappa="alone"
    becomes the model's calibrated coefficient of the sharpening alone,
    gain="alone" the gain calibrated for the network without the sharpening
    (the closure's own gain if none is stored), and network=False drops the
    model."""
    kw = dict(kw)
    if kw.get("kappa") == "alone":
        kw["kappa"] = float(np.asarray(model.stats.get("kappa_alone", 0.0)))
    if kw.get("gain") == "alone":
        kw["gain"] = float(np.asarray(model.stats.get("gain_alone", model.stats.get("gain", 1.0))))
    return (model if kw.pop("network", True) else None), kw


def evaluate(seeds, model, steps: int = 2000, nx: int = 2048, ny: int = 1024, log=None, variants=("closure",),
             specs=None):
    """Fine run (the reference) vs the plain coarse solver and closure variants
    (VARIANTS: limiter on/off, which corrections apply) from the restricted
    initial condition to the fine run's final time.  Errors are taken
    against the restricted fine state (the coarse grid's best attainable
    answer): the integral mixing width of the restricted state (mw_err_*),
    the L1 density error (l1rho_*) and the mean |C - C_fine| (l1c_*, C
    clipped to [0, 1]); the fine grid's own mixing width is
    reported as mw_fine.  specs: {tag: ClosureSolver keywords} in place of
    VARIANTS."""
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
        ccfg = coarse_config(s.cfg)
        Ufr = restrict(Uf)
        rho_f = Ufr[0]
        mw_fr = mixing_width(ghosted(Ufr).cpu().numpy(), ccfg)
        rec = {"seed": seed, "t_end": t_end, "mw_fine": mw_f, "mw_fine_restricted": mw_fr, "fine_wall_s": fine_wall}
        table = specs if specs is not None else VARIANTS
        runs = [("plain", None, {})] + [(v, model, table[v]) for v in variants]
        c_f = (Ufr[9] / Ufr[0]).clamp(0, 1)
        for tag, m, kw in runs:
            if m is not None:
                m, kw = _resolve(m, kw)
            cs = ClosureSolver(ccfg, m, **kw)
            cs.s.U = ghosted(fine["U0"])
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            cs.run(t_end)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t1
            Uc = cs.s.ctx.download()
            rho_c = Uc[0, 2:-2, 2:-2]
            rec[f"mw_{tag}"] = mixing_width(Uc.cpu().numpy(), ccfg)
            rec[f"mw_err_{tag}"] = abs(rec[f"mw_{tag}"] - mw_fr) / max(abs(mw_fr), 1e-30)
            rec[f"l1rho_{tag}"] = float((rho_c - rho_f).abs().mean() / rho_f.abs().mean())
            rec[f"l1c_{tag}"] = float(((Uc[9, 2:-2, 2:-2] / rho_c).clamp(0, 1) - c_f).abs().mean())
            rec[f"steps_{tag}"] = cs.s.step
            rec[f"wall_{tag}_s"] = wall
            rec[f"fallbacks_{tag}"] = cs.fallbacks
        rows.append(rec)
        if log:
            log(f"[closure] evaluation seed {seed}: " + ", ".join(f"{k}={v:.4g}" if isinstance(v, float) else f"{k}={v}"
                                                             for k, v in rec.items()))
    return rows


def fp64_fine_throughput(seed: int, steps: int = 2000, nx: int = 2048, ny: int = 1024) -> dict:
    """Cell-updates per second of the plain FP64 fine solve of seed's
    magnetic sieve (device loop, graph replay), the closure's reference."""
    torch = _L.require_gpu()
    from amhd_ics import magnetic_sieve
    from physics import prim_to_cons
    from solver import MHDSolver

    cfg, W = magnetic_sieve(nx, ny, precision="fp64", seed=seed, max_steps=steps)
    cfg.t_end = 1e30
    s = MHDSolver(cfg)
    s.U = prim_to_cons(W, cfg.gamma)
    s.advance(5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = s.advance(steps - 5)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"mcups": nx * ny * n / wall / 1e6, "wall_s_for_run": wall * steps / n, "steps": steps}


def evaluate_batched(seeds, model, steps: int = 2000, nx: int = 2048, ny: int = 1024, log=None,
                     reps: int = 3) -> dict:
    """The batched form: the plain coarse solver, the closure and (when the
    model holds its coefficient) the sharpening of C alone each run over all
    seeds at once (BatchSolver / BatchClosure, every instance to its own
    fine run's end time), scored per seed as evaluate() does.

    Wall time: the median of `reps` timed runs from the same initial states,
    after an untimed run of 2 steps that loads the kernels, as the FP64
    fine solve it is compared with is timed after 5 warm-up steps.  The runs
    are identical (the scores are taken from the last)."""
    torch = _L.require_gpu()
    from amhd_diag import mixing_width

    refs, cases = [], []
    for seed in seeds:
        fine = {}

        def grab0(s):
            fine["U0"] = restrict(s.ctx.download())

        s = fine_run(seed, nx, ny, steps, grab0, [0])
        Ufr = restrict(s.ctx.download())
        ccfg = coarse_config(s.cfg)
        ccfg.t_end, ccfg.max_steps = s.t, 10 ** 7
        refs.append((Ufr, mixing_width(ghosted(Ufr).cpu().numpy(), ccfg)))
        cases.append((ccfg, ghosted(fine["U0"]).cpu().numpy()))
    out = {"seeds": list(seeds), "instances": len(cases)}
    forms = [("plain", None, {}), ("closure", model, {})]
    ka = float(np.asarray(model.stats.get("kappa_alone", 0.0))) if model is not None else 0.0
    if ka > 0:
        forms.append(("sharpen", None, {"kappa": ka}))
        out["kappa_alone"] = ka
    for tag, m, kw in forms:
        warm = []
        for c, U in cases:
            wc = Config(**{f: getattr(c, f) for f in c.__dataclass_fields__})
            wc.max_steps = 2
            warm.append((wc, U))
        BatchClosure(warm, m, **kw).run()
        walls = []
        for _ in range(max(1, reps)):
            bc = BatchClosure(cases, m, **kw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            bc.run()
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
        out[f"wall_{tag}_s"] = float(np.median(walls))
        out[f"wall_{tag}_runs_s"] = [round(w, 4) for w in walls]
        mw_err, l1, l1c = [], [], []
        for k, (Ufr, mw_fr) in enumerate(refs):
            Uc = bc.b.download(k)
            mw_err.append(abs(mixing_width(Uc, cases[k][0]) - mw_fr) / max(abs(mw_fr), 1e-30))
            rho_c = Uc[0, 2:-2, 2:-2]
            rho_f = Ufr[0].cpu().numpy()
            l1.append(float(np.abs(rho_c - rho_f).mean() / np.abs(rho_f).mean()))
            c_f = np.clip(Ufr[9].cpu().numpy() / rho_f, 0, 1)
            l1c.append(float(np.abs(np.clip(Uc[9, 2:-2, 2:-2] / rho_c, 0, 1) - c_f).mean()))
        out[f"mw_err_{tag}_mean"] = float(np.mean(mw_err))
        out[f"l1rho_{tag}_mean"] = float(np.mean(l1))
        out[f"l1c_{tag}_mean"] = float(np.mean(l1c))
        out[f"fallbacks_{tag}"] = int(bc.fallbacks.sum())
        out[f"steps_{tag}"] = int(bc.b.steps.sum())
    if log:
        log("[closure] batched evaluation: " + ", ".join(f"{k}={v:.4g}" if isinstance(v, float) else f"{k}={v}"
                                                      for k, v in out.items() if k != "seeds"))
    return out


def evaluate_resolutions(seeds, factors=(4, 2), steps: int = 2000, nx: int = 2048, ny: int = 1024, reps: int = 3,
                         log=None) -> dict:
    """The plain solver (no network) at finer grids than the closure's: for
    each coarsening factor f (4: 512x256, 2: 1024x512), the seeds' fine
    initial states restricted by f advance as one plain FP32 batch to their
    fine runs' end times; the result, restricted to the closure's grid
    (FACTOR), is scored as evaluate_batched scores the closure (integral
    mixing width, L1 density error and mean |C - C_fine| against the
    restricted fine state).
    Wall time as in evaluate_batched: the median of `reps` timed runs after
    an untimed 2-step run.  This gives the plain solver's cost of reaching a
    given error, against which the closure's cost at its error compares."""
    torch = _L.require_gpu()
    from amhd_batch import BatchSolver
    from amhd_diag import mixing_width

    refs, fines = [], []
    for seed in seeds:
        fine = {}

        def grab0(s):
            fine["U0"] = s.ctx.download()

        s = fine_run(seed, nx, ny, steps, grab0, [0])
        Uf = s.ctx.download()
        ccfg = coarse_config(s.cfg)
        Ufr = restrict(Uf).cpu().numpy()
        refs.append((Ufr[0], mixing_width(ghosted(restrict(Uf)).cpu().numpy(), ccfg), ccfg,
                     np.clip(Ufr[9] / Ufr[0], 0, 1)))
        fines.append((s.cfg, fine["U0"], s.t))
    out = {"seeds": list(seeds)}
    for f in factors:
        cases = []
        for cfg, U0, t_end in fines:
            c = coarse_config(cfg, f)
            c.t_end, c.max_steps = t_end, 10 ** 7
            cases.append((c, ghosted(restrict(U0, f)).cpu().numpy()))
        warm = []
        for c, U in cases:
            wc = Config(**{k: getattr(c, k) for k in c.__dataclass_fields__})
            wc.max_steps = 2
            warm.append((wc, U))
        BatchSolver(warm).run(10)
        walls = []
        for _ in range(max(1, reps)):
            b = BatchSolver(cases)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            while b.run(64).sum() > 0:
                pass
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
        mw_err, l1, l1c = [], [], []
        for k, (rho_f, mw_fr, ccfg, c_f) in enumerate(refs):
            Uc = restrict(torch.from_numpy(b.download(k)), FACTOR // f).numpy()
            mw_err.append(abs(mixing_width(ghosted(torch.from_numpy(Uc).cuda()).cpu().numpy(), ccfg) - mw_fr)
                          / max(abs(mw_fr), 1e-30))
            l1.append(float(np.abs(Uc[0] - rho_f).mean() / np.abs(rho_f).mean()))
            l1c.append(float(np.abs(np.clip(Uc[9] / Uc[0], 0, 1) - c_f).mean()))
        tag = f"{nx // f}x{ny // f}"
        out[tag] = {"factor": f, "wall_s": float(np.median(walls)), "wall_runs_s": [round(w, 4) for w in walls],
                    "mw_err_mean": float(np.mean(mw_err)), "l1rho_mean": float(np.mean(l1)),
                    "l1c_mean": float(np.mean(l1c)), "steps": int(b.steps.sum())}
        if log:
            log(f"[closure] plain {tag}: " + ", ".join(f"{k}={v}" for k, v in out[tag].items()))
    return out


def calibrate_gain(model, seeds, gains, steps: int = 2000, log=None, kappa=None):
    """The correction gain that minimises the mean mixing-width error over
    `seeds` (training seeds: the held-out ones stay untouched); a network
    fitted by least squares under-predicts sharp corrections, and a scalar
    gain is the one-parameter remedy.  kappa: the sharpening coefficient
    during the calibration (None: the model's).  Stores the gain in
    model.stats["gain"]."""
    specs = {f"g{g:g}": dict(keep=0.5, gain=float(g), kappa=kappa) for g in gains}
    rows = evaluate(seeds, model, steps=steps, variants=tuple(specs), specs=specs)
    err = {t: sum(r[f"mw_err_{t}"] for r in rows) / len(rows) for t in specs}
    best = min(err, key=err.get)
    model.stats["gain"] = np.asarray(specs[best]["gain"], dtype=np.float32)
    if log:
        log(f"[closure] gain calibration on seeds {list(seeds)}: " +
            ", ".join(f"{t}={e:.4f}" for t, e in err.items()) + f" -> {best}")
    return float(specs[best]["gain"]), err


def calibrate_kappa(model, seeds, kappas, steps: int = 2000, log=None, alone: bool = False):
    """The coefficient of the limited anti-diffusion of C that minimises the
    mean mixing-width error over `seeds` (training seeds), with the model's
    corrections at their calibrated gain, or, alone=True, without the
    network (the sharpening by itself).  Stores it in
    model.stats["kappa"] (alone: model.stats["kappa_alone"])."""
    specs = {f"k{k:g}": (dict(network=False, kappa=float(k)) if alone else dict(keep=0.5, kappa=float(k)))
             for k in kappas}
    rows = evaluate(seeds, model, steps=steps, variants=tuple(specs), specs=specs)
    err = {t: sum(r[f"mw_err_{t}"] for r in rows) / len(rows) for t in specs}
    best = min(err, key=err.get)
    model.stats["kappa_alone" if alone else "kappa"] = np.asarray(specs[best]["kappa"], dtype=np.float64)
    if log:
        log(f"[closure] {'sharpening-alone ' if alone else ''}kappa calibration on seeds {list(seeds)}: " +
            ", ".join(f"{t}={e:.4f}" for t, e in err.items()) + f" -> {best}")
    return float(specs[best]["kappa"]), err


def calibrate_gain_kappa(model, seeds, gains, kappas, steps: int = 2000, log=None):
    """The (gain, kappa) pair of the grid gains x kappas that minimises the
    mean mixing-width error over `seeds` (training seeds): the gain scales
    the network's mixing flux and kappa the sharpening that follows it, and
    the two trade against each other.  Stores both in model.stats."""
    specs = {f"g{g:g}_k{k:g}": dict(keep=0.5, gain=float(g), kappa=float(k)) for g in gains for k in kappas}
    rows = evaluate(seeds, model, steps=steps, variants=tuple(specs), specs=specs)
    err = {t: sum(r[f"mw_err_{t}"] for r in rows) / len(rows) for t in specs}
    best = min(err, key=err.get)
    model.stats["gain"] = np.asarray(specs[best]["gain"], dtype=np.float32)
    model.stats["kappa"] = np.asarray(specs[best]["kappa"], dtype=np.float64)
    if log:
        log(f"[closure] gain x kappa calibration on seeds {list(seeds)}: " +
            ", ".join(f"{t}={e:.4f}" for t, e in err.items()) + f" -> {best}")
    return float(specs[best]["gain"]), float(specs[best]["kappa"]), err


def summarize(rows, tags) -> dict:
    """Means over the evaluated seeds."""
    out = {}
    for t in tags:
        for k in ("mw_err", "l1rho", "l1c", "fallbacks", "steps", "wall"):
            key = f"wall_{t}_s" if k == "wall" else f"{k}_{t}"
            out[key + "_mean"] = sum(r[key] for r in rows) / len(rows)
    return out
