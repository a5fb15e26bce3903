"""x-slab decomposition of one grid over several contexts (no counterpart
in the reference).

The multi-GPU layout of the single-instance solver is 1-D slabs along x,
the slow axis: each GPU owns rows [i0, i1) of every variable plane, and
before every SSP-RK3 stage it receives the two rows on each side from its
neighbours; the step's dt is the reference's formula (solver.py:206-223)
applied to the four CFL maxima taken over all slabs.  On one GPU this
module runs that scheme with one context per slab: the halo exchange is a
device copy between contexts (amhd_edge_rows into the neighbour's
amhd_set_halo buffer) where several GPUs would use RCCL send/recv, and the
maxima are combined on the host where several GPUs would all-reduce them.
The result equals the undivided solver's bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import amhd_lib as _L
from solver import Context, NVAR, solver_params


class SlabSolver:
    """One grid (a Config) split into n_slabs x-slabs, stepped in lock-step."""

    def __init__(self, cfg, n_slabs: int):
        torch = _L.require_gpu()
        if n_slabs < 1 or cfg.nx < 2 * n_slabs:
            raise ValueError(f"cannot split nx={cfg.nx} into {n_slabs} slabs of at least 2 rows")
        self.cfg = cfg
        self.lib = _L.load()
        self.n = n_slabs
        self.bounds = [(k * cfg.nx // n_slabs, (k + 1) * cfg.nx // n_slabs) for k in range(n_slabs)]
        self.periodic = cfg.get_bc_x() == "periodic"
        p = solver_params(cfg)
        self.ctx = [Context(i1 - i0, cfg.ny, cfg.precision, p) for i0, i1 in self.bounds]
        dtype = torch.float32 if self.ctx[0].dtype == _L.AMHD_F32 else torch.float64
        # halo[k] = (lo, hi) buffers of slab k: 10 planes x 2 rows x ny
        self.halo = []
        for k in range(n_slabs):
            has_lo = k > 0 or (self.periodic and n_slabs > 1)
            has_hi = k < n_slabs - 1 or (self.periodic and n_slabs > 1)
            lo = torch.zeros(NVAR, 2, cfg.ny, device="cuda", dtype=dtype) if has_lo else None
            hi = torch.zeros(NVAR, 2, cfg.ny, device="cuda", dtype=dtype) if has_hi else None
            self.halo.append((lo, hi))
            _L.check(self.lib.amhd_set_halo(self.ctx[k].h, _ptr(lo), _ptr(hi)), "amhd_set_halo")
        torch.cuda.synchronize()
        self.t = 0.0
        self.step = 0
        self.dt = 0.0

    # ------------------------------------------------------------ state
    def upload(self, U) -> None:
        """Ghosted (10, nx+4, ny+4) float64 state of the whole grid (host or
        device); each slab takes its rows (its uploaded ghosts are unused)."""
        for (i0, i1), ctx in zip(self.bounds, self.ctx):
            ctx.upload(U[:, i0:i1 + 4, :])

    def download(self) -> np.ndarray:
        """Interior (10, nx, ny) float64 state of the whole grid."""
        return np.concatenate([ctx.download_numpy()[:, 2:-2, 2:-2] for ctx in self.ctx], axis=1)

    def boundary_energy(self) -> float:
        """Cumulative boundary energy flux, summed over the slabs (only the
        end slabs hold a domain boundary)."""
        return float(sum(ctx.boundary_energy()[0] for ctx in self.ctx))

    # ------------------------------------------------------------ stepping
    def compute_dt(self):
        """dt and ch from the maxima over all slabs (solver.py:206-223)."""
        m = np.zeros(4)
        for ctx in self.ctx:
            mk = (C.c_double * 4)()
            _L.check(self.lib.amhd_cfl_maxima(ctx.h, mk), "amhd_cfl_maxima")
            m = np.maximum(m, np.array(mk))
        dt, ch = C.c_double(), C.c_double()
        mm = (C.c_double * 4)(*m)
        _L.check(self.lib.amhd_dt_from_maxima(mm, float(self.cfg.cfl), float(self.cfg.dx), float(self.cfg.dy),
                                              C.byref(dt), C.byref(ch)), "amhd_dt_from_maxima")
        return dt.value, ch.value

    def _exchange(self, stage: int) -> None:
        n = self.n
        # every slab's previous stage must be done reading its halo buffers
        # before the copies below (on the neighbours' streams) refill them
        self._sync()
        for k in range(n):
            lo, hi = self.halo[k]
            if lo is not None:      # rows i0-2, i0-1: the last two rows of the lower neighbour
                _L.check(self.lib.amhd_edge_rows(self.ctx[(k - 1) % n].h, stage, 1, _ptr(lo)), "amhd_edge_rows")
            if hi is not None:      # rows i1, i1+1: the first two rows of the upper neighbour
                _L.check(self.lib.amhd_edge_rows(self.ctx[(k + 1) % n].h, stage, 0, _ptr(hi)), "amhd_edge_rows")
        self._sync()

    def _sync(self) -> None:
        for ctx in self.ctx:
            _L.check(self.lib.amhd_synchronize(ctx.h), "amhd_synchronize")

    def step_ssprk3(self, dt: float, ch: float) -> None:
        """One SSP-RK3 step of every slab, halos refreshed before each stage."""
        cfg = self.cfg
        decay = float(np.exp(-cfg.glm_alpha * ch * dt / min(cfg.dx, cfg.dy))) if ch > 0 else 1.0
        for stage in (1, 2, 3):
            self._exchange(stage)
            for ctx in self.ctx:
                _L.check(self.lib.amhd_stage(ctx.h, stage, float(dt), float(ch), decay), "amhd_stage")
        self._sync()

    def advance(self, n_steps: int) -> int:
        """The loop of solver.py:475-482 (bounded by cfg.t_end, cfg.max_steps)."""
        done = 0
        while done < n_steps and self.t < self.cfg.t_end and self.step < self.cfg.max_steps:
            dt, ch = self.compute_dt()
            dt = min(dt, self.cfg.t_end - self.t)
            if dt <= 1e-16:
                break
            self.step_ssprk3(dt, ch)
            self.dt = dt
            self.t += dt
            self.step += 1
            done += 1
        return done


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None
