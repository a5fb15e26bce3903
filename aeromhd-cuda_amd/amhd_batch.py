"""Batched form: many independent MHD instances of uneven grid size and
physical parameters advanced together (no counterpart in the reference).

Every RK stage is ONE persistent launch over all instances (tasks =
(instance, y strip, x segment)); each instance keeps its own device-side
dt, glm_ch, psi decay, t, step count and state-buffer index, so a single
captured step is replayed from a hipGraph with no host round trip.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import amhd_lib as _L
from config import NVAR
from solver import _PRECISIONS, solver_params


class BatchSolver:
    """Instances from (Config, conserved U) pairs; all share one precision."""

    def __init__(self, cases, precision: str | None = None):
        _L.require_gpu()
        self.lib = _L.load()
        self.cfgs = [c for c, _ in cases]
        prec = precision or getattr(self.cfgs[0], "precision", "fp32")
        if prec not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        self.precision = prec
        n = len(cases)
        self.n = n
        nx = (C.c_int * n)(*[c.nx for c in self.cfgs])
        ny = (C.c_int * n)(*[c.ny for c in self.cfgs])
        params = (_L.AmhdParams * n)(*[solver_params(c) for c in self.cfgs])
        h = C.c_void_p()
        _L.check(self.lib.amhd_batch_create(C.byref(h), n, nx, ny, _PRECISIONS[prec], params), "amhd_batch_create")
        self.h = h
        self.t = np.zeros(n)
        self.steps = np.zeros(n, dtype=np.int32)
        for k, (_, U) in enumerate(cases):
            self.upload(k, U)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.amhd_batch_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def set_rows(self, rows: int) -> None:
        _L.check(self.lib.amhd_batch_set_rows(self.h, int(rows)), "amhd_batch_set_rows")

    def info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _L.check(self.lib.amhd_batch_info(self.h, C.byref(a), C.byref(b), C.byref(c)), "amhd_batch_info")
        return {"tasks": a.value, "resident_waves": b.value, "rows": c.value}

    def upload(self, k: int, U) -> None:
        Uh = np.ascontiguousarray(U, dtype=np.float64)
        cfg = self.cfgs[k]
        if Uh.shape != (NVAR, cfg.nx + 4, cfg.ny + 4):
            raise ValueError(f"instance {k}: U must have shape {(NVAR, cfg.nx + 4, cfg.ny + 4)}")
        _L.check(self.lib.amhd_batch_upload(self.h, int(k), Uh.ctypes.data_as(C.c_void_p), 0), "amhd_batch_upload")

    def download(self, k: int) -> np.ndarray:
        cfg = self.cfgs[k]
        U = np.empty((NVAR, cfg.nx + 4, cfg.ny + 4))
        _L.check(self.lib.amhd_batch_download(self.h, int(k), U.ctypes.data_as(C.c_void_p), 0), "amhd_batch_download")
        return U

    def run(self, n_steps: int, use_graph: bool = True) -> np.ndarray:
        """Up to n_steps steps of every instance (each bounded by its own
        t_end and max_steps); returns the steps taken per instance."""
        n = self.n
        cfl = (C.c_double * n)(*[c.cfl for c in self.cfgs])
        t_end = (C.c_double * n)(*[c.t_end for c in self.cfgs])
        mx = (C.c_int * n)(*[int(min(c.max_steps, 2 ** 31 - 1)) for c in self.cfgs])
        t = (C.c_double * n)(*self.t.tolist())
        st = (C.c_int * n)(*self.steps.tolist())
        s0 = self.steps.copy()
        _L.check(self.lib.amhd_batch_run(self.h, int(n_steps), cfl, t_end, mx, 1 if use_graph else 0, t, st),
                 "amhd_batch_run")
        self.t = np.array(t[:])
        self.steps = np.array(st[:], dtype=np.int32)
        return self.steps - s0

    @property
    def cell_count(self) -> int:
        return int(sum(c.nx * c.ny for c in self.cfgs))
