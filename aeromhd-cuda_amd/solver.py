"""Drop-in for the reference's MHDSolver (solver.py:69-520).

The solver state lives in a device context of libaeromhd_hip.so (interior
cells only, structure of arrays, FP64 or FP32).  The attributes and methods
the reference's callers use keep their names and meaning:

* ``U`` -- the ghosted conserved array (NVAR, nx+4, ny+4) as a float64
  device tensor.  Reading it materialises the ghosts exactly as
  ``apply_bc`` leaves them; assigning it uploads the interior.
* ``t``, ``step``, ``dt``, ``ng``, ``nx_tot``, ``ny_tot``, ``x``, ``y``,
  ``X``, ``Y``.
* ``initialize()``, ``apply_bc(U)``, ``enforce_scalar_bounds(U)``,
  ``compute_dt(W)``, ``compute_rhs(U)``, ``step_ssprk3()``, ``run()``,
  ``compute_diagnostics()``, ``save_snapshot()``.

Added: ``advance(n_steps)`` runs the device-resident time loop (dt, glm_ch
and the psi decay computed on the GPU, steps replayed from a hipGraph).
"""
from __future__ import annotations

import ctypes as C
import sys
import time

import numpy as np

import amhd_lib as _L
from config import NVAR, PSI, iRHO, iVX, iBY, iPR, iCLR  # noqa: F401
from physics import xp, cons_to_prim, prim_to_cons, mhd_rankine_hugoniot, _to_device, _ptr, _stream

_PRECISIONS = {"fp64": _L.AMHD_F64, "float64": _L.AMHD_F64, "fp32": _L.AMHD_F32, "float32": _L.AMHD_F32}


def to_numpy(a):
    """Device tensor -> NumPy array; NumPy input passes through (solver.py:59-63)."""
    try:
        import torch

        if isinstance(a, torch.Tensor):
            return a.detach().cpu().numpy()
    except ImportError:  # pragma: no cover
        pass
    return np.asarray(a)


def solver_params(cfg) -> _L.AmhdParams:
    """The amhd_params of a Config: BC kinds and the Powell gate of solver.py:304."""
    p = _L.AmhdParams()
    p.gamma = float(cfg.gamma)
    p.dx = float(cfg.dx)
    p.dy = float(cfg.dy)
    p.x_min = float(cfg.x_min)
    p.y_min = float(cfg.y_min)
    p.glm_alpha = float(cfg.glm_alpha)
    p.bc_x_periodic = 1 if cfg.get_bc_x() == "periodic" else 0
    p.bc_y_periodic = 1 if cfg.bc_y_type == "periodic" else 0
    p.powell = 1 if (cfg.powell_source and cfg.B_transverse > 0) else 0
    return p


class Context:
    """Owns one amhd context (device state buffers, stream, graphs)."""

    def __init__(self, nx: int, ny: int, precision: str, params: _L.AmhdParams):
        _L.require_gpu()
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {precision!r}")
        self.lib = _L.load()
        self.nx, self.ny = int(nx), int(ny)
        self.dtype = _PRECISIONS[precision]
        self.params = params
        h = C.c_void_p()
        _L.check(self.lib.amhd_create(C.byref(h), self.nx, self.ny, self.dtype, C.byref(params)), "amhd_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.amhd_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def set_rows(self, rows: int):
        _L.check(self.lib.amhd_set_rows(self.h, int(rows)), "amhd_set_rows")

    @property
    def rows(self) -> int:
        return int(self.lib.amhd_get_rows(self.h))

    def set_packing(self, max_groups: int):
        """Let up to max_groups (2, 4 or 8; 0 = off) segments of a narrow
        last y strip share one wavefront (default 0); the results do not
        depend on it."""
        _L.check(self.lib.amhd_set_packing(self.h, int(max_groups)), "amhd_set_packing")

    def set_balance(self, on: bool):
        """Size the segments by cost each step (default on); results do
        not depend on it."""
        _L.check(self.lib.amhd_set_balance(self.h, 1 if on else 0), "amhd_set_balance")

    def upload(self, U) -> None:
        torch = _L.require_gpu()
        if isinstance(U, torch.Tensor):
            Ud = U.detach().to(device="cuda", dtype=torch.float64).contiguous()
            torch.cuda.current_stream().synchronize()
            _L.check(self.lib.amhd_upload_state(self.h, _ptr(Ud), 1), "amhd_upload_state")
        else:
            Uh = np.ascontiguousarray(U, dtype=np.float64)
            _L.check(self.lib.amhd_upload_state(self.h, Uh.ctypes.data_as(C.c_void_p), 0), "amhd_upload_state")

    def download(self):
        torch = _L.require_gpu()
        U = torch.empty((NVAR, self.nx + 4, self.ny + 4), dtype=torch.float64, device="cuda")
        torch.cuda.current_stream().synchronize()
        _L.check(self.lib.amhd_download_state(self.h, _ptr(U), 1), "amhd_download_state")
        return U

    def download_numpy(self) -> np.ndarray:
        U = np.empty((NVAR, self.nx + 4, self.ny + 4), dtype=np.float64)
        _L.check(self.lib.amhd_download_state(self.h, U.ctypes.data_as(C.c_void_p), 0), "amhd_download_state")
        return U

    def compute_dt(self, cfl: float):
        dt, ch = C.c_double(), C.c_double()
        _L.check(self.lib.amhd_compute_dt(self.h, float(cfl), C.byref(dt), C.byref(ch)), "amhd_compute_dt")
        return dt.value, ch.value

    def step(self, dt: float, ch: float, decay: float) -> None:
        _L.check(self.lib.amhd_step_ssprk3(self.h, float(dt), float(ch), float(decay)), "amhd_step_ssprk3")

    def run(self, n_steps: int, cfl: float, t_end: float, max_steps: int, t: float, step: int,
            use_graph: bool = True):
        tt, ss, ld = C.c_double(t), C.c_int(step), C.c_double(0.0)
        _L.check(self.lib.amhd_run(self.h, int(n_steps), float(cfl), float(t_end), int(max_steps),
                                   1 if use_graph else 0, C.byref(tt), C.byref(ss), C.byref(ld)), "amhd_run")
        return tt.value, ss.value, ld.value

    def run_profiled(self, n_steps: int, cfl: float, t_end: float, max_steps: int, t: float, step: int):
        """Device loop with HIP events filled by every launch's dispatch; returns
        (t, step, last_dt, [stage-1/2/3 ms], ctl ms) summed over the steps."""
        tt, ss, ld = C.c_double(t), C.c_int(step), C.c_double(0.0)
        ms = (C.c_double * 3)()
        cms = C.c_double(0.0)
        _L.check(self.lib.amhd_run_profiled(self.h, int(n_steps), float(cfl), float(t_end), int(max_steps),
                                            C.byref(tt), C.byref(ss), C.byref(ld), ms, C.byref(cms)),
                 "amhd_run_profiled")
        return tt.value, ss.value, ld.value, list(ms), cms.value

    def boundary_energy(self):
        cum, last = C.c_double(), C.c_double()
        _L.check(self.lib.amhd_boundary_energy(self.h, C.byref(cum), C.byref(last)), "amhd_boundary_energy")
        return cum.value, last.value

    def reset_boundary_energy(self, value: float = 0.0) -> None:
        _L.check(self.lib.amhd_reset_boundary_energy(self.h, float(value)), "amhd_reset_boundary_energy")

    def global_sums(self):
        out = (C.c_double * 4)()
        _L.check(self.lib.amhd_global_sums(self.h, out), "amhd_global_sums")
        return tuple(out)

    def concentration_profile(self):
        mean = np.empty(self.nx, dtype=np.float64)
        var = np.empty(self.nx, dtype=np.float64)
        _L.check(self.lib.amhd_concentration_profile(self.h, mean.ctypes.data_as(C.POINTER(C.c_double)),
                                                     var.ctypes.data_as(C.POINTER(C.c_double))),
                 "amhd_concentration_profile")
        return mean, var

    def flow_diagnostics(self):
        """(row_ens, row_stag, col_pos): per-row sums of rho omega^2, per-row
        maxima of the stagnation pressure and per-column interface positions
        (NaN where none), reduced on the device."""
        ens = np.empty(self.nx, dtype=np.float64)
        stag = np.empty(self.nx, dtype=np.float64)
        pos = np.empty(self.ny, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        _L.check(self.lib.amhd_flow_diagnostics(self.h, ens.ctypes.data_as(dp), stag.ctypes.data_as(dp),
                                                pos.ctypes.data_as(dp)), "amhd_flow_diagnostics")
        return ens, stag, pos


class MHDSolver:
    """2-D ideal MHD: MUSCL/van Leer + HLLD, SSP-RK3, GLM + Powell
    (solver.py:69-107 state and grid)."""

    def __init__(self, config):
        self.cfg = config
        self.ng = 2
        self.nx_tot = config.nx + 2 * self.ng
        self.ny_tot = config.ny + 2 * self.ng
        self.x = xp.asarray(config.x_min + (np.arange(self.nx_tot) - self.ng + 0.5) * config.dx)
        self.y = xp.asarray(config.y_min + (np.arange(self.ny_tot) - self.ng + 0.5) * config.dy)
        self.X, self.Y = xp.meshgrid(self.x, self.y, indexing="ij")
        self.ctx = Context(config.nx, config.ny, getattr(config, "precision", "fp64"), solver_params(config))
        self._U_view = None      # last U handed out or assigned (authoritative when set)
        self.U = xp.zeros((NVAR, self.nx_tot, self.ny_tot))
        self.t = 0.0
        self.step = 0
        self.dt = 0.0
        self._glm_ch_frozen = 0.0

        self.diag_times = []
        self.diag_mixing_width_integral = []
        self.diag_mixing_width_thresh = []
        self.diag_mixedness = []
        self.diag_enstrophy = []
        self.diag_enstrophy_local = []
        self.diag_perturbation_amp = []
        self.diag_mode_amps = []
        self.diag_stag_pressure = []
        self.diag_divB_max = []
        self.diag_divB_L2 = []
        self.diag_energy_total = []
        self.diag_boundary_flux_cumulative = []
        self.snapshots = {}
        self._e0 = None          # energy of the first device_diagnostics call (drift reference)
        self._rho_light = 1.0
        self._rho_heavy = 1.0
        self._rh_results = None

    # ------------------------------------------------------------ state
    @property
    def U(self):
        """Ghosted conserved state as a float64 device tensor.  The tensor
        handed out stays authoritative until the next step (in-place edits
        are honoured), as the reference's array is."""
        if self._U_view is None:
            self._U_view = self.ctx.download()
        return self._U_view

    @U.setter
    def U(self, value):
        Ud, _ = _to_device(value)
        if tuple(Ud.shape) != (NVAR, self.nx_tot, self.ny_tot):
            raise ValueError(f"U must have shape {(NVAR, self.nx_tot, self.ny_tot)}, got {tuple(Ud.shape)}")
        self._U_view = Ud

    def _sync_state(self):
        """Push a handed-out / assigned U into the context before stepping."""
        if self._U_view is not None:
            self.ctx.upload(self._U_view)
            self._U_view = None

    @property
    def _cumulative_boundary_energy(self) -> float:
        return self.ctx.boundary_energy()[0]

    @_cumulative_boundary_energy.setter
    def _cumulative_boundary_energy(self, value: float):
        self.ctx.reset_boundary_energy(value)

    # ------------------------------------------------------------ physics
    def initialize(self):
        """RMI initial condition (solver.py:109-156): Rankine-Hugoniot
        post-shock gas behind x = interface_x - 0.3 and a tanh density
        interface perturbed by a sine of mode perturbation_mode."""
        from amhd_ics import rmi_single_mode_prim

        W, rh, rho_h = rmi_single_mode_prim(self.cfg)
        self._rh_results = rh
        self.U = prim_to_cons(W, self.cfg.gamma)
        self.t = 0.0
        self.step = 0
        self._cumulative_boundary_energy = 0.0
        self._e0 = None
        self._rho_light = 1.0
        self._rho_heavy = rho_h

    def apply_bc(self, U):
        """Boundary conditions in place (solver.py:158-197); returns U."""
        cfg = self.cfg
        p = solver_params(cfg)
        Ud, host = _to_device(U)
        _L.call("amhd_apply_bc", _ptr(Ud), cfg.nx, cfg.ny, p.bc_x_periodic, p.bc_y_periodic, _stream())
        return _write_back(U, Ud, host)

    def enforce_scalar_bounds(self, U):
        """rhoC = rho * clip(rhoC/rho, 0, 1) in place (solver.py:199-204)."""
        Ud, host = _to_device(U)
        _L.call("amhd_enforce_scalar_bounds", _ptr(Ud), Ud.numel() // NVAR, _stream())
        return _write_back(U, Ud, host)

    def compute_dt(self, W=None):
        """CFL time step over the primitive array W (solver.py:206-223); sets
        cfg.glm_ch.  W=None reads the solver's own state (fused path)."""
        cfg = self.cfg
        if W is None:
            self._sync_state()
            dt, ch = self.ctx.compute_dt(cfg.cfl)
        else:
            Wd, _ = _to_device(W)
            dtc, chc = C.c_double(), C.c_double()
            _L.call("amhd_compute_dt_w", _ptr(Wd), Wd.numel() // NVAR, float(cfg.gamma), float(cfg.cfl),
                    float(cfg.dx), float(cfg.dy), C.byref(dtc), C.byref(chc), _stream())
            dt, ch = dtc.value, chc.value
        cfg.glm_ch = ch
        return dt

    def compute_rhs(self, U):
        """-div(F) + Powell source (solver.py:268-308); applies BC to U in place."""
        torch = _L.require_gpu()
        Ud, host = _to_device(U)
        R = torch.empty_like(Ud)
        torch.cuda.current_stream().synchronize()
        _L.check(self.ctx.lib.amhd_compute_rhs(self.ctx.h, _ptr(Ud), _ptr(R), float(self._glm_ch_frozen)),
                 "amhd_compute_rhs")
        _write_back(U, Ud, host)
        return R.cpu().numpy() if host else R

    def step_ssprk3(self):
        """One SSP-RK3 step with self.dt and the frozen cfg.glm_ch
        (solver.py:310-333).  t and step are advanced by the caller."""
        dt = self.dt
        self._glm_ch_frozen = self.cfg.glm_ch
        ch = self._glm_ch_frozen
        decay = float(np.exp(-self.cfg.glm_alpha * ch * dt / min(self.cfg.dx, self.cfg.dy))) if ch > 0 else 1.0
        self._sync_state()
        self.ctx.step(dt, ch, decay)

    def advance(self, n_steps: int, use_graph: bool = True) -> int:
        """Device-resident loop of solver.py:475-482 for up to n_steps steps
        (bounded by cfg.t_end and cfg.max_steps); returns the steps taken."""
        self._sync_state()
        s0 = self.step
        self.t, self.step, last = self.ctx.run(n_steps, self.cfg.cfl, self.cfg.t_end, self.cfg.max_steps,
                                               self.t, self.step, use_graph)
        if self.step > s0:
            self.dt = last
        return self.step - s0

    # ------------------------------------------------------------ analysis
    def global_sums(self):
        """FP64 device sums of the current state: (mass, energy, max|divB|,
        L2 divB) with mass = sum(rho) dx dy, energy as solver.py:430 and
        divB as solver.py:419-428."""
        self._sync_state()
        return self.ctx.global_sums()

    def device_diagnostics(self) -> dict:
        """The diagnostics the time loop needs, reduced on the device in
        FP64 without downloading the state: total mass and energy
        (solver.py:430), max / L2 divB (solver.py:419-428), the integral and
        threshold mixing widths and the mixedness (solver.py:352-371, from
        the device concentration profile; the O(nx) trapezoid sums on the
        host), and the energy drift raw and corrected by the boundary flux
        (solver.py:507-511) against the first call's energy."""
        mass, energy, div_max, div_l2 = self.global_sums()
        cbar, cvar = self.ctx.concentration_profile()
        x = self.cfg.x_min + (np.arange(self.cfg.nx) + 0.5) * self.cfg.dx
        trap = lambda y, xx: float(np.sum(0.5 * (y[1:] + y[:-1]) * np.diff(xx))) if len(xx) > 1 else 0.0
        mw_integral = trap(cbar * (1.0 - cbar), x)
        mixed = np.where((cbar > 0.01) & (cbar < 0.99))[0]
        mw_thresh = float(x[mixed[-1]] - x[mixed[0]]) if len(mixed) > 1 else 0.0
        if len(mixed) > 1:
            num = trap(cvar[mixed], x[mixed])
            den = trap(np.maximum(cbar * (1.0 - cbar), 1e-14)[mixed], x[mixed])
            mixedness = float(np.clip(1.0 - num / max(den, 1e-14), 0.0, 1.0))
        else:
            mixedness = 0.0
        if self._e0 is None:
            self._e0 = energy
        bf = self._cumulative_boundary_energy
        e0 = self._e0
        return {"mass_total": mass, "energy_total": energy, "divB_max": div_max, "divB_L2": div_l2,
                "mw_integral": mw_integral, "mw_thresh": mw_thresh, "mixedness": mixedness,
                "energy_drift_raw_pct": abs(energy - e0) / max(abs(e0), 1e-14) * 100,
                "energy_drift_corrected_pct": abs(energy - bf - e0) / max(abs(e0), 1e-14) * 100,
                "boundary_flux": bf}

    def diagnostics_dict(self) -> dict:
        """Every diagnostic of solver.py:335-444 without downloading the
        state: row and column reductions in FP64 on the device
        (amhd_global_sums, amhd_concentration_profile,
        amhd_flow_diagnostics), then the O(nx + ny) rest (trapezoids, the
        interface's interpolation and FFT) on the host, with the steps of
        amhd_diag.diagnostics."""
        from amhd_diag import from_reductions

        self._sync_state()
        mass, energy, div_max, div_l2 = self.global_sums()
        cbar, cvar = self.ctx.concentration_profile()
        row_ens, row_stag, pos = self.ctx.flow_diagnostics()
        return from_reductions(self.cfg, cbar, cvar, row_ens, row_stag, pos, mass, energy, div_max, div_l2)

    def compute_diagnostics(self):
        """Diagnostics of solver.py:335-444, appended to the diag_* lists as
        the reference's loop does; the state stays on the device."""
        d = self.diagnostics_dict()
        self.diag_times.append(self.t)
        self.diag_mixing_width_integral.append(d["mw_integral"])
        self.diag_mixing_width_thresh.append(d["mw_thresh"])
        self.diag_mixedness.append(d["mixedness"])
        self.diag_enstrophy.append(d["enstrophy"])
        self.diag_enstrophy_local.append(d["enstrophy_local"])
        self.diag_perturbation_amp.append(d["pert_amp"])
        self.diag_mode_amps.append(d["mode_amps"].copy())
        self.diag_stag_pressure.append(d["stag"])
        self.diag_divB_max.append(d["divB_max"])
        self.diag_divB_L2.append(d["divB_L2"])
        self.diag_energy_total.append(d["energy_total"])
        self.diag_boundary_flux_cumulative.append(self._cumulative_boundary_energy)

    def save_snapshot(self, label=None):
        """Primitive interior copied to a host dict (solver.py:446-463)."""
        ng = self.ng
        W = to_numpy(cons_to_prim(self.U, self.cfg.gamma))
        key = label or f"t={self.t:.4f}"
        names = {"rho": 0, "p": 7, "vx": 1, "vy": 2, "Bx": 4, "By": 5, "Bz": 6, "C": 9}
        snap = {k: W[v, ng:-ng, ng:-ng].copy() for k, v in names.items()}
        snap.update(t=self.t, x=to_numpy(self.x[ng:-ng]), y=to_numpy(self.y[ng:-ng]))
        self.snapshots[key] = snap

    def run(self):
        """Time loop of solver.py:465-520: dt, step, diagnostics every
        diag_interval steps, snapshots near cfg.snapshot_times."""
        cfg = self.cfg
        print(f"\n--- Running (HIP): B_y = {cfg.B_transverse} ---")
        sys.stdout.flush()
        t0 = time.time()
        self.compute_diagnostics()
        self.save_snapshot(label="initial")
        first = True
        while self.t < cfg.t_end and self.step < cfg.max_steps:
            if first:
                dt_cfl = self.compute_dt(cons_to_prim(self.U, cfg.gamma))
                first = False
            else:
                dt_cfl = self.compute_dt()
            self.dt = min(dt_cfl, cfg.t_end - self.t)
            if self.dt <= 1e-16:
                break
            self.step_ssprk3()
            self.t += self.dt
            self.step += 1
            if self.step % cfg.diag_interval == 0:
                self.compute_diagnostics()
            for st in cfg.snapshot_times:
                lbl = f"t~{st:.2f}"
                if abs(self.t - st) < 1.5 * self.dt and lbl not in self.snapshots:
                    self.save_snapshot(label=lbl)
            if self.step % 200 == 0:
                print(f"  Step {self.step:5d} t={self.t:.4f} dt={self.dt:.2e} "
                      f"MW={self.diag_mixing_width_integral[-1]:.4f} "
                      f"divB={self.diag_divB_max[-1]:.1f} E={self.diag_energy_total[-1]:.2f}")
                sys.stdout.flush()
        self.compute_diagnostics()
        self.save_snapshot(label="final")
        elapsed = time.time() - t0
        if len(self.diag_energy_total) >= 2:
            e0, ef = self.diag_energy_total[0], self.diag_energy_total[-1]
            bf = self._cumulative_boundary_energy
            print(f"  Energy drift: raw={abs(ef - e0) / max(abs(e0), 1e-14) * 100:.2f}%, "
                  f"corrected={abs(ef - bf - e0) / max(abs(e0), 1e-14) * 100:.2f}%")
        print(f"  Done: {self.step} steps, {elapsed:.1f}s")
        sys.stdout.flush()


def _write_back(orig, Ud, host: bool):
    """Mirror an in-place device update into the caller's array."""
    if host:
        orig[...] = Ud.cpu().numpy()
        return orig
    if Ud.data_ptr() != orig.data_ptr():
        orig.copy_(Ud)
        return orig
    return Ud
