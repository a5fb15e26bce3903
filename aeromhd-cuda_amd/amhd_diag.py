"""Host diagnostics of a downloaded state (solver.py:335-444): mixing widths,
mixedness, interface mode amplitudes, enstrophy, stagnation pressure,
div B and total energy, in float64 NumPy.

``diagnostics`` restates the reference's computation on the whole field;
``from_reductions`` finishes the same diagnostics from the per-row and
per-column reductions the device makes (the solver's compute_diagnostics
path), sharing the O(nx + ny) steps with it.  ``mixing_width`` is the
integral mixing width alone (the measure the learned closure is evaluated
by).
"""
from __future__ import annotations

import numpy as np


def _c2p(U: np.ndarray, gamma: float) -> np.ndarray:
    W = np.empty_like(U)
    rho = np.maximum(U[0], 1e-12)
    ri = 1.0 / rho
    W[0] = rho
    W[1] = U[1] * ri
    W[2] = U[2] * ri
    W[3] = U[3] * ri
    W[4:7] = U[4:7]
    W[8] = U[8]
    W[9] = U[9] * ri
    ke = 0.5 * rho * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3])
    me = 0.5 * (U[4] * U[4] + U[5] * U[5] + U[6] * U[6])
    W[7] = np.maximum((gamma - 1) * (U[7] - ke - me), 1e-12)
    return W


def _trapezoid(y, x):
    return float(np.sum(0.5 * (y[1:] + y[:-1]) * np.diff(x))) if len(x) > 1 else 0.0


def mixing_width(U: np.ndarray, cfg, ng: int = 2) -> float:
    """Integral mixing width int C_bar (1 - C_bar) dx (solver.py:353-356)."""
    rho = np.maximum(U[0, ng:-ng, ng:-ng], 1e-12)
    c = np.clip(U[9, ng:-ng, ng:-ng] / rho, 0.0, 1.0)
    cbar = c.mean(axis=1)
    x = cfg.x_min + (np.arange(cfg.nx) + 0.5) * cfg.dx
    return _trapezoid(cbar * (1.0 - cbar), x)


def mixing_from_profile(cbar: np.ndarray, cvar: np.ndarray, x: np.ndarray):
    """Integral and threshold mixing widths and the mixedness from the row
    means and variances of C = clip(C, 0, 1) (solver.py:352-369)."""
    mw_integral = _trapezoid(cbar * (1.0 - cbar), x)
    mixed = np.where((cbar > 0.01) & (cbar < 0.99))[0]
    mw_thresh = float(x[mixed[-1]] - x[mixed[0]]) if len(mixed) > 1 else 0.0
    denom = np.maximum(cbar * (1.0 - cbar), 1e-14)
    if len(mixed) > 1:
        num = _trapezoid(cvar[mixed], x[mixed])
        den = _trapezoid(denom[mixed], x[mixed])
        mixedness = float(np.clip(1.0 - num / max(den, 1e-14), 0.0, 1.0))
    else:
        mixedness = 0.0
    return mw_integral, mw_thresh, mixedness


def interface_modes(pos: np.ndarray, y: np.ndarray, perturbation_mode: int):
    """Mode amplitudes of the interface position per column (NaN where none;
    solver.py:386-397): returns (mode_amps, pert_amp, valid, n_valid)."""
    ny_i = len(pos)
    valid = ~np.isnan(pos)
    nv = int(np.sum(valid))
    if nv > ny_i // 2:
        pi = np.interp(y, y[valid], pos[valid]) if nv < ny_i else pos
        modes = np.fft.rfft(pi - np.mean(pi))
        mode_amps = 2.0 * np.abs(modes) / ny_i
        pert_amp = float(mode_amps[perturbation_mode]) if perturbation_mode < len(mode_amps) else 0.0
    else:
        mode_amps = np.zeros(ny_i // 2 + 1)
        pert_amp = 0.0
    return mode_amps, pert_amp, valid, nv


def local_rows(pos: np.ndarray, valid: np.ndarray, nv: int, x: np.ndarray):
    """Rows [lo, hi) within 0.5 of the mean interface position, or None
    where the local enstrophy falls back to the global one
    (solver.py:401-410)."""
    if nv > len(pos) // 2:
        xc = float(np.mean(pos[valid]))
        lo = max(int(np.searchsorted(x, xc - 0.5)), 0)
        hi = min(int(np.searchsorted(x, xc + 0.5)), len(x))
        if hi - lo > 3:
            return lo, hi
    return None


def from_reductions(cfg, cbar, cvar, row_ens, row_stag, pos, mass, energy, div_max, div_l2) -> dict:
    """The diagnostics of solver.py:335-444 from per-row and per-column
    reductions made on the device (amhd_concentration_profile,
    amhd_flow_diagnostics, amhd_global_sums): the O(nx + ny) rest on the
    host, with the same steps as ``diagnostics``."""
    x = cfg.x_min + (np.arange(cfg.nx) + 0.5) * cfg.dx
    y = cfg.y_min + (np.arange(cfg.ny) + 0.5) * cfg.dy
    mw_integral, mw_thresh, mixedness = mixing_from_profile(cbar, cvar, x)
    mode_amps, pert_amp, valid, nv = interface_modes(pos, y, cfg.perturbation_mode)
    ens = float(np.sum(row_ens) / (cfg.nx * cfg.ny))
    rows = local_rows(pos, valid, nv, x)
    ens_local = float(np.sum(row_ens[rows[0]:rows[1]]) / ((rows[1] - rows[0]) * cfg.ny)) if rows else ens
    return {
        "mw_integral": mw_integral, "mw_thresh": mw_thresh, "mixedness": mixedness,
        "enstrophy": ens, "enstrophy_local": ens_local, "pert_amp": pert_amp, "mode_amps": mode_amps,
        "stag": float(np.max(row_stag)), "divB_max": float(div_max), "divB_L2": float(div_l2),
        "energy_total": float(energy), "mass_total": float(mass),
    }


def interface_positions(cc: np.ndarray, x: np.ndarray, dx: float) -> np.ndarray:
    """Interface x position per column from the clipped concentration
    (solver.py:372-384)."""
    nx_i, ny_i = cc.shape
    grad = np.abs(np.gradient(cc, dx, axis=0)) if nx_i > 1 else np.zeros_like(cc)
    pos = np.full(ny_i, np.nan)
    for j in range(ny_i):
        col = grad[:, j]
        if np.max(col) > 1e-6:
            k = int(np.argmax(col))
            hw = min(10, k, nx_i - k - 1)
            if hw > 0:
                wts = col[k - hw:k + hw + 1]
                ws = np.sum(wts)
                if ws > 1e-12:
                    pos[j] = np.sum(x[k - hw:k + hw + 1] * wts) / ws
    return pos


def diagnostics(U: np.ndarray, cfg, ng: int = 2) -> dict:
    """All diagnostics of solver.py:335-444 for the ghosted conserved U."""
    W = _c2p(U, cfg.gamma)
    sl = (slice(ng, -ng), slice(ng, -ng))
    rho, p, vx, vy = W[0][sl], W[7][sl], W[1][sl], W[2][sl]
    bx, by, bz, c = W[4][sl], W[5][sl], W[6][sl], W[9][sl]
    x = cfg.x_min + (np.arange(cfg.nx) + 0.5) * cfg.dx
    y = cfg.y_min + (np.arange(cfg.ny) + 0.5) * cfg.dy
    nx_i, ny_i = rho.shape

    cc = np.clip(c, 0.0, 1.0)
    cbar = cc.mean(axis=1)
    cps = np.mean((cc - cbar[:, None]) ** 2, axis=1)
    mw_integral, mw_thresh, mixedness = mixing_from_profile(cbar, cps, x)

    pos = interface_positions(cc, x, cfg.dx)
    mode_amps, pert_amp, valid, nv = interface_modes(pos, y, cfg.perturbation_mode)

    if nx_i > 1 and ny_i > 1:
        omega = np.gradient(vy, cfg.dx, axis=0) - np.gradient(vx, cfg.dy, axis=1)
    else:
        omega = np.zeros_like(vx)
    ens = float(np.mean(rho * omega ** 2))
    rows = local_rows(pos, valid, nv, x)
    ens_local = float(np.mean(rho[rows[0]:rows[1]] * omega[rows[0]:rows[1]] ** 2)) if rows else ens

    stag = float(np.max(p + 0.5 * rho * (vx ** 2 + vy ** 2) + 0.5 * (bx ** 2 + by ** 2 + bz ** 2)))
    div = np.zeros_like(bx)
    if nx_i > 2 and ny_i > 2:
        div[1:-1, 1:-1] = ((0.5 * (bx[1:-1, 1:-1] + bx[2:, 1:-1]) - 0.5 * (bx[:-2, 1:-1] + bx[1:-1, 1:-1])) / cfg.dx
                           + (0.5 * (by[1:-1, 1:-1] + by[1:-1, 2:]) - 0.5 * (by[1:-1, :-2] + by[1:-1, 1:-1])) / cfg.dy)
    energy = float(np.sum(U[7][sl]) * cfg.dx * cfg.dy)
    return {
        "mw_integral": mw_integral, "mw_thresh": mw_thresh, "mixedness": mixedness,
        "enstrophy": ens, "enstrophy_local": ens_local, "pert_amp": pert_amp, "mode_amps": mode_amps,
        "stag": stag, "divB_max": float(np.max(np.abs(div))), "divB_L2": float(np.sqrt(np.mean(div ** 2))),
        "energy_total": energy, "mass_total": float(np.sum(U[0][sl]) * cfg.dx * cfg.dy),
    }
