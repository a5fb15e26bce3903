"""Host diagnostics of a downloaded state (solver.py:335-444): mixing widths,
mixedness, interface mode amplitudes, enstrophy, stagnation pressure,
div B and total energy, in float64 NumPy.

``mixing_width`` is the integral mixing width alone (the measure the
learned closure is evaluated by).
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
    mw_integral = _trapezoid(cbar * (1.0 - cbar), x)
    mixed = np.where((cbar > 0.01) & (cbar < 0.99))[0]
    mw_thresh = float(x[mixed[-1]] - x[mixed[0]]) if len(mixed) > 1 else 0.0
    cps = np.mean((cc - cbar[:, None]) ** 2, axis=1)
    denom = np.maximum(cbar * (1.0 - cbar), 1e-14)
    if len(mixed) > 1:
        num = _trapezoid(cps[mixed], x[mixed])
        den = _trapezoid(denom[mixed], x[mixed])
        mixedness = float(np.clip(1.0 - num / max(den, 1e-14), 0.0, 1.0))
    else:
        mixedness = 0.0

    grad = np.abs(np.gradient(cc, cfg.dx, axis=0)) if nx_i > 1 else np.zeros_like(cc)
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
    valid = ~np.isnan(pos)
    nv = int(np.sum(valid))
    if nv > ny_i // 2:
        pi = np.interp(y, y[valid], pos[valid]) if nv < ny_i else pos
        modes = np.fft.rfft(pi - np.mean(pi))
        mode_amps = 2.0 * np.abs(modes) / ny_i
        pert_amp = float(mode_amps[cfg.perturbation_mode]) if cfg.perturbation_mode < len(mode_amps) else 0.0
    else:
        mode_amps = np.zeros(ny_i // 2 + 1)
        pert_amp = 0.0

    if nx_i > 1 and ny_i > 1:
        omega = np.gradient(vy, cfg.dx, axis=0) - np.gradient(vx, cfg.dy, axis=1)
    else:
        omega = np.zeros_like(vx)
    ens = float(np.mean(rho * omega ** 2))
    ens_local = ens
    if nv > ny_i // 2:
        xc = float(np.mean(pos[valid]))
        lo = max(int(np.searchsorted(x, xc - 0.5)), 0)
        hi = min(int(np.searchsorted(x, xc + 0.5)), nx_i)
        if hi - lo > 3:
            ens_local = float(np.mean(rho[lo:hi] * omega[lo:hi] ** 2))

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
