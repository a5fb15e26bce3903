"""Initial conditions of the benchmark and verification cases, built on the
host in float64 (primitive arrays of shape (NVAR, nx+4, ny+4), ghosts
included, as the reference's drivers build them).

* ``brio_wu``         main.py:39-59 (Brio-Wu shock tube)
* ``alfven_wave``     main.py:131-160 (linear Alfven wave, one period)
* ``contact``         main.py:220-237 (moving contact discontinuity)
* ``rmi_single_mode_prim``  solver.py:109-139 (RMI, single-mode interface)
* ``rmi_uniform_bx``  single-mode RMI, Mach 10, Atwood 0.5, uniform Bx at
                      beta = 10 (parallel field: the hydro jump is exact)
* ``magnetic_sieve``  RMI with By(y) = B0 (1 + 0.5 sin(2 pi y / lambda_B)),
                      a per-row Rankine-Hugoniot post-shock state and a
                      seeded multimode (k = 4..32, k^-2) interface
* ``batched_instances``  seeded draws of uneven RMI instances

The reference has no uniform-Bx, By(y) or multimode initial condition; those
three are extensions with the parameters stated in their docstrings.
"""
from __future__ import annotations

import functools
import math

import numpy as np

from config import Config, NVAR, iRHO, iVX, iVY, iBX, iBY, iPR, iCLR
from physics import mhd_rankine_hugoniot


def cell_centres(cfg, ng: int = 2):
    """x_i = x_min + (i - ng + 0.5) dx over the ghosted index range (solver.py:79-80)."""
    x = cfg.x_min + (np.arange(cfg.nx + 2 * ng) - ng + 0.5) * cfg.dx
    y = cfg.y_min + (np.arange(cfg.ny + 2 * ng) - ng + 0.5) * cfg.dy
    return x, y


def _prim(cfg) -> np.ndarray:
    return np.zeros((NVAR, cfg.nx + 4, cfg.ny + 4))


# ---------------------------------------------------------- verification ICs
def brio_wu(nx: int = 800, cfl: float = 0.30, t_end: float = 0.1, precision: str = "fp64"):
    """Brio-Wu: gamma = 2, Bx = 0.75; left (1, 1, By=1), right (0.125, 0.1,
    By=-1); extrapolation in x, periodic in y (main.py:39-59).  The reference
    test uses CFL 0.30; the benchmark configuration uses 0.4."""
    cfg = Config(nx=nx, ny=4, x_min=0.0, x_max=1.0, y_min=0.0, y_max=0.05, t_end=t_end, cfl=cfl, gamma=2.0,
                 mach=1.0, B_transverse=0.0, interface_x=0.5, perturbation_amp=0.0, density_ratio=1.0,
                 diag_interval=10000, snapshot_times=[t_end], powell_source=False, use_char_bc=False,
                 bc_x_type="extrapolation", bc_y_type="periodic", precision=precision)
    x, _ = cell_centres(cfg)
    left = (x < 0.5)[:, None]
    W = _prim(cfg)
    W[iRHO] = np.where(left, 1.0, 0.125)
    W[iPR] = np.where(left, 1.0, 0.1)
    W[iBX] = 0.75
    W[iBY] = np.where(left, 1.0, -1.0)
    W[iCLR] = np.where(left, 1.0, 0.0)
    return cfg, W


def alfven_wave(nx: int, precision: str = "fp64"):
    """Linear Alfven wave (main.py:131-160): rho0 = 1, p0 = 0.1, Bx0 = 1,
    By = 1e-6 sin(2 pi x), vy = -By/sqrt(rho0), gamma = 5/3, CFL 0.25,
    periodic in x and y, one period t = Lx / vA."""
    rho0, p0, bx0, amp, gamma, lx = 1.0, 0.1, 1.0, 1e-6, 5.0 / 3.0, 1.0
    period = lx / (bx0 / np.sqrt(rho0))
    cfg = Config(nx=nx, ny=4, x_min=0.0, x_max=lx, y_min=0.0, y_max=0.05, t_end=period, cfl=0.25, gamma=gamma,
                 mach=1.0, B_transverse=0.0, interface_x=0.5, perturbation_amp=0.0, density_ratio=1.0,
                 diag_interval=100000, snapshot_times=[period], powell_source=False, use_char_bc=False,
                 bc_x_type="periodic", bc_y_type="periodic", precision=precision)
    x, _ = cell_centres(cfg)
    kx = 2.0 * np.pi / lx
    W = _prim(cfg)
    W[iRHO] = rho0
    W[iPR] = p0
    W[iBX] = bx0
    W[iBY] = amp * np.sin(kx * x)[:, None] * np.ones(cfg.ny + 4)[None, :]
    W[iVY] = -amp * np.sin(kx * x)[:, None] * np.ones(cfg.ny + 4)[None, :] / np.sqrt(rho0)
    return cfg, W


def contact(precision: str = "fp64"):
    """Moving contact (main.py:220-237): rho 1 -> 3 at x = 0.5, p = 1,
    vx = 1, Bx = 0.5, gamma = 5/3, 400 x 4 cells, t = 0.2."""
    cfg = Config(nx=400, ny=4, x_min=0.0, x_max=1.0, y_min=0.0, y_max=0.01, t_end=0.2, cfl=0.30,
                 gamma=5.0 / 3.0, mach=1.0, B_transverse=0.0, interface_x=0.5, perturbation_amp=0.0,
                 density_ratio=1.0, diag_interval=100000, snapshot_times=[0.2], powell_source=False,
                 use_char_bc=False, bc_x_type="extrapolation", bc_y_type="periodic", precision=precision)
    x, _ = cell_centres(cfg)
    W = _prim(cfg)
    W[iRHO] = np.where((x < 0.5)[:, None], 1.0, 3.0)
    W[iPR] = 1.0
    W[iVX] = 1.0
    W[iBX] = 0.5
    W[iCLR] = np.where((x < 0.5)[:, None], 0.0, 1.0)
    return cfg, W


# ------------------------------------------------------------------ RMI ICs
def _rmi_fill(cfg, W, X, x_if, post_state, by_pre, rho_h):
    """Shocked gas behind x_shock = interface_x - 0.3, tanh interface ahead
    of it (solver.py:121-137)."""
    rho2, p2, vx2, by2 = post_state
    x_shock = cfg.interface_x - 0.3
    delta = cfg.interface_width * cfg.dx
    phi = 0.5 * (1.0 + np.tanh((X - x_if) / max(delta, 1e-10)))
    post = X < x_shock
    rho_pre = 1.0 * (1 - phi) + rho_h * phi
    W[iRHO] = np.where(post, rho2, rho_pre)
    W[iVX] = np.where(post, vx2, 0.0)
    W[iBY] = np.where(post, by2, by_pre)
    W[iPR] = np.where(post, p2, 1.0)
    W[iCLR] = np.where(post, 0.0, phi)
    return W


def rmi_single_mode_prim(cfg):
    """Primitive RMI field of MHDSolver.initialize (solver.py:109-139):
    rho1 = p1 = 1, By_pre = B_transverse, heavy gas rho1 * density_ratio.
    Returns (W, Rankine-Hugoniot results, rho_heavy)."""
    g, M = cfg.gamma, cfg.mach
    rho2, p2, vx2, by2, vs = mhd_rankine_hugoniot(g, M, 1.0, 1.0, cfg.B_transverse)
    x, y = cell_centres(cfg)
    X, Y = np.meshgrid(x, y, indexing="ij")
    ky = 2 * np.pi * cfg.perturbation_mode / (cfg.y_max - cfg.y_min)
    x_if = cfg.interface_x + cfg.perturbation_amp * np.sin(ky * Y)
    rho_h = cfg.density_ratio
    W = _rmi_fill(cfg, _prim(cfg), X, x_if, (rho2, p2, vx2, by2), cfg.B_transverse, rho_h)
    return W, {"rho2": rho2, "p2": p2, "vx2": vx2, "By2": by2, "vs": vs}, rho_h


def rmi_uniform_bx(nx: int = 1024, ny: int = 512, precision: str = "fp64", mach: float = 10.0,
                   atwood: float = 0.5, beta: float = 10.0, mode: int = 4, amp_over_lambda: float = 0.05,
                   phase: float = 0.0, max_steps: int = 2000):
    """Single-mode RMI with a uniform Bx (the 1024 x 512 benchmark case):
    domain [0, 6] x [0, 2], gamma = 5/3, Mach 10 planar shock at x = 1.2
    into an interface at x = 1.5 with density ratio (1+A)/(1-A) = 3,
    sinusoid of mode 4 (lambda = 0.5) and amplitude 0.05 lambda,
    Bx = sqrt(2 p1 / beta) = sqrt(0.2), By = 0 (B_transverse = 0, so the
    Powell source is off under the reference's gate), zero-gradient x,
    periodic y, CFL 0.3.  A field parallel to the shock normal leaves the
    hydro jump conditions exact."""
    ly = 2.0
    lam = ly / mode
    cfg = Config(nx=nx, ny=ny, x_min=0.0, x_max=6.0, y_min=0.0, y_max=ly, t_end=1e9, cfl=0.30,
                 max_steps=max_steps, gamma=5.0 / 3.0, mach=mach, interface_x=1.5,
                 perturbation_amp=amp_over_lambda * lam, perturbation_mode=mode,
                 density_ratio=(1 + atwood) / (1 - atwood), interface_width=2.0, B_transverse=0.0,
                 powell_source=True, use_char_bc=False, bc_x_type="extrapolation", bc_y_type="periodic",
                 diag_interval=10 ** 9, snapshot_times=[], precision=precision)
    rho2, p2, vx2, _, _ = mhd_rankine_hugoniot(cfg.gamma, mach, 1.0, 1.0, 0.0)
    x, y = cell_centres(cfg)
    X, Y = np.meshgrid(x, y, indexing="ij")
    x_if = cfg.interface_x + cfg.perturbation_amp * np.sin(2 * np.pi * mode * Y / ly + phase)
    W = _rmi_fill(cfg, _prim(cfg), X, x_if, (rho2, p2, vx2, 0.0), 0.0, cfg.density_ratio)
    W[iBX] = math.sqrt(2.0 * 1.0 / beta)
    return cfg, W


def sieve_interface(y: np.ndarray, ly: float, x0: float, rms: float, seed: int, kmin: int = 4, kmax: int = 32):
    """Seeded multimode interface x_if(y) = x0 + sum_k a_k sin(2 pi k y/Ly + phi_k),
    k = kmin..kmax, a_k = k^-2 * u_k with u_k ~ U[0.5, 1.5], phi_k ~ U[0, 2 pi),
    rescaled so that the rms over a period equals `rms` (sum a_k^2 / 2 = rms^2)."""
    rng = np.random.default_rng(seed)
    ks = np.arange(kmin, kmax + 1)
    amps = ks.astype(np.float64) ** -2 * rng.uniform(0.5, 1.5, size=ks.size)
    phases = rng.uniform(0.0, 2 * np.pi, size=ks.size)
    amps *= rms / math.sqrt(0.5 * float(np.sum(amps ** 2)))
    x_if = np.full_like(y, x0, dtype=np.float64)
    for k, a, ph in zip(ks, amps, phases):
        x_if = x_if + a * np.sin(2 * np.pi * k * y / ly + ph)
    return x_if, ks, amps, phases


def magnetic_sieve(nx: int = 2048, ny: int = 1024, precision: str = "fp32", seed: int = 0, mach: float = 10.0,
                   atwood: float = 0.5, beta: float = 10.0, max_steps: int = 2000):
    """Magnetic-sieve RMI (the 2048 x 1024 headline case): as
    ``rmi_uniform_bx`` except that the field is transverse and modulated,
    By(y) = B0 (1 + 0.5 sin(2 pi y / lambda_B)), B0 = sqrt(2/beta),
    lambda_B = lambda / 4 = 0.125, Bx = 0; the post-shock state of each row
    is the Rankine-Hugoniot solution for that row's By; the interface is the
    seeded multimode ``sieve_interface`` with rms 0.02 lambda = 0.01.
    B_transverse = B0 > 0 turns the Powell source on.  By(y) is not
    divergence-free; GLM and Powell act on that divergence from t = 0."""
    ly = 2.0
    mode = 4
    lam = ly / mode
    lam_b = lam / 4
    b0 = math.sqrt(2.0 / beta)
    cfg = Config(nx=nx, ny=ny, x_min=0.0, x_max=6.0, y_min=0.0, y_max=ly, t_end=1e9, cfl=0.30,
                 max_steps=max_steps, gamma=5.0 / 3.0, mach=mach, interface_x=1.5,
                 perturbation_amp=0.02 * lam, perturbation_mode=mode,
                 density_ratio=(1 + atwood) / (1 - atwood), interface_width=2.0, B_transverse=b0,
                 powell_source=True, use_char_bc=False, bc_x_type="extrapolation", bc_y_type="periodic",
                 diag_interval=10 ** 9, snapshot_times=[], precision=precision)
    x, y = cell_centres(cfg)
    by_pre = b0 * (1.0 + 0.5 * np.sin(2 * np.pi * y / lam_b))
    post = _row_jumps(cfg.gamma, mach, tuple(float(b) for b in by_pre))
    x_if, *_ = sieve_interface(y, ly, cfg.interface_x, 0.02 * lam, seed)
    X = np.broadcast_to(x[:, None], (x.size, y.size))
    rows = (post[:, 0][None, :], post[:, 1][None, :], post[:, 2][None, :], post[:, 3][None, :])
    W = _rmi_fill(cfg, _prim(cfg), X, x_if[None, :], rows, by_pre[None, :], cfg.density_ratio)
    return cfg, W


@functools.lru_cache(maxsize=8)
def _row_jumps(gamma: float, mach: float, by_rows: tuple) -> np.ndarray:
    """Post-shock (rho2, p2, vx2, By2) for each row's pre-shock By (the
    profile does not depend on the seed, so the solves are cached)."""
    return np.array([mhd_rankine_hugoniot(gamma, mach, 1.0, 1.0, b)[:4] for b in by_rows])


def batched_instances(n: int = 64, seed: int = 0, precision: str = "fp32", max_steps: int = 500,
                      nx_range=(256, 1024)):
    """Seeded draws of uneven single-mode RMI instances: nx ~ U{256..1024}
    rounded to a multiple of 64, ny = nx / 2; Mach ~ U[2, 12];
    Atwood ~ U[0.2, 0.8]; beta ~ log-U[1, 100] (uniform Bx = sqrt(2/beta));
    mode ~ U{1..8}; phase ~ U[0, 2 pi); amplitude 0.05 lambda."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(round(rng.uniform(*nx_range) / 64.0)) * 64
        nx = min(max(nx, nx_range[0]), nx_range[1])
        mach = float(rng.uniform(2.0, 12.0))
        atw = float(rng.uniform(0.2, 0.8))
        beta = float(10 ** rng.uniform(0.0, 2.0))
        mode = int(rng.integers(1, 9))
        phase = float(rng.uniform(0.0, 2 * np.pi))
        out.append(rmi_uniform_bx(nx=nx, ny=nx // 2, precision=precision, mach=mach, atwood=atw, beta=beta,
                                  mode=mode, phase=phase, max_steps=max_steps))
    return out


def cons_from_prim_host(W: np.ndarray, gamma: float) -> np.ndarray:
    """physics.py:304-320 on the host (IC construction without a device)."""
    U = np.empty_like(W)
    rho = np.maximum(W[0], 1e-12)
    U[0] = rho
    U[1] = rho * W[1]
    U[2] = rho * W[2]
    U[3] = rho * W[3]
    U[4:7] = W[4:7]
    U[8] = W[8]
    U[9] = rho * W[9]
    ke = 0.5 * rho * (W[1] * W[1] + W[2] * W[2] + W[3] * W[3])
    me = 0.5 * (W[4] * W[4] + W[5] * W[5] + W[6] * W[6])
    U[7] = W[7] / (gamma - 1) + ke + me
    return U
