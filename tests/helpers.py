"""Shared test helpers: seeded valid MHD states and the textbook ideal-MHD
flux (written from the conservation law, not from the oracle's code)."""
import numpy as np

NV = 10


def random_prim(rng, n, vmax=2.0, bmax=1.5):
    """n random physical primitive vectors (NV, n)."""
    W = np.empty((NV, n))
    W[0] = rng.uniform(0.1, 4.0, n)
    W[1:4] = rng.uniform(-vmax, vmax, (3, n))
    W[4:7] = rng.uniform(-bmax, bmax, (3, n))
    W[7] = rng.uniform(0.05, 3.0, n)
    W[8] = rng.uniform(-0.2, 0.2, n)
    W[9] = rng.uniform(0.0, 1.0, n)
    return W


def physical_flux_x(W, gamma, bx, psi_flux=0.0, bx_flux=0.0):
    """Ideal-MHD x-flux of the conserved vector [rho, m, B, E, psi, rhoC]
    with normal field bx (the face value)."""
    rho, u, v, w = W[0], W[1], W[2], W[3]
    by, bz, p, c = W[5], W[6], W[7], W[9]
    b2 = bx * bx + by * by + bz * bz
    pt = p + 0.5 * b2
    e = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * b2
    vb = u * bx + v * by + w * bz
    return np.array([
        rho * u, rho * u * u + pt - bx * bx, rho * u * v - bx * by, rho * u * w - bx * bz,
        psi_flux + 0 * u, by * u - bx * v, bz * u - bx * w, (e + pt) * u - bx * vb, bx_flux + 0 * u, rho * c * u,
    ])


def physical_flux_y(W, gamma, by, psi_flux=0.0, by_flux=0.0):
    """Ideal-MHD y-flux (normal field by)."""
    rho, u, v, w = W[0], W[1], W[2], W[3]
    bx, bz, p, c = W[4], W[6], W[7], W[9]
    b2 = bx * bx + by * by + bz * bz
    pt = p + 0.5 * b2
    e = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w) + 0.5 * b2
    vb = u * bx + v * by + w * bz
    return np.array([
        rho * v, rho * v * u - by * bx, rho * v * v + pt - by * by, rho * v * w - by * bz,
        bx * v - by * u, psi_flux + 0 * v, bz * v - by * w, (e + pt) * v - by * vb, by_flux + 0 * v, rho * c * v,
    ])


def smooth_state(nx, ny, seed=0, amp=0.2, bx0=0.5, by0=0.3):
    """Ghosted smooth periodic conserved state (NV, nx+4, ny+4) and its
    primitive field: sinusoidal perturbations of a uniform MHD state."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx + 4) - 2 + 0.5) / nx
    y = (np.arange(ny + 4) - 2 + 0.5) / ny
    X, Y = np.meshgrid(x, y, indexing="ij")
    W = np.zeros((NV, nx + 4, ny + 4))
    ph = rng.uniform(0, 2 * np.pi, 12)
    s = lambda k, a, b: np.sin(2 * np.pi * (a * X + b * Y) + ph[k])
    W[0] = 1.0 + amp * s(0, 1, 1)
    W[1] = 0.3 + amp * s(1, 1, 2)
    W[2] = -0.2 + amp * s(2, 2, 1)
    W[3] = 0.1 * s(3, 1, 0)
    W[4] = bx0 + 0.1 * s(4, 0, 1)
    W[5] = by0 + 0.1 * s(5, 1, 0)
    W[6] = 0.05 * s(6, 1, 1)
    W[7] = 1.0 + amp * s(7, 2, 2)
    W[8] = 0.01 * s(8, 1, 1)
    W[9] = 0.5 + 0.4 * s(9, 1, 1)
    return W


def prim_to_cons_np(W, gamma):
    U = np.empty_like(W)
    rho = np.maximum(W[0], 1e-12)
    U[0] = rho
    U[1:4] = rho * W[1:4]
    U[4:7] = W[4:7]
    U[8] = W[8]
    U[9] = rho * W[9]
    U[7] = W[7] / (gamma - 1) + 0.5 * rho * (W[1] ** 2 + W[2] ** 2 + W[3] ** 2) + 0.5 * (
        W[4] ** 2 + W[5] ** 2 + W[6] ** 2)
    return U


def transpose_state(U):
    """x <-> y mirror of a conserved/primitive array: swap axes, (mx, my),
    (Bx, By)."""
    T = np.ascontiguousarray(np.transpose(U, (0, 2, 1)))
    T[[1, 2]] = T[[2, 1]]
    T[[4, 5]] = T[[5, 4]]
    return T


def rel_linf(a, b):
    """Per-variable relative L-inf error, normalised by max|b| of the variable."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    out = []
    for v in range(a.shape[0]):
        scale = max(float(np.max(np.abs(b[v]))), 1e-300)
        out.append(float(np.max(np.abs(a[v] - b[v]))) / scale)
    return np.array(out)
