"""Writes tests/golden/hlld_ref_vectors.npz: inputs and outputs of the
reference's own HLLD kernel (physics.py:43-275), run on the GPU from
oracle/_ref/libhlld_ref.so (built by oracle/build_ref.py from the kernel
text, hipcc, -ffp-contract=off).

Cases (seeded): random valid faces at several ch, ch = 0 (supersonic and
star branches), degenerate faces that exercise every clamp of
physics.py:114-134,158-165 (den, SM, dS-SM, eps_rel, SAL/SAR), floors
(rho, p below 1e-12) and faces harvested from a Brio-Wu oracle state.

Run on a GPU machine:  python tests/golden/make_hlld_vectors.py [out.npz]
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "aeromhd-cuda_amd"), os.path.join(REPO, "oracle"), os.path.dirname(HERE)]


def ref_lib():
    lib = C.CDLL(os.path.join(REPO, "oracle", "_ref", "libhlld_ref.so"))
    vp, d, i = C.c_void_p, C.c_double, C.c_int
    lib.ref_hlld_flux_x.restype = i
    lib.ref_hlld_flux_x.argtypes = [vp, vp, vp, vp, i, d, d, d, d]
    return lib


def run_ref(lib, WL, WR, gamma, ch):
    WL = np.ascontiguousarray(WL, dtype=np.float64)
    WR = np.ascontiguousarray(WR, dtype=np.float64)
    n = WL.shape[1]
    F = np.empty_like(WL)
    s = np.empty(n)
    rc = lib.ref_hlld_flux_x(WL.ctypes.data, WR.ctypes.data, F.ctypes.data, s.ctypes.data, n, gamma, ch, 1e-12,
                             1e-12)
    if rc != 0:
        raise RuntimeError(f"reference kernel failed: {rc}")
    return F, s


def cases():
    from helpers import random_prim

    rng = np.random.default_rng(2026)
    out = {}
    WL, WR = random_prim(rng, 4096), random_prim(rng, 4096)
    out["random_ch5"] = (WL, WR, 5.0 / 3.0, 5.0)
    WL, WR = random_prim(rng, 2048, vmax=6.0), random_prim(rng, 2048, vmax=6.0)
    out["random_ch0"] = (WL, WR, 1.4, 0.0)
    out["random_ch0p5_g2"] = (random_prim(rng, 2048), random_prim(rng, 2048), 2.0, 0.5)
    # degenerate: tiny field, equal states, near-vacuum, floors
    WL, WR = random_prim(rng, 1024), random_prim(rng, 1024)
    WL[4:7] *= 1e-9
    WR[4:7] *= 1e-9
    WR[:, :256] = WL[:, :256]
    WL[0, 256:384] = 1e-14
    WR[7, 384:512] = -1.0
    WL[5:7, 512:640] = 0.0
    WR[5:7, 512:640] = 0.0
    out["degenerate_ch3"] = (WL, WR, 5.0 / 3.0, 3.0)
    WL2, WR2 = WL.copy(), WR.copy()
    out["degenerate_ch0"] = (WL2, WR2, 5.0 / 3.0, 0.0)
    # Brio-Wu oracle state after 200 steps: harvested faces
    import oracle
    from amhd_ics import brio_wu, cons_from_prim_host

    cfg, W = brio_wu(400, cfl=0.4)
    U, *_ = oracle.run_loop(cons_from_prim_host(W, cfg.gamma), cfg, 200)
    Wp = oracle.cons_to_prim(oracle.apply_bc(U, cfg), cfg.gamma)
    wl, wr = oracle.muscl_x(Wp)
    _, ch = oracle.compute_dt(Wp, cfg)
    out["brio_wu_faces"] = (wl.reshape(10, -1), wr.reshape(10, -1), cfg.gamma, ch)
    return out


# Every k-th face of each case is kept in the committed fixture (keeps it
# small; each kept face is still the reference kernel's own output).
KEEP_EVERY = {"random_ch5": 4, "random_ch0": 4, "random_ch0p5_g2": 4, "degenerate_ch3": 2, "degenerate_ch0": 2,
              "brio_wu_faces": 4}


def subsample(name, WL, WR, F, s):
    k = KEEP_EVERY.get(name, 1)
    return WL[:, ::k], WR[:, ::k], F[:, ::k], s[::k]


def main(path=None):
    path = path or os.path.join(HERE, "hlld_ref_vectors.npz")
    lib = ref_lib()
    arrays = {}
    for name, (WL, WR, gamma, ch) in cases().items():
        F, s = run_ref(lib, WL, WR, gamma, ch)
        WL, WR, F, s = subsample(name, WL, WR, F, s)
        arrays.update({f"{name}__WL": WL, f"{name}__WR": WR, f"{name}__F": F, f"{name}__smax": s,
                       f"{name}__gamma": np.float64(gamma), f"{name}__ch": np.float64(ch)})
        print(name, WL.shape[1], "faces", flush=True)
    np.savez_compressed(path, **arrays)
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)
