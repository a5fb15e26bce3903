"""CPU tests of the host side: the C-ABI library loads and exports every
symbol include/aeromhd.h declares; the Config mirror; the Powell gate and
BC mapping; the initial-condition generators (Rankine-Hugoniot closed form
and jump residuals, sieve interface, seeded draws)."""
import ctypes as C
import dataclasses
import math
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "aeromhd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(amhd_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    import amhd_lib

    path = amhd_lib.LIB_PATH
    assert os.path.exists(path), "libaeromhd_hip.so not built"
    lib = C.CDLL(path)
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(amhd_lib.SIGNATURES), "ctypes binding out of step with the header"
    lib.amhd_abi_version.restype = C.c_int
    assert lib.amhd_abi_version() == 1


def test_closure_source_checks_arguments_before_device_work():
    """amhd_closure_source refuses a network output that is not 16-byte
    aligned (its kernels read 4 channels per load), and a field order without
    rho first and rhoC last, before touching any memory; no device needed."""
    import numpy as np
    import amhd_lib

    lib = amhd_lib.load()
    fields = (C.c_int * 9)(0, 1, 2, 3, 4, 5, 6, 7, 9)
    fs = (C.c_float * 9)(*([1.0] * 9))
    dt = (C.c_double * 1)(1e-3)
    buf = np.zeros(64, dtype=np.float32)
    base = buf.ctypes.data + (-buf.ctypes.data) % 16
    P = C.c_void_p * 1
    st, out = P(base), P(base)

    def call(dF, flds):
        return lib.amhd_closure_source(C.c_void_p(dF), st, out, amhd_lib.AMHD_F32, flds, 1, 4, 4, fs, 0.1, 0.1, dt,
                                       1.4, 0.5, None, None, 0, None)

    assert call(base + 4, fields) == -1
    assert b"16-byte aligned" in lib.amhd_last_error()
    bad = (C.c_int * 9)(1, 0, 2, 3, 4, 5, 6, 7, 9)
    assert call(base, bad) == -1
    assert b"rho" in lib.amhd_last_error()


def test_closure_sharpen_checks_arguments_before_device_work():
    """amhd_closure_sharpen refuses a negative, NaN or infinite kappa, a bad
    dtype, a null state list and, with kappa > 0, a null workspace -- before
    any launch; kappa = 0 returns at once with nothing to do.  The host-side
    solvers refuse such a kappa too; no device needed."""
    import numpy as np
    import amhd_closure as cl
    import amhd_lib

    lib = amhd_lib.load()
    buf = np.zeros(64, dtype=np.float32)
    st = (C.c_void_p * 1)(buf.ctypes.data)
    work = C.c_void_p(buf.ctypes.data)

    def call(states, dtype, kappa, w):
        return lib.amhd_closure_sharpen(states, None, dtype, 1, 4, 4, kappa, w, None)

    for k in (-1.0, float("nan"), float("inf")):
        assert call(st, amhd_lib.AMHD_F32, k, work) == -1
        assert b"bad argument" in lib.amhd_last_error()
    assert call(st, 7, 0.1, work) == -1
    assert b"dtype" in lib.amhd_last_error()
    assert call(None, amhd_lib.AMHD_F32, 0.1, work) == -1
    assert call(st, amhd_lib.AMHD_F32, 0.1, None) == -1
    assert b"work" in lib.amhd_last_error()
    assert call(st, amhd_lib.AMHD_F32, 0.0, None) == 0
    from amhd_ics import magnetic_sieve

    cfg, _ = magnetic_sieve(nx=256, ny=128, precision="fp32", seed=0)
    for k in (-0.5, float("nan")):
        with pytest.raises(ValueError, match="kappa"):
            cl.ClosureSolver(cfg, None, kappa=k)
        with pytest.raises(ValueError, match="kappa"):
            cl.BatchClosure([(cfg, None)], None, kappa=k)


def test_params_struct_layout_matches_header():
    import amhd_lib

    assert C.sizeof(amhd_lib.AmhdParams) == 6 * 8 + 4 * 4


def test_config_mirrors_reference_fields():
    from config import Config, ConsVar, PrimVar, NVAR, FLOOR_RHO, FLOOR_PR

    names = [f.name for f in dataclasses.fields(Config)]
    ref = ["nx", "ny", "x_min", "x_max", "y_min", "y_max", "t_end", "cfl", "max_steps", "gamma", "mach",
           "interface_x", "perturbation_amp", "perturbation_mode", "density_ratio", "interface_width",
           "B_transverse", "glm_ch", "glm_alpha", "powell_source", "use_char_bc", "bc_x_type", "bc_y_type",
           "enable_smoothing", "diag_interval", "snapshot_times"]
    assert names[:len(ref)] == ref
    c = Config()
    assert (c.nx, c.ny, c.cfl, c.gamma, c.glm_alpha, c.perturbation_amp) == (400, 200, 0.30, 5.0 / 3.0, 1.5, 0.15)
    assert c.dx == 6.0 / 400 and c.dy == 2.0 / 200
    assert c.get_bc_x() == "characteristic"
    assert Config(use_char_bc=False).get_bc_x() == "extrapolation"
    assert Config(bc_x_type="periodic").get_bc_x() == "periodic"
    assert NVAR == 10 and FLOOR_RHO == FLOOR_PR == 1e-12
    assert [int(v) for v in ConsVar] == list(range(10)) and [int(v) for v in PrimVar] == list(range(10))


def test_solver_params_gate_and_bcs():
    from config import Config
    from solver import solver_params

    p = solver_params(Config(B_transverse=0.0))
    assert p.powell == 0 and p.bc_x_periodic == 0 and p.bc_y_periodic == 1
    p = solver_params(Config(B_transverse=0.5))
    assert p.powell == 1
    p = solver_params(Config(B_transverse=0.5, powell_source=False, bc_x_type="periodic", bc_y_type="outflow"))
    assert p.powell == 0 and p.bc_x_periodic == 1 and p.bc_y_periodic == 0


def test_rankine_hugoniot_hydro_closed_form():
    from physics import mhd_rankine_hugoniot

    g, M = 5.0 / 3.0, 10.0
    rho2, p2, vx2, by2, vs = mhd_rankine_hugoniot(g, M, 1.0, 1.0, 0.0)
    r = (g + 1) * M * M / ((g - 1) * M * M + 2)
    assert rho2 == pytest.approx(r) and by2 == 0.0
    assert vs == pytest.approx(M * math.sqrt(g))
    assert p2 == pytest.approx((2 * g * M * M - (g - 1)) / (g + 1))
    assert vx2 == pytest.approx(vs * (1 - 1 / r))


@pytest.mark.parametrize("by1", [0.2, math.sqrt(0.2), 0.8])
def test_rankine_hugoniot_mhd_jump_residuals(by1):
    """The perpendicular-shock solution satisfies mass, momentum, energy and
    flux-freezing jump conditions in the shock frame."""
    from physics import mhd_rankine_hugoniot

    g = 5.0 / 3.0
    rho2, p2, vx2, by2, vs = mhd_rankine_hugoniot(g, 4.0, 1.0, 1.0, by1)
    u1, u2 = vs, vs - vx2          # shock-frame speeds
    assert rho2 * u2 == pytest.approx(1.0 * u1, rel=1e-12)
    mom1 = 1.0 * u1 ** 2 + 1.0 + 0.5 * by1 ** 2
    mom2 = rho2 * u2 ** 2 + p2 + 0.5 * by2 ** 2
    assert mom2 == pytest.approx(mom1, rel=1e-12)
    assert by2 / rho2 == pytest.approx(by1 / 1.0, rel=1e-12)
    h1 = 0.5 * u1 ** 2 + g / (g - 1) * 1.0 / 1.0 + by1 ** 2 / 1.0
    h2 = 0.5 * u2 ** 2 + g / (g - 1) * p2 / rho2 + by2 ** 2 / rho2
    assert h2 == pytest.approx(h1, rel=1e-9)


def test_sieve_interface_rms_and_spectrum():
    from amhd_ics import sieve_interface

    y = (np.arange(4096) + 0.5) * 2.0 / 4096
    x_if, ks, amps, _ = sieve_interface(y, 2.0, 1.5, 0.01, seed=3)
    assert np.sqrt(np.mean((x_if - 1.5) ** 2)) == pytest.approx(0.01, rel=1e-6)
    assert ks[0] == 4 and ks[-1] == 32
    env = amps * ks.astype(float) ** 2
    assert env.max() / env.min() <= 3.0 + 1e-12      # k^-2 envelope, U[0.5,1.5] factors
    x2, *_ = sieve_interface(y, 2.0, 1.5, 0.01, seed=3)
    np.testing.assert_array_equal(x_if, x2)


def test_magnetic_sieve_ic_small():
    from amhd_ics import magnetic_sieve
    from physics import mhd_rankine_hugoniot

    cfg, W = magnetic_sieve(nx=64, ny=32, seed=1)
    assert cfg.B_transverse > 0 and cfg.precision == "fp32"
    b0 = math.sqrt(0.2)
    y = cfg.y_min + (np.arange(36) - 2 + 0.5) * cfg.dy
    by_pre = b0 * (1 + 0.5 * np.sin(2 * np.pi * y / 0.125))
    np.testing.assert_allclose(W[5, -1], by_pre)
    rho2, p2, vx2, by2, _ = mhd_rankine_hugoniot(cfg.gamma, 10.0, 1.0, 1.0, float(by_pre[7]))
    assert W[0, 0, 7] == rho2 and W[5, 0, 7] == by2 and W[7, 0, 7] == p2 and W[1, 0, 7] == vx2


def test_batched_draws_reproducible_and_in_range():
    from amhd_ics import batched_instances

    a = batched_instances(n=8, seed=4, max_steps=5, nx_range=(64, 256))
    b = batched_instances(n=8, seed=4, max_steps=5, nx_range=(64, 256))
    for (ca, wa), (cb, wb) in zip(a, b):
        assert (ca.nx, ca.ny, ca.mach) == (cb.nx, cb.ny, cb.mach)
        np.testing.assert_array_equal(wa, wb)
        assert ca.nx % 64 == 0 and ca.ny == ca.nx // 2 and 2 <= ca.mach <= 12
        assert 0.2 <= (ca.density_ratio - 1) / (ca.density_ratio + 1) <= 0.8


def test_closure_rejects_unsupported_boundaries():
    """The closure's kernels assume periodic y and non-periodic x boundaries
    (the RMI runs it is trained on); other configurations are refused before
    any device work."""
    from amhd_closure import check_closure_boundaries
    from config import Config

    check_closure_boundaries(Config(bc_x_type="extrapolation", bc_y_type="periodic"))
    check_closure_boundaries(Config(bc_x_type="auto", use_char_bc=True, bc_y_type="periodic"))
    for bx, by in (("periodic", "periodic"), ("extrapolation", "outflow"), ("characteristic", "extrapolation")):
        with pytest.raises(ValueError, match="periodic y"):
            check_closure_boundaries(Config(bc_x_type=bx, bc_y_type=by))
