"""amhd_compute_dt_w_noalloc (CFL dt and glm_ch of solver.py:206-223 with
its scratch in a file-scope device array) against amhd_compute_dt_w, bit
for bit, on two calls in a row, the second with a smaller input (the
array keeps the first call's partials, which must not reach the result)."""
import ctypes as C

import numpy as np
import pytest

from helpers import NV, random_prim

pytestmark = pytest.mark.gpu


def _dt(lib, name, W, n, gamma, cfl, dx, dy):
    dt, ch = C.c_double(), C.c_double()
    rc = getattr(lib, name)(C.c_void_p(W.data_ptr()), n, gamma, cfl, dx, dy, C.byref(dt), C.byref(ch), None)
    assert rc == 0, lib.amhd_last_error()
    return dt.value, ch.value


def test_noalloc_dt_equals_allocating_dt(torch_gpu):
    import amhd_lib

    torch = torch_gpu
    lib = amhd_lib.load()
    rng = np.random.default_rng(31)
    # first call: 600k cells (the grid is capped at 1024 partials, each
    # covering several blocks' worth of cells); the fastest cells sit near
    # the end so the maxima come from the last partials
    got = []
    for n, vmax, hot in ((600_000, 3.0, 599_990), (5_000, 1.0, 17)):
        Wh = random_prim(rng, n, vmax=vmax)
        Wh[1, hot] = 9.0 if n > 10_000 else 4.0
        if n > 10_000:
            Wh[0, 1234] = 1e-14      # a density below the floor
        W = torch.from_numpy(Wh).cuda()
        torch.cuda.synchronize()
        a = _dt(lib, "amhd_compute_dt_w", W, n, 5.0 / 3.0, 0.3, 1.0 / 512, 1.0 / 256)
        b = _dt(lib, "amhd_compute_dt_w_noalloc", W, n, 5.0 / 3.0, 0.3, 1.0 / 512, 1.0 / 256)
        print(f"n {n}: dt {a[0]!r} / {b[0]!r}, ch {a[1]!r} / {b[1]!r}", flush=True)
        assert a == b
        got.append(b)
    # the second call's maxima are far below the first's: a stale partial of
    # the first call reaching the second result would have kept ch
    assert got[1][1] < 1e-3 * got[0][1]


def test_noalloc_dt_rejects_bad_arguments(torch_gpu):
    import amhd_lib

    lib = amhd_lib.load()
    dt, ch = C.c_double(), C.c_double()
    assert lib.amhd_compute_dt_w_noalloc(None, 10, 1.4, 0.3, 0.1, 0.1, C.byref(dt), C.byref(ch), None) == -1
    W = torch_gpu.zeros((NV, 4), dtype=torch_gpu.float64, device="cuda")
    assert lib.amhd_compute_dt_w_noalloc(C.c_void_p(W.data_ptr()), 0, 1.4, 0.3, 0.1, 0.1, C.byref(dt), C.byref(ch),
                                         None) == -1
    assert b"bad argument" in lib.amhd_last_error()
