"""GPU: failures cross the C ABI as status codes, with a reason, and leave
nothing behind.  Device memory is exhausted on purpose with ordinary
allocations (contexts of a 1020^3 FP32 grid, 42 GB each); the failing
dg_create must return DG_ERR_ALLOC, say why through dg_last_error(NULL), and
release what it had allocated: after the failure, as many contexts as before
can be created again."""
import ctypes
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_create_reports_device_allocation_failure_and_leaks_nothing():
    import deepgreen
    L = deepgreen.load_library()
    cfg = deepgreen.Simulation(1020, 1020, 1020, 1.0 / 1020, [0] * 6, precision="fp32").config(L)
    made, rc = [], 0
    for _ in range(12):   # 12 x 42 GB is far beyond one GPU's memory
        h = ctypes.c_void_p()
        rc = L.dg_create(ctypes.byref(cfg), ctypes.byref(h))
        print(f"  dg_create #{len(made) + 1}: {rc}", flush=True)
        sys.stdout.flush()
        if rc != 0:
            assert not h.value
            break
        made.append(h)
    try:
        assert rc == -3, rc                      # DG_ERR_ALLOC
        assert len(made) >= 1
        msg = L.dg_last_error(None)
        assert b"device allocation of the state" in msg, msg
    finally:
        for h in made:
            L.dg_destroy(h)
    again = []
    try:
        for _ in range(len(made)):
            h = ctypes.c_void_p()
            assert L.dg_create(ctypes.byref(cfg), ctypes.byref(h)) == 0, L.dg_last_error(None)
            again.append(h)
    finally:
        for h in again:
            L.dg_destroy(h)
