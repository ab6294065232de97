"""CPU-side checks of the HIP library: it loads, exports every entry point
declared in include/deepgreen.h, and validates configurations before it
touches a device.  No compute call is made."""
import ctypes
import os
import re

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def header_entry_points():
    txt = open(os.path.join(ROOT, "include", "deepgreen.h")).read()
    return sorted(set(re.findall(r"\b(dg_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_entry_point():
    import deepgreen
    L = deepgreen.load_library()
    declared = header_entry_points()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(deepgreen.EXPORTS) == declared


def test_config_default_is_reference_constants():
    import deepgreen
    c = deepgreen.DgConfig()
    deepgreen.load_library().dg_config_default(ctypes.byref(c))
    assert c.gamma == 5.0 / 3.0 and c.cfl == 0.2 and c.g == -0.1  # globals.rs:3-7
    assert c.nghosts == 2 and list(c.bc) == [0] * 6 and c.limiter == 0


@pytest.mark.parametrize("mutate", [
    lambda c: setattr(c, "nghosts", 3),
    lambda c: c.n.__setitem__(0, 1),
    lambda c: c.bc.__setitem__(0, 7),
    lambda c: c.bc.__setitem__(1, 1),           # unpaired periodic face
    lambda c: setattr(c, "ds", 0.0),
    lambda c: setattr(c, "gamma", 1.0),
    lambda c: setattr(c, "precision", 5),
    lambda c: setattr(c, "limiter", 9),
    lambda c: setattr(c, "tile_y", 5),
    lambda c: (setattr(c, "math", 1), setattr(c, "ds", 1e-6)),   # fast FP32 outside its ds range
    lambda c: (setattr(c, "math", 1), setattr(c, "ds", 1e5)),
])
def test_invalid_configs_are_rejected_before_device_use(mutate):
    import deepgreen
    L = deepgreen.load_library()
    c = deepgreen.DgConfig()
    L.dg_config_default(ctypes.byref(c))
    mutate(c)
    h = ctypes.c_void_p()
    assert L.dg_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert not h.value


def test_null_arguments_return_errors():
    import threading

    import deepgreen
    L = deepgreen.load_library()
    seen = []
    t = threading.Thread(target=lambda: seen.append(L.dg_last_error(None)))
    t.start()
    t.join()
    assert seen == [b"null context"]          # no failed dg_create on that thread
    assert L.dg_create(None, None) == -1
    assert L.dg_last_error(None) == b"null argument"
    assert L.dg_step(None, None) == -1


def test_failed_create_explains_itself():
    """dg_last_error(NULL) gives the reason of the calling thread's last failed
    dg_create (no context exists to hold it)."""
    import deepgreen
    L = deepgreen.load_library()
    c = deepgreen.DgConfig()
    L.dg_config_default(ctypes.byref(c))
    c.bc[0] = 1                                   # x- outflow against a periodic x+
    h = ctypes.c_void_p()
    assert L.dg_create(ctypes.byref(c), ctypes.byref(h)) == -1 and not h.value
    assert L.dg_last_error(None) == b"invalid configuration: a periodic face needs a periodic opposite face"
    c.bc[0] = 0
    c.n[0] = c.n[1] = c.n[2] = 2048               # 32 GiB per field: beyond the 32-bit offsets
    assert L.dg_create(ctypes.byref(c), ctypes.byref(h)) == -1
    assert b"exceeds 4 GiB" in L.dg_last_error(None)
