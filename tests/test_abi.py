"""CPU checks of the C ABI: the library loads, exports every function the
public headers declare, and rejects invalid arguments before touching a GPU."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import silt_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(hdr).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(silt_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert "silt_step_sheet" in names and len(names) >= 5
    lib = C.CDLL(S.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert S.lib().silt_version().decode().startswith("silt_amd")


def test_invalid_arguments_are_rejected_without_a_device():
    n = 16
    a = np.zeros(n, np.float32)
    g = np.ones(n, np.uint8)
    L = S.lib()
    d = S.Drive.gravity(0.0)
    # NULL plane
    assert L.silt_step_sheet(None, S._ptr(a), S._ptr(a), S._ptr(a), S._ptr(a), S._ptr(g), n, 4, 0.2, 0.01, d) \
        == S.SILT_EINVAL
    # unknown drive
    bad = S.Drive(7, 0.0, 0.0, 0.0)
    assert L.silt_step_sheet(S._ptr(a), S._ptr(a), S._ptr(a), S._ptr(a), S._ptr(a), S._ptr(g), n, 4, 0.2, 0.01, bad) \
        == S.SILT_EINVAL


@pytest.mark.parametrize("dt,res,length,depth", [(0.0, 4, 16, 0.1), (-1.0, 4, 16, 0.1), (0.01, 0, 16, 0.1),
                                                 (0.01, 4, 15, 0.1), (0.01, 4, 16, 0.0)])
def test_reference_no_ops_return_zero_and_leave_planes(dt, res, length, depth):
    """dt <= 0, res == 0, short slices and a dry sheet are silent no-ops
    (src/fluid_core.rs:840-847, 873-880); none of them reaches the GPU."""
    h = np.full(16, depth, np.float32)
    u = np.full(16, 0.3, np.float32)
    v = np.zeros(16, np.float32)
    bed = np.zeros(16, np.float32)
    nu = np.full(16, 0.9, np.float32)
    gate = np.ones(16, np.uint8)
    before = [x.copy() for x in (h, u, v)]
    L = S.lib()
    rc = L.silt_step_sheet(S._ptr(h), S._ptr(u), S._ptr(v), S._ptr(bed), S._ptr(nu), S._ptr(gate), length, res, 0.2,
                           dt, S.Drive.gravity(0.0))
    assert rc == 0
    for x, y in zip((h, u, v), before):
        assert np.array_equal(x, y)
