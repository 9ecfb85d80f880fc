"""The rut stamp entry point that allocates nothing
(silt_rut_stamp_device_noalloc, include/silt_rut.h) against the one that
grows its own contact buffer (silt_rut_stamp_device): the same planes bit for
bit, over calls in a row that shrink, on one GPU's shared array.  Sizes over
the array's maximum are rejected before any work."""
import ctypes as C

import numpy as np
import pytest

import silt_amd as S
from hipmem import DeviceArray, synchronize
from silt_amd import drive as D

pytestmark = pytest.mark.gpu

DT = 1.0 / 30.0


def _grid(res, seed):
    rng = np.random.default_rng(seed)
    bed = (0.02 * rng.standard_normal(res * res)).astype(np.float32)
    return bed, np.zeros(res * res, np.float32), bed.copy()


def _both(res, cell, origin, contacts, planes):
    """Stamp the same contacts with both entry points into two copies of the
    planes; returns the planes of each."""
    (bed, placed, init) = planes
    a = [DeviceArray(x) for x in (bed, placed, init)]
    b = [DeviceArray(x) for x in (bed, placed, init)]
    S.rut_stamp_device(a[0].ptr, a[1].ptr, a[2].ptr, res, cell, origin, contacts, DT)
    S.rut_stamp_device_noalloc(b[0].ptr, b[1].ptr, b[2].ptr, res, cell, origin, contacts, DT)
    synchronize()
    return [x.numpy() for x in a[:2]], [x.numpy() for x in b[:2]]


def test_noalloc_stamp_equals_the_allocating_stamp_on_shrinking_calls():
    cases = [
        # (res, vehicles, contacts used): each call smaller than the one before
        (256, 6, 24),
        (256, 6, 7),
        (160, 2, 5),
    ]
    moved = 0.0
    for k, (res, vehicles, used) in enumerate(cases):
        cell = 0.2
        origin = (-0.5 * res * cell, -0.5 * res * cell)
        fleet = D.Fleet(seed=30 + k, n=vehicles, extent=res * cell, half=3.0)
        for _ in range(3 * k):
            fleet.advance(DT)
        contacts = fleet.contacts()[:used]
        planes = _grid(res, 40 + k)
        (ab, ap), (bb, bp) = _both(res, cell, origin, contacts, planes)
        assert np.array_equal(ab, bb), f"call {k}: bed differs"
        assert np.array_equal(ap, bp), f"call {k}: placed differs"
        moved = max(moved, float((planes[0] - bb).max()))
    assert moved > 0.005  # the contacts cut ruts


def test_noalloc_stamp_rejects_more_contacts_than_its_array():
    res, cell = 64, 0.2
    bed, placed, init = _grid(res, 1)
    d = [DeviceArray(x) for x in (bed, placed, init)]
    c = dict(x=0.0, z=0.0, hub_y=0.38, forward_x=0.0, forward_z=1.0, right_x=1.0, right_z=0.0, radius=0.38,
             width=0.28, sinkage=0.02, normal=3500.0, push=400.0, k_eq=30e3 / 0.28 + 1.5e6, n=1.0)
    n = S.NOALLOC_MAX_CONTACTS + 1
    arr = (S.TireContact * n)(*[S.TireContact(**c) for _ in range(n)])
    L = S._rut_lib()
    rc = L.silt_rut_stamp_device_noalloc(d[0].ptr, d[1].ptr, d[2].ptr, res, cell, -6.4, -6.4, arr, n, DT, None)
    assert rc == S.SILT_ETOOBIG
    synchronize()
    assert np.array_equal(d[0].numpy(), bed) and np.array_equal(d[1].numpy(), placed)
    assert C.sizeof(S.TireContact) == 14 * 4
