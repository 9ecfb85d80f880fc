"""GPU parity of the Bekker rut stamp against the oracle restatement, and the
rut feeding the sheet bed inside the FluidCore handle."""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu

R, B = 0.38, 0.28
K_EQ = 30e3 / B + 1.5e6


def _contacts(k):
    """Four tires at 3.5 kN each, rolling along +z, 8 cm per step."""
    out = []
    for x in (-0.7, 0.7):
        for dz in (-1.2, 1.2):
            out.append(dict(x=x, z=-2.0 + dz + 0.08 * k, hub_y=R, forward_x=0.0, forward_z=1.0, right_x=1.0,
                            right_z=0.0, radius=R, width=B, sinkage=0.02, normal=3500.0, push=400.0, k_eq=K_EQ, n=1.0))
    return out


def test_rut_stamp_matches_oracle_step_by_step():
    res, cell = 200, 0.2
    origin = (-20.0, -20.0)
    rng = np.random.default_rng(3)
    bed = (0.02 * rng.standard_normal(res * res)).astype(np.float32)
    placed = np.zeros(res * res, np.float32)
    init = bed.copy()
    ob, op = bed.copy(), placed.copy()
    db, dp, di = (DeviceArray(a) for a in (bed, placed, init))
    for k in range(10):
        cs = _contacts(k)
        for c in cs:
            O.stamp_tread(ob, op, init, res, cell, origin, c, 1.0 / 30.0)
        S.rut_stamp_device(db.ptr, dp.ptr, di.ptr, res, cell, origin, cs, 1.0 / 30.0)
        synchronize()
        assert np.array_equal(db.numpy(), ob), f"bed differs at step {k}"
        assert np.array_equal(dp.numpy(), op), f"placed differs at step {k}"
    assert (init - ob).max() > 0.01  # the tires cut a rut


def test_core_ruts_lower_the_sheet_bed_and_publish():
    core = S.FluidCore((0.0, 0.0))
    o = core.origin()
    n = 200 * 200
    core.set_ground(np.zeros(n, np.float32))
    core.pour_water(-4.0, 4.0, -4.0, 4.0, 0.05)
    for k in range(5):
        core.stamp_ruts(_contacts(k), 1.0 / 30.0)
        for c in _contacts(k):
            core.push_tire(x=c["x"], z=c["z"], forward_x=0.0, forward_z=1.0, radius=R, width=B, speed=2.0, slip=0.2)
        core.step(1.0 / 30.0)
    bed, placed, rut = core.ground()
    p = core.planes()
    assert np.allclose(p["bed"], bed + placed)
    assert rut.max() > 0.01 and np.allclose(rut, np.maximum(-bed, 0.0))
    gr, film, flow = core.publish()
    assert np.allclose(gr[0::4], bed + placed) and np.allclose(gr[1::4], rut)
