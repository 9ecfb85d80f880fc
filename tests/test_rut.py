"""Bekker rut coupling: oracle pins (CPU) and GPU parity of the rut stamp.

Pins: the rigid-wheel chord integral against its closed form for n = 1, the
bisection's resolution, the reference's caps, terramech's end points, and
volume conservation of the cut and its spoil."""
import math

import numpy as np
import pytest

import pyoracle as O

R, B = 0.38, 0.28  # WHEEL_RADIUS, WHEEL_WIDTH (src/vehicle.rs:15-16)
K_EQ = 30e3 / B + 1.5e6  # k_c / b + k_phi for k_c = 30 kN/m^2, k_phi = 1500 kN/m^3, n = 1


def segment_area(r, z):
    return r * r * math.acos((r - z) / r) - (r - z) * math.sqrt(2 * r * z - z * z)


@pytest.mark.parametrize("z", [0.005, 0.02, 0.05, 0.1, 0.2, 0.3])
def test_bekker_load_closed_form_for_n_one(z):
    """For n = 1 the load is k_eq * b * (area of the circular segment under the
    chord); bekker_load's 8 midpoint samples of the concave depth profile
    overestimate it by a few per cent at most (src/soil.rs:1224-1245)."""
    exact = K_EQ * B * segment_area(R, z)
    mid = O.bekker_load(z, R, B, K_EQ, 1.0, precision=64)
    assert exact <= mid <= exact * 1.03, (z, mid, exact)


def test_sinkage_bisection_brackets_the_load():
    """18 halvings of [0, 0.9 r] (src/soil.rs:932-946): the returned sinkage is
    within 0.9 r / 2^18 of the point where the chord load crosses the tire load."""
    step = 0.9 * R / 2 ** 18
    for load in (500.0, 3500.0, 20000.0):
        s = O.sinkage_with(load, R, B, K_EQ, 1.0, precision=64)
        assert O.bekker_load(s - step, R, B, K_EQ, 1.0, precision=64) < load
        assert O.bekker_load(s + step, R, B, K_EQ, 1.0, precision=64) >= load


def test_rut_depth_caps_and_threshold():
    """rut_depth_from_force (src/soil.rs:1036-1061): under 1 N no rut; the
    resultant of normal and push; never deeper than 0.85 r."""
    assert O.rut_depth_from_force(0.5, 0.5, R, B, K_EQ, 1.0) == 0.0
    a = O.rut_depth_from_force(3000.0, 4000.0, R, B, K_EQ, 1.0)
    b = O.rut_depth_from_force(5000.0, 0.0, R, B, K_EQ, 1.0)
    assert a == b > 0.0
    assert O.rut_depth_from_force(1e9, 0.0, R, B, 1000.0, 1.0) == pytest.approx(0.85 * R, rel=1e-6)


def test_terramech_end_points():
    assert O.terramech(0.0, 64) == pytest.approx((320000.0, 0.95))
    assert O.terramech(1.0, 64) == pytest.approx((48000.0, 0.72))
    assert O.softness_of(1.0, 0.09, 0.18, 64) == pytest.approx(1.0)


def _contact(x, z, load=3500.0):
    return dict(x=x, z=z, hub_y=R, forward_x=0.0, forward_z=1.0, right_x=1.0, right_z=0.0, radius=R, width=B,
                sinkage=0.02, normal=load, push=0.0, k_eq=K_EQ, n=1.0)


def test_stamp_moves_soil_into_the_berm_without_loss():
    """The cut volume lands on the berm band (deposit_spoil, src/soil.rs:2881-2946):
    bed + placed integrates to the same volume; the floor is no deeper than the
    Bekker depth; a straight contact cuts a mirror-symmetric rut."""
    res, cell = 64, 0.2
    origin = (-res * cell / 2, -res * cell / 2)
    bed = np.zeros(res * res, np.float64)
    placed = np.zeros_like(bed)
    initial = bed.copy()
    c = _contact(0.9, 0.0)
    forced = O.rut_depth_from_force(c["normal"], c["push"], R, B, K_EQ, 1.0, precision=64)
    total = 0.0
    for _ in range(30):
        total += O.stamp_tread(bed, placed, initial, res, cell, origin, c, 1.0 / 30.0, precision=64)
    assert total > 1e-4
    # deposit_spoil stops once less than 1e-8 m^3 is left to place.
    assert abs((bed + placed).sum()) * cell * cell <= 30 * 1e-8
    assert -bed.min() <= forced + 1e-9
    B2 = bed.reshape(res, res)
    i0 = int(round((c["x"] - origin[0]) / cell - 0.5))
    col = B2[:, i0]
    assert col.min() < -0.5 * forced
