"""Pins of the tracer restatement (oracle/core.c: foam and stirred silt on the
window, src/fluid_core.rs:6210-6750, 9676-9880) by closed forms and known
answers:

* breaking_foam at the Froude midpoint is smoothstep(1/2) = 1/2;
* sample_foam reproduces a bilinear field exactly;
* foam carried by a velocity of exactly one cell per step moves one cell and
  decays by exp(-dt / 1.8);
* add_foam's relaxation toward 0.85, and the crest foam of note_surface_foam;
* one silt column carried one cell's fraction downstream, the column sum
  kept;
* the settling of still water, exp(-dt / 4), and the dry cell that lands all;
* the slipping contact patch's lift, min(0.015 dt, 0.0025) per cell;
* shore_cell on a hand-worked bank cell."""
import math

import numpy as np
import pytest

import pyoracle as O

N = 200 * 200
DT = np.float32(1.0 / 30.0)


def _core(water_depth=0.1, u=0.0, v=0.0):
    c = O.Core((0.0, 0.0))
    c.map_water_bed(np.zeros(N, np.float32), np.full(N, water_depth, np.float32))
    p = c.planes()
    p["uw"][:] = u
    p["vw"][:] = v
    c.set_planes(p)
    return c


def test_breaking_foam_froude_midpoint_and_gates():
    h = 0.05
    speed = 1.15 * math.sqrt(9.81 * h)
    assert O.breaking_foam(speed, h) == pytest.approx(0.5, abs=2e-6)
    assert O.breaking_foam(0.85 * math.sqrt(9.81 * h) * 0.999, h) == 0.0
    assert O.breaking_foam(1.46 * math.sqrt(9.81 * h), h) == 1.0
    assert O.breaking_foam(5.0, 0.0039) == 0.0 and O.breaking_foam(0.05, 0.1) == 0.0


def test_sample_foam_is_exact_on_a_bilinear_field():
    res = 6
    X, Z = np.meshgrid(np.arange(res), np.arange(res))
    f = (0.05 + 0.02 * X + 0.03 * Z + 0.004 * X * Z).astype(np.float64).ravel()
    for x, z in ((1.25, 2.5), (3.75, 0.5), (0.0, 4.0)):
        want = 0.05 + 0.02 * x + 0.03 * z + 0.004 * x * z
        assert O.sample_foam(f, res, x, z, precision=64) == pytest.approx(want, abs=1e-12)
    assert O.sample_foam(f, res, -0.1, 1.0, precision=64) == 0.0


def test_foam_moves_one_cell_per_step_and_decays():
    """u dt / cell = 1: the semi-Lagrangian sample lands on the upstream
    cell centre, so the pattern shifts by one cell, times exp(-dt / 1.8).
    0.2 m/s in 0.1 m of water is Froude 0.2: no new foam."""
    dt = 1.0
    cell = 0.2
    c = _core(0.1, u=cell / dt)
    foam = np.zeros((200, 200), np.float32)
    foam[100, 50:53] = (0.3, 0.5, 0.2)
    c.set_tracers(foam=foam.ravel())
    c.tracer_step(dt)
    got, _, _, _ = c.tracers()
    got = got.reshape(200, 200)
    keep = np.float32(math.exp(-dt / 1.8))
    for k, want in enumerate((0.3, 0.5, 0.2)):
        assert got[100, 51 + k] == pytest.approx(want * keep, rel=1e-5)
    assert got[100, 50] == 0.0


def test_add_foam_relaxes_toward_its_ceiling_and_crest_foam():
    """A crest 11 mm above its neighbours' mean: production smoothstep(0.004,
    0.018, 0.011) = 1/2, foam 0.85 * 0.5 * min(dt / 0.35, 1)."""
    c = _core(0.05)
    p = c.planes()
    p["water"].reshape(200, 200)[120, 60] = 0.05 + 0.011
    c.set_planes(p)
    dt = np.float32(0.1)
    c.note_surface_foam(dt)  # no last pushes, no live foam or silt: nothing is visited
    foam, _, _, _ = c.tracers()
    assert foam.max() == 0.0
    # With a tire in the last step the window is visited.
    c.push_tire(x=-15.0, z=-15.0, forward_x=1.0, forward_z=0.0, radius=0.38, width=0.28, speed=3.0, slip=0.0)
    c.tracer_step(dt)
    c.set_tracers(foam=np.zeros(N, np.float32))
    c.note_surface_foam(dt)
    foam, _, _, _ = c.tracers()
    want = 0.85 * 0.5 * min(float(dt) / 0.35, 1.0)
    assert foam.reshape(200, 200)[120, 60] == pytest.approx(want, rel=1e-5)
    # Twice: cur + (0.85 - cur) * 0.5 * dt / 0.35.
    c.note_surface_foam(dt)
    foam2, _, _, _ = c.tracers()
    want2 = want + (0.85 - want) * 0.5 * float(dt) / 0.35
    assert foam2.reshape(200, 200)[120, 60] == pytest.approx(want2, rel=1e-5)


def test_silt_column_moves_downstream_and_the_sum_is_kept():
    """One cell of silt at u = 0.6 m/s: the face carries column * min(u dt /
    cell, 0.45) to the next cell, then both settle by exp(-dt / 4)."""
    c = _core(0.1, u=0.6)
    t = np.zeros((200, 200), np.float32)
    t[90, 90] = 0.01
    c.set_tracers(turbid=t.ravel())
    c.tracer_step(DT)
    _, got, stirred, settled = c.tracers()
    got = got.reshape(200, 200).astype(np.float64)
    frac = min(0.6 * float(DT) / 0.2, 0.45)
    keep = math.exp(-float(DT) / 4.0)
    assert got[90, 91] == pytest.approx(0.01 * frac * keep, rel=1e-5)
    assert got[90, 90] == pytest.approx(0.01 * (1 - frac) * keep, rel=1e-5)
    assert got.sum() + settled / 0.04 == pytest.approx(0.01, rel=1e-5)
    assert stirred == 0.0


def test_dry_cells_land_their_silt():
    c = _core(0.0)
    t = np.zeros(N, np.float32)
    t[5000] = 0.002
    c.set_tracers(turbid=t)
    c.tracer_step(DT)
    _, got, _, settled = c.tracers()
    assert got.max() == 0.0 and settled == pytest.approx(0.002 * 0.04, rel=1e-6)


def test_slipping_contact_patch_lifts_silt():
    """A tire spinning (speed 4 m/s, slip 0.5) in 0.1 m of still water:
    shear 2 m/s, tau = 35 * 4 = 140 Pa over tau_c 6, the rate caps at 0.015
    m/s, dh = min(0.015 dt, 0.0025) on each contact-patch cell."""
    c = _core(0.1)
    c.set_near_field(False)
    c.push_tire(x=0.1, z=0.1, forward_x=0.0, forward_z=1.0, radius=0.38, width=0.28, speed=4.0, slip=0.5)
    c.step(DT)
    _, got, stirred, _ = c.tracers()
    dh = min(0.015 * float(DT), 0.0025)
    lifted = got[got > 0]
    assert lifted.size > 0
    assert np.allclose(lifted, dh, rtol=1e-5)
    assert stirred == pytest.approx(lifted.size * dh * 0.04, rel=1e-5)


def test_shore_cell_known_answer():
    """A wet bank cell with a dry neighbour at +x, running at 0.5 m/s into
    it in 3 cm of water: into = 0.5, production max(breaking_foam(0.5,
    0.03), smoothstep(0.25, 1.1, 0.5)); the silt rate 0.006 (0.5 + 0.5 *
    0.35) under its 0.01 cap."""
    res = 5
    water = np.full(res * res, 0.03, np.float32)
    water[2 * res + 3] = 0.0
    uw = np.zeros(res * res, np.float32)
    vw = np.zeros(res * res, np.float32)
    idx = 2 * res + 2
    uw[idx] = 0.5
    prod, rate = O.shore_cell(water, uw, vw, res, idx)
    fr = 0.5 / math.sqrt(9.81 * 0.03)
    t_b = min(max((fr - 0.85) / 0.6, 0.0), 1.0)
    t_s = (0.5 - 0.25) / 0.85
    want = max(t_b * t_b * (3 - 2 * t_b), t_s * t_s * (3 - 2 * t_s))
    assert prod == pytest.approx(want, rel=1e-5)
    # strain |uw[+1] - uw[-1]| = 0; shear 0.5 > 0.45, tau = 25 * 0.25 = 6.25 > 3.
    assert rate == pytest.approx(0.006 * (0.5 + 0.5 * 0.35), rel=1e-5)
