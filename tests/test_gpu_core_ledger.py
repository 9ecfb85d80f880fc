"""The FluidCore ledger closed in FP64 step by step (silt_core_ledger_lines,
include/silt_core.h): every kernel of the step that changes a depth books
what it changed, so the gap (volume no line accounts for) stays at the FP64
summation noise of the ledger's own sums.  Covered: the nested 2 cm patch
with the near field (refine, lay-down, restrict, spray), the coarse-only
model, and the surrogate with judged steps whose lines are undone with
their planes.  The reference's tolerance is 1e-3 relative
(src/fluid_core.rs:11635-11638); the bound here is 1e-9 of the total per
step and over the run."""
import numpy as np
import pytest

import silt_amd as S
from silt_amd import drive as D
from silt_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 30.0


def _core(seed, path, model=None, judge_every=0):
    c = D.nested_scene_core(S, scenes, seed=seed, path=path)
    if model == "surrogate":
        m = S.Mlp(7)
        c.set_surrogate(m.params(), judge_every=judge_every, tol=1e-3)
        m.close()
    if model:
        c.set_tire_model(model)
    return c


def _drive_and_check(c, drv, steps, ponds=False):
    first = c.ledger_lines()
    for l in (0, 1):
        assert abs(first["gap_m3"][l]) <= 1e-15 * first["total_m3"][l]
    prev = first
    for k in range(steps):
        D.drive_step(c, drv, DT, ponds=ponds)
        L = c.ledger_lines()
        for l, name in enumerate(("water", "slurry")):
            tot = L["total_m3"][l]
            assert tot > 0.0
            step_gap = L["gap_m3"][l] - prev["gap_m3"][l]
            assert abs(step_gap) <= 1e-9 * tot, (name, k, step_gap, tot)
            assert abs(L["gap_m3"][l]) <= 1e-9 * tot, (name, k, L["gap_m3"][l], tot)
            # The lines are rounding-sized: the FP32 residue and the
            # handoffs at most 1e-6 of the total per step, the dry cut
            # non-negative and no more than 1e-6 m on each window cell.
            assert abs(L["residue_m3"][l] - prev["residue_m3"][l]) <= 1e-6 * tot, (name, k)
            if not ponds:  # the seat's line also holds water no floor takes back
                assert abs(L["handoff_m3"][l] - prev["handoff_m3"][l]) <= 1e-6 * tot, (name, k)
            cut = L["cut_m3"][l] - prev["cut_m3"][l]
            assert 0.0 <= cut <= 1e-6 * 200 * 200 * 0.04 * 9, (name, k, cut)
        prev = L
    return prev


@pytest.mark.parametrize("path", ["exact", "fast"])
def test_nested_steps_close_the_ledger_per_step(path):
    c = _core(31, path)
    try:
        last = _drive_and_check(c, D.Drive(seed=32), 40)
        led = c.ledger()
        # The nested patch ran: spray was peeled and the handoffs booked.
        assert led["spray_m3"] > 0.0
        assert last["handoff_m3"][0] != 0.0
        assert last["residue_m3"][0] != 0.0
    finally:
        c.close()


def test_steps_with_pond_and_seat_close_the_ledger_per_step():
    """ThroughDrive's order: each step is followed by pond_ruts and
    seat_rut_sheet (src/world.rs:4686-4694), which book what they move."""
    c = _core(39, "fast")
    try:
        first = c.ledger_lines()
        last = _drive_and_check(c, D.Drive(seed=40), 30, ponds=True)
        assert last["handoff_m3"][0] != first["handoff_m3"][0]
    finally:
        c.close()


def test_coarse_model_closes_the_ledger_per_step():
    c = _core(33, "fast", model="coarse")
    try:
        _drive_and_check(c, D.Drive(seed=34), 30)
    finally:
        c.close()


def test_surrogate_with_judged_steps_closes_the_ledger_per_step():
    """Every fourth step is judged: the surrogate step runs, its planes and
    lines are put back, the nested step runs in its place."""
    c = _core(35, "fast", model="surrogate", judge_every=4)
    try:
        _drive_and_check(c, D.Drive(seed=36), 24)
        assert c.surrogate_stats()["judged"] >= 2
    finally:
        c.close()


def test_a_caller_edit_moves_the_base_not_the_gap():
    """A pour adds volume no step made: the base moves to the new total and
    the gap starts again from zero."""
    c = _core(37, "fast")
    drv = D.Drive(seed=38)
    try:
        for _ in range(5):
            D.drive_step(c, drv, DT)
        before = c.ledger_lines()
        c.pour_water(-3.0, 3.0, 4.0, 8.0, 0.05)
        after = c.ledger_lines()
        added = after["total_m3"][0] - before["total_m3"][0]
        assert added > 0.1
        assert after["base_m3"][0] - before["base_m3"][0] == pytest.approx(added, rel=1e-9)
        for l in (0, 1):
            assert abs(after["gap_m3"][l]) <= 1e-15 * after["total_m3"][l]
        _drive_and_check(c, drv, 10)
    finally:
        c.close()
