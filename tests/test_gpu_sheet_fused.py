"""GPU parity of the fused streaming kernel.

* fused_exact (the streaming kernel with the reference's operation order) must
  equal the stage-kernel exact path value for value: this pins the streaming
  schedule (rings, ghosts, box masks, chunk and strip seams, ping-pong).
* fast (fused multiply-add, reciprocals) is held to the north-star tolerance
  against the FP64 oracle on one step from the same FP32 state:
  |gpu - oracle64| <= 1e-5 * max|oracle64| per field.
"""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from silt_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 1.0e-5


def _eq(a, b, what):
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    assert bad.size == 0, f"{what}: {bad.size} cells differ, first {bad[:5]} {a[bad[:5]]} vs {b[bad[:5]]}"


def _run_pair(state, bed, nu, gate, res, cell, dt, drive, steps, **opts):
    ex = [a.copy() for a in state]
    fu = [a.copy() for a in state]
    for step in range(steps):
        n1 = S.step_sheet(*ex, bed, nu, gate, res, cell, dt, drive, path="exact", **opts)
        n2 = S.step_sheet(*fu, bed, nu, gate, res, cell, dt, drive, path="fused_exact", **opts)
        assert n1 == n2
        for a, b, name in zip(fu, ex, "huv"):
            _eq(a, b, f"{name} step {step}")


def _window(seed, layer="water"):
    sc = scenes.window_scene(seed=seed)
    n = sc["res"] ** 2
    return sc, [sc[layer].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)]


def test_fused_exact_equals_stage_path_on_window_scene():
    sc, st = _window(1)
    n = sc["res"] ** 2
    _run_pair(st, sc["bed"], np.full(n, 0.9, np.float32), np.ones(n, np.uint8), sc["res"], sc["cell"],
              1.0 / 120.0, ("gravity", 0.08), 25)


def test_fused_exact_equals_stage_path_with_cell_viscosity_and_gate():
    sc, st = _window(2, "slurry")
    n = sc["res"] ** 2
    rng = np.random.default_rng(1)
    nu = (0.3 + 60.0 * rng.random(n)).astype(np.float32)
    gate = (rng.random(n) > 0.05).astype(np.uint8)
    st[0] += sc["water"]
    _run_pair(st, sc["bed"], nu, gate, sc["res"], sc["cell"], 1.0 / 60.0, ("gravity", 0.05), 15)


def test_fused_exact_on_split_boxes_small_boxes_and_wind():
    res, cell = 300, 0.2
    n = res * res
    H = np.zeros((res, res), np.float32)
    H[10:40, 20:50] = 0.08
    H[150:260, 100:290] = 0.05
    H[280:299, 0:12] = 0.03  # touches the grid edge
    rng = np.random.default_rng(5)
    bed = (0.05 * rng.standard_normal((res, res))).astype(np.float32).ravel()
    nu = np.full(n, 0.9, np.float32)
    gate = np.ones(n, np.uint8)
    assert len(O.work_boxes(H.ravel(), res, 1e-6)) == 3
    _run_pair([H.ravel(), np.zeros(n, np.float32), np.zeros(n, np.float32)], bed, nu, gate, res, cell, 1.0 / 30.0,
              ("gravity", 0.08), 12, confine="jacobi")
    wres, wcell = 64, 2.0 * (200 * 0.2 / 128)
    wn = wres * wres
    wh = np.zeros((wres, wres), np.float32)
    wh[24:36, 12:24] = 0.04
    _run_pair([wh.ravel(), np.zeros(wn, np.float32), np.zeros(wn, np.float32)], np.zeros(wn, np.float32),
              np.full(wn, 0.4, np.float32), np.ones(wn, np.uint8), wres, wcell, 1.0 / 30.0,
              ("wind", 1.2, -0.3, 0.35), 20, confine="jacobi")


def test_fused_exact_on_fine_sheet_with_several_substeps():
    """2 cm sheet, pace 2 (adaptive CFL on the wet rectangle): several substeps
    per step, so the ping-pong buffers and the odd-count copy back run."""
    res, cell = 160, 0.02
    n = res * res
    H = np.full((res, res), 0.02, np.float32)
    B = np.zeros((res, res), np.float32)
    B[24:48, 24:48] = (-0.008 * np.arange(24, 48, dtype=np.float32))[:, None]
    H[24:48, 24:48] = 0.06
    u = np.full(n, 0.02, np.float32)
    u.reshape(res, res)[60:80, 60:80] = 1.2
    nsub = O.step_sheet(H.ravel().copy(), u.copy(), np.zeros(n, np.float32), B.ravel(), np.ones(n, np.float32),
                        np.ones(n, np.uint8), res, cell, 1.0 / 30.0, ("gravity", 0.02), pace=2)
    assert nsub >= 3
    _run_pair([H.ravel(), u, np.zeros(n, np.float32)], B.ravel(), np.ones(n, np.float32), np.ones(n, np.uint8), res,
              cell, 1.0 / 30.0, ("gravity", 0.02), 4, pace=2)


def test_fused_exact_on_a_multi_strip_multi_chunk_grid():
    """A grid wider than several 235-column strips and taller than several
    row chunks, with a size that divides neither."""
    sc = scenes.large_scene(res=1000, seed=3, puddles=12)
    res, cell = sc["res"], sc["cell"]
    n = res * res
    _run_pair([sc["water"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)], sc["bed"],
              np.full(n, 0.9, np.float32), np.ones(n, np.uint8), res, cell, 1.0 / 120.0, ("gravity", 0.08), 3)


def _fast_vs_fp64(state, bed, nu, gate, res, cell, dt, drive, **opts):
    f = [a.copy() for a in state]
    o = [a.astype(np.float64) for a in state]
    S.step_sheet(*f, bed, nu, gate, res, cell, dt, drive, path="fast", **opts)
    okw = {k: v for k, v in opts.items() if k in ("pace",)}
    if opts.get("confine") == "jacobi":
        okw["confine"] = "jacobi"
    O.step_sheet(*o, bed.astype(np.float64), nu.astype(np.float64), gate, res, cell, dt, drive, precision=64, **okw)
    worst = {}
    for a, b, name in zip(f, o, "huv"):
        scale = float(np.abs(b).max())
        err = float(np.abs(a.astype(np.float64) - b).max())
        worst[name] = err / max(scale, 1e-30)
        assert err <= TOL * scale, f"{name}: max err {err:.3e} > {TOL} x {scale:.3e}"
    return worst


def test_fast_path_within_tolerance_of_fp64_oracle():
    sc, st = _window(1)
    n = sc["res"] ** 2
    # A moving state: run a few exact steps first so velocities are nonzero.
    for _ in range(20):
        S.step_sheet(*st, sc["bed"], np.full(n, 0.9, np.float32), np.ones(n, np.uint8), sc["res"], sc["cell"],
                     1.0 / 120.0, ("gravity", 0.08))
    _fast_vs_fp64(st, sc["bed"], np.full(n, 0.9, np.float32), np.ones(n, np.uint8), sc["res"], sc["cell"],
                  1.0 / 120.0, ("gravity", 0.08))
    sc2, st2 = _window(3, "slurry")
    for _ in range(20):
        S.step_sheet(*st2, sc2["bed"], np.full(n, 0.5, np.float32), np.ones(n, np.uint8), sc2["res"], sc2["cell"],
                     1.0 / 120.0, ("gravity", 0.05))
    _fast_vs_fp64(st2, sc2["bed"], np.full(n, 0.5, np.float32), np.ones(n, np.uint8), sc2["res"], sc2["cell"],
                  1.0 / 120.0, ("gravity", 0.05))


def test_fast_path_conserves_like_the_exact_path_over_many_steps():
    sc, st = _window(4)
    n = sc["res"] ** 2
    nu = np.full(n, 0.9, np.float32)
    gate = np.ones(n, np.uint8)
    ex = [a.copy() for a in st]
    fa = [a.copy() for a in st]
    v0 = float(st[0].astype(np.float64).sum())
    for _ in range(200):
        S.step_sheet(*ex, sc["bed"], nu, gate, sc["res"], sc["cell"], 1.0 / 120.0, ("gravity", 0.08))
        S.step_sheet(*fa, sc["bed"], nu, gate, sc["res"], sc["cell"], 1.0 / 120.0, ("gravity", 0.08), path="fast")
    ve = float(ex[0].astype(np.float64).sum())
    vf = float(fa[0].astype(np.float64).sum())
    # Volume moves only through the dry cut and the max(0) clamp in both.
    assert abs(ve - v0) < 1e-3 * v0 and abs(vf - v0) < 1e-3 * v0
    assert abs(vf - ve) < 1e-4 * v0
    # Fields stay close in the L1 sense after 200 steps (threshold flips aside).
    l1 = float(np.abs(fa[0] - ex[0]).sum()) / float(np.abs(ex[0]).sum())
    assert l1 < 1e-3, l1


def _pair_from(sc, path="fast"):
    n = sc["res"] * sc["res"]
    z = np.zeros(n, np.float32)
    p = S.SheetPair(sc["res"], sc["cell"], path=path)
    p.upload(sc["bed"], sc["water"], z, z, sc["slurry"], z, z, sc["nu_slurry"])
    return p


@pytest.mark.parametrize("path", ["fast", "fused_exact"])
def test_results_do_not_depend_on_the_chunk_height(path):
    """Tiles recompute their halo rows and columns with the same arithmetic,
    so the rows per workgroup leave every value unchanged."""
    sc = scenes.large_scene(res=700, seed=5, puddles=10)
    res = sc["res"]
    outs = []
    try:
        for rows in (0, 24, 97):
            S.set_chunk_rows(res, rows)
            p = _pair_from(sc, path)
            for _ in range(3):
                p.step(1.0 / 120.0)
            outs.append(p.download())
            p.close()
    finally:
        S.set_chunk_rows(res, 0)
    for other in outs[1:]:
        for a, b, name in zip(outs[0], other, ("water", "uw", "vw", "slurry", "us", "vs")):
            _eq(a, b, name)


def test_pair_autotune_restores_the_state_and_pins_the_fastest_height():
    sc = scenes.large_scene(res=600, seed=6, puddles=8)
    res = sc["res"]
    p = _pair_from(sc)
    try:
        for _ in range(2):
            p.step(1.0 / 120.0)
        before, led0 = p.download(), p.ledger()
        rows, ms = p.autotune(1.0 / 120.0, rows=[0, 32, 96], reps=2)
        assert rows in (0, 32, 96) and ms > 0.0
        assert S.chunk_rows(res) == rows
        for a, b, name in zip(before, p.download(), ("water", "uw", "vw", "slurry", "us", "vs")):
            _eq(a, b, f"{name} after autotune")
        led1 = p.ledger()
        assert led1.water.gap_m3 == led0.water.gap_m3 and led1.slurry.residue_m3 == led0.slurry.residue_m3
    finally:
        S.set_chunk_rows(res, 0)
        p.close()
