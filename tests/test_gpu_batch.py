"""GPU parity of the batch handle (include/silt_batch.h): windows of uneven
sides stepped together, one launch per substep index, replayed from a
captured graph, must give each window exactly what stepping it alone gives.

The single-window reference is the library's own step_sheet on the same
path, which the other GPU tests tie to the FP32 oracle (exact path) or to
the FP64 oracle (fast path)."""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S
from silt_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0


def _eq(a, b, what):
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    assert bad.size == 0, f"{what}: {bad.size} cells differ, first {bad[:5]} {a[bad[:5]]} vs {b[bad[:5]]}"


def _windows():
    """Four windows: a dry-shore puddle scene whose wet set splits into
    bodies, two small fully wet ones, and one on a film."""
    ws = scenes.batch_scene(n=3, seed=21, lo=40, hi=150)
    ws += scenes.batch_scene(n=1, seed=5, lo=260, hi=260, film=0.02)
    nus = [0.9, 0.5, 4.0, 0.9]
    vorts = [0.08, 0.05, 0.08, 0.02]
    return ws, nus, vorts


def _batch(ws, nus, vorts, path, graph=True):
    b = S.Batch([w["res"] for w in ws], [w["cell"] for w in ws], path=path)
    b.use_graph(graph)
    for k, w in enumerate(ws):
        n = w["res"] ** 2
        b.upload(k, w["bed"], w["water"], np.zeros(n, np.float32), np.zeros(n, np.float32), nus[k], vorts[k])
    return b


@pytest.mark.parametrize("path,single", [("fused_exact", "exact"), ("fast", "fast")])
def test_batch_equals_each_window_stepped_alone(path, single):
    ws, nus, vorts = _windows()
    b = _batch(ws, nus, vorts, path)
    state = []
    for w in ws:
        n = w["res"] ** 2
        state.append([w["water"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)])
    try:
        for step in range(4):
            b.step(DT)
            for k, w in enumerate(ws):
                n = w["res"] ** 2
                S.step_sheet(*state[k], w["bed"], np.full(n, nus[k], np.float32), np.ones(n, np.uint8), w["res"],
                             w["cell"], DT, ("gravity", vorts[k]), path=single)
                got = b.download(k)
                for a, e, name in zip(got, state[k], "huv"):
                    _eq(a, e, f"window {k} {name} step {step}")
        assert b.graphs_captured() >= 1
    finally:
        b.close()


def test_graph_replay_equals_direct_launches_and_the_ledger_closes():
    ws, nus, vorts = _windows()
    g, d = _batch(ws, nus, vorts, "fast", graph=True), _batch(ws, nus, vorts, "fast", graph=False)
    try:
        for _ in range(8):
            g.step(DT)
            d.step(DT)
        for k in range(len(ws)):
            for a, e, name in zip(g.download(k), d.download(k), "huv"):
                _eq(a, e, f"window {k} {name}")
            led = g.ledger(k)
            assert abs(led.gap_m3) <= 1e-9 * max(led.initial_m3, 1e-12), (k, led.gap_m3, led.initial_m3)
        # Replays reuse graphs: far fewer captures than steps.
        assert 1 <= g.graphs_captured() <= 4
    finally:
        g.close()
        d.close()


def test_batch_fused_exact_window_matches_the_fp32_oracle():
    """One window of the batch against the oracle itself (not only the
    library's single-window path)."""
    ws, nus, vorts = _windows()
    b = _batch(ws, nus, vorts, "fused_exact")
    k = 0
    w = ws[k]
    n = w["res"] ** 2
    st = [w["water"].copy(), np.zeros(n, np.float32), np.zeros(n, np.float32)]
    try:
        for step in range(3):
            b.step(DT)
            O.step_sheet(*st, w["bed"], np.full(n, nus[k], np.float32), np.ones(n, np.uint8), w["res"], w["cell"],
                         DT, ("gravity", vorts[k]))
            for a, e, name in zip(b.download(k), st, "huv"):
                _eq(a, e, f"{name} step {step}")
    finally:
        b.close()


def test_batch_autotune_restores_the_state():
    ws, nus, vorts = _windows()
    b = _batch(ws, nus, vorts, "fast")
    try:
        b.step(DT)
        before = [b.download(k) for k in range(len(ws))]
        rows, ms = b.autotune(DT, rows=[0, 16, 48], reps=2)
        assert rows in (0, 16, 48) and ms > 0.0
        for k in range(len(ws)):
            for a, e, name in zip(b.download(k), before[k], "huv"):
                _eq(a, e, f"window {k} {name} after autotune")
    finally:
        b.close()


def test_xcd_weights_leave_the_batch_results_unchanged():
    """Rows of every launch's tiles scaled per XCD give the same planes bit
    for bit as uniform chunks, through graph replays and the autotuner's
    second knob (which restores the state)."""
    ws, nus, vorts = _windows()
    a, b = _batch(ws, nus, vorts, "fast"), _batch(ws, nus, vorts, "fast")
    try:
        b.set_xcd_weights([0.6, 1.4, 0.8, 1.2, 1.0, 0.7, 1.3, 1.0])
        assert b.xcd_weights_on()
        for _ in range(6):
            a.step(DT)
            b.step(DT)
        for k in range(len(ws)):
            for x, y, name in zip(a.download(k), b.download(k), "huv"):
                _eq(y, x, f"window {k} {name} with XCD weights")
        before = [b.download(k) for k in range(len(ws))]
        b.autotune(DT, rows=[0], reps=2)
        for k in range(len(ws)):
            for x, y, name in zip(b.download(k), before[k], "huv"):
                _eq(x, y, f"window {k} {name} after the XCD knob")
        for _ in range(3):
            a.step(DT)
            b.step(DT)
        for k in range(len(ws)):
            for x, y, name in zip(a.download(k), b.download(k), "huv"):
                _eq(y, x, f"window {k} {name} after tuning")
    finally:
        a.close()
        b.close()


def test_wet_bodies_labelled_on_the_gpu_give_the_host_boxes():
    """Scattered puddles: the reference's host scan plans 4 split boxes,
    more than 8 (so the wet bounding box), 7, and bodies whose grown boxes
    merge into one.  The GPU labels the bodies (no plane read back); every
    window must step exactly as the host-planned single-window path."""
    rng = np.random.default_rng(8)
    ws = []
    shapes = [(120, 5, 0.8, 0.01), (200, 14, 0.6, 0.02), (200, 14, 0.6, 0.01), (90, 3, 2.0, 0.005)]
    for k, (res, blobs, sigma, cut) in enumerate(shapes):
        cell = 0.2
        origin = (-res * cell / 2.0, -res * cell / 2.0)
        bed = scenes.fbm_terrain(res, cell, origin, 40 + k).ravel()
        water = np.zeros(res * res, np.float32)
        for _ in range(blobs):
            cx, cz = rng.uniform(origin[0] + 2, -origin[0] - 2, size=2)
            water += scenes.gaussian(res, cell, origin, np.float32(cx), np.float32(cz), 0.05, sigma).ravel()
        water[water < cut] = 0.0
        ws.append({"res": res, "cell": cell, "bed": bed, "water": water.astype(np.float32)})
    assert [len(O.work_boxes(w["water"], w["res"], 1.0e-6)) for w in ws] == [4, 1, 7, 1]
    nus, vorts = [0.9] * 4, [0.08] * 4
    b = _batch(ws, nus, vorts, "fused_exact")
    state = [[w["water"].copy(), np.zeros(w["res"] ** 2, np.float32), np.zeros(w["res"] ** 2, np.float32)]
             for w in ws]
    try:
        for step in range(3):
            b.step(DT)
            assert b.last_fallbacks() == 0
            for k, w in enumerate(ws):
                n = w["res"] ** 2
                S.step_sheet(*state[k], w["bed"], np.full(n, 0.9, np.float32), np.ones(n, np.uint8), w["res"],
                             w["cell"], DT, ("gravity", 0.08), path="exact")
                for a, e, name in zip(b.download(k), state[k], "huv"):
                    _eq(a, e, f"window {k} {name} step {step}")
    finally:
        b.close()
