"""GPU work-box planning (plan_device.hip) against the oracle's
sheet_work_boxes (src/fluid_core.rs:1325-1447): the same set of boxes for
every plane of a batch planned together.  The list order of 2 to 8 boxes may
differ (boxes are at least 2 cells apart and stepped independently), so the
lists are compared as sets."""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S

pytestmark = pytest.mark.gpu

EPS = 1.0e-6


def _blobs(rng, res, count, rmax, speckle=0.0, holes=0.0):
    y, x = np.mgrid[:res, :res]
    h = np.zeros((res, res), np.float32)
    for _ in range(count):
        cx, cy = rng.uniform(0, res, 2)
        r = rng.uniform(0.5, rmax)
        h[(x - cx) ** 2 + (y - cy) ** 2 < r * r] = 0.1
    if speckle:
        h[rng.random((res, res)) < speckle] = 0.1
    if holes:
        h[rng.random((res, res)) < holes] = 0.0
    return h.ravel()


def _check(planes):
    got = S.plan_boxes(planes, 0.2, EPS, cap=16)
    for k, (h, g) in enumerate(zip(planes, got)):
        res = int(round(np.sqrt(h.size)))
        ref = O.work_boxes(h, res, EPS) if (h > EPS).any() else []
        assert g is not None, f"plane {k}: planned on the host"
        assert sorted(g) == sorted(ref), f"plane {k} (res {res}): {g} vs {ref}"


def test_random_wet_sets_give_the_reference_boxes():
    rng = np.random.default_rng(5)
    planes = []
    for _ in range(90):
        res = int(rng.integers(12, 170))
        planes.append(_blobs(rng, res, int(rng.integers(1, 30)), max(res / 6, 1.0),
                             speckle=0.02 if rng.random() < 0.3 else 0.0,
                             holes=0.1 if rng.random() < 0.3 else 0.0))
    # More than 64 sheets: two launches of the batched kernels.
    assert len(planes) > 64
    _check(planes)


def test_large_sheets_with_holes_and_long_runs():
    """Tile grids taller than one join block (32 tile rows): the two-phase
    joins; a mostly wet sheet with dry holes; a spiral whose body winds
    through every block."""
    rng = np.random.default_rng(6)
    planes = [_blobs(rng, 700, 12, 90.0), _blobs(rng, 520, 3, 300.0, holes=0.05)]
    res = 300
    sp = np.zeros((res, res), np.float32)
    for k in range(0, res // 2, 6):  # concentric rings joined by a corridor
        sp[k, k:res - k] = sp[res - 1 - k, k:res - k] = 0.1
        sp[k:res - k, k] = sp[k:res - k, res - 1 - k] = 0.1
    sp[:, res // 2] = 0.1
    planes.append(sp.ravel())
    _check(planes)


def test_separate_puddles_keep_their_own_boxes():
    """2 to 8 merged boxes (the list the kernels step), cluster counts past
    the first copy's 64 records, a dry plane and a solid rectangle."""
    rng = np.random.default_rng(7)
    planes = []
    res = 240
    for n in (2, 3, 5, 8, 9):
        h = np.zeros((res, res), np.float32)
        cs = rng.choice(np.arange(20, res - 20, 40), size=(n, 2))
        for cx, cy in cs:
            h[cy - 4:cy + 5, cx - 4:cx + 5] = 0.05
        planes.append(h.ravel())
    h = np.zeros((256, 256), np.float32)
    h[::8, ::8] = 0.01  # 1024 isolated cells: 1024 clusters, one merged box
    planes.append(h.ravel())
    planes.append(np.zeros(64 * 64, np.float32))
    sol = np.zeros((100, 100), np.float32)
    sol[10:40, 20:90] = 0.2
    planes.append(sol.ravel())
    counts = [len(O.work_boxes(p, int(round(np.sqrt(p.size))), EPS)) for p in planes[:5]]
    assert sum(2 <= c <= 8 for c in counts) >= 3, counts
    _check(planes)


def test_more_clusters_than_read_back_plan_on_the_host():
    h = np.zeros((528, 528), np.float32)
    h[::8, ::8] = 0.01  # 66^2 = 4356 clusters > 4096
    assert S.plan_boxes([h.ravel()], 0.2, EPS) == [None]
