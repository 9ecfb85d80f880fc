"""The CPU baseline build of the oracle (liboracle_par.so: AVX2 width, row
stripes of at most 4 threads, water || slurry) computes the serial oracle's
step bit for bit: the stripes split only passes whose rows are independent,
and the z transport rebuilds each row's sums in the serial order."""
import numpy as np

import pyoracle as O
from silt_amd import scenes


def _scene(kind):
    if kind == "film":
        sc = scenes.large_scene(res=300, seed=5, puddles=6)
        return sc["res"], sc["bed"], sc["water"].copy(), sc["slurry"].copy()
    res = 256
    sc = scenes.large_scene(res=res, seed=8, puddles=3)
    water = np.zeros((res, res), np.float32)
    for cx, cz in ((-10.0, -12.0), (14.0, 9.0), (-12.0, 15.0)):
        water += scenes.gaussian(res, 0.2, sc["origin"], cx, cz, 0.10, 3.0)
    water[water < 1e-6] = 0.0
    slurry = np.zeros((res, res), np.float32)
    slurry[60:200, 40:120] = 0.02
    return res, sc["bed"], water.ravel(), slurry.ravel()


def test_striped_pair_step_equals_the_serial_oracle():
    for kind in ("film", "puddles"):
        res, bed, hw, hs = _scene(kind)
        n = res * res
        rng = np.random.default_rng(3)
        uw = (0.2 * rng.standard_normal(n) * (hw > 0)).astype(np.float32)
        vw = (0.2 * rng.standard_normal(n) * (hw > 0)).astype(np.float32)
        zs = np.zeros(n, np.float32)
        nu_w = np.full(n, 0.9, np.float32)
        nu_s = np.full(n, 0.5, np.float32)
        gate = np.ones(n, np.uint8)
        ser = [hw.copy(), uw.copy(), vw.copy(), hs.copy(), zs.copy(), zs.copy()]
        par = [a.copy() for a in ser]
        for step in range(4):
            O.step_sheet(*ser[:3], bed, nu_w, gate, res, 0.2, 1.0 / 60.0, ("gravity", 0.08))
            O.step_sheet(*ser[3:], bed, nu_s, gate, res, 0.2, 1.0 / 60.0, ("gravity", 0.05))
            O.step_pair_par(par[:3], par[3:], bed, nu_w, nu_s, gate, res, 0.2, 1.0 / 60.0)
            for a, b, name in zip(par, ser, ("h", "u", "v", "hs", "us", "vs")):
                bad = np.flatnonzero(a != b)
                assert bad.size == 0, f"{kind} step {step} {name}: {bad.size} cells differ"
        assert np.abs(ser[0] - hw).max() > 1e-4  # water moved
