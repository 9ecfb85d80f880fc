"""The tire sub-grid network inside the FluidCore (include/silt_core.h tire
models, include/silt_mlp.h trainer).

* The training targets are a least-norm flux field whose divergence is the
  nested-minus-coarse depth difference: applying them to the coarse-only
  depths gives the nested ones (minus the region's mean difference), checked
  against a dense least-squares solve in numpy.
* Applying any fluxes moves volume between cells and never makes a depth
  negative: the volume balance closes by construction.
* With a zero tolerance the judge rejects every surrogate step, so the
  surrogate model must reproduce the nested model exactly (snapshot, spray
  and restore are exact).
* A network trained on the GPU on recorded nested steps predicts held-out
  steps' fluxes better than the coarse-only step (predicting zero) does."""
import numpy as np
import pytest

import silt_amd as S
from silt_amd import drive as D
from silt_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0
N = 200 * 200


def _lstsq_fluxes(d, nx, nz):
    """Least-norm outflows (E, W, N, S) with divergence d over an nx x nz grid."""
    faces = []
    for j in range(nz):
        for i in range(nx - 1):
            faces.append((j * nx + i, j * nx + i + 1))
    for j in range(nz - 1):
        for i in range(nx):
            faces.append((j * nx + i, (j + 1) * nx + i))
    D_ = np.zeros((nx * nz, len(faces)))
    for f, (a, b) in enumerate(faces):
        D_[a, f] = -1.0  # flux a -> b lowers a
        D_[b, f] = 1.0
    F = np.linalg.lstsq(D_, d - d.mean(), rcond=None)[0]
    out = np.zeros((nx * nz, 4))
    for f, (a, b) in enumerate(faces):
        if b == a + 1:
            out[a, 0], out[b, 1] = F[f], -F[f]
        else:
            out[a, 2], out[b, 3] = F[f], -F[f]
    return out


def test_targets_are_the_least_norm_fluxes_of_the_difference():
    rng = np.random.default_rng(4)
    i0, j0, nx, nz = 90, 101, 20, 17
    hc_w = rng.uniform(0.0, 0.1, N).astype(np.float32)
    hc_s = rng.uniform(0.0, 0.02, N).astype(np.float32)
    hn_w, hn_s = hc_w.copy(), hc_s.copy()
    for h, amp in ((hn_w, 3e-3), (hn_s, 5e-4)):
        for j in range(nz):
            for i in range(nx):
                h[(j0 + j) * 200 + i0 + i] += np.float32(amp * np.sin(0.7 * i) * np.cos(0.4 * j + 0.3))
    t = S.surrogate_targets(hn_w, hc_w, hn_s, hc_s, (i0, j0, nx, nz))
    for layer, (hn, hc) in enumerate(((hn_w, hc_w), (hn_s, hc_s))):
        d = np.array([float(hn[(j0 + q // nx) * 200 + i0 + q % nx]) - float(hc[(j0 + q // nx) * 200 + i0 + q % nx])
                      for q in range(nx * nz)])
        ref = _lstsq_fluxes(d, nx, nz) * 1000.0
        got = t[:, 4 * layer:4 * layer + 4].astype(np.float64)
        assert np.abs(got - ref).max() <= 1e-4 * np.abs(ref).max() + 1e-7, layer
        # The divergence of the targets is the difference (minus its mean).
        div = -got.sum(axis=1) / 1000.0
        assert np.abs(div - (d - d.mean())).max() < 1e-8


def test_applied_fluxes_conserve_volume_and_keep_depths_non_negative():
    rng = np.random.default_rng(5)
    i0, j0, nx, nz = 0, 180, 20, 20
    w = rng.uniform(0.0, 0.05, N).astype(np.float32)
    s = rng.uniform(0.0, 0.01, N).astype(np.float32)
    w[(j0 + 3) * 200 + 5] = 0.0  # a dry donor
    y = rng.normal(0.0, 20.0, (nx * nz, 8)).astype(np.float32)  # large: the limiter acts
    w2, s2 = S.surrogate_apply(w, s, y, (i0, j0, nx, nz))
    for a, b in ((w, w2), (s, s2)):
        assert abs(a.astype(np.float64).sum() - b.astype(np.float64).sum()) <= 1e-6 * a.astype(np.float64).sum()
        assert b.min() >= 0.0
        mask = np.ones((200, 200), bool)
        mask[j0:j0 + nz, i0:i0 + nx] = False
        assert np.array_equal(a.reshape(200, 200)[mask], b.reshape(200, 200)[mask])
    assert not np.array_equal(w, w2)


def _core(seed, model=None, params=None, judge_every=0, tol=1e-3):
    c = D.nested_scene_core(S, scenes, seed=seed, path="fast")
    if params is not None:
        c.set_surrogate(params, judge_every=judge_every, tol=tol)
    if model:
        c.set_tire_model(model)
    return c


def test_surrogate_with_zero_tolerance_is_the_nested_step():
    m = S.Mlp(3)
    params = m.params()
    m.close()
    a = _core(21)
    b = _core(21, "surrogate", params, judge_every=4, tol=0.0)
    da, db = D.Drive(seed=22), D.Drive(seed=22)
    try:
        for k in range(16):
            D.drive_step(a, da, DT)
            D.drive_step(b, db, DT)
            pa, pb = a.planes(), b.planes()
            for name in ("water", "slurry", "uw", "vw", "us", "vs"):
                assert np.array_equal(pa[name], pb[name]), (name, k)
        st = b.surrogate_stats()
        assert st["judged"] == 4 and st["fallbacks"] == 4 and st["last_judge_err_m"] > 0.0
        la, lb = a.ledger(), b.ledger()
        assert la == lb
    finally:
        a.close()
        b.close()


def test_surrogate_steps_close_the_ledger():
    m = S.Mlp(5)
    params = m.params()
    m.close()
    c = _core(23, "surrogate", params, judge_every=0)
    dv = D.Drive(seed=24)
    try:
        first = c.ledger()["total_m3"]
        for _ in range(30):
            D.drive_step(c, dv, DT)
        st = c.surrogate_stats()
        assert st["last_mode"] == 2 and st["last_tire_ms"] > 0.0
        # The surrogate's fluxes are conservative; the rest of the step is
        # the coarse sheet, which conserves like the nested one.
        assert abs(c.ledger()["total_m3"] - first) <= 1e-5 * first
    finally:
        c.close()


def test_trained_network_beats_the_coarse_only_step_on_held_out_steps():
    """Ten drives of 400 steps (4000 recorded steps, 1.6 M cell samples): the
    network trains on eight and must predict the other two drives' fluxes
    with a mean squared error under 0.85 of predicting zero (the coarse-only
    step)."""
    tr = S.Trainer(max_steps=4000, seed=2)
    try:
        for ep in range(10):
            core = _core(300 + ep)
            dv = D.Drive(seed=400 + ep)
            for _ in range(400):
                D.drive_step(core, dv, DT, step_fn=lambda c, d: tr.record(c, d))
            core.close()
        n = tr.steps()
        ntr = 3200
        assert n == 4000 and tr.samples() > 1500000
        for lr in (1e-3,) * 6 + (3e-4, 1e-4):
            tr.fit(1000, batch=1024, lr=lr, train_steps=ntr)
        mse_net, mse_zero = tr.evaluate(ntr, n - ntr)
        assert mse_net < 0.85 * mse_zero, (mse_net, mse_zero)
    finally:
        tr.close()
