"""GPU parity of the tire sub-grid network's MFMA kernels (include/silt_mlp.h)
against the CPU oracle (oracle/mlp.c, FP32 with the same BF16 rounding
points).

Tolerances: the GPU accumulates each product in a different order from the
oracle (MFMA blocks of 16 along K), so an FP32 sum differs in its last bits,
and a BF16-rounded activation or gradient can land one BF16 step (2^-8
relative) apart.  Outputs must agree within 1e-2 of their scale, gradient
blocks within 2e-2 relative in norm."""
import numpy as np
import pytest

import pyoracle as O
import silt_amd as S

pytestmark = pytest.mark.gpu

BLOCKS = (("w1", 0, 196608), ("b1", 196608, 196736), ("w2", 196736, 213120), ("b2", 213120, 213248),
          ("w3", 213248, 214272), ("b3", 214272, 214280))


def _net(seed):
    m = S.Mlp(seed)
    p = m.params()
    # Non-zero biases so their paths are exercised.
    rng = np.random.default_rng(seed)
    for _, a, b in BLOCKS[1::2]:
        p[a:b] = rng.normal(0, 0.1, b - a)
    m.set_params(p)
    return m, p


@pytest.mark.parametrize("n", [1, 100, 256])
def test_forward_matches_the_oracle(n):
    m, p = _net(3)
    try:
        rng = np.random.default_rng(n)
        x = rng.normal(0, 1, (n, 1536)).astype(np.float32)
        y = m.forward(x)
        ref = O.mlp_forward(p, x, bf16=True, precision=32)
        scale = max(np.abs(ref).max(), 1e-3)
        assert np.abs(y - ref).max() <= 1e-2 * scale, np.abs(y - ref).max() / scale
    finally:
        m.close()


@pytest.mark.parametrize("n", [37, 128])
def test_gradients_match_the_oracle(n):
    m, p = _net(5)
    try:
        rng = np.random.default_rng(100 + n)
        x = rng.normal(0, 1, (n, 1536)).astype(np.float32)
        t = rng.normal(0, 0.5, (n, 8)).astype(np.float32)
        sse, g = m.gradients(x, t)
        sse_ref, g_ref = O.mlp_gradients(p, x, t, bf16=True, precision=32)
        assert abs(sse - sse_ref) <= 1e-2 * sse_ref
        for name, a, b in BLOCKS:
            rel = np.linalg.norm(g[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-12)
            assert rel < 2e-2, (name, rel)
        # The weights are untouched by a gradient evaluation.
        np.testing.assert_array_equal(m.params(), p)
    finally:
        m.close()


def test_training_fits_a_smooth_target():
    """Adam on the GPU drives the loss of a fixed regression down by 20x."""
    m, _ = _net(7)
    try:
        rng = np.random.default_rng(9)
        x = rng.normal(0, 1, (512, 1536)).astype(np.float32)
        proj = rng.normal(0, 1.0 / np.sqrt(1536), (1536, 8)).astype(np.float32)
        t = np.tanh(x @ proj).astype(np.float32)
        first = m.train_batch(x, t, lr=1e-3)
        last = first
        for _ in range(300):
            last = m.train_batch(x, t, lr=1e-3)
        assert last < first / 20.0, (first, last)
    finally:
        m.close()


def test_one_launch_inference_matches_the_training_forward():
    """The tire steps' one-launch inference (all three layers, K summed in a
    permuted order) against the training forward and the oracle."""
    rng = np.random.default_rng(31)
    m = S.Mlp(17)
    try:
        for n in (1, 15, 16, 17, 400):
            x = rng.standard_normal((n, S.MLP_IN)).astype(np.float32)
            yf, yi = m.forward(x), m.infer(x)
            scale = max(float(np.abs(yf).max()), 1e-3)
            assert np.abs(yi - yf).max() <= 2e-2 * scale, n
        yo = O.mlp_forward(m.params(), x, bf16=True, precision=32)
        assert np.abs(yi - yo).max() <= 3e-2 * max(float(np.abs(yo).max()), 1e-3)
    finally:
        m.close()
