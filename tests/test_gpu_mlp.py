"""GPU: the surrogate perceptron's BF16 MFMA forward/backward against the
float64 oracle (with the device's bfloat16 rounding points), and training.

Tolerances: the device accumulates in float32 (the oracle in float64), and a
float32-vs-float64 difference can move a bfloat16 rounding by one unit
(2^-8 relative) for an occasional element; the comparisons are relative to
each array's scale: outputs and loss 1e-3, gradients 2e-2 per element of the
largest magnitude and 5e-3 in norm; outputs 1e-3 of the scale for 99.9 % of
the elements and 5e-3 for all (a flipped rounding of one hidden unit).
"""
import numpy as np
import pytest

import mlp_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("p_in,p_out,B", [(3, 3, 128), (13, 13, 128), (40, 7, 128),
                                          (13, 13, 1024)])
def test_forward_backward_matches_oracle(gpu_lib, p_in, p_out, B):
    """Batch 128 and 1024: the slice kernel's forward and backward passes
    over 8 and 64 slices of 16 rows, and the weight-gradient kernel's
    interleaved pairs of K steps (one pair for two waves, four pairs per
    wave) added in wave order."""
    rng = np.random.default_rng(p_in + B)
    width = 256
    m = gpu_lib.Mlp(p_in, p_out, width=width, batch=B, seed=3)
    try:
        flat = MO.init_flat(p_in, p_out, width, rng).astype(np.float32)
        m.set_params(flat)
        x = rng.uniform(-1, 1, (B, p_in)).astype(np.float32)
        y = rng.normal(size=(B, p_out)).astype(np.float32)
        got = m.forward_backward(x, y)
    finally:
        m.close()
    ref = MO.forward_backward(flat.astype(np.float64), x.astype(np.float64),
                              y.astype(np.float64), p_in, p_out, width, bf16=True)
    N3 = MO.pad64(2 * p_out)
    out, rout = got["out"].astype(np.float64), ref["out"]
    assert out.shape == (B, N3)
    err, scale = np.abs(out - rout), np.max(np.abs(rout))
    # a flipped bfloat16 rounding of one hidden unit moves an output by ~1e-3
    # of the scale; at batch 1024 a few elements see one
    assert np.quantile(err, 0.999) <= 1e-3 * scale and err.max() <= 5e-3 * scale, err.max()
    assert abs(got["loss"] - ref["loss"]) <= 1e-3 * abs(ref["loss"])
    g, rg = got["grads"].astype(np.float64), ref["grads"]
    o = 0
    for K, N in MO.shapes(p_in, p_out, width):
        for size in (K * N, N):
            a, b = g[o:o + size], rg[o:o + size]
            scale = np.max(np.abs(b)) + 1e-30
            assert np.max(np.abs(a - b)) <= 2e-2 * scale, (o, np.max(np.abs(a - b)), scale)
            assert np.linalg.norm(a - b) <= 5e-3 * np.linalg.norm(b) + 1e-30
            o += size
    # padding rows of the first layer and padding output columns stay at zero gradient
    K0 = MO.pad64(p_in)
    assert np.all(g[:K0 * width].reshape(K0, width)[p_in:] == 0.0)


def test_adam_keeps_bfloat16_copies_of_the_weights(gpu_lib):
    """The Adam kernel writes each updated weight to its bfloat16 copies (row-
    major and transposed, which the products read): after training steps the
    device's outputs equal the oracle's forward pass on the master weights."""
    rng = np.random.default_rng(8)
    p_in, p_out, B = 5, 4, 1024
    m = gpu_lib.Mlp(p_in, p_out, width=256, batch=B, seed=2)
    try:
        xs = rng.uniform(-1, 1, (4096, p_in))
        ys = rng.normal(size=(4096, p_out))
        m.train(xs, ys, steps=5, lr=1e-2, seed=4)
        flat = m.get_params()
        x = rng.uniform(-1, 1, (B, p_in)).astype(np.float32)
        y = rng.normal(size=(B, p_out)).astype(np.float32)
        got = m.forward_backward(x, y)
    finally:
        m.close()
    ref = MO.forward_backward(flat.astype(np.float64), x.astype(np.float64),
                              y.astype(np.float64), p_in, p_out, 256, bf16=True)
    err, scale = np.abs(got["out"] - ref["out"]), np.max(np.abs(ref["out"]))
    assert np.quantile(err, 0.999) <= 1e-3 * scale and err.max() <= 5e-3 * scale, err.max()


def test_training_fits_a_smooth_map(gpu_lib):
    """Heteroscedastic regression y = f(x) + s(x) eps: the trained mean is
    within a small fraction of the noise of f on fresh inputs, and the
    predicted log-variance tracks log s^2."""
    rng = np.random.default_rng(5)
    p_in, p_out, n = 3, 2, 1 << 16
    A = rng.normal(size=(p_in, p_out))

    def f(x):
        return np.tanh(x @ A) + 0.3 * x[:, :p_out] ** 2

    def sd(x):
        return 0.2 + 0.1 * np.abs(x[:, :1])

    x = rng.uniform(-1, 1, (n, p_in))
    y = f(x) + sd(x) * rng.normal(size=(n, p_out))
    m = gpu_lib.Mlp(p_in, p_out, width=256, batch=1024, seed=11)
    try:
        first = m.forward_backward(x[:1024], y[:1024])["loss"]
        last = m.train(x, y, steps=800, lr=2e-3, seed=3)
        xt = rng.uniform(-0.9, 0.9, (4096, p_in))
        pred = m.predict(xt).astype(np.float64)
    finally:
        m.close()
    assert last < first
    err = pred[:, :p_out] - f(xt)
    assert np.sqrt(np.mean(err ** 2)) < 0.05, np.sqrt(np.mean(err ** 2))
    lv = pred[:, p_out:]
    assert np.corrcoef(lv[:, 0], np.log(sd(xt)[:, 0] ** 2))[0, 1] > 0.8


@pytest.mark.parametrize("width,p_in,p_out", [(64, 2, 1), (192, 7, 40)])
def test_narrow_and_wide_output_shapes(gpu_lib, width, p_in, p_out):
    """Other compiled shapes of the slice kernel: width 64 (one column tile
    per wave) with one output, width 192 with 2 p_out = 80 outputs (N_3 = 128,
    two column tiles per wave in the output product), batch 64 (four
    slices; the weight-gradient kernel's waves 1-3 take no K step)."""
    rng = np.random.default_rng(width + p_out)
    B = 64
    m = gpu_lib.Mlp(p_in, p_out, width=width, batch=B, seed=3)
    try:
        flat = MO.init_flat(p_in, p_out, width, rng).astype(np.float32)
        m.set_params(flat)
        x = rng.uniform(-1, 1, (B, p_in)).astype(np.float32)
        y = rng.normal(size=(B, p_out)).astype(np.float32)
        got = m.forward_backward(x, y)
    finally:
        m.close()
    ref = MO.forward_backward(flat.astype(np.float64), x.astype(np.float64),
                              y.astype(np.float64), p_in, p_out, width, bf16=True)
    err, scale = np.abs(got["out"] - ref["out"]), np.max(np.abs(ref["out"]))
    assert err.max() <= 5e-3 * scale, err.max()
    assert abs(got["loss"] - ref["loss"]) <= 1e-3 * abs(ref["loss"])
    g, rg = got["grads"].astype(np.float64), ref["grads"]
    assert np.linalg.norm(g - rg) <= 5e-3 * np.linalg.norm(rg)


def test_width_above_the_fused_bound_is_refused(gpu_lib):
    from cusna_amd.abi import CusnaError

    with pytest.raises(CusnaError, match="SHAPE|status 2"):
        gpu_lib.Mlp(3, 3, width=320, batch=128, seed=1)
