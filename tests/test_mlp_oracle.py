"""CPU: the surrogate perceptron's oracle pinned by float64 central finite
differences (its backward pass) and by its bfloat16 rounding rule."""
import numpy as np

import mlp_oracle as MO


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p_in, p_out, width, B = 3, 2, 64, 5
    flat = MO.init_flat(p_in, p_out, width, rng, scale=1.5)
    x = rng.normal(size=(B, p_in))
    y = rng.normal(size=(B, p_out))
    r = MO.forward_backward(flat, x, y, p_in, p_out, width, bf16=False)
    g = r["grads"]
    # probe a spread of real (non-padding) parameters of every layer
    idx = []
    o = 0
    for (K, N), (k, nn) in zip(MO.shapes(p_in, p_out, width),
                               [(p_in, width), (width, width), (width, width),
                                (width, 2 * p_out)]):
        for _ in range(6):
            idx.append(o + rng.integers(k) * N + rng.integers(nn))
        o += K * N
        for _ in range(3):
            idx.append(o + rng.integers(nn))
        o += N
    h = 1e-6
    for i in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        num = (MO.forward_backward(fp, x, y, p_in, p_out, width, bf16=False)["loss"]
               - MO.forward_backward(fm, x, y, p_in, p_out, width, bf16=False)["loss"]) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])
    # padding parameters get exactly zero gradient
    K0 = MO.pad64(p_in)
    W0 = g[:K0 * width].reshape(K0, width)
    assert np.all(W0[p_in:] == 0.0)


def test_loss_closed_form():
    # zero network: mu = 0, lv = 0 -> loss = 0.5 * mean_b sum_a y^2
    p_in, p_out, width = 2, 3, 64
    flat = np.zeros(sum(K * N + N for K, N in MO.shapes(p_in, p_out, width)))
    rng = np.random.default_rng(1)
    x, y = rng.normal(size=(7, p_in)), rng.normal(size=(7, p_out))
    r = MO.forward_backward(flat, x, y, p_in, p_out, width, bf16=False)
    np.testing.assert_allclose(r["loss"], 0.5 * np.mean(np.sum(y * y, axis=1)), rtol=1e-14)


def test_bf16_rounding_rule():
    # nearest even on the 16 dropped bits of the float32 pattern
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -3.14159265, 65504.0])
    got = MO.to_bf16(vals)
    # 1 + 2^-8 is the midpoint of 1 and 1 + 2^-7 (odd last bit): ties go to 1;
    # 65504 rounds up into the next binade
    want = [1.0, 1.0, 1.0 + 2 ** -7, 1.0, -3.140625, 65536.0]
    np.testing.assert_array_equal(got, want)
