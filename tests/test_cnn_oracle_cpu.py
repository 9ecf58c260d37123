"""Pins the CPU oracle of the proposal network (oracle/cnn_oracle.py):
the forward convolution against a direct per-pixel sum, the data gradient as
the exact adjoint of the convolution, and every analytic gradient of the
whole network (weights, biases, through ReLU, pooling and loss) against
central finite differences."""
import numpy as np

import cnn_oracle as co


def _rand(rng, *shape, scale=1.0):
    return rng.standard_normal(shape) * scale


def test_forward_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = _rand(rng, 1, 5, 6, 32)
    w = _rand(rng, 9, 32, 32, scale=0.1)
    b = _rand(rng, 32)
    for dil in (1, 2):
        z = co.conv(x, w, b, dil)
        for (y, xx) in [(0, 0), (2, 3), (4, 5), (1, 0)]:
            want = b.copy()
            for tap in range(9):
                yy, xs = y + (tap // 3 - 1) * dil, xx + (tap % 3 - 1) * dil
                if 0 <= yy < 5 and 0 <= xs < 6:
                    want += x[0, yy, xs] @ w[tap]
            assert np.allclose(z[0, y, xx], want, rtol=1e-12, atol=1e-12)


def test_data_grad_is_the_adjoint():
    rng = np.random.default_rng(2)
    x = _rand(rng, 2, 7, 9, 32)
    dz = _rand(rng, 2, 7, 9, 32)
    w = _rand(rng, 9, 32, 32)
    for dil in (1, 3):
        lhs = float((co.conv(x, w, None, dil) * dz).sum())
        rhs = float((x * co.conv_data_grad(dz, w, dil)).sum())
        assert abs(lhs - rhs) <= 1e-9 * max(1.0, abs(lhs))


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    dils = (1, 2, 1)
    layers = [(_rand(rng, 9, 32, 32, scale=0.15), _rand(rng, 32, scale=0.1), d) for d in dils]
    x = _rand(rng, 2, 6, 7, 32)
    windows = [(0, 1, 1, 4, 5), (1, 0, 2, 7, 6), (1, 2, 0, 5, 3)]
    label = _rand(rng, 3, 2)
    weight = np.array([1.0, 0.5, 2.0])
    inv = 1.0 / 3.0
    net = co.Net(layers)
    l0, _, grads = net.loss_and_grads(x, windows, label, weight, inv)

    def loss_with(k, which, idx, delta):
        w, b, d = layers[k]
        w2, b2 = w.copy(), b.copy()
        (w2 if which == "w" else b2)[idx] += delta
        ls = list(layers)
        ls[k] = (w2, b2, d)
        return co.Net(ls).loss_and_grads(x, windows, label, weight, inv)[0]

    h = 1e-6
    for k in range(3):
        for _ in range(4):
            idx = tuple(int(rng.integers(0, s)) for s in (9, 32, 32))
            num = (loss_with(k, "w", idx, h) - loss_with(k, "w", idx, -h)) / (2 * h)
            assert abs(num - grads[k][0][idx]) <= 1e-6 * max(1.0, abs(num)), (k, idx)
        idx = (int(rng.integers(0, 32)),)
        num = (loss_with(k, "b", idx, h) - loss_with(k, "b", idx, -h)) / (2 * h)
        assert abs(num - grads[k][1][idx]) <= 1e-6 * max(1.0, abs(num)), (k, idx)
    assert l0 > 0


def test_bf16_rounding_ties_to_even():
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0], dtype=np.float32)
    got = co.bf16_round(vals)
    assert list(got) == [1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, 0.0]
