"""The FP64 oracle of the learned preconditioner network pinned by independent
checks: a direct loop convolution, central finite differences of the loss for
the hand-written gradients, and bfloat16 rounding against known bit patterns."""
import numpy as np

import lpc_oracle as L


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to the even significand (1)
    assert L.bf16_round(np.float32(1.0 + 2.0 ** -8)) == 1.0
    # 1 + 3*2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert L.bf16_round(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    assert L.bf16_round(np.float32(1.0 + 2.0 ** -7)) == 1.0 + 2.0 ** -7
    assert L.bf16_round(np.float32(-3.0)) == -3.0


def test_conv3_matches_direct_loops():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(6, 6, 6, 3))
    w = rng.normal(size=(27, 3, 2))
    y = L.conv3(x, w)
    z, yy, xx, co = 2, 1, 3, 1
    want = sum(x[z + dz, yy + dy, xx + dx, ci] * w[dx + 3 * dy + 9 * dz, ci, co]
               for dz in range(3) for dy in range(3) for dx in range(3) for ci in range(3))
    assert np.isclose(y[z, yy, xx, co], want, rtol=1e-13)


def test_gradients_match_finite_differences():
    # a reduced tile keeps the finite differences cheap: the same functions run on
    # any cube, so the check covers the 16^3 network's code path
    rng = np.random.default_rng(1)
    d = 5
    X = rng.normal(size=(d + 2, d + 2, d + 2, L.CIN0))
    w = L.init_weights(2)
    gO = rng.normal(size=(d, d, d))

    def loss(wv):
        o, _ = L.forward(X, wv)
        return float((o * gO).sum())

    _, cache = L.forward(X, w)
    g = L.backward(cache, gO, w)
    # probe weights of every layer, including the head
    idx = [3, 700, 1729, 5000, 15600, 20000, 29380, 29391]
    for i in idx:
        e = np.zeros_like(w)
        h = 1e-6 * max(1.0, abs(w[i]))
        e[i] = h
        fd = (loss(w + e) - loss(w - e)) / (2 * h)
        assert np.isclose(g[i], fd, rtol=2e-5, atol=1e-7), (i, g[i], fd)


def test_receptive_field_is_three_cells_past_the_halo():
    # three 3^3 layers: an output at depth z sees the input halo only if z < 3; a
    # slip in the tap offsets or the padding would widen or shift that reach
    rng = np.random.default_rng(5)
    X = rng.normal(size=(L.H, L.H, L.H, L.CIN0))
    w = L.init_weights(4)
    o, _ = L.forward(X, w)
    X2 = X.copy()
    X2[0, :, :, :] += 10.0      # change the halo plane z = -1 only
    o2, _ = L.forward(X2, w)
    assert np.array_equal(o[3:], o2[3:])
    assert not np.allclose(o[2], o2[2])
