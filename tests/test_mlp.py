"""The learned-guess network: the numpy oracle pinned by finite differences
and closed forms (CPU), then the HIP forward/backward kernels against the
oracle (GPU)."""
import numpy as np
import pytest

import oracle.mlp as om


def test_bf16_rounding_closed_forms():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 0.0, 65504.0], np.float32)
    r = om.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even keeps 1;
    # 1 + 3*2^-8 is a tie between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    assert r.tolist() == [1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 0.0, 65536.0]
    rng = np.random.default_rng(0)
    v = rng.standard_normal(10000).astype(np.float32)
    rel = np.abs(om.bf16(v) - v) / np.abs(v)
    assert rel.max() <= 2.0 ** -8 + 1e-12  # half an ulp of an 8-bit significand


def _small(seed=0):
    rng = np.random.default_rng(seed)
    shapes = om.shapes_of(5, 7, 2, 3)
    n = sum(a * b + c[0] for (a, b), c in shapes)
    p = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((11, 5))
    dy = rng.standard_normal((11, 3))
    return shapes, p, x, dy


def test_oracle_backward_matches_finite_differences():
    shapes, p, x, dy = _small()
    g, dx = om.backward(p, shapes, x, dy, bf=False)

    def f(pp, xx):
        return float(np.sum(dy * om.forward(pp, shapes, xx, bf=False)[-1]))

    h = 1e-6
    for i in range(0, p.size, 7):
        e = np.zeros_like(p)
        e[i] = h
        fd = (f(p + e, x) - f(p - e, x)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i
    for (r, c) in [(0, 0), (3, 2), (10, 4)]:
        e = np.zeros_like(x)
        e[r, c] = h
        fd = (f(p, x + e) - f(p, x - e)) / (2 * h)
        assert abs(fd - dx[r, c]) <= 1e-6 * max(1.0, abs(fd))


def test_oracle_forward_closed_form():
    # one hidden layer, identity-like weights: y = W1 tanh(W0 x + b0) + b1
    shapes = om.shapes_of(2, 2, 1, 1)
    W0 = np.array([[1.0, 0.0], [0.0, 2.0]])
    b0 = np.array([0.5, -0.25])
    W1 = np.array([[3.0, -1.0]])
    b1 = np.array([0.125])
    p = np.concatenate([W0.ravel(), b0, W1.ravel(), b1])
    x = np.array([[0.3, -0.7]])
    y = om.forward(p, shapes, x, bf=False)[-1]
    want = 3 * np.tanh(0.3 + 0.5) - np.tanh(-1.4 - 0.25) + 0.125
    assert abs(y[0, 0] - want) < 1e-14


def test_guess_head_gradient_and_warm_start():
    rng = np.random.default_rng(3)
    n = 9
    xt1 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xt2 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xtrue = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, n)
    # o = 0 reproduces the warm start; o = (0, 0, 1, 0) the linear extrapolation
    assert np.allclose(om.guess_head(np.zeros((n, 4)), xt1, xt2), xt1)
    assert np.allclose(om.guess_head(np.tile([0, 0, 1.0, 0], (n, 1)), xt1, xt2), 2 * xt1 - xt2)
    o = rng.standard_normal((n, 4)) * 0.3
    loss, do = om.guess_loss(o, xt1, xt2, xtrue, w)
    h = 1e-7
    for r in range(n):
        for c in range(4):
            e = np.zeros_like(o)
            e[r, c] = h
            fd = (om.guess_loss(o + e, xt1, xt2, xtrue, w)[0] - om.guess_loss(o - e, xt1, xt2, xtrue, w)[0]) / (2 * h)
            assert abs(fd - do[r, c]) <= 1e-6 * max(1.0, abs(fd))


def test_phase_correction_is_the_reference_warm_start():
    # exp(-j k (d - d_prev) . e): a pure phase, identity for d = d_prev
    rng = np.random.default_rng(4)
    e = rng.standard_normal((6, 3))
    x = rng.standard_normal(6) + 1j * rng.standard_normal(6)
    d = np.array([-1.0, 0, 0])
    assert np.allclose(om.phase_corrected(x, e, 2 * np.pi, d, d), x)
    d2 = np.array([-np.cos(0.1), -np.sin(0.1), 0])
    y = om.phase_corrected(x, e, 2 * np.pi, d2, d)
    assert np.allclose(np.abs(y), np.abs(x))
    assert np.allclose(y / x, np.exp(-1j * 2 * np.pi * e @ (d2 - d)))


# ------------------------------------------------------------------ GPU


def _gpu_net(seed=5):
    import lywell_amd as lw
    if lw.device_count() < 1:
        pytest.skip("no HIP device")
    net = lw.Mlp(16, 256, 3, 4, seed=seed)
    rng = np.random.default_rng(seed)
    p = net.get_params()
    # non-zero output layer so that every layer carries gradient
    p = p + rng.standard_normal(p.size).astype(np.float32) * 0.02
    net.set_params(p)
    return net, p, om.shapes_of(16, 256, 3, 4)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_gpu_forward_matches_oracle():
    net, p, shapes = _gpu_net()
    rng = np.random.default_rng(6)
    x = rng.standard_normal((1000, 16)).astype(np.float32)
    y = net.forward(x)
    y_o = om.forward(p, shapes, x, bf=True)[-1]
    print("forward rel err", rel(y, y_o))
    assert rel(y, y_o) < 2e-3
    # the bf16 rounding is what the oracle models: without it the error is larger
    assert rel(y, om.forward(p, shapes, x, bf=False)[-1]) > rel(y, y_o)


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [64, 5000])
def test_gpu_backward_matches_oracle(rows):
    net, p, shapes = _gpu_net()
    rng = np.random.default_rng(7)
    x = rng.standard_normal((rows, 16)).astype(np.float32)
    dy = rng.standard_normal((rows, 4)).astype(np.float32)
    g, dx = net.backward(x, dy)
    g_o, dx_o = om.backward(p, shapes, x, dy, bf=True)
    layers = om.unpack(g.astype(np.float64), shapes)
    layers_o = om.unpack(g_o, shapes)
    for l, ((gw, gb), (ow, ob)) in enumerate(zip(layers, layers_o)):
        print(f"layer {l}: dW rel err {rel(gw, ow):.2e}, db rel err {rel(gb, ob):.2e}")
        assert rel(gw, ow) < 5e-3, l
        assert rel(gb, ob) < 5e-3, l
    assert rel(dx, dx_o) < 5e-3
    # a transposed or missing term would be O(1) off
    assert rel(layers[0][0].T[:4, :4].ravel(), layers_o[0][0][:4, :4].ravel()) > 0.1
