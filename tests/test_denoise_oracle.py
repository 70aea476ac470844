"""Pins of the denoiser oracle (oracle/denoise_ref.py), CPU only.

  - conv against a brute-force loop over pixels, taps and channels;
  - max-pool routing, ties included (first maximum in scan order);
  - the loss against its closed form;
  - the whole-network gradient against central finite differences in
    float64, so a dropped term, a flipped tap or a transposed weight in the
    backward restatement fails here before the GPU is compared with it;
  - bf16 rounding against known values.
"""
import numpy as np
import pytest

import denoise_ref as dr


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 6, 16))
    w = rng.normal(size=(32, 9, 16))
    b = rng.normal(size=32)
    y = dr.conv(x, w, b)
    ref = np.zeros_like(y)
    for n in range(2):
        for h in range(5):
            for ww in range(6):
                for co in range(32):
                    s = b[co]
                    for t in range(9):
                        hh, xx = h + t // 3 - 1, ww + t % 3 - 1
                        if 0 <= hh < 5 and 0 <= xx < 6:
                            s += x[n, hh, xx] @ w[co, t]
                    ref[n, h, ww, co] = s
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_conv_backward_matches_brute_force():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 4, 5, 16))
    w = rng.normal(size=(16, 9, 16))
    g = rng.normal(size=(1, 4, 5, 16))
    dx, dw, db = dr.conv_backward(x, w, g)
    # brute force through the linear map y = A x: dx = A^T g
    ref_dx = np.zeros_like(x)
    ref_dw = np.zeros_like(w)
    for h in range(4):
        for ww in range(5):
            for t in range(9):
                hh, xx = h + t // 3 - 1, ww + t % 3 - 1
                if 0 <= hh < 4 and 0 <= xx < 5:
                    ref_dx[0, hh, xx] += w[:, t].T @ g[0, h, ww]
                    ref_dw[:, t] += np.outer(g[0, h, ww], x[0, hh, xx])
    np.testing.assert_allclose(dx, ref_dx, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(dw, ref_dw, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(db, g.sum(axis=(0, 1, 2)), rtol=1e-12)


def test_pool_routes_to_first_maximum():
    x = np.zeros((1, 2, 2, 2))
    x[0, :, :, 0] = [[1.0, 3.0], [3.0, 2.0]]   # tie: (0,1) before (1,0)
    x[0, :, :, 1] = [[5.0, 5.0], [5.0, 5.0]]   # all equal: (0,0)
    assert dr.pool(x).tolist() == [[[[3.0, 5.0]]]]
    g = dr.pool_backward(x, np.array([[[[7.0, 9.0]]]]))
    assert g[0, :, :, 0].tolist() == [[0.0, 7.0], [0.0, 0.0]]
    assert g[0, :, :, 1].tolist() == [[9.0, 0.0], [0.0, 0.0]]


def test_loss_closed_form():
    L1 = dr.LOSS_L1_TONEMAP
    t = np.full((1, 2, 2, 3), 3.0)                 # T(t) = 3/4
    z = np.zeros((1, 2, 2, 16))
    z[..., :3] = np.log(4.0)                       # 1 - exp(-z) = 3/4
    loss, dz = dr.loss_and_dz(z, t, kind=L1)
    assert loss == pytest.approx(0.0, abs=1e-15)
    z[..., :3] = 0.0                                # T = 0: |0 - 3/4|
    loss, dz = dr.loss_and_dz(z, t, kind=L1)
    assert loss == pytest.approx(0.75)
    assert np.allclose(dz[..., :3], -1.0 / 12.0) and not dz[..., 3:].any()
    z[..., :3] = -2.0                               # below zero the tone map is the identity
    loss, _ = dr.loss_and_dz(z, t, kind=L1)
    assert loss == pytest.approx(2.75)


def test_relative_l2_closed_form():
    t = np.full((1, 2, 2, 3), 3.0)
    z = np.zeros((1, 2, 2, 16))
    z[..., :3] = np.log(4.0)                       # y = expm1(z) = 3: no error
    loss, dz = dr.loss_and_dz(z, t)
    assert loss == pytest.approx(0.0, abs=1e-15) and not dz.any()
    z[..., :3] = 0.0                                # y = 0: (0 - 3)^2 / (0 + 0.01)
    loss, dz = dr.loss_and_dz(z, t)
    assert loss == pytest.approx(900.0)
    assert np.allclose(dz[..., :3], 2.0 * -3.0 / 0.01 / 12.0) and not dz[..., 3:].any()
    z[..., :3] = -2.0                               # y = z below zero; the denominator uses max(y, 0)
    loss, dz = dr.loss_and_dz(z, t)
    assert loss == pytest.approx(2500.0)
    assert np.allclose(dz[..., :3], 2.0 * -5.0 / 0.01 / 12.0)


def test_relative_l2_gradient_holds_the_denominator():
    """d/dz of (y(z) - t)^2 / c with c = max(y(z0), 0)^2 + eps frozen at z0,
    by finite differences, on both branches of y."""
    rng = np.random.default_rng(4)
    z0 = np.concatenate([rng.uniform(-1.5, -0.1, 20), rng.uniform(0.1, 3.0, 20)])
    t = rng.uniform(0.0, 5.0, 40)
    _, g = dr.rel_l2_elem(z0, t)
    c = np.maximum(np.where(z0 >= 0, np.expm1(z0), z0), 0.0) ** 2 + dr.REL_L2_EPS
    y = lambda z: np.where(z >= 0, np.expm1(z), z)
    h = 1e-6
    fd = ((y(z0 + h) - t) ** 2 - (y(z0 - h) - t) ** 2) / (2 * h) / c
    np.testing.assert_allclose(g, fd, rtol=1e-6)


def test_relative_l2_learns_the_mean_and_l1_the_median():
    """One prediction y fit to many noisy targets of a skewed (Monte Carlo
    like) distribution: the relative-L2 gradient, its denominator shared,
    vanishes at the targets' mean; the L1 tone-mapped loss is smallest at
    their median, below the mean. This is why relative L2 is the default."""
    rng = np.random.default_rng(8)
    t = rng.lognormal(-1.0, 1.2, 4001)
    mean, median = t.mean(), np.median(t)
    assert median < 0.6 * mean

    def grad_rel(y):
        return dr.rel_l2_elem(np.full_like(t, np.log1p(y)), t)[1].sum()
    lo, hi = 1e-3, 50.0                              # bisection on the gradient's sign
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if grad_rel(mid) < 0 else (lo, mid)
    assert lo == pytest.approx(mean, rel=1e-9)
    ys = np.linspace(0.01, 3.0 * mean, 3001)
    l1 = [np.abs((1 - 1 / (1 + y)) - t / (1 + t)).sum() for y in ys]
    assert ys[int(np.argmin(l1))] == pytest.approx(median, abs=2 * (ys[1] - ys[0]))


def test_bf16_rounding():
    assert dr.bf16(np.array([1.0, 1.00390625, 1.005859375, -2.0]))[:4].tolist() == [1.0, 1.0, 1.0078125, -2.0]
    assert np.isnan(dr.bf16(np.array([np.nan]))[0])


def _problem(levels, channels, n=2, h=8, w=8, seed=3):
    rng = np.random.default_rng(seed)
    net = dr.Net(levels, channels)
    net.params = rng.normal(scale=0.15, size=net.nparams)
    beauty = rng.uniform(0.0, 3.0, (n, h, w, 3))
    albedo = rng.uniform(0.0, 1.0, (n, h, w, 3))
    normal = rng.uniform(0.0, 1.0, (n, h, w, 3))
    target = rng.uniform(0.0, 3.0, (n, h, w, 3))
    return net, beauty, albedo, normal, target


@pytest.mark.parametrize("levels,channels,h,w", [(2, (16, 16), 8, 8), (3, (16, 32, 16), 7, 9)])
def test_gradient_matches_finite_differences(levels, channels, h, w):
    """Through the network with the L1 loss, whose gradient is the true
    derivative (the relative-L2 gradient holds its denominator, checked
    element by element above; the chain through the network is the same)."""
    net, b, a, nr, t = _problem(levels, channels, h=h, w=w)
    kind = dr.LOSS_L1_TONEMAP
    loss, g, _ = dr.loss_and_gradients(net, b, a, nr, t, kind=kind)
    rng = np.random.default_rng(11)
    idx = set(rng.choice(net.nparams, 40, replace=False).tolist())
    for i in range(len(net.convs)):  # at least one weight and the first bias of every convolution
        w0, b0 = net.offsets[i]
        cin = net.convs[i][0]
        idx.update([w0 + 4 * cin + 1, b0])
    eps = 1e-6
    checked = 0
    for i in sorted(idx):
        p = net.params.copy()
        p[i] += eps
        lp = dr.loss_and_gradients(net, b, a, nr, t, params=p, kind=kind)[0]
        p[i] -= 2 * eps
        lm = dr.loss_and_gradients(net, b, a, nr, t, params=p, kind=kind)[0]
        fd = (lp - lm) / (2 * eps)
        assert fd == pytest.approx(g[i], rel=1e-4, abs=1e-8), f"param {i}"
        checked += 1
    assert checked >= 40
    # padded input channels (9..15) carry no gradient
    w0, _ = net.offsets[0]
    wg = g[w0:w0 + net.convs[0][1] * 9 * 16].reshape(-1, 9, 16)
    assert not wg[:, :, 9:].any()


def test_halves_pack_mean_and_log_difference():
    a = np.array([[[[1.0, 3.0, 0.0]], [[100.0, 0.0, 0.5]]]])
    b = np.array([[[[3.0, 1.0, -1.0]], [[0.0, 100.0, 0.5]]]])
    x = dr.pack_input(a, np.zeros_like(a), np.zeros_like(a), 1, beauty_b=b)
    np.testing.assert_allclose(x[0, 0, 0, 0:3], np.log1p([2.0, 2.0, 0.0]))
    np.testing.assert_allclose(x[0, 0, 0, 9:12], [np.log(2.0) - np.log(4.0), np.log(4.0) - np.log(2.0), 0.0])
    # a firefly's log ratio is clipped to +-2
    np.testing.assert_allclose(x[0, 1, 0, 9:12], [2.0, -2.0, 0.0])
    assert not x[0, 0, 0, 12:].any()


def test_halves_network_with_zero_difference_weights_is_the_single_network():
    """Weights trained for one beauty image carry over: with zero weights on
    channels 9-11, the halves network on (a, b) equals the single-image
    network on their mean."""
    net, b, a, nr, _ = _problem(2, (16, 16))
    w0, b0 = net.offsets[0]
    net.params[w0:b0].reshape(net.convs[0][1], 9, 16)[:, :, 9:] = 0.0
    rng = np.random.default_rng(2)
    b2 = rng.uniform(0.0, 3.0, b.shape)
    z_halves, _ = dr.forward(net, b, a, nr, beauty_b=b2)
    z_mean, _ = dr.forward(net, 0.5 * (b + b2), a, nr)
    np.testing.assert_allclose(z_halves, z_mean, rtol=1e-12, atol=1e-12)
