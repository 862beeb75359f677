"""Pins the surrogate's CPU oracle (oracle/surrogate_oracle.py) without a GPU.

- conv3d against torch.nn.functional.conv3d on the CPU in FP64 (an independent
  implementation of the same cross-correlation);
- every adjoint (data and weight gradients, pooling and upsampling
  backward) against central differences of the oracle's own forward (the
  functions are linear in the perturbed argument, so the differences are
  exact up to rounding) and adjoint identities <A a, b> = <a, A^T b>;
- the network's backward against central differences of its forward
  (without rounding, two levels, away from ReLU kinks);
- Adam's first step against its closed form p - lr g / (|g| + eps);
- BF16 rounding against torch's float32 -> bfloat16 conversion.

A planted mistake in the oracle -- a flipped tap, a transposed weight, a
missing 1/8 in the pooling adjoint, a dropped bias-correction -- fails one
of these.
"""
import numpy as np
import pytest
import torch

from oracle import surrogate_oracle as so


def rnd(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def torch_conv(x, w, b):
    """conv3d through torch (NCDHW, weight [co][ci][kd][kh][kw])."""
    xt = torch.from_numpy(x).permute(0, 4, 1, 2, 3)
    co, _, ci = w.shape
    wt = torch.from_numpy(w.reshape(co, 3, 3, 3, ci)).permute(0, 4, 1, 2, 3).contiguous()
    y = torch.nn.functional.conv3d(xt, wt, torch.from_numpy(b), padding=1)
    return y.permute(0, 2, 3, 4, 1).numpy()


@pytest.mark.parametrize("shape,ci,co", [((1, 4, 5, 6), 3, 2), ((2, 3, 3, 3), 5, 4), ((1, 1, 2, 7), 2, 3)])
def test_conv3d_matches_torch(shape, ci, co):
    x = rnd(shape + (ci,), 1)
    w = rnd((co, 27, ci), 2)
    b = rnd((co,), 3)
    np.testing.assert_allclose(so.conv3d(x, w, b), torch_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv3d_tap_order_single_tap():
    """One non-zero tap t = 9 (dz+1) + 3 (dy+1) + (dx+1) reads x at v + o(t)."""
    x = rnd((1, 4, 4, 4, 1), 4)
    for t in (0, 5, 13, 22, 26):
        w = np.zeros((1, 27, 1))
        w[0, t, 0] = 1.0
        dz, dy, dx = t // 9 - 1, t // 3 % 3 - 1, t % 3 - 1
        y = so.conv3d(x, w)
        assert y[0, 1, 1, 1, 0] == x[0, 1 + dz, 1 + dy, 1 + dx, 0]


def _fd_check(f, arg, grad, idxs, h=1e-3):
    for idx in idxs:
        e = np.zeros_like(arg)
        e[idx] = h
        fd = (f(arg + e) - f(arg - e)) / (2 * h)
        assert abs(fd - grad[idx]) <= 1e-9 * max(1.0, abs(fd)), (idx, fd, grad[idx])


def test_conv3d_adjoints_by_finite_differences():
    x = rnd((2, 3, 4, 5, 3), 5)
    w = rnd((4, 27, 3), 6)
    b = rnd((4,), 7)
    R = rnd((2, 3, 4, 5, 4), 8)
    dx = so.conv3d_bwd_data(R, w)
    dw, db = so.conv3d_bwd_weight(x, R)
    rng = np.random.default_rng(9)
    pick = lambda shape: [tuple(int(rng.integers(0, s)) for s in shape) for _ in range(12)]  # noqa: E731
    _fd_check(lambda a: float((so.conv3d(a, w, b) * R).sum()), x, dx, pick(x.shape))
    _fd_check(lambda a: float((so.conv3d(x, a, b) * R).sum()), w, dw, pick(w.shape))
    _fd_check(lambda a: float((so.conv3d(x, w, a) * R).sum()), b, db, pick(b.shape))


def test_conv3d_adjoints_match_torch_autograd():
    x = rnd((1, 4, 4, 4, 3), 10)
    w = rnd((5, 27, 3), 11)
    R = rnd((1, 4, 4, 4, 5), 12)
    xt = torch.from_numpy(x).requires_grad_(True)
    wt = torch.from_numpy(w).requires_grad_(True)
    y = torch.nn.functional.conv3d(xt.permute(0, 4, 1, 2, 3),
                                   wt.reshape(5, 3, 3, 3, 3).permute(0, 4, 1, 2, 3), padding=1)
    (y.permute(0, 2, 3, 4, 1) * torch.from_numpy(R)).sum().backward()
    np.testing.assert_allclose(so.conv3d_bwd_data(R, w), xt.grad.numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(so.conv3d_bwd_weight(x, R)[0], wt.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_pool_and_upsample_adjoints():
    x = rnd((2, 4, 6, 8, 3), 13)
    y = rnd((2, 2, 3, 4, 3), 14)
    # pooling: the mean of each 2x2x2 block, checked on one block by hand
    assert np.isclose(so.pool2(x)[1, 1, 2, 3, 2], x[1, 2:4, 4:6, 6:8, 2].mean())
    # <pool2(x), y> = <x, up2(y) / 8>
    assert np.isclose((so.pool2(x) * y).sum(), (x * so.up2(y) / 8.0).sum())
    # <up2(y), x> = <y, up2_adjoint(x)>
    assert np.isclose((so.up2(y) * x).sum(), (y * so.up2_adjoint(x)).sum())
    assert so.up2(y)[0, 3, 5, 7, 1] == y[0, 1, 2, 3, 1]


def test_mse_gradient():
    out = rnd((1, 2, 2, 2, 6), 15)
    tgt = rnd((8, 3), 16)
    loss, g = so.mse(out, tgt, 3)
    assert np.all(g[..., 3:] == 0)
    _fd_check(lambda a: so.mse(a, tgt, 3)[0], out, g, [(0, 0, 1, 1, 0), (0, 1, 0, 1, 2), (0, 1, 1, 0, 1)], h=1e-4)


def test_adam_first_step_closed_form():
    p = rnd((50,), 17)
    g = rnd((50,), 18)
    lr, eps = 1e-2, 1e-8
    pn, m, v = so.adam(p, g, np.zeros(50), np.zeros(50), lr, 0.9, 0.999, eps, 1)
    np.testing.assert_allclose(pn, p - lr * g / (np.abs(g) + eps), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(19)
    f = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 5000),
                        # exact ties: the bit below the kept 16 set, nothing lower
                        (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)])
    ref = torch.from_numpy(f).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(so.bf16_round(f), ref)
    np.testing.assert_array_equal(so.from_bf16_bits(so.bf16_bits(f)), ref)


def _tiny_net(chans, cin, cout, seed):
    rng = np.random.default_rng(seed)
    params = {}
    for name, ci, co, _, relu in so.unet_layers(cin, cout, chans):
        params[name] = (rng.standard_normal((co, 27, ci)) * np.sqrt(2.0 / (27 * ci)), rng.standard_normal(co) * 0.1)
    return params


def test_unet_backward_by_finite_differences():
    chans = (3, 4)
    params = _tiny_net(chans, 2, 2, 20)
    x = rnd((1, 4, 4, 4, 2), 21)
    R = rnd((1, 4, 4, 4, 2), 22)

    def loss(p):
        out, _ = so.unet_forward(p, x, chans, rounding=False)
        return float((out * R).sum())

    out, act = so.unet_forward(params, x, chans, rounding=False)
    grads = so.unet_backward(params, act, R, chans, rounding=False)
    rng = np.random.default_rng(23)
    for name in params:
        w, b = params[name]
        dw, db = grads[name]
        for _ in range(4):
            idx = tuple(int(rng.integers(0, s)) for s in w.shape)
            h = 1e-6
            wp, wm = w.copy(), w.copy()
            wp[idx] += h
            wm[idx] -= h
            fd = (loss({**params, name: (wp, b)}) - loss({**params, name: (wm, b)})) / (2 * h)
            assert abs(fd - dw[idx]) <= 1e-5 * max(1.0, abs(fd)), (name, idx, fd, dw[idx])
        k = int(rng.integers(0, b.shape[0]))
        bp, bm = b.copy(), b.copy()
        bp[k] += 1e-6
        bm[k] -= 1e-6
        fd = (loss({**params, name: (w, bp)}) - loss({**params, name: (w, bm)})) / 2e-6
        assert abs(fd - db[k]) <= 1e-5 * max(1.0, abs(fd)), (name, "bias", fd, db[k])


def test_trilinear_reproduces_linear_fields_and_constants():
    """Away from the edges, interpolation of a linear coarse field is exact;
    a constant field is reproduced everywhere (the clamped edges too)."""
    F = 4
    ncz, ncy, ncx = 3, 4, 5
    k, j, i = np.meshgrid(np.arange(ncz), np.arange(ncy), np.arange(ncx), indexing="ij")
    # value at coarse centre (i F + 1.5, ...) of a linear function of position
    lin = lambda x, y, z: 0.3 * x - 0.7 * y + 1.1 * z + 2.0  # noqa: E731
    coarse = lin(i * F + 1.5, j * F + 1.5, k * F + 1.5)
    fine = so.trilinear_fine(coarse, F)
    z, y, x = np.meshgrid(np.arange(ncz * F), np.arange(ncy * F), np.arange(ncx * F), indexing="ij")
    core = (slice(2, ncz * F - 2), slice(2, ncy * F - 2), slice(2, ncx * F - 2))
    np.testing.assert_allclose(fine[core], lin(x, y, z)[core], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(so.trilinear_fine(np.full((ncz, ncy, ncx), 0.25), F), 0.25, rtol=1e-14)
    # the block average of the fine field of a linear function is its coarse value
    np.testing.assert_allclose(so.block_average(lin(x, y, z), F), coarse, rtol=1e-12)


def test_feq_moments_and_jump_shift():
    """feq has moments rho and rho u; a jump with a uniform change on a
    uniform equilibrium state moves every cell to feq(rho + drho, u + du)."""
    rho, ux, uy, uz = 1.02, 0.01, -0.003, 0.002
    f = so.feq(np.float64(rho), ux, uy, uz)
    assert np.isclose(f.sum(), rho)
    assert np.isclose((so.CX * f).sum(), rho * ux) and np.isclose((so.CY * f).sum(), rho * uy)
    assert np.isclose((so.CZ * f).sum(), rho * uz)
    shape = (8, 8, 8)
    f2 = np.array([np.full(shape, v) for v in f])
    d = np.array([0.001, 0.002, -0.001, 0.0005])
    delta = np.array([np.full((2, 2, 2), v) for v in d])
    out = so.jump_pulled(f2, delta, 4)
    want = so.feq(np.float64(rho + d[0]), ux + d[1], uy + d[2], uz + d[3])
    for q in range(19):
        np.testing.assert_allclose(out[q], want[q], rtol=1e-13)


def test_raster_counts_and_sums():
    pos = np.array([[0.4, 0.2, 0.1], [3.49, 0.0, 0.0], [3.51, 0.0, 0.0], [7.9, 7.9, 7.9]])
    frc = np.array([[1.0, 2.0, 3.0], [0.5, 0.5, 0.5], [1.0, 0.0, 0.0], [2.0, 2.0, 2.0]])
    r = so.raster(pos, frc, 4, (2, 2, 2))
    assert r[0, 0, 0, 0] == 2 and r[0, 0, 0, 1] == 1 and r[0, 1, 1, 1] == 1
    np.testing.assert_allclose(r[1:, 0, 0, 0], [1.5, 2.5, 3.5])
