"""Surrogate kernels (csrc/lbmib_surr.hip) against the CPU oracle
(oracle/surrogate_oracle.py), through the C-ABI of include/lbmib_surrogate.h.

Tolerances, written per element:
  BF16 outputs (activations, data gradients): |gpu - ref| <= 2^-8 (|ref| +
      a) + a, with a = 2^-16 * sum |terms| -- one rounding to BF16 (half an
      ulp, 2^-8 relative) of a value whose FP32 accumulation may differ from
      the exact sum by a few FP32 ulps of the sum of |terms|;
  FP32 weight gradients: |gpu - ref| <= 2^-18 * sum |terms| + 1e-30;
  pooling, upsampling and their backward: the same BF16 bound with the
      oracle's FP64 value;
  the whole network (16^3 voxels, four levels): the output and every
      parameter gradient within 3x the relative L2 distance from exact FP64
      math that the same BF16 rounding gives the oracle, plus 1e-3.
"""
import numpy as np
import pytest
import torch

from oracle import surrogate_oracle as so

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sg():
    import surrogate
    return surrogate


# Uploaded tensors stay referenced until the test ends: the kernels run
# asynchronously, and a tensor whose last reference goes away right after its
# pointer is taken would return its memory to torch's caching allocator for
# the next upload to overwrite before the kernel reads it.
_alive = []


@pytest.fixture(autouse=True)
def _keep_uploads_alive():
    yield
    torch.cuda.synchronize()
    _alive.clear()


def dev_bf16(a):
    t = torch.from_numpy(so.bf16_bits(a).view(np.int16)).to("cuda").view(torch.bfloat16)
    _alive.append(t)
    return t


def host_bf16(t):
    torch.cuda.synchronize()
    return so.from_bf16_bits(t.contiguous().view(torch.int16).cpu().numpy().view(np.uint16))


def dev_f32(a):
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")
    _alive.append(t)
    return t


def host_f32(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(np.float64)


def assert_bf16_close(gpu, ref, absum, what):
    a = 2.0 ** -16 * absum
    tol = 2.0 ** -8 * (np.abs(ref) + a) + a
    bad = np.abs(gpu - ref) > tol
    assert not bad.any(), f"{what}: {bad.sum()} of {bad.size} off, worst {np.abs(gpu - ref).max():.3e}"


CONV_CASES = [
    # (B, D, H, W, Cin, Cout): every pair of the network, grids that fill and
    # that do not fill the block tiles; samples of at most 512 voxels take
    # the padded-slab form, (1, 8, 16, 8, ...) the split over chunks
    (1, 8, 8, 8, 16, 32), (2, 4, 6, 8, 32, 32), (1, 16, 16, 16, 32, 64), (1, 8, 8, 8, 64, 64),
    (1, 4, 8, 8, 64, 128), (2, 4, 4, 4, 128, 128), (1, 4, 4, 8, 128, 256), (1, 2, 4, 4, 256, 256),
    (1, 4, 4, 4, 256, 128), (1, 8, 8, 8, 128, 64), (1, 8, 16, 8, 64, 32), (1, 8, 8, 8, 32, 16),
    (3, 5, 7, 3, 16, 48), (1, 64, 16, 16, 32, 32),
    (1, 16, 16, 16, 16, 32),  # two waves per tile splitting k, the last chunk half past 27 Cin
    (1, 16, 32, 64, 64, 32),  # 32768 voxels, wide input: the input box with 16-channel tiles
    (1, 16, 16, 16, 128, 64),  # 16^3 with 128 input channels: the padded slab of each z-plane
    (1, 32, 64, 64, 32, 32),  # 131072 voxels: the input box in LDS
]
# 65536 voxels x 64 -> 128 channels: the input box with eight channel blocks
# and 64-channel tiles (forward only: its oracle takes seconds per convolution)
WIDE_BOX_CASE = (1, 32, 32, 64, 64, 128)


@pytest.mark.parametrize("case", CONV_CASES + [WIDE_BOX_CASE], ids=lambda c: "x".join(map(str, c)))
def test_conv3d_fwd_matches_oracle(sg, case):
    B, D, H, W, ci, co = case
    print(f"conv fwd {case}", flush=True)
    rng = np.random.default_rng(hash(case) % 2**32)
    x = so.bf16_round(rng.standard_normal((B, D, H, W, ci)))
    w = so.bf16_round(rng.standard_normal((co, 27, ci)) / np.sqrt(27 * ci))
    b = rng.standard_normal(co).astype(np.float32).astype(np.float64) * 0.1
    L = sg._bind()
    for relu in (0, 1):
        y = torch.empty((B, D, H, W, co), dtype=torch.bfloat16, device="cuda")
        sg.check(L.lbmib_sg_conv3d_fwd(sg.ptr(dev_bf16(x)), sg.ptr(dev_bf16(w)), sg.ptr(dev_f32(b)), None, sg.ptr(y),
                                       B, D, H, W, ci, co, relu, sg.stream()), "conv fwd")
        ref = so.conv3d(x, w, b)
        if relu:
            ref = np.maximum(ref, 0.0)
        absum = so.conv3d(np.abs(x), np.abs(w), np.abs(b))
        assert_bf16_close(host_bf16(y), ref, absum, f"fwd relu={relu}")


@pytest.mark.parametrize("case", [(1, 4, 8, 8, 32, 64), (1, 8, 16, 8, 64, 32), (1, 16, 16, 16, 32, 64),
                                  (1, 32, 64, 64, 32, 32)],
                         ids=["volume", "split", "wave-split", "box"])
def test_conv3d_fwd_mask(sg, case):
    B, D, H, W, ci, co = case
    rng = np.random.default_rng(1)
    x = so.bf16_round(rng.standard_normal((B, D, H, W, ci)))
    w = so.bf16_round(rng.standard_normal((co, 27, ci)) / 30)
    mask = so.bf16_round(np.maximum(rng.standard_normal((B, D, H, W, co)), 0.0))
    L = sg._bind()
    y = torch.empty((B, D, H, W, co), dtype=torch.bfloat16, device="cuda")
    sg.check(L.lbmib_sg_conv3d_fwd(sg.ptr(dev_bf16(x)), sg.ptr(dev_bf16(w)), None, sg.ptr(dev_bf16(mask)), sg.ptr(y),
                                   B, D, H, W, ci, co, 0, sg.stream()), "conv fwd mask")
    ref = so.conv3d(x, w) * (mask > 0)
    assert_bf16_close(host_bf16(y), ref, so.conv3d(np.abs(x), np.abs(w)), "fwd mask")


@pytest.mark.parametrize("case", [c for c in CONV_CASES if c[4] >= 16 and c[5] in (16, 32, 64, 128, 256)],
                         ids=lambda c: "x".join(map(str, c)))
def test_conv3d_bwd_data_matches_oracle(sg, case):
    B, D, H, W, ci, co = case
    print(f"conv bwd data {case}", flush=True)
    rng = np.random.default_rng(hash(case) % 2**32 + 1)
    dy = so.bf16_round(rng.standard_normal((B, D, H, W, co)))
    w = so.bf16_round(rng.standard_normal((co, 27, ci)) / np.sqrt(27 * co))
    mask = so.bf16_round(np.maximum(rng.standard_normal((B, D, H, W, ci)), 0.0))
    L = sg._bind()
    wt = torch.empty((ci, 27, co), dtype=torch.bfloat16, device="cuda")
    for use_mask in (False, True):
        dx = torch.empty((B, D, H, W, ci), dtype=torch.bfloat16, device="cuda")
        sg.check(L.lbmib_sg_conv3d_bwd_data(sg.ptr(dev_bf16(dy)), sg.ptr(dev_bf16(w)), sg.ptr(wt),
                                            sg.ptr(dev_bf16(mask)) if use_mask else None, sg.ptr(dx), B, D, H, W, ci,
                                            co, sg.stream()), "conv bwd data")
        ref = so.conv3d_bwd_data(dy, w)
        if use_mask:
            ref = ref * (mask > 0)
        assert_bf16_close(host_bf16(dx), ref, so.conv3d_bwd_data(np.abs(dy), np.abs(w)), f"dgrad mask={use_mask}")
    # the flipped weight: wt[ci][26 - t][co] = w[co][t][ci]
    np.testing.assert_array_equal(host_bf16(wt), w[:, ::-1, :].transpose(2, 1, 0))


@pytest.mark.parametrize("case", CONV_CASES + [(1, 2, 2, 2, 256, 128), (1, 1, 3, 5, 32, 64)],
                         ids=lambda c: "x".join(map(str, c)))
def test_conv3d_bwd_weight_matches_oracle(sg, case):
    B, D, H, W, ci, co = case
    print(f"conv bwd weight {case}", flush=True)
    rng = np.random.default_rng(hash(case) % 2**32 + 2)
    x = so.bf16_round(rng.standard_normal((B, D, H, W, ci)))
    dy = so.bf16_round(rng.standard_normal((B, D, H, W, co)))
    L = sg._bind()
    dw = torch.empty((co, 27, ci), dtype=torch.float32, device="cuda")
    db = torch.empty(co, dtype=torch.float32, device="cuda")
    sg.check(L.lbmib_sg_conv3d_bwd_weight(sg.ptr(dev_bf16(x)), sg.ptr(dev_bf16(dy)), sg.ptr(dw), sg.ptr(db), B, D, H,
                                          W, ci, co, sg.stream()), "conv bwd weight")
    rw, rb = so.conv3d_bwd_weight(x, dy)
    aw, ab = so.conv3d_bwd_weight(np.abs(x), np.abs(dy))
    assert np.all(np.abs(host_f32(dw) - rw) <= 2.0 ** -18 * aw + 1e-30)
    assert np.all(np.abs(host_f32(db) - rb) <= 2.0 ** -18 * ab + 1e-30)


def test_pool_upsample_and_backward(sg):
    B, D, H, W, C = 2, 4, 6, 8, 32
    rng = np.random.default_rng(3)
    L = sg._bind()
    x = so.bf16_round(rng.standard_normal((B, D, H, W, C)))
    y = torch.empty((B, D // 2, H // 2, W // 2, C), dtype=torch.bfloat16, device="cuda")
    sg.check(L.lbmib_sg_pool2(sg.ptr(dev_bf16(x)), sg.ptr(y), B, D // 2, H // 2, W // 2, C, sg.stream()), "pool2")
    assert_bf16_close(host_bf16(y), so.pool2(x), 8 * so.pool2(np.abs(x)), "pool2")

    c = so.bf16_round(rng.standard_normal((B, D // 2, H // 2, W // 2, C)))
    e = so.bf16_round(rng.standard_normal((B, D, H, W, C)))
    s = torch.empty((B, D, H, W, C), dtype=torch.bfloat16, device="cuda")
    sg.check(L.lbmib_sg_up2_add(sg.ptr(dev_bf16(c)), sg.ptr(dev_bf16(e)), sg.ptr(s), B, D, H, W, C, sg.stream()),
             "up2_add")
    assert_bf16_close(host_bf16(s), so.up2(c) + e, so.up2(np.abs(c)) + np.abs(e), "up2_add")

    mask = so.bf16_round(np.maximum(rng.standard_normal((B, D, H, W, C)), 0))
    out = torch.empty((B, D, H, W, C), dtype=torch.bfloat16, device="cuda")
    sg.check(L.lbmib_sg_pool2_bwd(sg.ptr(dev_bf16(e)), sg.ptr(dev_bf16(c)), sg.ptr(dev_bf16(mask)), sg.ptr(out), B, D,
                                  H, W, C, sg.stream()), "pool2_bwd")
    assert_bf16_close(host_bf16(out), (e + so.up2(c) / 8.0) * (mask > 0), np.abs(e) + so.up2(np.abs(c)) / 8.0,
                      "pool2_bwd")

    cm = so.bf16_round(np.maximum(rng.standard_normal(c.shape), 0))
    gc = torch.empty(c.shape, dtype=torch.bfloat16, device="cuda")
    sg.check(L.lbmib_sg_up2_bwd(sg.ptr(dev_bf16(e)), sg.ptr(dev_bf16(cm)), sg.ptr(gc), B, D // 2, H // 2, W // 2, C,
                                sg.stream()), "up2_bwd")
    assert_bf16_close(host_bf16(gc), so.up2_adjoint(e) * (cm > 0), so.up2_adjoint(np.abs(e)), "up2_bwd")


def test_mse_adam_pack_unpack(sg):
    L = sg._bind()
    rng = np.random.default_rng(4)
    M, C, nl = 4096, 16, 4
    out = so.bf16_round(rng.standard_normal((M, C)))
    tgt = rng.standard_normal((M, nl)).astype(np.float32).astype(np.float64)
    dout = torch.empty((M, C), dtype=torch.bfloat16, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    sg.check(L.lbmib_sg_mse(sg.ptr(dev_bf16(out)), sg.ptr(dev_f32(tgt)), sg.ptr(dout), sg.ptr(loss), M, C, nl,
                            sg.stream()), "mse")
    rl, rg = so.mse(out, tgt, nl)
    assert abs(host_f32(loss)[0] - rl) <= 1e-5 * rl
    assert_bf16_close(host_bf16(dout), rg, np.abs(rg), "mse grad")

    n = 10000
    p, g = rng.standard_normal(n), rng.standard_normal(n)
    p32, g32 = p.astype(np.float32).astype(np.float64), g.astype(np.float32).astype(np.float64)
    dp, dg = dev_f32(p32), dev_f32(g32)
    dm, dv = torch.zeros(n, device="cuda"), torch.zeros(n, device="cuda")
    w16 = torch.empty(n, dtype=torch.bfloat16, device="cuda")
    rp, rm, rv = p32, np.zeros(n), np.zeros(n)
    for t in (1, 2, 3):
        sg.check(L.lbmib_sg_adam(sg.ptr(dp), sg.ptr(dg), sg.ptr(dm), sg.ptr(dv), sg.ptr(w16), n, 1e-3, 0.9, 0.999, 1e-8,
                                 t, sg.stream()), "adam")
        rp, rm, rv = so.adam(rp, g32, rm, rv, 1e-3, 0.9, 0.999, 1e-8, t)
    np.testing.assert_allclose(host_f32(dp), rp, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(host_bf16(w16), so.bf16_round(host_f32(dp)))

    nf, Cx = 4, 16
    f = rng.standard_normal((nf, M)).astype(np.float32)
    mean, istd, cval = [0.1, -0.2, 0.3, 0.0], [2.0, 3.0, 0.5, 1.0], [0.7, -1.5]
    x = torch.empty((M, Cx), dtype=torch.bfloat16, device="cuda")
    sg.check(L.lbmib_sg_pack(sg.ptr(dev_f32(f)), nf, sg.dvec(mean), sg.dvec(istd), 2, sg.dvec(cval), sg.ptr(x), M, Cx,
                             sg.stream()), "pack")
    ref = np.zeros((M, Cx))
    ref[:, :nf] = ((f.astype(np.float64) - np.array(mean)[:, None]) * np.array(istd)[:, None]).T
    ref[:, nf:nf + 2] = cval
    assert_bf16_close(host_bf16(x), ref, np.abs(ref) + 1.0, "pack")

    f1 = rng.standard_normal((nf, M)).astype(np.float32)
    tg = torch.empty((M, nf), dtype=torch.float32, device="cuda")
    sg.check(L.lbmib_sg_target(sg.ptr(dev_f32(f)), sg.ptr(dev_f32(f1)), nf, sg.dvec(mean), sg.dvec(istd), sg.ptr(tg), M,
                               sg.stream()), "target")
    rt = (((f1 - f).astype(np.float64) - np.array(mean)[:, None]) * np.array(istd)[:, None]).T
    np.testing.assert_allclose(host_f32(tg), rt, rtol=1e-6, atol=1e-6)

    d = torch.empty((nf, M), dtype=torch.float32, device="cuda")
    sg.check(L.lbmib_sg_unpack(sg.ptr(x), Cx, nf, sg.dvec(mean), sg.dvec(istd), sg.ptr(d), M, sg.stream()), "unpack")
    np.testing.assert_allclose(host_f32(d), (host_bf16(x)[:, :nf] * np.array(istd) + np.array(mean)).T, rtol=1e-6,
                               atol=1e-6)

    o2 = torch.zeros(2, dtype=torch.float64, device="cuda")
    sg.check(L.lbmib_sg_sqdiff(sg.ptr(dev_f32(f)), sg.ptr(dev_f32(f1)), f.size, sg.ptr(o2), sg.stream()), "sqdiff")
    torch.cuda.synchronize()
    r0 = ((f.astype(np.float64) - f1) ** 2).sum()
    r1 = (f1.astype(np.float64) ** 2).sum()
    np.testing.assert_allclose(o2.cpu().numpy(), [r0, r1], rtol=1e-12)


def test_network_forward_backward_matches_oracle(sg):
    """The whole encoder-decoder on 16^3 voxels (levels 16, 8, 4, 2): the
    GPU's output and every parameter gradient against the oracle network in
    exact FP64 (no rounding of stored tensors).  The GPU stores every
    activation and gradient in BF16, so it cannot match exact math; the bound
    is what the same rounding does to the oracle: the GPU's distance to the
    exact result may be at most 3x the rounding oracle's own distance plus
    1e-3 (relative L2).  An orchestration mistake -- a wrong mask, skip or
    layer input -- puts the GPU far outside that."""
    net = sg.Net(cin=16, cout=16, nout=4, seed=5, out_scale=0.05)
    params = {k: (so.bf16_round(net.p[k].cpu().numpy()), net.b[k].cpu().numpy().astype(np.float64)) for k in net.p}
    rng = np.random.default_rng(6)
    x = so.bf16_round(rng.standard_normal((1, 16, 16, 16, 16)))
    out = host_bf16(net.forward(dev_bf16(x)))
    exact, act_e = so.unet_forward(params, x, sg.CHANS, rounding=False)
    rounded, act_r = so.unet_forward(params, x, sg.CHANS)
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))  # noqa: E731
    e_gpu, e_orc = rel(out, exact), rel(rounded, exact)
    print(f"output: gpu {e_gpu:.2e}, rounding oracle {e_orc:.2e} (relative L2 to exact)")
    assert e_gpu <= 3 * e_orc + 1e-3
    dout = so.bf16_round(rng.standard_normal(out.shape) * 0.01)
    net.backward(dev_bf16(dout))
    g_exact = so.unet_backward(params, act_e, dout, sg.CHANS, rounding=False)
    g_round = so.unet_backward(params, act_r, dout, sg.CHANS)
    for k in net.p:
        for i, (gpu, what) in enumerate(((host_f32(net.g[k]), "w"), (host_f32(net.gb[k]), "b"))):
            e_gpu, e_orc = rel(gpu, g_exact[k][i]), rel(g_round[k][i], g_exact[k][i])
            print(f"{k}.{what}: gpu {e_gpu:.2e}, rounding oracle {e_orc:.2e}")
            assert e_gpu <= 3 * e_orc + 1e-3, (k, what, e_gpu, e_orc)


def test_training_reduces_the_loss(sg):
    """A fixed random field-to-field map on 16^3: 60 Adam steps of the GPU
    kernels bring the loss well below its starting value."""
    net = sg.Net(cin=16, cout=16, nout=4, seed=7)
    rng = np.random.default_rng(8)
    xs = [dev_bf16(rng.standard_normal((2, 16, 16, 16, 16))) for _ in range(4)]
    # target: a smoothed linear function of the first input channels
    tg = []
    for x in xs:
        xh = host_bf16(x)
        t = 0.5 * xh[..., :4] + 0.25 * np.roll(xh[..., 4:8], 1, axis=2)
        tg.append(dev_f32(t.reshape(-1, 4)))
    losses = []
    for it in range(60):
        loss = net.train_step(xs[it % 4], tg[it % 4], 2e-3)
        losses.append(float(loss.item()))
    print("loss", losses[0], "->", losses[-1])
    assert np.isfinite(losses).all()
    assert np.mean(losses[-8:]) < 0.5 * np.mean(losses[:4])
