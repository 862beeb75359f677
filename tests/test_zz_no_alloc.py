"""lbmib_sg_conv3d_bwd_weight_noalloc against lbmib_sg_conv3d_bwd_weight,
through the C-ABI: the same weight and bias gradients bit for bit, on two
calls in a row (the second with a smaller input, so the scratch array still
holds the first call's partials), and a shape beyond the scratch size is
rejected.  The entry point allocates nothing; its partials live in the
library's fixed device array."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _lib():
    import surrogate
    L = surrogate._bind()
    vp, i = ctypes.c_void_p, ctypes.c_int
    L.lbmib_sg_conv3d_bwd_weight_noalloc.argtypes = [vp, vp, vp, vp, i, i, i, i, i, i, vp]
    L.lbmib_sg_conv3d_bwd_weight_noalloc.restype = i
    return L


def _run(L, fn, x, dy, shape, ci, co):
    dw = torch.full((co, 27, ci), float("nan"), dtype=torch.float32, device="cuda")
    db = torch.full((co,), float("nan"), dtype=torch.float32, device="cuda")
    rc = fn(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(dy.data_ptr()), ctypes.c_void_p(dw.data_ptr()),
            ctypes.c_void_p(db.data_ptr()), *shape, ci, co, None)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return dw.cpu().numpy(), db.cpu().numpy()


def test_noalloc_weight_gradient_matches_allocating_entry_point():
    L = _lib()
    g = torch.Generator(device="cpu").manual_seed(3)
    cases = [((2, 32, 32, 32), 32, 64), ((1, 8, 8, 8), 64, 32)]  # second call smaller
    for shape, ci, co in cases:
        print(f"noalloc wgrad {shape} {ci}->{co}", flush=True)
        x = torch.randn(shape + (ci,), generator=g).to(torch.bfloat16).to("cuda")
        dy = torch.randn(shape + (co,), generator=g).to(torch.bfloat16).to("cuda")
        a_w, a_b = _run(L, L.lbmib_sg_conv3d_bwd_weight_noalloc, x, dy, shape, ci, co)
        r_w, r_b = _run(L, L.lbmib_sg_conv3d_bwd_weight, x, dy, shape, ci, co)
        assert np.isfinite(a_w).all() and np.isfinite(a_b).all()
        np.testing.assert_array_equal(a_w, r_w)
        np.testing.assert_array_equal(a_b, r_b)


def test_noalloc_rejects_unsupported_shapes():
    L = _lib()
    x = torch.zeros((1, 4, 4, 4, 32), dtype=torch.bfloat16, device="cuda")
    dw = torch.zeros(1, dtype=torch.float32, device="cuda")
    # 1024 output channels: beyond what the entry point accepts
    rc = L.lbmib_sg_conv3d_bwd_weight_noalloc(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                              ctypes.c_void_p(dw.data_ptr()), None, 1, 4, 4, 4, 32, 1024, None)
    assert rc == 1  # LBMIB_ERR_ARG
