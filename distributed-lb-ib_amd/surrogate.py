"""Learned surrogate of the LBM-IB step loop: a 3-D convolutional
encoder-decoder trained on the GPU with the library's own BF16-MFMA kernels
(include/lbmib_surrogate.h, csrc/lbmib_surr.hip).

Not in the reference: cube_pthread/lbm_ib.c has no learned model.  The
network maps the coarse-grained fluid state of compute_rho_and_u
(lbm_ib.c:2440-2538) and a raster of the sheet of moveFiberSheet
(lbm_ib.c:2540-2723) to the change of the coarse fluid state K steps of the
do_thread loop (lbm_ib.c:3115-3563) later; surrogate_jump.py applies it to
a running simulation.

PyTorch only allocates device memory and provides the stream handle: every
operation on the tensors is a kernel of liblbmib.so called through ctypes.
There is no CPU fallback.
"""
import ctypes

import numpy as np
import torch

import lbmib

CHANS = (32, 64, 128, 256)


def _bind():
    L = lbmib.load_library()
    if getattr(L, "_sg_bound", False):
        return L
    vp, i, l, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float
    dp = ctypes.POINTER(ctypes.c_double)
    L.lbmib_sg_conv3d_fwd.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, i, i, i, vp]
    L.lbmib_sg_conv3d_bwd_data.argtypes = [vp, vp, vp, vp, vp, i, i, i, i, i, i, vp]
    L.lbmib_sg_conv3d_bwd_weight.argtypes = [vp, vp, vp, vp, i, i, i, i, i, i, vp]
    L.lbmib_sg_conv3d_bwd_weight_noalloc.argtypes = [vp, vp, vp, vp, i, i, i, i, i, i, vp]
    L.lbmib_sg_pool2.argtypes = [vp, vp, i, i, i, i, i, vp]
    L.lbmib_sg_up2_add.argtypes = [vp, vp, vp, i, i, i, i, i, vp]
    L.lbmib_sg_pool2_bwd.argtypes = [vp, vp, vp, vp, i, i, i, i, i, vp]
    L.lbmib_sg_up2_bwd.argtypes = [vp, vp, vp, i, i, i, i, i, vp]
    L.lbmib_sg_mse.argtypes = [vp, vp, vp, vp, l, i, i, vp]
    L.lbmib_sg_adam.argtypes = [vp, vp, vp, vp, vp, l, f, f, f, f, i, vp]
    L.lbmib_sg_pack.argtypes = [vp, i, dp, dp, i, dp, vp, l, i, vp]
    L.lbmib_sg_target.argtypes = [vp, vp, i, dp, dp, vp, l, vp]
    L.lbmib_sg_unpack.argtypes = [vp, i, i, dp, dp, vp, l, vp]
    L.lbmib_sg_sqdiff.argtypes = [vp, vp, l, vp, vp]
    L.lbmib_sg_cast_bf16.argtypes = [vp, vp, l, vp]
    for n in SG_SYMBOLS:
        getattr(L, n).restype = i
    L._sg_bound = True
    return L


SG_SYMBOLS = ("lbmib_sg_conv3d_fwd", "lbmib_sg_conv3d_bwd_data", "lbmib_sg_conv3d_bwd_weight", "lbmib_sg_pool2",
              "lbmib_sg_up2_add", "lbmib_sg_pool2_bwd", "lbmib_sg_up2_bwd", "lbmib_sg_mse", "lbmib_sg_adam",
              "lbmib_sg_pack", "lbmib_sg_target", "lbmib_sg_unpack", "lbmib_sg_sqdiff", "lbmib_sg_cast_bf16",
              "lbmib_sg_conv3d_bwd_weight_noalloc")


def check(rc, where):
    if rc != 0:
        raise lbmib.LbmibError(rc, where)


def ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def dvec(vals):
    a = (ctypes.c_double * max(1, len(vals)))(*[float(v) for v in vals])
    return a


def layer_list(cin, cout, chans=CHANS):
    """(name, cin, cout, level, relu) of every convolution: two per encoder
    level, an upward convolution and a merge convolution per decoder level,
    one output convolution."""
    L = []
    prev = cin
    for lvl, c in enumerate(chans):
        L.append((f"enc{lvl}a", prev, c, lvl, True))
        L.append((f"enc{lvl}b", c, c, lvl, True))
        prev = c
    for lvl in range(len(chans) - 2, -1, -1):
        L.append((f"dec{lvl}up", chans[lvl + 1], chans[lvl], lvl + 1, True))
        L.append((f"dec{lvl}b", chans[lvl], chans[lvl], lvl, True))
    L.append(("out", chans[0], cout, 0, False))
    return L


def conv_flops(layers, B, D, H, W):
    """Multiply-add FLOPs (2 per MAC) of one forward pass."""
    tot = 0
    for _, ci, co, lvl, _ in layers:
        vox = B * (D >> lvl) * (H >> lvl) * (W >> lvl)
        tot += 2 * vox * 27 * ci * co
    return tot


class Net:
    """The encoder-decoder on the GPU: FP32 master weights, BF16 copies
    read by the convolutions, FP32 gradients and Adam moments.

    cin / cout: input / output channels (multiples of 16); nout: channels
    that the loss covers.  Weights are He-normal for the ReLU layers; the
    output layer starts at a small scale so that the untrained network's
    prediction is close to zero change."""

    def __init__(self, cin=16, cout=16, nout=4, chans=CHANS, seed=0, out_scale=1e-3):
        self.L = _bind()
        self.cin, self.cout, self.nout, self.chans = cin, cout, nout, tuple(chans)
        self.layers = layer_list(cin, cout, chans)
        rng = np.random.default_rng(seed)
        dev = torch.device("cuda")
        self.p, self.w16, self.g, self.m, self.v = {}, {}, {}, {}, {}
        self.b, self.gb, self.mb, self.vb = {}, {}, {}, {}
        self.wt = {}
        for name, ci, co, _, relu in self.layers:
            std = np.sqrt(2.0 / (27 * ci)) if relu else out_scale
            w = rng.standard_normal((co, 27, ci)) * std
            self.p[name] = torch.tensor(w, dtype=torch.float32, device=dev)
            self.b[name] = torch.zeros(co, dtype=torch.float32, device=dev)
            for d in (self.g, self.m, self.v):
                d[name] = torch.zeros_like(self.p[name])
            for d in (self.gb, self.mb, self.vb):
                d[name] = torch.zeros_like(self.b[name])
            self.w16[name] = torch.empty((co, 27, ci), dtype=torch.bfloat16, device=dev)
            self._cast(self.p[name], self.w16[name])
            self.wt[name] = torch.empty((ci, 27, co), dtype=torch.bfloat16, device=dev)
        self.t = 0
        self._bufs = {}
        self.timing = None  # {(op, layer): [event pairs]} while bench() times the layers

    def _cast(self, src, dst):
        check(self.L.lbmib_sg_cast_bf16(ptr(src), ptr(dst), src.numel(), stream()), "cast_bf16")

    # ------------------------------------------------------------ buffers
    def _buf(self, key, shape):
        b = self._bufs.get(key)
        if b is None or tuple(b.shape) != tuple(shape):
            b = torch.empty(shape, dtype=torch.bfloat16, device="cuda")
            self._bufs[key] = b
        return b

    def n_params(self):
        return sum(p.numel() + self.b[k].numel() for k, p in self.p.items())

    # ------------------------------------------------------------ forward
    def _timed(self, op, name, fn):
        if self.timing is None:
            return fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        self.timing.setdefault((op, name), []).append((e0, e1))

    def _conv(self, name, x, y, dims, relu):
        _, ci, co, _, _ = next(l for l in self.layers if l[0] == name)
        B, D, H, W = dims
        self._timed("fwd", name, lambda: check(
            self.L.lbmib_sg_conv3d_fwd(ptr(x), ptr(self.w16[name]), ptr(self.b[name]), None, ptr(y), B, D, H, W, ci,
                                       co, int(relu), stream()), "conv3d_fwd " + name))

    def forward(self, x):
        """x: BF16 [B, D, H, W, cin] on the GPU (D, H, W divisible by
        2^(levels-1)).  Returns the BF16 output [B, D, H, W, cout]; keeps the
        activations for backward()."""
        B, D, H, W, C = x.shape
        assert C == self.cin and x.dtype == torch.bfloat16
        n = len(self.chans)
        assert D % (1 << (n - 1)) == 0 and H % (1 << (n - 1)) == 0 and W % (1 << (n - 1)) == 0
        self.dims = (B, D, H, W)
        a = {"x": x}
        h = x
        for lvl, c in enumerate(self.chans):
            dims = (B, D >> lvl, H >> lvl, W >> lvl)
            if lvl > 0:
                pin = self._buf(f"pool{lvl}", dims + (self.chans[lvl - 1],))
                check(self.L.lbmib_sg_pool2(ptr(a[f"enc{lvl - 1}b"]), ptr(pin), *dims, self.chans[lvl - 1], stream()),
                      "pool2")
                a[f"pool{lvl}"] = h = pin
            ya = self._buf(f"enc{lvl}a", dims + (c,))
            self._conv(f"enc{lvl}a", h, ya, dims, True)
            yb = self._buf(f"enc{lvl}b", dims + (c,))
            self._conv(f"enc{lvl}b", ya, yb, dims, True)
            a[f"enc{lvl}a"], a[f"enc{lvl}b"] = ya, yb
            h = yb
        d = h
        for lvl in range(n - 2, -1, -1):
            cdims = (B, D >> (lvl + 1), H >> (lvl + 1), W >> (lvl + 1))
            fdims = (B, D >> lvl, H >> lvl, W >> lvl)
            c = self._buf(f"dec{lvl}up", cdims + (self.chans[lvl],))
            self._conv(f"dec{lvl}up", d, c, cdims, True)
            s = self._buf(f"dec{lvl}s", fdims + (self.chans[lvl],))
            check(self.L.lbmib_sg_up2_add(ptr(c), ptr(a[f"enc{lvl}b"]), ptr(s), *fdims, self.chans[lvl], stream()),
                  "up2_add")
            db = self._buf(f"dec{lvl}b", fdims + (self.chans[lvl],))
            self._conv(f"dec{lvl}b", s, db, fdims, True)
            a[f"dec{lvl}up"], a[f"dec{lvl}s"], a[f"dec{lvl}b"] = c, s, db
            d = db
        out = self._buf("out", (B, D, H, W, self.cout))
        self._conv("out", d, out, (B, D, H, W), False)
        a["out"] = out
        self.act = a
        return out

    # ----------------------------------------------------------- backward
    def _wgrad(self, name, x, dy, dims):
        _, ci, co, _, _ = next(l for l in self.layers if l[0] == name)
        self._timed("wgrad", name, lambda: check(
            self.L.lbmib_sg_conv3d_bwd_weight(ptr(x), ptr(dy), ptr(self.g[name]), ptr(self.gb[name]), *dims, ci, co,
                                              stream()), "conv3d_bwd_weight " + name))

    def _dgrad(self, name, dy, dx, mask, dims):
        _, ci, co, _, _ = next(l for l in self.layers if l[0] == name)
        self._timed("dgrad", name, lambda: check(
            self.L.lbmib_sg_conv3d_bwd_data(ptr(dy), ptr(self.w16[name]), ptr(self.wt[name]), ptr(mask), ptr(dx),
                                            *dims, ci, co, stream()), "conv3d_bwd_data " + name))

    def backward(self, dout):
        """Parameter gradients (FP32, overwritten) for the output gradient
        dout (BF16, the shape of forward()'s output)."""
        a = self.act
        B, D, H, W = self.dims
        n = len(self.chans)
        dims = lambda lvl: (B, D >> lvl, H >> lvl, W >> lvl)  # noqa: E731
        self._wgrad("out", a["dec0b"], dout, dims(0))
        g = self._buf("g0", dims(0) + (self.chans[0],))
        self._dgrad("out", dout, g, a["dec0b"], dims(0))
        dskip = {}
        for lvl in range(0, n - 1):
            c = self.chans[lvl]
            self._wgrad(f"dec{lvl}b", a[f"dec{lvl}s"], g, dims(lvl))
            ds = self._buf(f"ds{lvl}", dims(lvl) + (c,))
            self._dgrad(f"dec{lvl}b", g, ds, None, dims(lvl))
            dskip[lvl] = ds
            gc = self._buf(f"gc{lvl}", dims(lvl + 1) + (c,))
            check(self.L.lbmib_sg_up2_bwd(ptr(ds), ptr(a[f"dec{lvl}up"]), ptr(gc), *dims(lvl + 1), c, stream()),
                  "up2_bwd")
            below = a[f"dec{lvl + 1}b"] if lvl + 1 < n - 1 else a[f"enc{n - 1}b"]
            self._wgrad(f"dec{lvl}up", below, gc, dims(lvl + 1))
            g = self._buf(f"g{lvl + 1}", dims(lvl + 1) + (self.chans[lvl + 1],))
            self._dgrad(f"dec{lvl}up", gc, g, below, dims(lvl + 1))
        gpool = None
        for lvl in range(n - 1, -1, -1):
            c = self.chans[lvl]
            if lvl < n - 1:
                ge = self._buf(f"ge{lvl}", dims(lvl) + (c,))
                check(self.L.lbmib_sg_pool2_bwd(ptr(dskip[lvl]), ptr(gpool), ptr(a[f"enc{lvl}b"]), ptr(ge), *dims(lvl),
                                                c, stream()), "pool2_bwd")
                g = ge
            self._wgrad(f"enc{lvl}b", a[f"enc{lvl}a"], g, dims(lvl))
            ga = self._buf(f"ga{lvl}", dims(lvl) + (c,))
            self._dgrad(f"enc{lvl}b", g, ga, a[f"enc{lvl}a"], dims(lvl))
            inp = a["x"] if lvl == 0 else a[f"pool{lvl}"]
            self._wgrad(f"enc{lvl}a", inp, ga, dims(lvl))
            if lvl > 0:
                gpool = self._buf(f"gp{lvl}", dims(lvl) + (self.chans[lvl - 1],))
                self._dgrad(f"enc{lvl}a", ga, gpool, None, dims(lvl))

    # ----------------------------------------------------------- training
    def loss_grad(self, out, tgt):
        """MSE over the first nout channels against tgt (FP32 [M, nout]);
        returns (device loss tensor, BF16 output gradient)."""
        M = out.numel() // self.cout
        dout = self._buf("dout", tuple(out.shape))
        loss = self._bufs.get("loss")
        if loss is None:
            loss = self._bufs["loss"] = torch.zeros(1, dtype=torch.float32, device="cuda")
        check(self.L.lbmib_sg_mse(ptr(out), ptr(tgt), ptr(dout), ptr(loss), M, self.cout, self.nout, stream()), "mse")
        return loss, dout

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        for name in self.p:
            check(self.L.lbmib_sg_adam(ptr(self.p[name]), ptr(self.g[name]), ptr(self.m[name]), ptr(self.v[name]),
                                       ptr(self.w16[name]), self.p[name].numel(), lr, b1, b2, eps, self.t, stream()),
                  "adam")
            check(self.L.lbmib_sg_adam(ptr(self.b[name]), ptr(self.gb[name]), ptr(self.mb[name]), ptr(self.vb[name]),
                                       None, self.b[name].numel(), lr, b1, b2, eps, self.t, stream()), "adam bias")

    def train_step(self, x, tgt, lr):
        out = self.forward(x)
        loss, dout = self.loss_grad(out, tgt)
        self.backward(dout)
        self.adam_step(lr)
        return loss

    # ------------------------------------------------------------ weights
    def state(self, bf16=False):
        """Weights and biases as numpy arrays; bf16=True stores the weights as
        the BF16 bit patterns (uint16) the convolutions read, half the size
        and the same inference results."""
        st = {}
        torch.cuda.synchronize()
        for name in self.p:
            if bf16:
                st[name + ".w"] = self.w16[name].view(torch.int16).cpu().numpy().view(np.uint16)
            else:
                st[name + ".w"] = self.p[name].cpu().numpy()
            st[name + ".b"] = self.b[name].cpu().numpy()
        return st

    def load_state(self, st):
        for name in self.p:
            w = np.asarray(st[name + ".w"])
            if w.dtype == np.uint16:  # BF16 bit patterns
                w = (w.astype(np.uint32) << 16).view(np.float32)
            self.p[name].copy_(torch.from_numpy(np.ascontiguousarray(w, dtype=np.float32)))
            self.b[name].copy_(torch.from_numpy(np.asarray(st[name + ".b"], dtype=np.float32)))
            self._cast(self.p[name], self.w16[name])
        torch.cuda.synchronize()


def bench(batch=4, edge=64, warm=3, iters=10, predict_batch=1):
    """Times the network's kernels on synthetic data: the training step
    (forward, MSE, data and weight gradients, Adam) at batch x edge^3 and the
    forward pass at predict_batch x edge^3 (the coupling's prediction), with
    HIP events on the library's stream.  FLOPs are the convolutions' 2 per
    multiply-add: the forward pass F, a training step 3F (data and weight
    gradients each cost one forward's MACs).  Returns a dict."""
    net = Net()
    rng = np.random.default_rng(1)
    x = torch.tensor(rng.standard_normal((batch, edge, edge, edge, net.cin)), dtype=torch.bfloat16, device="cuda")
    tgt = torch.tensor(rng.standard_normal((batch * edge ** 3, net.nout)) * 1e-3, dtype=torch.float32, device="cuda")
    xp = x[:predict_batch].contiguous()
    out = {"batch": batch, "edge": edge, "params": net.n_params()}
    for name, fn, b, nflop in (("train_step", lambda: net.train_step(x, tgt, 1e-4), batch, 3),
                               ("forward", lambda: net.forward(xp), predict_batch, 1)):
        for _ in range(warm):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(iters):
            fn()
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1) / iters
        flops = nflop * conv_flops(net.layers, b, edge, edge, edge)
        out[name] = {"batch": b, "ms": round(ms, 4), "conv_gflop": round(flops / 1e9, 2),
                     "tflops": round(flops / (ms * 1e-3) / 1e12, 1)}
    # per layer and operation: one more pass of each with events around every
    # convolution call (the calls' own launches, split-K reductions included)
    for tag, fn, b in (("train", lambda: net.train_step(x, tgt, 1e-4), batch),
                       ("predict", lambda: net.forward(xp), predict_batch)):
        net.timing = {}
        for _ in range(iters):
            fn()
        torch.cuda.synchronize()
        rows = []
        for (op, lname), evs in net.timing.items():
            _, ci, co, lvl, _ = next(l for l in net.layers if l[0] == lname)
            ms = sum(e0.elapsed_time(e1) for e0, e1 in evs) / len(evs)
            fl = 2 * b * (edge >> lvl) ** 3 * 27 * ci * co
            rows.append({"op": op, "layer": lname, "cin": ci, "cout": co, "grid": edge >> lvl, "us": round(ms * 1e3, 1),
                         "tflops": round(fl / (ms * 1e-3) / 1e12, 1)})
        out[f"layers_{tag}"] = rows
        net.timing = None
    return out


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="time the surrogate network's kernels on synthetic data")
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--edge", type=int, default=64)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--predict-batch", type=int, default=1)
    args = ap.parse_args()
    print(json.dumps(bench(args.batch, args.edge, iters=args.iters, predict_batch=args.predict_batch)), flush=True)
