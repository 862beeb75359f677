"""Training data and GPU training of the surrogate (surrogate.Net).

Not in the reference (cube_pthread/lbm_ib.c has no learned model).
Trajectories come from the batched solver (lbmib.Batch): instances of the
256^3 workload's geometry -- inlet/outlet x, periodic y, walls z, one
104 x 104 sheet centred at (64, 128, 128) -- with tau, the sheet stiffness
Ks = Kb and the inflow speed drawn from SplitMix64(seed).  Every `every`
steps the coarse state of each instance (surrogate_jump.Coarse: rho, u of
the stored populations, node count, elastic force, on 64^3 at f = 4) is kept
on the GPU; a training pair
is (state at t, change of rho/u over K steps).  Training runs the library's
own forward, backward and Adam kernels.

Usage (from the package directory, on a GPU):
  python surrogate_train.py --trajectories 96 --steps 1000 --out surrogate_c5.npz
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import lbmib  # noqa: E402
import lbmib_configs  # noqa: E402
import surrogate as sg  # noqa: E402
import surrogate_jump as sj  # noqa: E402

TAU_RANGE = (0.55, 0.75)
K_RANGE = (5e-4, 2e-3)        # Ks = Kb, log-uniform
U_RANGE = (0.006, 0.014)      # inflow speed


def training_configs(n, seed, grid=256, rows=104, precision="f32"):
    """n instance configurations (Engine keyword arguments) with seeded tau,
    stiffness and inflow speed; the sheet jitter seed differs per instance."""
    u = lbmib.uniform_stream(seed)
    out = []
    for i in range(n):
        tau = TAU_RANGE[0] + (TAU_RANGE[1] - TAU_RANGE[0]) * next(u)
        k = 10.0 ** (math.log10(K_RANGE[0]) + (math.log10(K_RANGE[1]) - math.log10(K_RANGE[0])) * next(u))
        u_in = U_RANGE[0] + (U_RANGE[1] - U_RANGE[0]) * next(u)
        sheet = lbmib.Sheet.centred(rows, rows, 1.0, (grid / 4.0, grid / 2.0, grid / 2.0), k, k, jitter=0.1,
                                    seed=seed * 7919 + i)
        out.append(dict(nx=grid, ny=grid, nz=grid, bc=lbmib_configs.BASE_BC, precision=precision, tau=tau, rho0=1.0,
                        u0=(u_in, 0.0, 0.0), quirks=True, sheets=[sheet]))
    return out


class Trajectory:
    def __init__(self, snaps, tau, u_in):
        self.snaps, self.tau, self.u_in = snaps, tau, u_in  # snaps: [S][NF][nc] FP32 on the GPU


def generate(cfgs, steps, every, factor=4, batch=8, log=print):
    """Runs the instances on the batched solver, `batch` at a time, and
    returns one Trajectory per instance with snapshots at t = 1, 1 + every,
    ... (t = 1: after the first step)."""
    nsnap = steps // every + 1
    trajs = []
    for c0 in range(0, len(cfgs), batch):
        chunk = cfgs[c0:c0 + batch]
        t0 = time.time()
        with lbmib.Batch(chunk) as b:
            b.init()
            b.step(1)
            cos = [sj.Coarse(e, factor) for e in b.instances]
            snaps = [torch.empty((nsnap, sj.NF, cos[0].nc), dtype=torch.float32, device="cuda") for _ in chunk]
            for s in range(nsnap):
                for i, co in enumerate(cos):
                    co.read(snaps[i][s])
                if s + 1 < nsnap:
                    b.step(every)
            st = b.status()
            if st != 0:
                raise lbmib.LbmibError(st, "trajectory batch")
        for i, c in enumerate(chunk):
            trajs.append(Trajectory(snaps[i], c["tau"], c["u0"][0]))
        log(f"[data] trajectories {c0 + len(chunk)}/{len(cfgs)}: {len(chunk)} x {steps} steps in "
            f"{time.time() - t0:.1f} s")
    torch.cuda.synchronize()
    return trajs


def norm_from(trajs, lag, nsample=64, seed=0):
    """Per-plane mean and std of the inputs and of the K-step changes, from
    host copies of a sample of snapshots."""
    rng = np.random.default_rng(seed)
    xs, ds = [], []
    for _ in range(nsample):
        tr = trajs[int(rng.integers(len(trajs)))]
        i = int(rng.integers(tr.snaps.shape[0] - lag))
        a = tr.snaps[i].cpu().numpy().astype(np.float64)
        b = tr.snaps[i + lag, :sj.NOUT].cpu().numpy().astype(np.float64)
        xs.append(a)
        ds.append(b - a[:sj.NOUT])
    x = np.stack(xs, 1).reshape(sj.NF, -1)
    d = np.stack(ds, 1).reshape(sj.NOUT, -1)
    in_std = np.maximum(x.std(axis=1), 1e-12)
    out_std = np.maximum(d.std(axis=1), 1e-12)
    # the three velocity planes share one scale (the RMS of their spreads),
    # in the input and in the change, so that the loss weighs velocity errors
    # as the velocity L2 error does (a component that barely varies is not
    # blown up to unit size)
    in_std[1:4] = np.sqrt(np.mean(in_std[1:4] ** 2))
    out_std[1:4] = np.sqrt(np.mean(out_std[1:4] ** 2))
    out_mean = d.mean(axis=1)
    return sj.Norm(x.mean(axis=1), in_std, out_mean, out_std)


class Loader:
    """Packs batches of (input, target) pairs on the GPU with the library's
    pack / target kernels."""

    def __init__(self, trajs, norm, lag, batch, dims):
        self.trajs, self.norm, self.lag, self.B = trajs, norm, lag, batch
        self.nz, self.ny, self.nx = dims
        self.nc = self.nz * self.ny * self.nx
        self.x = torch.empty((batch, self.nz, self.ny, self.nx, sj.CIN), dtype=torch.bfloat16, device="cuda")
        self.tgt = torch.empty((batch * self.nc, sj.NOUT), dtype=torch.float32, device="cuda")
        self.L = sg._bind()

    def pairs(self):
        return [(t, i) for t in range(len(self.trajs)) for i in range(self.trajs[t].snaps.shape[0] - self.lag)]

    def fill(self, picks):
        s = sg.stream()
        istd_out = [1.0 / v for v in self.norm.out_std]
        xb = self.x.element_size() * self.nc * sj.CIN
        tb = self.tgt.element_size() * self.nc * sj.NOUT
        for b, (t, i) in enumerate(picks):
            tr = self.trajs[t]
            consts = sj.const_channels(tr.tau, tr.u_in)
            xp = ctypes.c_void_p(self.x.data_ptr() + b * xb)
            istd = [1.0 / v for v in self.norm.in_std]
            sg.check(self.L.lbmib_sg_pack(ctypes.c_void_p(tr.snaps[i].data_ptr()), sj.NF, sg.dvec(self.norm.in_mean),
                                          sg.dvec(istd), sj.NCONST, sg.dvec(consts), xp, self.nc, sj.CIN, s), "pack")
            sg.check(self.L.lbmib_sg_target(ctypes.c_void_p(tr.snaps[i].data_ptr()),
                                            ctypes.c_void_p(tr.snaps[i + self.lag].data_ptr()), sj.NOUT,
                                            sg.dvec(self.norm.out_mean), sg.dvec(istd_out),
                                            ctypes.c_void_p(self.tgt.data_ptr() + b * tb), self.nc, s), "target")
        return self.x, self.tgt


def validate(net, loader, pairs):
    """Relative L2 of the predicted change against the true change, over
    held-out pairs (normalised planes): sqrt(sum (pred - true)^2 / sum true^2);
    the zero prediction scores about 1."""
    num = den = 0.0
    L = sg._bind()
    out2 = torch.zeros(2, dtype=torch.float64, device="cuda")
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    dout = None
    for k in range(0, len(pairs) - loader.B + 1, loader.B):
        x, tgt = loader.fill(pairs[k:k + loader.B])
        out = net.forward(x)
        if dout is None:
            dout = torch.empty_like(out)
        n = tgt.numel()
        sg.check(L.lbmib_sg_mse(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(tgt.data_ptr()),
                                ctypes.c_void_p(dout.data_ptr()), ctypes.c_void_p(loss.data_ptr()),
                                loader.B * loader.nc, 16, sj.NOUT, sg.stream()), "mse")
        sg.check(L.lbmib_sg_sqdiff(ctypes.c_void_p(tgt.data_ptr()), None, n, ctypes.c_void_p(out2.data_ptr()),
                                   sg.stream()), "sqdiff")
        num += float(loss.item()) * n
        den += float(out2.cpu().numpy()[0])
    return math.sqrt(num / max(den, 1e-300))


def train(net, trajs, norm, lag, iters, batch, lr, dims, val_trajs=(), seed=0, log=print):
    loader = Loader(trajs, norm, lag, batch, dims)
    pairs = loader.pairs()
    rng = np.random.default_rng(seed)
    t0 = time.time()
    hist = []
    for it in range(iters):
        picks = [pairs[int(j)] for j in rng.integers(0, len(pairs), batch)]
        x, tgt = loader.fill(picks)
        # cosine decay to 5% of lr
        lr_t = lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * it / max(1, iters - 1))))
        loss = net.train_step(x, tgt, lr_t)
        if it % 50 == 0 or it == iters - 1:
            lv = float(loss.item())
            hist.append((it, lv))
            log(f"[train] iter {it} loss {lv:.4e} lr {lr_t:.2e} ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()
    res = {"iters": iters, "batch": batch, "seconds": time.time() - t0, "loss_history": hist}
    if val_trajs:
        vl = Loader(list(val_trajs), norm, lag, batch, dims)
        res["val_rel_l2"] = validate(net, vl, vl.pairs())
    return res


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--trajectories", type=int, default=96)
    p.add_argument("--val", type=int, default=4)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--every", type=int, default=25)
    p.add_argument("--ksteps", type=int, default=50)
    p.add_argument("--factor", type=int, default=4)
    p.add_argument("--batch", type=int, default=4)
    p.add_argument("--gen-batch", type=int, default=8)
    p.add_argument("--iters", type=int, default=1500)
    p.add_argument("--lr", type=float, default=1e-3)
    p.add_argument("--seed", type=int, default=2026)
    p.add_argument("--out", default=os.path.join(HERE, "surrogate_c5.npz"))
    p.add_argument("--log", default=None)
    a = p.parse_args()
    logf = open(a.log, "w") if a.log else None

    def log(msg):
        print(msg, flush=True)
        if logf:
            logf.write(msg + "\n")
            logf.flush()

    lag = a.ksteps // a.every
    cfgs = training_configs(a.trajectories + a.val, a.seed)
    trajs = generate(cfgs, a.steps, a.every, a.factor, a.gen_batch, log)
    train_t, val_t = trajs[:a.trajectories], trajs[a.trajectories:]
    norm = norm_from(train_t, lag)
    log("[norm] " + json.dumps({k: v.tolist() for k, v in norm.arrays().items()}))
    net = sg.Net(cin=sj.CIN, cout=16, nout=sj.NOUT, seed=a.seed)
    dims = (256 // a.factor,) * 3
    vl = Loader(list(val_t), norm, lag, a.batch, dims)
    before = validate(net, vl, vl.pairs()) if val_t else None
    res = train(net, train_t, norm, lag, a.iters, a.batch, a.lr, dims, val_t, a.seed, log)
    res.update({"trajectories": a.trajectories, "val_trajectories": a.val, "steps": a.steps, "every": a.every,
                "ksteps": a.ksteps, "factor": a.factor, "val_rel_l2_untrained": before,
                "params": net.n_params()})
    log("[result] " + json.dumps(res))
    sj.save_model(a.out, net, norm, a.factor, a.ksteps,
                  {"train_trajectories": a.trajectories, "train_steps": a.steps, "seed": a.seed})
    log(f"[saved] {a.out}")


if __name__ == "__main__":
    main()
