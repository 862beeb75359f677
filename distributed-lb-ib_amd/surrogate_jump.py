"""Surrogate-accelerated stepping of a running simulation: jumps of K steps
predicted by the encoder-decoder (surrogate.Net), each re-verified by the
plain solver, with fallback to plain steps when the verification fails.

Not in the reference: cube_pthread/lbm_ib.c runs every step of its do_thread
loop (lbm_ib.c:3115-3563).  One jump cycle here:
  1. the coarse state at t: rho/u of the stored populations (the state
     compute_rho_and_u, lbm_ib.c:2440-2538, reads one streaming later)
     averaged over f^3 blocks and the sheet raster (node count, elastic
     force sums) -- lbmib_coarse_state's stored planes / lbmib_coarse_sheets;
  2. the network predicts the change of the four fluid planes over K steps;
  3. lbmib_jump shifts every cell's equilibrium by the interpolated change,
     advects the sheet, keeps the pre-jump state aside and sets t += K;
  4. re-verification: R plain steps (they count as simulated steps); the
     change of the coarse velocity over them (the jump reports the jumped
     state's, and the read after the steps -- the next cycle's input -- the
     new one, both from passes the cycle makes anyway), per step, divided by
     the jump's own mean change per step, is the verification ratio.  A
     state the plain dynamics accept changes at about the rate the
     prediction implies; a jump that leaves it inconsistent (pressure waves,
     a broken non-equilibrium part) changes much faster.  Ratio <= tol: commit;
     otherwise lbmib_jump_revert and K + R plain steps from the checkpoint,
     which reproduce the plain trajectory bit for bit.

Every kernel runs on the engine's stream (with early=True the network's
next prediction runs on a stream of its own, ordered by events, beside the
verification steps); PyTorch only allocates memory.
"""
import ctypes

import numpy as np
import torch

import lbmib
import surrogate as sg

NF = 8        # feature planes: rho, ux, uy, uz, node count, elastic force x, y, z
NOUT = 4      # predicted planes: the change of rho, ux, uy, uz
NCONST = 2    # constant channels: viscosity and inflow speed
CIN = 16      # network input channels (NF + NCONST, zero padded)


def const_channels(tau, u_in):
    """Viscosity (tau - 1/2)/3 and inflow speed, centred and scaled to O(1)."""
    nu = (tau - 0.5) / 3.0
    return [(nu - 0.05) / 0.03, (u_in - 0.01) / 0.004]


class Norm:
    """Per-plane shift and scale of the inputs and of the predicted change."""

    def __init__(self, in_mean, in_std, out_mean, out_std):
        self.in_mean = [float(v) for v in in_mean]
        self.in_std = [float(v) for v in in_std]
        self.out_mean = [float(v) for v in out_mean]
        self.out_std = [float(v) for v in out_std]

    def arrays(self):
        return {"in_mean": np.array(self.in_mean), "in_std": np.array(self.in_std),
                "out_mean": np.array(self.out_mean), "out_std": np.array(self.out_std)}

    @classmethod
    def from_arrays(cls, a):
        return cls(a["in_mean"], a["in_std"], a["out_mean"], a["out_std"])


def load_model(path):
    """Net and Norm from an .npz written by surrogate_train (no pickle)."""
    with np.load(path, allow_pickle=False) as z:
        st = {k: z[k] for k in z.files}
    chans = tuple(int(c) for c in st.pop("chans"))
    factor = int(st.pop("factor"))
    ksteps = int(st.pop("ksteps"))
    norm = Norm.from_arrays(st)
    net = sg.Net(cin=CIN, cout=16, nout=NOUT, chans=chans)
    net.load_state(st)
    tol = float(st["verify_tol"]) if "verify_tol" in st else None
    return net, norm, factor, ksteps, tol


def save_model(path, net, norm, factor, ksteps, extra=None):
    """Weights as BF16 bit patterns plus the normalisation and the jump's
    constants, in a compressed .npz (numpy arrays only)."""
    st = net.state(bf16=True)
    st.update(norm.arrays())
    st["chans"] = np.array(net.chans)
    st["factor"] = np.array(factor)
    st["ksteps"] = np.array(ksteps)
    for k, v in (extra or {}).items():
        st[k] = np.asarray(v)
    np.savez_compressed(path, **st)


class Coarse:
    """Device feature planes [NF][nc] of one engine at coarsening factor f."""

    def __init__(self, engine, factor):
        self.e, self.f = engine, factor
        self.dims = (engine.nz // factor, engine.ny // factor, engine.nx // factor)
        self.nc = int(np.prod(self.dims))
        self.feat = torch.empty((NF, self.nc), dtype=torch.float32, device="cuda")

    def read(self, out=None):
        """Fills out (default: self.feat) with the engine's current state:
        planes 0-3 the coarse means of the stored moments (rho, and
        u = sum c f / rho on the updated cells: lbmib_coarse_state's stored
        planes, the convention lbmib_jump reports a jumped state in, read
        without the neighbours' populations), planes 4-7 the sheet raster."""
        t = self.feat if out is None else out
        self.e.coarse_moments(self.f, 0, t[0].data_ptr())
        self.e.coarse_sheets(self.f, t[4].data_ptr())
        return t

    def step_read(self):
        """One plain step whose fluid launch also writes planes 0-3 of the
        new state (lbmib_step_coarse: the same bits as a step and then
        read(), without the read's pass over the populations), then the
        sheet raster."""
        self.e.step_coarse(self.f, self.feat[0].data_ptr())
        self.e.coarse_sheets(self.f, self.feat[4].data_ptr())
        return self.feat


def pack_input(feat, norm, consts, x, nc, stream):
    """Network input [.., CIN] (BF16) of feature planes feat [NF][nc]."""
    L = sg._bind()
    istd = [1.0 / s for s in norm.in_std]
    sg.check(L.lbmib_sg_pack(ctypes.c_void_p(feat.data_ptr()), NF, sg.dvec(norm.in_mean), sg.dvec(istd), NCONST,
                             sg.dvec(consts), ctypes.c_void_p(x.data_ptr()), nc, CIN, stream), "pack")


class JumpRunner:
    """Steps one engine with surrogate jumps (see the module docstring).

    tol: acceptance bound of the verification ratio; 0 rejects every jump
    (all plain steps, bit-identical to lbmib_step)."""

    def __init__(self, engine, net, norm, tau, u_in, factor=4, ksteps=50, verify_steps=1, tol=4.0, graph=True,
                 early=False):
        self.e, self.net, self.norm = engine, net, norm
        self.K, self.R, self.tol = ksteps, verify_steps, tol
        self.consts = const_channels(tau, u_in)
        self.stream_handle = engine.stream_handle()
        self.stream = ctypes.c_void_p(self.stream_handle)
        self.ext = torch.cuda.ExternalStream(self.stream_handle)
        self.co = Coarse(engine, factor)
        nz, ny, nx = self.co.dims
        self.x = torch.empty((1, nz, ny, nx, CIN), dtype=torch.bfloat16, device="cuda")
        self.delta = torch.empty((NOUT, self.co.nc), dtype=torch.float32, device="cuda")
        # early: the next cycle's prediction reads the jumped state (its
        # planes 0-3 as lbmib_jump reports them, the raster of the moved
        # sheet) and runs on a stream of its own beside the verification
        # steps; the feature planes the network reads are then a buffer of
        # their own, into which the jump writes its d_after
        self.early = early
        if early:
            self.pfeat = torch.empty((NF, self.co.nc), dtype=torch.float32, device="cuda")
            self.pred_stream = torch.cuda.Stream()
            self.pred_done = torch.cuda.Event()
            self.pred_pending = False
        else:
            self.pfeat = self.co.feat
        # the jumped state's coarse planes (lbmib_jump's d_after, the
        # convention of the feature planes 0-3)
        self.after = self.pfeat[0:4] if early else torch.empty((4, self.co.nc), dtype=torch.float32,
                                                               device="cuda")
        self.sums = torch.zeros(4, dtype=torch.float64, device="cuda")
        # the verification sums reach the host through a side stream, so that
        # the engine's stream goes on with the next cycle's prediction while
        # the host waits for them (the page-locked buffer is only ever used
        # on that torch-owned stream)
        self.side = torch.cuda.Stream()
        self.host_sums = torch.empty(4, dtype=torch.float64, pin_memory=True)
        self.sums_ready = torch.cuda.Event()
        self.stats = {"jumps": 0, "accepted": 0, "rejected": 0, "plain_steps": 0, "ratios": []}
        self.have_feat = False
        self.pred_ready = False  # self.delta holds the network's prediction from self.co.feat
        self.use_graph, self.graph = graph, None

    def _predict_kernels(self):
        L = sg._bind()
        pack_input(self.pfeat, self.norm, self.consts, self.x, self.co.nc, self.stream)
        out = self.net.forward(self.x)
        sg.check(L.lbmib_sg_unpack(ctypes.c_void_p(out.data_ptr()), 16, NOUT, sg.dvec(self.norm.out_mean),
                                   sg.dvec(self.norm.out_std), ctypes.c_void_p(self.delta.data_ptr()), self.co.nc,
                                   self.stream), "unpack")

    def predict(self):
        """delta = the network's change of (rho, ux, uy, uz) from self.co.feat.

        The first call runs the kernels one by one (it also sizes the
        network's buffers and the library workspace); with graph=True the
        second call records the same launches -- packing, the convolutions,
        unpacking, all on the engine's stream, reading and writing only
        buffers that stay in place -- into a HIP graph, and every call from
        then on replays it."""
        with torch.cuda.stream(self.ext):
            if self.early and self.pfeat is not self.co.feat:
                self.pfeat.copy_(self.co.feat)
            if self.graph is not None:
                self.graph.replay()
            elif not self.use_graph or not getattr(self, "_eager_done", False):
                self._predict_kernels()
                self._eager_done = True
            else:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=self.ext):
                    self._predict_kernels()
                self.graph = g
                g.replay()
        return self.delta

    def prepare(self, predictor=None):
        """One-time set-up, outside any timed region: with the network
        (predictor None), two predictions -- the first sizes its buffers, the
        second records the graph -- a jump of the current state, reverted at
        once, which allocates the engine's checkpoint buffers, and one pass
        of the verification sums to the host.  The engine's state is left
        bit for bit as it was; the first cycle reads and predicts afresh."""
        with torch.cuda.stream(self.ext):
            self.co.read()
            if predictor is None:
                self.predict()
                self.predict()
            else:
                predictor(self)
            self.e.jump(self.co.f, self.co.feat.data_ptr(), self.delta.data_ptr(), self.K, self.after.data_ptr())
            self.e.jump_revert()
            # the verification sums and their way to the host, once
            self._sqdiff(self.co.feat[1:4], self.after[1:4], self.sums[0:2])
            self._sqdiff(self.delta[1:4], None, self.sums[2:4])
            self._sums_to_host()
        self.sums_ready.synchronize()
        self.have_feat = self.pred_ready = False

    def _sums_to_host(self):
        """Copies self.sums to the page-locked buffer on the side stream
        after the engine's stream has produced them (a pinned buffer used on
        the engine's stream would have torch's allocator record an event on
        that stream when the buffer is freed, which may be after the engine
        has destroyed it); self.sums_ready marks the copy's end."""
        done = torch.cuda.Event()
        done.record(self.ext)
        self.side.wait_event(done)
        with torch.cuda.stream(self.side):
            self.host_sums.copy_(self.sums, non_blocking=True)
            self.sums_ready.record(self.side)

    def _sqdiff(self, a, b, out):
        L = sg._bind()
        sg.check(L.lbmib_sg_sqdiff(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(0 if b is None else b.data_ptr()),
                                   a.numel(), ctypes.c_void_p(out.data_ptr()), self.stream), "sqdiff")

    def cycle(self, predictor=None, lookahead=False):
        """One jump of K steps plus R verification steps (K + R simulated
        steps), or K + R plain steps after a rejected jump.  predictor: a
        function filling self.delta instead of the network (None: the
        network).  lookahead (network only): another cycle follows, so its
        prediction from the verified state is enqueued before the host
        waits for the verification sums; a rejected jump discards it.
        Returns True if the jump was accepted."""
        with torch.cuda.stream(self.ext):
            return self._cycle(predictor, lookahead and predictor is None)

    def _cycle(self, predictor, lookahead):
        e = self.e
        if self.early and self.pred_pending:  # the side prediction has read pfeat and written delta
            self.ext.wait_event(self.pred_done)
            self.pred_pending = False
        if not self.have_feat:
            self.co.read()
            self.have_feat = True
            self.pred_ready = False
        if predictor is not None:
            predictor(self)
        elif not self.pred_ready:
            self.predict()
        self.pred_ready = False
        # the jump also reports the jumped state's coarse stored moments
        e.jump(self.co.f, self.co.feat.data_ptr(), self.delta.data_ptr(), self.K, self.after.data_ptr())
        # verification ratio: |du over R steps| / R  against  |du of the jump| / K
        self._sqdiff(self.delta[1:4], None, self.sums[2:4])
        early = lookahead and self.early and self.graph is not None
        if early:
            # the next change, from the jumped state, beside the verification
            e.coarse_sheets(self.co.f, self.pfeat[4].data_ptr())
            ready = torch.cuda.Event()
            ready.record(self.ext)
            self.pred_stream.wait_event(ready)
            with torch.cuda.stream(self.pred_stream):
                self.graph.replay()
                self.pred_done.record(self.pred_stream)
            self.pred_pending = True
        # R plain steps, the last of which also reads the next cycle's input,
        # which the verification compares with the jumped state
        if self.R > 1:
            e.step(self.R - 1)
        self.co.step_read()
        self._sqdiff(self.co.feat[1:4], self.after[1:4], self.sums[0:2])
        self._sums_to_host()
        if early:
            self.pred_ready = True
        elif lookahead:
            self.predict()  # the next cycle's change, from the verified state
            self.pred_ready = True
        self.sums_ready.synchronize()
        s = self.host_sums.numpy()
        step_rate = np.sqrt(s[0]) / self.R
        jump_rate = np.sqrt(s[2]) / self.K
        ratio = float(step_rate / max(jump_rate, 1e-30))
        self.stats["jumps"] += 1
        self.stats["ratios"].append(ratio)
        # (a non-finite prediction gives a NaN ratio and is rejected; device
        # errors are latched for good, so run() checks them once)
        if ratio <= self.tol or self.tol == float("inf"):
            e.jump_commit()
            self.stats["accepted"] += 1
            return True
        e.jump_revert()
        e.step(self.K + self.R - 1)
        self.co.step_read()
        self.stats["rejected"] += 1
        self.stats["plain_steps"] += self.K + self.R
        self.pred_ready = False
        return False

    def run(self, nsteps, predictor=None):
        """Advances nsteps simulated steps: whole jump cycles, then plain steps
        for the remainder."""
        done, span = 0, self.K + self.R
        while done + span <= nsteps:
            self.cycle(predictor, lookahead=done + 2 * span <= nsteps)
            done += span
        if self.early and self.pred_pending:  # no side work outlives the call
            self.ext.wait_event(self.pred_done)
            self.pred_pending = False
        if done < nsteps:
            self.e.step(nsteps - done)
            self.stats["plain_steps"] += nsteps - done
            self.have_feat = False
        self.stats["status"] = self.e.status()
        return self.stats


def zero_predictor(runner):
    """The method without the network: no change of the coarse state over
    the jump (the sheet still moves with the coarse velocity).  Called inside
    JumpRunner.cycle, on the engine's stream."""
    runner.delta.zero_()


def evaluate(net, norm, kw, warm, horizon, factor=4, ksteps=50, verify_steps=1, tol=4.0, baseline=True,
             log=print, early=False):
    """The 256^3 evaluation: from the same state after `warm` plain steps,
    `horizon` more steps (a) plain, (b) with surrogate jumps, (c) with jumps
    of the method without the network (zero change).  Returns timings
    (wall clock around each leg, synchronised) and the velocity error of
    (b) and (c) against (a): relative L2 over the whole field, and relative
    to the departure of (a) from the uniform inflow."""
    import time
    out = {"warm": warm, "horizon": horizon, "ksteps": ksteps, "verify_steps": verify_steps, "tol": tol,
           "factor": factor}
    u_in = kw["u0"][0]

    def fields(e):
        rho, vel = e.macroscopic()
        return vel

    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(warm)
        e.sync()
        t0 = time.time()
        e.step(horizon)
        e.sync()
        out["plain_s"] = time.time() - t0
        assert e.status() == 0
        ref = fields(e)
        ref_x = e.fibers(0)[..., :3]
    # the method without the network is shown with every jump accepted: its
    # zero change would fail any verification ratio (its own rate is zero)
    legs = [("surrogate", None, tol)] + ([("no_network", zero_predictor, float("inf"))] if baseline else [])
    for name, pred, leg_tol in legs:
        with lbmib.Engine(**kw) as e:
            e.init()
            e.step(warm)
            r = JumpRunner(e, net, norm, kw["tau"], u_in, factor, ksteps, verify_steps, leg_tol, early=early)
            r.prepare(pred)  # buffers and the prediction graph, outside the timed leg
            e.sync()
            t0 = time.time()
            st = r.run(horizon, pred)
            e.sync()
            dt = time.time() - t0
            status = e.status()
            vel = fields(e)
            x = e.fibers(0)[..., :3]
        pert = ref.copy()
        pert[0] -= u_in
        leg = {"seconds": dt, "status": status, "jumps": st["jumps"], "accepted": st["accepted"],
               "rejected": st["rejected"], "plain_steps": st["plain_steps"],
               "ratio_max": max(st["ratios"]) if st["ratios"] else None,
               "ratio_median": float(np.median(st["ratios"])) if st["ratios"] else None,
               "vel_rel_l2": float(np.linalg.norm(vel - ref) / np.linalg.norm(ref)),
               "vel_rel_l2_perturbation": float(np.linalg.norm(vel - ref) / np.linalg.norm(pert)),
               "fiber_pos_max_abs": float(np.abs(x - ref_x).max()),
               "speedup": out["plain_s"] / dt}
        out[name] = leg
        log(f"[eval] {name}: {leg}")
    return out


if __name__ == "__main__":
    import argparse
    import json
    import os

    import lbmib_configs

    ap = argparse.ArgumentParser(description="the surrogate on the 256^3 one-sheet workload: plain steps against "
                                             "surrogate jumps, with the jump method without the network beside")
    ap.add_argument("--model", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "models",
                                                    "surrogate_c5.npz"))
    ap.add_argument("--warm", type=int, default=100)
    ap.add_argument("--horizon", type=int, default=1020)
    ap.add_argument("--tol", type=float, default=5.0)
    ap.add_argument("--verify-steps", type=int, default=1)
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--early", action="store_true",
                    help="predict the next change from the jumped state, beside the verification step")
    args = ap.parse_args()
    net, norm, factor, ksteps, _ = load_model(args.model)
    kw, _ = lbmib_configs.c2(precision="f32")
    res = evaluate(net, norm, kw, args.warm, args.horizon, factor, ksteps, args.verify_steps, args.tol,
                   baseline=not args.no_baseline, early=args.early)
    print(json.dumps(res), flush=True)
