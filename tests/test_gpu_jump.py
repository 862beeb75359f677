"""Coupling of the surrogate to a running simulation (csrc/lbmib_jump.hip,
distributed-lb-ib_amd/surrogate_jump.py), through the C-ABI.

- lbmib_coarse_moments: FP64 engine against the block average of the FP64
  CPU oracle's rho/u (lbm_ib.c:2440-2538), <= 1e-6 relative (FP32 output of
  FP64 sums), and bit-consistent with lbmib_get_macroscopic's fields;
- lbmib_coarse_state's stored planes alone: the combined read's bit for bit;
- lbmib_step_coarse: a step and the stored planes' read, bit for bit;
- lbmib_coarse_sheets: node counts exact, force sums within FP32 rounding and
  the 2^-40 fixed-point quantum per node, against a numpy raster of the
  downloaded nodes;
- lbmib_jump: with a zero change the populations stay bit for bit and the
  nodes move by K times the interpolated coarse velocity (<= 1e-12); with a
  smooth random change the post-stream populations match the numpy
  restatement (oracle/surrogate_oracle.jump_pulled) to 1e-13 (F64) and to
  FP32 rounding (F32, whose shift is computed in FP32 on the deviations);
- lbmib_jump_revert: after a jump and several steps, revert plus plain steps
  give the plain trajectory bit for bit (F64, F32, BF16, with a sheet);
- the jump loop with a tolerance of 0 rejects every jump and reproduces
  lbmib_step bit for bit.
"""
import numpy as np
import pytest
import torch

from oracle import surrogate_oracle as so
from oracle.oracle import BC_INOUT, BC_PERIODIC, BC_WALL, Oracle

pytestmark = pytest.mark.gpu


def small_cfg(lbmib, precision, n=(64, 32, 32), rows=12, seed=3):
    nx, ny, nz = n
    sheet = lbmib.Sheet.centred(rows, rows, 1.0, (18.0, ny / 2.0, nz / 2.0), 2e-3, 2e-3, jitter=0.08, seed=seed)
    return dict(nx=nx, ny=ny, nz=nz, bc=(BC_INOUT, BC_PERIODIC, BC_WALL), precision=precision, tau=0.7, rho0=1.0,
                u0=(0.02, 0.0, 0.0), quirks=True, sheets=[sheet])


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(np.float64)


def coarse_of(e, f):
    out = torch.empty((4, e.nz // f, e.ny // f, e.nx // f), dtype=torch.float32, device="cuda")
    e.coarse_moments(f, out.data_ptr())
    e.sync()
    return host(out)


@pytest.mark.parametrize("factor", [2, 4])
def test_coarse_moments_match_oracle_block_average(lbmib, factor):
    kw = small_cfg(lbmib, "f64")
    o = Oracle(kw["nx"], kw["ny"], kw["nz"], kw["bc"], kw["tau"], kw["rho0"], kw["u0"], True)
    s = kw["sheets"][0]
    o.add_sheet(s.rows, s.cols, s.width, s.height, s.ks, s.kb, s.xyz)
    o.init()
    o.preshift(all_sheets=True)
    assert o.step(12) == 0
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(12)
        c = coarse_of(e, factor)
        rho, vel = e.macroscopic()
    ref = [so.block_average(o.rho(), factor)] + [so.block_average(o.vel()[a], factor) for a in range(3)]
    for a in range(4):
        assert np.abs(c[a] - ref[a]).max() <= 1e-6 * np.abs(ref[a]).max(), a
    # the same fields lbmib_get_macroscopic returns, averaged
    np.testing.assert_allclose(c[0], so.block_average(rho, factor), rtol=1e-6)


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16"])
def test_stored_planes_alone_match_the_combined_read(lbmib, precision):
    """lbmib_coarse_state with d_moments null (each cell's own populations
    only) gives the stored planes of the combined read bit for bit."""
    kw = small_cfg(lbmib, precision)
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(7)
        for f in (2, 4):
            shp = (4, e.nz // f, e.ny // f, e.nx // f)
            cm, both, alone = (torch.empty(shp, dtype=torch.float32, device="cuda") for _ in range(3))
            e.coarse_moments(f, cm.data_ptr(), both.data_ptr())
            e.coarse_moments(f, 0, alone.data_ptr())
            e.sync()
            np.testing.assert_array_equal(host(alone), host(both))
        with pytest.raises(lbmib.LbmibError):  # neither plane set
            e.coarse_moments(4, 0, 0)


@pytest.mark.parametrize("precision", ["f64", "f32", "f16", "bf16"])
@pytest.mark.parametrize("xbc", ["inout", "periodic"])
def test_step_coarse_matches_step_then_read(lbmib, precision, xbc):
    """lbmib_step_coarse (the fluid launch also reporting the stored moments;
    on a periodic x axis, the two-pass form) against lbmib_step and
    lbmib_coarse_state: the same populations, sheet and planes, bit for bit,
    from the first step (the populations' own moments) on."""
    kw = small_cfg(lbmib, precision)
    if xbc == "periodic":
        kw["bc"] = (BC_PERIODIC, BC_PERIODIC, BC_WALL)
    with lbmib.Engine(**kw) as a, lbmib.Engine(**kw) as b:
        a.init()
        b.init()
        for rep in range(3):
            for f in (4, 2):
                shp = (4, a.nz // f, a.ny // f, a.nx // f)
                pa, pb = (torch.empty(shp, dtype=torch.float32, device="cuda") for _ in range(2))
                a.step_coarse(f, pa.data_ptr())
                b.step(1)
                b.coarse_moments(f, 0, pb.data_ptr())
                a.sync()
                b.sync()
                np.testing.assert_array_equal(host(pa), host(pb))
            a.step(3)
            b.step(3)
        assert a.time == b.time
        np.testing.assert_array_equal(a.df(), b.df())
        np.testing.assert_array_equal(a.fibers(0), b.fibers(0))
        with pytest.raises(lbmib.LbmibError):  # factor 3
            a.step_coarse(3, pa.data_ptr())
        assert a.status() == 0


def test_coarse_sheets_raster(lbmib):
    kw = small_cfg(lbmib, "f32")
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(5)
        out = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        e.coarse_sheets(4, out.data_ptr())
        e.sync()
        got = host(out)
        nodes = e.fibers(0).reshape(-1, 12)
    ref = so.raster(nodes[:, :3], nodes[:, 9:12], 4, (8, 8, 16))
    np.testing.assert_array_equal(got[0], ref[0])
    assert got[0].sum() == nodes.shape[0]
    # FP32 planes of 64-bit fixed-point sums (2^-40 per node and component)
    assert np.all(np.abs(got[1:] - ref[1:]) <= 2.0 ** -23 * np.abs(ref[1:]) + ref[0] * 2.0 ** -40 + 1e-30)


def test_zero_jump_keeps_populations_and_advects_sheet(lbmib):
    kw = small_cfg(lbmib, "f64")
    K = 10
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(6)
        df0 = e.df()
        x0 = e.fibers(0)[..., :3].reshape(-1, 3)
        cm = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        e.coarse_moments(4, cm.data_ptr())
        zero = torch.zeros((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        before = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        after = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        e.coarse_moments(4, cm.data_ptr(), before.data_ptr())
        t0 = e.time
        e.jump(4, cm.data_ptr(), zero.data_ptr(), K, after.data_ptr())
        assert e.time == t0 + K
        np.testing.assert_array_equal(e.df(), df0)
        # the jump's stored-moment planes of an unchanged state are the read's
        np.testing.assert_array_equal(host(after), host(before))
        x1 = e.fibers(0)[..., :3].reshape(-1, 3)
        c = host(cm)
        want = x0 + K * np.stack([so.trilinear_at(c[1 + a], 4, x0) for a in range(3)], 1)
        np.testing.assert_allclose(x1, want, rtol=0, atol=1e-12)
        e.jump_commit()
        e.step(2)
        assert e.status() == 0


def test_jump_shift_matches_restatement(lbmib):
    """A smooth random change on an F64 state without sheets: the post-stream
    populations after the jump, away from the boundaries, against
    oracle/surrogate_oracle.jump_pulled."""
    kw = small_cfg(lbmib, "f64")
    kw["sheets"] = []
    kw["bc"] = (BC_PERIODIC, BC_PERIODIC, BC_PERIODIC)
    rng = np.random.default_rng(11)
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(3)
        f2 = e.df()
        cm = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        e.coarse_moments(4, cm.data_ptr())
        delta = (rng.standard_normal((4, 8, 8, 16)) * np.array([1e-3, 1e-3, 5e-4, 5e-4])[:, None, None, None])
        delta = delta.astype(np.float32).astype(np.float64)
        d_dev = dev(delta)
        after = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        e.jump(4, cm.data_ptr(), d_dev.data_ptr(), 4, after.data_ptr())
        got = e.df()
        reread = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        e.coarse_moments(4, cm.data_ptr(), reread.data_ptr())
        e.sync()
    want = so.jump_pulled(f2, delta, 4)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
    # the jumped state's stored moments (every cell is updated on this
    # all-periodic grid): h'_q(y) = f2'_q(y + c_q), block means of rho, u
    hq = np.array([np.roll(got[q], (-so.CZ[q], -so.CY[q], -so.CX[q]), axis=(0, 1, 2)) for q in range(19)])
    rho = hq.sum(0)
    mom = [rho] + [(c[:, None, None, None] * hq).sum(0) / rho for c in (so.CX, so.CY, so.CZ)]
    ref = np.array([so.block_average(m, 4) for m in mom])
    for a in range(4):
        assert np.abs(host(after)[a] - ref[a]).max() <= 1e-6 * np.abs(ref[a]).max(), a
        # the stored-moment read of the same state gives the same planes
        assert np.abs(host(reread)[a] - ref[a]).max() <= 1e-6 * np.abs(ref[a]).max(), a


def test_jump_shift_f32_matches_restatement(lbmib):
    """The FP32 deviation form computes the shift in FP32 on h = f - r
    (moments rho_r + sum h, rho_r u_r + sum c h; the change in difference
    form).  A first random jump and six steps make the state non-uniform;
    then, against the FP64 restatement applied to the FP32 state, each
    population lies within half an FP32 ulp of its stored deviation plus
    1e-6 of the largest shift (FP32 rounding of terms of that size), and the
    coarse reads (pulled and stored moments) within 1e-6 of the block means
    of the FP64 moments."""
    kw = small_cfg(lbmib, "f32")
    kw["sheets"] = []
    kw["bc"] = (BC_PERIODIC, BC_PERIODIC, BC_PERIODIC)
    rng = np.random.default_rng(12)
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(3)
        cm = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        e.coarse_moments(4, cm.data_ptr())
        d1 = dev(rng.standard_normal((4, 8, 8, 16)) * np.array([2e-3, 2e-3, 1e-3, 1e-3])[:, None, None, None])
        e.jump(4, cm.data_ptr(), d1.data_ptr(), 4)
        e.jump_commit()
        e.step(6)
        f2 = e.df()
        st = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        e.coarse_moments(4, cm.data_ptr(), st.data_ptr())
        pulled, stored = host(cm), host(st)
        delta = (rng.standard_normal((4, 8, 8, 16)) * np.array([1e-3, 1e-3, 5e-4, 5e-4])[:, None, None, None])
        delta = delta.astype(np.float32).astype(np.float64)
        d_dev = dev(delta)
        e.jump(4, cm.data_ptr(), d_dev.data_ptr(), 4)
        got = e.df()
        e.sync()
    h = np.array([np.roll(f2[q], (-so.CZ[q], -so.CY[q], -so.CX[q]), axis=(0, 1, 2)) for q in range(19)])
    for got_planes, pops in ((pulled, f2), (stored, h)):
        rho = pops.sum(0)
        mom = [rho] + [(c[:, None, None, None] * pops).sum(0) / rho for c in (so.CX, so.CY, so.CZ)]
        for a in range(4):
            ref = so.block_average(mom[a], 4)
            assert np.abs(got_planes[a] - ref).max() <= 1e-6 * np.abs(ref).max(), a
    want = so.jump_pulled(f2, delta, 4)
    r = so.feq(np.float64(kw["rho0"]), np.float64(kw["u0"][0]), 0.0, 0.0)[:, None, None, None]
    bound = 2.0 ** -24 * np.abs(want - r) + 1e-6 * np.abs(want - f2).max() + 1e-15
    err = np.abs(got - want)
    assert np.all(err <= bound), float((err / bound).max())


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16"])
def test_revert_gives_plain_trajectory(lbmib, precision):
    kw = small_cfg(lbmib, precision)
    rng = np.random.default_rng(5)
    delta = dev(rng.standard_normal((4, 8, 8, 16)) * 1e-3)
    with lbmib.Engine(**kw) as a, lbmib.Engine(**kw) as b:
        a.init()
        b.init()
        a.step(7)
        b.step(7)
        cm = torch.empty((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        a.coarse_moments(4, cm.data_ptr())
        a.jump(4, cm.data_ptr(), delta.data_ptr(), 20)
        a.step(3)        # verification steps overwrite both force slots
        a.jump_revert()
        assert a.time == b.time
        a.step(9)
        b.step(9)
        assert a.status() == 0 and b.status() == 0
        np.testing.assert_array_equal(a.df(), b.df())
        np.testing.assert_array_equal(a.fibers(0), b.fibers(0))
        np.testing.assert_array_equal(a.force(), b.force())


def test_rejecting_every_jump_reproduces_plain_steps(lbmib):
    import surrogate as sg
    import surrogate_jump as sj
    kw = small_cfg(lbmib, "f32")
    net = sg.Net(cin=sj.CIN, cout=16, nout=sj.NOUT, chans=(32, 64, 128, 256), seed=1)
    norm = sj.Norm([1.0] + [0.0] * 7, [1e-3] * 8, [0.0] * 4, [1e-4] * 4)
    with lbmib.Engine(**kw) as a, lbmib.Engine(**kw) as b:
        a.init()
        b.init()
        a.step(2)
        b.step(2)
        r = sj.JumpRunner(a, net, norm, kw["tau"], kw["u0"][0], factor=4, ksteps=10, verify_steps=2, tol=0.0)
        st = r.run(30)
        b.step(30)
        assert st["jumps"] == 2 and st["rejected"] == 2 and st["accepted"] == 0
        assert a.time == b.time
        np.testing.assert_array_equal(a.df(), b.df())
        np.testing.assert_array_equal(a.fibers(0), b.fibers(0))


def test_lookahead_prediction_gives_the_same_run(lbmib):
    """Cycles whose next prediction is enqueued before the host reads the
    verification sums (run()) reach the same bits as cycles that predict at
    their start, accepted and rejected jumps alike."""
    import surrogate as sg
    import surrogate_jump as sj
    kw = small_cfg(lbmib, "f32")
    net = sg.Net(cin=sj.CIN, cout=16, nout=sj.NOUT, seed=4, out_scale=0.05)
    norm = sj.Norm([1.0, 0.02, 0, 0, 0, 0, 0, 0], [1e-3, 3e-3, 3e-3, 3e-3, 1.0, 1e-4, 1e-4, 1e-4], [0.0] * 4,
                   [1e-4] * 4)
    for tol in (float("inf"), 0.0):
        with lbmib.Engine(**kw) as a, lbmib.Engine(**kw) as b:
            a.init()
            b.init()
            a.step(2)
            b.step(2)
            t0 = a.time
            ra = sj.JumpRunner(a, net, norm, kw["tau"], kw["u0"][0], factor=4, ksteps=10, verify_steps=2, tol=tol)
            rb = sj.JumpRunner(b, net, norm, kw["tau"], kw["u0"][0], factor=4, ksteps=10, verify_steps=2, tol=tol)
            sa = ra.run(36)
            for _ in range(3):
                rb.cycle(lookahead=False)
            assert sa["jumps"] == 3 and rb.stats["jumps"] == 3
            assert sa["accepted"] == rb.stats["accepted"] == (3 if tol > 0 else 0)
            assert a.time == b.time == t0 + 36
            np.testing.assert_array_equal(a.df(), b.df())
            np.testing.assert_array_equal(a.fibers(0), b.fibers(0))


def test_early_prediction_beside_the_verification(lbmib):
    """early=True: the next prediction reads the jumped state and runs on a
    stream of its own while the engine's stream verifies.  Two such runs
    give the same bits (the streams are ordered by events, not by timing),
    and with every jump rejected the discarded side predictions leave the
    plain trajectory bit for bit."""
    import surrogate as sg
    import surrogate_jump as sj
    kw = small_cfg(lbmib, "f32")
    net = sg.Net(cin=sj.CIN, cout=16, nout=sj.NOUT, seed=4, out_scale=0.05)
    norm = sj.Norm([1.0, 0.02, 0, 0, 0, 0, 0, 0], [1e-3, 3e-3, 3e-3, 3e-3, 1.0, 1e-4, 1e-4, 1e-4], [0.0] * 4,
                   [1e-4] * 4)
    for tol in (float("inf"), 0.0):
        print(f"early prediction, tol {tol}", flush=True)
        with lbmib.Engine(**kw) as a, lbmib.Engine(**kw) as b, lbmib.Engine(**kw) as c:
            for e in (a, b, c):
                e.init()
                e.step(2)
            t0 = a.time
            runs = []
            for e in (a, b):
                r = sj.JumpRunner(e, net, norm, kw["tau"], kw["u0"][0], factor=4, ksteps=10, verify_steps=2, tol=tol,
                                  early=True)
                r.prepare()
                runs.append(r.run(48))
            assert runs[0]["jumps"] == runs[1]["jumps"] == 4
            assert runs[0]["accepted"] == runs[1]["accepted"] == (4 if tol > 0 else 0)
            assert runs[0]["status"] == 0
            assert a.time == b.time == t0 + 48
            np.testing.assert_array_equal(a.df(), b.df())
            np.testing.assert_array_equal(a.fibers(0), b.fibers(0))
            if tol == 0.0:
                c.step(48)
                np.testing.assert_array_equal(a.df(), c.df())
                np.testing.assert_array_equal(a.fibers(0), c.fibers(0))


def test_coupling_rejects_bad_shapes_and_states(lbmib):
    kw = small_cfg(lbmib, "f32", n=(40, 32, 32))
    buf = torch.zeros((4, 1024), dtype=torch.float32, device="cuda")
    with lbmib.Engine(**kw) as e:
        e.init()
        with pytest.raises(lbmib.LbmibError):      # nx not a multiple of 64
            e.coarse_moments(4, buf.data_ptr())
    kw = small_cfg(lbmib, "f32")
    with lbmib.Engine(**kw) as e:
        e.init()
        cm = torch.zeros((4, 8, 8, 16), dtype=torch.float32, device="cuda")
        with pytest.raises(lbmib.LbmibError):      # no completed step yet
            e.jump(4, cm.data_ptr(), cm.data_ptr(), 10)
        e.step(1)
        with pytest.raises(lbmib.LbmibError):      # odd step count
            e.jump(4, cm.data_ptr(), cm.data_ptr(), 9)
        with pytest.raises(lbmib.LbmibError):      # nothing to revert
            e.jump_revert()
        with pytest.raises(lbmib.LbmibError):      # factor 3
            e.coarse_moments(3, cm.data_ptr())


def test_prediction_graph_replays_match_eager(lbmib):
    """The network's launches recorded in a HIP graph on the engine's stream:
    every replay, with the engine's own steps and synchronisations between
    replays, gives the same bits as the launches run one by one."""
    import surrogate as sg
    import surrogate_jump as sj
    kw = small_cfg(lbmib, "f32")
    net = sg.Net(cin=sj.CIN, cout=16, nout=sj.NOUT, seed=4, out_scale=0.05)
    norm = sj.Norm([1.0, 0.02, 0, 0, 0, 0, 0, 0], [1e-3, 3e-3, 3e-3, 3e-3, 1.0, 1e-4, 1e-4, 1e-4], [0.0] * 4,
                   [1e-4] * 4)
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(3)
        eager = sj.JumpRunner(e, net, norm, kw["tau"], kw["u0"][0], factor=4, ksteps=10, graph=False)
        graphed = sj.JumpRunner(e, net, norm, kw["tau"], kw["u0"][0], factor=4, ksteps=10, graph=True)
        for rep in range(4):
            eager.co.read()
            graphed.co.read()
            eager.predict()
            e.sync()                      # the engine's stream does not order with torch's
            want = eager.delta.cpu().numpy().copy()
            graphed.predict()
            e.sync()
            got = graphed.delta.cpu().numpy()
            assert (graphed.graph is not None) == (rep >= 1)
            np.testing.assert_array_equal(got, want)
            e.step(2)
            e.sync()

