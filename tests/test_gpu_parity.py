"""HIP path against the FP64 oracle and the reference's own fixtures.

Everything here runs the kernels of liblbmib.so through its C-ABI.
Tolerances:
  F64            <= 1e-12 relative (positions, fields) -- same operation order
                 as the oracle; only cos() and the order of FMA-free sums on
                 the device can differ by an ulp.
  F32            velocity <= 1e-5 relative to max|u| against the FP64 oracle.
  F16 / BF16     velocity <= 1e-3 / 1e-2 relative to max|u| (small grids);
                 on the 512x256x256 workload (2000 steps) against FP64,
                 storing f - feq(rho0, u0x): relative L2 <= 1e-3 / 2e-2 and
                 max <= 5e-3 / 1e-1; storing f - w: L2 <= 1e-2 / 5e-2 and
                 max <= 3e-2 / 2e-1.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, read_fiber_dump, reference_sheet_xyz
from oracle.oracle import BC_INOUT, BC_PERIODIC, BC_WALL, Oracle

pytestmark = pytest.mark.gpu


def oracle_for(kw, preshift=True, preshift_all=True):
    o = Oracle(kw["nx"], kw["ny"], kw["nz"], kw["bc"], kw["tau"], kw["rho0"], kw["u0"], kw.get("quirks", True))
    for s in kw.get("sheets", []):
        o.add_sheet(s.rows, s.cols, s.width, s.height, s.ks, s.kb, s.xyz)
    o.init()
    if preshift and kw.get("sheets"):
        o.preshift(all_sheets=preshift_all)
    return o


def active_slices(kw):
    sl = []
    for a, n in ((2, kw["nz"]), (1, kw["ny"]), (0, kw["nx"])):
        if kw["bc"][a] == BC_PERIODIC:
            sl.append(slice(0, n))
        else:
            sl.append(slice(2, n - 2))
    return tuple(sl)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def small_ib_config(lbmib, precision, bc=(BC_INOUT, BC_PERIODIC, BC_WALL), quirks=True, n=(40, 32, 36)):
    nx, ny, nz = n
    sheet = lbmib.Sheet.centred(10, 12, 1.0, (14.0, ny / 2.0, nz / 2.0), 2e-3, 2e-3, jitter=0.08, seed=11)
    return dict(nx=nx, ny=ny, nz=nz, bc=bc, precision=precision, tau=0.7, rho0=1.0,
                u0=(0.01, 0.0, 0.0), quirks=quirks, sheets=[sheet])


# ------------------------------------------------------------ fixtures replay

def test_f64_replays_reference_fixture_refA(lbmib, configs, fixtures_meta):
    """The compiled reference's own dumps (64^3, 26x26 sheet, 20 steps) from
    the HIP FP64 path."""
    meta = fixtures_meta["refA"]
    kw = configs.reference_mode(meta["grid"], meta["rows"], meta["cols"], meta["width"], meta["height"],
                                tuple(meta["origin"]), precision="f64")
    with lbmib.Engine(**kw) as e:
        e.init()
        done = 0
        for step in meta["fiber_steps"]:
            e.step(step - done)
            done = step
            assert e.status() == 0
            _, ref = read_fiber_dump(os.path.join(GOLDEN, f"refA_fiber_step{step}.bin"))
            got = e.fibers(0).reshape(-1, 12)
            print(f"step {step}: pos {np.abs(got[:, :3] - ref[:, :3]).max():.3e}  "
                  f"force {np.abs(got[:, 3:] - ref[:, 3:]).max():.3e} (max |F| {np.abs(ref[:, 3:]).max():.3e})")
            assert np.abs(got[:, :3] - ref[:, :3]).max() <= 1e-12
            assert np.abs(got[:, 3:] - ref[:, 3:]).max() <= 1e-9 * max(np.abs(ref[:, 3:]).max(), 1e-12) + 1e-20
        fx = np.load(os.path.join(GOLDEN, f"refA_fluid_step{meta['fluid_steps'][-1]}.npz"))
        X, Y, Z = fx["xyz"][:, 0], fx["xyz"][:, 1], fx["xyz"][:, 2]
        rho, vel = e.macroscopic()
        df = e.df()
        frc = e.force()
        assert rel(vel[:, Z, Y, X].T, fx["vel"]) < 1e-11
        assert np.abs(rho[Z, Y, X] - fx["rho"]).max() < 1e-13
        assert np.abs(df[:, Z, Y, X].T - fx["df2"]).max() < 1e-14
        assert np.abs(frc[:, Z, Y, X].T - fx["force"]).max() <= 1e-9 * np.abs(fx["force"]).max() + 1e-22


# ---------------------------------------------------------- F64 vs oracle

BC_CASES = [
    (BC_INOUT, BC_WALL, BC_WALL),
    (BC_INOUT, BC_PERIODIC, BC_WALL),
    (BC_PERIODIC, BC_PERIODIC, BC_PERIODIC),
    (BC_WALL, BC_WALL, BC_WALL),
    (BC_WALL, BC_PERIODIC, BC_PERIODIC),
]


@pytest.mark.parametrize("bc", BC_CASES)
def test_f64_fluid_matches_oracle(lbmib, bc):
    kw = dict(nx=20, ny=17, nz=14, bc=bc, precision="f64", tau=0.8, rho0=1.0, u0=(0.02, 0.003, -0.004),
              quirks=True, sheets=[])
    o = oracle_for(kw)
    with lbmib.Engine(**kw) as e:
        e.init()
        # before any step: df1 = feq on the active box
        sl = (slice(None),) + active_slices(kw)
        assert np.array_equal(e.df()[sl], o.df1()[sl])
        for k in range(3):
            e.step(2)
            o.step(2)
            assert rel(e.df()[sl], o.df1()[sl]) < 1e-14
            rho, vel = e.macroscopic()
            assert rel(rho, o.rho()) < 1e-14
            assert rel(vel, o.vel()) < 1e-12


@pytest.mark.parametrize("bc", [(BC_INOUT, BC_PERIODIC, BC_WALL), (BC_INOUT, BC_WALL, BC_WALL),
                                (BC_PERIODIC, BC_PERIODIC, BC_WALL)])
@pytest.mark.parametrize("quirks", [True, False])
def test_f64_ib_matches_oracle(lbmib, bc, quirks):
    kw = small_ib_config(lbmib, "f64", bc=bc, quirks=quirks)
    o = oracle_for(kw)
    with lbmib.Engine(**kw) as e:
        e.init()
        for chunk in (1, 4, 10):
            e.step(chunk)
            assert o.step(chunk) == 0
            assert e.status() == 0
            got, ref = e.fibers(0), o.nodes(0)
            assert np.abs(got[..., :3] - ref[..., :3]).max() < 1e-12
            assert rel(got[..., 3:], ref[..., 3:]) < 1e-9
            assert rel(e.force(), o.force()) < 1e-9
            rho, vel = e.macroscopic()
            assert rel(vel, o.vel()) < 1e-11
            assert rel(rho, o.rho()) < 1e-14


def test_named_entry_points_match_step(lbmib):
    """The reference's per-step sequence called one entry point at a time
    (do_thread order, lbm_ib.c:3124-3460) gives exactly lbmib_step."""
    kw = small_ib_config(lbmib, "f32")
    seq = ["compute_bendingforce", "compute_stretchingforce", "compute_elasticforce",
           "get_influentialdomain_fluid_grid_and_SpreadForce", "compute_eqlbrmdistrfuncDF1",
           "stream_distrfunc", "bounceback_rigidwalls", "compute_rho_and_u", "moveFiberSheet",
           "copy_inout_to_df2", "replace_old_DF", "periodicBC"]
    with lbmib.Engine(**kw) as a, lbmib.Engine(**kw) as b:
        a.init()
        b.init()
        for _ in range(6):
            for name in seq:
                a.call(name)
        b.step(6)
        assert a.time == b.time == 7
        assert np.array_equal(a.fibers(0), b.fibers(0))
        assert np.array_equal(a.df(), b.df())


# ----------------------------------------------------- reduced precisions

@pytest.mark.parametrize("dev_ref", ["eq0", "rest"])
@pytest.mark.parametrize("bc", BC_CASES)
def test_deviation_forms_fluid_match_oracle(lbmib, bc, dev_ref):
    """FP32 storage of f - r, for both reference populations r
    (feq(rho0, u0x) and w), on every boundary combination: a reflected
    population is stored relative to r of the opposite direction, so a wrong
    or missing bounce-back correction shows as an error of order u0 at the
    walls, far above the FP32 rounding."""
    kw = dict(nx=20, ny=17, nz=14, bc=bc, precision="f32", tau=0.8, rho0=1.0, u0=(0.02, 0.003, -0.004),
              quirks=True, sheets=[], dev_ref=dev_ref)
    o = oracle_for(kw)
    with lbmib.Engine(**kw) as e:
        e.init()
        sl = (slice(None),) + active_slices(kw)
        for k in range(3):
            e.step(4)
            o.step(4)
            rho, vel = e.macroscopic()
            err_f, err_u = rel(e.df()[sl], o.df1()[sl]), rel(vel, o.vel())
            assert err_f < 1e-6, (k, err_f)
            assert err_u < 1e-5, (k, err_u)
            assert rel(rho, o.rho()) < 1e-6
        print(f"{bc} {dev_ref}: df {err_f:.2e} velocity {err_u:.2e}")


@pytest.mark.parametrize("dev_ref", ["eq0", "rest"])
@pytest.mark.parametrize("precision,tol", [("f32", 1e-5), ("f16", 1e-3), ("bf16", 1e-2)])
def test_reduced_precision_error_vs_fp64_oracle(lbmib, precision, tol, dev_ref):
    kw = small_ib_config(lbmib, precision)
    kw["dev_ref"] = dev_ref
    o = oracle_for(kw)
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(40)
        assert o.step(40) == 0
        _, vel = e.macroscopic()
        err_u = rel(vel, o.vel())
        err_x = np.abs(e.fibers(0)[..., :3] - o.nodes(0)[..., :3]).max()
        print(f"{precision} {dev_ref}: velocity rel err {err_u:.3e}, position abs err {err_x:.3e}")
        assert err_u < tol
        assert err_x < 10 * tol


# ------------------------------------------------------------ invariants

@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 1e-6)])
@pytest.mark.parametrize("bc", [(BC_PERIODIC, BC_PERIODIC, BC_PERIODIC), (BC_WALL, BC_WALL, BC_WALL)])
def test_closed_box_conserves_mass(lbmib, precision, tol, bc):
    kw = small_ib_config(lbmib, precision, bc=bc, quirks=False, n=(32, 32, 32))
    with lbmib.Engine(**kw) as e:
        e.init()
        m0 = e.totals()[0]
        e.step(60)
        m1 = e.totals()[0]
        print(f"{precision} {bc}: mass drift {abs(m1 - m0) / m0:.3e}")
        assert abs(m1 - m0) / m0 < tol


def test_runs_are_bitwise_deterministic(lbmib):
    kw = small_ib_config(lbmib, "f32")
    outs = []
    for _ in range(2):
        with lbmib.Engine(**kw) as e:
            e.init()
            e.step(15)
            outs.append((e.fibers(0), e.df(), e.force()))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_fiber_leaving_grid_reports_error(lbmib):
    kw = small_ib_config(lbmib, "f32")
    s = kw["sheets"][0]
    s.xyz[..., 0] = 0.5  # support would start at x = -1 (lbm_ib.c:1458 exits here)
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(1)
        assert e.status() == lbmib.ERR_FIBER_OOB


def test_fiber_far_off_grid_reports_error(lbmib):
    """A node far beyond the grid: its support start is beyond the int range,
    so the test on it has to be made before the conversion (which saturates);
    the node is left out of spreading and moving and the error is latched."""
    kw = small_ib_config(lbmib, "f32")
    s = kw["sheets"][0]
    s.xyz[0, 0, 0] = 1e30
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(2)
        assert e.status() == lbmib.ERR_FIBER_OOB


VARIANTS = {  # (cells per thread, nontemporal mask, blocks per CU)
    "f64": [(1, 0, 0), (1, 1, 0), (2, 1, 0), (2, 3, 4), (3, 3, 0), (1, 3, 0), (1, 3, 4), (1, 1, 8)],
    "f32": [(1, 0, 0), (1, 1, 0), (2, 1, 0), (2, 3, 0), (2, 3, 4), (3, 3, 0), (3, 1, 2), (4, 1, 0), (1, 3, 0),
            (1, 3, 4), (1, 1, 8), (1, 2, 2)],
    "f16": [(1, 1, 0), (2, 3, 0), (2, 0, 2), (3, 3, 0), (4, 1, 0), (8, 1, 0), (8, 0, 0), (1, 3, 4)],
    "bf16": [(1, 1, 0), (2, 3, 0), (2, 1, 4), (3, 3, 0), (3, 3, 4), (4, 1, 0), (8, 1, 0), (1, 3, 4)],
}


@pytest.mark.parametrize("precision", sorted(VARIANTS))
def test_kernel_variants_are_bitwise_identical(lbmib, precision):
    """Every launch shape of the fused kernel computes the same bits (same
    per-cell arithmetic; only the memory access width, cache hints and the
    block-to-row mapping differ)."""
    kw = small_ib_config(lbmib, precision, n=(64, 32, 36))
    ref = None
    for v, nt, bpc in VARIANTS[precision]:
        with lbmib.Engine(**kw) as e:
            e.set_kernel_variant(v, nt, bpc)
            e.init()
            e.step(12)
            assert e.status() == 0
            out = (e.df(), e.fibers(0))
        if ref is None:
            ref = out
        else:
            for a, b in zip(out, ref):
                assert np.array_equal(a, b), (precision, v, nt, bpc)


@pytest.mark.parametrize("precision", ["f32", "bf16", "f64"])
def test_box_row_store_policy_is_bitwise_neutral(lbmib, precision):
    """The cache policy of the box rows' stores (lbmib_set_box_row_stores)
    changes no value: every row kernel (one cell per lane, x pairs, two rows
    per block) with the box rows cached or nontemporal gives the same bits."""
    kw = small_ib_config(lbmib, precision, n=(64, 32, 36))
    ref = None
    for v in ((1, 3, 0), (2, 3, 0), (3, 3, 0)):
        for cached in (0, 1):
            with lbmib.Engine(**kw) as e:
                e.set_kernel_variant(*v)
                e.set_box_row_stores(cached)
                e.init()
                e.step(12)
                assert e.status() == 0
                out = (e.df(), e.fibers(0))
            if ref is None:
                ref = out
            else:
                for a, b in zip(out, ref):
                    assert np.array_equal(a, b), (precision, v, cached)


@pytest.mark.parametrize("bc", [(BC_WALL, BC_PERIODIC, BC_WALL), (BC_PERIODIC, BC_WALL, BC_PERIODIC),
                                (BC_INOUT, BC_WALL, BC_WALL)])
@pytest.mark.parametrize("precision", ["f32", "bf16"])
def test_row_and_pair_paths_agree_on_every_boundary(lbmib, bc, precision):
    """The row kernels -- one cell per lane, x pairs, two rows per block; one
    block per row or persistent -- with the x-edge kernel completing wall /
    periodic x rows, give the same bits on every boundary combination."""
    kw = small_ib_config(lbmib, precision, bc=bc, n=(62, 30, 34))
    outs = []
    for v in ((1, 3, 0), (2, 3, 0), (2, 1, 3), (3, 3, 0), (3, 1, 2)):
        with lbmib.Engine(**kw) as e:
            e.set_kernel_variant(*v)
            e.init()
            e.step(15)
            assert e.status() == 0
            outs.append((e.df(), e.fibers(0)))
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)


LAYOUT_SHAPES = {  # launch shapes (cells per thread, nontemporal mask, blocks per CU) per precision
    "f64": [(1, 3, 0), (2, 3, 0), (3, 3, 0), (1, 3, 4)],
    "f32": [(1, 3, 0), (2, 3, 0), (3, 3, 0), (4, 1, 0), (1, 3, 4)],
    "f16": [(1, 3, 0), (2, 3, 0), (3, 3, 0), (8, 1, 0), (2, 3, 4)],
    "bf16": [(1, 3, 0), (2, 3, 0), (3, 3, 0), (4, 1, 0), (2, 3, 4)],
}


@pytest.mark.parametrize("bc", [(BC_INOUT, BC_PERIODIC, BC_WALL), (BC_WALL, BC_PERIODIC, BC_WALL),
                                (BC_PERIODIC, BC_WALL, BC_PERIODIC)])
@pytest.mark.parametrize("precision", sorted(LAYOUT_SHAPES))
def test_layouts_are_bitwise_identical(lbmib, bc, precision):
    """The rows layout (the 19 populations of each x row side by side) gives
    the same bits as the planes layout, for every launch shape, unsplit and
    split steps, and the materialised fields and force."""
    kw = small_ib_config(lbmib, precision, bc=bc, n=(64, 30, 34))
    ref = None
    ran = {"planes": 0, "rows": 0}
    for layout in ("planes", "rows"):
        runs = [(v, 0) for v in LAYOUT_SHAPES[precision]] + [(LAYOUT_SHAPES[precision][0], 1)]
        for v, split in runs:
            with lbmib.Engine(layout=layout, **kw) as e:
                try:
                    e.set_kernel_variant(*v)
                except lbmib.LbmibError:
                    continue  # a shape this grid and boundary set do not take (same for both layouts)
                e.set_overlap(split)
                e.init()
                e.step(14)
                assert e.status() == 0
                rho, vel = e.macroscopic()
                out = (e.df(), e.fibers(0), e.force(), rho, vel)
            ran[layout] += 1
            if ref is None:
                ref = out
                continue
            for a, b in zip(out, ref):
                assert np.array_equal(a, b), (layout, v, split)
    assert ran["planes"] == ran["rows"] >= 3, ran


def test_batch_mixes_layouts(lbmib):
    """A batch whose instances use either layout evolves each instance as it
    evolves alone."""
    kw = small_ib_config(lbmib, "f32", n=(62, 30, 34))
    cfgs = [dict(kw, layout="rows"), dict(kw, layout="planes")]
    with lbmib.Batch(cfgs) as b:
        b.init()
        b.step(12)
        assert b.status() == 0
        got = [(b.instances[i].df(), b.instances[i].fibers(0)) for i in range(2)]
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(12)
        ref = (e.df(), e.fibers(0))
    for g in got:
        for a, r in zip(g, ref):
            assert np.array_equal(a, r)


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16"])
def test_split_steps_are_bitwise_identical(lbmib, precision):
    """lbmib_step with the IB chain on a second stream (bulk / band split)
    gives the same bits as the single-stream step."""
    kw = small_ib_config(lbmib, precision, n=(64, 32, 36))
    outs = []
    for mode in (0, 1):
        with lbmib.Engine(**kw) as e:
            e.set_overlap(mode)
            e.init()
            e.step(25)
            assert e.status() == 0
            outs.append((e.df(), e.fibers(0), e.force()))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_split_steps_follow_a_moving_sheet(lbmib):
    """A fast inflow carries the sheet across several cells in x: split steps
    stay identical to unsplit steps."""
    sheet = lbmib.Sheet.centred(8, 8, 1.0, (14.0, 16.0, 18.0), 1e-3, 1e-3, jitter=0.05, seed=3)
    kw = dict(nx=96, ny=32, nz=36, bc=(BC_PERIODIC, BC_PERIODIC, BC_WALL), precision="f32", tau=0.8,
              rho0=1.0, u0=(0.08, 0.0, 0.0), quirks=False, sheets=[sheet])
    outs = []
    for mode in (0, 1):
        with lbmib.Engine(**kw) as e:
            e.set_overlap(mode)
            e.init()
            for _ in range(6):
                e.step(50)
            assert e.status() == 0
            outs.append((e.fibers(0), e.df()))
    x0 = kw["sheets"][0].xyz[..., 0].mean()
    print(f"sheet moved {outs[0][0][..., 0].mean() - x0:.2f} cells")
    assert outs[0][0][..., 0].mean() - x0 > 15.0
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_split_steps_follow_a_sheet_off_a_periodic_edge(lbmib):
    """A sheet whose support starts at y = 0 on a periodic y axis (the band
    then takes the whole axis, for the wrapped pulls) drifts along y: the band
    rows follow the support boxes every step, with results identical to
    unsplit steps."""
    sheet = lbmib.Sheet.centred(8, 8, 1.0, (16.0, 5.0, 18.0), 1e-3, 1e-3, jitter=0.05, seed=4)
    kw = dict(nx=32, ny=40, nz=36, bc=(BC_PERIODIC, BC_PERIODIC, BC_WALL), precision="f32", tau=0.8,
              rho0=1.0, u0=(0.0, 0.03, 0.0), quirks=False, sheets=[sheet])
    outs = []
    for mode in (0, 1):
        with lbmib.Engine(**kw) as e:
            e.set_overlap(mode)
            e.init()
            for _ in range(6):
                e.step(40)
            assert e.status() == 0
            outs.append((e.fibers(0), e.df()))
    y0 = kw["sheets"][0].xyz[..., 1].mean()
    print(f"sheet moved {outs[0][0][..., 1].mean() - y0:.2f} cells in y")
    assert outs[0][0][..., 1].mean() - y0 > 3.0
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_split_steps_two_sheets(lbmib):
    sheets = [lbmib.Sheet.centred(10, 10, 1.0, (xc, 16.0, 18.0), 2e-3, 2e-3, jitter=0.05, seed=5 + k)
              for k, xc in enumerate((14.0, 44.0))]
    kw = dict(nx=64, ny=32, nz=36, bc=(BC_INOUT, BC_PERIODIC, BC_WALL), precision="f32", tau=0.7, rho0=1.0,
              u0=(0.01, 0.0, 0.0), quirks=True, sheets=sheets)
    outs = []
    for mode in (0, 1):
        with lbmib.Engine(**kw) as e:
            e.set_overlap(mode)
            e.init()
            e.step(30)
            assert e.status() == 0
            outs.append((e.fibers(0), e.fibers(1), e.df()))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("overlap", [0, 1])
def test_f64_overlapping_sheets_match_oracle(lbmib, overlap):
    """Two sheets 1.5 cells apart in x with shifted y/z extents, so their
    support boxes overlap: cells in both get the contributions of both
    sheets in node order, the spreading candidates of one sheet's tiles hold
    the other's nodes, and the band boxes of split steps overlap.  FP64
    against the oracle, unsplit and split steps."""
    sheets = [lbmib.Sheet.centred(10, 10, 1.0, (14.0, 15.0, 17.0), 2e-3, 2e-3, jitter=0.05, seed=31),
              lbmib.Sheet.centred(8, 9, 1.0, (15.5, 18.0, 15.5), 3e-3, 1e-3, jitter=0.05, seed=32)]
    kw = dict(nx=40, ny=32, nz=36, bc=(BC_INOUT, BC_PERIODIC, BC_WALL), precision="f64", tau=0.7, rho0=1.0,
              u0=(0.01, 0.0, 0.0), quirks=True, sheets=sheets)
    o = oracle_for(kw)
    with lbmib.Engine(**kw) as e:
        e.set_overlap(overlap)
        e.init()
        for chunk in (1, 5, 14):
            e.step(chunk)
            assert o.step(chunk) == 0
            assert e.status() == 0
            for k in range(2):
                assert np.abs(e.fibers(k)[..., :3] - o.nodes(k)[..., :3]).max() < 1e-12, k
            assert rel(e.force(), o.force()) < 1e-9
            rho, vel = e.macroscopic()
            assert rel(vel, o.vel()) < 1e-11
            assert rel(rho, o.rho()) < 1e-14


# ------------------------------------------------- drop-in host layer (CLI)

def test_cli_replays_reference_fixture_refA(tmp_path, fixtures_meta):
    """The reference's command line (refA flags) through the host layer with
    the reference's per-step names (do_thread -> lbmib_* C-ABI) writes fiber
    dumps in the reference's binary format that match the compiled
    reference's own dumps."""
    import subprocess
    meta = fixtures_meta["refA"]
    exe = os.path.join(os.path.dirname(GOLDEN), "..", "distributed-lb-ib_amd", "lbmib")
    out = subprocess.run([os.path.abspath(exe)] + meta["args"] + ["-dump_fiber"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    print(out.stdout[-400:], out.stderr[-400:])
    assert out.returncode == 0
    assert "TOTAL TIME TAKEN IN Seconds:" in out.stdout
    for step in meta["fiber_steps"]:
        ij_got, got = read_fiber_dump(os.path.join(tmp_path, f"Fiber1_dump_step{step}.bin"))
        ij_ref, ref = read_fiber_dump(os.path.join(GOLDEN, f"refA_fiber_step{step}.bin"))
        assert np.array_equal(ij_got, ij_ref)
        assert np.abs(got[:, :3] - ref[:, :3]).max() <= 1e-12
        assert np.abs(got[:, 3:] - ref[:, 3:]).max() <= 1e-9 * max(np.abs(ref[:, 3:]).max(), 1e-12) + 1e-20


def test_threaded_reference_caller_replays_refA(tmp_path, fixtures_meta):
    """A caller written against the reference's API with 2x2x2 pthreads, each
    running do_thread with a barrier after every routine (lbm_ib.c:3115-3563,
    main() at :3836-3955), writes the compiled reference's binary fiber dumps."""
    import subprocess
    from test_host_writers import build_pthread_caller
    meta = fixtures_meta["refA"]
    exe = build_pthread_caller(tmp_path)
    x0, y0, z0 = meta["origin"]
    args = [str(meta["grid"]), str(meta["steps"]), "10", "2", "2", "2", str(meta["rows"]), str(meta["cols"]),
            str(meta["width"]), str(meta["height"]), str(x0), str(y0), str(z0)]
    out = subprocess.run([exe] + args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(out.stdout[-600:], out.stderr[-400:])
    assert out.returncode == 0
    assert f"time {meta['steps'] + 1}, status 0" in out.stdout
    for step in meta["fiber_steps"]:
        ij_got, got = read_fiber_dump(os.path.join(tmp_path, f"Fiber1_dump_step{step}.bin"))
        ij_ref, ref = read_fiber_dump(os.path.join(GOLDEN, f"refA_fiber_step{step}.bin"))
        assert np.array_equal(ij_got, ij_ref)
        assert np.abs(got[:, :3] - ref[:, :3]).max() <= 1e-12
        assert np.abs(got[:, 3:] - ref[:, 3:]).max() <= 1e-9 * max(np.abs(ref[:, 3:]).max(), 1e-12) + 1e-20


def test_cli_fluid_dump_matches_reference_dump(tmp_path, fixtures_meta):
    """-dump_fluid writes the reference's %.24f fluid-cube text
    (save_fluid_sub_grid, lbm_ib.c:332-392) on its box (lbm_ib.c:3511); at
    step 20 of refA every column -- dfeq, df1, df2, rho, velocity, force --
    matches the compiled reference's own Fluid0_dump_step20.dat
    (tests/golden/refA_fluid_step20.npz), parsed by the same parser."""
    import subprocess
    sys.path.insert(0, GOLDEN)
    from make_fixtures import parse_fluid
    meta = fixtures_meta["refA"]
    exe = os.path.join(os.path.dirname(GOLDEN), "..", "distributed-lb-ib_amd", "lbmib")
    out = subprocess.run([os.path.abspath(exe)] + meta["args"] + ["-dump_fluid", "-quiet"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-400:]
    step = meta["fluid_steps"][-1]
    got = parse_fluid(os.path.join(tmp_path, f"Fluid0_dump_step{step}.dat"))
    ref = np.load(os.path.join(GOLDEN, f"refA_fluid_step{step}.npz"))
    np.testing.assert_array_equal(got["xyz"], ref["xyz"])
    for k, tol in (("df1", 1e-12), ("df2", 1e-12), ("dfeq", 1e-12), ("rho", 1e-12), ("vel", 1e-10),
                   ("force", 1e-9)):
        scale = max(np.abs(ref[k]).max(), 1e-30)
        err = np.abs(got[k] - ref[k]).max() / scale
        print(f"fluid dump {k}: max relative difference {err:.2e}")
        assert err <= tol, (k, err)


def test_cli_fiber_text_dump_matches_reference(tmp_path, fixtures_meta):
    """-dump_fiber_text writes the reference's %.24f fiber text
    (save_fiber_sub_grid, lbm_ib.c:560-601); its values match the compiled
    reference's binary fiber dumps of the same steps."""
    import re
    import subprocess
    meta = fixtures_meta["refA"]
    exe = os.path.join(os.path.dirname(GOLDEN), "..", "distributed-lb-ib_amd", "lbmib")
    out = subprocess.run([os.path.abspath(exe)] + meta["args"] + ["-dump_fiber_text", "-quiet"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-400:]
    line = re.compile(r"\(\s*(\d+),\s*(\d+)\):\{([^}]*)\} \|\| ([^|]*) \|\| ([^|]*) \|\| (.*)$")
    for step in meta["fiber_steps"]:
        text = open(os.path.join(tmp_path, f"Fiber1_dump_step{step}.dat")).read().splitlines()
        assert text[0].startswith("Fiber_sheets[0] (i,j): {cord_x, cord_y, cord_z} || SF_x, SF_y, SF_z")
        rows = []
        for ln in text[1:]:
            m = line.match(ln)
            assert m, ln
            vals = [float(v) for g in m.groups()[2:] for v in g.split(",")]
            rows.append([int(m.group(1)), int(m.group(2))] + vals)
        rows = np.array(rows)
        ij_ref, ref = read_fiber_dump(os.path.join(GOLDEN, f"refA_fiber_step{step}.bin"))
        assert np.array_equal(rows[:, :2].astype(int), ij_ref)
        # text order: position, stretch, bend, elastic; binary: position, bend, stretch, elastic
        got = np.concatenate([rows[:, 2:5], rows[:, 8:11], rows[:, 5:8], rows[:, 11:14]], axis=1)
        assert np.abs(got[:, :3] - ref[:, :3]).max() <= 1e-12
        assert np.abs(got[:, 3:] - ref[:, 3:]).max() <= 1e-9 * max(np.abs(ref[:, 3:]).max(), 1e-12) + 1e-20


def test_cli_fieldview_dumps(tmp_path, fixtures_meta):
    """The refA run with the FieldView dumps on (the reference's USE_FV
    files, lbm_ib.c:3100-3112 and :3514-3535): names and grid files once;
    fluid velocity and fiber grid / elastic force files at step 1 and every
    dump steps.  The fluid velocity at step 20 matches the compiled
    reference's own fluid dump on its box (to the %f of the format), and the
    fiber grid matches the binary fiber dump of the same run."""
    import subprocess
    meta = fixtures_meta["refA"]
    n, rows, cols = meta["grid"], meta["rows"], meta["cols"]
    exe = os.path.join(os.path.dirname(GOLDEN), "..", "distributed-lb-ib_amd", "lbmib")
    out = subprocess.run([os.path.abspath(exe)] + meta["args"] + ["-dump_fv", "-dump_fiber", "-quiet"],
                         cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-400:]
    assert open(tmp_path / "FV_nameFluid.nam").read() == "u\nv\nw"
    assert open(tmp_path / "FV_nameFiber.nam").read() == "Afx\nAfy\nAfz"
    head, body = open(tmp_path / "FV_gridFluid.xyz").read().split("\n", 1)
    assert head == f"{n},{n},{n}" and body.count(",") == 3 * n ** 3

    def values(text):
        return np.array([float(v) for v in text.strip().rstrip(",").split(",")])

    for step in meta["fiber_steps"]:
        head, body = open(tmp_path / f"FV_fluid_vel_{step}.f").read().split("\n", 1)
        assert head == f"{n},{n},{n},3"
        vel = values(body).reshape(3, n, n, n)  # [comp, z, y, x]: x fastest
        head, body = open(tmp_path / f"FV_gridFiber{step}.xyz").read().split("\n", 1)
        assert head == f"{rows},{cols},1"
        grid = values(body).reshape(3, cols, rows)  # [comp, col, row]: row fastest
        _, rec = read_fiber_dump(os.path.join(tmp_path, f"Fiber1_dump_step{step}.bin"))
        pos = rec[:, :3].reshape(rows, cols, 3)
        assert np.abs(grid.transpose(2, 1, 0) - pos).max() <= 5e-7
        head, body = open(tmp_path / f"FV_fiber_elastic_force_{step}.f").read().split("\n", 1)
        assert head == f"{rows},{cols},1,3"
        if step == meta["fluid_steps"][-1]:
            fx = np.load(os.path.join(GOLDEN, f"refA_fluid_step{step}.npz"))
            X, Y, Z = fx["xyz"][:, 0], fx["xyz"][:, 1], fx["xyz"][:, 2]
            assert np.abs(vel[:, Z, Y, X].T - fx["vel"]).max() <= 5e-7


# ----------------------------------------------- oracle anchors at full size

def test_c2_full_size_f64_matches_oracle(lbmib, configs):
    """The 256^3 workload (104 x 104 sheet) in FP64 for 5 steps against the
    CPU oracle (lbm_ib.c:1673-2538 semantics): the truth that the full-size
    FP32 and 16-bit accuracy tests compare with is itself anchored at this
    size.  Positions within 1e-12, velocity within 1e-11 and density within
    1e-14 relative."""
    kw, _ = configs.c2(precision="f64")
    steps = 5
    o = oracle_for(kw)
    assert o.step(steps) == 0
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(steps)
        assert e.status() == 0
        got = e.fibers(0)
        rho, vel = e.macroscopic()
    ref = o.nodes(0)
    dx = np.abs(got[..., :3] - ref[..., :3]).max()
    eu, er = rel(vel, o.vel()), rel(rho, o.rho())
    print(f"256^3 f64 vs oracle after {steps} steps: positions {dx:.2e}, velocity {eu:.2e}, density {er:.2e}")
    assert dx < 1e-12
    assert eu < 1e-11
    assert er < 1e-14


# instances of the C4 ensemble the small oracle tests do not reach: tau near
# 0.55, stiffness near 1e-2, edges 32, 48 and 96
C4_PICKS = {3: "32^3, tau 0.83", 18: "48^3, tau 0.57", 29: "smallest tau (32^3, 0.553)",
            19: "largest K (96^3, 7.7e-3)"}


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_c4_instances_match_oracle(lbmib, configs, precision):
    """Four instances of the seeded C4 ensemble stepped together by the
    batched launches for 100 steps, each against its own CPU oracle run
    (lbm_ib.c:3115-3563): F64 positions within 1e-11, velocity within 1e-10
    and density within 1e-13 relative; F32 velocity within 1e-5 of max|u|
    and positions within 1e-6."""
    cfgs, _ = configs.c4(precision=precision)
    sel = [cfgs[i] for i in C4_PICKS]
    steps = 100
    with lbmib.Batch(sel) as b:
        b.init()
        b.step(steps)
        assert b.status() == 0
        for e, (i, what), cfg in zip(b.instances, C4_PICKS.items(), sel):
            o = oracle_for(cfg)
            assert o.step(steps) == 0
            got, ref = e.fibers(0), o.nodes(0)
            rho, vel = e.macroscopic()
            dx = np.abs(got[..., :3] - ref[..., :3]).max()
            eu, er = rel(vel, o.vel()), rel(rho, o.rho())
            print(f"c4 instance {i} ({what}) {precision}: positions {dx:.2e}, velocity {eu:.2e}, density {er:.2e}")
            if precision == "f64":
                assert dx < 1e-11 and eu < 1e-10 and er < 1e-13, (i, dx, eu, er)
            else:
                assert dx < 1e-6 and eu < 1e-5, (i, dx, eu)


# ------------------------------------------------------- batched ensembles

def _ensemble(lbmib, precision):
    """Five instances of uneven size: one wider than 64 cells (two waves per
    row), one without a sheet, different tau and stiffness."""
    out = []
    for k, (n, rows, tau, kk) in enumerate(((24, 8, 0.6, 1e-3), (40, 12, 0.8, 3e-3), (66, 14, 0.55, 1e-4),
                                            (32, 0, 0.7, 0.0), (48, 16, 0.95, 1e-2))):
        sheets = []
        if rows:
            sheets = [lbmib.Sheet.centred(rows, rows, 1.0, (n / 3.0, n / 2.0 + 0.3, n / 2.0 - 0.2), kk, kk,
                                          jitter=0.05, seed=20 + k)]
        out.append(dict(nx=n, ny=n + 2, nz=n - 2, bc=(BC_INOUT, BC_PERIODIC, BC_WALL), precision=precision,
                        tau=tau, rho0=1.0, u0=(0.01, 0.0, 0.0), quirks=True, sheets=sheets))
    return out


@pytest.mark.parametrize("rows", [1, 2])
@pytest.mark.parametrize("overlap", [1, 0])
@pytest.mark.parametrize("precision", ["f32", "bf16", "f64"])
def test_batch_matches_single_instances(lbmib, precision, overlap, rows):
    """Each member of a batched ensemble ends with the bits of the same
    simulation stepped alone, with the instances in two overlapping groups
    (the default) or in one, and one or two rows per lane in the fused
    launch."""
    cfgs = _ensemble(lbmib, precision)
    with lbmib.Batch(cfgs) as b:
        b.set_overlap(overlap)
        b.set_rows_per_lane(rows)
        b.init()
        b.step(5)
        b.step(7)
        assert b.status() == 0
        got = [(e.df(), e.fibers(0) if e.sheets else None) for e in b.instances]
    for cfg, (df, fib) in zip(cfgs, got):
        with lbmib.Engine(**cfg) as e:
            e.init()
            e.step(12)
            assert e.status() == 0
            assert np.array_equal(df, e.df())
            if fib is not None:
                assert np.array_equal(fib, e.fibers(0))


def test_batch_tile_shapes_are_bitwise_identical(lbmib):
    cfgs = _ensemble(lbmib, "f32")
    outs = []
    for wpb, nt, rows, cached in ((2, 3, 1, 0), (4, 1, 1, 1), (8, 0, 1, 1), (2, 3, 2, 1), (4, 1, 2, 0),
                                  (8, 3, 2, 1)):
        with lbmib.Batch(cfgs) as b:
            b.set_tile(wpb, nt)
            b.set_rows_per_lane(rows)
            b.set_box_row_stores(cached)
            b.init()
            b.step(10)
            assert b.status() == 0
            outs.append([e.df() for e in b.instances])
    for o in outs[1:]:
        for a, r in zip(o, outs[0]):
            assert np.array_equal(a, r)


def test_batch_row_pairs_need_even_ny(lbmib):
    cfgs = _ensemble(lbmib, "f32")[:2]
    cfgs[1]["ny"] += 1
    with lbmib.Batch(cfgs) as b:
        with pytest.raises(RuntimeError):
            b.set_rows_per_lane(2)
        with pytest.raises(RuntimeError):
            b.set_rows_per_lane(3)
        b.set_rows_per_lane(1)


def test_c4_workload_batch_matches_single_instances(lbmib, configs):
    """The 64-instance ensemble (48.3 M cells, two overlapping groups) for
    30 steps: instances of every edge length end with the bits of the same
    simulation stepped alone."""
    cfgs, _ = configs.c4("f32")
    pick = {}
    for i, c in enumerate(cfgs):
        pick.setdefault(c["nx"], i)  # the first instance of each edge
    with lbmib.Batch(cfgs) as b:
        b.init()
        b.step(30)
        assert b.status() == 0
        got = {i: (b.instances[i].df(), b.instances[i].fibers(0)) for i in pick.values()}
    for i, (df, fib) in sorted(got.items()):
        with lbmib.Engine(**cfgs[i]) as e:
            e.init()
            e.step(30)
            assert e.status() == 0
            assert np.array_equal(df, e.df()), i
            assert np.array_equal(fib, e.fibers(0)), i
    print(f"c4: instances {sorted(got)} (edges {sorted(pick)}) match their single runs")


def test_batch_rejects_unsupported_members(lbmib):
    cfgs = _ensemble(lbmib, "f32")
    cfgs[1] = dict(cfgs[1], bc=(BC_WALL, BC_PERIODIC, BC_WALL))
    with pytest.raises(lbmib.LbmibError):
        lbmib.Batch(cfgs)
    cfgs = _ensemble(lbmib, "f32")
    cfgs[2] = dict(cfgs[2], precision="f64")
    with pytest.raises(lbmib.LbmibError):
        lbmib.Batch(cfgs)


# -------------------------------------------- the benchmark workloads

def test_c1_workload_f64_matches_oracle(configs):
    """The CPU-runnable workload (64^3, 26x26 sheet, inlet/outlet x, periodic
    y, walls z, tau 0.7) for its full 200 steps: the FP64 GPU path against
    the FP64 oracle."""
    import lbmib
    kw, steps = configs.c1("f64")
    o = oracle_for(kw)
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(steps)
        assert o.step(steps) == 0
        assert e.status() == 0
        rho, vel = e.macroscopic()
        got = e.fibers(0)
    ref = o.nodes(0)
    print(f"c1 f64: velocity {rel(vel, o.vel()):.3e}, rho {rel(rho, o.rho()):.3e}, "
          f"positions {np.abs(got[..., :3] - ref[..., :3]).max():.3e}")
    assert rel(vel, o.vel()) < 1e-10
    assert rel(rho, o.rho()) < 1e-13
    assert np.abs(got[..., :3] - ref[..., :3]).max() < 1e-11


def _run(lbmib, kw, steps):
    with lbmib.Engine(**kw) as e:
        e.init()
        e.step(steps)
        assert e.status() == 0
        rho, vel = e.macroscopic()
        return rho, vel, [e.fibers(k)[..., :3] for k in range(len(kw["sheets"]))]


def l2rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_c2_workload_f32_within_1e5_of_f64(lbmib, configs):
    """256^3 with the 104x104 sheet for its full 1000 steps: FP32 storage
    (FP64 collision) against the FP64 path (itself pinned to the oracle) --
    velocity error relative to max|u| within 1e-5."""
    kw64, steps = configs.c2("f64")
    kw32, _ = configs.c2("f32")
    _, v64, x64 = _run(lbmib, kw64, steps)
    _, v32, x32 = _run(lbmib, kw32, steps)
    err = rel(v32, v64)
    print(f"c2 f32 vs f64: velocity {err:.3e} (L2 {l2rel(v32, v64):.3e}), "
          f"positions {np.abs(x32[0] - x64[0]).max():.3e}")
    assert err < 1e-5
    assert np.abs(x32[0] - x64[0]).max() < 1e-5


@pytest.fixture(scope="module")
def c3_f64(lbmib, configs):
    """The 512x256x256 two-sheet workload in FP64 for its full 2000 steps
    (the FP64 path runs the general per-cell kernel here: its arrays pass the
    4 GiB of the buffer-addressed path)."""
    kw64, steps = configs.c3("f64")
    return _run(lbmib, kw64, steps)


C3_TOL = {  # (precision, dev_ref): relative L2, max relative to the field's max
    ("f16", "eq0"): (1e-3, 5e-3), ("bf16", "eq0"): (2e-2, 1e-1),
    ("f16", "rest"): (1e-2, 3e-2), ("bf16", "rest"): (5e-2, 2e-1),
}


@pytest.mark.parametrize("precision,dev_ref", sorted(C3_TOL))
def test_c3_workload_16bit_vs_f64(lbmib, configs, c3_f64, precision, dev_ref):
    """512x256x256 with two 128x128 sheets, 2000 steps: 16-bit storage of
    f - r (r = feq(rho0, u0x) or w) against the FP64 path, per field:
    relative L2 error and max error relative to the field's max."""
    tol_l2, tol_max = C3_TOL[precision, dev_ref]
    r64, v64, x64 = c3_f64
    kw16, steps = configs.c3(precision)
    kw16["dev_ref"] = dev_ref
    r16, v16, x16 = _run(lbmib, kw16, steps)
    fields = {"ux": (v16[0], v64[0]), "uy": (v16[1], v64[1]), "uz": (v16[2], v64[2]),
              "rho-1": (r16 - 1.0, r64 - 1.0)}
    report = ", ".join(f"{k} L2 {l2rel(a, b):.2e} max {rel(a, b):.2e}" for k, (a, b) in fields.items())
    dx = max(np.abs(a - b).max() for a, b in zip(x16, x64))
    print(f"c3 {precision} {dev_ref} vs f64: velocity L2 {l2rel(v16, v64):.3e} max {rel(v16, v64):.3e}; "
          f"{report}; positions {dx:.3e}")
    assert l2rel(v16, v64) < tol_l2
    assert rel(v16, v64) < tol_max
