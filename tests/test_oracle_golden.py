"""The CPU oracle against the compiled reference's own dumps.

The fixtures under tests/golden/ were written by cube_pthread/lbm_ib.c itself
(built with -DDUMP_VERIFY, see tests/golden/make_fixtures.py).  The oracle in
reference mode must reproduce the fiber dumps bit for bit and the fluid dump
to the precision of its %.24f text.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, read_fiber_dump, reference_sheet_xyz
from oracle.oracle import BC_INOUT, BC_WALL, Oracle, reference_constants


def run_reference_mode(meta, stop_steps):
    tau, u_l, rho_l, ks, kb = reference_constants()
    n = meta["grid"]
    o = Oracle(n, n, n, (BC_INOUT, BC_WALL, BC_WALL), tau, rho_l, (u_l, 0.0, 0.0), quirks=True)
    o.add_sheet(meta["rows"], meta["cols"], meta["width"], meta["height"], ks, kb, reference_sheet_xyz(meta))
    o.init()
    o.preshift(False)
    done = 0
    for s in sorted(stop_steps):
        assert o.step(s - done) == 0
        done = s
        yield s, o


def test_reference_constants():
    tau, u_l, rho_l, ks, kb = reference_constants()
    # init_gv: nu = u L / Re = 0.001*20/150; tau = 0.5 (6 nu + 1)
    assert tau == pytest.approx(0.5004, abs=1e-15)
    assert kb == pytest.approx(0.005 * 1e-6 * 20.0 ** 4, rel=1e-15)
    assert ks == pytest.approx(20.0 * 1e-6 * 20.0, rel=1e-15)


@pytest.mark.parametrize("name", ["refA", "refB"])
def test_oracle_reproduces_reference_fiber_dumps(name, fixtures_meta):
    meta = fixtures_meta[name]
    for step, o in run_reference_mode(meta, meta["fiber_steps"]):
        ij, vals = read_fiber_dump(os.path.join(GOLDEN, f"{name}_fiber_step{step}.bin"))
        rows, cols = meta["rows"], meta["cols"]
        assert np.array_equal(ij[:, 0], np.repeat(np.arange(rows), cols))
        assert np.array_equal(ij[:, 1], np.tile(np.arange(cols), rows))
        mine = o.nodes(0).reshape(-1, 12)
        # positions, bending, stretching and elastic force: identical doubles
        assert np.array_equal(mine, vals), f"{name} step {step}: max diff {np.abs(mine - vals).max()}"


@pytest.mark.parametrize("name", ["refA", "refB"])
def test_oracle_reproduces_reference_fluid_dump(name, fixtures_meta):
    meta = fixtures_meta[name]
    step = meta["fluid_steps"][-1]
    fx = np.load(os.path.join(GOLDEN, f"{name}_fluid_step{step}.npz"))
    for s, o in run_reference_mode(meta, [step]):
        X, Y, Z = fx["xyz"][:, 0], fx["xyz"][:, 1], fx["xyz"][:, 2]
        # %.24f keeps 24 decimals, so the text is exact to < 1e-24 absolute
        tol = 1e-23
        np.testing.assert_allclose(o.vel()[:, Z, Y, X].T, fx["vel"], rtol=0, atol=tol)
        np.testing.assert_allclose(o.rho()[Z, Y, X], fx["rho"], rtol=0, atol=tol)
        np.testing.assert_allclose(o.force()[:, Z, Y, X].T, fx["force"], rtol=0, atol=tol)
        np.testing.assert_allclose(o.dfeq()[:, Z, Y, X].T, fx["dfeq"], rtol=0, atol=tol)
        np.testing.assert_allclose(o.df1()[:, Z, Y, X].T, fx["df1"], rtol=0, atol=tol)
        np.testing.assert_allclose(o.df2()[:, Z, Y, X].T, fx["df2"], rtol=0, atol=tol)


def test_fixture_has_live_forces(fixtures_meta):
    """The refB run is long enough for the sheet to deform: its dumps carry
    non-zero stretching and bending forces, so the force stencils (and the
    quirks) are really exercised by the fixture comparison."""
    ij, vals = read_fiber_dump(os.path.join(GOLDEN, "refB_fiber_step120.bin"))
    assert np.abs(vals[:, 3:6]).max() > 1e-12   # bending
    assert np.abs(vals[:, 6:9]).max() > 1e-9    # stretching
    assert np.abs(vals[:, 0] - vals[0, 0]).max() > 1e-6  # sheet no longer flat
