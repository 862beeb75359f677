"""Closed forms, invariants and hand-computed answers that pin the oracle.

Each check is chosen so that a plausible slip fails it: a dropped 4.5 (c.u)^2
term breaks the second moment of feq, a wrong Guo sign breaks the forcing
moments, a transposed stream offset breaks the brute-force permutation, the
bending x-slip and the stretching end sign show up in the hand stencils.
"""
import numpy as np
import pytest

from oracle.oracle import BC_INOUT, BC_PERIODIC, BC_WALL, Oracle

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
CY = np.array([0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, -1, 1, 1, -1])
CZ = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])


def test_lattice_tables():
    # opposite pairs, isotropy of the D3Q19 weights: Sum w = 1, Sum w c = 0,
    # Sum w c_a c_b = delta_ab / 3
    C = np.stack([CX, CY, CZ], 1)
    assert np.array_equal(C[OPP], -C)
    assert np.isclose(W.sum(), 1.0)
    assert np.allclose(W @ C, 0)
    assert np.allclose((W[:, None, None] * C[:, :, None] * C[:, None, :]).sum(0), np.eye(3) / 3)


def random_state(o, rng, umax=0.05):
    rho = o.rho()
    vel = o.vel()
    rho[...] = 1.0 + 0.1 * rng.standard_normal(rho.shape)
    vel[...] = umax * rng.standard_normal(vel.shape)


def test_equilibrium_moments():
    o = Oracle(10, 10, 10, (BC_PERIODIC, BC_PERIODIC, BC_PERIODIC), 0.8, 1.0, (0, 0, 0))
    o.init()
    rng = np.random.default_rng(1)
    random_state(o, rng)
    o.force()[...] = 0
    o.run_substep("collide")
    feq = o.dfeq().reshape(19, -1)
    rho = o.rho().reshape(-1)
    u = o.vel().reshape(3, -1)
    C = np.stack([CX, CY, CZ], 0).astype(float)
    np.testing.assert_allclose(feq.sum(0), rho, rtol=1e-14)
    np.testing.assert_allclose(C @ feq, rho * u, rtol=0, atol=1e-15)
    # second moment: rho (cs^2 delta_ab + u_a u_b) -- fails without 4.5 (c.u)^2
    for a in range(3):
        for b in range(3):
            m2 = (C[a] * C[b]) @ feq
            expect = rho * ((a == b) / 3.0 + u[a] * u[b])
            np.testing.assert_allclose(m2, expect, rtol=0, atol=1e-14)


def test_collision_fixed_point_and_guo_moments():
    tau = 0.7
    o = Oracle(8, 8, 8, (BC_PERIODIC, BC_PERIODIC, BC_PERIODIC), tau, 1.0, (0, 0, 0))
    o.init()
    rng = np.random.default_rng(2)
    random_state(o, rng)
    o.force()[...] = 0
    o.run_substep("collide")          # writes feq of (rho, u) into dfeq
    feq = o.dfeq().copy()
    o.df1()[...] = feq
    o.run_substep("collide")          # f = feq, F = 0: fixed point
    np.testing.assert_allclose(o.df1(), feq, rtol=1e-14, atol=1e-16)
    F = 1e-3 * rng.standard_normal(o.force().shape)
    o.force()[...] = F
    o.df1()[...] = feq
    o.run_substep("collide")
    S = (o.df1() - feq).reshape(19, -1)
    C = np.stack([CX, CY, CZ], 0).astype(float)
    # Guo source: zero mass, momentum (1 - 1/(2 tau)) F
    np.testing.assert_allclose(S.sum(0), 0, atol=1e-15)
    np.testing.assert_allclose(C @ S, (1 - 0.5 / tau) * F.reshape(3, -1), rtol=1e-12, atol=1e-15)


def brute_pull(df1, bc, lo, hi):
    """Expected post-stream/bounce-back populations, written independently as
    a pull: df2[q](x) = df1[q](x - c_q), or df1[opp q](x) if x - c_q lies
    beyond a wall."""
    nz, ny, nx = df1.shape[1:]
    out = np.zeros_like(df1)
    dims = (nx, ny, nz)
    for z in range(lo[2], hi[2] + 1):
        for y in range(lo[1], hi[1] + 1):
            for x in range(lo[0], hi[0] + 1):
                for q in range(19):
                    s = [x - CX[q], y - CY[q], z - CZ[q]]
                    bb = False
                    for a in range(3):
                        if bc[a] == BC_PERIODIC:
                            s[a] %= dims[a]
                        elif not (lo[a] <= s[a] <= hi[a]):
                            bb = True
                    out[q, z, y, x] = df1[OPP[q], z, y, x] if bb else df1[q, s[2], s[1], s[0]]
    return out


@pytest.mark.parametrize("bc", [(BC_PERIODIC, BC_PERIODIC, BC_PERIODIC), (BC_WALL, BC_WALL, BC_WALL),
                                (BC_PERIODIC, BC_PERIODIC, BC_WALL), (BC_WALL, BC_PERIODIC, BC_WALL)])
def test_stream_bounceback_is_the_pull_permutation(bc):
    n = (9, 8, 10)
    o = Oracle(*n, bc, 0.8, 1.0, (0, 0, 0))
    o.init()
    rng = np.random.default_rng(3)
    o.df1()[...] = rng.standard_normal(o.df1().shape)
    o.df2()[...] = 0
    o.run_substep("stream")
    o.run_substep("bounceback")
    lo = [0 if bc[a] == BC_PERIODIC else 2 for a in range(3)]
    hi = [n[a] - 1 if bc[a] == BC_PERIODIC else n[a] - 3 for a in range(3)]
    expect = brute_pull(o.df1(), bc, lo, hi)
    got = o.df2()
    sl = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
    assert np.array_equal(got[sl], expect[sl])


def test_inout_planes_and_swap():
    o = Oracle(12, 10, 10, (BC_INOUT, BC_WALL, BC_WALL), 0.8, 1.0, (0.02, 0, 0))
    o.init()
    rng = np.random.default_rng(4)
    o.df2()[...] = rng.standard_normal(o.df2().shape)
    o.run_substep("inout")
    tab = o.inlet().copy()
    # the outlet table is sampled at x = N-1 and equals the inlet at init
    assert np.array_equal(o.outlet(), tab)
    for x in (1, 2, 9, 10):
        assert np.array_equal(o.df2()[:, 1:9, 1:9, x], np.broadcast_to(tab[:, None, None], (19, 8, 8)))
    o.run_substep("swap")
    assert np.array_equal(o.df1()[:, 1:9, 1:9, 1:11], o.df2()[:, 1:9, 1:9, 1:11])


def test_moments_against_numpy():
    o = Oracle(10, 9, 8, (BC_PERIODIC, BC_PERIODIC, BC_PERIODIC), 0.8, 1.0, (0, 0, 0))
    o.init()
    rng = np.random.default_rng(5)
    f = W[:, None, None, None] * (1 + 0.05 * rng.standard_normal(o.df2().shape))
    F = 1e-3 * rng.standard_normal(o.force().shape)
    o.df2()[...] = f
    o.force()[...] = F
    o.run_substep("rho_u")
    rho = f.sum(0)
    C = np.stack([CX, CY, CZ], 0).astype(float)
    m = np.tensordot(C, f, axes=(1, 0))
    np.testing.assert_allclose(o.rho(), rho, rtol=1e-15)
    np.testing.assert_allclose(o.vel(), (m + 0.5 * F) / rho, rtol=1e-13, atol=1e-17)


def flat_sheet(rows, cols, ds=1.0, x0=10.0, y0=8.0, z0=8.0):
    xyz = np.zeros((rows, cols, 3))
    for r in range(rows):
        for c in range(cols):
            xyz[r, c] = (x0, y0 + ds * r, z0 + ds * c)
    return xyz


def sheet_oracle(xyz, ks=1.0, kb=1.0, quirks=True):
    rows, cols = xyz.shape[:2]
    o = Oracle(24, 24, 24, (BC_INOUT, BC_WALL, BC_WALL), 0.8, 1.0, (0, 0, 0), quirks=quirks)
    o.add_sheet(rows, cols, float(cols - 1), float(rows - 1), ks, kb, xyz)
    o.init()
    o.run_substep("bending")
    o.run_substep("stretching")
    o.run_substep("elastic")
    return o.nodes(0)


@pytest.mark.parametrize("quirks", [True, False])
def test_flat_sheet_is_force_free(quirks):
    nodes = sheet_oracle(flat_sheet(6, 7), quirks=quirks)
    assert np.abs(nodes[..., 3:]).max() < 1e-14


@pytest.mark.parametrize("quirks", [True, False])
def test_bending_hand_stencil(quirks):
    """Node (2,2) of a flat 6x6 sheet displaced by +1 in x.  Row 2 reads
    x = [0,0,1,0,0,0]; the one-sided/interior stencils of lbm_ib.c:1004-1054
    give row forces [-1, 4, -6, 4, q, 0] and column forces [-1, 4, -6, 4, -1, 0]
    (times Kb/ds^3), where q = 0 with the reference's x-slip at j = cols-2
    (it subtracts x_j instead of x_{j-2}) and q = -1 without it."""
    xyz = flat_sheet(6, 6, x0=0.0)
    xyz[..., 0] = 10.0
    xyz[2, 2, 0] += 1.0
    nodes = sheet_oracle(xyz, ks=0.0, kb=1.0, quirks=quirks)
    bx = nodes[..., 3]
    expect = np.zeros((6, 6))
    row = [-1, 4, -6, 4, 0 if quirks else -1, 0]
    col = [-1, 4, -6, 4, -1, 0]
    expect[2, :] += row
    expect[:, 2] += col
    # the flat background x = 10 contributes through the slip as well: the
    # slip stencil 2x5 - 5x4 + 4x3 - x4 = 0 for a flat line, so only row 2
    # differs
    np.testing.assert_allclose(bx, expect, atol=1e-12)
    assert np.abs(nodes[..., 4:6]).max() < 1e-12


@pytest.mark.parametrize("quirks", [True, False])
def test_stretching_hand_answer(quirks):
    """4x4 sheet, spacing 1, last column moved +0.5 in z.  Along rows the
    spring (2)-(3) is stretched by 0.5: node 2 is pulled +0.5 Ks in z; node 3
    is pulled back (-0.5 Ks) -- the reference's end tangent points the other
    way and pushes it out (+0.5 Ks).  Columns are unchanged."""
    xyz = flat_sheet(4, 4)
    xyz[:, 3, 2] += 0.5
    o_nodes = sheet_oracle(xyz, ks=1.0, kb=0.0, quirks=quirks)
    sz = o_nodes[..., 8]
    np.testing.assert_allclose(sz[:, 2], 0.5, atol=1e-14)
    np.testing.assert_allclose(sz[:, 3], 0.5 if quirks else -0.5, atol=1e-14)
    np.testing.assert_allclose(sz[:, :2], 0.0, atol=1e-14)
    assert np.abs(o_nodes[..., 6:8]).max() < 1e-14


def test_stretching_newton_third_law_without_quirks():
    rng = np.random.default_rng(6)
    xyz = flat_sheet(8, 9) + 0.1 * rng.standard_normal((8, 9, 3))
    nodes = sheet_oracle(xyz, ks=1.0, kb=0.0, quirks=False)
    assert np.abs(nodes[..., 6:9].sum((0, 1))).max() < 1e-13
    q = sheet_oracle(xyz, ks=1.0, kb=0.0, quirks=True)
    assert np.abs(q[..., 6:9].sum((0, 1))).max() > 1e-3  # the end-sign slip breaks it


def test_delta_partition_of_unity_spread_and_interpolation():
    """The 4-point cosine delta sums to one over its support for any
    position, so spreading conserves the total force and interpolating a
    uniform velocity returns it."""
    rng = np.random.default_rng(7)
    xyz = flat_sheet(5, 5, ds=0.7, x0=11.3) + 0.3 * rng.standard_normal((5, 5, 3))
    o = Oracle(24, 24, 24, (BC_INOUT, BC_WALL, BC_WALL), 0.8, 1.0, (0, 0, 0))
    o.add_sheet(5, 5, 2.8, 2.8, 0.0, 0.0, xyz)
    o.init()
    nodes = o.nodes(0)
    Fn = rng.standard_normal((5, 5, 3))
    nodes[..., 9:12] = Fn
    o.run_substep("spread")
    assert o.error == 0
    np.testing.assert_allclose(o.force().sum((1, 2, 3)), Fn.sum((0, 1)), rtol=1e-13)
    u = np.array([0.013, -0.004, 0.007])
    o.vel()[...] = u[:, None, None, None]
    before = nodes[..., :3].copy()
    o.run_substep("move")
    np.testing.assert_allclose(nodes[..., :3] - before, np.broadcast_to(u, before.shape), rtol=1e-12)


def test_fiber_out_of_grid_is_an_error():
    xyz = flat_sheet(4, 4, x0=0.5)  # support starts at floor(0.5-2)+1 = -1
    o = Oracle(16, 16, 16, (BC_INOUT, BC_WALL, BC_WALL), 0.8, 1.0, (0, 0, 0))
    o.add_sheet(4, 4, 3.0, 3.0, 0.0, 0.0, xyz)
    o.init()
    assert o.step(1) == 1  # ORC_ERR_FIBER_OOB replaces the reference's exit(1)


@pytest.mark.parametrize("bc", [(BC_PERIODIC, BC_PERIODIC, BC_PERIODIC), (BC_WALL, BC_WALL, BC_WALL)])
def test_closed_box_conserves_mass(bc):
    rng = np.random.default_rng(8)
    xyz = flat_sheet(6, 6, x0=8.0, y0=6.0, z0=6.0) + 0.1 * rng.standard_normal((6, 6, 3))
    o = Oracle(18, 18, 18, bc, 0.7, 1.0, (0.01, 0, 0), quirks=False)
    o.add_sheet(6, 6, 5.0, 5.0, 1e-2, 1e-2, xyz)
    o.init()
    lo = [0 if b == BC_PERIODIC else 2 for b in bc]
    hi = [17 if b == BC_PERIODIC else 15 for b in bc]
    sl = (slice(None), slice(lo[2], hi[2] + 1), slice(lo[1], hi[1] + 1), slice(lo[0], hi[0] + 1))
    m0 = o.df1()[sl].sum()
    assert o.step(40) == 0
    m1 = o.df1()[sl].sum()
    assert abs(m1 - m0) / m0 < 1e-13
