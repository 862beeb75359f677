"""Named workloads of the LBM-IB engine.

``reference_mode`` reproduces the hard-coded physics of the reference's
init_gv (lbm_ib.c:750-773): Re = 150, u_l = 0.001, L = 20, so
tau = 0.5 (6 nu + 1) = 0.5004, Kb_l = 0.005 u^2 L^4, Ks_l = 20 u^2 L, with
walls on y and z, inlet/outlet on x and the force quirks kept.

The ``c1``..``c3`` workloads are the benchmark configurations: inlet/outlet
along x, bounce-back walls on z, periodic y, D3Q19 BGK, rho0 = 1.
"""
import math

import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from lbmib import BC_INOUT, BC_PERIODIC, BC_WALL, Sheet, sheet_positions, uniform_stream  # noqa: E402


def reference_constants():
    """(tau, u_l, rho_l, Ks_l, Kb_l) evaluated in init_gv's order."""
    Re, rho_l, u_l, L_l = 1.5e2, 1.0e0, 0.001, 2.0e1
    nu_l = u_l * L_l / Re
    tau = 0.5 * (6.0 * nu_l + 1.0)
    Kb_l = 0.005 * rho_l * u_l * u_l * math.pow(L_l, 4)
    Ks_l = 2.0e1 * rho_l * u_l * u_l * L_l
    return tau, u_l, rho_l, Ks_l, Kb_l


def reference_mode(n, rows, cols, width, height, origin, precision="f64"):
    """Engine keyword arguments for the reference's own physics on an n^3 grid
    with one sheet generated as gen_fiber_shape does (no jitter)."""
    tau, u_l, rho_l, ks, kb = reference_constants()
    xyz = sheet_positions(rows, cols, width, height, *origin)
    sheet = Sheet(rows, cols, width, height, ks, kb, xyz)
    return dict(nx=n, ny=n, nz=n, bc=(BC_INOUT, BC_WALL, BC_WALL), precision=precision, tau=tau,
                rho0=rho_l, u0=(u_l, 0.0, 0.0), quirks=True, sheets=[sheet], preshift=True,
                preshift_all=False)


BASE_BC = (BC_INOUT, BC_PERIODIC, BC_WALL)


def c1(precision="f64", seed=1, quirks=True):
    """64^3, tau 0.7, u_in 0.01, one 26x26 sheet (spacing 1) centred at
    (16, 32, 32), Ks = Kb = 1e-3, jitter +-0.05, 200 steps."""
    sheet = Sheet.centred(26, 26, 1.0, (16.0, 32.0, 32.0), 1e-3, 1e-3, jitter=0.05, seed=seed)
    return dict(nx=64, ny=64, nz=64, bc=BASE_BC, precision=precision, tau=0.7, rho0=1.0,
                u0=(0.01, 0.0, 0.0), quirks=quirks, sheets=[sheet]), 200


def c2(precision="f32", seed=2, quirks=True):
    """256^3, tau 0.6, u_in 0.01, one 104x104 sheet centred at (64, 128, 128),
    Ks = Kb = 1e-3, jitter +-0.1, 1000 steps."""
    sheet = Sheet.centred(104, 104, 1.0, (64.0, 128.0, 128.0), 1e-3, 1e-3, jitter=0.1, seed=seed)
    return dict(nx=256, ny=256, nz=256, bc=BASE_BC, precision=precision, tau=0.6, rho0=1.0,
                u0=(0.01, 0.0, 0.0), quirks=quirks, sheets=[sheet]), 1000


def c3(precision="bf16", seed=3, quirks=True):
    """512x256x256, two 128x128 sheets centred at x = 128 and x = 256,
    u_in ~ U[0.005, 0.02] from the seed, 2000 steps."""
    u = uniform_stream(seed)
    u_in = 0.005 + 0.015 * next(u)
    sheets = [Sheet.centred(128, 128, 1.0, (xc, 128.0, 128.0), 1e-3, 1e-3, jitter=0.1, seed=seed + k + 1)
              for k, xc in enumerate((128.0, 256.0))]
    return dict(nx=512, ny=256, nz=256, bc=BASE_BC, precision=precision, tau=0.6, rho0=1.0,
                u0=(u_in, 0.0, 0.0), quirks=quirks, sheets=sheets), 2000


C4_EDGES = (32, 48, 64, 96, 128)


def c4(precision="f32", seed=4, instances=64, quirks=True):
    """Batched ensemble: `instances` independent cubic grids with the edge
    drawn from {32, 48, 64, 96, 128}, tau ~ U[0.55, 1.0], one sheet of 16-64
    nodes per side (spacing 1, at most edge - 8 so it fits between the
    walls), Ks = Kb ~ log-uniform [1e-4, 1e-2], centre x ~ U[edge/4, edge/2]
    and y, z placed uniformly where the sheet fits, jitter 0.05; inlet/outlet
    x (u_in 0.01), periodic y, walls z.  Every draw comes from SplitMix64(seed)
    in instance order.  500 steps."""
    u = uniform_stream(seed)
    out = []
    for i in range(instances):
        edge = C4_EDGES[min(int(next(u) * len(C4_EDGES)), len(C4_EDGES) - 1)]
        tau = 0.55 + 0.45 * next(u)
        hi = min(64, edge - 8)
        side = 16 + min(int(next(u) * (hi - 16 + 1)), hi - 16)
        k = 10.0 ** (-4.0 + 2.0 * next(u))
        half = (side - 1) / 2.0
        xc = edge / 4.0 + (edge / 4.0) * next(u)
        lo_c, hi_c = half + 4.0, edge - 1.0 - half - 4.0
        yc = lo_c + (hi_c - lo_c) * next(u)
        zc = lo_c + (hi_c - lo_c) * next(u)
        sheet = Sheet.centred(side, side, 1.0, (xc, yc, zc), k, k, jitter=0.05, seed=seed * 1000 + i)
        out.append(dict(nx=edge, ny=edge, nz=edge, bc=BASE_BC, precision=precision, tau=tau, rho0=1.0,
                        u0=(0.01, 0.0, 0.0), quirks=quirks, sheets=[sheet]))
    return out, 500


def scaled(cfg, n, rows, centre_x):
    """A small variant of a workload with the same physics, for parity tests."""
    kw = dict(cfg)
    kw["nx"] = kw["ny"] = kw["nz"] = n
    s0 = cfg["sheets"][0]
    kw["sheets"] = [Sheet.centred(rows, rows, 1.0, (centre_x, n / 2.0, n / 2.0), s0.ks, s0.kb, jitter=0.05, seed=7)]
    return kw
