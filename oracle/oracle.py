"""ctypes view of the FP64 CPU oracle (oracle/lbmib_oracle.c).

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only tests/,
``__graft_entry__.smoke()`` and the cpu_baseline leg of ``bench.py`` import
it; the product (the HIP library and its host layer) never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC_INOUT, BC_WALL, BC_PERIODIC = 0, 1, 2
FIBERNODE_FIELDS = ("x", "y", "z", "bend_x", "bend_y", "bend_z",
                    "stretch_x", "stretch_y", "stretch_z",
                    "elastic_x", "elastic_y", "elastic_z")

_lib = None


class _OwnedView(np.ndarray):
    """ndarray view into oracle memory that keeps the owning Oracle alive."""
    _owner = None


def build():
    """Compile liboracle.so (and nothing else) with the committed Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i, i, i, i, i, i, d, d, d, d, d, i]
        L.orc_destroy.argtypes = [vp]
        L.orc_add_sheet.restype = i
        L.orc_add_sheet.argtypes = [vp, i, i, d, d, d, d, dp]
        for name in ("orc_init", "orc_bending", "orc_stretching", "orc_elastic", "orc_spread",
                     "orc_collide", "orc_stream", "orc_bounceback", "orc_rho_u", "orc_move",
                     "orc_inout", "orc_swap"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = None
        L.orc_preshift.argtypes = [vp, i]
        L.orc_step.argtypes = [vp, i]
        L.orc_step.restype = i
        L.orc_ncells.argtypes = [vp]
        L.orc_ncells.restype = ctypes.c_long
        for name in ("orc_df1", "orc_df2", "orc_dfeq", "orc_rho", "orc_vel", "orc_force",
                     "orc_inlet", "orc_outlet"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = dp
        L.orc_sheet_nodes.argtypes = [vp, i]
        L.orc_sheet_nodes.restype = dp
        L.orc_time.argtypes = [vp]
        L.orc_time.restype = i
        L.orc_error.argtypes = [vp]
        L.orc_error.restype = i
        L.orc_cs2.argtypes = [vp]
        L.orc_cs2.restype = d
        L.orc_cs4.argtypes = [vp]
        L.orc_cs4.restype = d
        L.orc_reference_constants.argtypes = [dp]
        _lib = L
    return _lib


def reference_constants():
    """(tau, u_l, rho_l, Ks_l, Kb_l) of init_gv (lbm_ib.c:750-773)."""
    out = (ctypes.c_double * 5)()
    lib().orc_reference_constants(out)
    return tuple(out)


class Oracle:
    """Sequential FP64 LBM-IB state; arrays are zero-copy numpy views.

    Field layout: ``df1/df2/dfeq[q, z, y, x]``, ``rho[z, y, x]``,
    ``vel/force[d, z, y, x]`` (x fastest, as the GPU layout).
    """

    def __init__(self, nx, ny, nz, bc=(BC_INOUT, BC_WALL, BC_WALL), tau=0.6, rho0=1.0,
                 u0=(0.01, 0.0, 0.0), quirks=True):
        self.L = lib()
        self.shape = (nz, ny, nx)
        self.nx, self.ny, self.nz = nx, ny, nz
        self.h = self.L.orc_create(nx, ny, nz, bc[0], bc[1], bc[2], tau, rho0,
                                   u0[0], u0[1], u0[2], int(bool(quirks)))
        if not self.h:
            raise ValueError("oracle: bad grid/boundary arguments")
        self.sheet_dims = []

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_sheet(self, rows, cols, width, height, ks, kb, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(rows * cols * 3)
        rc = self.L.orc_add_sheet(self.h, rows, cols, width, height, ks, kb,
                                  xyz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc:
            raise RuntimeError(f"oracle add_sheet failed: {rc}")
        self.sheet_dims.append((rows, cols))

    def init(self):
        self.L.orc_init(self.h)

    def preshift(self, all_sheets=False):
        self.L.orc_preshift(self.h, int(all_sheets))

    def step(self, n=1):
        return self.L.orc_step(self.h, n)

    def _own(self, arr):
        v = arr.view(_OwnedView)
        v._owner = self
        return v

    def _view(self, fn, shape):
        ptr = getattr(self.L, fn)(self.h)
        return self._own(np.ctypeslib.as_array(ptr, shape=shape))

    @property
    def ncells(self):
        return self.nx * self.ny * self.nz

    def df1(self):
        return self._view("orc_df1", (19,) + self.shape)

    def df2(self):
        return self._view("orc_df2", (19,) + self.shape)

    def dfeq(self):
        return self._view("orc_dfeq", (19,) + self.shape)

    def rho(self):
        return self._view("orc_rho", self.shape)

    def vel(self):
        return self._view("orc_vel", (3,) + self.shape)

    def force(self):
        return self._view("orc_force", (3,) + self.shape)

    def inlet(self):
        return self._view("orc_inlet", (19,))

    def outlet(self):
        return self._view("orc_outlet", (19,))

    def nodes(self, s=0):
        """Fibernode records of sheet s as an array [rows, cols, 12]."""
        rows, cols = self.sheet_dims[s]
        ptr = self.L.orc_sheet_nodes(self.h, s)
        return self._own(np.ctypeslib.as_array(ptr, shape=(rows, cols, 12)))

    @property
    def time(self):
        return self.L.orc_time(self.h)

    @property
    def error(self):
        return self.L.orc_error(self.h)

    @property
    def cs2(self):
        return self.L.orc_cs2(self.h)

    @property
    def cs4(self):
        return self.L.orc_cs4(self.h)

    def run_substep(self, name):
        getattr(self.L, "orc_" + name)(self.h)
