"""Python binding of the MI355X LBM-IB engine (ctypes over liblbmib.so).

The engine's product surface is the C-ABI in include/lbmib.h; this module
mirrors it for tests and benchmarks.  Arrays cross the boundary as host
numpy buffers; all simulation state stays on the device.

There is no CPU fallback: if liblbmib.so is missing, or no GPU is present,
constructing an Engine raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbmib.so")

OK, ERR_ARG, ERR_HIP, ERR_ALLOC, ERR_FIBER_OOB, ERR_BIN_OVERFLOW, ERR_NONFINITE, ERR_STATE = range(8)
ERR_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_ALLOC", 4: "ERR_FIBER_OOB",
             5: "ERR_BIN_OVERFLOW", 6: "ERR_NONFINITE", 7: "ERR_STATE"}
BC_INOUT, BC_WALL, BC_PERIODIC = 0, 1, 2
PRECISIONS = {"f64": 0, "f32": 1, "f16": 2, "bf16": 3}
STORAGE_BYTES = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2}

# Entry points of include/lbmib.h that take only the context.
CTX_CALLS = (
    "init_eqlbrmdistrfuncDF0", "init_df1", "init_df_inout",
    "compute_bendingforce", "compute_stretchingforce", "compute_elasticforce", "compute_fiber_forces",
    "fiber_SpreadForce", "fluid_get_SpreadForce", "get_influentialdomain_fluid_grid_and_SpreadForce",
    "compute_eqlbrmdistrfuncDF1", "stream_distrfunc", "bounceback_rigidwalls", "copy_inout_to_df2",
    "periodicBC", "compute_rho_and_u", "fluid_SpreadVelocity", "fiber_get_SpreadVelocity",
    "moveFiberSheet", "replace_old_DF", "sync", "status",
)
# Every symbol the header declares (checked by the CPU test suite).
ALL_SYMBOLS = ("lbmib_create", "lbmib_destroy", "lbmib_set_fiber_positions", "lbmib_step",
               "lbmib_get_time", "lbmib_get_fibers", "lbmib_get_macroscopic", "lbmib_get_df",
               "lbmib_get_force", "lbmib_get_totals", "lbmib_get_band_box", "lbmib_time_steps",
               "lbmib_bytes_per_cell", "lbmib_probe_copy_bandwidth",
               "lbmib_set_kernel_variant", "lbmib_set_box_row_stores", "lbmib_set_overlap", "lbmib_timed_kernel_cells",
               "lbmib_batch_create", "lbmib_batch_destroy", "lbmib_batch_instance", "lbmib_batch_init",
               "lbmib_batch_step", "lbmib_batch_set_tile", "lbmib_batch_set_rows_per_lane", "lbmib_batch_set_box_row_stores",
               "lbmib_batch_set_overlap",
               "lbmib_batch_cells", "lbmib_batch_status",
               "lbmib_batch_time_steps", "lbmib_coarse_moments", "lbmib_coarse_sheets", "lbmib_jump",
               "lbmib_jump_commit", "lbmib_jump_revert", "lbmib_get_stream", "lbmib_coarse_state",
               "lbmib_step_coarse") + tuple("lbmib_" + n for n in CTX_CALLS)


class LbmibError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where}: {ERR_NAMES.get(code, code)}")
        self.code = code


class FluidParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("bc_x", ctypes.c_int), ("bc_y", ctypes.c_int), ("bc_z", ctypes.c_int),
                ("precision", ctypes.c_int), ("quirks", ctypes.c_int),
                ("tau", ctypes.c_double), ("rho0", ctypes.c_double),
                ("u0", ctypes.c_double * 3), ("g", ctypes.c_double), ("dev_ref", ctypes.c_int),
                ("layout", ctypes.c_int)]


class SheetParams(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int), ("cols", ctypes.c_int),
                ("width", ctypes.c_double), ("height", ctypes.c_double),
                ("ks", ctypes.c_double), ("kb", ctypes.c_double)]


_lib = None


def load_library(path=LIB_PATH):
    """Loads liblbmib.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build() or make -C distributed-lb-ib_amd)")
    # PyTorch's ROCm build carries its own HIP runtime under the same soname
    # as the system one liblbmib.so links; loaded first, it serves both, so a
    # process that also uses torch (the surrogate's device memory) sees one
    # runtime.  Loaded after liblbmib.so, torch would find no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.lbmib_create.argtypes = [ctypes.POINTER(FluidParams), i, ctypes.POINTER(SheetParams), ctypes.POINTER(vp)]
    L.lbmib_destroy.argtypes = [vp]
    L.lbmib_set_fiber_positions.argtypes = [vp, i, dp]
    L.lbmib_step.argtypes = [vp, i]
    L.lbmib_get_time.argtypes = [vp, ip]
    L.lbmib_get_fibers.argtypes = [vp, i, dp]
    L.lbmib_get_macroscopic.argtypes = [vp, dp, dp]
    L.lbmib_get_df.argtypes = [vp, dp]
    L.lbmib_get_force.argtypes = [vp, dp]
    L.lbmib_get_totals.argtypes = [vp, dp]
    L.lbmib_get_band_box.argtypes = [vp, i, ip]
    L.lbmib_time_steps.argtypes = [vp, i, i, dp, dp]
    L.lbmib_bytes_per_cell.argtypes = [vp, dp]
    L.lbmib_probe_copy_bandwidth.argtypes = [d, i, dp]
    L.lbmib_set_kernel_variant.argtypes = [vp, i, i, i]
    L.lbmib_set_box_row_stores.argtypes = [vp, i]
    L.lbmib_set_overlap.argtypes = [vp, i]
    L.lbmib_timed_kernel_cells.argtypes = [vp, dp]
    L.lbmib_batch_create.argtypes = [i, ctypes.POINTER(FluidParams), ip, ctypes.POINTER(SheetParams),
                                     ctypes.POINTER(vp)]
    L.lbmib_batch_destroy.argtypes = [vp]
    L.lbmib_batch_instance.argtypes = [vp, i, ctypes.POINTER(vp)]
    L.lbmib_batch_init.argtypes = [vp]
    L.lbmib_batch_step.argtypes = [vp, i]
    L.lbmib_batch_set_tile.argtypes = [vp, i, i]
    L.lbmib_batch_set_rows_per_lane.argtypes = [vp, i]
    L.lbmib_batch_set_box_row_stores.argtypes = [vp, i]
    L.lbmib_batch_set_overlap.argtypes = [vp, i]
    L.lbmib_batch_cells.argtypes = [vp, dp, ip]
    L.lbmib_batch_status.argtypes = [vp]
    L.lbmib_batch_time_steps.argtypes = [vp, i, i, dp, dp]
    L.lbmib_coarse_moments.argtypes = [vp, i, vp]
    L.lbmib_coarse_sheets.argtypes = [vp, i, vp]
    L.lbmib_jump.argtypes = [vp, i, vp, vp, i, vp]
    L.lbmib_coarse_state.argtypes = [vp, i, vp, vp]
    L.lbmib_step_coarse.argtypes = [vp, i, vp]
    L.lbmib_jump_commit.argtypes = [vp]
    L.lbmib_jump_revert.argtypes = [vp]
    L.lbmib_get_stream.argtypes = [vp, ctypes.POINTER(vp)]
    for n in CTX_CALLS:
        getattr(L, "lbmib_" + n).argtypes = [vp]
    for n in ALL_SYMBOLS:
        getattr(L, n).restype = i
    _lib = L
    return L


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ------------------------------------------------------------ sheet geometry

def splitmix64(seed):
    """SplitMix64 generator (Steele, Lea, Flood 2014); yields uint64 ints."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    mask = 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        yield z ^ (z >> 31)


def uniform_stream(seed):
    """Doubles in [0, 1) from the top 53 bits of SplitMix64."""
    for z in splitmix64(seed):
        yield (z >> 11) * (1.0 / 9007199254740992.0)


def sheet_positions(rows, cols, width, height, x0, y0, z0, jitter=0.0, seed=0):
    """Node coordinates of gen_fiber_shape (lbm_ib.c:899-906): the sheet is
    normal to x, rows run along y (spacing height/(rows-1)) and columns along
    z (spacing width/(cols-1)).  Optional jitter adds U(-jitter, jitter) to
    each coordinate, drawn in (row, col, xyz) order from SplitMix64(seed).
    Returns float64 [rows, cols, 3]."""
    zdist = width / (cols - 1)
    ydist = height / (rows - 1)
    xyz = np.empty((rows, cols, 3), dtype=np.float64)
    for r in range(rows):
        for c in range(cols):
            xyz[r, c, 0] = x0
            xyz[r, c, 1] = ydist * r + y0
            xyz[r, c, 2] = zdist * c + z0
    if jitter:
        u = uniform_stream(seed)
        flat = xyz.reshape(-1)
        for k in range(flat.size):
            flat[k] += jitter * (2.0 * next(u) - 1.0)
    return xyz


class Sheet:
    """One fiber sheet: parameters plus initial positions [rows, cols, 3]."""

    def __init__(self, rows, cols, width, height, ks, kb, xyz):
        self.rows, self.cols = rows, cols
        self.width, self.height = width, height
        self.ks, self.kb = ks, kb
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(rows, cols, 3)

    @classmethod
    def centred(cls, rows, cols, spacing, centre, ks, kb, jitter=0.0, seed=0):
        width = spacing * (cols - 1)
        height = spacing * (rows - 1)
        xc, yc, zc = centre
        xyz = sheet_positions(rows, cols, width, height, xc, yc - height / 2.0, zc - width / 2.0, jitter, seed)
        return cls(rows, cols, width, height, ks, kb, xyz)


# ------------------------------------------------------------------ engine

DEV_REFS = {"eq0": 0, "rest": 1}
# population layout in HBM: "rows" f[z][y][q][x], "planes" f[q][z][y][x]
LAYOUTS = {"rows": 0, "planes": 1}
DEFAULT_LAYOUT = "rows"


def fluid_params(nx, ny, nz, bc, precision, tau, rho0, u0, quirks, g=0.0, dev_ref="eq0", layout=DEFAULT_LAYOUT):
    return FluidParams(nx, ny, nz, bc[0], bc[1], bc[2], PRECISIONS[precision], int(bool(quirks)),
                       tau, rho0, (ctypes.c_double * 3)(*u0), g, DEV_REFS[dev_ref], LAYOUTS[layout])


class Engine:
    """One simulation on the GPU behind the C-ABI.

    ``precision``: "f64", "f32", "f16" or "bf16".  ``bc``: per-axis boundary
    kinds (x may be BC_INOUT).  ``dev_ref``: what the deviation forms store
    populations relative to, "eq0" (feq(rho0, u0)) or "rest" (w).
    ``layout``: population layout in HBM, "planes" or "rows" (same results).
    Field arrays returned by the getters are indexed [..., z, y, x].
    """

    def __init__(self, nx, ny, nz, *, bc=(BC_INOUT, BC_WALL, BC_WALL), precision="f32", tau=0.6,
                 rho0=1.0, u0=(0.01, 0.0, 0.0), quirks=True, sheets=(), g=0.0, preshift=True,
                 preshift_all=True, dev_ref="eq0", layout=DEFAULT_LAYOUT):
        self.L = load_library()
        self._owner = True
        self._describe(nx, ny, nz, precision, sheets)
        fp = fluid_params(nx, ny, nz, bc, precision, tau, rho0, u0, quirks, g, dev_ref, layout)
        sp = (SheetParams * max(1, len(self.sheets)))()
        for k, s in enumerate(self.sheets):
            sp[k] = SheetParams(s.rows, s.cols, s.width, s.height, s.ks, s.kb)
        h = ctypes.c_void_p()
        self._check(self.L.lbmib_create(ctypes.byref(fp), len(self.sheets), sp, ctypes.byref(h)), "lbmib_create")
        self.h = h
        self.upload_sheets(u0, preshift, preshift_all)

    def _describe(self, nx, ny, nz, precision, sheets):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shape = (nz, ny, nx)
        self.precision = precision
        self.sheets = list(sheets)

    @classmethod
    def _view(cls, L, handle, nx, ny, nz, precision, sheets):
        """A non-owning view of a context that a batch owns."""
        e = cls.__new__(cls)
        e.L = L
        e._owner = False
        e._describe(nx, ny, nz, precision, sheets)
        e.h = handle
        return e

    def upload_sheets(self, u0, preshift=True, preshift_all=True):
        for k, s in enumerate(self.sheets):
            xyz = s.xyz.copy()
            # init_gv shifts the sheet by u_l*dt before the first step
            # (lbm_ib.c:857-862); the reference does it for sheets[0] only.
            if preshift and (preshift_all or k == 0):
                xyz[..., 0] += u0[0] * 1
            self.set_positions(k, xyz)

    @staticmethod
    def _check(rc, where):
        if rc != OK:
            raise LbmibError(rc, where)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owner", True):
                self.L.lbmib_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def call(self, name):
        """Calls one context-only entry point, e.g. ``call("compute_rho_and_u")``."""
        self._check(getattr(self.L, "lbmib_" + name)(self.h), "lbmib_" + name)

    def set_positions(self, sheet, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self._check(self.L.lbmib_set_fiber_positions(self.h, sheet, _dptr(xyz)), "set_fiber_positions")

    def init(self):
        for n in ("init_eqlbrmdistrfuncDF0", "init_df1", "init_df_inout"):
            self.call(n)
        return self

    def step(self, n=1):
        self._check(self.L.lbmib_step(self.h, n), "lbmib_step")

    def sync(self):
        self._check(self.L.lbmib_sync(self.h), "lbmib_sync")

    def status(self):
        return self.L.lbmib_status(self.h)

    @property
    def time(self):
        t = ctypes.c_int()
        self._check(self.L.lbmib_get_time(self.h, ctypes.byref(t)), "get_time")
        return t.value

    def fibers(self, sheet=0):
        s = self.sheets[sheet]
        out = np.empty((s.rows, s.cols, 12), dtype=np.float64)
        self._check(self.L.lbmib_get_fibers(self.h, sheet, _dptr(out)), "get_fibers")
        return out

    def macroscopic(self):
        rho = np.empty(self.shape, dtype=np.float64)
        vel = np.empty((3,) + self.shape, dtype=np.float64)
        self._check(self.L.lbmib_get_macroscopic(self.h, _dptr(rho), _dptr(vel)), "get_macroscopic")
        return rho, vel

    def df(self):
        out = np.empty((19,) + self.shape, dtype=np.float64)
        self._check(self.L.lbmib_get_df(self.h, _dptr(out)), "get_df")
        return out

    def force(self):
        out = np.empty((3,) + self.shape, dtype=np.float64)
        self._check(self.L.lbmib_get_force(self.h, _dptr(out)), "get_force")
        return out

    def totals(self):
        out = np.empty(4, dtype=np.float64)
        self._check(self.L.lbmib_get_totals(self.h, _dptr(out)), "get_totals")
        return out

    def band_box(self, sheet=0):
        b = (ctypes.c_int * 6)()
        self._check(self.L.lbmib_get_band_box(self.h, sheet, b), "get_band_box")
        return tuple(b)

    def time_steps(self, nwarm, nsteps, main_kernel=True):
        tot = ctypes.c_double()
        main = ctypes.c_double()
        self._check(self.L.lbmib_time_steps(self.h, nwarm, nsteps, ctypes.byref(tot),
                                            ctypes.byref(main) if main_kernel else None), "time_steps")
        return tot.value, (main.value if main_kernel else None)

    def set_box_row_stores(self, cached):
        """1 (default): the rows meeting a support box are stored with the
        default cache policy, 0: nontemporally like the others (see
        lbmib_set_box_row_stores); same results."""
        self._check(self.L.lbmib_set_box_row_stores(self.h, int(cached)), "set_box_row_stores")

    def set_kernel_variant(self, cells_per_thread=0, nontemporal=-1, blocks_per_cu=0):
        """Fused-kernel launch shape (see lbmib_set_kernel_variant); all
        variants give identical results."""
        self._check(self.L.lbmib_set_kernel_variant(self.h, cells_per_thread, nontemporal, blocks_per_cu),
                    "set_kernel_variant")

    def set_overlap(self, mode):
        """-1 automatic (unsplit), 0 never split steps, 1 always (see
        lbmib_set_overlap)."""
        self._check(self.L.lbmib_set_overlap(self.h, mode), "set_overlap")

    def timed_kernel_cells(self):
        out = ctypes.c_double()
        self._check(self.L.lbmib_timed_kernel_cells(self.h, ctypes.byref(out)), "timed_kernel_cells")
        return out.value

    # ------------------------------------------------ surrogate coupling
    # d_* arguments are device addresses (ints, e.g. a torch tensor's
    # data_ptr()) of FP32 planes [c][nz/f][ny/f][nx/f]
    def coarse_moments(self, factor, d_out, d_stored=0):
        self._check(self.L.lbmib_coarse_state(self.h, factor, ctypes.c_void_p(d_out), ctypes.c_void_p(d_stored)),
                    "coarse_state")

    def step_coarse(self, factor, d_stored):
        """One step that also writes the coarse stored-moment planes of the
        new state to d_stored (lbmib_step_coarse)."""
        self._check(self.L.lbmib_step_coarse(self.h, factor, ctypes.c_void_p(d_stored)), "step_coarse")

    def coarse_sheets(self, factor, d_out):
        self._check(self.L.lbmib_coarse_sheets(self.h, factor, ctypes.c_void_p(d_out)), "coarse_sheets")

    def jump(self, factor, d_coarse, d_delta, ksteps, d_after=0):
        self._check(self.L.lbmib_jump(self.h, factor, ctypes.c_void_p(d_coarse), ctypes.c_void_p(d_delta), ksteps,
                                      ctypes.c_void_p(d_after)), "jump")

    def stream_handle(self):
        """The context's HIP stream as an int (for torch.cuda.ExternalStream)."""
        h = ctypes.c_void_p()
        self._check(self.L.lbmib_get_stream(self.h, ctypes.byref(h)), "get_stream")
        return h.value or 0

    def jump_commit(self):
        self._check(self.L.lbmib_jump_commit(self.h), "jump_commit")

    def jump_revert(self):
        self._check(self.L.lbmib_jump_revert(self.h), "jump_revert")

    def bytes_per_cell(self):
        b = ctypes.c_double()
        self._check(self.L.lbmib_bytes_per_cell(self.h, ctypes.byref(b)), "bytes_per_cell")
        return b.value


def probe_copy_bandwidth(nbytes=4e9, reps=5):
    """Best-of-reps GB/s of a plain 16-byte-per-lane streaming copy of
    nbytes (read + write), the measured ceiling for HBM-bound kernels."""
    L = load_library()
    out = ctypes.c_double()
    rc = L.lbmib_probe_copy_bandwidth(float(nbytes), int(reps), ctypes.byref(out))
    if rc != OK:
        raise LbmibError(rc, "probe_copy_bandwidth")
    return out.value


class Batch:
    """An ensemble of independent simulations stepped together: each kernel
    of a step runs as one launch over all instances (lbmib_batch_*).

    ``configs``: one dict of Engine keyword arguments per instance (nx, ny,
    nz, bc, precision, tau, rho0, u0, quirks, sheets, g).  ``instances`` are
    Engine views of the member contexts for readbacks and status.
    """

    def __init__(self, configs, preshift=True, preshift_all=True):
        self.L = load_library()
        n = len(configs)
        fps = (FluidParams * n)()
        nsh = (ctypes.c_int * n)()
        all_sheets = [s for c in configs for s in c.get("sheets", ())]
        sps = (SheetParams * max(1, len(all_sheets)))()
        for i, c in enumerate(configs):
            fps[i] = fluid_params(c["nx"], c["ny"], c["nz"], c.get("bc", (BC_INOUT, BC_WALL, BC_WALL)),
                                  c.get("precision", "f32"), c.get("tau", 0.6), c.get("rho0", 1.0),
                                  c.get("u0", (0.01, 0.0, 0.0)), c.get("quirks", True), c.get("g", 0.0),
                                  c.get("dev_ref", "eq0"), c.get("layout", DEFAULT_LAYOUT))
            nsh[i] = len(c.get("sheets", ()))
        for k, s in enumerate(all_sheets):
            sps[k] = SheetParams(s.rows, s.cols, s.width, s.height, s.ks, s.kb)
        h = ctypes.c_void_p()
        rc = self.L.lbmib_batch_create(n, fps, nsh, sps, ctypes.byref(h))
        if rc != OK:
            raise LbmibError(rc, "lbmib_batch_create")
        self.h = h
        self.instances = []
        for i, c in enumerate(configs):
            ch = ctypes.c_void_p()
            self._check(self.L.lbmib_batch_instance(self.h, i, ctypes.byref(ch)), "batch_instance")
            e = Engine._view(self.L, ch, c["nx"], c["ny"], c["nz"], c.get("precision", "f32"), c.get("sheets", ()))
            e.upload_sheets(c.get("u0", (0.01, 0.0, 0.0)), preshift, preshift_all)
            self.instances.append(e)

    _check = staticmethod(Engine._check)

    def close(self):
        if getattr(self, "h", None):
            for e in self.instances:
                e.h = None
            self.L.lbmib_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def init(self):
        self._check(self.L.lbmib_batch_init(self.h), "batch_init")
        return self

    def step(self, n=1):
        self._check(self.L.lbmib_batch_step(self.h, n), "batch_step")

    def status(self):
        return self.L.lbmib_batch_status(self.h)

    def set_tile(self, waves_per_block, nontemporal=-1):
        self._check(self.L.lbmib_batch_set_tile(self.h, waves_per_block, nontemporal), "batch_set_tile")

    def set_rows_per_lane(self, rows):
        """1 (default) or 2: each lane of the fused launch updates one cell of
        two neighbouring rows; same results."""
        self._check(self.L.lbmib_batch_set_rows_per_lane(self.h, int(rows)), "batch_set_rows_per_lane")

    def set_box_row_stores(self, cached):
        """See Engine.set_box_row_stores; same results."""
        self._check(self.L.lbmib_batch_set_box_row_stores(self.h, int(cached)), "batch_set_box_row_stores")

    def set_overlap(self, on):
        """Two instance groups whose IB kernels overlap each other's fluid
        update (1, the default) or one group (0); same results."""
        self._check(self.L.lbmib_batch_set_overlap(self.h, int(on)), "batch_set_overlap")

    def cells(self):
        c, m = ctypes.c_double(), ctypes.c_int()
        self._check(self.L.lbmib_batch_cells(self.h, ctypes.byref(c), ctypes.byref(m)), "batch_cells")
        return c.value, m.value

    def time_steps(self, nwarm, nsteps, fused=True):
        """(total ms of nsteps, mean ms of one batched fused launch or None)."""
        t, f = ctypes.c_double(), ctypes.c_double()
        self._check(self.L.lbmib_batch_time_steps(self.h, nwarm, nsteps, ctypes.byref(t),
                                                  ctypes.byref(f) if fused else None), "batch_time_steps")
        return t.value, (f.value if fused else None)
