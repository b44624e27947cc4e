"""ctypes binding of the C ABI in include/ffroom.h (libffroom.so).

The pybind11 module ``_ffroom_core`` is the reference-compatible surface;
this module is the thin foreign-function view of the same library, used by
the benchmark, the serving loop and the parity tests, and an example of the
binding a non-C++ host would write. Every call checks its status and raises
``FfrError`` with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libffroom.so")

FFR_OK = 0
FLUID, SOLID, INFLOW, OUTFLOW = 0, 1, 2, 3
SNAP_FIELDS, SNAP_CELLS = 0, 1
UPWIND1, QUICK, LAX_WENDROFF = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
POISSON_AUTO, POISSON_CG, POISSON_MGPCG = 0, 1, 2


class FfrError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class Fan(C.Structure):
    _fields_ = [("i_min", C.c_int), ("i_max", C.c_int), ("j_min", C.c_int), ("j_max", C.c_int),
                ("k_min", C.c_int), ("k_max", C.c_int), ("vel", C.c_double * 3)]


class Outflow(C.Structure):
    _fields_ = [("axis", C.c_int), ("side", C.c_int), ("a_min", C.c_int), ("a_max", C.c_int),
                ("b_min", C.c_int), ("b_max", C.c_int)]


class Opening(C.Structure):
    _fields_ = Outflow._fields_ + [("T_outside", C.c_double)]


_dp = C.POINTER(C.c_double)


class Mark(C.Structure):
    _fields_ = [("type", C.c_int), ("i_min", C.c_int), ("i_max", C.c_int), ("j_min", C.c_int),
                ("j_max", C.c_int), ("k_min", C.c_int), ("k_max", C.c_int)]


class SceneDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dxv", _dp), ("dyv", _dp), ("dzv", _dp), ("cell_type", C.POINTER(C.c_uint8)),
                ("fans", C.POINTER(Fan)), ("n_fans", C.c_int),
                ("outflows", C.POINTER(Outflow)), ("n_outflows", C.c_int),
                ("openings", C.POINTER(Opening)), ("n_openings", C.c_int),
                ("marks", C.POINTER(Mark)), ("n_marks", C.c_int)]


class Params(C.Structure):
    _fields_ = [("rho", C.c_double), ("nu", C.c_double), ("dt", C.c_double),
                ("max_steps", C.c_int), ("convergence_tol", C.c_double),
                ("poisson_max_iter", C.c_int), ("poisson_tol", C.c_double),
                ("advection", C.c_int), ("use_openmp", C.c_int), ("use_multigrid", C.c_int),
                ("thermal", C.c_int), ("T_initial", C.c_double), ("T_target", C.c_double),
                ("buoyancy", C.c_int), ("g_accel", C.c_double), ("beta", C.c_double),
                ("k_thermal", C.c_double), ("cp", C.c_double),
                ("precision", C.c_int), ("poisson_method", C.c_int), ("mg_levels", C.c_int),
                ("mg_bf16", C.c_int), ("mg_pre_sweeps", C.c_int), ("mg_level_sweeps", C.c_int),
                ("mg_coarse_sweeps", C.c_int),
                ("inner_tol", C.c_double), ("inner_max_iter", C.c_int), ("graphs", C.c_int),
                ("mg_tail_kb", C.c_int), ("warm_start", C.c_int), ("mg_tail_cells", C.c_int),
                ("mg_tail_groups", C.c_int)]


class StepResult(C.Structure):
    _fields_ = [("step", C.c_int), ("velocity_change", C.c_double),
                ("pressure_residual", C.c_double), ("divergence_max", C.c_double),
                ("converged", C.c_int), ("T_mean", C.c_double), ("poisson_iters", C.c_int),
                ("outer_iters", C.c_int), ("nonfinite", C.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveInfo(C.Structure):
    _fields_ = [("residual", C.c_double), ("iterations", C.c_int), ("outer_iterations", C.c_int)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int),
                ("xcds", C.c_int), ("hbm_bytes", C.c_size_t), ("l2_bytes", C.c_size_t),
                ("lds_bytes_per_cu", C.c_size_t), ("clock_khz", C.c_int)]


class GangStats(C.Structure):
    _fields_ = [("rounds", C.c_uint64), ("launches", C.c_uint64), ("ops", C.c_uint64),
                ("longest_ops", C.c_uint64), ("host_us", C.c_double)]


class ProfEntry(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_long), ("total_ms", C.c_double),
                ("bytes_per_launch", C.c_double)]


STEP_CALLBACK = C.CFUNCTYPE(None, C.POINTER(StepResult), C.c_void_p)

# name -> (restype, argtypes); every symbol declared in include/ffroom.h
_V, _I = None, C.c_int
_P = C.c_void_p
SIGNATURES = {
    "ffr_abi_version": (_I, []),
    "ffr_status_string": (C.c_char_p, [C.c_int]),
    "ffr_params_default": (_V, [C.POINTER(Params)]),
    "ffr_ctx_create": (_I, [C.c_int, C.c_size_t, C.POINTER(_P)]),
    "ffr_ctx_destroy": (_V, [_P]),
    "ffr_last_error": (C.c_char_p, [_P]),
    "ffr_device_query": (_I, [_P, C.POINTER(DeviceInfo)]),
    "ffr_ctx_bytes_in_use": (C.c_size_t, [_P]),
    "ffr_ctx_bytes_cached": (C.c_size_t, [_P]),
    "ffr_ctx_trim": (C.c_size_t, [_P]),
    "ffr_scene_mg_layout": (_I, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.c_uint)]),
    "ffr_scene_mg_tail_groups": (_I, [_P]),
    "ffr_scene_footprint": (C.c_size_t, [C.POINTER(SceneDesc), C.POINTER(Params)]),
    "ffr_scene_create": (_I, [_P, C.POINTER(SceneDesc), C.POINTER(Params), C.POINTER(_P)]),
    "ffr_scene_destroy": (_V, [_P]),
    "ffr_scene_set_params": (_I, [_P, C.POINTER(Params)]),
    "ffr_scene_upload": (_I, [_P, _dp, _dp, _dp, _dp, _dp]),
    "ffr_scene_download": (_I, [_P, _dp, _dp, _dp, _dp, _dp]),
    "ffr_scene_cell_type": (_I, [_P, C.POINTER(C.c_uint8)]),
    "ffr_snapshot_begin": (_I, [_P, C.c_int, C.POINTER(C.c_int)]),
    "ffr_lbm_attach": (_I, [_P, C.c_double, C.c_double, C.c_double, C.c_double]),
    "ffr_lbm_step": (_I, [_P, C.c_void_p]),
    "ffr_snapshot_end": (_I, [_P, C.c_int, C.POINTER(C.c_int), _dp, _dp, _dp, _dp, _dp]),
    "ffr_scene_step_count": (_I, [_P]),
    "ffr_scene_set_step_count": (_I, [_P, C.c_int]),
    "ffr_step": (_I, [_P, C.POINTER(StepResult)]),
    "ffr_run": (_I, [_P, C.c_int, STEP_CALLBACK, _P, C.POINTER(StepResult)]),
    "ffr_divergence_max": (_I, [_P, _dp]),
    "ffr_T_mean": (_I, [_P, _dp]),
    "ffr_apply_laplacian": (_I, [_P, _dp, _dp]),
    "ffr_poisson_solve": (_I, [_P, _dp, _dp, C.POINTER(SolveInfo)]),
    "ffr_build_rhs": (_I, [_P, _dp]),
    "ffr_apply_correction": (_I, [_P, _dp, _dp]),
    "ffr_poisson_solve_noalloc": (_I, [_P, _dp, _dp, C.POINTER(SolveInfo)]),
    "ffr_submit": (_I, [_P, _P, C.c_int, C.POINTER(C.c_uint64)]),
    "ffr_poll": (_I, [_P, C.c_uint64, C.POINTER(StepResult), C.POINTER(C.c_int)]),
    "ffr_drain": (_I, [_P]),
    "ffr_serve_config": (_I, [_P, C.c_int]),
    "ffr_serve_threads": (_I, [_P, C.c_int]),
    "ffr_serve_gangs": (_I, [_P, C.c_int]),
    "ffr_serve_gang_stats": (_I, [_P, C.c_void_p]),
    "ffr_progress": (_I, [_P, C.POINTER(C.c_int)]),
    "ffr_ticket_info": (_I, [_P, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "ffr_release": (_I, [_P, C.c_uint64]),
    "ffr_set_max_blocks": (_I, [C.c_int]),
    "ffr_get_max_blocks": (_I, []),
    "ffr_prof_enable": (_I, [_P, C.c_int]),
    "ffr_prof_reset": (_I, [_P]),
    "ffr_prof_read": (_I, [_P, C.POINTER(ProfEntry), C.c_int, C.POINTER(C.c_int)]),
    "ffr_prof_kernel": (_I, [_P, C.c_char_p, C.c_int, _dp, _dp]),
}

_lib = None


def lib():
    """The loaded libffroom.so with typed signatures (raises if not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run `python ff-room_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def default_params(**kw) -> Params:
    p = Params()
    lib().ffr_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


class Context:
    """One GPU context (stream, HBM budget, profiler)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._l = lib()
        h = C.c_void_p()
        rc = self._l.ffr_ctx_create(device, hbm_budget, C.byref(h))
        if rc != FFR_OK:
            raise FfrError(rc, self._l.ffr_last_error(None).decode())
        self.h = h

    def check(self, rc: int, what: str):
        if rc != FFR_OK:
            raise FfrError(rc, f"{what}: {self._l.ffr_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self._l.ffr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self) -> dict:
        d = DeviceInfo()
        self.check(self._l.ffr_device_query(self.h, C.byref(d)), "ffr_device_query")
        return {k: (getattr(d, k).decode() if isinstance(getattr(d, k), bytes) else getattr(d, k))
                for k, _ in d._fields_}

    def bytes_in_use(self) -> int:
        return self._l.ffr_ctx_bytes_in_use(self.h)

    def bytes_cached(self) -> int:
        """Device bytes of destroyed scenes the context keeps for new ones."""
        return self._l.ffr_ctx_bytes_cached(self.h)

    def trim(self) -> int:
        """Frees the cached blocks; returns the device bytes freed."""
        return self._l.ffr_ctx_trim(self.h)

    # ---- serving (continuous batching of many scenes) ----
    def serve_config(self, slots: int):
        self.check(self._l.ffr_serve_config(self.h, slots), "ffr_serve_config")

    def serve_threads(self, threads: int):
        """Host threads advancing the served scenes (ffr_serve_threads)."""
        self.check(self._l.ffr_serve_threads(self.h, threads), "ffr_serve_threads")

    def serve_gangs(self, gangs: int):
        """Merged launches over groups of scenes (0: every scene on its own stream)."""
        self.check(self._l.ffr_serve_gangs(self.h, gangs), "ffr_serve_gangs")

    def gang_stats(self) -> dict:
        """Totals over all gangs (ffr_gang_stats)."""
        st = GangStats()
        self.check(self._l.ffr_serve_gang_stats(self.h, C.byref(st)), "ffr_serve_gang_stats")
        return {k: getattr(st, k) for k, _ in GangStats._fields_}

    def submit(self, scene: "DeviceScene", max_steps: int) -> int:
        t = C.c_uint64(0)
        self.check(self._l.ffr_submit(self.h, scene.h, max_steps, C.byref(t)), "ffr_submit")
        return t.value

    def poll(self, ticket: int):
        """Advances every active ticket once; (done, last step result)."""
        r, done = StepResult(), C.c_int(0)
        self.check(self._l.ffr_poll(self.h, ticket, C.byref(r), C.byref(done)), "ffr_poll")
        return bool(done.value), r.as_dict()

    def progress(self) -> int:
        """Advances every active ticket once; the number of steps completed."""
        n = C.c_int(0)
        self.check(self._l.ffr_progress(self.h, C.byref(n)), "ffr_progress")
        return n.value

    def drain(self):
        self.check(self._l.ffr_drain(self.h), "ffr_drain")

    def ticket_info(self, ticket: int):
        steps, done = C.c_int(0), C.c_int(0)
        self.check(self._l.ffr_ticket_info(self.h, ticket, C.byref(steps), C.byref(done)),
                   "ffr_ticket_info")
        return steps.value, bool(done.value)

    def release(self, ticket: int):
        self.check(self._l.ffr_release(self.h, ticket), "ffr_release")

    def prof_enable(self, on: bool = True):
        self.check(self._l.ffr_prof_enable(self.h, int(on)), "ffr_prof_enable")

    def prof_reset(self):
        self.check(self._l.ffr_prof_reset(self.h), "ffr_prof_reset")

    def prof_read(self) -> dict:
        n = C.c_int(0)
        self.check(self._l.ffr_prof_read(self.h, None, 0, C.byref(n)), "ffr_prof_read")
        arr = (ProfEntry * max(1, n.value))()
        self.check(self._l.ffr_prof_read(self.h, arr, n.value, C.byref(n)), "ffr_prof_read")
        return {arr[i].name.decode(): dict(launches=arr[i].launches, total_ms=arr[i].total_ms,
                                           bytes_per_launch=arr[i].bytes_per_launch)
                for i in range(n.value)}


def make_desc(nx, ny, nz, dxv, dyv, dzv, flags, fans=(), outflows=(), openings=(), marks=()):
    """Returns (SceneDesc, keepalive): the arrays must outlive the descriptor.
    With flags None the library marks the cells on the device from `marks`
    [(type, i0, i1, j0, j1, k0, k1)] in order."""
    dxv = np.ascontiguousarray(dxv, dtype=np.float64)
    dyv = np.ascontiguousarray(dyv, dtype=np.float64)
    dzv = np.ascontiguousarray(dzv, dtype=np.float64)
    if flags is not None:
        flags = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
    ma = (Mark * max(1, len(marks)))()
    for n, m in enumerate(marks):
        (ma[n].type, ma[n].i_min, ma[n].i_max, ma[n].j_min, ma[n].j_max, ma[n].k_min,
         ma[n].k_max) = [int(x) for x in m]
    fa = (Fan * max(1, len(fans)))()
    for n, f in enumerate(fans):
        fa[n].i_min, fa[n].i_max, fa[n].j_min, fa[n].j_max, fa[n].k_min, fa[n].k_max = [int(x) for x in f[:6]]
        fa[n].vel[0], fa[n].vel[1], fa[n].vel[2] = [float(x) for x in f[6:9]]
    oa = (Outflow * max(1, len(outflows)))()
    for n, o in enumerate(outflows):
        oa[n].axis, oa[n].side, oa[n].a_min, oa[n].a_max, oa[n].b_min, oa[n].b_max = [int(x) for x in o[:6]]
    pa = (Opening * max(1, len(openings)))()
    for n, o in enumerate(openings):
        pa[n].axis, pa[n].side, pa[n].a_min, pa[n].a_max, pa[n].b_min, pa[n].b_max = [int(x) for x in o[:6]]
        pa[n].T_outside = float(o[6])
    d = SceneDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.dxv, d.dyv, d.dzv = _ptr(dxv), _ptr(dyv), _ptr(dzv)
    d.cell_type = flags.ctypes.data_as(C.POINTER(C.c_uint8)) if flags is not None else None
    d.fans, d.n_fans = fa, len(fans)
    d.outflows, d.n_outflows = oa, len(outflows)
    d.openings, d.n_openings = pa, len(openings)
    d.marks, d.n_marks = ma, (0 if flags is not None else len(marks))
    return d, (dxv, dyv, dzv, flags, fa, oa, pa, ma)


class DeviceScene:
    """A scene resident on the GPU (ffr_scene)."""

    def __init__(self, ctx: Context, nx, ny, nz, dxv, dyv, dzv, flags, fans=(), outflows=(),
                 openings=(), params: Optional[Params] = None, marks=()):
        self.ctx = ctx
        self.nx, self.ny, self.nz = nx, ny, nz
        self.params = params if params is not None else default_params()
        self.desc, self._keep = make_desc(nx, ny, nz, dxv, dyv, dzv, flags, fans, outflows, openings,
                                          marks)
        h = C.c_void_p()
        ctx.check(ctx._l.ffr_scene_create(ctx.h, C.byref(self.desc), C.byref(self.params),
                                          C.byref(h)), "ffr_scene_create")
        self.h = h

    @property
    def sizes(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        return dict(u=(nx + 1) * ny * nz, v=nx * (ny + 1) * nz, w=nx * ny * (nz + 1),
                    p=nx * ny * nz, T=nx * ny * nz)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):  # a destroyed context took its scenes along
                self.ctx._l.ffr_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: Params):
        self.ctx.check(self.ctx._l.ffr_scene_set_params(self.h, C.byref(params)),
                       "ffr_scene_set_params")
        self.params = params

    def upload(self, u=None, v=None, w=None, p=None, T=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
                for a in (u, v, w, p, T)]
        self.ctx.check(self.ctx._l.ffr_scene_upload(self.h, *[_ptr(a) for a in arrs]),
                       "ffr_scene_upload")

    def download(self) -> dict:
        out = {k: np.empty(n) for k, n in self.sizes.items()}
        self.ctx.check(self.ctx._l.ffr_scene_download(
            self.h, *[_ptr(out[k]) for k in ("u", "v", "w", "p", "T")]), "ffr_scene_download")
        return out

    def lbm_attach(self, tau: float, dt: float, convergence_tol: float, dx_ref: float):
        """A D3Q19 lattice at rest on this FP64 scene (ffr_lbm_attach)."""
        self.ctx.check(self.ctx._l.ffr_lbm_attach(self.h, tau, dt, convergence_tol, dx_ref),
                       "ffr_lbm_attach")

    def lbm_step(self) -> dict:
        r = StepResult()
        self.ctx.check(self.ctx._l.ffr_lbm_step(self.h, C.byref(r)), "ffr_lbm_step")
        return r.as_dict()

    def cell_type(self) -> np.ndarray:
        """The scene's cell flags (ffr_scene_cell_type), nx*ny*nz bytes."""
        out = np.empty(self.nx * self.ny * self.nz, dtype=np.uint8)
        self.ctx.check(self.ctx._l.ffr_scene_cell_type(
            self.h, out.ctypes.data_as(C.POINTER(C.c_uint8))), "ffr_scene_cell_type")
        return out

    def snapshot_begin(self, cells: bool = False) -> int:
        """Queues a copy of the current fields, or with ``cells`` of the
        FieldSnapshot arrays computed on the device (ffr_snapshot_begin);
        returns its slot."""
        slot = C.c_int(-1)
        kind = SNAP_CELLS if cells else SNAP_FIELDS
        self.ctx.check(self.ctx._l.ffr_snapshot_begin(self.h, kind, C.byref(slot)),
                       "ffr_snapshot_begin")
        self._snap_kind = getattr(self, "_snap_kind", {})
        self._snap_kind[slot.value] = kind
        return slot.value

    def snapshot_end(self, slot: int):
        """(step at capture, arrays) of a queued snapshot (ffr_snapshot_end):
        u, v, w, p, T for fields; vel (n, 3), speed, T for cells (flat)."""
        step = C.c_int(0)
        if getattr(self, "_snap_kind", {}).get(slot) == SNAP_CELLS:
            n = self.nx * self.ny * self.nz
            out = {"vel": np.empty((n, 3)), "speed": np.empty(n), "T": np.empty(n)}
            ptrs = [_ptr(out["vel"]), _ptr(out["speed"]), _ptr(out["T"]), None, None]
        else:
            out = {k: np.empty(n) for k, n in self.sizes.items()}
            ptrs = [_ptr(out[k]) for k in ("u", "v", "w", "p", "T")]
        self.ctx.check(self.ctx._l.ffr_snapshot_end(self.h, slot, C.byref(step), *ptrs),
                       "ffr_snapshot_end")
        return step.value, out

    def step(self) -> dict:
        r = StepResult()
        self.ctx.check(self.ctx._l.ffr_step(self.h, C.byref(r)), "ffr_step")
        return r.as_dict()

    def run(self, max_steps: int) -> dict:
        r = StepResult()
        self.ctx.check(self.ctx._l.ffr_run(self.h, max_steps, STEP_CALLBACK(0), None, C.byref(r)),
                       "ffr_run")
        return r.as_dict()

    def laplacian(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(x)
        self.ctx.check(self.ctx._l.ffr_apply_laplacian(self.h, _ptr(x), _ptr(out)),
                       "ffr_apply_laplacian")
        return out

    def poisson(self, rhs: np.ndarray):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        p = np.empty_like(rhs)
        info = SolveInfo()
        self.ctx.check(self.ctx._l.ffr_poisson_solve(self.h, _ptr(rhs), _ptr(p), C.byref(info)),
                       "ffr_poisson_solve")
        return p, dict(residual=info.residual, iterations=info.iterations,
                       outer_iterations=info.outer_iterations)

    def poisson_noalloc(self, rhs: np.ndarray):
        """The CG solve with the library's fixed scratch array
        (ffr_poisson_solve_noalloc): no device allocation."""
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        p = np.empty_like(rhs)
        info = SolveInfo()
        self.ctx.check(self.ctx._l.ffr_poisson_solve_noalloc(self.h, _ptr(rhs), _ptr(p),
                                                             C.byref(info)),
                       "ffr_poisson_solve_noalloc")
        return p, dict(residual=info.residual, iterations=info.iterations,
                       outer_iterations=info.outer_iterations)

    def build_rhs(self) -> np.ndarray:
        """The step's right-hand side from the current u, v, w taken as u*
        (ffr_build_rhs)."""
        out = np.empty(self.nx * self.ny * self.nz)
        self.ctx.check(self.ctx._l.ffr_build_rhs(self.h, _ptr(out)), "ffr_build_rhs")
        return out

    def apply_correction(self, p: np.ndarray) -> float:
        """Projection of the current u, v, w (taken as u*) with pressure p and
        the post-solve masks (ffr_apply_correction); returns max |u - u*|."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        dv = C.c_double()
        self.ctx.check(self.ctx._l.ffr_apply_correction(self.h, _ptr(p), C.byref(dv)),
                       "ffr_apply_correction")
        return dv.value

    def divergence_max(self) -> float:
        d = C.c_double()
        self.ctx.check(self.ctx._l.ffr_divergence_max(self.h, C.byref(d)), "ffr_divergence_max")
        return d.value

    def T_mean(self) -> float:
        d = C.c_double()
        self.ctx.check(self.ctx._l.ffr_T_mean(self.h, C.byref(d)), "ffr_T_mean")
        return d.value

    def mg_layout(self) -> dict:
        """The multigrid hierarchy: levels, tail_first (-1: no single-workgroup
        kernel), tail_lds_from and the cells of each level."""
        lv, tf, tl = C.c_int(), C.c_int(), C.c_int()
        cells = (C.c_uint * 12)()
        self.ctx.check(self.ctx._l.ffr_scene_mg_layout(self.h, C.byref(lv), C.byref(tf), C.byref(tl),
                                                       cells), "ffr_scene_mg_layout")
        groups = self.ctx._l.ffr_scene_mg_tail_groups(self.h)
        if groups < 0:
            self.ctx.check(groups, "ffr_scene_mg_tail_groups")
        return {"levels": lv.value, "tail_first": tf.value, "tail_lds_from": tl.value,
                "cells": [cells[i] for i in range(lv.value)], "tail_groups": groups}

    def prof_kernel(self, name: str, reps: int = 200):
        """(mean launch ms, algorithmic bytes) of `reps` back-to-back launches."""
        ms, b = C.c_double(), C.c_double()
        self.ctx.check(self.ctx._l.ffr_prof_kernel(self.h, name.encode(), reps, C.byref(ms),
                                                   C.byref(b)), "ffr_prof_kernel")
        return ms.value, b.value

    def step_count(self) -> int:
        return self.ctx._l.ffr_scene_step_count(self.h)


def header_symbols(header_path: str) -> Sequence[str]:
    """Function names declared in a C header (for the export check)."""
    import re
    txt = open(header_path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ffr_[A-Za-z0-9_]+)\s*\(", txt)) - {"ffr_step_callback"})
