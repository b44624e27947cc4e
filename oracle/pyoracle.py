"""ctypes access to the parity checkers (test infrastructure only).

Only tests, the smoke check and the CPU-baseline leg of the benchmark import
this module; the product never does.

* ``OracleSolver`` drives ``liboracle.so``, the FP64 C restatement of
  ff-room's ``FluidSolver::step`` (oracle.c).
* ``RefSolver`` drives ``_ref/libffref.so``, the reference's own C++ classes
  compiled from its sources (ref_driver.cpp); absent when the reference tree
  was not available at build time.

Both take the same ``SceneArrays`` (flags, spacings and boundary lists in the
reference's layouts) and the same parameter dict.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libffref.so")

FLUID, SOLID, INFLOW, OUTFLOW = 0, 1, 2, 3
UPWIND1, QUICK, LAX_WENDROFF = 0, 1, 2


@dataclass
class SceneArrays:
    """One scene in the reference's host layout (Grid.hpp, BoundaryManager.hpp)."""
    nx: int
    ny: int
    nz: int
    lx: float
    ly: float
    lz: float
    flags: np.ndarray                      # uint8 (nx*ny*nz,)
    fans: List[Tuple] = field(default_factory=list)      # (i0,i1,j0,j1,k0,k1,vx,vy,vz)
    outflows: List[Tuple] = field(default_factory=list)  # plain patches (axis,side,a0,a1,b0,b1)
    openings: List[Tuple] = field(default_factory=list)  # (axis,side,a0,a1,b0,b1,T_out)
    # face coordinates (xs, ys, zs) of a stretched grid, sizes n+1 per axis
    # (Grid::set_face_coords); None: uniform
    faces: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None

    def spacings(self):
        if self.faces is not None:  # Grid.cpp:44-46: d[i] = face[i+1] - face[i]
            return tuple(np.asarray(f, dtype=np.float64)[1:] - np.asarray(f, dtype=np.float64)[:-1]
                         for f in self.faces)
        # Grid.cpp:6-37: uniform spacings lx/nx etc.
        dx, dy, dz = self.lx / self.nx, self.ly / self.ny, self.lz / self.nz
        return (np.full(self.nx, dx), np.full(self.ny, dy), np.full(self.nz, dz))

    @property
    def shapes(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        return dict(u=(nx + 1) * ny * nz, v=nx * (ny + 1) * nz, w=nx * ny * (nz + 1),
                    p=nx * ny * nz, T=nx * ny * nz)


DEFAULT_PARAMS = dict(
    # FluidSolver.hpp:11-38 / PoissonSolver.hpp:14-18
    rho=1.2, nu=1.5e-5, dt=0.01, max_steps=500, convergence_tol=1.0,
    poisson_max_iter=1000, poisson_tol=1e-6, advection=UPWIND1, use_multigrid=0,
    thermal=0, T_initial=20.0, T_target=1e30, buoyancy=0, g_accel=9.81,
    beta=1.0 / 293.0, k_thermal=0.026, cp=1005.0,
)


class _OrcFan(C.Structure):
    _fields_ = [("i_min", C.c_int), ("i_max", C.c_int), ("j_min", C.c_int),
                ("j_max", C.c_int), ("k_min", C.c_int), ("k_max", C.c_int),
                ("vel", C.c_double * 3)]


class _OrcOutflow(C.Structure):
    _fields_ = [("axis", C.c_int), ("side", C.c_int), ("a_min", C.c_int),
                ("a_max", C.c_int), ("b_min", C.c_int), ("b_max", C.c_int)]


class _OrcOpening(C.Structure):
    _fields_ = [("axis", C.c_int), ("side", C.c_int), ("a_min", C.c_int),
                ("a_max", C.c_int), ("b_min", C.c_int), ("b_max", C.c_int),
                ("T_outside", C.c_double)]


_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


class _OrcState(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dxv", _dp), ("dyv", _dp), ("dzv", _dp), ("flags", _u8p),
                ("u", _dp), ("v", _dp), ("w", _dp), ("p", _dp), ("T", _dp),
                ("fans", C.POINTER(_OrcFan)), ("n_fans", C.c_int),
                ("outflows", C.POINTER(_OrcOutflow)), ("n_outflows", C.c_int),
                ("openings", C.POINTER(_OrcOpening)), ("n_openings", C.c_int)]


class _OrcParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("nu", C.c_double), ("dt", C.c_double),
                ("max_steps", C.c_int), ("convergence_tol", C.c_double),
                ("poisson_max_iter", C.c_int), ("poisson_tol", C.c_double),
                ("advection", C.c_int), ("use_multigrid", C.c_int), ("thermal", C.c_int),
                ("T_initial", C.c_double), ("T_target", C.c_double), ("buoyancy", C.c_int),
                ("g_accel", C.c_double), ("beta", C.c_double), ("k_thermal", C.c_double),
                ("cp", C.c_double), ("nthreads", C.c_int)]


class _OrcResult(C.Structure):
    _fields_ = [("step", C.c_int), ("velocity_change", C.c_double),
                ("pressure_residual", C.c_double), ("divergence_max", C.c_double),
                ("converged", C.c_int), ("T_mean", C.c_double), ("poisson_iters", C.c_int)]


def _ptr(a: np.ndarray, t=_dp):
    return a.ctypes.data_as(t)


_lib_oracle = None


def oracle_lib():
    global _lib_oracle
    if _lib_oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} not built (make -C oracle)")
        lib = C.CDLL(ORACLE_SO)
        lib.orc_step.argtypes = [C.POINTER(_OrcState), C.POINTER(_OrcParams),
                                 C.POINTER(C.c_int), C.POINTER(_OrcResult)]
        lib.orc_init.argtypes = [C.POINTER(_OrcState), C.POINTER(_OrcParams)]
        lib.orc_apply_laplacian.argtypes = [C.POINTER(_OrcState), _dp, _dp, C.c_int]
        lib.orc_poisson_cg.argtypes = [C.POINTER(_OrcState), _dp, C.c_int, C.c_double,
                                       C.c_int, C.POINTER(C.c_int)]
        lib.orc_poisson_cg.restype = C.c_double
        lib.orc_poisson_mg.argtypes = [C.POINTER(_OrcState), _dp, C.c_int, C.c_double,
                                       C.POINTER(C.c_int)]
        lib.orc_poisson_mg.restype = C.c_double
        lib.orc_poisson_rhs.argtypes = [C.POINTER(_OrcState), C.POINTER(_OrcParams), _dp]
        for name in ("orc_predictor", "orc_buoyancy", "orc_correction",
                     "orc_advect_temperature"):
            getattr(lib, name).argtypes = [C.POINTER(_OrcState), C.POINTER(_OrcParams)]
        for name in ("orc_apply_noslip", "orc_apply_inflow", "orc_apply_outflow",
                     "orc_apply_opening_temperature"):
            getattr(lib, name).argtypes = [C.POINTER(_OrcState)]
        lib.orc_divergence_max.argtypes = [C.POINTER(_OrcState)]
        lib.orc_divergence_max.restype = C.c_double
        lib.orc_T_mean.argtypes = [C.POINTER(_OrcState), C.POINTER(_OrcParams)]
        lib.orc_T_mean.restype = C.c_double
        _lib_oracle = lib
    return _lib_oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def _params_struct(params: dict, nthreads: int = 1) -> _OrcParams:
    p = dict(DEFAULT_PARAMS)
    p.update(params)
    s = _OrcParams()
    for k, _ in _OrcParams._fields_:
        if k == "nthreads":
            s.nthreads = nthreads
        else:
            setattr(s, k, p[k])
    return s


class OracleSolver:
    """FluidSolver semantics over the C restatement; fields live in numpy."""

    def __init__(self, scene: SceneArrays, params: dict, nthreads: int = 1, init=True):
        self.lib = oracle_lib()
        self.scene = scene
        self.params = dict(DEFAULT_PARAMS)
        self.params.update(params)
        self.nthreads = nthreads
        sh = scene.shapes
        self.u = np.zeros(sh["u"]); self.v = np.zeros(sh["v"]); self.w = np.zeros(sh["w"])
        self.p = np.zeros(sh["p"]); self.T = np.zeros(sh["T"])
        self.dxv, self.dyv, self.dzv = scene.spacings()
        self.flags = np.ascontiguousarray(scene.flags, dtype=np.uint8)
        self._fans = (_OrcFan * max(1, len(scene.fans)))()
        for n, f in enumerate(scene.fans):
            self._fans[n].i_min, self._fans[n].i_max = f[0], f[1]
            self._fans[n].j_min, self._fans[n].j_max = f[2], f[3]
            self._fans[n].k_min, self._fans[n].k_max = f[4], f[5]
            self._fans[n].vel[0], self._fans[n].vel[1], self._fans[n].vel[2] = f[6], f[7], f[8]
        # an opening is an outflow patch as well (BoundaryManager.cpp:132-135)
        self._all_outs = [tuple(o[:6]) for o in scene.openings] + list(scene.outflows)
        self._outs = (_OrcOutflow * max(1, len(self._all_outs)))()
        for n, o in enumerate(self._all_outs):
            (self._outs[n].axis, self._outs[n].side, self._outs[n].a_min,
             self._outs[n].a_max, self._outs[n].b_min, self._outs[n].b_max) = o
        self._opens = (_OrcOpening * max(1, len(scene.openings)))()
        for n, o in enumerate(scene.openings):
            (self._opens[n].axis, self._opens[n].side, self._opens[n].a_min,
             self._opens[n].a_max, self._opens[n].b_min, self._opens[n].b_max,
             self._opens[n].T_outside) = o
        self.step_count = C.c_int(0)
        self._pstruct = _params_struct(self.params, nthreads)
        if init:
            self.lib.orc_init(C.byref(self.state()), C.byref(self._pstruct))

    def state(self) -> _OrcState:
        s = _OrcState()
        s.nx, s.ny, s.nz = self.scene.nx, self.scene.ny, self.scene.nz
        s.dxv, s.dyv, s.dzv = _ptr(self.dxv), _ptr(self.dyv), _ptr(self.dzv)
        s.flags = _ptr(self.flags, _u8p)
        s.u, s.v, s.w, s.p, s.T = (_ptr(self.u), _ptr(self.v), _ptr(self.w),
                                   _ptr(self.p), _ptr(self.T))
        s.fans, s.n_fans = self._fans, len(self.scene.fans)
        s.outflows, s.n_outflows = self._outs, len(self._all_outs)
        s.openings, s.n_openings = self._opens, len(self.scene.openings)
        return s

    def step(self) -> dict:
        r = _OrcResult()
        st = self.state()
        self.lib.orc_step(C.byref(st), C.byref(self._pstruct), C.byref(self.step_count),
                          C.byref(r))
        return {k: getattr(r, k) for k, _ in _OrcResult._fields_}

    def laplacian(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        st = self.state()
        self.lib.orc_apply_laplacian(C.byref(st), _ptr(x), _ptr(out), self.nthreads)
        return out

    def poisson_cg(self, rhs: np.ndarray, max_iter: int, tol: float):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        it = C.c_int(0)
        st = self.state()
        res = self.lib.orc_poisson_cg(C.byref(st), _ptr(rhs), max_iter, tol, self.nthreads,
                                      C.byref(it))
        return res, it.value

    def poisson_mg(self, rhs: np.ndarray, max_iter: int, tol: float):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        it = C.c_int(0)
        st = self.state()
        res = self.lib.orc_poisson_mg(C.byref(st), _ptr(rhs), max_iter, tol, C.byref(it))
        return res, it.value

    def rhs(self) -> np.ndarray:
        out = np.zeros(self.scene.nx * self.scene.ny * self.scene.nz)
        st = self.state()
        self.lib.orc_poisson_rhs(C.byref(st), C.byref(self._pstruct), _ptr(out))
        return out

    def call(self, name: str):
        st = self.state()
        fn = getattr(self.lib, name)
        if len(fn.argtypes) == 2:
            fn(C.byref(st), C.byref(self._pstruct))
        else:
            fn(C.byref(st))

    def divergence_max(self) -> float:
        st = self.state()
        return self.lib.orc_divergence_max(C.byref(st))

    def fields(self) -> dict:
        return dict(u=self.u.copy(), v=self.v.copy(), w=self.w.copy(), p=self.p.copy(),
                    T=self.T.copy())

    def set_fields(self, **kw):
        for k, a in kw.items():
            getattr(self, k)[:] = a


class _RefParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("nu", C.c_double), ("dt", C.c_double),
                ("max_steps", C.c_int), ("convergence_tol", C.c_double),
                ("poisson_max_iter", C.c_int), ("poisson_tol", C.c_double),
                ("advection", C.c_int), ("use_multigrid", C.c_int), ("thermal", C.c_int),
                ("T_initial", C.c_double), ("T_target", C.c_double), ("buoyancy", C.c_int),
                ("g_accel", C.c_double), ("beta", C.c_double), ("k_thermal", C.c_double),
                ("cp", C.c_double), ("use_openmp", C.c_int)]


class _RefResult(C.Structure):
    _fields_ = [("step", C.c_int), ("velocity_change", C.c_double),
                ("pressure_residual", C.c_double), ("divergence_max", C.c_double),
                ("converged", C.c_int), ("T_mean", C.c_double)]


_lib_ref = None


def ref_lib():
    global _lib_ref
    if _lib_ref is None:
        if not ref_available():
            raise FileNotFoundError(f"{REF_SO} not built (make -C oracle ref)")
        lib = C.CDLL(REF_SO)
        lib.ref_create.argtypes = [C.c_int] * 3 + [C.c_double] * 3
        lib.ref_create.restype = C.c_void_p
        lib.ref_destroy.argtypes = [C.c_void_p]
        lib.ref_add_solid_box.argtypes = [C.c_void_p] + [C.c_int] * 6
        lib.ref_add_fan.argtypes = [C.c_void_p] + [C.c_int] * 6 + [C.c_double] * 3
        lib.ref_add_outflow.argtypes = [C.c_void_p] + [C.c_int] * 6
        lib.ref_add_opening.argtypes = [C.c_void_p] + [C.c_int] * 6 + [C.c_double]
        lib.ref_get_flags.argtypes = [C.c_void_p, _u8p]
        lib.ref_set_flags.argtypes = [C.c_void_p, _u8p]
        lib.ref_make_solver.argtypes = [C.c_void_p, C.POINTER(_RefParams)]
        lib.ref_set_face_coords.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.ref_step.argtypes = [C.c_void_p, C.POINTER(_RefResult)]
        lib.ref_run.argtypes = [C.c_void_p, C.POINTER(_RefResult)]
        lib.ref_get_fields.argtypes = [C.c_void_p] + [_dp] * 5
        lib.ref_set_fields.argtypes = [C.c_void_p] + [_dp] * 5
        lib.ref_divergence_max.argtypes = [C.c_void_p]
        lib.ref_divergence_max.restype = C.c_double
        lib.ref_poisson_solve.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_double,
                                          C.POINTER(C.c_int)]
        lib.ref_poisson_solve.restype = C.c_double
        for n in ("ref_apply_noslip", "ref_apply_inflow", "ref_apply_outflow",
                  "ref_apply_opening_temperature"):
            getattr(lib, n).argtypes = [C.c_void_p]
        _lib_ref = lib
    return _lib_ref


class RefSolver:
    """The reference's own FluidSolver (compiled from its sources)."""

    def __init__(self, scene: SceneArrays, params: dict, use_openmp: bool = False,
                 flags_override: Optional[np.ndarray] = None):
        self.lib = ref_lib()
        self.scene = scene
        self.h = self.lib.ref_create(scene.nx, scene.ny, scene.nz, scene.lx, scene.ly, scene.lz)
        if scene.faces is not None:
            self._faces = [np.ascontiguousarray(f, dtype=np.float64) for f in scene.faces]
            self.lib.ref_set_face_coords(self.h, *[_ptr(f) for f in self._faces])
        self._bc(scene)
        flags = scene.flags if flags_override is None else flags_override
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        self.lib.ref_set_flags(self.h, _ptr(f, _u8p))
        p = dict(DEFAULT_PARAMS)
        p.update(params)
        self.params = p
        ps = _RefParams()
        for k, _ in _RefParams._fields_:
            setattr(ps, k, int(use_openmp) if k == "use_openmp" else p[k])
        self._ps = ps
        self.lib.ref_make_solver(self.h, C.byref(ps))

    def _bc(self, scene: SceneArrays):
        # Registers the BC lists; flags are then overwritten with scene.flags so
        # both solvers start from identical cell types.
        for f in scene.fans:
            self.lib.ref_add_fan(self.h, *[int(x) for x in f[:6]], *[float(x) for x in f[6:9]])
        for o in scene.outflows:
            self.lib.ref_add_outflow(self.h, *[int(x) for x in o])
        for o in scene.openings:
            self.lib.ref_add_opening(self.h, *[int(x) for x in o[:6]], float(o[6]))

    def __del__(self):
        try:
            self.lib.ref_destroy(self.h)
        except Exception:
            pass

    def step(self) -> dict:
        r = _RefResult()
        self.lib.ref_step(self.h, C.byref(r))
        return {k: getattr(r, k) for k, _ in _RefResult._fields_}

    def run(self) -> dict:
        r = _RefResult()
        self.lib.ref_run(self.h, C.byref(r))
        return {k: getattr(r, k) for k, _ in _RefResult._fields_}

    def fields(self) -> dict:
        sh = self.scene.shapes
        out = {k: np.zeros(n) for k, n in sh.items()}
        self.lib.ref_get_fields(self.h, *[_ptr(out[k]) for k in ("u", "v", "w", "p", "T")])
        return out

    def set_fields(self, u, v, w, p, T):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, p, T)]
        self.lib.ref_set_fields(self.h, *[_ptr(a) for a in arrs])

    def poisson(self, rhs: np.ndarray, use_mg: bool, max_iter: int, tol: float):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        it = C.c_int(0)
        res = self.lib.ref_poisson_solve(self.h, _ptr(rhs), int(use_mg), max_iter, tol,
                                         C.byref(it))
        return res, it.value

    def divergence_max(self) -> float:
        return self.lib.ref_divergence_max(self.h)


def ref_flags_from_calls(nx, ny, nz, lx, ly, lz, solid_boxes=(), fans=(), outflows=(),
                         openings=()) -> np.ndarray:
    """Flags written by the reference BoundaryManager for the given calls, in
    the order solid boxes -> fans -> openings -> outflows (scene.py:89-106)."""
    lib = ref_lib()
    h = lib.ref_create(nx, ny, nz, lx, ly, lz)
    try:
        for b in solid_boxes:
            lib.ref_add_solid_box(h, *[int(x) for x in b])
        for f in fans:
            lib.ref_add_fan(h, *[int(x) for x in f[:6]], *[float(x) for x in f[6:9]])
        for o in openings:
            lib.ref_add_opening(h, *[int(x) for x in o[:6]], float(o[6]))
        for o in outflows:
            lib.ref_add_outflow(h, *[int(x) for x in o])
        out = np.zeros(nx * ny * nz, dtype=np.uint8)
        lib.ref_get_flags(h, _ptr(out, _u8p))
        return out
    finally:
        lib.ref_destroy(h)


# ---- D3Q19 lattice Boltzmann (LBMSolver) ------------------------------------
class _OrcLbm(C.Structure):
    _fields_ = [("tau", C.c_double), ("dt", C.c_double), ("dx_ref", C.c_double),
                ("convergence_tol", C.c_double), ("f", _dp), ("f2", _dp), ("ux", _dp),
                ("uy", _dp), ("uz", _dp)]


_lbm_sigs_done = False


def _lbm_sigs(lib):
    global _lbm_sigs_done
    if not _lbm_sigs_done:
        lib.orc_lbm_init.argtypes = [C.POINTER(_OrcState), C.POINTER(_OrcLbm)]
        lib.orc_lbm_step.argtypes = [C.POINTER(_OrcState), C.POINTER(_OrcLbm),
                                     C.POINTER(C.c_int), C.POINTER(_OrcResult)]
        _lbm_sigs_done = True


class OracleLBM:
    """LBMSolver semantics over the C restatement (oracle.c); the boundary
    lists as the scene registers them (openings, then plain outflows)."""

    def __init__(self, scene: SceneArrays, tau: float, dt: float, convergence_tol: float = 1e-6):
        self.base = OracleSolver(scene, {"dt": dt}, init=False)
        self.lib = self.base.lib
        _lbm_sigs(self.lib)
        nc = scene.nx * scene.ny * scene.nz
        self.f, self.f2 = np.zeros(19 * nc), np.zeros(19 * nc)
        self.ux, self.uy, self.uz = np.zeros(nc), np.zeros(nc), np.zeros(nc)
        self.L = _OrcLbm(tau, dt, scene.lx / scene.nx, convergence_tol, _ptr(self.f),
                         _ptr(self.f2), _ptr(self.ux), _ptr(self.uy), _ptr(self.uz))
        self._st = self.base.state()
        self.lib.orc_lbm_init(C.byref(self._st), C.byref(self.L))

    def step(self) -> dict:
        r = _OrcResult()
        self.lib.orc_lbm_step(C.byref(self._st), C.byref(self.L), C.byref(self.base.step_count),
                              C.byref(r))
        return {k: getattr(r, k) for k, _ in _OrcResult._fields_ if k != "poisson_iters"}

    def fields(self) -> dict:
        return self.base.fields()


class RefLBM:
    """The reference's own LBMSolver (compiled from its sources)."""

    def __init__(self, scene: SceneArrays, tau: float, dt: float, convergence_tol: float = 1e-6,
                 max_steps: int = 1000):
        self.lib = ref_lib()
        self.lib.ref_make_lbm.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double]
        self.lib.ref_lbm_step.argtypes = [C.c_void_p, C.POINTER(_RefResult)]
        self.scene = scene
        self.h = self.lib.ref_create(scene.nx, scene.ny, scene.nz, scene.lx, scene.ly, scene.lz)
        if scene.faces is not None:
            self._faces = [np.ascontiguousarray(f, dtype=np.float64) for f in scene.faces]
            self.lib.ref_set_face_coords(self.h, *[_ptr(f) for f in self._faces])
        for f in scene.fans:
            self.lib.ref_add_fan(self.h, *[int(x) for x in f[:6]], *[float(x) for x in f[6:9]])
        for o in scene.openings:
            self.lib.ref_add_opening(self.h, *[int(x) for x in o[:6]], float(o[6]))
        for o in scene.outflows:
            self.lib.ref_add_outflow(self.h, *[int(x) for x in o])
        f = np.ascontiguousarray(scene.flags, dtype=np.uint8)
        self.lib.ref_set_flags(self.h, _ptr(f, _u8p))
        self.lib.ref_make_lbm(self.h, tau, max_steps, convergence_tol, dt)

    def __del__(self):
        try:
            self.lib.ref_destroy(self.h)
        except Exception:
            pass

    def step(self) -> dict:
        r = _RefResult()
        self.lib.ref_lbm_step(self.h, C.byref(r))
        return {k: getattr(r, k) for k, _ in _RefResult._fields_}

    def fields(self) -> dict:
        sh = self.scene.shapes
        out = {k: np.zeros(n) for k, n in sh.items()}
        self.lib.ref_get_fields(self.h, *[_ptr(out[k]) for k in ("u", "v", "w", "p", "T")])
        return out
