"""ctypes binding of the device library ``libjrhip.so`` (declared in ``include/jrhip.h``).

The library is the only compute path: there is no CPU fallback.  A missing or
unloadable library raises ``DeviceUnavailable`` at first use.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from .mesh.atlas import Atlas

LIB_PATH = Path(__file__).with_name("libjrhip.so")
NOALLOC_HISTORY_MAX = 65536      # JR_NOALLOC_HISTORY_MAX of include/jrhip.h

FLUX_CODES = {"hllc": 0, "roe": 1}
LIMITER_CODES = {"minmod": 0, "none": 1, "van_albada": 2, "van_albada_smooth": 3}
FAR_FIELD_CODES = {"characteristic": 0, "characteristic_point_vortex": 1}

EXPORTS = (
    "jr_ctx_create", "jr_ctx_destroy", "jr_set_freestream", "jr_set_state", "jr_get_state", "jr_residual",
    "jr_residual_dev", "jr_tangent", "jr_tangent_dev", "jr_wave_speeds", "jr_precond_factor", "jr_precond_apply",
    "jr_newton_init", "jr_newton_block", "jr_get_scalars", "jr_set_cfl", "jr_forces", "jr_timer_start",
    "jr_timer_stop", "jr_synchronize", "jr_device_alloc", "jr_device_free", "jr_copy_to_device", "jr_copy_to_host",
    "jr_device_count", "jr_profile", "jr_set_graphs", "jr_probe_arm", "jr_probe_read", "jr_newton_block_async",
    "jr_newton_block_poll", "jr_newton_block_wait", "jr_ctx_bytes", "jr_mem_info", "jr_ctx_create3", "jr_forces3",
    "jr_set_geometry", "jr_set_state_case", "jr_newton_init_case", "jr_linear_solve", "jr_rk_block", "jr_last_error",
    "jr_newton_block_noalloc", "jr_wall_loads", "jr_tangent_freestream", "jr_forces_tangent", "jr_park_case",
    "jr_set_tile_tangent", "jr_forces3_tangent", "jr_wall_loads_tangent",
)


class DeviceUnavailable(RuntimeError):
    """The HIP library is missing, or no GPU is visible."""


class DeviceError(RuntimeError):
    """A call into the HIP library failed."""


_P = C.POINTER
_d, _i64, _i32, _u8 = _P(C.c_double), _P(C.c_int64), _P(C.c_int32), _P(C.c_uint8)


class GridDesc(C.Structure):
    _fields_ = [
        ("NI", C.c_int32), ("NJ", C.c_int32),
        ("area", _d), ("face_i", _d), ("face_j", _d), ("centres", _d), ("distance", _d),
        ("interior", _u8), ("neighbours", _i64),
        ("n_task", C.c_int32 * 2), ("task_dest", _i64 * 2), ("task_src", _i64 * 2), ("task_kind", _i32 * 2),
        ("task_normal", _d * 2),
        ("n_lines", C.c_int32), ("n_slots", C.c_int32),
        ("line_start", _i64), ("line_cells", _i64), ("line_up_face", _i64), ("line_up_sign", _i64),
        ("line_down_face", _i64), ("line_down_sign", _i64),
        ("n_walls", C.c_int32),
        ("wall_cell", _i64), ("wall_family", _i64), ("wall_face", _i64), ("wall_into", _d), ("wall_mid", _d),
        ("task_mid", _d * 2), ("vortex_x", C.c_double), ("vortex_y", C.c_double),
    ]


class GridDesc3(C.Structure):
    _fields_ = [
        ("NI", C.c_int32), ("NJ", C.c_int32), ("NK", C.c_int32),
        ("volume", _d), ("face", _d * 3), ("centres", _d), ("distance", _d),
        ("interior", _u8), ("neighbours", _i64),
        ("n_ghost", C.c_int32), ("ghost_dest", _i64), ("ghost_root", _i64), ("ghost_nops", _i64),
        ("ghost_ops", _i64), ("n_tasks", C.c_int32), ("task_kind", _i32), ("task_normal", _d),
        ("n_lines", C.c_int32), ("n_slots", C.c_int32),
        ("line_start", _i64), ("line_cells", _i64), ("line_up_face", _i64), ("line_up_sign", _i64),
        ("line_down_face", _i64), ("line_down_sign", _i64),
        ("n_walls", C.c_int32),
        ("wall_cell", _i64), ("wall_family", _i64), ("wall_face", _i64), ("wall_into", _d), ("wall_mid", _d),
    ]


class Numerics(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("kappa", C.c_double), ("flux", C.c_int32), ("limiter", C.c_int32),
        ("viscous", C.c_int32), ("sa", C.c_int32), ("sutherland", C.c_double), ("c_cr1", C.c_double),
        ("jacobian_exact", C.c_int32), ("mixed", C.c_int32), ("line_solver", C.c_int32),
        ("krylov_subspace", C.c_int32),
        ("krylov_restarts", C.c_int32), ("krylov_tol", C.c_double), ("krylov_success", C.c_double),
        ("cfl_growth", C.c_double), ("cfl_cutback", C.c_double), ("cfl_min", C.c_double), ("cfl_max", C.c_double),
        ("max_relative_update", C.c_double), ("far_field", C.c_int32),
    ]


_LIB = None


def library(path: str | os.PathLike | None = None):
    """Load ``libjrhip.so`` (once)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.is_file():
        raise DeviceUnavailable(f"{p} is not built; run __graft_entry__.build() or make -C jax-rans_amd/csrc")
    lib = C.CDLL(str(p))
    for name in EXPORTS:
        fn = getattr(lib, name)
        fn.restype = C.c_char_p if name == "jr_last_error" else C.c_int
    if path is None:
        _LIB = lib
    return lib


def mem_info() -> tuple[int, int]:
    """(free, total) bytes of device memory."""
    f, t = C.c_size_t(), C.c_size_t()
    _check(library().jr_mem_info(C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


def _check(rc: int) -> None:
    if rc != 0:
        raise DeviceError(library().jr_last_error().decode())


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def device_count() -> int:
    n = C.c_int(0)
    rc = library().jr_device_count(C.byref(n))
    return int(n.value) if rc == 0 else 0


class Context:
    """One device context: an atlas, its numerics, ``n_cases`` cases.

    Arrays handed in are copied by the library; the state stays on the device
    between ``newton_block`` calls.
    """

    def __init__(self, atlas: Atlas, *, gamma: float, flux: str, limiter: str, kappa: float = 0.0,
                 viscous: bool = False, sa: bool = False, sutherland: float | None = None, c_cr1: float = 0.0,
                 implicit=None, n_cases: int = 1, far_field: str = "characteristic"):
        lib = library()
        if device_count() < 1:
            raise DeviceUnavailable("no HIP device is visible")
        self.atlas, self.n_cases = atlas, int(n_cases)
        self.dim = getattr(atlas, "dim", 2)
        self.nvar = (5 if sa else 4) + (self.dim == 3)
        self.viscous = bool(viscous)
        self.shape = (self.n_cases, self.nvar) + tuple(atlas.shape)
        self._keep = []

        def arr(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a

        if far_field not in FAR_FIELD_CODES:
            raise DeviceError(f"unknown far-field model {far_field!r}; known: {sorted(FAR_FIELD_CODES)}")
        if far_field != "characteristic" and self.dim == 3:
            raise DeviceError(f"far field model {far_field!r} is defined in 2D only")
        self.far_field = far_field
        g = self._desc3(atlas, arr) if self.dim == 3 else self._desc2(atlas, arr, far_field != "characteristic")
        from .core.config import ImplicitConfig
        cfg = implicit if implicit is not None else ImplicitConfig()
        n = Numerics()
        n.gamma, n.kappa = gamma, kappa
        n.flux, n.limiter = FLUX_CODES[flux], LIMITER_CODES[limiter]
        n.viscous, n.sa = int(viscous), int(sa)
        n.sutherland = -1.0 if sutherland is None else float(sutherland)
        n.c_cr1 = c_cr1
        n.jacobian_exact = int(cfg.jacobian == "exact")
        n.mixed = int(cfg.precision == "mixed")
        n.line_solver = {"cyclic_reduction": 0, "thomas": 1}[cfg.line_solver]
        n.krylov_subspace, n.krylov_restarts = cfg.krylov_subspace, cfg.krylov_restarts
        n.krylov_tol, n.krylov_success = cfg.krylov_tol, cfg.krylov_success
        n.cfl_growth, n.cfl_cutback, n.cfl_min, n.cfl_max = cfg.cfl_growth, cfg.cfl_cutback, cfg.cfl_min, cfg.cfl_max
        n.max_relative_update = -1.0 if cfg.max_relative_update is None else cfg.max_relative_update
        n.far_field = FAR_FIELD_CODES[far_field]
        if cfg.preconditioner != "line_j":
            raise DeviceError("only the line_j preconditioner is implemented on the device")
        self.config = cfg
        h = C.c_void_p()
        create = lib.jr_ctx_create3 if self.dim == 3 else lib.jr_ctx_create
        _check(create(C.byref(g), C.byref(n), self.n_cases, C.byref(h)))
        self._h = h
        self._lib = lib

    @staticmethod
    def _desc2(atlas, arr, vortex: bool = False) -> GridDesc:
        g = GridDesc()
        g.NI, g.NJ = atlas.shape
        g.area = _ptr(arr(atlas.area, np.float64), C.c_double)
        g.face_i = _ptr(arr(atlas.face_i.reshape(2, -1), np.float64), C.c_double)
        g.face_j = _ptr(arr(atlas.face_j.reshape(2, -1), np.float64), C.c_double)
        g.centres = _ptr(arr(atlas.centres, np.float64), C.c_double)
        g.distance = _ptr(arr(np.where(np.isfinite(atlas.distance), atlas.distance, 1e300), np.float64), C.c_double)
        g.interior = _ptr(arr(atlas.interior, np.uint8), C.c_uint8)
        g.neighbours = _ptr(arr(atlas.neighbours, np.int64), C.c_int64)
        for s, st in enumerate(atlas.stages):
            g.n_task[s] = st.dest.size
            g.task_dest[s] = _ptr(arr(st.dest, np.int64), C.c_int64)
            g.task_src[s] = _ptr(arr(st.src, np.int64), C.c_int64)
            g.task_kind[s] = _ptr(arr(st.kind, np.int32), C.c_int32)
            g.task_normal[s] = _ptr(arr(st.normal, np.float64), C.c_double)
            if vortex:
                g.task_mid[s] = _ptr(arr(st.midpoint, np.float64), C.c_double)
        if vortex:
            from .mesh.atlas import quarter_chord
            g.vortex_x, g.vortex_y = quarter_chord(atlas.walls)
        Context._lines_walls(g, atlas, arr)
        return g

    @staticmethod
    def _desc3(atlas, arr) -> GridDesc3:
        g = GridDesc3()
        g.NI, g.NJ, g.NK = atlas.shape
        g.volume = _ptr(arr(atlas.volume, np.float64), C.c_double)
        for a in range(3):
            g.face[a] = _ptr(arr(atlas.faces[a].reshape(3, -1), np.float64), C.c_double)
        g.centres = _ptr(arr(atlas.centres, np.float64), C.c_double)
        g.distance = _ptr(arr(np.where(np.isfinite(atlas.distance), atlas.distance, 1e300), np.float64), C.c_double)
        g.interior = _ptr(arr(atlas.interior, np.uint8), C.c_uint8)
        g.neighbours = _ptr(arr(atlas.neighbours, np.int64), C.c_int64)
        ch = atlas.chains
        g.n_ghost = ch.dest.size
        g.ghost_dest = _ptr(arr(ch.dest, np.int64), C.c_int64)
        g.ghost_root = _ptr(arr(ch.root, np.int64), C.c_int64)
        g.ghost_nops = _ptr(arr(ch.n_ops, np.int64), C.c_int64)
        g.ghost_ops = _ptr(arr(ch.ops, np.int64), C.c_int64)
        g.n_tasks = ch.kind.size
        g.task_kind = _ptr(arr(ch.kind, np.int32), C.c_int32)
        g.task_normal = _ptr(arr(ch.normal, np.float64), C.c_double)
        Context._lines_walls(g, atlas, arr)
        return g

    @staticmethod
    def _lines_walls(g, atlas, arr) -> None:
        L = atlas.lines
        g.n_lines, g.n_slots = L.n_lines, L.cells.size
        g.line_start = _ptr(arr(L.start, np.int64), C.c_int64)
        g.line_cells = _ptr(arr(L.cells, np.int64), C.c_int64)
        g.line_up_face = _ptr(arr(L.up_face, np.int64), C.c_int64)
        g.line_up_sign = _ptr(arr(L.up_sign, np.int64), C.c_int64)
        g.line_down_face = _ptr(arr(L.down_face, np.int64), C.c_int64)
        g.line_down_sign = _ptr(arr(L.down_sign, np.int64), C.c_int64)
        W = atlas.walls
        g.n_walls = W.n
        g.wall_cell = _ptr(arr(W.cell, np.int64), C.c_int64)
        g.wall_family = _ptr(arr(W.family, np.int64), C.c_int64)
        g.wall_face = _ptr(arr(W.face, np.int64), C.c_int64)
        g.wall_into = _ptr(arr(W.into_fluid, np.float64), C.c_double)
        g.wall_mid = _ptr(arr(W.mid.T, np.float64), C.c_double)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.jr_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- set-up ---------------------------------------------------------------
    def set_freestream(self, q_inf: np.ndarray, mu_inf=None) -> None:
        """Freestream primitives per case: (rho, u, v, p[, nu ratio]), 3D (rho, u, v, w, p[, nu ratio])."""
        q = np.zeros((self.n_cases, 6 if self.dim == 3 else 5))
        qi = np.atleast_2d(np.asarray(q_inf, np.float64))
        q[:, :qi.shape[1]] = qi
        mu = None if mu_inf is None else np.ascontiguousarray(np.broadcast_to(mu_inf, (self.n_cases,)), np.float64)
        _check(self._lib.jr_set_freestream(self._h, _ptr(q, C.c_double),
                                           None if mu is None else _ptr(mu, C.c_double)))
        self.freestream = q

    def set_geometry(self, k: int, atlas) -> None:
        """Case ``k`` steps on ``atlas``: another grid of this context's
        topology (same index plans, other coordinates)."""
        keep = []

        def arr(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a

        g = self._desc2(atlas, arr)
        _check(self._lib.jr_set_geometry(self._h, int(k), C.byref(g)))

    def set_state_case(self, k: int, U: np.ndarray) -> None:
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape[1:]), np.float64)
        _check(self._lib.jr_set_state_case(self._h, int(k), _ptr(U, C.c_double)))

    def newton_init_case(self, k: int, cfl0: float) -> None:
        _check(self._lib.jr_newton_init_case(self._h, int(k), C.c_double(cfl0)))

    def set_state(self, U: np.ndarray) -> None:
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape), np.float64)
        _check(self._lib.jr_set_state(self._h, _ptr(U, C.c_double)))

    def get_state(self) -> np.ndarray:
        out = np.empty(self.shape)
        _check(self._lib.jr_get_state(self._h, _ptr(out, C.c_double)))
        return out

    # --- components -------------------------------------------------------------
    def residual(self, U: np.ndarray, first_order: bool = False) -> np.ndarray:
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape), np.float64)
        out = np.empty(self.shape)
        _check(self._lib.jr_residual(self._h, _ptr(U, C.c_double), _ptr(out, C.c_double), int(first_order)))
        return out

    def tangent(self, U, dU, first_order: bool = True, with_vdt: bool = False) -> np.ndarray:
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape), np.float64)
        dU = np.ascontiguousarray(np.broadcast_to(dU, self.shape), np.float64)
        out = np.empty(self.shape)
        _check(self._lib.jr_tangent(self._h, _ptr(U, C.c_double), _ptr(dU, C.c_double), _ptr(out, C.c_double),
                                    int(first_order), int(with_vdt)))
        return out

    def _freestream_tangent(self, dq_inf, dmu_inf):
        dq = np.zeros((self.n_cases, 6 if self.dim == 3 else 5))
        d = np.atleast_2d(np.asarray(dq_inf, np.float64))
        dq[:, :d.shape[1]] = d
        dmu = None if dmu_inf is None else np.ascontiguousarray(np.broadcast_to(dmu_inf, (self.n_cases,)), np.float64)
        return dq, dmu

    def tangent_freestream(self, U, dq_inf, dmu_inf=None, dU=None) -> np.ndarray:
        """The exact residual's tangent at ``U`` along the freestream's
        tangent ``dq_inf`` (per case, the layout of ``set_freestream``),
        ``dmu_inf`` and the state's ``dU`` (None: zero)."""
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape), np.float64)
        dU = None if dU is None else np.ascontiguousarray(np.broadcast_to(dU, self.shape), np.float64)
        dq, dmu = self._freestream_tangent(dq_inf, dmu_inf)
        out = np.empty(self.shape)
        _check(self._lib.jr_tangent_freestream(self._h, _ptr(U, C.c_double),
                                               None if dU is None else _ptr(dU, C.c_double), _ptr(dq, C.c_double),
                                               None if dmu is None else _ptr(dmu, C.c_double), _ptr(out, C.c_double)))
        return out

    def forces_tangent(self, U, dq_inf, dmu_inf=None, dU=None, length: float = 1.0, centre=(0.25, 0.0)) -> np.ndarray:
        """2D: the tangent of ``forces`` at ``U`` along ``(dU, dq_inf,
        dmu_inf)``: ``(n_cases, 8)`` = d(cl, cd, cm, cd_friction, fx, fy), 0, 0."""
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape), np.float64)
        dU = None if dU is None else np.ascontiguousarray(np.broadcast_to(dU, self.shape), np.float64)
        dq, dmu = self._freestream_tangent(dq_inf, dmu_inf)
        out = np.empty((self.n_cases, 8))
        _check(self._lib.jr_forces_tangent(self._h, _ptr(U, C.c_double), None if dU is None else _ptr(dU, C.c_double),
                                           _ptr(dq, C.c_double), None if dmu is None else _ptr(dmu, C.c_double),
                                           C.c_double(length), C.c_double(centre[0]), C.c_double(centre[1]),
                                           _ptr(out, C.c_double)))
        return out

    def forces3_tangent(self, U, dq_inf, dmu_inf=None, dU=None, area: float = 1.0, length: float = 1.0,
                        centre=(0.25, 0.0, 0.0)) -> np.ndarray:
        """3D: the tangent of ``forces3`` at ``U`` along ``(dU, dq_inf,
        dmu_inf)``: ``(n_cases, 8)`` = d(cl, cd, cm, cd_friction, fx, fy, fz, cy)."""
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape), np.float64)
        dU = None if dU is None else np.ascontiguousarray(np.broadcast_to(dU, self.shape), np.float64)
        dq, dmu = self._freestream_tangent(dq_inf, dmu_inf)
        out = np.empty((self.n_cases, 8))
        _check(self._lib.jr_forces3_tangent(self._h, _ptr(U, C.c_double), None if dU is None else _ptr(dU, C.c_double),
                                            _ptr(dq, C.c_double), None if dmu is None else _ptr(dmu, C.c_double),
                                            C.c_double(area), C.c_double(length), C.c_double(centre[0]),
                                            C.c_double(centre[2]), _ptr(out, C.c_double)))
        return out

    def wall_loads_tangent(self, U, dq_inf, dmu_inf=None, dU=None) -> np.ndarray:
        """3D: the tangent of ``wall_loads`` at ``U`` along ``(dU, dq_inf,
        dmu_inf)``: ``(n_cases, 4, n_walls)``."""
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape), np.float64)
        dU = None if dU is None else np.ascontiguousarray(np.broadcast_to(dU, self.shape), np.float64)
        dq, dmu = self._freestream_tangent(dq_inf, dmu_inf)
        out = np.empty((self.n_cases, 4, max(1, self.atlas.walls.n)))
        _check(self._lib.jr_wall_loads_tangent(self._h, _ptr(U, C.c_double),
                                               None if dU is None else _ptr(dU, C.c_double), _ptr(dq, C.c_double),
                                               None if dmu is None else _ptr(dmu, C.c_double),
                                               _ptr(out, C.c_double)))
        return out

    def wave_speeds(self, U):
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape), np.float64)
        if self.dim == 3:
            nf = sum(int(np.prod(f.shape[1:])) for f in self.atlas.faces)
        else:
            NI, NJ = self.atlas.shape
            nf = (NI + 1) * NJ + NI * (NJ + 1)
        lam = np.empty((self.n_cases, nf))
        lsum = np.empty((self.n_cases,) + tuple(self.atlas.shape))
        _check(self._lib.jr_wave_speeds(self._h, _ptr(U, C.c_double), _ptr(lam, C.c_double), _ptr(lsum, C.c_double)))
        return lam, lsum

    def precond_factor(self, U) -> None:
        U = np.ascontiguousarray(np.broadcast_to(U, self.shape), np.float64)
        _check(self._lib.jr_precond_factor(self._h, _ptr(U, C.c_double)))

    def precond_apply(self, r) -> np.ndarray:
        r = np.ascontiguousarray(np.broadcast_to(r, self.shape), np.float64)
        out = np.empty(self.shape)
        _check(self._lib.jr_precond_apply(self._h, _ptr(r, C.c_double), _ptr(out, C.c_double)))
        return out

    # --- Newton -------------------------------------------------------------------
    def newton_init(self, cfl0) -> None:
        c = np.ascontiguousarray(np.broadcast_to(cfl0, (self.n_cases,)), np.float64)
        _check(self._lib.jr_newton_init(self._h, _ptr(c, C.c_double)))

    def set_cfl(self, cfl) -> None:
        c = np.ascontiguousarray(np.broadcast_to(cfl, (self.n_cases,)), np.float64)
        _check(self._lib.jr_set_cfl(self._h, _ptr(c, C.c_double)))

    def linear_solve(self, b: np.ndarray) -> np.ndarray:
        """x with (V/dt + dR~/dU) x = b at the device state and CFL: the
        Newton step's GMRES solve, the state left as it is."""
        b = np.ascontiguousarray(np.broadcast_to(b, self.shape), np.float64)
        out = np.empty(self.shape)
        _check(self._lib.jr_linear_solve(self._h, _ptr(b, C.c_double), _ptr(out, C.c_double)))
        return out

    def rk_block(self, n_steps: int, alphas) -> np.ndarray:
        """``n_steps`` explicit multistage steps (stage coefficients
        ``alphas``, the CFL of ``newton_init`` / ``set_cfl``); returns the
        residual history ``(n_steps, n_cases)``."""
        a = np.ascontiguousarray(alphas, np.float64)
        hist = np.empty((n_steps, self.n_cases))
        _check(self._lib.jr_rk_block(self._h, int(n_steps), _ptr(a, C.c_double), int(a.size),
                                     _ptr(hist, C.c_double)))
        return hist

    def newton_block(self, n_steps: int) -> np.ndarray:
        hist = np.empty((n_steps, self.n_cases))
        _check(self._lib.jr_newton_block(self._h, int(n_steps), _ptr(hist, C.c_double)))
        return hist

    def newton_block_noalloc(self, n_steps: int) -> np.ndarray:
        """``newton_block`` through the entry point that allocates nothing:
        the history is staged in the library's static per-device array
        (``n_steps * n_cases`` at most ``NOALLOC_HISTORY_MAX``); the same
        results as ``newton_block``.  Calls that overlap on one device race."""
        hist = np.empty((n_steps, self.n_cases))
        _check(self._lib.jr_newton_block_noalloc(self._h, int(n_steps), _ptr(hist, C.c_double)))
        return hist

    def newton_block_async(self, n_steps: int, forces_ref=None) -> None:
        """Enqueue ``n_steps`` Newton steps and return at once (see ``wait``).
        ``forces_ref = (length, (x, y))`` also enqueues the force coefficients
        of the block's final state."""
        ref = None
        if forces_ref is not None:
            length, (x, y) = forces_ref
            ref = np.array([length, x, y], np.float64)
        _check(self._lib.jr_newton_block_async(self._h, int(n_steps), None if ref is None else _ptr(ref, C.c_double)))
        self._pending = int(n_steps)
        self._pending_forces = ref is not None

    def poll(self) -> bool:
        """Whether the pending asynchronous block has finished."""
        rc = self._lib.jr_newton_block_poll(self._h)
        if rc < 0:
            _check(rc)
        return rc == 1

    def wait(self):
        """Finish the pending asynchronous block: its residual history
        ``(n_steps, n_cases)``, and with ``forces_ref`` given to
        ``newton_block_async`` the coefficients ``(n_cases, 8)`` as well
        (a tuple then)."""
        hist = np.empty((getattr(self, "_pending", 0), self.n_cases))
        forces = np.empty((self.n_cases, 8)) if getattr(self, "_pending_forces", False) else None
        _check(self._lib.jr_newton_block_wait(self._h, _ptr(hist, C.c_double),
                                              None if forces is None else _ptr(forces, C.c_double)))
        self._pending, self._pending_forces = 0, False
        return hist if forces is None else (hist, forces)

    @property
    def device_bytes(self) -> int:
        out = C.c_size_t()
        _check(self._lib.jr_ctx_bytes(self._h, C.byref(out)))
        return int(out.value)

    def park_case(self, k: int, parked: bool = True) -> None:
        """Park (or release) case ``k``: a parked case's Newton steps skip its
        GMRES and leave its state as it is (an empty serving slot)."""
        _check(self._lib.jr_park_case(self._h, int(k), int(bool(parked))))

    def set_tile_tangent(self, enable: bool) -> None:
        """2D: the Krylov tangent by LDS tiles (default) or by the separate
        node, face and cell kernels (they agree to round-off)."""
        _check(self._lib.jr_set_tile_tangent(self._h, int(bool(enable))))

    def set_graphs(self, enable: bool) -> None:
        """Replay Newton blocks from a captured HIP graph, or launch every
        kernel on the context stream (the default)."""
        _check(self._lib.jr_set_graphs(self._h, int(bool(enable))))

    PROBE = ("precond_apply", "tangent_faces", "tangent_fill", "tangent_cells", "cgs2_dots", "cgs2_update_dots",
             "cgs2_update_norm", "precond_factor")

    def probe_block(self, n_steps: int, kernels=PROBE, max_launches: int = 16384) -> dict:
        """``n_steps`` Newton steps with every launch of ``kernels`` bracketed by
        HIP events on the library's stream: per kernel the launch count and the
        mean device milliseconds per launch."""
        mask = sum(1 << self.PROBE.index(k) for k in kernels)
        _check(self._lib.jr_probe_arm(self._h, C.c_uint(mask), int(max_launches)))
        self.newton_block(n_steps)
        ms = np.zeros(8)
        n = np.zeros(8, np.int32)
        _check(self._lib.jr_probe_read(self._h, _ptr(ms, C.c_double), _ptr(n, C.c_int)))
        return {k: {"launches": int(n[i]), "ms": float(ms[i] / n[i]) if n[i] else float("nan")}
                for i, k in enumerate(self.PROBE) if k in kernels}

    def scalars(self) -> np.ndarray:
        out = np.empty((self.n_cases, 9))
        _check(self._lib.jr_get_scalars(self._h, _ptr(out, C.c_double)))
        return out

    def forces3(self, area: float, length: float = 1.0, centre=(0.25, 0.0, 0.0)) -> np.ndarray:
        """3D: (cl, cd, cm, cd_friction, fx, fy, fz, cy) per case; moment about y through ``centre``."""
        out = np.empty((self.n_cases, 8))
        _check(self._lib.jr_forces3(self._h, C.c_double(area), C.c_double(length), C.c_double(centre[0]),
                                    C.c_double(centre[2]), _ptr(out, C.c_double)))
        return out

    def wall_loads(self) -> np.ndarray:
        """Per wall face of the device state (the atlas's wall order):
        ``(n_cases, 1 + dim, n_walls)`` = gauge pressure, then the viscous
        force on the body per component (zero when inviscid)."""
        out = np.empty((self.n_cases, self.dim + 1, max(1, self.atlas.walls.n)))
        _check(self._lib.jr_wall_loads(self._h, _ptr(out, C.c_double)))
        return out[:, :, :self.atlas.walls.n]

    def forces(self, length: float = 1.0, centre=(0.25, 0.0)) -> np.ndarray:
        out = np.empty((self.n_cases, 8))
        _check(self._lib.jr_forces(self._h, C.c_double(length), C.c_double(centre[0]), C.c_double(centre[1]),
                                   _ptr(out, C.c_double)))
        return out

    # --- timing ---------------------------------------------------------------------
    def timer_start(self) -> None:
        _check(self._lib.jr_timer_start(self._h))

    def timer_stop(self) -> float:
        ms = C.c_float()
        _check(self._lib.jr_timer_stop(self._h, C.byref(ms)))
        return float(ms.value)

    PROFILE = {"residual": 0, "tangent": 1, "precond_apply": 2, "residual_faces": 3, "tangent_faces": 4,
               "cgs2_update_dots": 5, "precond_factor": 6, "tangent_cells": 7}

    def profile(self, kernel: str, reps: int = 20) -> float:
        """Average milliseconds of one launch of ``kernel`` (HIP events, device stream)."""
        ms = C.c_float()
        _check(self._lib.jr_profile(self._h, self.PROFILE[kernel], int(reps), C.byref(ms)))
        return float(ms.value)

    def synchronize(self) -> None:
        _check(self._lib.jr_synchronize(self._h))


def context_for(atlas: Atlas, flow, numerics, implicit=None, n_cases: int = 1) -> Context:
    """A context for a case's flow and numerics (``FlowConditions``, ``NumericsConfig``)."""
    return Context(atlas, gamma=flow.gamma, flux=numerics.flux, limiter=numerics.limiter, kappa=numerics.kappa,
                   viscous=not flow.inviscid, sa=flow.turbulence is not None, sutherland=flow.sutherland_ratio,
                   c_cr1=0.3 if flow.turbulence == "sa_noft2_qcr" else 0.0, implicit=implicit, n_cases=n_cases,
                   far_field=numerics.far_field)
