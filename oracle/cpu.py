"""ctypes binding of the oracle's C++/OpenMP restatement (``oracle/cpu/newton2d.cpp``).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the CPU baseline of
the benchmark and its parity tests against the NumPy oracle load it.  Build:
``make -C oracle/cpu`` (``__graft_entry__.build()`` does it).
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "cpu" / "libnewton2d.so"
FLUX = {"hllc": 0, "roe": 1}
LIMITER = {"minmod": 0, "none": 1, "van_albada": 2, "van_albada_smooth": 3}

_d, _i64, _i32, _u8 = C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)


class Grid(C.Structure):
    _fields_ = [("NI", C.c_int32), ("NJ", C.c_int32), ("area", _d), ("face_i", _d), ("face_j", _d),
                ("centres", _d), ("distance", _d), ("interior", _u8), ("neighbours", _i64),
                ("n_task", C.c_int32 * 2), ("task_dest", _i64 * 2), ("task_src", _i64 * 2), ("task_kind", _i32 * 2),
                ("task_normal", _d * 2), ("n_lines", C.c_int32), ("n_slots", C.c_int32), ("line_start", _i64),
                ("line_cells", _i64), ("line_up_face", _i64), ("line_up_sign", _i64), ("line_down_face", _i64),
                ("line_down_sign", _i64)]


class Phys(C.Structure):
    _fields_ = [("gamma", C.c_double), ("kappa", C.c_double), ("flux", C.c_int32), ("limiter", C.c_int32),
                ("viscous", C.c_int32), ("sa", C.c_int32), ("sutherland", C.c_double), ("c_cr1", C.c_double),
                ("mu_inf", C.c_double), ("q_inf", C.c_double * 5), ("krylov_subspace", C.c_int32),
                ("krylov_restarts", C.c_int32), ("krylov_tol", C.c_double), ("krylov_success", C.c_double),
                ("cfl_growth", C.c_double), ("cfl_cutback", C.c_double), ("cfl_min", C.c_double),
                ("cfl_max", C.c_double), ("max_relative_update", C.c_double), ("exact", C.c_int32)]


_LIB = None


def library():
    global _LIB
    if _LIB is None:
        if not LIB.is_file():
            raise FileNotFoundError(f"{LIB} is not built (make -C oracle/cpu)")
        lib = C.CDLL(str(LIB))
        lib.oc_create.restype = C.c_void_p
        lib.oc_create.argtypes = [C.POINTER(Grid), C.POINTER(Phys)]
        for name in ("oc_destroy", "oc_residual", "oc_jvp", "oc_wave_speeds", "oc_precond_setup",
                     "oc_precond_apply", "oc_newton_step"):
            getattr(lib, name).restype = None
        lib.oc_destroy.argtypes = [C.c_void_p]
        lib.oc_residual.argtypes = [C.c_void_p, _d, _d, C.c_int]
        lib.oc_jvp.argtypes = [C.c_void_p, _d, _d, _d, C.c_int]
        lib.oc_wave_speeds.argtypes = [C.c_void_p, _d, _d]
        lib.oc_precond_setup.argtypes = [C.c_void_p, _d, C.c_double]
        lib.oc_precond_apply.argtypes = [C.c_void_p, _d, _d]
        lib.oc_newton_step.argtypes = [C.c_void_p, _d, _d, _d]
        lib.oc_threads.restype = C.c_int
        _LIB = lib
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Case:
    """One case of the oracle's setup (``oracle.residual.Setup``) and
    implicit settings on the C++ restatement."""

    def __init__(self, setup, implicit=None):
        lib = library()
        at = setup.atlas
        keep = []

        def arr(a, dtype):
            a = np.ascontiguousarray(a, dtype)
            keep.append(a)
            return a

        g = Grid()
        g.NI, g.NJ = at.shape
        g.area = _p(arr(at.area, np.float64), C.c_double)
        g.face_i = _p(arr(at.face_i.reshape(2, -1), np.float64), C.c_double)
        g.face_j = _p(arr(at.face_j.reshape(2, -1), np.float64), C.c_double)
        g.centres = _p(arr(at.centres, np.float64), C.c_double)
        g.distance = _p(arr(np.where(np.isfinite(at.distance), at.distance, 1e300), np.float64), C.c_double)
        g.interior = _p(arr(at.interior, np.uint8), C.c_uint8)
        g.neighbours = _p(arr(at.neighbours, np.int64), C.c_int64)
        for s, st in enumerate(at.stages):
            g.n_task[s] = st.dest.size
            g.task_dest[s] = _p(arr(st.dest, np.int64), C.c_int64)
            g.task_src[s] = _p(arr(st.src, np.int64), C.c_int64)
            g.task_kind[s] = _p(arr(st.kind, np.int32), C.c_int32)
            g.task_normal[s] = _p(arr(st.normal, np.float64), C.c_double)
        L = at.lines
        g.n_lines, g.n_slots = L.n_lines, L.cells.size
        for name, a in (("line_start", L.start), ("line_cells", L.cells), ("line_up_face", L.up_face),
                        ("line_up_sign", L.up_sign), ("line_down_face", L.down_face),
                        ("line_down_sign", L.down_sign)):
            setattr(g, name, _p(arr(a, np.int64), C.c_int64))
        ph = Phys()
        ph.gamma, ph.kappa = setup.gamma, setup.kappa
        ph.flux, ph.limiter = FLUX[setup.flux], LIMITER[setup.limiter]
        ph.viscous, ph.sa = int(setup.viscous), int(setup.sa)
        ph.sutherland = -1.0 if setup.sutherland is None else float(setup.sutherland)
        ph.c_cr1 = setup.c_cr1
        ph.mu_inf = setup.mu_inf if setup.viscous else 0.0
        q = np.zeros(5)
        q[:len(setup.q_inf)] = setup.q_inf
        for k in range(5):
            ph.q_inf[k] = q[k]
        if implicit is not None:
            ph.krylov_subspace, ph.krylov_restarts = implicit.krylov_subspace, implicit.krylov_restarts
            ph.krylov_tol, ph.krylov_success = implicit.krylov_tol, implicit.krylov_success
            ph.cfl_growth, ph.cfl_cutback = implicit.cfl_growth, implicit.cfl_cutback
            ph.cfl_min, ph.cfl_max = implicit.cfl_min, implicit.cfl_max
            ph.max_relative_update = -1.0 if implicit.max_relative_update is None else implicit.max_relative_update
            ph.exact = int(implicit.jacobian == "exact")
        self._keep = keep + [g, ph]
        self._lib = lib
        self._h = lib.oc_create(C.byref(g), C.byref(ph))
        self.shape = (setup.nvar,) + tuple(at.shape)
        self.nf = (at.shape[0] + 1) * at.shape[1] + at.shape[0] * (at.shape[1] + 1)

    def close(self):
        if self._h:
            self._lib.oc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _in(self, a):
        return np.ascontiguousarray(a, np.float64).reshape(self.shape)

    def residual(self, U, first_order=False):
        U = self._in(U)
        out = np.empty(self.shape)
        self._lib.oc_residual(self._h, _p(U, C.c_double), _p(out, C.c_double), int(first_order))
        return out

    def jvp(self, U, dU, first_order=True):
        U, dU = self._in(U), self._in(dU)
        out = np.empty(self.shape)
        self._lib.oc_jvp(self._h, _p(U, C.c_double), _p(dU, C.c_double), _p(out, C.c_double), int(first_order))
        return out

    def wave_speeds(self, U):
        U = self._in(U)
        out = np.empty(self.nf)
        self._lib.oc_wave_speeds(self._h, _p(U, C.c_double), _p(out, C.c_double))
        return out

    def precond_setup(self, U, cfl):
        U = self._in(U)
        self._lib.oc_precond_setup(self._h, _p(U, C.c_double), float(cfl))

    def precond_apply(self, r):
        r = self._in(r)
        out = np.empty(self.shape)
        self._lib.oc_precond_apply(self._h, _p(r, C.c_double), _p(out, C.c_double))
        return out

    def newton_step(self, U, state):
        """One Newton step on ``U`` (updated in place); ``state`` =
        ``[cfl, |R|_prev]`` (updated); returns ``(|R(U)|, accepted)``."""
        out = np.zeros(4)
        self._lib.oc_newton_step(self._h, _p(U, C.c_double), _p(state, C.c_double), _p(out, C.c_double))
        return float(out[0]), bool(out[1])


def threads() -> int:
    return int(library().oc_threads())
