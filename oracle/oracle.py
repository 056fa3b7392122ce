"""ctypes binding of the CPU oracle (liblsoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py. The product path never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblsoracle.so")


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("bandwidth", C.c_double), ("tol", C.c_double),
                ("lambda_eps", C.c_double), ("niter", C.c_int), ("inside_niter", C.c_int),
                ("mode", C.c_int), ("nthreads", C.c_int), ("max_fim_iters", C.c_int), ("warm", C.c_int),
                ("warm_margin", C.c_double), ("warm_res_tol", C.c_double), ("warm_abs_margin", C.c_double),
                ("cfl", C.c_int), ("pad", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("fim_iters", C.c_int), ("nb_count", C.c_int), ("patch_visits", C.c_int64),
                ("probes", C.c_int64), ("nb_verts", C.c_int64), ("n_seed_patches", C.c_int),
                ("warm_patches", C.c_int), ("warm_residual", C.c_double), ("warm_fallback", C.c_int),
                ("pad", C.c_int), ("dt_used", C.c_double)]


def build() -> str:
    subprocess.check_call(["make", "-s", "-C", _HERE, "liblsoracle.so"])
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.lso_create.restype = vp
        L.lso_create.argtypes = [C.c_int, C.c_int, C.c_int64, dp, C.c_int64, ip, ip, C.c_int]
        L.lso_destroy.argtypes = [vp]
        L.lso_set_phi.argtypes = [vp, dp]
        L.lso_set_velocity.argtypes = [vp, dp, C.c_int]
        L.lso_step.argtypes = [vp, C.POINTER(Params), C.POINTER(Stats)]
        L.lso_get_phi.argtypes = [vp, dp]
        L.lso_get_narrowband.argtypes = [vp, C.c_void_p]
        L.lso_num_copies.restype = C.c_int64
        L.lso_num_copies.argtypes = [vp]
        L.lso_get_layout.argtypes = [vp] + [C.c_void_p] * 5
        L.lso_tet_solve.restype = C.c_double
        L.lso_tet_solve.argtypes = [C.c_int, dp, dp, C.c_double]
        L.lso_tri_solve.restype = C.c_double
        L.lso_tri_solve.argtypes = [C.c_int] + [C.c_double] * 5
        L.lso_tet_element_update.argtypes = [C.c_int, dp, dp, C.c_double]
        L.lso_max_threads.restype = C.c_int
        L.lso_set_narrowband.argtypes = [vp, C.c_void_p, C.c_int]
        L.lso_get_ambiguous.argtypes = [vp, C.c_void_p]
        L.lso_get_coords.argtypes = [vp, C.c_void_p]
        L.lso_advect_all.argtypes = [vp, C.c_double]
        L.lso_set_warm.argtypes = [vp, C.c_void_p]
        L.lso_get_warm_residuals.argtypes = [vp, C.c_void_p]
        L.lso_set_cfl_coefficients.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.lso_compute_dt.argtypes = [vp, C.c_void_p]
        L.lso_get_labels.argtypes = [vp, C.c_void_p]
        L.lso_inradius.restype = C.c_double
        L.lso_inradius.argtypes = [C.c_int, dp]
        _lib = L
    return _lib


def make_params(dt, bandwidth, niter, tol=1e-5, lambda_eps=1e-5, inside_niter=1, mode=0, nthreads=0,
                max_fim_iters=0, warm=0, warm_margin=0.0, warm_res_tol=0.0, warm_abs_margin=0.0, cfl=0):
    return Params(dt, bandwidth, tol, lambda_eps, niter, inside_niter, mode, nthreads, max_fim_iters, warm,
                  warm_margin, warm_res_tol, warm_abs_margin, cfl, 0)


class Oracle:
    """One mesh + state, FP64 (precision=64) or the FP32 mirror (precision=32)."""

    def __init__(self, dim, xyz, cells, labels, nparts, precision=64):
        self.L = lib()
        self.dim = dim
        self.N = xyz.shape[0]
        self.E = cells.shape[0]
        self.nparts = int(nparts)
        self._xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self._cells = np.ascontiguousarray(cells, dtype=np.int32)
        self._labels = np.ascontiguousarray(labels, dtype=np.int32)
        self.h = self.L.lso_create(dim, precision, self.N, self._xyz.ravel(), self.E,
                                   self._cells.ravel(), self._labels, self.nparts)
        if not self.h:
            raise RuntimeError("lso_create failed")

    def close(self):
        if self.h:
            self.L.lso_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_phi(self, phi):
        self.L.lso_set_phi(self.h, np.ascontiguousarray(phi, dtype=np.float64))

    def set_velocity(self, vel, mode=0):
        """mode 0: each element's velocity; 1: the patch's first copy's; 2
        (triangles): the reference's 2D advection velocity from phi."""
        self.L.lso_set_velocity(self.h, np.ascontiguousarray(vel, dtype=np.float64).ravel(), mode)

    def step(self, params: Params) -> Stats:
        st = Stats()
        rc = self.L.lso_step(self.h, C.byref(params), C.byref(st))
        if rc != 0:
            raise RuntimeError(f"lso_step rc={rc}")
        return st

    def phi(self):
        out = np.empty(self.N, dtype=np.float64)
        self.L.lso_get_phi(self.h, out)
        return out

    def set_narrowband(self, ids):
        a = np.ascontiguousarray(ids, dtype=np.int32)
        self.L.lso_set_narrowband(self.h, a.ctypes.data, int(a.size))

    def ambiguous(self):
        out = np.zeros(self.N, dtype=np.uint8)
        self.L.lso_get_ambiguous(self.h, out.ctypes.data)
        return out.astype(bool)

    def set_warm(self, warm):
        """Warm-start distances (caller order) for the next step only."""
        self._warm = np.ascontiguousarray(warm, dtype=np.float64)
        self.L.lso_set_warm(self.h, self._warm.ctypes.data)

    def set_cfl_coefficients(self, ceik=None, ccurv=None):
        self._ceik = None if ceik is None else np.ascontiguousarray(ceik, dtype=np.float64)
        self._ccurv = None if ccurv is None else np.ascontiguousarray(ccurv, dtype=np.float64)
        self.L.lso_set_cfl_coefficients(self.h, None if self._ceik is None else self._ceik.ctypes.data,
                                        None if self._ccurv is None else self._ccurv.ctypes.data)

    def labels(self):
        """Vertex labels after the last step (0: seed ... 3: far), caller order."""
        out = np.empty(self.N, dtype=np.int32)
        self.L.lso_get_labels(self.h, out.ctypes.data)
        return out

    def compute_dt(self):
        """(CFL bound, largest inscribed radius) over the current narrowband."""
        out = np.zeros(2)
        self.L.lso_compute_dt(self.h, out.ctypes.data)
        return float(out[0]), float(out[1])

    def warm_residuals(self):
        """Per-vertex acceptance residuals F_v - DT_v of the last warm step."""
        out = np.zeros(self.N, dtype=np.float32)
        self.L.lso_get_warm_residuals(self.h, out.ctypes.data)
        return out

    def advect_all(self, dt):
        self.L.lso_advect_all(self.h, float(dt))

    def coords(self):
        F = self.L.lso_num_copies(self.h)
        nk = 6 if self.dim == 3 else 3
        out = np.empty(F * nk, np.float64)
        self.L.lso_get_coords(self.h, out.ctypes.data)
        return out.reshape(F, nk)

    def narrowband(self):
        n = self.L.lso_get_narrowband(self.h, None)
        ids = np.empty(max(n, 1), dtype=np.int32)
        self.L.lso_get_narrowband(self.h, ids.ctypes.data)
        return ids[:n].copy()

    def layout(self):
        F = self.L.lso_num_copies(self.h)
        nc = self.dim + 1
        vp = np.empty(self.N, np.int32)
        ep = np.empty(F, np.int32)
        vo = np.empty(self.nparts + 1, np.int32)
        eo = np.empty(self.nparts + 1, np.int32)
        el = np.empty(nc * F, np.int32)
        self.L.lso_get_layout(self.h, vp.ctypes.data, ep.ctypes.data, vo.ctypes.data,
                              eo.ctypes.data, el.ctypes.data)
        return {"vert_perm": vp, "copy_src": ep, "vert_off": vo, "ele_off": eo,
                "copy_verts": el.reshape(nc, F).T.copy()}


def tet_solve(T, dots, precision=64, lambda_eps=1e-5):
    return lib().lso_tet_solve(precision, np.asarray(T, np.float64), np.asarray(dots, np.float64), lambda_eps)


def tri_solve(TA, TB, AB, BC, AC, precision=64):
    return lib().lso_tri_solve(precision, TA, TB, AB, BC, AC)


def tet_element_update(L6, vals, precision=64, lambda_eps=1e-5):
    v = np.array(vals, dtype=np.float64)
    lib().lso_tet_element_update(precision, np.asarray(L6, np.float64), v, lambda_eps)
    return v


def inradius(dim, L):
    """Inscribed radius of an element from its local coordinates."""
    return lib().lso_inradius(dim, np.ascontiguousarray(L, np.float64))


def max_threads():
    return lib().lso_max_threads()
