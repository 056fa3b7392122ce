"""ctypes binding of liblevelset_hip.so (include/levelset_hip.h).

Loading fails loudly when the HIP library is missing: there is no CPU
fallback on the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblevelset_hip.so")

LS_F32, LS_F64 = 0, 1
LS_MODE_ADVECT, LS_MODE_MCF = 0, 1
LS_EXEC_PERSISTENT, LS_EXEC_KERNELS, LS_EXEC_GRAPH = 0, 1, 2
LS_WARM_NONE, LS_WARM_IDENTITY, LS_WARM_LEARNED = 0, 1, 2
LS_VEL_ELEMENT, LS_VEL_PATCH_FIRST, LS_VEL_PHI_2D = 0, 1, 2
LS_SWEEP_AUTO, LS_SWEEP_ELEMENTS, LS_SWEEP_TASKS, LS_SWEEP_UNSTAGED = 0, 1, 2, 4
LS_SWEEP_NO_PLANE_BOUND = 8
LS_LAYOUT_DEVICE, LS_LAYOUT_HOST = 0, 1


class StepParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("bandwidth", C.c_double), ("tol", C.c_double),
                ("lambda_eps", C.c_double), ("niter", C.c_int32), ("inside_niter", C.c_int32),
                ("mode", C.c_int32), ("max_fim_iters", C.c_int32), ("warm", C.c_int32),
                ("cfl", C.c_int32), ("warm_margin", C.c_double), ("warm_res_tol", C.c_double),
                ("warm_abs_margin", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("fim_iters", C.c_int32), ("nb_count", C.c_int32), ("patch_visits", C.c_int64),
                ("probes", C.c_int64), ("nb_verts", C.c_int64), ("n_seed_patches", C.c_int32),
                ("sweep_launches", C.c_int32), ("ms_seed", C.c_double), ("ms_fim", C.c_double),
                ("ms_nb", C.c_double), ("ms_advect", C.c_double), ("ms_total", C.c_double),
                ("ms_sweep_kernels", C.c_double), ("ms_advect_kernels", C.c_double),
                ("visit_copies", C.c_int64), ("visit_verts", C.c_int64), ("visit_halo", C.c_int64),
                ("probe_copies", C.c_int64), ("probe_verts", C.c_int64), ("probe_halo", C.c_int64),
                ("adv_copies", C.c_int64), ("adv_verts", C.c_int64), ("adv_halo", C.c_int64),
                ("seed_copies", C.c_int64), ("seed_verts", C.c_int64), ("seed_halo", C.c_int64),
                ("ms_step_kernel", C.c_double), ("exec_mode", C.c_int32), ("sweep_staged", C.c_int32),
                ("ms_ph_seed", C.c_double), ("ms_ph_visit", C.c_double), ("ms_ph_copyback", C.c_double),
                ("ms_ph_probe", C.c_double), ("ms_ph_next", C.c_double), ("ms_ph_nb", C.c_double),
                ("ms_ph_advect", C.c_double), ("elem_updates", C.c_int64), ("local_solves", C.c_int64),
                ("warm_residual", C.c_double), ("warm_patches", C.c_int32), ("warm_fallback", C.c_int32),
                ("sweep_tasks", C.c_int32), ("step_wg_per_cu", C.c_int32), ("dt_used", C.c_double),
                ("cfl_dt", C.c_double), ("cfl_max_rin", C.c_double), ("fast_len_roots", C.c_int32),
                ("plane_bound", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LayoutInfo(C.Structure):
    _fields_ = [("n_verts", C.c_int64), ("n_cells", C.c_int64), ("n_copies", C.c_int64),
                ("n_patches", C.c_int32), ("max_patch_copies", C.c_int32),
                ("max_patch_verts", C.c_int32), ("max_patch_halo", C.c_int32),
                ("max_vertex_slots", C.c_int32), ("dim", C.c_int32),
                ("n_halo", C.c_int64), ("n_slots", C.c_int64), ("n_adjacent", C.c_int64)]


EXPORTS = ["ls_create", "ls_destroy", "ls_last_error", "ls_set_mesh", "ls_set_precision", "ls_set_execution",
           "ls_set_sweep", "ls_partition_grid",
           "ls_partition_rcb", "ls_partition_labels", "ls_get_layout_info", "ls_get_layout", "ls_get_patch_lists",
           "ls_set_layout_builder",
           "ls_set_phi", "ls_set_velocity", "ls_step", "ls_get_phi", "ls_get_narrowband",
           "ls_set_mesh_batch", "ls_partition_batch", "ls_set_batch_params", "ls_get_batch_nb_counts",
           "ls_get_distances", "ls_mlp_set_params", "ls_mlp_get_params", "ls_mlp_predict", "ls_mlp_collect",
           "ls_mlp_dataset", "ls_mlp_gradient", "ls_mlp_train", "ls_set_warm",
           "ls_get_warm", "ls_compute_dt", "ls_set_cfl_coefficients",
           "ls_mlp_gradient_noalloc"]
MLP_PARAMS = 5892


def build(verbose=False) -> str:
    cmd = ["make", "-C", os.path.join(_HERE, "csrc")]
    if not verbose:
        cmd.insert(1, "-s")
    subprocess.check_call(cmd)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.ls_create.argtypes = [C.POINTER(vp), C.c_int]
        L.ls_destroy.argtypes = [vp]
        L.ls_destroy.restype = None
        L.ls_last_error.argtypes = [vp]
        L.ls_last_error.restype = C.c_char_p
        L.ls_set_mesh.argtypes = [vp, C.c_int, C.c_int64, vp, C.c_int64, vp]
        L.ls_set_precision.argtypes = [vp, C.c_int]
        L.ls_set_execution.argtypes = [vp, C.c_int]
        L.ls_set_sweep.argtypes = [vp, C.c_int]
        L.ls_set_layout_builder.argtypes = [vp, C.c_int]
        L.ls_get_patch_lists.argtypes = [vp] + [vp] * 7
        L.ls_partition_grid.argtypes = [vp] + [C.c_int] * 6
        L.ls_partition_rcb.argtypes = [vp, C.c_int]
        L.ls_partition_labels.argtypes = [vp, vp, C.c_int]
        L.ls_get_layout_info.argtypes = [vp, C.POINTER(LayoutInfo)]
        L.ls_get_layout.argtypes = [vp] + [vp] * 5
        L.ls_set_phi.argtypes = [vp, vp, C.c_int]
        L.ls_set_velocity.argtypes = [vp, vp, C.c_int]
        L.ls_step.argtypes = [vp, C.POINTER(StepParams), C.POINTER(Stats)]
        L.ls_get_phi.argtypes = [vp, vp, C.c_int]
        L.ls_get_narrowband.argtypes = [vp, vp, C.POINTER(C.c_int32)]
        L.ls_set_mesh_batch.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp]
        L.ls_partition_batch.argtypes = [vp, C.c_int]
        L.ls_set_batch_params.argtypes = [vp, vp, vp]
        L.ls_get_batch_nb_counts.argtypes = [vp, vp]
        L.ls_get_distances.argtypes = [vp, vp]
        L.ls_set_warm.argtypes = [vp, vp]
        L.ls_get_warm.argtypes = [vp, vp]
        L.ls_compute_dt.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.ls_set_cfl_coefficients.argtypes = [vp, vp, vp]
        L.ls_mlp_set_params.argtypes = [vp, vp, C.c_int]
        L.ls_mlp_get_params.argtypes = [vp, vp, C.c_int]
        L.ls_mlp_predict.argtypes = [vp]
        L.ls_mlp_collect.argtypes = [vp, C.c_int, C.c_int64]
        L.ls_mlp_dataset.argtypes = [vp, vp, vp, C.POINTER(C.c_int64)]
        L.ls_mlp_gradient.argtypes = [vp, vp, vp, C.c_int64, vp, C.POINTER(C.c_double)]
        L.ls_mlp_gradient_noalloc.argtypes = [vp, vp, vp, C.c_int64, vp, C.POINTER(C.c_double)]
        L.ls_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_uint64, C.POINTER(C.c_double)]
        _lib = L
    return _lib


class LevelSetError(RuntimeError):
    pass


def step_params(dt, bandwidth, niter, tol=1e-5, lambda_eps=1e-5, inside_niter=1, mode=LS_MODE_ADVECT,
                max_fim_iters=0, warm=LS_WARM_NONE, warm_margin=0.0, warm_res_tol=0.0, warm_abs_margin=0.0,
                cfl=0):
    return StepParams(dt, bandwidth, tol, lambda_eps, niter, inside_niter, mode, max_fim_iters, warm, cfl,
                      warm_margin, warm_res_tol, warm_abs_margin)


class Solver:
    """One mesh on one GPU through the C-ABI."""

    def __init__(self, device=0):
        self.L = lib()
        self.ctx = C.c_void_p()
        self._check(self.L.ls_create(C.byref(self.ctx), device))
        self.N = self.E = 0
        self.dim = 0

    def _check(self, rc):
        if rc != 0:
            msg = self.L.ls_last_error(self.ctx).decode() if self.ctx else ""
            raise LevelSetError(f"levelset rc={rc}: {msg}")

    def close(self):
        if self.ctx:
            self.L.ls_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mesh(self, dim, xyz, cells):
        self._xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self._cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.dim, self.N, self.E = dim, self._xyz.shape[0], self._cells.shape[0]
        self._check(self.L.ls_set_mesh(self.ctx, dim, self.N, self._xyz.ctypes.data, self.E,
                                       self._cells.ctypes.data))

    def set_execution(self, mode):
        self._check(self.L.ls_set_execution(self.ctx, int(mode)))

    def set_sweep(self, mode):
        self._check(self.L.ls_set_sweep(self.ctx, int(mode)))

    def set_layout_builder(self, builder):
        self._check(self.L.ls_set_layout_builder(self.ctx, int(builder)))

    def set_precision(self, dtype):
        self._check(self.L.ls_set_precision(self.ctx, int(dtype)))

    def partition_labels(self, labels, nparts):
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        self._check(self.L.ls_partition_labels(self.ctx, lab.ctypes.data, int(nparts)))

    def partition_rcb(self, nparts):
        self._check(self.L.ls_partition_rcb(self.ctx, int(nparts)))

    def partition_grid(self, L, W, H, bl, bw, bh):
        self._check(self.L.ls_partition_grid(self.ctx, L, W, H, bl, bw, bh))

    def layout_info(self):
        info = LayoutInfo()
        self._check(self.L.ls_get_layout_info(self.ctx, C.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def layout(self):
        info = self.layout_info()
        N, F, P, nc = info["n_verts"], info["n_copies"], info["n_patches"], info["dim"] + 1
        vp = np.empty(N, np.int32)
        cs = np.empty(F, np.int32)
        vo = np.empty(P + 1, np.int32)
        eo = np.empty(P + 1, np.int32)
        cv = np.empty(F * nc, np.int32)
        self._check(self.L.ls_get_layout(self.ctx, vp.ctypes.data, cs.ctypes.data, vo.ctypes.data,
                                         eo.ctypes.data, cv.ctypes.data))
        return {"vert_perm": vp, "copy_src": cs, "vert_off": vo, "ele_off": eo,
                "copy_verts": cv.reshape(F, nc)}

    def patch_lists(self):
        info = self.layout_info()
        N, F, P = info["n_verts"], info["n_copies"], info["n_patches"]
        out = {"halo_off": np.empty(P + 1, np.int32), "halo_ids": np.empty(info["n_halo"], np.int32),
               "copy_local": np.empty((F, 4), np.uint16), "vslot_off": np.empty(N + 1, np.int32),
               "vslots": np.empty(info["n_slots"], np.uint16), "padj_off": np.empty(P + 1, np.int32),
               "padj": np.empty(info["n_adjacent"], np.int32)}
        self._check(self.L.ls_get_patch_lists(self.ctx, *(out[k].ctypes.data for k in
                                                          ("halo_off", "halo_ids", "copy_local", "vslot_off",
                                                           "vslots", "padj_off", "padj"))))
        return out

    def set_phi(self, phi):
        a = np.ascontiguousarray(phi, dtype=np.float64)
        self._check(self.L.ls_set_phi(self.ctx, a.ctypes.data, LS_F64))

    def set_velocity(self, vel, mode=LS_VEL_ELEMENT):
        a = np.ascontiguousarray(vel, dtype=np.float64)
        self._check(self.L.ls_set_velocity(self.ctx, a.ctypes.data, int(mode)))

    def step(self, params: StepParams) -> Stats:
        st = Stats()
        self._check(self.L.ls_step(self.ctx, C.byref(params), C.byref(st)))
        return st

    def phi(self):
        out = np.empty(self.N, dtype=np.float64)
        self._check(self.L.ls_get_phi(self.ctx, out.ctypes.data, LS_F64))
        return out

    # -- batches of independent meshes ------------------------------------
    def set_mesh_batch(self, dim, meshes):
        """meshes: list of (xyz (n_m, 3), cells (e_m, dim+1)) with mesh-local
        vertex ids; they are concatenated in order."""
        voff = np.zeros(len(meshes) + 1, np.int64)
        coff = np.zeros(len(meshes) + 1, np.int64)
        for m, (x, c) in enumerate(meshes):
            voff[m + 1] = voff[m] + x.shape[0]
            coff[m + 1] = coff[m] + c.shape[0]
        if voff[-1] >= 2 ** 31:
            raise LevelSetError("batch too large for int32 vertex ids")
        self._xyz = np.ascontiguousarray(np.concatenate([x for x, _ in meshes]), dtype=np.float64)
        self._cells = np.concatenate([np.asarray(c, np.int32) + np.int32(voff[m]) for m, (_, c) in enumerate(meshes)])
        self._voff, self._coff = voff, coff
        self.dim, self.N, self.E = dim, int(voff[-1]), int(coff[-1])
        self.n_meshes = len(meshes)
        self._check(self.L.ls_set_mesh_batch(self.ctx, dim, self.n_meshes, voff.ctypes.data, self._xyz.ctypes.data,
                                             coff.ctypes.data, self._cells.ctypes.data))

    def partition_batch(self, patch_verts):
        self._check(self.L.ls_partition_batch(self.ctx, int(patch_verts)))

    def set_batch_params(self, dt, bandwidth):
        self._bdt = np.ascontiguousarray(dt, dtype=np.float64)
        self._bbw = np.ascontiguousarray(bandwidth, dtype=np.float64)
        self._check(self.L.ls_set_batch_params(self.ctx, self._bdt.ctypes.data, self._bbw.ctypes.data))

    def batch_nb_counts(self):
        out = np.zeros(self.n_meshes, np.int32)
        self._check(self.L.ls_get_batch_nb_counts(self.ctx, out.ctypes.data))
        return out

    def distances(self):
        out = np.empty(self.N, dtype=np.float64)
        self._check(self.L.ls_get_distances(self.ctx, out.ctypes.data))
        return out

    def set_warm(self, warm):
        """Predictor distances (caller order) for the next LS_WARM_LEARNED step."""
        a = np.ascontiguousarray(warm, dtype=np.float64)
        self._check(self.L.ls_set_warm(self.ctx, a.ctypes.data))

    def compute_dt(self):
        """(CFL bound, largest inscribed radius) over the current narrowband."""
        dt, rin = C.c_double(), C.c_double()
        self._check(self.L.ls_compute_dt(self.ctx, C.byref(dt), C.byref(rin)))
        return dt.value, rin.value

    def set_cfl_coefficients(self, ceik=None, ccurv=None):
        """Per-element eikonal and curvature coefficients of the CFL bound
        (None: the reference's 10 and 0)."""
        self._ceik = None if ceik is None else np.ascontiguousarray(ceik, dtype=np.float64)
        self._ccurv = None if ccurv is None else np.ascontiguousarray(ccurv, dtype=np.float64)
        self._check(self.L.ls_set_cfl_coefficients(
            self.ctx, None if self._ceik is None else self._ceik.ctypes.data,
            None if self._ccurv is None else self._ccurv.ctypes.data))

    def warm_values(self):
        """The predictor distances the context holds (caller order)."""
        out = np.empty(self.N, dtype=np.float64)
        self._check(self.L.ls_get_warm(self.ctx, out.ctypes.data))
        return out

    # -- learned warm start ------------------------------------------------
    def mlp_set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        assert a.size == MLP_PARAMS
        self._check(self.L.ls_mlp_set_params(self.ctx, a.ctypes.data, MLP_PARAMS))

    def mlp_params(self):
        out = np.empty(MLP_PARAMS, np.float32)
        self._check(self.L.ls_mlp_get_params(self.ctx, out.ctypes.data, MLP_PARAMS))
        return out

    def mlp_predict(self):
        self._check(self.L.ls_mlp_predict(self.ctx))

    def mlp_collect(self, phase, capacity=0):
        self._check(self.L.ls_mlp_collect(self.ctx, int(phase), int(capacity)))

    def mlp_dataset_size(self):
        n = C.c_int64(0)
        self._check(self.L.ls_mlp_dataset(self.ctx, None, None, C.byref(n)))
        return n.value

    def mlp_dataset(self):
        n = C.c_int64(0)
        self._check(self.L.ls_mlp_dataset(self.ctx, None, None, C.byref(n)))
        X = np.empty((n.value, 22), np.float32)
        Y = np.empty((n.value, 4), np.float32)
        if n.value:
            self._check(self.L.ls_mlp_dataset(self.ctx, X.ctypes.data, Y.ctypes.data, C.byref(n)))
        return X, Y

    def mlp_gradient(self, X, Y, noalloc=False):
        """(gradient of the mean loss, loss sum) over caller samples; noalloc
        uses ls_mlp_gradient_noalloc (no device allocation, n <= 16384)."""
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_double(0.0)
        fn = self.L.ls_mlp_gradient_noalloc if noalloc else self.L.ls_mlp_gradient
        self._check(fn(self.ctx, X.ctypes.data, Y.ctypes.data, X.shape[0], g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def mlp_train(self, epochs, batch, lr, seed=1):
        loss = C.c_double(0.0)
        self._check(self.L.ls_mlp_train(self.ctx, int(epochs), int(batch), float(lr), int(seed), C.byref(loss)))
        return loss.value

    def split(self, values):
        """Per-mesh views of a concatenated per-vertex array."""
        return [values[self._voff[m]:self._voff[m + 1]] for m in range(self.n_meshes)]

    def narrowband(self):
        n = C.c_int32(0)
        self._check(self.L.ls_get_narrowband(self.ctx, None, C.byref(n)))
        ids = np.empty(max(n.value, 1), np.int32)
        self._check(self.L.ls_get_narrowband(self.ctx, ids.ctypes.data, C.byref(n)))
        return ids[: n.value].copy()


def solver_for_case(case, device=0, precision=LS_F32, execution=LS_EXEC_PERSISTENT):
    s = Solver(device)
    s.set_mesh(case.dim, case.xyz, case.cells)
    s.set_precision(precision)
    s.set_execution(execution)
    s.partition_labels(case.labels, case.nparts)
    s.set_phi(case.phi0)
    s.set_velocity(case.vel)
    return s


def mlp_init_params(seed=0):
    """Initial network parameters (flat layout of ls_mlp_set_params):
    He-normal W1 and W2, W3 scaled by 0.1 so the first corrections are
    small, zero biases."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, np.sqrt(2.0 / 22), (22, 64))
    w2 = rng.normal(0, np.sqrt(2.0 / 64), (64, 64))
    w3 = rng.normal(0, 0.1 * np.sqrt(2.0 / 64), (64, 4))
    return np.concatenate([w1.ravel(), np.zeros(64), w2.ravel(), np.zeros(64), w3.ravel(),
                           np.zeros(4)]).astype(np.float32)


def solver_for_batch(batch, device=0, execution=LS_EXEC_GRAPH, patch_verts=None, labels=None, precision=LS_F32):
    """A Solver holding a cases.Batch; patches from recursive bisection
    (patch_verts) or from caller labels (labels, nparts)."""
    s = Solver(device)
    s.set_mesh_batch(batch.dim, [(m.xyz, m.cells) for m in batch.meshes])
    s.set_precision(precision)
    s.set_execution(execution)
    if labels is not None:
        s.partition_labels(labels[0], labels[1])
    else:
        s.partition_batch(patch_verts or batch.patch_verts)
    s.set_batch_params([m.dt for m in batch.meshes], [m.bandwidth for m in batch.meshes])
    s.set_phi(np.concatenate([m.phi0 for m in batch.meshes]))
    s.set_velocity(np.concatenate([m.vel for m in batch.meshes]))
    return s


def autotune_batch(batch, candidates, executions=(LS_EXEC_GRAPH, LS_EXEC_PERSISTENT, LS_EXEC_KERNELS),
                   trial_steps=3, device=0, log=None, precision=LS_F32):
    """Times `trial_steps` steps of the batch for every patching candidate and
    execution mode and returns the fastest as (candidate, execution, ms/step,
    table). A candidate is either an int (recursive-bisection patch size) or
    a (name, labels, nparts) tuple (caller block partition). Every trial
    starts from the batch's initial state, with the given storage precision."""
    table = []
    best = None
    for cand in candidates:
        for ex in executions:
            kw = {"patch_verts": cand} if isinstance(cand, int) else {"labels": (cand[1], cand[2])}
            name = cand if isinstance(cand, int) else cand[0]
            try:
                s = solver_for_batch(batch, device=device, execution=ex, precision=precision, **kw)
            except LevelSetError as e:  # a patch of this size exceeds the device's LDS
                if "rc=-4" not in str(e):
                    raise
                if log:
                    log(f"autotune: patches {name} skipped: {e}")
                break
            prm = step_params(1.0, 1.0, batch.niter, mode=batch.mode)
            s.step(prm)  # first step: graph capture and warm caches
            ms = 0.0
            for _ in range(trial_steps):
                ms += s.step(prm).ms_total
            s.close()
            ms /= trial_steps
            table.append((name, ex, ms))
            if log:
                log(f"autotune: patches {name} exec {ex}: {ms:.3f} ms/step")
            if best is None or ms < best[2]:
                best = (cand, ex, ms)
    return best[0], best[1], best[2], table
