#!/usr/bin/env python3
"""Benchmark: narrowband vertex-updates/s of the level-set timestep on one MI355X.

Workload: the 3D torus revolve -- a 64^3 Kuhn tet mesh (262,144 vertices,
1,500,282 tets), jitter 0.1h, torus (R=0.3, r=0.1), rigid rotation about z
(omega = 2 pi), CFL 0.5, band 6h, 256-vertex patches, 100 timesteps. A
vertex-update is one owned vertex of a narrowband patch receiving its final
phi for a step (the set CopyOutBack_levelset3d writes,
src/core3d/include/meshFIM_kernels.h:22-34). One bench "step" is one full
timestep (seeds, sign save, FIM, narrowband + sign restore, advection).

Storage is FP64 (the engine bit-exact with an FP64 host build); FIM
tolerance 1e-6. Besides the JSON line's throughput it reports:
  * roofline of one timestep's launches (the persistent seeds/FIM/narrowband
    kernel and the advection after it), from HIP-event durations and the
    algorithmic bytes of the patches they processed, with the FP64 compute
    rate beside it and the counter-measured HBM traffic from profiles/;
  * parity: the FP64 oracle run beside the engine from phi0 over the timed
    steps, compared on every finite vertex, in the band and by narrowband;
  * the CPU baseline: that oracle trajectory's time (OpenMP over patches,
    all host threads);
  * the FP32-storage engine on the same steps with its trajectory error, and
    the learned and identity warm starts against plain FIM.
--workload mcf / batch run the 192^3 mean-curvature flow and the 512-mesh
batch (FP32 storage), checked along a trajectory against the oracle's FP32
(bit for bit) and FP64 instances on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

HBM_PEAK_GBS = 8000.0


def step_bytes(st, w=4):
    """Algorithmic bytes of a whole timestep: seeds, FIM visits and probes,
    advection. w is the width of a stored real (4: FP32 storage, 8: FP64);
    connectivity is 16 B per element copy either way."""
    seeds = 16 * st.seed_copies + w * (st.seed_verts + st.seed_halo) + (2 * w + 1) * st.seed_verts
    return seeds + fim_bytes(st, w) + adv_bytes(st, w)


def fim_bytes(st, w=4):
    """Algorithmic bytes of FIM patch visits and probes: per visit
    (16 + 6w) E + w (V + H) + (w + 1) V (connectivity and six local
    coordinates per element copy, distances in, distances out + 1 B flag;
    40 E + 4 (V + H) + 5 V with FP32 storage); probes do not write distances."""
    visits = (16 + 6 * w) * st.visit_copies + w * (st.visit_verts + st.visit_halo) + (w + 1) * st.visit_verts
    probes = (16 + 6 * w) * st.probe_copies + w * (st.probe_verts + st.probe_halo) + 1 * st.probe_verts
    return visits + probes


def adv_bytes(st, w=4):
    """Advection per narrowband patch: (16 + 3w) E + 4w (V + H) + w V
    (connectivity + per-element velocity; coordinates and phi per vertex;
    phi out; 28 E + 16 (V + H) + 4 V with FP32 storage)."""
    return (16 + 3 * w) * st.adv_copies + 4 * w * (st.adv_verts + st.adv_halo) + w * st.adv_verts


def trajectory(case, oracle_mod, lsb, steps, engines, prm_kw):
    """GPU engines and the FP64 oracle each run from phi0, side by side, with
    no restart: after every step each engine's phi is compared with the
    oracle's over every vertex both hold finite (unreached vertices are
    +-LARGENUM on both sides; a vertex finite on one side only counts as a
    mismatch) and over the vertices inside the band, and the narrowband
    patch sets are compared. The oracle's time is the CPU baseline."""
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts, 64)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    op = oracle_mod.make_params(case.dt, case.bandwidth, case.niter, tol=prm_kw.get("tol", 1e-5))
    gp = lsb.step_params(case.dt, case.bandwidth, case.niter, **prm_kw)
    gs = {name: lsb.solver_for_case(case, precision=prec, execution=ex) for name, prec, ex in engines}
    res = {name: {"steps": 0, "linf": 0.0, "linf_in_band": 0.0, "finite_mismatch": 0, "nb_diff": 0,
                  "bit_mismatches": 0, "linf_per_step": []} for name in gs}
    cpu_s, cpu_v, done = 0.0, 0, 0
    for _ in range(steps):
        t0 = time.perf_counter()
        ost = o.step(op)
        cpu_s += time.perf_counter() - t0
        cpu_v += ost.nb_verts
        ophi, onb = o.phi(), o.narrowband()
        ofin = np.abs(ophi) < 1e8
        for name, g in gs.items():
            g.step(gp)
            gphi = g.phi()
            gfin = np.abs(gphi) < 1e8
            fin = ofin & gfin
            d = np.abs(gphi - ophi)
            inb = fin & (np.abs(ophi) <= case.bandwidth)
            r = res[name]
            step_linf = float(np.max(d[fin])) if fin.any() else 0.0
            r["linf"] = max(r["linf"], step_linf)
            r["linf_in_band"] = max(r["linf_in_band"], float(np.max(d[inb])) if inb.any() else 0.0)
            r["linf_per_step"].append(step_linf)
            r["finite_mismatch"] += int(np.count_nonzero(ofin != gfin))
            r["nb_diff"] += int(np.setxor1d(g.narrowband(), onb).size)
            r["bit_mismatches"] += int(np.count_nonzero(gphi != ophi))
            r["steps"] += 1
        done += 1
        if ost.nb_count == 0:
            break
    for g in gs.values():
        g.close()
    o.close()
    return res, {"cpu_s": cpu_s, "cpu_verts": cpu_v, "steps": done}


class SideBySide:
    """An engine checked along a trajectory against the oracle, each stepping
    from phi0 with no restart: bit for bit against the oracle instance of
    the engine's storage type, and by L-inf against the FP64 instance over
    every vertex both hold finite and inside the band (for an FP64 engine
    the two are one instance). One entry per mesh; the time of the instance
    of the engine's storage type is the CPU baseline."""

    def __init__(self, O, steps, engine_bits=32):
        self.O, self.steps, self.bits, self.meshes = O, steps, engine_bits, []
        self.cpu_s, self.cpu_v = 0.0, 0
        self.r = {"steps": 0, f"fp{engine_bits}_bit_mismatches": 0, f"fp{engine_bits}_nb_diff": 0,
                  "linf_vs_fp64": 0.0, "linf_vs_fp64_in_band": 0.0, "fp64_finite_set_mismatches": 0,
                  "fp64_nb_diff": 0, "linf_vs_fp64_per_step": []}

    def add(self, dim, xyz, cells, labels, nparts, phi0, vel, params, bandwidth):
        pair = []
        for bits in sorted({self.bits, 64}):
            o = self.O.Oracle(dim, xyz, cells, labels, nparts, bits)
            o.set_phi(phi0)
            o.set_velocity(vel)
            pair.append(o)
        self.meshes.append((pair[0], pair[-1], params, bandwidth))

    def step(self, phis, nbs):
        """One step of every oracle pair against the engine's phi and its
        narrowband, per mesh: the patch ids, or the narrowband's size where
        the engine numbers patches differently (batches)."""
        step_linf = 0.0
        for (o32, o64, prm, bw), phi, nb in zip(self.meshes, phis, nbs):
            t0 = time.perf_counter()
            st = o32.step(prm)
            self.cpu_s += time.perf_counter() - t0
            self.cpu_v += st.nb_verts
            if o64 is not o32:
                o64.step(prm)
            key = f"fp{self.bits}_"
            self.r[key + "bit_mismatches"] += int(np.count_nonzero(phi != o32.phi()) if self.bits == 64 else
                                                  np.count_nonzero(phi.astype(np.float32) != o32.phi().astype(np.float32)))
            self.r[key + "nb_diff"] += self._nb_diff(nb, o32.narrowband())
            ophi = o64.phi()
            ofin, gfin = np.abs(ophi) < 1e8, np.abs(phi) < 1e8
            fin = ofin & gfin
            d = np.abs(phi.astype(np.float64) - ophi)
            inb = fin & (np.abs(ophi) <= bw)
            if fin.any():
                step_linf = max(step_linf, float(np.max(d[fin])))
            if inb.any():
                self.r["linf_vs_fp64_in_band"] = max(self.r["linf_vs_fp64_in_band"], float(np.max(d[inb])))
            self.r["fp64_finite_set_mismatches"] += int(np.count_nonzero(ofin != gfin))
            self.r["fp64_nb_diff"] += self._nb_diff(nb, o64.narrowband())
        self.r["linf_vs_fp64"] = max(self.r["linf_vs_fp64"], step_linf)
        self.r["linf_vs_fp64_per_step"].append(float(f"{step_linf:.4g}"))
        self.r["steps"] += 1

    @staticmethod
    def _nb_diff(nb, onb):
        if np.isscalar(nb):
            return abs(int(nb) - int(onb.size))
        return int(np.setxor1d(nb, onb).size)

    def finish(self, sample):
        for o32, o64, _, _ in self.meshes:
            o32.close()
            if o64 is not o32:
                o64.close()
        cpu = {"value": self.cpu_v / self.cpu_s if self.cpu_s > 0 else None, "unit": "vertex-updates/s",
               "cores": self.O.max_threads(), "kind": "port", "sample": sample}
        par = dict(self.r, oracle=("FP64 host build of the same scheme, from phi0" if self.bits == 64 else
                                   "FP32-storage and FP64 host builds of the same scheme, each from phi0"),
                   tolerance=1e-5, meets_bar=self.r[f"fp{self.bits}_bit_mismatches"] == 0
                   and self.r["linf_vs_fp64"] <= 1e-5)
        return cpu, par


def profiled_traffic(name):
    """HBM bytes per launch that the committed rocprofv3 counter summary
    (profiles/<name>.md, FETCH_SIZE and WRITE_SIZE scaled by a 1 GiB copy)
    records, or None."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("HBM traffic per launch (bytes):"):
                return float(line.split(":")[1])
    return None


def timed_run(g, prm, steps, w=4):
    """Times `steps` timesteps; sums the per-step statistics (w: bytes per
    stored real, for the algorithmic byte counts)."""
    tot = {"verts": 0, "sweep_ms": 0.0, "sweep_launches": 0, "sweep_bytes": 0, "adv_ms": 0.0,
           "adv_bytes": 0, "adv_launches": 0, "ms_fim": 0.0, "ms_total": 0.0, "fim_iters": 0,
           "visits": 0, "probes": 0, "step_ms": 0.0, "step_bytes": 0, "exec_modes": set(),
           "step_wg_per_cu": set(), "solves": 0}
    t0 = time.perf_counter()
    for _ in range(steps):
        st = g.step(prm)
        tot["verts"] += st.nb_verts
        tot["sweep_ms"] += st.ms_sweep_kernels
        tot["sweep_launches"] += st.sweep_launches
        tot["sweep_bytes"] += fim_bytes(st, w)
        tot["adv_ms"] += st.ms_advect_kernels
        tot["adv_bytes"] += adv_bytes(st, w)
        tot["adv_launches"] += 1 if st.nb_count > 0 else 0
        tot["ms_fim"] += st.ms_fim
        tot["ms_total"] += st.ms_total
        tot["fim_iters"] += st.fim_iters
        tot["visits"] += st.patch_visits
        tot["probes"] += st.probes
        tot["step_ms"] += st.ms_step_kernel
        tot["step_bytes"] += step_bytes(st, w)
        tot["exec_modes"].add(st.exec_mode)
        tot["step_wg_per_cu"].add(st.step_wg_per_cu)
        tot["solves"] += st.local_solves
        tot["sweep_tasks"] = st.sweep_tasks
    tot["wall"] = time.perf_counter() - t0
    return tot


def roofline(tot, steps, kernels_mode):
    if kernels_mode:
        kname = "k_sweep (FIM patch sweep)"
        launches = tot["sweep_launches"]
        avg_ms = tot["sweep_ms"] / max(launches, 1)
        avg_bytes = tot["sweep_bytes"] / max(launches, 1)
    else:
        kname = "k_step + k_advect_dev + k_copy_phi_dev (one timestep: persistent seeds/FIM/narrowband, then advection)"
        launches = steps
        avg_ms = tot["step_ms"] / steps
        avg_bytes = tot["step_bytes"] / steps
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    return {"kernel": kname, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "avg_launch_ms": avg_ms,
            "avg_launch_algorithmic_bytes": avg_bytes, "launches": launches}


# FP64 operations of one tet corner solve in this build's gfx950 code
# (v_mul_f64 201, v_add_f64 173, v_fma_f64 130 counted twice; square roots
# and divisions excluded), and the MI355X FP64 vector peak.
FLOPS_PER_SOLVE = 634
FP64_PEAK_TFLOPS = 78.6


def compute_fraction(tot, steps, kernels_mode):
    """FP64 arithmetic rate of the local solves over the same kernel time as
    the roofline entry, against the FP64 vector peak."""
    ms = (tot["sweep_ms"] if kernels_mode else tot["step_ms"]) / steps
    flops = FLOPS_PER_SOLVE * tot["solves"] / steps
    achieved = flops / (ms * 1e-3) / 1e12
    return {"unit": "TFLOP/s", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "frac": achieved / FP64_PEAK_TFLOPS,
            "flops_per_launch": flops, "flops_per_solve": FLOPS_PER_SOLVE}


def batch_workload(args, lsb, cases):
    """512 uneven ellipsoid meshes in one launch per timestep (hipGraph), the
    patch size and execution mode chosen by the autotuner on a 64-mesh
    subset. CPU baseline and parity on a sample of the meshes."""
    b = cases.config4(steps=args.steps)
    prec = lsb.LS_F64 if args.precision == "f64" else lsb.LS_F32
    print(f"batch: {len(b.meshes)} meshes, {b.n_verts} vertices, {b.n_cells} tets", file=sys.stderr, flush=True)
    if args.batch_patch != "auto" and args.exec != "auto":
        names = {"graph": lsb.LS_EXEC_GRAPH, "persistent": lsb.LS_EXEC_PERSISTENT, "kernels": lsb.LS_EXEC_KERNELS}
        cand, ex, table = int(args.batch_patch), names[args.exec], []
    else:
        sub = cases.Batch(b.name + "_tune", 3, b.meshes[:64], patch_verts=b.patch_verts)
        cand, ex, ms, table = lsb.autotune_batch(sub, [64, 128, 256, 512], trial_steps=2, precision=prec,
                                                 log=lambda m: print(m, file=sys.stderr, flush=True))
    t0 = time.perf_counter()
    g = lsb.solver_for_batch(b, execution=ex, patch_verts=cand, precision=prec)
    setup_s = time.perf_counter() - t0
    info = g.layout_info()
    prm = lsb.step_params(1.0, 1.0, 5)
    for _ in range(args.warmup):
        g.step(prm)
    g.set_phi(np.concatenate([m.phi0 for m in b.meshes]))
    tot = timed_run(g, prm, args.steps, w=8 if prec == lsb.LS_F64 else 4)
    g.close()
    learned = None
    if not args.no_learned:
        learned = learned_batch(args, lsb, b, cand, ex, prec, tot["verts"] / tot["wall"])
    out = {"workload": f"{b.name}: {len(b.meshes)} Kuhn meshes, {b.n_verts} verts, {b.n_cells} tets, "
                       f"ellipsoids, affine velocity, CFL 0.5, band 6h, {args.steps} timesteps",
           "patch_verts": cand, "execution": {lsb.LS_EXEC_GRAPH: "graph", lsb.LS_EXEC_PERSISTENT: "persistent",
                                              lsb.LS_EXEC_KERNELS: "kernels"}[ex],
           "autotune_ms_per_step": {f"{c}/{e}": m for c, e, m in table} or None,
           "patches": info["n_patches"], "element_copies": info["n_copies"],
           "setup_s": setup_s, "learned": learned}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        sample = cases.Batch(b.name + "_sample", 3, b.meshes[:24], patch_verts=cand)
        gs = lsb.solver_for_batch(sample, execution=ex, patch_verts=cand, precision=prec)
        cmp = SideBySide(O, args.cpu_steps, engine_bits=64 if prec == lsb.LS_F64 else 32)
        for m in sample.meshes:
            nparts = -(-m.xyz.shape[0] // cand)
            cmp.add(3, m.xyz, m.cells, cases.partition_rcb(m.xyz, nparts), nparts, m.phi0, m.vel,
                    O.make_params(m.dt, m.bandwidth, 5), m.bandwidth)
        for _ in range(args.cpu_steps):
            gs.step(prm)
            cmp.step(gs.split(gs.phi()), [int(c) for c in gs.batch_nb_counts()])
        gs.close()
        cpu, parity = cmp.finish(f"first {args.cpu_steps} timesteps of the first 24 meshes from phi0, the "
                                 f"oracle's {args.precision.upper()} instance (the scheme of this engine), "
                                 "OpenMP over patches")
        cpu = (cpu, dict(parity, sample_meshes=24))
    return out, tot, ex == lsb.LS_EXEC_KERNELS, cpu


def mcf_workload(args, lsb, cases):
    """Mean-curvature flow of 8 spheres on a 192^3 Kuhn grid (41.8M tets)."""
    case = cases.config3(steps=args.steps)
    prec = lsb.LS_F64 if args.precision == "f64" else lsb.LS_F32
    print(f"mcf: {case.xyz.shape[0]} vertices, {case.cells.shape[0]} tets", file=sys.stderr, flush=True)
    names = {"graph": lsb.LS_EXEC_GRAPH, "persistent": lsb.LS_EXEC_PERSISTENT, "kernels": lsb.LS_EXEC_KERNELS}
    prm = lsb.step_params(case.dt, case.bandwidth, case.niter, mode=lsb.LS_MODE_MCF)
    # Patch size: block partitions timed over a few steps in the persistent
    # mode (ms per step: the vertex-update count grows with the patch size).
    patch_tune = {}
    if args.mcf_block == "auto":
        for blk in [(4, 4, 4), (8, 4, 4), (8, 8, 4), (8, 8, 8)]:
            case.labels, case.nparts = cases.partition_blocks(192, blk, 3)
            try:
                gt = lsb.solver_for_case(case, precision=prec, execution=lsb.LS_EXEC_PERSISTENT)
            except lsb.LevelSetError as e:
                print(f"mcf block {blk} skipped: {e}", file=sys.stderr, flush=True)
                continue
            gt.step(prm)
            patch_tune["x".join(map(str, blk))] = sum(gt.step(prm).ms_total for _ in range(3)) / 3
            gt.close()
            print(f"mcf block {blk}: {patch_tune['x'.join(map(str, blk))]:.2f} ms/step", file=sys.stderr, flush=True)
        best = min(patch_tune, key=patch_tune.get)
        block = tuple(int(x) for x in best.split("x"))
    else:
        block = tuple(int(x) for x in args.mcf_block.split("x"))
    case.labels, case.nparts = cases.partition_blocks(192, block, 3)
    t0 = time.perf_counter()
    g = lsb.solver_for_case(case, precision=prec)
    setup_s = time.perf_counter() - t0
    info = g.layout_info()
    tune = None
    if args.exec == "auto":
        tune = autotune_exec(lsb, case, names, steps=3, g=g, prm=prm)
        args.exec = min(tune, key=tune.get)
    ex = names[args.exec]
    g.set_execution(ex)
    for _ in range(args.warmup):
        g.step(prm)
    g.set_phi(case.phi0)
    tot = timed_run(g, prm, args.steps, w=8 if prec == lsb.LS_F64 else 4)
    g.close()
    learned = None
    if not args.no_learned:
        O = None
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
        learned = learned_run(args, lsb, case, ex, tot["verts"] / tot["wall"], prec, {"mode": lsb.LS_MODE_MCF}, O)
    out = {"workload": f"{case.name}: 192^3 Kuhn tets ({case.xyz.shape[0]} verts, {case.cells.shape[0]} tets), "
                       f"8 spheres, mean-curvature flow, dt = h^2/8, band 8h, {args.steps} timesteps",
           "execution": args.exec, "execution_trial_ms_per_step": tune, "patch_block": "x".join(map(str, block)),
           "patch_trial_ms_per_step": patch_tune or None, "patches": info["n_patches"],
           "element_copies": info["n_copies"], "setup_s": setup_s, "learned": learned}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        nsteps = min(args.cpu_steps, 4)
        g = lsb.solver_for_case(case, precision=prec, execution=ex)
        cmp = SideBySide(O, nsteps, engine_bits=64 if prec == lsb.LS_F64 else 32)
        cmp.add(case.dim, case.xyz, case.cells, case.labels, case.nparts, case.phi0, case.vel,
                O.make_params(case.dt, case.bandwidth, case.niter, mode=1), case.bandwidth)
        for _ in range(nsteps):
            g.step(prm)
            cmp.step([g.phi()], [g.narrowband()])
        g.close()
        cpu, parity = cmp.finish(f"first {nsteps} timesteps of the same workload from phi0, the oracle's "
                                 f"{args.precision.upper()} instance (the scheme of this engine), OpenMP over patches")
        cpu = (cpu, parity)
    return out, tot, ex == lsb.LS_EXEC_KERNELS, cpu


def autotune_exec(lsb, case, names, steps=6, g=None, prm=None, precision=None, prm_kw=None):
    """Mean GPU ms/step of each execution mode over a few steps from phi0
    (after one untimed step) on one solver; the bench runs the fastest."""
    own = g is None
    if own:
        g = lsb.solver_for_case(case, precision=lsb.LS_F32 if precision is None else precision)
    prm = prm or lsb.step_params(case.dt, case.bandwidth, case.niter, mode=case.mode, **(prm_kw or {}))
    out = {}
    for name, mode in names.items():
        g.set_execution(mode)
        g.set_phi(case.phi0)
        g.step(prm)
        out[name] = sum(g.step(prm).ms_total for _ in range(steps)) / steps
    if own:
        g.close()
    return out


WARM_MARGIN, WARM_ABS_H = 0.03, 0.25  # warm value = prediction * 1.03 + h / 4
PERFECT_ABS = 1e-5  # the best-case predictor: the plain step's own distances + 1e-5
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix peak
MLP_FWD_FLOPS = 2 * (22 * 64 + 64 * 64 + 64 * 4)  # per element copy (11,520)


def mlp_throughput(g, n_train, train_s, epochs):
    """The network's inference over the current narrowband: ls_mlp_predict
    (one launch, synchronised) timed over repeated calls, its algorithmic
    FLOPs (22-64-64-4 per element copy) against the BF16 MFMA peak; and the
    training rate end to end (forward + backward, about 3x the forward
    FLOPs per pair, host round trips per batch included)."""
    nb = g.narrowband()
    eo = g.layout()["ele_off"]
    copies = int(np.sum(eo[nb + 1] - eo[nb]))
    g.mlp_predict()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        g.mlp_predict()
    ms = 1e3 * (time.perf_counter() - t0) / reps
    tf = copies * MLP_FWD_FLOPS / (ms * 1e-3) / 1e12
    train_tf = epochs * n_train * 3 * MLP_FWD_FLOPS / train_s / 1e12 if train_s > 0 else None
    return {"predict": {"element_copies": copies, "ms": ms, "tflops": tf, "peak_tflops": BF16_MFMA_PEAK_TFLOPS,
                        "frac_of_bf16_mfma_peak": tf / BF16_MFMA_PEAK_TFLOPS,
                        "note": "host-timed launch + synchronisation; a small dense MLP per copy whose inputs "
                                "are gathered from the mesh, so memory and launch latency, not the MFMA rate, "
                                "bound it"},
            "train": {"pairs_x_epochs": epochs * n_train, "s": train_s, "tflops_end_to_end": train_tf}}


def learned_run(args, lsb, case, exec_mode, plain_value, precision, prm_kw, oracle_mod=None):
    """The learned warm start on the same workload and engine: pairs
    collected from plain steps on the GPU, GPU training (the pairs of the
    last two collection steps held out for scoring), then the timed steps
    with warm values from the network (prediction time included), and the
    identity predictor (|phi|) timed the same way. A warm step in which a
    vertex inside the band still holds a warm value below what its elements
    give it re-runs the plain method (counted as a fallback). With an oracle,
    the first accepted warm steps are compared with the FP64 plain step from
    the same input."""
    # the plain line's execution mode (warm starts need the persistent kernel)
    wexec = exec_mode if exec_mode != lsb.LS_EXEC_KERNELS else lsb.LS_EXEC_GRAPH
    g = lsb.solver_for_case(case, precision=precision, execution=wexec)
    g.mlp_set_params(lsb.mlp_init_params(5))
    prm = lsb.step_params(case.dt, case.bandwidth, case.niter, **prm_kw)
    g.step(prm)
    t0 = time.perf_counter()
    last = 0
    for _ in range(args.train_steps):  # room is kept for the held-out pairs
        n0 = g.mlp_dataset_size()
        if n0 + last > 10_500_000:
            break
        g.mlp_collect(0, 12_000_000)
        g.step(prm)
        g.mlp_collect(1)
        last = g.mlp_dataset_size() - n0
    t1 = time.perf_counter()
    loss = g.mlp_train(args.train_epochs, 4096, 2e-3, seed=7)
    t2 = time.perf_counter()
    n_train = g.mlp_dataset_size()
    for _ in range(2):  # held-out pairs: two later timesteps, never trained on
        g.mlp_collect(0, 12_000_000)
        g.step(prm)
        g.mlp_collect(1)
    X, Y = g.mlp_dataset()
    mlp = mlp_throughput(g, n_train, t2 - t1, args.train_epochs)
    held = slice(n_train, X.shape[0])
    _, sse = g.mlp_gradient(X[held], Y[held])
    trained_mse = sse / max(Y[held].size, 1)
    identity_mse = float(np.mean(Y[held].astype(np.float64) ** 2))

    def timed(warm):
        wprm = lsb.step_params(case.dt, case.bandwidth, case.niter, warm=warm, warm_margin=WARM_MARGIN,
                               warm_abs_margin=WARM_ABS_H * case.h, **prm_kw)
        g.set_phi(case.phi0)
        r = {"verts": 0, "fallbacks": 0, "iters": 0, "visits": 0}
        ts = time.perf_counter()
        for k in range(args.steps):
            if warm == lsb.LS_WARM_LEARNED and k > 0:
                g.mlp_predict()
            st = g.step(wprm if k > 0 else prm)
            r["verts"] += st.nb_verts
            r["fallbacks"] += st.warm_fallback
            r["iters"] += st.fim_iters
            r["visits"] += st.patch_visits
        r["wall"] = time.perf_counter() - ts
        return r

    def timed_perfect():
        # The best any predictor can do: warm values from the plain method's
        # own distances of the same step on the same input (a second context
        # computes them, outside the timed region), plus PERFECT_ABS.
        g2 = lsb.solver_for_case(case, precision=precision, execution=wexec)
        wprm = lsb.step_params(case.dt, case.bandwidth, case.niter, warm=lsb.LS_WARM_LEARNED, warm_margin=0.0,
                               warm_abs_margin=PERFECT_ABS, warm_res_tol=PERFECT_ABS, **prm_kw)
        g.set_phi(case.phi0)
        r = {"verts": 0, "fallbacks": 0, "iters": 0, "visits": 0, "wall": 0.0}
        for k in range(args.steps):
            if k > 0:
                g2.set_phi(g.phi())
                g2.step(prm)
                g.set_warm(g2.distances())
            ts = time.perf_counter()
            st = g.step(wprm if k > 0 else prm)
            r["wall"] += time.perf_counter() - ts
            r["verts"] += st.nb_verts
            r["fallbacks"] += st.warm_fallback
            r["iters"] += st.fim_iters
            r["visits"] += st.patch_visits
        g2.close()
        return r

    learned = timed(lsb.LS_WARM_LEARNED)
    ident = timed(lsb.LS_WARM_IDENTITY)
    perfect = timed_perfect()
    parity = None
    if oracle_mod is not None and args.learned_parity_steps > 0:
        # per step from the same input: the GPU's warm step against the FP64
        # oracle's plain step
        o = oracle_mod.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
        o.set_velocity(case.vel)
        g.set_phi(case.phi0)
        g.step(prm)
        errs, nbd, acc = [], 0, 0
        for _ in range(args.learned_parity_steps):
            phi, nb = g.phi(), g.narrowband()
            g.mlp_predict()
            st = g.step(lsb.step_params(case.dt, case.bandwidth, case.niter, warm=lsb.LS_WARM_LEARNED,
                                        warm_margin=WARM_MARGIN, warm_abs_margin=WARM_ABS_H * case.h, **prm_kw))
            o.set_phi(phi)
            o.set_narrowband(nb)
            o.step(oracle_mod.make_params(case.dt, case.bandwidth, case.niter, tol=prm_kw.get("tol", 1e-5),
                                          mode=prm_kw.get("mode", 0)))
            gphi, ophi = g.phi(), o.phi()
            fin = (np.abs(ophi) < 1e8) & (np.abs(gphi) < 1e8)
            band = fin & (np.abs(ophi) <= case.bandwidth)
            errs.append((float(np.max(np.abs(gphi - ophi)[band])), float(np.max(np.abs(gphi - ophi)[fin]))))
            nbd += int(np.setxor1d(g.narrowband(), o.narrowband()).size)
            acc += 1 - st.warm_fallback
        o.close()
        parity = {"steps": len(errs), "accepted_steps": acc, "linf_in_band_vs_fp64_plain": max(e[0] for e in errs),
                  "linf_all_finite_vs_fp64_plain": max(e[1] for e in errs), "nb_membership_diff": nbd,
                  "meets_bar": (max(e[0] for e in errs) <= 1e-5 and nbd == 0) if acc else None,
                  "why_not": ("the plain method's result depends on its schedule: slots lowered inside a patch "
                              "visit are never written back (redistance_kernels3d.h:440-449), so a warm start, "
                              "which changes which patches are active when, moves values inside the band by "
                              "~1e-4 (the plain method with 10 instead of 5 sweeps per visit moves them by "
                              "2.5e-4)") if acc else
                             "no checked warm step was accepted: each fell back and returned the plain bits"}
    g.close()
    value = learned["verts"] / learned["wall"]
    return {"value": value, "unit": "vertex-updates/s", "vs_plain_fim": value / plain_value,
            "execution": {lsb.LS_EXEC_GRAPH: "graph", lsb.LS_EXEC_PERSISTENT: "persistent"}[wexec],
            "ms_per_step": 1e3 * learned["wall"] / args.steps, "fim_iters_per_step": learned["iters"] / args.steps,
            "patch_visits_per_step": learned["visits"] / args.steps,
            "fallback_steps": learned["fallbacks"], "steps": args.steps,
            "identity_predictor": {"value": ident["verts"] / ident["wall"],
                                   "vs_plain_fim": ident["verts"] / ident["wall"] / plain_value,
                                   "fim_iters_per_step": ident["iters"] / args.steps,
                                   "patch_visits_per_step": ident["visits"] / args.steps,
                                   "fallback_steps": ident["fallbacks"]},
            "perfect_predictor": {"what": "warm values = the plain method's own distances of the same step "
                                          f"from the same input + {PERFECT_ABS:g}, accepted up to an undercut "
                                          f"of {PERFECT_ABS:g} (not a usable predictor: the ceiling of any warm "
                                          "start on this scheme); only the warm step is timed, per step",
                                  "value": perfect["verts"] / perfect["wall"],
                                  "vs_plain_fim": perfect["verts"] / perfect["wall"] / plain_value,
                                  "fim_iters_per_step": perfect["iters"] / args.steps,
                                  "patch_visits_per_step": perfect["visits"] / args.steps,
                                  "fallback_steps": perfect["fallbacks"]},
            "train_pairs": int(n_train), "held_out_pairs": int(X.shape[0] - n_train),
            "collect_s": t1 - t0, "train_s": t2 - t1, "train_epochs": args.train_epochs,
            "train_loss": loss, "held_out_mse": trained_mse, "held_out_identity_mse": identity_mse,
            "warm_margin": WARM_MARGIN, "warm_abs_margin_h": WARM_ABS_H, "mlp": mlp, "parity": parity}


def learned_batch(args, lsb, b, cand, ex, prec, plain_value):
    """The learned warm start on the whole batch (one network for all
    meshes): pairs from plain steps, GPU training, then the timed steps with
    warm values from the network and with the identity predictor, from the
    same initial state as the plain run. Each mesh accepts or rejects its
    warm start on its own (fallback_mesh_steps counts the meshes that re-ran
    plain, summed over the steps). The absolute margin is a quarter of
    the coarsest mesh spacing (one margin serves every mesh of a batch, and
    one that is small against a coarse mesh's spacing leaves warm values
    below the result there, which the acceptance check rejects)."""
    wex = ex if ex != lsb.LS_EXEC_KERNELS else lsb.LS_EXEC_GRAPH
    g = lsb.solver_for_batch(b, execution=wex, patch_verts=cand, precision=prec)
    g.mlp_set_params(lsb.mlp_init_params(5))
    prm = lsb.step_params(1.0, 1.0, 5)
    phi0 = np.concatenate([m.phi0 for m in b.meshes])
    g.step(prm)
    t0 = time.perf_counter()
    for _ in range(args.train_steps):
        g.mlp_collect(0, 12_000_000)
        g.step(prm)
        g.mlp_collect(1)
    t1 = time.perf_counter()
    loss = g.mlp_train(args.train_epochs, 4096, 2e-3, seed=7)
    t2 = time.perf_counter()
    abs_margin = WARM_ABS_H * max(m.h for m in b.meshes)

    def timed(warm):
        wprm = lsb.step_params(1.0, 1.0, 5, warm=warm, warm_margin=args.batch_warm_margin,
                               warm_abs_margin=abs_margin)
        g.set_phi(phi0)
        r = {"verts": 0, "fallbacks": 0, "iters": 0, "visits": 0}
        ts = time.perf_counter()
        for k in range(args.steps):
            if warm == lsb.LS_WARM_LEARNED and k > 0:
                g.mlp_predict()
            st = g.step(wprm if k > 0 else prm)
            r["verts"] += st.nb_verts
            r["fallbacks"] += st.warm_fallback
            r["iters"] += st.fim_iters
            r["visits"] += st.patch_visits
        r["wall"] = time.perf_counter() - ts
        return r

    def timed_perfect():
        # warm values from the plain step's own distances on the same input
        # (a second context, outside the timed region); see learned_run
        g2 = lsb.solver_for_batch(b, execution=wex, patch_verts=cand, precision=prec)
        wprm = lsb.step_params(1.0, 1.0, 5, warm=lsb.LS_WARM_LEARNED, warm_margin=0.0,
                               warm_abs_margin=PERFECT_ABS, warm_res_tol=PERFECT_ABS)
        g.set_phi(phi0)
        r = {"verts": 0, "fallbacks": 0, "iters": 0, "visits": 0, "wall": 0.0}
        for k in range(args.steps):
            if k > 0:
                g2.set_phi(g.phi())
                g2.step(prm)
                g.set_warm(g2.distances())
            ts = time.perf_counter()
            st = g.step(wprm if k > 0 else prm)
            r["wall"] += time.perf_counter() - ts
            r["verts"] += st.nb_verts
            r["fallbacks"] += st.warm_fallback
            r["iters"] += st.fim_iters
            r["visits"] += st.patch_visits
        g2.close()
        return r

    learned = timed(lsb.LS_WARM_LEARNED)
    ident = timed(lsb.LS_WARM_IDENTITY)
    perfect = timed_perfect()
    g.close()
    value = learned["verts"] / learned["wall"]
    return {"value": value, "vs_plain_fim": value / plain_value, "ms_per_step": 1e3 * learned["wall"] / args.steps,
            "fim_iters_per_step": learned["iters"] / args.steps,
            "patch_visits_per_step": learned["visits"] / args.steps,
            "fallback_mesh_steps": learned["fallbacks"], "mesh_steps": args.steps * len(b.meshes),
            "identity_predictor": {"value": ident["verts"] / ident["wall"],
                                   "vs_plain_fim": ident["verts"] / ident["wall"] / plain_value,
                                   "fim_iters_per_step": ident["iters"] / args.steps,
                                   "fallback_mesh_steps": ident["fallbacks"]},
            "perfect_predictor": {"what": "warm values = the plain method's own distances of the same step "
                                          f"from the same input + {PERFECT_ABS:g}, accepted up to an undercut "
                                          f"of {PERFECT_ABS:g} (the ceiling of any warm start on this scheme); "
                                          "only the warm step is timed, per step",
                                  "value": perfect["verts"] / perfect["wall"],
                                  "vs_plain_fim": perfect["verts"] / perfect["wall"] / plain_value,
                                  "fim_iters_per_step": perfect["iters"] / args.steps,
                                  "patch_visits_per_step": perfect["visits"] / args.steps,
                                  "fallback_mesh_steps": perfect["fallbacks"]},
            "collect_s": t1 - t0, "train_s": t2 - t1, "train_loss": loss, "warm_margin": args.batch_warm_margin,
            "warm_abs_margin": abs_margin}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed timesteps (default: the configured run, 100 for the torus, 50 for mcf and batch)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=8, help="timesteps of the CPU-baseline sample")
    ap.add_argument("--parity-steps", type=int, default=0,
                    help="timesteps of the trajectory compared with the FP64 oracle (0: the timed steps)")
    ap.add_argument("--precision", default="f64", choices=["f64", "f32"],
                    help="storage of the engine: f64 (default; bit-exact with the FP64 oracle) or f32")
    ap.add_argument("--tol", type=float, default=1e-6,
                    help="FIM tolerance of the torus workload (1e-6: the plain method's own result lies "
                         "within 1e-6 of the scheme's fixed point inside the band)")
    ap.add_argument("--no-fp32-line", action="store_true", help="skip the other storage precision's line")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--exec", default="auto", choices=["auto", "graph", "persistent", "kernels"],
                    help="auto (default): the fastest of graph (persistent step kernel replayed from a "
                         "hipGraph), persistent and kernels over a few trial steps")
    ap.add_argument("--batch-patch", default="auto",
                    help="batch patch size in vertices (with --exec other than auto: no autotuning), or auto")
    ap.add_argument("--mcf-block", default="auto",
                    help="MCF patch block LxWxH (e.g. 8x8x4), or auto: timed over 4x4x4..8x8x8")
    ap.add_argument("--sweep", default="auto", choices=["auto", "elements", "tasks"],
                    help="FIM sweep organisation (identical results)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned warm-start run")
    ap.add_argument("--train-steps", type=int, default=13, help="plain steps whose pairs train the network")
    ap.add_argument("--train-epochs", type=int, default=2)
    ap.add_argument("--batch-warm-margin", type=float, default=WARM_MARGIN,
                    help="relative warm-start margin of the batch workload's learned run")
    ap.add_argument("--learned-parity-steps", type=int, default=3,
                    help="learned warm steps compared with the FP64 oracle's plain step from the same input")
    ap.add_argument("--workload", default="torus", choices=["torus", "batch", "mcf"],
                    help="torus: 3D torus revolve, 64^3 (default); batch: 512 ellipsoid meshes; "
                         "mcf: 8-sphere mean-curvature flow, 192^3")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device (a reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 100 if args.workload == "torus" else 50
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU path is not built.",
              file=sys.stderr)
        sys.exit(2)

    ge.load_package()
    import levelset_amd.cases as cases
    import levelset_amd.levelset as lsb

    if args.workload != "torus":
        fn = batch_workload if args.workload == "batch" else mcf_workload
        cfg, tot, kmode, cpu = fn(args, lsb, cases)
        out = {"metric": "narrowband vertex-updates/s, learned warm-start vs plain FIM @ Linf<=1e-5",
               "value": tot["verts"] / tot["wall"], "unit": "vertex-updates/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": 1e3 * tot["wall"] / args.steps, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
               "config": dict(cfg, mode="plain_fim", gpu_event_ms_per_step=tot["ms_total"] / args.steps,
                              fim_iters_per_step=tot["fim_iters"] / args.steps,
                              patch_visits_per_step=tot["visits"] / args.steps,
                              exec_modes_used=sorted(tot["exec_modes"]),
                              step_wg_per_cu=sorted(tot["step_wg_per_cu"])),
               "roofline": roofline(tot, args.steps, kmode)}
        out["roofline"]["compute"] = compute_fraction(tot, args.steps, kmode)
        out["roofline"]["limiter"] = ("FP64 local-solve chains of the FIM sweeps at full occupancy (every CU "
                                      "holds patches); neither HBM nor the FP64 VALUs are saturated")
        if not kmode:
            prof_name = ("k_step_batch_512" if args.workload == "batch" else "k_step_mcf_spheres_192") + \
                ("_f64" if args.precision == "f64" else "")
            traffic = profiled_traffic(prof_name)
            if traffic is not None:
                out["roofline"]["traffic"] = traffic
                out["roofline"]["traffic_source"] = (f"profiles/{prof_name}.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                                     "copy-calibrated)")
        if cpu:
            out["cpu_baseline"], out["parity"] = cpu
        print(json.dumps(out))
        return

    case = cases.config2(steps=args.steps)
    names = {"graph": lsb.LS_EXEC_GRAPH, "persistent": lsb.LS_EXEC_PERSISTENT, "kernels": lsb.LS_EXEC_KERNELS}
    precisions = {"f64": lsb.LS_F64, "f32": lsb.LS_F32}
    prec = precisions[args.precision]
    prm_kw = {"tol": args.tol}
    tune = None
    if args.exec == "auto":
        tune = autotune_exec(lsb, case, names, precision=prec, prm_kw=prm_kw)
        args.exec = min(tune, key=tune.get)
    exec_mode = names[args.exec]
    sweep = {"auto": lsb.LS_SWEEP_AUTO, "elements": lsb.LS_SWEEP_ELEMENTS, "tasks": lsb.LS_SWEEP_TASKS}[args.sweep]
    prm = lsb.step_params(case.dt, case.bandwidth, case.niter, **prm_kw)

    def timed_torus(precision):
        g = lsb.solver_for_case(case, precision=precision, execution=exec_mode)
        g.set_sweep(sweep)
        for _ in range(args.warmup):
            g.step(prm)
        g.set_phi(case.phi0)  # the timed region is the configured run from phi0
        tot = timed_run(g, prm, args.steps, w=8 if precision == lsb.LS_F64 else 4)
        info = g.layout_info()
        g.close()
        return tot, info

    tot, info = timed_torus(prec)
    value = tot["verts"] / tot["wall"]
    kmode = exec_mode == lsb.LS_EXEC_KERNELS
    roof = roofline(tot, args.steps, kmode)
    prof_name = ("k_sweep_torus_revolve_64" if kmode else
                 "k_step_torus_revolve_64_f64" if prec == lsb.LS_F64 else "k_step_torus_revolve_64_graph")
    traffic = profiled_traffic(prof_name)
    roof["traffic"] = traffic
    roof["traffic_source"] = (f"profiles/{prof_name}.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)"
                              if traffic else None)
    roof["compute"] = compute_fraction(tot, args.steps, kmode)
    roof["limiter"] = (("latency: each FIM launch lasts its slowest patch visit (dependent FP64 solve "
                        "chains, four workgroup barriers per sweep), 14-260 active patches for 256 CUs; "
                        "neither HBM nor the FP64 VALUs are saturated") if kmode else
                       ("latency: each FIM iteration lasts its slowest patch visit or probe-then-visit "
                        "(dependent FP64 solve chains, four workgroup barriers per sweep), one grid barrier "
                        "per iteration, 14-260 active patches for 256 CUs; neither HBM nor the FP64 VALUs "
                        "are saturated"))

    out = {
        "metric": "narrowband vertex-updates/s, learned warm-start vs plain FIM @ Linf<=1e-5",
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tot["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": f"3d_torus_revolve_64: 64^3 Kuhn tets (262144 verts, 1500282 tets), "
                        f"256-vertex patches, band 6h, CFL 0.5, rigid rotation, {args.steps} timed timesteps "
                        f"from phi0 (the configured run is 100)",
            "mode": "plain_fim",
            "storage": args.precision,
            "fim_tol": args.tol,
            "execution": args.exec,
            "execution_trial_ms_per_step": tune,
            "sweep": "tasks" if tot.get("sweep_tasks") else "elements",
            "step_wg_per_cu": sorted(tot["step_wg_per_cu"]),
            "patches": info["n_patches"],
            "element_copies": info["n_copies"],
            "fim_iters_per_step": tot["fim_iters"] / args.steps,
            "patch_visits_per_step": tot["visits"] / args.steps,
            "probes_per_step": tot["probes"] / args.steps,
            "local_solves_per_step": tot["solves"] / args.steps,
            "gpu_event_ms_per_step": tot["ms_total"] / args.steps,
        },
        "roofline": roof,
    }

    # The other storage precision as its own labelled line (FP32 storage,
    # FP64 element arithmetic): faster, but the reference's hard advection
    # thresholds flip under FP32 rounding, so its trajectory leaves the bar.
    other = "f32" if args.precision == "f64" else "f64"
    if not args.no_fp32_line:
        tot2, _ = timed_torus(precisions[other])
        out["other_storage"] = {"storage": other, "value": tot2["verts"] / tot2["wall"],
                                "ms_per_step": 1e3 * tot2["wall"] / args.steps,
                                "fim_iters_per_step": tot2["fim_iters"] / args.steps}

    O = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
    if not args.no_learned:
        out["learned"] = learned_run(args, lsb, case, exec_mode, value, prec, prm_kw, O)

    if not args.no_cpu:
        engines = [(args.precision, prec, exec_mode)]
        if "other_storage" in out:
            engines.append((other, precisions[other], exec_mode))
        nsteps = args.parity_steps if args.parity_steps > 0 else args.steps
        res, cpu = trajectory(case, O, lsb, nsteps, engines, prm_kw)
        out["cpu_baseline"] = {
            "value": cpu["cpu_verts"] / cpu["cpu_s"],
            "unit": "vertex-updates/s",
            "cores": O.max_threads(),
            "kind": "port",
            "sample": f"first {cpu['steps']} timesteps of the same workload from phi0, the oracle's FP64 "
                      f"instance (the scheme of the headline engine), OpenMP over patches",
        }
        head = res[args.precision]
        out["parity"] = {
            "oracle": "FP64 host build of the same scheme, same schedule",
            "trajectory_steps": head["steps"],
            "linf_all_finite_vertices": head["linf"],
            "linf_in_band": head["linf_in_band"],
            "bit_mismatches": head["bit_mismatches"],
            "finite_set_mismatches": head["finite_mismatch"],
            "nb_membership_diff": head["nb_diff"],
            "tolerance": 1e-5,
            "meets_bar": head["linf"] <= 1e-5 and head["nb_diff"] == 0 and head["finite_mismatch"] == 0,
        }
        if other in res:
            r = res[other]
            out["other_storage"]["parity"] = {
                "trajectory_steps": r["steps"], "linf_all_finite_vertices": r["linf"],
                "linf_in_band": r["linf_in_band"], "nb_membership_diff": r["nb_diff"],
                "finite_set_mismatches": r["finite_mismatch"],
                "linf_per_step": [float(f"{x:.3e}") for x in r["linf_per_step"]],
                "meets_bar": r["linf"] <= 1e-5 and r["nb_diff"] == 0 and r["finite_mismatch"] == 0}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
