#!/usr/bin/env python3
"""Benchmark: narrowband vertex-updates/s of the level-set timestep on one MI355X.

Workload: the 3D torus revolve -- a 64^3 Kuhn tet mesh (262,144 vertices,
1,500,282 tets), jitter 0.1h, torus (R=0.3, r=0.1), rigid rotation about z
(omega = 2 pi), CFL 0.5, band 6h, 256-vertex patches, 100 timesteps. A
vertex-update is one owned vertex of a narrowband patch receiving its final
phi for a step (the set CopyOutBack_levelset3d writes,
src/core3d/include/meshFIM_kernels.h:22-34). One bench "step" is one full
timestep (seeds, sign save, FIM, narrowband + sign restore, advection).

Besides the JSON line's throughput it reports:
  * roofline of the dominant kernel (the FIM patch sweep), from HIP-event
    durations of its launches and the algorithmic bytes of the patches they
    visited;
  * the CPU baseline: the oracle's FP32 instance (same scheme, OpenMP over
    patches, all host threads) on a bounded sample of the same run;
  * parity: the GPU against the FP32 oracle (bit-exact) and the FP64 oracle
    (L-inf on phi, narrowband membership diff) on the first timesteps.
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


def step_bytes(st):
    """Algorithmic bytes of a whole timestep: seeds (16E+4(V+H)+9V per scanned
    patch), FIM visits and probes, narrowband + sign save/restore (4N+4P+10N),
    advection."""
    seeds = 16 * st.seed_copies + 4 * (st.seed_verts + st.seed_halo) + 9 * st.seed_verts
    return seeds + fim_bytes(st) + adv_bytes(st)


def fim_bytes(st):
    """Algorithmic bytes of FIM patch visits and probes: per visit
    40*E + 4*(V+H) + 5*V (16 B connectivity + 24 B local coordinates per
    element copy, distances in, distances out + 1 B flag); probes do not write
    distances: 40*E + 4*(V+H) + V."""
    visits = 40 * st.visit_copies + 4 * (st.visit_verts + st.visit_halo) + 5 * st.visit_verts
    probes = 40 * st.probe_copies + 4 * (st.probe_verts + st.probe_halo) + 1 * st.probe_verts
    return visits + probes


def adv_bytes(st):
    """Advection per narrowband patch: 28*E + 16*(V+H) + 4*V."""
    return 28 * st.adv_copies + 16 * (st.adv_verts + st.adv_halo) + 4 * st.adv_verts


def fp64_error(gphi, ophi, amb, band):
    """L-inf of the GPU's phi against the FP64 oracle's over finite values,
    without and with the vertices whose advection update hinges on a
    threshold decision inside the FP32-storage window (counted), and over the
    vertices inside the band (|phi| <= band; beyond it a patch's FIM stops
    when a sweep changes its values by less than the tolerance, a test that
    FP32 and FP64 storage can decide a sweep apart)."""
    fin = (np.abs(ophi) < 1e8) & (np.abs(gphi) < 1e8)
    d = np.abs(gphi - ophi)
    ok = fin & ~amb
    inb = ok & (np.abs(ophi) <= band)
    return {"linf_vs_fp64": float(np.max(d[ok])) if ok.any() else 0.0,
            "linf_vs_fp64_in_band": float(np.max(d[inb])) if inb.any() else 0.0,
            "linf_vs_fp64_incl_threshold_ambiguous": float(np.max(d[fin])) if fin.any() else 0.0,
            "threshold_ambiguous_verts": int(np.count_nonzero(fin & amb))}


def lockstep(case, oracle_mod, lsb, precision, steps):
    """FP32 oracle: runs beside the GPU from phi0 (bit-exact check, timed as the
    CPU baseline). FP64 oracle: restarted every step from the GPU's phi and
    narrowband, so each step is compared from identical inputs; vertices whose
    advection hinges on a threshold decision inside the FP32-storage window are
    reported separately."""
    g = lsb.solver_for_case(case)
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts, precision)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    gp = lsb.step_params(case.dt, case.bandwidth, case.niter)
    op = oracle_mod.make_params(case.dt, case.bandwidth, case.niter)
    res = {"steps": 0, "linf": 0.0, "linf_band": 0.0, "linf_all": 0.0, "ambiguous": 0, "nb_diff": 0, "bit_mismatch": 0,
           "cpu_s": 0.0, "cpu_verts": 0}
    gphi, gnb = case.phi0, None
    for _ in range(steps):
        if precision == 64:
            o.set_phi(gphi)
            if gnb is not None:
                o.set_narrowband(gnb)
        g.step(gp)
        t0 = time.perf_counter()
        ost = o.step(op)
        res["cpu_s"] += time.perf_counter() - t0
        res["cpu_verts"] += ost.nb_verts
        gphi, ophi = g.phi(), o.phi()
        gnb, onb = g.narrowband(), o.narrowband()
        res["nb_diff"] += int(np.setxor1d(gnb, onb).size)
        fin = (np.abs(ophi) < 1e8) & (np.abs(gphi) < 1e8)
        d = np.abs(gphi - ophi)
        res["linf_all"] = max(res["linf_all"], float(np.max(d[fin])) if fin.any() else 0.0)
        if precision == 64:
            amb = o.ambiguous()
            res["ambiguous"] = max(res["ambiguous"], int(amb.sum()))
            ok = fin & ~amb
            res["linf"] = max(res["linf"], float(np.max(d[ok])) if ok.any() else 0.0)
            inb = ok & (np.abs(ophi) <= case.bandwidth)
            res["linf_band"] = max(res["linf_band"], float(np.max(d[inb])) if inb.any() else 0.0)
        res["bit_mismatch"] += int(np.count_nonzero(gphi.astype(np.float32) != ophi.astype(np.float32)))
        res["steps"] += 1
        if ost.nb_count == 0:
            break
    g.close()
    o.close()
    return res


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


def timed_run(g, prm, steps):
    """Times `steps` timesteps; sums the per-step statistics."""
    tot = {"verts": 0, "sweep_ms": 0.0, "sweep_launches": 0, "sweep_bytes": 0, "adv_ms": 0.0,
           "adv_bytes": 0, "adv_launches": 0, "ms_fim": 0.0, "ms_total": 0.0, "fim_iters": 0,
           "visits": 0, "probes": 0, "step_ms": 0.0, "step_bytes": 0, "exec_modes": set(),
           "step_wg_per_cu": set()}
    t0 = time.perf_counter()
    for _ in range(steps):
        st = g.step(prm)
        tot["verts"] += st.nb_verts
        tot["sweep_ms"] += st.ms_sweep_kernels
        tot["sweep_launches"] += st.sweep_launches
        tot["sweep_bytes"] += fim_bytes(st)
        tot["adv_ms"] += st.ms_advect_kernels
        tot["adv_bytes"] += adv_bytes(st)
        tot["adv_launches"] += 1 if st.nb_count > 0 else 0
        tot["ms_fim"] += st.ms_fim
        tot["ms_total"] += st.ms_total
        tot["fim_iters"] += st.fim_iters
        tot["visits"] += st.patch_visits
        tot["probes"] += st.probes
        tot["step_ms"] += st.ms_step_kernel
        tot["step_bytes"] += step_bytes(st)
        tot["exec_modes"].add(st.exec_mode)
        tot["step_wg_per_cu"].add(st.step_wg_per_cu)
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


def batch_workload(args, lsb, cases):
    """512 uneven ellipsoid meshes in one launch per timestep (hipGraph), the
    patch size and execution mode chosen by the autotuner on a 64-mesh
    subset. CPU baseline and parity on a sample of the meshes."""
    b = cases.config4(steps=args.steps)
    print(f"batch: {len(b.meshes)} meshes, {b.n_verts} vertices, {b.n_cells} tets", file=sys.stderr, flush=True)
    if args.batch_patch != "auto" and args.exec != "auto":
        names = {"graph": lsb.LS_EXEC_GRAPH, "persistent": lsb.LS_EXEC_PERSISTENT, "kernels": lsb.LS_EXEC_KERNELS}
        cand, ex, table = int(args.batch_patch), names[args.exec], []
    else:
        sub = cases.Batch(b.name + "_tune", 3, b.meshes[:64], patch_verts=b.patch_verts)
        cand, ex, ms, table = lsb.autotune_batch(sub, [64, 128, 256, 512], trial_steps=2,
                                                 log=lambda m: print(m, file=sys.stderr, flush=True))
    t0 = time.perf_counter()
    g = lsb.solver_for_batch(b, execution=ex, patch_verts=cand)
    setup_s = time.perf_counter() - t0
    info = g.layout_info()
    prm = lsb.step_params(1.0, 1.0, 5)
    for _ in range(args.warmup):
        g.step(prm)
    g.set_phi(np.concatenate([m.phi0 for m in b.meshes]))
    tot = timed_run(g, prm, args.steps)
    g.close()
    out = {"workload": f"{b.name}: {len(b.meshes)} Kuhn meshes, {b.n_verts} verts, {b.n_cells} tets, "
                       f"ellipsoids, affine velocity, CFL 0.5, band 6h, {args.steps} timesteps",
           "patch_verts": cand, "execution": {lsb.LS_EXEC_GRAPH: "graph", lsb.LS_EXEC_PERSISTENT: "persistent",
                                              lsb.LS_EXEC_KERNELS: "kernels"}[ex],
           "autotune_ms_per_step": {f"{c}/{e}": m for c, e, m in table} or None,
           "patches": info["n_patches"], "element_copies": info["n_copies"],
           "setup_s": setup_s}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        sample = cases.Batch(b.name + "_sample", 3, b.meshes[:24], patch_verts=cand)
        gs = lsb.solver_for_batch(sample, execution=ex, patch_verts=cand)
        orc = []
        for m in sample.meshes:
            nparts = -(-m.xyz.shape[0] // cand)
            o = O.Oracle(3, m.xyz, m.cells, cases.partition_rcb(m.xyz, nparts), nparts, 32)
            o.set_phi(m.phi0)
            o.set_velocity(m.vel)
            orc.append(o)
        cpu_s, cpu_v, mism, nbd = 0.0, 0, 0, 0
        for _ in range(args.cpu_steps):
            gs.step(prm)
            counts = gs.batch_nb_counts()
            for k, (o, phi) in enumerate(zip(orc, gs.split(gs.phi()))):
                t0 = time.perf_counter()
                ost = o.step(O.make_params(sample.meshes[k].dt, sample.meshes[k].bandwidth, 5))
                cpu_s += time.perf_counter() - t0
                cpu_v += ost.nb_verts
                mism += int(np.count_nonzero(phi.astype(np.float32) != o.phi().astype(np.float32)))
                nbd += abs(int(counts[k]) - ost.nb_count)
        for o in orc:
            o.close()
        # One more step of every sample mesh checked against the FP64 oracle
        # from the GPU's state (the oracle scans every patch for seeds).
        o64s = []
        for m, phi in zip(sample.meshes, gs.split(gs.phi())):
            nparts = -(-m.xyz.shape[0] // cand)
            o = O.Oracle(3, m.xyz, m.cells, cases.partition_rcb(m.xyz, nparts), nparts, 64)
            o.set_velocity(m.vel)
            o.set_phi(phi)
            o64s.append(o)
        gs.step(prm)
        l64 = {"linf_vs_fp64": 0.0, "linf_vs_fp64_in_band": 0.0, "linf_vs_fp64_incl_threshold_ambiguous": 0.0,
               "threshold_ambiguous_verts": 0}
        for k, (o, phi) in enumerate(zip(o64s, gs.split(gs.phi()))):
            o.step(O.make_params(sample.meshes[k].dt, sample.meshes[k].bandwidth, 5))
            e = fp64_error(phi, o.phi(), o.ambiguous(), sample.meshes[k].bandwidth)
            l64 = {key: max(l64[key], e[key]) if key != "threshold_ambiguous_verts" else l64[key] + e[key]
                   for key in l64}
            o.close()
        gs.close()
        cpu = ({"value": cpu_v / cpu_s, "unit": "vertex-updates/s", "cores": O.max_threads(), "kind": "port",
                "sample": f"first {args.cpu_steps} timesteps of the first 24 meshes, oracle FP32 instance, "
                          "OpenMP over patches"},
               dict({"fp32_oracle_meshes": 24, "fp32_oracle_steps": args.cpu_steps, "fp32_bit_mismatches": mism,
                     "nb_count_diff": nbd, "fp64_oracle_steps": 1}, **l64, tolerance=1e-5))
    return out, tot, ex == lsb.LS_EXEC_KERNELS, cpu


def mcf_workload(args, lsb, cases):
    """Mean-curvature flow of 8 spheres on a 192^3 Kuhn grid (41.8M tets)."""
    case = cases.config3(steps=args.steps)
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
                gt = lsb.solver_for_case(case, execution=lsb.LS_EXEC_PERSISTENT)
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
    g = lsb.solver_for_case(case)
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
    tot = timed_run(g, prm, args.steps)
    g.close()
    out = {"workload": f"{case.name}: 192^3 Kuhn tets ({case.xyz.shape[0]} verts, {case.cells.shape[0]} tets), "
                       f"8 spheres, mean-curvature flow, dt = h^2/8, band 8h, {args.steps} timesteps",
           "execution": args.exec, "execution_trial_ms_per_step": tune, "patch_block": "x".join(map(str, block)),
           "patch_trial_ms_per_step": patch_tune or None, "patches": info["n_patches"],
           "element_copies": info["n_copies"], "setup_s": setup_s}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        g = lsb.solver_for_case(case, execution=ex)
        o = O.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 32)
        o.set_phi(case.phi0)
        o.set_velocity(case.vel)
        cpu_s, cpu_v, mism = 0.0, 0, 0
        for _ in range(args.cpu_steps):
            g.step(prm)
            t0 = time.perf_counter()
            ost = o.step(O.make_params(case.dt, case.bandwidth, case.niter, mode=1))
            cpu_s += time.perf_counter() - t0
            cpu_v += ost.nb_verts
            mism += int(np.count_nonzero(g.phi().astype(np.float32) != o.phi().astype(np.float32)))
        o.close()
        # One more step checked against the FP64 oracle from the GPU's state.
        o64 = O.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
        o64.set_velocity(case.vel)
        o64.set_phi(g.phi())
        o64.set_narrowband(g.narrowband())
        g.step(prm)
        o64.step(O.make_params(case.dt, case.bandwidth, case.niter, mode=1))
        l64 = fp64_error(g.phi(), o64.phi(), o64.ambiguous(), case.bandwidth)
        o64.close()
        g.close()
        cpu = ({"value": cpu_v / cpu_s, "unit": "vertex-updates/s", "cores": O.max_threads(), "kind": "port",
                "sample": f"first {args.cpu_steps} timesteps of the same workload, oracle FP32 instance"},
               dict({"fp32_oracle_steps": args.cpu_steps, "fp32_bit_mismatches": mism, "fp64_oracle_steps": 1},
                    **l64, tolerance=1e-5))
    return out, tot, ex == lsb.LS_EXEC_KERNELS, cpu


def autotune_exec(lsb, case, names, steps=6, g=None, prm=None):
    """Mean GPU ms/step of each execution mode over a few steps from phi0
    (after one untimed step) on one solver; the bench runs the fastest."""
    own = g is None
    if own:
        g = lsb.solver_for_case(case)
    prm = prm or lsb.step_params(case.dt, case.bandwidth, case.niter, mode=case.mode)
    out = {}
    for name, mode in names.items():
        g.set_execution(mode)
        g.set_phi(case.phi0)
        g.step(prm)
        out[name] = sum(g.step(prm).ms_total for _ in range(steps)) / steps
    if own:
        g.close()
    return out


def learned_run(args, lsb, case, exec_mode, plain_value):
    """The learned warm start on the same workload: pairs collected from
    plain steps, GPU training, then the timed steps with warm values from
    the network (prediction time included). A warm start whose fixed-point
    residual over the band exceeds 1e-5 re-runs the plain method."""
    # warm starts run in the persistent step kernel (here replayed from a graph)
    g = lsb.solver_for_case(case, execution=lsb.LS_EXEC_GRAPH)
    g.mlp_set_params(lsb.mlp_init_params(5))
    prm = lsb.step_params(case.dt, case.bandwidth, case.niter)
    g.step(prm)
    t0 = time.perf_counter()
    for _ in range(args.train_steps):
        g.mlp_collect(0, 12_000_000)
        g.step(prm)
        g.mlp_collect(1)
    X, Y = g.mlp_dataset()
    t1 = time.perf_counter()
    loss = g.mlp_train(args.train_epochs, 4096, 2e-3, seed=7)
    t2 = time.perf_counter()
    held = slice(0, min(200_000, X.shape[0]))
    _, sse = g.mlp_gradient(X[held], Y[held])  # the network's squared error on the GPU
    trained_mse = sse / Y[held].size
    identity_mse = float(np.mean(Y[held].astype(np.float64) ** 2))
    wprm = lsb.step_params(case.dt, case.bandwidth, case.niter, warm=lsb.LS_WARM_LEARNED, warm_margin=2e-3)
    g.set_phi(case.phi0)
    verts, fallbacks, iters = 0, 0, 0
    ts = time.perf_counter()
    for k in range(args.steps):
        if k > 0:
            g.mlp_predict()
        st = g.step(wprm)
        verts += st.nb_verts
        fallbacks += st.warm_fallback
        iters += st.fim_iters
    wall = time.perf_counter() - ts
    g.close()
    value = verts / wall
    return {"value": value, "unit": "vertex-updates/s", "vs_plain_fim": value / plain_value, "execution": "graph",
            "ms_per_step": 1e3 * wall / args.steps, "fim_iters_per_step": iters / args.steps,
            "fallback_steps": fallbacks, "steps": args.steps, "train_pairs": int(X.shape[0]),
            "collect_s": t1 - t0, "train_s": t2 - t1, "train_epochs": args.train_epochs,
            "train_loss": loss, "trained_mse": trained_mse, "identity_mse": identity_mse,
            "residual_tolerance": 1e-5, "warm_margin": 2e-3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-steps", type=int, default=8, help="timesteps of the CPU-baseline sample")
    ap.add_argument("--parity-steps", type=int, default=4, help="timesteps checked against the FP64 oracle")
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
    ap.add_argument("--workload", default="torus", choices=["torus", "batch", "mcf"],
                    help="torus: 3D torus revolve, 64^3 (default); batch: 512 ellipsoid meshes; "
                         "mcf: 8-sphere mean-curvature flow, 192^3")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device (a reference kernel for HBM byte counters)")
    args = ap.parse_args()
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
               "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": dict(cfg, mode="plain_fim", gpu_event_ms_per_step=tot["ms_total"] / args.steps,
                              fim_iters_per_step=tot["fim_iters"] / args.steps,
                              exec_modes_used=sorted(tot["exec_modes"]),
                              step_wg_per_cu=sorted(tot["step_wg_per_cu"])),
               "roofline": roofline(tot, args.steps, kmode)}
        if not kmode:
            prof_name = "k_step_batch_512" if args.workload == "batch" else "k_step_mcf_spheres_192"
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
    tune = None
    if args.exec == "auto":
        tune = autotune_exec(lsb, case, names)
        args.exec = min(tune, key=tune.get)
    exec_mode = names[args.exec]
    sweep = {"auto": lsb.LS_SWEEP_AUTO, "elements": lsb.LS_SWEEP_ELEMENTS, "tasks": lsb.LS_SWEEP_TASKS}[args.sweep]
    g = lsb.solver_for_case(case, execution=exec_mode)
    g.set_sweep(sweep)
    info = g.layout_info()
    prm = lsb.step_params(case.dt, case.bandwidth, case.niter)
    for _ in range(args.warmup):
        g.step(prm)
    g.set_phi(case.phi0)  # the timed region is the configured run from phi0

    tot = {"verts": 0, "sweep_ms": 0.0, "sweep_launches": 0, "sweep_bytes": 0, "adv_ms": 0.0,
           "adv_bytes": 0, "adv_launches": 0, "ms_fim": 0.0, "ms_total": 0.0, "fim_iters": 0,
           "visits": 0, "probes": 0, "step_ms": 0.0, "step_bytes": 0}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = g.step(prm)
        tot["verts"] += st.nb_verts
        tot["sweep_ms"] += st.ms_sweep_kernels
        tot["sweep_launches"] += st.sweep_launches
        tot["sweep_bytes"] += fim_bytes(st)
        tot["adv_ms"] += st.ms_advect_kernels
        tot["adv_bytes"] += adv_bytes(st)
        tot["adv_launches"] += 1 if st.nb_count > 0 else 0
        tot["ms_fim"] += st.ms_fim
        tot["ms_total"] += st.ms_total
        tot["fim_iters"] += st.fim_iters
        tot["visits"] += st.patch_visits
        tot["probes"] += st.probes
        tot["step_ms"] += st.ms_step_kernel
        tot["step_bytes"] += step_bytes(st)
        tot["sweep_tasks"] = st.sweep_tasks
        tot["step_wg_per_cu"] = st.step_wg_per_cu
    wall = time.perf_counter() - t0
    g.close()

    value = tot["verts"] / wall
    if exec_mode == lsb.LS_EXEC_KERNELS:
        kname = "k_sweep (FIM patch sweep)"
        launches = tot["sweep_launches"]
        avg_ms = tot["sweep_ms"] / max(launches, 1)
        avg_bytes = tot["sweep_bytes"] / max(launches, 1)
    else:
        kname = "k_step + k_advect_dev + k_copy_phi_dev (one timestep: persistent seeds/FIM/narrowband, then advection)"
        launches = args.steps
        avg_ms = tot["step_ms"] / args.steps
        avg_bytes = tot["step_bytes"] / args.steps
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    prof_name = "k_sweep_torus_revolve_64" if exec_mode == lsb.LS_EXEC_KERNELS else "k_step_torus_revolve_64_graph"
    traffic = profiled_traffic(prof_name)

    out = {
        "metric": "narrowband vertex-updates/s, learned warm-start vs plain FIM @ Linf<=1e-5",
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3d_torus_revolve_64: 64^3 Kuhn tets (262144 verts, 1500282 tets), "
                        "256-vertex patches, band 6h, CFL 0.5, rigid rotation, 100 timesteps",
            "mode": "plain_fim",
            "execution": args.exec,
            "execution_trial_ms_per_step": tune,
            "sweep": "tasks" if tot.get("sweep_tasks") else "elements",
            "step_wg_per_cu": tot.get("step_wg_per_cu", 0),
            "patches": info["n_patches"],
            "element_copies": info["n_copies"],
            "fim_iters_per_step": tot["fim_iters"] / args.steps,
            "patch_visits_per_step": tot["visits"] / args.steps,
            "probes_per_step": tot["probes"] / args.steps,
            "gpu_event_ms_per_step": tot["ms_total"] / args.steps,
            "fim_ms_per_step": tot["ms_fim"] / args.steps,
        },
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": f"profiles/{prof_name}.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                              "copy-calibrated)" if traffic else None,
            "avg_launch_ms": avg_ms,
            "avg_launch_algorithmic_bytes": avg_bytes,
            "launches": launches,
        },
    }

    if not args.no_learned:
        out["learned"] = learned_run(args, lsb, case, exec_mode, value)

    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        fp32 = lockstep(case, O, lsb, 32, args.cpu_steps)
        fp64 = lockstep(case, O, lsb, 64, args.parity_steps)
        out["cpu_baseline"] = {
            "value": fp32["cpu_verts"] / fp32["cpu_s"],
            "unit": "vertex-updates/s",
            "cores": O.max_threads(),
            "kind": "port",
            "sample": f"first {fp32['steps']} of the {args.steps} timesteps of the same workload, "
                      "oracle FP32 instance, OpenMP over patches",
        }
        out["parity"] = {
            "fp32_oracle_steps": fp32["steps"],
            "fp32_bit_mismatches": fp32["bit_mismatch"],
            "fp32_nb_membership_diff": fp32["nb_diff"],
            "fp64_oracle_steps": fp64["steps"],
            "linf_vs_fp64": fp64["linf"],
            "linf_vs_fp64_in_band": fp64["linf_band"],
            "linf_vs_fp64_incl_threshold_ambiguous": fp64["linf_all"],
            "threshold_ambiguous_verts_max_per_step": fp64["ambiguous"],
            "fp64_nb_membership_diff": fp64["nb_diff"],
            "tolerance": 1e-5,
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
