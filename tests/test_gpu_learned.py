"""Learned warm start on the GPU: the MFMA gradient against the network
oracle, the training pairs against features and distances computed on the
host, training, the warm-started FIM against the oracle, and an ablation of
what the trained network gains over the identity predictor and plain FIM."""
import numpy as np
import pytest

import levelset_amd.cases as cases
import levelset_amd.levelset as lsb

pytestmark = pytest.mark.gpu


def _case(n=24, steps=6):
    return cases.small3d(n=n, shape="torus", block=(8, 8, 4), steps=steps)


def _params(M, seed=5):
    p = M.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed)
    p[22 * 64:22 * 64 + 64] = rng.normal(0, 0.05, 64)  # b1
    p[-4:] = rng.normal(0, 0.05, 4)                     # b3
    return p.astype(np.float32)


def test_mlp_gradient_matches_oracle(mlp_oracle):
    M = mlp_oracle
    case = _case(n=12)
    s = lsb.solver_for_case(case)
    p = _params(M)
    s.mlp_set_params(p)
    np.testing.assert_array_equal(s.mlp_params(), p)
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1, (333, 22)).astype(np.float32)
    Y = rng.normal(0, 0.3, (333, 4)).astype(np.float32)
    g, loss = s.mlp_gradient(X, Y)
    s.close()
    ref_loss, ref = M.loss_and_grad(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64), bf16_ops=True)
    print(f"loss gpu {loss:.6f} oracle {ref_loss:.6f}")
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    o = 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        a, b = g[o:o + k], ref[o:o + k]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        print(f"{name}: |g_gpu - g_oracle| / |g_oracle| = {rel:.2e}")
        assert rel <= 2e-2, name
        o += k
    # and against central differences of the exact FP64 loss (BF16 operands
    # account for the difference)
    idx = np.random.default_rng(0).choice(M.N_PARAMS, 60, replace=False)
    fd = M.finite_difference(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64), idx, h=1e-4)
    assert np.linalg.norm(g[idx] - fd) <= 5e-2 * np.linalg.norm(fd)


def test_collected_pairs_match_host_features(mlp_oracle):
    M = mlp_oracle
    case = _case(n=16)
    s = lsb.solver_for_case(case)
    prm = lsb.step_params(case.dt, case.bandwidth, case.niter)
    s.step(prm)
    nb = s.narrowband()
    phi_before = s.phi()
    s.mlp_collect(0, 100000)
    s.step(prm)
    s.mlp_collect(1)
    X, Y = s.mlp_dataset()
    dist = s.distances()
    lay = s.layout()
    s.close()
    eo = lay["ele_off"]
    copies = np.concatenate([np.arange(eo[p], eo[p + 1]) for p in nb])
    assert X.shape[0] == copies.size
    tets = case.cells[lay["copy_src"][copies]]
    L = M.local_coords(case.xyz, tets)
    want, sc = M.features(phi_before[tets].astype(np.float32), L)
    np.testing.assert_allclose(X, want, rtol=2e-5, atol=2e-6)
    D = dist[tets]
    wy = np.where(D < 5e8, D / sc[:, None] - np.abs(want[:, :4]), 0.0)
    np.testing.assert_allclose(Y, wy, rtol=1e-4, atol=1e-4)


def _train(case, M, steps=5, epochs=12, seed=5, precision=lsb.LS_F32):
    s = lsb.solver_for_case(case, precision=precision)
    s.mlp_set_params(_params(M, seed))
    prm = lsb.step_params(case.dt, case.bandwidth, case.niter)
    s.step(prm)
    for _ in range(steps):
        s.mlp_collect(0, 4_000_000)
        s.step(prm)
        s.mlp_collect(1)
    X, Y = s.mlp_dataset()
    loss = s.mlp_train(epochs, 512, 2e-3, seed=7)
    return s, X, Y, loss


def test_training_beats_identity_predictor_on_held_out_steps(mlp_oracle):
    """Trained on the pairs of five timesteps, scored on the pairs of the
    two timesteps that follow (never seen in training)."""
    M = mlp_oracle
    case = _case(n=24, steps=8)
    s, X, Y, loss = _train(case, M)
    p = s.mlp_params()
    prm = lsb.step_params(case.dt, case.bandwidth, case.niter)
    n_train = X.shape[0]
    for _ in range(2):
        s.mlp_collect(0, 4_000_000)
        s.step(prm)
        s.mlp_collect(1)
    Xa, Ya = s.mlp_dataset()
    s.close()
    Xh, Yh = Xa[n_train:], Ya[n_train:]
    assert Xh.shape[0] > 0
    identity = float(np.mean(Yh.astype(np.float64) ** 2))  # the identity predictor outputs 0
    O, _ = M.forward(p.astype(np.float64), Xh.astype(np.float64), bf16_ops=True)
    trained = float(np.mean((O - Yh) ** 2))
    print(f"train pairs {n_train}, held-out pairs {Xh.shape[0]}: identity mse {identity:.3e}, "
          f"trained mse {trained:.3e} (gpu epoch loss {loss:.3e})")
    assert trained < identity


def _warm_lockstep(case, oracle_mod, precision, prm_kw, steps=4, warm_values=None):
    """GPU and oracle run the same warm-started steps from phi0: bit-exact
    phi, narrowband, schedule and acceptance decisions. warm_values(k, phi)
    gives caller warm distances for step k (LS_WARM_LEARNED through
    ls_set_warm), or None for the identity predictor."""
    g = lsb.solver_for_case(case, precision=precision)
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts, 64 if precision else 32)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    warm = lsb.LS_WARM_IDENTITY if warm_values is None else lsb.LS_WARM_LEARNED
    rows = []
    for k in range(steps):
        kw = dict(prm_kw)
        if k > 0:
            kw["warm"] = warm
            if warm_values is not None:
                W = warm_values(k, g.phi())
                g.set_warm(W)
                o.set_warm(W)
        gs = g.step(lsb.step_params(case.dt, case.bandwidth, case.niter, mode=case.mode, **kw))
        os_ = o.step(oracle_mod.make_params(case.dt, case.bandwidth, case.niter, mode=case.mode, **kw))
        rows.append((gs.fim_iters, gs.warm_patches, gs.warm_fallback, gs.warm_residual))
        assert (gs.warm_patches, gs.warm_fallback, gs.fim_iters, gs.patch_visits, gs.probes) == \
               (os_.warm_patches, os_.warm_fallback, os_.fim_iters, os_.patch_visits, os_.probes), k
        assert np.float32(gs.warm_residual) == np.float32(os_.warm_residual), k
        np.testing.assert_array_equal(g.narrowband(), o.narrowband(), err_msg=f"narrowband step {k}")
        if precision == lsb.LS_F64:
            np.testing.assert_array_equal(g.phi(), o.phi(), err_msg=f"phi step {k}")
        else:
            np.testing.assert_array_equal(g.phi().astype(np.float32), o.phi().astype(np.float32))
    g.close()
    o.close()
    print("iters / warm patches / fallback / residual per step:", rows)
    return rows


@pytest.mark.parametrize("precision", [lsb.LS_F32, lsb.LS_F64])
@pytest.mark.parametrize("margin,abs_h", [(0.0, 0.0), (0.02, 0.25)])
def test_identity_warm_start_bit_exact_with_oracle(oracle_mod, precision, margin, abs_h):
    case = _case(n=20)
    kw = {"tol": 1e-6, "warm_margin": margin, "warm_abs_margin": abs_h * case.h}
    rows = _warm_lockstep(case, oracle_mod, precision, kw, steps=5)
    if margin == 0.0:
        assert all(r[2] == 1 for r in rows[1:])  # |phi| alone undercuts: every warm step falls back
    else:
        assert all(r[2] == 0 for r in rows[1:])


def test_caller_warm_values_bit_exact_with_oracle(oracle_mod):
    """ls_set_warm: upper bounds are accepted, a planted undercut inside the
    band rejects the step; GPU and oracle agree on every bit and decision."""
    case = _case(n=20)

    def warm(k, phi):
        W = np.abs(phi) * 1.03 + 0.2 * case.h
        if k == 3:  # one vertex well inside the band and away from the front, far below
            W[int(np.argmin(np.abs(np.abs(phi) - 0.5 * case.bandwidth)))] = 0.0
        return W
    rows = _warm_lockstep(case, oracle_mod, lsb.LS_F64, {"tol": 1e-6}, steps=5, warm_values=warm)
    assert [r[2] for r in rows[1:]] == [0, 0, 1, 0]


def test_biased_network_falls_back_to_plain_bits(mlp_oracle):
    """A network whose output bias is -1 predicts |phi| - |AB| (below every
    distance): each warm step is rejected and returns the plain method's
    bits."""
    case = _case(n=20)
    p = _params(mlp_oracle)
    p[-4:] = -1.0
    runs = {}
    for warm in (lsb.LS_WARM_NONE, lsb.LS_WARM_LEARNED):
        g = lsb.solver_for_case(case, precision=lsb.LS_F64, execution=lsb.LS_EXEC_GRAPH)
        g.mlp_set_params(p)
        out = []
        for k in range(4):
            if warm and k > 0:
                g.mlp_predict()
            st = g.step(lsb.step_params(case.dt, case.bandwidth, case.niter, tol=1e-6,
                                        warm=warm if k > 0 else lsb.LS_WARM_NONE, warm_margin=0.02))
            out.append((g.phi(), g.narrowband(), st.warm_fallback))
        g.close()
        runs[warm] = out
    for (a, na, _), (b, nb_, fb) in zip(runs[lsb.LS_WARM_NONE][1:], runs[lsb.LS_WARM_LEARNED][1:]):
        assert fb == 1
        np.testing.assert_array_equal(na, nb_)
        np.testing.assert_array_equal(a, b)


def test_learned_warm_start_bit_exact_with_oracle_in_mcf_mode(mlp_oracle, oracle_mod):
    """Mean-curvature flow with the network's warm values: the oracle takes
    the GPU's predictions through set_warm and reproduces every step."""
    case = cases.mcf_sphere(n=20, block=(4, 4, 4), steps=5, seed=3)
    g = lsb.solver_for_case(case, precision=lsb.LS_F64)
    g.mlp_set_params(_params(mlp_oracle))
    o = oracle_mod.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    fallbacks = []
    for k in range(5):
        kw = {"tol": 1e-6, "mode": lsb.LS_MODE_MCF}
        if k > 0:
            g.mlp_predict()
            o.set_warm(g.warm_values())
            kw.update(warm=lsb.LS_WARM_LEARNED, warm_margin=0.03, warm_abs_margin=0.25 * case.h)
        gs = g.step(lsb.step_params(case.dt, case.bandwidth, case.niter, **kw))
        os_ = o.step(oracle_mod.make_params(case.dt, case.bandwidth, case.niter, **kw))
        fallbacks.append(gs.warm_fallback)
        assert (gs.warm_fallback, gs.fim_iters, gs.warm_patches) == (os_.warm_fallback, os_.fim_iters,
                                                                      os_.warm_patches), k
        np.testing.assert_array_equal(g.narrowband(), o.narrowband())
        np.testing.assert_array_equal(g.phi(), o.phi(), err_msg=f"phi step {k}")
    print("fallbacks per step:", fallbacks)
    g.close()
    o.close()


def test_warm_start_on_triangles_bit_exact_with_oracle(oracle_mod):
    """The warm start (identity predictor; the network itself is per tet) on
    the 2D circle workload, FP64 triangles."""
    import dataclasses
    case = dataclasses.replace(cases.config1(n=65), labels=None, nparts=None)
    lab = cases.partition_rcb(case.xyz, 64)
    case = dataclasses.replace(case, labels=lab, nparts=64)
    rows = _warm_lockstep(case, oracle_mod, lsb.LS_F64, {"tol": 1e-6, "warm_margin": 0.02,
                                                          "warm_abs_margin": 0.25 * case.h}, steps=5)
    assert rows[1][1] > 0


def test_learned_warm_start_ablation(mlp_oracle, oracle_mod):
    """Plain, identity and learned warm starts on the same steps. Each step
    is checked against two FP64 oracle runs from the same input: the same
    warm-started scheme (the identity predictor, or the network's predictions
    handed over as caller values), which the GPU must match bit for bit
    (L-inf 0, within the 1e-5 bar), and the plain method. An accepted warm
    step needs fewer FIM iterations than the plain method from the same input
    and keeps its narrowband; its values lie within the plain method's own
    schedule sensitivity (slots lowered inside a visit are never written
    back, so the active schedule moves values inside the band by ~1e-4); a
    rejected one returns the plain bits. Prints FIM iterations, visits and
    errors per mode."""
    M = mlp_oracle
    case = _case(n=24, steps=8)
    s, X, Y, _ = _train(case, M, precision=lsb.LS_F64)
    params = s.mlp_params()
    s.close()
    margins = {"warm_margin": 0.03, "warm_abs_margin": 0.25 * case.h}
    modes = {"plain": lsb.LS_WARM_NONE, "identity": lsb.LS_WARM_IDENTITY, "learned": lsb.LS_WARM_LEARNED}
    table = {}
    for name, mode in modes.items():
        g = lsb.solver_for_case(case, precision=lsb.LS_F64)
        g.mlp_set_params(params)
        o = oracle_mod.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
        o.set_velocity(case.vel)
        ow = oracle_mod.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
        ow.set_velocity(case.vel)
        phi, nb = case.phi0, None
        rows, warm_iters, plain_iters = [], 0, 0
        for k in range(5):
            wmode = mode if k > 0 else lsb.LS_WARM_NONE
            if wmode == lsb.LS_WARM_LEARNED:
                g.mlp_predict()
                ow.set_warm(g.warm_values())
            st = g.step(lsb.step_params(case.dt, case.bandwidth, case.niter, tol=1e-6, warm=wmode, **margins))
            gphi = g.phi()
            for orc in (o, ow):
                orc.set_phi(phi)
                if nb is not None:
                    orc.set_narrowband(nb)
            ost = o.step(oracle_mod.make_params(case.dt, case.bandwidth, case.niter, tol=1e-6))
            wst = ow.step(oracle_mod.make_params(case.dt, case.bandwidth, case.niter, tol=1e-6, warm=wmode, **margins))
            # the same scheme: bit-exact, same schedule and decision
            np.testing.assert_array_equal(gphi, ow.phi(), err_msg=f"{name} step {k}")
            assert (st.fim_iters, st.warm_fallback) == (wst.fim_iters, wst.warm_fallback), (name, k)
            ophi = o.phi()
            band = np.abs(ophi) <= case.bandwidth
            err = float(np.max(np.abs(gphi - ophi)[band]))
            rows.append((st.fim_iters, ost.fim_iters, st.patch_visits, st.warm_fallback, f"{err:.1e}"))
            np.testing.assert_array_equal(g.narrowband(), o.narrowband())
            if wmode == lsb.LS_WARM_NONE or st.warm_fallback:
                np.testing.assert_array_equal(gphi, ophi)
            else:
                assert err <= 1e-3, (name, k, err)
                warm_iters += st.fim_iters
                plain_iters += ost.fim_iters
            phi, nb = gphi, g.narrowband()
        g.close()
        o.close()
        ow.close()
        table[name] = rows
        print(name, "iters / plain iters / visits / fallback / in-band Linf vs plain per step:", rows)
        if mode != lsb.LS_WARM_NONE:
            assert plain_iters > 0, f"{name}: no warm step was accepted"
            assert warm_iters < plain_iters, (name, warm_iters, plain_iters)


def _tri_case():
    import dataclasses
    case = dataclasses.replace(cases.config1(n=65), labels=None, nparts=None)
    lab = cases.partition_rcb(case.xyz, 64)
    return dataclasses.replace(case, labels=lab, nparts=64)


def test_triangle_pairs_and_predictions_match_network_oracle(mlp_oracle):
    """The network on triangles (the 2D circle workload): collected inputs
    and targets against the host features, and the per-vertex warm values
    against the numpy forward pass with BF16 operands, averaged over each
    vertex's copies in its own patch."""
    M = mlp_oracle
    case = _tri_case()
    p = _params(M)
    s = lsb.solver_for_case(case, precision=lsb.LS_F64)
    s.mlp_set_params(p)
    prm = lsb.step_params(case.dt, case.bandwidth, case.niter)
    s.step(prm)
    nb = s.narrowband()
    phi_before = s.phi()
    s.mlp_collect(0, 100000)
    s.step(prm)
    s.mlp_collect(1)
    X, Y = s.mlp_dataset()
    dist = s.distances()
    lay = s.layout()
    eo = lay["ele_off"]
    copies = np.concatenate([np.arange(eo[q], eo[q + 1]) for q in nb])
    assert X.shape[0] == copies.size
    tris = case.cells[lay["copy_src"][copies]]
    want, sc = M.features_tri(phi_before[tris].astype(np.float32), M.local_coords_tri(case.xyz, tris))
    np.testing.assert_allclose(X, want, rtol=2e-5, atol=2e-6)
    D = dist[tris]
    wy = np.zeros((copies.size, 4))
    wy[:, :3] = np.where(D < 5e8, D / sc[:, None] - np.abs(want[:, :3]), 0.0)
    np.testing.assert_allclose(Y, wy, rtol=1e-4, atol=1e-4)

    # predictions from the current state over the current narrowband
    nb = s.narrowband()
    phi = s.phi()
    s.mlp_predict()
    W = s.warm_values()
    s.close()
    copies = np.concatenate([np.arange(eo[q], eo[q + 1]) for q in nb])
    cpatch = np.repeat(nb, [eo[q + 1] - eo[q] for q in nb])
    tris = case.cells[lay["copy_src"][copies]]
    Xn, scn = M.features_tri(phi[tris].astype(np.float32), M.local_coords_tri(case.xyz, tris))
    out = M.forward(p.astype(np.float64), Xn, bf16_ops=True)[0]
    pred = np.maximum(np.abs(phi[tris].astype(np.float32)) + scn[:, None] * out[:, :3], 0.0)
    acc = np.zeros(case.xyz.shape[0])
    cnt = np.zeros(case.xyz.shape[0])
    for k in range(3):
        own = case.labels[tris[:, k]] == cpatch
        np.add.at(acc, tris[own, k], pred[own, k])
        np.add.at(cnt, tris[own, k], 1)
    v = cnt > 0
    print(f"predicted vertices {int(v.sum())}, max |diff| {np.max(np.abs(W[v] - acc[v] / cnt[v])):.2e}")
    np.testing.assert_allclose(W[v], acc[v] / cnt[v], rtol=1e-4, atol=1e-6)


def test_learned_warm_start_on_triangles_bit_exact_with_oracle(mlp_oracle, oracle_mod):
    """The network's warm values on triangles: the oracle takes the GPU's
    predictions through set_warm and reproduces every step and decision."""
    case = _tri_case()
    g = lsb.solver_for_case(case, precision=lsb.LS_F64)
    g.mlp_set_params(_params(mlp_oracle))
    o = oracle_mod.Oracle(2, case.xyz, case.cells, case.labels, case.nparts, 64)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    rows = []
    for k in range(5):
        kw = {"tol": 1e-6}
        if k > 0:
            g.mlp_predict()
            o.set_warm(g.warm_values())
            kw.update(warm=lsb.LS_WARM_LEARNED, warm_margin=0.03, warm_abs_margin=0.25 * case.h)
        gs = g.step(lsb.step_params(case.dt, case.bandwidth, case.niter, **kw))
        os_ = o.step(oracle_mod.make_params(case.dt, case.bandwidth, case.niter, **kw))
        rows.append((gs.fim_iters, gs.warm_patches, gs.warm_fallback))
        assert (gs.warm_fallback, gs.fim_iters, gs.warm_patches) == (os_.warm_fallback, os_.fim_iters,
                                                                      os_.warm_patches), k
        np.testing.assert_array_equal(g.narrowband(), o.narrowband())
        np.testing.assert_array_equal(g.phi(), o.phi(), err_msg=f"phi step {k}")
    print("iters / warm patches / fallback per step:", rows)
    assert rows[1][1] > 0
    g.close()
    o.close()
