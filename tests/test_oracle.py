"""Pins the CPU oracle without the reference (which cannot be built here):
brute-force Hopf-Lax minimisation, textbook closed forms, invariants of the
scheme and convergence to analytic distances. CPU only."""
import numpy as np
import pytest

import levelset_amd.cases as cases


def _dots(A, B, C, D):
    AC, BC, CD = C - A, C - B, D - C
    return [AC @ AC, AC @ BC, AC @ CD, BC @ BC, BC @ CD, CD @ CD]


def _brute_tet(A, B, C, D, TA, TB, TC):
    """min over the face ABC of l1 TA + l2 TB + l3 TC + |D - (l1 A + l2 B + l3 C)|,
    by zooming grids over the simplex (the function is convex)."""
    def f(l1, l2):
        l3 = 1.0 - l1 - l2
        E = l1[:, None] * A[None] + l2[:, None] * B[None] + l3[:, None] * C[None]
        return l1 * TA + l2 * TB + l3 * TC + np.linalg.norm(D[None] - E, axis=1)
    c1, c2, w = 1.0 / 3, 1.0 / 3, 1.0
    best = np.inf
    for _ in range(30):
        g = np.linspace(-w, w, 61)
        l1, l2 = np.meshgrid(c1 + g, c2 + g, indexing="ij")
        l1, l2 = l1.ravel(), l2.ravel()
        ok = (l1 >= 0) & (l2 >= 0) & (l1 + l2 <= 1)
        l1, l2 = l1[ok], l2[ok]
        v = f(l1, l2)
        k = int(np.argmin(v))
        best, c1, c2 = min(best, float(v[k])), l1[k], l2[k]
        w *= 0.5
    return best


def test_tet_solve_matches_brute_force_hopf_lax(oracle_mod):
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(60):
        A, B, C = rng.normal(size=(3, 3))
        D = rng.normal(size=3) + np.array([0.0, 0.0, 1.5])
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        off = rng.normal()
        T = [abs(P @ n - off) for P in (A, B, C)]
        got = oracle_mod.tet_solve(T, _dots(A, B, C, D), 64, lambda_eps=1e-12)
        ref = _brute_tet(A, B, C, D, *T)
        worst = max(worst, abs(got - ref))
    assert worst < 1e-9


def test_tet_solve_zero_values_is_point_to_triangle_distance(oracle_mod):
    # T_A = T_B = T_C = 0: the solve is the distance from D to triangle ABC;
    # here the foot lies on an edge, so every candidate family is exercised.
    A, B, C = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    for D, expect in [(np.array([0.5, -0.3, 0.4]), np.hypot(0.3, 0.4)),
                      (np.array([-0.3, -0.4, 0.0]), 0.5),
                      (np.array([1.3, 0.0, 0.4]), 0.5)]:
        got = oracle_mod.tet_solve([0.0, 0.0, 0.0], _dots(A, B, C, D), 64, lambda_eps=1e-12)
        assert got == pytest.approx(expect, abs=1e-12)


def test_tet_solve_linear_field_with_interior_foot(oracle_mod):
    # T linear with |grad T| = 1 and its characteristic through D hitting the
    # face interior: the solve reproduces T(D) exactly.
    A, B, C = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    g = np.array([0.2, 0.1, 1.0])
    g /= np.linalg.norm(g)
    D = np.array([0.3, 0.3, 0.5])
    T = [P @ g for P in (A, B, C)]
    got = oracle_mod.tet_solve(T, _dots(A, B, C, D), 64, lambda_eps=1e-12)
    assert got == pytest.approx(D @ g, abs=1e-12)


def test_tet_solve_all_unreached_stays_unreached(oracle_mod):
    A, B, C, D = np.eye(3)[0], np.eye(3)[1], np.eye(3)[2], np.zeros(3)
    assert oracle_mod.tet_solve([1e9, 1e9, 1e9], _dots(A, B, C, D), 64) == 1e9


def test_tri_solve_matches_brute_force_and_segment_distance(oracle_mod):
    rng = np.random.default_rng(1)
    for _ in range(50):
        A, B, C = rng.normal(size=(3, 2))
        TA, TB = abs(rng.normal()), abs(rng.normal())
        AB, BC, AC = np.linalg.norm(B - A), np.linalg.norm(C - B), np.linalg.norm(C - A)
        f = lambda lam: lam * TA + (1 - lam) * TB + np.linalg.norm(C - (lam * A + (1 - lam) * B))
        lo, hi = 0.0, 1.0
        for _ in range(200):  # ternary search: f is convex on [0, 1]
            m1, m2 = lo + (hi - lo) / 3, hi - (hi - lo) / 3
            lo, hi = (lo, m2) if f(m1) <= f(m2) else (m1, hi)
        ref = min(f(0.5 * (lo + hi)), f(0.0), f(1.0))
        got = oracle_mod.tri_solve(TA, TB, AB, BC, AC, 64)
        assert abs(got - ref) < 1e-9
    # T_A = T_B = 0: distance from C to segment AB.
    A, B, C = np.array([0.0, 0.0]), np.array([2.0, 0.0]), np.array([0.7, 0.9])
    got = oracle_mod.tri_solve(0.0, 0.0, 2.0, np.linalg.norm(C - B), np.linalg.norm(C - A), 64)
    assert got == pytest.approx(0.9, abs=1e-12)


def test_element_update_is_monotone_and_reaches_a_fixed_point(oracle_mod):
    """Every corner takes min(own, solve from the other three), all reading
    the values from before the update: values never rise, and repeating the
    update reaches a fixed point within four applications (information
    crosses the element corner by corner), where it is idempotent."""
    xyz, cells, h = cases.kuhn_grid(3)
    rng = np.random.default_rng(2)
    for t in cells[:12]:
        P = xyz[t]
        AB, AC, AD = P[1] - P[0], P[2] - P[0], P[3] - P[0]
        X = AB / np.linalg.norm(AB)
        Z = np.cross(AB, AC)
        Z /= np.linalg.norm(Z)
        Y = np.cross(Z, X)
        L = [np.linalg.norm(AB), AC @ X, AC @ Y, AD @ X, AD @ Y, AD @ Z]
        v0 = rng.uniform(0, 1, 4)
        v0[rng.integers(4)] = 0.0  # a seed-like corner, so solves lower the others
        v1 = oracle_mod.tet_element_update(L, v0, 64)
        assert np.all(v1 <= v0)
        v2 = oracle_mod.tet_element_update(L, v1, 64)
        assert np.all(v2 <= v1)
        v = v2
        for _ in range(3):
            v = oracle_mod.tet_element_update(L, v, 64)
        again = oracle_mod.tet_element_update(L, v, 64)
        np.testing.assert_array_equal(again, v)


def _run(oracle_mod, case, steps, precision=64, inside=1, leps=1e-5):
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts, precision)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    prm = oracle_mod.make_params(case.dt, case.bandwidth, case.niter, inside_niter=inside, lambda_eps=leps)
    for _ in range(steps):
        o.step(prm)
    return o


def test_fim_converges_to_sphere_distance_first_order(oracle_mod):
    errs = []
    for n in (12, 20):
        case = cases.small3d(n=n, shape="sphere", block=(4, 4, 4))
        o = _run(oracle_mod, case, 1, inside=0)
        phi = o.phi()
        band = np.abs(case.phi0) < 4 * case.h
        errs.append(np.max(np.abs(phi[band] - case.phi0[band])))
    assert errs[1] < errs[0] and errs[1] < 0.02


def test_fim_result_is_a_discrete_fixed_point(oracle_mod):
    # Within the band every non-seed distance equals the min over its
    # elements of the local solves (up to the convergence tolerance).
    case = cases.small3d(n=10, shape="sphere", block=(5, 5, 5))
    o = _run(oracle_mod, case, 1, inside=0, leps=1e-12)
    T = np.abs(o.phi())
    xyz, cells = case.xyz, case.cells
    cand = np.full(T.size, np.inf)
    for t in cells:
        if np.any(T[t] > 1e8):
            continue
        P = xyz[t]
        AB, AC, AD = P[1] - P[0], P[2] - P[0], P[3] - P[0]
        X = AB / np.linalg.norm(AB)
        Z = np.cross(AB, AC)
        Z /= np.linalg.norm(Z)
        Y = np.cross(Z, X)
        L = [np.linalg.norm(AB), AC @ X, AC @ Y, AD @ X, AD @ Y, AD @ Z]
        upd = oracle_mod.tet_element_update(L, [1e9 if k == 3 else T[t[k]] for k in range(4)], 64, 1e-12)
        cand[t[3]] = min(cand[t[3]], upd[3])
        for k in range(3):
            vals = [T[t[j]] for j in range(4)]
            vals[k] = 1e9
            cand[t[k]] = min(cand[t[k]], oracle_mod.tet_element_update(L, vals, 64, 1e-12)[k])
    seeds = np.zeros(T.size, bool)
    for t in cells:
        v = case.phi0[t]
        if not (np.all(v >= 0) or np.all(v <= 0)):
            seeds[t] = True
    inner = (np.abs(case.phi0) < 2 * case.h) & ~seeds & np.isfinite(cand)
    d = cand - T
    assert inner.sum() > 50
    # Converged: no vertex could still be lowered by more than the tolerance.
    assert np.min(d[inner]) > -1e-5
    # A vertex may sit below every local solve from the final values only next
    # to a seed: inside a patch visit a seed's element slots can drop below
    # |phi| and feed its neighbours while the seed itself keeps |phi|
    # (redistance_kernels3d.h:406-449).
    near_seed = np.zeros(T.size, bool)
    for t in cells:
        if seeds[t].any():
            near_seed[t] = True
    low = inner & (d > 1e-5)
    assert np.all(near_seed[low])


def test_advection_translates_linear_data_exactly(oracle_mod):
    # Linear phi and constant sigma: every element has H = vol * sigma.grad(phi)
    # and sum(alpha~) = 1, so each vertex moves by dt * sigma.grad(phi)
    # (meshFIM_kernels.h:227-273, 397-406).
    xyz, cells, h = cases.kuhn_grid(8)
    xyz = cases.jitter_interior(xyz, h, 0.1, 7, 3)
    lab, P = cases.partition_blocks(8, (4, 4, 4), 3)
    g = np.array([0.3, -0.2, 0.1])
    sigma = np.array([0.5, 0.25, -0.4])
    o = oracle_mod.Oracle(3, xyz, cells, lab, P, 64)
    phi = xyz @ g + 0.05
    o.set_phi(phi)
    o.set_velocity(np.tile(sigma, (cells.shape[0], 1)))
    dt = 1e-3
    o.advect_all(dt)
    new = o.phi()
    interior = np.all((xyz > 1e-9) & (xyz < 1 - 1e-9), axis=1)
    np.testing.assert_allclose(new[interior], (phi - dt * (sigma @ g))[interior], atol=1e-12)
    # Vertices on the domain boundary either translate or, at inflow, receive
    # no distributed residual and keep their value (no inflow condition).
    b = ~interior
    moved = np.isclose(new[b], (phi - dt * (sigma @ g))[b], atol=1e-12)
    kept = new[b] == phi[b]
    assert np.all(moved | kept)


def test_advection_zero_velocity_is_identity_and_max_principle(oracle_mod):
    xyz, cells, h = cases.kuhn_grid(8)
    xyz = cases.jitter_interior(xyz, h, 0.1, 8, 3)
    lab, P = cases.partition_blocks(8, (4, 4, 4), 3)
    phi = cases.sdf_sphere(xyz, r=0.3)
    o = oracle_mod.Oracle(3, xyz, cells, lab, P, 64)
    o.set_phi(phi)
    o.set_velocity(np.zeros((cells.shape[0], 3)))
    o.advect_all(1e-2)
    np.testing.assert_array_equal(o.phi(), phi)
    # Rotation with a CFL-limited step: each new value is a convex combination
    # of old values in the vertex's star, so it stays within their range.
    vel = cases.velocity_rotation_z(xyz, cells)
    o.set_phi(phi)
    o.set_velocity(vel)
    dt = 0.5 * h / np.max(np.linalg.norm(vel, axis=1))
    o.advect_all(dt)
    new = o.phi()
    lo = np.full(phi.size, np.inf)
    hi = np.full(phi.size, -np.inf)
    for t in cells:
        for v in t:
            lo[v] = min(lo[v], phi[t].min())
            hi[v] = max(hi[v], phi[t].max())
    assert np.all(new >= lo - 1e-12) and np.all(new <= hi + 1e-12)


def test_advection_2d_translates_linear_data_exactly(oracle_mod):
    case = cases.config1()
    o = oracle_mod.Oracle(2, case.xyz, case.cells, case.labels, case.nparts, 64)
    g = np.array([0.6, 0.8, 0.0])
    phi = case.xyz @ g - 0.4
    o.set_phi(phi)
    o.set_velocity(case.vel)
    o.advect_all(case.dt)
    new = o.phi()
    interior = np.all((case.xyz[:, :2] > 1e-9) & (case.xyz[:, :2] < 1 - 1e-9), axis=1)
    np.testing.assert_allclose(new[interior], (phi - case.dt * (case.vel[0] @ g))[interior], atol=1e-12)


def test_fp32_instance_tracks_fp64_on_distances(oracle_mod):
    case = cases.small3d(n=12, shape="sphere", block=(4, 4, 4))
    a = _run(oracle_mod, case, 1, precision=32, inside=0).phi()
    b = _run(oracle_mod, case, 1, precision=64, inside=0).phi()
    fin = np.abs(b) < 1e8
    assert np.max(np.abs(a[fin] - b[fin])) < 1e-6


def test_narrowband_membership_is_the_brute_force_set(oracle_mod):
    case = cases.small3d(n=12, shape="sphere", block=(4, 4, 4))
    o = _run(oracle_mod, case, 1, inside=0)
    T = np.abs(o.phi())
    nb = set(o.narrowband().tolist())
    expect = {p for p in range(case.nparts) if T[case.labels == p].min() <= case.bandwidth}
    assert nb == expect


@pytest.mark.parametrize("dim", [2, 3])
def test_seed_labels_and_distances_match_brute_force(oracle_mod, dim):
    """Seeds (kernel_seedlabel, src/core3d/include/redistance_kernels3d.h:623-770):
    every corner of an element whose values are neither all >= 0 nor all
    <= 0, and every vertex with phi == 0; seeds keep DT = |phi|. Brute force
    from the signs of phi0 over every element (the first step scans every
    patch), with exact zeros planted on and off the front."""
    case = cases.small3d(n=10, shape="sphere", block=(4, 4, 4)) if dim == 3 else cases.config1()
    phi0 = case.phi0.copy()
    rng = np.random.default_rng(4)
    near = np.flatnonzero(np.abs(phi0) < 0.5 * case.h)
    far = np.flatnonzero(np.abs(phi0) > 3 * case.h)
    phi0[rng.choice(near, 3, replace=False)] = 0.0
    phi0[rng.choice(far, 3, replace=False)] = 0.0  # isolated zeros: seeds through the phi == 0 rule alone
    v = phi0[case.cells]
    mixed = ~((v >= 0).all(axis=1) | (v <= 0).all(axis=1))
    want = np.zeros(phi0.size, bool)
    want[case.cells[mixed].ravel()] = True
    want |= phi0 == 0.0
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts, 64)
    o.set_phi(phi0)
    o.set_velocity(case.vel)
    o.step(oracle_mod.make_params(0.0, case.bandwidth, case.niter))  # dt = 0: phi = signed DT
    got = o.labels() == 0
    dist = np.abs(o.phi())
    o.close()
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(dist[want], np.abs(phi0[want]))
    assert (~want).any() and np.all(dist[~want] > 0.0)
