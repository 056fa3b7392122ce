"""The oracle's CFL bound (kernel_compute_timestep2d,
src/core2d/include/meshFIM2d_kernels.h:405-475; meshFIM2d::compute_deltaT,
src/core2d/cuda/meshFIM2d.cu:609-674), pinned on the CPU:

* inscribed radii against closed forms (equilateral and 3-4-5 triangles,
  regular and Kuhn tetrahedra);
* the bound against a numpy brute force over the narrowband's elements,
  computed from the original vertex coordinates and velocities (not from the
  oracle's local frames), with random per-element coefficients;
* a step with cfl = 1 advects with min(dt, bound).
"""
import math

import numpy as np
import pytest

import levelset_amd.cases as cases


def _local_tet(P):
    A, B, C, D = (np.asarray(x, float) for x in P)
    X = (B - A) / np.linalg.norm(B - A)
    Z = np.cross(B - A, C - A)
    Z /= np.linalg.norm(Z)
    Y = np.cross(Z, X)
    return [np.linalg.norm(B - A), (C - A) @ X, (C - A) @ Y, (D - A) @ X, (D - A) @ Y, (D - A) @ Z]


def _local_tri(P):
    A, B, C = (np.asarray(x, float) for x in P)
    X = (B - A) / np.linalg.norm(B - A)
    n = np.cross(B - A, C - A)
    Y = np.cross(n, B - A)
    Y /= np.linalg.norm(Y)
    return [np.linalg.norm(B - A), (C - A) @ X, (C - A) @ Y]


def test_inscribed_radius_closed_forms(oracle_mod):
    a = 0.7
    tri_eq = [(0, 0, 0), (a, 0, 0), (a / 2, a * math.sqrt(3) / 2, 0)]
    assert oracle_mod.inradius(2, _local_tri(tri_eq)) == pytest.approx(a / (2 * math.sqrt(3)), rel=1e-14)
    tri_345 = [(0, 0, 0), (3, 0, 0), (0, 4, 0)]
    assert oracle_mod.inradius(2, _local_tri(tri_345)) == pytest.approx(1.0, rel=1e-14)
    s = 1.3  # regular tet with edge s * sqrt(2)
    reg = [(s, s, s), (s, -s, -s), (-s, s, -s), (-s, -s, s)]
    edge = 2 * s * math.sqrt(2)
    assert oracle_mod.inradius(3, _local_tet(reg)) == pytest.approx(edge / (2 * math.sqrt(6)), rel=1e-13)
    h = 0.25  # Kuhn (path) tet: legs h, h, h; volume h^3/6, faces h^2/2 (x2) and h^2 sqrt(2)/2 (x2)
    kuhn = [(0, 0, 0), (h, 0, 0), (h, h, 0), (h, h, h)]
    assert oracle_mod.inradius(3, _local_tet(kuhn)) == pytest.approx(h / (2 + 2 * math.sqrt(2)), rel=1e-13)


def _brute_force(case, patches_of_elem, nb, ceik, ccurv):
    """Least bound and largest inscribed radius over the elements with a copy
    in a narrowband patch, from global coordinates."""
    xyz, cells, d = case.xyz, case.cells, case.dim
    P = xyz[cells]
    if d == 3:
        a, b, c = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]
        vol = np.abs(np.einsum("ij,ij->i", np.cross(a, b), c)) / 6
        faces = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
        area = sum(np.linalg.norm(np.cross(P[:, j] - P[:, i], P[:, k] - P[:, i]), axis=1) / 2 for i, j, k in faces)
        rin = 3 * vol / area
    else:
        ar = np.abs(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])[:, 2]) / 2
        per = sum(np.linalg.norm(P[:, (i + 1) % 3] - P[:, i], axis=1) for i in range(3))
        rin = 2 * ar / per
    cadv = np.linalg.norm(case.vel[:, :d], axis=1)
    dt1 = np.where(np.abs(ccurv) > 1e-12, 4 * rin ** 2 / np.maximum(np.abs(ccurv), 1e-300) / 2 / d, 1e9)
    s = np.abs(ceik) + cadv
    dt2 = np.where(s > 1e-12, 2 * rin / np.maximum(s, 1e-300) / d, 1e9)
    sel = np.isin(patches_of_elem, nb).any(axis=1)
    return float(np.min(np.minimum(dt1, dt2)[sel])), float(np.max(rin[sel]))


@pytest.mark.parametrize("which", ["tet", "tri"])
def test_cfl_bound_matches_brute_force(oracle_mod, which):
    case = cases.small3d(n=12, shape="sphere", block=(4, 4, 4)) if which == "tet" else cases.config1()
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts, 64)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    rng = np.random.default_rng(7)
    E = case.cells.shape[0]
    ceik, ccurv = rng.uniform(-12, 12, E), rng.uniform(-2, 2, E)
    patches_of_elem = case.labels[case.cells]
    for coeffs in ((None, None), (ceik, ccurv)):
        o.set_cfl_coefficients(*coeffs)
        got_all = o.compute_dt()  # before the first step: every patch
        want_all = _brute_force(case, patches_of_elem, np.arange(case.nparts),
                                np.full(E, 10.0) if coeffs[0] is None else ceik,
                                np.zeros(E) if coeffs[1] is None else ccurv)
        np.testing.assert_allclose(got_all, want_all, rtol=1e-12)
    o.step(oracle_mod.make_params(case.dt, case.bandwidth, case.niter))
    nb = o.narrowband()
    got = o.compute_dt()
    want = _brute_force(case, patches_of_elem, nb, ceik, ccurv)
    print(f"{which}: bound {got[0]:.6e} max R {got[1]:.6e} over {nb.size} narrowband patches")
    np.testing.assert_allclose(got, want, rtol=1e-12)
    o.close()


def test_cfl_step_uses_the_bound(oracle_mod):
    """cfl = 1 advects with min(dt, bound of the new narrowband): identical to
    a plain step given that dt directly."""
    case = cases.small3d(n=12, shape="sphere", block=(4, 4, 4))
    big = 10 * case.dt  # far beyond the bound, so the bound decides
    runs = []
    for cfl in (1, 0):
        o = oracle_mod.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
        o.set_phi(case.phi0)
        o.set_velocity(case.vel)
        if cfl:
            st = o.step(oracle_mod.make_params(big, case.bandwidth, case.niter, cfl=1))
            bound = o.compute_dt()[0]  # the narrowband the step advected
            assert st.dt_used == bound < big
            dt_used = st.dt_used
        else:
            st = o.step(oracle_mod.make_params(dt_used, case.bandwidth, case.niter))
        runs.append(o.phi())
        o.close()
    np.testing.assert_array_equal(runs[0], runs[1])
