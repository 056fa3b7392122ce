"""Mean-curvature flow in the oracle against what the mathematics fixes.

* A sphere under phi_t = kappa |grad phi| shrinks as R(t)^2 = R0^2 - 4t, a
  circle as R(t)^2 = R0^2 - 2t: the zero level set of the oracle must follow.
  A sign error grows the sphere; a wrong mass normalisation (sum of volumes
  rather than the lumped mass, a factor d+1) changes the rate 4x.
* A plane has zero curvature: interior vertices do not move, because
  sum_e vol_e gradN_i = 0 at interior vertices for any constant normal.
* The configured timestep lies inside the explicit stability bound.
"""
import numpy as np
import pytest

import levelset_amd.cases as cases


def _run(oracle_mod, case, precision=64, steps=None):
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts, precision)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    prm = oracle_mod.make_params(case.dt, case.bandwidth, case.niter, mode=1)
    out = []
    for _ in range(steps or case.steps):
        o.step(prm)
        out.append(o.phi())
    o.close()
    return out


@pytest.mark.parametrize("precision", [32, 64])
def test_mcf_sphere_follows_closed_form(precision, oracle_mod):
    case = cases.mcf_sphere(n=24, steps=40)
    r0, h = case.meta["r0"], case.h
    phis = _run(oracle_mod, case, precision)
    prev = r0
    for k, phi in enumerate(phis):
        t = (k + 1) * case.dt
        exact = np.sqrt(r0 * r0 - 4.0 * t)
        got = cases.interface_radius(case.xyz, phi, (0.5, 0.5, 0.5), h, 3)
        assert abs(got - exact) <= 0.1 * h, (k, got, exact)
        assert got < prev
        prev = got
    # the radius dropped by more than a cell width, so the rate is tested
    assert r0 - prev > h


def test_mcf_circle_follows_closed_form(oracle_mod):
    case = cases.mcf_circle(n=33, steps=40)
    r0, h = case.meta["r0"], case.h
    phis = _run(oracle_mod, case)
    for k, phi in enumerate(phis):
        t = (k + 1) * case.dt
        exact = np.sqrt(r0 * r0 - 2.0 * t)
        got = cases.interface_radius(case.xyz, phi, (0.5, 0.5, 0.5), h, 2)
        assert abs(got - exact) <= 0.05 * h, (k, got, exact)
    assert r0 - got > 0.25 * h


def test_mcf_plane_interior_is_stationary(oracle_mod):
    n = 12
    xyz, cells, h = cases.kuhn_grid(n)
    phi0 = xyz[:, 0] - 0.5 + 0.3 * h  # a plane between vertex layers
    labels, P = cases.partition_blocks(n, (4, 4, 4), 3)
    o = oracle_mod.Oracle(3, xyz, cells, labels, P, 64)
    o.set_phi(phi0)
    o.set_velocity(np.zeros((cells.shape[0], 3)))
    o.step(oracle_mod.make_params(0.125 * h * h, 20 * h, 5, mode=1))
    phi = o.phi()
    o.close()
    interior = np.all((xyz > 1e-9) & (xyz < 1 - 1e-9), axis=1)
    assert np.max(np.abs(phi[interior] - phi0[interior])) < 1e-12


def test_mcf_step_is_explicitly_stable():
    # Largest eigenvalue of the lumped P1 Laplacian on a Kuhn grid: 12/h^2
    # (the 7-point stencil), so explicit Euler needs dt <= h^2/6.
    assert cases.MCF_DT_FACTOR <= 1.0 / 6.0


def test_mcf_large_step_diverges(oracle_mod):
    # dt = 0.5 h^2 is beyond the bound: the interface radius leaves the
    # closed form by more than a cell within a dozen steps.
    case = cases.mcf_sphere(n=24, steps=12)
    case.dt = 0.5 * case.h * case.h
    phi = _run(oracle_mod, case)[-1]
    exact = np.sqrt(case.meta["r0"] ** 2 - 4.0 * 12 * case.dt)
    assert abs(cases.interface_radius(case.xyz, phi, (0.5, 0.5, 0.5), case.h, 3) - exact) > case.h
