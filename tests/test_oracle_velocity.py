"""Velocity modes of the oracle's advection, pinned on the CPU.

Mode 2 restates the reference's 2D advection, which ignores the given
velocity and moves each triangle with (phi_A, phi_B) / |grad phi| in its
local frame (src/core2d/include/meshFIM2d_kernels.h:133-134,170-181). The
pin computes that velocity independently in numpy from the global
coordinates (local frame: X along AB, Y along (AB x AC) x AB; |grad phi| of
the linear interpolant), turns it into a global per-element velocity, and
checks that one advection pass in mode 0 with it matches mode 2.

Mode 1 (the patch's first element copy's velocity for the whole patch,
src/core3d/include/meshFIM_kernels.h:51-53) agrees with mode 0 on a uniform
field and differs on a rotation.
"""
import numpy as np

import levelset_amd.cases as cases


def _quirk_velocity(xyz, cells, phi):
    P = xyz[cells]
    A, B, C = P[:, 0], P[:, 1], P[:, 2]
    AB, AC = B - A, C - A
    X = AB / np.linalg.norm(AB, axis=1)[:, None]
    n = np.cross(AB, AC)
    Y = np.cross(n, AB)
    Y /= np.linalg.norm(Y, axis=1)[:, None]
    ph = phi[cells]
    # gradient of the linear interpolant in the plane: solve [AB; AC] g = dphi
    M = np.stack([AB[:, :2], AC[:, :2]], axis=1)
    d = np.stack([ph[:, 1] - ph[:, 0], ph[:, 2] - ph[:, 0]], axis=1)
    g = np.linalg.solve(M, d[:, :, None])[:, :, 0]
    norm = np.linalg.norm(g, axis=1)
    s0 = np.where(norm > 0, ph[:, 0] / np.where(norm > 0, norm, 1), 0.0)
    s1 = np.where(norm > 0, ph[:, 1] / np.where(norm > 0, norm, 1), 0.0)
    return s0[:, None] * X + s1[:, None] * Y


def test_phi_velocity_mode_matches_explicit_velocity(oracle_mod):
    case = cases.config1(n=33)
    phi = case.phi0
    runs = []
    for mode, vel in ((2, case.vel), (0, _quirk_velocity(case.xyz, case.cells, phi))):
        o = oracle_mod.Oracle(2, case.xyz, case.cells, case.labels, case.nparts, 64)
        o.set_phi(phi)
        o.set_velocity(vel, mode)
        o.advect_all(case.dt)
        runs.append(o.phi())
        o.close()
    moved = np.abs(runs[0] - phi) > 1e-9
    print(f"vertices moved: {int(moved.sum())} of {phi.size}; max diff {np.max(np.abs(runs[0] - runs[1])):.2e}")
    assert moved.sum() > phi.size // 2
    np.testing.assert_allclose(runs[0], runs[1], rtol=0, atol=1e-12)


def test_patch_first_velocity_mode(oracle_mod):
    """A uniform velocity field moves identically in modes 0 and 1; the
    rotation field differs across a patch, so mode 1 moves differently."""
    case = cases.small3d(n=10, shape="sphere", block=(4, 4, 4))
    u = np.tile([0.3, -0.2, 0.1], (case.cells.shape[0], 1))
    runs = {}
    for name, vel, mode in (("u0", u, 0), ("u1", u, 1), ("r0", case.vel, 0), ("r1", case.vel, 1)):
        o = oracle_mod.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
        o.set_phi(case.phi0)
        o.set_velocity(vel, mode)
        o.advect_all(case.dt)
        runs[name] = o.phi()
        o.close()
    np.testing.assert_array_equal(runs["u0"], runs["u1"])
    assert np.max(np.abs(runs["r0"] - runs["r1"])) > 1e-6
