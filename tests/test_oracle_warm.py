"""The oracle's warm start and its acceptance check (CPU).

A warm start sets the vertices of the previous narrowband that share no
element with a seed to prediction * (1 + margin) + abs_margin before the
FIM. The FIM only lowers values, so a prediction above the result is
finished by it, while one below it (an undercut) would survive: the
acceptance check must reject every undercut (the step then re-runs the plain
method and returns its bits) and pass upper bounds.

The plain method's result is not a function of phi alone: halo and seed
slots lowered inside a patch visit are never written back
(src/core3d/include/redistance_kernels3d.h:440-449), so the schedule (sweeps
per visit, which patches are active when) moves values inside the band by
~1e-4. test_plain_result_depends_on_the_schedule pins that; it is why an
accepted warm start agrees with the plain method only to that level.
"""
import numpy as np
import pytest

import levelset_amd.cases as cases

TOL = 1e-6   # FIM tolerance of these runs


@pytest.fixture(scope="module")
def state(oracle_mod):
    """A 24^3 torus after two plain timesteps: phi, narrowband, the plain
    step's result from there (tol 1e-6) and its distances (dt = 0)."""
    case = cases.small3d(n=24, shape="torus", block=(4, 4, 4), steps=4)
    o = oracle_mod.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    for _ in range(2):
        o.step(oracle_mod.make_params(case.dt, case.bandwidth, case.niter, tol=TOL))
    S, nb = o.phi(), o.narrowband()

    def run(warm=0, W=None, margin=0.0, abs_margin=0.0, tol=TOL, res_tol=0.0, dt=None, niter=None):
        o.set_phi(S)
        o.set_narrowband(nb)
        if W is not None:
            o.set_warm(W)
        st = o.step(oracle_mod.make_params(case.dt if dt is None else dt, case.bandwidth, niter or case.niter,
                                           tol=tol, warm=warm, warm_margin=margin, warm_abs_margin=abs_margin,
                                           warm_res_tol=res_tol))
        return o.phi(), o.narrowband(), st

    plain = run()
    dist = np.abs(run(dt=0.0)[0])  # dt = 0 leaves phi at the signed distances
    yield {"case": case, "run": run, "plain": plain, "dist": dist, "S": S}
    o.close()


def _band(state):
    return np.abs(state["plain"][0]) <= state["case"].bandwidth


def _seed_ring(state):
    """(seeds, vertices sharing an element with a seed but not seeds)."""
    cells, S = state["case"].cells, state["S"]
    sgn = np.sign(S[cells])
    mixed = (sgn.min(axis=1) < 0) & (sgn.max(axis=1) > 0)
    seed = np.zeros(S.size, bool)
    seed[cells[mixed].ravel()] = True
    seed |= S == 0.0
    ring = np.zeros(S.size, bool)
    ring[cells[np.any(seed[cells], axis=1)].ravel()] = True
    return seed, ring & ~seed


def test_upper_bound_warm_start_is_accepted(state):
    """|phi| (the identity predictor) with a 2 % + h/4 margin lies above the
    result everywhere here: accepted, same narrowband, no more FIM iterations
    (every patch of the previous band is active from the first iteration;
    on the 64^3 torus that saves 2 of 13)."""
    case = state["case"]
    phi, nb, st = state["run"](warm=1, margin=0.02, abs_margin=0.25 * case.h)
    pphi, pnb, pst = state["plain"]
    inb = _band(state)
    err = float(np.max(np.abs(phi - pphi)[inb]))
    print(f"warm iters {st.fim_iters} visits {st.patch_visits} vs plain {pst.fim_iters} / {pst.patch_visits}; "
          f"in-band L-inf vs plain {err:.2e}")
    assert st.warm_fallback == 0 and st.warm_patches > 0
    assert st.warm_residual == 0.0  # no vertex inside the band kept its warm value
    np.testing.assert_array_equal(nb, pnb)
    assert st.fim_iters <= pst.fim_iters
    assert err <= 1e-3  # the plain method's own schedule sensitivity (see below)


def test_plain_result_depends_on_the_schedule(state):
    """The plain method at a 1e-10 tolerance with 5 and with 10 sweeps per
    visit: both converged, yet they differ inside the band by far more than
    1e-5, because slots lowered inside a visit are never written back."""
    inb = _band(state)
    r5 = state["run"](dt=0.0, tol=1e-10, niter=5)[0]
    r10 = state["run"](dt=0.0, tol=1e-10, niter=10)[0]
    d = float(np.max(np.abs(r5 - r10)[inb]))
    print(f"niter 5 vs 10 at tol 1e-10: in-band L-inf {d:.2e}")
    assert d > 1e-5


@pytest.mark.parametrize("kind", ["zeroed", "biased_low", "one_vertex", "identity_no_margin"])
def test_undercut_warm_start_falls_back_to_plain_bits(state, kind):
    dist = state["dist"]
    pphi, pnb, pst = state["plain"]
    fin = dist < 1e8
    inb = _band(state)
    _, ring = _seed_ring(state)
    W = np.where(fin, dist + 1e-3, 1e9)
    warm = 2
    if kind == "zeroed":
        W = np.zeros_like(dist)
    elif kind == "biased_low":
        W = np.where(fin, 0.9 * dist, 1e9)
    elif kind == "one_vertex":
        # the deepest in-band vertex off the seed ring, 1e-3 below its value
        v = int(np.argmax(np.where(inb & fin & ~ring, dist, -1.0)))
        W[v] = dist[v] - 1e-3
    elif kind == "identity_no_margin":
        warm, W = 1, None
    phi, nb, st = state["run"](warm=warm, W=W)
    print(f"{kind}: residual {st.warm_residual:.3e} fallback {st.warm_fallback}")
    assert st.warm_fallback == 1
    assert st.warm_residual > TOL
    np.testing.assert_array_equal(nb, pnb)
    np.testing.assert_array_equal(phi, pphi)


def test_seed_ring_starts_cold(state):
    """Vertices next to a seed never take a warm value: planting zeros there
    changes nothing, bit for bit."""
    dist = state["dist"]
    _, ring = _seed_ring(state)
    W = np.where(dist < 1e8, dist + 1e-3, 1e9)
    ref = state["run"](warm=2, W=W)
    W2 = W.copy()
    W2[ring] = 0.0
    got = state["run"](warm=2, W=W2)
    assert ref[2].warm_fallback == 0 and got[2].warm_fallback == 0
    np.testing.assert_array_equal(got[0], ref[0])
    np.testing.assert_array_equal(got[1], ref[1])
