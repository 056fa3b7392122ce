"""The oracle against its committed golden vectors (tests/golden/*.npz,
written by tests/golden/make_golden.py): reruns must reproduce them bit for
bit, and the vectors must keep the properties the scheme guarantees."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
sys.path.insert(0, GOLDEN)

import make_golden  # noqa: E402


def _load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", ["pyramid", "torus", "circle2d"])
@pytest.mark.parametrize("precision", [32, 64])
def test_oracle_reproduces_golden(name, precision, oracle_mod):
    g = _load(name)
    case = make_golden.CASES[name]()
    np.testing.assert_array_equal(case["cells"], g["cells"])
    np.testing.assert_array_equal(case["phi0"], g["phi0"])
    phis, nbs, amb = make_golden.run_oracle(oracle_mod, case, precision)
    np.testing.assert_array_equal(phis, g[f"phi_fp{precision}"])
    np.testing.assert_array_equal(amb, g[f"amb_fp{precision}"])
    np.testing.assert_array_equal(np.concatenate(nbs), g[f"nb_fp{precision}"])


@pytest.mark.parametrize("name", ["pyramid", "torus", "circle2d"])
def test_golden_fp32_tracks_fp64(name):
    # Step 0 only: later steps start from different states. Vertices whose
    # advection update hinges on the 1e-8 thresholds within rounding are
    # excluded and counted.
    g = _load(name)
    a, b, amb = g["phi_fp32"][0], g["phi_fp64"][0], g["amb_fp64"][0]
    fin = np.abs(b) < 1e8
    assert np.array_equal(fin, np.abs(a) < 1e8)
    scale = max(1.0, float(np.max(np.abs(b[fin]))))
    ok = fin & ~amb
    assert float(np.max(np.abs(a[ok] - b[ok]))) <= 1e-5 * scale
    assert amb.sum() <= max(1, 0.05 * amb.size)


def test_golden_pyramid_double_root():
    # Vertex 6 of the pyramid is solved from the edge between vertices 4 and 5,
    # which both hold 0: the edge quadratic has a double root (discriminant
    # exactly 0), its foot is the edge midpoint and the value is
    # |(0,1,1) - (100,.5,.5)| = hypot(100, sqrt(.5)) before advection. Both
    # precisions must find it rather than fall back to the vertex candidate
    # hypot(100, 1).
    g = _load("pyramid")
    for prec in (32, 64):
        v = float(g[f"phi_fp{prec}"][0][6])
        assert abs(v + np.hypot(100.0, np.sqrt(0.5))) < 1e-4, prec


def test_torus_redistance_bounds(oracle_mod):
    # phi0 = x - midX on the reference's torus mesh. Redistancing (dt = 0)
    # yields distances through the mesh, which can only exceed the straight
    # distance |phi0| to the plane; vertices of elements the plane cuts are
    # seeds and keep |phi0|; signs are kept.
    case = make_golden.CASES["torus"]()
    for prec in (32, 64):
        o = oracle_mod.Oracle(3, case["xyz"], case["cells"], case["labels"], case["nparts"], prec)
        o.set_phi(case["phi0"])
        o.set_velocity(case["vel"])
        o.step(oracle_mod.make_params(0.0, 16.0, 5))
        r, p0 = o.phi(), case["phi0"]
        o.close()
        assert np.all(np.abs(r) >= np.abs(p0) * (1 - 1e-6) - 1e-6)
        assert np.array_equal(np.sign(r), np.sign(p0))
        near = np.abs(p0) < 0.5
        assert near.sum() > 20
        np.testing.assert_allclose(r[near], p0[near], rtol=1e-6, atol=1e-6)
