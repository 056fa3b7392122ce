"""The network oracle (oracle/mlp_oracle.py) against what fixes it: central
finite differences of the FP64 loss for every parameter's analytic gradient,
a closed-form feature check, and the local frame against the layout's."""
import numpy as np

import levelset_amd.cases as cases


def test_analytic_gradient_matches_finite_differences(mlp_oracle):
    M = mlp_oracle
    rng = np.random.default_rng(1)
    p = M.init_params(3).astype(np.float64)
    p[M.N_PARAMS - 4:] = rng.normal(0, 0.1, 4)       # nonzero b3
    p[22 * 64:22 * 64 + 64] = rng.normal(0, 0.1, 64)  # nonzero b1
    X = rng.normal(size=(9, 22))
    Y = rng.normal(size=(9, 4))
    loss, g = M.loss_and_grad(p, X, Y)
    assert abs(loss / 9 - M.mean_loss(p, X, Y)) < 1e-12
    fd = M.finite_difference(p, X, Y, range(M.N_PARAMS), h=1e-6)
    err = np.abs(fd - g)
    assert np.max(err) < 1e-6 * max(1.0, np.max(np.abs(g))), np.argmax(err)


def test_features_closed_form(mlp_oracle):
    L = np.array([[2.0, 0.5, 1.5, 0.25, 0.75, 1.25]])
    ph = np.array([[0.2, -0.4, 0.6, -0.8]])
    X, s = mlp_oracle.features(ph, L)
    assert s[0] == 2.0
    A = np.zeros(3)
    B, C, D = np.array([2.0, 0, 0]), np.array([0.5, 1.5, 0]), np.array([0.25, 0.75, 1.25])
    want = np.concatenate([ph[0], B - A, C - A, D - A, C - B, D - B, D - C]) / 2.0
    np.testing.assert_allclose(X[0], want, rtol=0, atol=1e-15)


def test_bf16_rounding_is_nearest_even(mlp_oracle):
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -9])
    np.testing.assert_array_equal(mlp_oracle.bf16(x), [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 1.0])


def test_local_frame_matches_layout(oracle_mod, mlp_oracle):
    case = cases.small3d(n=6, block=(3, 3, 3))
    o = oracle_mod.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
    lay = o.layout()
    L = o.coords()
    o.close()
    mine = mlp_oracle.local_coords(case.xyz, case.cells[lay["copy_src"]])
    np.testing.assert_allclose(mine, L, rtol=1e-12, atol=1e-14)


def test_triangle_features_closed_form(mlp_oracle):
    """A triangle copy fills the tet layout: its three corners' phi, edges AB,
    AC, BC, everything else zero, all divided by |AB|."""
    c = np.array([[2.0, 0.5, 1.5]])
    ph = np.array([[0.2, -0.4, 0.6]])
    X, s = mlp_oracle.features_tri(ph, c)
    assert s[0] == 2.0
    B, C = np.array([2.0, 0, 0]), np.array([0.5, 1.5, 0])
    z3 = np.zeros(3)
    want = np.concatenate([ph[0], [0.0], B, C, z3, C - B, z3, z3]) / 2.0
    np.testing.assert_allclose(X[0], want, rtol=0, atol=1e-15)


def test_triangle_local_frame_matches_layout(oracle_mod, mlp_oracle):
    case = cases.config1()
    o = oracle_mod.Oracle(2, case.xyz, case.cells, case.labels, case.nparts, 64)
    lay = o.layout()
    L = o.coords()
    o.close()
    mine = mlp_oracle.local_coords_tri(case.xyz, case.cells[lay["copy_src"]])
    np.testing.assert_allclose(mine, L, rtol=1e-12, atol=1e-14)
