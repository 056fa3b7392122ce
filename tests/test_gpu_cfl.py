"""CFL bound on the GPU (ls_compute_dt, ls_set_cfl_coefficients and steps
with cfl = 1) against the oracle's restatement of kernel_compute_timestep2d
(src/core2d/include/meshFIM2d_kernels.h:405-475): the same FP64 formulas
over the same stored local coordinates and velocities, so the bound, the
timestep it gives and the advected phi agree bit for bit."""
import numpy as np
import pytest

import levelset_amd.cases as cases
import levelset_amd.levelset as lsb

pytestmark = pytest.mark.gpu


def _pair(case, oracle_mod, precision, execution=lsb.LS_EXEC_PERSISTENT):
    g = lsb.solver_for_case(case, precision=precision, execution=execution)
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts,
                          64 if precision == lsb.LS_F64 or case.dim == 2 else 32)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    return g, o


@pytest.mark.parametrize("which,precision", [("tet", lsb.LS_F64), ("tet", lsb.LS_F32), ("tri", lsb.LS_F64)])
def test_compute_dt_bit_exact_with_oracle(oracle_mod, which, precision):
    case = cases.small3d(n=16, shape="torus", block=(4, 4, 4)) if which == "tet" else cases.config1()
    g, o = _pair(case, oracle_mod, precision)
    rng = np.random.default_rng(3)
    E = case.cells.shape[0]
    coeff = (rng.uniform(-12, 12, E), rng.uniform(-2, 2, E))
    for k in range(3):
        for c in ((None, None), coeff):
            g.set_cfl_coefficients(*c)
            o.set_cfl_coefficients(*c)
            got, want = g.compute_dt(), o.compute_dt()
            print(f"{which} step {k}: bound {got[0]:.9e} max R {got[1]:.9e}")
            assert got == want
        g.step(lsb.step_params(case.dt, case.bandwidth, case.niter))
        o.step(oracle_mod.make_params(case.dt, case.bandwidth, case.niter))
    g.close()
    o.close()


@pytest.mark.parametrize("execution", [lsb.LS_EXEC_PERSISTENT, lsb.LS_EXEC_GRAPH, lsb.LS_EXEC_KERNELS])
def test_cfl_steps_bit_exact_with_oracle(oracle_mod, execution):
    """A timestep far beyond the bound: with cfl = 1 every step advects with
    the new narrowband's bound instead."""
    case = cases.small3d(n=20, shape="torus", block=(8, 8, 4))
    g, o = _pair(case, oracle_mod, lsb.LS_F64, execution)
    big = 20 * case.dt
    for k in range(4):
        gs = g.step(lsb.step_params(big, case.bandwidth, case.niter, cfl=1))
        os_ = o.step(oracle_mod.make_params(big, case.bandwidth, case.niter, cfl=1))
        print(f"step {k}: dt used {gs.dt_used:.6e} (bound {gs.cfl_dt:.6e}, max R {gs.cfl_max_rin:.6e})")
        assert gs.exec_mode == execution
        assert gs.dt_used == os_.dt_used == gs.cfl_dt < big
        assert (gs.cfl_dt, gs.cfl_max_rin) == g.compute_dt()
        np.testing.assert_array_equal(g.narrowband(), o.narrowband())
        np.testing.assert_array_equal(g.phi(), o.phi(), err_msg=f"phi step {k}")
    g.close()
    o.close()


def test_cfl_with_per_mesh_timesteps_is_refused():
    b = cases.small_batch(n_meshes=3)
    s = lsb.solver_for_batch(b, execution=lsb.LS_EXEC_PERSISTENT)
    with pytest.raises(lsb.LevelSetError, match="per mesh"):
        s.step(lsb.step_params(1.0, 1.0, 5, cfl=1))
    s.close()
