"""GPU mean-curvature flow (LS_MODE_MCF) against the oracle and the closed
form. FP32 storage and the FP64 engines are bit-exact against the oracle
instance of the same storage type; FP32 stays within 1e-5 of the FP64
oracle step by step; a shrinking sphere follows R(t)^2 = R0^2 - 4t."""
import numpy as np
import pytest

import levelset_amd.cases as cases
import levelset_amd.levelset as lsb

pytestmark = pytest.mark.gpu


def _pair(case, oracle_mod, engine, precision, execution=lsb.LS_EXEC_PERSISTENT):
    g = lsb.solver_for_case(case, precision=engine, execution=execution)
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts, precision)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    return g, o


@pytest.mark.parametrize("execution", [lsb.LS_EXEC_PERSISTENT, lsb.LS_EXEC_KERNELS, lsb.LS_EXEC_GRAPH])
def test_mcf_fp32_bit_exact(execution, oracle_mod):
    case = cases.mcf_sphere(n=16, block=(4, 4, 4), seed=7)
    g, o = _pair(case, oracle_mod, lsb.LS_F32, 32, execution)
    gp = lsb.step_params(case.dt, case.bandwidth, case.niter, mode=lsb.LS_MODE_MCF)
    op = oracle_mod.make_params(case.dt, case.bandwidth, case.niter, mode=1)
    for k in range(6):
        gs, os_ = g.step(gp), o.step(op)
        print(f"mcf step {k}: nb {gs.nb_count} iters {gs.fim_iters}")
        np.testing.assert_array_equal(g.narrowband(), o.narrowband(), err_msg=f"step {k}")
        assert gs.fim_iters == os_.fim_iters
        np.testing.assert_array_equal(g.phi().astype(np.float32), o.phi().astype(np.float32), err_msg=f"step {k}")
    g.close()
    o.close()


def test_mcf_fp64_engine_and_2d_bit_exact(oracle_mod):
    for case, engine in ((cases.mcf_sphere(n=14, block=(4, 4, 4), seed=3), lsb.LS_F64),
                         (cases.mcf_circle(n=33), lsb.LS_F64)):
        g, o = _pair(case, oracle_mod, engine, 64)
        gp = lsb.step_params(case.dt, case.bandwidth, case.niter, mode=lsb.LS_MODE_MCF)
        op = oracle_mod.make_params(case.dt, case.bandwidth, case.niter, mode=1)
        for k in range(5):
            g.step(gp)
            o.step(op)
            np.testing.assert_array_equal(g.narrowband(), o.narrowband(), err_msg=f"{case.name} step {k}")
            np.testing.assert_array_equal(g.phi(), o.phi(), err_msg=f"{case.name} step {k}")
        g.close()
        o.close()


def test_mcf_fp32_vs_fp64_oracle_per_step(oracle_mod):
    case = cases.mcf_sphere(n=20, block=(4, 4, 4), seed=5)
    g = lsb.solver_for_case(case)
    o = oracle_mod.Oracle(3, case.xyz, case.cells, case.labels, case.nparts, 64)
    o.set_velocity(case.vel)
    gp = lsb.step_params(case.dt, case.bandwidth, case.niter, mode=lsb.LS_MODE_MCF)
    op = oracle_mod.make_params(case.dt, case.bandwidth, case.niter, mode=1)
    phi, nb = case.phi0, None
    for k in range(5):
        o.set_phi(phi)
        if nb is not None:
            o.set_narrowband(nb)
        g.step(gp)
        o.step(op)
        gphi, ophi = g.phi(), o.phi()
        fin = np.abs(ophi) < 1e8
        err = float(np.max(np.abs(gphi - ophi)[fin]))
        print(f"mcf step {k}: Linf vs FP64 {err:.2e}")
        assert err <= 1e-5
        phi, nb = gphi, g.narrowband()
    g.close()
    o.close()


def test_mcf_sphere_closed_form_on_gpu():
    case = cases.mcf_sphere(n=40, steps=60)
    g = lsb.solver_for_case(case)
    gp = lsb.step_params(case.dt, case.bandwidth, case.niter, mode=lsb.LS_MODE_MCF)
    r0, h = case.meta["r0"], case.h
    for k in range(case.steps):
        g.step(gp)
        if k % 10 == 9:
            t = (k + 1) * case.dt
            got = cases.interface_radius(case.xyz, g.phi(), (0.5, 0.5, 0.5), h, 3)
            exact = np.sqrt(r0 * r0 - 4.0 * t)
            print(f"step {k}: R {got:.5f} exact {exact:.5f} ({(got - exact) / h:+.3f} h)")
            assert abs(got - exact) <= 0.1 * h
    g.close()
