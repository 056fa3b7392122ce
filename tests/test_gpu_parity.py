"""GPU parity: the HIP path through the C-ABI against the oracle.

* FP32 tets: bit-exact against the oracle's FP32 instance (same operation
  order, FP64 discriminant) -- phi, narrowband membership and FIM schedule
  statistics; within 1e-5 L-inf of the FP64 oracle.
* FP64 triangles: bit-exact against the FP64 oracle.
"""
import numpy as np
import pytest

import levelset_amd.cases as cases
import levelset_amd.levelset as lsb

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-5  # L-inf tolerance on phi of the FP32 GPU path against the FP64 oracle


def _run_pair(case, steps, oracle_mod, precision):
    g = lsb.solver_for_case(case)
    o = oracle_mod.Oracle(case.dim, case.xyz, case.cells, case.labels, case.nparts, precision)
    o.set_phi(case.phi0)
    o.set_velocity(case.vel)
    gp = lsb.step_params(case.dt, case.bandwidth, case.niter)
    op = oracle_mod.make_params(case.dt, case.bandwidth, case.niter)
    out = []
    for k in range(steps):
        gs = g.step(gp)
        os_ = o.step(op)
        out.append((k, gs, os_, g.phi(), o.phi(), g.narrowband(), o.narrowband()))
        if os_.nb_count == 0:
            break
    g.close()
    o.close()
    return out


@pytest.mark.parametrize("shape,n,block", [("sphere", 12, (4, 4, 4)), ("torus", 20, (8, 8, 4)),
                                           ("plane", 16, (4, 4, 4))])
def test_tet_fp32_bit_exact(shape, n, block, oracle_mod):
    case = cases.small3d(n=n, shape=shape, block=block, steps=4)
    for k, gs, os_, gphi, ophi, gnb, onb in _run_pair(case, 4, oracle_mod, 32):
        print(f"{shape} step {k}: gpu iters {gs.fim_iters} visits {gs.patch_visits} nb {gs.nb_count}")
        np.testing.assert_array_equal(gnb, onb, err_msg=f"narrowband step {k}")
        assert (gs.fim_iters, gs.patch_visits, gs.probes, gs.nb_verts) == \
               (os_.fim_iters, os_.patch_visits, os_.probes, os_.nb_verts), f"schedule step {k}"
        np.testing.assert_array_equal(gphi.astype(np.float32), ophi.astype(np.float32),
                                      err_msg=f"phi step {k}")


@pytest.mark.parametrize("shape,n,block", [("sphere", 12, (4, 4, 4)), ("torus", 20, (8, 8, 4))])
def test_tet_fp32_vs_fp64_oracle(shape, n, block, oracle_mod):
    case = cases.small3d(n=n, shape=shape, block=block, steps=3)
    for k, gs, os_, gphi, ophi, gnb, onb in _run_pair(case, 3, oracle_mod, 64):
        np.testing.assert_array_equal(gnb, onb, err_msg=f"narrowband step {k}")
        finite = np.abs(ophi) < 1e8
        np.testing.assert_array_equal(np.abs(gphi) < 1e8, finite)
        err = np.max(np.abs(gphi[finite] - ophi[finite]))
        print(f"{shape} step {k}: Linf vs FP64 = {err:.3e}")
        assert err <= TOL_FP64


def test_tri_fp64_bit_exact_config1(oracle_mod):
    case = cases.config1()
    res = _run_pair(case, case.steps, oracle_mod, 64)
    assert len(res) == case.steps
    for k, gs, os_, gphi, ophi, gnb, onb in res:
        np.testing.assert_array_equal(gnb, onb)
        assert (gs.fim_iters, gs.patch_visits, gs.probes) == (os_.fim_iters, os_.patch_visits, os_.probes)
        np.testing.assert_array_equal(gphi, ophi, err_msg=f"phi step {k}")
