"""The patch layout built on the GPU (radix sorts) against the host builder:
every array identical on grid, bisection and batch partitions of tet and
triangle meshes, identical timesteps from either layout, and the same
errors for invalid partitions."""
import os

import numpy as np
import pytest

import levelset_amd.cases as cases
import levelset_amd.levelset as lsb

pytestmark = pytest.mark.gpu

MESHES = os.path.join(os.path.dirname(__file__), "golden", "meshes")


def _tetgen_torus():
    xyz, tets = cases.read_tetgen(os.path.join(MESHES, "torus"))
    tets = cases.reorient_tets(xyz, tets)
    return xyz, tets


def _layouts(dim, xyz, cells, labels, nparts):
    out = []
    for builder in (lsb.LS_LAYOUT_HOST, lsb.LS_LAYOUT_DEVICE):
        s = lsb.Solver(0)
        s.set_mesh(dim, xyz, cells)
        s.set_layout_builder(builder)
        s.partition_labels(labels, nparts)
        out.append((s.layout_info(), s.layout(), s.patch_lists(), s))
    return out


def _assert_same(host, dev):
    hi, hl, hp, _ = host
    di, dl, dp, _ = dev
    assert hi == di
    for k in hl:
        np.testing.assert_array_equal(dl[k], hl[k], err_msg=k)
    for k in hp:
        np.testing.assert_array_equal(dp[k], hp[k], err_msg=k)


CASES = {
    "torus24_blocks": lambda: (3,) + _case_arrays(cases.small3d(n=24, shape="torus", block=(8, 8, 4), steps=2)),
    "sphere12_rcb7": lambda: (3,) + _rcb(cases.small3d(n=12, shape="sphere", block=(4, 4, 4), steps=2), 7),
    "tetgen_torus_rcb": lambda: (3,) + _rcb_arrays(*_tetgen_torus(), 16),
    "tri17_config1": lambda: (2,) + _case_arrays(cases.config1()),
    "tri33_rcb": lambda: (2,) + _rcb(cases.mcf_circle(n=33, steps=2), 13),
    "one_patch": lambda: (3,) + _rcb(cases.small3d(n=8, shape="sphere", block=(4, 4, 4), steps=2), 1),
}


def _case_arrays(case):
    return case.xyz, case.cells, case.labels, case.nparts


def _rcb(case, nparts):
    return _rcb_arrays(case.xyz, case.cells, nparts)


def _rcb_arrays(xyz, cells, nparts):
    return xyz, cells, cases.partition_rcb(xyz, nparts), nparts


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_layout_matches_host(name):
    dim, xyz, cells, labels, nparts = CASES[name]()
    host, dev = _layouts(dim, xyz, cells, labels, nparts)
    print(f"{name}: {host[0]}")
    _assert_same(host, dev)
    for _, _, _, s in (host, dev):
        s.close()


def test_device_layout_batch_matches_host():
    b = cases.small_batch()
    solvers = []
    for builder in (lsb.LS_LAYOUT_HOST, lsb.LS_LAYOUT_DEVICE):
        s = lsb.Solver(0)
        s.set_mesh_batch(b.dim, [(m.xyz, m.cells) for m in b.meshes])
        s.set_layout_builder(builder)
        s.partition_batch(b.patch_verts)
        solvers.append(s)
    host, dev = [(s.layout_info(), s.layout(), s.patch_lists(), s) for s in solvers]
    _assert_same(host, dev)
    for s in solvers:
        s.close()


def test_steps_identical_from_either_layout():
    case = cases.small3d(n=24, shape="torus", block=(8, 8, 4), steps=3)
    phis = []
    for builder in (lsb.LS_LAYOUT_HOST, lsb.LS_LAYOUT_DEVICE):
        s = lsb.Solver(0)
        s.set_mesh(3, case.xyz, case.cells)
        s.set_layout_builder(builder)
        s.partition_labels(case.labels, case.nparts)
        s.set_phi(case.phi0)
        s.set_velocity(case.vel)
        out = []
        for _ in range(3):
            s.step(lsb.step_params(case.dt, case.bandwidth, case.niter))
            out.append(s.phi())
        s.close()
        phis.append(out)
    for a, b in zip(*phis):
        np.testing.assert_array_equal(a, b)


def test_device_layout_errors_match_host():
    case = cases.small3d(n=16, shape="sphere", block=(4, 4, 4), steps=1)
    msgs = []
    for builder in (lsb.LS_LAYOUT_HOST, lsb.LS_LAYOUT_DEVICE):
        s = lsb.Solver(0)
        s.set_mesh(3, case.xyz, case.cells)
        s.set_layout_builder(builder)
        # 15^3 * 6 = 20250 copies in one patch: more than 65535 corner slots.
        with pytest.raises(RuntimeError) as ei:
            s.partition_labels(np.zeros(case.xyz.shape[0], np.int32), 1)
        msgs.append(str(ei.value))
        # and the context still works with a valid partition
        s.partition_labels(case.labels, case.nparts)
        s.close()
    print(msgs)
    assert msgs[0] == msgs[1]
    assert "65535" in msgs[0]
