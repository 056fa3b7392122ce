"""Patch layout: the product's builder against the oracle's independent
restatement (stable sorts over SoA arrays) and against brute-force
definitions of each ordering. Runs on the CPU (host-only context)."""
import numpy as np
import pytest

import levelset_amd.cases as cases
import levelset_amd.levelset as lsb


def _host_solver(dim, xyz, cells):
    s = lsb.Solver(device=-1)
    s.set_mesh(dim, xyz, cells)
    return s


def _small_cases():
    xyz, cells, h = cases.kuhn_grid(9)
    xyz = cases.jitter_interior(xyz, h, 0.1, 3, 3)
    lab, P = cases.partition_blocks(9, (4, 4, 4), 3)
    yield 3, xyz, cells, lab, P
    lab2 = cases.partition_rcb(xyz, 7)
    yield 3, xyz, cells, lab2, 7
    c1 = cases.config1()
    yield 2, c1.xyz, c1.cells, c1.labels, c1.nparts


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_layout_matches_oracle(idx, oracle_mod):
    dim, xyz, cells, lab, P = list(_small_cases())[idx]
    s = _host_solver(dim, xyz, cells)
    s.partition_labels(lab, P)
    got = s.layout()
    o = oracle_mod.Oracle(dim, xyz, cells, lab, P, 64)
    ref = o.layout()
    for k in ["vert_perm", "copy_src", "vert_off", "ele_off", "copy_verts"]:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_layout_brute_force_definitions():
    dim, xyz, cells, lab, P = list(_small_cases())[1]
    s = _host_solver(dim, xyz, cells)
    s.partition_labels(lab, P)
    L = s.layout()
    # Vertex renumbering: patches contiguous, original ids ascending inside a patch.
    vp = L["vert_perm"]
    assert sorted(vp.tolist()) == list(range(xyz.shape[0]))
    plab = lab[vp]
    assert np.all(np.diff(plab) >= 0)
    for p in range(P):
        seg = vp[L["vert_off"][p]:L["vert_off"][p + 1]]
        assert np.all(np.diff(seg) > 0)
    # Element copies: one per distinct patch of an element's vertices.
    ncopy = np.array([len(set(lab[c].tolist())) for c in cells])
    assert L["copy_src"].size == ncopy.sum()
    iperm = np.empty_like(vp)
    iperm[vp] = np.arange(vp.size)
    for p in range(P):
        es, ee = L["ele_off"][p], L["ele_off"][p + 1]
        src = L["copy_src"][es:ee]
        expect = np.nonzero([p in set(lab[c].tolist()) for c in cells])[0]
        np.testing.assert_array_equal(src, expect)
        np.testing.assert_array_equal(L["copy_verts"][es:ee], iperm[cells[src]])


def test_rcb_partition_matches_python_restatement():
    xyz, cells, h = cases.kuhn_grid(8)
    xyz = cases.jitter_interior(xyz, h, 0.1, 4, 3)
    s = _host_solver(3, xyz, cells)
    s.partition_rcb(11)
    L = s.layout()
    lab_py = cases.partition_rcb(xyz, 11)
    # Same labels => same renumbering.
    o_order = np.lexsort((np.arange(xyz.shape[0]), lab_py))
    np.testing.assert_array_equal(L["vert_perm"], o_order)
    sizes = np.bincount(lab_py, minlength=11)
    assert sizes.max() - sizes.min() <= 1


def test_grid_partition_matches_block_rule():
    xyz, cells, h = cases.kuhn_grid(10)
    s = _host_solver(3, xyz, cells)
    s.partition_grid(10, 10, 10, 4, 4, 3)
    L = s.layout()
    lab, P = cases.partition_blocks(10, (4, 4, 3), 3)
    assert L["vert_off"].size == P + 1
    np.testing.assert_array_equal(L["vert_perm"], np.lexsort((np.arange(lab.size), lab)))


def test_c_abi_exports_every_declared_symbol():
    import ctypes
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "levelset_hip.h")).read()
    names = re.findall(r"^\s*(?:int|void|const char\*)\s+(ls_\w+)\s*\(", hdr, re.M)
    assert len(names) >= 14
    lib = ctypes.CDLL(lsb.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(lsb.EXPORTS) <= set(names)


def test_errors_are_returned_not_raised_across_abi():
    s = lsb.Solver(device=-1)
    with pytest.raises(lsb.LevelSetError, match="set the mesh first"):
        s.partition_rcb(4)
    xyz, cells, h = cases.kuhn_grid(4)
    s.set_mesh(3, xyz, cells)
    with pytest.raises(lsb.LevelSetError, match="out of range"):
        s.partition_labels(np.full(xyz.shape[0], 9, np.int32), 2)
    vel = np.zeros((cells.shape[0], 3))
    with pytest.raises(lsb.LevelSetError, match="triangle meshes"):
        s.set_velocity(vel, lsb.LS_VEL_PHI_2D)
    with pytest.raises(lsb.LevelSetError, match="unknown velocity mode"):
        s.set_velocity(vel, 7)
    bad = cells.copy()
    bad[0, 0] = 10 ** 6
    with pytest.raises(lsb.LevelSetError, match="out of range"):
        s.set_mesh(3, xyz, bad)


def test_python_structs_match_c_header(tmp_path):
    """ctypes mirrors of ls_step_params / ls_stats / ls_layout_info have the C
    sizes and field offsets of include/levelset_hip.h."""
    import ctypes as C
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    structs = {"ls_step_params": lsb.StepParams, "ls_stats": lsb.Stats, "ls_layout_info": lsb.LayoutInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "levelset_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    subprocess.check_call(["gcc", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)])
    got = {}
    for line in subprocess.check_output([str(exe)], text=True).splitlines():
        cname, field, val = line.split()
        got[(cname, field)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
    assert lsb.MLP_PARAMS == 5892
