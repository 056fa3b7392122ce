"""Batch API on the host (device -1: layout only, no GPU): patch counts,
patches confined to their mesh, rejected inputs."""
import numpy as np
import pytest

import levelset_amd.cases as cases
import levelset_amd.levelset as lsb


def _meshes(sides=(5, 7, 4)):
    out = []
    for n in sides:
        xyz, cells, _ = cases.kuhn_grid(n)
        out.append((xyz, cells))
    return out


def test_partition_batch_confines_patches_to_meshes():
    meshes = _meshes()
    s = lsb.Solver(device=-1)
    s.set_mesh_batch(3, meshes)
    s.partition_batch(40)
    info = s.layout_info()
    assert info["n_patches"] == sum(-(-x.shape[0] // 40) for x, _ in meshes)
    lay = s.layout()
    mesh_of_vertex = np.repeat(np.arange(len(meshes)), [x.shape[0] for x, _ in meshes])
    vo = lay["vert_off"]
    for p in range(info["n_patches"]):
        owners = mesh_of_vertex[lay["vert_perm"][vo[p]:vo[p + 1]]]
        assert np.unique(owners).size == 1
    # every element copy's vertices belong to the copy's own mesh
    mesh_of_cell = np.repeat(np.arange(len(meshes)), [c.shape[0] for _, c in meshes])
    cv = lay["copy_verts"]
    orig = lay["vert_perm"][cv]
    assert np.all(mesh_of_vertex[orig] == mesh_of_cell[lay["copy_src"]][:, None])
    s.close()


def test_batch_partition_matches_single_mesh_rcb():
    meshes = _meshes((6, 5))
    s = lsb.Solver(device=-1)
    s.set_mesh_batch(3, meshes)
    s.partition_batch(32)
    lay = s.layout()
    n0 = meshes[0][0].shape[0]
    p0 = -(-n0 // 32)
    lab = cases.partition_rcb(meshes[0][0], p0)
    vo = lay["vert_off"]
    for p in range(p0):
        ids = np.sort(lay["vert_perm"][vo[p]:vo[p + 1]])
        np.testing.assert_array_equal(ids, np.flatnonzero(lab == p))
    s.close()


def test_batch_rejects_cross_mesh_cells_and_patches():
    meshes = _meshes((4, 4))
    s = lsb.Solver(device=-1)
    bad = [(meshes[0][0], meshes[0][1]), (meshes[1][0], meshes[1][1] - 64)]  # second mesh names the first's ids
    with pytest.raises(lsb.LevelSetError, match="another mesh"):
        s.set_mesh_batch(3, bad)
    s.set_mesh_batch(3, meshes)
    labels = np.zeros(128, np.int32)  # one patch over both meshes
    with pytest.raises(lsb.LevelSetError, match="spans meshes"):
        s.partition_labels(labels, 1)
    with pytest.raises(lsb.LevelSetError):
        s.partition_batch(0)
    s.close()


def test_batch_params_require_a_batch():
    xyz, cells, _ = cases.kuhn_grid(4)
    s = lsb.Solver(device=-1)
    s.set_mesh(3, xyz, cells)
    with pytest.raises(lsb.LevelSetError, match="batch"):
        s.partition_batch(16)
    s.close()


def test_ellipsoid_distance_properties():
    rng = np.random.default_rng(3)
    axes = np.array([0.2, 0.3, 0.1])
    p = rng.uniform(-0.5, 0.5, size=(4000, 3))
    p[0] = 0.0
    d = cases.ellipsoid_distance(p, axes)
    assert np.all(np.isfinite(d))
    assert abs(d[0] + 0.1) < 1e-15                      # centre: the shortest semi-axis
    # closest point on the surface and normal to it
    ap = np.abs(p)
    inside = np.sum((ap / axes) ** 2, axis=1) < 1
    assert np.array_equal(inside, d < 0)
    # a sphere reduces to |p| - r
    np.testing.assert_allclose(cases.ellipsoid_distance(p, (0.3, 0.3, 0.3)), np.linalg.norm(p, axis=1) - 0.3,
                               atol=1e-14)
    # never farther than any sampled surface point
    u = rng.normal(size=(200000, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    surf = u * axes
    for i in range(0, 4000, 97):
        assert abs(d[i]) <= np.min(np.linalg.norm(surf - p[i], axis=1)) + 1e-12


def test_config4_shape():
    b = cases.config4(n_meshes=16, workers=1)
    sides = b.meta["sides"]
    n = sides ** 3
    assert np.all(n >= 0.8 * 2000) and np.all(n <= 1.2 * 200000)
    assert b.n_verts == int(np.sum(n))
    for m in b.meshes:
        assert np.all(np.isfinite(m.phi0)) and m.dt > 0 and m.bandwidth > 0
