"""The drop-in LevelSet class end to end: the Example1/Example2 drivers
(reference CLI, src/examples/example1.cu and example2.cu) run on the
reference's test meshes, and every result<step>.vtk they write is compared
with the oracle run from the same inputs: the mesh as the Python readers load
it, the partition the drop-in chooses (recursive bisection with the
Partition_METIS patch count), and LevelSet's default initial values and
velocities (src/LevelSet.cu:176-231).

Tet meshes run the FP32 engine (compared with the FP32 oracle); triangle
meshes run the FP64 engine (compared with the FP64 oracle). VTK files print
values with 12 decimals; the oracle's values are printed the same way and
the two texts must agree exactly."""
import os
import subprocess

import numpy as np
import pytest

import levelset_amd.cases as cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESHES = os.path.join(ROOT, "tests", "golden", "meshes")
BIN = os.path.join(ROOT, "sci-solver_level-set_amd", "bin")


def _read_vtk(path):
    with open(path) as f:
        toks = f.read().split()
    i = toks.index("POINTS")
    nv = int(toks[i + 1])
    pts = np.array(toks[i + 3:i + 3 + 3 * nv], dtype=np.float64).reshape(nv, 3)
    i = toks.index("CELLS")
    nt, tot = int(toks[i + 1]), int(toks[i + 2])
    raw = np.array(toks[i + 3:i + 3 + tot], dtype=np.int64).reshape(nt, -1)
    i = toks.index("LOOKUP_TABLE")
    vals = np.array(toks[i + 2:i + 2 + nv], dtype=np.float64)
    return pts, raw[:, 1:].astype(np.int32), vals


def _run_example(exe, args, cwd):
    print("running", exe, " ".join(args), flush=True)
    r = subprocess.run([os.path.join(BIN, exe)] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    print(r.stdout + r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    files = sorted((f for f in os.listdir(cwd) if f.startswith("result") and f.endswith(".vtk")),
                   key=lambda f: int(f[6:-4]))
    assert files, r.stdout + r.stderr
    return [_read_vtk(os.path.join(cwd, f)) for f in files], r.stdout


def _oracle_steps(oracle_mod, dim, xyz, cells, phi0, vel, nparts, nsteps, dt, bw, prec):
    labels = cases.partition_rcb(xyz, nparts)
    o = oracle_mod.Oracle(dim, xyz, cells, labels, nparts, prec)
    o.set_phi(phi0.astype(np.float64))
    o.set_velocity(vel)
    prm = oracle_mod.make_params(dt, bw, 5 if dim == 3 else 10)
    out = [phi0.astype(np.float64)]
    for _ in range(nsteps):
        st = o.step(prm)
        if st.nb_count == 0:
            break
        out.append(o.phi())
    o.close()
    return out


def _printed(a):
    """Values as the VTK writer prints float data: %.12f of the float32 value."""
    return np.array([float("%.12f" % v) for v in np.asarray(a, dtype=np.float32).astype(np.float64).ravel()])


def _compare(results, expected):
    assert len(results) == len(expected)
    for k, ((_, _, vals), ref) in enumerate(zip(results, expected)):
        exp = _printed(ref)
        print(f"step {k}: max |vtk - oracle| = {float(np.max(np.abs(vals - exp))):.2e}")
        np.testing.assert_array_equal(vals, exp, err_msg=f"step {k}")


@pytest.mark.parametrize("mesh,steps,metis", [("pyramid", 2, 16), ("sphere334", 4, 16), ("torus", 3, 32)])
def test_example1_tet_default_x(mesh, steps, metis, oracle_mod, tmp_path):
    res, _ = _run_example("Example1", ["-i", os.path.join(MESHES, mesh + ".node"), "-n", str(steps),
                                       "-m", str(metis)], str(tmp_path))
    xyz, tets = cases.read_tetgen(os.path.join(MESHES, mesh))
    tets = cases.reorient_tets(xyz, tets)
    pts, cells, _ = res[0]
    np.testing.assert_array_equal(pts.ravel(), _printed(xyz))
    np.testing.assert_array_equal(cells, tets)
    phi0, vel = cases.levelset_default_x(xyz, tets.shape[0], 3)
    exp = _oracle_steps(oracle_mod, 3, xyz, tets, phi0, vel, cases.metis_parts(xyz.shape[0], metis, 3),
                        steps, 1.0, 16.0, 32)
    _compare(res, exp)


def test_example1_tet_grid_partition(oracle_mod, tmp_path):
    # -p 1: GraphPartition_Square over an nside^3 vertex grid; a Kuhn grid
    # written as a tetgen mesh stands in for a structured input file.
    n, b = 8, 4
    xyz, tets, _ = cases.kuhn_grid(n)
    xyz = (xyz * 10.0).astype(np.float32).astype(np.float64)
    base = str(tmp_path / "grid")
    with open(base + ".node", "w") as f:
        f.write(f"{n ** 3} 3 0 0\n")
        for i, p in enumerate(xyz):
            f.write(f"{i} {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
    with open(base + ".ele", "w") as f:
        f.write(f"{tets.shape[0]} 4 0\n")
        for i, t in enumerate(tets):
            f.write(f"{i} {t[0]} {t[1]} {t[2]} {t[3]}\n")
    run = tmp_path / "run"
    run.mkdir()
    res, _ = _run_example("Example1", ["-i", base, "-n", "3", "-p", "1", "-d", str(n), "-b", str(b),
                                       "-w", "4"], str(run))
    labels, nparts = cases.partition_blocks(n, (b, b, b), 3)
    phi0, vel = cases.levelset_default_x(xyz, tets.shape[0], 3)
    o = oracle_mod.Oracle(3, xyz, tets, labels, nparts, 32)
    o.set_phi(phi0.astype(np.float64))
    o.set_velocity(vel)
    prm = oracle_mod.make_params(1.0, 4.0, 5)
    exp = [phi0.astype(np.float64)]
    for _ in range(3):
        if o.step(prm).nb_count == 0:
            break
        exp.append(o.phi())
    o.close()
    _compare(res, exp)


@pytest.mark.parametrize("mesh,steps,metis", [("SquareMesh_size16.ply", 2, 64), ("sphere_266verts.ply", 2, 16)])
def test_example2_tri_default_x(mesh, steps, metis, oracle_mod, tmp_path):
    res, _ = _run_example("Example2", ["-i", os.path.join(MESHES, mesh), "-n", str(steps), "-m", str(metis)],
                          str(tmp_path))
    xyz, tris = cases.read_ply(os.path.join(MESHES, mesh))
    pts, cells, _ = res[0]
    np.testing.assert_array_equal(cells, tris)
    phi0, vel = cases.levelset_default_x(xyz, tris.shape[0], 2)
    exp = _oracle_steps(oracle_mod, 2, xyz, tris, phi0, vel, cases.metis_parts(xyz.shape[0], metis, 2),
                        steps, 1.0, 16.0, 64)
    _compare(res, exp)


def test_example_zero_narrowband_ends_early(tmp_path):
    # A negative bandwidth admits no patch, so the narrowband is empty after
    # the first redistance and only the initial values are stored
    # (src/core3d/cuda/meshFIM3d.cu:655-659).
    res, out = _run_example("Example1", ["-i", os.path.join(MESHES, "sphere334"), "-n", "5", "-w", "-1"],
                            str(tmp_path))
    assert "Ending at timestep 0" in out
    assert len(res) == 1
