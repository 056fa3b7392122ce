"""Synthetic level-set workloads: meshes, initial signed distances, velocities.

Mesh conventions follow the reference's own fixtures:
  * triangle grids split along the (i,j)->(i+1,j+1) diagonal, as the faces
    ``3 1 18 2`` / ``3 1 17 18`` of src/test/test_data/SquareMesh_size16.ply:267-268;
  * tetrahedral grids split into the 6 Kuhn tetrahedra of each cube, vertex index
    k*n*n + j*n + i (x fastest), the ordering GraphPartition_Square assumes
    (src/core3d/cuda/meshFIM3d.cu:316-326).
Velocities are one vector per element, as LevelSet::initializeAdvection takes
them (src/LevelSet.h:35, size checked at src/LevelSet.cu:98).
"""
from __future__ import annotations

import itertools
import math
import os
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Case:
    name: str
    dim: int
    xyz: np.ndarray          # (N, 3) float64
    cells: np.ndarray        # (E, dim+1) int32
    phi0: np.ndarray         # (N,) float64
    vel: np.ndarray          # (E, 3) float64, one vector per element
    dt: float
    steps: int
    bandwidth: float
    h: float
    labels: np.ndarray       # (N,) int32 patch labels
    nparts: int
    mode: int = 0            # 0 advection, 1 mean-curvature flow
    niter: int = 5
    meta: dict = field(default_factory=dict)


def tri_grid(n: int):
    """n x n vertices on [0,1]^2, cells split along the (i,j)->(i+1,j+1) diagonal."""
    h = 1.0 / (n - 1)
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    xyz = np.stack([i.ravel() * h, j.ravel() * h, np.zeros(n * n)], axis=1)
    ci, cj = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    v = (cj * n + ci).ravel()
    t1 = np.stack([v, v + n + 1, v + 1], axis=1)
    t2 = np.stack([v, v + n, v + n + 1], axis=1)
    cells = np.empty((2 * v.size, 3), dtype=np.int32)
    cells[0::2] = t1
    cells[1::2] = t2
    return xyz, cells, h


def kuhn_grid(n: int):
    """n^3 vertices on [0,1]^3, each cube split into 6 Kuhn tetrahedra, positively oriented."""
    h = 1.0 / (n - 1)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    xyz = np.stack([i.ravel() * h, j.ravel() * h, k.ravel() * h], axis=1)
    m = n - 1
    ck, cj, ci = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    base = (ck * n * n + cj * n + ci).ravel().astype(np.int64)
    step = np.array([1, n, n * n], dtype=np.int64)
    tets = []
    for perm in itertools.permutations(range(3)):
        v0 = base
        v1 = v0 + step[perm[0]]
        v2 = v1 + step[perm[1]]
        v3 = base + step.sum()
        t = np.stack([v0, v1, v2, v3], axis=1)
        # Orientation of this permutation's path tet: odd permutations flip sign.
        sgn = _perm_sign(perm)
        if sgn < 0:
            t = t[:, [0, 2, 1, 3]]
        tets.append(t)
    cells = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    return xyz, cells, h


def _perm_sign(p):
    s = 1
    p = list(p)
    for a in range(len(p)):
        for b in range(a + 1, len(p)):
            if p[a] > p[b]:
                s = -s
    return s


def jitter_interior(xyz: np.ndarray, h: float, amp: float, seed: int, dim: int) -> np.ndarray:
    """Move every interior vertex by U(-amp*h, amp*h) per axis (boundary vertices stay)."""
    rng = np.random.default_rng(seed)
    d = rng.uniform(-amp * h, amp * h, size=(xyz.shape[0], 3))
    eps = 1e-12
    interior = np.ones(xyz.shape[0], dtype=bool)
    for a in range(dim):
        interior &= (xyz[:, a] > eps) & (xyz[:, a] < 1.0 - eps)
    d[~interior] = 0.0
    if dim == 2:
        d[:, 2] = 0.0
    return xyz + d


def tet_volumes(xyz, cells):
    a = xyz[cells[:, 1]] - xyz[cells[:, 0]]
    b = xyz[cells[:, 2]] - xyz[cells[:, 0]]
    c = xyz[cells[:, 3]] - xyz[cells[:, 0]]
    return np.einsum("ij,ij->i", np.cross(a, b), c) / 6.0


def sdf_circle(xyz, center=(0.5, 0.5), r=0.3):
    return np.hypot(xyz[:, 0] - center[0], xyz[:, 1] - center[1]) - r


def sdf_sphere(xyz, center=(0.5, 0.5, 0.5), r=0.3):
    return np.linalg.norm(xyz - np.asarray(center)[None, :], axis=1) - r


def sdf_torus(xyz, R=0.3, r=0.1, center=(0.5, 0.5, 0.5)):
    p = xyz - np.asarray(center)[None, :]
    q = np.hypot(p[:, 0], p[:, 1]) - R
    return np.hypot(q, p[:, 2]) - r


def sdf_plane(xyz, normal, offset):
    nrm = np.asarray(normal, dtype=np.float64)
    nrm = nrm / np.linalg.norm(nrm)
    return xyz @ nrm - offset


def sdf_spheres_union(xyz, centers, radii):
    out = np.full(xyz.shape[0], np.inf)
    for c, r in zip(centers, radii):
        out = np.minimum(out, np.linalg.norm(xyz - c[None, :], axis=1) - r)
    return out


def ellipsoid_distance(p, axes, bisect=16, newton=6):
    """Signed distance from points p (N,3), in the ellipsoid's own frame, to the
    ellipsoid x^2/a^2 + y^2/b^2 + z^2/c^2 = 1 (negative inside), in FP64.

    The closest point is x_k = a_k^2 p_k / (a_k^2 + t) with t the root of
    f(t) = sum (a_k p_k / (a_k^2 + t))^2 - 1 on t > -min(a)^2, where f is
    decreasing and convex: bisection narrows the bracket, then Newton from
    its left end (f > 0) converges monotonically. Interior points with no
    root there (on the plane of the shortest axis, e.g. the centre) take the
    degenerate closest point x_k = a_k^2 p_k / (a_k^2 - a_min^2) (k not the
    shortest axis), x_min = a_min sqrt(1 - sum (x_k / a_k)^2)."""
    a = np.abs(np.asarray(axes, dtype=np.float64))
    ap = np.abs(np.asarray(p, dtype=np.float64))
    kmin = int(np.argmin(a))
    amin2 = float(a[kmin] ** 2)
    inside = np.sum((ap / a[None, :]) ** 2, axis=1) < 1.0
    a2 = (a * a)[None, :]

    def f(t):
        return np.sum((a[None, :] * ap / (a2 + t[:, None])) ** 2, axis=1) - 1.0

    def df(t):
        return -2.0 * np.sum((a[None, :] * ap) ** 2 / (a2 + t[:, None]) ** 3, axis=1)

    eps = 1e-12 * amin2
    lo = np.where(inside, -amin2 + eps, 0.0)
    hi = np.where(inside, 0.0, np.maximum(np.linalg.norm(ap * a[None, :], axis=1), 1e-300))
    degenerate = inside & ~(f(lo) > 0.0)
    for _ in range(bisect):
        mid = 0.5 * (lo + hi)
        pos = f(mid) > 0.0
        lo = np.where(pos, mid, lo)
        hi = np.where(pos, hi, mid)
    t = lo
    for _ in range(newton):
        d = df(t)
        step = np.where(d < 0.0, -f(t) / np.where(d < 0.0, d, -1.0), 0.0)
        t = np.minimum(t + np.maximum(step, 0.0), hi)
    x = a2 * ap / (a2 + t[:, None])
    if degenerate.any():
        shortest = (a2 - amin2)[0] <= 0.0  # every axis as short as the shortest
        den = np.where(shortest, 1.0, a2[0] - amin2)[None, :]
        xd = np.where(shortest[None, :], 0.0, a2 * ap[degenerate] / den)
        rest = np.sum((xd / a[None, :]) ** 2, axis=1)
        xd[:, kmin] = a[kmin] * np.sqrt(np.maximum(0.0, 1.0 - rest))
        x[degenerate] = xd
    dist = np.linalg.norm(x - ap, axis=1)
    return np.where(inside, -dist, dist)


def centroids(xyz, cells):
    return xyz[cells].mean(axis=1)


def velocity_rotation_z(xyz, cells, omega=2 * math.pi, center=(0.5, 0.5)):
    c = centroids(xyz, cells)
    v = np.zeros((cells.shape[0], 3))
    v[:, 0] = -omega * (c[:, 1] - center[1])
    v[:, 1] = omega * (c[:, 0] - center[0])
    return v


def partition_blocks(n: int, block, dim: int) -> tuple[np.ndarray, int]:
    """Grid block partition (GraphPartition_Square, src/core3d/cuda/meshFIM3d.cu:269-352)."""
    if dim == 2:
        bl, bw = block
        nbl, nbw = -(-n // bl), -(-n // bw)
        j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        lab = (j // bw) * nbl + (i // bl)
        return lab.ravel().astype(np.int32), nbl * nbw
    bl, bw, bh = block
    nbl, nbw, nbh = -(-n // bl), -(-n // bw), -(-n // bh)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    lab = (k // bh) * nbl * nbw + (j // bw) * nbl + (i // bl)
    return lab.ravel().astype(np.int32), nbl * nbw * nbh


def partition_rcb(xyz: np.ndarray, nparts: int) -> np.ndarray:
    """Recursive coordinate bisection, the same rule as the C++ partitioner:
    split along the longest bounding-box axis at n*floor(k/2)/k, ties by id."""
    labels = np.zeros(xyz.shape[0], dtype=np.int32)

    def split(ids, k, base):
        if k <= 1 or ids.size <= 1:
            labels[ids] = base
            return
        pts = xyz[ids]
        ext = pts.max(axis=0) - pts.min(axis=0)
        axis = int(np.argmax(ext))
        kl = k // 2
        m = (ids.size * kl) // k
        order = np.lexsort((ids, pts[:, axis]))
        split(ids[order[:m]], kl, base)
        split(ids[order[m:]], k - kl, base + kl)

    split(np.arange(xyz.shape[0]), nparts, 0)
    return labels


def config1(n: int = 17, seed: int = 1) -> Case:
    """2D circle: 17x17 tri grid, jitter 0.1h, circle r=0.3, sigma=(0.5,0), CFL 0.5,
    10 steps, band 4h, ~64-vertex patches (P = N // 64, src/core2d/cuda/meshFIM2d.cu:380-382)."""
    xyz, cells, h = tri_grid(n)
    xyz = jitter_interior(xyz, h, 0.1, seed, 2)
    phi0 = sdf_circle(xyz)
    vel = np.zeros((cells.shape[0], 3))
    vel[:, 0] = 0.5
    P = max(2, xyz.shape[0] // 64)
    labels = partition_rcb(xyz, P)
    return Case("2d_circle_17", 2, xyz, cells, phi0, vel, dt=0.5 * h / 0.5, steps=10,
                bandwidth=4 * h, h=h, labels=labels, nparts=P, niter=10)


def config2(n: int = 64, seed: int = 2, block=(8, 8, 4), steps: int = 100) -> Case:
    """3D torus revolve: n^3 Kuhn grid, jitter 0.1h, torus R=0.3 r=0.1, rigid rotation
    about z (omega = 2 pi) at element centroids, CFL 0.5, band 6h, 256-vertex patches."""
    xyz, cells, h = kuhn_grid(n)
    xyz = jitter_interior(xyz, h, 0.1, seed, 3)
    phi0 = sdf_torus(xyz)
    vel = velocity_rotation_z(xyz, cells)
    vmax = float(np.max(np.linalg.norm(vel, axis=1)))
    labels, P = partition_blocks(n, block, 3)
    return Case(f"3d_torus_revolve_{n}", 3, xyz, cells, phi0, vel, dt=0.5 * h / vmax,
                steps=steps, bandwidth=6 * h, h=h, labels=labels, nparts=P, niter=5,
                meta={"vmax": vmax, "block": block})


def small3d(n: int = 12, seed: int = 5, block=(4, 4, 4), shape: str = "sphere", steps: int = 3) -> Case:
    """A small tet case for parity tests (sphere or plane, rotation velocity)."""
    xyz, cells, h = kuhn_grid(n)
    xyz = jitter_interior(xyz, h, 0.1, seed, 3)
    if shape == "sphere":
        phi0 = sdf_sphere(xyz, r=0.3)
    elif shape == "torus":
        phi0 = sdf_torus(xyz)
    else:
        phi0 = sdf_plane(xyz, (1.0, 0.5, 0.25), 0.45)
    vel = velocity_rotation_z(xyz, cells)
    vmax = float(np.max(np.linalg.norm(vel, axis=1)))
    labels, P = partition_blocks(n, block, 3)
    return Case(f"3d_{shape}_{n}", 3, xyz, cells, phi0, vel, dt=0.5 * h / vmax, steps=steps,
                bandwidth=6 * h, h=h, labels=labels, nparts=P, niter=5)


# Explicit mean-curvature flow with P1 elements and a lumped mass is stable
# for dt up to about h^2/6 on Kuhn grids (the largest eigenvalue of the
# lumped P1 Laplacian there is 12/h^2); 0.5 h^2 diverges within ten steps.
MCF_DT_FACTOR = 0.125


def mcf_sphere(n: int = 24, r0: float = 0.3, block=(8, 8, 8), steps: int = 24, seed: int = 0) -> Case:
    """A shrinking sphere under mean-curvature flow (closed form R(t)^2 = R0^2 - 4t)."""
    xyz, cells, h = kuhn_grid(n)
    if seed:
        xyz = jitter_interior(xyz, h, 0.05, seed, 3)
    labels, P = partition_blocks(n, block, 3)
    return Case(f"3d_mcf_sphere_{n}", 3, xyz, cells, sdf_sphere(xyz, r=r0), np.zeros((cells.shape[0], 3)),
                dt=MCF_DT_FACTOR * h * h, steps=steps, bandwidth=8 * h, h=h, labels=labels, nparts=P,
                mode=1, niter=5, meta={"r0": r0})


def mcf_circle(n: int = 33, r0: float = 0.3, block=(8, 8), steps: int = 40) -> Case:
    """A shrinking circle under curve-shortening flow (R(t)^2 = R0^2 - 2t)."""
    xyz, cells, h = tri_grid(n)
    labels, P = partition_blocks(n, block, 2)
    return Case(f"2d_mcf_circle_{n}", 2, xyz, cells, sdf_circle(xyz, r=r0), np.zeros((cells.shape[0], 3)),
                dt=MCF_DT_FACTOR * h * h, steps=steps, bandwidth=8 * h, h=h, labels=labels, nparts=P,
                mode=1, niter=10, meta={"r0": r0})


def config3(n: int = 192, seed: int = 3, block=(8, 8, 4), steps: int = 50) -> Case:
    """Mean-curvature flow of 8 spheres: n^3 Kuhn grid, jitter 0.05h, centres
    U[0.25,0.75]^3, radii U[0.08,0.15], phi0 = min of the sphere distances,
    band 8h, dt = h^2/8 (explicitly stable), 256-vertex block patches."""
    xyz, cells, h = kuhn_grid(n)
    xyz = jitter_interior(xyz, h, 0.05, seed, 3)
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.25, 0.75, size=(8, 3))
    radii = rng.uniform(0.08, 0.15, size=8)
    phi0 = sdf_spheres_union(xyz, centres, radii)
    labels, P = partition_blocks(n, block, 3)
    return Case(f"3d_mcf_spheres_{n}", 3, xyz, cells, phi0, np.zeros((cells.shape[0], 3)),
                dt=MCF_DT_FACTOR * h * h, steps=steps, bandwidth=8 * h, h=h, labels=labels, nparts=P,
                mode=1, niter=5, meta={"centres": centres, "radii": radii, "block": block})


def interface_radius(xyz, phi, centre, h, dim):
    """Mean radius of the zero level set of a distance-like phi about centre,
    from the vertices within h of it (radius = |x - c| - phi)."""
    r = np.linalg.norm(xyz[:, :dim] - np.asarray(centre)[:dim], axis=1)
    near = np.abs(phi) < h
    return float(np.mean(r[near] - phi[near]))


# --- batches of independent meshes ---------------------------------------


@dataclass
class Batch:
    name: str
    dim: int
    meshes: list             # list of Case, one per mesh (labels unused: the batch is partitioned whole)
    patch_verts: int = 256
    niter: int = 5
    mode: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def n_verts(self):
        return sum(m.xyz.shape[0] for m in self.meshes)

    @property
    def n_cells(self):
        return sum(m.cells.shape[0] for m in self.meshes)


def random_rotation(rng):
    """Haar-distributed rotation (QR of a Gaussian matrix, signs fixed)."""
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))[None, :]
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def ellipsoid_mesh(n: int, rng, cfl: float = 0.5, band_h: float = 6.0) -> Case:
    """n^3 Kuhn grid on [0,1]^3; phi0 = exact signed distance to an ellipsoid
    centred at (0.5,0.5,0.5) with semi-axes U[0.1,0.35] and a Haar rotation;
    affine velocity A x + b (entries U[-1,1]) at element centroids; CFL dt."""
    xyz, cells, h = kuhn_grid(n)
    axes = rng.uniform(0.1, 0.35, size=3)
    Q = random_rotation(rng)
    A = rng.uniform(-1.0, 1.0, size=(3, 3))
    b = rng.uniform(-1.0, 1.0, size=3)
    local = (xyz - 0.5) @ Q  # coordinates in the ellipsoid frame (columns of Q are its axes)
    phi0 = ellipsoid_distance(local, axes)
    vel = centroids(xyz, cells) @ A.T + b[None, :]
    vmax = float(np.max(np.linalg.norm(vel, axis=1)))
    return Case(f"ellipsoid_{n}", 3, xyz, cells, phi0, vel, dt=cfl * h / vmax, steps=0, bandwidth=band_h * h,
                h=h, labels=np.zeros(xyz.shape[0], np.int32), nparts=1, niter=5,
                meta={"axes": axes, "rotation": Q, "A": A, "b": b, "vmax": vmax})


def _ellipsoid_mesh_seeded(args):
    n, seed = args
    return ellipsoid_mesh(n, np.random.default_rng(seed))


def config4(n_meshes: int = 512, seed: int = 4, lo: int = 2000, hi: int = 200000, steps: int = 50,
            patch_verts: int = 256, workers: int = 0) -> Batch:
    """512 Kuhn meshes with vertex counts log-uniform in [2k, 200k] (side
    round(N^(1/3))), random ellipsoids and affine velocities, CFL 0.5, band 6h.
    Mesh m draws its ellipsoid and velocity from its own seed (seed, m), so
    the batch is the same whatever the number of worker processes."""
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_meshes))
    sides = np.maximum(3, np.round(np.cbrt(counts)).astype(int))
    jobs = [(int(n), [seed, m]) for m, n in enumerate(sides)]
    workers = workers or min(8, os.cpu_count() or 1)
    if workers > 1 and n_meshes > 8:
        import concurrent.futures as cf
        with cf.ProcessPoolExecutor(max_workers=workers) as ex:
            meshes = list(ex.map(_ellipsoid_mesh_seeded, jobs, chunksize=4))
    else:
        meshes = [_ellipsoid_mesh_seeded(j) for j in jobs]
    return Batch(f"batch_{n_meshes}_ellipsoids", 3, meshes, patch_verts=patch_verts, niter=5, mode=0,
                 meta={"sides": sides, "steps": steps, "seed": seed})


def small_batch(n_meshes: int = 6, seed: int = 11, lo: int = 300, hi: int = 6000) -> Batch:
    """A few uneven ellipsoid meshes for tests."""
    return config4(n_meshes, seed=seed, lo=lo, hi=hi, steps=4, patch_verts=128)


# --- mesh files -------------------------------------------------------------
# Python readers for the formats the C++ LevelSet class loads (tetgen
# .node/.ele, ASCII PLY). Coordinates are rounded to float32, as the mesh
# classes store them (src/examples/example1.cu:88-95 works on float points).

def _tetgen_rows(path):
    rows = []
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].split()
            if line:
                rows.append(line)
    return rows


def read_tetgen(basename: str):
    """(xyz float32-valued float64 (N,3), tets int32 (E,4)); indices start at the
    first point index of the .node file (0 or 1)."""
    nrows = _tetgen_rows(basename + ".node")
    nv = int(nrows[0][0])
    pts = np.array([[float(v) for v in r[1:4]] for r in nrows[1:1 + nv]], dtype=np.float32)
    first = int(nrows[1][0])
    erows = _tetgen_rows(basename + ".ele")
    ne = int(erows[0][0])
    tets = np.array([[int(v) for v in r[1:5]] for r in erows[1:1 + ne]], dtype=np.int64) - first
    return pts.astype(np.float64), tets.astype(np.int32)


def read_ply(path: str):
    """ASCII PLY: (xyz float32-valued float64 (N,3), triangles int32 (E,3)); polygons
    are split into fans; a file whose faces name vertices 1..N is read as
    1-based, and faces naming a vertex the file does not hold are dropped, as
    the C++ reader does."""
    with open(path) as f:
        lines = f.read().split("\n")
    nv = nf = 0
    k = 0
    while lines[k].strip() != "end_header":
        w = lines[k].split()
        if w[:2] == ["element", "vertex"]:
            nv = int(w[2])
        if w[:2] == ["element", "face"]:
            nf = int(w[2])
        k += 1
    k += 1
    pts = np.array([[float(v) for v in lines[k + i].split()[:3]] for i in range(nv)], dtype=np.float32)
    faces = []
    for i in range(nf):
        w = [int(float(v)) for v in lines[k + nv + i].split()]
        poly = w[1:1 + w[0]]
        for q in range(2, len(poly)):
            faces.append([poly[0], poly[q - 1], poly[q]])
    faces = np.array(faces, dtype=np.int64).reshape(-1, 3)
    if faces.size and faces.min() == 1 and faces.max() == nv:  # 1-based file
        faces -= 1
    faces = faces[((faces >= 0) & (faces < nv)).all(axis=1)]  # drop faces naming missing vertices
    return pts.astype(np.float64), faces.astype(np.int32)


def reorient_tets(xyz, tets):
    """Exchange corners 1 and 2 of every tet with negative signed volume."""
    t = tets.copy()
    a, b, c, d = (xyz[t[:, i]] for i in range(4))
    neg = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a) < 0.0
    t[neg, 1], t[neg, 2] = tets[neg, 2], tets[neg, 1]
    return t


def levelset_default_x(xyz, ncells, dim):
    """Initial values and velocities LevelSet::solveLevelSet fills in when the
    caller gives none (src/LevelSet.cu:176-231): phi0 = x - midX (float32),
    velocity (rangeX/100, 0, 0) per tet in 3D and zero in 2D."""
    x = xyz[:, 0].astype(np.float32)
    mn, mx = x.min(), x.max()
    mid = float(np.float32(mn + mx)) / 2.0
    rng = float(np.float32(mx - mn))
    phi0 = (x.astype(np.float64) - mid).astype(np.float32)
    vel = np.zeros((ncells, 3))
    if dim == 3:
        vel[:, 0] = float(np.float32(rng / 100.0))
    return phi0, vel


def metis_parts(n_verts: int, metis_size: int, dim: int) -> int:
    """Patch count of Partition_METIS: ceil(n/size) in 3D
    (src/core3d/cuda/meshFIM3d.cu:361), n//size with a minimum of 2 in 2D
    (src/core2d/cuda/meshFIM2d.cu:380-382)."""
    m = max(1, metis_size)
    p = -(-n_verts // m) if dim == 3 else max(2, n_verts // m)
    return max(1, min(p, n_verts))
