"""The supporting-plane bound the GPU sweeps use to skip tet corner solves
(ElemSolve<T, 4>::plane_mask in csrc/ls_kernels.hip) must never exceed the
corner's solve by more than its margin, or a skipped solve could have lowered
the value. Checked here against the oracle's tet solve on regular, sliver and
needle tets with seed-like, random and unreached values; the bound's FP32
evaluation is restated in numpy with the same operations. CPU only."""
import numpy as np

BIG = 1e9  # an unreached value


def _local(P):
    """Local coordinates (A at the origin, B on x, C in the xy plane)."""
    AB, AC, AD = P[1] - P[0], P[2] - P[0], P[3] - P[0]
    X = AB / np.linalg.norm(AB)
    Z = np.cross(AB, AC)
    Z /= np.linalg.norm(Z)
    Y = np.cross(Z, X)
    return np.array([np.linalg.norm(AB), AC @ X, AC @ Y, AD @ X, AD @ Y, AD @ Z])


def _corner_bound(T, qk, q, n, r, wrong_sign=False):
    """Bound for one corner from the face values T[0..2], squared distances
    qk from the corner to the face corners, squared face edges q = (q01, q02,
    q12), lengths n and reciprocal lengths r, in FP32."""
    f = np.float32
    s = f(-1.0) if wrong_sign else f(1.0)
    G01 = f(0.5) * (qk[0] + qk[1] - s * q[0])
    G02 = f(0.5) * (qk[0] + qk[2] - s * q[1])
    G12 = f(0.5) * (qk[1] + qk[2] - s * q[2])
    b0 = min(T[0] + n[0], T[1] + G01 * r[0], T[2] + G02 * r[0])
    b1 = min(T[1] + n[1], T[0] + G01 * r[1], T[2] + G12 * r[1])
    b2 = min(T[2] + n[2], T[0] + G02 * r[2], T[1] + G12 * r[2])
    return f(max(b0, b1, b2))


def plane_bounds(L, v, wrong_sign=False):
    """The four corners' bounds, as the GPU evaluates them."""
    f = np.float32
    L = L.astype(f)
    P = np.array([[0, 0, 0], [L[0], 0, 0], [L[1], L[2], 0], [L[3], L[4], L[5]]], dtype=f)
    q = {(i, j): f(((P[i] - P[j]) ** 2).sum()) for i in range(4) for j in range(4) if i < j}
    qq = lambda i, j: q[(min(i, j), max(i, j))]
    T = v.astype(f)
    out = []
    for k in range(4):
        c = [j for j in range(4) if j != k]
        qk = [qq(k, j) for j in c]
        nk = [np.sqrt(x) for x in qk]
        rk = [f(1.0) / x for x in nk]
        fq = (qq(c[0], c[1]), qq(c[0], c[2]), qq(c[1], c[2]))
        out.append(_corner_bound([T[j] for j in c], qk, fq, nk, rk, wrong_sign))
    return np.array(out, dtype=np.float64)


def _solves(oracle_mod, L, v):
    """Each corner's solve from the other three (own value unreached)."""
    out = np.empty(4)
    for k in range(4):
        w = v.copy()
        w[k] = BIG
        out[k] = oracle_mod.tet_element_update(L, w, 64)[k]
    return out


def _tets(rng, n):
    for i in range(n):
        P = rng.normal(size=(4, 3))
        kind = i % 3
        if kind == 1:  # sliver: D close to the plane of ABC
            nrm = np.cross(P[1] - P[0], P[2] - P[0])
            nrm /= np.linalg.norm(nrm)
            P[3] -= ((P[3] - P[0]) @ nrm - 1e-3) * nrm
        elif kind == 2:  # needle
            P[1] = P[0] + 1e-2 * (P[1] - P[0])
        yield P


def _values(rng, P):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    v = np.abs(P @ n - rng.normal())
    mode = rng.integers(4)
    if mode == 1:
        v[rng.integers(4)] = 0.0
    elif mode == 2:
        v[rng.integers(4)] = BIG
    elif mode == 3:
        v = rng.uniform(0, 3, 4)
    return v


def _violations(oracle_mod, wrong_sign=False, n=300):
    rng = np.random.default_rng(7)
    bad = pruned = total = 0
    for P in _tets(rng, n):
        L = _local(P)
        scale = np.abs(L).sum()
        v = _values(rng, P)
        s = _solves(oracle_mod, L, v)
        b = plane_bounds(L, v, wrong_sign)
        for k in range(4):
            Mo = max(v[j] for j in range(4) if j != k)
            margin = 1e-5 * (Mo + scale + 1.0)
            total += 1
            if b[k] - margin > s[k]:
                bad += 1
            # the bound leaves the corner out: its solve cannot go below b - margin
            if b[k] - margin > min(s[k], BIG) - 1.0:
                pruned += 1
    return bad, pruned, total


def test_plane_bound_never_exceeds_the_solve(oracle_mod):
    bad, _, total = _violations(oracle_mod)
    assert total == 1200 and bad == 0


def test_plane_bound_with_a_planted_sign_error_is_caught(oracle_mod):
    # the law of cosines with the wrong sign overestimates the projections
    bad, _, _ = _violations(oracle_mod, wrong_sign=True)
    assert bad > 0


def test_plane_bound_is_the_vertex_distance_when_one_face_corner_is_a_seed(oracle_mod):
    # Face corner c0 at 0, the other two unreached: the solve is |c0 - D|
    # (the nearest face point along the value's growth is c0 itself), and
    # the bound toward c0 reaches it, up to FP32 rounding.
    rng = np.random.default_rng(3)
    for P in _tets(rng, 30):
        L = _local(P)
        for k in range(4):
            c = [j for j in range(4) if j != k]
            v = np.full(4, BIG)
            v[c[0]] = 0.0
            s = _solves(oracle_mod, L, v)[k]
            Pl = np.array([[0, 0, 0], [L[0], 0, 0], [L[1], L[2], 0], [L[3], L[4], L[5]]])
            d = np.linalg.norm(Pl[c[0]] - Pl[k])
            assert abs(s - d) <= 1e-9 * (1 + d)
            b = plane_bounds(L, v)[k]
            assert abs(b - d) <= 1e-5 * (1 + d), (b, d)
