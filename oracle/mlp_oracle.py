"""CPU oracle of the learned warm start's network (test infrastructure only).

A numpy FP64 restatement of the 22-64-64-4 ReLU MLP the GPU runs with BF16
MFMA: the per-copy features (tets, and triangles in the same layout), the
forward pass, the mean squared error and its analytic gradient. `bf16=True` rounds every matrix-product operand to
BF16 (round to nearest even) as the GPU's MFMA inputs are, so GPU results
can be compared at FP32-accumulation tolerance; `bf16=False` is the exact
FP64 network that central finite differences pin.

Parameter vector layout: W1 (22x64, row = input), b1 (64), W2 (64x64),
b2 (64), W3 (64x4), b3 (4); 5892 values.
"""
import numpy as np

N_IN, N_HID, N_OUT = 22, 64, 4
SHAPES = [("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_OUT)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def init_params(seed=0):
    """He-normal weights, zero biases, W3 scaled down so the initial
    correction is small."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_HID)), "b1": np.zeros(N_HID),
         "W2": rng.normal(0, np.sqrt(2.0 / N_HID), (N_HID, N_HID)), "b2": np.zeros(N_HID),
         "W3": rng.normal(0, 0.1 * np.sqrt(2.0 / N_HID), (N_HID, N_OUT)), "b3": np.zeros(N_OUT)}
    return pack(d).astype(np.float32)


def bf16(x):
    """Round to the nearest BF16 value (ties to even), as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def features(ph, L):
    """Features of copies (rows): ph (n,4) corner phi, L (n,6) local
    coordinates A=(0,0,0), B=(L0,0,0), C=(L1,L2,0), D=(L3,L4,L5).
    Returns (X (n,22), s (n,)) with everything divided by s = |AB|."""
    ph = np.asarray(ph, np.float64)
    L = np.asarray(L, np.float64)
    s = L[:, 0]
    z = np.zeros_like(s)
    AB = np.stack([L[:, 0], z, z], 1)
    AC = np.stack([L[:, 1], L[:, 2], z], 1)
    AD = np.stack([L[:, 3], L[:, 4], L[:, 5]], 1)
    X = np.concatenate([ph, AB, AC, AD, AC - AB, AD - AB, AD - AC], 1) / s[:, None]
    return X, s


def features_tri(ph, c):
    """Features of triangle copies: ph (n,3) corner phi, c (n,3) local
    coordinates A=(0,0), B=(c0,0), C=(c1,c2), in the tet layout with the
    fourth corner's phi and the edges AD, BD, CD zero."""
    ph = np.asarray(ph, np.float64)
    c = np.asarray(c, np.float64)
    s = c[:, 0]
    z = np.zeros_like(s)
    AB = np.stack([c[:, 0], z, z], 1)
    AC = np.stack([c[:, 1], c[:, 2], z], 1)
    Z3 = np.zeros((s.size, 3))
    X = np.concatenate([ph, z[:, None], AB, AC, Z3, AC - AB, Z3, Z3], 1) / s[:, None]
    return X, s


def local_coords_tri(xyz, tris):
    """Local coordinates of triangles (rows A,B,C): |AB|, AC.X, AC.Y with X
    along AB and Y the in-plane normal to it towards C."""
    A, B, Cc = (np.asarray(xyz, np.float64)[tris[:, i]] for i in range(3))
    ab = B - A
    lab = np.linalg.norm(ab, axis=1)
    ex = ab / lab[:, None]
    ac = Cc - A
    acx = np.einsum("ij,ij->i", ac, ex)
    yv = ac - acx[:, None] * ex
    return np.stack([lab, acx, np.linalg.norm(yv, axis=1)], 1)


def forward(p, X, bf16_ops=False):
    w = unpack(p)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    Z1 = r(X) @ r(w["W1"]) + w["b1"]
    H1 = np.maximum(Z1, 0.0)
    Z2 = r(H1) @ r(w["W2"]) + w["b2"]
    H2 = np.maximum(Z2, 0.0)
    O = r(H2) @ r(w["W3"]) + w["b3"]
    return O, (X, H1, H2)


def loss_and_grad(p, X, Y, bf16_ops=False):
    """Sum of squared errors and the gradient of their mean over samples
    (sum over the 4 outputs), by backpropagation."""
    w = unpack(p)
    r = bf16 if bf16_ops else (lambda a: np.asarray(a, np.float64))
    n = X.shape[0]
    O, (X0, H1, H2) = forward(p, X, bf16_ops)
    D = O - np.asarray(Y, np.float64)
    loss = float(np.sum(D * D))
    dO = 2.0 * D / n
    g = {"W3": r(H2).T @ r(dO), "b3": dO.sum(0)}
    dH2 = (r(dO) @ r(w["W3"]).T) * (H2 > 0)
    g["W2"] = r(H1).T @ r(dH2)
    g["b2"] = dH2.sum(0)
    dH1 = (r(dH2) @ r(w["W2"]).T) * (H1 > 0)
    g["W1"] = r(X0).T @ r(dH1)
    g["b1"] = dH1.sum(0)
    return loss, pack(g)


def mean_loss(p, X, Y):
    O, _ = forward(p, X)
    return float(np.mean(np.sum((O - Y) ** 2, axis=1)))


def finite_difference(p, X, Y, idx, h=1e-6):
    """Central differences of the FP64 mean loss for parameters idx."""
    p = np.asarray(p, np.float64)
    out = np.empty(len(idx))
    for k, i in enumerate(idx):
        a = p.copy()
        b = p.copy()
        a[i] += h
        b[i] -= h
        out[k] = (mean_loss(a, X, Y) - mean_loss(b, X, Y)) / (2 * h)
    return out


def local_coords(xyz, tets):
    """Local coordinates of tets (rows A,B,C,D): B on the x axis, C in the
    xy plane (the frame the layout stores)."""
    A, B, Cc, D = (np.asarray(xyz, np.float64)[tets[:, i]] for i in range(4))
    ab = B - A
    lab = np.linalg.norm(ab, axis=1)
    ex = ab / lab[:, None]
    ac = Cc - A
    acx = np.einsum("ij,ij->i", ac, ex)
    yv = ac - acx[:, None] * ex
    acy = np.linalg.norm(yv, axis=1)
    ey = yv / acy[:, None]
    ez = np.cross(ex, ey)
    ad = D - A
    return np.stack([lab, acx, acy, np.einsum("ij,ij->i", ad, ex), np.einsum("ij,ij->i", ad, ey),
                     np.einsum("ij,ij->i", ad, ez)], 1)
