"""numpy restatement of the learned classifier's definitions (csrc/htx_learn.h)
used as the checker of the GPU kernels: face features (FP32 arithmetic in the
device's operation order), the MLP forward pass in FP64 and with the bf16
rounding of the matrix-core inference, and the training loss gradient."""
import numpy as np

F = np.float32
CLS_IN, CLS_H = 12, 64
GELU_A = 0.3989422804014327
GELU_B = -0.037625858331402064
GELU_X0 = 1.8799712059732503


def map_index(g, n, bc):
    if bc == "periodic":
        return np.mod(g, n)
    return np.clip(g, 0, n - 1)


def stencils(q, bc, direction, flag_faces=False):
    """Conserved state of the six stencil cells of every face: shape
    (6, 4, ny, nx); direction 0 = x faces (cell (i, j) | (i+1, j)).  With
    flag_faces, the faces of the per-step flag arrays instead: also the face
    one cell past each boundary, (6, 4, ny, nx + 2) for x faces i = -1 .. nx
    and (6, 4, ny + 2, nx) for y faces j = -1 .. ny."""
    nf, ny, nx = q.shape
    a_i = np.arange(-1, nx + 1) if flag_faces and direction == 0 else np.arange(nx)
    a_j = np.arange(-1, ny + 1) if flag_faces and direction == 1 else np.arange(ny)
    J, I = np.meshgrid(a_j, a_i, indexing="ij")
    out = np.empty((6, 4) + J.shape)
    for m in range(6):
        if direction == 0:
            gi, gj = map_index(I - 2 + m, nx, bc), J
        else:
            gi, gj = I, map_index(J - 2 + m, ny, bc)
        out[m] = q[:4, gj, gi]
    return out


def teno_contrast(e1, e2, e3, e4, eps):
    """TENO contrast of a window from its four FP32 differences."""
    c1 = F(13.0) / F(12.0)
    c_0, b_0 = e4 - e3, (e4 - F(3.0) * e3) * F(0.5)
    c_1, b_1 = e3 - e2, -(e2 + e3) * F(0.5)
    c_2, b_2 = e2 - e1, (F(3.0) * e2 - e1) * F(0.5)
    s0 = b_0 * b_0 + c1 * c_0 * c_0
    s1 = b_1 * b_1 + c1 * c_1 * c_1
    s2 = b_2 * b_2 + c1 * c_2 * c_2
    m11, m22, m12 = F(1727.0) / F(1260.0), F(2029.0) / F(168.0), F(-2479.0) / F(672.0)
    m13, m14, m23 = F(5123.0) / F(1680.0), F(-8209.0) / F(10080.0), F(-12157.0) / F(1120.0)
    t1 = m11 * e1 + m12 * e2 + m13 * e3 + m14 * e4
    t2 = m12 * e1 + m22 * e2 + m23 * e3 + m13 * e4
    t3 = m13 * e1 + m23 * e2 + m22 * e3 + m12 * e4
    t4 = m14 * e1 + m13 * e2 + m12 * e3 + m11 * e4
    b5 = e1 * t1 + e2 * t2 + e3 * t3 + e4 * t4
    tau = np.abs(b5 - (s0 + s2 + F(4.0) * s1) * (F(1.0) / F(6.0)))
    lo = np.minimum(np.minimum(s0, s1), s2) + eps
    hi = np.maximum(np.maximum(s0, s1), s2) + eps
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        x, y = tau / hi, tau / lo
        inv = np.where(np.isfinite(x), (F(1.0) + x) / (F(1.0) + y), F(1.0))
    r = F(1.0) - inv.astype(np.float32)
    return np.where(r > 0, r, F(0.0))


def F32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32)


def stencil_contrast(d, eps):
    """d = the five FP32 differences of a 6-cell stencil."""
    return np.maximum(teno_contrast(d[0], d[1], d[2], d[3], eps), teno_contrast(d[1], d[2], d[3], d[4], eps))


def log_amp(a):
    return np.log2(a + F(1e-12)) * F(0.30102999566398120 / 12.0) + F(1.0)


def features(q, gamma, bc, direction, flag_faces=False):
    """The 12 inputs of every face of one family, shape (ny, nx, 12) (the
    face grid of stencils() with flag_faces)."""
    st = stencils(q, bc, direction, flag_faces)
    return stencil_features(st, gamma, direction)


def stencil_features(st, gamma, direction, floor=0.0):
    """The 12 inputs of the faces of stencils st (6, 4, ...).  floor > 0:
    differences within floor x their scale (density, pressure, sound speed
    of the face) count as zero (the features of states advanced in FP32 bulk
    arithmetic)."""
    g1 = gamma - 1.0
    r = st[:, 0]
    qn, qt = (st[:, 1], st[:, 2]) if direction == 0 else (st[:, 2], st[:, 1])
    un, ut = qn / r, qt / r
    p = (st[:, 3] - 0.5 * (qn * un + qt * ut)) * g1
    dr = [F32(r[k + 1] - r[k]) for k in range(5)]
    dp = [F32(p[k + 1] - p[k]) for k in range(5)]
    du = [F32(un[k + 1] - un[k]) for k in range(5)]
    dv = [F32(ut[k + 1] - ut[k]) for k in range(5)]
    if floor > 0.0:
        fr = F32(0.5 * (np.abs(r[2]) + np.abs(r[3])))
        fp = F32(0.5 * (np.abs(p[2]) + np.abs(p[3])))
        zr, zp = F(floor) * fr, F(floor) * fp
        zc = F(floor) * np.sqrt(F(gamma) * fp * (F(1.0) / (fr + F(1e-30))))
        zero = lambda d, z: [np.where(np.abs(x) <= z, F(0.0), x) for x in d]
        dr, dp, du, dv = zero(dr, zr), zero(dp, zp), zero(du, zc), zero(dv, zc)
    sr = sp = np.zeros(r.shape[1:], dtype=np.float32)
    ar = ap = au = at = np.zeros(r.shape[1:], dtype=np.float32)
    for k in range(5):
        sr = sr + np.abs(dr[k])
        sp = sp + np.abs(dp[k])
        ar, ap = np.maximum(ar, np.abs(dr[k])), np.maximum(ap, np.abs(dp[k]))
        au, at = np.maximum(au, np.abs(du[k])), np.maximum(at, np.abs(dv[k]))
    rf = F32(0.5 * (np.abs(r[2]) + np.abs(r[3])))
    pf = F32(0.5 * (np.abs(p[2]) + np.abs(p[3])))
    irf, ipf = F(1.0) / (rf + F(1e-30)), F(1.0) / (pf + F(1e-30))
    cf2 = F(gamma) * pf * irf
    icf = F(1.0) / (np.sqrt(cf2) + F(1e-30))
    nr = F(1.0) / (sr + F(1e-6) * rf + F(1e-30))
    npp = F(1.0) / (sp + F(1e-6) * pf + F(1e-30))
    x = np.empty(r.shape[1:] + (12,), dtype=np.float32)
    x[..., 0] = dr[1] * nr
    x[..., 1] = dr[2] * nr
    x[..., 2] = dp[1] * npp
    x[..., 3] = dp[2] * npp
    x[..., 4] = log_amp(ar * irf)
    x[..., 5] = log_amp(ap * ipf)
    x[..., 6] = log_amp(au * icf)
    x[..., 7] = log_amp(at * icf)
    x[..., 8] = stencil_contrast(dr, F(1e-12) * rf * rf + F(1e-30))
    x[..., 9] = stencil_contrast(dp, F(1e-12) * pf * pf + F(1e-30))
    ev = F(1e-12) * cf2 + F(1e-30)
    x[..., 10] = stencil_contrast(du, ev)
    x[..., 11] = stencil_contrast(dv, ev)
    return x


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    W1 = p[o:o + 768].reshape(12, 64); o += 768
    b1 = p[o:o + 64]; o += 64
    W2 = p[o:o + 4096].reshape(64, 64); o += 4096
    b2 = p[o:o + 64]; o += 64
    w3 = p[o:o + 64]; o += 64
    return W1, b1, W2, b2, w3, p[o]


def gelu(x):
    t = np.clip(x, -GELU_X0, GELU_X0)
    return x * (0.5 + t * (GELU_A + GELU_B * t * t))


def gelu_d(x):
    t = np.clip(x, -GELU_X0, GELU_X0)
    phi = 0.5 + t * (GELU_A + GELU_B * t * t)
    inside = (x > -GELU_X0) & (x < GELU_X0)
    return phi + np.where(inside, x * (GELU_A + 3.0 * GELU_B * t * t), 0.0)


def forward(params, X):
    """FP64 logits of feature rows X (n, 12)."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    h1 = gelu(X.astype(np.float64) @ W1 + b1)
    h2 = gelu(h1 @ W2 + b2)
    return h2 @ w3 + b3


def bf16(a):
    """Round to bf16 (nearest even) and return as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(params, X):
    """Logits as the matrix-core inference forms them: weights, features and
    the first hidden layer rounded to bf16, sums in high precision, biases,
    activations and the output layer in FP32-range arithmetic."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    h1 = gelu(bf16(X) @ bf16(W1) + b1)
    h2 = gelu(bf16(h1) @ bf16(W2) + b2)
    return h2 @ w3 + b3


def loss_grad(params, X, y, troubled_weight):
    """Class-balanced weighted logistic loss over the whole set and its
    gradient (FP64), with the trainer's weighting."""
    W1, b1, W2, b2, w3, b3 = unpack(params)
    n = len(y)
    fpos = y.mean()
    fneg = 1.0 - fpos
    wpos, wneg = 0.5 / fpos, 0.5 / fneg * troubled_weight
    norm = 1.0 / (fpos * wpos + fneg * wneg) / n
    w = np.where(y > 0, wpos, wneg) * norm
    X = X.astype(np.float64)
    z1 = X @ W1 + b1
    h1 = gelu(z1)
    z2 = h1 @ W2 + b2
    h2 = gelu(z2)
    z3 = h2 @ w3 + b3
    sg = 1.0 / (1.0 + np.exp(-z3))
    loss = np.sum(w * (np.maximum(z3, 0) + np.log1p(np.exp(-np.abs(z3))) - y * z3))
    g3 = w * (sg - y)
    gw3 = h2.T @ g3
    gb3 = g3.sum()
    G2 = np.outer(g3, w3) * gelu_d(z2)
    gW2 = h1.T @ G2
    gb2 = G2.sum(0)
    G1 = (G2 @ W2.T) * gelu_d(z1)
    gW1 = X.T @ G1
    gb1 = G1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
