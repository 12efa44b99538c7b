"""CPU FP64 restatement of the learned pose proposer (forward, loss, backward).

TEST INFRASTRUCTURE ONLY -- imported by tests/ as the checker of the library's
BF16 MFMA kernels.  `quantize=True` rounds to BF16 at the points where the
library stores BF16 (inputs, weights, inter-layer activations, the pooled
vector, the output gradient fed to the GEMMs); `quantize=False` is the plain
FP64 network, whose backward is pinned by finite differences.

The network and loss (no counterpart in the reference):
  frame per cloud: centroid c, principal axes V = [e1, e2, e1 x e2] (descending
  covariance eigenvalues; e1, e2 signed so the third moment along them >= 0)
  x = [V^T (xyz - c) / 10, features]                        (per point, 8)
  H1 = relu(x W1 + b1), H2 = relu(H1 W2 + b2), H3 = relu(H2 W3 + b3)
  g = max over points (first index on ties); z = [g_src, g_tgt]
  z1 = relu(z W4 + b4); o = z1 W5 + b5 (64)
  hypothesis k: raw quaternion o[7k:7k+4] (rotation in the canonical frames:
  R = V_tgt rot(q) V_src^T), translation residual o[7k+4:7k+7], confidence
  logits o[56:64]
  L_k = |q^_k - s_k q*|^2 + |o_t,k - d*|^2, s_k = sign(q^_k . q*),
  q* = quat(V_tgt^T R* V_src),
  d* = (t* - (c_tgt - R* c_src)) / 10, k* = argmin L_k,
  loss = sum_k w_k L_k + 0.1 CE(softmax(conf), k*), w_k* = 0.95, else 0.05/7,
  averaged over pairs.
"""
from __future__ import annotations

import numpy as np

K = 8
SCALE = 10.0
SIZES = [(8, 64), (64, 128), (128, 256), (512, 256), (256, 64)]


def n_params():
    return sum(a * b + b for a, b in SIZES)


def unpack(w):
    """Flat vector -> [(W, b)] per layer (row-major W, input index first)."""
    out, o = [], 0
    for a, b in SIZES:
        W = np.asarray(w[o:o + a * b], np.float64).reshape(a, b)
        o += a * b
        bb = np.asarray(w[o:o + b], np.float64)
        o += b
        out.append((W, bb))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def bf16(x):
    """Round to BF16 (nearest even), returned as float64."""
    x32 = np.ascontiguousarray(np.asarray(x, np.float32))
    u = x32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def quat_to_rot(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def rot_to_quat(R):
    """Unit quaternion (w >= 0) of a rotation."""
    R = np.asarray(R, np.float64)
    tr = np.trace(R)
    if tr > 0:
        s = np.sqrt(tr + 1.0) * 2
        q = np.array([0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s])
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        q = np.array([(R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s])
    elif R[1, 1] > R[2, 2]:
        s = np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        q = np.array([(R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s])
    else:
        s = np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        q = np.array([(R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s])
    q /= np.linalg.norm(q)
    return q if q[0] >= 0 else -q


def principal_frame(xyz):
    """Centroid and canonical axes of one cloud (n, 3)."""
    c = xyz.mean(axis=0)
    y = xyz - c
    ev, V = np.linalg.eigh(y.T @ y)
    V = V[:, np.argsort(-ev)].copy()
    for k in range(2):
        if np.sum((y @ V[:, k]) ** 3) < 0:
            V[:, k] = -V[:, k]
    V[:, 2] = np.cross(V[:, 0], V[:, 1])
    return c, V


def forward(w, a, b, quantize=True):
    """a, b: (n_pairs, npts, 8).  Returns o (n_pairs, 64) and a cache."""
    Q = bf16 if quantize else (lambda v: np.asarray(v, np.float64))
    L = unpack(w)
    Wq = [Q(W) for W, _ in L]
    bs = [np.asarray(np.asarray(bb, np.float32), np.float64) if quantize else bb for _, bb in L]
    clouds = np.stack([a, b], axis=1).reshape(-1, a.shape[1], 8).astype(np.float64)  # [pair][cloud]
    frames = [principal_frame(cl[:, :3]) for cl in clouds]
    cen = np.array([f[0] for f in frames])
    Vs = np.array([f[1] for f in frames])
    X = np.concatenate([np.einsum("cpi,cik->cpk", clouds[:, :, :3] - cen[:, None, :], Vs) / SCALE,
                        clouds[:, :, 3:]], axis=2)
    X = Q(X)
    H1 = Q(np.maximum(X @ Wq[0] + bs[0], 0.0))
    H2 = Q(np.maximum(H1 @ Wq[1] + bs[1], 0.0))
    H3 = np.maximum(H2 @ Wq[2] + bs[2], 0.0)
    arg = H3.argmax(axis=1)                       # (clouds, 256)
    g = np.take_along_axis(H3, arg[:, None, :], axis=1)[:, 0, :]
    n_pairs = a.shape[0]
    z = Q(g.reshape(n_pairs, 512))
    z1 = Q(np.maximum(z @ Wq[3] + bs[3], 0.0))
    o = z1 @ Wq[4] + bs[4]
    cache = dict(Wq=Wq, X=X, H1=H1, H2=H2, arg=arg, g=g, z=z, z1=z1, cen=cen, V=Vs, quantize=quantize)
    return o, cache


def loss_grad_out(o, targets, cen, V):
    """Mean loss and d loss / d o (n_pairs, 64)."""
    n = o.shape[0]
    dout = np.zeros_like(o)
    total = 0.0
    for p in range(n):
        Rs = quat_to_rot(np.asarray(targets[p, :4], np.float64))
        qs = rot_to_quat(V[2 * p + 1].T @ Rs @ V[2 * p])
        cs, ct = cen[2 * p], cen[2 * p + 1]
        dstar = (np.asarray(targets[p, 4:7], np.float64) - (ct - Rs @ cs)) / SCALE
        Ls, qh, sg, nr = [], [], [], []
        for k in range(K):
            r = o[p, 7 * k:7 * k + 7]
            nk = np.linalg.norm(r[:4]) + 1e-8
            qk = r[:4] / nk
            s = -1.0 if qk @ qs < 0 else 1.0
            Ls.append(np.sum((qk - s * qs) ** 2) + np.sum((r[4:] - dstar) ** 2))
            qh.append(qk)
            sg.append(s)
            nr.append(nk)
        best = int(np.argmin(Ls))
        cl = o[p, 56:64]
        lse = cl.max() + np.log(np.sum(np.exp(cl - cl.max())))
        loss = 0.1 * (lse - cl[best])
        for k in range(K):
            wk = 0.95 if k == best else 0.05 / 7
            loss += wk * Ls[k]
            gq = 2 * wk * (qh[k] - sg[k] * qs)
            dout[p, 7 * k:7 * k + 4] = (gq - (gq @ qh[k]) * qh[k]) / nr[k] / n
            dout[p, 7 * k + 4:7 * k + 7] = 2 * wk * (o[p, 7 * k + 4:7 * k + 7] - dstar) / n
            dout[p, 56 + k] = 0.1 * (np.exp(cl[k] - lse) - (1.0 if k == best else 0.0)) / n
        total += loss
    return total / n, dout


def loss(w, a, b, targets, quantize=False):
    o, c = forward(w, a, b, quantize)
    return loss_grad_out(o, targets, c["cen"], c["V"])[0]


def gradient(w, a, b, targets, quantize=True):
    """Mean loss and its gradient (flat, parameter order)."""
    Q = bf16 if quantize else (lambda v: np.asarray(v, np.float64))
    o, c = forward(w, a, b, quantize)
    L, dout = loss_grad_out(o, targets, c["cen"], c["V"])
    Wq, z, z1 = c["Wq"], c["z"], c["z1"]
    dq = Q(dout)
    dW5 = z1.T @ dq
    db5 = dout.sum(axis=0)
    dz1 = Q((dq @ Wq[4].T) * (z1 > 0))
    dW4 = z.T @ dz1
    db4 = dz1.sum(axis=0)
    dz = dz1 @ Wq[3].T                              # (pairs, 512)
    g, arg, H1, H2, X = c["g"], c["arg"], c["H1"], c["H2"], c["X"]
    n_clouds = g.shape[0]
    v = np.where(g > 0, dz.reshape(n_clouds, 256), 0.0)  # (clouds, 256)
    cl = np.repeat(np.arange(n_clouds), 256)
    ch = np.tile(np.arange(256), n_clouds)
    pt = arg.reshape(-1)
    vr = v.reshape(-1)
    H2r = H2[cl, pt]                                 # (R, 128)
    H1r = H1[cl, pt]
    Xr = X[cl, pt]
    dW3 = np.zeros((128, 256))
    np.add.at(dW3.T, ch, vr[:, None] * H2r)
    db3 = v.sum(axis=0)
    dH2 = Q(vr[:, None] * Wq[2][:, ch].T * (H2r > 0))
    dW2 = H1r.T @ dH2
    db2 = dH2.sum(axis=0)
    dH1 = Q((dH2 @ Wq[1].T) * (H1r > 0))
    dW1 = Xr.T @ dH1
    db1 = dH1.sum(axis=0)
    grad = pack([(dW1, db1), (dW2, db2), (dW3, db3), (dW4, db4), (dW5, db5)])
    return L, grad


def propose(w, a, b, quantize=True):
    """Decoded proposals: (n_pairs, 8, 7) unit quaternion (w >= 0) + translation,
    and softmax confidences (n_pairs, 8)."""
    o, c = forward(w, a, b, quantize)
    n = o.shape[0]
    qt = np.zeros((n, K, 7))
    conf = np.zeros((n, K))
    for p in range(n):
        cs, ct = c["cen"][2 * p], c["cen"][2 * p + 1]
        Vs, Vt = c["V"][2 * p], c["V"][2 * p + 1]
        for k in range(K):
            r = o[p, 7 * k:7 * k + 7]
            R = Vt @ quat_to_rot(r[:4]) @ Vs.T
            qt[p, k, :4] = rot_to_quat(R)
            qt[p, k, 4:] = ct - R @ cs + SCALE * r[4:]
        cl = o[p, 56:]
        e = np.exp(cl - cl.max())
        conf[p] = e / e.sum()
    return qt, conf, o
