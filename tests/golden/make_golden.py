"""Writes the golden vectors under tests/golden/ from library routines
(numpy, scipy, PyYAML) and the reference's data files -- independently of the
oracle's code, so the oracle is checked against them, not against itself.

Run in the build container (needs scipy and PyYAML; /root/reference only for
the icosahedral table):  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import re

import numpy as np
import scipy.linalg
import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]], dtype=np.float64)


def twist_hat(xi):
    """4x4 se(3) matrix of xi = [w; v]."""
    X = np.zeros((4, 4))
    X[:3, :3] = skew(xi[:3])
    X[:3, 3] = xi[3:]
    return X


# ---------------------------------------------------------------------------
def params_golden():
    """Each reference YAML parsed by PyYAML (directive lines dropped, first
    document only): the keys present and their values."""
    out = {}
    for fn in sorted(os.listdir(os.path.join(HERE, "yaml"))):
        with open(os.path.join(HERE, "yaml", fn)) as f:
            text = "".join(l for l in f if not l.lstrip().startswith("%"))
        docs = [d for d in yaml.safe_load_all(text) if d is not None]
        d = docs[0] if docs else {}
        out[fn] = {k: (v if not isinstance(v, bool) else int(v)) for k, v in d.items()}
    return out


def cubic_golden(rng):
    cases = []
    for _ in range(40):
        c = rng.normal(size=4) * 10 ** rng.uniform(-3, 3, size=4)
        cases.append(c)
    # constructed roots: three real, a double root, a tight complex pair
    for roots in ([0.5, 2.0, -1.0], [1e-3, 1e-3, 5.0], [0.01, 0.02, 0.03], [-2.0, 1e-6, 3.0]):
        cases.append(np.poly(roots) * 3.7)
    for re_, im_ in ((0.2, 1e-7), (0.2, 1e-3), (1.5, 2.0)):
        cases.append(np.poly([re_ + 1j * im_, re_ - 1j * im_, 0.7]).real * 0.5)
    out = []
    for c in cases:
        r = np.roots(c)
        out.append({"coef": list(map(float, c)), "re": list(map(float, r.real)), "im": list(map(float, r.imag))})
    return out


def se3_golden(rng):
    out = []
    for k in range(30):
        w = rng.normal(size=3)
        if k % 5 == 0:
            w *= 1e-8  # small-angle branch
        xi = np.concatenate([w, rng.normal(size=3)])
        xi /= np.linalg.norm(xi)
        dt = float(10 ** rng.uniform(-5, 0.3))
        E = scipy.linalg.expm(twist_hat(xi * dt))
        L = scipy.linalg.logm(E).real
        tangent = np.concatenate([L[:3, 3], [L[2, 1], L[0, 2], L[1, 0]]])  # (upsilon, omega)
        out.append({"xi": list(map(float, xi)), "dt": dt, "T": E[:3, :4].ravel().tolist(),
                    "log": list(map(float, tangent)), "log_norm": float(np.linalg.norm(tangent))})
    return out


def _tiny_scene(rng, n, m, F, K):
    x = rng.uniform(-0.6, 0.6, size=(n, 3))
    y = x[rng.integers(0, n, size=m)] + rng.normal(scale=0.15, size=(m, 3))
    fx = rng.uniform(0, 1, size=(n, F))
    fy = np.clip(fx[rng.integers(0, n, size=m)] + rng.normal(scale=0.05, size=(m, F)), 0, 1)
    sx = rng.dirichlet(np.ones(K) * 0.3, size=n)
    sy = rng.dirichlet(np.ones(K) * 0.3, size=m)
    return x.astype(np.float32), y.astype(np.float32), fx.astype(np.float32), fy.astype(np.float32), \
        sx.astype(np.float32), sy.astype(np.float32)


def assoc_golden(rng):
    """Kernel definition of fill_in_A_mat_gpu evaluated in float64 by brute
    force: a_ij = sigma^2 exp(-|x-y'|^2/2l_i^2) c_sigma^2 exp(-|fx-fy|^2/2c_ell^2)
    s_sigma^2 exp(-|sx-sy|^2/2s_ell^2), gates by the sp_thres-derived radii,
    l_i = (|x_i|/500+1) ell, y' = T^-1 y, first K passing targets by index."""
    cases = []
    for case in range(4):
        n, m, F, K = 12, 40, 5, 4
        x, y, fx, fy, sx, sy = _tiny_scene(rng, n, m, F, K)
        ang = rng.normal(size=3) * 0.05
        T = np.eye(4)
        T[:3, :3] = scipy.linalg.expm(skew(ang))
        T[:3, 3] = rng.normal(size=3) * 0.05
        T = T.astype(np.float32).astype(np.float64)
        P = dict(sigma=0.1, sp_thres=1e-3, c_ell=0.3, c_sigma=0.8, s_ell=0.4, s_sigma=0.9,
                 is_using_intensity=1 if case != 1 else 0, is_using_semantics=1 if case >= 2 else 0)
        ell = [0.3, 0.5, 0.4, 0.6][case]
        Kcap = [4, 6, 3, 50][case]
        Ti = np.linalg.inv(T)
        yp = y.astype(np.float64) @ Ti[:3, :3].T + Ti[:3, 3]
        pairs = []
        for i in range(n):
            l = (np.linalg.norm(x[i].astype(np.float64)) / 500 + 1) * ell
            d2t = -2 * l * l * np.log(P["sp_thres"] / P["sigma"] ** 2)
            d2ct = -2 * P["c_ell"] ** 2 * np.log(P["sp_thres"] / P["c_sigma"] ** 2)
            d2st = -2 * P["s_ell"] ** 2 * np.log(P["sp_thres"] / P["s_sigma"] ** 2)
            stored = 0
            for j in range(m):
                if stored == Kcap:
                    break
                d2 = float(np.sum((yp[j] - x[i]) ** 2))
                if not d2 < d2t:
                    continue
                a = P["sigma"] ** 2 * np.exp(-d2 / (2 * l * l))
                margin = abs(d2 - d2t) / d2t
                if P["is_using_intensity"]:
                    d2c = float(np.sum((fx[i].astype(np.float64) - fy[j]) ** 2))
                    if not d2c < d2ct:
                        continue
                    margin = min(margin, abs(d2c - d2ct) / d2ct)
                    a *= P["c_sigma"] ** 2 * np.exp(-d2c / (2 * P["c_ell"] ** 2))
                if P["is_using_semantics"]:
                    d2s = float(np.sum((sx[i].astype(np.float64) - sy[j]) ** 2))
                    if not d2s < d2st:
                        continue
                    margin = min(margin, abs(d2s - d2st) / d2st)
                    a *= P["s_sigma"] ** 2 * np.exp(-d2s / (2 * P["s_ell"] ** 2))
                if a > P["sp_thres"]:
                    pairs.append([i, j, float(a), float(min(margin, abs(a - P["sp_thres"]) / P["sp_thres"]))])
                    stored += 1
        cases.append({"params": P, "ell": ell, "K": Kcap, "T": T.ravel().tolist(),
                      "x": x.tolist(), "y": y.tolist(), "fx": fx.tolist(), "fy": fy.tolist(),
                      "sx": sx.tolist(), "sy": sy.tolist(), "pairs": pairs})
    return cases


def flow_step_golden(rng):
    """Flow sums and step-size Taylor coefficients of fixed pairs.

    flow:  omega = sum a (x cross y)/c,  v = sum a (y - x)/d.
    step:  B..E are the s^1..s^4 coefficients of
           sum_ij a_ij exp(-(|x_i - w_j(s)|^2 - |x_i - z_j|^2) / (2 l^2))
           with the path w_j(s) = z_j - sum_{k=1..4} s^k xi^k z_j
           (xi^k z = homogeneous twist power applied to [z;1]), truncated at
           order 4 -- computed here by numpy polynomial arithmetic."""
    from numpy.polynomial import polynomial as Pn
    cases = []
    for case in range(3):
        n, m = 5, 7
        x = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
        z = rng.uniform(-1, 1, size=(m, 3)).astype(np.float32)
        K = 4
        counts = rng.integers(0, K + 1, size=n)
        cols = np.full((n, K), -1, np.int64)
        vals = np.zeros((n, K), np.float32)
        for i in range(n):
            js = np.sort(rng.choice(m, size=counts[i], replace=False))
            cols[i, :counts[i]] = js
            vals[i, :counts[i]] = rng.uniform(1e-3, 5e-3, size=counts[i])
        c_, d_ = 7.0, 7.0
        om = np.zeros(3)
        v = np.zeros(3)
        for i in range(n):
            for s in range(counts[i]):
                j = cols[i, s]
                a = float(vals[i, s])
                om += a * np.cross(x[i].astype(np.float64), z[j].astype(np.float64))
                v += a * (z[j].astype(np.float64) - x[i])
        flow = np.concatenate([om / c_, v / d_])
        xi = flow / np.linalg.norm(flow)
        ell = [0.25, 0.5, 0.375][case]  # exact in float32 (the API takes a float ell)
        cc = 1.0 / (2 * ell * ell)
        H = twist_hat(xi)
        coef = np.zeros(5)
        for i in range(n):
            for s in range(counts[i]):
                j = cols[i, s]
                a = float(vals[i, s])
                zh = np.append(z[j].astype(np.float64), 1.0)
                terms = [np.linalg.matrix_power(H, k) @ zh for k in range(1, 5)]
                # w(s) = z - sum s^k xi^k z, per coordinate a polynomial in s
                diff_poly = []  # x - w(s)
                for d in range(3):
                    p = np.zeros(5)
                    p[0] = float(x[i][d]) - float(z[j][d])
                    for k in range(4):
                        p[k + 1] = terms[k][d]
                    diff_poly.append(p)
                sq = np.zeros(1)
                for d in range(3):
                    sq = Pn.polyadd(sq, Pn.polymul(diff_poly[d], diff_poly[d]))
                sq = np.pad(sq, (0, max(0, 5 - len(sq))))[:5]
                f = -cc * (sq - np.array([sq[0], 0, 0, 0, 0]))
                # exp(f) truncated at order 4 (f has no constant term)
                e = np.array([1.0, 0, 0, 0, 0])
                term = np.array([1.0, 0, 0, 0, 0])
                for k in range(1, 5):
                    term = Pn.polymul(term, f)[:5] / k
                    term = np.pad(term, (0, 5 - len(term)))
                    e = e + term
                coef += a * e
        cases.append({"x": x.tolist(), "z": z.tolist(), "counts": counts.tolist(), "cols": cols.tolist(),
                      "vals": vals.tolist(), "c": c_, "d": d_, "flow": flow.tolist(), "ell": ell,
                      "bcde": coef[1:5].tolist()})
    return cases


def icosahedral_table():
    path = os.path.join(REF, "src/cvo/CvoGPU.cpp")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    mats = []
    for m in re.finditer(r"pose\[(\d+)\]<<([^;]+);", text):
        vals = [float(v) for v in m.group(2).split(",")]
        mats.append((int(m.group(1)), vals))
    mats.sort()
    return [v for _, v in mats]


# ---------------------------------------------------------------------------
def _rand_pose(rng, tscale=3.0):
    w = rng.normal(size=3)
    w *= rng.uniform(0.1, 2.5) / np.linalg.norm(w)
    T = np.eye(4)
    T[:3, :3] = scipy.linalg.expm(skew(w))
    T[:3, 3] = rng.normal(size=3) * tscale
    return T


def _plus(T, d):
    """T Exp([rho; phi]) by the matrix exponential of the twist."""
    return T @ scipy.linalg.expm(twist_hat(np.r_[d[3:], d[:3]]))


def _residual(T1, T2, p1, p2, a):
    diff = (T1[:3, :3] @ p1 + T1[:3, 3]) - (T2[:3, :3] @ p2 + T2[:3, 3])
    return a * float(diff @ diff)


def ba_golden(rng):
    """Bundle-adjustment pieces by finite differences and dense assembly:
    the retraction T Exp([rho; phi]) (expm), its 12x6 Jacobian at 0, the pair
    residual r = a |T1 p1 - T2 p2|^2 with the Jacobians the reference supplies
    (d r / d vec(T) divided by 2a: IRLS_Cost_CPU.hpp:339-351), and the normal
    equations H = sum j j^T, g = sum r j, F = sum r^2 / 2 of a tiny window."""
    out = {"plus": [], "plus_jac": [], "pair": [], "normal": None}
    for scale in (1e-9, 1e-4, 0.05, 0.7, 2.0):
        T = _rand_pose(rng)
        d = rng.normal(size=6)
        d *= scale / np.linalg.norm(d)
        out["plus"].append({"T": T[:3, :4].ravel().tolist(), "d": d.tolist(), "out": _plus(T, d)[:3, :4].ravel().tolist()})
    for _ in range(3):
        T = _rand_pose(rng)
        h = 1e-6
        J = np.zeros((12, 6))
        for c in range(6):
            e = np.zeros(6)
            e[c] = h
            J[:, c] = ((_plus(T, e) - _plus(T, -e))[:3, :4].ravel()) / (2 * h)
        out["plus_jac"].append({"T": T[:3, :4].ravel().tolist(), "J": J.ravel().tolist()})
    for _ in range(4):
        T1, T2 = _rand_pose(rng), _rand_pose(rng)
        p1, p2 = rng.normal(size=3) * 4, rng.normal(size=3) * 4
        a = float(rng.uniform(0.01, 1.0))
        r = _residual(T1, T2, p1, p2, a)
        J1, J2 = np.zeros(12), np.zeros(12)
        h = 1e-5
        for k in range(12):
            E = np.zeros((4, 4))
            E[k // 4, k % 4] = h
            J1[k] = (_residual(T1 + E, T2, p1, p2, a) - _residual(T1 - E, T2, p1, p2, a)) / (2 * h) / (2 * a)
            J2[k] = (_residual(T1, T2 + E, p1, p2, a) - _residual(T1, T2 - E, p1, p2, a)) / (2 * h) / (2 * a)
        out["pair"].append({"T1": T1[:3, :4].ravel().tolist(), "T2": T2[:3, :4].ravel().tolist(),
                            "p1": p1.tolist(), "p2": p2.tolist(), "a": a, "r": r,
                            "J1": J1.tolist(), "J2": J2.tolist()})
    # normal equations of a 3-frame window, 12 pairs on frame pairs (0,1),(1,2),(0,2)
    poses = [_rand_pose(rng, 1.0) for _ in range(3)]
    fr = [(0, 1), (1, 2), (0, 2)]
    pairs = []
    for k in range(12):
        f1, f2 = fr[k % 3]
        p1 = rng.normal(size=3).astype(np.float32)
        p2 = rng.normal(size=3).astype(np.float32)
        a = np.float32(rng.uniform(0.05, 1.0))
        pairs.append((f1, f2, p1, p2, a))
    n = 18
    H = np.zeros((n, n))
    g = np.zeros(n)
    F = 0.0
    h = 1e-6
    for f1, f2, p1, p2, a in pairs:
        p1d, p2d, ad = p1.astype(np.float64), p2.astype(np.float64), float(a)
        r = _residual(poses[f1], poses[f2], p1d, p2d, ad)
        js = {}
        for f in (f1, f2):
            j = np.zeros(6)
            for c in range(6):
                e = np.zeros(6)
                e[c] = h
                Pp = [P.copy() for P in poses]
                Pm = [P.copy() for P in poses]
                Pp[f] = _plus(poses[f], e)
                Pm[f] = _plus(poses[f], -e)
                j[c] = (_residual(Pp[f1], Pp[f2], p1d, p2d, ad) - _residual(Pm[f1], Pm[f2], p1d, p2d, ad)) / (2 * h) / (2 * ad)
            js[f] = j
        F += 0.5 * r * r
        for u in (f1, f2):
            g[6 * u:6 * u + 6] += r * js[u]
            for w in (f1, f2):
                H[6 * u:6 * u + 6, 6 * w:6 * w + 6] += np.outer(js[u], js[w])
    out["normal"] = {"poses": [P[:3, :4].ravel().tolist() for P in poses],
                     "pair_frames": [[f1, f2] for f1, f2, _, _, _ in pairs],
                     "p1": [p1.tolist() for _, _, p1, _, _ in pairs], "p2": [p2.tolist() for _, _, _, p2, _ in pairs],
                     "a": [float(a) for *_, a in pairs], "H": H.ravel().tolist(), "g": g.tolist(), "F": F}
    return out


def main():
    rng = np.random.default_rng(20261202)
    json.dump(params_golden(), open(os.path.join(HERE, "params_yaml.json"), "w"), indent=1, sort_keys=True)
    json.dump(cubic_golden(rng), open(os.path.join(HERE, "cubic_roots.json"), "w"), indent=1)
    json.dump(se3_golden(rng), open(os.path.join(HERE, "se3_exp_log.json"), "w"), indent=1)
    json.dump(assoc_golden(rng), open(os.path.join(HERE, "assoc_tiny.json"), "w"))
    json.dump(flow_step_golden(rng), open(os.path.join(HERE, "flow_step.json"), "w"), indent=1)
    json.dump(ba_golden(np.random.default_rng(20261203)), open(os.path.join(HERE, "ba_pieces.json"), "w"))
    ico = icosahedral_table()
    if ico is not None:
        json.dump(ico, open(os.path.join(HERE, "icosahedral_rotations.json"), "w"))
    print("golden vectors written")


if __name__ == "__main__":
    main()
