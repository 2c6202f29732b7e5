"""Train the J predictor of the learned-J filtered mode and write its weights.

The filtered solve's final pass corrects x_j = d'_j - a'_j x_0 - c'_j x_N on
every interior row of a partition; the reference truncates the a-term past
row J and the c-term before row m-1-J, with one J per partition from the
analytic bound cal_J (filtered_tdma.cpp:63-115, filtered_tdma_kernels.cuh:
220-261). Here a small MLP (5 -> 32 -> 32 -> 1, ReLU) predicts J per (system,
partition) from

    rho          max over the partition's rows 1..m-1 of |a/b|, |c/b| (<= 1)
    lg|d_0|      first right-hand-side row of the partition
    lg|d_m-1|    last right-hand-side row
    lg max|d|    largest right-hand-side row
    lg eps_

where lg x = e + f for |x| = (1 + f) 2^e (the piecewise-linear log2 read
off the float64 bit pattern; lg 0 = -1100): exact on any machine, so the
kernel and the CPU restatement compute the same features bit for bit.

and is trained on the smallest J whose dropped terms all stay below eps_
(the "minimal J observed in exact solves"): with the partition's modified
Thomas coefficients a'_j, c'_j (rows 0 and 1 divided by b, full a- and
c-recurrences) and the exact x_0, x_N of the whole system,

    J_min = min J such that |a'_j x_0| <= eps_ for every j > J and
            |c'_j x_N| <= eps_ for every j < m-1-J.

The kernel checks the first dropped term of its predicted J against eps_
and falls back to the full correction when it exceeds it, so the MLP only
decides how much arithmetic is skipped, never the accuracy bound.

    python -m filtered_tdma_amd.learned_j.train     # writes mlp_j.json and
                                                    # csrc/ftdma_mlpj_weights.h

numpy only, deterministic (seeded); a few seconds on a laptop core.
"""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
H = 32


def lg(x):
    """Piecewise-linear log2 of |x| from its float64 bit pattern."""
    x = np.abs(np.asarray(x, dtype=np.float64))
    m, e = np.frexp(x)  # x = m 2^e, m in [0.5, 1)
    v = (e - 1).astype(np.float64) + (2.0 * m - 1.0)
    return np.where(x == 0.0, -1100.0, v)


def thomas(a, b, c, d):
    """Exact solve, vectorized over the trailing axis (rows first)."""
    n = d.shape[0]
    cp = np.empty_like(d)
    dp = np.empty_like(d)
    cp[0] = c[0] / b[0]
    dp[0] = d[0] / b[0]
    for j in range(1, n):
        r = 1.0 / (b[j] - a[j] * cp[j - 1])
        cp[j] = r * c[j]
        dp[j] = r * (d[j] - a[j] * dp[j - 1])
    x = np.empty_like(d)
    x[-1] = dp[-1]
    for j in range(n - 2, -1, -1):
        x[j] = dp[j] - cp[j] * x[j + 1]
    return x


def modified(a, b, c, d):
    """a', c' of one partition (rows 0, 1 divided by b; forward a-fill and
    c; backward a and c recurrences), vectorized over the trailing axis."""
    m = d.shape[0]
    A = np.empty_like(d)
    C = np.empty_like(d)
    for j in range(2):
        A[j] = a[j] / b[j]
        C[j] = c[j] / b[j]
    for j in range(2, m):
        inv = 1.0 / (b[j] - a[j] * C[j - 1])
        C[j] = inv * c[j]
        A[j] = -inv * a[j] * A[j - 1]
    for j in range(m - 3, 0, -1):
        A[j] = A[j] - C[j] * A[j + 1]
        C[j] = -C[j] * C[j + 1]
    return A, C


def channel_modes(rng, n, n_sys):
    """Wall-normal pressure systems of a channel (the coefficients of
    cases.channel_poisson): a tanh-stretched grid of random stretching,
    a_k = 1/(dz_k dmz_k), c_k = 1/(dz_k dmz_{k+1}) (0 at the walls),
    b = -(a + c) - lambda with lambda = kx^2 + ky^2 of random modes (about
    one system in 16 has lambda = 0, the mode that is pinned in a channel
    solve and left singular here only in its row 0, which is made 1)."""
    from filtered_tdma_amd.cases import tanh_faces

    xf = tanh_faces(n, rng.uniform(0.5, 3.0), 2.0)
    xc = np.empty(n + 2)
    xc[1:n + 1] = 0.5 * (xf[:-1] + xf[1:])
    xc[0] = xf[0] - (xc[1] - xf[0])
    xc[n + 1] = xf[n] + (xf[n] - xc[n])
    dz = np.empty(n + 2)
    dz[1:n + 1] = xf[1:] - xf[:-1]
    dz[0], dz[n + 1] = dz[1], dz[n]
    dmz = np.empty(n + 2)
    dmz[1:] = xc[1:] - xc[:-1]
    dmz[0] = dmz[1]
    k = np.arange(n)
    a = 1.0 / (dz[k + 1] * dmz[k + 1])
    c = 1.0 / (dz[k + 1] * dmz[k + 2])
    a[0] = 0.0
    c[n - 1] = 0.0
    lam = np.exp(rng.uniform(np.log(1e-2), np.log(1e6), n_sys))
    lam[rng.uniform(size=n_sys) < 1.0 / 16] = 0.0
    A = np.broadcast_to(a[:, None], (n, n_sys)).copy()
    C = np.broadcast_to(c[:, None], (n, n_sys)).copy()
    B = -(a + c)[:, None] - lam[None, :]
    pin = lam == 0.0
    A[0, pin], B[0, pin], C[0, pin] = 0.0, 1.0, 0.0
    return A, B, C


def samples(rng, count):
    """Systems of P partitions: Toeplitz (a = c = -r, b = 1 + 2r) with
    log-uniform r and normal right-hand sides, the same with r uniform over a
    random range and uniform right-hand sides (the standalone batched
    workload's kind), non-symmetric walls-like (a != c), general dominant
    rows, and channel wall-normal modes; RHS of random scale; eps_
    log-uniform."""
    X, Y, S, Ms = [], [], [], []
    while len(Y) < count:
        n_sys = 512
        m = int(rng.choice([8, 16, 32, 48, 64, 128, 256, 512]))
        P = int(rng.choice([2, 4, 8] if m <= 64 else [2, 4]))
        n = m * P
        kind = rng.integers(5)
        if kind == 3:
            lo = np.exp(rng.uniform(np.log(0.05), np.log(50.0)))
            r = rng.uniform(lo, lo * np.exp(rng.uniform(np.log(2.0), np.log(100.0))), n_sys)
            a = np.broadcast_to(-r, (n, n_sys)).copy()
            c = a.copy()
            b = 1.0 + 2.0 * r + np.zeros((n, n_sys))
        elif kind == 4:
            a, b, c = channel_modes(rng, n, n_sys)
        elif kind == 0:
            r = np.exp(rng.uniform(np.log(0.05), np.log(500.0), n_sys))
            a = np.broadcast_to(-r, (n, n_sys)).copy()
            c = a.copy()
            b = 1.0 + 2.0 * r + np.zeros((n, n_sys))
        elif kind == 1:
            r = np.exp(rng.uniform(np.log(0.05), np.log(500.0), (n, n_sys)))
            s = rng.uniform(0.5, 1.5, (n, n_sys))
            a, c = -r * s, -r / s
            b = (np.abs(a) + np.abs(c)) * rng.uniform(1.0001, 1.5, (n, n_sys))
        else:
            a = rng.uniform(-1, 1, (n, n_sys))
            c = rng.uniform(-1, 1, (n, n_sys))
            b = (np.abs(a) + np.abs(c)) * rng.uniform(1.01, 3.0, (n, n_sys))
            b *= np.where(rng.uniform(size=(n, n_sys)) < 0.5, -1.0, 1.0)
        scale = 10.0 ** rng.uniform(-8, 4, n_sys)
        d = rng.standard_normal((n, n_sys)) * scale
        if kind == 3:
            d = rng.uniform(-1.0, 1.0, (n, n_sys)) * scale
        elif rng.uniform() < 0.3:  # smooth right-hand sides
            t = np.linspace(0, 1, n)[:, None]
            d = np.sin(np.pi * t * rng.uniform(1, 8, n_sys)) * scale
        eps = 10.0 ** rng.uniform(-14, -3, n_sys)
        x = thomas(a, b, c, d)
        for p in range(P):
            sl = slice(p * m, (p + 1) * m)
            A, C = modified(a[sl], b[sl], c[sl], d[sl])
            x0, xN = x[p * m], x[(p + 1) * m - 1]
            ta = np.abs(A * x0) > eps  # a-terms above eps_ (rows 1..m-2 matter)
            tc = np.abs(C * xN) > eps
            jmin = np.zeros(n_sys)
            rows = np.arange(m)[:, None]
            # a-term at row j dropped when j > J: need J >= last row with ta
            la = np.where(ta[1:m - 1], rows[1:m - 1], 0).max(axis=0)
            # c-term at row j dropped when j < m-1-J: need J >= m-1-(first row with tc)
            fc = np.where(tc[1:m - 1], rows[1:m - 1], m - 1).min(axis=0)
            jmin = np.maximum(la, (m - 1) - fc)
            rho = np.maximum(np.abs(a[sl][1:] / b[sl][1:]), np.abs(c[sl][1:] / b[sl][1:])).max(axis=0)
            dd = d[sl]
            feats = np.stack([np.minimum(rho, 1.0), lg(dd[0]), lg(dd[-1]),
                              lg(np.abs(dd).max(axis=0)), lg(eps)], axis=1)
            feats = feats.astype(np.float32).astype(np.float64)
            X.append(feats)
            Y.append(np.minimum(jmin, m - 1).astype(np.float64))
            # J_min = m - 2 means "the whole partition": the depth the
            # features call for may be larger, and the kernel clamps J to
            # m - 1, so such a target is only a lower bound
            S.append(jmin >= m - 2)
            Ms.append(np.full(n_sys, m - 1.0))
        X_all = np.concatenate(X)
        if X_all.shape[0] >= count:
            break
    X = np.concatenate(X)[:count]
    Y = np.concatenate(Y)[:count]
    S = np.concatenate(S)[:count]
    Mx = np.concatenate(Ms)[:count]
    return X, Y, S, Mx


# rows added to the rounded prediction (also applied by the kernel)
MARGIN = 1

# feature normalisation baked into the first layer
MU = np.array([0.25, -6.6, -6.6, -5.0, -28.2])
SD = np.array([0.15, 13.3, 13.3, 13.3, 10.6])


def train(seed=0, count=120000, epochs=60):
    rng = np.random.default_rng(seed)
    X, Y, S, Mx = samples(rng, count)
    Xn = (X - MU) / SD
    Yn = Y / 64.0
    W1 = rng.standard_normal((5, H)) * np.sqrt(2 / 5)
    b1 = np.zeros(H)
    W2 = rng.standard_normal((H, H)) * np.sqrt(2 / H)
    b2 = np.zeros(H)
    W3 = rng.standard_normal((H, 1)) * np.sqrt(1 / H)
    b3 = np.zeros(1)
    params = [W1, b1, W2, b2, W3, b3]
    mom = [np.zeros_like(p) for p in params]
    vel = [np.zeros_like(p) for p in params]
    lr, beta1, beta2, t = 3e-3, 0.9, 0.999, 0
    bs = 512
    for ep in range(epochs):
        perm = rng.permutation(len(Yn))
        for k in range(0, len(Yn), bs):
            idx = perm[k:k + bs]
            x, y, sat = Xn[idx], Yn[idx][:, None], S[idx][:, None]
            z1 = x @ W1 + b1
            h1 = np.maximum(z1, 0)
            z2 = h1 @ W2 + b2
            h2 = np.maximum(z2, 0)
            out = h2 @ W3 + b3
            # asymmetric loss: predicting below J_min costs a fallback
            # (full correction), above it only extra multiply-adds
            err = out - y
            w = np.where(err < 0, 12.0, np.where(sat, 0.0, 1.0))
            g = 2 * w * err / len(idx)
            gW3 = h2.T @ g
            gb3 = g.sum(0)
            g2 = (g @ W3.T) * (z2 > 0)
            gW2 = h1.T @ g2
            gb2 = g2.sum(0)
            g1 = (g2 @ W2.T) * (z1 > 0)
            gW1 = x.T @ g1
            gb1 = g1.sum(0)
            t += 1
            for i, gr in enumerate([gW1, gb1, gW2, gb2, gW3, gb3]):
                mom[i] = beta1 * mom[i] + (1 - beta1) * gr
                vel[i] = beta2 * vel[i] + (1 - beta2) * gr * gr
                mh = mom[i] / (1 - beta1 ** t)
                vh = vel[i] / (1 - beta2 ** t)
                params[i] -= lr * mh / (np.sqrt(vh) + 1e-8)
        lr *= 0.95
    W1, b1, W2, b2, W3, b3 = params
    # fold the normalisation and the output scale into the weights
    W1f = W1 / SD[:, None]
    b1f = b1 - (MU / SD) @ W1
    W3f = W3 * 64.0
    b3f = b3 * 64.0
    pred = np.maximum(np.maximum(X @ W1f + b1f, 0) @ W2 + b2, 0) @ W3f + b3f
    jp = np.minimum(np.maximum(np.ceil(pred[:, 0]) + MARGIN, 0), Mx)  # the kernel's clamp
    # the analytic bound per system (cal_J's rhs-bound formula with this
    # system's rho and boundary maxima) on the same samples, for comparison
    rho = np.clip(X[:, 0], 1e-12, 0.4999999)
    lam = (1 + np.sqrt(1 - 4 * rho * rho)) / 2
    q = rho / lam
    mb = np.maximum(2.0 ** X[:, 1], 2.0 ** X[:, 2])
    Bv = q * (2 + q) / ((1 - q) * (1 - 2 * rho)) * mb
    with np.errstate(divide="ignore", invalid="ignore"):
        jc = np.floor(np.log(2.0 ** X[:, 4] / Bv) / np.log(q)) + 1
    jc = np.where((X[:, 0] >= 0.5) | ~np.isfinite(jc), Mx, np.clip(jc, 0, Mx))
    stats = {"samples": int(len(Y)), "mean_J_min": float(Y.mean()),
             "mean_J_pred": float(jp.mean()),
             "mean_J_cal": float(jc.mean()),
             "under_predicted_fraction": float((jp < Y).mean()),
             "cal_under_fraction": float((jc < Y).mean())}
    return [W1f, b1f, W2, b2, W3f, b3f], stats


def write(params, stats):
    W1, b1, W2, b2, W3, b3 = (np.asarray(p, dtype=np.float32) for p in params)
    with open(os.path.join(HERE, "mlp_j.json"), "w") as f:
        json.dump({"layers": [5, H, H, 1], "hidden": H, "W1": W1.tolist(), "b1": b1.tolist(),
                   "W2": W2.tolist(), "b2": b2.tolist(), "W3": W3[:, 0].tolist(),
                   "b3": float(b3[0]), "margin": MARGIN, "train": stats}, f)

    def arr(name, a):
        vals = ", ".join(f"{v:.9e}f" for v in np.asarray(a, dtype=np.float32).reshape(-1))
        return f"__constant__ float {name}[{a.size}] = {{{vals}}};\n"

    with open(os.path.join(CSRC, "ftdma_mlpj_weights.h"), "w") as f:
        f.write("// Generated by filtered_tdma_amd/learned_j/train.py: weights of the J\n"
                f"// predictor (5 -> {H} -> {H} -> 1, ReLU, FP32). Features: rho, lg|d_0|,\n"
                "// lg|d_m-1|, lg max|d|, lg eps_ with lg the piecewise-linear log2 of the\n"
                "// float64 bit pattern (normalisation folded into W1, b1; output in rows).\n"
                "#pragma once\n")
        f.write(f"constexpr int kMlpH = {H};\n")
        # per first-layer unit k, 8 + H floats: b1[k], W1[0..4][k], two zeros,
        # then row k of W2, so one unit's weights arrive in three wide loads
        K = np.zeros((H, 8 + H), dtype=np.float32)
        K[:, 0] = b1
        K[:, 1:6] = W1.T
        K[:, 8:] = W2
        f.write("// kMlpK[k][0] = b1[k], [1..5] = W1[0..4][k], [8 + l] = W2[k][l]\n")
        f.write(arr("kMlpK", K))
        f.write(arr("kMlpB2", b2))
        f.write(arr("kMlpW3", W3[:, 0]))
        f.write(f"__constant__ float kMlpB3[1] = {{{float(b3[0]):.9e}f}};\n")
        f.write(f"constexpr int kMlpMargin = {MARGIN};\n")


if __name__ == "__main__":
    params, stats = train()
    write(params, stats)
    print(json.dumps(stats))
