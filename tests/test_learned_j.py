"""Pins the oracle's learned-J filtered mode (ftdo_filtered_learned) and the
J network's weights before they check the HIP kernel.

  * with weights that force a constant J, the mode reduces to the reference's
    v1 solve with that J (ftdo_filtered with J_force), bit for bit;
  * an eps_ no term can meet forces the fallback everywhere, which reduces it
    to the full correction (J = m-1);
  * the fallback flag agrees with |a'_k x_0|, |c'_k x_N| > eps_ recomputed by
    an independent modified-Thomas restatement (the training script's) and
    the exact solution;
  * the network agrees with a float64 matrix evaluation of the same weights
    away from integer boundaries; lg agrees with log2 where it is exact;
  * the kernel's weight header holds the same float32 values as the JSON.
"""
import os
import re

import numpy as np
import pytest

import oracle as O
from filtered_tdma_amd import cases
from filtered_tdma_amd.learned_j import train as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "filtered_tdma_amd", "csrc", "ftdma_mlpj_weights.h")


def const_weights(J0, H=32, margin=1):
    """All weights zero, b3 = J0 - margin - 0.5: the network answers J0."""
    w = np.zeros(H * H + 8 * H + 1, dtype=np.float32)
    w[-1] = np.float32(J0 - margin - 0.5)
    return w, H, margin


@pytest.mark.parametrize("n_sys,n_row,P,cyclic", [(37, 64, 4, False), (20, 131, 8, True),
                                                  (9, 24, 2, False)])
@pytest.mark.parametrize("J0", [0, 1, 3, 100])
def test_constant_J_equals_v1_with_that_J(n_sys, n_row, P, cyclic, J0):
    A, B, C, D = cases.general_batch(n_sys, n_row, seed=7)
    x, J = O.filtered_learned(A, B, C, D, P, 1e300, eps_direct=True, cyclic=cyclic,
                              weights=const_weights(J0))
    ms = [O.para_range(0, n_row - 1, P, p) for p in range(P)]
    ms = [e - s + 1 for s, e in ms]
    Jp = [min(J0, m - 1) for m in ms]
    # no fallback with eps_ = 1e300; the clamp to [0, m-1] is applied per partition
    assert (J >> 16).max() == 0
    for p in range(P):
        assert (J[p] == Jp[p]).all()
    ref = O.filtered(A, B, C, D, P, 1, 1e300, eps_direct=True, cyclic=cyclic, J_force=Jp)
    np.testing.assert_array_equal(x, ref)


def test_unreachable_eps_falls_back_to_full_correction():
    n_sys, n_row, P = 50, 96, 4
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 0.1, 5.0, seed=3)
    x, J = O.filtered_learned(A, B, C, D, P, 1e-300, eps_direct=True, weights=const_weights(0))
    assert ((J >> 16) == 1).all() and ((J & 0xFFFF) == 23).all()
    ref = O.filtered(A, B, C, D, P, 1, 1e-300, eps_direct=True, J_force=[23] * P)
    np.testing.assert_array_equal(x, ref)


def test_fallback_flag_matches_the_dropped_terms():
    """Independent restatement: the first dropped a- and c-terms from
    modified-Thomas coefficients (training script) and the exact x."""
    n_sys, n_row, P = 400, 128, 4
    m = n_row // P
    A, B, C, D, _ = cases.toeplitz_batch(n_sys, n_row, 0.05, 3.0, seed=11)
    D = D * (10.0 ** np.linspace(-6, 2, n_sys))[None, :]
    eps = 1e-9
    w = O.mlp_weights()
    x, J = O.filtered_learned(A, B, C, D, P, eps, eps_direct=True, weights=w)
    xe = T.thomas(A, B, C, D)
    n_checked = 0
    for p in range(P):
        sl = slice(p * m, (p + 1) * m)
        Am, Cm = T.modified(A[sl], B[sl], C[sl], D[sl])
        x0, xN = xe[p * m], xe[(p + 1) * m - 1]
        feats = np.stack([np.minimum(np.maximum(np.abs(A[sl][1:] / B[sl][1:]),
                                                np.abs(C[sl][1:] / B[sl][1:])).max(0), 1.0),
                          T.lg(D[sl][0]), T.lg(D[sl][-1]), T.lg(np.abs(D[sl]).max(0)),
                          np.full(n_sys, T.lg(eps))], axis=1).astype(np.float32)
        for i in range(n_sys):
            Jl = min(max(O.mlp_J(feats[i], w), 0), m - 1)
            kL = min(Jl, m - 2) + 1
            kR = max(m - 1 - Jl, 1) - 1
            ta = abs(Am[kL, i] * x0[i]) if kL <= m - 2 else 0.0
            tc = abs(Cm[kR, i] * xN[i]) if kR >= 1 else 0.0
            t = max(ta, tc)
            if abs(t - eps) < 1e-6 * eps:
                continue  # too close to call between two restatements
            fb = int(t > eps)
            assert (J[p, i] >> 16) == fb, (p, i, t, eps)
            assert (J[p, i] & 0xFFFF) == (m - 1 if fb else Jl)
            n_checked += 1
    assert n_checked > 0.99 * P * n_sys
    # accuracy: the learned solve stays close to the exact one
    err = np.max(np.abs(x - xe), axis=0)
    assert np.all(err <= 1e3 * eps), err.max()


def test_network_matches_float64_evaluation():
    w, H, margin = O.mlp_weights()
    o = np.cumsum([0, 5 * H, H, H * H, H, H])
    W1 = w[o[0]:o[1]].reshape(5, H).astype(np.float64)
    b1 = w[o[1]:o[2]].astype(np.float64)
    W2 = w[o[2]:o[3]].reshape(H, H).astype(np.float64)
    b2 = w[o[3]:o[4]].astype(np.float64)
    W3 = w[o[4]:o[5]].astype(np.float64)
    b3 = float(w[o[5]])
    rng = np.random.default_rng(5)
    F = np.stack([rng.uniform(0, 1, 300), rng.uniform(-40, 20, 300), rng.uniform(-40, 20, 300),
                  rng.uniform(-40, 20, 300), rng.uniform(-50, -10, 300)], 1).astype(np.float32)
    out = np.maximum(np.maximum(F @ W1 + b1, 0) @ W2 + b2, 0) @ W3 + b3
    checked = 0
    for f, o in zip(F, out):
        if abs(o - np.round(o)) < 1e-3:
            continue
        assert O.mlp_J(f, (w, H, margin)) == int(np.ceil(o)) + margin
        checked += 1
    assert checked > 250


def test_lg_feature():
    for k in range(-1070, 1020, 37):
        assert O.lg(2.0 ** k) == float(k)
        assert O.lg(-(2.0 ** k)) == float(k)
    assert O.lg(0.0) == -1100.0
    assert O.lg(1.5) == 0.5 and O.lg(3.0) == 1.5
    xs = 10.0 ** np.random.default_rng(2).uniform(-300, 300, 2000)
    for x in xs:
        v = O.lg(x)
        tol = 1.2e-7 * abs(v) + 1e-6  # the float32 rounding of the feature
        assert np.log2(x) - 0.0861 - tol <= v <= np.log2(x) + tol
    np.testing.assert_array_equal(np.float32(T.lg(xs)), np.array([O.lg(x) for x in xs],
                                                                   dtype=np.float32))


def test_header_weights_equal_json():
    w, H, margin = O.mlp_weights()
    text = open(HEADER).read()

    def arr(name):
        body = re.search(name + r"\[\d+\] = \{([^}]*)\}", text).group(1)
        return np.array([np.float32(float(v.strip().rstrip("f"))) for v in body.split(",")],
                        dtype=np.float32)

    K = arr("kMlpK").reshape(H, 8 + H)  # per unit: b1, W1 column, 0, 0, W2 row
    assert (K[:, 6:8] == 0).all()
    got = np.concatenate([K[:, 1:6].T.reshape(-1), K[:, 0], K[:, 8:].reshape(-1),
                          arr("kMlpB2"), arr("kMlpW3"), arr("kMlpB3")])
    np.testing.assert_array_equal(got, w)
    assert re.search(r"kMlpMargin = (\d+);", text).group(1) == str(margin)
    assert re.search(r"kMlpH = (\d+);", text).group(1) == str(H)


def test_training_reproduces_the_committed_weights():
    """The weights in mlp_j.json (and so in the kernel's header) are what
    learned_j/train.py produces: seeded, deterministic."""
    params, stats = T.train()
    w, H, margin = O.mlp_weights()
    flat = np.concatenate([np.asarray(params[0], np.float32).reshape(-1),
                           np.asarray(params[1], np.float32), np.asarray(params[2], np.float32).reshape(-1),
                           np.asarray(params[3], np.float32), np.asarray(params[4], np.float32).reshape(-1),
                           np.asarray(params[5], np.float32).reshape(-1)])
    assert H == T.H and margin == T.MARGIN
    np.testing.assert_array_equal(flat, w)
