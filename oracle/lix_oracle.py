"""lix_oracle.py -- CPU restatement of the learned index's arithmetic (TEST INFRASTRUCTURE ONLY).

The learned index has no counterpart in the reference; what it must reproduce is the
reference's suffix order and interval sizes (FMDIndex::count over backward search,
src/FMDIndex.cpp:163-176).  This module states, in numpy:

  * suffix keys: the 16-symbol prefix of each suffix packed 3 bits per symbol
    (past the end = 0 < '\\0' = 1 < A = 2 < C = 3 < G = 4 < T = 5 < N = 6 < other = 7),
    first symbol most significant, so keys in SA order are non-decreasing;
  * exact lower bounds lb(x) = first SA row whose key is >= x (numpy.searchsorted);
  * the piecewise-linear root and the model inputs (segment, in-segment coordinate t,
    symbols 1..15 as (sym - 3.5) / 2);
  * the 2-layer MLP (16 -> 64 ReLU -> 1) forward and backward of the mean squared error
    in float64, and the bf16 inference arithmetic (bf16-rounded inputs and first-layer
    weights, round to nearest even).

Checks that pin it (tests/test_lix_oracle.py): the analytic gradient against central
finite differences, keys against a per-suffix Python loop, bf16 rounding against
torch.bfloat16, lower bounds against brute force.  Only tests import this module.
"""
import numpy as np

SYM3 = np.full(256, 7, np.uint64)
for _b, _s in ((0, 1), (1, 2), (2, 3), (4, 4), (8, 5), (15, 6)):
    SYM3[_b] = _s

N_PARAMS = 16 * 64 + 64 + 64 + 1


def suffix_keys(bd, sa):
    """key of suffix sa[i] of text bd for every i (vectorized over 16 symbol positions)"""
    bd = np.asarray(bd, np.uint8)
    sa = np.asarray(sa, np.int64)
    n = len(bd)
    k = np.zeros(len(sa), np.uint64)
    for t in range(16):
        p = sa + t
        sym = np.where(p < n, SYM3[bd[np.minimum(p, n - 1)]], np.uint64(0))
        k = (k << np.uint64(3)) | sym
    return k


def key_of(bases):
    """key of 16 nt16 bases (a k-mer), as the device packs a read window"""
    k = 0
    for b in bases:
        k = (k << 3) | int(SYM3[int(b)])
    return k


def lower_bounds(ct, keys):
    return np.searchsorted(ct, np.asarray(keys, np.uint64), side="left").astype(np.int64)


def plr(kkey, krank, keys):
    """segment j (largest knot key <= x, 0 below all), t, base and width of each key"""
    keys = np.asarray(keys, np.uint64)
    S = len(kkey) - 1
    j = np.searchsorted(kkey[:S], keys, side="right").astype(np.int64) - 1
    j = np.clip(j, 0, S - 1)
    k0, k1 = kkey[j], kkey[j + 1]
    r0, r1 = krank[j].astype(np.float64), krank[j + 1].astype(np.float64)
    num = (keys - k0).astype(np.float64)
    den = (k1 - k0).astype(np.float64)
    ok = (k1 > k0) & (keys > k0)
    t = np.where(ok, num / np.where(den > 0, den, 1.0), 0.0)
    t = np.minimum(t, 1.0)
    width = np.maximum(r1 - r0, 1.0)
    base = r0 + t * (r1 - r0)
    return j, t, base, width


def features(keys, t):
    keys = np.asarray(keys, np.uint64)
    f = np.zeros((len(keys), 16), np.float32)
    f[:, 0] = t.astype(np.float32)
    for k in range(1, 16):
        s = ((keys >> np.uint64(3 * (15 - k))) & np.uint64(7)).astype(np.float32)
        f[:, k] = (s - np.float32(3.5)) * np.float32(0.5)
    return f


def unpack(W):
    W = np.asarray(W, np.float64)
    W1 = W[:1024].reshape(16, 64)
    b1 = W[1024:1088]
    W2 = W[1088:1152]
    b2 = W[1152]
    return W1, b1, W2, b2


def mlp_forward(W, X):
    W1, b1, W2, b2 = unpack(W)
    X = np.asarray(X, np.float64)
    A = X @ W1 + b1
    H = np.maximum(A, 0.0)
    return H @ W2 + b2, A, H


def mse_loss(W, X, T):
    out, _, _ = mlp_forward(W, X)
    d = out - np.asarray(T, np.float64)
    return float(np.mean(d * d))


def mse_grad(W, X, T):
    """gradient of mean((out - T)^2) with respect to the 1153 parameters"""
    X = np.asarray(X, np.float64)
    out, A, H = mlp_forward(W, X)
    B = len(X)
    dout = 2.0 * (out - np.asarray(T, np.float64)) / B
    W1, b1, W2, b2 = unpack(W)
    g = np.zeros(N_PARAMS, np.float64)
    g[1088:1152] = H.T @ dout
    g[1152] = dout.sum()
    dA = np.outer(dout, W2) * (A > 0)
    g[1024:1088] = dA.sum(0)
    g[:1024] = (X.T @ dA).reshape(-1)
    return g


def bf16(x):
    """float32 -> bfloat16 (round to nearest even) -> float32"""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32)


def infer_bf16(W, X):
    """model output with bf16 inputs and first-layer weights, float64 sums"""
    W1, b1, W2, b2 = unpack(W)
    Xb = bf16(np.asarray(X, np.float32)).astype(np.float64)
    W1b = bf16(W1.astype(np.float32)).astype(np.float64)
    H = np.maximum(Xb @ W1b + b1, 0.0)
    return H @ W2 + b2
