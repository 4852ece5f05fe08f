/*
 * msgs_lix.hip -- learned index over the FMD index's suffix order.
 *
 * Keys.  CT[i] is the 16-symbol prefix of suffix SA[i] packed 3 bits per symbol
 * (past the end < '\0' < A < C < G < T < N, first symbol most significant), so CT is
 * non-decreasing in i and lb(x) = first i with CT[i] >= x is the SA rank of the
 * 16-mer x: the first row of its interval, count(x) = lb(x+1) - lb(x).
 *
 * Model.  A piecewise-linear root over S knots taken at rank quantiles (knot j is the
 * key of row j*n/S and its exact lower bound), and a 2-layer MLP (16 -> 64 ReLU -> 1)
 * that predicts the residual of the root inside a segment from the in-segment
 * coordinate and the key's later symbols.  Trained in fp32 on (key, rank) pairs
 * sampled from the index by the kernels below (forward, backward with fixed-order
 * reductions, Adam); inference runs the first layer as bf16 MFMA
 * (v_mfma_f32_16x16x16 bf16) with fp32 accumulation.
 *
 * Exactness.  E_j is the largest |prediction - lb| over the distinct keys present in
 * CT that fall in segment j, measured with the same inference code (the recorded max
 * error; E = max_j E_j is reported).  A query in segment j searches [pred-E_j, pred+E_j] and
 * checks CT[lo-1] < x <= CT[lo]; when the check fails (keys absent from the index can
 * fall outside the bound) it falls back to a binary search over all of CT.  Results
 * are therefore exact lower bounds; tests judge them against FMD backward search.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "msgs_capi_impl.h"
#include "msgs_lix.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint64_t sym3(uint8_t b) {
    switch (b) {
    case 0: return 1; case 1: return 2; case 2: return 3; case 4: return 4; case 8: return 5; case 15: return 6;
    default: return 7;
    }
}

/* fp32 -> bf16, round to nearest even */
__device__ __forceinline__ short toBf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float fromBf16(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }

/* key of every SA row: the 16 text bytes at SA[i] from three aligned words (the BD text is
 * padded by 64 bytes), not 16 byte loads at a random address */
__global__ void k_ct(DevIndex ix, uint64_t* ct) {
    const uint8_t* bd = ix.bd;
    const int64_t n = ix.n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = saAt(ix, i);
        const uint64_t* w = (const uint64_t*)(bd + (p & ~(int64_t)7));
        const uint64_t w0 = w[0], w1 = w[1], w2 = w[2];
        const uint32_t sh = 8u * (uint32_t)(p & 7);
        const uint64_t lo = sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0;
        const uint64_t hi = sh ? (w1 >> sh) | (w2 << (64 - sh)) : w1;
        uint64_t k = 0;
        for (int t = 0; t < 16; ++t) {
            const uint8_t b = (uint8_t)((t < 8 ? lo : hi) >> (8 * (t & 7)));
            k = (k << 3) | (p + t < n ? sym3(b) : 0);
        }
        ct[i] = k;
    }
}

__device__ __forceinline__ int64_t lbFull(const uint64_t* ct, int64_t n, uint64_t x) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (ct[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ void k_knots(const uint64_t* ct, int64_t n, int32_t S, uint64_t* kKey, int64_t* kRank) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= S; j += (int64_t)gridDim.x * blockDim.x) {
        if (j == S) { kKey[S] = ~0ull; kRank[S] = n; continue; }
        const uint64_t x = ct[j * n / S];
        kKey[j] = x;
        kRank[j] = lbFull(ct, n, x);
    }
}

/* kBucket[b] = first knot j with kKey[j] >= b << (48 - kBucketBits) */
__global__ void k_buckets(const uint64_t* kKey, int32_t S, uint32_t* kBucket) {
    const int64_t nb = ((int64_t)1 << kBucketBits) + 1;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t x = (uint64_t)b << (48 - kBucketBits);
        int32_t lo = 0, hi = S;   /* kKey[S] = all ones >= x */
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (kKey[mid] < x) lo = mid + 1; else hi = mid;
        }
        kBucket[b] = (uint32_t)lo;
    }
}

/* segment of x: largest j < S with kKey[j] <= x (0 when x is below every knot); the
 * first knot above x lies between the bucket bounds of x's top bits */
__device__ __forceinline__ int32_t segOf(const LixView& m, uint64_t x) {
    const uint64_t b = x >> (48 - kBucketBits);
    if (b >= ((uint64_t)1 << kBucketBits)) return m.S - 1;   /* above every 48-bit key */
    int32_t lo = (int32_t)m.kBucket[b], hi = (int32_t)m.kBucket[b + 1];
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (m.kKey[mid] > x) hi = mid; else lo = mid + 1;
    }
    return lo > 0 ? lo - 1 : 0;
}

struct Plr { double base, width, t; };

__device__ __forceinline__ Plr plrOf(const LixView& m, uint64_t x, int32_t j) {
    const uint64_t k0 = m.kKey[j], k1 = m.kKey[j + 1];
    const double r0 = (double)m.kRank[j], r1 = (double)m.kRank[j + 1];
    double t = 0;
    if (k1 > k0 && x > k0) t = (double)(x - k0) / (double)(k1 - k0);
    if (t > 1) t = 1;
    Plr p;
    p.width = r1 - r0 > 1 ? r1 - r0 : 1;
    p.base = r0 + t * (r1 - r0);
    p.t = t;
    return p;
}

/* the 16 model inputs of key x */
__device__ __forceinline__ void features(uint64_t x, double t, float* f) {
    f[0] = (float)t;
    for (int k = 1; k < 16; ++k) {
        const uint32_t s = (uint32_t)((x >> (3 * (15 - k))) & 7);
        f[k] = ((float)s - 3.5f) * 0.5f;
    }
}

/* fp32 reference forward of the MLP (training and the VALU check) */
__device__ __forceinline__ float mlpFwd(const float* W, const float* f, float* h) {
    const float* W1 = W;
    const float* b1 = W + 16 * 64;
    const float* W2 = b1 + 64;
    const float b2 = W2[64];
    float out = b2;
    for (int u = 0; u < 64; ++u) {
        float a = b1[u];
        for (int k = 0; k < 16; ++k) a += f[k] * W1[k * 64 + u];
        a = a > 0 ? a : 0;
        if (h) h[u] = a;
        out += a * W2[u];
    }
    return out;
}

/* Wave-cooperative bf16 MFMA inference: every lane brings one 16-feature row
 * (zeros when idle) and gets the MLP output of its row.  Must be called by all 64
 * lanes of the wave.  sFeat: this wave's 64x16 bf16 LDS tile; W1b: 16x64 bf16 in LDS. */
__device__ float mlpMfma(const LixView& m, const float* f, const short* W1b) {
    const int lane = threadIdx.x & 63;
    const float* b1 = m.W + 16 * 64;
    const float* W2 = b1 + 64;
    const int g = lane >> 4, c = lane & 15;
    float acc[4][4];
    for (int mt = 0; mt < 4; ++mt)
        for (int j = 0; j < 4; ++j) acc[mt][j] = 0.f;
    for (int mt = 0; mt < 4; ++mt) {
        /* A: row mt*16 + c (that lane's query), features 4g .. 4g+3 */
        bf16x4 A;
        const int src = mt * 16 + c;
        for (int k = 0; k < 16; ++k) {
            const float v = __shfl(f[k], src, 64);
            if ((k >> 2) == g) A[k & 3] = toBf16(v);
        }
        for (int nt = 0; nt < 4; ++nt) {
            /* B: features 4g .. 4g+3 of hidden unit nt*16 + c */
            bf16x4 B;
            for (int j = 0; j < 4; ++j) B[j] = W1b[(4 * g + j) * 64 + nt * 16 + c];
            f32x4 D = {0.f, 0.f, 0.f, 0.f};
            D = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, B, D, 0, 0, 0);
            /* D[j] = hidden unit nt*16 + c of row mt*16 + 4g + j */
            const int u = nt * 16 + c;
            for (int j = 0; j < 4; ++j) {
                float h = D[j] + b1[u];
                h = h > 0 ? h : 0;
                acc[mt][j] += h * W2[u];
            }
        }
    }
    /* sum over the 16 hidden columns held by the lanes of one group */
    for (int mt = 0; mt < 4; ++mt)
        for (int j = 0; j < 4; ++j) {
            float v = acc[mt][j];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            v += __shfl_xor(v, 8, 64);
            acc[mt][j] = v;
        }
    /* row r = mt*16 + 4*g + j sits in group g; lane `lane` takes row `lane` */
    const int mt = lane >> 4, gg = (lane & 15) >> 2, jj = lane & 3;
    float mine = 0.f;
    for (int t = 0; t < 4; ++t)
        for (int j = 0; j < 4; ++j) {
            const float v = __shfl(acc[t][j], gg * 16, 64);
            if (t == mt && j == jj) mine = v;
        }
    return mine + m.W[16 * 64 + 64 + 64];
}

__device__ __forceinline__ void loadW1(const LixView& m, short* W1b) {
    for (int i = threadIdx.x; i < 16 * 64; i += blockDim.x) W1b[i] = toBf16(m.W[i]);
    __syncthreads();
}

/* prediction of lb(x) (rounded) with the MFMA path; useMlp = 0 gives the root alone;
 * seg receives the key's segment */
__device__ __forceinline__ int64_t predict(const LixView& m, uint64_t x, bool active, const short* W1b, bool useMlp,
                                           int32_t* seg) {
    double base = 0, width = 1;
    float f[16];
    for (int k = 0; k < 16; ++k) f[k] = 0.f;
    *seg = 0;
    if (active) {
        const int32_t j = segOf(m, x);
        *seg = j;
        Plr p = plrOf(m, x, j);
        base = p.base; width = p.width;
        features(x, p.t, f);
    }
    double r = 0;
    if (useMlp) r = (double)mlpMfma(m, f, W1b);
    double y = base + r * width;
    if (y < 0) y = 0;
    if (y > (double)m.n) y = (double)m.n;
    return (int64_t)llround(y);
}

/* exact lb(x) by a search bounded by the recorded error, with the full-search fallback */
__device__ __forceinline__ int64_t lbBounded(const LixView& m, uint64_t x, int64_t pred, int64_t E, int32_t* probes,
                                             int32_t* fellBack) {
    int64_t lo = pred - E, hi = pred + E + 1;
    if (lo < 0) lo = 0;
    if (hi > m.n) hi = m.n;
    const int64_t L0 = lo, H0 = hi;
    int32_t pr = 0;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        pr++;
        if (m.ct[mid] < x) lo = mid + 1; else hi = mid;
    }
    /* rows [L0, lo) hold keys < x and rows [lo, H0) keys >= x; check the two ends */
    bool ok = true;
    if (lo == L0 && lo > 0) { pr++; ok = m.ct[lo - 1] < x; }
    if (ok && lo == H0 && lo < m.n) { pr++; ok = m.ct[lo] >= x; }
    *fellBack = 0;
    if (!ok) {
        *fellBack = 1;
        int64_t a = 0, b = m.n;
        while (a < b) {
            int64_t mid = (a + b) >> 1;
            pr++;
            if (m.ct[mid] < x) a = mid + 1; else b = mid;
        }
        lo = a;
    }
    *probes = pr;
    return lo;
}

/* ---- training ---- */

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void k_sample(LixView m, int64_t M, uint64_t seed, float* X, float* T) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = (int64_t)(mix64(seed ^ (uint64_t)i * 0x2545f4914f6cdd1dull) % (uint64_t)m.n);
        const uint64_t x = m.ct[r];
        const int64_t y = lbFull(m.ct, m.n, x);
        const int32_t j = segOf(m, x);
        Plr p = plrOf(m, x, j);
        float f[16];
        features(x, p.t, f);
        for (int k = 0; k < 16; ++k) X[i * 16 + k] = f[k];
        T[i] = (float)(((double)y - p.base) / p.width);
    }
}

/* forward over a mini-batch: hidden activations, dL/dout, and per-sample squared error */
__global__ void k_fwd(const float* W, const float* X, const float* T, int64_t off, int32_t B, float* H, float* Dout,
                      float* Err) {
    int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    const float* f = X + (off + s) * 16;
    float h[64];
    const float out = mlpFwd(W, f, h);
    for (int u = 0; u < 64; ++u) H[s * 64 + u] = h[u];
    const float d = out - T[off + s];
    Dout[s] = 2.f * d / (float)B;
    Err[s] = d * d;
}

/* gradients: one 256-thread block per parameter; thread t sums samples t, t+256, ...
 * in order, then a fixed pairwise tree in LDS (deterministic: the summation order does
 * not depend on scheduling).  Block kLixParams - 1 (b2) also reduces the loss. */
__global__ void __launch_bounds__(256) k_grad(const float* W, const float* X, int64_t off, int32_t B, const float* H,
                                              const float* Dout, const float* Err, float* G, float* loss) {
    __shared__ float sA[256], sE[256];
    const int32_t id = blockIdx.x;
    const int t = threadIdx.x;
    const float* W2 = W + 16 * 64 + 64;
    float a = 0.f, e = 0.f;
    if (id < 16 * 64) {
        const int k = id / 64, u = id % 64;
        const float w2 = W2[u];
        for (int32_t s = t; s < B; s += 256) {
            const float dh = H[s * 64 + u] > 0 ? Dout[s] * w2 : 0.f;
            a += X[(off + s) * 16 + k] * dh;
        }
    } else if (id < 16 * 64 + 64) {
        const int u = id - 16 * 64;
        const float w2 = W2[u];
        for (int32_t s = t; s < B; s += 256) a += H[s * 64 + u] > 0 ? Dout[s] * w2 : 0.f;
    } else if (id < 16 * 64 + 128) {
        const int u = id - 16 * 64 - 64;
        for (int32_t s = t; s < B; s += 256) a += H[s * 64 + u] * Dout[s];
    } else {
        for (int32_t s = t; s < B; s += 256) { a += Dout[s]; e += Err[s]; }
    }
    sA[t] = a;
    sE[t] = e;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (t < w) { sA[t] += sA[t + w]; sE[t] += sE[t + w]; }
        __syncthreads();
    }
    if (t == 0) {
        G[id] = sA[0];
        if (id == kLixParams - 1 && loss) *loss = sE[0] / (float)B;
    }
}

__global__ void k_adam(float* W, const float* G, float* Mo, float* Vo, int32_t P, float lr, float b1t, float b2t) {
    int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float g = G[i];
    const float m = 0.9f * Mo[i] + 0.1f * g;
    const float v = 0.999f * Vo[i] + 0.001f * g * g;
    Mo[i] = m;
    Vo[i] = v;
    const float mh = m / (1.f - b1t), vh = v / (1.f - b2t);
    W[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

/* max |prediction - lb| over distinct keys, per segment and overall: [0] MLP (MFMA
 * path), [1] root alone */
__global__ void __launch_bounds__(256) k_maxerr(LixView m, uint32_t* eSeg, uint32_t* eSegRoot, unsigned long long* err) {
    __shared__ short W1b[16 * 64];
    loadW1(m, W1b);
    unsigned long long e0 = 0, e1 = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t rounds = (m.n + stride - 1) / stride;
    for (int64_t r = 0; r < rounds; ++r) {
        const int64_t i = r * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const bool active = i < m.n && (i == 0 || m.ct[i] != m.ct[i - 1]);
        const uint64_t x = active ? m.ct[i] : 0;
        int32_t j = 0, j0 = 0;
        const int64_t p = predict(m, x, active, W1b, true, &j);
        if (active) {
            const int64_t p0 = predict(m, x, true, W1b, false, &j0);
            const unsigned long long a = (unsigned long long)(p > i ? p - i : i - p);
            const unsigned long long b = (unsigned long long)(p0 > i ? p0 - i : i - p0);
            e0 = max(e0, a);
            e1 = max(e1, b);
            atomicMax(&eSeg[j], (uint32_t)min(a, 0xffffffffull));
            atomicMax(&eSegRoot[j0], (uint32_t)min(b, 0xffffffffull));
        }
    }
    atomicMax(&err[0], e0);
    atomicMax(&err[1], e1);
}

/* useMlp: 1 model + per-segment bound, 0 root + its per-segment bound, -1 search of the
 * whole table */
__global__ void __launch_bounds__(256) k_query(LixView m, const uint64_t* keys, int64_t nq, int32_t useMlp,
                                               int64_t* lb, int32_t* probes, int32_t* fell) {
    __shared__ short W1b[16 * 64];
    loadW1(m, W1b);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t rounds = (nq + stride - 1) / stride;
    for (int64_t r = 0; r < rounds; ++r) {
        const int64_t i = r * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const bool active = i < nq;
        const uint64_t x = active ? keys[i] : 0;
        int32_t j = 0;
        const int64_t p = predict(m, x, active, W1b, useMlp > 0, &j);
        if (active) {
            int32_t pr = 0, fb = 0;
            const int64_t E = useMlp > 0 ? (int64_t)m.eSeg[j] : (useMlp == 0 ? (int64_t)m.eSegRoot[j] : m.n);
            lb[i] = lbBounded(m, x, p, E, &pr, &fb);
            if (probes) probes[i] = pr;
            if (fell) fell[i] = fb;
        }
    }
}

/* 16-mer key of read bases s[0..16), or ~0 when any base is ambiguous */
__device__ __forceinline__ uint64_t readKey(const uint8_t* s) {
    uint64_t k = 0;
    for (int t = 0; t < 16; ++t) {
        const uint8_t b = s[t];
        if (!(b == 1 || b == 2 || b == 4 || b == 8)) return ~0ull;
        k = (k << 3) | sym3(b);
    }
    return k;
}

/*
 * Form-2 front end: the learned index decides, before the SMEM search, what the search
 * would output for two kinds of reads.  need[r] codes (bits 0-1): 0 = the read emits no
 * SMEM (skip it), 1 = search it, 2 = its SMEM list is exact[r] alone, 3 = undecided
 * (searched unless a later pass decides it); bit 2 marks searches to start first (the
 * read's middle 16-mer occurs nowhere, the sign of a long search), and order[] lists the
 * reads in that order (k_order_*).
 *
 * Exact reads (k_exact, two lanes per read).  Let h = L/2 and suppose every base is
 * A/C/G/T, the read occurs exactly once in the BD text, and so do read[0, h+1) and
 * read[h, L).  Then SMEM_LIST::findAllSMEMS (src/SMEM.cpp:186-215) runs a single
 * findFwdBackSMEMS from h (src/SMEM.cpp:64-117): the forward phase reaches L with size 1
 * and pushes [h, L) (its gate: the logP of [h, L) plus the logliks of read[0, h) <= T);
 * the backward phase keeps [x, L) first with size 1 down to x = 0 (every longer candidate
 * list entry [x, k), k > h, has size 1 by x = 1 and is dropped as a repeat of that size),
 * and the step past the read start ('\0' is ambiguous, so size 0) emits [0, L) if its logP
 * is <= T and L >= minLen; then from = -1 and to = L+1 end the driver loop.  All
 * occurrences of read[0, h+1) and of the read lie in the key block of read[0, 16), all
 * occurrences of read[h, L) in the key block of read[h, h+16): lane 0 counts the first
 * two by comparing the text at each row of its block, lane 1 the third.  The reverse row
 * is ISA of the mirrored occurrence (forward and reverse halves of a chrom's 2(len+1) BD
 * symbols).  Blocks longer than kBlock rows, any other base and any failed gate leave
 * the read to the search.
 *
 * Reads without SMEMs (k_window1 / k_window2).  An SMEM the search emits is a substring
 * Y of the read with |Y| >= minLen that occurs in the BD text and whose logP (a sum of the
 * per-base logliks, src/SMEM.cpp:26-40) is <= T (the e-value gate, src/SMEM.cpp:99,152).
 * All logliks are negative, so a Y starting at a contains W(a) = read[a, c) with c the
 * first end such that c - a >= minLen and the left-to-right sum is <= T + delta (delta
 * covers any summation order: |error| < 1e-11 for reads of a few hundred bases).  If no
 * W(a) occurs, the read emits no SMEM and its search can be skipped.  W(a) occurs iff
 * some SA row whose 16-symbol key equals W's first 16 bases (found with the learned
 * index) continues in the text with W's remaining bases.  Windows with a base other than
 * A/C/G/T, and key blocks longer than kBlock rows, keep the read (conservative).
 * k_window1 tries eight windows from the middle of each undecided read (eight lanes per
 * read); reads it leaves open go to a list that k_window2 finishes with a wavefront per
 * read.  Requires minLen >= 16.
 */
constexpr int kBlock = 16;
constexpr int kFilterLanes = 8;   /* lanes (windows in flight) per read in k_window1 */

__global__ void __launch_bounds__(256, 4) k_exact(LixView m, DevIndex ix, const uint8_t* seqs, const int64_t* offs,
                                               int64_t nReads, int32_t minLen, double T, int32_t gate1,
                                               uint8_t* need, SmemRec* exact, unsigned long long* stats) {
    __shared__ short W1b[16 * 64];
    loadW1(m, W1b);
    const int lane = threadIdx.x & 63;
    const int sub = lane & 1;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nWaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const double lA = ix.ll[1], lC = ix.ll[2], lG = ix.ll[4], lT = ix.ll[8];
    unsigned long long nq = 0, nExact = 0;
    for (int64_t r0 = wave * 32; r0 < nReads; r0 += nWaves * 32) {
        const int64_t rid = r0 + (lane >> 1);
        const bool mine = rid < nReads;
        const int64_t off = mine ? offs[rid] : 0;
        const int64_t L = mine ? offs[rid + 1] - off : 0;
        const int64_t h = L / 2;
        const bool xr = mine && L >= 32 && L >= minLen;
        const uint64_t xkey = xr ? readKey(seqs + off + (sub == 0 ? 0 : h)) : ~0ull;
        const bool xact = xr && xkey != ~0ull;
        int32_t xj = 0;
        const int64_t xp = predict(m, xkey, xact, W1b, true, &xj);
        int32_t nPart = -1, nFull = 0;
        int64_t fRow = -1, fPos = -1;
        bool emptyBlock = false;
        if (xact) {
            ++nq;
            int32_t pr = 0, fb = 0;
            int64_t r = lbBounded(m, xkey, xp, (int64_t)m.eSeg[xj], &pr, &fb);
            emptyBlock = r >= m.n || m.ct[r] != xkey;
            const int64_t s0 = sub == 0 ? 0 : h, len = L - s0, part = sub == 0 ? h + 1 : L - h;
            nPart = 0;
            for (int k = 0; r < m.n && m.ct[r] == xkey; ++r, ++k) {
                if (k == kBlock) { nPart = -1; break; }
                const int64_t pos = saAt(ix, r);
                const int64_t run = fwdMatch(ix.bd, pos, seqs, off + s0, len);
                if (run >= part) nPart++;
                if (run == len) { nFull++; fRow = r; fPos = pos; }
                if (nPart > 1) break;   /* not unique: the read goes to the search either way */
            }
        }
        const int32_t nSuffix = __shfl(nPart, lane + 1);
        /* the read's middle 16-mer occurs nowhere: a read with errors there or from outside
         * the database, likely among the longest searches */
        const bool midEmpty = __shfl((int)emptyBlock, lane + 1) != 0;
        bool xok = xact && sub == 0 && nPart == 1 && nFull == 1 && nSuffix == 1;
        if (xok) {
            /* logP in the search's order: the base at h, the forward steps to L-1, then the
             * backward steps down to 0; the forward push gate adds read[0, h) in increasing
             * order to the forward logP (src/SMEM.cpp:26-40, src/SMEM.h:140-177) */
            auto llb = [&](int64_t i) {
                const uint8_t v = seqs[off + i];
                return v == 1 ? lA : (v == 2 ? lC : (v == 4 ? lG : lT));
            };
            double lf = 0;
            lf += llb(h);
            for (int64_t i = h + 1; i < L; ++i) lf += llb(i);
            if (gate1) {
                double ext = lf;
                for (int64_t i = 0; i < h; ++i) ext += llb(i);
                xok = ext <= T;
            }
            double lb = lf;
            for (int64_t i = h - 1; i >= 0; --i) lb += llb(i);
            xok = xok && lb <= T;
            if (xok) {
                int32_t lo = 0, hi = ix.nChrom;
                while (hi - lo > 1) {
                    const int32_t md = (lo + hi) >> 1;
                    if (ix.chromBD[md] <= fPos) lo = md; else hi = md;
                }
                const int64_t cs = ix.chromBD[lo], cl = ix.chromBD[lo + 1] - cs;
                const int64_t mir = 2 * cs + cl - 1 - fPos - L;
                SmemRec x;
                x.from = 0; x.to = (int32_t)L;
                x.fwd = fRow; x.rev = isaAt(ix, mir); x.size = 1; x.logP = lb;
                exact[rid] = x;
            }
        }
        if (mine && sub == 0) {
            need[rid] = xok ? 2 : (midEmpty ? 7 : 3);
            nExact += xok ? 1 : 0;
        }
    }
    if (stats) {
        atomicAdd(&stats[4], nq);
        if (nExact) atomicAdd(&stats[8], nExact);
    }
}

/* window W(a) of a read: its end c (-1 when no such window) and the key of its first 16
 * bases; `keep` when a base other than A/C/G/T lies in it (conservative) */
__device__ __forceinline__ void windowOf(const uint8_t* seqs, int64_t off, int64_t L, int64_t a, int32_t minLen,
                                         double Tm, double lA, double lC, double lG, double lT, int64_t& c,
                                         uint64_t& key, bool& keep) {
    double x = 0;
    int64_t wb = -1;
    uint64_t w = 0;
    c = -1;
    key = 0;
    for (int64_t b = a; b < L; ++b) {
        const int64_t gp = off + b;
        if ((gp & ~(int64_t)7) != wb) { wb = gp & ~(int64_t)7; w = *(const uint64_t*)(seqs + wb); }
        const uint32_t v = (uint32_t)(w >> (8 * (gp & 7))) & 0xffu;
        if (!(v == 1 || v == 2 || v == 4 || v == 8)) { keep = true; return; }
        if (b < a + 16) key = (key << 3) | sym3((uint8_t)v);
        x += v == 1 ? lA : (v == 2 ? lC : (v == 4 ? lG : lT));
        if (b + 1 - a >= minLen && x <= Tm) { c = b + 1; return; }
    }
}

/* whether window [a, c) with key `key` occurs (or the check gives up: block too long) */
__device__ __forceinline__ bool windowOccurs(const LixView& m, const DevIndex& ix, const uint8_t* seqs, int64_t off,
                                             int64_t a, int64_t c, uint64_t key, int64_t p, int32_t seg) {
    int32_t pr = 0, fb = 0;
    int64_t r = lbBounded(m, key, p, (int64_t)m.eSeg[seg], &pr, &fb);
    const int64_t wl = c - a;
    for (int k = 0; r < m.n && m.ct[r] == key; ++r, ++k) {
        if (k == kBlock) return true;
        const int64_t pos = saAt(ix, r);
        if (pos + wl > ix.n) continue;
        if (fwdMatch(ix.bd, pos + 16, seqs, off + a + 16, wl - 16) == wl - 16) return true;
    }
    return false;
}

/* window a of round `base` + sub: the windows of a read from the middle outwards */
__device__ __forceinline__ int64_t windowStart(int64_t t, int64_t nw) {
    const int64_t mid = nw / 2;
    return (t & 1) ? mid - (t + 1) / 2 : mid + t / 2;
}

__global__ void __launch_bounds__(256, 3) k_window1(LixView m, DevIndex ix, const uint8_t* seqs, const int64_t* offs,
                                                 int64_t nReads, int32_t minLen, double T, uint8_t* need,
                                                 int64_t* open, unsigned long long* nOpen,
                                                 unsigned long long* stats) {
    __shared__ short W1b[16 * 64];
    loadW1(m, W1b);
    const int lane = threadIdx.x & 63;
    const int g = lane / kFilterLanes, sub = lane % kFilterLanes;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nWaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    constexpr int kGroups = 64 / kFilterLanes;
    const double lA = ix.ll[1], lC = ix.ll[2], lG = ix.ll[4], lT = ix.ll[8];
    const double Tm = T + 1e-9;
    unsigned long long nq = 0, nSkip = 0;
    for (int64_t r0 = wave * kGroups; r0 < nReads; r0 += nWaves * kGroups) {
        const int64_t rid = r0 + g;
        const uint8_t nd0 = rid < nReads ? need[rid] : 0;
        const bool mine = (nd0 & 3) == 3;
        const int64_t off = mine ? offs[rid] : 0;
        const int64_t L = mine ? offs[rid + 1] - off : 0;
        const int64_t nw = L >= minLen ? L - minLen + 1 : 0;
        bool keep = false, active = false;
        int64_t a = 0, c = -1;
        uint64_t key = 0;
        if (mine && sub < nw) {
            a = windowStart(sub, nw);
            windowOf(seqs, off, L, a, minLen, Tm, lA, lC, lG, lT, c, key, keep);
            active = !keep && c >= 0;
        }
        int32_t j = 0;
        const int64_t p = predict(m, key, active, W1b, true, &j);
        if (active) {
            ++nq;
            keep = windowOccurs(m, ix, seqs, off, a, c, key, p, j);
        }
        const unsigned long long bal = __ballot(keep);
        const bool found = ((bal >> (g * kFilterLanes)) & ((1ull << kFilterLanes) - 1)) != 0;
        if (mine && sub == 0) {
            if (found) need[rid] = (uint8_t)((nd0 & 4) | 1);
            else if (nw <= kFilterLanes) { need[rid] = 0; ++nSkip; }
            else open[atomicAdd(nOpen, 1ull)] = rid;
        }
    }
    if (stats) {
        atomicAdd(&stats[4], nq);
        if (nSkip) atomicAdd(&stats[9], nSkip);
    }
}

__global__ void __launch_bounds__(256, 3) k_window2(LixView m, DevIndex ix, const uint8_t* seqs, const int64_t* offs,
                                                 int32_t minLen, double T, uint8_t* need, const int64_t* open,
                                                 const unsigned long long* nOpen, unsigned long long* stats) {
    __shared__ short W1b[16 * 64];
    loadW1(m, W1b);
    const int lane = threadIdx.x & 63;
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nWaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const double lA = ix.ll[1], lC = ix.ll[2], lG = ix.ll[4], lT = ix.ll[8];
    const double Tm = T + 1e-9;
    const int64_t nList = (int64_t)*nOpen;
    unsigned long long nq = 0, nSkip = 0;
    for (int64_t w = wave; w < nList; w += nWaves) {
        const int64_t rid = open[w];
        const int64_t off = offs[rid];
        const int64_t L = offs[rid + 1] - off;
        const int64_t nw = L - minLen + 1;   /* > kFilterLanes: k_window1 listed it */
        bool found = false;
        /* rounds of 64 windows after the eight k_window1 tried */
        for (int64_t base = kFilterLanes; !found && base < nw; base += 64) {
            const int64_t t = base + lane;
            bool keep = false, active = false;
            int64_t a = 0, c = -1;
            uint64_t key = 0;
            if (t < nw) {
                a = windowStart(t, nw);
                windowOf(seqs, off, L, a, minLen, Tm, lA, lC, lG, lT, c, key, keep);
                active = !keep && c >= 0;
            }
            int32_t j = 0;
            const int64_t p = predict(m, key, active, W1b, true, &j);
            if (active) {
                ++nq;
                keep = windowOccurs(m, ix, seqs, off, a, c, key, p, j);
            }
            found = __ballot(keep) != 0;
        }
        if (lane == 0) {
            need[rid] = found ? (uint8_t)((need[rid] & 4) | 1) : 0;
            nSkip += found ? 0 : 1;
        }
    }
    if (stats) {
        atomicAdd(&stats[4], nq);
        if (nSkip) atomicAdd(&stats[9], nSkip);
    }
}

/* the search order: reads marked to start first, then the others (order inside each
 * part is whatever the atomics give; results do not depend on it) */
__global__ void __launch_bounds__(256) k_order_count(const uint8_t* need, int64_t n, unsigned long long* cnt) {
    __shared__ unsigned int s;
    if (threadIdx.x == 0) s = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (need[i] & 4)) atomicAdd(&s, 1u);
    __syncthreads();
    if (threadIdx.x == 0 && s) atomicAdd(&cnt[0], (unsigned long long)s);
}

__global__ void __launch_bounds__(256) k_order_scatter(const uint8_t* need, int64_t n, unsigned long long* cnt,
                                                       int32_t* order) {
    __shared__ unsigned int s[2];
    __shared__ unsigned long long base[2];
    if (threadIdx.x < 2) s[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int part = -1;
    unsigned int r = 0;
    if (i < n) { part = (need[i] & 4) ? 0 : 1; r = atomicAdd(&s[part], 1u); }
    __syncthreads();
    if (threadIdx.x < 2 && s[threadIdx.x]) base[threadIdx.x] = atomicAdd(&cnt[1 + threadIdx.x], (unsigned long long)s[threadIdx.x]);
    __syncthreads();
    if (part >= 0) order[(part == 0 ? 0 : cnt[0]) + base[part] + r] = (int32_t)i;
}

/* raw model outputs (before rounding) for checks: mode 0 bf16 MFMA, mode 1 the same
 * arithmetic on VALU (bf16-rounded features and first-layer weights, fp32 sums), mode 2
 * the root alone */
__global__ void __launch_bounds__(256) k_predict(LixView m, const uint64_t* keys, int64_t nq, int32_t mode,
                                                 double* out) {
    __shared__ short W1b[16 * 64];
    loadW1(m, W1b);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t rounds = (nq + stride - 1) / stride;
    for (int64_t r = 0; r < rounds; ++r) {
        const int64_t i = r * stride + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const bool active = i < nq;
        const uint64_t x = active ? keys[i] : 0;
        float f[16];
        for (int k = 0; k < 16; ++k) f[k] = 0.f;
        double base = 0, width = 1;
        if (active) {
            Plr p = plrOf(m, x, segOf(m, x));
            base = p.base; width = p.width;
            features(x, p.t, f);
        }
        float y;
        if (mode == 0) {
            y = mlpMfma(m, f, W1b);
        } else if (mode == 2) {
            y = 0.f;
        } else {
            const float* b1 = m.W + 16 * 64;
            const float* W2 = b1 + 64;
            y = 0.f;
            for (int u = 0; u < 64; ++u) {
                float a = 0.f;
                for (int k = 0; k < 16; ++k) a += fromBf16(toBf16(f[k])) * fromBf16(W1b[k * 64 + u]);
                a += b1[u];
                a = a > 0 ? a : 0;
                y += a * W2[u];
            }
            y += W2[64];
        }
        if (active) out[i] = base + (double)y * width;
    }
}

inline unsigned gridFor(int64_t n, int t = 256) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + t - 1) / t, 1 << 16));
}

}  // namespace

/* ------------------------------------------------------------------ host side */

#define LCK(x, msg)                                                        \
    do {                                                                   \
        hipError_t e_ = (x);                                               \
        if (e_ != hipSuccess) return fail(c, MSGS_E_HIP, std::string(msg) + ": " + hipGetErrorString(e_)); \
    } while (0)

LixView LixState::view() const {
    LixView v;
    v.ct = dCt; v.n = n; v.kKey = dKey; v.kRank = dRank; v.S = S; v.W = dW; v.kBucket = dBucket; v.eSeg = dESeg;
    v.eSegRoot = dESegRoot;
    return v;
}

void LixState::release() {
    void* p[] = {dCt, dKey, dRank, dW, dBucket, dESeg, dESegRoot};
    for (void* x : p) if (x) (void)hipFree(x);
    dCt = nullptr; dKey = nullptr; dRank = nullptr; dW = nullptr; dBucket = nullptr; dESeg = nullptr; dESegRoot = nullptr;
    trained = false;
}

static void initWeights(std::vector<float>& W, uint64_t seed) {
    W.assign(kLixParams, 0.f);
    uint64_t s = seed ? seed : 1;
    auto u = [&]() {   /* xorshift64*, uniform in (0,1) */
        s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
        return ((s * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0) + 1e-12;
    };
    auto gauss = [&]() { return std::sqrt(-2 * std::log(u())) * std::cos(6.283185307179586 * u()); };
    for (int i = 0; i < 16 * 64; ++i) W[i] = (float)(gauss() * std::sqrt(2.0 / 16));
    for (int i = 0; i < 64; ++i) W[16 * 64 + 64 + i] = (float)(gauss() * 0.01);
}

extern "C" int msgs_lix_train(msgs_ctx* c, int32_t segments, int64_t samples, int32_t steps, int32_t batch,
                              double lr, uint64_t seed, double* info) {
    if (!c || !c->db.n) return fail(c, MSGS_E_STATE, "no database");
    if (segments < 2 || segments > (1 << 22) || samples < 1 || steps < 0 || batch < 1 || batch > 65536)
        return fail(c, MSGS_E_ARG, "bad learned-index settings");
    LixState& L = c->lix;
    L.release();
    const int64_t n = c->db.n;
    hipStream_t st = c->x().stream;
    L.n = n;
    L.S = segments;
    LCK(hipMalloc(&L.dCt, n * sizeof(uint64_t)), "CT");
    LCK(hipMalloc(&L.dKey, (segments + 1) * sizeof(uint64_t)), "knots");
    LCK(hipMalloc(&L.dRank, (segments + 1) * sizeof(int64_t)), "knots");
    LCK(hipMalloc(&L.dW, kLixParams * sizeof(float)), "weights");
    LCK(hipMalloc(&L.dBucket, (((size_t)1 << kBucketBits) + 1) * sizeof(uint32_t)), "buckets");
    LCK(hipMalloc(&L.dESeg, segments * sizeof(uint32_t)), "bounds");
    LCK(hipMalloc(&L.dESegRoot, segments * sizeof(uint32_t)), "bounds");
    hipEvent_t e0 = c->x().ev0, e1 = c->x().ev1;
    LCK(hipEventRecord(e0, st), "event");
    k_ct<<<gridFor(n), 256, 0, st>>>(c->db.view(), L.dCt);
    k_knots<<<gridFor(segments + 1), 256, 0, st>>>(L.dCt, n, segments, L.dKey, L.dRank);
    k_buckets<<<gridFor(((int64_t)1 << kBucketBits) + 1), 256, 0, st>>>(L.dKey, segments, L.dBucket);
    std::vector<float> W;
    initWeights(W, seed);
    LCK(hipMemcpyAsync(L.dW, W.data(), kLixParams * 4, hipMemcpyHostToDevice, st), "W");
    hipError_t e = hipSuccess;
    float* X = c->x().bufs.get<float>("lx.X", samples * 16, e);
    float* T = c->x().bufs.get<float>("lx.T", samples, e);
    float* H = c->x().bufs.get<float>("lx.H", (size_t)batch * 64, e);
    float* Dout = c->x().bufs.get<float>("lx.D", batch, e);
    float* Err = c->x().bufs.get<float>("lx.E", batch, e);
    float* G = c->x().bufs.get<float>("lx.G", kLixParams, e);
    float* Mo = c->x().bufs.get<float>("lx.M", kLixParams, e);
    float* Vo = c->x().bufs.get<float>("lx.V", kLixParams, e);
    float* dLoss = c->x().bufs.get<float>("lx.loss", 1, e);
    unsigned long long* dErr = c->x().bufs.get<unsigned long long>("lx.err", 2, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "training buffers");
    LCK(hipMemsetAsync(Mo, 0, kLixParams * 4, st), "adam");
    LCK(hipMemsetAsync(Vo, 0, kLixParams * 4, st), "adam");
    k_sample<<<gridFor(samples), 256, 0, st>>>(L.view(), samples, seed, X, T);
    LCK(hipGetLastError(), "sample");
    float b1t = 1.f, b2t = 1.f;
    float firstLoss = 0, lastLoss = 0;
    for (int32_t s = 0; s < steps; ++s) {
        const int64_t off = ((int64_t)s * batch) % std::max<int64_t>(1, samples - batch + 1);
        const int32_t B = (int32_t)std::min<int64_t>(batch, samples - off);
        k_fwd<<<gridFor(B), 256, 0, st>>>(L.dW, X, T, off, B, H, Dout, Err);
        k_grad<<<kLixParams, 256, 0, st>>>(L.dW, X, off, B, H, Dout, Err, G, dLoss);
        b1t *= 0.9f; b2t *= 0.999f;
        k_adam<<<gridFor(kLixParams), 256, 0, st>>>(L.dW, G, Mo, Vo, kLixParams, (float)lr, b1t, b2t);
        if (s == 0 || s == steps - 1) {
            float lv = 0;
            LCK(hipMemcpyAsync(&lv, dLoss, 4, hipMemcpyDeviceToHost, st), "loss");
            LCK(hipStreamSynchronize(st), "train");
            if (s == 0) firstLoss = lv;
            lastLoss = lv;
        }
    }
    LCK(hipEventRecord(e1, st), "event");
    LCK(hipMemsetAsync(dErr, 0, 16, st), "err");
    LCK(hipMemsetAsync(L.dESeg, 0, segments * sizeof(uint32_t), st), "bounds");
    LCK(hipMemsetAsync(L.dESegRoot, 0, segments * sizeof(uint32_t), st), "bounds");
    k_maxerr<<<gridFor(std::min<int64_t>(n, (int64_t)c->nCU * 2048)), 256, 0, st>>>(L.view(), L.dESeg, L.dESegRoot, dErr);
    unsigned long long he[2] = {0, 0};
    LCK(hipMemcpyAsync(he, dErr, 16, hipMemcpyDeviceToHost, st), "err");
    LCK(hipStreamSynchronize(st), "max error");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    L.E = (int64_t)he[0];
    L.Eplr = (int64_t)he[1];
    L.trained = true;
    if (info) {
        info[0] = (double)L.E; info[1] = (double)L.Eplr; info[2] = firstLoss; info[3] = lastLoss; info[4] = ms;
    }
    return MSGS_OK;
}

extern "C" int msgs_lix_weights(msgs_ctx* c, float* get, const float* set) {
    if (!c || !c->lix.dW) return fail(c, MSGS_E_STATE, "no learned index");
    if (get) LCK(hipMemcpy(get, c->lix.dW, kLixParams * 4, hipMemcpyDeviceToHost), "weights");
    if (set) LCK(hipMemcpy(c->lix.dW, set, kLixParams * 4, hipMemcpyHostToDevice), "weights");
    return MSGS_OK;
}

/* gradient of the mean squared error over B rows of features X (B x 16) and targets T at
 * weights W, by the training kernels (for checks against an independent computation) */
extern "C" int msgs_lix_grad(msgs_ctx* c, const float* W, const float* X, const float* T, int32_t B, float* grad,
                             float* loss) {
    if (!c || !W || !X || !T || !grad || B < 1 || B > 65536) return fail(c, MSGS_E_ARG, "bad gradient arguments");
    hipError_t e = hipSuccess;
    hipStream_t st = c->x().stream;
    float* dW = c->x().bufs.get<float>("gc.W", kLixParams, e);
    float* dX = c->x().bufs.get<float>("gc.X", (size_t)B * 16, e);
    float* dT = c->x().bufs.get<float>("gc.T", B, e);
    float* H = c->x().bufs.get<float>("gc.H", (size_t)B * 64, e);
    float* Dout = c->x().bufs.get<float>("gc.D", B, e);
    float* Err = c->x().bufs.get<float>("gc.E", B, e);
    float* G = c->x().bufs.get<float>("gc.G", kLixParams, e);
    float* dL = c->x().bufs.get<float>("gc.L", 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "gradient buffers");
    LCK(hipMemcpyAsync(dW, W, kLixParams * 4, hipMemcpyHostToDevice, st), "W");
    LCK(hipMemcpyAsync(dX, X, (size_t)B * 16 * 4, hipMemcpyHostToDevice, st), "X");
    LCK(hipMemcpyAsync(dT, T, (size_t)B * 4, hipMemcpyHostToDevice, st), "T");
    k_fwd<<<gridFor(B), 256, 0, st>>>(dW, dX, dT, 0, B, H, Dout, Err);
    k_grad<<<kLixParams, 256, 0, st>>>(dW, dX, 0, B, H, Dout, Err, G, dL);
    LCK(hipMemcpyAsync(grad, G, kLixParams * 4, hipMemcpyDeviceToHost, st), "G");
    if (loss) LCK(hipMemcpyAsync(loss, dL, 4, hipMemcpyDeviceToHost, st), "loss");
    LCK(hipStreamSynchronize(st), "gradient");
    return MSGS_OK;
}

/* exact lower bounds of 16-mer keys; useMlp 0 uses the root model alone with its own
 * bound, and useMlp -1 a plain binary search (the method without a learned index) */
extern "C" int msgs_lix_query(msgs_ctx* c, const uint64_t* keys, int64_t nq, int32_t useMlp, int64_t* lb,
                              int32_t* probes, int32_t* fell) {
    if (!c || !c->lix.trained) return fail(c, MSGS_E_STATE, "learned index not trained");
    if (nq <= 0) return MSGS_OK;
    hipError_t e = hipSuccess;
    hipStream_t st = c->x().stream;
    uint64_t* dK = c->x().bufs.get<uint64_t>("lq.k", nq, e);
    int64_t* dL = c->x().bufs.get<int64_t>("lq.l", nq, e);
    int32_t* dP = c->x().bufs.get<int32_t>("lq.p", nq, e);
    int32_t* dF = c->x().bufs.get<int32_t>("lq.f", nq, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "query buffers");
    LCK(hipMemcpyAsync(dK, keys, nq * 8, hipMemcpyHostToDevice, st), "keys");
    const LixState& L = c->lix;
    k_query<<<gridFor(nq), 256, 0, st>>>(L.view(), dK, nq, useMlp > 0 ? 1 : (useMlp == 0 ? 0 : -1), dL, dP, dF);
    LCK(hipMemcpyAsync(lb, dL, nq * 8, hipMemcpyDeviceToHost, st), "lb");
    if (probes) LCK(hipMemcpyAsync(probes, dP, nq * 4, hipMemcpyDeviceToHost, st), "probes");
    if (fell) LCK(hipMemcpyAsync(fell, dF, nq * 4, hipMemcpyDeviceToHost, st), "fell");
    LCK(hipStreamSynchronize(st), "query");
    return MSGS_OK;
}

int runLixFilter(msgs_ctx* c, const DevReads& r, int32_t minLen, double T, int32_t gate1, uint8_t** needOut,
                 SmemRec** exactOut, unsigned long long* dStats) {
    hipError_t e = hipSuccess;
    uint8_t* need = c->x().bufs.get<uint8_t>("lx.need", r.n + 1, e);
    SmemRec* exact = c->x().bufs.get<SmemRec>("lx.exact", r.n + 1, e);
    int64_t* open = c->x().bufs.get<int64_t>("lx.open", r.n + 1, e);
    unsigned long long* nOpen = c->x().bufs.get<unsigned long long>("lx.nopen", 1, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "filter buffer");
    if (r.n) {
        const int64_t cap = (int64_t)std::max(1, c->nCU) * 32;
        k_exact<<<(unsigned)std::min<int64_t>((r.n + 127) / 128, cap), 256, 0, c->x().stream>>>(
            c->lix.view(), c->db.view(), r.seq, r.offs, r.n, minLen, T, gate1, need, exact, dStats);
        if (c->readFilter) {
            LCK(hipMemsetAsync(nOpen, 0, 8, c->x().stream), "filter list");
            k_window1<<<(unsigned)std::min<int64_t>((r.n + 31) / 32, cap), 256, 0, c->x().stream>>>(
                c->lix.view(), c->db.view(), r.seq, r.offs, r.n, minLen, T, need, open, nOpen, dStats);
            k_window2<<<(unsigned)cap, 256, 0, c->x().stream>>>(c->lix.view(), c->db.view(), r.seq, r.offs, minLen, T,
                                                              need, open, nOpen, dStats);
        }
    }
    LCK(hipGetLastError(), "filter");
    *needOut = need;
    *exactOut = exact;
    return MSGS_OK;
}

int runLixOrder(msgs_ctx* c, int64_t n, const uint8_t* need, int32_t** orderOut,
                const unsigned long long** splitOut) {
    hipError_t e = hipSuccess;
    int32_t* order = c->x().bufs.get<int32_t>("lx.order", n + 1, e);
    unsigned long long* cnt = c->x().bufs.get<unsigned long long>("lx.ocnt", 4, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "search order");
    if (n) {
        LCK(hipMemsetAsync(cnt, 0, 4 * 8, c->x().stream), "order counters");
        const unsigned blocks = (unsigned)((n + 255) / 256);
        k_order_count<<<blocks, 256, 0, c->x().stream>>>(need, n, cnt);
        k_order_scatter<<<blocks, 256, 0, c->x().stream>>>(need, n, cnt, order);
    }
    LCK(hipGetLastError(), "order");
    *orderOut = order;
    if (splitOut) *splitOut = cnt;
    return MSGS_OK;
}

extern "C" int msgs_lix_predict(msgs_ctx* c, const uint64_t* keys, int64_t nq, int32_t mode, double* out) {
    if (!c || !c->lix.trained) return fail(c, MSGS_E_STATE, "learned index not trained");
    if (nq <= 0) return MSGS_OK;
    hipError_t e = hipSuccess;
    uint64_t* dK = c->x().bufs.get<uint64_t>("lp.k", nq, e);
    double* dO = c->x().bufs.get<double>("lp.o", nq, e);
    if (e != hipSuccess) return fail(c, MSGS_E_NOMEM, "predict buffers");
    LCK(hipMemcpyAsync(dK, keys, nq * 8, hipMemcpyHostToDevice, c->x().stream), "keys");
    k_predict<<<gridFor(nq), 256, 0, c->x().stream>>>(c->lix.view(), dK, nq, mode, dO);
    LCK(hipMemcpyAsync(out, dO, nq * 8, hipMemcpyDeviceToHost, c->x().stream), "out");
    LCK(hipStreamSynchronize(c->x().stream), "predict");
    return MSGS_OK;
}

/* the knot table (S+1 keys and lower bounds) and the key table size, for checks */
extern "C" int msgs_lix_knots(msgs_ctx* c, uint64_t* keys, int64_t* ranks, int32_t* S) {
    if (!c || !c->lix.trained) return fail(c, MSGS_E_STATE, "learned index not trained");
    if (S) *S = c->lix.S;
    if (keys) LCK(hipMemcpy(keys, c->lix.dKey, (c->lix.S + 1) * 8, hipMemcpyDeviceToHost), "knots");
    if (ranks) LCK(hipMemcpy(ranks, c->lix.dRank, (c->lix.S + 1) * 8, hipMemcpyDeviceToHost), "knots");
    return MSGS_OK;
}
