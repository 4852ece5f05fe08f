/*
 * msgs_lix.h -- learned index state (see msgs_lix.hip).
 */
#ifndef MSGS_LIX_H
#define MSGS_LIX_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* W1[16][64], b1[64], W2[64], b2 */
constexpr int kLixParams = 16 * 64 + 64 + 64 + 1;
/* the root's bucket table indexes keys by their top 24 bits (first 8 symbols) */
constexpr int kBucketBits = 24;

struct LixView {
    const uint64_t* ct;     /* 16-symbol key of each SA row (non-decreasing) */
    int64_t n;
    const uint64_t* kKey;   /* S+1 knot keys (last = all ones) */
    const int64_t* kRank;   /* S+1 knot lower bounds (last = n) */
    int32_t S;
    const float* W;         /* MLP parameters */
    const uint32_t* kBucket;    /* 2^24 + 1: first knot with key >= bucket b << 24 */
    const uint32_t* eSeg;       /* per segment: max |prediction - lb| of its keys, MLP */
    const uint32_t* eSegRoot;   /* the same for the root alone */
};

struct LixState {
    uint64_t* dCt = nullptr;
    uint64_t* dKey = nullptr;
    int64_t* dRank = nullptr;
    float* dW = nullptr;
    uint32_t* dBucket = nullptr;
    uint32_t* dESeg = nullptr;
    uint32_t* dESegRoot = nullptr;
    int64_t n = 0;
    int32_t S = 0;
    int64_t E = 0, Eplr = 0;   /* recorded max errors: MLP, root alone */
    bool trained = false;
    LixView view() const;
    void release();
};

#endif
