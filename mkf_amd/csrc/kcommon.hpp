// Device helpers shared by the kernel translation units.
#pragma once

#include "ntt.hpp"
#include "tables.hpp"

namespace mkhip {

extern __shared__ u64 g_lds[];

template <int LOGN>
__device__ __forceinline__ void load_row_cached(u64 (&v)[16], const u64 *__restrict__ src, int tid) {
    constexpr int T = (1 << LOGN) / 16;
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = src[tid + i * T];
}

template <int LOGN>
__device__ __forceinline__ void store_row(const u64 (&v)[16], u64 *__restrict__ dst, int tid) {
    constexpr int T = (1 << LOGN) / 16;
#pragma unroll
    for (int i = 0; i < 16; ++i) dst[tid + i * T] = v[i];
}

// Row r of a (level+1+K)-row QP polynomial is ring modulus r (Q) or L + r - level - 1 (P).
__device__ __forceinline__ int ring_index(int r, int level, int L) { return r <= level ? r : L + (r - level - 1); }

// 128-bit accumulate of a * b.
__device__ __forceinline__ void mac128(u64 &hi, u64 &lo, u64 a, u64 b) {
    u64 l = a * b;
    u64 h = mulhi64(a, b);
    lo += l;
    hi += h + (lo < l ? 1 : 0);
}

// Lazy Montgomery reduction of a 128-bit accumulator as in multSum
// (basis_extension.go:623-645): hi - hi(lo * pinv * p) + p.
__device__ __forceinline__ u64 mont_reduce128(u64 hi, u64 lo, u64 p, u64 pinv) {
    return hi - mulhi64(lo * pinv, p) + p;
}

// a * b mod q for two canonical operands: MRed(MForm(a), b).
__device__ __forceinline__ u64 mulmod_m(u64 a, u64 b, const ModConst &m) {
    return mred(mred(a, m.r2, m.q, m.qinv), b, m.q, m.qinv);
}

constexpr int kMaxK = 4;

// ModUp of one coefficient from the P basis (the K rows at src + (nl + p) * N) to
// Q row j (modupExact with paramsPtoQ, basis_extension.go:337-357, 537-646):
// y_p = MRed(x_p, (P/p_p)^-1 R), v = trunc(sum float64(y_p)/float64(p_p)) in
// ascending p, result = sum_p y_p (P/p_p) + v (-P) mod q_j, canonical.
struct PUp {
    u64 y[kMaxK];
    u64 v;
};

__device__ __forceinline__ PUp modup_p_prepare(const DevTables &T, const u64 *__restrict__ src, int nl, size_t c) {
    PUp u;
    double vi = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxK; ++k) {
        u.y[k] = 0;
        if (k < T.K) {
            const ModConst mp = T.mod[T.L + k];
            u.y[k] = mred(src[(size_t)(nl + k) * T.N + c], T.md_y[k], mp.q, mp.qinv);
            vi += (double)u.y[k] / (double)mp.q;
        }
    }
    u.v = (u64)vi;
    return u;
}

__device__ __forceinline__ u64 modup_p_to_row(const DevTables &T, const PUp &u, int j, const ModConst &mq) {
    u64 hi = 0, lo = 0;
#pragma unroll
    for (int k = 0; k < kMaxK; ++k)
        if (k < T.K) mac128(hi, lo, u.y[k], T.md_qstar[j * T.K + k]);
    const u64 z = mont_reduce128(hi, lo, mq.q, mq.qinv) + T.md_vt[j * (T.K + 1) + u.v];
    return reduce64(z, mq.q, mq.mu);
}

}  // namespace mkhip
