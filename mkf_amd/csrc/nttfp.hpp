// Forward and inverse NTT of one RNS limb in FP64 arithmetic, for moduli q < 2^50.
//
// The transform is the one of ntt.hpp (same twiddle order, same passes and LDS
// exchanges); only the butterfly arithmetic differs.  Residues are held as
// doubles: integers of magnitude below 2^53 are exact, and for q < 2^50 every
// value stays within +-5.7q < 2^53.  The product y w (|y| < 6q, 0 <= w < q) is
// split exactly into h + l with one FMA (h = fl(y w), l = y w - h), the quotient
// estimate qt = rint(h / q) comes from one multiplication by fl(1/q), and
// r = (h - qt q) + l is computed exactly by a second FMA and an addition
// (|h - qt q| < 2^52 is an integer, so the FMA result is exact), giving
// r = y w mod q with |r| < 2.6q (the estimate h fl(1/q) is within 3 2^-53 |y w / q|
// of y w / q, at most 2.1 here, plus 0.5 from rint).  Butterflies are x' = u + r,
// y' = u - r, and u is reduced to |u| <= q/2 before every other stage, so values
// alternate below 3.1q and 5.7q: under 2^53 for q < 2^50.
// Results are reduced to canonical residues, so the output words equal the
// integer transform's bit for bit.  On gfx950 these are full-rate FP64 VALU
// operations, where the integer form needs several quarter-rate 32-bit
// multiplies per butterfly.
#pragma once

#include "ntt.hpp"

namespace mkhip {

// Moduli whose limbs may take the FP64 path.
__device__ __forceinline__ bool fp_ok(u64 q) { return q < (1ull << 50); }

struct FpMod {
    double q, qinv;  // q and fl(1 / q)
};

__device__ __forceinline__ double fp_mulmod(double a, double w, const FpMod &m) {
    const double h = a * w;
    const double l = __builtin_fma(a, w, -h);
    const double qt = __builtin_rint(h * m.qinv);
    return __builtin_fma(-qt, m.q, h) + l;
}

// |result| <= q/2 (plus an ulp-level excess), congruent to x, for |x| < 2^52
__device__ __forceinline__ double fp_reduce(double x, const FpMod &m) {
    return __builtin_fma(-__builtin_rint(x * m.qinv), m.q, x);
}

// exact for x < 2^53
__device__ __forceinline__ double fp_from_u64(u64 x) {
    return __builtin_fma((double)(u32)(x >> 32), 4294967296.0, (double)(u32)x);
}

// canonical residue of x (|x| < 2^52), as a double in [0, q)
__device__ __forceinline__ double fp_canon(double x, const FpMod &m) {
    double r = fp_reduce(x, m);
    r = r < 0.0 ? r + m.q : r;
    return r >= m.q ? r - m.q : r;
}

// an integer-valued double in [0, 2^53) as a word
__device__ __forceinline__ u64 fp_word(double r) {
    const double hi = __builtin_floor(r * (1.0 / 4294967296.0));
    const double lo = __builtin_fma(-hi, 4294967296.0, r);
    return ((u64)(u32)hi << 32) | (u64)(u32)lo;
}

template <bool UNIFORM>
__device__ __forceinline__ double twiddle_fp(const double *__restrict__ tw, int idx) {
    if (UNIFORM) {  // scalar load, as twiddle<true>
        typedef __attribute__((address_space(4))) const double cd;
        return ((cd *)tw)[__builtin_amdgcn_readfirstlane(idx)];
    }
    typedef __attribute__((address_space(1))) const double gd;
    return ((gd *)tw)[idx];
}

template <int LOGN, int S0, int NS>
struct FwdPassFp : FwdPass<LOGN, S0, NS> {
    typedef FwdPass<LOGN, S0, NS> B;
    // stage S0 + u reduces u when (S0 + u + par) is odd.  GATHER: per-lane twiddles
    // loaded together at the start (more registers, one memory wait per pass).
    template <bool GATHER = true>
    __device__ static __forceinline__ void run(double (&v)[16], int tid, const double *__restrict__ tw,
                                               const FpMod &m, int hh, int par) {
        constexpr int R = B::R, C = B::C, TL = B::TL, T = B::T;
        constexpr bool UNI = TL % 64 == 0;
        constexpr bool PRE = !UNI && GATHER;
        // per-lane twiddles: all of the pass's are gathered up front, so the pass
        // waits for memory once instead of once per stage
        double twl[PRE ? C * (R - 1) : 1];
        if constexpr (PRE) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int b = (tid + c * T) / TL;
#pragma unroll
                for (int u = 0; u < NS; ++u)
#pragma unroll
                    for (int grp = 0; grp < (1 << u); ++grp)
                        twl[c * (R - 1) + (1 << u) - 1 + grp] =
                            twiddle_fp<false>(tw, (1 << (S0 + u)) * hh + b * (1 << u) + grp);
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int col = tid + c * T;
            const int b = UNI ? __builtin_amdgcn_readfirstlane(col / TL) : col / TL;
#pragma unroll
            for (int u = 0; u < NS; ++u) {
                __builtin_amdgcn_sched_barrier(0);
                const int half = R >> (u + 1);
                const bool red = ((S0 + u + par) & 1) != 0;
#pragma unroll
                for (int grp = 0; grp < (1 << u); ++grp) {
                    const double w = PRE ? twl[c * (R - 1) + (1 << u) - 1 + grp]
                                         : twiddle_fp<UNI>(tw, (1 << (S0 + u)) * hh + b * (1 << u) + grp);
#pragma unroll
                    for (int kk = 0; kk < half; ++kk) {
                        const int k0 = grp * 2 * half + kk;
                        double &x = v[c * R + k0], &y = v[c * R + k0 + half];
                        const double r = fp_mulmod(y, w, m);
                        const double uu = red ? fp_reduce(x, m) : x;
                        x = uu + r;
                        y = uu - r;
                    }
                }
            }
        }
    }
};

template <class From, class To>
__device__ __forceinline__ void lds_exchange_fp(double (&v)[16], u64 *s, int tid) {
    static_assert(From::pad_ok && To::pad_ok, "padded LDS layout");
    double *d = reinterpret_cast<double *>(s);
    double *w = d + lds_pad(From::base(tid));
#pragma unroll
    for (int i = 0; i < 16; ++i) w[lds_pad(From::off(i))] = v[i];
    __syncthreads();
    const double *r = d + lds_pad(To::base(tid));
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = r[lds_pad(To::off(i))];
    __syncthreads();
}

// Forward NTT in FP64; v in Linear layout (|v| < 2q), Quad layout out, values in
// (-6q, 6q) congruent to the transform (canonicalize with fp_canon).  hh as in
// ntt_fwd; par = 1 shifts the reduction schedule by one stage (used after a
// first stage computed outside).
template <int LOGN>
struct FpPlan {
    typedef NttPlan<LOGN> P;
    typedef FwdPassFp<LOGN, 0, 4> F0;
    typedef FwdPassFp<LOGN, 4, 4> F1;
    typedef FwdPassFp<LOGN, 8, 4> F2;
    typedef FwdPassFp<LOGN, 12, LOGN - 12> F3;
    static_assert(F0::TL == P::F0::TL && F1::TL == P::F1::TL && F2::TL == P::F2::TL && F3::TL == P::F3::TL &&
                      F3::R == P::F3::R,
                  "the passes of the integer transform");
};

// The passes that use LDS (all but the last); LDS is free again when this returns.
template <int LOGN, bool GATHER = true>
__device__ __forceinline__ void ntt_fwd_fp_head(double (&v)[16], u64 *s, int tid, const double *__restrict__ tw,
                                                const FpMod &m, int hh, int par) {
    typedef FpPlan<LOGN> Q;
    typedef NttPlan<LOGN> P;
    Q::F0::template run<GATHER>(v, tid, tw, m, hh, par);
    lds_exchange_fp<typename P::F0, typename P::F1>(v, s, tid);
    Q::F1::template run<GATHER>(v, tid, tw, m, hh, par);
    lds_exchange_fp<typename P::F1, typename P::F2>(v, s, tid);
    Q::F2::template run<GATHER>(v, tid, tw, m, hh, par);
    lds_exchange_fp<typename P::F2, typename P::F3>(v, s, tid);
}

template <int LOGN, bool GATHER = true>
__device__ __forceinline__ void ntt_fwd_fp_tail(double (&v)[16], int tid, const double *__restrict__ tw,
                                                const FpMod &m, int hh, int par) {
    FpPlan<LOGN>::F3::template run<GATHER>(v, tid, tw, m, hh, par);
}

template <int LOGN, bool GATHER = true>
__device__ __forceinline__ void ntt_fwd_fp(double (&v)[16], u64 *s, int tid, const double *__restrict__ tw,
                                           const FpMod &m, int hh = 1, int par = 0) {
    ntt_fwd_fp_head<LOGN, GATHER>(v, s, tid, tw, m, hh, par);
    ntt_fwd_fp_tail<LOGN, GATHER>(v, tid, tw, m, hh, par);
}

// Inverse passes in FP64 (Gentleman-Sande): x' = reduce(x + y) with |x'| <= q/2,
// y' = (x - y) w mod q with |y'| < 2.6q, so every stage starts from values below
// 2.6q.
template <int LOGN, int U0, int NS>
struct InvPassFp : InvPass<LOGN, U0, NS> {
    typedef InvPass<LOGN, U0, NS> B;
    __device__ static __forceinline__ void run(double (&v)[16], int tid, const double *__restrict__ itw,
                                               const FpMod &m, int hh) {
        constexpr int R = B::R, C = B::C, TF = B::TF, T = B::T, N = B::N;
        constexpr bool UNI = TF % 64 == 0;
        // per-lane twiddles gathered up front (one memory wait per pass); slot of
        // (stage u, group grp): R - (R >> u) + grp
        double twl[UNI ? 1 : C * (R - 1)];
        if constexpr (!UNI) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int b = (tid + c * T) / TF;
#pragma unroll
                for (int u = 0; u < NS; ++u) {
                    const int ngrp = R >> (u + 1);
#pragma unroll
                    for (int grp = 0; grp < ngrp; ++grp)
                        twl[c * (R - 1) + R - (R >> u) + grp] =
                            twiddle_fp<false>(itw, (N >> (U0 + u + 1)) * hh + b * ngrp + grp);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int col = tid + c * T;
            const int b = UNI ? __builtin_amdgcn_readfirstlane(col / TF) : col / TF;
#pragma unroll
            for (int u = 0; u < NS; ++u) {
                __builtin_amdgcn_sched_barrier(0);
                const int half = 1 << u;
                const int mm = N >> (U0 + u + 1);
                const int ngrp = R / (2 * half);
#pragma unroll
                for (int grp = 0; grp < ngrp; ++grp) {
                    const double w = UNI ? twiddle_fp<true>(itw, mm * hh + b * ngrp + grp)
                                         : twl[c * (R - 1) + R - (R >> u) + grp];
#pragma unroll
                    for (int kk = 0; kk < half; ++kk) {
                        const int k0 = grp * 2 * half + kk;
                        double &x = v[c * R + k0], &y = v[c * R + k0 + half];
                        const double sum = x + y, dif = x - y;
                        x = fp_reduce(sum, m);
                        y = fp_mulmod(dif, w, m);
                    }
                }
            }
        }
    }
};

// Inverse NTT in FP64 of words below 4q (Quad layout in, Linear layout out).  With
// scale the N^-1 (or t N^-1) factor `ninv` is applied and the words returned are
// canonical; without it (a half of a split limb) they are canonical residues of
// the values before the last stage.
// The same on residues already held as doubles (|d| < 2.6q); words out.
template <int LOGN>
__device__ __forceinline__ void ntt_inv_fp_d(double (&d)[16], u64 (&v)[16], u64 *s, int tid,
                                             const double *__restrict__ itw, const FpMod &m, double ninv, int hh,
                                             bool scale) {
    typedef NttPlan<LOGN> P;
    typedef InvPassFp<LOGN, 0, LOGN - 12> I0;
    typedef InvPassFp<LOGN, LOGN - 12, 4> I1;
    typedef InvPassFp<LOGN, LOGN - 8, 4> I2;
    typedef InvPassFp<LOGN, LOGN - 4, 4> I3;
    I0::run(d, tid, itw, m, hh);
    lds_exchange_fp<typename P::I0, typename P::I1>(d, s, tid);
    I1::run(d, tid, itw, m, hh);
    lds_exchange_fp<typename P::I1, typename P::I2>(d, s, tid);
    I2::run(d, tid, itw, m, hh);
    lds_exchange_fp<typename P::I2, typename P::I3>(d, s, tid);
    I3::run(d, tid, itw, m, hh);
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = fp_word(fp_canon(scale ? fp_mulmod(d[i], ninv, m) : d[i], m));
}

template <int LOGN>
__device__ __forceinline__ void ntt_inv_fp(u64 (&v)[16], u64 *s, int tid, const double *__restrict__ itw,
                                           const FpMod &m, double ninv, int hh, bool scale) {
    typedef NttPlan<LOGN> P;
    typedef InvPassFp<LOGN, 0, LOGN - 12> I0;
    typedef InvPassFp<LOGN, LOGN - 12, 4> I1;
    typedef InvPassFp<LOGN, LOGN - 8, 4> I2;
    typedef InvPassFp<LOGN, LOGN - 4, 4> I3;
    static_assert(I0::TF == P::I0::TF && I1::TF == P::I1::TF && I2::TF == P::I2::TF && I3::TF == P::I3::TF &&
                      I0::R == P::I0::R,
                  "the passes of the integer transform");
    double d[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) d[i] = fp_from_u64(v[i]);
    I0::run(d, tid, itw, m, hh);
    lds_exchange_fp<typename P::I0, typename P::I1>(d, s, tid);
    I1::run(d, tid, itw, m, hh);
    lds_exchange_fp<typename P::I1, typename P::I2>(d, s, tid);
    I2::run(d, tid, itw, m, hh);
    lds_exchange_fp<typename P::I2, typename P::I3>(d, s, tid);
    I3::run(d, tid, itw, m, hh);
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = fp_word(fp_canon(scale ? fp_mulmod(d[i], ninv, m) : d[i], m));
}

}  // namespace mkhip
