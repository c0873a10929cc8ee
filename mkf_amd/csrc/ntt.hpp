// LDS-resident negacyclic NTT / INTT of one RNS limb for N = 2^13 and 2^14 (and the
// register-split N = 2^15 variant) on one workgroup of N/16 threads.
//
// Transform definition (what must match bit for bit): the reference's ring NTT
// (Lattigo v2.3.0 NTTLazy / InvNTTLazy, restated in
// hardware/u55c/hls/kernels/par1_mkf_ckks_kernel.cpp:92-157): forward stage s
// (t = N >> (s+1)) pairs j and j+t of group g = j / 2t with twiddle psi[2^s + g]
// (Cooley-Tukey, natural order in, bit-reversed order out); inverse stage u
// (t = 2^u) uses psiinv[(N >> (u+1)) + g] (Gentleman-Sande) and finishes with a
// multiplication by N^-1.  psi[brv(j)] = psi^j for a primitive 2N-th root psi.
//
// Implementation: every thread holds 16 coefficients in registers.  Stages are
// processed in passes of up to 4 (radix-16 in registers); between passes the
// limb is exchanged through LDS (N * 8 bytes plus one pad word per 16: 136 KiB
// at N = 2^14).  Butterflies
// are Harvey's lazy forms with Shoup twiddles whose quotient is approximated
// (shoup_lazy4: product in [0, 4q)); values stay in [0, 8q) forward and
// [0, 4q) inverse (q < 2^61) and are reduced to canonical residues on the way out.
#pragma once

#include "modarith.hpp"

namespace mkhip {

// ---- butterflies ---------------------------------------------------------

// q4 = 4q.  In: x, y in [0, 8q).  Out: x, y in [0, 8q).
// The sum u + t is formed inside the product (shoup_lazy4_add) and the
// difference as 2u + 4q - (u + t): one shift-add and one subtraction.
__device__ __forceinline__ void fwd_bfly(u64 &x, u64 &y, u64 w, u64 wp, u64 q4, u64 q) {
    const u64 u = x >= q4 ? x - q4 : x;           // [0, 4q)
    const u64 s = shoup_lazy4_add(y, w, wp, q, u);  // u + t, t in [0, 4q): [0, 8q)
    x = s;
    y = (u << 1) + q4 - s;                        // u + 4q - t in (0, 8q)
}

// Unreduced form for moduli with headroom (q < 2^57): no conditional subtraction;
// each stage adds at most 4q to the bound (x + t, x + 4q - t with t < 4q), so 15
// stages from inputs < 8q stay below 68q < 2^64.
__device__ __forceinline__ void fwd_bfly_lazy(u64 &x, u64 &y, u64 w, u64 wp, u64 q4, u64 q) {
    const u64 u = x;
    const u64 s = shoup_lazy4_add(y, w, wp, q, u);  // u + t, t in [0, 4q)
    x = s;
    y = (u << 1) + q4 - s;                          // u + 4q - t
}

// The moduli whose forward transforms run unreduced (and end with reduce_big).
__device__ __forceinline__ bool fwd_lazy_ok(u64 q) { return q >= (1ull << 32) && q < (1ull << 57); }

// In: x, y in [0, 4q).  Out: x, y in [0, 4q).
__device__ __forceinline__ void inv_bfly(u64 &x, u64 &y, u64 w, u64 wp, u64 q4, u64 q) {
    u64 u = x, v = y;
    u64 s = u + v;
    x = s >= q4 ? s - q4 : s;              // [0, 4q)
    y = shoup_lazy4(u + q4 - v, w, wp, q); // [0, 4q)
}

// Unreduced inverse butterfly: x, y < B (c = B, 2B <= 2^64, B >= 2q).  Out:
// x + y < 2B and (x + B - y) w lazily in [0, 4q) <= 2B -- no conditional
// subtraction.
__device__ __forceinline__ void inv_bfly_lazy(u64 &x, u64 &y, u64 w, u64 wp, u64 c, u64 q) {
    const u64 u = x, v = y;
    x = u + v;
    y = shoup_lazy4(u + c - v, w, wp, q);
}

// The moduli whose inverse transforms may run unreduced (ntt_inv with mu != 0):
// reduce_big needs q >= 2^32, and a pass of up to four stages from values below
// q must stay below 2^64.
__device__ __forceinline__ bool inv_lazy_ok(u64 q) { return q >= (1ull << 32) && q < (1ull << 57); }

// [0, 8q) -> [0, q)
__device__ __forceinline__ u64 reduce_8q(u64 x, u64 q) {
    const u64 q4 = q << 2, q2 = q << 1;
    x = x >= q4 ? x - q4 : x;
    x = x >= q2 ? x - q2 : x;
    return x >= q ? x - q : x;
}

// Twiddle (w, w') at index idx.  When idx is the same on every lane of the wave
// the pair is moved to scalar registers (readfirstlane), so it occupies no VGPRs
// while the butterflies of the stage run.
// A uniform index is read through the constant address space, which compiles to a
// scalar load (waited on with lgkmcnt, so it is not held up behind the wave's
// outstanding vector-memory operations).
template <bool UNIFORM>
__device__ __forceinline__ ulonglong2 twiddle(const ulonglong2 *__restrict__ tw, int idx) {
    ulonglong2 w;
    if (UNIFORM) {
        typedef __attribute__((address_space(4))) const u64 cu64;
        const cu64 *p = (const cu64 *)tw + 2 * __builtin_amdgcn_readfirstlane(idx);
        w.x = p[0];
        w.y = p[1];
    } else {
        typedef __attribute__((address_space(1))) const u64 gu64c;
        const gu64c *p = (const gu64c *)tw + 2 * idx;
        w.x = p[0];
        w.y = p[1];
    }
    return w;
}

// Closed form of the padded-LDS condition (the CPU tests check it
// exhaustively) for a pass of line length L (TL or TF) and radix R on T lanes:
// with L a power of two dividing T and T R a multiple of 16, base mod 16 is below
// min(L, 16) (plus multiples of R L when R L < 16) and off mod 16 is a multiple of
// L below 16 minus that bound, so base mod 16 + off mod 16 never reaches 16.
__host__ __device__ constexpr bool pad_split_ok(int L, int T, int R) {
    return L > 0 && (L & (L - 1)) == 0 && T % L == 0 && (T * R) % 16 == 0;
}

// ---- pass geometry -------------------------------------------------------
// A forward pass (S0, NS) covers stages S0..S0+NS-1.  Its "columns" are the sets
// of 2^NS coefficients that interact within the pass: column col has base
// b * 2^NS * TL + r (b = col / TL, r = col % TL, TL = t of the pass's last stage)
// and stride TL.  Thread tid owns columns tid + c * T (c < 16 / 2^NS).

template <int LOGN, int S0, int NS>
struct FwdPass {
    static constexpr int N = 1 << LOGN;
    static constexpr int T = N / 16;
    static constexpr int TL = N >> (S0 + NS);
    static constexpr int R = 1 << NS;
    static constexpr int C = 16 / R;
    __host__ __device__ static constexpr int elem(int tid, int i) {
        const int c = i / R, k = i % R;
        const int col = tid + c * T;
        const int b = col / TL, r = col % TL;
        return b * R * TL + r + k * TL;
    }
    // elem = base(tid) + off(i) (TL divides T): the lane's part and a constant
    __host__ __device__ static constexpr int base(int tid) { return (tid / TL) * R * TL + tid % TL; }
    __host__ __device__ static constexpr int off(int i) { return (i / R) * T * R + (i % R) * TL; }
    static constexpr bool pad_ok = pad_split_ok(TL, T, R);
    // hh: twiddle multiplier -- 1 for a whole transform; 2 + h for half h of a
    // transform of twice this size (its stage s twiddles start at 2^(s+1) + h 2^s)
    template <bool LAZY = false>
    __device__ static __forceinline__ void run(u64 (&v)[16], int tid, const ulonglong2 *__restrict__ tw, u64 q,
                                               int hh) {
        const u64 q4 = q << 2;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int col = tid + c * T;
            // b is the same on every lane of a wave when TL is a multiple of 64:
            // the twiddles then come through scalar loads
            const int b = (TL % 64 == 0) ? __builtin_amdgcn_readfirstlane(col / TL) : col / TL;
            // no scheduling barrier between the stages: the scheduler interleaves
            // the end of one stage with the start of the next (measured faster in
            // the decompositions; the inverse passes keep theirs)
#pragma unroll
            for (int u = 0; u < NS; ++u) {
                const int half = R >> (u + 1);
#pragma unroll
                for (int grp = 0; grp < (1 << u); ++grp) {
                    const ulonglong2 w = twiddle<(TL % 64 == 0)>(tw, (1 << (S0 + u)) * hh + b * (1 << u) + grp);
#pragma unroll
                    for (int kk = 0; kk < half; ++kk) {
                        const int k0 = grp * 2 * half + kk;
                        if (LAZY)
                            fwd_bfly_lazy(v[c * R + k0], v[c * R + k0 + half], w.x, w.y, q4, q);
                        else
                            fwd_bfly(v[c * R + k0], v[c * R + k0 + half], w.x, w.y, q4, q);
                    }
                }
            }
        }
    }
};

// An inverse pass (U0, NS) covers inverse stages U0..U0+NS-1 (t = 2^U0 .. 2^(U0+NS-1)).
template <int LOGN, int U0, int NS>
struct InvPass {
    static constexpr int STAGES = NS;
    static constexpr int N = 1 << LOGN;
    static constexpr int T = N / 16;
    static constexpr int TF = 1 << U0;
    static constexpr int R = 1 << NS;
    static constexpr int C = 16 / R;
    __host__ __device__ static constexpr int elem(int tid, int i) {
        const int c = i / R, k = i % R;
        const int col = tid + c * T;
        const int b = col / TF, r = col % TF;
        return b * R * TF + r + k * TF;
    }
    __host__ __device__ static constexpr int base(int tid) { return (tid / TF) * R * TF + tid % TF; }
    __host__ __device__ static constexpr int off(int i) { return (i / R) * T * R + (i % R) * TF; }
    static constexpr bool pad_ok = pad_split_ok(TF, T, R);
    // LAZY: unreduced butterflies (inv_bfly_lazy); the values enter below bnd,
    // stage u's below bnd 2^u, and leave below bnd 2^NS (bnd 2^(NS-1) <= 2^63)
    template <bool LAZY = false>
    __device__ static __forceinline__ void run(u64 (&v)[16], int tid, const ulonglong2 *__restrict__ itw, u64 q,
                                               int hh, u64 bnd = 0) {
        const u64 q4 = q << 2;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const int col = tid + c * T;
            const int b = (TF % 64 == 0) ? __builtin_amdgcn_readfirstlane(col / TF) : col / TF;
#pragma unroll
            for (int u = 0; u < NS; ++u) {
                __builtin_amdgcn_sched_barrier(0);
                const int half = 1 << u;
                const int m = N >> (U0 + u + 1);
                const int ngrp = R / (2 * half);
#pragma unroll
                for (int grp = 0; grp < ngrp; ++grp) {
                    const ulonglong2 w = twiddle<(TF % 64 == 0)>(itw, m * hh + b * ngrp + grp);
#pragma unroll
                    for (int kk = 0; kk < half; ++kk) {
                        const int k0 = grp * 2 * half + kk;
                        if (LAZY)
                            inv_bfly_lazy(v[c * R + k0], v[c * R + k0 + half], w.x, w.y, bnd << u, q);
                        else
                            inv_bfly(v[c * R + k0], v[c * R + k0 + half], w.x, w.y, q4, q);
                    }
                }
            }
        }
    }
};

// LDS word of element idx: one pad word after every 16, so that the strided
// patterns of the passes (lanes 2, 32 or 64 words apart within a wave) start on
// different banks (measured on MI355X: no bank-conflict cycles in the half-limb
// transforms, 0.11 per active LDS cycle at 2^14 -- against 1.2 with one pad word
// per 32).  Every pass layout splits as elem = base(tid) + off(i) with
// base mod 16 + off mod 16 < 16, so lds_pad(elem) = lds_pad(base) + lds_pad(off):
// one address register per lane and compile-time offsets for the 16 values
// (pad_ok asserts the closed form of this condition at compile time; the CPU
// tests check every lane).
__host__ __device__ constexpr int lds_pad(int idx) { return idx + (idx >> 4); }

// LDS words a workgroup holding ns elements uses.
__host__ __device__ constexpr int lds_words(int ns) { return ns + ns / 16; }

template <class From, class To>
__device__ __forceinline__ void lds_exchange(u64 (&v)[16], u64 *s, int tid) {
    static_assert(From::pad_ok && To::pad_ok, "padded LDS layout");
    u64 *w = s + lds_pad(From::base(tid));
#pragma unroll
    for (int i = 0; i < 16; ++i) w[lds_pad(From::off(i))] = v[i];
    __syncthreads();
    const u64 *r = s + lds_pad(To::base(tid));
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = r[lds_pad(To::off(i))];
    __syncthreads();
}

// Natural coalesced layout: element tid + i * T.
template <int LOGN>
struct Linear {
    static constexpr int T = (1 << LOGN) / 16;
    __device__ static __forceinline__ int elem(int tid, int i) { return tid + i * T; }
};

// ---- whole transforms ------------------------------------------------------
// Input: v[i] = coefficient tid + i*T (values < 8q).  Output: v[i] = NTT value at
// index tid + i*T, canonical.  s: N words of LDS.

template <int LOGN>
struct NttPlan;

template <>
struct NttPlan<14> {
    typedef FwdPass<14, 0, 4> F0;
    typedef FwdPass<14, 4, 4> F1;
    typedef FwdPass<14, 8, 4> F2;
    typedef FwdPass<14, 12, 2> F3;
    typedef InvPass<14, 0, 2> I0;
    typedef InvPass<14, 2, 4> I1;
    typedef InvPass<14, 6, 4> I2;
    typedef InvPass<14, 10, 4> I3;
};

template <>
struct NttPlan<13> {
    typedef FwdPass<13, 0, 4> F0;
    typedef FwdPass<13, 4, 4> F1;
    typedef FwdPass<13, 8, 4> F2;
    typedef FwdPass<13, 12, 1> F3;
    typedef InvPass<13, 0, 1> I0;
    typedef InvPass<13, 1, 4> I1;
    typedef InvPass<13, 5, 4> I2;
    typedef InvPass<13, 9, 4> I3;
};

// "Quad" layout: the layout of the last forward pass, which is also the layout
// of the first inverse pass.  Lane tid holds, for c < 16/R, the R consecutive
// elements (tid + c N/16) R + k (R = 4 at N = 2^14, 2 at N = 2^13), so a wave
// reads or writes 64 R consecutive words per c.  Forward transforms leave their
// output in this layout and inverse transforms take their input in it: the
// exchange through LDS to the natural layout is skipped on both sides, and the
// kernels read and write their rows at these element positions instead.
template <int LOGN>
struct Quad {
    typedef typename NttPlan<LOGN>::F3 F;
    typedef typename NttPlan<LOGN>::I0 I;
    static constexpr int R = F::R;
    static constexpr int T = (1 << LOGN) / 16;
    static_assert(F::TL == 1 && I::TF == 1 && I::R == R, "last forward / first inverse pass must match");
    __device__ static __forceinline__ int elem(int tid, int i) { return (tid + (i / R) * T) * R + (i % R); }
};

// Forward NTT; v in Linear layout in, Quad layout out; returns canonical values.
// With hh = 2 + h it is the transform of half h of a limb of 2^(LOGN+1) after
// that limb's first stage (fwd_first_stage).
// LAZY (2^32 <= q < 2^57, mu = floor(2^64 / q)): unreduced butterflies, one
// reduction per value at the end (reduce_big); with canon = false that reduction
// is left out and the values stay below 68q (< 2^63.1), for consumers that take
// them as Montgomery operands (see lazy_digits_ok).
template <int LOGN, bool LAZY = false>
__device__ __forceinline__ void ntt_fwd(u64 (&v)[16], u64 *s, int tid, const ulonglong2 *__restrict__ tw,
                                        u64 q, int hh = 1, u64 mu = 0, bool canon = true) {
    typedef NttPlan<LOGN> P;
    static_assert(P::F0::TL == Linear<LOGN>::T, "first pass must be the linear layout");
    P::F0::template run<LAZY>(v, tid, tw, q, hh);
    lds_exchange<typename P::F0, typename P::F1>(v, s, tid);
    P::F1::template run<LAZY>(v, tid, tw, q, hh);
    lds_exchange<typename P::F1, typename P::F2>(v, s, tid);
    P::F2::template run<LAZY>(v, tid, tw, q, hh);
    lds_exchange<typename P::F2, typename P::F3>(v, s, tid);
    P::F3::template run<LAZY>(v, tid, tw, q, hh);
    if (LAZY) {
        if (canon) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                v[i] = reduce_big(v[i], q, mu);
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = reduce_8q(v[i], q);
    }
}

// Forward NTT choosing the unreduced form when the modulus allows it (a
// wave-uniform branch: one limb has one modulus).
template <int LOGN>
__device__ __forceinline__ void ntt_fwd_any(u64 (&v)[16], u64 *s, int tid, const ulonglong2 *__restrict__ tw,
                                            u64 q, u64 mu, int hh = 1) {
    if (fwd_lazy_ok(q))
        ntt_fwd<LOGN, true>(v, s, tid, tw, q, hh, mu);
    else
        ntt_fwd<LOGN, false>(v, s, tid, tw, q, hh);
}

// Inverse NTT including the N^-1 scaling; v < 4q in Quad layout; returns
// canonical values in Linear layout.  With hh = 2 + h and scale = false it runs
// every stage of a limb of 2^(LOGN+1) that stays inside half h and leaves the
// values in [0, 4q); inv_last_stage finishes the limb.
// Unreduced inverse (mu = floor(2^64 / q) given and inv_lazy_ok(q)): the
// Gentleman-Sande sums are not reduced; the bound of the values is tracked
// per pass (wave-uniform) and, when the pass's last stage could exceed 2^64,
// the values are first brought to [0, q) (reduce_big).  The scaling takes any
// value below 2^64 (exact Shoup quotient); without it (split limbs) the values
// are reduced at the end, so that the finishing stage gets them below 4q.
template <class Pass>
__device__ __forceinline__ void inv_pass_lazy(u64 (&v)[16], int tid, const ulonglong2 *__restrict__ itw, u64 q,
                                              u64 mu, int hh, u64 &bnd) {
    if (bnd > (~0ull >> Pass::STAGES)) {  // bnd 2^(STAGES-1) > 2^63
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = reduce_big(v[i], q, mu);
        // 4q, not q: a stage's differences leave below 4q whatever its inputs,
        // so a bound below 2q would not hold after the pass's first stage
        bnd = q << 2;
    }
    Pass::template run<true>(v, tid, itw, q, hh, bnd);
    bnd <<= Pass::STAGES;
}

template <int LOGN>
__device__ __forceinline__ void ntt_inv(u64 (&v)[16], u64 *s, int tid, const ulonglong2 *__restrict__ itw,
                                        u64 q, u64 ninv, u64 ninvp, int hh = 1, bool scale = true, u64 mu = 0) {
    typedef NttPlan<LOGN> P;
    static_assert(P::I3::TF * 16 == (1 << LOGN), "last inverse pass must be the linear layout");
    if (mu != 0 && inv_lazy_ok(q)) {
        u64 bnd = q << 2;
        inv_pass_lazy<typename P::I0>(v, tid, itw, q, mu, hh, bnd);
        lds_exchange<typename P::I0, typename P::I1>(v, s, tid);
        inv_pass_lazy<typename P::I1>(v, tid, itw, q, mu, hh, bnd);
        lds_exchange<typename P::I1, typename P::I2>(v, s, tid);
        inv_pass_lazy<typename P::I2>(v, tid, itw, q, mu, hh, bnd);
        lds_exchange<typename P::I2, typename P::I3>(v, s, tid);
        inv_pass_lazy<typename P::I3>(v, tid, itw, q, mu, hh, bnd);
        if (!scale) {
#pragma unroll
            for (int i = 0; i < 16; ++i) v[i] = reduce_big(v[i], q, mu);
        }
    } else {
        P::I0::run(v, tid, itw, q, hh);
        lds_exchange<typename P::I0, typename P::I1>(v, s, tid);
        P::I1::run(v, tid, itw, q, hh);
        lds_exchange<typename P::I1, typename P::I2>(v, s, tid);
        P::I2::run(v, tid, itw, q, hh);
        lds_exchange<typename P::I2, typename P::I3>(v, s, tid);
        P::I3::run(v, tid, itw, q, hh);
    }
    if (scale) {
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = shoup(v[i], ninv, ninvp, q);
    }
}

// ---- limbs of 2^15: two workgroups per limb, one per half -----------------
// Forward: the first stage (t = N/2, twiddle tw[1]) pairs element j of the
// lower half with element j of the upper half; half h keeps x + w y (h = 0) or
// x - w y (h = 1), after which each half is a transform of N/2 (hh = 2 + h).
// lo / hi: this lane's elements of the two halves (values < 8q); out in [0, 8q).
__device__ __forceinline__ void fwd_first_stage(u64 (&v)[16], const u64 (&lo)[16], const u64 (&hi)[16],
                                                const ulonglong2 *__restrict__ tw, u64 q, int h) {
    const ulonglong2 w = twiddle<true>(tw, 1);
    const u64 q4 = q << 2;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        u64 x = lo[i], y = hi[i];
        fwd_bfly(x, y, w.x, w.y, q4, q);
        v[i] = h ? y : x;
    }
}

// fwd_first_stage for the unreduced transforms (2^32 <= q < 2^57): only the kept
// result is formed, with no conditional subtraction; out < lo + 4q (< 12q), after
// which the 14 or fewer remaining unreduced stages stay below 68q < 2^64.
__device__ __forceinline__ void fwd_first_stage_lazy(u64 (&v)[16], const u64 (&lo)[16], const u64 (&hi)[16],
                                                     const ulonglong2 *__restrict__ tw, u64 q, int h) {
    const ulonglong2 w = twiddle<true>(tw, 1);
    // half 1 keeps lo + 4q - t = lo + (t ^ ~0) + 4q + 1, half 0 lo + t: one form
    // with a uniform mask, so that each value is formed as soon as its product is
    const u64 mask = 0 - (u64)(__builtin_amdgcn_readfirstlane(h) & 1);
    const u64 add = mask & ((q << 2) + 1);
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = lo[i] + (shoup_lazy4(hi[i], w.x, w.y, q) ^ mask) + add;
}

// The same stage keeping both results: lo <- x + w y, hi <- x - w y.
__device__ __forceinline__ void fwd_first_stage_both(u64 (&lo)[16], u64 (&hi)[16], const ulonglong2 *__restrict__ tw,
                                                     u64 q) {
    const ulonglong2 w = twiddle<true>(tw, 1);
    const u64 q4 = q << 2;
#pragma unroll
    for (int i = 0; i < 16; ++i) fwd_bfly(lo[i], hi[i], w.x, w.y, q4, q);
}

// Inverse: after each half ran its N/2 stages (values in [0, 4q)), the last stage
// (t = N/2, twiddle itw[1]) and the N^-1 scaling: x' = (x + y) N^-1,
// y' = (x - y) (w N^-1), canonical.  wn = w N^-1 mod q with its Shoup companion.
__device__ __forceinline__ void inv_last_stage(u64 &x, u64 &y, u64 wn, u64 wnp, u64 ninv, u64 ninvp, u64 q) {
    const u64 q4 = q << 2;
    const u64 sum = x + y;
    const u64 dif = x + q4 - y;
    x = shoup(sum, ninv, ninvp, q);
    y = shoup(dif, wn, wnp, q);
}

}  // namespace mkhip
