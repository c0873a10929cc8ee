// Device helpers shared by the kernel translation units.
#pragma once

#include "ntt.hpp"
#include "nttfp.hpp"
#include "tables.hpp"

namespace mkhip {

extern __shared__ u64 g_lds[];

// Pointers read from descriptor tables are generic to the compiler; every HBM
// access goes through these global-address-space views so that it compiles to
// global (not flat) memory instructions with a scalar base.
typedef __attribute__((address_space(1))) u64 gu64;
__device__ __forceinline__ const gu64 *G(const u64 *p) { return (const gu64 *)p; }
__device__ __forceinline__ gu64 *G(u64 *p) { return (gu64 *)p; }

// Scalar base of a row: p must be the same on every lane of the wave (all call
// sites pass a per-workgroup row pointer); the per-lane offset is then a single
// 32-bit register shared by all rows.
__device__ __forceinline__ const gu64 *uniform_ptr(const u64 *p) {
    const unsigned long long a = (unsigned long long)p;
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)a), hi = __builtin_amdgcn_readfirstlane((u32)(a >> 32));
    return (const gu64 *)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ gu64 *uniform_ptr(u64 *p) { return (gu64 *)uniform_ptr((const u64 *)p); }

template <int LOGN>
__device__ __forceinline__ void load_row_cached(u64 (&v)[16], const u64 *__restrict__ src, int tid) {
    constexpr int T = (1 << LOGN) / 16;
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = uniform_ptr(src + i * T)[tid];
}

// Rows are stored non-temporally: every row a kernel writes is read back by a
// later kernel, after far more data than the L2 holds has streamed through, so
// keeping it in L2 only evicts what the running kernels reuse (twiddles, keys,
// input limbs shared by several workgroups).  Measured on the headline batch:
// 4 % more throughput, most of it in the external products.
template <int LOGN>
__device__ __forceinline__ void store_row(const u64 (&v)[16], u64 *__restrict__ dst, int tid) {
    constexpr int T = (1 << LOGN) / 16;
#pragma unroll
    for (int i = 0; i < 16; ++i) __builtin_nontemporal_store(v[i], uniform_ptr(dst + i * T) + tid);
}

// Asynchronous copy of one row of N = 2^LOGN words (natural order) into LDS words
// 0..N-1, without registers: global_load_lds_dwordx4 writes lane l's 16 bytes at
// the wave's LDS base + 16 l.  Block b of 128 words goes to LDS words 128 b.. ;
// each of the N/1024 waves copies 8 blocks.  Complete after s_waitcnt vmcnt(0) in
// every wave and a barrier.
template <int LOGN>
__device__ __forceinline__ void row_to_lds_async(const u64 *row, int tid) {
    typedef __attribute__((address_space(1))) void gvoid;
    typedef __attribute__((address_space(3))) void lvoid;
    constexpr int NW = (1 << LOGN) / 16 / 64;  // waves of the workgroup
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int b = j * NW + wave;
        __builtin_amdgcn_global_load_lds((gvoid *)(row + b * 128 + lane * 2), (lvoid *)(g_lds + b * 128), 16, 0, 0);
    }
}

// 16-byte global accesses (two consecutive words)
typedef unsigned long long u64v2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) u64v2 gv2;

// Rows in the Quad layout of the transforms (Quad<LOGN> in ntt.hpp): v[i] is
// element Quad::elem(tid, i); the R consecutive words of each column move as
// 16-byte accesses.
template <int LOGN>
__device__ __forceinline__ void load_row_quad(u64 (&v)[16], const u64 *__restrict__ src, int tid) {
    typedef Quad<LOGN> Qd;
    const gu64 *s = uniform_ptr(src);
#pragma unroll
    for (int i = 0; i < 16; i += 2) {
        const u64v2 x = *(const gv2 *)(s + Qd::elem(tid, i));
        v[i] = x.x;
        v[i + 1] = x.y;
    }
}

template <int LOGN>
__device__ __forceinline__ void store_row_quad(const u64 (&v)[16], u64 *__restrict__ dst, int tid) {
    typedef Quad<LOGN> Qd;
    gu64 *d = uniform_ptr(dst);
#pragma unroll
    for (int i = 0; i < 16; i += 2) {
        u64v2 x;
        x.x = v[i];
        x.y = v[i + 1];
        __builtin_nontemporal_store(x, (gv2 *)(d + Qd::elem(tid, i)));
    }
}

// Wave-uniform copy of an index the compiler cannot prove uniform (one derived from
// blockIdx through a run-time division, which it computes on the VALU): the
// descriptor loads and row addresses that depend on it then stay in scalar
// registers instead of being rebuilt with readfirstlane before every access.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

// A wave-uniform pointer made opaque to the optimizer: table loads that use it
// stay where they are written instead of being hoisted out of an enclosing loop
// and held in registers across its iterations.
template <class P>
__device__ __forceinline__ P *opaque(P *p) {
    const unsigned long long a = (unsigned long long)p;
    u32 lo = __builtin_amdgcn_readfirstlane((u32)a), hi = __builtin_amdgcn_readfirstlane((u32)(a >> 32));
    asm volatile("" : "+s"(lo), "+s"(hi));
    return (P *)(((unsigned long long)hi << 32) | lo);
}
// likewise for the lane index, so that the LDS exchange addresses derived from it
// are recomputed per transform rather than kept live across a loop
__device__ __forceinline__ int opaque_lane(int t) {
    asm volatile("" : "+v"(t));
    return t;
}

// ---- limb geometry -------------------------------------------------------------
// One workgroup of TH = NS/16 threads transforms NS = min(N, 2^14) elements in
// registers + LDS.  A limb of 2^15 is split over H = 2 workgroups (halves): the
// forward transform's first stage is computed by both from the two halves, the
// inverse leaves its last stage to k_intt_finish.  The elements a workgroup
// holds are h NS + tid + e TH (e < 16).
// SPLIT = true at LOGN = 14 is the same halving for kernels that choose it (two
// workgroups of half the LDS per limb, so that two run on one CU).
template <int LOGN, bool SPLIT = (LOGN == 15)>
struct Geo {
    static constexpr bool split = SPLIT;
    static constexpr int LS = split ? LOGN - 1 : LOGN;
    static constexpr int H = split ? 2 : 1;
    static constexpr int N = 1 << LOGN;
    static constexpr int NS = 1 << LS;
    static constexpr int TH = NS / 16;
};

// MODE: -1 picks the unreduced butterflies at run time when the modulus allows
// them; 1 / 0 force the unreduced / reduced form (the caller checked fwd_lazy_ok).
// canon = false: the unreduced form (MODE 1) leaves its output below 68q.
template <int LOGN, int MODE = -1>
__device__ __forceinline__ void fwd_ntt_mode(u64 (&v)[16], int tid, const ulonglong2 *tw, const ModConst &m,
                                             int hh, bool canon = true) {
    if constexpr (MODE == 1)
        ntt_fwd<LOGN, true>(v, g_lds, tid, tw, m.q, hh, m.mu, canon);
    else if constexpr (MODE == 0)
        ntt_fwd<LOGN, false>(v, g_lds, tid, tw, m.q, hh);
    else
        ntt_fwd_any<LOGN>(v, g_lds, tid, tw, m.q, m.mu, hh);
}

// Forward transform of this workgroup's part of a limb whose element c is f(c)
// (values < 8q); v receives the NTT values at h NS + tid + e TH, canonical.
template <int LOGN, int MODE = -1, bool SPLIT = (LOGN == 15), class F>
__device__ __forceinline__ void fwd_transform(u64 (&v)[16], F f, int tid, int h, const ulonglong2 *tw,
                                              const ModConst &m) {
    typedef Geo<LOGN, SPLIT> Gm;
    if constexpr (Gm::split) {
        u64 lo[16], hi[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            lo[e] = f(tid + e * Gm::TH);
            hi[e] = f(Gm::NS + tid + e * Gm::TH);
        }
        if constexpr (MODE == 1)
            fwd_first_stage_lazy(v, lo, hi, tw, m.q, h);
        else
            fwd_first_stage(v, lo, hi, tw, m.q, h);
        fwd_ntt_mode<Gm::LS, MODE>(v, tid, tw, m, 2 + h);
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = f(tid + e * Gm::TH);
        fwd_ntt_mode<LOGN, MODE>(v, tid, tw, m, 1);
    }
}

// fwd_transform of a stored limb (row: its N words, values < 3q when red, else
// < 8q).  The reduction is a wave-uniform branch over all values, not a
// per-value select, and the loads use the row's scalar base.
template <int LOGN, int MODE = -1, bool SPLIT = (LOGN == 15)>
__device__ __forceinline__ void fwd_transform_row(u64 (&v)[16], const u64 *row, bool red, int tid, int h,
                                                  const ulonglong2 *tw, const ModConst &m, bool canon = true) {
    typedef Geo<LOGN, SPLIT> Gm;
    if constexpr (Gm::split) {
        u64 lo[16], hi[16];
        load_row_cached<Gm::LS>(lo, row, tid);
        load_row_cached<Gm::LS>(hi, row + Gm::NS, tid);
        if (red) {
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                lo[e] = reduce64(lo[e], m.q, m.mu);
                hi[e] = reduce64(hi[e], m.q, m.mu);
            }
        }
        if constexpr (MODE == 1)
            fwd_first_stage_lazy(v, lo, hi, tw, m.q, h);
        else
            fwd_first_stage(v, lo, hi, tw, m.q, h);
        fwd_ntt_mode<Gm::LS, MODE>(v, tid, tw, m, 2 + h, canon);
    } else {
        load_row_cached<LOGN>(v, row, tid);
        if (red) {
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = reduce64(v[e], m.q, m.mu);
        }
        fwd_ntt_mode<LOGN, MODE>(v, tid, tw, m, 1, canon);
    }
}

// fwd_transform_row in FP64 (modulus q < 2^50): v receives canonical residues as
// doubles, in the Quad layout.  Inputs below 2q (red: below 2^64, reduced first).
__device__ __forceinline__ FpMod fp_mod(const ModConst &m) {
    FpMod f;
    f.q = m.qd;
    f.qinv = m.qinvd;
    return f;
}

// fwd_transform in FP64: element c of the limb is f(c) (a word below 2q).
template <int LOGN, bool SPLIT = (LOGN == 15), class F>
__device__ __forceinline__ void fwd_transform_fp(double (&v)[16], F f, int tid, int h, const double *twd,
                                                 const ModConst &m) {
    typedef Geo<LOGN, SPLIT> Gm;
    const FpMod fm = fp_mod(m);
    if constexpr (Gm::split) {
        const double w = twiddle_fp<true>(twd, 1);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const double r = fp_mulmod(fp_from_u64(f(Gm::NS + tid + e * Gm::TH)), w, fm);
            const double x = fp_from_u64(f(tid + e * Gm::TH));
            v[e] = h ? x - r : x + r;
        }
        ntt_fwd_fp<Gm::LS>(v, g_lds, tid, twd, fm, 2 + h, 1);
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = fp_from_u64(f(tid + e * Gm::TH));
        ntt_fwd_fp<LOGN>(v, g_lds, tid, twd, fm, 1, 0);
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = fp_canon(v[e], fm);
}

template <int LOGN, bool SPLIT = (LOGN == 15)>
__device__ __forceinline__ void fwd_transform_row_fp(double (&v)[16], const u64 *row, bool red, int tid, int h,
                                                     const double *twd, const ModConst &m) {
    typedef Geo<LOGN, SPLIT> Gm;
    const FpMod fm = fp_mod(m);
    if constexpr (Gm::split) {
        u64 lo[16], hi[16];
        load_row_cached<Gm::LS>(lo, row, tid);
        load_row_cached<Gm::LS>(hi, row + Gm::NS, tid);
        if (red) {
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                lo[e] = reduce64(lo[e], m.q, m.mu);
                hi[e] = reduce64(hi[e], m.q, m.mu);
            }
        }
        // first stage (t = N/2, twiddle tw[1]): half h keeps x + w y or x - w y
        const double w = twiddle_fp<true>(twd, 1);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const double r = fp_mulmod(fp_from_u64(hi[e]), w, fm);
            const double x = fp_from_u64(lo[e]);
            v[e] = h ? x - r : x + r;
        }
        ntt_fwd_fp<Gm::LS>(v, g_lds, tid, twd, fm, 2 + h, 1);
    } else {
        u64 x[16];
        load_row_cached<LOGN>(x, row, tid);
        if (red) {
#pragma unroll
            for (int e = 0; e < 16; ++e) x[e] = reduce64(x[e], m.q, m.mu);
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = fp_from_u64(x[e]);
        ntt_fwd_fp<LOGN>(v, g_lds, tid, twd, fm, 1, 0);
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = fp_canon(v[e], fm);
}

// Inverse transform of this workgroup's elements (v[e] = element h NS + tid + e TH,
// values < 4q) with the N^-1 (or t N^-1) scaling given; for a split limb the last
// stage and the scaling are left to k_intt_finish.
// mu = floor(2^64 / q): the unreduced butterflies where the modulus allows them.
template <int LOGN, bool SPLIT = (LOGN == 15)>
__device__ __forceinline__ void inv_transform(u64 (&v)[16], int tid, int h, const ulonglong2 *itw, u64 q,
                                              u64 ninv, u64 ninvp, u64 mu = 0) {
    typedef Geo<LOGN, SPLIT> Gm;
    if constexpr (Gm::split)
        ntt_inv<Gm::LS>(v, g_lds, tid, itw, q, ninv, ninvp, 2 + h, false, mu);
    else
        ntt_inv<LOGN>(v, g_lds, tid, itw, q, ninv, ninvp, 1, true, mu);
}

// inv_transform choosing the FP64 arithmetic for moduli below 2^50 (a wave-uniform
// branch).  tscale: t N^-1 instead of N^-1.  Same words as the integer form.
template <int LOGN, bool SPLIT = (LOGN == 15)>
__device__ __forceinline__ void inv_transform_sel(u64 (&v)[16], int tid, int h, const DevTables &T, int ri,
                                                  const ModConst &m, bool tscale) {
    typedef Geo<LOGN, SPLIT> Gm;
    if (fp_ok(m.q) && T.fp64) {
        const double *itwd = opaque(T.itwd + (size_t)ri * Gm::N);
        if constexpr (Gm::split)
            ntt_inv_fp<Gm::LS>(v, g_lds, tid, itwd, fp_mod(m), 0.0, 2 + h, false);
        else
            ntt_inv_fp<LOGN>(v, g_lds, tid, itwd, fp_mod(m), fp_from_u64(tscale ? m.tninv : m.ninv), 1, true);
    } else {
        inv_transform<LOGN, SPLIT>(v, tid, h, opaque(T.itw + (size_t)ri * Gm::N), m.q, tscale ? m.tninv : m.ninv,
                                   tscale ? m.tninvp : m.ninvp, T.inv_lazy ? m.mu : 0);
    }
}

// Row r of a (level+1+K)-row QP polynomial is ring modulus r (Q) or L + r - level - 1 (P).
__device__ __forceinline__ int ring_index(int r, int level, int L) { return r <= level ? r : L + (r - level - 1); }

// 128-bit accumulate of a * b.
__device__ __forceinline__ void mac128(u64 &hi, u64 &lo, u64 a, u64 b) {
    u64 ph, pl;
    mul_wide(a, b, ph, pl);
    lo += pl;
    hi += ph + (lo < pl ? 1 : 0);
}

// Lazy Montgomery reduction of a 128-bit accumulator as in multSum
// (basis_extension.go:623-645): hi - hi(lo * pinv * p) + p.
__device__ __forceinline__ u64 mont_reduce128(u64 hi, u64 lo, u64 p, u64 pinv) {
    return hi - mulhi64(lo * pinv, p) + p;
}

// T * 2^-64 mod q, canonical, for T = sum of at most 32 products of canonical
// residues (q < 2^61: T < 32 q^2, so hi < 4q is brought below q first).  Equal to
// the sum of the products' MRed values mod q.
__device__ __forceinline__ u64 redc_sum(u64 hi, u64 lo, const ModConst &m) {
    const u64 q2 = m.q << 1;
    hi = hi >= q2 ? hi - q2 : hi;
    hi = hi >= m.q ? hi - m.q : hi;
    const u64 r = mont_reduce128(hi, lo, m.q, m.qinv);
    return r >= m.q ? r - m.q : r;
}

// External-product rows (Quad layout: value e is element Quad::elem(tid, e)): for
// the 16 values of this lane,
//   a0[e] = sum_i MRed(k0_i[e], d_i[e]) mod q      (i < ndig, canonical)
// and with NK = 2 the same with k1, written straight to park1 (the caller's output
// row, where it waits while a0 is transformed).  d_i = d + i dstride, k_i = k + i
// kstride.  Digits are walked in the outer loop (all 16 values of a row per
// digit, four loads in flight per chunk).
// LAZY (for 2^32 <= q and ndig q < 2^63, see ext_lazy_ok): the products'
// MRed values are summed unreduced in (0, 2q) each and brought to canonical once,
// after the last digit.
template <int LOGN, int NK, bool LAZY = false>
__device__ __forceinline__ void ext_rows(u64 (&a0)[16], u64 *park1, const u64 *d, size_t dstride, const u64 *k0,
                                         const u64 *k1, size_t kstride, int ndig, int tid, const ModConst &m) {
    typedef Quad<LOGN> Qd;
    u64 a1[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) a0[e] = a1[e] = 0;
    for (int i = 0; i < ndig; ++i) {
        const u64 *dr = d + i * dstride, *k0r = k0 + i * kstride, *k1r = NK > 1 ? k1 + i * kstride : nullptr;
#pragma unroll
        for (int c = 0; c < 16; c += 4) {
            u64 x[4], y0[4], y1[4];
#pragma unroll
            for (int e = 0; e < 4; e += 2) {
                const int off = Qd::elem(tid, c + e);
                const u64v2 xd = *(const gv2 *)(uniform_ptr(dr) + off), yd = *(const gv2 *)(uniform_ptr(k0r) + off);
                x[e] = xd.x;
                x[e + 1] = xd.y;
                y0[e] = yd.x;
                y0[e + 1] = yd.y;
                if (NK > 1) {
                    const u64v2 zd = *(const gv2 *)(uniform_ptr(k1r) + off);
                    y1[e] = zd.x;
                    y1[e + 1] = zd.y;
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                if constexpr (LAZY) {
                    a0[c + e] += mred_lazy(y0[e], x[e], m.q, m.qinv);
                    if (NK > 1) a1[c + e] += mred_lazy(y1[e], x[e], m.q, m.qinv);
                } else {
                    a0[c + e] = add_mod(a0[c + e], mred(y0[e], x[e], m.q, m.qinv), m.q);
                    if (NK > 1) a1[c + e] = add_mod(a1[c + e], mred(y1[e], x[e], m.q, m.qinv), m.q);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if constexpr (LAZY) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            a0[e] = reduce_big(a0[e], m.q, m.mu);
            if (NK > 1) a1[e] = reduce_big(a1[e], m.q, m.mu);
        }
    }
    if constexpr (NK > 1) store_row_quad<LOGN>(a1, park1, tid);
}

// ext_rows with carry-save accumulation (Acc3): the 16 values are walked in
// chunks (four values for one key, two for two keys) with the digits inside, so
// that only a chunk's three-word sums are live; each sum of ndig products key * digit is reduced once
// (redc128), to the same canonical residue as the sum of the products' MRed
// values.  For moduli where those sums fit (ext_cs_ok).
template <int LOGN, int NK>
__device__ __forceinline__ void ext_rows_cs(u64 (&a0)[16], u64 *park1, const u64 *d, size_t dstride,
                                            const u64 *k0, const u64 *k1, size_t kstride, int ndig, int tid,
                                            const ModConst &m) {
    typedef Quad<LOGN> Qd;
    // values per chunk: the sums of one key for four values, or of two keys for two
    constexpr int CH = NK > 1 ? 2 : 4;
    // finished values wait in this lane's LDS words (the LDS is free until the
    // inverse transform), so that no more than a chunk is live in registers
    u64 *park0 = g_lds;
#pragma unroll
    for (int c = 0; c < 16; c += CH) {
        Acc3 s0[CH], s1[CH];
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            acc3_zero(s0[e]);
            if (NK > 1) acc3_zero(s1[e]);
        }
        // digits two at a time (all their loads issued before the first product),
        // then an odd last one
        auto digits = [&](int i, auto npair) {
            constexpr int NP = decltype(npair)::value;
            __builtin_amdgcn_sched_barrier(0);
            u64v2 xd[NP][CH / 2], yd[NP][CH / 2], zd[NP][CH / 2];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const gu64 *dr = uniform_ptr(d + (i + j) * dstride), *k0r = uniform_ptr(k0 + (i + j) * kstride);
#pragma unroll
                for (int e = 0; e < CH; e += 2) {
                    const int off = Qd::elem(tid, c + e);
                    // one key: a chunk covers whole 32-byte segments of the digit
                    // rows, each read once (non-temporal, as they were stored);
                    // two keys: a chunk reads half a segment, whose other half the
                    // next chunk reads from L2 (and the Galois plans read one
                    // decomposition for several rotations)
                    if constexpr (NK == 1)
                        xd[j][e / 2] = __builtin_nontemporal_load((const gv2 *)(dr + off));
                    else
                        xd[j][e / 2] = *(const gv2 *)(dr + off);
                    yd[j][e / 2] = *(const gv2 *)(k0r + off);
                    if (NK > 1) zd[j][e / 2] = *(const gv2 *)(uniform_ptr(k1 + (i + j) * kstride) + off);
                }
            }
#pragma unroll
            for (int j = 0; j < NP; ++j)
#pragma unroll
                for (int e = 0; e < CH; e += 2) {
                    acc3_mac(s0[e], yd[j][e / 2].x, xd[j][e / 2].x);
                    acc3_mac(s0[e + 1], yd[j][e / 2].y, xd[j][e / 2].y);
                    if (NK > 1) {
                        acc3_mac(s1[e], zd[j][e / 2].x, xd[j][e / 2].x);
                        acc3_mac(s1[e + 1], zd[j][e / 2].y, xd[j][e / 2].y);
                    }
                }
        };
        int i = 0;
        for (; i + 1 < ndig; i += 2) digits(i, std::integral_constant<int, 2>());
        if (i < ndig) digits(i, std::integral_constant<int, 1>());
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            u64 hi, lo;
            acc3_fold(s0[e], hi, lo);
            park0[(c + e) * Qd::T + tid] = redc128(hi, lo, m.q, m.qinv);
        }
        if constexpr (NK > 1) {
            gu64 *p1 = uniform_ptr(park1);
#pragma unroll
            for (int e = 0; e < CH; e += 2) {
                u64 hi, lo;
                u64v2 w;
                acc3_fold(s1[e], hi, lo);
                w.x = redc128(hi, lo, m.q, m.qinv);
                acc3_fold(s1[e + 1], hi, lo);
                w.y = redc128(hi, lo, m.q, m.qinv);
                *(gv2 *)(p1 + Qd::elem(tid, c + e)) = w;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) a0[e] = park0[e * Qd::T + tid];
    __syncthreads();  // before the transform's first exchange writes other lanes' words
}

// Whether ext_rows_cs's chunks cover whole 32-byte segments of the rows (one
// key: four values; two keys: two values, a whole segment only where the
// transform's Quad layout holds pairs, at 2^13 per workgroup).  Two keys over
// half segments measured slower than the reduced products (MKF switching-key
// MACs at 2^14 1.25 -> 1.9 ms; hoisted rotations 1.26 -> 1.5 ms), one key or
// whole segments faster (KKLSS external products 1.51 -> 1.30 and 1.16 -> 1.04 ms).
template <int LS, int NK>
__device__ constexpr bool ext_cs_geometry() {
    return NK == 1 || Quad<LS>::R == 2;
}

// Whether ext_rows_cs may take a row of modulus q: ndig products of a canonical
// key word and a digit word -- below 68q when the digit rows are kept unreduced
// (lazy_digits_ok), else canonical -- fit the carry-save sum, which also keeps
// the sum's high word below q for redc128.
__device__ __forceinline__ bool ext_cs_ok(u64 q, int ndig);

// Whether ext_rows may accumulate lazily: ndig sums of values below 2q stay
// below 2^64, and reduce_big's quotient estimate needs q >= 2^32.
__device__ __forceinline__ bool ext_lazy_ok(u64 q, int ndig) {
    const u64 n = (u64)ndig;
    return q >= ((u64)1 << 32) && mulhi64(n, q) == 0 && n * q < ((u64)1 << 63);
}

// Whether a digit row of modulus q may be stored and used unreduced (below 68q,
// the bound of the unreduced forward transform): its consumers take it as the
// second operand of a Montgomery product with a canonical first operand
// (mred_lazy, mred, mulmod_m), which stays exact for operands below 2^63.1 and
// moduli below 2^61 -- 68q < 2^63.1 holds for every q the unreduced transform
// takes (fwd_lazy_ok: q < 2^57).
__device__ __forceinline__ bool lazy_digits_ok(u64 q) { return fwd_lazy_ok(q); }

__device__ __forceinline__ bool ext_cs_ok(u64 q, int ndig) {
    return ndig > 0 && acc3_fits((u64)ndig, q - 1, lazy_digits_ok(q) ? 68 * q : q - 1);
}

// ext_rows in FP64 (modulus below 2^50): a0[e] = sum_i MRed(k0_i[e], d_i[e]) as a
// double congruent residue (|a0| < 2.6q), computed as (sum_i k0_i d_i) 2^-64; with
// NK = 2 the k1 sum is written to park1 as canonical words.
template <int LOGN, int NK>
__device__ __forceinline__ void ext_rows_fp(double (&a0)[16], u64 *park1, const u64 *d, size_t dstride,
                                            const u64 *k0, const u64 *k1, size_t kstride, int ndig, int tid,
                                            const ModConst &m) {
    typedef Quad<LOGN> Qd;
    const FpMod fm = fp_mod(m);
    double a1[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) a0[e] = a1[e] = 0.0;
    for (int i = 0; i < ndig; ++i) {
        const u64 *dr = d + i * dstride, *k0r = k0 + i * kstride, *k1r = NK > 1 ? k1 + i * kstride : nullptr;
        // products below 0.6q: reduce every fourth digit
        const bool red = (i & 3) == 3;
        // CH values of every operand are loaded before they are used (more loads in
        // flight per wave); CH is bounded by the registers the accumulators leave
        constexpr int CH = NK > 1 ? 8 : 16;
#pragma unroll
        for (int c = 0; c < 16; c += CH) {
            u64 x[CH], y0[CH], y1[CH];
#pragma unroll
            for (int e = 0; e < CH; e += 2) {
                const int off = Qd::elem(tid, c + e);
                const u64v2 xd = *(const gv2 *)(uniform_ptr(dr) + off), yd = *(const gv2 *)(uniform_ptr(k0r) + off);
                x[e] = xd.x;
                x[e + 1] = xd.y;
                y0[e] = yd.x;
                y0[e + 1] = yd.y;
                if (NK > 1) {
                    const u64v2 zd = *(const gv2 *)(uniform_ptr(k1r) + off);
                    y1[e] = zd.x;
                    y1[e + 1] = zd.y;
                }
            }
#pragma unroll
            for (int e = 0; e < CH; ++e) {
                const double xv = fp_from_u64(x[e]);
                a0[c + e] += fp_mulmod(fp_from_u64(y0[e]), xv, fm);
                if (NK > 1) a1[c + e] += fp_mulmod(fp_from_u64(y1[e]), xv, fm);
                if (red) {
                    a0[c + e] = fp_reduce(a0[c + e], fm);
                    if (NK > 1) a1[c + e] = fp_reduce(a1[c + e], fm);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) a0[e] = fp_mulmod(fp_reduce(a0[e], fm), m.rinvd, fm);
    if constexpr (NK > 1) {
        u64 w[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) w[e] = fp_word(fp_canon(fp_mulmod(fp_reduce(a1[e], fm), m.rinvd, fm), fm));
        store_row_quad<LOGN>(w, park1, tid);
    }
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

// Word c of row `row` (modulus index mod) of a QP polynomial.  unf: the row's
// inverse transform ran as two half-limb transforms and left the last stage
// (t = N/2) and the N^-1 scaling undone (values below 4q) -- they are done here,
// from the word and its partner c ^ N/2, as k_intt_finish would: the lower word
// becomes (x + y) N^-1, the upper (x - y) psiinv[1] N^-1 (x lower, y upper).
__device__ __forceinline__ u64 row_word(const DevTables &T, const u64 *__restrict__ src, int row, int mod, size_t c,
                                        bool unf) {
    const u64 x = G(src)[(size_t)row * T.N + c];
    if (!unf) return x;
    const size_t half = (size_t)T.N >> 1;
    const u64 y = G(src)[(size_t)row * T.N + (c ^ half)];
    const ModConst &mm = T.mod[mod];
    return c < half ? shoup(x + y, mm.ninv, mm.ninvp, mm.q) : shoup(y + (mm.q << 2) - x, mm.wn, mm.wnp, mm.q);
}

__device__ __forceinline__ PUp modup_p_prepare(const DevTables &T, const u64 *__restrict__ src, int nl, size_t c,
                                               bool unf = false) {
    PUp u;
    double vi = 0.0;
#pragma unroll
    for (int k = 0; k < kMaxK; ++k) {
        u.y[k] = 0;
        if (k < T.K) {
            const ModConst mp = T.mod[T.L + k];
            u.y[k] = mred(row_word(T, src, nl + k, T.L + k, c, unf), T.md_y[k], mp.q, mp.qinv);
            vi += (double)u.y[k] / (double)mp.q;
        }
    }
    u.v = (u64)vi;
    return u;
}

// The same residue below 3q, not reduced (for sums that are reduced once).
__device__ __forceinline__ u64 modup_p_to_row_lazy(const DevTables &T, const PUp &u, int j, const ModConst &mq) {
    u64 hi = 0, lo = 0;
#pragma unroll
    for (int k = 0; k < kMaxK; ++k)
        if (k < T.K) mac128(hi, lo, u.y[k], T.md_qstar[j * T.K + k]);
    return mont_reduce128(hi, lo, mq.q, mq.qinv) + T.md_vt[j * (T.K + 1) + u.v];  // (0, 2q) + [0, q)
}

__device__ __forceinline__ u64 modup_p_to_row(const DevTables &T, const PUp &u, int j, const ModConst &mq) {
    return reduce64(modup_p_to_row_lazy(T, u, j, mq), mq.q, mq.mu);
}

}  // namespace mkhip
