// 64-bit modular arithmetic for moduli q < 2^61 on CDNA4 (host + device).
//
// Conventions follow the reference's ring arithmetic (Lattigo v2.3.0 as used by
// software/kklss/mkrlwe and software/mkf/mkf_cpu):
//   mred(x, y)   = x * y * 2^-64 mod q        (Lattigo ring.MRed; keys are stored
//                                               "in Montgomery form", x * 2^64 mod q)
//   shoup(a, w)  = a * w mod q, lazily in [0, 2q), with the precomputed companion
//                  w' = floor(w * 2^64 / q); used for NTT twiddles and constants.
// All functions are exact integer arithmetic.  The NTT kernels keep values in
// [0, 8q) (Harvey's lazy butterflies with an approximate Shoup quotient), which
// needs 8q <= 2^64, i.e. q < 2^61.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mkhip {

typedef uint64_t u64;
typedef uint32_t u32;

#define MKHIP_HD __host__ __device__ __forceinline__

MKHIP_HD u64 mulhi64(u64 a, u64 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (u64)(((unsigned __int128)a * b) >> 64);
#endif
}

// Full 128-bit product a * b.  On the device: four 32 x 32 -> 64 multiply-adds
// (computing the low and high halves separately would take seven).
MKHIP_HD void mul_wide(u64 a, u64 b, u64 &hi, u64 &lo) {
#if defined(__HIP_DEVICE_COMPILE__)
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 p00 = (u64)a0 * b0;
    const u64 t = (u64)a0 * b1 + (p00 >> 32);
    const u64 u = (u64)a1 * b0 + (u32)t;
    lo = (u << 32) | (u32)p00;
    hi = (u64)a1 * b1 + (t >> 32) + (u >> 32);
#else
    const unsigned __int128 p = (unsigned __int128)a * b;
    lo = (u64)p;
    hi = (u64)(p >> 64);
#endif
}

// a * b * 2^-64 mod q, canonical, for a * b < q * 2^64 (Lattigo MRed).
MKHIP_HD u64 mred(u64 a, u64 b, u64 q, u64 qinv) {
    u64 hi, lo;
    mul_wide(a, b, hi, lo);
    u64 m = lo * qinv;
    u64 mqhi = mulhi64(m, q);
    u64 r = hi - mqhi + q;
    return r >= q ? r - q : r;
}

// a * w mod q in [0, 2q) for any a < 2^64 and w < q, w' = floor(w 2^64 / q).
MKHIP_HD u64 shoup_lazy(u64 a, u64 w, u64 wp, u64 q) {
    u64 qh = mulhi64(a, wp);
    return a * w - qh * q;
}

// hi64(a * b) - e with e in {0, 1, 2}: the a_lo * b_lo partial product and the
// carries out of the low 32 bits of the cross products are dropped, which leaves
// one 64-bit and two 32-bit multiplies (instead of four 64-bit partials).
__device__ __forceinline__ u64 mulhi64_approx(u64 a, u64 b) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 cross = (u64)__umulhi(a0, b1) + (u64)__umulhi(a1, b0);
    return (u64)a1 * b1 + cross;
}

// a * w mod q in [0, 4q) for any a < 2^64 and w < q, w' = floor(w 2^64 / q):
// the exact Shoup quotient gives [0, 2q); the approximate one is at most 2 lower.
// The result a w - qh q is formed mod 2^64 as a w + qh (2^64 - q): the two low
// partial products chain through one multiply-add each (v_mad_u64_u32 takes the
// other as its addend) and the four cross products land in the high word, so no
// 64-bit subtraction (and its borrow chain) is needed.
__device__ __forceinline__ u64 shoup_lazy4(u64 a, u64 w, u64 wp, u64 q) {
    const u64 qh = mulhi64_approx(a, wp);
    const u64 nq = 0 - q;
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), w0 = (u32)w, w1 = (u32)(w >> 32);
    const u32 h0 = (u32)qh, h1 = (u32)(qh >> 32), n0 = (u32)nq, n1 = (u32)(nq >> 32);
    const u64 lo = (u64)h0 * n0 + (u64)a0 * w0;
    // the cross products stay 32-bit multiplies summed by two three-input adds
    // (opaque to the combiner, which would otherwise widen them into 64-bit
    // multiply-adds and a 64-bit add)
    u32 c0 = a0 * w1, c1 = a1 * w0, c2 = h0 * n1, c3 = h1 * n0;
    asm("" : "+v"(c0), "+v"(c1), "+v"(c2), "+v"(c3));
    const u32 hi = (u32)(lo >> 32) + c0 + c1 + c2 + c3;
    return ((u64)hi << 32) | (u32)lo;
}

// u + shoup_lazy4(a, w, wp, q) mod 2^64, with u taken as the addend of the first
// multiply-add of the product's low word (so that a forward butterfly's sum
// x + w y costs no addition of its own).
__device__ __forceinline__ u64 shoup_lazy4_add(u64 a, u64 w, u64 wp, u64 q, u64 u) {
    const u64 qh = mulhi64_approx(a, wp);
    const u64 nq = 0 - q;
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), w0 = (u32)w, w1 = (u32)(w >> 32);
    const u32 h0 = (u32)qh, h1 = (u32)(qh >> 32), n0 = (u32)nq, n1 = (u32)(nq >> 32);
    const u64 lo = (u64)h0 * n0 + ((u64)a0 * w0 + u);
    u32 c0 = a0 * w1, c1 = a1 * w0, c2 = h0 * n1, c3 = h1 * n0;
    asm("" : "+v"(c0), "+v"(c1), "+v"(c2), "+v"(c3));
    const u32 hi = (u32)(lo >> 32) + c0 + c1 + c2 + c3;
    return ((u64)hi << 32) | (u32)lo;
}

MKHIP_HD u64 shoup(u64 a, u64 w, u64 wp, u64 q) {
    u64 r = shoup_lazy(a, w, wp, q);
    return r >= q ? r - q : r;
}

// x mod q for any x < 2^64, mu = floor(2^64 / q) (Barrett with one correction).
MKHIP_HD u64 reduce64(u64 x, u64 q, u64 mu) {
    u64 r = x - mulhi64(x, mu) * q;  // r < 2q + q for q < 2^63
    r = r >= q ? r - q : r;
    return r >= q ? r - q : r;
}

// x mod q for any x < 2^64 and 2^32 <= q (so that mu = floor(2^64 / q) < 2^32):
// the quotient estimate hi32(x_hi * mu) is at most 2 below floor(x / q), so
// x - est q < 3q and two conditional subtractions finish.  One 32-bit multiply
// for the estimate and one multiply-add (x + est (2^64 - q)) for the remainder.
__device__ __forceinline__ u64 reduce_big(u64 x, u64 q, u64 mu) {
    const u32 est = __umulhi((u32)(x >> 32), (u32)mu);
    const u64 nq = 0 - q;
    u64 r = (u64)est * (u32)nq + x;
    r += (u64)(est * (u32)(nq >> 32)) << 32;
    const u64 q2 = q << 1;
    r = r >= q2 ? r - q2 : r;
    return r >= q ? r - q : r;
}

// (u64)a * b + c: one v_mad_u64_u32.
__device__ __forceinline__ u64 mad32(u32 a, u32 b, u64 c) { return (u64)a * b + c; }

// a * b * 2^-64 mod q in (0, 2q) for a, b < q < 2^61 (MRed without its final
// conditional subtraction: hi(ab) - hi(m q) + q with m = lo(ab) q^-1 mod 2^64).
// Both 128-bit products are formed from 32-bit multiply-adds whose addends are
// the running partial sums, so that no 64-bit addend has to be assembled from a
// 32-bit half and a zero: with a1, b1, q1 < 2^29 no partial sum exceeds 2^64
// (the middle sum of m q, which can, is split at bit 32).
__device__ __forceinline__ u64 mred_lazy(u64 a, u64 b, u64 q, u64 qinv) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 p00 = (u64)a0 * b0;
    const u64 mid = mad32(a0, b1, mad32(a1, b0, p00 >> 32));  // a0 b1 + a1 b0 + hi32(a0 b0)
    const u64 hi = mad32(a1, b1, mid >> 32);                  // hi64(a b)
    const u32 l0 = (u32)p00, l1 = (u32)mid;                   // lo64(a b)
    const u32 q0 = (u32)q, q1 = (u32)(q >> 32), i0 = (u32)qinv, i1 = (u32)(qinv >> 32);
    const u64 mm = (u64)l0 * i0;
    const u32 m0 = (u32)mm, m1 = (u32)(mm >> 32) + l0 * i1 + l1 * i0;
    const u64 s = mad32(m1, q0, ((u64)m0 * q0) >> 32);        // m1 q0 + hi32(m0 q0) < 2^64
    const u64 t = mad32(m0, q1, (u32)s);                      // m0 q1 + lo32(s) < 2^62
    const u64 mqh = mad32(m1, q1, s >> 32) + (t >> 32);       // hi64(m q)
    return hi + q - mqh;
}

// Carry-save sum of 128-bit products.  With a = a1 2^32 + a0 and b = b1 2^32 + b0,
// each product adds lo32(a0 b0) to L, hi32(a0 b0) + a0 b1 + a1 b0 to M and a1 b1
// to H, so that the sum of the products is L + 2^32 M + 2^64 H with no carry
// passed between the words: one multiply and three multiply-adds per product
// (a Montgomery reduction per product costs about three times that).  Exact
// while M + hi32(L) stays below 2^64 (acc3_fits).
struct Acc3 {
    u64 L, M, H;
};

__device__ __forceinline__ void acc3_zero(Acc3 &A) { A.L = A.M = A.H = 0; }

__device__ __forceinline__ void acc3_mac(Acc3 &A, u64 a, u64 b) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 p00 = (u64)a0 * b0;
    A.L += (u32)p00;
    A.M = mad32(a0, b1, mad32(a1, b0, A.M + (p00 >> 32)));
    A.H = mad32(a1, b1, A.H);
}

// Whether n products of a <= amax and b <= bmax fit: each adds below
// 2^32 (1 + a1 + b1) to M and below 2^32 to L.
MKHIP_HD bool acc3_fits(u64 n, u64 amax, u64 bmax) {
    const unsigned __int128 per = ((unsigned __int128)((amax >> 32) + (bmax >> 32) + 1) << 32) + 1;
    return (unsigned __int128)n * per < ((unsigned __int128)1 << 64);
}

// The accumulated sum as (hi, lo), exact.
__device__ __forceinline__ void acc3_fold(const Acc3 &A, u64 &hi, u64 &lo) {
    const u64 m = A.M + (A.L >> 32);
    lo = (m << 32) | (u32)A.L;
    hi = A.H + (m >> 32);
}

// (hi 2^64 + lo) 2^-64 mod q, canonical, for hi < q (Lattigo's MRed on a 128-bit
// value: lo - lo(m q) = 0 with m = lo q^-1, so the quotient is hi - hi(m q)).
__device__ __forceinline__ u64 redc128(u64 hi, u64 lo, u64 q, u64 qinv) {
    const u64 r = hi + q - mulhi64(lo * qinv, q);
    return r >= q ? r - q : r;
}

MKHIP_HD u64 add_mod(u64 a, u64 b, u64 q) {
    u64 s = a + b;
    return s >= q ? s - q : s;
}

MKHIP_HD u64 sub_mod(u64 a, u64 b, u64 q) { return a >= b ? a - b : a + q - b; }

// Per-modulus constants resident in device memory.
struct ModConst {
    u64 q;      // modulus
    u64 qinv;   // q^-1 mod 2^64 (Montgomery)
    u64 mu;     // floor(2^64 / q) (Barrett)
    u64 r2;     // 2^128 mod q: mred(x, r2) = x * 2^64 mod q (MForm)
    u64 r1, r1p;  // 2^64 mod q and its Shoup companion: shoup(x, r1, r1p) = MForm(x) for any x < 2^64
    u64 ninv;   // N^-1 mod q (canonical) for this modulus at the context's N
    u64 ninvp;  // Shoup companion of ninv
    u64 tninv;  // t * N^-1 mod q (BFV: Quantize's MulScalar(t) folded into the INTT)
    u64 tninvp; // Shoup companion of tninv
    u64 wn, wnp;    // psiinv[1] * N^-1 (last inverse stage of a split limb) + companion
    u64 twn, twnp;  // t * psiinv[1] * N^-1 + companion
    double qd, qinvd;  // q and fl(1 / q) for the FP64 transforms (q < 2^50)
    double rinvd;      // 2^-64 mod q (exact as a double for q < 2^53): MRed's factor
};

// host helpers (exact, 128-bit); host-only functions, parsed in both compilation passes
inline u64 h_mulmod(u64 a, u64 b, u64 q) { return (u64)(((unsigned __int128)a * b) % q); }
inline u64 h_powmod(u64 x, u64 e, u64 q) {
    u64 r = 1 % q, b = x % q;
    while (e) {
        if (e & 1) r = h_mulmod(r, b, q);
        b = h_mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}
inline u64 h_inv(u64 a, u64 q) { return h_powmod(a % q, q - 2, q); }
inline u64 h_mform(u64 a, u64 q) { return (u64)((((unsigned __int128)(a % q)) << 64) % q); }
inline u64 h_shoup_companion(u64 w, u64 q) { return (u64)((((unsigned __int128)w) << 64) / q); }
inline u64 h_qinv(u64 q) {
    u64 x = q;  // Newton iteration for q^-1 mod 2^64
    for (int i = 0; i < 6; ++i) x *= 2 - q * x;
    return x;
}


}  // namespace mkhip
