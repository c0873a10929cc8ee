// Device-resident constant tables of one parameter set, and the work descriptors
// the kernels consume.  Shared by host (mkhip.hip) and device code.
#pragma once

#include "modarith.hpp"

namespace mkhip {

constexpr int kMaxParties = 32;   // components per ciphertext side handled by one descriptor
constexpr int kMaxModuli = 48;    // Q + P (+ BFV auxiliary) moduli of one context
constexpr int kMaxAlpha = 4;      // primes per gadget digit

// Pointers into one device allocation holding every table of a context.
struct DevTables {
    const ModConst *mod;      // [nmod] : Q[0..L) then P[0..K) then (BFV) QMul[0..L)
    const ulonglong2 *tw;     // [nmod][N] forward twiddles (w, Shoup w'), Lattigo index order
    const ulonglong2 *itw;    // [nmod][N] inverse twiddles
    const double *twd;        // [nmod][N] forward twiddles as doubles (exact for q < 2^50)
    const double *itwd;       // [nmod][N] inverse twiddles as doubles
    // ModDown Q u P -> Q (basis_extension.go:192-232 with paramsPtoQ[K-1]):
    const u64 *md_y;          // [K]       MForm((P/p_i)^-1 mod p_i)
    const u64 *md_qstar;      // [L][K]    MForm((P/p_i) mod q_j)
    const u64 *md_vt;         // [L][K+1]  v * (-P) mod q_j
    const u64 *md_pinv;       // [L]       MForm(-P^-1 mod q_j)
    // gadget digits with nd >= 2 primes (Decomposer, basis_extension.go:368-424):
    // index e = (digit * kMaxAlpha + nd) ; rows r over Q_all u P_all (L + K)
    const u64 *dc_y;          // [beta][kMaxAlpha+1][kMaxAlpha]        MForm((D/q_l)^-1 mod q_l)
    const u64 *dc_qstar;      // [beta][kMaxAlpha+1][L+K][kMaxAlpha]    MForm((D/q_l) mod m_r)
    const u64 *dc_vt;         // [beta][kMaxAlpha+1][L+K][kMaxAlpha+1]  v * (-D) mod m_r
    // rescale (DivRoundByLastModulus) : for dropped modulus index l and j < l
    const u64 *rs_inv;        // [L][L]  MForm(q_l^-1 mod q_j)
    const u64 *rs_half;       // [L][L]  (q_l >> 1) mod q_j
    // BFV (R = Q u QMul, mkbfv/basis_extension.go); null in CKKS contexts.
    // QMul limb j has ring index L + K + j.
    const u64 *bq_qm;         // [L]        MForm(QMul mod q_i)
    const u64 *bq_up_y;       // [L]        Q -> QMul : MForm((Q/q_i)^-1 mod q_i)
    const u64 *bq_up_qstar;   // [L][L]     [j][i] MForm((Q/q_i) mod qm_j)
    const u64 *bq_up_vt;      // [L][L+1]   v * (-Q) mod qm_j
    const u64 *bq_dn_y;       // [L]        QMul -> Q : MForm((QMul/qm_j)^-1 mod qm_j)
    const u64 *bq_dn_qstar;   // [L][L]     [i][j] MForm((QMul/qm_j) mod q_i)
    const u64 *bq_dn_vt;      // [L][L+1]   v * (-QMul) mod q_i
    const u64 *bq_rsc;        // [L]        qm_j - MForm(Q^-1 mod qm_j)   (Rescale)
    const u64 *bq_qz;         // [L]        q_i - MForm(QMul^-1 mod q_i)  (Quantize)
    int logN, N, L, K, alpha, beta_max, nmod;
    int bfv;                  // 1 when the QMul chain is present
    int xcd;                  // 1: XCD-contiguous workgroup order (xcd_remap), a tuning knob
    int decomp_order;         // fused decomposition grid order (a tuning knob)
    int fp64;                 // FP64 transforms for moduli below 2^50 (a tuning knob)
    int split14;              // N = 2^14 one-prime decompositions over half-limb workgroups (a tuning knob)
    int split14_ext;          // N = 2^14 external products over half-limb workgroups (a tuning knob)
    int digit_pre;            // MKF CRT digits: digit-only ModUp part once per digit (a tuning knob)
    int xcd_ext;              // 1: external-product kernels (k_ext, k_mac2) in XCD-contiguous
                              // order, a tuning knob separate from xcd
    int ext_cs;               // 1: external-product MACs summed in carry-save form (Acc3)
                              // where the sums fit, a tuning knob
    int inv_lazy;             // 1: unreduced inverse transforms for 2^32 <= q < 2^57 (ntt_inv)
};

// N = 2^14 external products over half-limb workgroups: a plan that emits KEXT
// stages then finishes their inverse transforms with k_intt_finish, as at 2^15.
inline bool ext_split(const DevTables &T) { return T.logN == 14 && T.split14_ext != 0; }


// ring index of limb l of an R = Q u QMul polynomial
__host__ __device__ __forceinline__ int r_index(int l, int L, int K) { return l < L ? l : K + l; }

// primes of gadget digit i of components decomposed at level src (Decomposer's
// decompLvl + 2, basis_extension.go:440-451)
__host__ __device__ __forceinline__ int digit_primes(int i, int src, int alpha) {
    return (src > alpha * (i + 1) - 1) ? alpha : (src % alpha) + 1;
}

__device__ __forceinline__ int dc_base(const DevTables &T, int digit, int nd) {
    return digit * (kMaxAlpha + 1) + nd;
}

// ---- kernels ---------------------------------------------------------------

enum KernelId {
    KNTT_FWD = 0,        // k_ntt_batch<LOGN, false>
    KNTT_INV = 1,        // k_ntt_batch<LOGN, true>
    KDECOMP_MAC = 2,     // k_decomp_mac<LOGN>     alpha = 1: decomposition fused with the x/y MAC
    KTENSOR = 3,         // k_tensor<LOGN>
    KEXT1 = 4,           // k_ext<LOGN, 1>
    KMODDOWN = 5,        // k_moddown<false>
    KDECOMP = 6,         // k_decomp<LOGN, false>  decomposition only
    KEXT2 = 7,           // k_ext<LOGN, 2>         two keys sharing one decomposition
    KMODDOWN_FINAL = 8,  // k_moddown<true>        + accumulation into the output + rescale
    KMAC_XY = 9,         // k_mac_xy               key-stationary x / y accumulation
    KMAC2 = 10,          // k_mac2<LOGN>           MKF switch: two-output key MAC (+ INTT)
    KMD_NTT = 11,        // k_moddown_ntt<LOGN>    NTT-domain ModDown, summed over a fold
    KKARA = 12,          // k_karatsuba            MKF tensor (Karatsuba, NTT domain)
    KRESCALE_NTT = 13,   // k_rescale_ntt<LOGN>    NTT-domain rescale + materialized copies
    KBFV_LIFT = 14,      // k_bfv_lift             ModUpQtoR / Rescale into R
    KBFV_TENSOR = 15,    // k_tensor<LOGN, true>   tensor over R, t x, INTT
    KBFV_QUANTIZE = 16,  // k_quantize             ModDown R -> Q (+ accumulate)
    KINTT_FINISH = 17,   // k_intt_finish          last inverse stage of split (2^15) limbs
    KAUTO = 18,          // k_automorph            Galois automorphism (coefficient or NTT domain)
    KSMALL_NTT = 19,     // k_small_ntt<LOGN>      small signed polynomial -> residues -> NTT (-> MForm)
    KKEYOP = 20,         // k_keyop                elementwise key / encryption combination
    KLINEAR = 21,        // k_linear               CKKS linear operations of the polynomial evaluator
    KDIGIT_PRE = 22,     // k_digit_pre            digit-only part of the CRT ModUp, once per digit
};

// One NTT-domain row (ring modulus `mod`) of a CKKS linear operation; the constant
// is c_lo on slots [0, N/2) and c_hi on [N/2, N) (a + b psi^(N/2), a - b psi^(N/2)).
//   LIN_COPY     out = a
//   LIN_ADD      out = a + b
//   LIN_SUB      out = a - b
//   LIN_MULC     out = MRed(a, c)            c in Montgomery form
//   LIN_MULCADD  out = b + MRed(a, c)        c in Montgomery form (out may be b)
//   LIN_ADDC     out = a + c                 c a canonical residue
//   LIN_ZERO     out = 0
enum LinOp { LIN_COPY = 0, LIN_ADD, LIN_SUB, LIN_MULC, LIN_MULCADD, LIN_ADDC, LIN_ZERO };
struct LinDesc {
    const u64 *a, *b;
    u64 *out;
    u64 c_lo, c_hi;
    int mod, op;
};

// A small signed polynomial (secret, error, the encryption's ternary v) reduced
// modulo ring modulus `mod`, transformed, optionally put in Montgomery form.
struct SmallNttDesc {
    const int64_t *src;  // [N]
    u64 *dst;            // [N]
    int mod, mform;
};

// One row of a key or ciphertext, elementwise mod q (ring modulus `mod`):
//   op 0: out = x + g pmod + y + e - MRed(ka, kb)
//   op 1: out = q - (x + g pmod + y + e + MRed(ka, kb))   (Lattigo NegLvl: a zero becomes q)
//   op 2: out = x + g pmod + y + e + MRed(ka, kb)
// Null operands are absent terms; e is a small signed polynomial.
struct KeyOpDesc {
    const u64 *x, *g, *ka, *kb, *y;
    const int64_t *e;
    u64 *out;
    u64 pmod;
    int mod, op;
};

// Galois automorphism X -> X^gal of nl limbs, src -> dst (no aliasing).
// Coefficient domain: Lattigo ring.Permute (restated inline at
// mkrlwe/keyswitch.go:283-303), negation written q - x.  NTT domain: the slot
// permutation of ring.PermuteNTTIndex.
struct AutoDesc {
    const u64 *src;
    u64 *dst;
    u64 gal, gal_inv;  // Galois element (odd, < 2N) and its inverse mod 2N
    int nl, ntt;
};

// One limb of 2^15 whose halves ran every inverse stage but the last: apply it
// in place with the N^-1 (tscale: t N^-1) scaling.
struct FinishDesc {
    u64 *row;
    int mod, tscale;
};

// ---- BFV descriptors ----------------------------------------------------------

// Lift of one coefficient-domain Q poly (L limbs) into R (2L limbs):
// mode 0 ModUpQtoR, mode 1 Rescale (mkbfv/basis_extension.go:50-66, 88-104).
struct BfvLiftDesc {
    const u64 *src;
    u64 *dst;
    int mode;
    int pad;
};

// Quantize's ModDown of a coefficient-domain R poly into Q (basis_extension.go:69-85,
// the t x and INTT already applied): dst (=|+=) (x_Q - ModUp(x_QMul)) QMul^-1.
struct BfvQuantDesc {
    const u64 *src;
    u64 *dst;
    int accumulate;
    int pad;
};

// ---- MKF (mkf_cpu) descriptors ----------------------------------------------

// One hoisted key switch without ModDown (Lattigo KeyswitchHoistedNoModDown, called
// from mkf_cpu/common.go:82-96): acc_o[r] = sum_b swk[b][o][r] (.) D[b][r]
// for o = 0, 1; the P rows (every row with intt_all) are returned to the
// coefficient domain.  swk layout [beta][2][L+K][N].
struct Mac2Desc {
    const u64 *D;
    const u64 *swk;
    u64 *acc[2];
    int level, intt_all;
};

// NTT-domain ModDown summed over nacc switches (ModDownQPtoQNTT,
// basis_extension.go:239-288, folded as in mkf_cpu/common.go:109-126):
// dst[j] = base[j] + sum_a (acc_a[j] - NTT_j(ModUp(acc_a P rows))) P^-1.
struct MdNttDesc {
    const u64 *acc[kMaxParties];
    const u64 *base;       // NTT domain, may be null
    u64 *dst;              // level+1 limbs
    int nacc, level;
};

// Karatsuba tensor (mkf_cpu/ckks.go:120-131), NTT domain, level+1 limbs.
struct KaraDesc {
    const u64 *l0, *l1, *r0, *r1;
    u64 *mu0, *mu1, *mu2;
};

// Rescale by the last modulus in the NTT domain (nres = 1) or a plain copy
// (nres = 0), written to nout destinations (the k-fold materialization of Gamma,
// mkf_cpu/common.go:128-139).  xlast: the last limb already in the coefficient domain.
struct RescaleNttDesc {
    const u64 *x;
    const u64 *xlast;
    u64 *out[kMaxParties + 1];
    int nout, level, nres;
    int pad;
};

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8).  Remapping b to a logical index that gives every XCD one contiguous
// range keeps workgroups that share data (an input limb decomposed into every
// row, a key row used by every op) inside one XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    constexpr int kXcd = 8;
    const int full = n - n % kXcd;
    if (b >= full) return b;
    return (b % kXcd) * (full / kXcd) + b / kXcd;
}

// Logical index of this workgroup: XCD-contiguous when the context's knob is on.
__device__ __forceinline__ int block_index(const DevTables &T) {
    return T.xcd ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
}

// The same for the external-product kernels, whose knob is tuned separately (their
// adjacent workgroups share the x / y or key rows of one op, the decompositions'
// share input limbs).
__device__ __forceinline__ int block_index_ext(const DevTables &T) {
    return T.xcd_ext ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
}

// ---- descriptors -----------------------------------------------------------

// One hoisted gadget decomposition side: k component polys (coefficient domain,
// level+1 limbs each, row stride N) decomposed into D[k][beta][rows][N]; when key
// pointers are given the kernel also accumulates xy[beta][rows] = MForm(sum_p key_p (.) D_p).
struct DecompDesc {
    const u64 *comp[kMaxParties];
    const u64 *key[kMaxParties];   // KKLSS d_i / b_j (MForm, beta x (L+K) x N); null => no MAC
    u64 *D;                        // [k][beta][rows][N]
    u64 *xy;                       // [beta][rows][N] (MForm), may be null
    int k, level;
    int src_level;                 // level the components were decomposed at (their
                                   // limb count - 1), above `level` for an operand
                                   // above the product's level: digit i spans limbs
                                   // i alpha .. min(i alpha + alpha - 1, src_level);
                                   // a value below `level` means `level`
    u64 *pre;                      // [k][beta][alpha+1][N]: y_l rows and v of the CRT digits
                                   // (k_digit_pre), or null: the rows compute them
    int lazy_out;                  // 1: rows of moduli with lazy_digits_ok may be stored
                                   // unreduced (below 68q; every plan-internal consumer is a
                                   // Montgomery product), 0: canonical (the public entry points)
};

// Generic batched limb NTT job.
struct NttJob {
    const u64 *src;
    u64 *dst;
    int mod;       // ring index
    int reduce;    // bit 0: src values may exceed the modulus (reduce mod q first);
                   // bit 1 (inverse only): scale by t N^-1 instead of N^-1
};

// Tensor output component: out = a0 (.) b_u + a_u (.) b0 (either term may be absent),
// or out = a0 (.) b0 for the c0 component.  Inputs NTT domain, output coefficient domain.
struct TensorDesc {
    const u64 *a0, *b0, *au, *bu;  // au / bu null when the party is absent on that side
    u64 *out;
    int is_c0;
    int ustride;  // limb l of au / bu at l * ustride * N (0 or 1: contiguous)
};

// Key-stationary x / y accumulation over a batch: for one side of every op,
// xy[beta][rows] = MForm(sum_p key_p (.) D[p]) (keyswitch_hoisted.go:84-123).
// One thread owns one coefficient of one (digit, row) and walks every op of the
// batch, so each key word is fetched once per batch (L1/L2 reuse across ops).
struct MacDesc {
    const u64 *D;                   // [k][beta][rows][N]
    u64 *xy;                        // [beta][L+K][N] (full-level key layout)
    const u64 *key[kMaxParties];    // [beta][L+K][N] per component (MForm)
    int k, level;
};

// External-product MAC over beta digits followed by an INTT of every row:
// acc_o[r] = INTT( sum_i key_o[i][r] (.) D[i][r] ) for o < nkey.
struct ExtDesc {
    const u64 *D;          // [beta][rows][N]
    const u64 *key[2];     // [beta][L+K][N] each (key row index = ring index)
    u64 *acc[2];           // [rows][N]
    int nkey, level;
};

// ModDown of nsrc accumulators (rows = level+1+K, coefficient domain) into dst
// (level+1 limbs): dst (=|+=) sum_s ModDown(src_s).  With `final`, the result
// plus `base` is rescaled nres times and written to out (level+1-nres limbs).
struct ModDownDesc {
    const u64 *src[kMaxParties];
    u64 *dst;              // level+1 limbs
    const u64 *base;       // added before writing (may be null)
    u64 *out;              // final output (when nres >= 0 and final)
    int nsrc, level, accumulate, final_, nres;
    int unfinished;        // 1: the sources' inverse transforms ran on half-limb workgroups
                           // and left their last stage and N^-1 scaling to this kernel
    int pad[2];
};

}  // namespace mkhip
