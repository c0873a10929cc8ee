/*
 * mkref_poly.c -- CPU oracle, part 5: the CKKS linear operations the MKF
 * polynomial evaluator applies between its products
 * (software/mkf/mkf_cpu/true_mkf_ckks_polynomial.go, true_mkf_ckks.go:187-251).
 *
 * TEST INFRASTRUCTURE ONLY (see mkref.h).  MKF ciphertexts are NTT-domain
 * [c0, c_1..c_k] with (level+1) limbs per component.  The MKF evaluator hands
 * them to Lattigo v2.3.0's ckks.Evaluator as degree-k ciphertexts; the
 * algorithms restated here are that evaluator's, whose text the reference
 * carries in its KKLSS fork (software/kklss/mkckks/evaluator.go, utils.go) and
 * in its constant-layout tests (mkf_cpu/p2gn16_same_run_physical_operand_
 * observer_overlay_test.go:193-206).
 *
 * A constant c = a + b i is added or multiplied in the NTT domain as the vector
 * [a + b psi^(N/2)] on the first N/2 slots and [a - b psi^(N/2)] on the last N/2
 * (psi^(N/2) = NttPsi[1], a square root of -1): outside the NTT domain that is
 * a + b X^(N/2), and X^(N/2) maps to i in every CKKS slot.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mkref.h"

typedef unsigned __int128 u128;

static u64 cred(u64 a, u64 q) { return a >= q ? a - q : a; }

/* 2^e mod q */
static u64 pow2_mod(int e, u64 q) {
    u64 r = 1 % q, b = 2 % q;
    for (; e > 0; e >>= 1, b = (u64)(((u128)b * b) % q))
        if (e & 1) r = (u64)(((u128)r * b) % q);
    return r;
}

/* x = m 2^e with m an integer below 2^53 (x finite, x != 0) */
static void split_double(double x, u64 *m, int *e) {
    int E;
    double f = frexp(fabs(x), &E); /* |x| = f 2^E, f in [0.5, 1) */
    *m = (u64)ldexp(f, 53);
    *e = E - 53;
}

/* floor(y) mod q for a finite y >= 0 given as a double */
static u64 floor_double_mod(double y, u64 q) {
    if (y < 18446744073709551616.0) return (u64)y % q;
    u64 m;
    int e;
    split_double(y, &m, &e); /* e > 11: y is an integer */
    return (u64)(((u128)(m % q) * pow2_mod(e, q)) % q);
}

/*
 * scaleUpExact (mkckks/utils.go:59-84): |value| n in float64, plus 0.5 in a
 * 53-bit big.Float (the float64 sum), truncated to an integer, reduced mod q;
 * for a negative value the result is q - r (q itself when r = 0, as written).
 */
u64 mkref_scale_up_exact(double value, double n, u64 q) {
    int neg = value < 0;
    double x = neg ? -n * value : n * value;
    double y = x + 0.5;
    u64 r = floor_double_mod(y, q);
    return neg ? q - r : r;
}

/*
 * Residue of trunc(a b) mod q, the product taken exactly (a 128-bit big.Float
 * holds the product of two float64 mantissas exactly), truncated toward zero
 * (big.Float.Int) and reduced with big.Int.Mod (Euclidean, in [0, q)):
 * true_mkf_ckks_polynomial.go:258-264 and the interfaceMod of *big.Int
 * (mkckks/utils.go:31-56).
 */
u64 mkref_trunc_product_mod(double a, double b, u64 q) {
    if (a == 0 || b == 0) return 0;
    u64 ma, mb;
    int ea, eb;
    split_double(a, &ma, &ea);
    split_double(b, &mb, &eb);
    u128 p = (u128)ma * mb; /* below 2^106 */
    int e = ea + eb;
    u64 r;
    if (e >= 0) {
        r = (u64)((((p % q) * pow2_mod(e, q))) % q);
    } else {
        u128 t = -e >= 128 ? 0 : p >> -e;
        r = (u64)(t % q);
    }
    int neg = (a < 0) != (b < 0);
    return neg && r ? q - r : r;
}

/* Euclidean residue of a signed integer (big.Int.Mod) */
u64 mkref_i64_mod(int64_t x, u64 q) {
    if (x >= 0) return (u64)x % q;
    u64 r = (u64)(-(x + 1)) % q; /* |x| - 1, no overflow at INT64_MIN */
    r = (r + 1) % q;
    return r ? q - r : 0;
}

/*
 * Constants of AddConst / MultByConst (mkckks/evaluator.go:118-160): per limb
 * 0..level, the first-half constant re + MRed(im, NttPsi[1]) and the second-half
 * constant re + (q - MRed(im, NttPsi[1])), re and im scaled by scaleUpExact with
 * `scale` (the ciphertext scale for AddConst; 1 or Q[level] for MultByConst,
 * getConstAndScale :41-95).  Returned reduced into [0, q): the two places the
 * reference uses them (CRed(x + c), MForm(c)) treat q as 0.
 */
void mkref_const_halves(const mkref_ctx *c, int level, double re, double im, double scale, u64 *pos, u64 *neg) {
    const mkref_ring *R = &c->ringQP;
    for (int i = 0; i <= level; i++) {
        u64 q = R->q[i], qinv = R->qinv[i], psi1 = R->psi[(size_t)i * R->N + 1];
        u64 sr = 0, si = 0, sc = 0;
        if (re != 0) {
            sr = mkref_scale_up_exact(re, scale, q);
            sc = sr;
        }
        if (im != 0) {
            si = mkref_mred(mkref_scale_up_exact(im, scale, q), psi1, q, qinv);
            sc = cred(sc + si, q);
        }
        pos[i] = sc % q;
        neg[i] = (im != 0 ? cred(sr + (q - si), q) : sc) % q;
    }
}

/*
 * Constants of MultByGaussianIntegerAndAdd from the residues of its real and
 * imaginary integers (Lattigo ckks.Evaluator; the reference restates the layout
 * in p2gn16_same_run_physical_operand_observer_overlay_test.go:193-206).
 */
void mkref_gauss_halves(const mkref_ctx *c, int level, const u64 *re_res, const u64 *im_res, u64 *pos, u64 *neg) {
    const mkref_ring *R = &c->ringQP;
    for (int i = 0; i <= level; i++) {
        u64 q = R->q[i], qinv = R->qinv[i], psi1 = R->psi[(size_t)i * R->N + 1];
        u64 it = mkref_mred(im_res[i], psi1, q, qinv);
        pos[i] = cred(re_res[i] + it, q);
        neg[i] = cred(re_res[i] + (q - it), q) % q;
    }
}

/*
 * out = in * C, or out += in * C (MultByGaussianIntegerAndAdd), over limbs
 * 0..level of ncomp NTT-domain components: z = MRed(x, MForm(C)) resp.
 * CRed(z + MRed(x, MForm(C))), C = pos on slots [0, N/2), neg on [N/2, N)
 * (mkckks/evaluator.go:162-197).  in / out hold in_limbs / out_limbs limbs per
 * component.
 */
void mkref_mul_const_ntt(const mkref_ctx *c, int level, int ncomp, const u64 *in, int in_limbs, const u64 *pos,
                         const u64 *neg, u64 *out, int out_limbs, int accumulate) {
    const mkref_ring *R = &c->ringQP;
    int N = c->N;
    for (int u = 0; u < ncomp; u++)
        for (int i = 0; i <= level; i++) {
            u64 q = R->q[i], qinv = R->qinv[i];
            u64 cp = mkref_mform(pos[i], q), cn = mkref_mform(neg[i], q);
            const u64 *x = in + ((size_t)u * in_limbs + i) * N;
            u64 *z = out + ((size_t)u * out_limbs + i) * N;
            for (int j = 0; j < N; j++) {
                u64 t = mkref_mred(x[j], j < N / 2 ? cp : cn, q, qinv);
                z[j] = accumulate ? cred(z[j] + t, q) : t;
            }
        }
}

/* AddConst on c0 (limbs 0..level, in place): CRed(x + C), C as above */
void mkref_add_const_ntt(const mkref_ctx *c, int level, u64 *c0, const u64 *pos, const u64 *neg) {
    const mkref_ring *R = &c->ringQP;
    int N = c->N;
    for (int i = 0; i <= level; i++) {
        u64 q = R->q[i];
        u64 *x = c0 + (size_t)i * N;
        for (int j = 0; j < N; j++) x[j] = cred(x[j] + (j < N / 2 ? pos[i] : neg[i]), q);
    }
}
