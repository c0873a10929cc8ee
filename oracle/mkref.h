/*
 * mkref.h -- CPU oracle for multi-key mul+relin (MK-CKKS / MK-BFV).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (KKLSS mkrlwe/mkckks/mkbfv Go packages, the MKF mkf_cpu Go code,
 * and the Lattigo v2.3.0 ring primitives those call).  It exists so that the
 * tests, the smoke check and the CPU-baseline leg of bench.py can check the HIP
 * library against it.  Nothing in the shipped library links, imports or calls
 * this code.
 *
 * Conventions (all restated from the reference, see the .c file for file:line):
 *   - polynomials are limb-major: p[limb * N + coeff], uint64 little endian;
 *   - every value handed in or out is a canonical residue in [0, q_limb);
 *   - keys and CRS are in NTT + Montgomery form (x * 2^64 mod q);
 *   - KKLSS ciphertexts are coefficient domain; MKF-CKKS ciphertexts are NTT.
 */
#ifndef MKREF_H
#define MKREF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint64_t u64;

/* ---- scalar modular arithmetic (Lattigo ring conventions) ---- */
u64 mkref_mred_params(u64 q);                 /* q^-1 mod 2^64                         */
u64 mkref_mred(u64 x, u64 y, u64 q, u64 qinv); /* x*y*2^-64 mod q, canonical            */
u64 mkref_mred_lazy(u64 x, u64 y, u64 q, u64 qinv); /* same, in [0, 2q)                */
u64 mkref_mform(u64 a, u64 q);                 /* a*2^64 mod q                          */
u64 mkref_modexp(u64 x, u64 e, u64 q);
u64 mkref_primitive_root(u64 q, int rule);     /* rule 0: smallest g >= 3; 1: g >= 2    */
/* CT butterfly of the forward NTT: (u + MRed(v,w), u - MRed(v,w)) mod q, canonical */
void mkref_butterfly(u64 u, u64 v, u64 w, u64 q, u64 qinv, u64 *sum, u64 *diff);

/* ---- NTT-friendly prime generation: primes q = 1 mod 2N near 2^bits ---- */
int mkref_gen_primes(int bits, int logN, int count, int downward_from_top, u64 *out,
                     const u64 *exclude, int n_exclude);

/* ---- one modulus chain ("ring") ---- */
typedef struct {
    int logN, N, n;   /* degree and number of moduli */
    u64 *q, *qinv;    /* moduli and Montgomery params */
    u64 *psi;         /* n*N, Lattigo NttPsi: psi^j at bitrev(j), Montgomery form */
    u64 *psiinv;      /* n*N, Lattigo NttPsiInv */
    u64 *ninv;        /* MForm(N^-1 mod q) */
    u64 *root;        /* primitive root used per modulus */
} mkref_ring;

int  mkref_ring_init(mkref_ring *r, int logN, const u64 *q, int n, int root_rule);
void mkref_ring_free(mkref_ring *r);

/* NTT / INTT of one limb in place (canonical in, canonical out) */
void mkref_ntt_limb(const mkref_ring *r, int limb, u64 *a);
void mkref_intt_limb(const mkref_ring *r, int limb, u64 *a);
void mkref_ntt_limb_tables(int logN, u64 q, u64 qinv, const u64 *psi, u64 *a);
void mkref_intt_limb_tables(int logN, u64 q, u64 qinv, const u64 *psiinv, u64 ninv, u64 *a);

/* ---- FastBasisExtender (mkrlwe/basis_extension.go) ---- */
typedef struct {
    int nin, nout;
    u64 *qoverqiinvqi;   /* [nin]        (Q/q_i)^-1 in Montgomery form, mod q_i   */
    u64 *qoverqimodp;    /* [nout][nin]  Q/q_i mod p_j in Montgomery form           */
    u64 *vtimesqmodp;    /* [nout][nin+1] v*Q mod p_j (canonical)                  */
    u64 *qin, *qinv_in;  /* [nin]  input moduli and their MRed params               */
    u64 *pinv_out;       /* [nout] MRed params of the output moduli                  */
} mkref_modup;

int  mkref_modup_params(const u64 *Q, int nq, const u64 *P, int np, mkref_modup *m);
void mkref_modup_free(mkref_modup *m);
/* exact basis extension of N coefficients from basis Q (in[nq][N]) to P (out[np][N]) */
void mkref_modup_exact(int N, const u64 *in, const u64 *Q, int nq,
                       u64 *out, const u64 *P, int np, const mkref_modup *m);
/* same, keeping multSum's unreduced outputs (congruent, possibly >= p_j) */
void mkref_modup_exact_raw(int N, const u64 *in, const u64 *Q, int nq,
                           u64 *out, const u64 *P, int np, const mkref_modup *m);
/* ModDown: out_Q = round-ish((x_QP - ModUp(x_P)) / P) mod Q, coefficient domain */
void mkref_moddown(int N, const u64 *xq, const u64 *Q, int nq,
                   const u64 *xp, const u64 *P, int np, u64 *outq);

/* ---- generic RNS context used by the scheme restatements ---- */
typedef struct {
    int logN, N;
    int L, K;              /* number of Q and P moduli at the top level           */
    int gamma;             /* KKLSS gamma (2), MKF/Lattigo (1)                    */
    mkref_ring ringQP;     /* moduli Q[0..L) then P[0..K)                          */
} mkref_ctx;

int  mkref_ctx_init(mkref_ctx *c, int logN, const u64 *Q, int L, const u64 *P, int K,
                    int gamma, int root_rule);
void mkref_ctx_free(mkref_ctx *c);
int  mkref_alpha(const mkref_ctx *c);            /* K / gamma                  */
int  mkref_beta(const mkref_ctx *c, int levelQ); /* ceil((levelQ+1)/alpha)     */

/*
 * Gadget digit i of a coefficient-domain poly a (levelQ+1 limbs), basis-extended
 * to Q[0..levelQ] u P and NTT'd row by row.  out is (levelQ+1+K) x N.
 * Restates mkrlwe Decomposer.DecomposeAndSplit + KeySwitcher.DecomposeSingleNTT.
 */
void mkref_decompose_digit(const mkref_ctx *c, int levelQ, int digit, const u64 *a, u64 *out);

/* KKLSS Decompose: all beta digits, out is beta x (levelQ+1+K) x N */
void mkref_decompose(const mkref_ctx *c, int levelQ, const u64 *a, u64 *out);
/* the digits of a poly at src_level >= levelQ (a: src_level+1 limbs; digit i =
 * limbs i alpha .. min(i alpha + alpha - 1, src_level)), the first Beta(levelQ)
 * of them, with the rows of level levelQ (Q[0..levelQ] u P) */
void mkref_decompose_digit_src(const mkref_ctx *c, int levelQ, int src_level, int digit, const u64 *a, u64 *out);
void mkref_decompose_src(const mkref_ctx *c, int levelQ, int src_level, const u64 *a, u64 *out);

/* KKLSS ExternalProductHoisted: out (levelQ+1 limbs, coefficient) =
 * ModDown(INTT(sum_beta a_hoisted[b] (.) key[b])).  key rows for a given digit are
 * laid out with stride (L+K) limbs (the key's full level), a_hoisted with (levelQ+1+K). */
void mkref_external_product_hoisted(const mkref_ctx *c, int levelQ, const u64 *a_hoisted,
                                    const u64 *key, u64 *out);
/* the same over an explicit digit count */
void mkref_external_product_n(const mkref_ctx *c, int levelQ, int ndig, const u64 *a_hoisted,
                              const u64 *key, u64 *out);
/* x += MRed(key, D) over ndig digits; x is a full-level switching key [ndig][L+K][N] */
void mkref_mac_keyed_n(const mkref_ctx *c, int levelQ, int ndig, const u64 *key, const u64 *D, u64 *x);
/* MForm of the rows 0..levelQ and P of a full-level switching key */
void mkref_mform_qp(const mkref_ctx *c, int levelQ, u64 *x, int ndigits);

/*
 * KKLSS MK-CKKS MulRelinNew(op0, op1, rlkSet) (mkckks/evaluator.go:419) without the
 * rescale.  op0 = [c0, c_ids0...] each (level+1) limbs; op1 likewise.  ids are
 * party indices into keys (b,d,v per party, each beta x (L+K) x N) .  The output
 * holds [c0, c_union...] with the union of ids in ascending order; returns the
 * number of output components.  Output is (level+1) limbs, coefficient domain.
 */
int mkref_kklss_mulrelin(const mkref_ctx *c, int level,
                         const u64 *op0, const int *ids0, int k0,
                         const u64 *op1, const int *ids1, int k1,
                         const u64 *const *rlk_b, const u64 *const *rlk_d,
                         const u64 *const *rlk_v, const u64 *crs_u,
                         u64 *out, int *ids_out);
/* the same with operands at their own levels lv0, lv1 >= level (components of
 * lv + 1 limbs), each decomposed at its own level (evaluator.go:433-439), the
 * product taken at `level` */
int mkref_kklss_mulrelin_lv(const mkref_ctx *c, int level, int lv0, int lv1,
                            const u64 *op0, const int *ids0, int k0,
                            const u64 *op1, const int *ids1, int k1,
                            const u64 *const *rlk_b, const u64 *const *rlk_d,
                            const u64 *const *rlk_v, const u64 *crs_u,
                            u64 *out, int *ids_out);

/* Rescale as KKLSS Evaluator.Rescale: number of divisions chosen by the float64
 * scale loop; returns nbRescales and writes the new scale. */
int  mkref_rescale_count(const u64 *Q, int level, double scale, double min_scale, double *new_scale);
/* DivRoundByLastModulus repeated n times on a coefficient-domain poly (level+1 limbs) */
void mkref_div_round_last(const mkref_ctx *c, int level, int n, u64 *poly);

/* ---- MKF (mkf_cpu) restatement ---- */
/* Lattigo rlwe.KeySwitcher switch of one NTT- or coefficient-domain poly with a
 * switching key (beta x 2 x (L+K) x N, layout [digit][output][limb][coeff]) */
void mkref_mkf_switch_one(const mkref_ctx *c, int levelQ, const u64 *in, int in_ntt,
                          const u64 *swk, int ntt_output, u64 *out0, u64 *out1);
/* MKF-CKKS CKKSCase.Run: fold both sides, Karatsuba, relin, (no rescale here) */
void mkref_mkf_ckks_mulrelin(const mkref_ctx *c, int level, int k,
                             const u64 *left, const u64 *right,  /* [c0, c_1..c_k] NTT */
                             const u64 *const *fold_keys, const u64 *relin_key,
                             u64 *out0, u64 *out1);
void mkref_div_round_last_ntt(const mkref_ctx *c, int level, int n, u64 *poly);
/* the same with each operand folded at its own level, tensor at the smaller one */
void mkref_mkf_ckks_mulrelin2(const mkref_ctx *c, int level_l, int level_r, int k,
                              const u64 *left, const u64 *right,
                              const u64 *const *fold_keys, const u64 *relin_key,
                              u64 *out0, u64 *out1);

/* ---- CKKS linear operations of the MKF polynomial evaluator, in mkref_poly.c ---- */
u64  mkref_scale_up_exact(double value, double n, u64 q);  /* mkckks/utils.go:59-84          */
u64  mkref_trunc_product_mod(double a, double b, u64 q);   /* trunc(a*b) exact, mod q        */
u64  mkref_i64_mod(int64_t x, u64 q);                       /* Euclidean residue              */
/* AddConst / MultByConst constants of the two NTT halves, limbs 0..level */
void mkref_const_halves(const mkref_ctx *c, int level, double re, double im, double scale, u64 *pos, u64 *neg);
/* MultByGaussianIntegerAndAdd constants from the residues of the real / imaginary integers */
void mkref_gauss_halves(const mkref_ctx *c, int level, const u64 *re_res, const u64 *im_res, u64 *pos, u64 *neg);
void mkref_mul_const_ntt(const mkref_ctx *c, int level, int ncomp, const u64 *in, int in_limbs, const u64 *pos,
                         const u64 *neg, u64 *out, int out_limbs, int accumulate);
void mkref_add_const_ntt(const mkref_ctx *c, int level, u64 *c0, const u64 *pos, const u64 *neg);

/* ---- Galois key switches (mkrlwe Rotate/Conjugate, MKF RotateNew), in mkref_rot.c ---- */
u64  mkref_galois_element(int logN, long long k);   /* 5^(k mod 2N) mod 2N          */
u64  mkref_galois_conjugate(int logN);             /* 2N - 1                        */
u64  mkref_galois_inverse(int logN, u64 g);        /* g^(2N-1) mod 2N               */
/* coefficient-domain automorphism of limbs 0..level (negation written q - x) */
void mkref_permute(const mkref_ctx *c, int level, const u64 *in, u64 gal, u64 *out);
/* NTT-slot permutation: NTT(Permute(a))[i] = NTT(a)[index[i]] */
void mkref_permute_ntt_index(int logN, u64 gal, u64 *index);
/* InvNTT, Permute, NTT of limbs 0..level */
void mkref_automorphism_ntt(const mkref_ctx *c, int level, const u64 *in, u64 gal, u64 *out);
/* ExternalProduct as the reference writes it (one decomposition buffer reused) */
void mkref_external_product_as_written(const mkref_ctx *c, int levelQ, const u64 *a, const u64 *key, u64 *out);
/* KKLSS RotateHoisted / Rotate / Conjugate: in and out [1+k][level+1][N] coefficient domain */
void mkref_kklss_rotate_hoisted(const mkref_ctx *c, int level, int k, const u64 *in, const u64 *D,
                                const u64 *const *rk, const u64 *a, u64 gal, u64 *out);
void mkref_kklss_rotate(const mkref_ctx *c, int level, int k, const u64 *in, const u64 *const *rk, const u64 *a,
                        u64 gal, u64 *out);
void mkref_kklss_conjugate(const mkref_ctx *c, int level, int k, const u64 *in, const u64 *const *ck,
                           const u64 *a, u64 *out);
/* MKF RotateNew: in [1+k][level+1][N] NTT; outputs sigma(C0) + sum u0_i and sum u1_i (NTT) */
void mkref_mkf_ckks_rotate(const mkref_ctx *c, int level, int k, const u64 *in, const u64 *const *keys, u64 gal,
                           u64 *out0, u64 *out1);

/* ---- key generation and encryption (mkrlwe keygen.go / encryptor.go), in mkref_keygen.c ----
 * Small polynomials (secrets, errors, the encryption's v) are int64 [N] inputs; keys
 * and secrets are full-level QP polys [L+K][N] (switching keys [beta][L+K][N]). */
void mkref_gen_secret(const mkref_ctx *c, const int64_t *s, u64 *sk);
void mkref_gen_public_key(const mkref_ctx *c, const u64 *sk, const int64_t *e, const u64 *a, u64 *pk0, u64 *pk1);
void mkref_gen_switching_key(const mkref_ctx *c, const u64 *skin, const int64_t *errs, u64 *swk);
void mkref_gen_relin_key(const mkref_ctx *c, const u64 *sk, const u64 *r, const int64_t *errs, const u64 *a,
                         const u64 *u, u64 *b, u64 *d, u64 *v);
void mkref_gen_rotation_key(const mkref_ctx *c, int rotidx, const u64 *sk, const int64_t *errs, const u64 *a,
                            u64 *rk);
void mkref_gen_conjugation_key(const mkref_ctx *c, const u64 *sk, const int64_t *errs, const u64 *a, u64 *ck);
void mkref_encrypt(const mkref_ctx *c, int level, const u64 *pk0, const u64 *pk1, const int64_t *v,
                   const int64_t *e0, const int64_t *e1, const u64 *pt, int pt_ntt, int ntt_out, u64 *c0, u64 *c1);

/* ---- MK-BFV (software/kklss/mkbfv, software/mkf/mkf_cpu/bfv.go), in mkref_bfv.c ---- */
/* R = Q u QMul (mkbfv/params.go:36-81): the QP context plus the QMul chain and
 * the plaintext modulus t.  gamma 2 for KKLSS (the decomposition of R polys is
 * restated for alpha = 1: K = 2 or 3, the reference's parameter sets), gamma 1
 * for MKF (Lattigo rlwe decomposition, alpha = K). */
typedef struct {
    mkref_ctx qp;          /* Q[0..L) u P[0..K)                           */
    mkref_ring ringR;      /* Q[0..L) then QMul[0..L)                   */
    u64 qmul[32];
    u64 t;
} mkref_bfv;

int  mkref_bfv_init(mkref_bfv *b, int logN, const u64 *Q, int L, const u64 *QMul, const u64 *P, int K,
                    u64 t, int gamma, int root_rule);
void mkref_bfv_free(mkref_bfv *b);
/* FastBasisExtender.ModUpQtoR (mkbfv/basis_extension.go:50-66): in [L][N] -> out [2L][N] */
void mkref_bfv_modup_q_to_r(const mkref_bfv *b, const u64 *in, u64 *out);
/* FastBasisExtender.Rescale (:88-104): floor(QMul x / Q) in R, in [L][N] -> out [2L][N] */
void mkref_bfv_rescale_q_to_r(const mkref_bfv *b, const u64 *in, u64 *out);
/* FastBasisExtender.Quantize (:69-85): NTT-domain R poly -> ModDown_QMul(t x) [L][N] */
void mkref_bfv_quantize(const mkref_bfv *b, const u64 *in_ntt, u64 *out);
/* KeySwitcher.DecomposeBFV (mkbfv/keyswitch.go:58-83): 2L digits [2L][L+K][N] */
void mkref_bfv_decompose(const mkref_bfv *b, const u64 *aR, u64 *out);
/* MulRelinNew (mkbfv/evaluator.go:85-152 + keyswitch_hoisted.go:50-252).
 * op0/op1: [1+k][L][N] coefficient domain.  Party keys: bkey/dkey [2L][L+K][N]
 * (branch 0's digits then branch 1's), vkey [L][L+K][N] (branch 0's v).
 * out: [1+ku][L][N] with the union of ids ascending in ids_out; returns ku. */
int mkref_kklss_bfv_mulrelin(const mkref_bfv *b,
                             const u64 *op0, const int *ids0, int k0,
                             const u64 *op1, const int *ids1, int k1,
                             const u64 *const *bkey, const u64 *const *dkey,
                             const u64 *const *vkey, const u64 *crs_u,
                             u64 *out, int *ids_out);
/* MKF MK-BFV BFVCase.Run (mkf_cpu/bfv.go:212-240) on a gamma = 1 context:
 * left/right [1+k][L][N] coefficient domain, fold keys / relin key Lattigo
 * switching keys [beta][2][L+K][N]; out0 = c0, out1 = Gamma ([L][N], coefficient) */
void mkref_mkf_bfv_mulrelin(const mkref_bfv *b, int k, const u64 *left, const u64 *right,
                            const u64 *const *fold_keys, const u64 *relin_key, u64 *out0, u64 *out1);
mkref_bfv *mkref_bfv_new(int logN, const u64 *Q, int L, const u64 *QMul, const u64 *P, int K, u64 t,
                         int gamma, int root_rule);
void mkref_bfv_delete(mkref_bfv *b);
void mkref_bfv_ntt_r(const mkref_bfv *b, int limb, u64 *a, int inverse);

/* ---- build-side deterministic PRNG (xoshiro256** seeded via splitmix64) ---- */
typedef struct { u64 s[4]; } mkref_prng;
void mkref_prng_seed(mkref_prng *g, u64 seed);
u64  mkref_prng_next(mkref_prng *g);
void mkref_sample_uniform(mkref_prng *g, const u64 *q, int nlimbs, int N, u64 *out);
void mkref_sample_ternary(mkref_prng *g, int N, double p_zero, int64_t *out);
void mkref_sample_gaussian(mkref_prng *g, int N, double sigma, int bound, int64_t *out);
/* small signed integer poly -> limbs (canonical) */
void mkref_small_to_rns(const int64_t *s, int N, const u64 *q, int n, u64 *out);

/* ---- heap handles for the ctypes test harness ---- */
mkref_ctx *mkref_ctx_new(int logN, const u64 *Q, int L, const u64 *P, int K, int gamma, int root_rule);
void mkref_ctx_delete(mkref_ctx *c);
/* copies the ring's psi / psiinv table of QP ring index `limb` (N words) */
void mkref_ctx_tables(const mkref_ctx *c, int limb, u64 *psi, u64 *psiinv, u64 *root, u64 *ninv);
void mkref_ctx_ntt(const mkref_ctx *c, int limb, u64 *a, int inverse);
/* exposes mkref_modup_params for (Q -> P): writes qoverqiinvqi[nq], qoverqimodp[np*nq],
 * vtimesqmodp[np*(nq+1)] */
void mkref_modup_tables(const u64 *Q, int nq, const u64 *P, int np, u64 *a, u64 *b, u64 *v);
void mkref_modup_exact_simple(int N, const u64 *in, const u64 *Q, int nq, u64 *out, const u64 *P, int np);

#ifdef __cplusplus
}
#endif

/* MK-BFV relinearization keys (mkbfv/keygen.go:24-162) and MKF keys
 * (mkf_cpu/common.go:32-51), in mkref_keygen.c */
void mkref_gen_bfv_relin_key(const mkref_ctx *c, const u64 *gm1, const u64 *gm2, const u64 *sk, const u64 *r,
                             const int64_t *errs, const u64 *a1, const u64 *a2, const u64 *u, u64 *b, u64 *d,
                             u64 *v);
void mkref_gen_mkf_switching_key(const mkref_ctx *c, const u64 *skin, const u64 *skout, const u64 *a,
                                 const int64_t *errs, u64 *swk);
void mkref_gen_mkf_relin_key(const mkref_ctx *c, const u64 *sk, const u64 *a, const int64_t *errs, u64 *swk);
void mkref_aggregate_secret(const mkref_ctx *c, int k, const u64 *sks, u64 *out);

#endif
