/*
 * mkref_keygen.c -- CPU oracle for key generation and public-key encryption.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's key
 * generator and encryptor, used by the tests as the checker of the HIP library.
 * The random choices of the reference (secrets, errors, the encryption's
 * ternary v) come from its crypto PRNG; here they are inputs (small signed
 * integer polynomials), so the arithmetic is deterministic and checkable.
 *
 *   software/kklss/mkrlwe/keygen.go:44-56     genSecretKeyFromSampler
 *   software/kklss/mkrlwe/keygen.go:88-110    GenPublicKey
 *   software/kklss/mkrlwe/keygen.go:123-134   GenGaussianError
 *   software/kklss/mkrlwe/keygen.go:137-187   GenRelinearizationKey
 *   software/kklss/mkrlwe/keygen.go:190-230   GenRotationKey
 *   software/kklss/mkrlwe/keygen.go:240-268   GenConjugationKey
 *   software/kklss/mkrlwe/keygen.go:270-327   GenSwitchingKey
 *   software/kklss/mkrlwe/encryptor.go:55-119 Encryptor.Encrypt
 *   software/kklss/mkbfv/keygen.go:24-162     MK-BFV GenRelinearizationKey and
 *                                             GenBFVSwitchingKey (the gadget values
 *                                             G_br,i mod q come in as a table: the
 *                                             caller computes them with exact big
 *                                             integers, as the reference does)
 *   software/mkf/mkf_cpu/common.go:32-51      aggregateSecret / generateOnlineKeys
 *                                             over Lattigo rlwe GenSwitchingKey and
 *                                             GenRelinearizationKey(s, 1)
 *
 * Lattigo v2.3.0 primitives they call (absent; restated as published):
 * ExtendBasisSmallNormAndCenter (a small coefficient's residue in every modulus of
 * the extended basis: here the signed integer reduced mod each modulus),
 * MulCoeffsMontgomery (MRed), MForm / InvMForm, MulScalarBigint (x (P mod q) mod
 * q), SubLvl / AddLvl (canonical), NegLvl (q - x, unreduced: a zero becomes q),
 * PermuteNTTWithIndexLvl.
 */
#include <stdlib.h>
#include <string.h>

#include "mkref.h"

typedef unsigned __int128 u128;

static u64 addm(u64 a, u64 b, u64 q) {
    u64 s = a + b;
    return s >= q ? s - q : s;
}
static u64 subm(u64 a, u64 b, u64 q) { return a >= b ? a - b : a + q - b; }
static u64 small_res(int64_t v, u64 q) {
    u64 r = v >= 0 ? (u64)v % q : q - ((u64)(-v) % q);
    return r == q ? 0 : r;
}

/* ring index of row r of a full QP polynomial: rows 0..L-1 are Q, L..L+K-1 are P */
static u64 qp_mod(const mkref_ctx *c, int r) { return c->ringQP.q[r]; }

/* NTT(small) over all L+K rows, optionally MForm */
static void small_ntt_qp(const mkref_ctx *c, const int64_t *s, int mform, u64 *out) {
    const int N = c->N, R = c->L + c->K;
    for (int r = 0; r < R; r++) {
        u64 q = qp_mod(c, r);
        u64 *row = out + (size_t)r * N;
        for (int i = 0; i < N; i++) row[i] = small_res(s[i], q);
        mkref_ntt_limb(&c->ringQP, r, row);
        if (mform)
            for (int i = 0; i < N; i++) row[i] = mkref_mform(row[i], q);
    }
}

/* genSecretKeyFromSampler: sampled Q part, centered extension to P, NTT, MForm */
void mkref_gen_secret(const mkref_ctx *c, const int64_t *s, u64 *sk) { small_ntt_qp(c, s, 1, sk); }

/* GenPublicKey: pk0 = NTT(e) - MRed(sk, a), pk1 = a (CRS[0] digit 0), over QP */
void mkref_gen_public_key(const mkref_ctx *c, const u64 *sk, const int64_t *e, const u64 *a, u64 *pk0, u64 *pk1) {
    const int N = c->N, R = c->L + c->K;
    small_ntt_qp(c, e, 0, pk0);
    memcpy(pk1, a, sizeof(u64) * (size_t)R * N);
    for (int r = 0; r < R; r++) {
        u64 q = qp_mod(c, r), qinv = c->ringQP.qinv[r];
        for (int i = 0; i < N; i++) {
            size_t x = (size_t)r * N + i;
            pk0[x] = subm(pk0[x], mkref_mred(sk[x], a[x], q, qinv), q);
        }
    }
}

/* P mod q for every row's modulus */
static void p_mod_q(const mkref_ctx *c, u64 *out) {
    const int R = c->L + c->K;
    for (int r = 0; r < R; r++) {
        u64 q = qp_mod(c, r), p = 1 % q;
        for (int k = 0; k < c->K; k++) p = (u64)((u128)p * (c->ringQP.q[c->L + k] % q) % q);
        out[r] = p;
    }
}

/* GenSwitchingKey: swk_i = MForm(NTT(e_i)); digit i's Q limbs += P * skIn (CRed) */
void mkref_gen_switching_key(const mkref_ctx *c, const u64 *skin, const int64_t *errs, u64 *swk) {
    const int N = c->N, L = c->L, K = c->K, alpha = mkref_alpha(c), beta = mkref_beta(c, L - 1);
    const size_t ksz = (size_t)(L + K) * N;
    u64 pm[128];
    p_mod_q(c, pm);
    for (int i = 0; i < beta; i++) {
        u64 *d = swk + i * ksz;
        small_ntt_qp(c, errs + (size_t)i * N, 1, d);
        for (int j = 0; j < alpha; j++) {
            int idx = i * alpha + j;
            if (idx >= L) break;
            u64 q = qp_mod(c, idx);
            for (int x = 0; x < N; x++) {
                u64 ps = (u64)((u128)skin[(size_t)idx * N + x] * pm[idx] % q);
                d[(size_t)idx * N + x] = addm(d[(size_t)idx * N + x], ps, q);
            }
        }
    }
}

/* out_i -= MRed(a_i, s) over every row of every digit (MulCoeffsMontgomeryAndSub) */
static void sub_mred(const mkref_ctx *c, u64 *out, const u64 *a, const u64 *s, int beta) {
    const int N = c->N, R = c->L + c->K;
    const size_t ksz = (size_t)R * N;
    for (int i = 0; i < beta; i++)
        for (int r = 0; r < R; r++) {
            u64 q = qp_mod(c, r), qinv = c->ringQP.qinv[r];
            for (int x = 0; x < N; x++) {
                size_t o = i * ksz + (size_t)r * N + x;
                out[o] = subm(out[o], mkref_mred(a[o], s[(size_t)r * N + x], q, qinv), q);
            }
        }
}

/*
 * GenRelinearizationKey(sk, r): errs [3][beta][N] are the Gaussian errors of b, d
 * and v in that order; a = CRS[0], u = CRS[-1] (each [beta][L+K][N]).
 *   b_i = MForm(NTT(e_i) - InvMForm(MRed(a_i, sk)))
 *   d_i = GenSwitchingKey(sk)_i - MRed(a_i, r)
 *   v_i = Neg(GenSwitchingKey(r)_i + MRed(u_i, sk))
 */
void mkref_gen_relin_key(const mkref_ctx *c, const u64 *sk, const u64 *r, const int64_t *errs, const u64 *a,
                         const u64 *u, u64 *b, u64 *d, u64 *v) {
    const int N = c->N, L = c->L, K = c->K, R = L + K, beta = mkref_beta(c, L - 1);
    const size_t ksz = (size_t)R * N;
    u64 *tmp = malloc(sizeof(u64) * ksz);
    for (int i = 0; i < beta; i++) {
        small_ntt_qp(c, errs + (size_t)i * N, 0, tmp);
        for (int rr = 0; rr < R; rr++) {
            u64 q = qp_mod(c, rr), qinv = c->ringQP.qinv[rr];
            for (int x = 0; x < N; x++) {
                size_t o = i * ksz + (size_t)rr * N + x, s = (size_t)rr * N + x;
                u64 m = mkref_mred(a[o], sk[s], q, qinv);
                m = mkref_mred(m, 1, q, qinv); /* InvMForm */
                b[o] = mkref_mform(subm(tmp[s], m, q), q);
            }
        }
    }
    mkref_gen_switching_key(c, sk, errs + (size_t)beta * N, d);
    sub_mred(c, d, a, r, beta);
    mkref_gen_switching_key(c, r, errs + (size_t)2 * beta * N, v);
    for (int i = 0; i < beta; i++)
        for (int rr = 0; rr < R; rr++) {
            u64 q = qp_mod(c, rr), qinv = c->ringQP.qinv[rr];
            for (int x = 0; x < N; x++) {
                size_t o = i * ksz + (size_t)rr * N + x;
                u64 y = addm(v[o], mkref_mred(u[o], sk[(size_t)rr * N + x], q, qinv), q);
                v[o] = q - y; /* NegLvl */
            }
        }
    free(tmp);
}

/* PermuteNTTWithIndexLvl over all L+K rows */
static void permute_ntt_qp(const mkref_ctx *c, const u64 *in, u64 gal, u64 *out) {
    const int N = c->N, R = c->L + c->K;
    u64 *idx = malloc(sizeof(u64) * N);
    mkref_permute_ntt_index(c->logN, gal, idx);
    for (int r = 0; r < R; r++)
        for (int i = 0; i < N; i++) out[(size_t)r * N + i] = in[(size_t)r * N + idx[i]];
    free(idx);
}

/* GenRotationKey(rotidx, sk): skOut = sk permuted by the inverse Galois element;
 * rk = GenSwitchingKey(sk); rk_i -= MRed(a_i, skOut), a = CRS[rotidx] */
void mkref_gen_rotation_key(const mkref_ctx *c, int rotidx, const u64 *sk, const int64_t *errs, const u64 *a,
                            u64 *rk) {
    const int N = c->N, R = c->L + c->K, beta = mkref_beta(c, c->L - 1);
    while (rotidx < 0) rotidx += N / 2;
    u64 gal = mkref_galois_inverse(c->logN, mkref_galois_element(c->logN, rotidx));
    u64 *so = malloc(sizeof(u64) * (size_t)R * N);
    permute_ntt_qp(c, sk, gal, so);
    mkref_gen_switching_key(c, sk, errs, rk);
    sub_mred(c, rk, a, so, beta);
    free(so);
}

/* GenConjugationKey(sk): skOut = sk permuted by 2N-1; ck = GenSwitchingKey(skOut);
 * ck_i -= MRed(a_i, sk), a = CRS[-2] */
void mkref_gen_conjugation_key(const mkref_ctx *c, const u64 *sk, const int64_t *errs, const u64 *a, u64 *ck) {
    const int N = c->N, R = c->L + c->K, beta = mkref_beta(c, c->L - 1);
    u64 *so = malloc(sizeof(u64) * (size_t)R * N);
    permute_ntt_qp(c, sk, mkref_galois_conjugate(c->logN), so);
    mkref_gen_switching_key(c, so, errs, ck);
    sub_mred(c, ck, a, sk, beta);
    free(so);
}

/*
 * Encryptor.Encrypt at `level` with public key (pk0, pk1) (their Q rows): v ternary,
 * e0 / e1 Gaussian, pt the plaintext (level+1 limbs, coefficient domain unless
 * pt_ntt).  ntt_out selects the ciphertext domain (ctOut.Value["0"].IsNTT).
 */
void mkref_encrypt(const mkref_ctx *c, int level, const u64 *pk0, const u64 *pk1, const int64_t *v,
                   const int64_t *e0, const int64_t *e1, const u64 *pt, int pt_ntt, int ntt_out, u64 *c0, u64 *c1) {
    const int N = c->N;
    u64 *vm = malloc(sizeof(u64) * N), *t = malloc(sizeof(u64) * N);
    for (int j = 0; j <= level; j++) {
        u64 q = qp_mod(c, j), qinv = c->ringQP.qinv[j];
        for (int i = 0; i < N; i++) vm[i] = small_res(v[i], q);
        mkref_ntt_limb(&c->ringQP, j, vm);
        for (int i = 0; i < N; i++) vm[i] = mkref_mform(vm[i], q);
        u64 *o0 = c0 + (size_t)j * N, *o1 = c1 + (size_t)j * N;
        const u64 *p = pt + (size_t)j * N;
        for (int i = 0; i < N; i++) {
            o0[i] = mkref_mred(vm[i], pk0[(size_t)j * N + i], q, qinv);
            o1[i] = mkref_mred(vm[i], pk1[(size_t)j * N + i], q, qinv);
        }
        if (ntt_out) {
            for (int i = 0; i < N; i++) t[i] = small_res(e1[i], q);
            mkref_ntt_limb(&c->ringQP, j, t);
            for (int i = 0; i < N; i++) o1[i] = addm(o1[i], t[i], q);
            for (int i = 0; i < N; i++) t[i] = small_res(e0[i], q);
            if (!pt_ntt) {
                for (int i = 0; i < N; i++) t[i] = addm(t[i], p[i], q);
                mkref_ntt_limb(&c->ringQP, j, t);
                for (int i = 0; i < N; i++) o0[i] = addm(o0[i], t[i], q);
            } else {
                mkref_ntt_limb(&c->ringQP, j, t);
                for (int i = 0; i < N; i++) o0[i] = addm(addm(o0[i], t[i], q), p[i], q);
            }
        } else {
            mkref_intt_limb(&c->ringQP, j, o0);
            mkref_intt_limb(&c->ringQP, j, o1);
            for (int i = 0; i < N; i++) {
                o0[i] = addm(o0[i], small_res(e0[i], q), q);
                o1[i] = addm(o1[i], small_res(e1[i], q), q);
            }
            if (!pt_ntt) {
                for (int i = 0; i < N; i++) o0[i] = addm(o0[i], p[i], q);
            } else {
                memcpy(t, p, sizeof(u64) * N);
                mkref_intt_limb(&c->ringQP, j, t);
                for (int i = 0; i < N; i++) o0[i] = addm(o0[i], t[i], q);
            }
        }
    }
    free(t);
    free(vm);
}

/* ---------------------------------------------------------------------------
 * KKLSS MK-BFV relinearization keys (mkbfv/keygen.go:24-162).  errs [5][beta][N]
 * are the errors of b_1, b_2, d_1, d_2, v; gm1 / gm2 [beta][L+K] are G_br,i mod q_r
 * of GenBFVSwitchingKey (:91-162); a1 = CRS[0], a2 = CRS[-3], u = CRS[-1].
 *   b_br = MForm(e - InvMForm(MRed(a_br, s)))                  (:46-63)
 *   d_br = MForm(InvMForm(s) G_br,i + e) - MRed(a_br, r)       (:65-73)
 *   v    = NegLvl(GenSwitchingKey(r) + MRed(u, s))             (:75-85)
 * b and d hold branch 1 then branch 2 (2 beta digits each).
 */
void mkref_gen_bfv_relin_key(const mkref_ctx *c, const u64 *gm1, const u64 *gm2, const u64 *sk, const u64 *r,
                             const int64_t *errs, const u64 *a1, const u64 *a2, const u64 *u, u64 *b, u64 *d,
                             u64 *v) {
    const int N = c->N, L = c->L, K = c->K, R = L + K, beta = mkref_beta(c, L - 1);
    const size_t ksz = (size_t)R * N;
    u64 *tmp = malloc(sizeof(u64) * ksz);
    for (int br = 0; br < 2; br++) {
        const u64 *a = br ? a2 : a1, *gm = br ? gm2 : gm1;
        for (int i = 0; i < beta; i++) {
            /* b_br */
            small_ntt_qp(c, errs + ((size_t)br * beta + i) * N, 0, tmp);
            for (int rr = 0; rr < R; rr++) {
                u64 q = qp_mod(c, rr), qinv = c->ringQP.qinv[rr];
                for (int x = 0; x < N; x++) {
                    size_t o = i * ksz + (size_t)rr * N + x, sx = (size_t)rr * N + x;
                    u64 m = mkref_mred(mkref_mred(a[o], sk[sx], q, qinv), 1, q, qinv);
                    b[br * beta * ksz + o] = mkref_mform(subm(tmp[sx], m, q), q);
                }
            }
            /* d_br */
            small_ntt_qp(c, errs + ((size_t)(2 + br) * beta + i) * N, 0, tmp);
            for (int rr = 0; rr < R; rr++) {
                u64 q = qp_mod(c, rr), qinv = c->ringQP.qinv[rr];
                u64 g = gm[(size_t)i * R + rr] % q;
                for (int x = 0; x < N; x++) {
                    size_t o = i * ksz + (size_t)rr * N + x, sx = (size_t)rr * N + x;
                    u64 splain = mkref_mred(sk[sx], 1, q, qinv);
                    u64 y = addm((u64)((u128)splain * g % q), tmp[sx], q);
                    y = mkref_mform(y, q);
                    d[br * beta * ksz + o] = subm(y, mkref_mred(a[o], r[sx], q, qinv), q);
                }
            }
        }
    }
    mkref_gen_switching_key(c, r, errs + (size_t)4 * beta * N, v);
    for (int i = 0; i < beta; i++)
        for (int rr = 0; rr < R; rr++) {
            u64 q = qp_mod(c, rr), qinv = c->ringQP.qinv[rr];
            for (int x = 0; x < N; x++) {
                size_t o = i * ksz + (size_t)rr * N + x;
                u64 y = addm(v[o], mkref_mred(u[o], sk[(size_t)rr * N + x], q, qinv), q);
                v[o] = q - y; /* NegLvl */
            }
        }
    free(tmp);
}

/* ---------------------------------------------------------------------------
 * Lattigo rlwe GenSwitchingKey(skIn, skOut) as the MKF keys use it
 * (mkf_cpu/common.go:44-51): swk [beta][2][L+K][N] with
 *   swk[i][0] = MForm(NTT(e_i)) + P skIn on digit i's Q limbs - MRed(a_i, skOut)
 *   swk[i][1] = a_i    (uniform, taken as already in NTT and Montgomery form)
 * a [beta][L+K][N], errs [beta][N].
 */
void mkref_gen_mkf_switching_key(const mkref_ctx *c, const u64 *skin, const u64 *skout, const u64 *a,
                                 const int64_t *errs, u64 *swk) {
    const int N = c->N, R = c->L + c->K, beta = mkref_beta(c, c->L - 1);
    const size_t ksz = (size_t)R * N;
    u64 *k0 = malloc(sizeof(u64) * ksz * beta);
    mkref_gen_switching_key(c, skin, errs, k0);
    sub_mred(c, k0, a, skout, beta);
    for (int i = 0; i < beta; i++) {
        memcpy(swk + (2 * (size_t)i) * ksz, k0 + (size_t)i * ksz, sizeof(u64) * ksz);
        memcpy(swk + (2 * (size_t)i + 1) * ksz, a + (size_t)i * ksz, sizeof(u64) * ksz);
    }
    free(k0);
}

/* GenRelinearizationKey(s, 1): the switching key from s^2 (MRed(s, s) of the MForm
 * secret is MForm(s^2)) to s */
void mkref_gen_mkf_relin_key(const mkref_ctx *c, const u64 *sk, const u64 *a, const int64_t *errs, u64 *swk) {
    const int N = c->N, R = c->L + c->K;
    u64 *s2 = malloc(sizeof(u64) * (size_t)R * N);
    for (int rr = 0; rr < R; rr++) {
        u64 q = qp_mod(c, rr), qinv = c->ringQP.qinv[rr];
        for (int x = 0; x < N; x++) {
            size_t o = (size_t)rr * N + x;
            s2[o] = mkref_mred(sk[o], sk[o], q, qinv);
        }
    }
    mkref_gen_mkf_switching_key(c, s2, sk, a, errs, swk);
    free(s2);
}

/* aggregateSecret (common.go:32-42): sum of the k secrets over Q u P */
void mkref_aggregate_secret(const mkref_ctx *c, int k, const u64 *sks, u64 *out) {
    const int N = c->N, R = c->L + c->K;
    memset(out, 0, sizeof(u64) * (size_t)R * N);
    for (int i = 0; i < k; i++)
        for (int rr = 0; rr < R; rr++) {
            u64 q = qp_mod(c, rr);
            for (int x = 0; x < N; x++) {
                size_t o = (size_t)rr * N + x;
                out[o] = addm(out[o], sks[(size_t)i * R * N + o], q);
            }
        }
}
