/*
 * mkref_rot.c -- CPU oracle for the Galois (rotation / conjugation) key switches.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's rotation
 * and conjugation paths, used by the tests as the checker of the HIP library.
 *
 *   software/kklss/mkrlwe/keyswitch.go:243-307          KeySwitcher.Rotate
 *   software/kklss/mkrlwe/keyswitch.go:309-344          KeySwitcher.Conjugate
 *   software/kklss/mkrlwe/keyswitch_hoisted.go:195-262  KeySwitcher.RotateHoisted
 *   software/kklss/mkrlwe/keyswitch.go:82-130           KeySwitcher.ExternalProduct
 *   software/mkf/mkf_cpu/true_mkf_ckks.go:306-353       mkfAutomorphismNTT, RotateNew
 *
 * Lattigo v2.3.0 (absent) supplies GaloisElementForColumnRotationBy,
 * GaloisElementForRowRotation, InverseGaloisElement, ring.Permute and
 * ring.PermuteNTTIndex; their published algorithms are restated here:
 *   galEl(k)  = GaloisGen^(k mod 2N) mod 2N, GaloisGen = 5
 *   row rot.  = 2N - 1
 *   inverse   = galEl^(2N-1) mod 2N
 *   Permute   : coefficient i -> position (i*galEl) mod N, negated when bit logN
 *               of i*galEl is set, negation written as q - x (so 0 -> q); the
 *               reference inlines the same loop at keyswitch.go:283-303.
 *   PermuteNTTIndex: out[i] = in[index[i]], index[i] = brev(((galEl*(2 brev(i)+1))
 *               mod 2N - 1) / 2): slot i of the bit-reversed negacyclic NTT is the
 *               evaluation at psi^(2 brev(i)+1), and (sigma_g a)(psi^e) = a(psi^(g e)).
 */
#include <stdlib.h>
#include <string.h>

#include "mkref.h"

static u64 addm(u64 a, u64 b, u64 q) {
    u64 s = a + b;
    return s >= q ? s - q : s;
}

static u64 brev(u64 x, int bits) {
    u64 r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
    return r;
}

u64 mkref_galois_element(int logN, long long k) {
    u64 twoN = (u64)2 << logN;
    u64 e = (u64)k & (twoN - 1);
    u64 r = 1, b = 5 % twoN;
    while (e) {
        if (e & 1) r = (r * b) & (twoN - 1);
        b = (b * b) & (twoN - 1);
        e >>= 1;
    }
    return r;
}

u64 mkref_galois_conjugate(int logN) { return ((u64)2 << logN) - 1; }

u64 mkref_galois_inverse(int logN, u64 g) {
    u64 twoN = (u64)2 << logN;
    u64 e = twoN - 1, r = 1, b = g & (twoN - 1);
    while (e) {
        if (e & 1) r = (r * b) & (twoN - 1);
        b = (b * b) & (twoN - 1);
        e >>= 1;
    }
    return r;
}

/* ring.Permute restricted to limbs 0..level (keyswitch.go:283-303): out must not alias in */
void mkref_permute(const mkref_ctx *c, int level, const u64 *in, u64 gal, u64 *out) {
    const int N = c->N, logN = c->logN;
    const u64 mask = (u64)N - 1;
    for (u64 i = 0; i < (u64)N; i++) {
        u64 raw = i * gal;
        u64 idx = raw & mask;
        u64 neg = (raw >> logN) & 1;
        for (int j = 0; j <= level; j++) {
            u64 q = c->ringQP.q[j];
            u64 x = in[(size_t)j * N + i];
            out[(size_t)j * N + idx] = (x * (neg ^ 1)) | ((q - x) * neg);
        }
    }
}

/* ring.PermuteNTTIndex */
void mkref_permute_ntt_index(int logN, u64 gal, u64 *index) {
    const u64 N = (u64)1 << logN, mask = 2 * N - 1;
    for (u64 i = 0; i < N; i++) {
        u64 e = 2 * brev(i, logN) + 1;
        u64 t = (((gal * e) & mask) - 1) >> 1;
        index[i] = brev(t, logN);
    }
}

/* mkfAutomorphismNTT (true_mkf_ckks.go:306-319): InvNTT, Permute, NTT of limbs 0..level */
void mkref_automorphism_ntt(const mkref_ctx *c, int level, const u64 *in, u64 gal, u64 *out) {
    const int N = c->N;
    size_t psz = (size_t)(level + 1) * N;
    u64 *coef = malloc(sizeof(u64) * psz);
    memcpy(coef, in, sizeof(u64) * psz);
    for (int l = 0; l <= level; l++) mkref_intt_limb(&c->ringQP, l, coef + (size_t)l * N);
    mkref_permute(c, level, coef, gal, out);
    for (int l = 0; l <= level; l++) mkref_ntt_limb(&c->ringQP, l, out + (size_t)l * N);
    free(coef);
}

/*
 * KeySwitcher.ExternalProduct as written (keyswitch.go:82-130): every digit is
 * decomposed into the same buffer (:104-106) before the MAC loop (:110-117), so
 * the MAC multiplies every key digit by the last digit's decomposition.  Kept to
 * document how the non-hoisted Rotate / Conjugate of the reference behave; the
 * library follows the hoisted external product instead.
 */
void mkref_external_product_as_written(const mkref_ctx *c, int levelQ, const u64 *a, const u64 *key, u64 *out) {
    const int N = c->N, K = c->K;
    const int beta = mkref_beta(c, levelQ), rows = levelQ + 1 + K;
    const size_t dsz = (size_t)rows * N, ksz = (size_t)(c->L + K) * N;
    u64 *last = malloc(sizeof(u64) * dsz);
    for (int b = 0; b < beta; b++) mkref_decompose_digit(c, levelQ, b, a, last);
    u64 *rep = malloc(sizeof(u64) * beta * dsz);
    for (int b = 0; b < beta; b++) memcpy(rep + b * dsz, last, sizeof(u64) * dsz);
    (void)ksz;
    mkref_external_product_n(c, levelQ, beta, rep, key, out);
    free(rep);
    free(last);
}

/*
 * KeySwitcher.RotateHoisted (keyswitch_hoisted.go:197-262) for one rotation:
 *   in  [1+k][level+1][N] coefficient domain, D [k][beta][level+1+K][N] its
 *   hoisted form (mkckks.Evaluator.HoistedForm, evaluator.go:546-556),
 *   rk[i] the rotation key of party i for this rotation, a = CRS[rotidx];
 *   out = sigma_gal(c0 + sum_i Ext(D_i, rk_i), Ext(D_1, a), ..., Ext(D_k, a)).
 * With k = 0 the c0 copy is permuted unchanged (no AddLvl touches it).
 */
void mkref_kklss_rotate_hoisted(const mkref_ctx *c, int level, int k, const u64 *in, const u64 *D,
                                const u64 *const *rk, const u64 *a, u64 gal, u64 *out) {
    const int N = c->N, nl = level + 1;
    const size_t psz = (size_t)nl * N, hsz = (size_t)mkref_beta(c, level) * (nl + c->K) * N;
    u64 *tmp = malloc(sizeof(u64) * (1 + k) * psz);
    u64 *e = malloc(sizeof(u64) * psz);
    memcpy(tmp, in, sizeof(u64) * psz);
    for (int i = 0; i < k; i++) {
        mkref_external_product_hoisted(c, level, D + i * hsz, rk[i], e);
        for (size_t x = 0; x < psz; x++) tmp[x] = addm(tmp[x], e[x], c->ringQP.q[x / N]);
        mkref_external_product_hoisted(c, level, D + i * hsz, a, tmp + (1 + i) * psz);
    }
    for (int p = 0; p <= k; p++) mkref_permute(c, level, tmp + p * psz, gal, out + p * psz);
    free(e);
    free(tmp);
}

/* KeySwitcher.Rotate (keyswitch.go:243-307) with the hoisted external product:
 * the decomposition is computed here. */
void mkref_kklss_rotate(const mkref_ctx *c, int level, int k, const u64 *in, const u64 *const *rk, const u64 *a,
                        u64 gal, u64 *out) {
    const int N = c->N, nl = level + 1;
    const size_t psz = (size_t)nl * N, hsz = (size_t)mkref_beta(c, level) * (nl + c->K) * N;
    u64 *D = malloc(sizeof(u64) * (k ? k : 1) * hsz);
    for (int i = 0; i < k; i++) mkref_decompose(c, level, in + (1 + i) * psz, D + i * hsz);
    mkref_kklss_rotate_hoisted(c, level, k, in, D, rk, a, gal, out);
    free(D);
}

/*
 * KeySwitcher.Conjugate (keyswitch.go:309-344) with the hoisted external product:
 * every component is permuted by 2N-1 first, then c0 += Ext(D(c_i'), ck_i) and
 * c_i = Ext(D(c_i'), CRS[-2]).  The reference's ring.Permute covers the ring's
 * whole chain; restricted to limbs 0..level here.
 */
void mkref_kklss_conjugate(const mkref_ctx *c, int level, int k, const u64 *in, const u64 *const *ck,
                           const u64 *a, u64 *out) {
    const int N = c->N, nl = level + 1;
    const size_t psz = (size_t)nl * N, hsz = (size_t)mkref_beta(c, level) * (nl + c->K) * N;
    const u64 gal = mkref_galois_conjugate(c->logN);
    for (int p = 0; p <= k; p++) mkref_permute(c, level, in + p * psz, gal, out + p * psz);
    u64 *D = malloc(sizeof(u64) * hsz);
    u64 *e = malloc(sizeof(u64) * psz);
    for (int i = 0; i < k; i++) {
        mkref_decompose(c, level, out + (1 + i) * psz, D);
        mkref_external_product_hoisted(c, level, D, ck[i], e);
        for (size_t x = 0; x < psz; x++) out[x] = addm(out[x], e[x], c->ringQP.q[x / N]);
        mkref_external_product_hoisted(c, level, D, a, out + (1 + i) * psz);
    }
    free(e);
    free(D);
}

/*
 * paperMKFCKKSEvaluator.RotateNew (true_mkf_ckks.go:323-353): in [1+k][level+1][N]
 * NTT domain; keys[i] the Lattigo switching key sigma_gal(s_i) -> s_E of party i
 * ([beta][2][L+K][N]); out0 = sigma(C0) + sum_i u0_i, out1 = sum_i u1_i (NTT).
 */
void mkref_mkf_ckks_rotate(const mkref_ctx *c, int level, int k, const u64 *in, const u64 *const *keys, u64 gal,
                           u64 *out0, u64 *out1) {
    const int N = c->N, nl = level + 1;
    const size_t psz = (size_t)nl * N;
    u64 *perm = malloc(sizeof(u64) * psz);
    u64 *u0 = malloc(sizeof(u64) * psz), *u1 = malloc(sizeof(u64) * psz);
    mkref_automorphism_ntt(c, level, in, gal, out0);
    memset(out1, 0, sizeof(u64) * psz);
    for (int i = 0; i < k; i++) {
        mkref_automorphism_ntt(c, level, in + (1 + i) * psz, gal, perm);
        mkref_mkf_switch_one(c, level, perm, 1, keys[i], 1, u0, u1);
        for (size_t x = 0; x < psz; x++) {
            u64 q = c->ringQP.q[x / N];
            out0[x] = addm(out0[x], u0[x], q);
            out1[x] = addm(out1[x], u1[x], q);
        }
    }
    free(u1);
    free(u0);
    free(perm);
}
