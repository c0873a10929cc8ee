/*
 * mkref_bfv.c -- CPU oracle, KKLSS MK-BFV mul + relin.  TEST INFRASTRUCTURE ONLY
 * (see mkref.h).
 *
 * Reference files restated (paths relative to the reference repository):
 *   software/kklss/mkbfv/params.go           R = Q u QMul, gamma 2 over Q u P
 *   software/kklss/mkbfv/basis_extension.go  ModUpQtoR, Quantize, Rescale
 *   software/kklss/mkbfv/keyswitch.go        DecomposeBFV
 *   software/kklss/mkbfv/keyswitch_hoisted.go ExternalProductBFVHoisted,
 *                                             MulAndRelinBFVHoisted
 *   software/kklss/mkbfv/evaluator.go        MulRelinNew / mulRelinHoisted
 *
 * One intermediate here is not reduced: the basis extensions of ModUpQtoR and
 * Rescale return multSum's raw value (congruent, possibly >= the modulus,
 * mkrlwe/basis_extension.go:581), and DecomposeBFV's copy path copies that raw
 * limb into the rows of every other modulus before their NTT
 * (mkrlwe/basis_extension.go:443-451), where its integer value (not its
 * residue) is reduced.  The restatement therefore keeps those raw values.
 */
#include "mkref.h"

#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static u64 b_addmod(u64 a, u64 b, u64 q) { u64 s = a + b; return s >= q ? s - q : s; }
static u64 b_mulmod(u64 a, u64 b, u64 q) { return (u64)(((u128)a * b) % q); }
static u64 b_inv(u64 a, u64 q) { return mkref_modexp(a % q, q - 2, q); }

static int b_find(const int *ids, int k, int id) {
    for (int i = 0; i < k; i++) if (ids[i] == id) return i;
    return -1;
}

static int b_union(const int *a, int ka, const int *b, int kb, int *out) {
    int n = 0;
    for (int i = 0; i < ka; i++) if (b_find(out, n, a[i]) < 0) out[n++] = a[i];
    for (int i = 0; i < kb; i++) if (b_find(out, n, b[i]) < 0) out[n++] = b[i];
    for (int i = 1; i < n; i++) {
        int v = out[i], j = i - 1;
        while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; j--; }
        out[j + 1] = v;
    }
    return n;
}

int mkref_bfv_init(mkref_bfv *b, int logN, const u64 *Q, int L, const u64 *QMul, const u64 *P, int K,
                   u64 t, int gamma, int root_rule) {
    memset(b, 0, sizeof(*b));
    if (L <= 0 || L > 32 || K <= 0) return -1;
    if (gamma != 1 && gamma != 2) return -1;
    if (gamma == 2 && K / 2 != 1) return -1;  /* KKLSS: alpha = K / gamma = 1 */
    if (mkref_ctx_init(&b->qp, logN, Q, L, P, K, gamma, root_rule)) return -1;
    u64 r[64];
    for (int i = 0; i < L; i++) { r[i] = Q[i]; r[L + i] = QMul[i]; b->qmul[i] = QMul[i]; }
    b->t = t;
    return mkref_ring_init(&b->ringR, logN, r, 2 * L, root_rule);
}

void mkref_bfv_free(mkref_bfv *b) {
    mkref_ctx_free(&b->qp);
    mkref_ring_free(&b->ringR);
}

/* ModUpQtoR (basis_extension.go:50-66): Q limbs copied, QMul limbs = the raw
 * output of convQQMul.ModUpQtoP (mkrlwe modUpExact with paramsQtoP[L-1]) */
void mkref_bfv_modup_q_to_r(const mkref_bfv *b, const u64 *in, u64 *out) {
    int N = b->qp.N, L = b->qp.L;
    const u64 *Q = b->qp.ringQP.q;
    memcpy(out, in, sizeof(u64) * (size_t)L * N);
    mkref_modup m;
    mkref_modup_params(Q, L, b->qmul, L, &m);
    mkref_modup_exact_raw(N, in, Q, L, out + (size_t)L * N, b->qmul, L, &m);
    mkref_modup_free(&m);
}

/*
 * Rescale (basis_extension.go:88-104):
 *   X = MRed(x, MForm(QMul)) = x QMul mod q_i                           (:95)
 *   m = ModDownQPtoP(X, 0): MRed(ModUp_{Q->QMul}(X) + 2p - 0, p - MForm(Q^-1))
 *       (mkrlwe/basis_extension.go:295-335, constants genModDownParams :34-54,
 *        indexed by levelP = levelQ here)                                    (:96-97)
 *   R = [ModUpPtoQ(m) raw, m]                                              (:98-103)
 */
void mkref_bfv_rescale_q_to_r(const mkref_bfv *b, const u64 *in, u64 *out) {
    int N = b->qp.N, L = b->qp.L;
    const u64 *Q = b->qp.ringQP.q, *QM = b->qmul;
    u64 *X = malloc(sizeof(u64) * (size_t)L * N);
    u64 *Y = malloc(sizeof(u64) * (size_t)L * N);
    for (int i = 0; i < L; i++) {
        u64 qi = Q[i], qinv = mkref_mred_params(qi);
        u64 qm = 1;
        for (int j = 0; j < L; j++) qm = b_mulmod(qm, QM[j] % qi, qi);
        u64 qmm = mkref_mform(qm, qi);
        for (int c = 0; c < N; c++) X[(size_t)i * N + c] = mkref_mred(in[(size_t)i * N + c], qmm, qi, qinv);
    }
    mkref_modup m;
    mkref_modup_params(Q, L, QM, L, &m);
    mkref_modup_exact_raw(N, X, Q, L, Y, QM, L, &m);
    mkref_modup_free(&m);
    u64 *mq = out + (size_t)L * N;
    for (int j = 0; j < L; j++) {
        u64 pj = QM[j], pinv = mkref_mred_params(pj);
        u64 qinvp = 1;
        for (int i = 0; i < L; i++) qinvp = b_mulmod(qinvp, b_inv(Q[i] % pj, pj), pj);
        u64 cst = pj - mkref_mform(qinvp, pj);
        for (int c = 0; c < N; c++) mq[(size_t)j * N + c] = mkref_mred(Y[(size_t)j * N + c] + 2 * pj, cst, pj, pinv);
    }
    mkref_modup_params(QM, L, Q, L, &m);
    mkref_modup_exact_raw(N, mq, QM, L, out, Q, L, &m);
    mkref_modup_free(&m);
    free(X);
    free(Y);
}

/* Quantize (basis_extension.go:69-85): MulScalar(t) and InvNTT over R, then
 * convQQMul.ModDownQPtoQ(Q part, QMul part) (mkrlwe/basis_extension.go:192-232) */
void mkref_bfv_quantize(const mkref_bfv *b, const u64 *in_ntt, u64 *out) {
    int N = b->qp.N, L = b->qp.L;
    const mkref_ring *R = &b->ringR;
    u64 *x = malloc(sizeof(u64) * (size_t)2 * L * N);
    for (int l = 0; l < 2 * L; l++) {
        u64 q = R->q[l], tq = b->t % q;
        for (int c = 0; c < N; c++) x[(size_t)l * N + c] = b_mulmod(in_ntt[(size_t)l * N + c], tq, q);
        mkref_intt_limb(R, l, x + (size_t)l * N);
    }
    mkref_moddown(N, x, R->q, L, x + (size_t)L * N, b->qmul, L, out);
    free(x);
}

/*
 * DecomposeBFV (mkbfv/keyswitch.go:58-83): kswRP.DecomposeSingleNTT at levelQ =
 * L - 1 over the R u P parameters, digit i = 0..2 beta - 1.  With alpha = 1 every
 * digit takes the copy path of DecomposeAndSplit: R limb i (unreduced) is copied
 * into rows Q[0..L) and P[0..K), which are NTT'd.  Digits 0..beta-1 fill ad1
 * (the Q limbs), beta..2 beta-1 fill ad2 (the QMul limbs).
 */
void mkref_bfv_decompose(const mkref_bfv *b, const u64 *aR, u64 *out) {
    const mkref_ctx *c = &b->qp;
    if (c->gamma != 2) return;
    int N = c->N, L = c->L, K = c->K, rows = L + K;
    for (int d = 0; d < 2 * L; d++) {
        const u64 *src = aR + (size_t)d * N;
        for (int r = 0; r < rows; r++) {
            u64 q = c->ringQP.q[r];
            u64 *row = out + ((size_t)d * rows + r) * N;
            for (int x = 0; x < N; x++) row[x] = src[x] % q;
            mkref_ntt_limb(&c->ringQP, r, row);
        }
    }
}

/* NTT over R of an R poly in place (2L limbs) */
static void ntt_r(const mkref_bfv *b, u64 *a) {
    int N = b->qp.N;
    for (int l = 0; l < b->ringR.n; l++) {
        u64 *row = a + (size_t)l * N;
        for (int x = 0; x < N; x++) row[x] %= b->ringR.q[l];
        mkref_ntt_limb(&b->ringR, l, row);
    }
}

/*
 * mkbfv Evaluator.MulRelinNew (evaluator.go:85-91) -> mulRelinHoisted (:121-152)
 * -> KeySwitcher.MulAndRelinBFVHoisted (keyswitch_hoisted.go:50-252).
 */
int mkref_kklss_bfv_mulrelin(const mkref_bfv *b,
                             const u64 *op0, const int *ids0, int k0,
                             const u64 *op1, const int *ids1, int k1,
                             const u64 *const *bkey, const u64 *const *dkey,
                             const u64 *const *vkey, const u64 *crs_u,
                             u64 *out, int *ids_out) {
    const mkref_ctx *c = &b->qp;
    const mkref_ring *R = &b->ringR;
    if (c->gamma != 2) return -1;
    int N = c->N, L = c->L, K = c->K, level = L - 1;
    int rows = L + K, beta = L;  /* alpha = 1 */
    size_t psz = (size_t)L * N, rsz = 2 * psz, dsz = (size_t)rows * N;
    size_t h2 = (size_t)2 * beta * dsz, hsz = (size_t)beta * dsz;
    int ku = b_union(ids0, k0, ids1, k1, ids_out);

    /* lifts of every component (evaluator.go:127-135) */
    u64 *r0 = malloc(sizeof(u64) * rsz * (1 + k0));
    u64 *r1 = malloc(sizeof(u64) * rsz * (1 + k1));
    for (int i = 0; i <= k0; i++) mkref_bfv_modup_q_to_r(b, op0 + i * psz, r0 + i * rsz);
    for (int j = 0; j <= k1; j++) mkref_bfv_rescale_q_to_r(b, op1 + j * psz, r1 + j * rsz);

    /* hoisted 2-branch decompositions (:140-146) */
    u64 *D0 = malloc(sizeof(u64) * h2 * (k0 ? k0 : 1));
    u64 *D1 = malloc(sizeof(u64) * h2 * (k1 ? k1 : 1));
    for (int i = 0; i < k0; i++) mkref_bfv_decompose(b, r0 + (1 + i) * rsz, D0 + i * h2);
    for (int j = 0; j < k1; j++) mkref_bfv_decompose(b, r1 + (1 + j) * rsz, D1 + j * h2);

    /* x = [x1; x2] = MForm(sum_i [d1; d2]_i (.) D0_i), y likewise with [b1; b2]
     * (keyswitch_hoisted.go:100-148), as 2 beta-digit switching keys */
    size_t xsz = (size_t)2 * beta * rows * N;
    u64 *x = calloc(xsz, sizeof(u64)), *y = calloc(xsz, sizeof(u64));
    for (int i = 0; i < k0; i++) mkref_mac_keyed_n(c, level, 2 * beta, dkey[ids0[i]], D0 + i * h2, x);
    mkref_mform_qp(c, level, x, 2 * beta);
    for (int j = 0; j < k1; j++) mkref_mac_keyed_n(c, level, 2 * beta, bkey[ids1[j]], D1 + j * h2, y);
    mkref_mform_qp(c, level, y, 2 * beta);

    /* tensor over R and Quantize (:154-200) */
    u64 *p1 = malloc(sizeof(u64) * rsz), *p2 = malloc(sizeof(u64) * rsz);
    u64 *p3 = malloc(sizeof(u64) * rsz), *p4 = malloc(sizeof(u64) * rsz);
    memcpy(p1, r0, sizeof(u64) * rsz);
    memcpy(p2, r1, sizeof(u64) * rsz);
    ntt_r(b, p1);
    ntt_r(b, p2);
    for (int l = 0; l < 2 * L; l++) {
        u64 q = R->q[l], qinv = R->qinv[l];
        for (int e = 0; e < N; e++) {
            size_t o = (size_t)l * N + e;
            p1[o] = mkref_mform(p1[o], q);
            p3[o] = mkref_mred(p1[o], p2[o], q, qinv);
            p2[o] = mkref_mform(p2[o], q);
        }
    }
    mkref_bfv_quantize(b, p3, out);
    for (int u = 0; u < ku; u++) {
        int id = ids_out[u];
        int i0 = b_find(ids0, k0, id), i1 = b_find(ids1, k1, id);
        for (size_t e = 0; e < rsz; e++) p3[e] = 0;
        if (i1 >= 0) {
            memcpy(p4, r1 + (1 + i1) * rsz, sizeof(u64) * rsz);
            ntt_r(b, p4);
            for (size_t e = 0; e < rsz; e++) {
                int l = (int)(e / N);
                p3[e] = mkref_mred(p1[e], p4[e], R->q[l], R->qinv[l]);
            }
        }
        if (i0 >= 0) {
            memcpy(p4, r0 + (1 + i0) * rsz, sizeof(u64) * rsz);
            ntt_r(b, p4);
            for (size_t e = 0; e < rsz; e++) {
                int l = (int)(e / N);
                p3[e] = b_addmod(p3[e], mkref_mred(p2[e], p4[e], R->q[l], R->qinv[l]), R->q[l]);
            }
        }
        mkref_bfv_quantize(b, p3, out + (size_t)(1 + u) * psz);
    }

    /* ctOut_j += ExtBFV(D1_j, x1, x2) (:205-218) */
    u64 *t = malloc(sizeof(u64) * psz), *t2 = malloc(sizeof(u64) * psz);
    u64 *Dt = malloc(sizeof(u64) * hsz);
    for (int j = 0; j < k1; j++) {
        int u = b_find(ids_out, ku, ids1[j]);
        mkref_external_product_n(c, level, 2 * beta, D1 + j * h2, x, t);
        u64 *o = out + (size_t)(1 + u) * psz;
        for (size_t e = 0; e < psz; e++) o[e] = b_addmod(o[e], t[e], c->ringQP.q[e / N]);
    }
    /* ctOut_0 += Ext(D(ExtBFV(D0_i, y)), v_i), ctOut_i += Ext(., u) (:223-248) */
    for (int i = 0; i < k0; i++) {
        int u = b_find(ids_out, ku, ids0[i]);
        mkref_external_product_n(c, level, 2 * beta, D0 + i * h2, y, t);
        mkref_decompose(c, level, t, Dt);
        mkref_external_product_n(c, level, beta, Dt, vkey[ids0[i]], t2);
        for (size_t e = 0; e < psz; e++) out[e] = b_addmod(out[e], t2[e], c->ringQP.q[e / N]);
        mkref_external_product_n(c, level, beta, Dt, crs_u, t2);
        u64 *o = out + (size_t)(1 + u) * psz;
        for (size_t e = 0; e < psz; e++) o[e] = b_addmod(o[e], t2[e], c->ringQP.q[e / N]);
    }

    free(r0); free(r1); free(D0); free(D1); free(x); free(y);
    free(p1); free(p2); free(p3); free(p4); free(t); free(t2); free(Dt);
    return ku + 1;
}

/*
 * MKF MK-BFV, BFVCase.Run (software/mkf/mkf_cpu/bfv.go:212-240) over a gamma = 1
 * context (Lattigo rlwe decomposition, alpha = K):
 *   fold both sides with one coefficient-domain key switch per party
 *     (common.go:79-126, nttOutput = false: InvNTT of every row + ModDownQPtoQ);
 *   modUpLeft / rescaleRight into R and NTT (bfv.go:154-175; Lattigo
 *     ring.FastBasisExtender, the algorithm restated in mkref_bfv_modup_q_to_r /
 *     mkref_bfv_rescale_q_to_r);
 *   tensorKaratsuba over R (:179-192); quantize each product (:194-210);
 *   relinearize mu2: out0 = p0 + switch(p2).0, Gamma = p1 + switch(p2).1 (:233-237).
 * left / right: [1+k][L][N] coefficient domain ([C0, components...]).
 * The k-fold materialization of Gamma (:238) is the caller's copy.
 */
void mkref_mkf_bfv_mulrelin(const mkref_bfv *b, int k, const u64 *left, const u64 *right,
                            const u64 *const *fold_keys, const u64 *relin_key, u64 *out0, u64 *out1) {
    const mkref_ctx *c = &b->qp;
    const mkref_ring *R = &b->ringR;
    int N = c->N, L = c->L, level = L - 1;
    size_t psz = (size_t)L * N, rsz = 2 * psz;
    u64 *f[2][2], *s0 = malloc(sizeof(u64) * psz), *s1 = malloc(sizeof(u64) * psz);
    const u64 *side[2] = {left, right};
    for (int s = 0; s < 2; s++) {
        f[s][0] = malloc(sizeof(u64) * psz);
        f[s][1] = calloc(psz, sizeof(u64));
        memcpy(f[s][0], side[s], sizeof(u64) * psz);
        for (int i = 0; i < k; i++) {
            mkref_mkf_switch_one(c, level, side[s] + (size_t)(1 + i) * psz, 0, fold_keys[i], 0, s0, s1);
            for (size_t e = 0; e < psz; e++) {
                u64 q = c->ringQP.q[e / N];
                f[s][0][e] = b_addmod(f[s][0][e], s0[e], q);
                f[s][1][e] = b_addmod(f[s][1][e], s1[e], q);
            }
        }
    }
    u64 *in[2][2];
    for (int j = 0; j < 2; j++) {
        in[0][j] = malloc(sizeof(u64) * rsz);
        in[1][j] = malloc(sizeof(u64) * rsz);
        mkref_bfv_modup_q_to_r(b, f[0][j], in[0][j]);
        mkref_bfv_rescale_q_to_r(b, f[1][j], in[1][j]);
        ntt_r(b, in[0][j]);
        ntt_r(b, in[1][j]);
    }
    u64 *mu[3];
    for (int i = 0; i < 3; i++) mu[i] = malloc(sizeof(u64) * rsz);
    for (size_t e = 0; e < rsz; e++) {
        int l = (int)(e / N);
        u64 q = R->q[l], qinv = R->qinv[l];
        u64 m0 = mkref_mred(mkref_mform(in[0][0][e], q), in[1][0][e], q, qinv);
        u64 m2 = mkref_mred(mkref_mform(in[0][1][e], q), in[1][1][e], q, qinv);
        u64 ls = mkref_mform(b_addmod(in[0][0][e], in[0][1][e], q), q);
        u64 rs = b_addmod(in[1][0][e], in[1][1][e], q);
        u64 m1 = mkref_mred(ls, rs, q, qinv);
        m1 = (m1 + 2 * q - m0 - m2) % q;
        mu[0][e] = m0;
        mu[1][e] = m1;
        mu[2][e] = m2;
    }
    u64 *p[3];
    for (int i = 0; i < 3; i++) {
        p[i] = malloc(sizeof(u64) * psz);
        mkref_bfv_quantize(b, mu[i], p[i]);
    }
    mkref_mkf_switch_one(c, level, p[2], 0, relin_key, 0, s0, s1);
    for (size_t e = 0; e < psz; e++) {
        u64 q = c->ringQP.q[e / N];
        out0[e] = b_addmod(p[0][e], s0[e], q);
        out1[e] = b_addmod(p[1][e], s1[e], q);
    }
    for (int s = 0; s < 2; s++) { free(f[s][0]); free(f[s][1]); free(in[s][0]); free(in[s][1]); }
    for (int i = 0; i < 3; i++) { free(mu[i]); free(p[i]); }
    free(s0);
    free(s1);
}

mkref_bfv *mkref_bfv_new(int logN, const u64 *Q, int L, const u64 *QMul, const u64 *P, int K, u64 t,
                         int gamma, int root_rule) {
    mkref_bfv *b = malloc(sizeof(*b));
    if (!b) return NULL;
    if (mkref_bfv_init(b, logN, Q, L, QMul, P, K, t, gamma, root_rule)) {
        free(b);
        return NULL;
    }
    return b;
}

void mkref_bfv_delete(mkref_bfv *b) {
    if (!b) return;
    mkref_bfv_free(b);
    free(b);
}

void mkref_bfv_ntt_r(const mkref_bfv *b, int limb, u64 *a, int inverse) {
    int N = b->qp.N;
    for (int x = 0; x < N; x++) a[x] %= b->ringR.q[limb];
    if (inverse) mkref_intt_limb(&b->ringR, limb, a);
    else mkref_ntt_limb(&b->ringR, limb, a);
}
