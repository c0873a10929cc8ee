/*
 * mkref.c -- CPU oracle: plain-C restatement of the reference multi-key
 * mul+relin path.  TEST INFRASTRUCTURE ONLY (see mkref.h).
 *
 * Reference files restated (paths relative to the reference repository):
 *   software/kklss/mkrlwe/basis_extension.go   FastBasisExtender, Decomposer
 *   software/kklss/mkrlwe/keyswitch.go          Decompose, DecomposeSingleNTT
 *   software/kklss/mkrlwe/keyswitch_hoisted.go  ExternalProductHoisted, MulAndRelinHoisted
 *   software/kklss/mkckks/evaluator.go          MulRelinNew, Rescale
 *   software/mkf/mkf_cpu/common.go, ckks.go     switchOne, fold, CKKSCase.Run
 *   hardware/u55c/hls/kernels/par1_mkf_ckks_kernel.cpp  ntt_limb (restatement of the
 *                                               Lattigo v2.3.0 lazy CT/GS butterflies)
 *   hardware/u55c/hls/kernels/par1_montgomery.hpp        mred64 (Lattigo MRed)
 *
 * Every intermediate of the reference is congruent to the canonical value
 * computed here (the reference keeps some values lazily in [0,2q) or [0,4q)
 * and reduces them at the next MRed/CRed); the final outputs of the reference
 * are canonical, so canonical intermediates give identical output bits.  The
 * one place where a non-modular quantity enters is the float64 overflow
 * estimate of the exact basis extension, which is reproduced operation for
 * operation (basis_extension.go:499-517, 560-578).
 */
#include "mkref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* scalar arithmetic                                                         */
/* ------------------------------------------------------------------------ */

/* Lattigo ring.MRedParams: q^(2^63-1) = q^-1 mod 2^64 (par1_montgomery.hpp:5-6) */
u64 mkref_mred_params(u64 q) {
    u64 qinv = 1, x = q;
    for (int i = 0; i < 63; i++) {
        qinv *= x;
        x *= x;
    }
    return qinv;
}

/* Lattigo ring.MRed: x*y*2^-64 mod q (par1_montgomery.hpp:7-26) */
u64 mkref_mred(u64 x, u64 y, u64 q, u64 qinv) {
    u128 xy = (u128)x * y;
    u64 hi = (u64)(xy >> 64), lo = (u64)xy;
    u64 m = lo * qinv;
    u64 mqhi = (u64)(((u128)m * q) >> 64);
    u64 r = hi - mqhi + q;
    if (r >= q) r -= q;
    return r;
}

/* Lattigo ring.MRedConstant (par1_montgomery.hpp:30-41) */
u64 mkref_mred_lazy(u64 x, u64 y, u64 q, u64 qinv) {
    u128 xy = (u128)x * y;
    u64 hi = (u64)(xy >> 64), lo = (u64)xy;
    u64 m = lo * qinv;
    u64 mqhi = (u64)(((u128)m * q) >> 64);
    return hi - mqhi + q;
}

/* Lattigo ring.MForm: a*2^64 mod q */
u64 mkref_mform(u64 a, u64 q) { return (u64)((((u128)(a % q)) << 64) % q); }

static u64 mulmod(u64 a, u64 b, u64 q) { return (u64)(((u128)a * b) % q); }
static u64 addmod(u64 a, u64 b, u64 q) { u64 s = a + b; return s >= q ? s - q : s; }
static u64 submod(u64 a, u64 b, u64 q) { return a >= b ? a - b : a + q - b; }

u64 mkref_modexp(u64 x, u64 e, u64 q) {
    u64 r = 1 % q, b = x % q;
    while (e) {
        if (e & 1) r = mulmod(r, b, q);
        b = mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}

static u64 invmod(u64 a, u64 q) { return mkref_modexp(a % q, q - 2, q); }

/* deterministic Miller-Rabin for 64-bit integers */
static int is_prime(u64 n) {
    if (n < 2) return 0;
    static const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (int i = 0; i < 12; i++) {
        if (n % small[i] == 0) return n == small[i];
    }
    u64 d = n - 1;
    int s = 0;
    while ((d & 1) == 0) { d >>= 1; s++; }
    for (int i = 0; i < 12; i++) {
        u64 x = mkref_modexp(small[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int r = 1; r < s; r++) {
            x = mulmod(x, x, n);
            if (x == n - 1) { comp = 0; break; }
        }
        if (comp) return 0;
    }
    return 1;
}

/* distinct prime factors of n by trial division (n - 1 of an NTT prime is
 * 2^k * m with m < 2^48, so the loop is bounded by ~2^24 iterations) */
static int prime_factors(u64 n, u64 *f) {
    int c = 0;
    if ((n & 1) == 0) {
        f[c++] = 2;
        while ((n & 1) == 0) n >>= 1;
    }
    for (u64 p = 3; p * p <= n; p += 2) {
        if (n % p == 0) {
            f[c++] = p;
            while (n % p == 0) n /= p;
        }
    }
    if (n > 1) f[c++] = n;
    return c;
}

/*
 * Primitive root of Z_q^*.  The reference tables come from Lattigo v2.3.0's
 * ring.NewRing, which searches the smallest generator starting the loop with
 * g = 2 followed by g++ before the first test (so candidates start at 3).
 * That rule is rule 0 here.  No in-tree file pins it (the NTT tables are only
 * exported at run time by software/mkf/mkf_cpu/fpga_constants_test.go:118-122),
 * so the rule is a parameter.
 */
u64 mkref_primitive_root(u64 q, int rule) {
    u64 f[64];
    int nf = prime_factors(q - 1, f);
    for (u64 g = (rule == 0 ? 3 : 2);; g++) {
        int ok = 1;
        for (int i = 0; i < nf; i++) {
            if (mkref_modexp(g, (q - 1) / f[i], q) == 1) { ok = 0; break; }
        }
        if (ok) return g;
    }
}

void mkref_butterfly(u64 u, u64 v, u64 w, u64 q, u64 qinv, u64 *sum, u64 *diff) {
    u64 t = mkref_mred(v, w, q, qinv);
    *sum = addmod(u, t, q);
    *diff = submod(u, t, q);
}

int mkref_gen_primes(int bits, int logN, int count, int downward_from_top, u64 *out,
                     const u64 *exclude, int n_exclude) {
    /* primes q = 1 mod 2^(logN+1), scanning downward from 2^bits + 1 */
    u64 step = (u64)1 << (logN + 1);
    u64 x = ((u64)1 << bits) + 1;
    (void)downward_from_top;
    int c = 0;
    while (c < count) {
        x -= step;
        if (x < ((u64)1 << (bits - 1))) return -1;
        if (!is_prime(x)) continue;
        int skip = 0;
        for (int i = 0; i < n_exclude; i++) if (exclude[i] == x) skip = 1;
        for (int i = 0; i < c; i++) if (out[i] == x) skip = 1;
        if (skip) continue;
        out[c++] = x;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ring / NTT                                                                */
/* ------------------------------------------------------------------------ */

static u64 bitrev(u64 x, int bits) {
    u64 r = 0;
    for (int i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

int mkref_ring_init(mkref_ring *r, int logN, const u64 *q, int n, int root_rule) {
    memset(r, 0, sizeof(*r));
    r->logN = logN;
    r->N = 1 << logN;
    r->n = n;
    r->q = malloc(sizeof(u64) * n);
    r->qinv = malloc(sizeof(u64) * n);
    r->ninv = malloc(sizeof(u64) * n);
    r->root = malloc(sizeof(u64) * n);
    r->psi = malloc(sizeof(u64) * n * r->N);
    r->psiinv = malloc(sizeof(u64) * n * r->N);
    if (!r->q || !r->qinv || !r->ninv || !r->root || !r->psi || !r->psiinv) return -1;
    for (int i = 0; i < n; i++) {
        u64 qi = q[i];
        if (((qi - 1) % (2 * (u64)r->N)) != 0) return -2;
        r->q[i] = qi;
        r->qinv[i] = mkref_mred_params(qi);
        r->ninv[i] = mkref_mform(invmod((u64)r->N, qi), qi);
        u64 g = mkref_primitive_root(qi, root_rule);
        r->root[i] = g;
        u64 power = (qi - 1) / (2 * (u64)r->N);
        u64 psi_m = mkref_mform(mkref_modexp(g, power, qi), qi);
        u64 psiinv_m = mkref_mform(mkref_modexp(g, (qi - 1) - power, qi), qi);
        u64 *P = r->psi + (size_t)i * r->N, *Pi = r->psiinv + (size_t)i * r->N;
        P[0] = mkref_mform(1, qi);
        Pi[0] = mkref_mform(1, qi);
        for (u64 j = 1; j < (u64)r->N; j++) {
            u64 prev = bitrev(j - 1, logN), next = bitrev(j, logN);
            P[next] = mkref_mred(P[prev], psi_m, qi, r->qinv[i]);
            Pi[next] = mkref_mred(Pi[prev], psiinv_m, qi, r->qinv[i]);
        }
    }
    return 0;
}

void mkref_ring_free(mkref_ring *r) {
    free(r->q); free(r->qinv); free(r->ninv); free(r->root); free(r->psi); free(r->psiinv);
    memset(r, 0, sizeof(*r));
}

/*
 * Forward NTT, Lattigo NTTLazy + exact final reduction, as restated by
 * par1_mkf_ckks_kernel.cpp:92-123: stage m (1,2,..N/2) pairs p, p+t in group g
 * with twiddle psi[m+g]; butterfly u' = u mod 4q, v' = MRed(v, w),
 * (u'+v', u'+2q-v').  Input is reduced mod q first (the reference's lazy
 * butterflies reach the same canonical output for any input below 2^63).
 */
void mkref_ntt_limb_tables(int logN, u64 q, u64 qinv, const u64 *psi, u64 *a) {
    int N = 1 << logN;
    u64 two_q = q << 1, four_q = q << 2;
    for (int i = 0; i < N; i++) a[i] %= q;
    int t = N >> 1;
    for (int m = 1; m < N; m <<= 1, t >>= 1) {
        for (int g = 0; g < m; g++) {
            u64 w = psi[m + g];
            int j1 = 2 * g * t;
            for (int j = j1; j < j1 + t; j++) {
                u64 u = a[j];
                if (u >= four_q) u -= four_q;
                u64 v = mkref_mred_lazy(a[j + t], w, q, qinv);
                a[j] = u + v;
                a[j + t] = u + two_q - v;
            }
        }
    }
    for (int i = 0; i < N; i++) a[i] %= q;
}

/* Inverse NTT, Lattigo InvNTTLazy with exact final MRed by N^-1
 * (par1_mkf_ckks_kernel.cpp:124-157). */
void mkref_intt_limb_tables(int logN, u64 q, u64 qinv, const u64 *psiinv, u64 ninv, u64 *a) {
    int N = 1 << logN;
    u64 two_q = q << 1, four_q = q << 2;
    for (int i = 0; i < N; i++) a[i] %= q;
    int t = 1;
    for (int m = N >> 1; m != 0; m >>= 1, t <<= 1) {
        for (int g = 0; g < m; g++) {
            u64 w = psiinv[m + g];
            int j1 = 2 * g * t;
            for (int j = j1; j < j1 + t; j++) {
                u64 u = a[j], v = a[j + t];
                u64 x = u + v;
                if (x >= two_q) x -= two_q;
                a[j] = x;
                a[j + t] = mkref_mred_lazy(u + four_q - v, w, q, qinv);
            }
        }
    }
    for (int i = 0; i < N; i++) a[i] = mkref_mred(a[i], ninv, q, qinv);
}

void mkref_ntt_limb(const mkref_ring *r, int limb, u64 *a) {
    mkref_ntt_limb_tables(r->logN, r->q[limb], r->qinv[limb], r->psi + (size_t)limb * r->N, a);
}

void mkref_intt_limb(const mkref_ring *r, int limb, u64 *a) {
    mkref_intt_limb_tables(r->logN, r->q[limb], r->qinv[limb], r->psiinv + (size_t)limb * r->N,
                           r->ninv[limb], a);
}

/* ------------------------------------------------------------------------ */
/* FastBasisExtender                                                          */
/* ------------------------------------------------------------------------ */

/* basisextenderparameters, basis_extension.go:83-153 */
int mkref_modup_params(const u64 *Q, int nq, const u64 *P, int np, mkref_modup *m) {
    m->nin = nq;
    m->nout = np;
    m->qoverqiinvqi = malloc(sizeof(u64) * nq);
    m->qoverqimodp = malloc(sizeof(u64) * nq * np);
    m->vtimesqmodp = malloc(sizeof(u64) * (nq + 1) * np);
    m->qin = malloc(sizeof(u64) * nq);
    m->qinv_in = malloc(sizeof(u64) * nq);
    m->pinv_out = malloc(sizeof(u64) * np);
    if (!m->qoverqiinvqi || !m->qoverqimodp || !m->vtimesqmodp || !m->qin || !m->qinv_in || !m->pinv_out)
        return -1;
    for (int i = 0; i < nq; i++) { m->qin[i] = Q[i]; m->qinv_in[i] = mkref_mred_params(Q[i]); }
    for (int j = 0; j < np; j++) m->pinv_out[j] = mkref_mred_params(P[j]);
    for (int i = 0; i < nq; i++) {
        u64 qi = Q[i];
        /* qiStar = prod_{j!=i} Q_j mod q_i, then (qiStar)^-1 in Montgomery form */
        u64 star = 1;
        for (int j = 0; j < nq; j++) if (j != i) star = mulmod(star, Q[j] % qi, qi);
        m->qoverqiinvqi[i] = mkref_mform(invmod(star, qi), qi);
        for (int j = 0; j < np; j++) {
            u64 pj = P[j];
            u64 s = 1;
            for (int u = 0; u < nq; u++) if (u != i) s = mulmod(s, Q[u] % pj, pj);
            m->qoverqimodp[(size_t)j * nq + i] = mkref_mform(s, pj);
        }
    }
    for (int j = 0; j < np; j++) {
        u64 pj = P[j];
        u64 qmodp = 1;
        for (int i = 0; i < nq; i++) qmodp = mulmod(qmodp, Q[i] % pj, pj);
        u64 v = (pj - qmodp) % pj;
        u64 *vt = m->vtimesqmodp + (size_t)j * (nq + 1);
        vt[0] = 0;
        for (int i = 1; i <= nq; i++) vt[i] = addmod(vt[i - 1], v, pj);
    }
    return 0;
}

void mkref_modup_free(mkref_modup *m) {
    free(m->qoverqiinvqi); free(m->qoverqimodp); free(m->vtimesqmodp);
    free(m->qin); free(m->qinv_in); free(m->pinv_out);
    memset(m, 0, sizeof(*m));
}

/*
 * y_i = MRed(x_i, (Q/q_i)^-1 R), v = (u64)(sum_i float64(y_i)/float64(q_i)) in
 * ascending i (reconstructRNS, basis_extension.go:537-579), then
 * out_j = sum_i y_i (Q/q_i) + v (-Q) mod p_j (multSum, :582-646).
 * With raw != 0 the value multSum returns is kept as is: hi - hi(lo*pinv*p) + p
 * + vt[v], congruent to out_j but possibly >= p_j (the reference's "[0, 2q-1]").
 */
static void modup_coeff_impl(const u64 *x, int stride, const u64 *Q, int nq, const u64 *P, int np,
                             const mkref_modup *m, u64 *out, int ostride, int raw) {
    u64 y[64];
    double vi = 0.0;
    for (int i = 0; i < nq; i++) {
        u64 qi = Q[i];
        y[i] = mkref_mred(x[(size_t)i * stride], m->qoverqiinvqi[i], qi, m->qinv_in[i]);
        vi += (double)y[i] / (double)qi;
    }
    u64 v = (u64)vi;
    for (int j = 0; j < np; j++) {
        u64 pj = P[j];
        u128 acc = 0;
        for (int i = 0; i < nq; i++) acc += (u128)y[i] * m->qoverqimodp[(size_t)j * nq + i];
        /* lazy Montgomery reduction of the 128-bit sum (multSum :623-645) */
        u64 hi = (u64)(acc >> 64), lo = (u64)acc;
        u64 mq = (u64)(((u128)(lo * m->pinv_out[j]) * pj) >> 64);
        u64 r = hi - mq + pj + m->vtimesqmodp[(size_t)j * (nq + 1) + v];
        out[(size_t)j * ostride] = raw ? r : r % pj;
    }
}

static void modup_coeff(const u64 *x, int stride, const u64 *Q, int nq, const u64 *P, int np,
                        const mkref_modup *m, u64 *out, int ostride) {
    modup_coeff_impl(x, stride, Q, nq, P, np, m, out, ostride, 0);
}

void mkref_modup_exact(int N, const u64 *in, const u64 *Q, int nq,
                       u64 *out, const u64 *P, int np, const mkref_modup *m) {
    for (int c = 0; c < N; c++) modup_coeff(in + c, N, Q, nq, P, np, m, out + c, N);
}

void mkref_modup_exact_raw(int N, const u64 *in, const u64 *Q, int nq,
                           u64 *out, const u64 *P, int np, const mkref_modup *m) {
    for (int c = 0; c < N; c++) modup_coeff_impl(in + c, N, Q, nq, P, np, m, out + c, N, 1);
}

/* ModDownQPtoQ, basis_extension.go:192-232: out = (x_Q - ModUpPtoQ(x_P)) * P^-1 */
void mkref_moddown(int N, const u64 *xq, const u64 *Q, int nq,
                   const u64 *xp, const u64 *P, int np, u64 *outq) {
    mkref_modup m;
    mkref_modup_params(P, np, Q, nq, &m);
    u64 *pool = malloc(sizeof(u64) * (size_t)nq * N);
    mkref_modup_exact(N, xp, P, np, pool, Q, nq, &m);
    for (int i = 0; i < nq; i++) {
        u64 qi = Q[i];
        /* genModDownParams (:34-54): MForm(prod_j p_j^-1 mod q_i); the loop uses q - that */
        u64 pinv = 1;
        for (int j = 0; j < np; j++) pinv = mulmod(pinv, invmod(P[j] % qi, qi), qi);
        u64 negpinv_m = qi - mkref_mform(pinv, qi);
        u64 qinv = mkref_mred_params(qi);
        for (int c = 0; c < N; c++) {
            u64 x = xq[(size_t)i * N + c] % qi, y = pool[(size_t)i * N + c];
            outq[(size_t)i * N + c] = mkref_mred(y + 2 * qi - x, negpinv_m, qi, qinv);
        }
    }
    free(pool);
    mkref_modup_free(&m);
}

/* ------------------------------------------------------------------------ */
/* context                                                                    */
/* ------------------------------------------------------------------------ */

int mkref_ctx_init(mkref_ctx *c, int logN, const u64 *Q, int L, const u64 *P, int K,
                   int gamma, int root_rule) {
    memset(c, 0, sizeof(*c));
    c->logN = logN;
    c->N = 1 << logN;
    c->L = L;
    c->K = K;
    c->gamma = gamma;
    u64 qp[128];
    for (int i = 0; i < L; i++) qp[i] = Q[i];
    for (int i = 0; i < K; i++) qp[L + i] = P[i];
    return mkref_ring_init(&c->ringQP, logN, qp, L + K, root_rule);
}

void mkref_ctx_free(mkref_ctx *c) { mkref_ring_free(&c->ringQP); }

/* mkrlwe/params.go:63-71 */
int mkref_alpha(const mkref_ctx *c) { return c->K / c->gamma; }
int mkref_beta(const mkref_ctx *c, int levelQ) {
    int a = mkref_alpha(c);
    return (levelQ + 1 + a - 1) / a;
}

static const u64 *Qs(const mkref_ctx *c) { return c->ringQP.q; }
static const u64 *Ps(const mkref_ctx *c) { return c->ringQP.q + c->L; }
/* ring index of row r of a (levelQ+1+K)-row QP poly */
static int qp_ring_index(const mkref_ctx *c, int levelQ, int r) {
    return r <= levelQ ? r : c->L + (r - levelQ - 1);
}

/* ------------------------------------------------------------------------ */
/* gadget decomposition                                                        */
/* ------------------------------------------------------------------------ */

/*
 * Decomposer.DecomposeAndSplit (basis_extension.go:428-535) for digit `digit`,
 * followed by the NTT of every row (keyswitch.go:22-32).  The digit's integer
 * value a mod D (D = product of the digit's primes) is reconstructed by the
 * exact basis extension and written in every row of Q[0..levelQ] u P.
 *
 * The digit is formed from the limbs of a polynomial at level src_level >=
 * levelQ (digit i = limbs i alpha .. min(i alpha + alpha - 1, src_level)) while
 * the output has the rows of level levelQ: an operand decomposed at its own
 * level (mkckks/evaluator.go:433-439) and read at the product's lower level,
 * where MulAndRelinHoisted uses rows 0..level and the P rows of the first
 * Beta(level) digits (keyswitch_hoisted.go:84-123).  src_level = levelQ is the
 * ordinary decomposition.
 */
void mkref_decompose_digit_src(const mkref_ctx *c, int levelQ, int src_level, int digit, const u64 *a, u64 *out) {
    int N = c->N, K = c->K;
    int alpha = mkref_alpha(c);
    int rows = levelQ + 1 + K;
    int start = digit * alpha;
    int decompLvl;
    if (src_level > alpha * (digit + 1) - 1) decompLvl = alpha - 2;
    else decompLvl = (src_level % alpha) - 1;

    if (decompLvl == -1) {
        /* copy path (:443-451): the limb's residue is copied unreduced into every row */
        for (int r = 0; r < rows; r++) memcpy(out + (size_t)r * N, a + (size_t)start * N, sizeof(u64) * N);
    } else {
        int nd = decompLvl + 2;
        /* modUpParams[gamma*alpha-2][digit][decompLvl] = basisextenderparameters(
         * Q[start..start+nd), Q_all u P_all)  (:376-421) */
        u64 out_mod[128];
        int nall = c->L + K;
        for (int i = 0; i < c->L; i++) out_mod[i] = Qs(c)[i];
        for (int i = 0; i < K; i++) out_mod[c->L + i] = Ps(c)[i];
        mkref_modup m;
        mkref_modup_params(Qs(c) + start, nd, out_mod, nall, &m);
        u64 tmp[128];
        for (int x = 0; x < N; x++) {
            modup_coeff(a + (size_t)start * N + x, N, Qs(c) + start, nd, out_mod, nall, &m, tmp, 1);
            for (int r = 0; r <= levelQ; r++) out[(size_t)r * N + x] = tmp[r];
            for (int j = 0; j < K; j++) out[(size_t)(levelQ + 1 + j) * N + x] = tmp[c->L + j];
            /* digit rows keep the input residue (:487); the reference later recomputes
             * them from the same extension, which is congruent (D = 0 mod q_j there).
             * Rows of the digit above levelQ are not part of the output. */
            for (int i = 0; i < nd && start + i <= levelQ; i++)
                out[(size_t)(start + i) * N + x] = a[(size_t)(start + i) * N + x];
        }
        mkref_modup_free(&m);
    }
    for (int r = 0; r < rows; r++) {
        int ri = qp_ring_index(c, levelQ, r);
        u64 *row = out + (size_t)r * N;
        for (int x = 0; x < N; x++) row[x] %= c->ringQP.q[ri];
        mkref_ntt_limb(&c->ringQP, ri, row);
    }
}

void mkref_decompose_digit(const mkref_ctx *c, int levelQ, int digit, const u64 *a, u64 *out) {
    mkref_decompose_digit_src(c, levelQ, levelQ, digit, a, out);
}

/* KeySwitcher.Decompose at src_level, read at levelQ (see mkref_decompose_digit_src) */
void mkref_decompose_src(const mkref_ctx *c, int levelQ, int src_level, const u64 *a, u64 *out) {
    int beta = mkref_beta(c, levelQ);
    size_t dsz = (size_t)(levelQ + 1 + c->K) * c->N;
    for (int i = 0; i < beta; i++) mkref_decompose_digit_src(c, levelQ, src_level, i, a, out + i * dsz);
}

/* KeySwitcher.Decompose (keyswitch.go:50-76), input coefficient domain */
void mkref_decompose(const mkref_ctx *c, int levelQ, const u64 *a, u64 *out) {
    int beta = mkref_beta(c, levelQ);
    size_t dsz = (size_t)(levelQ + 1 + c->K) * c->N;
    for (int i = 0; i < beta; i++) mkref_decompose_digit(c, levelQ, i, a, out + i * dsz);
}

/* row r (of a levelQ-level QP poly) of a key entry stored at the key's full level */
static const u64 *key_row(const mkref_ctx *c, const u64 *key_digit, int levelQ, int r) {
    int ri = qp_ring_index(c, levelQ, r);
    return key_digit + (size_t)ri * c->N;
}

/* ExternalProductHoisted (keyswitch_hoisted.go:11-45) */
void mkref_external_product_hoisted(const mkref_ctx *c, int levelQ, const u64 *a_hoisted,
                                    const u64 *key, u64 *out) {
    mkref_external_product_n(c, levelQ, mkref_beta(c, levelQ), a_hoisted, key, out);
}

/* the same over ndig digits (the BFV products run over the 2 beta digits of an
 * R poly, mkbfv/keyswitch_hoisted.go:14-46) */
void mkref_external_product_n(const mkref_ctx *c, int levelQ, int beta, const u64 *a_hoisted,
                              const u64 *key, u64 *out) {
    int N = c->N, K = c->K, L = c->L;
    int rows = levelQ + 1 + K;
    size_t dsz = (size_t)rows * N, ksz = (size_t)(L + K) * N;
    u64 *acc = calloc((size_t)rows * N, sizeof(u64));
    for (int r = 0; r < rows; r++) {
        int ri = qp_ring_index(c, levelQ, r);
        u64 q = c->ringQP.q[ri], qinv = c->ringQP.qinv[ri];
        for (int b = 0; b < beta; b++) {
            const u64 *kr = key_row(c, key + b * ksz, levelQ, r);
            const u64 *ar = a_hoisted + b * dsz + (size_t)r * N;
            for (int x = 0; x < N; x++)
                acc[(size_t)r * N + x] = addmod(acc[(size_t)r * N + x], mkref_mred(kr[x], ar[x], q, qinv), q);
        }
        mkref_intt_limb(&c->ringQP, ri, acc + (size_t)r * N);
    }
    mkref_moddown(N, acc, Qs(c), levelQ + 1, acc + (size_t)(levelQ + 1) * N, Ps(c), K, out);
    free(acc);
}

/* ------------------------------------------------------------------------ */
/* KKLSS MK-CKKS mul + relin                                                   */
/* ------------------------------------------------------------------------ */

static int find_id(const int *ids, int k, int id) {
    for (int i = 0; i < k; i++) if (ids[i] == id) return i;
    return -1;
}

static int union_ids(const int *a, int ka, const int *b, int kb, int *out) {
    int n = 0;
    for (int i = 0; i < ka; i++) if (find_id(out, n, a[i]) < 0) out[n++] = a[i];
    for (int i = 0; i < kb; i++) if (find_id(out, n, b[i]) < 0) out[n++] = b[i];
    for (int i = 1; i < n; i++) {  /* ascending order */
        int v = out[i], j = i - 1;
        while (j >= 0 && out[j] > v) { out[j + 1] = out[j]; j--; }
        out[j + 1] = v;
    }
    return n;
}

/* x += MRed(key, D) over ndig digits of QP rows (MulCoeffsMontgomeryAndAddLvl).
 * x is a switching key like the reference's pools (keyswitch_hoisted.go:60-70):
 * full-level layout [ndig][L+K][N], rows addressed by ring index. */
void mkref_mac_keyed_n(const mkref_ctx *c, int levelQ, int ndig, const u64 *key, const u64 *D, u64 *x) {
    int N = c->N, rows = levelQ + 1 + c->K;
    size_t dsz = (size_t)rows * N, ksz = (size_t)(c->L + c->K) * N;
    for (int b = 0; b < ndig; b++)
        for (int r = 0; r < rows; r++) {
            int ri = qp_ring_index(c, levelQ, r);
            u64 q = c->ringQP.q[ri], qinv = c->ringQP.qinv[ri];
            const u64 *kr = key_row(c, key + b * ksz, levelQ, r);
            const u64 *dr = D + b * dsz + (size_t)r * N;
            u64 *xr = x + b * ksz + (size_t)ri * N;
            for (int i = 0; i < N; i++) xr[i] = addmod(xr[i], mkref_mred(kr[i], dr[i], q, qinv), q);
        }
}

static void mac_keyed(const mkref_ctx *c, int levelQ, const u64 *key, const u64 *D, u64 *x) {
    mkref_mac_keyed_n(c, levelQ, mkref_beta(c, levelQ), key, D, x);
}

/* MFormLvl over the used rows of a full-layout switching key */
void mkref_mform_qp(const mkref_ctx *c, int levelQ, u64 *x, int ndigits) {
    int N = c->N, rows = levelQ + 1 + c->K;
    size_t ksz = (size_t)(c->L + c->K) * N;
    for (int b = 0; b < ndigits; b++)
        for (int r = 0; r < rows; r++) {
            int ri = qp_ring_index(c, levelQ, r);
            u64 q = c->ringQP.q[ri];
            u64 *xr = x + b * ksz + (size_t)ri * N;
            for (int i = 0; i < N; i++) xr[i] = mkref_mform(xr[i], q);
        }
}

/*
 * mkckks.Evaluator.MulRelinNew (evaluator.go:419-446) -> MulRelinHoistedNew
 * (:561-569) -> mulRelinHoisted (:575-586) -> KeySwitcher.MulAndRelinHoisted
 * (keyswitch_hoisted.go:49-194).  The rescale (:585) is separate.
 */
/*
 * Operands at their own levels lv0, lv1 >= level (components of lv + 1 limbs):
 * each is decomposed at its own level (evaluator.go:433-439) and the product is
 * taken at `level` (:577-581), which reads the operands' limbs 0..level.
 */
int mkref_kklss_mulrelin_lv(const mkref_ctx *c, int level, int lv0, int lv1,
                            const u64 *op0_in, const int *ids0, int k0,
                            const u64 *op1_in, const int *ids1, int k1,
                            const u64 *const *rlk_b, const u64 *const *rlk_d,
                            const u64 *const *rlk_v, const u64 *crs_u,
                            u64 *out, int *ids_out) {
    int N = c->N, K = c->K;
    int nl = level + 1, rows = nl + K, beta = mkref_beta(c, level);
    size_t psz = (size_t)nl * N, dsz = (size_t)rows * N, hsz = (size_t)beta * dsz;
    size_t psz0 = (size_t)(lv0 + 1) * N, psz1 = (size_t)(lv1 + 1) * N;
    int ku = union_ids(ids0, k0, ids1, k1, ids_out);
    const mkref_ring *R = &c->ringQP;
    /* the operands' limbs 0..level, contiguous */
    u64 *op0 = malloc(sizeof(u64) * psz * (1 + k0)), *op1 = malloc(sizeof(u64) * psz * (1 + k1));
    for (int i = 0; i <= k0; i++) memcpy(op0 + (size_t)i * psz, op0_in + (size_t)i * psz0, sizeof(u64) * psz);
    for (int j = 0; j <= k1; j++) memcpy(op1 + (size_t)j * psz, op1_in + (size_t)j * psz1, sizeof(u64) * psz);

    u64 *D0 = malloc(sizeof(u64) * hsz * (k0 ? k0 : 1));
    u64 *D1 = malloc(sizeof(u64) * hsz * (k1 ? k1 : 1));
    size_t xsz = (size_t)beta * (c->L + K) * N;   /* full-level switching keys */
    u64 *x = calloc(xsz, sizeof(u64)), *y = calloc(xsz, sizeof(u64));
    u64 *pool0 = malloc(sizeof(u64) * psz), *pool1 = malloc(sizeof(u64) * psz);
    u64 *pool2 = malloc(sizeof(u64) * psz), *t = malloc(sizeof(u64) * psz);
    u64 *Dt = malloc(sizeof(u64) * hsz);

    /* hoisted decompositions (evaluator.go:432-438) */
    for (int i = 0; i < k0; i++) mkref_decompose_src(c, level, lv0, op0_in + (size_t)(1 + i) * psz0, D0 + i * hsz);
    for (int j = 0; j < k1; j++) mkref_decompose_src(c, level, lv1, op1_in + (size_t)(1 + j) * psz1, D1 + j * hsz);

    /* x = sum d_i (.) D(op0_i), y = sum b_j (.) D(op1_j), then MForm (:84-123) */
    for (int i = 0; i < k0; i++) mac_keyed(c, level, rlk_d[ids0[i]], D0 + i * hsz, x);
    mkref_mform_qp(c, level, x, beta);
    for (int j = 0; j < k1; j++) mac_keyed(c, level, rlk_b[ids1[j]], D1 + j * hsz, y);
    mkref_mform_qp(c, level, y, beta);

    /* tensor (:126-153) */
    memset(out, 0, sizeof(u64) * psz * (1 + ku));
    memcpy(pool0, op0, sizeof(u64) * psz);
    memcpy(pool1, op1, sizeof(u64) * psz);
    for (int l = 0; l < nl; l++) {
        u64 q = R->q[l], qinv = R->qinv[l];
        mkref_ntt_limb(R, l, pool0 + (size_t)l * N);
        mkref_ntt_limb(R, l, pool1 + (size_t)l * N);
        u64 *p0 = pool0 + (size_t)l * N, *p1 = pool1 + (size_t)l * N, *o = out + (size_t)l * N;
        for (int i = 0; i < N; i++) {
            p0[i] = mkref_mform(p0[i], q);
            o[i] = mkref_mred(p0[i], p1[i], q, qinv);
            p1[i] = mkref_mform(p1[i], q);
        }
    }
    for (int i = 0; i < k0; i++) {
        int u = find_id(ids_out, ku, ids0[i]);
        u64 *o = out + (size_t)(1 + u) * psz;
        memcpy(pool2, op0 + (size_t)(1 + i) * psz, sizeof(u64) * psz);
        for (int l = 0; l < nl; l++) {
            mkref_ntt_limb(R, l, pool2 + (size_t)l * N);
            for (int x2 = 0; x2 < N; x2++)
                o[(size_t)l * N + x2] = mkref_mred(pool1[(size_t)l * N + x2], pool2[(size_t)l * N + x2], R->q[l], R->qinv[l]);
        }
    }
    for (int j = 0; j < k1; j++) {
        int u = find_id(ids_out, ku, ids1[j]);
        int in0 = find_id(ids0, k0, ids1[j]) >= 0;
        u64 *o = out + (size_t)(1 + u) * psz;
        memcpy(pool2, op1 + (size_t)(1 + j) * psz, sizeof(u64) * psz);
        for (int l = 0; l < nl; l++) {
            mkref_ntt_limb(R, l, pool2 + (size_t)l * N);
            for (int x2 = 0; x2 < N; x2++) {
                u64 v = mkref_mred(pool0[(size_t)l * N + x2], pool2[(size_t)l * N + x2], R->q[l], R->qinv[l]);
                o[(size_t)l * N + x2] = in0 ? addmod(o[(size_t)l * N + x2], v, R->q[l]) : v;
            }
        }
    }
    for (int u = 0; u <= ku; u++)
        for (int l = 0; l < nl; l++) mkref_intt_limb(R, l, out + (size_t)u * psz + (size_t)l * N);

    /* ctOut_j += Ext(D(op1_j), x) (:155-165) */
    for (int j = 0; j < k1; j++) {
        int u = find_id(ids_out, ku, ids1[j]);
        mkref_external_product_hoisted(c, level, D1 + j * hsz, x, t);
        u64 *o = out + (size_t)(1 + u) * psz;
        for (size_t e = 0; e < psz; e++) o[e] = addmod(o[e], t[e], R->q[e / N]);
    }
    /* nested products (:167-193) */
    for (int i = 0; i < k0; i++) {
        int u = find_id(ids_out, ku, ids0[i]);
        mkref_external_product_hoisted(c, level, D0 + i * hsz, y, t);
        mkref_decompose(c, level, t, Dt);
        mkref_external_product_hoisted(c, level, Dt, rlk_v[ids0[i]], pool2);
        for (size_t e = 0; e < psz; e++) out[e] = addmod(out[e], pool2[e], R->q[e / N]);
        mkref_external_product_hoisted(c, level, Dt, crs_u, pool2);
        u64 *o = out + (size_t)(1 + u) * psz;
        for (size_t e = 0; e < psz; e++) o[e] = addmod(o[e], pool2[e], R->q[e / N]);
    }

    free(D0); free(D1); free(x); free(y); free(pool0); free(pool1); free(pool2); free(t); free(Dt);
    free(op0); free(op1);
    return ku + 1;
}

int mkref_kklss_mulrelin(const mkref_ctx *c, int level,
                         const u64 *op0, const int *ids0, int k0,
                         const u64 *op1, const int *ids1, int k1,
                         const u64 *const *rlk_b, const u64 *const *rlk_d,
                         const u64 *const *rlk_v, const u64 *crs_u,
                         u64 *out, int *ids_out) {
    return mkref_kklss_mulrelin_lv(c, level, level, level, op0, ids0, k0, op1, ids1, k1, rlk_b, rlk_d, rlk_v,
                                   crs_u, out, ids_out);
}

/* ------------------------------------------------------------------------ */
/* rescale                                                                     */
/* ------------------------------------------------------------------------ */

/* scale loop of mkckks Evaluator.Rescale (evaluator.go:383-388) */
int mkref_rescale_count(const u64 *Q, int level, double scale, double min_scale, double *new_scale) {
    int nb = 0;
    double s = scale;
    while (level - nb >= 0 && s / (double)Q[level - nb] >= min_scale / 2) {
        s /= (double)Q[level - nb];
        nb++;
    }
    *new_scale = s;
    return nb;
}

/* Lattigo DivRoundByLastModulusLvl (called from evaluator.go:391), restated in
 * par1_kklss_ckks_kernel.cpp:227-246: out_i = (x_i + h - [x_L + h]_{q_L}) q_L^-1 */
static void div_round_last_once(const mkref_ctx *c, int level, u64 *poly) {
    int N = c->N;
    const u64 *Q = Qs(c);
    u64 qL = Q[level], h = qL >> 1;
    u64 *last = poly + (size_t)level * N;
    for (int x = 0; x < N; x++) last[x] = addmod(last[x], h, qL);
    for (int i = 0; i < level; i++) {
        u64 qi = Q[i];
        u64 inv = invmod(qL % qi, qi);
        u64 hi = h % qi;
        u64 *p = poly + (size_t)i * N;
        for (int x = 0; x < N; x++) {
            u64 tmod = last[x] % qi;
            u64 num = submod(addmod(p[x], hi, qi), tmod, qi);
            p[x] = mulmod(num, inv, qi);
        }
    }
}

void mkref_div_round_last(const mkref_ctx *c, int level, int n, u64 *poly) {
    for (int k = 0; k < n; k++) div_round_last_once(c, level - k, poly);
}

/* NTT-domain variant (Lattigo DivRoundByLastModulusManyNTTLvl; par1_mkf_ckks_kernel.cpp:415-453) */
void mkref_div_round_last_ntt(const mkref_ctx *c, int level, int n, u64 *poly) {
    const mkref_ring *R = &c->ringQP;
    int N = c->N;
    for (int l = 0; l <= level; l++) mkref_intt_limb(R, l, poly + (size_t)l * N);
    mkref_div_round_last(c, level, n, poly);
    for (int l = 0; l <= level - n; l++) mkref_ntt_limb(R, l, poly + (size_t)l * N);
}

/* ------------------------------------------------------------------------ */
/* MKF (mkf_cpu)                                                                */
/* ------------------------------------------------------------------------ */

/*
 * switchOne (mkf_cpu/common.go:79-107): Lattigo rlwe.KeySwitcher.DecomposeNTT,
 * KeyswitchHoistedNoModDown, then ModDownQPtoQNTT (nttOutput) or
 * InvNTTLazy + ModDownQPtoQ.  Switching key layout [digit][output][L+K][N].
 */
void mkref_mkf_switch_one(const mkref_ctx *c, int levelQ, const u64 *in, int in_ntt,
                          const u64 *swk, int ntt_output, u64 *out0, u64 *out1) {
    const mkref_ring *R = &c->ringQP;
    int N = c->N, K = c->K, L = c->L;
    int nl = levelQ + 1, rows = nl + K, beta = mkref_beta(c, levelQ);
    size_t psz = (size_t)nl * N, dsz = (size_t)rows * N, ksz = (size_t)(L + K) * N;
    u64 *coef = malloc(sizeof(u64) * psz);
    memcpy(coef, in, sizeof(u64) * psz);
    if (in_ntt)
        for (int l = 0; l < nl; l++) mkref_intt_limb(R, l, coef + (size_t)l * N);
    u64 *D = malloc(sizeof(u64) * beta * dsz);
    for (int b = 0; b < beta; b++) mkref_decompose_digit(c, levelQ, b, coef, D + b * dsz);
    u64 *acc[2];
    for (int o = 0; o < 2; o++) {
        acc[o] = calloc(dsz, sizeof(u64));
        for (int b = 0; b < beta; b++)
            for (int r = 0; r < rows; r++) {
                int ri = qp_ring_index(c, levelQ, r);
                u64 q = R->q[ri], qinv = R->qinv[ri];
                const u64 *kr = swk + ((size_t)b * 2 + o) * ksz + (size_t)ri * N;
                const u64 *dr = D + b * dsz + (size_t)r * N;
                u64 *ar = acc[o] + (size_t)r * N;
                for (int x = 0; x < N; x++) ar[x] = addmod(ar[x], mkref_mred(kr[x], dr[x], q, qinv), q);
            }
        /* INTT every row, ModDown in the coefficient domain; for NTT output the Q
         * part is transformed back (ModDownQPtoQNTT computes the same residues:
         * (x - NTT(ModUp(INTT(x_P)))) P^-1 == NTT((INTT(x) - ModUp(INTT(x_P))) P^-1)) */
        for (int r = 0; r < rows; r++) mkref_intt_limb(R, qp_ring_index(c, levelQ, r), acc[o] + (size_t)r * N);
        u64 *dst = o == 0 ? out0 : out1;
        mkref_moddown(N, acc[o], Qs(c), nl, acc[o] + psz, Ps(c), K, dst);
        if (ntt_output)
            for (int l = 0; l < nl; l++) mkref_ntt_limb(R, l, dst + (size_t)l * N);
        free(acc[o]);
    }
    free(coef);
    free(D);
}

/* CKKSCase.Run (mkf_cpu/ckks.go:102-139) up to, not including, the rescale;
 * also paperMKFCKKSEvaluator.MulRelinNew (true_mkf_ckks.go:253-303), which folds
 * each operand at its own level (common.go:109-126: switchOne at in.Level) and
 * forms the tensor at the smaller level.  out0/out1 have min(level_l, level_r)+1
 * limbs. */
void mkref_mkf_ckks_mulrelin2(const mkref_ctx *c, int level_l, int level_r, int k,
                              const u64 *left, const u64 *right,
                              const u64 *const *fold_keys, const u64 *relin_key,
                              u64 *out0, u64 *out1) {
    const mkref_ring *R = &c->ringQP;
    int N = c->N, lv[2] = {level_l, level_r};
    int level = level_l < level_r ? level_l : level_r, nl = level + 1;
    int nlmax = (level_l > level_r ? level_l : level_r) + 1;
    size_t psz = (size_t)nl * N, pmax = (size_t)nlmax * N;
    u64 *lf[2], *rf[2], *s0 = malloc(sizeof(u64) * pmax), *s1 = malloc(sizeof(u64) * pmax);
    const u64 *side[2] = {left, right};
    u64 **dst[2] = {lf, rf};
    for (int s = 0; s < 2; s++) {
        /* fold (common.go:109-126) at the operand's level */
        size_t ps = (size_t)(lv[s] + 1) * N;
        dst[s][0] = malloc(sizeof(u64) * ps);
        dst[s][1] = calloc(ps, sizeof(u64));
        memcpy(dst[s][0], side[s], sizeof(u64) * ps);
        for (int i = 0; i < k; i++) {
            mkref_mkf_switch_one(c, lv[s], side[s] + (size_t)(1 + i) * ps, 1, fold_keys[i], 1, s0, s1);
            for (size_t e = 0; e < ps; e++) {
                u64 q = R->q[e / N];
                dst[s][0][e] = addmod(dst[s][0][e], s0[e], q);
                dst[s][1][e] = addmod(dst[s][1][e], s1[e], q);
            }
        }
    }
    /* Karatsuba tensor (ckks.go:120-131) on limbs 0..level */
    u64 *mu2 = malloc(sizeof(u64) * psz);
    for (size_t e = 0; e < psz; e++) {
        int l = (int)(e / N);
        u64 q = R->q[l], qinv = R->qinv[l];
        u64 m0 = mkref_mred(mkref_mform(lf[0][e], q), rf[0][e], q, qinv);
        u64 m2 = mkref_mred(mkref_mform(lf[1][e], q), rf[1][e], q, qinv);
        u64 ls = mkref_mform(addmod(lf[0][e], lf[1][e], q), q);
        u64 rs = addmod(rf[0][e], rf[1][e], q);
        u64 m1 = mkref_mred(ls, rs, q, qinv);
        m1 = submod(submod(m1, m0, q), m2, q);
        out0[e] = m0;
        out1[e] = m1;
        mu2[e] = m2;
    }
    /* relinearize mu2 once (ckks.go:134-138) */
    mkref_mkf_switch_one(c, level, mu2, 1, relin_key, 1, s0, s1);
    for (size_t e = 0; e < psz; e++) {
        u64 q = R->q[e / N];
        out0[e] = addmod(out0[e], s0[e], q);
        out1[e] = addmod(out1[e], s1[e], q);
    }
    free(mu2); free(s0); free(s1);
    for (int s = 0; s < 2; s++) { free(dst[s][0]); free(dst[s][1]); }
}

void mkref_mkf_ckks_mulrelin(const mkref_ctx *c, int level, int k,
                             const u64 *left, const u64 *right,
                             const u64 *const *fold_keys, const u64 *relin_key,
                             u64 *out0, u64 *out1) {
    mkref_mkf_ckks_mulrelin2(c, level, level, k, left, right, fold_keys, relin_key, out0, out1);
}

/* ------------------------------------------------------------------------ */
/* PRNG and samplers (build-side convention; the reference uses a crypto PRNG) */
/* ------------------------------------------------------------------------ */

static u64 splitmix64(u64 *x) {
    u64 z = (*x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void mkref_prng_seed(mkref_prng *g, u64 seed) {
    u64 x = seed;
    for (int i = 0; i < 4; i++) g->s[i] = splitmix64(&x);
}

static u64 rotl(u64 x, int k) { return (x << k) | (x >> (64 - k)); }

u64 mkref_prng_next(mkref_prng *g) {
    u64 *s = g->s;
    u64 r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}

void mkref_sample_uniform(mkref_prng *g, const u64 *q, int nlimbs, int N, u64 *out) {
    for (int l = 0; l < nlimbs; l++) {
        u64 qi = q[l], mask = 1;
        while (mask < qi) mask = (mask << 1) | 1;
        for (int i = 0; i < N; i++) {
            u64 v;
            do { v = mkref_prng_next(g) & mask; } while (v >= qi);
            out[(size_t)l * N + i] = v;
        }
    }
}

static double uniform01(mkref_prng *g) { return (double)(mkref_prng_next(g) >> 11) * (1.0 / 9007199254740992.0); }

void mkref_sample_ternary(mkref_prng *g, int N, double p_zero, int64_t *out) {
    for (int i = 0; i < N; i++) {
        double u = uniform01(g);
        if (u < p_zero) out[i] = 0;
        else out[i] = (mkref_prng_next(g) & 1) ? 1 : -1;
    }
}

void mkref_sample_gaussian(mkref_prng *g, int N, double sigma, int bound, int64_t *out) {
    for (int i = 0; i < N; i++) {
        for (;;) {
            double u1 = uniform01(g), u2 = uniform01(g);
            if (u1 <= 0) continue;
            double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2) * sigma;
            long v = lround(z);
            if (v > bound || v < -bound) continue;
            out[i] = v;
            break;
        }
    }
}

void mkref_small_to_rns(const int64_t *s, int N, const u64 *q, int n, u64 *out) {
    for (int l = 0; l < n; l++)
        for (int i = 0; i < N; i++) {
            int64_t v = s[i];
            out[(size_t)l * N + i] = v >= 0 ? (u64)v % q[l] : q[l] - ((u64)(-v) % q[l]);
            if (out[(size_t)l * N + i] == q[l]) out[(size_t)l * N + i] = 0;
        }
}

/* ------------------------------------------------------------------------ */
/* heap handles for the ctypes test harness                                    */
/* ------------------------------------------------------------------------ */

mkref_ctx *mkref_ctx_new(int logN, const u64 *Q, int L, const u64 *P, int K, int gamma, int root_rule) {
    mkref_ctx *c = malloc(sizeof(mkref_ctx));
    if (!c) return NULL;
    if (mkref_ctx_init(c, logN, Q, L, P, K, gamma, root_rule) != 0) {
        free(c);
        return NULL;
    }
    return c;
}

void mkref_ctx_delete(mkref_ctx *c) {
    if (!c) return;
    mkref_ctx_free(c);
    free(c);
}

void mkref_ctx_tables(const mkref_ctx *c, int limb, u64 *psi, u64 *psiinv, u64 *root, u64 *ninv) {
    const mkref_ring *R = &c->ringQP;
    if (psi) memcpy(psi, R->psi + (size_t)limb * R->N, sizeof(u64) * R->N);
    if (psiinv) memcpy(psiinv, R->psiinv + (size_t)limb * R->N, sizeof(u64) * R->N);
    if (root) *root = R->root[limb];
    if (ninv) *ninv = R->ninv[limb];
}

void mkref_ctx_ntt(const mkref_ctx *c, int limb, u64 *a, int inverse) {
    if (inverse) mkref_intt_limb(&c->ringQP, limb, a);
    else mkref_ntt_limb(&c->ringQP, limb, a);
}

void mkref_modup_tables(const u64 *Q, int nq, const u64 *P, int np, u64 *a, u64 *b, u64 *v) {
    mkref_modup m;
    mkref_modup_params(Q, nq, P, np, &m);
    memcpy(a, m.qoverqiinvqi, sizeof(u64) * nq);
    memcpy(b, m.qoverqimodp, sizeof(u64) * nq * np);
    memcpy(v, m.vtimesqmodp, sizeof(u64) * (nq + 1) * np);
    mkref_modup_free(&m);
}

void mkref_modup_exact_simple(int N, const u64 *in, const u64 *Q, int nq, u64 *out, const u64 *P, int np) {
    mkref_modup m;
    mkref_modup_params(Q, nq, P, np, &m);
    mkref_modup_exact(N, in, Q, nq, out, P, np, &m);
    mkref_modup_free(&m);
}
