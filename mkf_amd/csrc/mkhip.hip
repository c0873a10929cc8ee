// Host side of the C ABI (include/mkhip.h): parameter tables, resident keys,
// KKLSS MK-CKKS mul+relin plans (descriptor tables + workspace + hipGraph).
//
// The orchestration follows mkrlwe.KeySwitcher.MulAndRelinHoisted
// (software/kklss/mkrlwe/keyswitch_hoisted.go:49-194) preceded by the hoisted
// decompositions of mkckks.Evaluator.MulRelinNew (mkckks/evaluator.go:419-446)
// and followed by Evaluator.Rescale (:360-401), as eight batched launches:
//
//   1 k_decomp    D(op0_i), D(op1_j) for every party, x = MForm(sum d_i D0_i),
//                 y = MForm(sum b_j D1_j)                                 (:84-123)
//   2 k_ntt       NTT of every input component                           (:126-150)
//   3 k_tensor    c0 c0', c0 c'_j + c_j c0', INTT                          (:126-153)
//   4 k_ext       Ext(D1_j, x) and Ext(D0_i, y) MACs + INTT               (:155-170)
//   5 k_moddown   ctOut_j += ModDown(.), t_i = ModDown(.)
//   6 k_decomp    D(t_i)                                                  (:175)
//   7 k_ext       Ext(D(t_i), v_i) and Ext(D(t_i), u) + INTT               (:177-193)
//   8 k_moddown   ctOut_0 += sum_i ModDown(.), ctOut_i += ModDown(.), rescale
#include "../../include/mkhip.h"
#include "tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

namespace mkhip {
hipError_t launch_kernel(int what, const DevTables &T, const void *descs, int n, int a, int b, hipStream_t st,
                         int c = 1);
}

using namespace mkhip;

namespace {

typedef unsigned __int128 u128;

// ---- number theory on the host -------------------------------------------

u64 bitrev(u64 x, int bits) {
    u64 r = 0;
    for (int i = 0; i < bits; ++i) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

bool is_prime64(u64 n) {
    if (n < 2) return false;
    static const u64 small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (u64 p : small)
        if (n % p == 0) return n == p;
    u64 d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        ++s;
    }
    for (u64 a : small) {
        u64 x = h_powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s; ++r) {
            x = h_mulmod(x, x, n);
            if (x == n - 1) {
                comp = false;
                break;
            }
        }
        if (comp) return false;
    }
    return true;
}

std::vector<u64> distinct_factors(u64 n) {
    std::vector<u64> f;
    if ((n & 1) == 0) {
        f.push_back(2);
        while ((n & 1) == 0) n >>= 1;
    }
    for (u64 p = 3; p * p <= n; p += 2) {
        if (n % p == 0) {
            f.push_back(p);
            while (n % p == 0) n /= p;
        }
        if (n > 1 && is_prime64(n)) break;
    }
    if (n > 1) f.push_back(n);
    return f;
}

// Smallest generator of Z_q^*, starting at 3 (Lattigo v2.3.0 ring.NewRing's search;
// flag MKHIP_ROOT_RULE_SMALLEST2 starts at 2).
u64 primitive_root(u64 q, bool from2) {
    std::vector<u64> f = distinct_factors(q - 1);
    for (u64 g = from2 ? 2 : 3;; ++g) {
        bool ok = true;
        for (u64 p : f)
            if (h_powmod(g, (q - 1) / p, q) == 1) {
                ok = false;
                break;
            }
        if (ok) return g;
    }
}

// basisextenderparameters (basis_extension.go:83-153) for input basis Qin and
// output moduli Pout: y[i] = MForm((Q/q_i)^-1 mod q_i), qstar[j][i] = MForm(Q/q_i mod p_j),
// vt[j][v] = v * (-Q) mod p_j.
void modup_params(const std::vector<u64> &Qin, const std::vector<u64> &Pout, std::vector<u64> &y,
                  std::vector<u64> &qstar, std::vector<u64> &vt) {
    const size_t n = Qin.size(), m = Pout.size();
    y.assign(n, 0);
    qstar.assign(m * n, 0);
    vt.assign(m * (n + 1), 0);
    for (size_t i = 0; i < n; ++i) {
        u64 qi = Qin[i], s = 1;
        for (size_t j = 0; j < n; ++j)
            if (j != i) s = h_mulmod(s, Qin[j] % qi, qi);
        y[i] = h_mform(h_inv(s, qi), qi);
        for (size_t j = 0; j < m; ++j) {
            u64 pj = Pout[j], t = 1;
            for (size_t u = 0; u < n; ++u)
                if (u != i) t = h_mulmod(t, Qin[u] % pj, pj);
            qstar[j * n + i] = h_mform(t, pj);
        }
    }
    for (size_t j = 0; j < m; ++j) {
        u64 pj = Pout[j], qm = 1;
        for (size_t i = 0; i < n; ++i) qm = h_mulmod(qm, Qin[i] % pj, pj);
        u64 v = (pj - qm) % pj;
        for (size_t i = 1; i <= n; ++i) vt[j * (n + 1) + i] = (vt[j * (n + 1) + i - 1] + v) % pj;
    }
}

int hip_status(hipError_t e) { return e == hipSuccess ? MKHIP_OK : (e == hipErrorOutOfMemory ? MKHIP_ERR_OOM : MKHIP_ERR_DEVICE); }

#define HIP_TRY(expr)                                  \
    do {                                               \
        hipError_t _e = (expr);                        \
        if (_e != hipSuccess) return hip_status(_e);   \
    } while (0)

}  // namespace

// ---------------------------------------------------------------------------

struct PartyKeys {
    u64 *b = nullptr, *d = nullptr, *v = nullptr;
};

struct mkhip_ctx {
    int logN = 0, N = 0, L = 0, K = 0, gamma = 1, alpha = 0, beta_max = 0;
    u64 t = 0;
    int flags = 0;
    std::vector<u64> Q, P;
    void *d_tables = nullptr;
    DevTables T;
    hipStream_t stream = nullptr;
    u64 *d_crs_u = nullptr;
    std::map<int, PartyKeys> parties;
    size_t swk_words = 0;
    mkhip_plan *single = nullptr;  // cached plan of the host-pointer entry point
    u64 *single_io = nullptr;      // its device operand buffers
    size_t single_io_words = 0;
};

struct mkhip_plan {
    mkhip_ctx *ctx = nullptr;
    void *d_mem = nullptr;  // descriptors + workspace
    size_t bytes = 0;
    int beta = 0, rows = 0, nl = 0;
    // stage descriptor offsets (device pointers) and counts
    const void *dec1 = nullptr, *mac = nullptr, *ntt = nullptr, *ten = nullptr, *ext1 = nullptr, *md1 = nullptr,
               *dec2 = nullptr, *ext2 = nullptr, *md2 = nullptr;
    int n_dec1 = 0, n_mac = 0, n_ntt = 0, n_ten = 0, n_ext1 = 0, n_md1 = 0, n_dec2 = 0, n_ext2 = 0, n_md2 = 0;
    int kmax1 = 1, kmax2 = 1;  // largest party count of the decompositions of stage 0 / 6
    bool graph = false;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    int sig[4 + 2 * MKHIP_MAX_PARTIES] = {0};  // shape signature (single-op cache)
};

extern "C" const char *mkhip_status_string(int s) {
    switch (s) {
    case MKHIP_OK: return "ok";
    case MKHIP_ERR_ARG: return "invalid argument";
    case MKHIP_ERR_PARAMS: return "invalid parameters (moduli not NTT-friendly for N, or unsupported N)";
    case MKHIP_ERR_GADGET: return "gadget decomposition impossible: alpha = K/gamma is 0";
    case MKHIP_ERR_UNKNOWN_PARTY: return "there is no relinearization key with given id";
    case MKHIP_ERR_LEVEL: return "operands and output have different levels";
    case MKHIP_ERR_NO_KEY: return "CRS or evaluation key missing";
    case MKHIP_ERR_OOM: return "device memory exhausted";
    case MKHIP_ERR_DEVICE: return "HIP runtime error";
    case MKHIP_ERR_SCALE: return "cannot Rescale (minScale <= 0, scale 0 or level 0)";
    case MKHIP_ERR_SCHEME: return "operation not valid for this context";
    case MKHIP_ERR_TOO_MANY: return "too many parties";
    default: return "unknown status";
    }
}

extern "C" int mkhip_ctx_create(mkhip_ctx **out, int logN, const uint64_t *q, int L, const uint64_t *p, int K,
                                int gamma, uint64_t t, const uint64_t *qmul, int flags) {
    (void)qmul;
    if (!out || !q || (K > 0 && !p) || L <= 0 || K < 0 || gamma <= 0) return MKHIP_ERR_ARG;
    *out = nullptr;
    if (logN != 13 && logN != 14) return MKHIP_ERR_PARAMS;
    if (L > 16 || K > 4 || L + K > kMaxModuli) return MKHIP_ERR_PARAMS;
    const int N = 1 << logN;
    for (int i = 0; i < L + K; ++i) {
        u64 m = i < L ? q[i] : p[i - L];
        if (m >= (1ULL << 61) || ((m - 1) % (2 * (u64)N)) != 0 || !is_prime64(m)) return MKHIP_ERR_PARAMS;
    }
    const int alpha = K / gamma;
    if (K > 0 && alpha == 0) return MKHIP_ERR_GADGET;
    if (alpha > kMaxAlpha) return MKHIP_ERR_PARAMS;

    mkhip_ctx *c = new mkhip_ctx();
    c->logN = logN;
    c->N = N;
    c->L = L;
    c->K = K;
    c->gamma = gamma;
    c->alpha = alpha;
    c->beta_max = alpha ? (L + alpha - 1) / alpha : 0;
    c->t = t;
    c->flags = flags;
    c->Q.assign(q, q + L);
    c->P.assign(p, p + K);
    c->swk_words = (size_t)c->beta_max * (L + K) * N;
    const int nmod = L + K;
    std::vector<u64> QP(c->Q);
    QP.insert(QP.end(), c->P.begin(), c->P.end());

    // ---- host tables ----
    std::vector<ModConst> mods(nmod);
    std::vector<ulonglong2> tw((size_t)nmod * N), itw((size_t)nmod * N);
    for (int i = 0; i < nmod; ++i) {
        const u64 qi = QP[i];
        ModConst &m = mods[i];
        m.q = qi;
        m.qinv = h_qinv(qi);
        m.mu = (u64)((((u128)1) << 64) / qi);
        m.r2 = (u64)((((u128)1 << 64) % qi) * (u128)(((u128)1 << 64) % qi) % qi);
        m.ninv = h_inv((u64)N, qi);
        m.ninvp = h_shoup_companion(m.ninv, qi);
        const u64 g = primitive_root(qi, (flags & MKHIP_ROOT_RULE_SMALLEST2) != 0);
        const u64 psi = h_powmod(g, (qi - 1) / (2 * (u64)N), qi);
        const u64 psiinv = h_inv(psi, qi);
        u64 a = 1, b = 1;
        for (int j = 0; j < N; ++j) {
            const u64 idx = bitrev((u64)j, logN);
            tw[(size_t)i * N + idx] = make_ulonglong2(a, h_shoup_companion(a, qi));
            itw[(size_t)i * N + idx] = make_ulonglong2(b, h_shoup_companion(b, qi));
            a = h_mulmod(a, psi, qi);
            b = h_mulmod(b, psiinv, qi);
        }
    }
    // ModDown P -> Q
    std::vector<u64> md_y, md_qstar, md_vt, md_pinv(L);
    if (K > 0) {
        modup_params(c->P, c->Q, md_y, md_qstar, md_vt);
        for (int j = 0; j < L; ++j) {
            u64 pinv = 1;
            for (int k = 0; k < K; ++k) pinv = h_mulmod(pinv, h_inv(c->P[k] % c->Q[j], c->Q[j]), c->Q[j]);
            md_pinv[j] = (c->Q[j] - h_mform(pinv, c->Q[j])) % c->Q[j];
        }
    }
    // gadget digits with nd >= 2 primes
    const int beta = c->beta_max > 0 ? c->beta_max : 1;
    const size_t ne = (size_t)beta * (kMaxAlpha + 1);
    std::vector<u64> dc_y(ne * kMaxAlpha, 0), dc_qstar(ne * nmod * kMaxAlpha, 0), dc_vt(ne * nmod * (kMaxAlpha + 1), 0);
    for (int i = 0; i < c->beta_max; ++i) {
        for (int nd = 2; nd <= alpha; ++nd) {
            if (i * alpha + nd > L) continue;
            std::vector<u64> Qin(c->Q.begin() + i * alpha, c->Q.begin() + i * alpha + nd), y, qs, vt;
            modup_params(Qin, QP, y, qs, vt);
            const size_t e = (size_t)i * (kMaxAlpha + 1) + nd;
            for (int l = 0; l < nd; ++l) dc_y[e * kMaxAlpha + l] = y[l];
            for (int r = 0; r < nmod; ++r) {
                for (int l = 0; l < nd; ++l) dc_qstar[(e * nmod + r) * kMaxAlpha + l] = qs[(size_t)r * nd + l];
                for (int v = 0; v <= nd; ++v) dc_vt[(e * nmod + r) * (kMaxAlpha + 1) + v] = vt[(size_t)r * (nd + 1) + v];
            }
        }
    }
    // rescale
    std::vector<u64> rs_inv((size_t)L * L, 0), rs_half((size_t)L * L, 0);
    for (int l = 1; l < L; ++l)
        for (int j = 0; j < l; ++j) {
            rs_inv[(size_t)l * L + j] = h_mform(h_inv(c->Q[l] % c->Q[j], c->Q[j]), c->Q[j]);
            rs_half[(size_t)l * L + j] = (c->Q[l] >> 1) % c->Q[j];
        }

    // ---- one device allocation ----
    struct Blob {
        std::vector<char> host;
        size_t add(const void *src, size_t bytes) {
            size_t off = (host.size() + 255) & ~(size_t)255;
            host.resize(off + bytes);
            if (bytes) std::memcpy(host.data() + off, src, bytes);
            return off;
        }
    } blob;
    size_t o_mod = blob.add(mods.data(), mods.size() * sizeof(ModConst));
    size_t o_tw = blob.add(tw.data(), tw.size() * sizeof(ulonglong2));
    size_t o_itw = blob.add(itw.data(), itw.size() * sizeof(ulonglong2));
    size_t o_mdy = blob.add(md_y.data(), md_y.size() * 8);
    size_t o_mdq = blob.add(md_qstar.data(), md_qstar.size() * 8);
    size_t o_mdv = blob.add(md_vt.data(), md_vt.size() * 8);
    size_t o_mdp = blob.add(md_pinv.data(), md_pinv.size() * 8);
    size_t o_dcy = blob.add(dc_y.data(), dc_y.size() * 8);
    size_t o_dcq = blob.add(dc_qstar.data(), dc_qstar.size() * 8);
    size_t o_dcv = blob.add(dc_vt.data(), dc_vt.size() * 8);
    size_t o_rsi = blob.add(rs_inv.data(), rs_inv.size() * 8);
    size_t o_rsh = blob.add(rs_half.data(), rs_half.size() * 8);
    hipError_t e = hipMalloc(&c->d_tables, blob.host.size());
    if (e == hipSuccess) e = hipMemcpy(c->d_tables, blob.host.data(), blob.host.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        if (c->d_tables) hipFree(c->d_tables);
        delete c;
        return hip_status(e);
    }
    char *base = (char *)c->d_tables;
    DevTables &T = c->T;
    T.mod = (const ModConst *)(base + o_mod);
    T.tw = (const ulonglong2 *)(base + o_tw);
    T.itw = (const ulonglong2 *)(base + o_itw);
    T.md_y = (const u64 *)(base + o_mdy);
    T.md_qstar = (const u64 *)(base + o_mdq);
    T.md_vt = (const u64 *)(base + o_mdv);
    T.md_pinv = (const u64 *)(base + o_mdp);
    T.dc_y = (const u64 *)(base + o_dcy);
    T.dc_qstar = (const u64 *)(base + o_dcq);
    T.dc_vt = (const u64 *)(base + o_dcv);
    T.rs_inv = (const u64 *)(base + o_rsi);
    T.rs_half = (const u64 *)(base + o_rsh);
    T.logN = logN;
    T.N = N;
    T.L = L;
    T.K = K;
    T.alpha = alpha;
    T.beta_max = c->beta_max;
    T.nmod = nmod;
    *out = c;
    return MKHIP_OK;
}

extern "C" void mkhip_plan_destroy(mkhip_plan *plan) {
    if (!plan) return;
    if (plan->ge) hipGraphExecDestroy(plan->ge);
    if (plan->g) hipGraphDestroy(plan->g);
    if (plan->d_mem) hipFree(plan->d_mem);
    delete plan;
}

extern "C" void mkhip_ctx_destroy(mkhip_ctx *c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    mkhip_plan_destroy(c->single);
    if (c->single_io) hipFree(c->single_io);
    for (auto &kv : c->parties) {
        hipFree(kv.second.b);
        hipFree(kv.second.d);
        hipFree(kv.second.v);
    }
    if (c->d_crs_u) hipFree(c->d_crs_u);
    if (c->d_tables) hipFree(c->d_tables);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int mkhip_ctx_info(const mkhip_ctx *c, int *alpha, int *beta, size_t *swk_words) {
    if (!c) return MKHIP_ERR_ARG;
    if (alpha) *alpha = c->alpha;
    if (beta) *beta = c->beta_max;
    if (swk_words) *swk_words = c->swk_words;
    return MKHIP_OK;
}

static int upload_new(mkhip_ctx *c, const uint64_t *src, u64 **dst) {
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, c->swk_words * 8));
    hipError_t e = hipMemcpy(p, src, c->swk_words * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(p);
        return hip_status(e);
    }
    *dst = (u64 *)p;
    return MKHIP_OK;
}

extern "C" int mkhip_set_crs(mkhip_ctx *c, int index, const uint64_t *swk) {
    if (!c || !swk) return MKHIP_ERR_ARG;
    if (index != -1) return MKHIP_ERR_ARG;
    if (c->d_crs_u) {
        hipFree(c->d_crs_u);
        c->d_crs_u = nullptr;
    }
    return upload_new(c, swk, &c->d_crs_u);
}

extern "C" int mkhip_register_party(mkhip_ctx *c, int party, const uint64_t *b, const uint64_t *d, const uint64_t *v) {
    if (!c || !b || !d || !v || party < 0) return MKHIP_ERR_ARG;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    mkhip_unregister_party(c, party);
    PartyKeys k;
    int s;
    if ((s = upload_new(c, b, &k.b)) != MKHIP_OK) return s;
    if ((s = upload_new(c, d, &k.d)) != MKHIP_OK) {
        hipFree(k.b);
        return s;
    }
    if ((s = upload_new(c, v, &k.v)) != MKHIP_OK) {
        hipFree(k.b);
        hipFree(k.d);
        return s;
    }
    c->parties[party] = k;
    return MKHIP_OK;
}

extern "C" int mkhip_unregister_party(mkhip_ctx *c, int party) {
    if (!c) return MKHIP_ERR_ARG;
    auto it = c->parties.find(party);
    if (it == c->parties.end()) return MKHIP_ERR_UNKNOWN_PARTY;
    hipStreamSynchronize(c->stream);
    hipFree(it->second.b);
    hipFree(it->second.d);
    hipFree(it->second.v);
    c->parties.erase(it);
    return MKHIP_OK;
}

extern "C" int mkhip_rescale_count(const mkhip_ctx *c, int level, double scale, double min_scale, int *nres,
                                   double *new_scale) {
    if (!c || !nres || level < 0 || level >= c->L) return MKHIP_ERR_ARG;
    // Evaluator.Rescale error checks (mkckks/evaluator.go:366-376)
    if (min_scale <= 0 || scale == 0 || level == 0) return MKHIP_ERR_SCALE;
    double s = scale;
    int nb = 0;
    while (level - nb >= 0 && s / (double)c->Q[level - nb] >= min_scale / 2) {
        s /= (double)c->Q[level - nb];
        ++nb;
    }
    if (nb > level) nb = level;  // never drop every limb
    *nres = nb;
    if (new_scale) *new_scale = s;
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// device buffers

extern "C" int mkhip_dev_alloc(mkhip_ctx *c, size_t bytes, void **p) {
    if (!c || !p) return MKHIP_ERR_ARG;
    HIP_TRY(hipMalloc(p, bytes ? bytes : 1));
    return MKHIP_OK;
}
extern "C" int mkhip_dev_free(mkhip_ctx *c, void *p) {
    if (!c) return MKHIP_ERR_ARG;
    HIP_TRY(hipFree(p));
    return MKHIP_OK;
}
extern "C" int mkhip_h2d(mkhip_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return MKHIP_ERR_ARG;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}
extern "C" int mkhip_d2h(mkhip_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c) return MKHIP_ERR_ARG;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}
extern "C" int mkhip_sync(mkhip_ctx *c) {
    if (!c) return MKHIP_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// KKLSS MK-CKKS plan

namespace {

struct Carve {
    size_t off = 0;
    size_t take(size_t bytes) {
        size_t o = (off + 255) & ~(size_t)255;
        off = o + bytes;
        return o;
    }
};

int find_id(const int *ids, int k, int id) {
    for (int i = 0; i < k; ++i)
        if (ids[i] == id) return i;
    return -1;
}

// stage s (0..7) of a KKLSS plan
hipError_t plan_stage(mkhip_plan *pl, int s) {
    mkhip_ctx *c = pl->ctx;
    const DevTables &T = c->T;
    hipStream_t st = c->stream;
    switch (s) {
    case 0: return launch_kernel(KDECOMP, T, pl->dec1, pl->n_dec1, pl->beta, pl->rows, st, pl->kmax1);
    case 1: return launch_kernel(KMAC_XY, T, pl->mac, pl->n_mac, pl->beta, pl->rows, st);
    case 2: return launch_kernel(KNTT_FWD, T, pl->ntt, pl->n_ntt, 0, 0, st);
    case 3: return launch_kernel(KTENSOR, T, pl->ten, pl->n_ten, pl->nl, 0, st);
    case 4: return launch_kernel(KEXT1, T, pl->ext1, pl->n_ext1, pl->beta, pl->rows, st);
    case 5: return launch_kernel(KMODDOWN, T, pl->md1, pl->n_md1, 0, 0, st);
    case 6: return launch_kernel(KDECOMP, T, pl->dec2, pl->n_dec2, pl->beta, pl->rows, st, pl->kmax2);
    case 7: return launch_kernel(KEXT2, T, pl->ext2, pl->n_ext2, pl->beta, pl->rows, st);
    case 8: return launch_kernel(KMODDOWN_FINAL, T, pl->md2, pl->n_md2, 0, 0, st);
    default: return hipErrorInvalidValue;
    }
}

int plan_launch_raw(mkhip_plan *pl) {
    for (int s = 0; s < MKHIP_PLAN_STAGES; ++s) HIP_TRY(plan_stage(pl, s));
    return MKHIP_OK;
}

}  // namespace

extern "C" int mkhip_plan_kklss_ckks(mkhip_ctx *c, int nops, const mkhip_kklss_op *ops, int use_graph,
                                     mkhip_plan **out) {
    if (!c || !ops || nops <= 0 || !out) return MKHIP_ERR_ARG;
    *out = nullptr;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    if (!c->d_crs_u) return MKHIP_ERR_NO_KEY;
    const int level = ops[0].level;
    const int N = c->N, K = c->K;
    for (int o = 0; o < nops; ++o) {
        const mkhip_kklss_op &op = ops[o];
        if (!op.op0 || !op.op1 || !op.out) return MKHIP_ERR_ARG;
        if (op.level != level) return MKHIP_ERR_LEVEL;
        if (op.level < 0 || op.level >= c->L) return MKHIP_ERR_LEVEL;
        if (op.k0 < 0 || op.k1 < 0 || op.k0 > MKHIP_MAX_PARTIES || op.k1 > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
        if (op.nres < 0 || op.nres > op.level) return MKHIP_ERR_SCALE;
        for (int i = 0; i < op.k0; ++i)
            if (!c->parties.count(op.ids0[i])) return MKHIP_ERR_UNKNOWN_PARTY;
        for (int i = 0; i < op.k1; ++i)
            if (!c->parties.count(op.ids1[i])) return MKHIP_ERR_UNKNOWN_PARTY;
    }
    const int nl = level + 1, rows = nl + K;
    const int beta = (nl + c->alpha - 1) / c->alpha;
    const size_t limb = (size_t)N * 8, poly = (size_t)nl * limb, qp = (size_t)rows * limb, hoist = (size_t)beta * qp;

    // ---- sizes ----
    Carve cv;
    std::vector<DecompDesc> dec1, dec2;
    std::vector<MacDesc> mac;
    std::vector<NttJob> ntt;
    std::vector<TensorDesc> ten;
    std::vector<ExtDesc> ext1, ext2;
    std::vector<ModDownDesc> md1, md2;
    struct OpWs {
        size_t D0, D1, x, y, tnt, ct, accx, accy, t, Dt, accv, accu;
        int ku;
        int ids[2 * MKHIP_MAX_PARTIES];
    };
    std::vector<OpWs> ws(nops);
    for (int o = 0; o < nops; ++o) {
        const mkhip_kklss_op &op = ops[o];
        OpWs &w = ws[o];
        // union of ids, ascending
        int ku = 0;
        for (int i = 0; i < op.k0; ++i)
            if (find_id(w.ids, ku, op.ids0[i]) < 0) w.ids[ku++] = op.ids0[i];
        for (int i = 0; i < op.k1; ++i)
            if (find_id(w.ids, ku, op.ids1[i]) < 0) w.ids[ku++] = op.ids1[i];
        std::sort(w.ids, w.ids + ku);
        w.ku = ku;
        if (ku > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
        w.D0 = cv.take(op.k0 * hoist);
        w.D1 = cv.take(op.k1 * hoist);
        w.x = cv.take(hoist);
        w.y = cv.take(hoist);
        w.tnt = cv.take((2 + op.k0 + op.k1) * poly);
        w.ct = cv.take((1 + ku) * poly);
        w.accx = cv.take(op.k1 * qp);
        w.accy = cv.take(op.k0 * qp);
        w.t = cv.take(op.k0 * poly);
        w.Dt = cv.take(op.k0 * hoist);
        w.accv = cv.take(op.k0 * qp);
        w.accu = cv.take(op.k0 * qp);
    }
    const size_t ws_bytes = cv.off;
    // descriptor counts
    int n_ntt = 0, n_ten = 0, n_ext1 = 0, n_md1 = 0, n_ext2 = 0, n_md2 = 0;
    for (int o = 0; o < nops; ++o) {
        n_ntt += (2 + ops[o].k0 + ops[o].k1) * nl;
        n_ten += 1 + ws[o].ku;
        n_ext1 += ops[o].k0 + ops[o].k1;
        n_md1 += ops[o].k0 + ops[o].k1;
        n_ext2 += ops[o].k0;
        n_md2 += 1 + ws[o].ku;
    }
    Carve dv;
    dv.off = ws_bytes;
    const size_t o_dec1 = dv.take(sizeof(DecompDesc) * 2 * nops);
    const size_t o_mac = dv.take(sizeof(MacDesc) * 2 * nops);
    const size_t o_ntt = dv.take(sizeof(NttJob) * n_ntt);
    const size_t o_ten = dv.take(sizeof(TensorDesc) * n_ten);
    const size_t o_ext1 = dv.take(sizeof(ExtDesc) * std::max(n_ext1, 1));
    const size_t o_md1 = dv.take(sizeof(ModDownDesc) * std::max(n_md1, 1));
    const size_t o_dec2 = dv.take(sizeof(DecompDesc) * nops);
    const size_t o_ext2 = dv.take(sizeof(ExtDesc) * std::max(n_ext2, 1));
    const size_t o_md2 = dv.take(sizeof(ModDownDesc) * n_md2);

    mkhip_plan *pl = new mkhip_plan();
    pl->ctx = c;
    pl->bytes = dv.off;
    hipError_t e = hipMalloc(&pl->d_mem, pl->bytes);
    if (e != hipSuccess) {
        delete pl;
        return hip_status(e);
    }
    char *B = (char *)pl->d_mem;
    auto P = [&](size_t off) { return (u64 *)(B + off); };

    for (int o = 0; o < nops; ++o) {
        const mkhip_kklss_op &op = ops[o];
        const OpWs &w = ws[o];
        // 1: decompositions of both sides; x / y accumulations (key-stationary)
        DecompDesc s0 = {}, s1 = {};
        MacDesc m0 = {}, m1 = {};
        for (int p = 0; p < op.k0; ++p) {
            s0.comp[p] = op.op0 + (size_t)(1 + p) * nl * N;
            m0.key[p] = c->parties[op.ids0[p]].d;
        }
        s0.D = P(w.D0);
        s0.k = op.k0;
        s0.level = level;
        m0.D = P(w.D0);
        m0.xy = P(w.x);
        m0.k = op.k0;
        m0.level = level;
        for (int p = 0; p < op.k1; ++p) {
            s1.comp[p] = op.op1 + (size_t)(1 + p) * nl * N;
            m1.key[p] = c->parties[op.ids1[p]].b;
        }
        s1.D = P(w.D1);
        s1.k = op.k1;
        s1.level = level;
        m1.D = P(w.D1);
        m1.xy = P(w.y);
        m1.k = op.k1;
        m1.level = level;
        dec1.push_back(s0);
        dec1.push_back(s1);
        mac.push_back(m0);
        mac.push_back(m1);
        pl->kmax1 = std::max(pl->kmax1, std::max(op.k0, op.k1));
        pl->kmax2 = std::max(pl->kmax2, op.k0);
        // 2: NTT of every input component: op0 c0, op0 comps, op1 c0, op1 comps
        u64 *tnt = P(w.tnt);
        const int nc0 = 1 + op.k0, nc1 = 1 + op.k1;
        for (int comp = 0; comp < nc0 + nc1; ++comp) {
            const u64 *src = comp < nc0 ? op.op0 + (size_t)comp * nl * N : op.op1 + (size_t)(comp - nc0) * nl * N;
            for (int l = 0; l < nl; ++l) {
                NttJob j;
                j.src = src + (size_t)l * N;
                j.dst = tnt + ((size_t)comp * nl + l) * N;
                j.mod = l;
                j.reduce = 0;
                ntt.push_back(j);
            }
        }
        const u64 *a0 = tnt, *b0 = tnt + (size_t)nc0 * nl * N;
        // 3: tensor into ct (coefficient domain)
        u64 *ct = P(w.ct);
        {
            TensorDesc t0 = {};
            t0.a0 = a0;
            t0.b0 = b0;
            t0.out = ct;
            t0.is_c0 = 1;
            ten.push_back(t0);
        }
        for (int u = 0; u < w.ku; ++u) {
            const int id = w.ids[u];
            const int p0 = find_id(op.ids0, op.k0, id), p1 = find_id(op.ids1, op.k1, id);
            TensorDesc td = {};
            td.a0 = a0;
            td.b0 = b0;
            td.au = p0 >= 0 ? tnt + (size_t)(1 + p0) * nl * N : nullptr;
            td.bu = p1 >= 0 ? tnt + (size_t)(nc0 + 1 + p1) * nl * N : nullptr;
            td.out = ct + (size_t)(1 + u) * nl * N;
            td.is_c0 = 0;
            ten.push_back(td);
        }
        // 4/5: Ext(D1_j, x) -> ct_j ; Ext(D0_i, y) -> t_i
        for (int p = 0; p < op.k1; ++p) {
            ExtDesc ed = {};
            ed.D = P(w.D1) + (size_t)p * beta * rows * N;
            ed.key[0] = P(w.x);
            ed.acc[0] = P(w.accx) + (size_t)p * rows * N;
            ed.nkey = 1;
            ed.level = level;
            ext1.push_back(ed);
            ModDownDesc md = {};
            md.src[0] = ed.acc[0];
            md.nsrc = 1;
            md.dst = ct + (size_t)(1 + find_id(w.ids, w.ku, op.ids1[p])) * nl * N;
            md.level = level;
            md.accumulate = 1;
            md1.push_back(md);
        }
        for (int p = 0; p < op.k0; ++p) {
            ExtDesc ed = {};
            ed.D = P(w.D0) + (size_t)p * beta * rows * N;
            ed.key[0] = P(w.y);
            ed.acc[0] = P(w.accy) + (size_t)p * rows * N;
            ed.nkey = 1;
            ed.level = level;
            ext1.push_back(ed);
            ModDownDesc md = {};
            md.src[0] = ed.acc[0];
            md.nsrc = 1;
            md.dst = P(w.t) + (size_t)p * nl * N;
            md.level = level;
            md.accumulate = 0;
            md1.push_back(md);
        }
        // 6: D(t_i)
        DecompDesc s2 = {};
        for (int p = 0; p < op.k0; ++p) s2.comp[p] = P(w.t) + (size_t)p * nl * N;
        s2.D = P(w.Dt);
        s2.k = op.k0;
        s2.level = level;
        dec2.push_back(s2);
        // 7: Ext(D(t_i), v_i), Ext(D(t_i), u)
        for (int p = 0; p < op.k0; ++p) {
            ExtDesc ed = {};
            ed.D = P(w.Dt) + (size_t)p * beta * rows * N;
            ed.key[0] = c->parties[op.ids0[p]].v;
            ed.key[1] = c->d_crs_u;
            ed.acc[0] = P(w.accv) + (size_t)p * rows * N;
            ed.acc[1] = P(w.accu) + (size_t)p * rows * N;
            ed.nkey = 2;
            ed.level = level;
            ext2.push_back(ed);
        }
        // 8: final ModDowns + rescale into the output
        {
            ModDownDesc md = {};
            for (int p = 0; p < op.k0; ++p) md.src[p] = P(w.accv) + (size_t)p * rows * N;
            md.nsrc = op.k0;
            md.dst = ct;
            md.level = level;
            md.accumulate = 1;
            md.final_ = 1;
            md.nres = op.nres;
            md.out = op.out;
            md2.push_back(md);
        }
        for (int u = 0; u < w.ku; ++u) {
            const int p0 = find_id(op.ids0, op.k0, w.ids[u]);
            ModDownDesc md = {};
            if (p0 >= 0) {
                md.src[0] = P(w.accu) + (size_t)p0 * rows * N;
                md.nsrc = 1;
            }
            md.dst = ct + (size_t)(1 + u) * nl * N;
            md.level = level;
            md.accumulate = 1;
            md.final_ = 1;
            md.nres = op.nres;
            md.out = op.out + (size_t)(1 + u) * (nl - op.nres) * N;
            md2.push_back(md);
        }
    }
    // ---- upload descriptors ----
    std::vector<char> hd(pl->bytes - ws_bytes);
    auto put = [&](size_t off, const void *src, size_t bytes) {
        if (bytes) std::memcpy(hd.data() + (off - ws_bytes), src, bytes);
    };
    put(o_dec1, dec1.data(), dec1.size() * sizeof(DecompDesc));
    put(o_mac, mac.data(), mac.size() * sizeof(MacDesc));
    put(o_ntt, ntt.data(), ntt.size() * sizeof(NttJob));
    put(o_ten, ten.data(), ten.size() * sizeof(TensorDesc));
    put(o_ext1, ext1.data(), ext1.size() * sizeof(ExtDesc));
    put(o_md1, md1.data(), md1.size() * sizeof(ModDownDesc));
    put(o_dec2, dec2.data(), dec2.size() * sizeof(DecompDesc));
    put(o_ext2, ext2.data(), ext2.size() * sizeof(ExtDesc));
    put(o_md2, md2.data(), md2.size() * sizeof(ModDownDesc));
    e = hipMemcpy(B + ws_bytes, hd.data(), hd.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        mkhip_plan_destroy(pl);
        return hip_status(e);
    }
    pl->beta = beta;
    pl->rows = rows;
    pl->nl = nl;
    pl->dec1 = B + o_dec1;
    pl->n_dec1 = (int)dec1.size();
    pl->mac = B + o_mac;
    pl->n_mac = nops;
    pl->ntt = B + o_ntt;
    pl->n_ntt = (int)ntt.size();
    pl->ten = B + o_ten;
    pl->n_ten = (int)ten.size();
    pl->ext1 = B + o_ext1;
    pl->n_ext1 = (int)ext1.size();
    pl->md1 = B + o_md1;
    pl->n_md1 = (int)md1.size();
    pl->dec2 = B + o_dec2;
    pl->n_dec2 = (int)dec2.size();
    pl->ext2 = B + o_ext2;
    pl->n_ext2 = (int)ext2.size();
    pl->md2 = B + o_md2;
    pl->n_md2 = (int)md2.size();

    if (use_graph) {
        // first launch eagerly (sets kernel attributes), then capture the sequence
        int s = plan_launch_raw(pl);
        if (s == MKHIP_OK) s = hip_status(hipStreamSynchronize(c->stream));
        if (s == MKHIP_OK) s = hip_status(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        if (s == MKHIP_OK) {
            int s2 = plan_launch_raw(pl);
            hipError_t e2 = hipStreamEndCapture(c->stream, &pl->g);
            s = s2 != MKHIP_OK ? s2 : hip_status(e2);
        }
        if (s == MKHIP_OK) s = hip_status(hipGraphInstantiate(&pl->ge, pl->g, nullptr, nullptr, 0));
        if (s != MKHIP_OK) {
            mkhip_plan_destroy(pl);
            return s;
        }
        pl->graph = true;
    }
    *out = pl;
    return MKHIP_OK;
}

extern "C" int mkhip_plan_launch(mkhip_plan *pl) {
    if (!pl) return MKHIP_ERR_ARG;
    if (pl->graph) {
        HIP_TRY(hipGraphLaunch(pl->ge, pl->ctx->stream));
        return MKHIP_OK;
    }
    return plan_launch_raw(pl);
}

extern "C" int mkhip_plan_time(mkhip_plan *pl, int iters, double *ms) {
    if (!pl || iters <= 0 || !ms) return MKHIP_ERR_ARG;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, pl->ctx->stream));
    for (int i = 0; i < iters; ++i) {
        int s = mkhip_plan_launch(pl);
        if (s != MKHIP_OK) return s;
    }
    HIP_TRY(hipEventRecord(b, pl->ctx->stream));
    HIP_TRY(hipEventSynchronize(b));
    float f = 0;
    HIP_TRY(hipEventElapsedTime(&f, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms = f / iters;
    return MKHIP_OK;
}

extern "C" size_t mkhip_plan_workspace_bytes(const mkhip_plan *pl) { return pl ? pl->bytes : 0; }

// Average device time of each of the plan's launches (HIP events on the context
// stream around each kernel, eager launches), in ms.
extern "C" int mkhip_plan_stage_times(mkhip_plan *pl, int iters, double *ms) {
    if (!pl || iters <= 0 || !ms) return MKHIP_ERR_ARG;
    constexpr int S = MKHIP_PLAN_STAGES;
    hipStream_t st = pl->ctx->stream;
    hipEvent_t ev[S + 1];
    for (int i = 0; i <= S; ++i) HIP_TRY(hipEventCreate(&ev[i]));
    double acc[S] = {0};
    for (int it = 0; it < iters; ++it) {
        HIP_TRY(hipEventRecord(ev[0], st));
        for (int s = 0; s < S; ++s) {
            HIP_TRY(plan_stage(pl, s));
            HIP_TRY(hipEventRecord(ev[s + 1], st));
        }
        HIP_TRY(hipEventSynchronize(ev[S]));
        for (int s = 0; s < S; ++s) {
            float f = 0;
            HIP_TRY(hipEventElapsedTime(&f, ev[s], ev[s + 1]));
            acc[s] += f;
        }
    }
    for (int i = 0; i <= S; ++i) (void)hipEventDestroy(ev[i]);
    for (int s = 0; s < S; ++s) ms[s] = acc[s] / iters;
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// host-pointer entry point

extern "C" int mkhip_kklss_ckks_mulrelin(mkhip_ctx *c, const uint64_t *op0, int level0, double scale0,
                                         const int *ids0, int k0, const uint64_t *op1, int level1, double scale1,
                                         const int *ids1, int k1, double min_scale, uint64_t *out, int *ids_out,
                                         int *ku_out, int *level_out, double *scale_out) {
    if (!c || !op0 || !op1 || !out || (k0 && !ids0) || (k1 && !ids1)) return MKHIP_ERR_ARG;
    if (k0 < 0 || k1 < 0) return MKHIP_ERR_ARG;
    if (k0 > MKHIP_MAX_PARTIES || k1 > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
    if (level0 < 0 || level1 < 0 || level0 >= c->L || level1 >= c->L) return MKHIP_ERR_LEVEL;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    // DropLevel to the common level (evaluator.go:577-581).  With alpha > 1 the
    // reference decomposes each operand at its own level, which this path does not model.
    const int level = std::min(level0, level1);
    if (level0 != level1 && c->alpha != 1) return MKHIP_ERR_LEVEL;
    // mulRelinHoisted ignores Rescale's error return (evaluator.go:585): a
    // rejected rescale (level 0, minScale <= 0) leaves the product unrescaled.
    int nres = 0;
    double ns = scale0 * scale1;
    int s = mkhip_rescale_count(c, level, scale0 * scale1, min_scale, &nres, &ns);
    if (s == MKHIP_ERR_SCALE) {
        nres = 0;
        ns = scale0 * scale1;
    } else if (s != MKHIP_OK) {
        return s;
    }
    const int N = c->N, nl = level + 1;
    int ids[2 * MKHIP_MAX_PARTIES], ku = 0;
    for (int i = 0; i < k0; ++i)
        if (find_id(ids, ku, ids0[i]) < 0) ids[ku++] = ids0[i];
    for (int i = 0; i < k1; ++i)
        if (find_id(ids, ku, ids1[i]) < 0) ids[ku++] = ids1[i];
    if (ku > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
    std::sort(ids, ids + ku);
    const size_t in0 = (size_t)(1 + k0) * nl * N, in1 = (size_t)(1 + k1) * nl * N, ow = (size_t)(1 + ku) * nl * N;
    const size_t words = in0 + in1 + ow;
    // shape signature of the cached plan
    int sig[4 + 2 * MKHIP_MAX_PARTIES] = {0};
    sig[0] = k0;
    sig[1] = k1;
    sig[2] = level;
    sig[3] = nres;
    for (int i = 0; i < k0; ++i) sig[4 + i] = ids0[i];
    for (int i = 0; i < k1; ++i) sig[4 + MKHIP_MAX_PARTIES + i] = ids1[i];
    if (!c->single || std::memcmp(sig, c->single->sig, sizeof(sig)) != 0 || c->single_io_words < words) {
        mkhip_plan_destroy(c->single);
        c->single = nullptr;
        if (c->single_io_words < words) {
            if (c->single_io) hipFree(c->single_io);
            c->single_io = nullptr;
            c->single_io_words = 0;
            HIP_TRY(hipMalloc(&c->single_io, words * 8));
            c->single_io_words = words;
        }
        mkhip_kklss_op op = {};
        op.op0 = c->single_io;
        op.op1 = c->single_io + in0;
        op.out = c->single_io + in0 + in1;
        for (int i = 0; i < k0; ++i) op.ids0[i] = ids0[i];
        for (int i = 0; i < k1; ++i) op.ids1[i] = ids1[i];
        op.k0 = k0;
        op.k1 = k1;
        op.level = level;
        op.nres = nres;
        mkhip_plan *pl = nullptr;
        s = mkhip_plan_kklss_ckks(c, 1, &op, 0, &pl);
        if (s != MKHIP_OK) return s;
        std::memcpy(pl->sig, sig, sizeof(sig));
        c->single = pl;
    }
    // stage the operands at the common level (limbs 0..level of each component)
    for (int comp = 0; comp <= k0; ++comp)
        HIP_TRY(hipMemcpyAsync(c->single_io + (size_t)comp * nl * N, op0 + (size_t)comp * (level0 + 1) * N,
                               (size_t)nl * N * 8, hipMemcpyHostToDevice, c->stream));
    for (int comp = 0; comp <= k1; ++comp)
        HIP_TRY(hipMemcpyAsync(c->single_io + in0 + (size_t)comp * nl * N, op1 + (size_t)comp * (level1 + 1) * N,
                               (size_t)nl * N * 8, hipMemcpyHostToDevice, c->stream));
    s = mkhip_plan_launch(c->single);
    if (s != MKHIP_OK) return s;
    const size_t outw = (size_t)(1 + ku) * (nl - nres) * N;
    HIP_TRY(hipMemcpyAsync(out, c->single_io + in0 + in1, outw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (ids_out)
        for (int i = 0; i < ku; ++i) ids_out[i] = ids[i];
    if (ku_out) *ku_out = ku;
    if (level_out) *level_out = level - nres;
    if (scale_out) *scale_out = ns;
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// RNS primitives, host buffers (the per-primitive operations of the reference's
// rnsbench, software/mkf/cmd/rnsbench/main.go:179-201: ntt_q / intt_q,
// decompose_q_to_qp, moddown_qp_to_q_one_poly, external_product_hoisted)

namespace {

struct DevScratch {
    void *p = nullptr;
    ~DevScratch() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int mkhip_ntt_limbs(mkhip_ctx *c, const uint64_t *in, uint64_t *out, const int *ring_idx, int nlimbs,
                               int inverse) {
    if (!c || !in || !out || !ring_idx || nlimbs <= 0) return MKHIP_ERR_ARG;
    const int N = c->N;
    for (int i = 0; i < nlimbs; ++i)
        if (ring_idx[i] < 0 || ring_idx[i] >= c->L + c->K) return MKHIP_ERR_ARG;
    const size_t words = (size_t)nlimbs * N;
    DevScratch s;
    HIP_TRY(hipMalloc(&s.p, words * 8 * 2 + sizeof(NttJob) * nlimbs + 256));
    u64 *din = (u64 *)s.p, *dout = din + words;
    NttJob *djobs = (NttJob *)(dout + words);
    std::vector<NttJob> jobs(nlimbs);
    for (int i = 0; i < nlimbs; ++i) {
        jobs[i].src = din + (size_t)i * N;
        jobs[i].dst = dout + (size_t)i * N;
        jobs[i].mod = ring_idx[i];
        jobs[i].reduce = 1;
    }
    HIP_TRY(hipMemcpyAsync(din, in, words * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(djobs, jobs.data(), sizeof(NttJob) * nlimbs, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_kernel(inverse ? KNTT_INV : KNTT_FWD, c->T, djobs, nlimbs, 0, 0, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, words * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

// KKLSS KeySwitcher.Decompose (mkrlwe/keyswitch.go:50-76): coefficient-domain a
// (level+1 limbs) -> beta digits x (level+1+K) rows, NTT domain.
extern "C" int mkhip_decompose(mkhip_ctx *c, const uint64_t *a, int level, uint64_t *out) {
    if (!c || !a || !out) return MKHIP_ERR_ARG;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    const int N = c->N, nl = level + 1, rows = nl + c->K, beta = (nl + c->alpha - 1) / c->alpha;
    const size_t inw = (size_t)nl * N, outw = (size_t)beta * rows * N;
    DevScratch s;
    HIP_TRY(hipMalloc(&s.p, (inw + outw) * 8 + sizeof(DecompDesc) + 256));
    u64 *din = (u64 *)s.p, *dout = din + inw;
    DecompDesc *dd = (DecompDesc *)(dout + outw);
    DecompDesc h = {};
    h.comp[0] = din;
    h.D = dout;
    h.k = 1;
    h.level = level;
    HIP_TRY(hipMemcpyAsync(din, a, inw * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dd, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_kernel(KDECOMP, c->T, dd, 1, beta, rows, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, outw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

// FastBasisExtender.ModDownQPtoQ (basis_extension.go:192-232): x is (level+1+K)
// rows, coefficient domain; out is level+1 limbs.
extern "C" int mkhip_moddown(mkhip_ctx *c, const uint64_t *x, int level, uint64_t *out) {
    if (!c || !x || !out) return MKHIP_ERR_ARG;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    if (c->K == 0) return MKHIP_ERR_PARAMS;
    const int N = c->N, nl = level + 1, rows = nl + c->K;
    const size_t inw = (size_t)rows * N, outw = (size_t)nl * N;
    DevScratch s;
    HIP_TRY(hipMalloc(&s.p, (inw + outw) * 8 + sizeof(ModDownDesc) + 256));
    u64 *din = (u64 *)s.p, *dout = din + inw;
    ModDownDesc *dd = (ModDownDesc *)(dout + outw);
    ModDownDesc h = {};
    h.src[0] = din;
    h.nsrc = 1;
    h.dst = dout;
    h.level = level;
    HIP_TRY(hipMemcpyAsync(din, x, inw * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dd, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_kernel(KMODDOWN, c->T, dd, 1, 0, 0, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, outw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

// KeySwitcher.ExternalProductHoisted (keyswitch_hoisted.go:11-45): a_hoisted is
// beta x (level+1+K) x N (NTT), key beta x (L+K) x N (NTT, MForm); out level+1 limbs.
extern "C" int mkhip_external_product(mkhip_ctx *c, const uint64_t *a_hoisted, const uint64_t *key, int level,
                                      uint64_t *out) {
    if (!c || !a_hoisted || !key || !out) return MKHIP_ERR_ARG;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    const int N = c->N, nl = level + 1, rows = nl + c->K, beta = (nl + c->alpha - 1) / c->alpha;
    const size_t aw = (size_t)beta * rows * N, kw = c->swk_words, accw = (size_t)rows * N, outw = (size_t)nl * N;
    DevScratch s;
    HIP_TRY(hipMalloc(&s.p, (aw + kw + accw + outw) * 8 + sizeof(ExtDesc) + sizeof(ModDownDesc) + 512));
    u64 *da = (u64 *)s.p, *dk = da + aw, *dacc = dk + kw, *dout = dacc + accw;
    ExtDesc *de = (ExtDesc *)(dout + outw);
    ModDownDesc *dm = (ModDownDesc *)((char *)de + ((sizeof(ExtDesc) + 255) & ~(size_t)255));
    ExtDesc he = {};
    he.D = da;
    he.key[0] = dk;
    he.acc[0] = dacc;
    he.nkey = 1;
    he.level = level;
    ModDownDesc hm = {};
    hm.src[0] = dacc;
    hm.nsrc = 1;
    hm.dst = dout;
    hm.level = level;
    HIP_TRY(hipMemcpyAsync(da, a_hoisted, aw * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dk, key, kw * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(de, &he, sizeof(he), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dm, &hm, sizeof(hm), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_kernel(KEXT1, c->T, de, 1, beta, rows, c->stream));
    HIP_TRY(launch_kernel(KMODDOWN, c->T, dm, 1, 0, 0, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, outw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}
