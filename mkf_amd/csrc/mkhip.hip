// Host side of the C ABI (include/mkhip.h): parameter tables (moduli, twiddles,
// basis-extension and rescale constants), resident keys, device buffers, and the
// RNS primitives on host buffers.  Plans live in mkhip_plan.hip, the KKLSS path
// in mkhip_kklss.hip, the MKF path in mkhip_mkf.hip.
#include "mkhip_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

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

}  // namespace

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
    case MKHIP_ERR_SCRATCH: return "work larger than the allocation-free entry point's fixed scratch";
    default: return "unknown status";
    }
}

extern "C" int mkhip_ctx_create(mkhip_ctx **out, int logN, const uint64_t *q, int L, const uint64_t *p, int K,
                                int gamma, uint64_t t, const uint64_t *qmul, int flags) {
    if (!out || !q || (K > 0 && !p) || L <= 0 || K < 0 || gamma <= 0) return MKHIP_ERR_ARG;
    *out = nullptr;
    if (logN < 13 || logN > 15) return MKHIP_ERR_PARAMS;
    if (L > 16 || K > 4 || L + K > kMaxModuli) return MKHIP_ERR_PARAMS;
    const bool bfv = qmul != nullptr;
    const int N = 1 << logN;
    const int nmod = L + K + (bfv ? L : 0);
    std::vector<u64> all(q, q + L);
    all.insert(all.end(), p, p + K);
    if (bfv) all.insert(all.end(), qmul, qmul + L);
    for (int i = 0; i < nmod; ++i) {
        const u64 m = all[i];
        if (m >= (1ULL << 61) || ((m - 1) % (2 * (u64)N)) != 0 || !is_prime64(m)) return MKHIP_ERR_PARAMS;
        for (int j = 0; j < i; ++j)
            if (all[j] == m) return MKHIP_ERR_PARAMS;
    }
    const int alpha = K / gamma;
    if (K > 0 && alpha == 0) return MKHIP_ERR_GADGET;
    if (alpha > kMaxAlpha) return MKHIP_ERR_PARAMS;
    // MK-BFV (mkbfv/params.go:36-81, mkf_cpu/bfv.go:105-137): R = Q u QMul and a
    // plaintext modulus t.  KKLSS-BFV evaluations additionally need gamma 2 with
    // alpha = 1 (checked where they are planned); MKF-BFV needs gamma 1.
    if (bfv && (t < 2 || (gamma != 1 && gamma != 2))) return MKHIP_ERR_PARAMS;

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
    c->bfv = bfv;
    if (bfv) c->QMul.assign(qmul, qmul + L);
    c->swk_words = (size_t)c->beta_max * (L + K) * N;
    c->mkf_swk_words = 2 * c->swk_words;
    std::vector<u64> QP(c->Q);
    QP.insert(QP.end(), c->P.begin(), c->P.end());

    // ---- host tables ----
    std::vector<ModConst> mods(nmod);
    std::vector<ulonglong2> tw((size_t)nmod * N), itw((size_t)nmod * N);
    std::vector<double> twd((size_t)nmod * N), itwd((size_t)nmod * N);
    for (int i = 0; i < nmod; ++i) {
        const u64 qi = all[i];
        ModConst &m = mods[i];
        m.q = qi;
        m.qinv = h_qinv(qi);
        m.mu = (u64)((((u128)1) << 64) / qi);
        m.r2 = (u64)((((u128)1 << 64) % qi) * (u128)(((u128)1 << 64) % qi) % qi);
        m.r1 = (u64)(((u128)1 << 64) % qi);
        m.r1p = h_shoup_companion(m.r1, qi);
        m.ninv = h_inv((u64)N, qi);
        m.ninvp = h_shoup_companion(m.ninv, qi);
        m.tninv = h_mulmod(t % qi, m.ninv, qi);
        m.tninvp = h_shoup_companion(m.tninv, qi);
        const u64 g = primitive_root(qi, (flags & MKHIP_ROOT_RULE_SMALLEST2) != 0);
        const u64 psi = h_powmod(g, (qi - 1) / (2 * (u64)N), qi);
        const u64 psiinv = h_inv(psi, qi);
        c->psi_half.push_back(h_powmod(psi, (u64)N / 2, qi));
        u64 a = 1, b = 1;
        for (int j = 0; j < N; ++j) {
            const u64 idx = bitrev((u64)j, logN);
            tw[(size_t)i * N + idx] = make_ulonglong2(a, h_shoup_companion(a, qi));
            itw[(size_t)i * N + idx] = make_ulonglong2(b, h_shoup_companion(b, qi));
            twd[(size_t)i * N + idx] = (double)a;
            itwd[(size_t)i * N + idx] = (double)b;
            a = h_mulmod(a, psi, qi);
            b = h_mulmod(b, psiinv, qi);
        }
        // last inverse stage of a limb split over two workgroups: twiddle itw[1]
        const u64 w1 = itw[(size_t)i * N + 1].x;
        m.wn = h_mulmod(w1, m.ninv, qi);
        m.wnp = h_shoup_companion(m.wn, qi);
        m.twn = h_mulmod(w1, m.tninv, qi);
        m.twnp = h_shoup_companion(m.twn, qi);
        m.qd = (double)qi;
        m.qinvd = 1.0 / (double)qi;
        m.rinvd = (double)h_inv((u64)((((u128)1) << 64) % qi), qi);
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
    const int nqp = L + K;  // digit rows span Q u P
    std::vector<u64> dc_y(ne * kMaxAlpha, 0), dc_qstar(ne * nqp * kMaxAlpha, 0), dc_vt(ne * nqp * (kMaxAlpha + 1), 0);
    for (int i = 0; i < c->beta_max; ++i) {
        for (int nd = 2; nd <= alpha; ++nd) {
            if (i * alpha + nd > L) continue;
            std::vector<u64> Qin(c->Q.begin() + i * alpha, c->Q.begin() + i * alpha + nd), y, qs, vt;
            modup_params(Qin, QP, y, qs, vt);
            const size_t e = (size_t)i * (kMaxAlpha + 1) + nd;
            for (int l = 0; l < nd; ++l) dc_y[e * kMaxAlpha + l] = y[l];
            for (int r = 0; r < nqp; ++r) {
                for (int l = 0; l < nd; ++l) dc_qstar[(e * nqp + r) * kMaxAlpha + l] = qs[(size_t)r * nd + l];
                for (int v = 0; v <= nd; ++v) dc_vt[(e * nqp + r) * (kMaxAlpha + 1) + v] = vt[(size_t)r * (nd + 1) + v];
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

    // BFV basis extensions between Q and QMul (mkbfv/basis_extension.go:21-104)
    std::vector<u64> bq_qm, bq_up_y, bq_up_qs, bq_up_vt, bq_dn_y, bq_dn_qs, bq_dn_vt, bq_rsc, bq_qz;
    if (bfv) {
        const std::vector<u64> &QM = c->QMul;
        modup_params(c->Q, QM, bq_up_y, bq_up_qs, bq_up_vt);
        modup_params(QM, c->Q, bq_dn_y, bq_dn_qs, bq_dn_vt);
        bq_qm.resize(L);
        bq_rsc.resize(L);
        bq_qz.resize(L);
        for (int i = 0; i < L; ++i) {
            u64 qm = 1, qzi = 1, rsc = 1;
            for (int j = 0; j < L; ++j) {
                qm = h_mulmod(qm, QM[j] % c->Q[i], c->Q[i]);
                qzi = h_mulmod(qzi, h_inv(QM[j] % c->Q[i], c->Q[i]), c->Q[i]);
                rsc = h_mulmod(rsc, h_inv(c->Q[j] % QM[i], QM[i]), QM[i]);
            }
            bq_qm[i] = h_mform(qm, c->Q[i]);
            bq_qz[i] = (c->Q[i] - h_mform(qzi, c->Q[i])) % c->Q[i];
            bq_rsc[i] = (QM[i] - h_mform(rsc, QM[i])) % QM[i];
        }
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
    size_t o_twd = blob.add(twd.data(), twd.size() * sizeof(double));
    size_t o_itwd = blob.add(itwd.data(), itwd.size() * sizeof(double));
    size_t o_mdy = blob.add(md_y.data(), md_y.size() * 8);
    size_t o_mdq = blob.add(md_qstar.data(), md_qstar.size() * 8);
    size_t o_mdv = blob.add(md_vt.data(), md_vt.size() * 8);
    size_t o_mdp = blob.add(md_pinv.data(), md_pinv.size() * 8);
    size_t o_dcy = blob.add(dc_y.data(), dc_y.size() * 8);
    size_t o_dcq = blob.add(dc_qstar.data(), dc_qstar.size() * 8);
    size_t o_dcv = blob.add(dc_vt.data(), dc_vt.size() * 8);
    size_t o_rsi = blob.add(rs_inv.data(), rs_inv.size() * 8);
    size_t o_rsh = blob.add(rs_half.data(), rs_half.size() * 8);
    size_t o_bq[9];
    const std::vector<u64> *bqv[9] = {&bq_qm, &bq_up_y, &bq_up_qs, &bq_up_vt, &bq_dn_y,
                                      &bq_dn_qs, &bq_dn_vt, &bq_rsc, &bq_qz};
    for (int i = 0; i < 9; ++i) o_bq[i] = blob.add(bqv[i]->data(), bqv[i]->size() * 8);
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
    T.twd = (const double *)(base + o_twd);
    T.itwd = (const double *)(base + o_itwd);
    T.md_y = (const u64 *)(base + o_mdy);
    T.md_qstar = (const u64 *)(base + o_mdq);
    T.md_vt = (const u64 *)(base + o_mdv);
    T.md_pinv = (const u64 *)(base + o_mdp);
    T.dc_y = (const u64 *)(base + o_dcy);
    T.dc_qstar = (const u64 *)(base + o_dcq);
    T.dc_vt = (const u64 *)(base + o_dcv);
    T.rs_inv = (const u64 *)(base + o_rsi);
    T.rs_half = (const u64 *)(base + o_rsh);
    const u64 **bqt[9] = {&T.bq_qm, &T.bq_up_y, &T.bq_up_qstar, &T.bq_up_vt, &T.bq_dn_y,
                          &T.bq_dn_qstar, &T.bq_dn_vt, &T.bq_rsc, &T.bq_qz};
    for (int i = 0; i < 9; ++i) *bqt[i] = bfv ? (const u64 *)(base + o_bq[i]) : nullptr;
    T.bfv = bfv ? 1 : 0;
    T.xcd = 1;
    T.fp64 = 0;  // integer transforms unless the FP64 knob is set
    T.split14 = 0;
    T.split14_ext = 0;
    T.digit_pre = 0;
    T.xcd_ext = 1;
    T.ext_cs = 1;
    T.inv_lazy = 1;
    T.logN = logN;
    T.N = N;
    T.L = L;
    T.K = K;
    T.alpha = alpha;
    T.beta_max = c->beta_max;
    T.nmod = nmod;
    // look up this device's allocation-free scratch once, with the device current
    // (a failure is kept and reported by the allocation-free entry point only)
    void *scratch_ptr = nullptr;
    (void)noalloc_scratch(&scratch_ptr);
    *out = c;
    return MKHIP_OK;
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
    for (auto &kv : c->mkf_groups) {
        for (u64 *f : kv.second.fold) hipFree(f);
        hipFree(kv.second.relin);
    }
    if (c->d_crs_u) hipFree(c->d_crs_u);
    for (auto &kv : c->crs) hipFree(kv.second);
    for (auto &kv : c->rot_keys) hipFree(kv.second);
    for (auto &kv : c->conj_keys) hipFree(kv.second);
    for (auto &kv : c->mkf_groups)
        for (auto &r : kv.second.rot)
            for (u64 *p : r.second) hipFree(p);
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

void mkhip::drop_cached_plan(mkhip_ctx *c) {
    if (!c->single) return;
    (void)hipStreamSynchronize(c->stream);
    mkhip_plan_destroy(c->single);
    c->single = nullptr;
}

int mkhip::upload_key(mkhip_ctx *c, const uint64_t *src, u64 **dst) {
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
    if (index == -1) {
        if (c->d_crs_u) {
            drop_cached_plan(c);
            hipStreamSynchronize(c->stream);
            hipFree(c->d_crs_u);
            c->d_crs_u = nullptr;
        }
        return upload_key(c, swk, &c->d_crs_u);
    }
    // CRS[0] (the a of relinearization and public keys), rotation CRS (index =
    // rotidx in [1, N/2)), conjugation CRS (-2) or the second a of the MK-BFV
    // relinearization keys (-3, mkbfv/keygen.go:37)
    if (index != -2 && index != -3 && (index < 0 || index >= c->N / 2)) return MKHIP_ERR_ARG;
    u64 *p = nullptr;
    int s = upload_key(c, swk, &p);
    if (s != MKHIP_OK) return s;
    auto it = c->crs.find(index);
    if (it != c->crs.end()) {
        drop_cached_plan(c);
        hipStreamSynchronize(c->stream);
        hipFree(it->second);
    }
    c->crs[index] = p;
    return MKHIP_OK;
}

extern "C" int mkhip_register_party(mkhip_ctx *c, int party, const uint64_t *b, const uint64_t *d, const uint64_t *v) {
    if (!c || !b || !d || !v || party < 0) return MKHIP_ERR_ARG;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    if (c->bfv) return MKHIP_ERR_SCHEME;
    if (c->parties.count(party)) mkhip_unregister_party(c, party);
    PartyKeys k;
    int s;
    if ((s = upload_key(c, b, &k.b)) != MKHIP_OK) return s;
    if ((s = upload_key(c, d, &k.d)) != MKHIP_OK) {
        hipFree(k.b);
        return s;
    }
    if ((s = upload_key(c, v, &k.v)) != MKHIP_OK) {
        hipFree(k.b);
        hipFree(k.d);
        return s;
    }
    c->parties[party] = k;
    return MKHIP_OK;
}

// Two switching keys laid out back to back (the two BFV gadget branches).
static int upload_pair(mkhip_ctx *c, const uint64_t *x0, const uint64_t *x1, u64 **dst) {
    void *p = nullptr;
    const size_t w = c->swk_words;
    HIP_TRY(hipMalloc(&p, 2 * w * 8));
    hipError_t e = hipMemcpy(p, x0, w * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy((u64 *)p + w, x1, w * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(p);
        return hip_status(e);
    }
    *dst = (u64 *)p;
    return MKHIP_OK;
}

extern "C" int mkhip_register_bfv_party(mkhip_ctx *c, int party, const uint64_t *b0, const uint64_t *d0,
                                        const uint64_t *v0, const uint64_t *b1, const uint64_t *d1) {
    if (!c || !b0 || !d0 || !v0 || !b1 || !d1 || party < 0) return MKHIP_ERR_ARG;
    if (!c->bfv || c->gamma != 2 || c->alpha != 1) return MKHIP_ERR_SCHEME;
    if (c->parties.count(party)) mkhip_unregister_party(c, party);
    PartyKeys k;
    k.bfv = true;
    int s;
    if ((s = upload_pair(c, b0, b1, &k.b)) != MKHIP_OK) return s;
    if ((s = upload_pair(c, d0, d1, &k.d)) != MKHIP_OK) {
        hipFree(k.b);
        return s;
    }
    if ((s = upload_key(c, v0, &k.v)) != MKHIP_OK) {
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
    drop_cached_plan(c);
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
// Device-to-device copy on the context stream, by a plain copy kernel (16-byte
// words; callers keeping ciphertexts resident use it to stage operands).
extern "C" int mkhip_d2d(mkhip_ctx *c, void *dst, const void *src, size_t bytes) {
    if (!c || (bytes && (!dst || !src))) return MKHIP_ERR_ARG;
    HIP_TRY(launch_copy(dst, src, bytes, c->stream));
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

// Split (2^15) inverse transforms: the last stage and the scaling of the nrows
// rows at base + r N (ring modulus mods[r]); synchronous, nothing below 2^15.
int mkhip::finish_rows(mkhip_ctx *c, u64 *base, const int *mods, int nrows, int tscale, bool split) {
    if ((c->logN != 15 && !split) || nrows <= 0) return MKHIP_OK;
    std::vector<FinishDesc> f(nrows);
    for (int r = 0; r < nrows; ++r) {
        f[r].row = base + (size_t)r * c->N;
        f[r].mod = mods[r];
        f[r].tscale = tscale;
    }
    DevScratch s;
    HIP_TRY(hipMalloc(&s.p, sizeof(FinishDesc) * nrows));
    HIP_TRY(hipMemcpyAsync(s.p, f.data(), sizeof(FinishDesc) * nrows, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_kernel(KINTT_FINISH, c->T, s.p, nrows, 0, 0, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

extern "C" int mkhip_ntt_limbs(mkhip_ctx *c, const uint64_t *in, uint64_t *out, const int *ring_idx, int nlimbs,
                               int inverse) {
    if (!c || !in || !out || !ring_idx || nlimbs <= 0) return MKHIP_ERR_ARG;
    const int N = c->N;
    for (int i = 0; i < nlimbs; ++i)
        if (ring_idx[i] < 0 || ring_idx[i] >= c->T.nmod) return MKHIP_ERR_ARG;
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
    if (inverse) {
        const int st = finish_rows(c, dout, ring_idx, nlimbs, 0);
        if (st != MKHIP_OK) return st;
    }
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
    {
        std::vector<int> mods(rows);
        for (int r = 0; r < rows; ++r) mods[r] = r < nl ? r : c->L + r - nl;
        const int st = finish_rows(c, dacc, mods.data(), rows, 0, ext_split(c->T));
        if (st != MKHIP_OK) return st;
    }
    HIP_TRY(launch_kernel(KMODDOWN, c->T, dm, 1, 0, 0, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, outw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}
