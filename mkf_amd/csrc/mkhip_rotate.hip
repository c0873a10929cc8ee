// Galois key switches: KKLSS Rotate / RotateHoisted / Conjugate and the MKF
// party-specific RotateNew.  Key registration, plan construction and the
// host-pointer entry points.
//
// KKLSS (software/kklss/mkrlwe/keyswitch.go:243-344, keyswitch_hoisted.go:195-262),
// launches batched over every op of a plan:
//   conj_permute  conjugations: sigma_{2N-1} of every input component (:324-327)
//   decompose     D(c_i) of the input -- one decomposition serves every rotation of
//                 the op, the hoisting of RotateHoisted -- and of each conjugated input
//   ext           Ext(D(c_i), rk_i) and Ext(D(c_i), CRS[r]) per rotation and party,
//                 accumulators back to the coefficient domain (:268-275, :331-343)
//   moddown       c0 + sum_i ModDown(.) and ModDown(.) per party
//   permute       rotations: sigma_gal of every component (:278-306)
//
// MKF (software/mkf/mkf_cpu/true_mkf_ckks.go:323-353):
//   permute       sigma_gal of every NTT-domain component: one slot permutation in
//                 place of InvNTT, Permute, NTT (:306-319), the same residues
//   intt_in, decompose, mac_fold, moddown_fold   the fold key switch
//                 (common.go:79-126) of the permuted components under the rotation
//                 keys, sigma(C0) as the base
//   materialize   the k copies of the second output (:344-346)
#include "mkhip_internal.hpp"

#include <algorithm>

using namespace mkhip;

namespace {

u64 pow_mod2n(u64 b, u64 e, u64 mask) {
    u64 r = 1;
    b &= mask;
    while (e) {
        if (e & 1) r = (r * b) & mask;
        b = (b * b) & mask;
        e >>= 1;
    }
    return r;
}

// GaloisElementForColumnRotationBy (Lattigo v2.3.0 rlwe.Parameters): 5^(k mod 2N) mod 2N
u64 galois_element(int logN, long long k) {
    const u64 mask = ((u64)2 << logN) - 1;
    return pow_mod2n(5, (u64)k & mask, mask);
}

// InverseGaloisElement: g^(2N-1) mod 2N
u64 galois_inverse(int logN, u64 g) {
    const u64 mask = ((u64)2 << logN) - 1;
    return pow_mod2n(g, mask, mask);
}

// mkckks / mkbfv Evaluator.rotate's normalisation into [0, N/2) (evaluator.go:498-505)
int norm_rot(int N, long long r) {
    const long long h = N / 2;
    r %= h;
    if (r < 0) r += h;
    return (int)r;
}

int find_id(const int *ids, int k, int id) {
    for (int i = 0; i < k; ++i)
        if (ids[i] == id) return i;
    return -1;
}

int replace_key(mkhip_ctx *c, u64 *&slot, u64 *fresh) {
    if (slot) {
        drop_cached_plan(c);
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(slot);
    }
    slot = fresh;
    return MKHIP_OK;
}

// The host entry points stage operands in c->single_io; growing it invalidates
// the cached plan that points into it.
int ensure_io(mkhip_ctx *c, size_t words) {
    if (c->single_io_words >= words) return MKHIP_OK;
    mkhip_plan_destroy(c->single);
    c->single = nullptr;
    if (c->single_io) (void)hipFree(c->single_io);
    c->single_io = nullptr;
    c->single_io_words = 0;
    HIP_TRY(hipMalloc(&c->single_io, words * 8));
    c->single_io_words = words;
    return MKHIP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// keys

extern "C" int mkhip_add_rotation_key(mkhip_ctx *c, int party, int rotidx, const uint64_t *swk) {
    if (!c || !swk || party < 0) return MKHIP_ERR_ARG;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    const int r = norm_rot(c->N, rotidx);
    if (r == 0) return MKHIP_ERR_ARG;
    u64 *p = nullptr;
    int s = upload_key(c, swk, &p);
    if (s != MKHIP_OK) return s;
    return replace_key(c, c->rot_keys[std::make_pair(party, r)], p);
}

extern "C" int mkhip_del_rotation_key(mkhip_ctx *c, int party, int rotidx) {
    if (!c) return MKHIP_ERR_ARG;
    auto it = c->rot_keys.find(std::make_pair(party, norm_rot(c->N, rotidx)));
    if (it == c->rot_keys.end()) return MKHIP_ERR_NO_KEY;
    drop_cached_plan(c);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(it->second);
    c->rot_keys.erase(it);
    return MKHIP_OK;
}

extern "C" int mkhip_add_conjugation_key(mkhip_ctx *c, int party, const uint64_t *swk) {
    if (!c || !swk || party < 0) return MKHIP_ERR_ARG;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    u64 *p = nullptr;
    int s = upload_key(c, swk, &p);
    if (s != MKHIP_OK) return s;
    return replace_key(c, c->conj_keys[party], p);
}

extern "C" int mkhip_del_conjugation_key(mkhip_ctx *c, int party) {
    if (!c) return MKHIP_ERR_ARG;
    auto it = c->conj_keys.find(party);
    if (it == c->conj_keys.end()) return MKHIP_ERR_NO_KEY;
    drop_cached_plan(c);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(it->second);
    c->conj_keys.erase(it);
    return MKHIP_OK;
}

extern "C" int mkhip_register_mkf_rotation(mkhip_ctx *c, int group, int rotation, const uint64_t *const *keys) {
    if (!c || !keys) return MKHIP_ERR_ARG;
    auto it = c->mkf_groups.find(group);
    if (it == c->mkf_groups.end()) return MKHIP_ERR_NO_KEY;
    MkfGroup &g = it->second;
    for (int i = 0; i < g.k; ++i)
        if (!keys[i]) return MKHIP_ERR_ARG;
    const u64 gal = galois_element(c->logN, rotation);
    const size_t bytes = c->mkf_swk_words * 8;
    std::vector<u64 *> up;
    hipError_t e = hipSuccess;
    for (int i = 0; i < g.k && e == hipSuccess; ++i) {
        void *p = nullptr;
        e = hipMalloc(&p, bytes);
        if (e == hipSuccess) {
            up.push_back((u64 *)p);
            e = hipMemcpy(p, keys[i], bytes, hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) {
        for (u64 *p : up) (void)hipFree(p);
        return hip_status(e);
    }
    auto old = g.rot.find(gal);
    if (old != g.rot.end()) {
        drop_cached_plan(c);
        (void)hipStreamSynchronize(c->stream);
        for (u64 *p : old->second) (void)hipFree(p);
    }
    g.rot[gal] = up;
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// KKLSS plan

int mkhip::plan_kklss_rotate(mkhip_ctx *c, int nops, const mkhip_rotate_op *ops, int use_graph, mkhip_plan **out,
                             const PlanMem &pm) {
    if (!c || !ops || nops <= 0 || !out) return MKHIP_ERR_ARG;
    *out = nullptr;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    const int level = ops[0].level;
    const int N = c->N, K = c->K;
    for (int o = 0; o < nops; ++o) {
        const mkhip_rotate_op &op = ops[o];
        if (!op.in || !op.out) return MKHIP_ERR_ARG;
        if (op.level != level || level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
        if (op.k < 0 || op.k > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
        if (op.nrot <= 0 || op.nrot > MKHIP_MAX_ROTATIONS) return MKHIP_ERR_ARG;
        for (int i = 0; i < op.k; ++i)
            if (find_id(op.ids, i, op.ids[i]) >= 0) return MKHIP_ERR_ARG;  // an id set
        for (int r = 0; r < op.nrot; ++r) {
            const int rot = op.rot[r];
            if (rot != MKHIP_CONJUGATE && (rot < 1 || rot >= N / 2)) return MKHIP_ERR_ARG;
            if (!c->crs.count(rot)) return MKHIP_ERR_NO_KEY;
            for (int i = 0; i < op.k; ++i) {
                const bool have = rot == MKHIP_CONJUGATE ? c->conj_keys.count(op.ids[i]) > 0
                                                         : c->rot_keys.count(std::make_pair(op.ids[i], rot)) > 0;
                if (!have) return MKHIP_ERR_NO_KEY;  // keys.go:151-160, :218-225
            }
        }
    }
    const int nl = level + 1, rows = nl + K;
    const int beta = (nl + c->alpha - 1) / c->alpha;
    const size_t limb = (size_t)N * 8, poly = (size_t)nl * limb, qp = (size_t)rows * limb, hoist = (size_t)beta * qp;

    struct RotWs {
        size_t pc = 0, Dc = 0, acc = 0, tmp = 0;
    };
    struct OpWs {
        bool din = false;  // the input is decomposed (a rotation with parties)
        size_t Din = 0;
        RotWs r[MKHIP_MAX_ROTATIONS];
    };
    std::vector<OpWs> ws(nops);
    Carve cv;
    for (int o = 0; o < nops; ++o) {
        const mkhip_rotate_op &op = ops[o];
        const int k = op.k;
        if (k == 0) continue;  // no key switch: the permutations read and write the op's buffers
        bool rot_any = false;
        for (int r = 0; r < op.nrot; ++r) rot_any |= op.rot[r] != MKHIP_CONJUGATE;
        if (rot_any) {
            ws[o].din = true;
            ws[o].Din = cv.take(k * hoist);
        }
        for (int r = 0; r < op.nrot; ++r) {
            RotWs &w = ws[o].r[r];
            if (op.rot[r] == MKHIP_CONJUGATE) {
                w.pc = cv.take((1 + k) * poly);
                w.Dc = cv.take(k * hoist);
            } else {
                w.tmp = cv.take((1 + k) * poly);
            }
            w.acc = cv.take(2 * k * qp);
        }
    }
    mkhip_plan *pl = nullptr;
    int st = plan_create(c, cv.off, &pl, pm);
    if (st != MKHIP_OK) return st;
    char *B = (char *)pl->d_mem;
    auto P = [&](size_t off) { return (u64 *)(B + off); };

    const u64 conj_gal = ((u64)2 << c->logN) - 1;  // GaloisElementForRowRotation, its own inverse
    std::vector<AutoDesc> pre, post;
    std::vector<DecompDesc> dec;
    std::vector<ExtDesc> ext;
    std::vector<ModDownDesc> md;
    int kmax = 1;
    for (int o = 0; o < nops; ++o) {
        const mkhip_rotate_op &op = ops[o];
        const OpWs &W = ws[o];
        const int k = op.k;
        kmax = std::max(kmax, k);
        auto comp = [&](const u64 *base, int p) { return base + (size_t)p * nl * N; };
        auto compw = [&](u64 *base, int p) { return base + (size_t)p * nl * N; };
        if (W.din) {
            DecompDesc d = {};
            d.lazy_out = 1;  // digits feed Montgomery products only
            for (int p = 0; p < k; ++p) d.comp[p] = comp(op.in, 1 + p);
            d.D = P(W.Din);
            d.k = k;
            d.level = level;
            dec.push_back(d);
        }
        for (int r = 0; r < op.nrot; ++r) {
            const RotWs &w = W.r[r];
            const bool conj = op.rot[r] == MKHIP_CONJUGATE;
            u64 *dst = op.out + (size_t)r * (1 + k) * nl * N;
            const u64 *Dsrc = nullptr;
            const u64 *base0 = op.in;  // c0 the key switches add to
            if (conj) {
                u64 *pc = k ? P(w.pc) : dst;  // with no party the permutation is the result
                for (int p = 0; p <= k; ++p) {
                    AutoDesc a = {};
                    a.src = comp(op.in, p);
                    a.dst = compw(pc, p);
                    a.gal = a.gal_inv = conj_gal;
                    a.nl = nl;
                    pre.push_back(a);
                }
                if (k) {
                    DecompDesc d = {};
                    d.lazy_out = 1;
                    for (int p = 0; p < k; ++p) d.comp[p] = comp(pc, 1 + p);
                    d.D = P(w.Dc);
                    d.k = k;
                    d.level = level;
                    dec.push_back(d);
                    Dsrc = P(w.Dc);
                    base0 = pc;
                }
            } else {
                Dsrc = k ? P(W.Din) : nullptr;
            }
            // key-switch outputs: rotations into tmp (permuted afterwards), conjugations into dst
            u64 *ks = conj ? dst : (k ? P(w.tmp) : nullptr);
            if (k) {
                const u64 *a = c->crs[op.rot[r]];
                ModDownDesc m0 = {};
                for (int p = 0; p < k; ++p) {
                    ExtDesc e = {};
                    e.D = Dsrc + (size_t)p * beta * rows * N;
                    e.key[0] = conj ? c->conj_keys[op.ids[p]] : c->rot_keys[std::make_pair(op.ids[p], op.rot[r])];
                    e.key[1] = a;
                    e.acc[0] = P(w.acc) + (size_t)(2 * p) * rows * N;
                    e.acc[1] = P(w.acc) + (size_t)(2 * p + 1) * rows * N;
                    e.nkey = 2;
                    e.level = level;
                    ext.push_back(e);
                    m0.src[p] = e.acc[0];
                    ModDownDesc mi = {};
                    mi.src[0] = e.acc[1];
                    mi.nsrc = 1;
                    mi.dst = compw(ks, 1 + p);
                    mi.level = level;
                    md.push_back(mi);
                }
                m0.nsrc = k;
                m0.base = base0;
                m0.dst = ks;
                m0.level = level;
                md.push_back(m0);
            }
            if (!conj) {
                const u64 g = galois_element(c->logN, op.rot[r]);
                for (int p = 0; p <= k; ++p) {
                    AutoDesc a = {};
                    a.src = k ? compw(ks, p) : comp(op.in, p);
                    a.dst = compw(dst, p);
                    a.gal = g;
                    a.gal_inv = galois_inverse(c->logN, g);
                    a.nl = nl;
                    post.push_back(a);
                }
            }
        }
    }
    DescBlob blob;
    auto &S = pl->stages;
    Finish fin(c);
    if (!pre.empty()) S.push_back(make_stage(KAUTO, blob.add(pre), (int)pre.size(), 0, 0, 1, "conj_permute"));
    if (!dec.empty()) S.push_back(make_stage(KDECOMP, blob.add(dec), (int)dec.size(), beta, rows, kmax, "decompose"));
    if (!ext.empty()) {
        S.push_back(make_stage(KEXT2, blob.add(ext), (int)ext.size(), beta, rows, 1, "ext"));
        // on half-limb workgroups the last inverse stage is left to the ModDowns,
        // the only readers of the accumulators
        if (Finish(c, ext_split(c->T)).on)
            for (ModDownDesc &m : md) m.unfinished = 1;
    }
    if (!md.empty()) S.push_back(make_stage(KMODDOWN, blob.add(md), (int)md.size(), 0, 0, 1, "moddown"));
    if (!post.empty()) S.push_back(make_stage(KAUTO, blob.add(post), (int)post.size(), 0, 0, 1, "permute"));
    st = plan_finish(pl, blob, use_graph);
    if (st != MKHIP_OK) return st;
    *out = pl;
    return MKHIP_OK;
}

extern "C" int mkhip_plan_kklss_rotate(mkhip_ctx *c, int nops, const mkhip_rotate_op *ops, int use_graph,
                                       mkhip_plan **out) {
    return plan_kklss_rotate(c, nops, ops, use_graph, out, PlanMem());
}

// ---------------------------------------------------------------------------
// KKLSS host entry points

namespace {

// one Galois key switch (rotation rot or MKHIP_CONJUGATE) of the staged input,
// through the cached single-op plan
int run_single(mkhip_ctx *c, const u64 *in, u64 *out, int level, const int *ids, int k, int rot) {
    std::vector<int> sig = {4, k, level, rot};
    sig.insert(sig.end(), ids, ids + k);
    sig.push_back((int)(in - c->single_io));
    sig.push_back((int)(out - c->single_io));
    if (!c->single || c->single->sig != sig) {
        mkhip_plan_destroy(c->single);
        c->single = nullptr;
        mkhip_rotate_op op = {};
        op.in = in;
        op.out = out;
        for (int i = 0; i < k; ++i) op.ids[i] = ids[i];
        op.k = k;
        op.level = level;
        op.nrot = 1;
        op.rot[0] = rot;
        mkhip_plan *pl = nullptr;
        int s = mkhip_plan_kklss_rotate(c, 1, &op, 0, &pl);
        if (s != MKHIP_OK) return s;
        pl->sig = sig;
        c->single = pl;
    }
    return mkhip_plan_launch(c->single);
}

int galois_host(mkhip_ctx *c, const uint64_t *in, int level, const int *ids, int k, const std::vector<int> &steps,
                uint64_t *out) {
    const size_t words = (size_t)(1 + k) * (level + 1) * c->N;
    int s = ensure_io(c, 2 * words);
    if (s != MKHIP_OK) return s;
    HIP_TRY(hipMemcpyAsync(c->single_io, in, words * 8, hipMemcpyHostToDevice, c->stream));
    // the steps ping-pong between the two halves of the staging buffer
    int cur = 0;
    for (int st : steps) {
        s = run_single(c, c->single_io + cur * words, c->single_io + (1 - cur) * words, level, ids, k, st);
        if (s != MKHIP_OK) return s;
        cur = 1 - cur;
    }
    HIP_TRY(hipMemcpyAsync(out, c->single_io + cur * words, words * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

int check_host_args(mkhip_ctx *c, const uint64_t *in, int level, const int *ids, int k, uint64_t *out) {
    if (!c || !in || !out || (k && !ids) || k < 0) return MKHIP_ERR_ARG;
    if (k > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    return MKHIP_OK;
}

}  // namespace

extern "C" int mkhip_kklss_rotate(mkhip_ctx *c, const uint64_t *in, int level, const int *ids, int k, int rotidx,
                                  int hoisted, uint64_t *out) {
    int s = check_host_args(c, in, level, ids, k, out);
    if (s != MKHIP_OK) return s;
    const int r = norm_rot(c->N, rotidx);
    if (r == 0) {  // evaluator.go:507-510: a copy
        std::memmove(out, in, (size_t)(1 + k) * (level + 1) * c->N * 8);
        return MKHIP_OK;
    }
    std::vector<int> steps;
    if (c->crs.count(r)) {
        steps.push_back(r);
    } else if (hoisted) {
        return MKHIP_ERR_NO_KEY;  // evaluator.go:619-621
    } else {  // evaluator.go:519-528: rotations by 1, 2, 4, ... of the binary expansion
        for (int kk = 1, rr = r; rr > 0; kk *= 2, rr /= 2)
            if (rr % 2) steps.push_back(kk);
    }
    return galois_host(c, in, level, ids, k, steps, out);
}

extern "C" int mkhip_kklss_conjugate(mkhip_ctx *c, const uint64_t *in, int level, const int *ids, int k,
                                     uint64_t *out) {
    int s = check_host_args(c, in, level, ids, k, out);
    if (s != MKHIP_OK) return s;
    return galois_host(c, in, level, ids, k, std::vector<int>{MKHIP_CONJUGATE}, out);
}

// ---------------------------------------------------------------------------
// MKF

int mkhip::plan_mkf_rotate(mkhip_ctx *c, int nops, const mkhip_mkf_rotate_op *ops, int use_graph, mkhip_plan **out,
                           const PlanMem &pm) {
    if (!c || !ops || nops <= 0 || !out) return MKHIP_ERR_ARG;
    *out = nullptr;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    if (c->gamma != 1 || c->bfv) return MKHIP_ERR_SCHEME;
    const int level = ops[0].level;
    const int N = c->N, K = c->K;
    std::vector<const std::vector<u64 *> *> keys(nops);
    for (int o = 0; o < nops; ++o) {
        const mkhip_mkf_rotate_op &op = ops[o];
        if (!op.in || !op.out) return MKHIP_ERR_ARG;
        if (op.level != level || level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
        auto g = c->mkf_groups.find(op.group);
        if (g == c->mkf_groups.end()) return MKHIP_ERR_NO_KEY;
        auto r = g->second.rot.find(galois_element(c->logN, op.rotation));
        if (r == g->second.rot.end()) return MKHIP_ERR_NO_KEY;  // true_mkf_ckks.go:327-330
        keys[o] = &r->second;
    }
    const int nl = level + 1, rows = nl + K;
    const int beta = (nl + c->alpha - 1) / c->alpha;
    const size_t limb = (size_t)N * 8, poly = (size_t)nl * limb, qp = (size_t)rows * limb, hoist = (size_t)beta * qp;
    struct OpWs {
        size_t perm, coef, D, facc;
        int k;
    };
    std::vector<OpWs> ws(nops);
    Carve cv;
    int kmax = 1;
    for (int o = 0; o < nops; ++o) {
        OpWs &w = ws[o];
        w.k = (int)keys[o]->size();
        kmax = std::max(kmax, w.k);
        w.perm = cv.take((1 + w.k) * poly);
        w.coef = cv.take(w.k * poly);
        w.D = cv.take(w.k * hoist);
        w.facc = cv.take(2 * w.k * qp);
    }
    mkhip_plan *pl = nullptr;
    int st = plan_create(c, cv.off, &pl, pm);
    if (st != MKHIP_OK) return st;
    char *B = (char *)pl->d_mem;
    auto P = [&](size_t off) { return (u64 *)(B + off); };

    std::vector<AutoDesc> perm;
    std::vector<NttJob> intt;
    std::vector<DecompDesc> dec;
    std::vector<Mac2Desc> mac;
    std::vector<MdNttDesc> mdn;
    std::vector<RescaleNttDesc> copies;
    for (int o = 0; o < nops; ++o) {
        const mkhip_mkf_rotate_op &op = ops[o];
        const OpWs &w = ws[o];
        const int k = w.k;
        const u64 g = galois_element(c->logN, op.rotation);
        u64 *pc = P(w.perm);
        for (int p = 0; p <= k; ++p) {
            AutoDesc a = {};
            a.src = op.in + (size_t)p * nl * N;
            a.dst = pc + (size_t)p * nl * N;
            a.gal = g;
            a.gal_inv = galois_inverse(c->logN, g);
            a.nl = nl;
            a.ntt = 1;
            perm.push_back(a);
        }
        DecompDesc d = {};
        d.lazy_out = 1;  // digits feed Montgomery products only
        MdNttDesc m0 = {}, m1 = {};
        for (int p = 0; p < k; ++p) {
            u64 *coef = P(w.coef) + (size_t)p * nl * N;
            for (int l = 0; l < nl; ++l) {
                NttJob j;
                j.src = pc + ((size_t)(1 + p) * nl + l) * N;
                j.dst = coef + (size_t)l * N;
                j.mod = l;
                j.reduce = 0;
                intt.push_back(j);
            }
            d.comp[p] = coef;
            Mac2Desc m = {};
            m.D = P(w.D) + (size_t)p * beta * rows * N;
            m.swk = (*keys[o])[p];
            m.acc[0] = P(w.facc) + (size_t)(2 * p) * rows * N;
            m.acc[1] = P(w.facc) + (size_t)(2 * p + 1) * rows * N;
            m.level = level;
            mac.push_back(m);
            m0.acc[p] = m.acc[0];
            m1.acc[p] = m.acc[1];
        }
        d.D = P(w.D);
        d.k = k;
        d.level = level;
        dec.push_back(d);
        m0.nacc = m1.nacc = k;
        m0.base = pc;  // sigma(C0)
        m0.dst = op.out;
        m1.dst = op.out + (size_t)nl * N;
        m0.level = m1.level = level;
        mdn.push_back(m0);
        mdn.push_back(m1);
        if (k > 1) {
            RescaleNttDesc r = {};
            r.x = m1.dst;
            r.xlast = m1.dst;
            for (int p = 1; p < k; ++p) r.out[p - 1] = op.out + (size_t)(1 + p) * nl * N;
            r.nout = k - 1;
            r.level = level;
            r.nres = 0;
            copies.push_back(r);
        }
    }
    DescBlob blob;
    auto &S = pl->stages;
    Finish fin(c);
    S.push_back(make_stage(KAUTO, blob.add(perm), (int)perm.size(), 0, 0, 1, "permute"));
    S.push_back(make_stage(KNTT_INV, blob.add(intt), (int)intt.size(), 0, 0, 1, "intt_in"));
    fin.jobs(intt);
    fin.emit(S, blob, "intt_in");
    S.push_back(make_stage(KDECOMP, blob.add(dec), (int)dec.size(), beta, rows, kmax, "decompose"));
    S.push_back(make_stage(KMAC2, blob.add(mac), (int)mac.size(), beta, rows, 1, "mac_fold"));
    fin.mac2(mac, c->L, c->K, N);
    fin.emit(S, blob, "mac_fold");
    S.push_back(make_stage(KMD_NTT, blob.add(mdn), (int)mdn.size(), nl, 0, 1, "moddown_fold"));
    if (!copies.empty())
        S.push_back(make_stage(KRESCALE_NTT, blob.add(copies), (int)copies.size(), nl, 0, 1, "materialize"));
    st = plan_finish(pl, blob, use_graph);
    if (st != MKHIP_OK) return st;
    *out = pl;
    return MKHIP_OK;
}

extern "C" int mkhip_plan_mkf_rotate(mkhip_ctx *c, int nops, const mkhip_mkf_rotate_op *ops, int use_graph,
                                     mkhip_plan **out) {
    return plan_mkf_rotate(c, nops, ops, use_graph, out, PlanMem());
}

extern "C" int mkhip_mkf_ckks_rotate(mkhip_ctx *c, int group, const uint64_t *in, int level, int rotation,
                                     uint64_t *out) {
    if (!c || !in || !out) return MKHIP_ERR_ARG;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    auto it = c->mkf_groups.find(group);
    if (it == c->mkf_groups.end()) return MKHIP_ERR_NO_KEY;
    const int k = it->second.k;
    const size_t words = (size_t)(1 + k) * (level + 1) * c->N;
    int s = ensure_io(c, 2 * words);
    if (s != MKHIP_OK) return s;
    const u64 g = galois_element(c->logN, rotation);
    std::vector<int> sig = {5, group, k, level, (int)g};
    if (!c->single || c->single->sig != sig) {
        mkhip_plan_destroy(c->single);
        c->single = nullptr;
        mkhip_mkf_rotate_op op = {};
        op.in = c->single_io;
        op.out = c->single_io + words;
        op.group = group;
        op.level = level;
        op.rotation = rotation;
        mkhip_plan *pl = nullptr;
        s = mkhip_plan_mkf_rotate(c, 1, &op, 0, &pl);
        if (s != MKHIP_OK) return s;
        pl->sig = sig;
        c->single = pl;
    }
    HIP_TRY(hipMemcpyAsync(c->single_io, in, words * 8, hipMemcpyHostToDevice, c->stream));
    s = mkhip_plan_launch(c->single);
    if (s != MKHIP_OK) return s;
    HIP_TRY(hipMemcpyAsync(out, c->single_io + words, words * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}
