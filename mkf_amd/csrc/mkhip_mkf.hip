// MKF MK-CKKS mul + relin (software/mkf/mkf_cpu/ckks.go:102-151): fold each side
// to the aggregate secret with one hoisted key switch per party
// (common.go:79-126), Karatsuba tensor, one relinearization of mu2, NTT-domain
// rescale, and the k-fold materialization of Gamma (common.go:128-139).
//
// Launches (each batched over all ops of the plan):
//   intt_in        components -> coefficient domain (DecomposeNTT's InvNTT)
//   decompose      gadget digits of every component, NTT'd (alpha = K, CRT when K > 1)
//   mac_fold       KeyswitchHoistedNoModDown: two accumulators per component
//   moddown_fold   ModDownQPtoQNTT of every switch, summed with c0 (fold)
//   karatsuba      mu0, mu1, mu2
//   intt_mu2, decompose_mu2, mac_relin, moddown_relin   switch(mu2) added to mu0, mu1
//   intt_last      last limb of the result (rescale)
//   rescale        DivRoundByLastModulus (NTT domain) into c0 and k copies of Gamma
#include "mkhip_internal.hpp"

#include <algorithm>

using namespace mkhip;

extern "C" int mkhip_register_mkf_group(mkhip_ctx *c, int group, int k, const uint64_t *const *fold_keys,
                                        const uint64_t *relin_key) {
    if (!c || !fold_keys || !relin_key || k <= 0 || group < 0) return MKHIP_ERR_ARG;
    if (k > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    for (int i = 0; i < k; ++i)
        if (!fold_keys[i]) return MKHIP_ERR_ARG;
    mkhip_unregister_mkf_group(c, group);
    MkfGroup g;
    g.k = k;
    const size_t bytes = c->mkf_swk_words * 8;
    auto upload = [&](const uint64_t *src, u64 **dst) -> int {
        void *p = nullptr;
        HIP_TRY(hipMalloc(&p, bytes));
        *dst = (u64 *)p;
        HIP_TRY(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
        return MKHIP_OK;
    };
    int s = MKHIP_OK;
    for (int i = 0; i < k && s == MKHIP_OK; ++i) {
        u64 *p = nullptr;
        s = upload(fold_keys[i], &p);
        if (p) g.fold.push_back(p);
    }
    if (s == MKHIP_OK) s = upload(relin_key, &g.relin);
    if (s != MKHIP_OK) {
        for (u64 *p : g.fold) (void)hipFree(p);
        if (g.relin) (void)hipFree(g.relin);
        return s;
    }
    c->mkf_groups[group] = g;
    return MKHIP_OK;
}

extern "C" int mkhip_unregister_mkf_group(mkhip_ctx *c, int group) {
    if (!c) return MKHIP_ERR_ARG;
    auto it = c->mkf_groups.find(group);
    if (it == c->mkf_groups.end()) return MKHIP_ERR_NO_KEY;
    drop_cached_plan(c);
    (void)hipStreamSynchronize(c->stream);
    for (u64 *p : it->second.fold) (void)hipFree(p);
    (void)hipFree(it->second.relin);
    for (auto &r : it->second.rot)
        for (u64 *p : r.second) (void)hipFree(p);
    c->mkf_groups.erase(it);
    return MKHIP_OK;
}

int mkhip::plan_mkf_ckks(mkhip_ctx *c, int nops, const mkhip_mkf_op *ops, int use_graph, mkhip_plan **out,
                       const PlanMem &pm) {
    if (!c || !ops || nops <= 0 || !out) return MKHIP_ERR_ARG;
    return plan_mkf_ckks_lv(c, nops, ops, ops[0].level, ops[0].level, use_graph, out, pm);
}

// As above with the left and right operands at levels lv_l, lv_r >= op.level: each
// is folded at its own level (common.go:109-126 switches at in.Level), the tensor
// and relinearization run at op.level (true_mkf_ckks.go:265-290).
int mkhip::plan_mkf_ckks_lv(mkhip_ctx *c, int nops, const mkhip_mkf_op *ops, int lv_l, int lv_r, int use_graph,
                            mkhip_plan **out, const PlanMem &pm) {
    if (!c || !ops || nops <= 0 || !out) return MKHIP_ERR_ARG;
    *out = nullptr;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    if (c->gamma != 1) return MKHIP_ERR_SCHEME;  // Lattigo rlwe decomposition: alpha = K
    const int level = ops[0].level, nres = ops[0].nres;
    const int N = c->N, K = c->K;
    if (lv_l < level || lv_r < level || lv_l >= c->L || lv_r >= c->L) return MKHIP_ERR_LEVEL;
    for (int o = 0; o < nops; ++o) {
        const mkhip_mkf_op &op = ops[o];
        if (!op.left || !op.right || !op.out) return MKHIP_ERR_ARG;
        if (op.level != level || level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
        if (op.nres != nres || nres < 0 || nres > 1 || nres > level) return MKHIP_ERR_SCALE;
        if (!c->mkf_groups.count(op.group)) return MKHIP_ERR_NO_KEY;
    }
    const bool same = lv_l == lv_r;
    const int lvs[2] = {lv_l, lv_r};
    const int nl = level + 1, rows = nl + K;
    const int beta = (nl + c->alpha - 1) / c->alpha;
    const int nlmax = std::max(lv_l, lv_r) + 1, rowsmax = nlmax + K;
    const int betamax = (nlmax + c->alpha - 1) / c->alpha;
    const size_t limb = (size_t)N * 8, poly = (size_t)nl * limb, qp = (size_t)rows * limb, hoist = (size_t)beta * qp;
    const size_t polymax = (size_t)nlmax * limb, qpmax = (size_t)rowsmax * limb, hoistmax = (size_t)betamax * qpmax;

    struct OpWs {
        size_t coef, D, facc, fold, mu, mu2c, Dr, racc, fin, last, pre, pre2;
        int k;
    };
    std::vector<OpWs> ws(nops);
    Carve cv;
    int kmax = 1;
    // CRT digits (alpha >= 2): the digit-only part of the ModUp once per digit
    const bool crt = c->alpha > 1 && c->T.digit_pre;
    const size_t preblk = (size_t)(c->alpha + 1) * limb;
    for (int o = 0; o < nops; ++o) {
        OpWs &w = ws[o];
        const int k = c->mkf_groups[ops[o].group].k;
        w.k = k;
        kmax = std::max(kmax, k);
        w.pre = crt ? cv.take(2 * k * betamax * preblk) : 0;
        w.pre2 = crt ? cv.take(beta * preblk) : 0;
        w.coef = cv.take(2 * k * polymax);
        w.D = cv.take(2 * k * hoistmax);
        w.facc = cv.take(2 * k * 2 * qpmax);
        w.fold = cv.take(4 * polymax);
        w.mu = cv.take(3 * poly);
        w.mu2c = cv.take(poly);
        w.Dr = cv.take(hoist);
        w.racc = cv.take(2 * qp);
        w.fin = cv.take(2 * poly);
        w.last = cv.take(2 * limb);
    }
    mkhip_plan *pl = nullptr;
    int st = plan_create(c, cv.off, &pl, pm);
    if (st != MKHIP_OK) return st;
    char *B = (char *)pl->d_mem;
    auto P = [&](size_t off) { return (u64 *)(B + off); };

    // fold descriptors per side when the sides' levels differ (their grids differ)
    std::vector<NttJob> intt_in[2], intt_mu2, intt_last;
    std::vector<DecompDesc> dec[2], dec_mu2;
    std::vector<Mac2Desc> mac_fold[2], mac_relin;
    std::vector<MdNttDesc> md_fold[2], md_relin;
    std::vector<KaraDesc> kara;
    std::vector<RescaleNttDesc> resc;
    for (int o = 0; o < nops; ++o) {
        const mkhip_mkf_op &op = ops[o];
        const OpWs &w = ws[o];
        const MkfGroup &g = c->mkf_groups[op.group];
        const int k = w.k;
        const u64 *side[2] = {op.left, op.right};
        auto coef = [&](int s, int p) { return P(w.coef) + ((size_t)s * k + p) * nlmax * N; };
        auto facc = [&](int s, int p, int oo) { return P(w.facc) + (((size_t)s * k + p) * 2 + oo) * rowsmax * N; };
        auto fold = [&](int s, int oo) { return P(w.fold) + ((size_t)s * 2 + oo) * nlmax * N; };
        auto mu = [&](int i) { return P(w.mu) + (size_t)i * nl * N; };
        for (int s = 0; s < 2; ++s) {
            const int sv = same ? 0 : s;
            const int ls = lvs[s], nls = ls + 1, rs = nls + K, bs = (nls + c->alpha - 1) / c->alpha;
            DecompDesc dd = {};
            dd.lazy_out = 1;  // digits feed Montgomery products only
            dd.D = P(w.D) + (size_t)s * k * betamax * rowsmax * N;
            if (crt) dd.pre = P(w.pre) + (size_t)s * k * betamax * (c->alpha + 1) * N;
            for (int p = 0; p < k; ++p) {
                for (int l = 0; l < nls; ++l) {
                    NttJob j;
                    j.src = side[s] + ((size_t)(1 + p) * nls + l) * N;
                    j.dst = coef(s, p) + (size_t)l * N;
                    j.mod = l;
                    j.reduce = 0;
                    intt_in[sv].push_back(j);
                }
                dd.comp[p] = coef(s, p);
                Mac2Desc md = {};
                md.D = dd.D + (size_t)p * bs * rs * N;
                md.swk = g.fold[p];
                md.acc[0] = facc(s, p, 0);
                md.acc[1] = facc(s, p, 1);
                md.level = ls;
                mac_fold[sv].push_back(md);
            }
            dd.k = k;
            dd.level = ls;
            dec[sv].push_back(dd);
            for (int oo = 0; oo < 2; ++oo) {
                MdNttDesc m = {};
                for (int p = 0; p < k; ++p) m.acc[p] = facc(s, p, oo);
                m.nacc = k;
                m.base = oo == 0 ? side[s] : nullptr;  // fold: out[0] starts from c0
                m.dst = fold(s, oo);
                m.level = ls;
                md_fold[sv].push_back(m);
            }
        }
        KaraDesc kd = {};
        kd.l0 = fold(0, 0);
        kd.l1 = fold(0, 1);
        kd.r0 = fold(1, 0);
        kd.r1 = fold(1, 1);
        kd.mu0 = mu(0);
        kd.mu1 = mu(1);
        kd.mu2 = mu(2);
        kara.push_back(kd);
        for (int l = 0; l < nl; ++l) {
            NttJob j;
            j.src = mu(2) + (size_t)l * N;
            j.dst = P(w.mu2c) + (size_t)l * N;
            j.mod = l;
            j.reduce = 0;
            intt_mu2.push_back(j);
        }
        DecompDesc d2 = {};
        d2.lazy_out = 1;
        d2.comp[0] = P(w.mu2c);
        d2.D = P(w.Dr);
        if (crt) d2.pre = P(w.pre2);
        d2.k = 1;
        d2.level = level;
        dec_mu2.push_back(d2);
        Mac2Desc mr = {};
        mr.D = P(w.Dr);
        mr.swk = g.relin;
        mr.acc[0] = P(w.racc);
        mr.acc[1] = P(w.racc) + (size_t)rows * N;
        mr.level = level;
        mac_relin.push_back(mr);
        for (int oo = 0; oo < 2; ++oo) {
            MdNttDesc m = {};
            m.acc[0] = mr.acc[oo];
            m.nacc = 1;
            m.base = mu(oo);
            m.dst = P(w.fin) + (size_t)oo * nl * N;
            m.level = level;
            md_relin.push_back(m);
            if (nres == 1) {
                NttJob j;
                j.src = m.dst + (size_t)level * N;
                j.dst = P(w.last) + (size_t)oo * N;
                j.mod = level;
                j.reduce = 0;
                intt_last.push_back(j);
            }
            RescaleNttDesc rd = {};
            rd.x = m.dst;
            rd.xlast = P(w.last) + (size_t)oo * N;
            rd.level = level;
            rd.nres = nres;
            const size_t olimbs = (size_t)(nl - nres) * N;
            if (oo == 0) {
                rd.out[0] = op.out;
                rd.nout = 1;
            } else {
                for (int p = 0; p < k; ++p) rd.out[p] = op.out + (size_t)(1 + p) * olimbs;
                rd.nout = k;
            }
            resc.push_back(rd);
        }
    }
    DescBlob blob;
    auto &S = pl->stages;
    Finish fin(c);
    const int nsides = same ? 1 : 2;
    for (int sv = 0; sv < nsides; ++sv) {
        S.push_back(make_stage(KNTT_INV, blob.add(intt_in[sv]), (int)intt_in[sv].size(), 0, 0, 1, "intt_in"));
        fin.jobs(intt_in[sv]);
        fin.emit(S, blob, "intt_in");
    }
    for (int sv = 0; sv < nsides; ++sv) {
        const int nls = lvs[sv] + 1, rs = nls + K, bs = (nls + c->alpha - 1) / c->alpha;
        const size_t off = blob.add(dec[sv]);
        if (crt) S.push_back(make_stage(KDIGIT_PRE, off, (int)dec[sv].size(), bs, 0, kmax, "digit_pre"));
        S.push_back(make_stage(KDECOMP, off, (int)dec[sv].size(), bs, rs, kmax, "decompose"));
    }
    for (int sv = 0; sv < nsides; ++sv) {
        const int nls = lvs[sv] + 1, rs = nls + K, bs = (nls + c->alpha - 1) / c->alpha;
        S.push_back(make_stage(KMAC2, blob.add(mac_fold[sv]), (int)mac_fold[sv].size(), bs, rs, 1, "mac_fold"));
        fin.mac2(mac_fold[sv], c->L, c->K, N);
        fin.emit(S, blob, "mac_fold");
    }
    for (int sv = 0; sv < nsides; ++sv)
        S.push_back(make_stage(KMD_NTT, blob.add(md_fold[sv]), (int)md_fold[sv].size(), lvs[sv] + 1, 0, 1,
                               "moddown_fold"));
    S.push_back(make_stage(KKARA, blob.add(kara), (int)kara.size(), nl, 0, 1, "karatsuba"));
    S.push_back(make_stage(KNTT_INV, blob.add(intt_mu2), (int)intt_mu2.size(), 0, 0, 1, "intt_mu2"));
    fin.jobs(intt_mu2);
    fin.emit(S, blob, "intt_mu2");
    {
        const size_t off = blob.add(dec_mu2);
        if (crt) S.push_back(make_stage(KDIGIT_PRE, off, (int)dec_mu2.size(), beta, 0, 1, "digit_pre_mu2"));
        S.push_back(make_stage(KDECOMP, off, (int)dec_mu2.size(), beta, rows, 1, "decompose_mu2"));
    }
    S.push_back(make_stage(KMAC2, blob.add(mac_relin), (int)mac_relin.size(), beta, rows, 1, "mac_relin"));
    fin.mac2(mac_relin, c->L, c->K, N);
    fin.emit(S, blob, "mac_relin");
    S.push_back(make_stage(KMD_NTT, blob.add(md_relin), (int)md_relin.size(), nl, 0, 1, "moddown_relin"));
    if (nres == 1) {
        S.push_back(make_stage(KNTT_INV, blob.add(intt_last), (int)intt_last.size(), 0, 0, 1, "intt_last"));
        fin.jobs(intt_last);
        fin.emit(S, blob, "intt_last");
    }
    S.push_back(make_stage(KRESCALE_NTT, blob.add(resc), (int)resc.size(), nl - nres, 0, 1, "rescale"));
    st = plan_finish(pl, blob, use_graph);
    if (st != MKHIP_OK) return st;
    *out = pl;
    return MKHIP_OK;
}

extern "C" int mkhip_plan_mkf_ckks(mkhip_ctx *c, int nops, const mkhip_mkf_op *ops, int use_graph,
                                     mkhip_plan **out) {
    return plan_mkf_ckks(c, nops, ops, use_graph, out, PlanMem());
}

// MKF CKKSCase.Run on host buffers: left/right [1+k][level+1][N] NTT domain;
// out [1+k][level_out+1][N]: c0 then k copies of Gamma.  The rescale error of
// Lattigo's Evaluator.Rescale is returned (ckks.go:141-149).
extern "C" int mkhip_mkf_ckks_mulrelin(mkhip_ctx *c, int group, const uint64_t *left, const uint64_t *right,
                                       int level, double scale_l, double scale_r, double min_scale, uint64_t *out,
                                       int *level_out, double *scale_out) {
    if (!c || !left || !right || !out) return MKHIP_ERR_ARG;
    if (level < 0 || level >= c->L) return MKHIP_ERR_LEVEL;
    auto it = c->mkf_groups.find(group);
    if (it == c->mkf_groups.end()) return MKHIP_ERR_NO_KEY;
    const int k = it->second.k;
    int nres = 0;
    double ns = scale_l * scale_r;
    int s = mkhip_rescale_count(c, level, scale_l * scale_r, min_scale, &nres, &ns);
    if (s != MKHIP_OK) return s;
    if (nres > 1) return MKHIP_ERR_SCALE;
    const int N = c->N, nl = level + 1;
    const size_t in = (size_t)(1 + k) * nl * N, ow = (size_t)(1 + k) * nl * N, words = 2 * in + ow;
    std::vector<int> sig = {1, group, k, level, nres};
    if (!c->single || c->single->sig != sig || c->single_io_words < words) {
        mkhip_plan_destroy(c->single);
        c->single = nullptr;
        if (c->single_io_words < words) {
            if (c->single_io) (void)hipFree(c->single_io);
            c->single_io = nullptr;
            c->single_io_words = 0;
            HIP_TRY(hipMalloc(&c->single_io, words * 8));
            c->single_io_words = words;
        }
        mkhip_mkf_op op = {};
        op.left = c->single_io;
        op.right = c->single_io + in;
        op.out = c->single_io + 2 * in;
        op.group = group;
        op.level = level;
        op.nres = nres;
        mkhip_plan *pl = nullptr;
        s = mkhip_plan_mkf_ckks(c, 1, &op, 0, &pl);
        if (s != MKHIP_OK) return s;
        pl->sig = sig;
        c->single = pl;
    }
    HIP_TRY(hipMemcpyAsync(c->single_io, left, in * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->single_io + in, right, in * 8, hipMemcpyHostToDevice, c->stream));
    s = mkhip_plan_launch(c->single);
    if (s != MKHIP_OK) return s;
    const size_t outw = (size_t)(1 + k) * (nl - nres) * N;
    HIP_TRY(hipMemcpyAsync(out, c->single_io + 2 * in, outw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (level_out) *level_out = level - nres;
    if (scale_out) *scale_out = ns;
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// MKF MK-BFV (software/mkf/mkf_cpu/bfv.go:212-252, BFVCase.Run) on a BFV context
// with gamma = 1.  Ciphertexts stay in the coefficient domain (bfv.go:69-85), so
// the fold switches return to it (common.go:97-103: InvNTT of every row, then
// ModDownQPtoQ).  Launches:
//   decompose       gadget digits of every component of both sides
//   mac_fold        KeyswitchHoistedNoModDown, every row back to coefficients
//   moddown_fold    ModDown of every switch summed with C0 (fold, common.go:109-126)
//   lift            modUpLeft / rescaleRight into R = Q u QMul (bfv.go:154-175)
//   tensor_ntt      NTT over R of the four lifted polys (in place)
//   tensor          mu0 = L0 R0, mu1 = L0 R1 + L1 R0, mu2 = L1 R1, then t INTT
//                   (tensorKaratsuba :179-192; mu1 equals the Karatsuba middle term
//                   residue for residue)
//   quantize        ModDown QMul -> Q of the three products (:194-210)
//   decompose_relin, mac_relin, moddown_relin  switch(p2) added to p0, p1 (:233-237)
//   materialize     Gamma copied to the k components (common.go:128-139)

int mkhip::plan_mkf_bfv(mkhip_ctx *c, int nops, const mkhip_mkf_op *ops, int use_graph, mkhip_plan **out,
                       const PlanMem &pm) {
    if (!c || !ops || nops <= 0 || !out) return MKHIP_ERR_ARG;
    *out = nullptr;
    if (!c->bfv || c->gamma != 1) return MKHIP_ERR_SCHEME;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    const int L = c->L, K = c->K, N = c->N, level = L - 1;
    for (int o = 0; o < nops; ++o) {
        const mkhip_mkf_op &op = ops[o];
        if (!op.left || !op.right || !op.out) return MKHIP_ERR_ARG;
        if (op.level != level) return MKHIP_ERR_LEVEL;
        if (op.nres != 0) return MKHIP_ERR_SCHEME;
        if (!c->mkf_groups.count(op.group)) return MKHIP_ERR_NO_KEY;
    }
    const int nl = L, rows = nl + K, nr = 2 * L;
    const int beta = (nl + c->alpha - 1) / c->alpha;
    const size_t limb = (size_t)N * 8, poly = (size_t)nl * limb, rpoly = (size_t)nr * limb;
    const size_t qp = (size_t)rows * limb, hoist = (size_t)beta * qp;

    struct OpWs {
        size_t D, facc, fold, lift, tq, p, Dr, racc;
        int k;
    };
    std::vector<OpWs> ws(nops);
    Carve cv;
    int kmax = 1;
    for (int o = 0; o < nops; ++o) {
        OpWs &w = ws[o];
        const int k = c->mkf_groups[ops[o].group].k;
        w.k = k;
        kmax = std::max(kmax, k);
        w.D = cv.take(2 * k * hoist);
        w.facc = cv.take(2 * k * 2 * qp);
        w.fold = cv.take(4 * poly);
        w.lift = cv.take(4 * rpoly);
        w.tq = cv.take(3 * rpoly);
        w.p = cv.take(3 * poly);
        w.Dr = cv.take(hoist);
        w.racc = cv.take(2 * qp);
    }
    mkhip_plan *pl = nullptr;
    int st = plan_create(c, cv.off, &pl, pm);
    if (st != MKHIP_OK) return st;
    char *B = (char *)pl->d_mem;
    auto P = [&](size_t off) { return (u64 *)(B + off); };

    std::vector<DecompDesc> dec, dec_relin;
    std::vector<Mac2Desc> mac_fold, mac_relin;
    std::vector<ModDownDesc> md_fold, md_relin;
    std::vector<BfvLiftDesc> lift;
    std::vector<NttJob> ntt;
    std::vector<TensorDesc> ten;
    std::vector<BfvQuantDesc> quant;
    std::vector<RescaleNttDesc> mat;
    for (int o = 0; o < nops; ++o) {
        const mkhip_mkf_op &op = ops[o];
        const OpWs &w = ws[o];
        const MkfGroup &g = c->mkf_groups[op.group];
        const int k = w.k;
        const u64 *side[2] = {op.left, op.right};
        auto facc = [&](int s, int p, int oo) { return P(w.facc) + (((size_t)s * k + p) * 2 + oo) * rows * N; };
        auto fold = [&](int s, int oo) { return P(w.fold) + ((size_t)s * 2 + oo) * nl * N; };
        auto lifted = [&](int s, int oo) { return P(w.lift) + ((size_t)s * 2 + oo) * nr * N; };
        auto prod = [&](int i) { return P(w.p) + (size_t)i * nl * N; };
        for (int s = 0; s < 2; ++s) {
            DecompDesc dd = {};
            dd.lazy_out = 1;  // digits feed Montgomery products only
            for (int p = 0; p < k; ++p) {
                dd.comp[p] = side[s] + (size_t)(1 + p) * nl * N;
                Mac2Desc md = {};
                md.D = P(w.D) + ((size_t)s * k + p) * beta * rows * N;
                md.swk = g.fold[p];
                md.acc[0] = facc(s, p, 0);
                md.acc[1] = facc(s, p, 1);
                md.level = level;
                md.intt_all = 1;
                mac_fold.push_back(md);
            }
            dd.D = P(w.D) + (size_t)s * k * beta * rows * N;
            dd.k = k;
            dd.level = level;
            dec.push_back(dd);
            for (int oo = 0; oo < 2; ++oo) {
                ModDownDesc m = {};
                for (int p = 0; p < k; ++p) m.src[p] = facc(s, p, oo);
                m.nsrc = k;
                m.base = oo == 0 ? side[s] : nullptr;  // fold: out[0] starts from C0
                m.dst = fold(s, oo);
                m.level = level;
                md_fold.push_back(m);
                BfvLiftDesc ld = {};
                ld.src = fold(s, oo);
                ld.dst = lifted(s, oo);
                ld.mode = s;  // left: ModUpQtoR, right: Rescale
                lift.push_back(ld);
                for (int l = 0; l < nr; ++l) {
                    NttJob j;
                    j.src = j.dst = lifted(s, oo) + (size_t)l * N;
                    j.mod = r_index(l, L, K);
                    j.reduce = 1;
                    ntt.push_back(j);
                }
            }
        }
        for (int i = 0; i < 3; ++i) {
            TensorDesc td = {};
            td.out = P(w.tq) + (size_t)i * nr * N;
            if (i == 1) {  // L0 R1 + L1 R0
                td.a0 = lifted(0, 0);
                td.b0 = lifted(1, 0);
                td.au = lifted(0, 1);
                td.bu = lifted(1, 1);
            } else {
                td.a0 = lifted(0, i / 2);
                td.b0 = lifted(1, i / 2);
                td.is_c0 = 1;
            }
            ten.push_back(td);
            BfvQuantDesc qd = {};
            qd.src = td.out;
            qd.dst = prod(i);
            quant.push_back(qd);
        }
        DecompDesc d2 = {};
        d2.lazy_out = 1;
        d2.comp[0] = prod(2);
        d2.D = P(w.Dr);
        d2.k = 1;
        d2.level = level;
        dec_relin.push_back(d2);
        Mac2Desc mr = {};
        mr.D = P(w.Dr);
        mr.swk = g.relin;
        mr.acc[0] = P(w.racc);
        mr.acc[1] = P(w.racc) + (size_t)rows * N;
        mr.level = level;
        mr.intt_all = 1;
        mac_relin.push_back(mr);
        for (int oo = 0; oo < 2; ++oo) {
            ModDownDesc m = {};
            m.src[0] = mr.acc[oo];
            m.nsrc = 1;
            m.base = prod(oo);
            m.dst = op.out + (size_t)oo * nl * N;  // c0, then the first copy of Gamma
            m.level = level;
            md_relin.push_back(m);
        }
        if (k > 1) {
            RescaleNttDesc rd = {};
            rd.x = op.out + (size_t)nl * N;
            rd.level = level;
            rd.nres = 0;
            for (int p = 1; p < k; ++p) rd.out[p - 1] = op.out + (size_t)(1 + p) * nl * N;
            rd.nout = k - 1;
            mat.push_back(rd);
        }
    }
    DescBlob blob;
    auto &S = pl->stages;
    Finish fin(c);
    S.push_back(make_stage(KDECOMP, blob.add(dec), (int)dec.size(), beta, rows, kmax, "decompose"));
    S.push_back(make_stage(KMAC2, blob.add(mac_fold), (int)mac_fold.size(), beta, rows, 1, "mac_fold"));
    fin.mac2(mac_fold, L, K, N);
    fin.emit(S, blob, "mac_fold");
    S.push_back(make_stage(KMODDOWN, blob.add(md_fold), (int)md_fold.size(), 0, 0, 1, "moddown_fold"));
    S.push_back(make_stage(KBFV_LIFT, blob.add(lift), (int)lift.size(), 0, 0, 1, "lift"));
    S.push_back(make_stage(KNTT_FWD, blob.add(ntt), (int)ntt.size(), 0, 0, 1, "tensor_ntt"));
    S.push_back(make_stage(KBFV_TENSOR, blob.add(ten), (int)ten.size(), nr, 0, 1, "tensor"));
    for (const TensorDesc &t : ten)
        for (int l = 0; l < nr; ++l) fin.row(t.out + (size_t)l * N, r_index(l, L, K), 1);
    fin.emit(S, blob, "tensor");
    S.push_back(make_stage(KBFV_QUANTIZE, blob.add(quant), (int)quant.size(), 0, 0, 1, "quantize"));
    S.push_back(make_stage(KDECOMP, blob.add(dec_relin), (int)dec_relin.size(), beta, rows, 1, "decompose_relin"));
    S.push_back(make_stage(KMAC2, blob.add(mac_relin), (int)mac_relin.size(), beta, rows, 1, "mac_relin"));
    fin.mac2(mac_relin, L, K, N);
    fin.emit(S, blob, "mac_relin");
    S.push_back(make_stage(KMODDOWN, blob.add(md_relin), (int)md_relin.size(), 0, 0, 1, "moddown_relin"));
    if (!mat.empty())
        S.push_back(make_stage(KRESCALE_NTT, blob.add(mat), (int)mat.size(), nl, 0, 1, "materialize"));
    st = plan_finish(pl, blob, use_graph);
    if (st != MKHIP_OK) return st;
    *out = pl;
    return MKHIP_OK;
}

extern "C" int mkhip_plan_mkf_bfv(mkhip_ctx *c, int nops, const mkhip_mkf_op *ops, int use_graph,
                                     mkhip_plan **out) {
    return plan_mkf_bfv(c, nops, ops, use_graph, out, PlanMem());
}

// MKF BFVCase.Run on host buffers: left/right [1+k][L][N] coefficient domain;
// out [1+k][L][N]: c0 then k copies of Gamma.
extern "C" int mkhip_mkf_bfv_mulrelin(mkhip_ctx *c, int group, const uint64_t *left, const uint64_t *right,
                                      uint64_t *out) {
    if (!c || !left || !right || !out) return MKHIP_ERR_ARG;
    if (!c->bfv || c->gamma != 1) return MKHIP_ERR_SCHEME;
    auto it = c->mkf_groups.find(group);
    if (it == c->mkf_groups.end()) return MKHIP_ERR_NO_KEY;
    const int k = it->second.k, N = c->N, nl = c->L;
    const size_t in = (size_t)(1 + k) * nl * N, words = 3 * in;
    std::vector<int> sig = {3, group, k};
    if (!c->single || c->single->sig != sig || c->single_io_words < words) {
        mkhip_plan_destroy(c->single);
        c->single = nullptr;
        if (c->single_io_words < words) {
            if (c->single_io) (void)hipFree(c->single_io);
            c->single_io = nullptr;
            c->single_io_words = 0;
            HIP_TRY(hipMalloc(&c->single_io, words * 8));
            c->single_io_words = words;
        }
        mkhip_mkf_op op = {};
        op.left = c->single_io;
        op.right = c->single_io + in;
        op.out = c->single_io + 2 * in;
        op.group = group;
        op.level = nl - 1;
        op.nres = 0;
        mkhip_plan *pl = nullptr;
        int s = mkhip_plan_mkf_bfv(c, 1, &op, 0, &pl);
        if (s != MKHIP_OK) return s;
        pl->sig = sig;
        c->single = pl;
    }
    HIP_TRY(hipMemcpyAsync(c->single_io, left, in * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->single_io + in, right, in * 8, hipMemcpyHostToDevice, c->stream));
    int s = mkhip_plan_launch(c->single);
    if (s != MKHIP_OK) return s;
    HIP_TRY(hipMemcpyAsync(out, c->single_io + 2 * in, in * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}
