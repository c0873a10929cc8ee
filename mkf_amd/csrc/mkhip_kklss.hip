// KKLSS MK-CKKS mul + relin + rescale: plan construction and the host-pointer
// entry point.
//
// The orchestration follows mkrlwe.KeySwitcher.MulAndRelinHoisted
// (software/kklss/mkrlwe/keyswitch_hoisted.go:49-194) preceded by the hoisted
// decompositions of mkckks.Evaluator.MulRelinNew (mkckks/evaluator.go:419-446)
// and followed by Evaluator.Rescale (:360-401), as nine batched launches:
//
//   decompose      D(op0_i), D(op1_j) for every party                      (:432-438)
//   mac_xy         x = MForm(sum d_i D0_i), y = MForm(sum b_j D1_j)         (:84-123)
//   tensor_ntt     NTT of every input component                           (:126-150)
//   tensor         c0 c0', c0 c'_j + c_j c0', INTT                          (:126-153)
//   ext_xy         Ext(D1_j, x), Ext(D0_i, y): MACs + INTT                  (:155-170)
//   moddown_xy     ctOut_j += ModDown(.), t_i = ModDown(.)
//   decompose_t    D(t_i)                                                  (:175)
//   ext_vu         Ext(D(t_i), v_i), Ext(D(t_i), u) + INTT                   (:177-193)
//   moddown_rescale ctOut_0 += sum_i ModDown(.), ctOut_i += ModDown(.), rescale
#include "mkhip_internal.hpp"

#include <algorithm>

using namespace mkhip;

namespace {

int find_id(const int *ids, int k, int id) {
    for (int i = 0; i < k; ++i)
        if (ids[i] == id) return i;
    return -1;
}

}  // namespace

int mkhip::plan_kklss_ckks(mkhip_ctx *c, int nops, const mkhip_kklss_op *ops, int use_graph, mkhip_plan **out,
                       const PlanMem &pm) {
    if (!c || !ops || nops <= 0 || !out) return MKHIP_ERR_ARG;
    *out = nullptr;
    if (c->alpha == 0) return MKHIP_ERR_GADGET;
    if (c->bfv) return MKHIP_ERR_SCHEME;
    if (!c->d_crs_u) return MKHIP_ERR_NO_KEY;
    const int level = ops[0].level;
    const int N = c->N, K = c->K;
    for (int o = 0; o < nops; ++o) {
        const mkhip_kklss_op &op = ops[o];
        if (!op.op0 || !op.op1 || !op.out) return MKHIP_ERR_ARG;
        if (op.level != level) return MKHIP_ERR_LEVEL;
        if (op.level < 0 || op.level >= c->L) return MKHIP_ERR_LEVEL;
        if (op.level0 >= c->L || op.level1 >= c->L) return MKHIP_ERR_LEVEL;
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
    const size_t keyw = (size_t)beta * (c->L + K) * limb;

    // ---- workspace layout ----
    Carve cv;
    struct OpWs {
        size_t D0, D1, x, y, tnt, ct, accx, accy, t, Dt, accv, accu;
        int ku;
        int ids[2 * MKHIP_MAX_PARTIES];
    };
    std::vector<OpWs> ws(nops);
    for (int o = 0; o < nops; ++o) {
        const mkhip_kklss_op &op = ops[o];
        OpWs &w = ws[o];
        int ku = 0;  // union of ids, ascending
        for (int i = 0; i < op.k0; ++i)
            if (find_id(w.ids, ku, op.ids0[i]) < 0) w.ids[ku++] = op.ids0[i];
        for (int i = 0; i < op.k1; ++i)
            if (find_id(w.ids, ku, op.ids1[i]) < 0) w.ids[ku++] = op.ids1[i];
        std::sort(w.ids, w.ids + ku);
        w.ku = ku;
        if (ku > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
        w.D0 = cv.take(op.k0 * hoist);
        w.D1 = cv.take(op.k1 * hoist);
        w.x = cv.take(keyw);  // full-level switching keys (k_ext reads them like keys)
        w.y = cv.take(keyw);
        w.tnt = cv.take((c->alpha == 1 ? 2 : 2 + op.k0 + op.k1) * poly);  // alpha 1: the c0s only
        w.ct = cv.take((1 + ku) * poly);
        w.accx = cv.take(op.k1 * qp);
        w.accy = cv.take(op.k0 * qp);
        w.t = cv.take(op.k0 * poly);
        w.Dt = cv.take(op.k0 * hoist);
        w.accv = cv.take(op.k0 * qp);
        w.accu = cv.take(op.k0 * qp);
    }
    mkhip_plan *pl = nullptr;
    int st = plan_create(c, cv.off, &pl, pm);
    if (st != MKHIP_OK) return st;
    char *B = (char *)pl->d_mem;
    auto P = [&](size_t off) { return (u64 *)(B + off); };

    std::vector<DecompDesc> dec1, dec2;
    std::vector<MacDesc> mac;
    std::vector<NttJob> ntt;
    std::vector<TensorDesc> ten;
    std::vector<ExtDesc> ext1, ext2;
    std::vector<std::vector<ExtDesc> > ext2p(MKHIP_MAX_PARTIES);  // by party slot: key rows shared across ops
    std::vector<ModDownDesc> md1, md2;
    int kmax1 = 1, kmax2 = 1;
    for (int o = 0; o < nops; ++o) {
        const mkhip_kklss_op &op = ops[o];
        const OpWs &w = ws[o];
        // each operand at its own level (its components' limb count); decomposed
        // there, read at the product's level
        const int lv0 = std::max(op.level0, level), lv1 = std::max(op.level1, level);
        const size_t cs0 = (size_t)(lv0 + 1) * N, cs1 = (size_t)(lv1 + 1) * N;
        // decompositions of both sides; x / y accumulations (key-stationary)
        DecompDesc s0 = {}, s1 = {};
        MacDesc m0 = {}, m1 = {};
        s0.lazy_out = s1.lazy_out = 1;  // digits feed Montgomery products only
        s0.src_level = lv0;
        s1.src_level = lv1;
        for (int p = 0; p < op.k0; ++p) {
            s0.comp[p] = op.op0 + (size_t)(1 + p) * cs0;
            m0.key[p] = s0.key[p] = c->parties[op.ids0[p]].d;
        }
        s0.D = P(w.D0);
        m0.D = s0.D;
        s0.k = m0.k = op.k0;
        s0.level = m0.level = level;
        m0.xy = s0.xy = P(w.x);
        for (int p = 0; p < op.k1; ++p) {
            s1.comp[p] = op.op1 + (size_t)(1 + p) * cs1;
            m1.key[p] = s1.key[p] = c->parties[op.ids1[p]].b;
        }
        s1.D = P(w.D1);
        m1.D = s1.D;
        s1.k = m1.k = op.k1;
        s1.level = m1.level = level;
        m1.xy = s1.xy = P(w.y);
        dec1.push_back(s0);
        dec1.push_back(s1);
        mac.push_back(m0);
        mac.push_back(m1);
        kmax1 = std::max(kmax1, std::max(op.k0, op.k1));
        kmax2 = std::max(kmax2, op.k0);
        // NTT of every input component: op0 c0, op0 comps, op1 c0, op1 comps.  With
        // one prime per digit, limb l of a decomposed component transformed mod q_l
        // is row l of its digit l (keyswitch.go:22-32: the digit is the limb itself),
        // so only the c0s are transformed here and the tensor reads the other
        // components from the decomposition's diagonal rows.
        u64 *tnt = P(w.tnt);
        const int nc0 = 1 + op.k0, nc1 = 1 + op.k1;
        const bool diag = c->alpha == 1;
        for (int comp = 0; comp < nc0 + nc1; ++comp) {
            if (diag && comp != 0 && comp != nc0) continue;
            const u64 *src = comp < nc0 ? op.op0 + (size_t)comp * cs0 : op.op1 + (size_t)(comp - nc0) * cs1;
            for (int l = 0; l < nl; ++l) {
                NttJob j;
                j.src = src + (size_t)l * N;
                j.dst = tnt + ((size_t)(diag ? (comp == 0 ? 0 : 1) : comp) * nl + l) * N;
                j.mod = l;
                j.reduce = 0;
                ntt.push_back(j);
            }
        }
        const u64 *a0 = tnt, *b0 = tnt + (size_t)(diag ? 1 : nc0) * nl * N;
        // tensor into ct (coefficient domain)
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
            if (diag) {  // limb l at D[p][digit l][row l]
                td.au = p0 >= 0 ? P(w.D0) + (size_t)p0 * beta * rows * N : nullptr;
                td.bu = p1 >= 0 ? P(w.D1) + (size_t)p1 * beta * rows * N : nullptr;
                td.ustride = rows + 1;
            } else {
                td.au = p0 >= 0 ? tnt + (size_t)(1 + p0) * nl * N : nullptr;
                td.bu = p1 >= 0 ? tnt + (size_t)(nc0 + 1 + p1) * nl * N : nullptr;
            }
            td.out = ct + (size_t)(1 + u) * nl * N;
            td.is_c0 = 0;
            ten.push_back(td);
        }
        // Ext(D1_j, x) -> ct_j ; Ext(D0_i, y) -> t_i
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
        // D(t_i)
        DecompDesc s2 = {};
        s2.lazy_out = 1;
        for (int p = 0; p < op.k0; ++p) s2.comp[p] = P(w.t) + (size_t)p * nl * N;
        s2.D = P(w.Dt);
        s2.k = op.k0;
        s2.level = level;
        dec2.push_back(s2);
        // Ext(D(t_i), v_i), Ext(D(t_i), u)
        for (int p = 0; p < op.k0; ++p) {
            ExtDesc ed = {};
            ed.D = P(w.Dt) + (size_t)p * beta * rows * N;
            ed.key[0] = c->parties[op.ids0[p]].v;
            ed.key[1] = c->d_crs_u;
            ed.acc[0] = P(w.accv) + (size_t)p * rows * N;
            ed.acc[1] = P(w.accu) + (size_t)p * rows * N;
            ed.nkey = 2;
            ed.level = level;
            ext2p[p].push_back(ed);
        }
        // final ModDowns + rescale into the output
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
    for (auto &v : ext2p) ext2.insert(ext2.end(), v.begin(), v.end());
    DescBlob blob;
    auto &S = pl->stages;
    if (c->alpha == 1 && c->tune_fused_decomp) {  // one prime per digit: decomposition fused with the MAC
        S.push_back(make_stage(KDECOMP_MAC, blob.add(dec1), nops, beta, rows, 1, "decompose_mac"));
    } else {
        S.push_back(make_stage(KDECOMP, blob.add(dec1), (int)dec1.size(), beta, rows, kmax1, "decompose"));
        S.push_back(make_stage(KMAC_XY, blob.add(mac), nops, beta, rows, 1, "mac_xy"));
    }
    Finish fin(c);
    S.push_back(make_stage(KNTT_FWD, blob.add(ntt), (int)ntt.size(), 0, 0, 1, "tensor_ntt"));
    S.push_back(make_stage(KTENSOR, blob.add(ten), (int)ten.size(), nl, 0, 1, "tensor"));
    for (const TensorDesc &t : ten)
        for (int l = 0; l < nl; ++l) fin.row(t.out + (size_t)l * N, l);
    fin.emit(S, blob, "tensor");
    // external products on half-limb workgroups (N = 2^15, or the knob at 2^14) leave
    // their last inverse stage and N^-1 scaling to the ModDowns, the only readers
    // of their rows (no finish pass over the accumulators)
    if (Finish(c, ext_split(c->T)).on) {
        for (ModDownDesc &m : md1) m.unfinished = 1;
        for (ModDownDesc &m : md2) m.unfinished = 1;
    }
    S.push_back(make_stage(KEXT1, blob.add(ext1), (int)ext1.size(), beta, rows, 1, "ext_xy"));
    S.push_back(make_stage(KMODDOWN, blob.add(md1), (int)md1.size(), 0, 0, 1, "moddown_xy"));
    S.push_back(make_stage(KDECOMP, blob.add(dec2), (int)dec2.size(), beta, rows, kmax2, "decompose_t"));
    S.push_back(make_stage(KEXT2, blob.add(ext2), (int)ext2.size(), beta, rows, 1, "ext_vu"));
    S.push_back(make_stage(KMODDOWN_FINAL, blob.add(md2), (int)md2.size(), 0, 0, 1, "moddown_rescale"));
    st = plan_finish(pl, blob, use_graph);
    if (st != MKHIP_OK) return st;
    *out = pl;
    return MKHIP_OK;
}

extern "C" int mkhip_plan_kklss_ckks(mkhip_ctx *c, int nops, const mkhip_kklss_op *ops, int use_graph,
                                     mkhip_plan **out) {
    return plan_kklss_ckks(c, nops, ops, use_graph, out, PlanMem());
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
    // The product is taken at the common level (evaluator.go:577-581) and each
    // operand is decomposed at its own level (:433-439).  With one prime per digit
    // the digits do not depend on the level, so the operands are staged at the
    // common level; with alpha > 1 a digit can span limbs above it, so every limb
    // of each operand is staged.
    const int level = std::min(level0, level1);
    const int lv0 = c->alpha == 1 ? level : level0, lv1 = c->alpha == 1 ? level : level1;
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
    const size_t in0 = (size_t)(1 + k0) * (lv0 + 1) * N, in1 = (size_t)(1 + k1) * (lv1 + 1) * N;
    const size_t ow = (size_t)(1 + ku) * nl * N;
    const size_t words = in0 + in1 + ow;
    // shape signature of the cached plan
    std::vector<int> sig = {0, k0, k1, level, nres, lv0, lv1};
    sig.insert(sig.end(), ids0, ids0 + k0);
    sig.insert(sig.end(), ids1, ids1 + k1);
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
        op.level0 = lv0;
        op.level1 = lv1;
        mkhip_plan *pl = nullptr;
        s = mkhip_plan_kklss_ckks(c, 1, &op, 0, &pl);
        if (s != MKHIP_OK) return s;
        pl->sig = sig;
        c->single = pl;
    }
    // stage limbs 0..lv of each component
    for (int comp = 0; comp <= k0; ++comp)
        HIP_TRY(hipMemcpyAsync(c->single_io + (size_t)comp * (lv0 + 1) * N, op0 + (size_t)comp * (level0 + 1) * N,
                               (size_t)(lv0 + 1) * N * 8, hipMemcpyHostToDevice, c->stream));
    for (int comp = 0; comp <= k1; ++comp)
        HIP_TRY(hipMemcpyAsync(c->single_io + in0 + (size_t)comp * (lv1 + 1) * N,
                               op1 + (size_t)comp * (level1 + 1) * N, (size_t)(lv1 + 1) * N * 8,
                               hipMemcpyHostToDevice, c->stream));
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
