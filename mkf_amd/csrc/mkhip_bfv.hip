// KKLSS MK-BFV mul + relin: plan construction, the host-pointer entry point and
// the R-basis primitives.
//
// The orchestration follows mkbfv.Evaluator.mulRelinHoisted
// (software/kklss/mkbfv/evaluator.go:121-152) and
// KeySwitcher.MulAndRelinBFVHoisted (mkbfv/keyswitch_hoisted.go:50-252), as
// eleven batched launches:
//
//   lift          op0 -> ModUpQtoR, op1 -> Rescale, every component   (evaluator.go:127-135)
//   decompose     DecomposeBFV: 2 beta digits of every party's R poly  (:140-146, keyswitch.go:58-83)
//   mac_xy        [x1;x2] = MForm(sum [d1;d2] D0), [y1;y2] likewise   (keyswitch_hoisted.go:100-148)
//   tensor_ntt    NTT over R of every lifted component (in place)      (:156-157, 172-194)
//   tensor        products over R, t x, INTT                           (:154-200)
//   quantize      ModDown R -> Q into ctOut                            (basis_extension.go:69-85)
//   ext_xy        ExtBFV(D1_j, x), ExtBFV(D0_i, y): MACs + INTT         (:205-237)
//   moddown_xy    ctOut_j += ModDown(.), t_i = ModDown(.)
//   decompose_t   D(t_i) (KKLSS gadget over Q)                         (:237)
//   ext_vu        Ext(D(t_i), v_i), Ext(D(t_i), u) + INTT                (:239-247)
//   moddown_final ctOut_0 += sum_i ModDown(.), ctOut_i += ModDown(.)
//
// With alpha = 1 (the only case a BFV context accepts) the decomposition of an
// R poly copies R limb d into every Q u P row (the copy path of DecomposeAndSplit),
// so the existing decomposition kernel serves with 2 beta = 2L digits.
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

int mkhip::plan_kklss_bfv(mkhip_ctx *c, int nops, const mkhip_kklss_op *ops, int use_graph, mkhip_plan **out,
                       const PlanMem &pm) {
    if (!c || !ops || nops <= 0 || !out) return MKHIP_ERR_ARG;
    *out = nullptr;
    if (!c->bfv || c->gamma != 2 || c->alpha != 1) return MKHIP_ERR_SCHEME;
    if (!c->d_crs_u) return MKHIP_ERR_NO_KEY;
    const int L = c->L, K = c->K, N = c->N;
    const int level = L - 1;  // BFV ciphertexts carry every Q limb
    for (int o = 0; o < nops; ++o) {
        const mkhip_kklss_op &op = ops[o];
        if (!op.op0 || !op.op1 || !op.out) return MKHIP_ERR_ARG;
        if (op.level != level) return MKHIP_ERR_LEVEL;
        if (op.nres != 0) return MKHIP_ERR_SCHEME;
        if (op.k0 < 0 || op.k1 < 0 || op.k0 > MKHIP_MAX_PARTIES || op.k1 > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
        for (int i = 0; i < op.k0; ++i) {
            auto it = c->parties.find(op.ids0[i]);
            if (it == c->parties.end() || !it->second.bfv) return MKHIP_ERR_UNKNOWN_PARTY;
        }
        for (int i = 0; i < op.k1; ++i) {
            auto it = c->parties.find(op.ids1[i]);
            if (it == c->parties.end() || !it->second.bfv) return MKHIP_ERR_UNKNOWN_PARTY;
        }
    }
    const int nl = L, rows = L + K, beta = L, nr = 2 * L;
    const size_t limb = (size_t)N * 8, poly = (size_t)nl * limb, rpoly = (size_t)nr * limb;
    const size_t qp = (size_t)rows * limb, hoist = (size_t)beta * qp, hoist2 = 2 * hoist;
    const size_t keyw2 = (size_t)2 * beta * (L + K) * limb;

    Carve cv;
    struct OpWs {
        size_t lift, D0, D1, x, y, tq, ct, accx, accy, t, Dt, accv, accu;
        int ku;
        int ids[2 * MKHIP_MAX_PARTIES];
    };
    std::vector<OpWs> ws(nops);
    for (int o = 0; o < nops; ++o) {
        const mkhip_kklss_op &op = ops[o];
        OpWs &w = ws[o];
        int ku = 0;
        for (int i = 0; i < op.k0; ++i)
            if (find_id(w.ids, ku, op.ids0[i]) < 0) w.ids[ku++] = op.ids0[i];
        for (int i = 0; i < op.k1; ++i)
            if (find_id(w.ids, ku, op.ids1[i]) < 0) w.ids[ku++] = op.ids1[i];
        std::sort(w.ids, w.ids + ku);
        w.ku = ku;
        if (ku > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
        w.lift = cv.take((2 + op.k0 + op.k1) * rpoly);  // lifted components; NTT'd in place later
        w.D0 = cv.take(op.k0 * hoist2);
        w.D1 = cv.take(op.k1 * hoist2);
        w.x = cv.take(keyw2);
        w.y = cv.take(keyw2);
        w.tq = cv.take((1 + ku) * rpoly);
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

    std::vector<BfvLiftDesc> lift;
    std::vector<DecompDesc> dec1, dec2;
    std::vector<MacDesc> mac;
    std::vector<NttJob> ntt;
    std::vector<TensorDesc> ten;
    std::vector<BfvQuantDesc> quant;
    std::vector<ExtDesc> ext1, ext2;
    std::vector<std::vector<ExtDesc> > ext2p(MKHIP_MAX_PARTIES);  // by party slot: key rows shared across ops
    std::vector<ModDownDesc> md1, md2;
    int kmax1 = 1, kmax2 = 1;
    for (int o = 0; o < nops; ++o) {
        const mkhip_kklss_op &op = ops[o];
        const OpWs &w = ws[o];
        const int nc0 = 1 + op.k0, nc1 = 1 + op.k1;
        u64 *lf = P(w.lift);
        auto lifted = [&](int side, int comp) { return lf + ((size_t)(side ? nc0 + comp : comp) * nr) * N; };
        for (int comp = 0; comp < nc0 + nc1; ++comp) {
            BfvLiftDesc d = {};
            const int side = comp < nc0 ? 0 : 1, cc = side ? comp - nc0 : comp;
            d.src = (side ? op.op1 : op.op0) + (size_t)cc * nl * N;
            d.dst = lifted(side, cc);
            d.mode = side;
            lift.push_back(d);
        }
        DecompDesc s0 = {}, s1 = {};
        MacDesc m0 = {}, m1 = {};
        s0.lazy_out = s1.lazy_out = 1;  // digits feed Montgomery products only
        for (int p = 0; p < op.k0; ++p) {
            s0.comp[p] = lifted(0, 1 + p);
            m0.key[p] = s0.key[p] = c->parties[op.ids0[p]].d;
        }
        s0.D = P(w.D0);
        m0.D = s0.D;
        s0.k = m0.k = op.k0;
        s0.level = m0.level = level;
        m0.xy = s0.xy = P(w.x);
        for (int p = 0; p < op.k1; ++p) {
            s1.comp[p] = lifted(1, 1 + p);
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
        // NTT over R of every lifted component, in place (raw limbs reduced first)
        for (int comp = 0; comp < nc0 + nc1; ++comp)
            for (int l = 0; l < nr; ++l) {
                NttJob j;
                j.src = lf + ((size_t)comp * nr + l) * N;
                j.dst = lf + ((size_t)comp * nr + l) * N;
                j.mod = r_index(l, L, K);
                j.reduce = 1;
                ntt.push_back(j);
            }
        const u64 *a0 = lifted(0, 0), *b0 = lifted(1, 0);
        u64 *tq = P(w.tq), *ct = P(w.ct);
        for (int u = 0; u <= w.ku; ++u) {
            TensorDesc td = {};
            td.a0 = a0;
            td.b0 = b0;
            td.out = tq + (size_t)u * nr * N;
            if (u == 0) {
                td.is_c0 = 1;
            } else {
                const int id = w.ids[u - 1];
                const int p0 = find_id(op.ids0, op.k0, id), p1 = find_id(op.ids1, op.k1, id);
                td.au = p0 >= 0 ? lifted(0, 1 + p0) : nullptr;
                td.bu = p1 >= 0 ? lifted(1, 1 + p1) : nullptr;
            }
            ten.push_back(td);
            BfvQuantDesc qd = {};
            qd.src = td.out;
            qd.dst = ct + (size_t)u * nl * N;
            quant.push_back(qd);
        }
        // ExtBFV(D1_j, [x1;x2]) -> ct_j ; ExtBFV(D0_i, [y1;y2]) -> t_i
        for (int p = 0; p < op.k1; ++p) {
            ExtDesc ed = {};
            ed.D = P(w.D1) + (size_t)p * 2 * beta * rows * N;
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
            ed.D = P(w.D0) + (size_t)p * 2 * beta * rows * N;
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
        DecompDesc s2 = {};
        s2.lazy_out = 1;
        for (int p = 0; p < op.k0; ++p) s2.comp[p] = P(w.t) + (size_t)p * nl * N;
        s2.D = P(w.Dt);
        s2.k = op.k0;
        s2.level = level;
        dec2.push_back(s2);
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
        {
            ModDownDesc md = {};
            for (int p = 0; p < op.k0; ++p) md.src[p] = P(w.accv) + (size_t)p * rows * N;
            md.nsrc = op.k0;
            md.dst = ct;
            md.level = level;
            md.accumulate = 1;
            md.final_ = 1;
            md.nres = 0;
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
            md.nres = 0;
            md.out = op.out + (size_t)(1 + u) * nl * N;
            md2.push_back(md);
        }
    }
    for (auto &v : ext2p) ext2.insert(ext2.end(), v.begin(), v.end());
    DescBlob blob;
    auto &S = pl->stages;
    S.push_back(make_stage(KBFV_LIFT, blob.add(lift), (int)lift.size(), 0, 0, 1, "lift"));
    if (c->tune_fused_decomp) {  // decomposition fused with the x / y MAC
        S.push_back(make_stage(KDECOMP_MAC, blob.add(dec1), nops, 2 * beta, rows, 1, "decompose_mac"));
    } else {
        S.push_back(make_stage(KDECOMP, blob.add(dec1), (int)dec1.size(), 2 * beta, rows, kmax1, "decompose"));
        S.push_back(make_stage(KMAC_XY, blob.add(mac), nops, 2 * beta, rows, 1, "mac_xy"));
    }
    Finish fin(c);
    S.push_back(make_stage(KNTT_FWD, blob.add(ntt), (int)ntt.size(), 0, 0, 1, "tensor_ntt"));
    S.push_back(make_stage(KBFV_TENSOR, blob.add(ten), (int)ten.size(), nr, 0, 1, "tensor"));
    for (const TensorDesc &t : ten)
        for (int l = 0; l < nr; ++l) fin.row(t.out + (size_t)l * N, l < c->L ? l : K + l, 1);
    fin.emit(S, blob, "tensor");
    S.push_back(make_stage(KBFV_QUANTIZE, blob.add(quant), (int)quant.size(), 0, 0, 1, "quantize"));
    // external products on half-limb workgroups (N = 2^15, or the knob at 2^14) leave
    // their last inverse stage and N^-1 scaling to the ModDowns, the only readers
    // of their rows (no finish pass over the accumulators)
    if (Finish(c, ext_split(c->T)).on) {
        for (ModDownDesc &m : md1) m.unfinished = 1;
        for (ModDownDesc &m : md2) m.unfinished = 1;
    }
    S.push_back(make_stage(KEXT1, blob.add(ext1), (int)ext1.size(), 2 * beta, rows, 1, "ext_xy"));
    S.push_back(make_stage(KMODDOWN, blob.add(md1), (int)md1.size(), 0, 0, 1, "moddown_xy"));
    S.push_back(make_stage(KDECOMP, blob.add(dec2), (int)dec2.size(), beta, rows, kmax2, "decompose_t"));
    S.push_back(make_stage(KEXT2, blob.add(ext2), (int)ext2.size(), beta, rows, 1, "ext_vu"));
    S.push_back(make_stage(KMODDOWN_FINAL, blob.add(md2), (int)md2.size(), 0, 0, 1, "moddown_final"));
    st = plan_finish(pl, blob, use_graph);
    if (st != MKHIP_OK) return st;
    *out = pl;
    return MKHIP_OK;
}

extern "C" int mkhip_plan_kklss_bfv(mkhip_ctx *c, int nops, const mkhip_kklss_op *ops, int use_graph,
                                     mkhip_plan **out) {
    return plan_kklss_bfv(c, nops, ops, use_graph, out, PlanMem());
}

// ---------------------------------------------------------------------------
// host-pointer entry point (mkbfv.Evaluator.MulRelinNew, evaluator.go:85-91)

extern "C" int mkhip_kklss_bfv_mulrelin(mkhip_ctx *c, const uint64_t *op0, const int *ids0, int k0,
                                        const uint64_t *op1, const int *ids1, int k1, uint64_t *out, int *ids_out,
                                        int *ku_out) {
    if (!c || !op0 || !op1 || !out || (k0 && !ids0) || (k1 && !ids1) || k0 < 0 || k1 < 0) return MKHIP_ERR_ARG;
    if (k0 > MKHIP_MAX_PARTIES || k1 > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
    if (!c->bfv || c->gamma != 2 || c->alpha != 1) return MKHIP_ERR_SCHEME;
    const int N = c->N, nl = c->L, level = c->L - 1;
    int ids[2 * MKHIP_MAX_PARTIES], ku = 0;
    for (int i = 0; i < k0; ++i)
        if (find_id(ids, ku, ids0[i]) < 0) ids[ku++] = ids0[i];
    for (int i = 0; i < k1; ++i)
        if (find_id(ids, ku, ids1[i]) < 0) ids[ku++] = ids1[i];
    if (ku > MKHIP_MAX_PARTIES) return MKHIP_ERR_TOO_MANY;
    std::sort(ids, ids + ku);
    const size_t in0 = (size_t)(1 + k0) * nl * N, in1 = (size_t)(1 + k1) * nl * N, ow = (size_t)(1 + ku) * nl * N;
    const size_t words = in0 + in1 + ow;
    std::vector<int> sig = {2, k0, k1};
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
        op.nres = 0;
        mkhip_plan *pl = nullptr;
        int s = mkhip_plan_kklss_bfv(c, 1, &op, 0, &pl);
        if (s != MKHIP_OK) return s;
        pl->sig = sig;
        c->single = pl;
    }
    HIP_TRY(hipMemcpyAsync(c->single_io, op0, in0 * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->single_io + in0, op1, in1 * 8, hipMemcpyHostToDevice, c->stream));
    int s = mkhip_plan_launch(c->single);
    if (s != MKHIP_OK) return s;
    HIP_TRY(hipMemcpyAsync(out, c->single_io + in0 + in1, ow * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (ids_out)
        for (int i = 0; i < ku; ++i) ids_out[i] = ids[i];
    if (ku_out) *ku_out = ku;
    return MKHIP_OK;
}

// ---------------------------------------------------------------------------
// R-basis primitives (mkbfv FastBasisExtender, basis_extension.go:50-104;
// KeySwitcher.DecomposeBFV, keyswitch.go:58-83), host buffers

namespace {

struct Scratch {
    void *p = nullptr;
    ~Scratch() {
        if (p) (void)hipFree(p);
    }
};

int lift_one(mkhip_ctx *c, const uint64_t *in, uint64_t *out, int mode) {
    if (!c || !in || !out) return MKHIP_ERR_ARG;
    if (!c->bfv) return MKHIP_ERR_SCHEME;
    const int N = c->N, L = c->L;
    const size_t inw = (size_t)L * N, outw = 2 * inw;
    Scratch s;
    HIP_TRY(hipMalloc(&s.p, (inw + outw) * 8 + sizeof(BfvLiftDesc) + 256));
    u64 *din = (u64 *)s.p, *dout = din + inw;
    BfvLiftDesc *dd = (BfvLiftDesc *)(dout + outw);
    BfvLiftDesc h = {};
    h.src = din;
    h.dst = dout;
    h.mode = mode;
    HIP_TRY(hipMemcpyAsync(din, in, inw * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dd, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_bfv(KBFV_LIFT, c->T, dd, 1, 0, 0, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, outw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

}  // namespace

extern "C" int mkhip_bfv_modup_q_to_r(mkhip_ctx *c, const uint64_t *in, uint64_t *out) {
    return lift_one(c, in, out, 0);
}

extern "C" int mkhip_bfv_rescale_q_to_r(mkhip_ctx *c, const uint64_t *in, uint64_t *out) {
    return lift_one(c, in, out, 1);
}

extern "C" int mkhip_bfv_quantize(mkhip_ctx *c, const uint64_t *in_ntt, uint64_t *out) {
    if (!c || !in_ntt || !out) return MKHIP_ERR_ARG;
    if (!c->bfv) return MKHIP_ERR_SCHEME;
    const int N = c->N, L = c->L, K = c->K, nr = 2 * L;
    const size_t rw = (size_t)nr * N, ow = (size_t)L * N;
    Scratch s;
    HIP_TRY(hipMalloc(&s.p, (2 * rw + ow) * 8 + sizeof(NttJob) * nr + sizeof(BfvQuantDesc) + 512));
    u64 *din = (u64 *)s.p, *dx = din + rw, *dout = dx + rw;
    NttJob *djobs = (NttJob *)(dout + ow);
    BfvQuantDesc *dq = (BfvQuantDesc *)(djobs + nr);
    std::vector<NttJob> jobs(nr);
    for (int l = 0; l < nr; ++l) {
        jobs[l].src = din + (size_t)l * N;
        jobs[l].dst = dx + (size_t)l * N;
        jobs[l].mod = r_index(l, L, K);
        jobs[l].reduce = 1 | 2;  // t N^-1: MulScalar(t) then InvNTT
    }
    BfvQuantDesc h = {};
    h.src = dx;
    h.dst = dout;
    HIP_TRY(hipMemcpyAsync(din, in_ntt, rw * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(djobs, jobs.data(), sizeof(NttJob) * nr, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dq, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_kernel(KNTT_INV, c->T, djobs, nr, 0, 0, c->stream));
    {
        std::vector<int> mods(nr);
        for (int l = 0; l < nr; ++l) mods[l] = r_index(l, L, K);
        const int st = finish_rows(c, dx, mods.data(), nr, 1);
        if (st != MKHIP_OK) return st;
    }
    HIP_TRY(launch_bfv(KBFV_QUANTIZE, c->T, dq, 1, 0, 0, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, ow * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}

extern "C" int mkhip_bfv_decompose(mkhip_ctx *c, const uint64_t *aR, uint64_t *out) {
    if (!c || !aR || !out) return MKHIP_ERR_ARG;
    if (!c->bfv || c->gamma != 2 || c->alpha != 1) return MKHIP_ERR_SCHEME;
    const int N = c->N, L = c->L, rows = L + c->K;
    const size_t inw = (size_t)2 * L * N, outw = (size_t)2 * L * rows * N;
    Scratch s;
    HIP_TRY(hipMalloc(&s.p, (inw + outw) * 8 + sizeof(DecompDesc) + 256));
    u64 *din = (u64 *)s.p, *dout = din + inw;
    DecompDesc *dd = (DecompDesc *)(dout + outw);
    DecompDesc h = {};
    h.comp[0] = din;
    h.D = dout;
    h.k = 1;
    h.level = L - 1;
    HIP_TRY(hipMemcpyAsync(din, aR, inw * 8, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dd, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_kernel(KDECOMP, c->T, dd, 1, 2 * L, rows, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout, outw * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MKHIP_OK;
}
