// Plans: a list of kernel launches over descriptor tables in device memory,
// replayed eagerly or as one hipGraph.  Shared by the KKLSS, MKF and BFV paths.
#include "mkhip_internal.hpp"

#include <algorithm>

using namespace mkhip;

namespace mkhip {

hipError_t plan_stage(mkhip_plan *pl, int s, hipStream_t st) {
    const PlanStage &S = pl->stages[s];
    if (!st) st = pl->ctx->stream;
    const DevTables &T = pl->T;  // knobs as built
    if (S.what >= KMAC2 && S.what <= KRESCALE_NTT) return launch_mkf(S.what, T, S.descs, S.n, S.a, S.b, st);
    if (S.what >= KBFV_LIFT && S.what <= KBFV_QUANTIZE) return launch_bfv(S.what, T, S.descs, S.n, S.a, S.b, st);
    if (S.what == KDECOMP_MAC) return launch_fused(S.what, T, S.descs, S.n, S.a, S.b, st);
    if (S.what == KAUTO) return launch_galois(S.what, T, S.descs, S.n, st);
    if (S.what == KLINEAR) return launch_linear(T, S.descs, S.n, st);
    return launch_kernel(S.what, T, S.descs, S.n, S.a, S.b, st, S.c);
}

int plan_launch_raw(mkhip_plan *pl, hipStream_t st) {
    for (int s = 0; s < (int)pl->stages.size(); ++s) HIP_TRY(plan_stage(pl, s, st));
    return MKHIP_OK;
}

int plan_create(mkhip_ctx *c, size_t ws_bytes, mkhip_plan **out, const PlanMem &pm) {
    *out = nullptr;
    mkhip_plan *pl = new mkhip_plan();
    pl->ctx = c;
    pl->T = c->T;
    pl->mem = pm;
    pl->bytes = ws_bytes;
    if (pm.query) {
        if (pm.ws_need) *pm.ws_need = ws_bytes;
    } else if (pm.ws) {
        if (pm.ws_bytes < ws_bytes) {
            delete pl;
            return MKHIP_ERR_OOM;
        }
        pl->d_mem = pm.ws;
    } else {
        hipError_t e = hipMalloc(&pl->d_mem, std::max<size_t>(ws_bytes, 256));
        if (e != hipSuccess) {
            delete pl;
            return hip_status(e);
        }
    }
    *out = pl;
    return MKHIP_OK;
}

int plan_finish(mkhip_plan *&pl, const DescBlob &blob, int use_graph) {
    mkhip_ctx *c = pl->ctx;
    const PlanMem &pm = pl->mem;
    const size_t db = std::max<size_t>(blob.host.size(), 256);
    if (pm.query) {
        if (pm.desc_need) *pm.desc_need = db;
        delete pl;
        pl = nullptr;
        return MKHIP_OK;
    }
    hipError_t e = hipSuccess;
    if (pm.desc) {
        if (pm.desc_bytes < db) {
            mkhip_plan_destroy(pl);
            pl = nullptr;
            return MKHIP_ERR_OOM;
        }
        pl->d_desc = pm.desc;
        const void *src = nullptr;
        if (pm.desc_host && pm.desc_host_bytes >= blob.host.size()) {
            std::memcpy(pm.desc_host, blob.host.data(), blob.host.size());
            src = pm.desc_host;
        } else {
            pl->host_desc = blob.host;
            src = pl->host_desc.data();
        }
        e = hipMemcpyAsync(pl->d_desc, src, blob.host.size(), hipMemcpyHostToDevice, pm.stream ? pm.stream : c->stream);
    } else {
        e = hipMalloc(&pl->d_desc, db);
        if (e == hipSuccess) e = hipMemcpy(pl->d_desc, blob.host.data(), blob.host.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        mkhip_plan_destroy(pl);
        pl = nullptr;
        return hip_status(e);
    }
    pl->bytes += blob.host.size();
    pl->desc_bytes = blob.host.size();
    for (PlanStage &s : pl->stages) s.descs = (const char *)pl->d_desc + reinterpret_cast<size_t>(s.descs);
    if (use_graph) {
        int st = plan_capture(pl);
        if (st != MKHIP_OK) {
            mkhip_plan_destroy(pl);
            pl = nullptr;
            return st;
        }
    }
    return MKHIP_OK;
}

int plan_capture(mkhip_plan *pl) {
    mkhip_ctx *c = pl->ctx;
    // first launch eagerly (sets kernel attributes), then capture the sequence
    int st = plan_launch_raw(pl);
    if (st == MKHIP_OK) st = hip_status(hipStreamSynchronize(c->stream));
    if (st == MKHIP_OK) st = hip_status(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    if (st == MKHIP_OK) {
        int s2 = plan_launch_raw(pl);
        hipError_t e2 = hipStreamEndCapture(c->stream, &pl->g);
        st = s2 != MKHIP_OK ? s2 : hip_status(e2);
    }
    if (st == MKHIP_OK) st = hip_status(hipGraphInstantiate(&pl->ge, pl->g, nullptr, nullptr, 0));
    if (st == MKHIP_OK) pl->graph = true;
    return st;
}

int plan_build(mkhip_ctx *c, int scheme, int nops, const void *ops, int use_graph, mkhip_plan **out,
               const PlanMem &pm) {
    switch (scheme) {
    case MKHIP_SCHEME_KKLSS_CKKS:
        return plan_kklss_ckks(c, nops, (const mkhip_kklss_op *)ops, use_graph, out, pm);
    case MKHIP_SCHEME_KKLSS_BFV:
        return plan_kklss_bfv(c, nops, (const mkhip_kklss_op *)ops, use_graph, out, pm);
    case MKHIP_SCHEME_MKF_CKKS:
        return plan_mkf_ckks(c, nops, (const mkhip_mkf_op *)ops, use_graph, out, pm);
    case MKHIP_SCHEME_MKF_BFV:
        return plan_mkf_bfv(c, nops, (const mkhip_mkf_op *)ops, use_graph, out, pm);
    case MKHIP_SCHEME_KKLSS_ROTATE:
        return plan_kklss_rotate(c, nops, (const mkhip_rotate_op *)ops, use_graph, out, pm);
    case MKHIP_SCHEME_MKF_ROTATE:
        return plan_mkf_rotate(c, nops, (const mkhip_mkf_rotate_op *)ops, use_graph, out, pm);
    default:
        return MKHIP_ERR_ARG;
    }
}

}  // namespace mkhip

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
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = f / iters;
    return MKHIP_OK;
}

extern "C" size_t mkhip_plan_workspace_bytes(const mkhip_plan *pl) { return pl ? pl->bytes : 0; }

extern "C" int mkhip_plan_stage_count(const mkhip_plan *pl) { return pl ? (int)pl->stages.size() : 0; }

extern "C" const char *mkhip_plan_stage_name(const mkhip_plan *pl, int s) {
    if (!pl || s < 0 || s >= (int)pl->stages.size()) return "";
    return pl->stages[s].name.c_str();
}

// Average device time of each launch of the plan (HIP events on the context
// stream around each kernel, eager launches), in ms; ms holds stage_count doubles.
extern "C" int mkhip_plan_stage_times(mkhip_plan *pl, int iters, double *ms) {
    if (!pl || iters <= 0 || !ms) return MKHIP_ERR_ARG;
    const int S = (int)pl->stages.size();
    hipStream_t st = pl->ctx->stream;
    std::vector<hipEvent_t> ev(S + 1);
    for (int i = 0; i <= S; ++i) HIP_TRY(hipEventCreate(&ev[i]));
    std::vector<double> acc(S, 0.0);
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

extern "C" void mkhip_plan_destroy(mkhip_plan *plan) {
    if (!plan) return;
    if (plan->ge) (void)hipGraphExecDestroy(plan->ge);
    if (plan->g) (void)hipGraphDestroy(plan->g);
    if (plan->d_mem && !plan->mem.ws) (void)hipFree(plan->d_mem);
    if (plan->d_desc && !plan->mem.desc) (void)hipFree(plan->d_desc);
    delete plan;
}

// ---------------------------------------------------------------------------
// tuning knobs and the autotuner

extern "C" int mkhip_set_tuning(mkhip_ctx *c, int knob, int value) {
    if (!c) return MKHIP_ERR_ARG;
    switch (knob) {
    case MKHIP_TUNE_FUSED_DECOMP:
        c->tune_fused_decomp = value ? 1 : 0;
        break;
    case MKHIP_TUNE_XCD_REMAP:
        c->T.xcd = value ? 1 : 0;
        break;
    case MKHIP_TUNE_DECOMP_ORDER:
        if (value < 0 || value > 2) return MKHIP_ERR_ARG;
        c->T.decomp_order = value;
        break;
    case MKHIP_TUNE_FP64:
        c->T.fp64 = value ? 1 : 0;
        break;
    case MKHIP_TUNE_SPLIT_LIMB:
        c->T.split14 = value ? 1 : 0;
        break;
    case MKHIP_TUNE_SPLIT_EXT:
        c->T.split14_ext = value ? 1 : 0;
        break;
    case MKHIP_TUNE_DIGIT_PRE:
        c->T.digit_pre = value ? 1 : 0;
        break;
    case MKHIP_TUNE_XCD_EXT:
        c->T.xcd_ext = value ? 1 : 0;
        break;
    case MKHIP_TUNE_EXT_CS:
        c->T.ext_cs = value ? 1 : 0;
        break;
    case MKHIP_TUNE_INV_LAZY:
        c->T.inv_lazy = value ? 1 : 0;
        break;
    default:
        return MKHIP_ERR_ARG;
    }
    // the cached single-op plan was built with the old setting
    mkhip_plan_destroy(c->single);
    c->single = nullptr;
    return MKHIP_OK;
}

extern "C" int mkhip_get_tuning(const mkhip_ctx *c, int knob, int *value) {
    if (!c || !value) return MKHIP_ERR_ARG;
    switch (knob) {
    case MKHIP_TUNE_FUSED_DECOMP:
        *value = c->tune_fused_decomp;
        return MKHIP_OK;
    case MKHIP_TUNE_XCD_REMAP:
        *value = c->T.xcd;
        return MKHIP_OK;
    case MKHIP_TUNE_DECOMP_ORDER:
        *value = c->T.decomp_order;
        return MKHIP_OK;
    case MKHIP_TUNE_FP64:
        *value = c->T.fp64;
        return MKHIP_OK;
    case MKHIP_TUNE_SPLIT_LIMB:
        *value = c->T.split14;
        return MKHIP_OK;
    case MKHIP_TUNE_SPLIT_EXT:
        *value = c->T.split14_ext;
        return MKHIP_OK;
    case MKHIP_TUNE_DIGIT_PRE:
        *value = c->T.digit_pre;
        return MKHIP_OK;
    case MKHIP_TUNE_XCD_EXT:
        *value = c->T.xcd_ext;
        return MKHIP_OK;
    case MKHIP_TUNE_EXT_CS:
        *value = c->T.ext_cs;
        return MKHIP_OK;
    case MKHIP_TUNE_INV_LAZY:
        *value = c->T.inv_lazy;
        return MKHIP_OK;
    default:
        return MKHIP_ERR_ARG;
    }
}

// Every combination of the knobs that apply to the scheme is planned and timed
// (graph launches, HIP events) on the caller's ops; the fastest is kept.
extern "C" int mkhip_autotune(mkhip_ctx *c, int scheme, int nops, const void *ops, int iters, int *ncand,
                              double *ms) {
    if (!c || !ops || nops <= 0 || iters <= 0) return MKHIP_ERR_ARG;
    const bool kklss = scheme == MKHIP_SCHEME_KKLSS_CKKS || scheme == MKHIP_SCHEME_KKLSS_BFV;
    // candidates: (fused, order) for alpha = 1 KKLSS plans -- unfused, fused order 0,
    // fused order 1, and at N = 2^14 the fused orders over half-limb workgroups --
    // each with and without the XCD remap
    struct Cand {
        int fused, order, xcd, split, split_ext, pre, xe;
    };
    std::vector<Cand> cands;
    const int old_fused = c->tune_fused_decomp, old_xcd = c->T.xcd, old_order = c->T.decomp_order;
    const int old_split = c->T.split14, old_split_ext = c->T.split14_ext, old_pre = c->T.digit_pre;
    const int old_xe = c->T.xcd_ext;
    const bool mkf_crt = scheme == MKHIP_SCHEME_MKF_CKKS && c->alpha > 1;
    for (int xcd = 1; xcd >= 0; --xcd) {
        if (kklss && c->alpha == 1) {
            cands.push_back({0, old_order, xcd, old_split, old_split_ext, old_pre, old_xe});
            cands.push_back({1, 0, xcd, 0, old_split_ext, old_pre, old_xe});
            cands.push_back({1, 1, xcd, 0, old_split_ext, old_pre, old_xe});
            cands.push_back({1, 2, xcd, 0, old_split_ext, old_pre, old_xe});
            if (c->logN == 14) {
                cands.push_back({1, 0, xcd, 1, old_split_ext, old_pre, old_xe});
                cands.push_back({1, 1, xcd, 1, old_split_ext, old_pre, old_xe});
                cands.push_back({1, 2, xcd, 1, old_split_ext, old_pre, old_xe});
            }
        } else if (mkf_crt) {  // MKF CKKS with CRT digits: with and without the digit pass
            cands.push_back({old_fused, old_order, xcd, old_split, old_split_ext, 0, old_xe});
            cands.push_back({old_fused, old_order, xcd, old_split, old_split_ext, 1, old_xe});
        } else {
            cands.push_back({old_fused, old_order, xcd, old_split, old_split_ext, old_pre, old_xe});
        }
    }
    auto apply = [&](const Cand &cd) {
        mkhip_set_tuning(c, MKHIP_TUNE_FUSED_DECOMP, cd.fused);
        mkhip_set_tuning(c, MKHIP_TUNE_DECOMP_ORDER, cd.order);
        mkhip_set_tuning(c, MKHIP_TUNE_XCD_REMAP, cd.xcd);
        mkhip_set_tuning(c, MKHIP_TUNE_SPLIT_LIMB, cd.split);
        mkhip_set_tuning(c, MKHIP_TUNE_SPLIT_EXT, cd.split_ext);
        mkhip_set_tuning(c, MKHIP_TUNE_DIGIT_PRE, cd.pre);
        mkhip_set_tuning(c, MKHIP_TUNE_XCD_EXT, cd.xe);
    };
    auto time_cand = [&](const Cand &cd, double *t) {
        apply(cd);
        mkhip_plan *pl = nullptr;
        int s = plan_build(c, scheme, nops, ops, 1, &pl, PlanMem());
        if (s == MKHIP_OK) s = mkhip_plan_launch(pl);  // warm-up
        if (s == MKHIP_OK) s = mkhip_plan_time(pl, iters, t);
        mkhip_plan_destroy(pl);
        return s;
    };
    // the decomposition candidates, then the best of them with each knob of the
    // external products flipped in turn (XCD order; at N = 2^14 half-limb
    // workgroups, whose last inverse stage the ModDowns then do), kept when faster
    int best = -1, n = 0;
    double best_ms = 0;
    const int nfirst = (int)cands.size();
    const int nflip = (kklss && c->alpha == 1 && c->logN == 14) ? 2 : 1;
    for (int ci = 0; ci < nfirst + nflip; ++ci) {
        if (ci >= nfirst) {
            Cand flip = cands[best];
            if (ci == nfirst)
                flip.xe = 1 - flip.xe;
            else
                flip.split_ext = 1 - flip.split_ext;
            cands.push_back(flip);
        }
        double t = 0;
        if (ci >= nfirst) {
            // a flip is compared with the best re-timed just before it, so that a
            // lucky first timing of the best (or a drift of the clocks since)
            // does not decide a choice between forms a few % apart
            const int s = time_cand(cands[best], &best_ms);
            if (s != MKHIP_OK) {
                apply({old_fused, old_order, old_xcd, old_split, old_split_ext, old_pre, old_xe});
                return s;
            }
        }
        const int s = time_cand(cands[ci], &t);
        if (s != MKHIP_OK) {
            apply({old_fused, old_order, old_xcd, old_split, old_split_ext, old_pre, old_xe});
            return s;
        }
        if (ms && n < 16) ms[n] = t;
        if (best < 0 || t < best_ms) {
            best = ci;
            best_ms = t;
        }
        ++n;
    }
    apply(cands[best]);
    if (ncand) *ncand = n;
    return MKHIP_OK;
}
