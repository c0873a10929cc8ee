// Uneven batches: many mul+relin ops of differing party count k, ring degree N
// and scheme, run by ONE hipGraph launch.
//
// Every context (one ring degree / parameter set) contributes a branch.  A
// branch's ops are cut, in order, into chunks whose plan workspace fits the
// branch's budget; the chunks run one after another in one shared workspace
// arena, so device memory is bounded by the budget and not by the batch size.
// Branches are independent and are captured as parallel paths of the graph
// (fork / join through events), so small-ring work fills the CUs that a
// large-ring stage leaves idle.  Within a chunk every stage walks flattened
// descriptor tables that already hold each op's own k, so ops of different k
// share every launch.
#include "mkhip_internal.hpp"

#include <algorithm>

using namespace mkhip;

struct mkhip_batch {
    struct Branch {
        mkhip_ctx *c = nullptr;
        int scheme = 0;
        hipStream_t st = nullptr;
        hipEvent_t join = nullptr;
        void *arena = nullptr;
        size_t arena_bytes = 0;
        std::vector<mkhip_plan *> plans;
        std::vector<int> chunk_ops;
    };
    std::vector<Branch> br;
    hipStream_t st0 = nullptr;
    hipEvent_t fork = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
};

namespace {

size_t op_size(int scheme) {
    switch (scheme) {
    case MKHIP_SCHEME_KKLSS_CKKS:
    case MKHIP_SCHEME_KKLSS_BFV:
        return sizeof(mkhip_kklss_op);
    case MKHIP_SCHEME_KKLSS_ROTATE:
        return sizeof(mkhip_rotate_op);
    case MKHIP_SCHEME_MKF_ROTATE:
        return sizeof(mkhip_mkf_rotate_op);
    default:
        return sizeof(mkhip_mkf_op);
    }
}

// workspace bytes of a plan over ops [0, n)
int query_ws(mkhip_ctx *c, int scheme, const char *ops, int n, size_t *ws) {
    PlanMem q;
    q.query = true;
    size_t d = 0;
    q.ws_need = ws;
    q.desc_need = &d;
    mkhip_plan *pl = nullptr;
    return plan_build(c, scheme, n, ops, 0, &pl, q);
}

void destroy_branch(mkhip_batch::Branch &b) {
    for (mkhip_plan *p : b.plans) mkhip_plan_destroy(p);
    b.plans.clear();
    if (b.arena) (void)hipFree(b.arena);
    if (b.join) (void)hipEventDestroy(b.join);
    if (b.st) (void)hipStreamDestroy(b.st);
    b = mkhip_batch::Branch();
}

}  // namespace

extern "C" int mkhip_batch_create(mkhip_batch **out) {
    if (!out) return MKHIP_ERR_ARG;
    *out = nullptr;
    mkhip_batch *b = new mkhip_batch();
    hipError_t e = hipStreamCreateWithFlags(&b->st0, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e != hipSuccess) {
        mkhip_batch_destroy(b);
        return hip_status(e);
    }
    *out = b;
    return MKHIP_OK;
}

extern "C" int mkhip_batch_add(mkhip_batch *b, mkhip_ctx *c, int scheme, int nops, const void *ops,
                               size_t ws_budget, int *nchunks) {
    if (!b || !c || !ops || nops <= 0) return MKHIP_ERR_ARG;
    if (b->ge) return MKHIP_ERR_ARG;  // already finalized
    if (scheme < MKHIP_SCHEME_KKLSS_CKKS || scheme > MKHIP_SCHEME_MKF_ROTATE) return MKHIP_ERR_ARG;
    const char *o = (const char *)ops;
    const size_t osz = op_size(scheme);
    // chunks: the longest prefix of the remaining ops whose workspace fits the
    // budget (workspace grows with the op count: binary search), at least one op
    std::vector<std::pair<int, int> > chunks;
    size_t arena = 0;
    for (int start = 0; start < nops;) {
        const char *base = o + start * osz;
        int lo = 1, hi = nops - start;
        size_t ws_lo = 0;
        int s = query_ws(c, scheme, base, lo, &ws_lo);
        if (s != MKHIP_OK) return s;
        if (ws_budget) {
            size_t ws_hi = 0;
            s = query_ws(c, scheme, base, hi, &ws_hi);
            if (s != MKHIP_OK) return s;
            if (ws_hi <= ws_budget) {
                lo = hi;
                ws_lo = ws_hi;
            } else {
                while (hi - lo > 1) {  // ws(lo) fits (or lo = 1), ws(hi) does not
                    const int mid = (lo + hi) / 2;
                    size_t w = 0;
                    s = query_ws(c, scheme, base, mid, &w);
                    if (s != MKHIP_OK) return s;
                    if (w <= ws_budget) {
                        lo = mid;
                        ws_lo = w;
                    } else {
                        hi = mid;
                    }
                }
            }
        } else {
            lo = nops - start;
            s = query_ws(c, scheme, base, lo, &ws_lo);
            if (s != MKHIP_OK) return s;
        }
        chunks.push_back({start, lo});
        arena = std::max(arena, ws_lo);
        start += lo;
    }
    mkhip_batch::Branch br;
    br.c = c;
    br.scheme = scheme;
    hipError_t e = hipStreamCreateWithFlags(&br.st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&br.join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&br.arena, std::max<size_t>(arena, 256));
    if (e != hipSuccess) {
        destroy_branch(br);
        return hip_status(e);
    }
    br.arena_bytes = arena;
    for (const auto &ch : chunks) {
        PlanMem pm;
        pm.ws = br.arena;
        pm.ws_bytes = arena;
        mkhip_plan *pl = nullptr;
        int s = plan_build(c, scheme, ch.second, o + ch.first * osz, 0, &pl, pm);
        if (s != MKHIP_OK) {
            destroy_branch(br);
            return s;
        }
        br.plans.push_back(pl);
        br.chunk_ops.push_back(ch.second);
    }
    if (nchunks) *nchunks = (int)chunks.size();
    b->br.push_back(br);
    return MKHIP_OK;
}

extern "C" int mkhip_batch_finalize(mkhip_batch *b) {
    if (!b || b->br.empty()) return MKHIP_ERR_ARG;
    if (b->ge) return MKHIP_OK;
    // one eager pass (kernel attributes are set on first use, outside capture)
    for (auto &br : b->br) {
        for (mkhip_plan *p : br.plans) {
            int s = plan_launch_raw(p, br.st);
            if (s != MKHIP_OK) return s;
        }
        HIP_TRY(hipStreamSynchronize(br.st));
    }
    HIP_TRY(hipStreamBeginCapture(b->st0, hipStreamCaptureModeThreadLocal));
    int st = MKHIP_OK;
    hipError_t e = hipEventRecord(b->fork, b->st0);
    for (auto &br : b->br) {
        if (e != hipSuccess || st != MKHIP_OK) break;
        e = hipStreamWaitEvent(br.st, b->fork, 0);
        for (mkhip_plan *p : br.plans)
            if (e == hipSuccess && st == MKHIP_OK) st = plan_launch_raw(p, br.st);
        if (e == hipSuccess) e = hipEventRecord(br.join, br.st);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->st0, br.join, 0);
    }
    hipError_t e2 = hipStreamEndCapture(b->st0, &b->g);
    if (st != MKHIP_OK) return st;
    HIP_TRY(e);
    HIP_TRY(e2);
    HIP_TRY(hipGraphInstantiate(&b->ge, b->g, nullptr, nullptr, 0));
    return MKHIP_OK;
}

extern "C" int mkhip_batch_submit(mkhip_batch *b) {
    if (!b || !b->ge) return MKHIP_ERR_ARG;
    HIP_TRY(hipGraphLaunch(b->ge, b->st0));
    return MKHIP_OK;
}

extern "C" int mkhip_batch_wait(mkhip_batch *b) {
    if (!b) return MKHIP_ERR_ARG;
    HIP_TRY(hipStreamSynchronize(b->st0));
    return MKHIP_OK;
}

extern "C" int mkhip_batch_time(mkhip_batch *b, int iters, double *ms) {
    if (!b || !b->ge || iters <= 0 || !ms) return MKHIP_ERR_ARG;
    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    HIP_TRY(hipEventRecord(t0, b->st0));
    for (int i = 0; i < iters; ++i) HIP_TRY(hipGraphLaunch(b->ge, b->st0));
    HIP_TRY(hipEventRecord(t1, b->st0));
    HIP_TRY(hipEventSynchronize(t1));
    float f = 0;
    HIP_TRY(hipEventElapsedTime(&f, t0, t1));
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    *ms = f / iters;
    return MKHIP_OK;
}

extern "C" size_t mkhip_batch_workspace_bytes(const mkhip_batch *b) {
    if (!b) return 0;
    size_t s = 0;
    for (const auto &br : b->br) {
        s += br.arena_bytes;
        for (const mkhip_plan *p : br.plans) s += p->desc_bytes;
    }
    return s;
}

extern "C" int mkhip_batch_branch_info(const mkhip_batch *b, int branch, int *nchunks, size_t *arena_bytes) {
    if (!b || branch < 0 || branch >= (int)b->br.size()) return MKHIP_ERR_ARG;
    if (nchunks) *nchunks = (int)b->br[branch].plans.size();
    if (arena_bytes) *arena_bytes = b->br[branch].arena_bytes;
    return MKHIP_OK;
}

extern "C" void mkhip_batch_destroy(mkhip_batch *b) {
    if (!b) return;
    if (b->st0) (void)hipStreamSynchronize(b->st0);
    if (b->ge) (void)hipGraphExecDestroy(b->ge);
    if (b->g) (void)hipGraphDestroy(b->g);
    for (auto &br : b->br) destroy_branch(br);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->st0) (void)hipStreamDestroy(b->st0);
    delete b;
}
