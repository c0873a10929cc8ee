// KKLSS MK-CKKS mul+relin+rescale that allocates nothing: the plan's workspace and
// descriptor tables live in a fixed device array instead of memory the library
// allocates per plan.  For callers that cannot call hipMalloc (tight serving
// loops, harnesses that forbid device allocation).
//
// The array `scratch` below is a file-scope device array: each GPU that loads the
// library holds its own copy until the process ends.  Its name in host code is
// only a handle for hipGetSymbolAddress; the device address of this GPU's copy
// is looked up once per device (at context creation, with that device current)
// and used for every kernel argument and copy.  Every call on one GPU shares that
// GPU's copy and leaves its contents behind; the kernels never read workspace
// they have not written in the same call, so nothing needs clearing, and calls
// that overlap on one GPU (two threads, or two contexts' streams) race.
#include "mkhip_internal.hpp"

#include <mutex>

using namespace mkhip;

namespace {

constexpr size_t kScratchBytes = MKHIP_NOALLOC_SCRATCH_BYTES;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];
int g_state[kMaxDevices];  // 0: not looked up yet, 1: g_ptr valid, < 0: status of the failed lookup

}  // namespace

namespace mkhip {

// Device address of the current device's scratch (looked up on the first call on
// that device).
int noalloc_scratch(void **ptr) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return MKHIP_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_state[dev] == 0) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_state[dev] = e != hipSuccess ? hip_status(e) : MKHIP_ERR_DEVICE;
        } else {
            g_ptr[dev] = p;
            g_state[dev] = 1;
        }
    }
    if (g_state[dev] != 1) return g_state[dev];
    *ptr = g_ptr[dev];
    return MKHIP_OK;
}

}  // namespace mkhip

extern "C" int mkhip_kklss_ckks_mulrelin_noalloc(mkhip_ctx *c, int nops, const mkhip_kklss_op *ops) {
    if (!c || nops <= 0 || !ops) return MKHIP_ERR_ARG;
    void *base = nullptr;
    int st = noalloc_scratch(&base);
    if (st != MKHIP_OK) return st;
    // sizes first (nothing is built or allocated by a query)
    size_t ws = 0, desc = 0;
    PlanMem q;
    q.query = true;
    q.ws_need = &ws;
    q.desc_need = &desc;
    mkhip_plan *pl = nullptr;
    st = plan_kklss_ckks(c, nops, ops, 0, &pl, q);
    if (st != MKHIP_OK) return st;
    const size_t ws_al = (ws + 255) & ~(size_t)255;
    if (ws_al > kScratchBytes || desc > kScratchBytes - ws_al) return MKHIP_ERR_SCRATCH;
    PlanMem pm;
    pm.ws = base;
    pm.ws_bytes = ws_al;
    pm.desc = (char *)base + ws_al;
    pm.desc_bytes = kScratchBytes - ws_al;
    pm.stream = c->stream;
    st = plan_kklss_ckks(c, nops, ops, 0, &pl, pm);
    if (st != MKHIP_OK) return st;
    st = plan_launch_raw(pl);
    // the descriptor upload reads the plan's host copy: wait before the plan goes
    const hipError_t e = hipStreamSynchronize(c->stream);
    mkhip_plan_destroy(pl);  // the workspace and descriptors are the scratch: nothing is freed
    if (st != MKHIP_OK) return st;
    return e == hipSuccess ? MKHIP_OK : hip_status(e);
}
