// Serving: a request loop with continuous batching over contexts whose parties'
// keys stay resident in HBM.
//
// Clients enqueue single mul+relin requests (host operands) and poll or wait on a
// ticket.  A scheduler thread owns the device side:
//   * admission by device memory -- every batch takes one block (operands,
//     outputs, plan workspace, descriptors) from a fixed device pool; a batch is
//     formed only from requests whose block fits the largest free extent, so
//     requests wait while memory is held by batches in flight;
//   * continuous batching -- whenever one of the server's streams is free, the
//     class (context) with the oldest waiting request is batched: its waiting
//     requests, oldest first, up to max_batch ops and the free memory;
//   * overlap -- batches on different streams run concurrently on the device
//     (their operand copies, kernels and result copies each on the batch's own
//     stream, so one batch's copies run beside another's kernels), and requests
//     that arrive while others are on the GPU start as soon as a stream and
//     memory free up, not when the GPU drains.  Descriptor tables go up from
//     pinned staging, so the scheduler thread never waits on a copy.  (One
//     upload and one download stream shared in admission order was measured
//     slower on MI355X: the host's enqueue latency between copies then leaves
//     the link idle, where copies on several streams overlap it.)
// Results do not depend on the mix: every op of a batch reads its own operands
// and keys and writes its own output (the plans share no accumulator across ops).
#include "mkhip_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

using namespace mkhip;

namespace {

typedef std::chrono::steady_clock Clock;

// First-fit allocator over [0, size) with coalescing of freed extents.
struct Pool {
    std::map<size_t, size_t> free_;  // offset -> bytes
    size_t size = 0;
    void init(size_t n) {
        size = n;
        free_.clear();
        free_[0] = n;
    }
    static size_t align(size_t n) { return (n + 255) & ~(size_t)255; }
    size_t largest() const {
        size_t m = 0;
        for (const auto &f : free_) m = std::max(m, f.second);
        return m;
    }
    bool take(size_t n, size_t *off) {
        n = align(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, o = it->first + n;
                free_.erase(it);
                if (rest) free_[o] = rest;
                return true;
            }
        }
        return false;
    }
    void give(size_t off, size_t n) {
        n = align(n);
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Req {
    uint64_t ticket = 0;
    mkhip_request r;
    Clock::time_point t_in, t_start, t_done;
    int status = 1;  // 1 pending, 0 ok, < 0 error
    size_t in0_w = 0, in1_w = 0, out_w = 0;
    // words of the request's operands and result the batch block holds (0 for
    // device-resident requests, served in place)
    size_t io_w() const { return r.on_device ? 0 : in0_w + in1_w + out_w; }
};

struct Class {
    mkhip_ctx *c = nullptr;
    int scheme = 0;
    std::deque<Req *> wait;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    void *desc_host = nullptr;          // pinned staging of the descriptor tables
    size_t desc_host_bytes = 0;
    bool busy = false;
    mkhip_plan *plan = nullptr;
    size_t off = 0, bytes = 0;
    std::vector<Req *> reqs;
};

bool is_kklss(int scheme) { return scheme == MKHIP_SCHEME_KKLSS_CKKS || scheme == MKHIP_SCHEME_KKLSS_BFV; }

int union_size(const mkhip_request &r) {
    int ids[2 * MKHIP_MAX_PARTIES], n = 0;
    for (int i = 0; i < r.k0; ++i) ids[n++] = r.ids0[i];
    for (int i = 0; i < r.k1; ++i) ids[n++] = r.ids1[i];
    std::sort(ids, ids + n);
    return (int)(std::unique(ids, ids + n) - ids);
}

}  // namespace

struct mkhip_server {
    mkhip_server_config cfg;
    std::vector<Class> cls;
    std::vector<Slot> slots;
    void *d_pool = nullptr;
    Pool pool;
    std::mutex mu;
    std::condition_variable cv_in, cv_done;
    std::deque<Req *> arrivals;
    std::unordered_map<uint64_t, Req *> reqs;
    uint64_t next_ticket = 1;
    std::thread th;
    std::atomic<bool> stop{false};
    bool started = false;
    // statistics
    uint64_t n_done = 0, n_batches = 0, max_batch_seen = 0, ops_in_batches = 0;
    int max_inflight = 0;
};

namespace {

// Builds the ops of reqs[0..n) with operands placed from io_base; returns the
// scheme's op array (as bytes).
std::vector<char> make_ops(const Class &C, const std::vector<Req *> &reqs, size_t n, u64 *io_base) {
    const mkhip_ctx *c = C.c;
    std::vector<char> v;
    u64 *p = io_base;
    auto adv = [&](size_t w) {
        u64 *r = p;
        p = p ? p + w : nullptr;
        return r ? r : (u64 *)(uintptr_t)256;  // placeholder for size queries
    };
    // a host request's operands and result in the batch block, a device request's
    // where they are
    auto place = [&](const Req *q, u64 **op0, u64 **op1, u64 **out) {
        if (q->r.on_device) {
            *op0 = (u64 *)q->r.in0;
            *op1 = (u64 *)q->r.in1;
            *out = q->r.out;
        } else {
            *op0 = adv(q->in0_w);
            *op1 = adv(q->in1_w);
            *out = adv(q->out_w);
        }
    };
    if (is_kklss(C.scheme)) {
        v.resize(n * sizeof(mkhip_kklss_op));
        mkhip_kklss_op *ops = (mkhip_kklss_op *)v.data();
        for (size_t i = 0; i < n; ++i) {
            const Req *q = reqs[i];
            mkhip_kklss_op o = {};
            u64 *o0, *o1, *oo;
            place(q, &o0, &o1, &oo);
            o.op0 = o0;
            o.op1 = o1;
            o.out = oo;
            for (int j = 0; j < q->r.k0; ++j) o.ids0[j] = q->r.ids0[j];
            for (int j = 0; j < q->r.k1; ++j) o.ids1[j] = q->r.ids1[j];
            o.k0 = q->r.k0;
            o.k1 = q->r.k1;
            o.level = c->L - 1;
            o.nres = q->r.nres;
            ops[i] = o;
        }
    } else {
        v.resize(n * sizeof(mkhip_mkf_op));
        mkhip_mkf_op *ops = (mkhip_mkf_op *)v.data();
        for (size_t i = 0; i < n; ++i) {
            const Req *q = reqs[i];
            mkhip_mkf_op o = {};
            u64 *o0, *o1, *oo;
            place(q, &o0, &o1, &oo);
            o.left = o0;
            o.right = o1;
            o.out = oo;
            o.group = q->r.group;
            o.level = c->L - 1;
            o.nres = q->r.nres;
            ops[i] = o;
        }
    }
    return v;
}

// Device bytes of a batch over reqs[0..n): operands + outputs + workspace + descriptors.
int batch_bytes(const Class &C, const std::vector<Req *> &reqs, size_t n, size_t *io, size_t *ws, size_t *desc) {
    size_t w = 0;
    for (size_t i = 0; i < n; ++i) w += reqs[i]->io_w();
    *io = w * 8;
    std::vector<char> ops = make_ops(C, reqs, n, nullptr);
    PlanMem q;
    q.query = true;
    q.ws_need = ws;
    q.desc_need = desc;
    mkhip_plan *pl = nullptr;
    return plan_build(C.c, C.scheme, (int)n, ops.data(), 0, &pl, q);
}

void finish_slot(mkhip_server *S, Slot &sl, int status) {
    const Clock::time_point now = Clock::now();
    mkhip_plan_destroy(sl.plan);
    sl.plan = nullptr;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->pool.give(sl.off, sl.bytes);
        for (Req *q : sl.reqs) {
            q->t_done = now;
            q->status = status;
            ++S->n_done;
        }
    }
    sl.reqs.clear();
    sl.busy = false;
    S->cv_done.notify_all();
}

// Launches one batch of class C on slot sl; false if nothing was admitted.
bool launch_batch(mkhip_server *S, Class &C, Slot &sl) {
    std::vector<Req *> cand(C.wait.begin(), C.wait.begin() + std::min<size_t>(C.wait.size(), S->cfg.max_batch));
    size_t largest;
    {
        std::lock_guard<std::mutex> g(S->mu);
        largest = S->pool.largest();
    }
    // the longest FIFO prefix whose block fits the largest free extent; a prefix
    // that a request cannot be planned in (say, a party unregistered after the
    // enqueue) is shortened past it, so that only that request fails
    size_t n = cand.size(), io = 0, ws = 0, desc = 0;
    int front_status = MKHIP_OK;
    while (n > 0) {
        const int qs = batch_bytes(C, cand, n, &io, &ws, &desc);
        if (qs != MKHIP_OK) {
            if (n == 1) {
                front_status = qs;
                n = 0;
                break;
            }
            --n;
            continue;
        }
        if (Pool::align(io) + Pool::align(ws) + Pool::align(desc) <= largest) break;
        n = n > 8 ? n * 3 / 4 : n - 1;
    }
    if (n == 0) {
        // the oldest request fails alone when it cannot be planned, or when its
        // block does not fit even the whole pool with nothing in flight
        bool idle = true;
        for (const Slot &o : S->slots) idle = idle && !o.busy;
        if (!C.wait.empty() && (front_status != MKHIP_OK || idle)) {
            Req *q = C.wait.front();
            C.wait.pop_front();
            std::lock_guard<std::mutex> g(S->mu);
            q->t_done = Clock::now();
            q->status = front_status != MKHIP_OK ? front_status : MKHIP_ERR_OOM;
            ++S->n_done;
            S->cv_done.notify_all();
            return true;
        }
        return false;
    }
    const size_t bytes = Pool::align(io) + Pool::align(ws) + Pool::align(desc);
    size_t off = 0;
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (!S->pool.take(bytes, &off)) return false;
    }
    char *base = (char *)S->d_pool + off;
    u64 *d_io = (u64 *)base;
    void *d_ws = base + Pool::align(io);
    void *d_desc = base + Pool::align(io) + Pool::align(ws);
    sl.reqs.assign(cand.begin(), cand.begin() + n);
    for (size_t i = 0; i < n; ++i) C.wait.pop_front();
    sl.off = off;
    sl.bytes = bytes;
    sl.busy = true;
    const Clock::time_point now = Clock::now();
    for (Req *q : sl.reqs) q->t_start = now;
    std::vector<char> ops = make_ops(C, sl.reqs, n, d_io);
    int st = MKHIP_OK;
    // pinned staging for the descriptor tables (grown, never per batch in steady state)
    if (sl.desc_host_bytes < desc) {
        if (sl.desc_host) (void)hipHostFree(sl.desc_host);
        sl.desc_host = nullptr;
        sl.desc_host_bytes = 0;
        const size_t cap = std::max<size_t>(desc, 1 << 20);
        if (hipHostMalloc(&sl.desc_host, cap, hipHostMallocDefault) == hipSuccess) sl.desc_host_bytes = cap;
    }
    // operands in (host requests; device requests are read in place)
    u64 *p = d_io;
    for (Req *q : sl.reqs) {
        if (q->r.on_device) continue;
        if (st == MKHIP_OK)
            st = hip_status(hipMemcpyAsync(p, q->r.in0, q->in0_w * 8, hipMemcpyHostToDevice, sl.st));
        p += q->in0_w;
        if (st == MKHIP_OK)
            st = hip_status(hipMemcpyAsync(p, q->r.in1, q->in1_w * 8, hipMemcpyHostToDevice, sl.st));
        p += q->in1_w + q->out_w;
    }
    PlanMem pm;
    pm.ws = d_ws;
    pm.ws_bytes = ws;
    pm.desc = d_desc;
    pm.desc_bytes = desc;
    pm.stream = sl.st;
    pm.desc_host = sl.desc_host;
    pm.desc_host_bytes = sl.desc_host_bytes;
    if (st == MKHIP_OK) st = plan_build(C.c, C.scheme, (int)n, ops.data(), 0, &sl.plan, pm);
    if (st == MKHIP_OK) st = plan_launch_raw(sl.plan, sl.st);
    // results out
    p = d_io;
    for (Req *q : sl.reqs) {
        if (q->r.on_device) continue;
        p += q->in0_w + q->in1_w;
        if (st == MKHIP_OK) st = hip_status(hipMemcpyAsync(q->r.out, p, q->out_w * 8, hipMemcpyDeviceToHost, sl.st));
        p += q->out_w;
    }
    if (st == MKHIP_OK) st = hip_status(hipEventRecord(sl.done, sl.st));
    if (st != MKHIP_OK) {
        (void)hipStreamSynchronize(sl.st);
        finish_slot(S, sl, st);
        return true;
    }
    int inflight = 0;
    for (const Slot &o : S->slots) inflight += o.busy;
    std::lock_guard<std::mutex> g(S->mu);
    ++S->n_batches;
    S->ops_in_batches += n;
    S->max_batch_seen = std::max<uint64_t>(S->max_batch_seen, n);
    S->max_inflight = std::max(S->max_inflight, inflight);
    return true;
}

void scheduler(mkhip_server *S) {
    while (!S->stop.load()) {
        bool progress = false;
        // arrivals -> class queues
        {
            std::unique_lock<std::mutex> g(S->mu);
            bool any_busy = false, any_wait = false;
            for (const Slot &sl : S->slots) any_busy = any_busy || sl.busy;
            for (const Class &C : S->cls) any_wait = any_wait || !C.wait.empty();
            if (S->arrivals.empty() && !any_busy && !any_wait)
                S->cv_in.wait_for(g, std::chrono::milliseconds(5));
            while (!S->arrivals.empty()) {
                Req *q = S->arrivals.front();
                S->arrivals.pop_front();
                S->cls[q->r.cls].wait.push_back(q);
                progress = true;
            }
        }
        // completions
        for (Slot &sl : S->slots) {
            if (!sl.busy) continue;
            const hipError_t e = hipEventQuery(sl.done);
            if (e == hipSuccess) {
                finish_slot(S, sl, MKHIP_OK);
                progress = true;
            } else if (e != hipErrorNotReady) {
                (void)hipStreamSynchronize(sl.st);
                finish_slot(S, sl, hip_status(e));
                progress = true;
            }
        }
        // admissions: free streams take the class with the oldest waiting request
        for (Slot &sl : S->slots) {
            if (sl.busy) continue;
            Class *best = nullptr;
            for (Class &C : S->cls)
                if (!C.wait.empty() && (!best || C.wait.front()->t_in < best->wait.front()->t_in)) best = &C;
            if (!best) break;
            if (launch_batch(S, *best, sl)) progress = true;
            else break;  // memory is held by batches in flight: wait for a completion
        }
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    // drain
    for (Slot &sl : S->slots)
        if (sl.busy) {
            (void)hipStreamSynchronize(sl.st);
            finish_slot(S, sl, MKHIP_OK);
        }
}

}  // namespace

extern "C" int mkhip_server_create(const mkhip_server_config *cfg, mkhip_server **out) {
    if (!cfg || !out || cfg->streams <= 0 || cfg->max_batch <= 0 || cfg->mem_bytes == 0) return MKHIP_ERR_ARG;
    *out = nullptr;
    mkhip_server *S = new mkhip_server();
    S->cfg = *cfg;
    hipError_t e = hipMalloc(&S->d_pool, cfg->mem_bytes);
    if (e != hipSuccess) {
        delete S;
        return hip_status(e);
    }
    S->pool.init(cfg->mem_bytes);
    S->slots.resize(cfg->streams);
    for (Slot &sl : S->slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        mkhip_server_destroy(S);
        return hip_status(e);
    }
    *out = S;
    return MKHIP_OK;
}

extern "C" int mkhip_server_add_class(mkhip_server *S, mkhip_ctx *c, int scheme, int *cls) {
    if (!S || !c || S->started) return MKHIP_ERR_ARG;
    if (scheme < MKHIP_SCHEME_KKLSS_CKKS || scheme > MKHIP_SCHEME_MKF_BFV) return MKHIP_ERR_ARG;
    if ((scheme == MKHIP_SCHEME_KKLSS_BFV || scheme == MKHIP_SCHEME_MKF_BFV) != c->bfv) return MKHIP_ERR_SCHEME;
    Class C;
    C.c = c;
    C.scheme = scheme;
    S->cls.push_back(C);
    if (cls) *cls = (int)S->cls.size() - 1;
    return MKHIP_OK;
}

extern "C" int mkhip_server_start(mkhip_server *S) {
    if (!S || S->started || S->cls.empty()) return MKHIP_ERR_ARG;
    S->started = true;
    S->th = std::thread(scheduler, S);
    return MKHIP_OK;
}

extern "C" int mkhip_server_enqueue(mkhip_server *S, const mkhip_request *r, uint64_t *ticket) {
    if (!S || !r || !ticket || !S->started) return MKHIP_ERR_ARG;
    if (r->cls < 0 || r->cls >= (int)S->cls.size()) return MKHIP_ERR_ARG;
    if (!r->in0 || !r->in1 || !r->out) return MKHIP_ERR_ARG;
    const Class &C = S->cls[r->cls];
    const mkhip_ctx *c = C.c;
    const bool ckks = C.scheme == MKHIP_SCHEME_KKLSS_CKKS || C.scheme == MKHIP_SCHEME_MKF_CKKS;
    if (r->nres < 0 || (!ckks && r->nres != 0) || (ckks && r->nres > c->L - 1)) return MKHIP_ERR_SCALE;
    Req *q = new Req();
    q->r = *r;
    const size_t poly = (size_t)c->L * c->N, opoly = (size_t)(c->L - r->nres) * c->N;
    if (is_kklss(C.scheme)) {
        if (r->k0 < 0 || r->k1 < 0 || r->k0 > MKHIP_MAX_PARTIES || r->k1 > MKHIP_MAX_PARTIES) {
            delete q;
            return MKHIP_ERR_TOO_MANY;
        }
        for (int i = 0; i < r->k0; ++i)
            if (!c->parties.count(r->ids0[i])) {
                delete q;
                return MKHIP_ERR_UNKNOWN_PARTY;
            }
        for (int i = 0; i < r->k1; ++i)
            if (!c->parties.count(r->ids1[i])) {
                delete q;
                return MKHIP_ERR_UNKNOWN_PARTY;
            }
        const int ku = union_size(*r);
        if (ku > MKHIP_MAX_PARTIES) {
            delete q;
            return MKHIP_ERR_TOO_MANY;
        }
        q->in0_w = (size_t)(1 + r->k0) * poly;
        q->in1_w = (size_t)(1 + r->k1) * poly;
        q->out_w = (size_t)(1 + ku) * opoly;
    } else {
        auto it = c->mkf_groups.find(r->group);
        if (it == c->mkf_groups.end()) {
            delete q;
            return MKHIP_ERR_NO_KEY;
        }
        const int k = it->second.k;
        q->in0_w = q->in1_w = (size_t)(1 + k) * poly;
        q->out_w = (size_t)(1 + k) * opoly;
    }
    q->t_in = Clock::now();
    {
        std::lock_guard<std::mutex> g(S->mu);
        q->ticket = S->next_ticket++;
        S->reqs[q->ticket] = q;
        S->arrivals.push_back(q);
    }
    S->cv_in.notify_one();
    *ticket = q->ticket;
    return MKHIP_OK;
}

namespace {
void report(const Req *q, int *status, double *latency_ms, double *queue_ms) {
    if (status) *status = q->status;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(q->t_done - q->t_in).count();
    if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(q->t_start - q->t_in).count();
}
}  // namespace

extern "C" int mkhip_server_poll(mkhip_server *S, uint64_t ticket, int *done, int *status, double *latency_ms,
                                 double *queue_ms) {
    if (!S || !done) return MKHIP_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return MKHIP_ERR_ARG;
    Req *q = it->second;
    *done = q->status != 1;
    if (*done) {
        report(q, status, latency_ms, queue_ms);
        S->reqs.erase(it);
        delete q;
    }
    return MKHIP_OK;
}

extern "C" int mkhip_server_wait(mkhip_server *S, uint64_t ticket, int *status, double *latency_ms,
                                 double *queue_ms) {
    if (!S) return MKHIP_ERR_ARG;
    std::unique_lock<std::mutex> g(S->mu);
    auto it = S->reqs.find(ticket);
    if (it == S->reqs.end()) return MKHIP_ERR_ARG;
    Req *q = it->second;
    S->cv_done.wait(g, [&] { return q->status != 1; });
    report(q, status, latency_ms, queue_ms);
    S->reqs.erase(it);
    delete q;
    return MKHIP_OK;
}

extern "C" int mkhip_server_stats(mkhip_server *S, mkhip_server_stats_t *st) {
    if (!S || !st) return MKHIP_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    st->completed = S->n_done;
    st->batches = S->n_batches;
    st->max_batch = S->max_batch_seen;
    st->mean_batch = S->n_batches ? (double)S->ops_in_batches / S->n_batches : 0.0;
    st->max_inflight = S->max_inflight;
    st->pool_bytes = S->pool.size;
    st->pool_free_largest = S->pool.largest();
    return MKHIP_OK;
}

extern "C" void mkhip_server_destroy(mkhip_server *S) {
    if (!S) return;
    if (S->started) {
        S->stop.store(true);
        S->cv_in.notify_all();
        if (S->th.joinable()) S->th.join();
    }
    for (auto &kv : S->reqs) delete kv.second;
    for (Slot &sl : S->slots) {
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.st) (void)hipStreamDestroy(sl.st);
        if (sl.desc_host) (void)hipHostFree(sl.desc_host);
    }

    if (S->d_pool) (void)hipFree(S->d_pool);
    delete S;
}

extern "C" int mkhip_host_alloc(size_t bytes, void **p) {
    if (!p || bytes == 0) return MKHIP_ERR_ARG;
    HIP_TRY(hipHostMalloc(p, bytes, hipHostMallocDefault));
    return MKHIP_OK;
}

extern "C" int mkhip_host_free(void *p) {
    if (!p) return MKHIP_ERR_ARG;
    HIP_TRY(hipHostFree(p));
    return MKHIP_OK;
}
