// Host-side internals shared by the C ABI translation units: the context, the
// plan (a list of kernel launches over descriptor tables, optionally captured in
// a hipGraph), status helpers.
#pragma once

#include "../../include/mkhip.h"
#include "tables.hpp"

#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace mkhip {
// kernels.hip
hipError_t launch_kernel(int what, const DevTables &T, const void *descs, int n, int a, int b, hipStream_t st,
                         int c = 1);
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st);
// kernels_mkf.hip
hipError_t launch_mkf(int what, const DevTables &T, const void *descs, int n, int a, int b, hipStream_t st);
// kernels_bfv.hip
hipError_t launch_bfv(int what, const DevTables &T, const void *descs, int n, int a, int b, hipStream_t st);
// kernels_fused.hip
hipError_t launch_fused(int what, const DevTables &T, const void *descs, int n, int a, int b, hipStream_t st);
// kernels_galois.hip
hipError_t launch_galois(int what, const DevTables &T, const void *descs, int n, hipStream_t st);
// kernels_keygen.hip
hipError_t launch_keygen(int what, const DevTables &T, const void *descs, int n, hipStream_t st);
hipError_t launch_linear(const DevTables &T, const void *descs, int n, hipStream_t st);

// mkhip_noalloc.hip: device address of the current device's allocation-free
// scratch, looked up once per device
int noalloc_scratch(void **ptr);

inline int hip_status(hipError_t e) {
    return e == hipSuccess ? MKHIP_OK : (e == hipErrorOutOfMemory ? MKHIP_ERR_OOM : MKHIP_ERR_DEVICE);
}
}  // namespace mkhip

#define HIP_TRY(expr)                                          \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return mkhip::hip_status(_e);    \
    } while (0)

// A party's KKLSS relinearization key (b, d, v), each beta x (L+K) x N; for a
// BFV party b and d hold both gadget branches (2 beta digits: branch 0, then 1).
struct PartyKeys {
    mkhip::u64 *b = nullptr, *d = nullptr, *v = nullptr;
    bool bfv = false;
};

// MKF keys of one party group: fold keys s_i -> s_aggregate and the
// relinearization key of s_aggregate (mkf_cpu/common.go:32-51), each
// [beta][2][L+K][N].
struct MkfGroup {
    int k = 0;
    std::vector<mkhip::u64 *> fold;
    mkhip::u64 *relin = nullptr;
    // party-specific Galois keys sigma(s_i) -> s_aggregate by Galois element
    // (true_mkf_ckks.go:323-333), k keys each
    std::map<mkhip::u64, std::vector<mkhip::u64 *> > rot;
};

// One launch of a plan: kernel id + descriptor table + launch geometry.
struct PlanStage {
    int what = 0;
    const void *descs = nullptr;
    int n = 0, a = 0, b = 0, c = 1;
    std::string name;
};

struct mkhip_ctx {
    int logN = 0, N = 0, L = 0, K = 0, gamma = 1, alpha = 0, beta_max = 0;
    mkhip::u64 t = 0;
    int flags = 0;
    std::vector<mkhip::u64> Q, P, QMul;  // QMul: BFV auxiliary chain (empty for CKKS)
    std::vector<mkhip::u64> psi_half;  // psi^(N/2) per ring modulus (a square root of -1; Lattigo NttPsi[1])
    bool bfv = false;
    void *d_tables = nullptr;
    mkhip::DevTables T;
    hipStream_t stream = nullptr;
    mkhip::u64 *d_crs_u = nullptr;
    std::map<int, mkhip::u64 *> crs;             // CRS of rotations (index = rotidx) and conjugation (-2)
    std::map<std::pair<int, int>, mkhip::u64 *> rot_keys;  // (party, rotidx) -> rotation key
    std::map<int, mkhip::u64 *> conj_keys;       // party -> conjugation key
    std::map<int, PartyKeys> parties;
    std::map<int, MkfGroup> mkf_groups;
    size_t swk_words = 0;       // beta x (L+K) x N
    size_t mkf_swk_words = 0;   // beta x 2 x (L+K) x N
    int tune_fused_decomp = 1;     // MKHIP_TUNE_FUSED_DECOMP (alpha = 1 KKLSS plans)
    mkhip_plan *single = nullptr;  // cached plan of the host-pointer entry point
    mkhip::u64 *single_io = nullptr;
    size_t single_io_words = 0;
};

namespace mkhip {
// Where a plan's device memory comes from.  By default the plan allocates its
// workspace and descriptor tables.  With ws / desc set it uses the caller's
// buffers (which it never frees; desc is filled by an asynchronous copy on
// `stream` from a host copy the plan keeps).  With query set nothing is allocated
// or built: the builder reports the sizes it needs and returns no plan.
struct PlanMem {
    void *ws = nullptr;
    size_t ws_bytes = 0;
    void *desc = nullptr;
    size_t desc_bytes = 0;
    hipStream_t stream = nullptr;
    bool query = false;
    size_t *ws_need = nullptr, *desc_need = nullptr;
    // pinned host staging for the descriptor upload into desc (at least the
    // descriptor bytes): the copy is then asynchronous for the calling thread
    void *desc_host = nullptr;
    size_t desc_host_bytes = 0;
};
}  // namespace mkhip

struct mkhip_plan {
    mkhip_ctx *ctx = nullptr;
    // the context's tables and tuning knobs as they were when the plan was built:
    // every stage launches with this copy, so a knob changed later (which may
    // change a stage's geometry, e.g. half-limb external products that leave their
    // last inverse stage to a finish stage) cannot reach an existing plan
    mkhip::DevTables T;
    void *d_mem = nullptr;   // workspace
    void *d_desc = nullptr;  // descriptor tables
    size_t bytes = 0;        // workspace + descriptor bytes
    size_t desc_bytes = 0;   // descriptor bytes
    mkhip::PlanMem mem;      // external buffers / size query
    std::vector<char> host_desc;  // source of the asynchronous upload into mem.desc
    std::vector<PlanStage> stages;
    bool graph = false;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    std::vector<int> sig;  // shape signature (single-op cache)
};

namespace mkhip {

// 256-byte aligned bump allocator over one buffer.
struct Carve {
    size_t off = 0;
    size_t take(size_t bytes) {
        size_t o = (off + 255) & ~(size_t)255;
        off = o + bytes;
        return o;
    }
};

// Host copy of the descriptor tables of a plan, uploaded after the workspace.
struct DescBlob {
    std::vector<char> host;
    template <class D>
    size_t add(const std::vector<D> &v) {
        size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + std::max<size_t>(v.size(), 1) * sizeof(D));
        if (!v.empty()) std::memcpy(host.data() + off, v.data(), v.size() * sizeof(D));
        return off;
    }
};

// mkhip.hip: last stage + scaling of split inverse transforms (see Finish below)
int finish_rows(mkhip_ctx *c, u64 *base, const int *mods, int nrows, int tscale, bool split = false);
// Launch stage s (on `st`, default the context stream) / every stage.
hipError_t plan_stage(mkhip_plan *pl, int s, hipStream_t st = nullptr);
int plan_launch_raw(mkhip_plan *pl, hipStream_t st = nullptr);
// Capture every stage of a finished plan into one hipGraph (after one eager launch).
int plan_capture(mkhip_plan *pl);
// New plan whose workspace (ws_bytes) is allocated or taken from pm; *out is null
// on failure.
int plan_create(mkhip_ctx *c, size_t ws_bytes, mkhip_plan **out, const PlanMem &pm = PlanMem());
// Upload the descriptor tables, turn each stage's blob offset (stage.descs holds
// the offset as an integer on entry) into a device pointer, and optionally
// capture the whole launch sequence into a hipGraph.  Destroys the plan on
// failure, and after reporting the sizes of a query (pl is then null).
int plan_finish(mkhip_plan *&pl, const DescBlob &blob, int use_graph);
// Plan builders (the C entry points call them with a default PlanMem).
int plan_kklss_ckks(mkhip_ctx *c, int nops, const mkhip_kklss_op *ops, int use_graph, mkhip_plan **out,
                    const PlanMem &pm);
int plan_kklss_bfv(mkhip_ctx *c, int nops, const mkhip_kklss_op *ops, int use_graph, mkhip_plan **out,
                   const PlanMem &pm);
int plan_mkf_ckks(mkhip_ctx *c, int nops, const mkhip_mkf_op *ops, int use_graph, mkhip_plan **out,
                  const PlanMem &pm);
int plan_mkf_ckks_lv(mkhip_ctx *c, int nops, const mkhip_mkf_op *ops, int lv_l, int lv_r, int use_graph,
                     mkhip_plan **out, const PlanMem &pm);
int plan_mkf_bfv(mkhip_ctx *c, int nops, const mkhip_mkf_op *ops, int use_graph, mkhip_plan **out,
                 const PlanMem &pm);
// Upload one switching key (swk_words words) into a new device allocation.
int upload_key(mkhip_ctx *c, const uint64_t *src, u64 **dst);
// Drop the plan cached by the host-pointer entry points: it holds key pointers,
// so every key removal or replacement calls this before freeing the old key.
void drop_cached_plan(mkhip_ctx *c);
int plan_kklss_rotate(mkhip_ctx *c, int nops, const mkhip_rotate_op *ops, int use_graph, mkhip_plan **out,
                      const PlanMem &pm);
int plan_mkf_rotate(mkhip_ctx *c, int nops, const mkhip_mkf_rotate_op *ops, int use_graph, mkhip_plan **out,
                    const PlanMem &pm);
// Builder of a scheme (MKHIP_SCHEME_*) over ops of the matching struct type.
int plan_build(mkhip_ctx *c, int scheme, int nops, const void *ops, int use_graph, mkhip_plan **out,
               const PlanMem &pm);
inline PlanStage make_stage(int what, size_t blob_off, int n, int a, int b, int c, const char *name) {
    PlanStage s;
    s.what = what;
    s.descs = reinterpret_cast<const void *>(blob_off);
    s.n = n;
    s.a = a;
    s.b = b;
    s.c = c;
    s.name = name;
    return s;
}

// Limbs of 2^15 are inverse-transformed as two halves that leave the last stage
// and the N^-1 (or t N^-1) scaling to a KINTT_FINISH launch over the rows the
// inverse transform wrote.  The collector records those rows (nothing for
// smaller rings) and emits the stage "<after>_finish" right after its producer.
struct Finish {
    bool on = false;
    std::vector<FinishDesc> d;
    explicit Finish(const mkhip_ctx *c) : on(c->logN == 15) {}
    // split: the stage's kernel ran on half-limb workgroups at this ring size too
    Finish(const mkhip_ctx *c, bool split) : on(c->logN == 15 || split) {}
    void row(u64 *r, int mod, int tscale = 0) {
        if (!on) return;
        FinishDesc f;
        f.row = r;
        f.mod = mod;
        f.tscale = tscale;
        d.push_back(f);
    }
    // rows 0..nl+K-1 of a QP polynomial at level nl-1 (Q rows, then P rows)
    void qp(u64 *base, int nl, int L, int K, size_t N, bool p_only = false) {
        for (int r = p_only ? nl : 0; r < nl + K; ++r) row(base + (size_t)r * N, r < nl ? r : L + r - nl);
    }
    // outputs of a batched inverse NTT
    void jobs(const std::vector<NttJob> &v) {
        for (const NttJob &j : v) row(j.dst, j.mod, (j.reduce & 2) ? 1 : 0);
    }
    // the rows k_mac2 inverse-transforms: P rows, or every row with intt_all
    void mac2(const std::vector<Mac2Desc> &v, int L, int K, size_t N) {
        for (const Mac2Desc &m : v)
            for (int o = 0; o < 2; ++o) qp(m.acc[o], m.level + 1, L, K, N, !m.intt_all);
    }
    void emit(std::vector<PlanStage> &S, DescBlob &blob, const std::string &after) {
        if (!on || d.empty()) return;
        S.push_back(make_stage(KINTT_FINISH, blob.add(d), (int)d.size(), 0, 0, 1, (after + "_finish").c_str()));
        d.clear();
    }
};

}  // namespace mkhip
