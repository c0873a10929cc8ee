// Fused KKLSS kernel for gadgets with one prime per digit (alpha = 1), where a
// gadget digit of a coefficient-domain poly is its limb i copied into every row
// (DecomposeAndSplit's copy path, mkrlwe/basis_extension.go:443-451):
//
//   k_decomp_mac    hoisted decomposition of every party's component fused with
//                   the x / y accumulation (mkrlwe/keyswitch_hoisted.go:84-123;
//                   for BFV the two-branch form, mkbfv/keyswitch_hoisted.go:100-148):
//                   each digit row is transformed once, stored for the later
//                   external products and multiplied into the accumulator while
//                   still in registers, so the digits are not read back.
//
// The grid is ordered so that the workgroups running together on one XCD are
// consecutive ops of the same (side, digit, row): they read the same key rows,
// which then come from that XCD's L2.
#include "kcommon.hpp"

namespace mkhip {

// blockIdx -> logical ((side * beta + i) * rows + r) * nops + o (order 0: the ops
// sharing a key row are adjacent) or ((side * beta + i) * nops + o) * rows + r
// (order 1: the rows sharing an input limb are adjacent); descs[2 o + side].
// Order 2 (XCD-tiled): each XCD takes a contiguous range of
// (((side * beta + i) * nrb + rb) * nops + o) * rbs + r', row r = rb rbs + r' --
// the rows are cut into nrb blocks of rbs, so an XCD walks every op of one
// (side, digit, row block): the op's input limbs serve the block's rbs rows from
// its L2 and the block's key rows (rbs x k rows, about 2.5 MB at k = 4) stay
// there across the ops, so that inputs are fetched about nrb times and keys about
// once per launch.
template <int LOGN, int MODE, bool SPLIT>
__device__ __forceinline__ void decomp_mac_row(const DecompDesc &Dd, int i, int r, int beta, int rows, int h,
                                               bool reduce, size_t krow, int tid, const ulonglong2 *tw,
                                               const ModConst &m) {
    typedef Geo<LOGN, SPLIT> Gm;
    constexpr int N = Gm::N;
    constexpr int TH = Gm::TH;
    const size_t hoff = (size_t)h * Gm::NS;
    u64 acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0;
    for (int p = 0; p < Dd.k; ++p) {
        u64 v[16];
        __builtin_amdgcn_sched_barrier(0);
        // values below 3 q_src may exceed the transform's bound 8m: reduce then.
        // MODE 1: the digit row is stored and multiplied unreduced (below 68q,
        // lazy_digits_ok); its consumers are Montgomery products
        fwd_transform_row<LOGN, MODE, SPLIT>(v, Dd.comp[p] + (size_t)i * N, reduce, opaque_lane(tid), h, opaque(tw),
                                             m, MODE != 1);
        __builtin_amdgcn_sched_barrier(0);
        store_row_quad<Gm::LS>(v, Dd.D + (((size_t)p * beta + i) * rows + r) * N + hoff, tid);
        const gu64 *key = uniform_ptr(Dd.key[p] + krow);
#pragma unroll
        for (int e = 0; e < 16; e += 2) {  // key words in the transform's Quad layout
            const u64v2 kw = *(const gv2 *)(key + Quad<Gm::LS>::elem(tid, e));
            if constexpr (MODE == 1) {
                // q < 2^57: digits below 68q < 2^63.1 are valid mred_lazy operands;
                // products in (0, 2q) summed unreduced, below 64 q for at most 32
                // parties
                acc[e] += mred_lazy(kw.x, v[e], m.q, m.qinv);
                acc[e + 1] += mred_lazy(kw.y, v[e + 1], m.q, m.qinv);
            } else {
                acc[e] = add_mod(acc[e], mred(kw.x, v[e], m.q, m.qinv), m.q);
                acc[e + 1] = add_mod(acc[e + 1], mred(kw.y, v[e + 1], m.q, m.qinv), m.q);
            }
        }
    }
    // MForm of the sum (below 64 q): a Shoup multiplication by 2^64 mod q, exact
    // for any operand below 2^64 and a third the instructions of a Montgomery
    // product by 2^128 mod q
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = shoup(acc[e], m.r1, m.r1p, m.q);
    store_row_quad<Gm::LS>(acc, Dd.xy + krow, tid);
}

// decomp_mac_row for a modulus below 2^50: FP64 transform and MAC.  The stored
// digit rows and the accumulator are the same canonical words as the integer
// form's: x = MForm(sum_p MRed(key_p, D_p)) = sum_p key_p D_p mod q.
template <int LOGN, bool SPLIT>
__device__ __forceinline__ void decomp_mac_row_fp(const DecompDesc &Dd, int i, int r, int beta, int rows, int h,
                                                  bool reduce, size_t krow, int tid, const double *twd,
                                                  const ModConst &m) {
    typedef Geo<LOGN, SPLIT> Gm;
    constexpr int N = Gm::N;
    const size_t hoff = (size_t)h * Gm::NS;
    const FpMod fm = fp_mod(m);
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = 0.0;
    // Unsplit limbs: each party's digit row is copied into LDS asynchronously while
    // the previous party's row finishes its transform and MAC, then read from LDS.
    if constexpr (!Gm::split)
        if (Dd.k > 0) row_to_lds_async<LOGN>(Dd.comp[0] + (size_t)i * N, tid);
    for (int p = 0; p < Dd.k; ++p) {
        double v[16];
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (Gm::split) {
            fwd_transform_row_fp<LOGN, SPLIT>(v, Dd.comp[p] + (size_t)i * N, reduce, opaque_lane(tid), h, opaque(twd),
                                              m);
        } else {
            const FpMod fm0 = fp_mod(m);
            const int t = opaque_lane(tid);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                u64 x = g_lds[t + e * Gm::TH];
                if (reduce) x = reduce64(x, m.q, m.mu);
                v[e] = fp_from_u64(x);
            }
            __syncthreads();
            // the accumulator is live across the transform: twiddles gathered per
            // stage only (registers)
            ntt_fwd_fp_head<LOGN, false>(v, g_lds, t, opaque(twd), fm0, 1, 0);
            if (p + 1 < Dd.k) row_to_lds_async<LOGN>(Dd.comp[p + 1] + (size_t)i * N, t);
            ntt_fwd_fp_tail<LOGN, false>(v, t, opaque(twd), fm0, 1, 0);
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = fp_canon(v[e], fm0);
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            u64 w[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) w[e] = fp_word(v[e]);
            store_row_quad<Gm::LS>(w, Dd.D + (((size_t)p * beta + i) * rows + r) * N + hoff, tid);
        }
        const gu64 *key = uniform_ptr(Dd.key[p] + krow);
        // products below 0.9q each: the accumulator is reduced every fourth party
        const bool red = (p & 3) == 3;
#pragma unroll
        for (int e = 0; e < 16; e += 2) {  // key words in the transform's Quad layout
            const u64v2 kw = *(const gv2 *)(key + Quad<Gm::LS>::elem(tid, e));
            acc[e] += fp_mulmod(fp_from_u64(kw.x), v[e], fm);
            acc[e + 1] += fp_mulmod(fp_from_u64(kw.y), v[e + 1], fm);
            if (red) {
                acc[e] = fp_reduce(acc[e], fm);
                acc[e + 1] = fp_reduce(acc[e + 1], fm);
            }
        }
    }
    u64 out[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) out[e] = fp_word(fp_canon(acc[e], fm));
    store_row_quad<Gm::LS>(out, Dd.xy + krow, tid);
}

// SPLIT (LOGN = 14): a limb over two workgroups of 512 threads and 64 KiB of LDS,
// at most 128 registers a lane, so that two workgroups share a CU
template <int LOGN, bool SPLIT>
__global__ void __launch_bounds__((Geo<LOGN, SPLIT>::TH), (LOGN == 14 && SPLIT ? 4 : 1))
    k_decomp_mac(DevTables T, const DecompDesc *__restrict__ descs, int nops, int beta, int rows, int order, int rbs) {
    typedef Geo<LOGN, SPLIT> Gm;
    constexpr int N = Gm::N;
    const int tid = threadIdx.x;
    int bid = order == 2 ? xcd_remap(blockIdx.x, gridDim.x) : block_index(T);
    const int h = Gm::split ? bid % 2 : 0;
    bid /= Gm::H;
    int o, r;
    if (order == 2) {
        const int rin = bid % rbs;
        bid /= rbs;
        o = bid % nops;
        bid /= nops;
        const int nrb = (rows + rbs - 1) / rbs;
        r = (bid % nrb) * rbs + rin;
        bid /= nrb;
        if (r >= rows) return;  // the last row block's unused slots (whole workgroup)
    } else if (order == 0) {
        o = bid % nops;
        bid /= nops;
        r = bid % rows;
        bid /= rows;
    } else {
        r = bid % rows;
        bid /= rows;
        o = bid % nops;
        bid /= nops;
    }
    o = uni(o);
    r = uni(r);
    const int i = uni(bid % beta);
    const int side = uni(bid / beta);
    const DecompDesc &Dd = descs[2 * o + side];
    const int ri = ring_index(r, Dd.level, T.L);
    const ModConst m = T.mod[ri];
    const ulonglong2 *tw = T.tw + (size_t)ri * N;
    const size_t hoff = (size_t)h * Gm::NS;
    const size_t krow = ((size_t)i * (T.L + T.K) + ri) * N + hoff;
    const bool reduce = T.mod[r_index(i, T.L, T.K)].q > 2 * m.q;
    // one instance of the loop per butterfly form (a wave-uniform branch), so that
    // neither form's registers are allocated around the other's
    if (fp_ok(m.q) && T.fp64)
        decomp_mac_row_fp<LOGN, SPLIT>(Dd, i, r, beta, rows, h, reduce, krow, tid, T.twd + (size_t)ri * N, m);
    else if (fwd_lazy_ok(m.q))
        decomp_mac_row<LOGN, 1, SPLIT>(Dd, i, r, beta, rows, h, reduce, krow, tid, tw, m);
    else
        decomp_mac_row<LOGN, 0, SPLIT>(Dd, i, r, beta, rows, h, reduce, krow, tid, tw, m);
}

template <int LOGN, bool SPLIT = (LOGN == 15)>
static hipError_t launch_fused_n(int what, const DevTables &T, const void *descs, int n, int a, int b,
                                 hipStream_t st) {
    typedef Geo<LOGN, SPLIT> Gm;
    constexpr size_t lds = (size_t)lds_words(Gm::NS) * sizeof(u64);
    static bool init = false;
    if (!init) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_decomp_mac<LOGN, SPLIT>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        init = true;
    }
    if (what != KDECOMP_MAC) return hipErrorInvalidValue;
    // n ops (2 descriptors each), a = beta, b = rows; order 2 cuts the rows into
    // blocks of at most 5 (their key rows fit one XCD's 4 MB L2 at k = 4)
    int rbs = b, nrb = 1;
    if (T.decomp_order == 2) {
        nrb = (b + 4) / 5;
        rbs = (b + nrb - 1) / nrb;
    }
    hipLaunchKernelGGL((k_decomp_mac<LOGN, SPLIT>), dim3(n * 2 * a * nrb * rbs * Gm::H), dim3(Gm::TH), lds, st, T,
                       (const DecompDesc *)descs, n, a, b, T.decomp_order, rbs);
    return hipGetLastError();
}

hipError_t launch_fused(int what, const DevTables &T, const void *descs, int n, int a, int b, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (T.alpha != 1) return hipErrorInvalidValue;
    if (T.logN == 13) return launch_fused_n<13>(what, T, descs, n, a, b, st);
    if (T.logN == 14)
        return T.split14 ? launch_fused_n<14, true>(what, T, descs, n, a, b, st)
                         : launch_fused_n<14, false>(what, T, descs, n, a, b, st);
    if (T.logN == 15) return launch_fused_n<15>(what, T, descs, n, a, b, st);
    return hipErrorInvalidValue;
}

}  // namespace mkhip
