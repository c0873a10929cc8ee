// CDNA4 kernels of the multi-key mul+relin path.
//
//   k_ntt_batch      batched limb NTT / INTT (tensor inputs, MKF switches)
//   k_decomp         gadget decomposition + NTT of every digit row, fused with the
//                    KKLSS x/y accumulation (mkrlwe/keyswitch.go:22-76,
//                    keyswitch_hoisted.go:84-123)
//
// The tensor product, the external products and the ModDowns are in
// kernels_ext.hip.
//
// Every kernel keeps one RNS limb of one polynomial per workgroup (N/16 threads,
// 16 coefficients per thread) or one coefficient per thread, reads and writes
// limb-major rows coalesced, and produces canonical residues.
#include "kcommon.hpp"

#include <algorithm>

namespace mkhip {

// kernels_ext.hip: KTENSOR, KEXT1, KEXT2, KMODDOWN, KMODDOWN_FINAL
hipError_t launch_ext_kernel(int what, const DevTables &T, const void *descs, int n, int a, int b,
                             hipStream_t st);

// ---------------------------------------------------------------------------
// batched NTT ; blockIdx.x = job * H + half

template <int LOGN, bool INV>
__global__ void __launch_bounds__(Geo<LOGN>::TH) k_ntt_batch(DevTables T, const NttJob *__restrict__ jobs) {
    typedef Geo<LOGN> Gm;
    const int tid = threadIdx.x;
    const int h = Gm::split ? (int)(blockIdx.x % 2) : 0;
    const NttJob j = jobs[blockIdx.x / Gm::H];
    const ModConst m = T.mod[j.mod];
    const bool red = j.reduce & 1;
    u64 v[16];
    if (INV) {
        load_row_quad<Gm::LS>(v, j.src + (size_t)h * Gm::NS, tid);
        if (red) {
#pragma unroll
            for (int i = 0; i < 16; ++i) v[i] = reduce64(v[i], m.q, m.mu);
        }
        inv_transform_sel<LOGN>(v, tid, h, T, j.mod, m, (j.reduce & 2) != 0);
    } else {
        const ulonglong2 *tw = T.tw + (size_t)j.mod * Gm::N;
        if constexpr (Gm::split) {
            // In place, a half may not be overwritten while the other workgroup still
            // reads it for the first stage: workgroup 0 then transforms both halves
            // (the upper one parked in its destination meanwhile) and workgroup 1 exits.
            if (j.src == j.dst) {
                if (h) return;
                u64 hi[16];
                load_row_cached<Gm::LS>(v, j.src, tid);
                load_row_cached<Gm::LS>(hi, j.src + Gm::NS, tid);
                if (red) {
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        v[e] = reduce64(v[e], m.q, m.mu);
                        hi[e] = reduce64(hi[e], m.q, m.mu);
                    }
                }
                fwd_first_stage_both(v, hi, tw, m.q);
                store_row<Gm::LS>(hi, j.dst + Gm::NS, tid);
                ntt_fwd_any<Gm::LS>(v, g_lds, tid, tw, m.q, m.mu, 2);
                store_row_quad<Gm::LS>(v, j.dst, tid);
                load_row_cached<Gm::LS>(v, j.dst + Gm::NS, tid);
                ntt_fwd_any<Gm::LS>(v, g_lds, opaque_lane(tid), opaque(tw), m.q, m.mu, 3);
                store_row_quad<Gm::LS>(v, j.dst + Gm::NS, tid);
                return;
            }
        }
        if (fwd_lazy_ok(m.q))
            fwd_transform_row<LOGN, 1>(v, j.src, red, tid, h, tw, m);
        else
            fwd_transform_row<LOGN, 0>(v, j.src, red, tid, h, tw, m);
    }
    // inverse output: natural layout; forward output: Quad layout
    if (INV)
        store_row<Gm::LS>(v, j.dst + (size_t)h * Gm::NS, tid);
    else
        store_row_quad<Gm::LS>(v, j.dst + (size_t)h * Gm::NS, tid);
}

// ---------------------------------------------------------------------------
// gadget digit of one coefficient into modulus `m` (ring index ri)

// Exact basis extension of the digit's nd residues to modulus m: y_l = MRed(x_l,
// (D/q_l)^-1 R), v = trunc(sum_l float64(y_l)/float64(q_l)) accumulated in
// ascending l, out = sum_l y_l (D/q_l) - v D mod m  (basis_extension.go:471-533).
__device__ __forceinline__ u64 digit_modup(const DevTables &T, const u64 *__restrict__ src, int c, int start,
                                           int nd, int ebase, int ri, const ModConst &m) {
    const int N = T.N;
    const u64 *yc = T.dc_y + (size_t)ebase * kMaxAlpha;
    const u64 *qs = T.dc_qstar + ((size_t)ebase * (T.L + T.K) + ri) * kMaxAlpha;
    const u64 *vt = T.dc_vt + ((size_t)ebase * (T.L + T.K) + ri) * (kMaxAlpha + 1);
    u64 hi = 0, lo = 0;
    double vi = 0.0;
    for (int l = 0; l < nd; ++l) {
        const ModConst ml = T.mod[start + l];
        const u64 x = G(src)[(size_t)(start + l) * N + c];
        const u64 y = mred(x, yc[l], ml.q, ml.qinv);
        vi += (double)y / (double)ml.q;
        mac128(hi, lo, y, qs[l]);
    }
    const u64 v = (u64)vi;
    u64 r = mont_reduce128(hi, lo, m.q, m.qinv) + vt[v];
    return reduce64(r, m.q, m.mu);
}

// digit_modup with the digit-only part (y_l, v) read from k_digit_pre's rows
__device__ __forceinline__ u64 digit_modup_pre(const DevTables &T, const u64 *__restrict__ pre, int c, int nd,
                                               int ebase, int ri, const ModConst &m) {
    const int N = T.N;
    const u64 *qs = T.dc_qstar + ((size_t)ebase * (T.L + T.K) + ri) * kMaxAlpha;
    const u64 *vt = T.dc_vt + ((size_t)ebase * (T.L + T.K) + ri) * (kMaxAlpha + 1);
    u64 hi = 0, lo = 0;
    for (int l = 0; l < nd; ++l) mac128(hi, lo, G(pre)[(size_t)l * N + c], qs[l]);
    const u64 v = G(pre)[(size_t)T.alpha * N + c];
    u64 r = mont_reduce128(hi, lo, m.q, m.qinv) + vt[v];
    return reduce64(r, m.q, m.mu);
}

// The part of digit_modup every target row of a CRT digit repeats -- y_l =
// MRed(x_l, (D/q_l)^-1 R) and v = trunc(sum_l y_l / q_l), same operations in the
// same order -- once per (component, digit, coefficient): rows l < nd of the
// digit's block hold y_l, row alpha holds v.  grid (N / 256, descs x kmax x beta).
__global__ void __launch_bounds__(256) k_digit_pre(DevTables T, const DecompDesc *__restrict__ descs, int kmax,
                                                   int beta) {
    const int N = T.N;
    const int c = blockIdx.x * 256 + threadIdx.x;
    int b = blockIdx.y;
    const int i = b % beta;
    b /= beta;
    const int p = b % kmax;
    const DecompDesc &Dd = descs[b / kmax];
    if (p >= Dd.k || !Dd.pre) return;
    const int alpha = T.alpha, start = i * alpha;
    const int nd = digit_primes(i, max(Dd.src_level, Dd.level), alpha);
    if (nd == 1) return;  // a one-prime digit is copied, not extended
    const int ebase = i * (kMaxAlpha + 1) + nd;
    const u64 *yc = T.dc_y + (size_t)ebase * kMaxAlpha;
    u64 *pre = Dd.pre + ((size_t)p * beta + i) * (alpha + 1) * N;
    const u64 *src = Dd.comp[p];
    double vi = 0.0;
    for (int l = 0; l < nd; ++l) {
        const ModConst ml = T.mod[start + l];
        const u64 y = mred(G(src)[(size_t)(start + l) * N + c], yc[l], ml.q, ml.qinv);
        vi += (double)y / (double)ml.q;
        G(pre)[(size_t)l * N + c] = y;
    }
    G(pre)[(size_t)alpha * N + c] = (u64)vi;
}

// One digit row: the digit's value (copy path / the digit's own rows: the residue
// itself, reduced mod m only when it may exceed the transform's input bound 8m --
// source values are below 3 q_src: canonical, or multSum's raw outputs in a BFV
// lift; otherwise the exact CRT ModUp), forward-transformed, stored at dst.
template <int LOGN, bool CRT, int MODE, bool SPLIT = (LOGN == 15)>
__device__ __forceinline__ void decomp_row(const DevTables &T, const u64 *src, u64 *dst, int start, int nd,
                                           bool own_row, int i, int r, int ri, int h, int tid,
                                           const ulonglong2 *tw, const ModConst &m, const u64 *pre, bool canon) {
    typedef Geo<LOGN, SPLIT> Gm;
    constexpr int N = Gm::N;
    u64 v[16];
    if (!CRT || nd == 1 || own_row) {
        const int sl = own_row ? r : start;
        const bool red = T.mod[r_index(sl, T.L, T.K)].q > 2 * m.q;
        fwd_transform_row<LOGN, MODE, SPLIT>(v, src + (size_t)sl * N, red, tid, h, tw, m, canon);
    } else {
        const int ebase = i * (kMaxAlpha + 1) + nd;
        if (pre)
            fwd_transform<LOGN, MODE, SPLIT>(
                v, [&](int c) { return digit_modup_pre(T, pre, c, nd, ebase, ri, m); }, tid, h, tw, m);
        else
            fwd_transform<LOGN, MODE, SPLIT>(
                v, [&](int c) { return digit_modup(T, src, c, start, nd, ebase, ri, m); }, tid, h, tw, m);
    }
    store_row_quad<Gm::LS>(v, dst, tid);
}

// decomp_row in FP64 for a modulus below 2^50 (the CRT rows' ModUp stays integer;
// its canonical output is transformed in FP64).
template <int LOGN, bool CRT, bool SPLIT = (LOGN == 15)>
__device__ __forceinline__ void decomp_fp(const DevTables &T, const u64 *src, u64 *dst, int start, int nd,
                                          bool own_row, int i, int r, int ri, int h, int tid, const ModConst &m,
                                          const u64 *pre) {
    typedef Geo<LOGN, SPLIT> Gm;
    constexpr int N = Gm::N;
    const double *twd = T.twd + (size_t)ri * N;
    double v[16];
    if (!CRT || nd == 1 || own_row) {
        const int sl = own_row ? r : start;
        const bool red = T.mod[r_index(sl, T.L, T.K)].q > 2 * m.q;
        fwd_transform_row_fp<LOGN, SPLIT>(v, src + (size_t)sl * N, red, tid, h, twd, m);
    } else {
        const int ebase = i * (kMaxAlpha + 1) + nd;
        if (pre)
            fwd_transform_fp<LOGN, SPLIT>(
                v, [&](int c) { return digit_modup_pre(T, pre, c, nd, ebase, ri, m); }, tid, h, twd, m);
        else
            fwd_transform_fp<LOGN, SPLIT>(
                v, [&](int c) { return digit_modup(T, src, c, start, nd, ebase, ri, m); }, tid, h, twd, m);
    }
    u64 w[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) w[e] = fp_word(v[e]);
    store_row_quad<Gm::LS>(w, dst, tid);
}

// decomp_row for an unsplit limb (N <= 2^14).
template <int LOGN, bool CRT, int MODE>
__device__ __forceinline__ void decomp_flat(const DevTables &T, const u64 *src, u64 *dst, int start, int nd,
                                            bool own_row, int i, int r, int ri, int tid, const ulonglong2 *tw,
                                            const ModConst &m, const u64 *pre, bool canon) {
    constexpr int N = 1 << LOGN;
    u64 v[16];
    if (!CRT || nd == 1 || own_row) {
        const int sl = own_row ? r : start;
        load_row_cached<LOGN>(v, src + (size_t)sl * N, tid);
        if (T.mod[r_index(sl, T.L, T.K)].q > 2 * m.q) {
#pragma unroll
            for (int e = 0; e < 16; ++e) v[e] = reduce64(v[e], m.q, m.mu);
        }
    } else {
        const int ebase = i * (kMaxAlpha + 1) + nd;
        if (pre)
            for (int e = 0; e < 16; ++e) v[e] = digit_modup_pre(T, pre, tid + e * (N / 16), nd, ebase, ri, m);
        else
            for (int e = 0; e < 16; ++e) v[e] = digit_modup(T, src, tid + e * (N / 16), start, nd, ebase, ri, m);
    }
    fwd_ntt_mode<LOGN, MODE>(v, tid, tw, m, 1, canon);
    store_row_quad<LOGN>(v, dst, tid);
}

// One workgroup per (desc, component p, digit i, row r[, half]): logical block
// (((desc * kmax + p) * beta + i) * rows + r) * H + h, XCD-remapped so the rows of
// one (component, digit) -- which all read the same input limbs -- share an L2.
// SPLIT at LOGN = 14: a limb over two workgroups of half the LDS (two per CU)
template <int LOGN, bool CRT, bool SPLIT = (LOGN == 15)>
__global__ void __launch_bounds__((Geo<LOGN, SPLIT>::TH), (LOGN == 14 && SPLIT ? 4 : 1))
    k_decomp(DevTables T, const DecompDesc *__restrict__ descs, int kmax, int beta, int rows) {
    typedef Geo<LOGN, SPLIT> Gm;
    constexpr int N = Gm::N;
    const int tid = threadIdx.x;
    int bid = block_index(T);
    const int h = Gm::split ? bid % 2 : 0;
    bid /= Gm::H;
    const int r = uni(bid % rows);
    const int i = uni((bid / rows) % beta);
    const int p = uni((bid / (rows * beta)) % kmax);
    const int d = uni(bid / (rows * beta * kmax));
    const DecompDesc &Dd = descs[d];
    if (p >= Dd.k) return;
    const int level = Dd.level;
    const int ri = ring_index(r, level, T.L);
    const ModConst m = T.mod[ri];
    const int alpha = T.alpha;
    const int start = i * alpha;
    const int nd = digit_primes(i, max(Dd.src_level, level), alpha);
    const bool own_row = (r <= level) && (r >= start) && (r < start + nd);
    const u64 *src = Dd.comp[p];
    const ulonglong2 *tw = T.tw + (size_t)ri * N;
    u64 *dst = Dd.D + (((size_t)p * beta + i) * rows + r) * N + (size_t)h * Gm::NS;
    const u64 *pre = (CRT && Dd.pre) ? Dd.pre + ((size_t)p * beta + i) * (alpha + 1) * N : nullptr;
    // unreduced digit rows (below 68q) for the plans' Montgomery consumers
    const bool canon = !(Dd.lazy_out && lazy_digits_ok(m.q));
    if (fp_ok(m.q) && T.fp64) {
        decomp_fp<LOGN, CRT, SPLIT>(T, src, dst, start, nd, own_row, i, r, ri, h, tid, m, pre);
        return;
    }
    if constexpr (!Gm::split) {
        const bool crt_row = CRT && !(nd == 1 || own_row);
        if (!crt_row && fwd_lazy_ok(m.q))
            decomp_flat<LOGN, CRT, 1>(T, src, dst, start, nd, own_row, i, r, ri, tid, tw, m, pre, canon);
        else
            decomp_flat<LOGN, CRT, 0>(T, src, dst, start, nd, own_row, i, r, ri, tid, tw, m, pre, true);
        return;
    }
    if (CRT && !(nd == 1 || own_row)) {  // CRT digit rows: one ModUp per value, then the transform
        decomp_row<LOGN, true, 0, SPLIT>(T, src, dst, start, nd, own_row, i, r, ri, h, tid, tw, m, pre, true);
    } else if (fwd_lazy_ok(m.q)) {
        decomp_row<LOGN, false, 1, SPLIT>(T, src, dst, start, nd, own_row, i, r, ri, h, tid, tw, m, pre, canon);
    } else {
        decomp_row<LOGN, false, 0, SPLIT>(T, src, dst, start, nd, own_row, i, r, ri, h, tid, tw, m, pre, true);
    }
}

// Key-stationary x / y accumulation: grid (N/256, beta*rows, 2 sides); the
// thread of coefficient c of (digit i, row r) walks all nops ops of the batch.
// x / y are written as full-level switching keys ([beta][L+K][N], row = ring
// index), the layout k_ext reads keys in.
__global__ void __launch_bounds__(256)
    k_mac_xy(DevTables T, const MacDesc *__restrict__ descs, int nops, int beta, int rows) {
    const int N = T.N;
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y / rows, r = blockIdx.y % rows;
    const int side = blockIdx.z;
    const size_t kstride = (size_t)(T.L + T.K) * N;
    for (int o = 0; o < nops; ++o) {
        const MacDesc &Md = descs[2 * o + side];
        const int ri = ring_index(r, Md.level, T.L);
        const ModConst m = T.mod[ri];
        const size_t koff = (size_t)i * kstride + (size_t)ri * N + c;
        const size_t doff = ((size_t)i * rows + r) * N + c;
        const size_t dstride = (size_t)beta * rows * N;
        u64 hi = 0, lo = 0;
        for (int p = 0; p < Md.k; ++p) mac128(hi, lo, G(Md.key[p])[koff], G(Md.D)[p * dstride + doff]);
        // MForm(sum_p MRed(key_p, D_p)); full-level key layout, like the keys
        G(Md.xy)[koff] = shoup(redc_sum(hi, lo, m), m.r1, m.r1p, m.q);
    }
}

// ---------------------------------------------------------------------------
// last inverse stage of split limbs ; grid (N/2/256, ndesc), thread = pair (c, c + N/2)

__global__ void __launch_bounds__(256) k_intt_finish(DevTables T, const FinishDesc *__restrict__ descs) {
    const int half = T.N / 2;
    const int c = blockIdx.x * 256 + threadIdx.x;
    const FinishDesc &Dd = descs[blockIdx.y];
    const ModConst m = T.mod[Dd.mod];
    gu64 *row = G(Dd.row);
    u64 x = row[c], y = row[c + half];
    if (Dd.tscale)
        inv_last_stage(x, y, m.twn, m.twnp, m.tninv, m.tninvp, m.q);
    else
        inv_last_stage(x, y, m.wn, m.wnp, m.ninv, m.ninvp, m.q);
    row[c] = x;
    row[c + half] = y;
}

// ---------------------------------------------------------------------------
// host launchers

template <class F>
static hipError_t allow_lds(F kernel, size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

template <int LOGN>
static hipError_t launch_all(int what, const DevTables &T, const void *descs, int n, int a, int b, int c,
                             hipStream_t st) {
    typedef Geo<LOGN> Gm;
    constexpr size_t lds = (size_t)lds_words(Gm::NS) * sizeof(u64);
    constexpr int H = Gm::H;
    static bool init = false;
    if (!init) {
        hipError_t e;
        if ((e = allow_lds(k_ntt_batch<LOGN, false>, lds)) != hipSuccess) return e;
        if ((e = allow_lds(k_ntt_batch<LOGN, true>, lds)) != hipSuccess) return e;
        if ((e = allow_lds(k_decomp<LOGN, false>, lds)) != hipSuccess) return e;
        if ((e = allow_lds(k_decomp<LOGN, true>, lds)) != hipSuccess) return e;
        init = true;
    }
    const dim3 block(Gm::TH);
    switch (what) {
    case KNTT_FWD:
        hipLaunchKernelGGL((k_ntt_batch<LOGN, false>), dim3(n * H), block, lds, st, T, (const NttJob *)descs);
        break;
    case KNTT_INV:
        hipLaunchKernelGGL((k_ntt_batch<LOGN, true>), dim3(n * H), block, lds, st, T, (const NttJob *)descs);
        break;
    case KDECOMP:  // n descs, a = beta, b = rows, c = kmax
        if constexpr (LOGN == 14) {
            // one-prime digits (copy path) over half-limb workgroups when tuned so
            if (T.split14 && T.alpha == 1) {
                static bool init_split = false;
                constexpr size_t lds2 = lds / 2;
                if (!init_split) {
                    hipError_t e = allow_lds(k_decomp<14, false, true>, lds2);
                    if (e != hipSuccess) return e;
                    init_split = true;
                }
                hipLaunchKernelGGL((k_decomp<14, false, true>), dim3(n * c * a * b * 2), dim3(Gm::TH / 2), lds2, st,
                                   T, (const DecompDesc *)descs, c, a, b);
                break;
            }
        }
        if (T.alpha > 1)
            hipLaunchKernelGGL((k_decomp<LOGN, true>), dim3(n * c * a * b * H), block, lds, st, T,
                               (const DecompDesc *)descs, c, a, b);
        else
            hipLaunchKernelGGL((k_decomp<LOGN, false>), dim3(n * c * a * b * H), block, lds, st, T,
                               (const DecompDesc *)descs, c, a, b);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Plain device-to-device copy (16-byte words, grid-stride).
__global__ void __launch_bounds__(256) k_copy(const ulonglong2 *__restrict__ src, ulonglong2 *__restrict__ dst,
                                             size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    if (n) {
        const int grid = (int)std::min<size_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, st, (const ulonglong2 *)src, (ulonglong2 *)dst, n);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const size_t tail = bytes - n * 16;
    if (tail) return hipMemcpyAsync((char *)dst + n * 16, (const char *)src + n * 16, tail, hipMemcpyDeviceToDevice, st);
    return hipSuccess;
}

// what: KernelId (tables.hpp).  NTT: n jobs; decomp: n descs x c (kmax) x a (beta) x
// b (rows); tensor: n descs x a (limbs); ext: n descs x b (rows), a = beta; moddown:
// n descs; mac_xy: n ops (2 descs each), a = beta, b = rows.
hipError_t launch_kernel(int what, const DevTables &T, const void *descs, int n, int a, int b, hipStream_t st,
                         int c) {
    if (n <= 0) return hipSuccess;
    if (what == KDIGIT_PRE) {  // n descs, a = beta, c = kmax
        hipLaunchKernelGGL(k_digit_pre, dim3(T.N / 256, n * c * a), dim3(256), 0, st, T, (const DecompDesc *)descs,
                           c, a);
        return hipGetLastError();
    }
    if (what == KMAC_XY) {
        hipLaunchKernelGGL(k_mac_xy, dim3(T.N / 256, a * b, 2), dim3(256), 0, st, T, (const MacDesc *)descs, n,
                           a, b);
        return hipGetLastError();
    }
    if (what == KINTT_FINISH) {  // n descs
        hipLaunchKernelGGL(k_intt_finish, dim3(T.N / 512, n), dim3(256), 0, st, T, (const FinishDesc *)descs);
        return hipGetLastError();
    }
    if (what == KTENSOR || what == KEXT1 || what == KEXT2 || what == KMODDOWN || what == KMODDOWN_FINAL)
        return launch_ext_kernel(what, T, descs, n, a, b, st);
    switch (T.logN) {
    case 13:
        return launch_all<13>(what, T, descs, n, a, b, c, st);
    case 14:
        return launch_all<14>(what, T, descs, n, a, b, c, st);
    case 15:
        return launch_all<15>(what, T, descs, n, a, b, c, st);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace mkhip
