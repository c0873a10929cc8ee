// CDNA4 kernels of the multi-key mul+relin path that follow the decompositions:
//
//   k_tensor         tensor product and INTT (keyswitch_hoisted.go:126-153)
//   k_ext            external-product MAC over the digits + INTT of every row
//                    (keyswitch_hoisted.go:11-45)
//   k_moddown        ModDown Q u P -> Q, accumulation into the output, and the
//                    fused rescale (basis_extension.go:192-232, mkckks/evaluator.go:360-401)
#include "kcommon.hpp"

namespace mkhip {

// ---------------------------------------------------------------------------
// tensor product + INTT ; logical block = (desc * nl + limb) * H + half

template <int LOGN>
__global__ void __launch_bounds__(Geo<LOGN>::TH)
    k_tensor(DevTables T, const TensorDesc *__restrict__ descs, int nl) {
    typedef Geo<LOGN> Gm;
    constexpr int N = Gm::N;
    const int tid = threadIdx.x;
    int bid = block_index(T);
    const int h = Gm::split ? bid % 2 : 0;
    bid /= Gm::H;
    const int l = uni(bid % nl);
    const TensorDesc &Dd = descs[uni(bid / nl)];
    const ModConst m = T.mod[l];
    const size_t off = (size_t)l * N + (size_t)h * Gm::NS;
    const size_t uoff = (size_t)l * (Dd.ustride > 1 ? Dd.ustride : 1) * N + (size_t)h * Gm::NS;
    u64 v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int qe = Quad<Gm::LS>::elem(tid, e);  // the inverse transform's input layout
        const size_t c = off + qe, cu = uoff + qe;
        u64 x;
        if (Dd.is_c0) {
            x = mulmod_m(G(Dd.a0)[c], G(Dd.b0)[c], m);
        } else {
            x = 0;
            if (Dd.au) x = mulmod_m(G(Dd.b0)[c], G(Dd.au)[cu], m);
            if (Dd.bu) x = add_mod(x, mulmod_m(G(Dd.a0)[c], G(Dd.bu)[cu], m), m.q);
        }
        v[e] = x;
    }
    inv_transform_sel<LOGN>(v, tid, h, T, l, m, false);
    store_row<Gm::LS>(v, Dd.out + off, tid);
}

// ---------------------------------------------------------------------------
// external product MAC + INTT ; logical block = row * ndesc + desc (descriptors
// that share key rows are adjacent, so their workgroups share an XCD's L2)

// NKEY keys share one hoisted decomposition D (the nested products with v_i and
// u, keyswitch_hoisted.go:177-193): one pass over D feeds both accumulators; the
// second is parked in its output row while the first is transformed.
// SPLIT at LOGN = 14: half-limb workgroups (two per CU); the last inverse stage
// is left to k_intt_finish as for 2^15
template <int LOGN, int NKEY, bool SPLIT = (LOGN == 15)>
__global__ void __launch_bounds__((Geo<LOGN, SPLIT>::TH), (LOGN == 14 && SPLIT ? 4 : 1))
    k_ext(DevTables T, const ExtDesc *__restrict__ descs, int beta, int rows) {
    typedef Geo<LOGN, SPLIT> Gm;
    constexpr int N = Gm::N;
    const int tid = threadIdx.x;
    int bid = block_index_ext(T);
    const int h = Gm::split ? bid % 2 : 0;
    bid /= Gm::H;
    const int ndesc = gridDim.x / (rows * Gm::H);
    const int r = uni(bid / ndesc);
    const ExtDesc &Dd = descs[uni(bid % ndesc)];
    const int ri = ring_index(r, Dd.level, T.L);
    const ModConst m = T.mod[ri];
    const size_t kstride = (size_t)(T.L + T.K) * N;
    const size_t hoff = (size_t)h * Gm::NS;
    u64 a0[16];
    u64 *o1 = NKEY > 1 ? Dd.acc[1] + (size_t)r * N + hoff : nullptr;
    if (fp_ok(m.q) && T.fp64) {
        double d0[16];
        ext_rows_fp<Gm::LS, NKEY>(d0, o1, Dd.D + (size_t)r * N + hoff, (size_t)rows * N,
                                  Dd.key[0] + (size_t)ri * N + hoff,
                                  NKEY > 1 ? Dd.key[1] + (size_t)ri * N + hoff : nullptr, kstride, beta, tid, m);
        const double *itwd = opaque(T.itwd + (size_t)ri * N);
        const int t = opaque_lane(tid);
        if constexpr (Gm::split)
            ntt_inv_fp_d<Gm::LS>(d0, a0, g_lds, t, itwd, fp_mod(m), 0.0, 2 + h, false);
        else
            ntt_inv_fp_d<LOGN>(d0, a0, g_lds, t, itwd, fp_mod(m), fp_from_u64(m.ninv), 1, true);
    } else {
        const u64 *dr = Dd.D + (size_t)r * N + hoff, *k0 = Dd.key[0] + (size_t)ri * N + hoff;
        const u64 *k1 = NKEY > 1 ? Dd.key[1] + (size_t)ri * N + hoff : nullptr;
        if (ext_cs_geometry<Gm::LS, NKEY>() && T.ext_cs && ext_cs_ok(m.q, beta))
            ext_rows_cs<Gm::LS, NKEY>(a0, o1, dr, (size_t)rows * N, k0, k1, kstride, beta, tid, m);
        else if (ext_lazy_ok(m.q, beta))
            ext_rows<Gm::LS, NKEY, true>(a0, o1, dr, (size_t)rows * N, k0, k1, kstride, beta, tid, m);
        else
            ext_rows<Gm::LS, NKEY>(a0, o1, dr, (size_t)rows * N, k0, k1, kstride, beta, tid, m);
        // opaque(): the second transform reloads its twiddles instead of keeping the
        // first one's live across it
        inv_transform_sel<LOGN, SPLIT>(a0, opaque_lane(tid), h, T, ri, m, false);
    }
    store_row<Gm::LS>(a0, Dd.acc[0] + (size_t)r * N + hoff, tid);
    if (NKEY > 1) {
        load_row_quad<Gm::LS>(a0, o1, tid);
        inv_transform_sel<LOGN, SPLIT>(a0, opaque_lane(tid), h, T, ri, m, false);
        store_row<Gm::LS>(a0, o1, tid);
    }
}

// ---------------------------------------------------------------------------
// ModDown (+ accumulate) (+ rescale) ; one coefficient per thread, grid (N/256, ndesc)

constexpr int kMaxQ = 16;

template <bool FINAL>
__global__ void __launch_bounds__(256) k_moddown(DevTables T, const ModDownDesc *__restrict__ descs) {
    const int N = T.N, L = T.L;
    const int c = blockIdx.x * 256 + threadIdx.x;
    const ModDownDesc &Dd = descs[blockIdx.y];
    const int nl = Dd.level + 1;
    u64 res[kMaxQ];
#pragma unroll
    for (int j = 0; j < kMaxQ; ++j) res[j] = (Dd.accumulate && j < nl) ? G(Dd.dst)[(size_t)j * N + c] : 0;
    const bool unf = Dd.unfinished != 0;
    for (int s = 0; s < Dd.nsrc; ++s) {
        const u64 *src = Dd.src[s];
        const PUp up = modup_p_prepare(T, src, nl, c, unf);
#pragma unroll
        for (int j = 0; j < kMaxQ; ++j) {
            if (j < nl) {
                const ModConst mq = T.mod[j];
                // ModUp below 3q, x canonical: the difference is in (q, 5q), a valid
                // Montgomery operand, so ModUp needs no reduction of its own
                const u64 z = modup_p_to_row_lazy(T, up, j, mq);
                const u64 x = row_word(T, src, j, j, c, unf);
                const u64 md = mred(z + 2 * mq.q - x, T.md_pinv[j], mq.q, mq.qinv);
                res[j] = add_mod(res[j], md, mq.q);
            }
        }
    }
    if (Dd.base) {
#pragma unroll
        for (int j = 0; j < kMaxQ; ++j)
            if (j < nl) res[j] = add_mod(res[j], G(Dd.base)[(size_t)j * N + c], T.mod[j].q);
    }
    if (!FINAL) {
#pragma unroll
        for (int j = 0; j < kMaxQ; ++j)
            if (j < nl) G(Dd.dst)[(size_t)j * N + c] = res[j];
        return;
    }
    // DivRoundByLastModulus nres times: out_j = (x_j + h - [x_l + h]_{q_l}) q_l^-1
    int top = nl;
    for (int n = 0; n < Dd.nres; ++n) {
        const int l = top - 1;
        const ModConst ml = T.mod[l];
        u64 xl = 0;
#pragma unroll
        for (int j = 0; j < kMaxQ; ++j)
            if (j == l) xl = res[j];
        const u64 t = add_mod(xl, ml.q >> 1, ml.q);
#pragma unroll
        for (int j = 0; j < kMaxQ; ++j) {
            if (j < l) {
                const ModConst mj = T.mod[j];
                const u64 tj = reduce64(t, mj.q, mj.mu);
                const u64 num = sub_mod(add_mod(res[j], T.rs_half[l * L + j], mj.q), tj, mj.q);
                res[j] = mred(num, T.rs_inv[l * L + j], mj.q, mj.qinv);
            }
        }
        top = l;
    }
#pragma unroll
    for (int j = 0; j < kMaxQ; ++j)
        if (j < top) G(Dd.out)[(size_t)j * N + c] = res[j];
}

// ---------------------------------------------------------------------------
// host launchers

template <class F>
static hipError_t allow_lds(F kernel, size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

template <int LOGN>
static hipError_t launch_ext_n(int what, const DevTables &T, const void *descs, int n, int a, int b,
                               hipStream_t st) {
    typedef Geo<LOGN> Gm;
    constexpr size_t lds = (size_t)lds_words(Gm::NS) * sizeof(u64);
    constexpr int H = Gm::H;
    static bool init = false;
    if (!init) {
        hipError_t e;
        if ((e = allow_lds(k_tensor<LOGN>, lds)) != hipSuccess) return e;
        if ((e = allow_lds(k_ext<LOGN, 1>, lds)) != hipSuccess) return e;
        if ((e = allow_lds(k_ext<LOGN, 2>, lds)) != hipSuccess) return e;
        init = true;
    }
    const dim3 block(Gm::TH);
    switch (what) {
    case KTENSOR:
        hipLaunchKernelGGL((k_tensor<LOGN>), dim3(n * a * H), block, lds, st, T, (const TensorDesc *)descs, a);
        break;
    case KEXT1:
    case KEXT2:
        if constexpr (LOGN == 14) {
            if (ext_split(T)) {
                static bool init_split = false;
                constexpr size_t lds2 = lds / 2;
                if (!init_split) {
                    hipError_t e;
                    if ((e = allow_lds(k_ext<14, 1, true>, lds2)) != hipSuccess) return e;
                    if ((e = allow_lds(k_ext<14, 2, true>, lds2)) != hipSuccess) return e;
                    init_split = true;
                }
                if (what == KEXT1)
                    hipLaunchKernelGGL((k_ext<14, 1, true>), dim3(n * b * 2), dim3(Gm::TH / 2), lds2, st, T,
                                       (const ExtDesc *)descs, a, b);
                else
                    hipLaunchKernelGGL((k_ext<14, 2, true>), dim3(n * b * 2), dim3(Gm::TH / 2), lds2, st, T,
                                       (const ExtDesc *)descs, a, b);
                break;
            }
        }
        if (what == KEXT1)
            hipLaunchKernelGGL((k_ext<LOGN, 1>), dim3(n * b * H), block, lds, st, T, (const ExtDesc *)descs, a, b);
        else
            hipLaunchKernelGGL((k_ext<LOGN, 2>), dim3(n * b * H), block, lds, st, T, (const ExtDesc *)descs, a, b);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// tensor: n descs x a (limbs); ext: n descs x b (rows), a = beta; moddown: n descs.
hipError_t launch_ext_kernel(int what, const DevTables &T, const void *descs, int n, int a, int b,
                             hipStream_t st) {
    if (what == KMODDOWN || what == KMODDOWN_FINAL) {
        if (what == KMODDOWN)
            hipLaunchKernelGGL((k_moddown<false>), dim3(T.N / 256, n), dim3(256), 0, st, T,
                               (const ModDownDesc *)descs);
        else
            hipLaunchKernelGGL((k_moddown<true>), dim3(T.N / 256, n), dim3(256), 0, st, T,
                               (const ModDownDesc *)descs);
        return hipGetLastError();
    }
    switch (T.logN) {
    case 13:
        return launch_ext_n<13>(what, T, descs, n, a, b, st);
    case 14:
        return launch_ext_n<14>(what, T, descs, n, a, b, st);
    case 15:
        return launch_ext_n<15>(what, T, descs, n, a, b, st);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace mkhip
