// CDNA4 kernels of the KKLSS MK-BFV path (software/kklss/mkbfv):
//
//   k_bfv_lift    ModUpQtoR / Rescale of a Q poly into R = Q u QMul
//                 (basis_extension.go:50-66, 88-104), one coefficient per thread
//   k_tensor_r    tensor product over R, MulScalar(t) folded into the INTT's
//                 N^-1 (keyswitch_hoisted.go:154-200, basis_extension.go:76-77)
//   k_quantize    Quantize's ModDown R -> Q (basis_extension.go:79-84), one
//                 coefficient per thread, optionally accumulating
//
// The basis extensions keep multSum's unreduced outputs where the reference
// does (mkrlwe/basis_extension.go:581-646): the R limbs a lift writes are later
// copied across moduli by the decomposition, which reduces their integer value.
#include "kcommon.hpp"

namespace mkhip {

constexpr int kMaxL = 16;

// Exact basis extension of one coefficient (ModUp of modUpExact): y_i =
// MRed(x_i, ystar_i), v = trunc(sum_i float64(y_i)/float64(q_i)) in ascending
// i (reconstructRNS, :537-579); out_j = hi - hi(lo pinv p) + p + vt[v]
// (multSum, :582-646), returned unreduced.
struct Up {
    u64 y[kMaxL];
    u64 v;
};

__device__ __forceinline__ Up up_prepare(const DevTables &T, const u64 (&x)[kMaxL], int n, int ring0,
                                         const u64 *__restrict__ ystar) {
    Up u;
    double vi = 0.0;
#pragma unroll
    for (int i = 0; i < kMaxL; ++i) {
        u.y[i] = 0;
        if (i < n) {
            const ModConst m = T.mod[ring0 + i];
            u.y[i] = mred(x[i], ystar[i], m.q, m.qinv);
            vi += (double)u.y[i] / (double)m.q;
        }
    }
    u.v = (u64)vi;
    return u;
}

__device__ __forceinline__ u64 up_raw(const Up &u, int n, const u64 *__restrict__ qstar_row,
                                      const u64 *__restrict__ vt_row, const ModConst &m) {
    u64 hi = 0, lo = 0;
#pragma unroll
    for (int i = 0; i < kMaxL; ++i)
        if (i < n) mac128(hi, lo, u.y[i], qstar_row[i]);
    return mont_reduce128(hi, lo, m.q, m.qinv) + vt_row[u.v];
}

// grid (N/256, ndesc): mode 0 ModUpQtoR, mode 1 Rescale.
__global__ void __launch_bounds__(256) k_bfv_lift(DevTables T, const BfvLiftDesc *__restrict__ descs) {
    const int N = T.N, L = T.L, K = T.K;
    const int c = blockIdx.x * 256 + threadIdx.x;
    const BfvLiftDesc &D = descs[blockIdx.y];
    const int qm0 = L + K;  // ring index of QMul limb 0
    u64 x[kMaxL];
#pragma unroll
    for (int i = 0; i < kMaxL; ++i) {
        x[i] = 0;
        if (i < L) {
            const u64 xi = G(D.src)[(size_t)i * N + c];
            if (D.mode == 0) {
                G(D.dst)[(size_t)i * N + c] = xi;
                x[i] = xi;
            } else {
                const ModConst m = T.mod[i];
                x[i] = mred(xi, T.bq_qm[i], m.q, m.qinv);  // x QMul mod q_i
            }
        }
    }
    const Up u = up_prepare(T, x, L, 0, T.bq_up_y);
    if (D.mode == 0) {
#pragma unroll
        for (int j = 0; j < kMaxL; ++j) {
            if (j < L) {
                const ModConst m = T.mod[qm0 + j];
                G(D.dst)[(size_t)(L + j) * N + c] = up_raw(u, L, T.bq_up_qstar + (size_t)j * L,
                                                        T.bq_up_vt + (size_t)j * (L + 1), m);
            }
        }
        return;
    }
    // Rescale: ModDownQPtoP with a zero QMul part, then ModUpPtoQ
    u64 mq[kMaxL];
#pragma unroll
    for (int j = 0; j < kMaxL; ++j) {
        mq[j] = 0;
        if (j < L) {
            const ModConst m = T.mod[qm0 + j];
            const u64 z = up_raw(u, L, T.bq_up_qstar + (size_t)j * L, T.bq_up_vt + (size_t)j * (L + 1), m);
            mq[j] = mred(z + 2 * m.q, T.bq_rsc[j], m.q, m.qinv);
            G(D.dst)[(size_t)(L + j) * N + c] = mq[j];
        }
    }
    const Up w = up_prepare(T, mq, L, qm0, T.bq_dn_y);
#pragma unroll
    for (int i = 0; i < kMaxL; ++i) {
        if (i < L) {
            const ModConst m = T.mod[i];
            G(D.dst)[(size_t)i * N + c] = up_raw(w, L, T.bq_dn_qstar + (size_t)i * L, T.bq_dn_vt + (size_t)i * (L + 1), m);
        }
    }
}

// grid (N/256, ndesc): dst (=|+=) (x_Q - ModUp_{QMul->Q}(x_QMul)) QMul^-1 via
// MRed(y + 2q - x, q - MForm(QMul^-1)) (mkrlwe/basis_extension.go:192-232).
__global__ void __launch_bounds__(256) k_quantize(DevTables T, const BfvQuantDesc *__restrict__ descs) {
    const int N = T.N, L = T.L, K = T.K;
    const int c = blockIdx.x * 256 + threadIdx.x;
    const BfvQuantDesc &D = descs[blockIdx.y];
    u64 x[kMaxL];
#pragma unroll
    for (int j = 0; j < kMaxL; ++j) x[j] = j < L ? G(D.src)[(size_t)(L + j) * N + c] : 0;
    const Up u = up_prepare(T, x, L, L + K, T.bq_dn_y);
#pragma unroll
    for (int i = 0; i < kMaxL; ++i) {
        if (i < L) {
            const ModConst m = T.mod[i];
            const u64 y = up_raw(u, L, T.bq_dn_qstar + (size_t)i * L, T.bq_dn_vt + (size_t)i * (L + 1), m);
            const u64 xq = G(D.src)[(size_t)i * N + c];
            u64 r = mred(y + 2 * m.q - xq, T.bq_qz[i], m.q, m.qinv);
            if (D.accumulate) r = add_mod(r, G(D.dst)[(size_t)i * N + c], m.q);
            G(D.dst)[(size_t)i * N + c] = r;
        }
    }
}

// Tensor over R ; logical block = (desc * 2L + limb) * H + half.  Inputs: NTT-domain
// R polys; output: t * INTT(product), coefficient domain, R layout.
template <int LOGN>
__global__ void __launch_bounds__(Geo<LOGN>::TH)
    k_tensor_r(DevTables T, const TensorDesc *__restrict__ descs, int nl) {
    typedef Geo<LOGN> Gm;
    constexpr int N = Gm::N;
    const int tid = threadIdx.x;
    int bid = block_index(T);
    const int h = Gm::split ? bid % 2 : 0;
    bid /= Gm::H;
    const int l = uni(bid % nl);
    const TensorDesc &Dd = descs[uni(bid / nl)];
    const int ri = r_index(l, T.L, T.K);
    const ModConst m = T.mod[ri];
    const size_t off = (size_t)l * N + (size_t)h * Gm::NS;
    u64 v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const size_t c = off + Quad<Gm::LS>::elem(tid, e);  // the inverse transform's input layout
        u64 x;
        if (Dd.is_c0) {
            x = mulmod_m(G(Dd.a0)[c], G(Dd.b0)[c], m);
        } else {
            x = 0;
            if (Dd.au) x = mulmod_m(G(Dd.b0)[c], G(Dd.au)[c], m);
            if (Dd.bu) x = add_mod(x, mulmod_m(G(Dd.a0)[c], G(Dd.bu)[c], m), m.q);
        }
        v[e] = x;
    }
    inv_transform<LOGN>(v, tid, h, T.itw + (size_t)ri * N, m.q, m.tninv, m.tninvp, T.inv_lazy ? m.mu : 0);
    store_row<Gm::LS>(v, Dd.out + off, tid);
}

template <int LOGN>
static hipError_t launch_tensor_r(const DevTables &T, const void *descs, int n, int nl, hipStream_t st) {
    typedef Geo<LOGN> Gm;
    constexpr size_t lds = (size_t)lds_words(Gm::NS) * sizeof(u64);
    static bool init = false;
    if (!init) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_tensor_r<LOGN>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        init = true;
    }
    hipLaunchKernelGGL((k_tensor_r<LOGN>), dim3(n * nl * Gm::H), dim3(Gm::TH), lds, st, T,
                       (const TensorDesc *)descs, nl);
    return hipGetLastError();
}

// what: KBFV_LIFT / KBFV_QUANTIZE (n descs), KBFV_TENSOR (n descs x a limbs).
hipError_t launch_bfv(int what, const DevTables &T, const void *descs, int n, int a, int b, hipStream_t st) {
    (void)b;
    if (n <= 0) return hipSuccess;
    if (!T.bfv || T.L > kMaxL) return hipErrorInvalidValue;
    switch (what) {
    case KBFV_LIFT:
        hipLaunchKernelGGL(k_bfv_lift, dim3(T.N / 256, n), dim3(256), 0, st, T, (const BfvLiftDesc *)descs);
        return hipGetLastError();
    case KBFV_QUANTIZE:
        hipLaunchKernelGGL(k_quantize, dim3(T.N / 256, n), dim3(256), 0, st, T, (const BfvQuantDesc *)descs);
        return hipGetLastError();
    case KBFV_TENSOR:
        if (T.logN == 13) return launch_tensor_r<13>(T, descs, n, a, st);
        if (T.logN == 14) return launch_tensor_r<14>(T, descs, n, a, st);
        if (T.logN == 15) return launch_tensor_r<15>(T, descs, n, a, st);
        return hipErrorInvalidValue;
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace mkhip
