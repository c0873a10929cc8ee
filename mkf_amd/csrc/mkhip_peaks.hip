// Peak-rate probes of the device the library runs on: HBM copy bandwidth and the
// in-register issue rates of the integer primitives the transforms are built on.
// They turn the kernels' measured rates into fractions of what this GPU does,
// rather than of vendor figures.  No counterpart in the reference (measurement).
//
//   MKHIP_PROBE_COPY        device-to-device copy (k_copy), bytes read + written / s
//   MKHIP_PROBE_COPY_STREAM the same with a streaming copy: four 16-byte loads in
//                           flight per lane before their stores, non-temporal
//   MKHIP_PROBE_MUL32       v_mul_lo_u32 per second (32 x 32 -> low 32)
//   MKHIP_PROBE_MAD64       v_mad_u64_u32 per second (32 x 32 + 64 -> 64)
//   MKHIP_PROBE_BFLY_LAZY   forward NTT butterflies per second, unreduced form (the
//                           transforms' form for moduli below 2^57)
//   MKHIP_PROBE_BFLY_RED    forward NTT butterflies per second, reduced form (moduli
//                           up to 2^61, e.g. the 60-bit P primes)
//   MKHIP_PROBE_MRED        Montgomery products MRed(a, b) plus modular add per second
//                           (the key-switch MACs)
//
// Every probe keeps its values in registers (no memory traffic but one store per
// thread at the end, kept live by a data-dependent condition), with 8 or 16
// independent chains per lane so that issue, not latency, bounds it.
#include "mkhip_internal.hpp"
#include "kcommon.hpp"

using namespace mkhip;

namespace {

constexpr int kProbeThreads = 256;

// n 16-byte elements, n a multiple of 4 * 256 * gridDim.x is not required: the
// tail is copied one element per lane
typedef unsigned long long pv2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(kProbeThreads) k_probe_copy(const pv2 *__restrict__ src, pv2 *__restrict__ dst,
                                                              size_t n) {
    constexpr int U = 4;
    const size_t stride = (size_t)gridDim.x * kProbeThreads * U;
    size_t i = (size_t)blockIdx.x * kProbeThreads * U + threadIdx.x;
    for (; i + (U - 1) * kProbeThreads < n; i += stride) {
        pv2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(src + i + u * kProbeThreads);
#pragma unroll
        for (int u = 0; u < U; ++u) __builtin_nontemporal_store(v[u], dst + i + u * kProbeThreads);
    }
    for (; i < n; i += kProbeThreads) dst[i] = src[i];
}

__global__ void __launch_bounds__(kProbeThreads) k_probe_mul32(u32 m, int iters, u32 *sink) {
    u32 x[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) x[j] = threadIdx.x * 8 + j + 1;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] *= m;
    }
    u32 s = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) s ^= x[j];
    if (s == 0x9e3779b9u) sink[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kProbeThreads) k_probe_mad64(u32 m, int iters, u64 *sink) {
    u64 a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = ((u64)threadIdx.x << 32) | (j + 1);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = (u64)(u32)(a[j] >> 32) * m + a[j];
    }
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) s ^= a[j];
    if (s == 0x9e3779b97f4a7c15ull) sink[blockIdx.x] = s;
}

// 16 values a lane, 8 butterflies a step with a wave-uniform twiddle, as in a
// radix-16 pass of the transforms.
template <bool LAZY>
__global__ void __launch_bounds__(kProbeThreads) k_probe_bfly(u64 w, u64 wp, u64 q, int iters, u64 *sink) {
    u64 v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = (threadIdx.x * 16 + j) % q;
    const u64 q4 = q << 2;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (LAZY)
                fwd_bfly_lazy(v[j], v[j + 8], w, wp, q4, q);
            else
                fwd_bfly(v[j], v[j + 8], w, wp, q4, q);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) s ^= v[j];
    if (s == 0x9e3779b97f4a7c15ull) sink[blockIdx.x] = s;
}

// Unreduced inverse butterflies (inv_bfly_lazy) on 16 values a lane, the values
// brought below q every twelfth step: the bound 4q 2^11 of the last step stays
// below 2^63 for the 50-bit probe modulus, as in the transforms, which reduce
// once in their 13 or 14 stages
__global__ void __launch_bounds__(kProbeThreads) k_probe_ibfly(u64 w, u64 wp, u64 q, u64 mu, int iters,
                                                               u64 *sink) {
    u64 v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = (threadIdx.x * 16 + j) % q;
    u64 bnd = q << 2;
    for (int it = 0; it < iters; ++it) {
        if (it % 12 == 0) {
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = reduce_big(v[j], q, mu);
            bnd = q << 2;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) inv_bfly_lazy(v[j], v[j + 8], w, wp, bnd, q);
        bnd <<= 1;
        __builtin_amdgcn_sched_barrier(0);
    }
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) s ^= v[j];
    if (s == 0x9e3779b97f4a7c15ull) sink[blockIdx.x] = s;
}

// Products of a key word and a digit word summed in carry-save form (acc3_mac),
// eight sums a lane, folded and reduced every eighth step as the external
// products do for beta = 8
__global__ void __launch_bounds__(kProbeThreads) k_probe_acc3(u64 b, u64 q, u64 qinv, int iters, u64 *sink) {
    u64 a[8];
    Acc3 acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        a[j] = (threadIdx.x * 8 + j + 1) % q;
        acc3_zero(acc[j]);
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 8; ++j) acc3_mac(acc[j], a[j], b + it);
        if ((it & 7) == 7) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                u64 hi, lo;
                acc3_fold(acc[j], hi, lo);
                a[j] = redc128(hi, lo, q, qinv);
                acc3_zero(acc[j]);
            }
        }
    }
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) s ^= a[j] ^ acc[j].L ^ acc[j].M ^ acc[j].H;
    if (s == 0x9e3779b97f4a7c15ull) sink[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kProbeThreads) k_probe_mred(u64 b, u64 q, u64 qinv, int iters, u64 *sink) {
    u64 acc[8], a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        a[j] = (threadIdx.x * 8 + j + 1) % q;
        acc[j] = 0;
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            acc[j] = add_mod(acc[j], mred(a[j], b, q, qinv), q);
            a[j] = acc[j];
        }
    }
    u64 s = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) s ^= acc[j];
    if (s == 0x9e3779b97f4a7c15ull) sink[blockIdx.x] = s;
}

}  // namespace

extern "C" int mkhip_probe_peak(mkhip_ctx *c, int kind, size_t size, int iters, double *rate) {
    if (!c || !rate || iters <= 0) return MKHIP_ERR_ARG;
    hipDevice_t dev;
    HIP_TRY(hipGetDevice(&dev));
    int ncu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    // enough workgroups for every SIMD of every CU several times over
    const int grid = std::max(ncu, 1) * 32;
    void *a = nullptr, *b = nullptr;
    double work = 0;  // units per launch
    const bool copy = kind == MKHIP_PROBE_COPY || kind == MKHIP_PROBE_COPY_STREAM;
    if (copy) {
        if (size < (1u << 20)) return MKHIP_ERR_ARG;
        size &= ~(size_t)15;
        HIP_TRY(hipMalloc(&a, size));
        hipError_t e = hipMalloc(&b, size);
        if (e != hipSuccess) {
            (void)hipFree(a);
            return hip_status(e);
        }
        work = 2.0 * (double)size;
    } else {
        HIP_TRY(hipMalloc(&a, (size_t)grid * 8));
    }
    const int inner = copy ? 1 : (size ? (int)std::min<size_t>(size, 1 << 20) : 4096);
    const double lanes = (double)grid * kProbeThreads;
    const u64 q50 = 0x3ffffffb80001ull, q60 = 0xffffffffffc0001ull;
    auto launch = [&](hipStream_t st) -> hipError_t {
        switch (kind) {
        case MKHIP_PROBE_COPY:
            return launch_copy(b, a, size, st);
        case MKHIP_PROBE_COPY_STREAM:
            hipLaunchKernelGGL(k_probe_copy, dim3(std::max(ncu, 1) * 8), dim3(kProbeThreads), 0, st,
                               (const pv2 *)a, (pv2 *)b, size / 16);
            break;
        case MKHIP_PROBE_MUL32:
            hipLaunchKernelGGL(k_probe_mul32, dim3(grid), dim3(kProbeThreads), 0, st, 0x9e3779b1u, inner, (u32 *)a);
            break;
        case MKHIP_PROBE_MAD64:
            hipLaunchKernelGGL(k_probe_mad64, dim3(grid), dim3(kProbeThreads), 0, st, 0x9e3779b1u, inner, (u64 *)a);
            break;
        case MKHIP_PROBE_BFLY_LAZY:
        case MKHIP_PROBE_BFLY_RED: {
            const u64 q = kind == MKHIP_PROBE_BFLY_LAZY ? q50 : q60;
            const u64 w = h_powmod(3, 12345, q);
            if (kind == MKHIP_PROBE_BFLY_LAZY)
                hipLaunchKernelGGL(k_probe_bfly<true>, dim3(grid), dim3(kProbeThreads), 0, st, w,
                                   h_shoup_companion(w, q), q, inner, (u64 *)a);
            else
                hipLaunchKernelGGL(k_probe_bfly<false>, dim3(grid), dim3(kProbeThreads), 0, st, w,
                                   h_shoup_companion(w, q), q, inner, (u64 *)a);
            break;
        }
        case MKHIP_PROBE_MRED:
            hipLaunchKernelGGL(k_probe_mred, dim3(grid), dim3(kProbeThreads), 0, st, h_mform(12345, q50), q50,
                               h_qinv(q50), inner, (u64 *)a);
            break;
        case MKHIP_PROBE_ACC3:
            hipLaunchKernelGGL(k_probe_acc3, dim3(grid), dim3(kProbeThreads), 0, st, h_mform(12345, q50), q50,
                               h_qinv(q50), inner, (u64 *)a);
            break;
        case MKHIP_PROBE_IBFLY_LAZY: {
            const u64 w = h_powmod(3, 12345, q50);
            hipLaunchKernelGGL(k_probe_ibfly, dim3(grid), dim3(kProbeThreads), 0, st, w, h_shoup_companion(w, q50),
                               q50, (u64)((((unsigned __int128)1) << 64) / q50), inner, (u64 *)a);
            break;
        }
        default:
            return hipErrorInvalidValue;
        }
        return hipGetLastError();
    };
    switch (kind) {
    case MKHIP_PROBE_COPY:
    case MKHIP_PROBE_COPY_STREAM:
        break;
    case MKHIP_PROBE_MUL32:
    case MKHIP_PROBE_MAD64:
    case MKHIP_PROBE_MRED:
    case MKHIP_PROBE_ACC3:
        work = lanes * inner * 8;
        break;
    case MKHIP_PROBE_BFLY_LAZY:
    case MKHIP_PROBE_BFLY_RED:
    case MKHIP_PROBE_IBFLY_LAZY:
        work = lanes * inner * 8;
        break;
    default:
        (void)hipFree(a);
        return MKHIP_ERR_ARG;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = MKHIP_OK;
    float ms = 0;
    hipError_t e = launch(c->stream);  // warm-up
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (int i = 0; i < iters && e == hipSuccess; ++i) e = launch(c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) st = hip_status(e);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (st != MKHIP_OK) return st;
    if (ms <= 0) return MKHIP_ERR_DEVICE;
    *rate = work * iters / (ms * 1e-3);
    return MKHIP_OK;
}
