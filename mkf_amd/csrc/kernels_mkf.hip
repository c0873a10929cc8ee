// CDNA4 kernels of the MKF construction (software/mkf/mkf_cpu): aggregate-secret
// fold by key switching, Karatsuba tensor, one relinearization, NTT-domain
// rescale, k-fold materialization of Gamma.  The decomposition and the batched
// NTT come from kernels.hip.
//
//   k_mac2          two-output switching-key MAC (KeyswitchHoistedNoModDown), P rows INTT'd
//   k_moddown_ntt   ModDownQPtoQNTT of several switches, summed (fold)
//   k_karatsuba     mu0 = l0 r0, mu2 = l1 r1, mu1 = (l0+l1)(r0+r1) - mu0 - mu2
//   k_rescale_ntt   DivRoundByLastModulus in the NTT domain + copies
#include "kcommon.hpp"

namespace mkhip {

// logical block = (desc * rows + row) * H + half (XCD-remapped)
template <int LOGN>
__global__ void __launch_bounds__(Geo<LOGN>::TH)
    k_mac2(DevTables T, const Mac2Desc *__restrict__ descs, int beta, int rows) {
    typedef Geo<LOGN> Gm;
    constexpr int N = Gm::N;
    const int tid = threadIdx.x;
    int bid = block_index_ext(T);
    const int h = Gm::split ? bid % 2 : 0;
    bid /= Gm::H;
    const int r = uni(bid % rows);
    const Mac2Desc &Dd = descs[uni(bid / rows)];
    const int ri = ring_index(r, Dd.level, T.L);
    const ModConst m = T.mod[ri];
    const size_t kstride = (size_t)(T.L + T.K) * N;
    const size_t hoff = (size_t)h * Gm::NS;
    u64 a0[16];
    u64 *o0 = Dd.acc[0] + (size_t)r * N + hoff, *o1 = Dd.acc[1] + (size_t)r * N + hoff;
    // swk digit b: output 0 at (2b) kstride, output 1 at (2b+1) kstride
    const u64 *dr = Dd.D + (size_t)r * N + hoff, *k0 = Dd.swk + (size_t)ri * N + hoff, *k1 = k0 + kstride;
    if (ext_cs_geometry<Gm::LS, 2>() && T.ext_cs && ext_cs_ok(m.q, beta))
        ext_rows_cs<Gm::LS, 2>(a0, o1, dr, (size_t)rows * N, k0, k1, 2 * kstride, beta, tid, m);
    else if (ext_lazy_ok(m.q, beta))
        ext_rows<Gm::LS, 2, true>(a0, o1, dr, (size_t)rows * N, k0, k1, 2 * kstride, beta, tid, m);
    else
        ext_rows<Gm::LS, 2>(a0, o1, dr, (size_t)rows * N, k0, k1, 2 * kstride, beta, tid, m);
    if (r > Dd.level || Dd.intt_all) {
        inv_transform_sel<LOGN>(a0, opaque_lane(tid), h, T, ri, m, false);
        store_row<Gm::LS>(a0, o0, tid);
        load_row_quad<Gm::LS>(a0, o1, tid);
        inv_transform_sel<LOGN>(a0, opaque_lane(tid), h, T, ri, m, false);
        store_row<Gm::LS>(a0, o1, tid);
    } else {  // NTT-domain rows stay in the MAC's Quad layout
        store_row_quad<Gm::LS>(a0, o0, tid);
    }
}

// logical block = (desc * (level+1) + Q row j) * H + half
template <int LOGN>
__global__ void __launch_bounds__(Geo<LOGN>::TH)
    k_moddown_ntt(DevTables T, const MdNttDesc *__restrict__ descs, int nl) {
    typedef Geo<LOGN> Gm;
    constexpr int N = Gm::N;
    const int tid = threadIdx.x;
    int bid = block_index(T);
    const int h = Gm::split ? bid % 2 : 0;
    bid /= Gm::H;
    const int j = uni(bid % nl);
    const MdNttDesc &Dd = descs[uni(bid / nl)];
    const ModConst m = T.mod[j];
    const ulonglong2 *tw = T.tw + (size_t)j * N;
    const u64 pinv = T.md_pinv[j];
    const size_t hoff = (size_t)h * Gm::NS;
    // sum_a (x_a - NTT(ModUp(x_a,P))) P^-1 = (sum_a x_a - NTT(sum_a ModUp(x_a,P))) P^-1
    // residue for residue (the NTT is linear mod q_j): one transform per row.
    u64 z[16], xs[16];
    auto f = [&](int c) {
        u64 acc = 0;
        for (int a = 0; a < Dd.nacc; ++a) {
            const PUp up = modup_p_prepare(T, Dd.acc[a], nl, c);
            acc = add_mod(acc, modup_p_to_row(T, up, j, m), m.q);
        }
        return acc;
    };
    if (fwd_lazy_ok(m.q))
        fwd_transform<LOGN, 1>(z, f, tid, h, tw, m);
    else
        fwd_transform<LOGN, 0>(z, f, tid, h, tw, m);
#pragma unroll
    for (int e = 0; e < 16; ++e) xs[e] = 0;
    for (int a = 0; a < Dd.nacc; ++a) {
        u64 xa[16];
        load_row_quad<Gm::LS>(xa, Dd.acc[a] + (size_t)j * N + hoff, tid);
#pragma unroll
        for (int e = 0; e < 16; ++e) xs[e] = add_mod(xs[e], xa[e], m.q);
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) z[e] = mred(z[e] + 2 * m.q - xs[e], pinv, m.q, m.qinv);
    if (Dd.base) {
        load_row_quad<Gm::LS>(xs, Dd.base + (size_t)j * N + hoff, tid);
#pragma unroll
        for (int e = 0; e < 16; ++e) z[e] = add_mod(z[e], xs[e], m.q);
    }
    store_row_quad<Gm::LS>(z, Dd.dst + (size_t)j * N + hoff, tid);
}

// grid (N/256, ndesc * nl), one coefficient per thread
__global__ void __launch_bounds__(256) k_karatsuba(DevTables T, const KaraDesc *__restrict__ descs, int nl) {
    const int N = T.N;
    const int l = blockIdx.y % nl;
    const KaraDesc &Dd = descs[blockIdx.y / nl];
    const size_t c = (size_t)l * N + blockIdx.x * 256 + threadIdx.x;
    const ModConst m = T.mod[l];
    const u64 l0 = G(Dd.l0)[c], l1 = G(Dd.l1)[c], r0 = G(Dd.r0)[c], r1 = G(Dd.r1)[c];
    const u64 mu0 = mulmod_m(l0, r0, m);
    const u64 mu2 = mulmod_m(l1, r1, m);
    u64 mu1 = mulmod_m(add_mod(l0, l1, m.q), add_mod(r0, r1, m.q), m);
    mu1 = sub_mod(sub_mod(mu1, mu0, m.q), mu2, m.q);
    G(Dd.mu0)[c] = mu0;
    G(Dd.mu1)[c] = mu1;
    G(Dd.mu2)[c] = mu2;
}

// logical block = (desc * nout_limbs + row j) * H + half ; nout_limbs = level + 1 - nres
template <int LOGN>
__global__ void __launch_bounds__(Geo<LOGN>::TH)
    k_rescale_ntt(DevTables T, const RescaleNttDesc *__restrict__ descs, int nlo) {
    typedef Geo<LOGN> Gm;
    constexpr int N = Gm::N;
    const int tid = threadIdx.x;
    int bid = block_index(T);
    const int h = Gm::split ? bid % 2 : 0;
    bid /= Gm::H;
    const int j = uni(bid % nlo);
    const RescaleNttDesc &Dd = descs[uni(bid / nlo)];
    const ModConst m = T.mod[j];
    const size_t hoff = (size_t)h * Gm::NS;
    u64 v[16];
    if (Dd.nres == 1) {
        // out_j = (x_j - NTT_j(([x_L + h]_{q_L} - h) mod q_j)) q_L^-1, h = q_L >> 1
        const int l = Dd.level;
        const u64 qL = T.mod[l].q;
        const u64 hq = qL >> 1;
        const u64 hj = T.rs_half[l * T.L + j];
        const gu64 *xl = G(Dd.xlast);
        u64 w[16];
        auto f = [&](int c) {
            const u64 t = add_mod(xl[c], hq, qL);
            return sub_mod(reduce64(t, m.q, m.mu), hj, m.q);
        };
        if (fwd_lazy_ok(m.q))
            fwd_transform<LOGN, 1>(w, f, tid, h, T.tw + (size_t)j * N, m);
        else
            fwd_transform<LOGN, 0>(w, f, tid, h, T.tw + (size_t)j * N, m);
        // x_j is loaded after the transform: it is not held across it
        load_row_quad<Gm::LS>(v, Dd.x + (size_t)j * N + hoff, tid);
        const u64 inv = T.rs_inv[l * T.L + j];
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = mred(sub_mod(v[e], w[e], m.q), inv, m.q, m.qinv);
    } else {
        load_row_quad<Gm::LS>(v, Dd.x + (size_t)j * N + hoff, tid);
    }
    for (int o = 0; o < Dd.nout; ++o) store_row_quad<Gm::LS>(v, Dd.out[o] + (size_t)j * N + hoff, tid);
}

// ---------------------------------------------------------------------------

template <class F>
static hipError_t allow_lds_mkf(F kernel, size_t bytes) {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes);
}

template <int LOGN>
static hipError_t launch_mkf_n(int what, const DevTables &T, const void *descs, int n, int a, int b,
                               hipStream_t st) {
    typedef Geo<LOGN> Gm;
    constexpr size_t lds = (size_t)lds_words(Gm::NS) * sizeof(u64);
    constexpr int H = Gm::H;
    static bool init = false;
    if (!init) {
        hipError_t e;
        if ((e = allow_lds_mkf(k_mac2<LOGN>, lds)) != hipSuccess) return e;
        if ((e = allow_lds_mkf(k_moddown_ntt<LOGN>, lds)) != hipSuccess) return e;
        if ((e = allow_lds_mkf(k_rescale_ntt<LOGN>, lds)) != hipSuccess) return e;
        init = true;
    }
    const dim3 block(Gm::TH);
    switch (what) {
    case KMAC2:  // n descs, a = beta, b = rows
        hipLaunchKernelGGL((k_mac2<LOGN>), dim3(n * b * H), block, lds, st, T, (const Mac2Desc *)descs, a, b);
        break;
    case KMD_NTT:  // n descs, a = level + 1
        hipLaunchKernelGGL((k_moddown_ntt<LOGN>), dim3(n * a * H), block, lds, st, T, (const MdNttDesc *)descs,
                           a);
        break;
    case KRESCALE_NTT:  // n descs, a = output limbs
        hipLaunchKernelGGL((k_rescale_ntt<LOGN>), dim3(n * a * H), block, lds, st, T,
                           (const RescaleNttDesc *)descs, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_mkf(int what, const DevTables &T, const void *descs, int n, int a, int b, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (what == KKARA) {  // n descs, a = level + 1
        hipLaunchKernelGGL(k_karatsuba, dim3(T.N / 256, n * a), dim3(256), 0, st, T, (const KaraDesc *)descs, a);
        return hipGetLastError();
    }
    switch (T.logN) {
    case 13:
        return launch_mkf_n<13>(what, T, descs, n, a, b, st);
    case 14:
        return launch_mkf_n<14>(what, T, descs, n, a, b, st);
    case 15:
        return launch_mkf_n<15>(what, T, descs, n, a, b, st);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace mkhip
