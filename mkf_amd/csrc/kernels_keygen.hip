// Key generation and public-key encryption kernels (the step before the
// mul+relin path: software/kklss/mkrlwe/keygen.go, encryptor.go).
//
//   k_small_ntt   a small signed polynomial (secret, Gaussian error, ternary v) as
//                 residues of one ring modulus, forward-transformed, optionally in
//                 Montgomery form -- ExtendBasisSmallNormAndCenter + NTTLvl (+ MFormLvl)
//   k_keyop       the elementwise combinations of keygen.go / encryptor.go: MRed
//                 products, the gadget term P s on a digit's limbs, errors, sums and
//                 Lattigo's negation
#include "kcommon.hpp"

namespace mkhip {

__device__ __forceinline__ u64 small_residue(int64_t v, u64 q) {
    // |v| is small (a secret or error coefficient): one remainder each way
    const u64 r = v >= 0 ? (u64)v % q : q - ((u64)(-v) % q);
    return r == q ? 0 : r;
}

template <int LOGN>
__global__ void __launch_bounds__(Geo<LOGN>::TH) k_small_ntt(DevTables T, const SmallNttDesc *__restrict__ descs) {
    typedef Geo<LOGN> Gm;
    const int tid = threadIdx.x;
    int bid = block_index(T);
    const int h = Gm::split ? bid % 2 : 0;
    bid /= Gm::H;
    const SmallNttDesc &Dd = descs[bid];
    const ModConst m = T.mod[Dd.mod];
    const int64_t *src = Dd.src;
    auto f = [&](int c) { return small_residue(src[c], m.q); };
    u64 v[16];
    if (fp_ok(m.q) && T.fp64) {
        double d[16];
        fwd_transform_fp<LOGN>(d, f, tid, h, T.twd + (size_t)Dd.mod * Gm::N, m);
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = fp_word(d[e]);
    } else {
        fwd_transform<LOGN>(v, f, tid, h, T.tw + (size_t)Dd.mod * Gm::N, m);
    }
    if (Dd.mform) {
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = mred(v[e], m.r2, m.q, m.qinv);
    }
    store_row_quad<Gm::LS>(v, Dd.dst + (size_t)h * Gm::NS, tid);
}

// one coefficient per thread; grid (N / 256, ndesc)
__global__ void __launch_bounds__(256) k_keyop(DevTables T, const KeyOpDesc *__restrict__ descs) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    const KeyOpDesc &Dd = descs[blockIdx.y];
    const ModConst m = T.mod[Dd.mod];
    u64 s = Dd.x ? G(Dd.x)[c] : 0;
    if (Dd.g) s = add_mod(s, mulmod_m(G(Dd.g)[c], Dd.pmod, m), m.q);
    if (Dd.y) s = add_mod(s, G(Dd.y)[c], m.q);
    if (Dd.e) s = add_mod(s, small_residue(Dd.e[c], m.q), m.q);
    if (Dd.ka) {
        const u64 p = mred(G(Dd.ka)[c], G(Dd.kb)[c], m.q, m.qinv);
        s = Dd.op == 0 ? sub_mod(s, p, m.q) : add_mod(s, p, m.q);
    }
    G(Dd.out)[c] = Dd.op == 1 ? m.q - s : s;
}

template <int LOGN>
static hipError_t launch_small_ntt(const DevTables &T, const void *descs, int n, hipStream_t st) {
    typedef Geo<LOGN> Gm;
    constexpr size_t lds = (size_t)lds_words(Gm::NS) * sizeof(u64);
    static bool init = false;
    if (!init) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_small_ntt<LOGN>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        init = true;
    }
    hipLaunchKernelGGL((k_small_ntt<LOGN>), dim3(n * Gm::H), dim3(Gm::TH), lds, st, T, (const SmallNttDesc *)descs);
    return hipGetLastError();
}

hipError_t launch_keygen(int what, const DevTables &T, const void *descs, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (what == KKEYOP) {
        hipLaunchKernelGGL(k_keyop, dim3(T.N / 256, n), dim3(256), 0, st, T, (const KeyOpDesc *)descs);
        return hipGetLastError();
    }
    if (what != KSMALL_NTT) return hipErrorInvalidValue;
    switch (T.logN) {
    case 13:
        return launch_small_ntt<13>(T, descs, n, st);
    case 14:
        return launch_small_ntt<14>(T, descs, n, st);
    case 15:
        return launch_small_ntt<15>(T, descs, n, st);
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace mkhip
