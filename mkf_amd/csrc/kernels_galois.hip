// Galois automorphisms X -> X^gal of RNS polynomials (rotations and
// conjugation).  One output coefficient per thread, gather form: every limb of
// the descriptor is read at the same source index, so the index arithmetic is
// done once per thread and the stores are coalesced.
//
// Coefficient domain (Lattigo ring.Permute, inlined by the reference at
// software/kklss/mkrlwe/keyswitch.go:283-303): input coefficient i goes to
// (i gal) mod N, negated when bit logN of i gal is set, the negation written as
// q - x (so a zero becomes q).  Output o therefore reads i' = o gal^-1 mod 2N:
// source i' mod N, negated iff i' >= N (then i gal = o + N mod 2N).
//
// NTT domain (Lattigo ring.PermuteNTTIndex): slot o of the bit-reversed
// negacyclic NTT holds the evaluation at psi^(2 brev(o) + 1), and
// (sigma_gal a)(psi^e) = a(psi^(gal e)), so slot o reads the slot whose exponent
// is gal (2 brev(o) + 1) mod 2N.
#include "kcommon.hpp"

namespace mkhip {

__global__ void __launch_bounds__(256) k_automorph(DevTables T, const AutoDesc *__restrict__ descs) {
    const int N = T.N, logN = T.logN;
    const unsigned o = blockIdx.x * 256 + threadIdx.x;
    const AutoDesc &Dd = descs[blockIdx.y];
    const u64 mask2 = 2 * (u64)N - 1;
    unsigned src;
    bool neg = false;
    if (Dd.ntt) {
        const u64 e = 2 * (u64)(__brev(o) >> (32 - logN)) + 1;
        const unsigned t = (unsigned)(((Dd.gal * e) & mask2) >> 1);
        src = __brev(t) >> (32 - logN);
    } else {
        const u64 ip = ((u64)o * Dd.gal_inv) & mask2;
        src = (unsigned)(ip & (u64)(N - 1));
        neg = ip >= (u64)N;
    }
    const gu64 *in = G(Dd.src);
    gu64 *out = G(Dd.dst);
    for (int j = 0; j < Dd.nl; ++j) {
        const u64 x = in[(size_t)j * N + src];
        out[(size_t)j * N + o] = neg ? T.mod[j].q - x : x;
    }
}

// n descriptors; grid (N / 256, n)
hipError_t launch_galois(int what, const DevTables &T, const void *descs, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (what != KAUTO) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_automorph, dim3(T.N / 256, n), dim3(256), 0, st, T, (const AutoDesc *)descs);
    return hipGetLastError();
}

}  // namespace mkhip
