// Elementwise CKKS linear operations between the products of the MKF polynomial
// evaluator (software/mkf/mkf_cpu/true_mkf_ckks_polynomial.go over Lattigo's
// ckks.Evaluator: Add, Sub, AddConst, MultByConst, MultByGaussianIntegerAndAdd;
// the algorithms are in software/kklss/mkckks/evaluator.go:118-330).  One thread
// per coefficient of one row; every operand a canonical residue.  HBM-bound: one
// to three words per coefficient.
#include "kcommon.hpp"

namespace mkhip {

// grid (N / 256, descs in this launch)
__global__ void __launch_bounds__(256) k_linear(DevTables T, const LinDesc *__restrict__ descs) {
    const LinDesc &Dd = descs[blockIdx.y];
    const int c = blockIdx.x * 256 + threadIdx.x;
    const ModConst m = T.mod[Dd.mod];
    const u64 k = c < (T.N >> 1) ? Dd.c_lo : Dd.c_hi;
    u64 r = 0;
    switch (Dd.op) {
    case LIN_COPY:
        r = G(Dd.a)[c];
        break;
    case LIN_ADD:
        r = add_mod(G(Dd.a)[c], G(Dd.b)[c], m.q);
        break;
    case LIN_SUB:
        r = sub_mod(G(Dd.a)[c], G(Dd.b)[c], m.q);
        break;
    case LIN_MULC:
        r = mred(G(Dd.a)[c], k, m.q, m.qinv);
        break;
    case LIN_MULCADD:
        r = add_mod(G(Dd.b)[c], mred(G(Dd.a)[c], k, m.q, m.qinv), m.q);
        break;
    case LIN_ADDC:
        r = add_mod(G(Dd.a)[c], k, m.q);
        break;
    default:
        break;
    }
    G(Dd.out)[c] = r;
}

hipError_t launch_linear(const DevTables &T, const void *descs, int n, hipStream_t st) {
    const LinDesc *d = (const LinDesc *)descs;
    for (int off = 0; off < n; off += 65535) {
        const int cnt = n - off < 65535 ? n - off : 65535;
        hipLaunchKernelGGL(k_linear, dim3(T.N / 256, cnt), dim3(256), 0, st, T, d + off);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace mkhip
