"""mkf_amd: MI355X-native multi-key mul+relin (MK-CKKS / MK-BFV).

Python ctypes bindings over the C ABI in include/mkhip.h (library
mkf_amd/libmkhip.so, built by mkf_amd/build.py).  There is no CPU fallback:
if the HIP library is missing or no GPU is present, calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmkhip.so")
MAX_PARTIES = 32

_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
intp = ctypes.POINTER(ctypes.c_int)
dblp = ctypes.POINTER(ctypes.c_double)

STATUS = {
    0: "MKHIP_OK", -1: "MKHIP_ERR_ARG", -2: "MKHIP_ERR_PARAMS", -3: "MKHIP_ERR_GADGET",
    -4: "MKHIP_ERR_UNKNOWN_PARTY", -5: "MKHIP_ERR_LEVEL", -6: "MKHIP_ERR_NO_KEY", -7: "MKHIP_ERR_OOM",
    -8: "MKHIP_ERR_DEVICE", -9: "MKHIP_ERR_SCALE", -10: "MKHIP_ERR_SCHEME", -11: "MKHIP_ERR_TOO_MANY",
    -12: "MKHIP_ERR_SCRATCH",
}
NOALLOC_SCRATCH_BYTES = 256 << 20

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = [
    "mkhip_status_string", "mkhip_ctx_create", "mkhip_ctx_destroy", "mkhip_ctx_info", "mkhip_set_crs",
    "mkhip_register_party", "mkhip_unregister_party", "mkhip_kklss_ckks_mulrelin", "mkhip_dev_alloc",
    "mkhip_dev_free", "mkhip_h2d", "mkhip_d2h", "mkhip_sync", "mkhip_plan_kklss_ckks", "mkhip_plan_launch",
    "mkhip_plan_time", "mkhip_plan_destroy", "mkhip_plan_workspace_bytes", "mkhip_rescale_count",
    "mkhip_ntt_limbs", "mkhip_decompose", "mkhip_moddown", "mkhip_external_product",
    "mkhip_plan_stage_times", "mkhip_plan_stage_count", "mkhip_plan_stage_name",
    "mkhip_register_mkf_group", "mkhip_unregister_mkf_group", "mkhip_mkf_ckks_mulrelin", "mkhip_plan_mkf_ckks",
    "mkhip_register_bfv_party", "mkhip_kklss_bfv_mulrelin", "mkhip_plan_kklss_bfv", "mkhip_bfv_modup_q_to_r",
    "mkhip_bfv_rescale_q_to_r", "mkhip_bfv_quantize", "mkhip_bfv_decompose", "mkhip_mkf_bfv_mulrelin",
    "mkhip_plan_mkf_bfv", "mkhip_set_tuning", "mkhip_get_tuning", "mkhip_autotune", "mkhip_d2d",
    "mkhip_batch_create", "mkhip_batch_add", "mkhip_batch_finalize", "mkhip_batch_submit", "mkhip_batch_wait",
    "mkhip_batch_time", "mkhip_batch_workspace_bytes", "mkhip_batch_branch_info", "mkhip_batch_destroy",
    "mkhip_server_create", "mkhip_server_add_class", "mkhip_server_start", "mkhip_server_enqueue",
    "mkhip_server_poll", "mkhip_server_wait", "mkhip_server_stats", "mkhip_server_destroy", "mkhip_host_alloc",
    "mkhip_host_free", "mkhip_add_rotation_key", "mkhip_del_rotation_key", "mkhip_add_conjugation_key",
    "mkhip_del_conjugation_key", "mkhip_kklss_rotate", "mkhip_kklss_conjugate", "mkhip_plan_kklss_rotate",
    "mkhip_register_mkf_rotation", "mkhip_mkf_ckks_rotate", "mkhip_plan_mkf_rotate",
    "mkhip_gen_secret", "mkhip_gen_public_key", "mkhip_gen_relin_key", "mkhip_register_party_keygen",
    "mkhip_gen_bfv_relin_key", "mkhip_register_bfv_party_keygen", "mkhip_gen_mkf_switching_key",
    "mkhip_gen_mkf_relin_key", "mkhip_register_mkf_group_keygen",
    "mkhip_gen_rotation_key", "mkhip_gen_conjugation_key", "mkhip_encrypt",
    "mkhip_plan_mkf_poly", "mkhip_mkf_poly_query", "mkhip_mkf_ckks_evaluate_poly", "mkhip_mkf_ckks_add",
    "mkhip_mkf_ckks_add_const", "mkhip_mkf_ckks_mult_by_const", "mkhip_mkf_ckks_mult_by_gaussian_integer_and_add",
    "mkhip_mkf_ckks_rescale", "mkhip_mkf_ckks_mulrelin_levels", "mkhip_probe_peak",
    "mkhip_kklss_ckks_mulrelin_noalloc",
]
BASIS_STANDARD, BASIS_CHEBYSHEV = 0, 1
TUNE_FUSED_DECOMP, TUNE_XCD_REMAP, TUNE_DECOMP_ORDER, TUNE_FP64, TUNE_SPLIT_LIMB, TUNE_SPLIT_EXT, TUNE_DIGIT_PRE, \
    TUNE_XCD_EXT, TUNE_EXT_CS, TUNE_INV_LAZY = 1, 2, 3, 4, 5, 6, 7, 8, 9, 10
PROBE_COPY, PROBE_MUL32, PROBE_MAD64, PROBE_BFLY_LAZY, PROBE_BFLY_RED, PROBE_MRED, PROBE_COPY_STREAM, PROBE_ACC3, \
    PROBE_IBFLY_LAZY = 0, 1, 2, 3, 4, 5, 6, 7, 8
SCHEME_IDS = {"kklss": 0, "bfv": 1, "mkf": 2, "mkf_bfv": 3, "rotate": 4, "mkf_rotate": 5}
MAX_ROTATIONS = 16
CONJUGATE = -2

# kernel run by each named plan stage
STAGE_KERNELS = {
    "decompose": "k_decomp", "mac_xy": "k_mac_xy", "tensor_ntt": "k_ntt_batch<fwd>", "tensor": "k_tensor",
    "ext_xy": "k_ext<1>", "moddown_xy": "k_moddown<false>", "decompose_t": "k_decomp", "ext_vu": "k_ext<2>",
    "moddown_rescale": "k_moddown<true>", "intt_in": "k_ntt_batch<inv>", "mac_fold": "k_mac2",
    "moddown_fold": "k_moddown_ntt", "karatsuba": "k_karatsuba", "intt_mu2": "k_ntt_batch<inv>",
    "decompose_mu2": "k_decomp", "mac_relin": "k_mac2", "moddown_relin": "k_moddown_ntt",
    "intt_last": "k_ntt_batch<inv>", "rescale": "k_rescale_ntt",
    "lift": "k_bfv_lift", "quantize": "k_quantize", "moddown_final": "k_moddown<true>",
    "decompose_relin": "k_decomp", "materialize": "k_rescale_ntt", "decompose_mac": "k_decomp_mac",
    "conj_permute": "k_automorph", "permute": "k_automorph", "ext": "k_ext<2>", "moddown": "k_moddown<false>",
    "linear": "k_linear", "rescale_intt": "k_ntt_batch<inv>", "rescale_div": "k_rescale_ntt",
    "digit_pre": "k_digit_pre", "digit_pre_mu2": "k_digit_pre",
}
# the BFV plans' tensor stage runs the R-basis kernel; MKF-BFV's ModDowns are coefficient-domain
BFV_STAGE_KERNELS = dict(STAGE_KERNELS, tensor="k_tensor_r")
MKF_BFV_STAGE_KERNELS = dict(BFV_STAGE_KERNELS, moddown_fold="k_moddown<false>", moddown_relin="k_moddown<false>")


def stage_kernel(name, table=STAGE_KERNELS):
    """Kernel a plan stage runs; at N = 2^15 a stage that inverse-transforms limbs is
    followed by '<stage>_finish' (last stage + N^-1 of the split limbs)."""
    return "k_intt_finish" if name.endswith("_finish") else table[name]


class MkhipError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().mkhip_status_string(status).decode()
        super().__init__("%s: %s (%s)" % (what, STATUS.get(status, status), msg))


class ServerConfig(ctypes.Structure):
    _fields_ = [("streams", ctypes.c_int), ("max_batch", ctypes.c_int), ("mem_bytes", ctypes.c_size_t)]


class Request(ctypes.Structure):
    _fields_ = [
        ("cls", ctypes.c_int), ("in0", ctypes.c_void_p), ("in1", ctypes.c_void_p), ("out", ctypes.c_void_p),
        ("ids0", ctypes.c_int * MAX_PARTIES), ("ids1", ctypes.c_int * MAX_PARTIES),
        ("k0", ctypes.c_int), ("k1", ctypes.c_int), ("group", ctypes.c_int), ("nres", ctypes.c_int),
        ("on_device", ctypes.c_int),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("max_batch", ctypes.c_uint64),
        ("mean_batch", ctypes.c_double), ("max_inflight", ctypes.c_int), ("pool_bytes", ctypes.c_size_t),
        ("pool_free_largest", ctypes.c_size_t),
    ]


class KKLSSOp(ctypes.Structure):
    _fields_ = [
        ("op0", ctypes.c_void_p), ("op1", ctypes.c_void_p), ("out", ctypes.c_void_p),
        ("ids0", ctypes.c_int * MAX_PARTIES), ("ids1", ctypes.c_int * MAX_PARTIES),
        ("k0", ctypes.c_int), ("k1", ctypes.c_int), ("level", ctypes.c_int), ("nres", ctypes.c_int),
        ("level0", ctypes.c_int), ("level1", ctypes.c_int),
    ]


class MKFOp(ctypes.Structure):
    _fields_ = [
        ("left", ctypes.c_void_p), ("right", ctypes.c_void_p), ("out", ctypes.c_void_p),
        ("group", ctypes.c_int), ("level", ctypes.c_int), ("nres", ctypes.c_int),
    ]


class RotateOp(ctypes.Structure):
    _fields_ = [
        ("in_", ctypes.c_void_p), ("out", ctypes.c_void_p), ("ids", ctypes.c_int * MAX_PARTIES),
        ("k", ctypes.c_int), ("level", ctypes.c_int), ("nrot", ctypes.c_int), ("rot", ctypes.c_int * 16),
    ]


class MKFRotateOp(ctypes.Structure):
    _fields_ = [
        ("in_", ctypes.c_void_p), ("out", ctypes.c_void_p),
        ("group", ctypes.c_int), ("level", ctypes.c_int), ("rotation", ctypes.c_int),
    ]


def lib():
    """Load libmkhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
                               % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


class PolyStruct(ctypes.Structure):
    _fields_ = [("coeffs_re", ctypes.POINTER(ctypes.c_double)), ("coeffs_im", ctypes.POINTER(ctypes.c_double)),
                ("ncoeffs", ctypes.c_int), ("max_deg", ctypes.c_int), ("lead", ctypes.c_int),
                ("basis", ctypes.c_int)]


class Polynomial:
    """ckks.Polynomial as the evaluator reads it: coefficients (complex), MaxDeg,
    Lead, Basis (BASIS_STANDARD or BASIS_CHEBYSHEV); a, b the Chebyshev interval the
    caller maps its input from."""

    def __init__(self, coeffs, max_deg=None, lead=True, basis=BASIS_STANDARD, a=-1.0, b=1.0):
        self.coeffs = [complex(c) for c in coeffs]
        self.max_deg = len(self.coeffs) - 1 if max_deg is None else int(max_deg)
        self.lead, self.basis, self.a, self.b = bool(lead), basis, a, b

    def struct(self):
        """(PolyStruct, arrays it points to)."""
        re = np.ascontiguousarray([c.real for c in self.coeffs], np.float64)
        im = np.ascontiguousarray([c.imag for c in self.coeffs], np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        st = PolyStruct(re.ctypes.data_as(dp), im.ctypes.data_as(dp), len(self.coeffs), self.max_deg,
                        int(self.lead), int(self.basis))
        return st, (re, im)


def _declare(L):
    I, V, S = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t
    U, D = ctypes.c_uint64, ctypes.c_double
    L.mkhip_status_string.restype = ctypes.c_char_p
    L.mkhip_status_string.argtypes = [I]
    L.mkhip_ctx_create.argtypes = [ctypes.POINTER(V), I, u64p, I, u64p, I, I, U, u64p, I]
    L.mkhip_ctx_destroy.argtypes = [V]
    L.mkhip_ctx_destroy.restype = None
    L.mkhip_ctx_info.argtypes = [V, intp, intp, ctypes.POINTER(S)]
    L.mkhip_set_crs.argtypes = [V, I, u64p]
    L.mkhip_register_party.argtypes = [V, I, u64p, u64p, u64p]
    L.mkhip_unregister_party.argtypes = [V, I]
    L.mkhip_kklss_ckks_mulrelin.argtypes = [V, u64p, I, D, intp, I, u64p, I, D, intp, I, D, u64p, intp,
                                            intp, intp, dblp]
    L.mkhip_dev_alloc.argtypes = [V, S, ctypes.POINTER(V)]
    L.mkhip_dev_free.argtypes = [V, V]
    L.mkhip_h2d.argtypes = [V, V, V, S]
    L.mkhip_d2h.argtypes = [V, V, V, S]
    L.mkhip_d2d.argtypes = [V, V, V, S]
    L.mkhip_probe_peak.argtypes = [V, I, S, I, dblp]
    L.mkhip_kklss_ckks_mulrelin_noalloc.argtypes = [V, I, ctypes.POINTER(KKLSSOp)]
    L.mkhip_sync.argtypes = [V]
    L.mkhip_plan_kklss_ckks.argtypes = [V, I, ctypes.POINTER(KKLSSOp), I, ctypes.POINTER(V)]
    L.mkhip_plan_launch.argtypes = [V]
    L.mkhip_plan_time.argtypes = [V, I, dblp]
    L.mkhip_plan_destroy.argtypes = [V]
    L.mkhip_plan_destroy.restype = None
    L.mkhip_plan_workspace_bytes.argtypes = [V]
    L.mkhip_plan_workspace_bytes.restype = S
    L.mkhip_rescale_count.argtypes = [V, I, D, D, intp, dblp]
    L.mkhip_ntt_limbs.argtypes = [V, u64p, u64p, intp, I, I]
    L.mkhip_decompose.argtypes = [V, u64p, I, u64p]
    L.mkhip_moddown.argtypes = [V, u64p, I, u64p]
    L.mkhip_external_product.argtypes = [V, u64p, u64p, I, u64p]
    L.mkhip_plan_stage_times.argtypes = [V, I, dblp]
    L.mkhip_plan_stage_count.argtypes = [V]
    L.mkhip_plan_stage_name.argtypes = [V, I]
    L.mkhip_plan_stage_name.restype = ctypes.c_char_p
    L.mkhip_register_mkf_group.argtypes = [V, I, I, ctypes.POINTER(u64p), u64p]
    L.mkhip_unregister_mkf_group.argtypes = [V, I]
    L.mkhip_mkf_ckks_mulrelin.argtypes = [V, I, u64p, u64p, I, D, D, D, u64p, intp, dblp]
    L.mkhip_plan_mkf_ckks.argtypes = [V, I, ctypes.POINTER(MKFOp), I, ctypes.POINTER(V)]
    L.mkhip_add_rotation_key.argtypes = [V, I, I, u64p]
    L.mkhip_del_rotation_key.argtypes = [V, I, I]
    L.mkhip_add_conjugation_key.argtypes = [V, I, u64p]
    L.mkhip_del_conjugation_key.argtypes = [V, I]
    L.mkhip_kklss_rotate.argtypes = [V, u64p, I, intp, I, I, I, u64p]
    L.mkhip_kklss_conjugate.argtypes = [V, u64p, I, intp, I, u64p]
    L.mkhip_plan_kklss_rotate.argtypes = [V, I, ctypes.POINTER(RotateOp), I, ctypes.POINTER(V)]
    L.mkhip_register_mkf_rotation.argtypes = [V, I, I, ctypes.POINTER(u64p)]
    L.mkhip_mkf_ckks_rotate.argtypes = [V, I, u64p, I, I, u64p]
    L.mkhip_plan_mkf_rotate.argtypes = [V, I, ctypes.POINTER(MKFRotateOp), I, ctypes.POINTER(V)]
    i64p = ctypes.POINTER(ctypes.c_int64)
    L.mkhip_gen_secret.argtypes = [V, i64p, u64p]
    L.mkhip_gen_public_key.argtypes = [V, u64p, i64p, u64p, u64p, u64p]
    L.mkhip_gen_relin_key.argtypes = [V, u64p, u64p, i64p, u64p, u64p, u64p]
    L.mkhip_register_party_keygen.argtypes = [V, I, u64p, u64p, i64p]
    L.mkhip_gen_bfv_relin_key.argtypes = [V, u64p, u64p, i64p, u64p, u64p, u64p]
    L.mkhip_register_bfv_party_keygen.argtypes = [V, I, u64p, u64p, i64p]
    L.mkhip_gen_mkf_switching_key.argtypes = [V, u64p, u64p, u64p, i64p, u64p]
    L.mkhip_gen_mkf_relin_key.argtypes = [V, u64p, u64p, i64p, u64p]
    L.mkhip_register_mkf_group_keygen.argtypes = [V, I, I, u64p, u64p, i64p]
    L.mkhip_gen_rotation_key.argtypes = [V, I, u64p, i64p, u64p]
    L.mkhip_gen_conjugation_key.argtypes = [V, u64p, i64p, u64p]
    L.mkhip_encrypt.argtypes = [V, u64p, u64p, I, i64p, i64p, i64p, u64p, I, I, u64p, u64p]
    PP = ctypes.POINTER(PolyStruct)
    L.mkhip_plan_mkf_poly.argtypes = [V, I, I, ctypes.POINTER(u64p), I, D, PP, D, ctypes.POINTER(u64p), I,
                                      ctypes.POINTER(V), intp, ctypes.POINTER(D)]
    L.mkhip_mkf_poly_query.argtypes = [V, I, I, D, PP, D, intp, ctypes.POINTER(D), intp]
    L.mkhip_mkf_ckks_evaluate_poly.argtypes = [V, I, u64p, I, D, PP, D, u64p, intp, ctypes.POINTER(D)]
    L.mkhip_mkf_ckks_add.argtypes = [V, I, u64p, I, D, u64p, I, D, I, u64p, intp, ctypes.POINTER(D)]
    L.mkhip_mkf_ckks_add_const.argtypes = [V, I, u64p, I, D, D, D, u64p]
    L.mkhip_mkf_ckks_mult_by_const.argtypes = [V, I, u64p, I, D, D, D, u64p, ctypes.POINTER(D)]
    L.mkhip_mkf_ckks_mult_by_gaussian_integer_and_add.argtypes = [V, I, u64p, I, ctypes.c_int64, ctypes.c_int64,
                                                                  u64p, I]
    L.mkhip_mkf_ckks_rescale.argtypes = [V, I, u64p, I, D, D, u64p, intp, ctypes.POINTER(D)]
    L.mkhip_mkf_ckks_mulrelin_levels.argtypes = [V, I, u64p, I, u64p, I, u64p]
    L.mkhip_register_bfv_party.argtypes = [V, I, u64p, u64p, u64p, u64p, u64p]
    L.mkhip_kklss_bfv_mulrelin.argtypes = [V, u64p, intp, I, u64p, intp, I, u64p, intp, intp]
    L.mkhip_plan_kklss_bfv.argtypes = [V, I, ctypes.POINTER(KKLSSOp), I, ctypes.POINTER(V)]
    for f in ("mkhip_bfv_modup_q_to_r", "mkhip_bfv_rescale_q_to_r", "mkhip_bfv_quantize", "mkhip_bfv_decompose"):
        getattr(L, f).argtypes = [V, u64p, u64p]
    L.mkhip_mkf_bfv_mulrelin.argtypes = [V, I, u64p, u64p, u64p]
    L.mkhip_set_tuning.argtypes = [V, I, I]
    L.mkhip_get_tuning.argtypes = [V, I, intp]
    L.mkhip_autotune.argtypes = [V, I, I, V, I, intp, dblp]
    L.mkhip_plan_mkf_bfv.argtypes = [V, I, ctypes.POINTER(MKFOp), I, ctypes.POINTER(V)]
    L.mkhip_batch_create.argtypes = [ctypes.POINTER(V)]
    L.mkhip_batch_add.argtypes = [V, V, I, I, V, S, intp]
    for f in ("mkhip_batch_finalize", "mkhip_batch_submit", "mkhip_batch_wait"):
        getattr(L, f).argtypes = [V]
    L.mkhip_batch_time.argtypes = [V, I, dblp]
    L.mkhip_batch_workspace_bytes.argtypes = [V]
    L.mkhip_batch_workspace_bytes.restype = S
    L.mkhip_batch_branch_info.argtypes = [V, I, intp, ctypes.POINTER(S)]
    L.mkhip_batch_destroy.argtypes = [V]
    L.mkhip_batch_destroy.restype = None
    L.mkhip_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(V)]
    L.mkhip_server_add_class.argtypes = [V, V, I, intp]
    L.mkhip_server_start.argtypes = [V]
    L.mkhip_server_enqueue.argtypes = [V, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]
    L.mkhip_server_poll.argtypes = [V, ctypes.c_uint64, intp, intp, dblp, dblp]
    L.mkhip_server_wait.argtypes = [V, ctypes.c_uint64, intp, dblp, dblp]
    L.mkhip_server_stats.argtypes = [V, ctypes.POINTER(ServerStats)]
    L.mkhip_server_destroy.argtypes = [V]
    L.mkhip_server_destroy.restype = None
    L.mkhip_host_alloc.argtypes = [S, ctypes.POINTER(V)]
    L.mkhip_host_free.argtypes = [V]


def _p(a):
    if a.dtype != np.uint64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous uint64 array")
    return a.ctypes.data_as(u64p)


def _u64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


def check(status, what):
    if status != 0:
        raise MkhipError(status, what)


class Context:
    """Parameter set + device tables + resident keys (mkhip_ctx)."""

    def __init__(self, logN, Q, P, gamma=2, t=0, qmul=None, flags=0):
        """qmul (L primes) and t make a KKLSS MK-BFV context over R = Q u QMul."""
        self.logN, self.N = logN, 1 << logN
        self.Q, self.P = [int(q) for q in Q], [int(p) for p in P]
        self.L, self.K, self.gamma = len(self.Q), len(self.P), gamma
        self.QMul = [int(q) for q in qmul] if qmul is not None else None
        self.t = int(t)
        h = ctypes.c_void_p()
        q = _u64(self.Q)
        p = _u64(self.P or [0])
        qm = _u64(self.QMul) if self.QMul is not None else None
        check(lib().mkhip_ctx_create(ctypes.byref(h), logN, _p(q), self.L, _p(p), self.K, gamma, self.t,
                                     _p(qm) if qm is not None else None, flags), "mkhip_ctx_create")
        self._h = h
        a, b, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
        check(lib().mkhip_ctx_info(h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(w)), "mkhip_ctx_info")
        self.alpha, self.beta, self.swk_words = a.value, b.value, w.value

    def close(self):
        if getattr(self, "_h", None):
            lib().mkhip_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def beta_at(self, level):
        return (level + 1 + self.alpha - 1) // self.alpha

    # ---- keys ----
    def set_crs_u(self, u):
        u = _u64(u)
        assert u.size == self.swk_words
        check(lib().mkhip_set_crs(self._h, -1, _p(u)), "mkhip_set_crs")

    def register_party(self, pid, b, d, v):
        b, d, v = _u64(b), _u64(d), _u64(v)
        for x in (b, d, v):
            assert x.size == self.swk_words, (x.size, self.swk_words)
        check(lib().mkhip_register_party(self._h, pid, _p(b), _p(d), _p(v)), "mkhip_register_party")

    def unregister_party(self, pid):
        check(lib().mkhip_unregister_party(self._h, pid), "mkhip_unregister_party")

    def set_crs(self, index, a):
        """CRS of a rotation (index = rotidx in [1, N/2)) or of conjugation (CONJUGATE)."""
        a = _u64(a)
        assert a.size == self.swk_words
        check(lib().mkhip_set_crs(self._h, index, _p(a)), "mkhip_set_crs")

    def add_rotation_key(self, pid, rotidx, key):
        key = _u64(key)
        assert key.size == self.swk_words
        check(lib().mkhip_add_rotation_key(self._h, pid, rotidx, _p(key)), "mkhip_add_rotation_key")

    def del_rotation_key(self, pid, rotidx):
        check(lib().mkhip_del_rotation_key(self._h, pid, rotidx), "mkhip_del_rotation_key")

    def add_conjugation_key(self, pid, key):
        key = _u64(key)
        assert key.size == self.swk_words
        check(lib().mkhip_add_conjugation_key(self._h, pid, _p(key)), "mkhip_add_conjugation_key")

    def del_conjugation_key(self, pid):
        check(lib().mkhip_del_conjugation_key(self._h, pid), "mkhip_del_conjugation_key")

    def kklss_rotate(self, ct, ids, rotidx, hoisted=False):
        """KKLSS Evaluator.RotateNew / RotateHoistedNew; ct (1+k, level+1, N) coefficient domain."""
        ct = _u64(ct)
        k = len(ids)
        assert ct.shape[0] == 1 + k
        out = np.zeros_like(ct)
        i = (ctypes.c_int * max(k, 1))(*ids)
        check(lib().mkhip_kklss_rotate(self._h, _p(ct), ct.shape[1] - 1, i, k, rotidx, int(hoisted), _p(out)),
              "mkhip_kklss_rotate")
        return out

    def kklss_conjugate(self, ct, ids):
        ct = _u64(ct)
        k = len(ids)
        assert ct.shape[0] == 1 + k
        out = np.zeros_like(ct)
        i = (ctypes.c_int * max(k, 1))(*ids)
        check(lib().mkhip_kklss_conjugate(self._h, _p(ct), ct.shape[1] - 1, i, k, _p(out)), "mkhip_kklss_conjugate")
        return out

    # ---- evaluator ----
    def rescale_count(self, level, scale, min_scale):
        n, s = ctypes.c_int(), ctypes.c_double()
        check(lib().mkhip_rescale_count(self._h, level, scale, min_scale, ctypes.byref(n), ctypes.byref(s)),
              "mkhip_rescale_count")
        return n.value, s.value

    def kklss_ckks_mulrelin(self, op0, ids0, op1, ids1, scale0=1.0, scale1=1.0, min_scale=1.0):
        """KKLSS MK-CKKS MulRelinNew with rescale.  op: (1+k, level+1, N) uint64.

        Returns (out, ids_out, level_out, scale_out)."""
        op0, op1 = _u64(op0), _u64(op1)
        k0, k1 = len(ids0), len(ids1)
        assert op0.shape[0] == 1 + k0 and op1.shape[0] == 1 + k1
        lvl0, lvl1 = op0.shape[1] - 1, op1.shape[1] - 1
        ku = len(set(ids0) | set(ids1))
        nl = min(lvl0, lvl1) + 1
        out = np.zeros((1 + ku) * nl * self.N, np.uint64)
        i0 = (ctypes.c_int * max(k0, 1))(*ids0)
        i1 = (ctypes.c_int * max(k1, 1))(*ids1)
        iout = (ctypes.c_int * max(ku, 1))()
        kuo, lo, so = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib().mkhip_kklss_ckks_mulrelin(self._h, _p(op0), lvl0, scale0, i0, k0, _p(op1), lvl1, scale1, i1,
                                              k1, min_scale, _p(out), iout, ctypes.byref(kuo), ctypes.byref(lo),
                                              ctypes.byref(so)), "mkhip_kklss_ckks_mulrelin")
        nlo = lo.value + 1
        out = out[:(1 + kuo.value) * nlo * self.N].reshape(1 + kuo.value, nlo, self.N)
        return out, list(iout)[:kuo.value], lo.value, so.value

    # ---- KKLSS MK-BFV ----
    def register_bfv_party(self, pid, bkey, dkey, vkey):
        """bkey / dkey: both gadget branches stacked, (2 beta, L+K, N); vkey (beta, L+K, N)."""
        b, d, v = _u64(bkey), _u64(dkey), _u64(vkey)
        w = self.swk_words
        assert b.size == 2 * w and d.size == 2 * w and v.size == w
        b, d = b.reshape(2, w), d.reshape(2, w)
        b0, b1, d0, d1 = (np.ascontiguousarray(x) for x in (b[0], b[1], d[0], d[1]))
        check(lib().mkhip_register_bfv_party(self._h, pid, _p(b0), _p(d0), _p(v), _p(b1), _p(d1)),
              "mkhip_register_bfv_party")

    def kklss_bfv_mulrelin(self, op0, ids0, op1, ids1):
        """mkbfv MulRelinNew: op (1+k, L, N) coefficient domain -> (out, ids_out)."""
        op0, op1 = _u64(op0), _u64(op1)
        k0, k1 = len(ids0), len(ids1)
        assert op0.shape == (1 + k0, self.L, self.N) and op1.shape == (1 + k1, self.L, self.N)
        ku = len(set(ids0) | set(ids1))
        out = np.zeros((1 + ku, self.L, self.N), np.uint64)
        i0 = (ctypes.c_int * max(k0, 1))(*ids0)
        i1 = (ctypes.c_int * max(k1, 1))(*ids1)
        iout = (ctypes.c_int * max(ku, 1))()
        kuo = ctypes.c_int()
        check(lib().mkhip_kklss_bfv_mulrelin(self._h, _p(op0), i0, k0, _p(op1), i1, k1, _p(out), iout,
                                             ctypes.byref(kuo)), "mkhip_kklss_bfv_mulrelin")
        return out, list(iout)[:kuo.value]

    def bfv_modup_q_to_r(self, a):
        out = np.zeros((2 * self.L, self.N), np.uint64)
        check(lib().mkhip_bfv_modup_q_to_r(self._h, _p(_u64(a)), _p(out)), "mkhip_bfv_modup_q_to_r")
        return out

    def bfv_rescale_q_to_r(self, a):
        out = np.zeros((2 * self.L, self.N), np.uint64)
        check(lib().mkhip_bfv_rescale_q_to_r(self._h, _p(_u64(a)), _p(out)), "mkhip_bfv_rescale_q_to_r")
        return out

    def bfv_quantize(self, a_ntt):
        out = np.zeros((self.L, self.N), np.uint64)
        check(lib().mkhip_bfv_quantize(self._h, _p(_u64(a_ntt)), _p(out)), "mkhip_bfv_quantize")
        return out

    def bfv_decompose(self, aR):
        out = np.zeros((2 * self.L, self.L + self.K, self.N), np.uint64)
        check(lib().mkhip_bfv_decompose(self._h, _p(_u64(aR)), _p(out)), "mkhip_bfv_decompose")
        return out

    # ---- MKF ----
    def register_mkf_group(self, group, fold_keys, relin_key):
        """fold_keys: k switching keys [beta][2][L+K][N]; relin_key likewise."""
        keys = [_u64(f) for f in fold_keys]
        relin = _u64(relin_key)
        for x in keys + [relin]:
            assert x.size == 2 * self.swk_words, (x.size, 2 * self.swk_words)
        arr = (u64p * len(keys))(*[_p(f) for f in keys])
        check(lib().mkhip_register_mkf_group(self._h, group, len(keys), arr, _p(relin)),
              "mkhip_register_mkf_group")

    # ---- key generation / encryption (small polynomials are int64 [N]) ----
    def _i64(self, x):
        a = np.ascontiguousarray(np.asarray(x, dtype=np.int64))
        return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))

    def gen_secret(self, s_small):
        a, p = self._i64(s_small)
        out = np.zeros((self.L + self.K, self.N), np.uint64)
        check(lib().mkhip_gen_secret(self._h, p, _p(out)), "mkhip_gen_secret")
        return out

    def gen_public_key(self, sk, e_small, a0):
        a, p = self._i64(e_small)
        pk0 = np.zeros((self.L + self.K, self.N), np.uint64)
        pk1 = np.zeros_like(pk0)
        check(lib().mkhip_gen_public_key(self._h, _p(_u64(sk)), p, _p(_u64(a0)), _p(pk0), _p(pk1)),
              "mkhip_gen_public_key")
        return pk0, pk1

    def gen_relin_key(self, sk, r, errs):
        a, p = self._i64(errs)
        shp = (self.beta, self.L + self.K, self.N)
        b, d, v = (np.zeros(shp, np.uint64) for _ in range(3))
        check(lib().mkhip_gen_relin_key(self._h, _p(_u64(sk)), _p(_u64(r)), p, _p(b), _p(d), _p(v)),
              "mkhip_gen_relin_key")
        return b, d, v

    def register_party_keygen(self, pid, sk, r, errs):
        a, p = self._i64(errs)
        check(lib().mkhip_register_party_keygen(self._h, pid, _p(_u64(sk)), _p(_u64(r)), p),
              "mkhip_register_party_keygen")

    def gen_bfv_relin_key(self, sk, r, errs):
        """KKLSS MK-BFV GenRelinearizationKey (mkbfv/keygen.go:24-162): errs (5, beta, N)
        of b_1, b_2, d_1, d_2, v; CRS[0], CRS[-3] and u uploaded.  Returns b, d (both
        branches, (2 beta, L+K, N)) and v ((beta, L+K, N))."""
        a, p = self._i64(errs)
        shp = (self.beta, self.L + self.K, self.N)
        b, d = np.zeros((2,) + shp, np.uint64), np.zeros((2,) + shp, np.uint64)
        v = np.zeros(shp, np.uint64)
        check(lib().mkhip_gen_bfv_relin_key(self._h, _p(_u64(sk)), _p(_u64(r)), p, _p(b), _p(d), _p(v)),
              "mkhip_gen_bfv_relin_key")
        return b.reshape((2 * self.beta,) + shp[1:]), d.reshape((2 * self.beta,) + shp[1:]), v

    def register_bfv_party_keygen(self, pid, sk, r, errs):
        a, p = self._i64(errs)
        check(lib().mkhip_register_bfv_party_keygen(self._h, pid, _p(_u64(sk)), _p(_u64(r)), p),
              "mkhip_register_bfv_party_keygen")

    def gen_mkf_switching_key(self, s_in, s_out, a_uniform, errs):
        """Lattigo GenSwitchingKey(s_in, s_out) as MKF uses it (mkf_cpu/common.go:44-51):
        a_uniform (beta, L+K, N), errs (beta, N); returns (beta, 2, L+K, N)."""
        a, p = self._i64(errs)
        out = np.zeros((self.beta, 2, self.L + self.K, self.N), np.uint64)
        check(lib().mkhip_gen_mkf_switching_key(self._h, _p(_u64(s_in)), _p(_u64(s_out)), _p(_u64(a_uniform)), p,
                                                _p(out)), "mkhip_gen_mkf_switching_key")
        return out

    def gen_mkf_relin_key(self, s, a_uniform, errs):
        a, p = self._i64(errs)
        out = np.zeros((self.beta, 2, self.L + self.K, self.N), np.uint64)
        check(lib().mkhip_gen_mkf_relin_key(self._h, _p(_u64(s)), _p(_u64(a_uniform)), p, _p(out)),
              "mkhip_gen_mkf_relin_key")
        return out

    def register_mkf_group_keygen(self, group, sks, a_uniform, errs):
        """Fold keys s_i -> sum s_i and the relinearization key, generated and registered
        in HBM: sks (k, L+K, N), a_uniform (k+1, beta, L+K, N), errs (k+1, beta, N)."""
        sks = _u64(sks)
        a, p = self._i64(errs)
        check(lib().mkhip_register_mkf_group_keygen(self._h, group, sks.shape[0], _p(sks), _p(_u64(a_uniform)), p),
              "mkhip_register_mkf_group_keygen")

    def gen_rotation_key(self, rotidx, sk, errs):
        a, p = self._i64(errs)
        out = np.zeros((self.beta, self.L + self.K, self.N), np.uint64)
        check(lib().mkhip_gen_rotation_key(self._h, rotidx, _p(_u64(sk)), p, _p(out)), "mkhip_gen_rotation_key")
        return out

    def gen_conjugation_key(self, sk, errs):
        a, p = self._i64(errs)
        out = np.zeros((self.beta, self.L + self.K, self.N), np.uint64)
        check(lib().mkhip_gen_conjugation_key(self._h, _p(_u64(sk)), p, _p(out)), "mkhip_gen_conjugation_key")
        return out

    def encrypt(self, pk0, pk1, v, e0, e1, pt, level=None, pt_ntt=False, ntt_out=False):
        level = self.L - 1 if level is None else level
        va, vp = self._i64(v)
        ea, ep = self._i64(e0)
        fa, fp = self._i64(e1)
        pt = np.ascontiguousarray(_u64(pt)[:level + 1])
        c0 = np.zeros((level + 1, self.N), np.uint64)
        c1 = np.zeros_like(c0)
        check(lib().mkhip_encrypt(self._h, _p(_u64(pk0)), _p(_u64(pk1)), level, vp, ep, fp, _p(pt), int(pt_ntt),
                                  int(ntt_out), _p(c0), _p(c1)), "mkhip_encrypt")
        return c0, c1

    def register_mkf_rotation(self, group, rotation, keys):
        keys = [_u64(x) for x in keys]
        arr = (u64p * len(keys))(*[_p(x) for x in keys])
        check(lib().mkhip_register_mkf_rotation(self._h, group, rotation, arr), "mkhip_register_mkf_rotation")

    def mkf_ckks_rotate(self, group, ct, rotation):
        """MKF RotateNew; ct (1+k, level+1, N) NTT domain -> (1+k, level+1, N)."""
        ct = _u64(ct)
        out = np.zeros_like(ct)
        check(lib().mkhip_mkf_ckks_rotate(self._h, group, _p(ct), ct.shape[1] - 1, rotation, _p(out)),
              "mkhip_mkf_ckks_rotate")
        return out

    def unregister_mkf_group(self, group):
        check(lib().mkhip_unregister_mkf_group(self._h, group), "mkhip_unregister_mkf_group")

    def mkf_ckks_mulrelin(self, group, left, right, scale_l=1.0, scale_r=1.0, min_scale=1.0):
        """MKF CKKSCase.Run: left/right (1+k, level+1, N) NTT domain -> (out, level_out, scale_out)."""
        left, right = _u64(left), _u64(right)
        k = left.shape[0] - 1
        level = left.shape[1] - 1
        out = np.zeros((1 + k) * (level + 1) * self.N, np.uint64)
        lo, so = ctypes.c_int(), ctypes.c_double()
        check(lib().mkhip_mkf_ckks_mulrelin(self._h, group, _p(left), _p(right), level, scale_l, scale_r,
                                            min_scale, _p(out), ctypes.byref(lo), ctypes.byref(so)),
              "mkhip_mkf_ckks_mulrelin")
        nlo = lo.value + 1
        return out[:(1 + k) * nlo * self.N].reshape(1 + k, nlo, self.N), lo.value, so.value

    def mkf_bfv_mulrelin(self, group, left, right):
        """MKF BFVCase.Run: left/right (1+k, L, N) coefficient domain -> out (1+k, L, N)."""
        left, right = _u64(left), _u64(right)
        out = np.zeros_like(left)
        check(lib().mkhip_mkf_bfv_mulrelin(self._h, group, _p(left), _p(right), _p(out)),
              "mkhip_mkf_bfv_mulrelin")
        return out

    # ---- MKF polynomial evaluator and linear operations (NTT domain, (1+k, level+1, N)) ----
    def mkf_poly_query(self, group, level, scale, poly, target_scale):
        st, keep = poly.struct()
        lo, so, nm = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        check(lib().mkhip_mkf_poly_query(self._h, group, level, scale, ctypes.byref(st), target_scale,
                                         ctypes.byref(lo), ctypes.byref(so), ctypes.byref(nm)),
              "mkhip_mkf_poly_query")
        return lo.value, so.value, nm.value

    def mkf_evaluate_poly(self, group, ct, scale, poly, target_scale):
        """EvaluatePoly on one ciphertext -> (out, level_out, scale_out)."""
        ct = _u64(ct)
        lo, so, _ = self.mkf_poly_query(group, ct.shape[1] - 1, scale, poly, target_scale)
        out = np.zeros((ct.shape[0], lo + 1, self.N), np.uint64)
        st, keep = poly.struct()
        l2, s2 = ctypes.c_int(), ctypes.c_double()
        check(lib().mkhip_mkf_ckks_evaluate_poly(self._h, group, _p(ct), ct.shape[1] - 1, scale, ctypes.byref(st),
                                                 target_scale, _p(out), ctypes.byref(l2), ctypes.byref(s2)),
              "mkhip_mkf_ckks_evaluate_poly")
        return out, l2.value, s2.value

    def mkf_add(self, a, scale_a, b, scale_b, sub=False):
        a, b = _u64(a), _u64(b)
        k, lv = a.shape[0] - 1, min(a.shape[1], b.shape[1]) - 1
        out = np.zeros((1 + k, lv + 1, self.N), np.uint64)
        lo, so = ctypes.c_int(), ctypes.c_double()
        check(lib().mkhip_mkf_ckks_add(self._h, k, _p(a), a.shape[1] - 1, scale_a, _p(b), b.shape[1] - 1, scale_b,
                                       int(sub), _p(out), ctypes.byref(lo), ctypes.byref(so)), "mkhip_mkf_ckks_add")
        return out, lo.value, so.value

    def mkf_add_const(self, a, scale, c):
        a = _u64(a)
        c = complex(c)
        out = np.zeros_like(a)
        check(lib().mkhip_mkf_ckks_add_const(self._h, a.shape[0] - 1, _p(a), a.shape[1] - 1, scale, c.real, c.imag,
                                             _p(out)), "mkhip_mkf_ckks_add_const")
        return out

    def mkf_mult_by_const(self, a, scale, c):
        a = _u64(a)
        c = complex(c)
        out = np.zeros_like(a)
        so = ctypes.c_double()
        check(lib().mkhip_mkf_ckks_mult_by_const(self._h, a.shape[0] - 1, _p(a), a.shape[1] - 1, scale, c.real,
                                                 c.imag, _p(out), ctypes.byref(so)), "mkhip_mkf_ckks_mult_by_const")
        return out, so.value

    def mkf_gauss_add(self, a, re, im, acc):
        """acc + a (re + im i) for int64 re, im (MultByGaussianIntegerAndAdd)."""
        a = _u64(a)
        acc = _u64(acc).copy()
        check(lib().mkhip_mkf_ckks_mult_by_gaussian_integer_and_add(self._h, a.shape[0] - 1, _p(a), a.shape[1] - 1,
                                                                    int(re), int(im), _p(acc), acc.shape[1] - 1),
              "mkhip_mkf_ckks_mult_by_gaussian_integer_and_add")
        return acc

    def mkf_rescale(self, a, scale, min_scale):
        a = _u64(a)
        out = np.zeros(a.size, np.uint64)
        lo, so = ctypes.c_int(), ctypes.c_double()
        check(lib().mkhip_mkf_ckks_rescale(self._h, a.shape[0] - 1, _p(a), a.shape[1] - 1, scale, min_scale,
                                           _p(out), ctypes.byref(lo), ctypes.byref(so)), "mkhip_mkf_ckks_rescale")
        n = a.shape[0] * (lo.value + 1) * self.N  # [1+k][level_out+1][N], packed
        return out[:n].reshape(a.shape[0], lo.value + 1, self.N), lo.value, so.value

    def mkf_mulrelin_levels(self, group, left, right):
        """MulRelinNew of the paper evaluator: operands at their own levels, no rescale."""
        left, right = _u64(left), _u64(right)
        lv = min(left.shape[1], right.shape[1]) - 1
        out = np.zeros((left.shape[0], lv + 1, self.N), np.uint64)
        check(lib().mkhip_mkf_ckks_mulrelin_levels(self._h, group, _p(left), left.shape[1] - 1, _p(right),
                                                   right.shape[1] - 1, _p(out)), "mkhip_mkf_ckks_mulrelin_levels")
        return out

    def kklss_ckks_mulrelin_noalloc(self, ops):
        """KKLSS MK-CKKS mul+relin+rescale of device-resident ops (dicts as for Plan)
        with no device allocation: workspace and descriptors in the library's fixed
        scratch (mkhip_kklss_ckks_mulrelin_noalloc).  Returns when the outputs are
        written."""
        arr = _marshal_ops("kklss", ops)
        check(lib().mkhip_kklss_ckks_mulrelin_noalloc(self._h, len(ops), arr), "mkhip_kklss_ckks_mulrelin_noalloc")

    def probe_peak(self, kind, size=0, iters=5):
        """Measured peak rate of this GPU (mkhip_probe_peak): bytes/s for PROBE_COPY
        (two buffers of `size` bytes), operations/s for the register probes."""
        r = ctypes.c_double()
        check(lib().mkhip_probe_peak(self._h, kind, int(size), int(iters), ctypes.byref(r)), "mkhip_probe_peak")
        return r.value

    # ---- tuning ----
    def set_tuning(self, knob, value):
        check(lib().mkhip_set_tuning(self._h, knob, int(value)), "mkhip_set_tuning")

    def get_tuning(self, knob):
        v = ctypes.c_int()
        check(lib().mkhip_get_tuning(self._h, knob, ctypes.byref(v)), "mkhip_get_tuning")
        return v.value

    def autotune(self, ops, scheme="kklss", iters=3):
        """Time every applicable knob combination on this batch and keep the fastest;
        returns the candidate times (ms per launch)."""
        arr = _marshal_ops(scheme, ops)
        n = ctypes.c_int()
        ms = (ctypes.c_double * 16)()
        check(lib().mkhip_autotune(self._h, SCHEME_IDS[scheme], len(ops), ctypes.cast(arr, ctypes.c_void_p),
                                   iters, ctypes.byref(n), ms), "mkhip_autotune")
        return list(ms)[:n.value]

    # ---- primitives ----
    def ntt(self, rows, ring_idx, inverse=False):
        rows = _u64(rows)
        ri = (ctypes.c_int * len(ring_idx))(*ring_idx)
        out = np.zeros_like(rows)
        check(lib().mkhip_ntt_limbs(self._h, _p(rows), _p(out), ri, len(ring_idx), int(inverse)),
              "mkhip_ntt_limbs")
        return out

    def decompose(self, a, level=None):
        a = _u64(a)
        level = a.shape[0] - 1 if level is None else level
        out = np.zeros((self.beta_at(level), level + 1 + self.K, self.N), np.uint64)
        check(lib().mkhip_decompose(self._h, _p(a), level, _p(out)), "mkhip_decompose")
        return out

    def moddown(self, x, level=None):
        x = _u64(x)
        level = x.shape[0] - 1 - self.K if level is None else level
        out = np.zeros((level + 1, self.N), np.uint64)
        check(lib().mkhip_moddown(self._h, _p(x), level, _p(out)), "mkhip_moddown")
        return out

    def external_product(self, a_hoisted, key, level=None):
        a_hoisted, key = _u64(a_hoisted), _u64(key)
        level = a_hoisted.shape[1] - 1 - self.K if level is None else level
        out = np.zeros((level + 1, self.N), np.uint64)
        check(lib().mkhip_external_product(self._h, _p(a_hoisted), _p(key), level, _p(out)),
              "mkhip_external_product")
        return out

    # ---- device memory ----
    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().mkhip_dev_alloc(self._h, nbytes, ctypes.byref(p)), "mkhip_dev_alloc")
        return p.value

    def free(self, ptr):
        check(lib().mkhip_dev_free(self._h, ptr), "mkhip_dev_free")

    def h2d(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        check(lib().mkhip_h2d(self._h, dptr, arr.ctypes.data, arr.nbytes), "mkhip_h2d")

    def d2h(self, arr, dptr):
        check(lib().mkhip_d2h(self._h, arr.ctypes.data, dptr, arr.nbytes), "mkhip_d2h")
        return arr

    def d2d(self, dst, src, nbytes):
        check(lib().mkhip_d2d(self._h, dst, src, nbytes), "mkhip_d2d")

    def sync(self):
        check(lib().mkhip_sync(self._h), "mkhip_sync")


def _marshal_ops(scheme, ops):
    """dict ops -> the C array the plan builders take (mkhip_kklss_op[] or mkhip_mkf_op[])."""
    if scheme in ("kklss", "bfv"):
        arr = (KKLSSOp * len(ops))()
        for i, o in enumerate(ops):
            a = arr[i]
            a.op0, a.op1, a.out = o["op0"], o["op1"], o["out"]
            for j, v in enumerate(o["ids0"]):
                a.ids0[j] = v
            for j, v in enumerate(o["ids1"]):
                a.ids1[j] = v
            a.k0, a.k1, a.level, a.nres = len(o["ids0"]), len(o["ids1"]), o["level"], o["nres"]
            a.level0, a.level1 = o.get("level0", 0), o.get("level1", 0)
        return arr
    if scheme == "rotate":
        arr = (RotateOp * len(ops))()
        for i, o in enumerate(ops):
            a = arr[i]
            a.in_, a.out = o["in"], o["out"]
            for j, v in enumerate(o["ids"]):
                a.ids[j] = v
            a.k, a.level, a.nrot = len(o["ids"]), o["level"], len(o["rot"])
            for j, v in enumerate(o["rot"]):
                a.rot[j] = v
        return arr
    if scheme == "mkf_rotate":
        arr = (MKFRotateOp * len(ops))()
        for i, o in enumerate(ops):
            a = arr[i]
            a.in_, a.out, a.group, a.level, a.rotation = o["in"], o["out"], o["group"], o["level"], o["rotation"]
        return arr
    if scheme in ("mkf", "mkf_bfv"):
        arr = (MKFOp * len(ops))()
        for i, o in enumerate(ops):
            a = arr[i]
            a.left, a.right, a.out = o["left"], o["right"], o["out"]
            a.group, a.level, a.nres = o["group"], o["level"], o["nres"]
        return arr
    raise ValueError(scheme)


class Plan:
    """A batched mul+relin over device-resident operands (one hipGraph per launch).

    scheme "kklss": ops are dicts {op0, op1, out, ids0, ids1, level, nres};
    scheme "bfv":   the same dicts, level = L-1 and nres = 0 (KKLSS MK-BFV);
    scheme "mkf":   ops are dicts {left, right, out, group, level, nres};
    scheme "mkf_bfv": the same dicts, level = L-1 and nres = 0 (MKF MK-BFV);
    scheme "rotate": Galois key switches, dicts {in, out, ids, level, rot} with rot a
                     list of normalised rotidx (CRS present) or CONJUGATE;
    scheme "mkf_rotate": dicts {in, out, group, level, rotation}."""

    def __init__(self, ctx, ops, use_graph=True, scheme="kklss"):
        h = ctypes.c_void_p()
        fn = {"kklss": "mkhip_plan_kklss_ckks", "bfv": "mkhip_plan_kklss_bfv", "mkf": "mkhip_plan_mkf_ckks",
              "mkf_bfv": "mkhip_plan_mkf_bfv", "rotate": "mkhip_plan_kklss_rotate",
              "mkf_rotate": "mkhip_plan_mkf_rotate"}[scheme]
        arr = _marshal_ops(scheme, ops)
        check(getattr(lib(), fn)(ctx.handle, len(ops), arr, int(use_graph), ctypes.byref(h)), fn)
        self._h = h
        self.ctx = ctx
        self.scheme = scheme

    @property
    def stages(self):
        n = lib().mkhip_plan_stage_count(self._h)
        return [lib().mkhip_plan_stage_name(self._h, i).decode() for i in range(n)]

    def launch(self):
        check(lib().mkhip_plan_launch(self._h), "mkhip_plan_launch")

    def time(self, iters):
        ms = ctypes.c_double()
        check(lib().mkhip_plan_time(self._h, iters, ctypes.byref(ms)), "mkhip_plan_time")
        return ms.value

    def stage_times(self, iters):
        ms = (ctypes.c_double * lib().mkhip_plan_stage_count(self._h))()
        check(lib().mkhip_plan_stage_times(self._h, iters, ms), "mkhip_plan_stage_times")
        return list(ms)

    @property
    def workspace_bytes(self):
        return lib().mkhip_plan_workspace_bytes(self._h)

    def close(self):
        if getattr(self, "_h", None):
            lib().mkhip_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PolyPlan(Plan):
    """EvaluatePoly over nct device-resident ciphertexts as one plan: ins / outs are
    device addresses of [1+k][level+1][N] / [1+k][level_out+1][N] words."""

    def __init__(self, ctx, group, ins, level, scale, poly, target_scale, outs, use_graph=True):
        h = ctypes.c_void_p()
        st, keep = poly.struct()
        n = len(ins)
        ia = (u64p * n)(*[ctypes.cast(ctypes.c_void_p(int(x)), u64p) for x in ins])
        oa = (u64p * n)(*[ctypes.cast(ctypes.c_void_p(int(x)), u64p) for x in outs])
        lo, so = ctypes.c_int(), ctypes.c_double()
        check(lib().mkhip_plan_mkf_poly(ctx.handle, group, n, ia, level, scale, ctypes.byref(st), target_scale, oa,
                                        int(use_graph), ctypes.byref(h), ctypes.byref(lo), ctypes.byref(so)),
              "mkhip_plan_mkf_poly")
        self._h = h
        self.ctx = ctx
        self.scheme = "mkf_poly"
        self.level_out, self.scale_out = lo.value, so.value


class Batch:
    """An uneven batch run by one hipGraph launch (mkhip_batch).

    add(ctx, ops, scheme, ws_budget) appends the ops of one context as a branch
    (any mix of party counts; ops as for Plan); the branch's ops are cut into
    chunks whose workspace fits ws_budget bytes and run in sequence in one
    arena, the branches run concurrently.  finalize() captures the graph."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().mkhip_batch_create(ctypes.byref(h)), "mkhip_batch_create")
        self._h = h
        self._keep = []  # contexts must outlive the batch

    def add(self, ctx, ops, scheme="kklss", ws_budget=0):
        arr = _marshal_ops(scheme, ops)
        n = ctypes.c_int()
        check(lib().mkhip_batch_add(self._h, ctx.handle, SCHEME_IDS[scheme], len(ops),
                                    ctypes.cast(arr, ctypes.c_void_p), int(ws_budget), ctypes.byref(n)),
              "mkhip_batch_add")
        self._keep.append(ctx)
        return n.value

    def finalize(self):
        check(lib().mkhip_batch_finalize(self._h), "mkhip_batch_finalize")

    def submit(self):
        check(lib().mkhip_batch_submit(self._h), "mkhip_batch_submit")

    def wait(self):
        check(lib().mkhip_batch_wait(self._h), "mkhip_batch_wait")

    def time(self, iters):
        ms = ctypes.c_double()
        check(lib().mkhip_batch_time(self._h, iters, ctypes.byref(ms)), "mkhip_batch_time")
        return ms.value

    @property
    def workspace_bytes(self):
        return lib().mkhip_batch_workspace_bytes(self._h)

    def branch_info(self, i):
        n, b = ctypes.c_int(), ctypes.c_size_t()
        check(lib().mkhip_batch_branch_info(self._h, i, ctypes.byref(n), ctypes.byref(b)), "mkhip_batch_branch_info")
        return n.value, b.value

    def close(self):
        if getattr(self, "_h", None):
            lib().mkhip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """A uint64 numpy array over pinned host memory (mkhip_host_alloc)."""

    def __init__(self, shape):
        n = int(np.prod(shape))
        p = ctypes.c_void_p()
        check(lib().mkhip_host_alloc(max(n, 1) * 8, ctypes.byref(p)), "mkhip_host_alloc")
        self._p = p
        self.a = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint64)), shape=(max(n, 1),))[:n]
        self.a = self.a.reshape(shape)

    @property
    def ptr(self):
        return self._p.value

    def close(self):
        if getattr(self, "_p", None):
            self.a = None
            lib().mkhip_host_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (mkhip_server).

    add_class(ctx, scheme) registers a context (keys resident) as a request class;
    start() launches the scheduler thread; enqueue() takes host operands (numpy
    uint64 arrays that must stay alive until the request completed) and returns a
    ticket; poll()/wait() report (status, latency_ms, queue_ms)."""

    def __init__(self, streams=4, max_batch=64, mem_bytes=8 << 30):
        cfg = ServerConfig(streams, max_batch, int(mem_bytes))
        h = ctypes.c_void_p()
        check(lib().mkhip_server_create(ctypes.byref(cfg), ctypes.byref(h)), "mkhip_server_create")
        self._h = h
        self._keep = []
        self._schemes = []

    def add_class(self, ctx, scheme="kklss"):
        c = ctypes.c_int()
        check(lib().mkhip_server_add_class(self._h, ctx.handle, SCHEME_IDS[scheme], ctypes.byref(c)),
              "mkhip_server_add_class")
        self._keep.append(ctx)
        self._schemes.append(scheme)
        return c.value

    def start(self):
        check(lib().mkhip_server_start(self._h), "mkhip_server_start")

    def enqueue(self, cls, in0, in1, out, ids0=(), ids1=(), group=0, nres=0):
        for a in (in0, in1, out):
            if a.dtype != np.uint64 or not a.flags["C_CONTIGUOUS"]:
                raise TypeError("expected C-contiguous uint64 arrays")
        return self._enqueue(cls, in0.ctypes.data, in1.ctypes.data, out.ctypes.data, ids0, ids1, group, nres, 0)

    def enqueue_device(self, cls, in0, in1, out, ids0=(), ids1=(), group=0, nres=0):
        """A request whose operands and result are device buffers (ctx.alloc
        pointers): served in place, with no host copies."""
        return self._enqueue(cls, int(in0), int(in1), int(out), ids0, ids1, group, nres, 1)

    def _enqueue(self, cls, in0, in1, out, ids0, ids1, group, nres, on_device):
        r = Request()
        r.cls, r.in0, r.in1, r.out, r.on_device = cls, in0, in1, out, on_device
        for j, v in enumerate(ids0):
            r.ids0[j] = v
        for j, v in enumerate(ids1):
            r.ids1[j] = v
        r.k0, r.k1, r.group, r.nres = len(ids0), len(ids1), group, nres
        t = ctypes.c_uint64()
        check(lib().mkhip_server_enqueue(self._h, ctypes.byref(r), ctypes.byref(t)), "mkhip_server_enqueue")
        return t.value

    def poll(self, ticket):
        d, st, lat, qd = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(lib().mkhip_server_poll(self._h, ticket, ctypes.byref(d), ctypes.byref(st), ctypes.byref(lat),
                                      ctypes.byref(qd)), "mkhip_server_poll")
        return (st.value, lat.value, qd.value) if d.value else None

    def wait(self, ticket):
        st, lat, qd = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        check(lib().mkhip_server_wait(self._h, ticket, ctypes.byref(st), ctypes.byref(lat), ctypes.byref(qd)),
              "mkhip_server_wait")
        return st.value, lat.value, qd.value

    def stats(self):
        s = ServerStats()
        check(lib().mkhip_server_stats(self._h, ctypes.byref(s)), "mkhip_server_stats")
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            lib().mkhip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
