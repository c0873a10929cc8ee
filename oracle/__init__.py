"""CPU oracle for multi-key mul+relin -- TEST INFRASTRUCTURE ONLY.

ctypes bindings over ``oracle/libmkref.so`` (built from mkref.c by ``make -C
oracle``), a plain-C restatement of the reference's KKLSS/MKF mul+relin path
(see mkref.c for the reference file:line each function follows).  Only the
tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
use this module, and only as the checker; the shipped library never does.

Parity status of each restated function (what pins it):
  mred / mred_params        64 RTL known-answer vectors (tb_mred64_pipeline.sv)
  butterfly                 128 RTL known-answer vectors (tb_ntt_butterfly_x8.sv)
  ntt / intt                invariants: INTT(NTT(x)) = x, NTT product = negacyclic
                            schoolbook product; N^-1 table vs HLS N_INV_*.  The
                            choice of primitive root (Lattigo ring.NewRing) is
                            pinned by no in-tree file: parity unpinned for the
                            absolute NTT-domain values of keys.
  modup params              HLS tables P2Q_*, DECOMP_*, P_INV_Q, Q2M_*, M2Q_*
  modup_exact / moddown     big-integer CRT closed form (exact when the float64
                            estimate is exact) and brute force on small moduli
  decompose / ext. product  gadget identity: Ext(D(a), g*s) ~ a*s*P/P, noise bound
  kklss / mkf mul+relin     decrypt-and-compare tests (mkckks_test.go:352-358 bound)
"""
import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmkref.so")
_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
i64p = ctypes.POINTER(ctypes.c_int64)
intp = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    U = ctypes.c_uint64
    I = ctypes.c_int
    L.mkref_mred_params.restype = U
    L.mkref_mred_params.argtypes = [U]
    L.mkref_mred.restype = U
    L.mkref_mred.argtypes = [U, U, U, U]
    L.mkref_mred_lazy.restype = U
    L.mkref_mred_lazy.argtypes = [U, U, U, U]
    L.mkref_mform.restype = U
    L.mkref_mform.argtypes = [U, U]
    L.mkref_primitive_root.restype = U
    L.mkref_primitive_root.argtypes = [U, I]
    L.mkref_butterfly.argtypes = [U, U, U, U, U, u64p, u64p]
    L.mkref_gen_primes.restype = I
    L.mkref_gen_primes.argtypes = [I, I, I, I, u64p, u64p, I]
    L.mkref_ctx_new.restype = ctypes.c_void_p
    L.mkref_ctx_new.argtypes = [I, u64p, I, u64p, I, I, I]
    L.mkref_ctx_delete.argtypes = [ctypes.c_void_p]
    L.mkref_ctx_tables.argtypes = [ctypes.c_void_p, I, u64p, u64p, u64p, u64p]
    L.mkref_ctx_ntt.argtypes = [ctypes.c_void_p, I, u64p, I]
    L.mkref_modup_tables.argtypes = [u64p, I, u64p, I, u64p, u64p, u64p]
    L.mkref_modup_exact_simple.argtypes = [I, u64p, u64p, I, u64p, u64p, I]
    L.mkref_moddown.argtypes = [I, u64p, u64p, I, u64p, u64p, I, u64p]
    L.mkref_decompose.argtypes = [ctypes.c_void_p, I, u64p, u64p]
    L.mkref_decompose_digit.argtypes = [ctypes.c_void_p, I, I, u64p, u64p]
    L.mkref_external_product_hoisted.argtypes = [ctypes.c_void_p, I, u64p, u64p, u64p]
    L.mkref_kklss_mulrelin.restype = I
    L.mkref_kklss_mulrelin.argtypes = [ctypes.c_void_p, I, u64p, intp, I, u64p, intp, I,
                                       ctypes.POINTER(u64p), ctypes.POINTER(u64p),
                                       ctypes.POINTER(u64p), u64p, u64p, intp]
    L.mkref_kklss_mulrelin_lv.restype = I
    L.mkref_kklss_mulrelin_lv.argtypes = [ctypes.c_void_p, I, I, I, u64p, intp, I, u64p, intp, I,
                                          ctypes.POINTER(u64p), ctypes.POINTER(u64p),
                                          ctypes.POINTER(u64p), u64p, u64p, intp]
    L.mkref_decompose_src.argtypes = [ctypes.c_void_p, I, I, u64p, u64p]
    L.mkref_rescale_count.restype = I
    L.mkref_rescale_count.argtypes = [u64p, I, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(ctypes.c_double)]
    L.mkref_div_round_last.argtypes = [ctypes.c_void_p, I, I, u64p]
    L.mkref_div_round_last_ntt.argtypes = [ctypes.c_void_p, I, I, u64p]
    L.mkref_mkf_switch_one.argtypes = [ctypes.c_void_p, I, u64p, I, u64p, I, u64p, u64p]
    L.mkref_mkf_ckks_mulrelin.argtypes = [ctypes.c_void_p, I, I, u64p, u64p,
                                          ctypes.POINTER(u64p), u64p, u64p, u64p]
    from . import poly as _poly
    _poly.declare(L)
    V = ctypes.c_void_p
    L.mkref_bfv_new.restype = V
    L.mkref_bfv_new.argtypes = [I, u64p, I, u64p, u64p, I, U, I, I]
    L.mkref_mkf_bfv_mulrelin.argtypes = [V, I, u64p, u64p, ctypes.POINTER(u64p), u64p, u64p, u64p]
    L.mkref_bfv_delete.argtypes = [V]
    L.mkref_bfv_modup_q_to_r.argtypes = [V, u64p, u64p]
    L.mkref_bfv_rescale_q_to_r.argtypes = [V, u64p, u64p]
    L.mkref_bfv_quantize.argtypes = [V, u64p, u64p]
    L.mkref_bfv_decompose.argtypes = [V, u64p, u64p]
    L.mkref_bfv_ntt_r.argtypes = [V, I, u64p, I]
    L.mkref_kklss_bfv_mulrelin.restype = I
    L.mkref_kklss_bfv_mulrelin.argtypes = [V, u64p, intp, I, u64p, intp, I, ctypes.POINTER(u64p),
                                           ctypes.POINTER(u64p), ctypes.POINTER(u64p), u64p, u64p, intp]
    L.mkref_galois_element.restype = U
    L.mkref_galois_element.argtypes = [I, ctypes.c_longlong]
    L.mkref_galois_conjugate.restype = U
    L.mkref_galois_conjugate.argtypes = [I]
    L.mkref_galois_inverse.restype = U
    L.mkref_galois_inverse.argtypes = [I, U]
    L.mkref_permute.argtypes = [V, I, u64p, U, u64p]
    L.mkref_permute_ntt_index.argtypes = [I, U, u64p]
    L.mkref_automorphism_ntt.argtypes = [V, I, u64p, U, u64p]
    L.mkref_external_product_as_written.argtypes = [V, I, u64p, u64p, u64p]
    L.mkref_kklss_rotate_hoisted.argtypes = [V, I, I, u64p, u64p, ctypes.POINTER(u64p), u64p, U, u64p]
    L.mkref_kklss_rotate.argtypes = [V, I, I, u64p, ctypes.POINTER(u64p), u64p, U, u64p]
    L.mkref_kklss_conjugate.argtypes = [V, I, I, u64p, ctypes.POINTER(u64p), u64p, u64p]
    L.mkref_mkf_ckks_rotate.argtypes = [V, I, I, u64p, ctypes.POINTER(u64p), U, u64p, u64p]
    L.mkref_gen_secret.argtypes = [V, i64p, u64p]
    L.mkref_gen_public_key.argtypes = [V, u64p, i64p, u64p, u64p, u64p]
    L.mkref_gen_switching_key.argtypes = [V, u64p, i64p, u64p]
    L.mkref_gen_relin_key.argtypes = [V, u64p, u64p, i64p, u64p, u64p, u64p, u64p, u64p]
    L.mkref_gen_bfv_relin_key.argtypes = [V, u64p, u64p, u64p, u64p, i64p, u64p, u64p, u64p, u64p, u64p, u64p]
    L.mkref_gen_mkf_switching_key.argtypes = [V, u64p, u64p, u64p, i64p, u64p]
    L.mkref_gen_mkf_relin_key.argtypes = [V, u64p, u64p, i64p, u64p]
    L.mkref_aggregate_secret.argtypes = [V, I, u64p, u64p]
    L.mkref_gen_rotation_key.argtypes = [V, I, u64p, i64p, u64p, u64p]
    L.mkref_gen_conjugation_key.argtypes = [V, u64p, i64p, u64p, u64p]
    L.mkref_encrypt.argtypes = [V, I, u64p, u64p, i64p, i64p, i64p, u64p, I, I, u64p, u64p]
    L.mkref_prng_seed.argtypes = [ctypes.c_void_p, U]
    L.mkref_prng_next.restype = U
    L.mkref_prng_next.argtypes = [ctypes.c_void_p]
    L.mkref_sample_uniform.argtypes = [ctypes.c_void_p, u64p, I, I, u64p]
    L.mkref_sample_ternary.argtypes = [ctypes.c_void_p, I, ctypes.c_double, i64p]
    L.mkref_sample_gaussian.argtypes = [ctypes.c_void_p, I, ctypes.c_double, I, i64p]
    L.mkref_small_to_rns.argtypes = [i64p, I, u64p, I, u64p]


def _p(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(u64p)


def _u64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


def _ip(a):
    assert a.dtype == np.int64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(i64p)


# ---- scalar helpers -------------------------------------------------------

def mred_params(q):
    return int(lib().mkref_mred_params(q))


def mred(x, y, q, qinv=None):
    return int(lib().mkref_mred(x, y, q, mred_params(q) if qinv is None else qinv))


def mform(a, q):
    return int(lib().mkref_mform(a, q))


def primitive_root(q, rule=0):
    return int(lib().mkref_primitive_root(q, rule))


def butterfly(u, v, w, q, qinv):
    s = ctypes.c_uint64()
    d = ctypes.c_uint64()
    lib().mkref_butterfly(u, v, w, q, qinv, ctypes.byref(s), ctypes.byref(d))
    return s.value, d.value


def gen_primes(bits, logN, count, exclude=()):
    out = np.zeros(count, dtype=np.uint64)
    ex = _u64(list(exclude) or [0])
    rc = lib().mkref_gen_primes(bits, logN, count, 1, _p(out), _p(ex), len(exclude))
    if rc != 0:
        raise ValueError("not enough primes")
    return [int(v) for v in out]


def modup_tables(Q, P):
    Q = _u64(Q)
    P = _u64(P)
    a = np.zeros(len(Q), np.uint64)
    b = np.zeros(len(Q) * len(P), np.uint64)
    v = np.zeros((len(Q) + 1) * len(P), np.uint64)
    lib().mkref_modup_tables(_p(Q), len(Q), _p(P), len(P), _p(a), _p(b), _p(v))
    return a, b.reshape(len(P), len(Q)), v.reshape(len(P), len(Q) + 1)


def modup_exact(x, Q, P):
    x = _u64(x)
    N = x.shape[-1]
    Qa, Pa = _u64(Q), _u64(P)
    out = np.zeros((len(P), N), np.uint64)
    lib().mkref_modup_exact_simple(N, _p(x), _p(Qa), len(Q), _p(out), _p(Pa), len(P))
    return out


def moddown(xq, xp, Q, P):
    xq, xp = _u64(xq), _u64(xp)
    N = xq.shape[-1]
    Qa, Pa = _u64(Q), _u64(P)
    out = np.zeros((len(Q), N), np.uint64)
    lib().mkref_moddown(N, _p(xq), _p(Qa), len(Q), _p(xp), _p(Pa), len(P), _p(out))
    return out


class Prng:
    """xoshiro256** seeded by splitmix64 (the build's deterministic key/input source)."""

    def __init__(self, seed):
        self._state = (ctypes.c_uint64 * 4)()
        lib().mkref_prng_seed(ctypes.cast(self._state, ctypes.c_void_p), seed)

    @property
    def _h(self):
        return ctypes.cast(self._state, ctypes.c_void_p)

    def uniform(self, moduli, N):
        q = _u64(moduli)
        out = np.zeros((len(moduli), N), np.uint64)
        lib().mkref_sample_uniform(self._h, _p(q), len(moduli), N, _p(out))
        return out

    def ternary(self, N, p_zero=0.5):
        out = np.zeros(N, np.int64)
        lib().mkref_sample_ternary(self._h, N, p_zero, out.ctypes.data_as(i64p))
        return out

    def gaussian(self, N, sigma=3.2, bound=19):
        out = np.zeros(N, np.int64)
        lib().mkref_sample_gaussian(self._h, N, sigma, bound, out.ctypes.data_as(i64p))
        return out


def small_to_rns(s, moduli):
    s = np.ascontiguousarray(np.asarray(s, np.int64))
    q = _u64(moduli)
    out = np.zeros((len(moduli), s.shape[0]), np.uint64)
    lib().mkref_small_to_rns(s.ctypes.data_as(i64p), s.shape[0], _p(q), len(moduli), _p(out))
    return out


class Ctx:
    """RNS context: Q chain, P chain, gamma (2 = KKLSS, 1 = MKF / Lattigo rlwe)."""

    def __init__(self, logN, Q, P, gamma=2, root_rule=0):
        self.logN, self.N = logN, 1 << logN
        self.Q, self.P = [int(q) for q in Q], [int(p) for p in P]
        self.L, self.K, self.gamma = len(Q), len(P), gamma
        q, p = _u64(self.Q), _u64(self.P or [0])
        self._h = lib().mkref_ctx_new(logN, _p(q), self.L, _p(p), self.K, gamma, root_rule)
        if not self._h:
            raise ValueError("mkref_ctx_new failed (moduli not NTT friendly?)")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mkref_ctx_delete(self._h)
            self._h = None

    @property
    def QP(self):
        return self.Q + self.P

    @property
    def alpha(self):
        return self.K // self.gamma

    def beta(self, level=None):
        level = self.L - 1 if level is None else level
        a = self.alpha
        return (level + 1 + a - 1) // a

    def tables(self, limb):
        psi = np.zeros(self.N, np.uint64)
        psiinv = np.zeros(self.N, np.uint64)
        root = ctypes.c_uint64()
        ninv = ctypes.c_uint64()
        lib().mkref_ctx_tables(self._h, limb, _p(psi), _p(psiinv), ctypes.byref(root), ctypes.byref(ninv))
        return psi, psiinv, root.value, ninv.value

    def ntt(self, a, limb, inverse=False):
        a = _u64(a).copy()
        lib().mkref_ctx_ntt(self._h, limb, _p(a), 1 if inverse else 0)
        return a

    def ntt_poly(self, a, inverse=False, rows=None):
        """NTT every limb of a (nlimbs, N) poly; rows gives ring indices per limb."""
        a = _u64(a).copy()
        rows = list(range(a.shape[0])) if rows is None else rows
        for i, r in enumerate(rows):
            lib().mkref_ctx_ntt(self._h, r, _p(a[i]), 1 if inverse else 0)
        return a

    def decompose(self, a, level=None):
        level = self.L - 1 if level is None else level
        a = _u64(a)
        out = np.zeros((self.beta(level), level + 1 + self.K, self.N), np.uint64)
        lib().mkref_decompose(self._h, level, _p(a), _p(out))
        return out

    def decompose_digit(self, a, digit, level=None):
        level = self.L - 1 if level is None else level
        a = _u64(a)
        out = np.zeros((level + 1 + self.K, self.N), np.uint64)
        lib().mkref_decompose_digit(self._h, level, digit, _p(a), _p(out))
        return out

    def external_product(self, a_hoisted, key, level=None):
        level = self.L - 1 if level is None else level
        out = np.zeros((level + 1, self.N), np.uint64)
        lib().mkref_external_product_hoisted(self._h, level, _p(_u64(a_hoisted)), _p(_u64(key)), _p(out))
        return out

    def kklss_mulrelin(self, op0, ids0, op1, ids1, rlk_b, rlk_d, rlk_v, crs_u, level=None):
        """KKLSS MK-CKKS mul+relin before rescale; op = (1+k, level+1, N). Returns (out, ids)."""
        level = op0.shape[1] - 1 if level is None else level
        op0, op1 = _u64(op0[:, :level + 1]), _u64(op1[:, :level + 1])
        keys = [_u64(k) for k in (list(rlk_b) + list(rlk_d) + list(rlk_v))]
        n = len(rlk_b)
        arr = (u64p * n)
        b = arr(*[_p(k) for k in keys[:n]])
        d = arr(*[_p(k) for k in keys[n:2 * n]])
        v = arr(*[_p(k) for k in keys[2 * n:]])
        u = _u64(crs_u)
        k0, k1 = len(ids0), len(ids1)
        i0 = (ctypes.c_int * max(k0, 1))(*ids0)
        i1 = (ctypes.c_int * max(k1, 1))(*ids1)
        ku = len(set(ids0) | set(ids1))
        out = np.zeros((1 + ku, level + 1, self.N), np.uint64)
        iout = (ctypes.c_int * ku)()
        nc = lib().mkref_kklss_mulrelin(self._h, level, _p(op0), i0, k0, _p(op1), i1, k1,
                                        b, d, v, _p(u), _p(out), iout)
        assert nc == ku + 1
        return out, list(iout)

    def kklss_mulrelin_levels(self, op0, ids0, op1, ids1, rlk_b, rlk_d, rlk_v, crs_u):
        """KKLSS MK-CKKS mul+relin before rescale of operands at their own levels
        (op = (1+k, lv+1, N)): each decomposed at its own level, the product at the
        lower one (mkckks/evaluator.go:433-439, 577-581).  Returns (out, ids)."""
        lv0, lv1 = op0.shape[1] - 1, op1.shape[1] - 1
        level = min(lv0, lv1)
        op0, op1 = _u64(op0), _u64(op1)
        keys = [_u64(k) for k in (list(rlk_b) + list(rlk_d) + list(rlk_v))]
        n = len(rlk_b)
        arr = (u64p * n)
        b = arr(*[_p(k) for k in keys[:n]])
        d = arr(*[_p(k) for k in keys[n:2 * n]])
        v = arr(*[_p(k) for k in keys[2 * n:]])
        u = _u64(crs_u)
        k0, k1 = len(ids0), len(ids1)
        i0 = (ctypes.c_int * max(k0, 1))(*ids0)
        i1 = (ctypes.c_int * max(k1, 1))(*ids1)
        ku = len(set(ids0) | set(ids1))
        out = np.zeros((1 + ku, level + 1, self.N), np.uint64)
        iout = (ctypes.c_int * ku)()
        nc = lib().mkref_kklss_mulrelin_lv(self._h, level, lv0, lv1, _p(op0), i0, k0, _p(op1), i1, k1,
                                           b, d, v, _p(u), _p(out), iout)
        assert nc == ku + 1
        return out, list(iout)

    def decompose_src(self, a, level):
        """Digits of a (src_level + 1 limbs) at its own level, read at `level`."""
        a = _u64(a)
        out = np.zeros((self.beta(level), level + 1 + self.K, self.N), np.uint64)
        lib().mkref_decompose_src(self._h, level, a.shape[0] - 1, _p(a), _p(out))
        return out

    def rescale_count(self, level, scale, min_scale):
        q = _u64(self.Q)
        ns = ctypes.c_double()
        nb = lib().mkref_rescale_count(_p(q), level, scale, min_scale, ctypes.byref(ns))
        return nb, ns.value

    def div_round_last(self, poly, n=1, ntt=False):
        poly = _u64(poly).copy()
        level = poly.shape[0] - 1
        fn = lib().mkref_div_round_last_ntt if ntt else lib().mkref_div_round_last
        fn(self._h, level, n, _p(poly))
        return poly[:level + 1 - n].copy()

    def mkf_switch_one(self, a, swk, in_ntt=True, ntt_output=True, level=None):
        level = a.shape[0] - 1 if level is None else level
        o0 = np.zeros((level + 1, self.N), np.uint64)
        o1 = np.zeros((level + 1, self.N), np.uint64)
        lib().mkref_mkf_switch_one(self._h, level, _p(_u64(a)), int(in_ntt), _p(_u64(swk)),
                                   int(ntt_output), _p(o0), _p(o1))
        return o0, o1

    def mkf_ckks_mulrelin(self, left, right, fold_keys, relin_key):
        k = left.shape[0] - 1
        level = left.shape[1] - 1
        keys = [_u64(f) for f in fold_keys]
        arr = (u64p * k)(*[_p(f) for f in keys])
        o0 = np.zeros((level + 1, self.N), np.uint64)
        o1 = np.zeros((level + 1, self.N), np.uint64)
        lib().mkref_mkf_ckks_mulrelin(self._h, level, k, _p(_u64(left)), _p(_u64(right)), arr,
                                      _p(_u64(relin_key)), _p(o0), _p(o1))
        return o0, o1

    def mkf_ckks_mulrelin2(self, left, right, fold_keys, relin_key):
        """MulRelinNew of the MKF polynomial evaluator: operands at their own levels
        (shapes [1+k][level+1][N]); outputs at the smaller level."""
        left, right = _u64(left), _u64(right)
        k = left.shape[0] - 1
        ll, lr = left.shape[1] - 1, right.shape[1] - 1
        level = min(ll, lr)
        keys = [_u64(f) for f in fold_keys]
        arr = (u64p * k)(*[_p(f) for f in keys])
        o0 = np.zeros((level + 1, self.N), np.uint64)
        o1 = np.zeros((level + 1, self.N), np.uint64)
        lib().mkref_mkf_ckks_mulrelin2(self._h, ll, lr, k, _p(left), _p(right), arr, _p(_u64(relin_key)),
                                       _p(o0), _p(o1))
        return o0, o1


    # ---- Galois key switches (mkref_rot.c) ----

    def galois_element(self, k):
        return galois_element(self.logN, k)

    def permute(self, a, gal):
        a = _u64(a)
        out = np.zeros_like(a)
        lib().mkref_permute(self._h, a.shape[0] - 1, _p(a), gal, _p(out))
        return out

    def automorphism_ntt(self, a, gal):
        a = _u64(a)
        out = np.zeros_like(a)
        lib().mkref_automorphism_ntt(self._h, a.shape[0] - 1, _p(a), gal, _p(out))
        return out

    def external_product_as_written(self, a, key):
        a = _u64(a)
        out = np.zeros_like(a)
        lib().mkref_external_product_as_written(self._h, a.shape[0] - 1, _p(a), _p(_u64(key)), _p(out))
        return out

    @staticmethod
    def _keyarr(keys):
        keys = [_u64(k) for k in keys]
        return keys, (u64p * max(len(keys), 1))(*[_p(k) for k in keys])

    def kklss_rotate_hoisted(self, ct, D, rks, a, gal):
        ct = _u64(ct)
        keep, arr = self._keyarr(rks)
        out = np.zeros_like(ct)
        D = _u64(D) if len(rks) else np.zeros(1, np.uint64)
        lib().mkref_kklss_rotate_hoisted(self._h, ct.shape[1] - 1, ct.shape[0] - 1, _p(ct), _p(D), arr,
                                         _p(_u64(a)), gal, _p(out))
        return out

    def kklss_rotate(self, ct, rks, a, gal):
        ct = _u64(ct)
        keep, arr = self._keyarr(rks)
        out = np.zeros_like(ct)
        lib().mkref_kklss_rotate(self._h, ct.shape[1] - 1, ct.shape[0] - 1, _p(ct), arr, _p(_u64(a)), gal, _p(out))
        return out

    def kklss_conjugate(self, ct, cks, a):
        ct = _u64(ct)
        keep, arr = self._keyarr(cks)
        out = np.zeros_like(ct)
        lib().mkref_kklss_conjugate(self._h, ct.shape[1] - 1, ct.shape[0] - 1, _p(ct), arr, _p(_u64(a)), _p(out))
        return out

    def kklss_rotate_new(self, ct, ids, rotidx, rot_keys, crs, hoisted=False):
        """mkckks/mkbfv Evaluator.RotateNew / RotateHoistedNew orchestration
        (mkckks/evaluator.go:496-529, :598-624; mkbfv/evaluator.go:164-194):
        rotidx normalised into [0, N/2); 0 copies; a rotation with a CRS entry runs
        once; otherwise (RotateNew only) the binary expansion of rotidx is applied
        as rotations by 1, 2, 4, ...  rot_keys[(party, r)], crs[r].  Raises
        KeyError where the reference panics."""
        half = self.N // 2
        r = rotidx % half
        if r == 0:
            return _u64(ct).copy()
        steps = [r] if r in crs else None
        if steps is None:
            if hoisted:
                raise KeyError("hoisted rotation needs a precomputed rotation key")
            steps, k, rr = [], 1, r
            while rr > 0:
                if rr % 2:
                    steps.append(k)
                rr //= 2
                k *= 2
        cur = _u64(ct)
        for st in steps:
            rks = [rot_keys[(pid, st)] for pid in ids]
            cur = self.kklss_rotate(cur, rks, crs[st], galois_element(self.logN, st))
        return cur

    def mkf_ckks_rotate(self, ct, keys, gal):
        ct = _u64(ct)
        keep, arr = self._keyarr(keys)
        o0 = np.zeros_like(ct[0])
        o1 = np.zeros_like(ct[0])
        lib().mkref_mkf_ckks_rotate(self._h, ct.shape[1] - 1, ct.shape[0] - 1, _p(ct), arr, gal, _p(o0), _p(o1))
        return o0, o1


class KeyGen:
    """Key generation / encryption restated from mkrlwe (keygen.go, encryptor.go) over an
    oracle Ctx; every random choice is an input (int64 small polynomials)."""

    def __init__(self, ctx):
        self.ctx = ctx

    def _i64(self, x):
        return np.ascontiguousarray(np.asarray(x, dtype=np.int64))

    def secret(self, s_small):
        c = self.ctx
        out = np.zeros((c.L + c.K, c.N), np.uint64)
        lib().mkref_gen_secret(c._h, _ip(self._i64(s_small)), _p(out))
        return out

    def public_key(self, sk, e_small, a0):
        c = self.ctx
        pk0 = np.zeros((c.L + c.K, c.N), np.uint64)
        pk1 = np.zeros_like(pk0)
        lib().mkref_gen_public_key(c._h, _p(_u64(sk)), _ip(self._i64(e_small)), _p(_u64(a0)), _p(pk0), _p(pk1))
        return pk0, pk1

    def switching_key(self, skin, errs):
        c = self.ctx
        out = np.zeros((c.beta(), c.L + c.K, c.N), np.uint64)
        lib().mkref_gen_switching_key(c._h, _p(_u64(skin)), _ip(self._i64(errs)), _p(out))
        return out

    def relin_key(self, sk, r, errs, a, u):
        c = self.ctx
        shp = (c.beta(), c.L + c.K, c.N)
        b, d, v = np.zeros(shp, np.uint64), np.zeros(shp, np.uint64), np.zeros(shp, np.uint64)
        lib().mkref_gen_relin_key(c._h, _p(_u64(sk)), _p(_u64(r)), _ip(self._i64(errs)), _p(_u64(a)), _p(_u64(u)),
                                  _p(b), _p(d), _p(v))
        return b, d, v

    def bfv_gadget(self, Q, QMul, P, t):
        """GenBFVSwitchingKey's gadget values G_br,i (mkbfv/keygen.go:91-162), with
        exact integers as the reference computes them, reduced mod every row's
        modulus: ([beta][L+K] for branch 1, the same for branch 2)."""
        c = self.ctx
        QP = [int(q) for q in c.QP]
        alpha = c.alpha
        Qp, QMp, Pp = math.prod(Q), math.prod(QMul), math.prod(P)
        QQ = Qp * QMp
        out = []
        for primes in (Q, QMul):
            tab = np.zeros((c.beta(), len(QP)), np.uint64)
            for i in range(c.beta()):
                Qi = math.prod(primes[i * alpha:(i + 1) * alpha])
                Gi = (QQ // Qi) * t * pow(QQ // Qi, -1, Qi) * Pp // QMp
                tab[i] = [Gi % q for q in QP]
            out.append(tab)
        return out

    def bfv_relin_key(self, Q, QMul, P, t, sk, r, errs, a1, a2, u):
        """KKLSS MK-BFV GenRelinearizationKey (mkbfv/keygen.go:24-162) over this Q u P
        context: errs (5, beta, N) of b_1, b_2, d_1, d_2, v.  Returns b, d (2 beta
        digits: branch 1 then 2) and v."""
        c = self.ctx
        gm1, gm2 = self.bfv_gadget(Q, QMul, P, t)
        shp = (c.beta(), c.L + c.K, c.N)
        b, d = np.zeros((2,) + shp, np.uint64), np.zeros((2,) + shp, np.uint64)
        v = np.zeros(shp, np.uint64)
        lib().mkref_gen_bfv_relin_key(c._h, _p(gm1), _p(gm2), _p(_u64(sk)), _p(_u64(r)), _ip(self._i64(errs)),
                                      _p(_u64(a1)), _p(_u64(a2)), _p(_u64(u)), _p(b), _p(d), _p(v))
        return b.reshape((2 * shp[0],) + shp[1:]), d.reshape((2 * shp[0],) + shp[1:]), v

    def mkf_switching_key(self, s_in, s_out, a, errs):
        """Lattigo GenSwitchingKey(s_in, s_out) (mkf_cpu/common.go:44-51): (beta, 2, L+K, N)."""
        c = self.ctx
        out = np.zeros((c.beta(), 2, c.L + c.K, c.N), np.uint64)
        lib().mkref_gen_mkf_switching_key(c._h, _p(_u64(s_in)), _p(_u64(s_out)), _p(_u64(a)),
                                          _ip(self._i64(errs)), _p(out))
        return out

    def mkf_relin_key(self, s, a, errs):
        c = self.ctx
        out = np.zeros((c.beta(), 2, c.L + c.K, c.N), np.uint64)
        lib().mkref_gen_mkf_relin_key(c._h, _p(_u64(s)), _p(_u64(a)), _ip(self._i64(errs)), _p(out))
        return out

    def aggregate_secret(self, sks):
        c = self.ctx
        sks = _u64(sks)
        out = np.zeros((c.L + c.K, c.N), np.uint64)
        lib().mkref_aggregate_secret(c._h, sks.shape[0], _p(sks), _p(out))
        return out

    def mkf_group_keys(self, sks, a, errs):
        """generateOnlineKeys (common.go:44-51): fold keys s_i -> sum s_i, then the
        relinearization key of the sum; a (k+1, beta, L+K, N), errs (k+1, beta, N)."""
        s = self.aggregate_secret(sks)
        fold = [self.mkf_switching_key(sks[i], s, a[i], errs[i]) for i in range(len(sks))]
        return fold, self.mkf_relin_key(s, a[len(sks)], errs[len(sks)])

    def rotation_key(self, rotidx, sk, errs, a):
        c = self.ctx
        out = np.zeros((c.beta(), c.L + c.K, c.N), np.uint64)
        lib().mkref_gen_rotation_key(c._h, rotidx, _p(_u64(sk)), _ip(self._i64(errs)), _p(_u64(a)), _p(out))
        return out

    def conjugation_key(self, sk, errs, a):
        c = self.ctx
        out = np.zeros((c.beta(), c.L + c.K, c.N), np.uint64)
        lib().mkref_gen_conjugation_key(c._h, _p(_u64(sk)), _ip(self._i64(errs)), _p(_u64(a)), _p(out))
        return out

    def encrypt(self, pk0, pk1, v, e0, e1, pt, level=None, pt_ntt=False, ntt_out=False):
        c = self.ctx
        level = c.L - 1 if level is None else level
        c0 = np.zeros((level + 1, c.N), np.uint64)
        c1 = np.zeros_like(c0)
        pt = _u64(pt)[:level + 1]
        lib().mkref_encrypt(c._h, level, _p(_u64(pk0)), _p(_u64(pk1)), _ip(self._i64(v)), _ip(self._i64(e0)),
                            _ip(self._i64(e1)), _p(np.ascontiguousarray(pt)), int(pt_ntt), int(ntt_out), _p(c0), _p(c1))
        return c0, c1


def galois_element(logN, k):
    """GaloisElementForColumnRotationBy: 5^(k mod 2N) mod 2N."""
    return int(lib().mkref_galois_element(logN, k))


def galois_conjugate(logN):
    return int(lib().mkref_galois_conjugate(logN))


def galois_inverse(logN, g):
    return int(lib().mkref_galois_inverse(logN, g))


def permute_ntt_index(logN, gal):
    out = np.zeros(1 << logN, np.uint64)
    lib().mkref_permute_ntt_index(logN, gal, _p(out))
    return out


class BfvCtx:
    """MK-BFV context: Q u P, QMul (len L), t.  gamma 2: KKLSS (alpha = K/2 must
    be 1); gamma 1: MKF (alpha = K).  ``qp`` is the matching Q u P context."""

    def __init__(self, logN, Q, QMul, P, t, gamma=2, root_rule=0):
        self.logN, self.N = logN, 1 << logN
        self.Q, self.QMul, self.P = [int(q) for q in Q], [int(q) for q in QMul], [int(p) for p in P]
        self.L, self.K, self.t, self.gamma = len(Q), len(P), int(t), gamma
        if len(QMul) != len(Q):
            raise ValueError("len(QMul) must equal len(Q)")
        if gamma == 2 and self.K // 2 != 1:
            raise ValueError("the KKLSS BFV restatement covers alpha = K/2 = 1")
        self.qp = Ctx(logN, Q, P, gamma=gamma, root_rule=root_rule)
        q, qm, p = _u64(self.Q), _u64(self.QMul), _u64(self.P)
        self._h = lib().mkref_bfv_new(logN, _p(q), self.L, _p(qm), _p(p), self.K, self.t, gamma, root_rule)
        if not self._h:
            raise ValueError("mkref_bfv_new failed")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mkref_bfv_delete(self._h)
            self._h = None

    @property
    def R(self):
        return self.Q + self.QMul

    @property
    def QP(self):
        return self.Q + self.P

    def beta(self):
        return self.L

    def modup_q_to_r(self, a):
        out = np.zeros((2 * self.L, self.N), np.uint64)
        lib().mkref_bfv_modup_q_to_r(self._h, _p(_u64(a)), _p(out))
        return out

    def rescale_q_to_r(self, a):
        out = np.zeros((2 * self.L, self.N), np.uint64)
        lib().mkref_bfv_rescale_q_to_r(self._h, _p(_u64(a)), _p(out))
        return out

    def quantize(self, a_ntt):
        out = np.zeros((self.L, self.N), np.uint64)
        lib().mkref_bfv_quantize(self._h, _p(_u64(a_ntt)), _p(out))
        return out

    def decompose(self, aR):
        out = np.zeros((2 * self.L, self.L + self.K, self.N), np.uint64)
        lib().mkref_bfv_decompose(self._h, _p(_u64(aR)), _p(out))
        return out

    def ntt_r(self, a, inverse=False):
        a = _u64(a).copy()
        for l in range(a.shape[0]):
            lib().mkref_bfv_ntt_r(self._h, l, _p(a[l]), 1 if inverse else 0)
        return a

    def mulrelin(self, op0, ids0, op1, ids1, bkeys, dkeys, vkeys, crs_u):
        """op = (1+k, L, N) coefficient domain; keys indexed by party id:
        bkeys/dkeys (2L, L+K, N), vkeys (L, L+K, N).  Returns (out, ids)."""
        n = len(bkeys)
        kb = [_u64(k) for k in bkeys]
        kd = [_u64(k) for k in dkeys]
        kv = [_u64(k) for k in vkeys]
        arr = (u64p * n)
        b = arr(*[_p(k) for k in kb])
        d = arr(*[_p(k) for k in kd])
        v = arr(*[_p(k) for k in kv])
        k0, k1 = len(ids0), len(ids1)
        i0 = (ctypes.c_int * max(k0, 1))(*ids0)
        i1 = (ctypes.c_int * max(k1, 1))(*ids1)
        ku = len(set(ids0) | set(ids1))
        out = np.zeros((1 + ku, self.L, self.N), np.uint64)
        iout = (ctypes.c_int * ku)()
        nc = lib().mkref_kklss_bfv_mulrelin(self._h, _p(_u64(op0)), i0, k0, _p(_u64(op1)), i1, k1,
                                            b, d, v, _p(_u64(crs_u)), _p(out), iout)
        assert nc == ku + 1
        return out, list(iout)

    def mkf_mulrelin(self, left, right, fold_keys, relin_key):
        """MKF BFVCase.Run: left/right (1+k, L, N) coefficient domain -> (c0, Gamma)."""
        k = left.shape[0] - 1
        keys = [_u64(f) for f in fold_keys]
        arr = (u64p * k)(*[_p(f) for f in keys])
        o0 = np.zeros((self.L, self.N), np.uint64)
        o1 = np.zeros((self.L, self.N), np.uint64)
        lib().mkref_mkf_bfv_mulrelin(self._h, k, _p(_u64(left)), _p(_u64(right)), arr, _p(_u64(relin_key)),
                                     _p(o0), _p(o1))
        return o0, o1


from .poly import CHEBYSHEV, STANDARD, Ct, MKFError, MKFEvaluator, Polynomial  # noqa: E402,F401
