"""MKF CKKS polynomial evaluator -- TEST INFRASTRUCTURE ONLY (part of the CPU oracle).

A restatement of paperMKFCKKSEvaluator's public operations and EvaluatePoly
(software/mkf/mkf_cpu/true_mkf_ckks.go:176-303, true_mkf_ckks_polynomial.go:18-273)
over the C oracle (mkref_poly.c for the linear operations, mkref.c for the
products and the rescale).  With ``ctx=None`` it runs on levels and scales only,
as the reference's metadata-only level test does
(p2gn16_lattigo_level_semantics_test.go:16-77).

Parity status: the linear operations are pinned by closed forms outside the NTT
domain (AddConst adds a + b X^(N/2) scaled; MultByGaussianIntegerAndAdd adds the
negacyclic product with a + b X^(N/2)); EvaluatePoly by decrypt-and-compare
against the polynomial evaluated on the slots, and its level schedule by the
reference's frozen level (degree-59 Chebyshev activation on the P2GN16 chain ends
at level 7).  ckks.IsNegligbleThreshold (1e-14, Lattigo v2.3.0) is restated as
published; no in-tree file fixes it: parity unpinned for coefficients within a
few orders of magnitude of it.
"""
import ctypes
import math

import numpy as np

NEGLIGIBLE = 1e-14  # Lattigo v2.3.0 ckks.IsNegligbleThreshold
STANDARD, CHEBYSHEV = 0, 1


class MKFError(Exception):
    """An error value of the MKF evaluator (the Go code returns these as error)."""


class Polynomial:
    """The ckks.Polynomial fields EvaluatePoly reads: Coeffs, MaxDeg, Lead, Basis
    (A, B are carried for the caller's change of variable only)."""

    def __init__(self, coeffs, max_deg=None, lead=True, basis=STANDARD, a=-1.0, b=1.0):
        self.coeffs = [complex(c) for c in coeffs]
        self.max_deg = len(self.coeffs) - 1 if max_deg is None else int(max_deg)
        self.lead = bool(lead)
        self.basis = basis
        self.a, self.b = a, b

    def degree(self):
        return len(self.coeffs) - 1

    def depth(self):
        return int(self.degree()).bit_length()

    def copy(self):
        return Polynomial(list(self.coeffs), self.max_deg, self.lead, self.basis, self.a, self.b)


class Ct:
    """paperMKFCKKSCiphertext: NTT-domain [c0, c_1..c_k] as one array
    [1+k][level+1][N] (None when only levels and scales are tracked)."""
    __slots__ = ("data", "level", "scale")

    def __init__(self, data, level, scale):
        self.data, self.level, self.scale = data, level, scale


def not_negligible(c):
    return abs(c.real) > NEGLIGIBLE or abs(c.imag) > NEGLIGIBLE


def _odd_or_even(coeffs):
    """paperMKFOddOrEven (true_mkf_ckks_polynomial.go:41-62); zeroes the other parity."""
    odd = even = True
    for i, c in enumerate(coeffs):
        nz = not_negligible(c)
        odd = odd and not (i & 1 == 0 and nz)
        even = even and not (i & 1 == 1 and nz)
        if not odd and not even:
            break
    if odd or even:
        for i in range(1 if even else 0, len(coeffs), 2):
            coeffs[i] = 0j
    return odd, even


def _split(p, split):
    """paperMKFSplitCoefficients (:160-182)."""
    rem = Polynomial([0j] * split, basis=p.basis)
    rem.lead = False
    if p.max_deg == p.degree():
        rem.max_deg = split - 1
    else:
        rem.max_deg = p.max_deg - (p.degree() - split + 1)
    rem.coeffs[:] = p.coeffs[:split]
    quo = Polynomial([0j] * (p.degree() - split + 1), max_deg=p.max_deg, basis=p.basis)
    quo.coeffs[0] = p.coeffs[split]
    if p.basis == STANDARD:
        quo.coeffs[1:] = p.coeffs[split + 1:]
    else:
        offset = 1
        for index in range(split + 1, p.degree() + 1):
            quo.coeffs[index - split] = 2 * p.coeffs[index]
            rem.coeffs[split - offset] -= p.coeffs[index]
            offset += 1
    quo.lead = p.lead
    return quo, rem


class MKFEvaluator:
    """paperMKFCKKSEvaluator over the oracle: ``ctx`` an oracle Ctx (gamma = 1) with the
    group's fold keys and relinearization key, or None with ``Q`` for levels and scales."""

    def __init__(self, ctx, k, fold_keys=None, relin_key=None, Q=None):
        self.ctx, self.k = ctx, k
        self.fold_keys, self.relin_key = fold_keys, relin_key
        self.Q = list(ctx.Q) if ctx is not None else list(Q)
        self.N = ctx.N if ctx is not None else None
        self.mul_calls = 0
        self.poly_calls = 0

    # ---------------------------------------------------------- helpers
    def _lib(self):
        from . import lib
        return lib()

    def _halves(self, level, re, im, scale):
        from . import _p
        pos = np.zeros(level + 1, np.uint64)
        neg = np.zeros(level + 1, np.uint64)
        self._lib().mkref_const_halves(self.ctx._h, level, float(re), float(im), float(scale), _p(pos), _p(neg))
        return pos, neg

    def _mul_halves(self, src, level, pos, neg, out=None):
        """out (+)= src * C over limbs 0..level of every component (mkref_mul_const_ntt)."""
        from . import _p, _u64
        src = _u64(src)
        acc = out is not None
        if out is None:
            out = np.zeros((src.shape[0], level + 1, self.N), np.uint64)
        self._lib().mkref_mul_const_ntt(self.ctx._h, level, src.shape[0], _p(src), src.shape[1], _p(pos), _p(neg),
                                        _p(out), out.shape[1], 1 if acc else 0)
        return out

    # ---------------------------------------------------------- public operations
    def validate(self, ct):
        """validateState (true_mkf_ckks.go:161-174)."""
        if ct is None or ct.level < 0 or not ct.scale > 0:
            raise MKFError("invalid paper MKF ciphertext")
        if ct.data is not None and ct.data.shape[0] != 1 + self.k:
            raise MKFError("component count mismatch: got %d want %d" % (ct.data.shape[0] - 1, self.k))

    def new_zero(self, level, scale):
        """newZero (:176-185)."""
        data = None if self.ctx is None else np.zeros((1 + self.k, level + 1, self.N), np.uint64)
        return Ct(data, level, scale)

    def _const_and_scale(self, level, constant):
        """getConstAndScale (mkckks/evaluator.go:41-95)."""
        scale = 1.0
        if isinstance(constant, complex):
            re, im = constant.real, constant.imag
            if re != 0 and re - float(int(re)) != 0:
                scale = float(self.Q[level])
            if im != 0 and im - float(int(im)) != 0:
                scale = float(self.Q[level])
        elif isinstance(constant, float):
            re, im = constant, 0.0
            if re != 0 and re - float(int(re)) != 0:
                scale = float(self.Q[level])
        else:  # uint64, int64, int
            re, im = float(constant), 0.0
        return re, im, scale

    def mult_by_const(self, a, constant):
        """MultByConstNew (true_mkf_ckks.go:218-225 over mkckks/evaluator.go:118-199)."""
        self.validate(a)
        re, im, scale = self._const_and_scale(a.level, constant)
        data = None
        if self.ctx is not None:
            pos, neg = self._halves(a.level, re, im, scale)
            data = self._mul_halves(a.data, a.level, pos, neg)
        return Ct(data, a.level, a.scale * scale)

    def add_const(self, a, constant):
        """AddConstNew (:209-216): c0 += the constant scaled by the ciphertext scale."""
        self.validate(a)
        if isinstance(constant, complex):
            re, im = constant.real, constant.imag
        else:
            re, im = float(constant), 0.0
        data = None
        if self.ctx is not None:
            from . import _p
            pos, neg = self._halves(a.level, re, im, a.scale)
            data = a.data.copy()
            c0 = np.ascontiguousarray(data[0])
            self._lib().mkref_add_const_ntt(self.ctx._h, a.level, _p(c0), _p(pos), _p(neg))
            data[0] = c0
        return Ct(data, a.level, a.scale)

    def _add_sub(self, a, b, sub):
        """AddNew / SubNew (:187-207) over evaluateInPlace (mkckks/evaluator.go:201-305):
        the operand of smaller scale is first multiplied by floor(ratio) when that is
        above 1; the result has the larger scale and the smaller level."""
        self.validate(a)
        self.validate(b)
        level = min(a.level, b.level)
        sa, sb = a.scale, b.scale
        ta, tb = a, b
        if sb > sa and math.floor(sb / sa) > 1:
            ta = self.mult_by_const(Ct(a.data, a.level, sa), float(math.floor(sb / sa)))
        elif sa > sb and math.floor(sa / sb) > 1:
            tb = self.mult_by_const(Ct(b.data, b.level, sb), float(math.floor(sa / sb)))
        data = None
        if self.ctx is not None:
            Q = np.array(self.Q[:level + 1], np.uint64)[None, :, None]
            x = ta.data[:, :level + 1]
            y = tb.data[:, :level + 1]
            data = (x + Q - y) % Q if sub else (x + y) % Q
        return Ct(data, level, max(sa, sb))

    def add(self, a, b):
        return self._add_sub(a, b, False)

    def sub(self, a, b):
        return self._add_sub(a, b, True)

    def mult_by_gaussian_integer_and_add(self, a, re, im, target):
        """multByGaussianIntegerAndAdd (true_mkf_ckks_polynomial.go:64-74): target +=
        a * (re + im i) for Gaussian integers given as (coefficient, scale) pairs whose
        exact product is truncated (:258-264), on target's level."""
        self.validate(a)
        self.validate(target)
        level = min(a.level, target.level)
        data = None
        if self.ctx is not None:
            from . import _p
            L = self._lib()
            rr = np.array([L.mkref_trunc_product_mod(re[0], re[1], q) for q in self.Q[:level + 1]], np.uint64)
            ri = np.array([L.mkref_trunc_product_mod(im[0], im[1], q) for q in self.Q[:level + 1]], np.uint64)
            pos = np.zeros(level + 1, np.uint64)
            neg = np.zeros(level + 1, np.uint64)
            L.mkref_gauss_halves(self.ctx._h, level, _p(rr), _p(ri), _p(pos), _p(neg))
            data = target.data.copy()
            out = np.ascontiguousarray(data)
            self._mul_halves(a.data, level, pos, neg, out=out)
            data = out
        return Ct(data, target.level, target.scale)

    def rescale(self, a, min_scale):
        """RescaleNew (:227-237) over Evaluator.Rescale (mkckks/evaluator.go:360-407)."""
        self.validate(a)
        if min_scale <= 0:
            raise MKFError("cannot Rescale: minScale is 0")
        if a.scale == 0:
            raise MKFError("cannot Rescale: ciphertext scale is 0")
        if a.level == 0:
            raise MKFError("cannot Rescale: input Ciphertext already at level 0")
        s, nb = a.scale, 0
        while s / float(self.Q[a.level - nb]) >= min_scale / 2:
            s /= float(self.Q[a.level - nb])
            nb += 1
            if a.level - nb < 0:  # the reference indexes below Q[0] next
                raise MKFError("cannot Rescale: scale divides past Q[0]")
        if nb == 0:
            return Ct(None if a.data is None else a.data.copy(), a.level, a.scale)
        data = None
        if self.ctx is not None:
            data = np.stack([self.ctx.div_round_last(a.data[c], nb, ntt=True) for c in range(1 + self.k)])
        return Ct(data, a.level - nb, s)

    def drop_level(self, a, levels):
        """DropLevelNew (:239-251)."""
        self.validate(a)
        if levels < 0 or a.level - levels < 0:
            raise MKFError("invalid level drop %d from level %d" % (levels, a.level))
        lv = a.level - levels
        return Ct(None if a.data is None else a.data[:, :lv + 1].copy(), lv, a.scale)

    def mulrelin(self, a, b):
        """MulRelinNew (:253-303): fold each operand at its own level, Karatsuba tensor
        and one relinearization at the smaller level; c0 and k copies of Gamma."""
        self.validate(a)
        self.validate(b)
        level = min(a.level, b.level)
        data = None
        if self.ctx is not None:
            o0, o1 = self.ctx.mkf_ckks_mulrelin2(a.data, b.data, self.fold_keys, self.relin_key)
            data = np.stack([o0] + [o1] * self.k)
        self.mul_calls += 1
        return Ct(data, level, a.scale * b.scale)

    # ---------------------------------------------------------- EvaluatePoly
    def evaluate_poly(self, x, poly_in, target_scale):
        """EvaluatePoly (true_mkf_ckks_polynomial.go:79-111)."""
        self.validate(x)
        if poly_in is None or len(poly_in.coeffs) == 0:
            raise MKFError("polynomial is empty")
        poly = poly_in.copy()
        if x.level < poly.depth():
            raise MKFError("%d levels < %d polynomial depth" % (x.level, poly.depth()))
        powers = {1: Ct(None if x.data is None else x.data.copy(), x.level, x.scale)}
        log_degree = int(poly.degree()).bit_length()
        log_split = log_degree >> 1
        odd, even = _odd_or_even(poly.coeffs)
        for degree in range(2, 1 << log_split):
            if not (even or odd) or (degree & 1 == 0 and even) or (degree & 1 == 1 and odd):
                self._power_basis(degree, powers, target_scale, poly.basis)
        for index in range(log_split, log_degree):
            self._power_basis(1 << index, powers, target_scale, poly.basis)
        result = self._recurse(target_scale, log_split, log_degree, poly, powers)
        result.scale = target_scale
        self.poly_calls += 1
        return result

    def _power_basis(self, degree, powers, scale, basis):
        """computePowerBasis (:113-157)."""
        if degree in powers:
            return
        sub_degree = 0
        if degree & (degree - 1) == 0:
            left, right = degree // 2, degree // 2
        else:
            kk = int(math.ceil(math.log2(float(degree)))) - 1
            left = (1 << kk) - 1
            right = degree + 1 - (1 << kk)
            if basis == CHEBYSHEV:
                sub_degree = int(abs(float(left) - float(right)))
        self._power_basis(left, powers, scale, basis)
        self._power_basis(right, powers, scale, basis)
        p = self.rescale(self.mulrelin(powers[left], powers[right]), scale)
        if basis == CHEBYSHEV:
            p = self.add(p, p)
            if sub_degree == 0:
                p = self.add_const(p, -1)
            else:
                self._power_basis(sub_degree, powers, scale, basis)
                p = self.sub(p, powers[sub_degree])
        powers[degree] = p

    def _recurse(self, target_scale, log_split, log_degree, coeffs, powers):
        """recursePolynomial (:184-228)."""
        if coeffs.degree() < (1 << log_split):
            if (coeffs.lead and log_split > 1
                    and coeffs.max_deg % (1 << (log_split + 1)) > (1 << (log_split - 1))):
                log_degree = int(coeffs.degree()).bit_length()
                log_split = log_degree >> 1
                return self._recurse(target_scale, log_split, log_degree, coeffs, powers)
            return self._from_power_basis(target_scale, coeffs, powers)
        next_power = 1 << log_split
        while next_power < (coeffs.degree() >> 1) + 1:
            next_power <<= 1
        quo, rem = _split(coeffs, next_power)
        if next_power not in powers:
            raise MKFError("power %d was not computed" % next_power)
        level = powers[next_power].level - 1
        if quo.max_deg >= (1 << (log_degree - 1)) and quo.lead:
            level += 1
        qi = float(self.Q[level])
        left = self._recurse(target_scale * qi / powers[next_power].scale, log_split, log_degree, quo, powers)
        right = self._recurse(target_scale, log_split, log_degree, rem, powers)
        if left.level > right.level:
            while left.level != right.level + 1:
                left = self.drop_level(left, 1)
        left = self.mulrelin(left, powers[next_power])
        if left.level > right.level:
            left = self.rescale(left, target_scale)
            return self.add(left, right)
        left = self.add(left, right)
        return self.rescale(left, target_scale)

    def _from_power_basis(self, target_scale, coeffs, powers):
        """evaluateFromPowerBasis (:230-273)."""
        min_degree = 0
        for degree in range(coeffs.degree(), 0, -1):
            if not_negligible(coeffs.coeffs[degree]):
                min_degree = degree
                break
        constant = coeffs.coeffs[0]
        if min_degree == 0:
            result = self.new_zero(powers[1].level, target_scale)
            if not_negligible(constant):
                return self.add_const(result, constant)
            return result
        min_degree = coeffs.degree()
        if min_degree not in powers:
            raise MKFError("power %d was not computed" % min_degree)
        qi = float(self.Q[powers[min_degree].level])
        result = self.new_zero(powers[min_degree].level, target_scale * qi)
        if not_negligible(constant):
            result = self.add_const(result, constant)
        for degree in range(coeffs.degree(), 0, -1):
            c = coeffs.coeffs[degree]
            if not not_negligible(c):
                continue
            if degree not in powers:
                raise MKFError("power %d was not computed" % degree)
            cs = target_scale * qi / powers[degree].scale
            result = self.mult_by_gaussian_integer_and_add(powers[degree], (c.real, cs), (c.imag, cs), result)
        return self.rescale(result, target_scale)


def declare(L):
    """ctypes prototypes of mkref_poly.c and the two-level product."""
    U, I, D = ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    V = ctypes.c_void_p
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.mkref_scale_up_exact.restype = U
    L.mkref_scale_up_exact.argtypes = [D, D, U]
    L.mkref_trunc_product_mod.restype = U
    L.mkref_trunc_product_mod.argtypes = [D, D, U]
    L.mkref_i64_mod.restype = U
    L.mkref_i64_mod.argtypes = [ctypes.c_int64, U]
    L.mkref_const_halves.argtypes = [V, I, D, D, D, u64p, u64p]
    L.mkref_gauss_halves.argtypes = [V, I, u64p, u64p, u64p, u64p]
    L.mkref_mul_const_ntt.argtypes = [V, I, I, u64p, I, u64p, u64p, u64p, I, I]
    L.mkref_add_const_ntt.argtypes = [V, I, u64p, u64p, u64p]
    L.mkref_mkf_ckks_mulrelin2.argtypes = [V, I, I, I, u64p, u64p, ctypes.POINTER(u64p), u64p, u64p, u64p]
