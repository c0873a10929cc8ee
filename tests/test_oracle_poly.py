"""Pin the MKF polynomial-evaluator part of the CPU oracle (mkref_poly.c, oracle/poly.py).

* Exact integer maps: scaleUpExact (mkckks/utils.go:59-84) and the truncated exact
  product of EvaluatePoly's constants (true_mkf_ckks_polynomial.go:258-264) against
  Python's exact rationals; the Euclidean residue against Python's %.
* Closed forms outside the NTT domain: AddConst(a + b i) adds round(a D) at X^0 and
  round(b D) at X^(N/2) of c0 only; MultByConst / MultByGaussianIntegerAndAdd are
  the negacyclic product with a + b X^(N/2); Add / Sub scale the operand of smaller
  scale by floor(ratio) (mkckks/evaluator.go:201-305).
* Decrypt-and-compare: products of operands at different levels, and EvaluatePoly
  on standard / Chebyshev, odd / mixed-parity and complex polynomials, against the
  polynomial evaluated on the CKKS slots (true_mkf_ckks_test.go:283-300 checks its
  degree-3 case within 2e-5).
* The level schedule: the reference's frozen level for the degree-59 Chebyshev
  activation on the P2GN16 chain (p2gn16_lattigo_level_semantics_test.go:45-51) and
  its three-prime packing rescale (:56-73).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import mkgen as G
import oracle as O


def _ctx(logN=10, nq=6, k=2, seed=11):
    Q = [int(q) for q in O.gen_primes(60, logN, 1)] + [int(q) for q in O.gen_primes(40, logN, nq)]
    P = O.gen_primes(60, logN, 2, exclude=Q)
    ctx = O.Ctx(logN, Q, P, gamma=1)
    keys = G.MKFKeys(ctx, k, seed=seed)
    return ctx, keys


def _encrypt(ctx, keys, g, msg, level=None):
    """[c0, c_1..c_k] NTT domain with party 0 holding msg (integer coefficients)."""
    Q = ctx.Q
    k = len(keys.small)
    parts = [keys.encrypt(p, O.small_to_rns(msg if p == 0 else np.zeros(ctx.N, np.int64), Q), g)
             for p in range(k)]
    c0 = parts[0][0]
    for p in parts[1:]:
        c0 = G.addmod(c0, p[0], Q)
    ct = np.stack([c0] + [p[1] for p in parts])
    if level is not None:
        ct = ct[:, :level + 1].copy()
    return ct


def _phase(ctx, keys, ct):
    """c0 + sum_i c_i s_i with each party's own secret, centered coefficients."""
    nl = ct.shape[1]
    Q = ctx.Q[:nl]
    acc = ct[0].copy()
    for i, s in enumerate(keys.small):
        si = ctx.ntt_poly(O.small_to_rns(s, Q))
        acc = G.addmod(acc, G.mulmod(ct[1 + i], si, Q), Q)
    return G.centered_crt(ctx.ntt_poly(acc, inverse=True), Q)


def _slots(coeffs):
    """Values of a real coefficient vector at the CKKS slot roots zeta^(4j+1), zeta = e^(i pi / N)."""
    N = len(coeffs)
    n = np.arange(N)
    y = np.fft.ifft(np.asarray(coeffs, complex) * np.exp(1j * np.pi * n / N)) * N
    return y[0::2]


def _message(g, N, D):
    """A message of a few terms (slot magnitudes below 0.3) as integer coefficients."""
    m = np.zeros(N)
    m[0] = 0.12
    m[3] = -0.07
    m[N // 2] = 0.05
    m[N // 2 + 5] = 0.04
    return np.round(m * D).astype(np.int64), m


# ------------------------------------------------------------------ exact integer maps

def test_scale_up_exact():
    L = O.lib()
    q = (1 << 40) - 87 * 2048 + 1
    assert L.mkref_scale_up_exact(3.0, float(1 << 30), q) == (3 << 30) % q
    assert L.mkref_scale_up_exact(2.5, 1.0, q) == 3        # + 0.5, then truncation
    assert L.mkref_scale_up_exact(2.4999, 1.0, q) == 2
    assert L.mkref_scale_up_exact(-2.5, 1.0, q) == q - 3   # sign applied after rounding |x|
    assert L.mkref_scale_up_exact(-0.2, 1.0, q) == q       # q - 0, as written
    assert L.mkref_scale_up_exact(1.0, float(2 ** 70), q) == pow(2, 70, q)
    rng = np.random.default_rng(1)
    for _ in range(200):
        v = float(rng.uniform(-1, 1)) * 10.0 ** int(rng.integers(-3, 3))
        n = float(2 ** int(rng.integers(20, 60)))
        x = Fraction(abs(v) * n)           # the float64 product, exact
        y = Fraction(float(x) + 0.5)       # rounded to 53 bits like the big.Float sum
        r = math.floor(y) % q
        assert L.mkref_scale_up_exact(v, n, q) == (q - r if v < 0 else r)


def test_trunc_product_mod_and_signed_residue():
    L = O.lib()
    rng = np.random.default_rng(2)
    for q in (int(O.gen_primes(40, 10, 1)[0]), int(O.gen_primes(60, 10, 1)[0])):
        for _ in range(300):
            a = float(rng.normal()) * 10.0 ** int(rng.integers(-6, 6))
            b = float(2.0 ** int(rng.integers(-10, 90))) * float(rng.uniform(0.5, 2))
            if rng.integers(2):
                b = -b
            exact = Fraction(a) * Fraction(b)
            t = math.trunc(exact)
            assert L.mkref_trunc_product_mod(a, b, q) == t % q, (a, b)
        for x in (-17, 9, 0, -1, -(1 << 62), (1 << 62) + 5, -q, q):
            assert L.mkref_i64_mod(x, q) == x % q


# ------------------------------------------------------------------ closed forms

def test_add_const_and_mult_by_const_closed_forms():
    ctx, keys = _ctx(nq=3)
    g = O.Prng(21)
    N, k = ctx.N, 2
    ev = O.MKFEvaluator(ctx, k, keys.fold, keys.relin)
    level = 2
    Q = ctx.Q[:level + 1]
    data = np.stack([g.uniform(Q, N) for _ in range(1 + k)])
    D = float(2 ** 38)
    x = O.Ct(data, level, D)
    y = ev.add_const(x, complex(0.3, -0.7))
    assert y.scale == D and y.level == level
    assert np.array_equal(y.data[1:], data[1:])
    diff = G.centered_crt(ctx.ntt_poly(G.submod(y.data[0], data[0], Q), inverse=True), Q)
    want = [0] * N
    want[0] = math.floor(0.3 * D + 0.5)
    want[N // 2] = -math.floor(0.7 * D + 0.5)
    assert diff == want
    # a fractional constant is scaled by Q[level]; an integer one by 1
    z = ev.mult_by_const(x, complex(-0.25, 0.5))
    assert z.scale == D * Q[level]
    ca = [0] * N
    ca[0] = -math.floor(0.25 * Q[level] + 0.5)
    ca[N // 2] = math.floor(0.5 * Q[level] + 0.5)
    for c in range(1 + k):
        got = G.centered_crt(ctx.ntt_poly(z.data[c], inverse=True), Q)
        src = G.centered_crt(ctx.ntt_poly(data[c], inverse=True), Q)
        Qp = Q[0] * Q[1] * Q[2]
        exp = [((v + Qp // 2) % Qp) - Qp // 2 for v in G.negacyclic_mul(src, ca)]
        assert got == exp, c
    w = ev.mult_by_const(x, 3)
    assert w.scale == D
    Qa = np.array(Q, np.uint64)[None, :, None]
    assert np.array_equal(w.data, data * np.uint64(3) % Qa)


def test_gaussian_integer_and_add_closed_form():
    """MultByGaussianIntegerAndAdd(-17 + 9 i) -- the reference's observer inputs
    (p2gn16_same_run_physical_operand_observer_overlay_test.go:649-650) -- and a
    truncated-product constant."""
    ctx, keys = _ctx(nq=3)
    g = O.Prng(22)
    N, k = ctx.N, 2
    ev = O.MKFEvaluator(ctx, k, keys.fold, keys.relin)
    level = 2
    Q = ctx.Q[:level + 1]
    Qp = Q[0] * Q[1] * Q[2]
    a = np.stack([g.uniform(ctx.Q[:level + 2], N) for _ in range(1 + k)])  # one limb above the target
    t = np.stack([g.uniform(Q, N) for _ in range(1 + k)])
    for (re, im) in (((-17.0, 1.0), (9.0, 1.0)), ((0.37, 2.0 ** 45), (-1.5e-3, 2.0 ** 45))):
        out = ev.mult_by_gaussian_integer_and_add(O.Ct(a, level + 1, 1.0), re, im, O.Ct(t, level, 5.0))
        assert out.level == level and out.scale == 5.0
        cr = math.trunc(Fraction(re[0]) * Fraction(re[1]))
        ci = math.trunc(Fraction(im[0]) * Fraction(im[1]))
        cc = [0] * N
        cc[0], cc[N // 2] = cr, ci
        for c in range(1 + k):
            got = G.centered_crt(ctx.ntt_poly(G.submod(out.data[c], t[c], Q), inverse=True), Q)
            src = G.centered_crt(ctx.ntt_poly(a[c, :level + 1], inverse=True), Q)
            exp = [((v + Qp // 2) % Qp) - Qp // 2 for v in G.negacyclic_mul(src, cc)]
            assert got == exp, (re, im, c)


def test_add_sub_scale_matching():
    ctx, keys = _ctx(nq=3)
    g = O.Prng(23)
    N, k = ctx.N, 2
    ev = O.MKFEvaluator(ctx, k, keys.fold, keys.relin)
    Q3 = ctx.Q[:3]
    a = np.stack([g.uniform(ctx.Q[:4], N) for _ in range(1 + k)])
    b = np.stack([g.uniform(Q3, N) for _ in range(1 + k)])
    sa, sb = 2.0 ** 30, 2.0 ** 30 * 1000.7
    s = ev.sub(O.Ct(a, 3, sa), O.Ct(b, 2, sb))     # a is scaled by floor(1000.7) = 1000
    assert s.level == 2 and s.scale == sb
    Qo = np.array(Q3, object)[None, :, None]
    want = ((a[:, :3].astype(object) * 1000 % Qo + Qo - b.astype(object)) % Qo).astype(np.uint64)
    assert np.array_equal(s.data, want)
    Qa = np.array(Q3, np.uint64)[None, :, None]
    u = ev.add(O.Ct(a, 3, sa), O.Ct(b, 2, sa * 1.5))  # floor(1.5) = 1: no scaling
    assert np.array_equal(u.data, (a[:, :3] + b) % Qa) and u.scale == sa * 1.5
    with pytest.raises(O.MKFError):
        ev.rescale(O.Ct(b[:, :1], 0, sa), sa)
    with pytest.raises(O.MKFError):
        ev.drop_level(O.Ct(b, 2, sa), 3)


# ------------------------------------------------------------------ products and EvaluatePoly

def test_mulrelin_of_operands_at_different_levels():
    ctx, keys = _ctx(nq=4, k=3, seed=31)
    g = O.Prng(32)
    N = ctx.N
    D = float(ctx.Q[-1])
    ev = O.MKFEvaluator(ctx, 3, keys.fold, keys.relin)
    mi, mf = _message(g, N, D)
    x = O.Ct(_encrypt(ctx, keys, g, mi), 4, D)
    y = O.Ct(_encrypt(ctx, keys, g, mi, level=2), 2, D)
    p = ev.mulrelin(x, y)
    assert p.level == 2 and p.scale == D * D
    # equal levels: the CKKSCase.Run product of the same operands
    o0, o1 = ctx.mkf_ckks_mulrelin(x.data[:, :3], y.data, keys.fold, keys.relin)
    q0 = ev.mulrelin(O.Ct(x.data[:, :3].copy(), 2, D), y)
    assert np.array_equal(q0.data[0], o0) and np.array_equal(q0.data[1], o1)
    # folding x at level 4 is not folding it at level 2, but both decrypt to x*y
    assert not np.array_equal(p.data, q0.data)
    for r in (p, q0):
        got = np.array(_phase(ctx, keys, r.data), float) / (D * D)
        exp = _slots(mf) ** 2
        assert np.max(np.abs(_slots(got) - exp)) < 1e-6


def _run_poly(poly, logN=10, nq=6, k=2, seed=41, level=None):
    ctx, keys = _ctx(logN=logN, nq=nq, k=k, seed=seed)
    g = O.Prng(seed + 1)
    D = float(2 ** 40)
    ev = O.MKFEvaluator(ctx, k, keys.fold, keys.relin)
    mi, mf = _message(g, ctx.N, D)
    lv = ctx.L - 1 if level is None else level
    x = O.Ct(_encrypt(ctx, keys, g, mi, level=lv), lv, D)
    out = ev.evaluate_poly(x, poly, D)
    got = _slots(np.array(_phase(ctx, keys, out.data), float) / D)
    z = _slots(mf)
    if poly.basis == O.CHEBYSHEV:
        exp = np.polynomial.chebyshev.chebval(z, np.array(poly.coeffs))
    else:
        exp = np.polynomial.polynomial.polyval(z, np.array(poly.coeffs))
    return ev, out, np.max(np.abs(got - exp))


def test_evaluate_poly_degree3_standard():
    """The reference's degree-3 case (true_mkf_ckks_test.go:283): 0.1 + 0.2x - 0.3x^2 + 0.4x^3."""
    ev, out, err = _run_poly(O.Polynomial([0.1, 0.2, -0.3, 0.4]))
    assert err < 2e-5
    assert out.level == 6 - 2 and ev.mul_calls == 2


def test_evaluate_poly_chebyshev_mixed_parity_degree15():
    c = np.polynomial.chebyshev.chebinterpolate(lambda t: np.maximum(t - 0.2, 0) * 0.8 + 0.1 * t * t, 15)
    ev, out, err = _run_poly(O.Polynomial(c, basis=O.CHEBYSHEV))
    assert err < 1e-4
    assert out.level == 6 - 4


def test_evaluate_poly_odd_chebyshev_and_complex_standard():
    c = np.polynomial.chebyshev.chebinterpolate(lambda t: np.sin(2 * t), 9)
    c[0::2] = 0
    ev, out, err = _run_poly(O.Polynomial(c, basis=O.CHEBYSHEV))
    assert err < 1e-4
    ev2, out2, err2 = _run_poly(O.Polynomial([0.25 - 0.5j, 1j, 0.3 + 0.1j, 0, -0.2j, 0.15]))
    assert err2 < 1e-4


def test_evaluate_poly_small_degrees():
    """Degree 0 (a constant: no product), degree 1, zero top coefficients, an even
    Chebyshev polynomial."""
    for coeffs, basis in (([0.7], O.STANDARD), ([0.1, -0.5], O.STANDARD), ([0.1, 0.2, 0.0, 0.0], O.STANDARD),
                          ([0.3, 0.0, -0.4], O.CHEBYSHEV)):
        ev, out, err = _run_poly(O.Polynomial(coeffs, basis=basis), nq=3)
        assert err < 1e-4, coeffs
        assert out.level == 3 - O.Polynomial(coeffs).depth(), coeffs


def test_evaluate_poly_errors():
    ctx, keys = _ctx(nq=2)
    ev = O.MKFEvaluator(ctx, 2, keys.fold, keys.relin)
    x = O.Ct(np.zeros((3, 3, ctx.N), np.uint64), 2, 2.0 ** 40)
    with pytest.raises(O.MKFError, match="polynomial depth"):
        ev.evaluate_poly(x, O.Polynomial([0.0] * 8 + [1.0]), 2.0 ** 40)  # degree 8: depth 4
    with pytest.raises(O.MKFError, match="empty"):
        ev.evaluate_poly(x, O.Polynomial([]), 2.0 ** 40)


# ------------------------------------------------------------------ the level schedule

# P2GN16 persistent chain: the security candidate's Q without its guard prime
# (p2g_n16_security_candidate_test.go:19-38, guarded_compact_p2g_n16_optin_test.go:57-58)
P2GN16_Q = [0xffffffffffc0001,
            0x3fffffffd60001, 0x3fffffffca0001, 0x3fffffff360001, 0x3fffffff2a0001,
            0x3fffffff000001, 0x3ffffffefa0001, 0x3ffffffef40001, 0x3ffffffeaa0001,
            0x3ffffffe920001, 0x3ffffffe800001, 0x3ffffffe320001, 0x3ffffffe080001,
            0x3ffffffdcc0001]


def test_level_semantics_p2gn16():
    ev = O.MKFEvaluator(None, 1, Q=P2GN16_Q)
    scale = float(1 << 54)
    top = len(P2GN16_Q) - 1
    relu = np.polynomial.chebyshev.chebinterpolate(lambda t: np.maximum(t, 0.0), 59)
    poly = O.Polynomial(relu, basis=O.CHEBYSHEV)
    x = O.Ct(None, top, scale)
    x = ev.mult_by_const(x, 2 / (poly.b - poly.a))
    x = ev.add_const(x, (-poly.a - poly.b) / (poly.b - poly.a))
    before = x.level
    x = ev.rescale(x, scale)
    assert x.level == before and x.scale == scale
    out = ev.evaluate_poly(x, poly, scale)
    assert out.level == 7
    # five masks at the fifth root of the top three active primes: one Rescale drops three
    lv = top - 1
    mask = (float(P2GN16_Q[lv]) * float(P2GN16_Q[lv - 1]) * float(P2GN16_Q[lv - 2])) ** (1.0 / 5.0)
    s = scale
    for _ in range(5):
        s *= mask
    packed = ev.rescale(O.Ct(None, lv, s), scale)
    assert lv - packed.level == 3 and packed.level == 9
