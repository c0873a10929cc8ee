"""GPU parity of the MKF polynomial evaluator and its linear operations with the CPU
oracle (bit-exact words, equal levels and scales): paperMKFCKKSEvaluator's
EvaluatePoly, MulRelinNew of operands at different levels, Add/Sub with scale
matching, AddConst, MultByConst, MultByGaussianIntegerAndAdd and Rescale
(software/mkf/mkf_cpu/true_mkf_ckks_polynomial.go:18-273, true_mkf_ckks.go:187-303).
The oracle's polynomial evaluation is pinned by the decrypt tests and the frozen
level schedule of test_oracle_poly.py."""
import numpy as np
import pytest

import mkgen as G
import oracle as O
from params import config

pytestmark = pytest.mark.gpu


def _chain(name):
    if name == "q40":  # N = 2^13: one 60-bit prime, six 40-bit primes (scale 2^40), two 60-bit P
        Q = [int(q) for q in O.gen_primes(60, 13, 1)] + [int(q) for q in O.gen_primes(40, 13, 6)]
        P = [int(p) for p in O.gen_primes(60, 13, 2, exclude=Q)]
        return 13, Q, P, float(2 ** 40)
    if name == "q40n15":  # N = 2^15 (limbs over two workgroups): one 60-bit and three 40-bit primes
        Q = [int(q) for q in O.gen_primes(60, 15, 1)] + [int(q) for q in O.gen_primes(40, 15, 3)]
        P = [int(p) for p in O.gen_primes(60, 15, 2, exclude=Q)]
        return 15, Q, P, float(2 ** 40)
    logN, Q, P = config("P1")  # the reference's KKLSS / MKF parameter literal, 52-bit scale primes
    return logN, [int(q) for q in Q], [int(p) for p in P], float(2 ** 52)


@pytest.fixture(scope="module", params=["q40", "P1", "q40n15"])
def env(request):
    import mkf_amd

    logN, Q, P, D = _chain(request.param)
    ctx = mkf_amd.Context(logN, Q, P, gamma=1)
    ref = O.Ctx(logN, Q, P, gamma=1)
    k = 2
    keys = G.MKFKeys(ref, k, seed=71)
    ctx.register_mkf_group(5, keys.fold, keys.relin)
    ev = O.MKFEvaluator(ref, k, keys.fold, keys.relin)
    yield request.param, ctx, ref, keys, ev, D
    ctx.close()


def _ct(ref, g, k, level):
    return np.stack([g.uniform(ref.Q[:level + 1], ref.N) for _ in range(1 + k)])


def _encrypt(ref, keys, g, D):
    N = ref.N
    m = np.zeros(N)
    m[0], m[3], m[N // 2], m[N // 2 + 5] = 0.12, -0.07, 0.05, 0.04
    mi = np.round(m * D).astype(np.int64)
    parts = [keys.encrypt(p, O.small_to_rns(mi if p == 0 else np.zeros(N, np.int64), ref.Q), g)
             for p in range(len(keys.small))]
    c0 = parts[0][0]
    for p in parts[1:]:
        c0 = G.addmod(c0, p[0], ref.Q)
    return np.stack([c0] + [p[1] for p in parts])


def test_linear_operations(env):
    name, ctx, ref, keys, ev, D = env
    g = O.Prng(72)
    k = 2
    L = ref.L
    a, b = _ct(ref, g, k, L - 1), _ct(ref, g, k, L - 3)
    for sa, sb in ((D, D * 1000.7), (D * 3.5, D), (D, D * 1.25)):
        for sub in (False, True):
            exp = (ev.sub if sub else ev.add)(O.Ct(a, L - 1, sa), O.Ct(b, L - 3, sb))
            got, lo, so = ctx.mkf_add(a, sa, b, sb, sub=sub)
            assert np.array_equal(got, exp.data) and lo == exp.level and so == exp.scale, (sa, sb, sub)
    for c in (complex(0.3, -0.7), complex(-1.0, 0.0), complex(0.0, 2.5), complex(5.0, 0.0)):
        exp = ev.add_const(O.Ct(a, L - 1, D), c)
        assert np.array_equal(ctx.mkf_add_const(a, D, c), exp.data), c
        exp = ev.mult_by_const(O.Ct(a, L - 1, D), c)
        got, so = ctx.mkf_mult_by_const(a, D, c)
        assert np.array_equal(got, exp.data) and so == exp.scale, c
    for re, im in ((-17, 9), (123456789, 0), (0, -5), (-(1 << 62), 1 << 61)):
        exp = ev.mult_by_gaussian_integer_and_add(O.Ct(a, L - 1, D), (float(re), 1.0), (float(im), 1.0),
                                                  O.Ct(b, L - 3, D))
        assert np.array_equal(ctx.mkf_gauss_add(a, re, im, b), exp.data), (re, im)
    for s, ms in ((D * D, D), (D * D * D, D), (D, D)):
        exp = ev.rescale(O.Ct(a, L - 1, s), ms)
        got, lo, so = ctx.mkf_rescale(a, s, ms)
        assert lo == exp.level and so == exp.scale and np.array_equal(got, exp.data), (s, ms)


def test_mulrelin_of_operands_at_different_levels(env):
    name, ctx, ref, keys, ev, D = env
    g = O.Prng(73)
    L = ref.L
    x = _encrypt(ref, keys, g, D)
    for ll, lr in ((L - 1, L - 3), (L - 2, L - 1), (L - 2, L - 2)):
        exp = ev.mulrelin(O.Ct(x[:, :ll + 1].copy(), ll, D), O.Ct(x[:, :lr + 1].copy(), lr, D))
        got = ctx.mkf_mulrelin_levels(5, x[:, :ll + 1], x[:, :lr + 1])
        assert np.array_equal(got, exp.data), (ll, lr)


def _polys(name):
    import mkf_amd

    relu = np.polynomial.chebyshev.chebinterpolate(lambda t: np.maximum(t - 0.2, 0) * 0.8 + 0.1 * t * t, 15)
    odd = np.polynomial.chebyshev.chebinterpolate(lambda t: np.sin(2 * t), 9)
    odd[0::2] = 0
    cases = [([0.1, 0.2, -0.3, 0.4], mkf_amd.BASIS_STANDARD)]  # true_mkf_ckks_test.go:283
    if name == "q40":
        cases += [(list(relu), mkf_amd.BASIS_CHEBYSHEV), (list(odd), mkf_amd.BASIS_CHEBYSHEV),
                  ([0.7], mkf_amd.BASIS_STANDARD), ([0.1, -0.5], mkf_amd.BASIS_STANDARD),
                  ([0.1, 0.2, 0.0, 0.0], mkf_amd.BASIS_STANDARD), ([0.3, 0.0, -0.4], mkf_amd.BASIS_CHEBYSHEV),
                  ([0.25 - 0.5j, 1j, 0.3 + 0.1j, 0, -0.2j, 0.15], mkf_amd.BASIS_STANDARD)]
    return cases


def test_evaluate_poly(env):
    import mkf_amd

    name, ctx, ref, keys, ev, D = env
    g = O.Prng(74)
    x = _encrypt(ref, keys, g, D)
    lv = ref.L - 1
    for coeffs, basis in _polys(name):
        op = O.Polynomial(coeffs, basis=basis)
        mp = mkf_amd.Polynomial(coeffs, basis=basis)
        ev.mul_calls = 0
        exp = ev.evaluate_poly(O.Ct(x, lv, D), op, D)
        lq, sq, nm = ctx.mkf_poly_query(5, lv, D, mp, D)
        assert (lq, sq, nm) == (exp.level, exp.scale, ev.mul_calls), coeffs
        got, lo, so = ctx.mkf_evaluate_poly(5, x, D, mp, D)
        assert lo == exp.level and so == exp.scale
        assert np.array_equal(got, exp.data), coeffs


def test_poly_plan_batch_and_replays(env):
    """nct ciphertexts through one plan (one hipGraph); every replay gives the words of
    the single-ciphertext evaluation."""
    import mkf_amd

    name, ctx, ref, keys, ev, D = env
    g = O.Prng(75)
    lv = ref.L - 1
    coeffs, basis = _polys(name)[-1]
    mp = mkf_amd.Polynomial(coeffs, basis=basis)
    xs = [_encrypt(ref, keys, g, D) for _ in range(3)]
    want = [ctx.mkf_evaluate_poly(5, x, D, mp, D) for x in xs]
    lo = want[0][1]
    ins = [ctx.alloc(x.nbytes) for x in xs]
    for d, x in zip(ins, xs):
        ctx.h2d(d, x)
    shape = (3, lo + 1, ref.N)
    outs = [ctx.alloc(3 * (lo + 1) * ref.N * 8) for _ in xs]
    plan = mkf_amd.PolyPlan(ctx, 5, ins, lv, D, mp, D, outs)
    assert plan.level_out == lo and plan.scale_out == want[0][2]
    for rep in range(3):
        for d in outs:
            ctx.h2d(d, np.zeros(shape, np.uint64))
        plan.launch()
        ctx.sync()
        for i, d in enumerate(outs):
            got = ctx.d2h(np.zeros(shape, np.uint64), d)
            assert np.array_equal(got, want[i][0]), (rep, i)
    plan.close()
    for d in ins + outs:
        ctx.free(d)


def test_poly_errors(env):
    import mkf_amd

    name, ctx, ref, keys, ev, D = env
    x = _ct(ref, O.Prng(76), 2, 2)
    with pytest.raises(mkf_amd.MkhipError) as e:  # "levels < polynomial depth"
        ctx.mkf_evaluate_poly(5, x, D, mkf_amd.Polynomial([0.0] * 8 + [1.0]), D)
    assert e.value.status == -5
    with pytest.raises(mkf_amd.MkhipError) as e:
        ctx.mkf_evaluate_poly(5, x, D, mkf_amd.Polynomial([]), D)
    assert e.value.status == -1
    with pytest.raises(mkf_amd.MkhipError) as e:  # Rescale at level 0
        ctx.mkf_rescale(x[:, :1], D * D, D)
    assert e.value.status == -9
