"""GPU parity of the Galois key switches against the CPU oracle (bit-exact).

KKLSS Rotate / RotateHoisted / Conjugate (mkrlwe/keyswitch.go:243-344,
keyswitch_hoisted.go:195-262) with RotateNew's normalisation and power-of-two
fallback (mkckks/evaluator.go:496-529), on CKKS and BFV contexts; MKF
party-specific RotateNew (mkf_cpu/true_mkf_ckks.go:323-353).  Keys, CRS and
ciphertexts are uniform residues (bit-exact parity needs no valid encryption;
the decrypt-level checks are in test_oracle_rotation.py).  Zeros planted in the
inputs exercise the reference's "q - x" negation (a negated zero is written q)
and the decomposition of such q-valued coefficients.
"""
import numpy as np
import pytest

import oracle as O
from params import bfv_config, config

pytestmark = pytest.mark.gpu


def _mk():
    import mkf_amd

    return mkf_amd


@pytest.fixture(scope="module", params=["small", "c2", "n15"])
def kk(request):
    mk = _mk()
    logN, Q, P = config(request.param)
    ctx = mk.Context(logN, Q, P, gamma=2)
    ref = O.Ctx(logN, Q, P, gamma=2)
    yield request.param, ctx, ref
    ctx.close()


def _galois_keys(ctx, ref, parties, rots, seed, conj=False):
    g = O.Prng(seed)
    beta = ref.beta()

    def swk():
        return np.stack([g.uniform(ref.QP, ref.N) for _ in range(beta)])

    crs, rk, ck = {}, {}, {}
    for r in rots:
        crs[r] = swk()
        ctx.set_crs(r, crs[r])
        for p in parties:
            rk[(p, r)] = swk()
            ctx.add_rotation_key(p, r, rk[(p, r)])
    if conj:
        crs[-2] = swk()
        ctx.set_crs(-2, crs[-2])
        for p in parties:
            ck[p] = swk()
            ctx.add_conjugation_key(p, ck[p])
    return crs, rk, ck


def _ct(ref, k, seed, level=None, zeros=False):
    g = O.Prng(seed)
    nl = ref.L if level is None else level + 1
    ct = np.stack([g.uniform(ref.Q[:nl], ref.N) for _ in range(1 + k)])
    if zeros:  # runs of zeros: i gal < N keeps 3..39 positive, N/4.. lands negated under 5
        ct[:, :, 3:40] = 0
        ct[:, :, ref.N // 4:ref.N // 4 + 37] = 0
    return ct


def test_kklss_rotate_parity(kk):
    name, ctx, ref = kk
    ids = [0, 3]
    crs, rk, _ = _galois_keys(ctx, ref, ids, [1, 4], seed=101)
    ct = _ct(ref, 2, seed=102)
    for r in (1, 4):
        gal = O.galois_element(ref.logN, r)
        exp = ref.kklss_rotate(ct, [rk[(p, r)] for p in ids], crs[r], gal)
        assert np.array_equal(ctx.kklss_rotate(ct, ids, r), exp), r
        assert np.array_equal(ctx.kklss_rotate(ct, ids, r, hoisted=True), exp), r
        # rotidx normalised into [0, N/2) (evaluator.go:498-505)
        assert np.array_equal(ctx.kklss_rotate(ct, ids, r - ref.N // 2), exp), r


def test_kklss_rotate_below_top_level(kk):
    name, ctx, ref = kk
    ids = [1]
    crs, rk, _ = _galois_keys(ctx, ref, ids, [2], seed=103)
    level = ref.L - 2
    ct = _ct(ref, 1, seed=104, level=level)
    exp = ref.kklss_rotate(ct, [rk[(1, 2)]], crs[2], O.galois_element(ref.logN, 2))
    assert np.array_equal(ctx.kklss_rotate(ct, ids, 2), exp)


def test_kklss_rotate_power_of_two_fallback(kk):
    name, ctx, ref = kk
    mk = _mk()
    ids = [0, 2]
    crs, rk, _ = _galois_keys(ctx, ref, ids, [1, 2, 4], seed=105)
    ct = _ct(ref, 2, seed=106)
    exp = ref.kklss_rotate_new(ct, ids, 7, rk, crs)
    assert np.array_equal(ctx.kklss_rotate(ct, ids, 7), exp)
    assert np.array_equal(ctx.kklss_rotate(ct, ids, 0), ct)
    with pytest.raises(mk.MkhipError) as e:  # RotateHoistedNew panics without the key
        ctx.kklss_rotate(ct, ids, 7, hoisted=True)
    assert e.value.status == -6


def test_kklss_conjugate_parity_with_negated_zeros(kk):
    name, ctx, ref = kk
    ids = [0, 1]
    crs, _, ck = _galois_keys(ctx, ref, ids, [], seed=107, conj=True)
    ct = _ct(ref, 2, seed=108, zeros=True)
    exp = ref.kklss_conjugate(ct, [ck[p] for p in ids], crs[-2])
    assert np.array_equal(ctx.kklss_conjugate(ct, ids), exp)


def test_kklss_rotate_without_parties_keeps_q(kk):
    """k = 0: the permuted c0 is the result, negated zeros written as q."""
    name, ctx, ref = kk
    crs, _, _ = _galois_keys(ctx, ref, [], [1, 2], seed=109)
    ct = _ct(ref, 0, seed=110, zeros=True)
    exp = ref.kklss_rotate_new(ct, [], 3, {}, crs)
    got = ctx.kklss_rotate(ct, [], 3)
    assert np.array_equal(got, exp)
    assert any(int(got[0][j].max()) == ref.Q[j] for j in range(ref.L))  # the q values survive


def test_kklss_rotation_plan_hoisted_batch(kk):
    """One decomposition per op serves every rotation; a conjugation in the same
    op; graph replays agree with the oracle."""
    name, ctx, ref = kk
    mk = _mk()
    ids = [0, 1]
    rots = [1, 2]
    crs, rk, ck = _galois_keys(ctx, ref, ids, rots, seed=111, conj=True)
    nl, N, k = ref.L, ref.N, len(ids)
    ops, bufs, exp = [], [], []
    for b in range(2):
        ct = _ct(ref, k, seed=112 + b)
        din, dout = ctx.alloc(ct.nbytes), ctx.alloc(3 * ct.nbytes)
        ctx.h2d(din, ct)
        bufs += [din, dout]
        ops.append(dict(**{"in": din}, out=dout, ids=ids, level=nl - 1, rot=[1, mk.CONJUGATE, 2]))
        e = [ref.kklss_rotate(ct, [rk[(p, 1)] for p in ids], crs[1], O.galois_element(ref.logN, 1)),
             ref.kklss_conjugate(ct, [ck[p] for p in ids], crs[-2]),
             ref.kklss_rotate(ct, [rk[(p, 2)] for p in ids], crs[2], O.galois_element(ref.logN, 2))]
        exp.append(np.stack(e))
    plan = mk.Plan(ctx, ops, use_graph=True, scheme="rotate")
    assert plan.stages[0] == "conj_permute" and plan.stages[-1] == "permute"
    for rep in range(2):
        plan.launch()
        ctx.sync()
        for b in range(2):
            got = ctx.d2h(np.zeros((3, 1 + k, nl, N), np.uint64), ops[b]["out"])
            assert np.array_equal(got, exp[b]), (rep, b)
    plan.close()
    for p in bufs:
        ctx.free(p)


def test_kklss_rotate_errors(kk):
    name, ctx, ref = kk
    mk = _mk()
    ct = _ct(ref, 1, seed=113)
    crs, _, _ = _galois_keys(ctx, ref, [], [8], seed=114)
    with pytest.raises(mk.MkhipError) as e:  # no rotation key for party 5 (keys.go:151-160)
        ctx.kklss_rotate(ct, [5], 8)
    assert e.value.status == -6
    ctx.add_rotation_key(5, 8, crs[8])
    ctx.kklss_rotate(ct, [5], 8)
    ctx.del_rotation_key(5, 8)
    with pytest.raises(mk.MkhipError):
        ctx.kklss_rotate(ct, [5], 8)


def test_kklss_rotate_on_a_bfv_context():
    """mkbfv.Evaluator.RotateNew runs the same mkrlwe key switch over Q."""
    mk = _mk()
    logN, Q, QM, P, t = bfv_config("bsmall")
    ctx = mk.Context(logN, Q, P, gamma=2, t=t, qmul=QM)
    ref = O.Ctx(logN, Q, P, gamma=2)
    ids = [0, 1]
    crs, rk, _ = _galois_keys(ctx, ref, ids, [1, 2], seed=115)
    ct = _ct(ref, 2, seed=116)
    exp = ref.kklss_rotate_new(ct, ids, 3, rk, crs)
    assert np.array_equal(ctx.kklss_rotate(ct, ids, 3), exp)
    ctx.close()


# ---------------------------------------------------------------- MKF

@pytest.fixture(scope="module", params=["c1", "c2", "n15"])
def mkf(request):
    mk = _mk()
    logN, Q, P = config(request.param)
    ctx = mk.Context(logN, Q, P, gamma=1)
    ref = O.Ctx(logN, Q, P, gamma=1)
    yield request.param, ctx, ref
    ctx.close()


def _mkf_swk(ref, g):
    return np.stack([np.stack([g.uniform(ref.QP, ref.N), g.uniform(ref.QP, ref.N)]) for _ in range(ref.beta())])


def test_mkf_rotate_parity(mkf):
    name, ctx, ref = mkf
    k = 2 if name == "c1" else 4
    g = O.Prng(201)
    ctx.register_mkf_group(9, [_mkf_swk(ref, g) for _ in range(k)], _mkf_swk(ref, g))
    ct = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + k)])
    for rotation in (1, 5, -3):
        keys = [_mkf_swk(ref, g) for _ in range(k)]
        ctx.register_mkf_rotation(9, rotation, keys)
        o0, o1 = ref.mkf_ckks_rotate(ct, keys, O.galois_element(ref.logN, rotation))
        got = ctx.mkf_ckks_rotate(9, ct, rotation)
        assert np.array_equal(got[0], o0), rotation
        for p in range(k):
            assert np.array_equal(got[1 + p], o1), (rotation, p)
    mk = _mk()
    with pytest.raises(mk.MkhipError) as e:  # true_mkf_ckks.go:327-330
        ctx.mkf_ckks_rotate(9, ct, 2)
    assert e.value.status == -6
    ctx.unregister_mkf_group(9)


def test_mkf_rotation_plan(mkf):
    name, ctx, ref = mkf
    mk = _mk()
    k = 2 if name == "c1" else 3
    g = O.Prng(202)
    ctx.register_mkf_group(4, [_mkf_swk(ref, g) for _ in range(k)], _mkf_swk(ref, g))
    keys = [_mkf_swk(ref, g) for _ in range(k)]
    ctx.register_mkf_rotation(4, 6, keys)
    gal = O.galois_element(ref.logN, 6)
    nl, N = ref.L - 1, ref.N
    ops, bufs, exp = [], [], []
    for b in range(2):
        ct = np.stack([g.uniform(ref.Q[:nl], N) for _ in range(1 + k)])
        din, dout = ctx.alloc(ct.nbytes), ctx.alloc(ct.nbytes)
        ctx.h2d(din, ct)
        bufs += [din, dout]
        ops.append(dict(**{"in": din}, out=dout, group=4, level=nl - 1, rotation=6))
        o0, o1 = ref.mkf_ckks_rotate(ct, keys, gal)
        exp.append(np.stack([o0] + [o1] * k))
    plan = mk.Plan(ctx, ops, use_graph=True, scheme="mkf_rotate")
    for rep in range(2):
        plan.launch()
        ctx.sync()
        for b in range(2):
            got = ctx.d2h(np.zeros((1 + k, nl, N), np.uint64), ops[b]["out"])
            assert np.array_equal(got, exp[b]), (rep, b)
    plan.close()
    for p in bufs:
        ctx.free(p)
    ctx.unregister_mkf_group(4)
