"""GPU parity of the MKF MK-CKKS path (software/mkf/mkf_cpu/ckks.go:102-151)
against the CPU oracle, bit-exact: fold (one key switch per party, common.go:79-126),
Karatsuba tensor, relinearization, NTT-domain rescale and the k-fold
materialization of Gamma.  Keys and ciphertexts are uniform residues."""
import numpy as np
import pytest

import oracle as O
from params import config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["c1", "c2", "n15"])
def setup(request):
    import mkf_amd

    logN, Q, P = config(request.param)
    ctx = mkf_amd.Context(logN, Q, P, gamma=1)  # Lattigo rlwe decomposition: alpha = K
    ref = O.Ctx(logN, Q, P, gamma=1)
    yield request.param, ctx, ref
    ctx.close()


def _keys(ref, k, seed):
    g = O.Prng(seed)
    beta = ref.beta()

    def swk():
        return np.stack([np.stack([g.uniform(ref.QP, ref.N), g.uniform(ref.QP, ref.N)]) for _ in range(beta)])

    return [swk() for _ in range(k)], swk()


def test_crt_decomposition_parity(setup):
    """alpha = K digits (CRT with the float64 estimate when K > 1)."""
    name, ctx, ref = setup
    g = O.Prng(41)
    a = g.uniform(ref.Q, ref.N)
    for level in (ref.L - 1, ref.L - 2):
        assert np.array_equal(ctx.decompose(a[:level + 1], level), ref.decompose(a[:level + 1], level)), level


def test_mkf_ckks_mulrelin_parity(setup):
    name, ctx, ref = setup
    k = 2 if name == "c1" else 4
    fold, relin = _keys(ref, k, seed=42)
    ctx.register_mkf_group(7, fold, relin)
    g = O.Prng(43)
    left = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + k)])
    right = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + k)])
    scale = float(ref.Q[-1])
    out, level, _ = ctx.mkf_ckks_mulrelin(7, left, right, scale, scale, scale)
    assert level == ref.L - 2
    o0, o1 = ref.mkf_ckks_mulrelin(left, right, fold, relin)
    r0 = ref.div_round_last(o0, 1, ntt=True)
    r1 = ref.div_round_last(o1, 1, ntt=True)
    assert np.array_equal(out[0], r0)
    for p in range(k):
        assert np.array_equal(out[1 + p], r1), p
    ctx.unregister_mkf_group(7)


def test_mkf_batched_plan_matches_oracle(setup):
    import mkf_amd

    name, ctx, ref = setup
    k = 2 if name == "c1" else 4
    fold, relin = _keys(ref, k, seed=44)
    ctx.register_mkf_group(3, fold, relin)
    g = O.Prng(45)
    nl, N, B = ref.L, ref.N, 3
    ops, bufs, exp = [], [], []
    for _ in range(B):
        left = np.stack([g.uniform(ref.Q, N) for _ in range(1 + k)])
        right = np.stack([g.uniform(ref.Q, N) for _ in range(1 + k)])
        dl, dr, do = ctx.alloc(left.nbytes), ctx.alloc(right.nbytes), ctx.alloc((1 + k) * (nl - 1) * N * 8)
        ctx.h2d(dl, left)
        ctx.h2d(dr, right)
        bufs += [dl, dr, do]
        ops.append(dict(left=dl, right=dr, out=do, group=3, level=nl - 1, nres=1))
        o0, o1 = ref.mkf_ckks_mulrelin(left, right, fold, relin)
        r0, r1 = ref.div_round_last(o0, 1, ntt=True), ref.div_round_last(o1, 1, ntt=True)
        exp.append(np.stack([r0] + [r1] * k))
    plan = mkf_amd.Plan(ctx, ops, use_graph=True, scheme="mkf")
    for rep in range(2):
        plan.launch()
        ctx.sync()
        for b in range(B):
            got = ctx.d2h(np.zeros((1 + k, nl - 1, N), np.uint64), ops[b]["out"])
            assert np.array_equal(got, exp[b]), (rep, b)
    plan.close()
    for p in bufs:
        ctx.free(p)
    ctx.unregister_mkf_group(3)


def test_digit_pass_matches_oracle(setup):
    """With the digit-only part of the CRT ModUp computed once per digit
    (k_digit_pre), mul+relin and a product of operands at different levels keep
    the oracle's words."""
    import mkf_amd

    name, ctx, ref = setup
    k = 2
    fold, relin = _keys(ref, k, seed=47)
    ctx.register_mkf_group(11, fold, relin)
    g = O.Prng(48)
    left = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + k)])
    right = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + k)])
    scale = float(ref.Q[-1])
    o0, o1 = ref.mkf_ckks_mulrelin(left, right, fold, relin)
    p0, p1 = ref.mkf_ckks_mulrelin2(left, right[:, :ref.L - 1].copy(), fold, relin)
    ctx.set_tuning(mkf_amd.TUNE_DIGIT_PRE, 1)
    try:
        out, level, _ = ctx.mkf_ckks_mulrelin(11, left, right, scale, scale, scale)
        lv = ctx.mkf_mulrelin_levels(11, left, right[:, :ref.L - 1])
    finally:
        ctx.set_tuning(mkf_amd.TUNE_DIGIT_PRE, 0)
    assert np.array_equal(out[0], ref.div_round_last(o0, 1, ntt=True))
    assert np.array_equal(out[1], ref.div_round_last(o1, 1, ntt=True))
    assert np.array_equal(lv[0], p0) and np.array_equal(lv[1], p1)
    ctx.unregister_mkf_group(11)
