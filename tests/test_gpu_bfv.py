"""GPU parity of the KKLSS MK-BFV path (software/kklss/mkbfv) against the CPU
oracle, bit-exact: the R-basis lifts (ModUpQtoR with multSum's unreduced
limbs, Rescale), Quantize, the two-branch decomposition, mul+relin through the
host-pointer entry point and batched graph plans.  Keys and ciphertexts are
uniform residues except in the decrypt test."""
import numpy as np
import pytest

import mkgen as G
import oracle as O
from params import bfv_config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["bsmall", "b3", "b15"])
def setup(request):
    import mkf_amd

    logN, Q, QM, P, t = bfv_config(request.param)
    ctx = mkf_amd.Context(logN, Q, P, gamma=2, t=t, qmul=QM)
    ref = O.BfvCtx(logN, Q, QM, P, t)
    yield request.param, ctx, ref
    ctx.close()


def _keys(ref, n, seed):
    g = O.Prng(seed)
    L, QP, N = ref.L, ref.QP, ref.N
    b = [np.stack([g.uniform(QP, N) for _ in range(2 * L)]) for _ in range(n)]
    d = [np.stack([g.uniform(QP, N) for _ in range(2 * L)]) for _ in range(n)]
    v = [np.stack([g.uniform(QP, N) for _ in range(L)]) for _ in range(n)]
    u = np.stack([g.uniform(QP, N) for _ in range(L)])
    return b, d, v, u


def test_lifts_match_oracle_including_raw_limbs(setup):
    name, ctx, ref = setup
    g = O.Prng(31)
    x = g.uniform(ref.Q, ref.N)
    assert np.array_equal(ctx.bfv_modup_q_to_r(x), ref.modup_q_to_r(x))
    assert np.array_equal(ctx.bfv_rescale_q_to_r(x), ref.rescale_q_to_r(x))


def test_quantize_and_decompose_match_oracle(setup):
    name, ctx, ref = setup
    g = O.Prng(32)
    xr = g.uniform(ref.R, ref.N)
    assert np.array_equal(ctx.bfv_quantize(xr), ref.quantize(xr))
    aR = ref.modup_q_to_r(g.uniform(ref.Q, ref.N))      # carries raw (>= q) limbs
    assert np.array_equal(ctx.bfv_decompose(aR), ref.decompose(aR))


@pytest.mark.parametrize("shape", [([0, 1], [1, 2]), ([0], [1]), ([0, 1, 2, 3], [0, 1, 2, 3])])
def test_bfv_mulrelin_parity(setup, shape):
    name, ctx, ref = setup
    ids0, ids1 = shape
    if name in ("b3", "b15") and shape[0] != [0, 1, 2, 3]:
        pytest.skip("party-set variants run at the small shape")
    b, d, v, u = _keys(ref, 4, seed=33)
    for pid in range(4):
        ctx.register_bfv_party(pid, b[pid], d[pid], v[pid])
    ctx.set_crs_u(u)
    g = O.Prng(34)
    op0 = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + len(ids0))])
    op1 = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + len(ids1))])
    out, ids = ctx.kklss_bfv_mulrelin(op0, ids0, op1, ids1)
    exp, eids = ref.mulrelin(op0, ids0, op1, ids1, b, d, v, u)
    assert ids == eids
    for c in range(exp.shape[0]):
        assert np.array_equal(out[c], exp[c]), c


def test_bfv_split_limb_matches_oracle(setup):
    """The half-limb decomposition workgroups (N = 2^14 tuning knob; other sizes
    ignore it) give the oracle's words for the two-branch BFV mul+relin."""
    import mkf_amd

    name, ctx, ref = setup
    b, d, v, u = _keys(ref, 3, seed=35)
    for pid in range(3):
        ctx.register_bfv_party(pid, b[pid], d[pid], v[pid])
    ctx.set_crs_u(u)
    g = O.Prng(36)
    op0 = np.stack([g.uniform(ref.Q, ref.N) for _ in range(3)])
    op1 = np.stack([g.uniform(ref.Q, ref.N) for _ in range(3)])
    exp, eids = ref.mulrelin(op0, [0, 1], op1, [1, 2], b, d, v, u)
    ctx.set_tuning(mkf_amd.TUNE_SPLIT_LIMB, 1)
    try:
        out, ids = ctx.kklss_bfv_mulrelin(op0, [0, 1], op1, [1, 2])
    finally:
        ctx.set_tuning(mkf_amd.TUNE_SPLIT_LIMB, 0)
    assert ids == eids and np.array_equal(out, exp)


def test_bfv_batched_plan_matches_oracle(setup):
    import mkf_amd

    name, ctx, ref = setup
    b, d, v, u = _keys(ref, 3, seed=35)
    for pid in range(3):
        ctx.register_bfv_party(pid, b[pid], d[pid], v[pid])
    ctx.set_crs_u(u)
    g = O.Prng(36)
    L, N, B = ref.L, ref.N, 3
    shapes = [([0, 1], [1, 2]), ([2], [0, 1]), ([0, 1, 2], [0, 1, 2])]
    ops, bufs, exp = [], [], []
    for ids0, ids1 in shapes[:B]:
        op0 = np.stack([g.uniform(ref.Q, N) for _ in range(1 + len(ids0))])
        op1 = np.stack([g.uniform(ref.Q, N) for _ in range(1 + len(ids1))])
        ku = len(set(ids0) | set(ids1))
        d0, d1, do = ctx.alloc(op0.nbytes), ctx.alloc(op1.nbytes), ctx.alloc((1 + ku) * L * N * 8)
        ctx.h2d(d0, op0)
        ctx.h2d(d1, op1)
        bufs += [d0, d1, do]
        ops.append(dict(op0=d0, op1=d1, out=do, ids0=ids0, ids1=ids1, level=L - 1, nres=0))
        exp.append(ref.mulrelin(op0, ids0, op1, ids1, b, d, v, u)[0])
    plan = mkf_amd.Plan(ctx, ops, use_graph=True, scheme="bfv")
    for rep in range(3):
        plan.launch()
        ctx.sync()
        for i in range(B):
            got = ctx.d2h(np.zeros(exp[i].shape, np.uint64), ops[i]["out"])
            assert np.array_equal(got, exp[i]), (rep, i)
        ctx.h2d(bufs[2], np.zeros(exp[0].shape, np.uint64))  # eager work between replays
    plan.close()
    for p in bufs:
        ctx.free(p)


def test_bfv_gpu_mulrelin_decrypts_exactly():
    """With real keys the GPU product decrypts to m0 m1 mod t (mkbfv_test.go:407-413)."""
    import mkf_amd

    logN, Q, QM, P, t = bfv_config("bsmall")
    bctx = O.BfvCtx(logN, Q, QM, P, t)
    keys = G.BFVKeys(bctx, 2, seed=37)
    ctx = mkf_amd.Context(logN, Q, P, gamma=2, t=t, qmul=QM)
    for pid in range(2):
        ctx.register_bfv_party(pid, keys.bkey[pid], keys.dkey[pid], keys.vkey[pid])
    ctx.set_crs_u(keys.u)
    g = O.Prng(38)
    N = bctx.N
    m0 = [int(x) for x in g.uniform([t], N)[0]]
    m1 = [int(x) for x in g.uniform([t], N)[0]]
    c0 = np.stack(keys.encrypt(0, keys.encode(m0), g))
    c1 = np.stack(keys.encrypt(1, keys.encode(m1), g))
    out, ids = ctx.kklss_bfv_mulrelin(c0, [0], c1, [1])
    ctx.close()
    assert ids == [0, 1]
    exp = [v % t for v in G.negacyclic_mul(m0, m1)]
    assert keys.decode(out, ids) == exp


# ---------------------------------------------------------------- MKF MK-BFV

@pytest.fixture(scope="module", params=["bsmall", "b3", "b15"])
def mkf_setup(request):
    import mkf_amd

    logN, Q, QM, P, t = bfv_config(request.param)
    ctx = mkf_amd.Context(logN, Q, P, gamma=1, t=t, qmul=QM)  # MKF: Lattigo rlwe, alpha = K
    ref = O.BfvCtx(logN, Q, QM, P, t, gamma=1)
    yield request.param, ctx, ref
    ctx.close()


def _mkf_keys(ref, k, seed):
    g = O.Prng(seed)
    beta = ref.qp.beta()

    def swk():
        return np.stack([np.stack([g.uniform(ref.QP, ref.N), g.uniform(ref.QP, ref.N)]) for _ in range(beta)])

    return [swk() for _ in range(k)], swk()


def test_mkf_bfv_mulrelin_parity(mkf_setup):
    name, ctx, ref = mkf_setup
    k = 2 if name == "bsmall" else 4
    fold, relin = _mkf_keys(ref, k, seed=51)
    ctx.register_mkf_group(5, fold, relin)
    g = O.Prng(52)
    left = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + k)])
    right = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + k)])
    out = ctx.mkf_bfv_mulrelin(5, left, right)
    c0, gamma = ref.mkf_mulrelin(left, right, fold, relin)
    assert np.array_equal(out[0], c0)
    for p in range(k):
        assert np.array_equal(out[1 + p], gamma), p
    ctx.unregister_mkf_group(5)


def test_mkf_bfv_batched_plan_matches_oracle(mkf_setup):
    import mkf_amd

    name, ctx, ref = mkf_setup
    k = 3
    fold, relin = _mkf_keys(ref, k, seed=53)
    ctx.register_mkf_group(6, fold, relin)
    g = O.Prng(54)
    L, N, B = ref.L, ref.N, 2
    ops, bufs, exp = [], [], []
    for _ in range(B):
        left = np.stack([g.uniform(ref.Q, N) for _ in range(1 + k)])
        right = np.stack([g.uniform(ref.Q, N) for _ in range(1 + k)])
        dl, dr, do = ctx.alloc(left.nbytes), ctx.alloc(right.nbytes), ctx.alloc(left.nbytes)
        ctx.h2d(dl, left)
        ctx.h2d(dr, right)
        bufs += [dl, dr, do]
        ops.append(dict(left=dl, right=dr, out=do, group=6, level=L - 1, nres=0))
        c0, gm = ref.mkf_mulrelin(left, right, fold, relin)
        exp.append(np.stack([c0] + [gm] * k))
    plan = mkf_amd.Plan(ctx, ops, use_graph=True, scheme="mkf_bfv")
    for rep in range(3):
        plan.launch()
        ctx.sync()
        for b in range(B):
            got = ctx.d2h(np.zeros((1 + k, L, N), np.uint64), ops[b]["out"])
            assert np.array_equal(got, exp[b]), (rep, b)
        ctx.h2d(ops[0]["out"], np.zeros((1 + k, L, N), np.uint64))  # eager work between replays
    plan.close()
    for p in bufs:
        ctx.free(p)
    ctx.unregister_mkf_group(6)
