"""GPU parity of key generation and public-key encryption with the CPU oracle
(bit-exact): secret, public, relinearization, rotation and conjugation keys
(mkrlwe/keygen.go:44-327) and Encryptor.Encrypt in both output domains
(encryptor.go:55-119).  Secrets, errors and v are seeded small polynomials; the
CRS entries uniform MForm residues.  The oracle's key equations are pinned by the
decrypt tests of test_oracle_keygen.py."""
import numpy as np
import pytest

import mkgen as G
import oracle as O
from params import config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["small", "c2", "n15"])
def kg(request):
    import mkf_amd

    logN, Q, P = config(request.param)
    ctx = mkf_amd.Context(logN, Q, P, gamma=2)
    ref = O.Ctx(logN, Q, P, gamma=2)
    yield request.param, ctx, ref, O.KeyGen(ref)
    ctx.close()


def _crs(ref, g):
    return np.stack([G.mform(g.uniform(ref.QP, ref.N), ref.QP) for _ in range(ref.beta())])


def test_secret_and_public_key(kg):
    name, ctx, ref, okg = kg
    g = O.Prng(301)
    s = g.ternary(ref.N)
    sk = ctx.gen_secret(s)
    assert np.array_equal(sk, okg.secret(s))
    a = _crs(ref, g)
    e = g.gaussian(ref.N)
    pk0, pk1 = ctx.gen_public_key(sk, e, a[0])
    e0, e1 = okg.public_key(sk, e, a[0])
    assert np.array_equal(pk0, e0) and np.array_equal(pk1, e1)


def test_relin_key_and_registration(kg):
    name, ctx, ref, okg = kg
    g = O.Prng(302)
    N = ref.N
    a, u = _crs(ref, g), _crs(ref, g)
    ctx.set_crs(0, a)
    ctx.set_crs_u(u)
    sks, rs, errs = [], [], []
    for _ in range(2):
        sks.append(okg.secret(g.ternary(N)))
        rs.append(okg.secret(g.ternary(N)))
        errs.append(np.stack([g.gaussian(N) for _ in range(3 * ref.beta())]))
    for p in range(2):
        got = ctx.gen_relin_key(sks[p], rs[p], errs[p])
        exp = okg.relin_key(sks[p], rs[p], errs[p], a, u)
        for x, y in zip(got, exp):
            assert np.array_equal(x, y), p
    # keys generated in HBM serve mul + relin like keys uploaded from the host
    ops = [np.stack([g.uniform(ref.Q, N) for _ in range(3)]) for _ in range(2)]
    scale = float(ref.Q[-1])
    for p in range(2):
        ctx.register_party(10 + p, *okg.relin_key(sks[p], rs[p], errs[p], a, u))
        ctx.register_party_keygen(20 + p, sks[p], rs[p], errs[p])
    want = ctx.kklss_ckks_mulrelin(ops[0], [10, 11], ops[1], [10, 11], scale, scale, scale)[0]
    got = ctx.kklss_ckks_mulrelin(ops[0], [20, 21], ops[1], [20, 21], scale, scale, scale)[0]
    assert np.array_equal(got, want)


def test_galois_keys(kg):
    name, ctx, ref, okg = kg
    g = O.Prng(303)
    N = ref.N
    sk = okg.secret(g.ternary(N))
    for r in (1, 5):
        a = _crs(ref, g)
        ctx.set_crs(r, a)
        errs = np.stack([g.gaussian(N) for _ in range(ref.beta())])
        assert np.array_equal(ctx.gen_rotation_key(r, sk, errs), okg.rotation_key(r, sk, errs, a)), r
    a = _crs(ref, g)
    ctx.set_crs(-2, a)
    errs = np.stack([g.gaussian(N) for _ in range(ref.beta())])
    assert np.array_equal(ctx.gen_conjugation_key(sk, errs), okg.conjugation_key(sk, errs, a))


@pytest.mark.parametrize("ntt_out,pt_ntt", [(False, False), (True, False), (False, True), (True, True)])
def test_encrypt(kg, ntt_out, pt_ntt):
    name, ctx, ref, okg = kg
    g = O.Prng(304)
    N = ref.N
    sk = okg.secret(g.ternary(N))
    pk0, pk1 = okg.public_key(sk, g.gaussian(N), _crs(ref, g)[0])
    pt = np.stack([g.uniform([q], N)[0] for q in ref.Q])
    v, e0, e1 = g.ternary(N), g.gaussian(N), g.gaussian(N)
    for level in (ref.L - 1, ref.L - 2):
        got = ctx.encrypt(pk0, pk1, v, e0, e1, pt, level=level, pt_ntt=pt_ntt, ntt_out=ntt_out)
        exp = okg.encrypt(pk0, pk1, v, e0, e1, pt, level=level, pt_ntt=pt_ntt, ntt_out=ntt_out)
        assert np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1]), level


@pytest.mark.parametrize("name", ["bsmall", "b3"])
def test_bfv_relin_key_and_registration(name):
    """KKLSS MK-BFV relinearization keys (mkbfv/keygen.go:24-162: both gadget
    branches of d, b over CRS[0] and CRS[-3], v) equal the oracle's, and keys
    generated and registered in HBM serve mul + relin like uploaded ones."""
    import mkf_amd

    from params import bfv_config

    logN, Q, QM, P, t = bfv_config(name)
    ctx = mkf_amd.Context(logN, Q, P, gamma=2, t=t, qmul=QM)
    ref = O.Ctx(logN, Q, P, gamma=2)
    okg = O.KeyGen(ref)
    try:
        g = O.Prng(311)
        N, beta = ref.N, ref.beta()
        a1, a2, u = _crs(ref, g), _crs(ref, g), _crs(ref, g)
        ctx.set_crs(0, a1)
        ctx.set_crs(-3, a2)
        ctx.set_crs_u(u)
        sks, rs, errs = [], [], []
        for p in range(2):
            sks.append(okg.secret(g.ternary(N)))
            rs.append(okg.secret(g.ternary(N)))
            errs.append(np.stack([np.stack([g.gaussian(N) for _ in range(beta)]) for _ in range(5)]))
            got = ctx.gen_bfv_relin_key(sks[p], rs[p], errs[p])
            exp = okg.bfv_relin_key(Q, QM, P, t, sks[p], rs[p], errs[p], a1, a2, u)
            for x, y in zip(got, exp):
                assert np.array_equal(x, y), (name, p)
            ctx.register_bfv_party(10 + p, *exp)
            ctx.register_bfv_party_keygen(20 + p, sks[p], rs[p], errs[p])
        ops = [np.stack([g.uniform(Q, N) for _ in range(3)]) for _ in range(2)]
        want = ctx.kklss_bfv_mulrelin(ops[0], [10, 11], ops[1], [10, 11])[0]
        got = ctx.kklss_bfv_mulrelin(ops[0], [20, 21], ops[1], [20, 21])[0]
        assert np.array_equal(got, want)
    finally:
        ctx.close()


@pytest.mark.parametrize("name", ["small", "c2"])
def test_mkf_keys_and_group_registration(name):
    """MKF keys (mkf_cpu/common.go:32-51 over Lattigo GenSwitchingKey /
    GenRelinearizationKey(s, 1)) equal the oracle's; a group generated and
    registered in HBM (aggregate secret, fold keys, relinearization key) serves
    MKF mul + relin like the same keys uploaded."""
    import mkf_amd

    logN, Q, P = config(name)
    ctx = mkf_amd.Context(logN, Q, P, gamma=1)
    ref = O.Ctx(logN, Q, P, gamma=1)
    okg = O.KeyGen(ref)
    try:
        g = O.Prng(312)
        N, QP, beta, k = ref.N, ref.QP, ref.beta(), 3
        sks = np.stack([okg.secret(g.ternary(N)) for _ in range(k)])
        a = np.stack([np.stack([g.uniform(QP, N) for _ in range(beta)]) for _ in range(k + 1)])
        errs = np.stack([np.stack([g.gaussian(N) for _ in range(beta)]) for _ in range(k + 1)])
        assert np.array_equal(ctx.gen_mkf_switching_key(sks[0], sks[1], a[0], errs[0]),
                              okg.mkf_switching_key(sks[0], sks[1], a[0], errs[0]))
        assert np.array_equal(ctx.gen_mkf_relin_key(sks[2], a[1], errs[1]), okg.mkf_relin_key(sks[2], a[1], errs[1]))
        fold, relin = okg.mkf_group_keys(sks, a, errs)
        ctx.register_mkf_group(1, fold, relin)
        ctx.register_mkf_group_keygen(2, sks, a, errs)
        left = np.stack([g.uniform(Q, N) for _ in range(1 + k)])
        right = np.stack([g.uniform(Q, N) for _ in range(1 + k)])
        want = ctx.mkf_ckks_mulrelin(1, left, right)[0]
        got = ctx.mkf_ckks_mulrelin(2, left, right)[0]
        assert np.array_equal(got, want)
    finally:
        ctx.close()
