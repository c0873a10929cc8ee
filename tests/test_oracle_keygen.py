"""Pin the key-generation / encryption part of the CPU oracle (mkref_keygen.c).

* Closed forms: the secret is MForm(NTT(s)); the public key satisfies
  pk0 + MRed(sk, pk1) = NTT(e) (keygen.go:88-110); a switching key's digit i holds
  MForm(NTT(e_i)) plus P s on the digit's Q limbs only (keygen.go:270-327).
* Decrypt-and-compare: an encryption decrypts to the plaintext plus the encryption
  noise (encryptor.go:55-119, decryptor.go:26-66), in both output domains; relin
  keys made by GenRelinearizationKey drive the KKLSS mul + relin to the product
  (mkckks_test.go:352-358's check); rotation and conjugation keys drive the
  Galois key switches to sigma(m) (mkckks_test.go:443-505).  Planted mistakes
  (swapped secrets) fail.
"""
import numpy as np

import mkgen as G
import oracle as O


def _setup(logN=9, L=3, K=2, seed=7):
    Q, P = O.gen_primes(50, 15, L), O.gen_primes(60, 15, K)
    ctx = O.Ctx(logN, Q, P, gamma=2)
    return ctx, O.KeyGen(ctx), O.Prng(seed)


def _crs(ctx, g):
    QP = ctx.QP
    return np.stack([G.mform(g.uniform(QP, ctx.N), QP) for _ in range(ctx.beta())])


def _errs(g, n, N):
    return np.stack([g.gaussian(N) for _ in range(n)])


def _decrypt(ctx, c0, c1_by_party, sks):
    """c0 + sum_i c_i s_i: NTT, MRed with MForm(NTT(s)), INTT (decryptor.go:26-43)."""
    nl = c0.shape[0]
    Q = ctx.Q[:nl]
    acc = c0.copy()
    for c1, sk in zip(c1_by_party, sks):
        prod = np.stack([np.array([O.mred(int(x), int(y), Q[j]) for x, y in zip(ctx.ntt(c1[j], j), sk[j])],
                                  np.uint64) for j in range(nl)])
        acc = G.addmod(acc, ctx.ntt_poly(prod, inverse=True), Q)
    return acc


def test_secret_and_public_key_closed_forms():
    ctx, kg, g = _setup()
    s = g.ternary(ctx.N)
    sk = kg.secret(s)
    QP = ctx.QP
    assert np.array_equal(ctx.ntt_poly(G.inv_mform(sk, QP), inverse=True), O.small_to_rns(s, QP))
    a = _crs(ctx, g)
    e = g.gaussian(ctx.N)
    pk0, pk1 = kg.public_key(sk, e, a[0])
    assert np.array_equal(pk1, a[0])
    mred = np.stack([np.array([O.mred(int(x), int(y), QP[r]) for x, y in zip(sk[r], pk1[r])], np.uint64)
                     for r in range(len(QP))])
    assert np.array_equal(G.addmod(pk0, mred, QP), ctx.ntt_poly(O.small_to_rns(e, QP)))


def test_switching_key_gadget_rows():
    ctx, kg, g = _setup()
    QP, N = ctx.QP, ctx.N
    sk = kg.secret(g.ternary(N))
    errs = _errs(g, ctx.beta(), N)
    swk = kg.switching_key(sk, errs)
    Pprod = 1
    for p in ctx.P:
        Pprod *= p
    for i in range(ctx.beta()):
        base = G.mform(ctx.ntt_poly(O.small_to_rns(errs[i], QP)), QP)
        diff = G.submod(swk[i], base, QP)
        for r in range(len(QP)):
            want = [int(x) * (Pprod % QP[r]) % QP[r] for x in sk[r]] if r == i else [0] * N
            assert [int(x) for x in diff[r]] == want, (i, r)


def test_encryption_decrypts_to_the_plaintext():
    ctx, kg, g = _setup(seed=8)
    QP, Q, N = ctx.QP, ctx.Q, ctx.N
    s = g.ternary(N)
    sk = kg.secret(s)
    a = _crs(ctx, g)
    pk0, pk1 = kg.public_key(sk, g.gaussian(N), a[0])
    m = (np.array([int(x) for x in g.uniform([1 << 20], N)[0]]) - (1 << 19)) << 20
    pt = O.small_to_rns(m, Q)
    v, e0, e1 = g.ternary(N), g.gaussian(N), g.gaussian(N)
    c0, c1 = kg.encrypt(pk0, pk1, v, e0, e1, pt)
    got = G.centered_crt(_decrypt(ctx, c0, [c1], [sk]), Q)
    assert max(abs(a_ - int(b_)) for a_, b_ in zip(got, m)) < 2 ** 16
    # the NTT-domain ciphertext is the transform of the same encryption
    n0, n1 = kg.encrypt(pk0, pk1, v, e0, e1, pt, ntt_out=True)
    assert np.array_equal(n0, ctx.ntt_poly(c0)) and np.array_equal(n1, ctx.ntt_poly(c1))
    # and a plaintext given in the NTT domain
    p0, p1 = kg.encrypt(pk0, pk1, v, e0, e1, ctx.ntt_poly(pt), pt_ntt=True)
    assert np.array_equal(p0, c0) and np.array_equal(p1, c1)
    # planted mistake: decrypting with another secret fails
    bad = G.centered_crt(_decrypt(ctx, c0, [c1], [kg.secret(g.ternary(N))]), Q)
    assert max(abs(a_ - int(b_)) for a_, b_ in zip(bad, m)) > 2 ** 100


def test_generated_relin_keys_drive_mulrelin_to_the_product():
    ctx, kg, g = _setup(logN=10, L=4, seed=9)
    QP, Q, N = ctx.QP, ctx.Q, ctx.N
    a, u = _crs(ctx, g), _crs(ctx, g)
    sks, pks, rlk = [], [], []
    for _ in range(2):
        s, r = g.ternary(N), g.ternary(N)
        sk, rk = kg.secret(s), kg.secret(r)
        sks.append(sk)
        pks.append(kg.public_key(sk, g.gaussian(N), a[0]))
        rlk.append(kg.relin_key(sk, rk, _errs(g, 3 * ctx.beta(), N), a, u))
    D = 1 << 40

    def small():
        return (np.array([int(v) for v in g.uniform([15], N)[0]]) - 7).astype(np.int64)

    xs, ys = [small(), small()], [small(), small()]

    def ct(ms):
        parts = [kg.encrypt(pks[p][0], pks[p][1], g.ternary(N), g.gaussian(N), g.gaussian(N),
                            O.small_to_rns(m * D, Q)) for p, m in enumerate(ms)]
        return np.stack([G.addmod(parts[0][0], parts[1][0], Q), parts[0][1], parts[1][1]])

    op0, op1 = ct(xs), ct(ys)
    out, ids = ctx.kklss_mulrelin(op0, [0, 1], op1, [0, 1], [k[0] for k in rlk], [k[1] for k in rlk],
                                  [k[2] for k in rlk], u)
    got = G.centered_crt(_decrypt(ctx, out[0], [out[1], out[2]], sks), Q)
    X = [int(v) * D for v in xs[0] + xs[1]]
    Y = [int(v) * D for v in ys[0] + ys[1]]
    exp = G.negacyclic_mul(X, Y)
    assert max(abs(a_ - b_) for a_, b_ in zip(got, exp)) < 2 ** 62  # product ~2^99
    # planted mistake: the parties' relinearization keys swapped
    bad, _ = ctx.kklss_mulrelin(op0, [0, 1], op1, [0, 1], [rlk[1][0], rlk[0][0]], [rlk[1][1], rlk[0][1]],
                                [rlk[1][2], rlk[0][2]], u)
    badv = G.centered_crt(_decrypt(ctx, bad[0], [bad[1], bad[2]], sks), Q)
    assert max(abs(a_ - b_) for a_, b_ in zip(badv, exp)) > 2 ** 120


def test_generated_galois_keys_rotate_and_conjugate():
    ctx, kg, g = _setup(seed=10)
    QP, Q, N = ctx.QP, ctx.Q, ctx.N
    a0 = _crs(ctx, g)
    crs = {1: _crs(ctx, g), 2: _crs(ctx, g), -2: _crs(ctx, g)}
    sks, pks = [], []
    for _ in range(2):
        sk = kg.secret(g.ternary(N))
        sks.append(sk)
        pks.append(kg.public_key(sk, g.gaussian(N), a0[0]))
    m = [((np.array([int(x) for x in g.uniform([15], N)[0]]) - 7) << 30) for _ in range(2)]
    parts = [kg.encrypt(pks[p][0], pks[p][1], g.ternary(N), g.gaussian(N), g.gaussian(N), O.small_to_rns(m[p], Q))
             for p in range(2)]
    ct = np.stack([G.addmod(parts[0][0], parts[1][0], Q), parts[0][1], parts[1][1]])
    msum = m[0] + m[1]
    for r in (1, 2):
        rks = [kg.rotation_key(r, sk, _errs(g, ctx.beta(), N), crs[r]) for sk in sks]
        gal = O.galois_element(ctx.logN, r)
        out = ctx.kklss_rotate(ct, rks, crs[r], gal)
        got = G.centered_crt(_decrypt(ctx, out[0], [out[1], out[2]], sks), Q)
        assert max(abs(a_ - int(b_)) for a_, b_ in zip(got, G.galois_small(msum, gal))) < 2 ** 20
    cks = [kg.conjugation_key(sk, _errs(g, ctx.beta(), N), crs[-2]) for sk in sks]
    out = ctx.kklss_conjugate(ct, cks, crs[-2])
    got = G.centered_crt(_decrypt(ctx, out[0], [out[1], out[2]], sks), Q)
    gal = O.galois_conjugate(ctx.logN)
    assert max(abs(a_ - int(b_)) for a_, b_ in zip(got, G.galois_small(msum, gal))) < 2 ** 20
