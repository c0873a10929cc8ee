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


# ---------------------------------------------------------------- MK-BFV and MKF keys

def _bfv_ctx(logN=9, L=3, seed=0, parties=2):
    import math  # noqa: F401

    Q = O.gen_primes(50, 15, L)
    QM = O.gen_primes(50, 15, L, exclude=Q)
    P = O.gen_primes(60, 15, 2)
    bctx = O.BfvCtx(logN, Q, QM, P, 65537)
    return bctx, G.BFVKeys(bctx, parties, seed=seed)


def test_bfv_gadget_closed_forms():
    """GenBFVSwitchingKey's gadget (mkbfv/keygen.go:91-162) with one prime per digit,
    from exact integers, equals the closed forms the library evaluates in 64-bit
    arithmetic: branch 1 is t P QMul^-1 mod q_i on row i and 0 elsewhere (Q QMul / q_i
    is divisible by QMul); branch 2 is -(Q t T' P mod m) m^-1 mod q_r on every row
    (m = qm_i, T' = (Q QMul / m)^-1 mod m; Q t T' P vanishes mod every prime of Q u P)."""
    import math

    bctx, _ = _bfv_ctx(logN=4, parties=0)
    ctx = bctx.qp
    g1, g2 = O.KeyGen(ctx).bfv_gadget(bctx.Q, bctx.QMul, bctx.P, bctx.t)
    QP = [int(q) for q in ctx.QP]
    Qp, QMp, Pp, t = math.prod(bctx.Q), math.prod(bctx.QMul), math.prod(bctx.P), bctx.t
    for i, qi in enumerate(bctx.Q):
        for r, q in enumerate(QP):
            want = (t * Pp * pow(QMp, -1, qi)) % qi if r == i else 0
            assert int(g1[i][r]) == want, (i, r)
    for i, m in enumerate(bctx.QMul):
        tp = pow(Qp * (QMp // m), -1, m)
        ym = (Qp * t * tp * Pp) % m
        for r, q in enumerate(QP):
            assert int(g2[i][r]) == ((-ym) * pow(m, -1, q)) % q, (i, r)


def _bfv_key_restated(ctx, G_br, s_plain, r_plain, e_b, e_d, a_m):
    """b_br, d_br of one branch from the key equations, with explicit errors."""
    QP = ctx.QP
    b, d = [], []
    for i in range(ctx.beta()):
        a = G.inv_mform(a_m[i], QP)
        eb = ctx.ntt_poly(O.small_to_rns(e_b[i], QP))
        b.append(G.mform(G.submod(eb, G.mulmod(a, s_plain, QP), QP), QP))
        ed = ctx.ntt_poly(O.small_to_rns(e_d[i], QP))
        gs = np.stack([np.array((s_plain[j].astype(object) * (G_br[i] % q)) % q, np.uint64)
                       for j, q in enumerate(QP)])
        d.append(G.mform(G.submod(G.addmod(gs, ed, QP), G.mulmod(a, r_plain, QP), QP), QP))
    return np.stack(b), np.stack(d)


def test_bfv_relin_key_matches_the_key_equations():
    """The oracle's MK-BFV keys equal the key equations evaluated with exact
    integers (BFVKeys' restatement, explicit errors), v included."""
    bctx, keys = _bfv_ctx(logN=8, parties=1)
    ctx = bctx.qp
    okg = O.KeyGen(ctx)
    g = O.Prng(71)
    N, beta = ctx.N, ctx.beta()
    sk, r = okg.secret(g.ternary(N)), okg.secret(g.ternary(N))
    errs = np.stack([_errs(g, beta, N) for _ in range(5)])
    b, d, v = okg.bfv_relin_key(bctx.Q, bctx.QMul, bctx.P, bctx.t, sk, r, errs, keys.a, keys.a2, keys.u)
    QP = ctx.QP
    s_plain, r_plain = G.inv_mform(sk, QP), G.inv_mform(r, QP)
    b1, d1 = _bfv_key_restated(ctx, keys.G1, s_plain, r_plain, errs[0], errs[2], keys.a)
    b2, d2 = _bfv_key_restated(ctx, keys.G2, s_plain, r_plain, errs[1], errs[3], keys.a2)
    assert np.array_equal(b, np.concatenate([b1, b2])) and np.array_equal(d, np.concatenate([d1, d2]))
    # v = -(swk(r) + MRed(u, s)): the mkrlwe equation (relin_key's v with the same errors)
    vk = okg.relin_key(sk, r, np.concatenate([errs[0], errs[2], errs[4]]), keys.a, keys.u)[2]
    assert np.array_equal(v, vk)


def test_bfv_keys_from_the_oracle_decrypt_exactly():
    """MK-BFV relinearization keys made by the oracle's key generator (fresh r and
    errors, the parties' secrets and CRS) drive mul + relin to m0 m1 mod t exactly
    (mkbfv_test.go:407-413); the branches swapped fail."""
    bctx, keys = _bfv_ctx(logN=9, L=3, seed=3, parties=2)
    ctx = bctx.qp
    okg = O.KeyGen(ctx)
    g = O.Prng(72)
    N, t, beta = ctx.N, bctx.t, ctx.beta()
    for p in range(2):
        r = okg.secret(g.ternary(N))
        errs = np.stack([_errs(g, beta, N) for _ in range(5)])
        keys.bkey[p], keys.dkey[p], keys.vkey[p] = okg.bfv_relin_key(bctx.Q, bctx.QMul, bctx.P, t, keys.sk[p], r,
                                                                     errs, keys.a, keys.a2, keys.u)
    m0 = [int(v) for v in g.uniform([t], N)[0]]
    m1 = [int(v) for v in g.uniform([t], N)[0]]
    c0 = keys.encrypt(0, keys.encode(m0), g)
    c1 = keys.encrypt(1, keys.encode(m1), g)
    out, ids = bctx.mulrelin(np.stack(c0), [0], np.stack(c1), [1], keys.bkey, keys.dkey, keys.vkey, keys.u)
    exp = [v % t for v in G.negacyclic_mul(m0, m1)]
    assert keys.decode(out, ids) == exp
    swapped = [np.concatenate([x[beta:], x[:beta]]) for x in keys.dkey]
    out2, _ = bctx.mulrelin(np.stack(c0), [0], np.stack(c1), [1], keys.bkey, swapped, keys.vkey, keys.u)
    assert keys.decode(out2, ids) != exp


def test_mkf_switching_key_matches_the_key_equation():
    """Lattigo GenSwitchingKey as MKF uses it: ([MForm(e + P s_in g_b) - a s_out], [a])
    with a taken as stored (MKFKeys' restatement with explicit a and e)."""
    logN = 8
    Q, P = O.gen_primes(50, 15, 4), O.gen_primes(60, 15, 2)
    ctx = O.Ctx(logN, Q, P, gamma=1)
    okg = O.KeyGen(ctx)
    g = O.Prng(73)
    N, QP, beta = ctx.N, ctx.QP, ctx.beta()
    s_in, s_out = okg.secret(g.ternary(N)), okg.secret(g.ternary(N))
    a = np.stack([g.uniform(QP, N) for _ in range(beta)])
    errs = _errs(g, beta, N)
    swk = okg.mkf_switching_key(s_in, s_out, a, errs)
    Pprod = int(np.prod([int(p) for p in P], dtype=object))
    si, so = G.inv_mform(s_in, QP), G.inv_mform(s_out, QP)
    for b in range(beta):
        e = ctx.ntt_poly(O.small_to_rns(errs[b], QP))
        v0 = G.submod(e, G.mulmod(G.inv_mform(a[b], QP), so, QP), QP)
        for j in range(ctx.alpha):
            idx = b * ctx.alpha + j
            if idx < ctx.L:
                q = QP[idx]
                v0[idx] = np.array((v0[idx].astype(object) + si[idx].astype(object) * (Pprod % q)) % q, np.uint64)
        assert np.array_equal(swk[b][0], G.mform(v0, QP)), b
        assert np.array_equal(swk[b][1], a[b]), b


def test_mkf_keys_from_the_oracle_decrypt():
    """Fold and relinearization keys made by the oracle (generateOnlineKeys over the
    parties' secrets) drive MKF CKKSCase.Run to the product under the aggregate
    secret (ckks.go:155-165's check as a coefficient bound); a relinearization key
    of another secret fails."""
    logN = 10
    Q, P = O.gen_primes(50, 15, 4), O.gen_primes(60, 15, 2)
    ctx = O.Ctx(logN, Q, P, gamma=1)
    k = 3
    keys = G.MKFKeys(ctx, k, seed=75)
    okg = O.KeyGen(ctx)
    g = O.Prng(76)
    N, QP, beta = ctx.N, ctx.QP, ctx.beta()
    sks = np.stack([okg.secret(x) for x in keys.small])
    a = np.stack([np.stack([g.uniform(QP, N) for _ in range(beta)]) for _ in range(k + 1)])
    errs = np.stack([_errs(g, beta, N) for _ in range(k + 1)])
    fold, relin = okg.mkf_group_keys(sks, a, errs)
    D = 1 << 40

    def small():
        return (np.array([int(v) for v in g.uniform([15], N)[0]]) - 7).astype(np.int64)

    def multi(xs):
        parts = [keys.encrypt(p, O.small_to_rns(x * D, Q), g) for p, x in enumerate(xs)]
        c0 = parts[0][0]
        for p in parts[1:]:
            c0 = G.addmod(c0, p[0], Q)
        return np.stack([c0] + [p[1] for p in parts])

    xs, ys = [small() for _ in range(k)], [small() for _ in range(k)]
    left, right = multi(xs), multi(ys)
    X = [int(v) * D for v in np.sum(np.stack(xs), axis=0)]
    Y = [int(v) * D for v in np.sum(np.stack(ys), axis=0)]
    exp = G.negacyclic_mul(X, Y)
    o0, o1 = ctx.mkf_ckks_mulrelin(left, right, fold, relin)
    got = G.centered_crt(keys.decrypt(o0, o1), Q)
    assert max(abs(a_ - b_) for a_, b_ in zip(got, exp)) < 2 ** 62
    wrong = okg.mkf_relin_key(sks[0], a[k], errs[k])  # s_0^2 -> s_0, not the aggregate's
    o0, o1 = ctx.mkf_ckks_mulrelin(left, right, fold, wrong)
    got = G.centered_crt(keys.decrypt(o0, o1), Q)
    assert max(abs(a_ - b_) for a_, b_ in zip(got, exp)) > 2 ** 100
