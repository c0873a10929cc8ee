"""Pin the Galois (rotation / conjugation) part of the CPU oracle.

* Galois elements: closed forms of GaloisElementForColumnRotationBy (5^k mod 2N,
  order N/2), row rotation 2N-1, and inverses.
* ring.Permute: a ring homomorphism of Z_q[X]/(X^N+1) (the permutation of a
  negacyclic product is the product of the permutations), composition
  sigma_g sigma_h = sigma_gh, and the reference's "q - x" negation (a negated 0
  is written as q, keyswitch.go:300).
* PermuteNTTIndex: NTT(Permute(a)) = NTT(a)[index].
* KKLSS Rotate / Conjugate and MKF RotateNew: decrypt-and-compare against
  sigma(m), with keys made by the key equations of mkrlwe/keygen.go:190-268 and
  true_mkf_ckks.go:379-390 (mkckks_test.go:443-458 checks the same identity on
  slots); RotateNew's power-of-two fallback equals the direct rotation.
* The reference's non-hoisted ExternalProduct (keyswitch.go:82-130) reuses one
  decomposition buffer for every digit: equal to the hoisted product with one
  digit, not a key switch with more (documents why the library follows the
  hoisted form).
"""
import numpy as np
import pytest

import mkgen as G
import oracle as O


def _small_ctx(logN=8, L=3, K=2, gamma=2):
    Q, P = O.gen_primes(50, 15, L), O.gen_primes(60, 15, K)
    return O.Ctx(logN, Q, P, gamma=gamma)


def test_galois_element_closed_forms():
    for logN in (4, 10, 14):
        N = 1 << logN
        assert O.galois_element(logN, 0) == 1
        assert O.galois_element(logN, 1) == 5
        assert O.galois_element(logN, 3) == 125 % (2 * N)
        for k in (1, 2, 7, N // 4 - 1):
            g = O.galois_element(logN, k)
            assert g == pow(5, k, 2 * N)
            assert O.galois_element(logN, k + N // 2) == g          # 5 has order N/2 mod 2N
            assert (g * O.galois_element(logN, -k)) % (2 * N) == 1  # right rotation inverts
            assert (g * O.galois_inverse(logN, g)) % (2 * N) == 1
        assert O.galois_conjugate(logN) == 2 * N - 1
        # the N/2 column rotations are distinct and never the conjugation
        els = {O.galois_element(logN, k) for k in range(N // 2)}
        assert len(els) == N // 2 and 2 * N - 1 not in els


def test_permute_is_a_ring_homomorphism():
    ctx = _small_ctx()
    g = O.Prng(11)
    N, Q = ctx.N, ctx.Q
    a, b = g.uniform(Q, N), g.uniform(Q, N)
    for gal in (O.galois_element(8, 1), O.galois_element(8, 5), O.galois_conjugate(8)):
        ab = ctx.ntt_poly(G.mulmod(ctx.ntt_poly(a), ctx.ntt_poly(b), Q), inverse=True)
        pa, pb = ctx.permute(a, gal), ctx.permute(b, gal)
        lhs = ctx.permute(ab, gal)
        rhs = ctx.ntt_poly(G.mulmod(ctx.ntt_poly(pa), ctx.ntt_poly(pb), Q), inverse=True)
        red = lambda x: np.stack([x[i] % np.uint64(q) for i, q in enumerate(Q)])
        assert np.array_equal(red(lhs), red(rhs))
        # brute-force substitution X -> X^gal on the integer residues of limb 0
        x0 = [int(v) for v in a[0]]
        sub = G.galois_small(x0, gal)
        assert np.array_equal(np.array([int(v) % Q[0] for v in sub], np.uint64), red(pa)[0])


def test_permute_composition_and_negated_zero():
    ctx = _small_ctx()
    g = O.Prng(12)
    N, Q = ctx.N, ctx.Q
    a = g.uniform(Q, N)
    g1, g2 = O.galois_element(8, 3), O.galois_element(8, 9)
    red = lambda x: np.stack([x[i] % np.uint64(q) for i, q in enumerate(Q)])
    both = ctx.permute(red(ctx.permute(a, g1)), g2)
    assert np.array_equal(red(both), red(ctx.permute(a, (g1 * g2) % (2 * N))))
    # a zero sent to a negated position is written q, exactly as the reference does
    out = ctx.permute(np.zeros_like(a), g1)
    for i in range(N):
        neg = (i * g1) % (2 * N) >= N
        assert all(int(out[j][(i * g1) % N]) == (Q[j] if neg else 0) for j in range(len(Q)))
    # a q input (a previous negated zero) is negated to 0 and kept as q otherwise
    back = ctx.permute(out, g1)
    for i in range(N):
        neg = (i * g1) % (2 * N) >= N
        for j in range(len(Q)):
            x = int(out[j][i])
            assert int(back[j][(i * g1) % N]) == (Q[j] - x if neg else x)
    assert int(back.max()) > 0 and (back == 0).any()


def test_permute_ntt_index_matches_coefficient_permute():
    ctx = _small_ctx()
    g = O.Prng(13)
    a = g.uniform(ctx.Q, ctx.N)
    for k in (1, 2, 5, -1):
        gal = O.galois_element(ctx.logN, k)
        idx = O.permute_ntt_index(ctx.logN, gal).astype(np.int64)
        assert sorted(idx.tolist()) == list(range(ctx.N))
        A = ctx.ntt_poly(a)
        assert np.array_equal(ctx.ntt_poly(ctx.permute(a, gal)), A[:, idx])
        assert np.array_equal(ctx.automorphism_ntt(A, gal), A[:, idx])


def _kklss_setup(parties=2, logN=9, L=3, seed=20):
    ctx = _small_ctx(logN=logN, L=L)
    keys = G.KKLSSKeys(ctx, parties, seed=seed)
    g = O.Prng(seed + 1)
    Dl = 1 << 30
    N, Q = ctx.N, ctx.Q
    msgs = [((np.array([int(v) for v in g.uniform([15], N)[0]]) - 7) * Dl).astype(np.int64) for _ in range(parties)]
    parts = [keys.encrypt(p, O.small_to_rns(m, Q), g) for p, m in enumerate(msgs)]
    c0 = parts[0][0]
    for p in parts[1:]:
        c0 = G.addmod(c0, p[0], Q)
    ct = np.stack([c0] + [p[1] for p in parts])
    return ctx, keys, g, ct, np.sum(np.stack(msgs), axis=0)


def _err(ctx, keys, out, ids, want):
    got = G.centered_crt(keys.decrypt(out, ids), ctx.Q[:out.shape[1]])
    return max(abs(a - int(b)) for a, b in zip(got, want))


def test_kklss_rotate_decrypts_to_the_rotated_message():
    ctx, keys, g, ct, m = _kklss_setup()
    gal_of = lambda r: O.galois_element(ctx.logN, r)
    rot, _, crs = G.kklss_galois_keys(keys, gal_of, g, rotations=(1, 2, 4))
    for r in (1, 2, 4):
        out = ctx.kklss_rotate(ct, [rot[(0, r)], rot[(1, r)]], crs[r], gal_of(r))
        assert _err(ctx, keys, out, [0, 1], G.galois_small(m, gal_of(r))) < 2 ** 20
    # planted mistake: the other party's key breaks decryption
    bad = ctx.kklss_rotate(ct, [rot[(1, 1)], rot[(0, 1)]], crs[1], gal_of(1))
    assert _err(ctx, keys, bad, [0, 1], G.galois_small(m, gal_of(1))) > 2 ** 100
    # hoisted form: same result from a precomputed decomposition
    D = np.stack([ctx.decompose(ct[1 + i]) for i in range(2)])
    hout = ctx.kklss_rotate_hoisted(ct, D, [rot[(0, 2)], rot[(1, 2)]], crs[2], gal_of(2))
    assert np.array_equal(hout, ctx.kklss_rotate(ct, [rot[(0, 2)], rot[(1, 2)]], crs[2], gal_of(2)))


def test_kklss_rotate_new_power_of_two_fallback():
    ctx, keys, g, ct, m = _kklss_setup(seed=30)
    gal_of = lambda r: O.galois_element(ctx.logN, r)
    rot, _, crs = G.kklss_galois_keys(keys, gal_of, g, rotations=(1, 2, 4))
    out = ctx.kklss_rotate_new(ct, [0, 1], 7, rot, crs)
    assert _err(ctx, keys, out, [0, 1], G.galois_small(m, gal_of(7))) < 2 ** 22
    # normalisation: -N/2 + 3 and 3 + N/2 rotate like 3
    r3 = ctx.kklss_rotate_new(ct, [0, 1], 3, rot, crs)
    assert np.array_equal(ctx.kklss_rotate_new(ct, [0, 1], 3 - ctx.N // 2, rot, crs), r3)
    assert np.array_equal(ctx.kklss_rotate_new(ct, [0, 1], 3 + ctx.N // 2, rot, crs), r3)
    assert np.array_equal(ctx.kklss_rotate_new(ct, [0, 1], 0, rot, crs), ct)
    with pytest.raises(KeyError):
        ctx.kklss_rotate_new(ct, [0, 1], 3, rot, crs, hoisted=True)


def test_kklss_conjugate_decrypts_to_the_conjugated_message():
    ctx, keys, g, ct, m = _kklss_setup(seed=40)
    a = np.stack([G.mform(g.uniform(ctx.QP, ctx.N), ctx.QP) for _ in range(keys.beta)])
    _, conj, _ = G.kklss_galois_keys(keys, None, g, conj_crs=a)
    out = ctx.kklss_conjugate(ct, [conj[0], conj[1]], a)
    gal = O.galois_conjugate(ctx.logN)
    assert _err(ctx, keys, out, [0, 1], G.galois_small(m, gal)) < 2 ** 20
    # below the top level
    out1 = ctx.kklss_conjugate(ct[:, :2], [conj[0], conj[1]], a)
    assert _err(ctx, keys, out1, [0, 1], G.galois_small(m, gal)) < 2 ** 20


def test_reference_external_product_reuses_the_last_digit():
    ctx, keys, g, ct, m = _kklss_setup(seed=50)
    Q, QP = ctx.Q, ctx.QP
    # gadget key of s: key_j = MForm(e_j + g_j P s), so Ext(a, key) ~ a s
    key = np.stack([G.mform(keys._gadget(g, keys.sk[0], j), QP) for j in range(keys.beta)])
    top = ct[1]
    lvl0 = ct[1][:1]
    # one digit: identical to the hoisted product
    assert np.array_equal(ctx.external_product_as_written(lvl0, key),
                          ctx.external_product(ctx.decompose(lvl0, level=0), key, level=0))
    # several digits: it is sum_b key_b (.) D_last(a), not sum_b key_b (.) D_b(a)
    D = ctx.decompose(top)
    hoisted = ctx.external_product(D, key)
    written = ctx.external_product_as_written(top, key)
    assert np.array_equal(written, ctx.external_product(np.stack([D[-1]] * D.shape[0]), key))
    s = keys._plain(keys.sk[0])[:len(Q)]
    want = G.centered_crt(ctx.ntt_poly(G.mulmod(ctx.ntt_poly(top), s, Q), inverse=True), Q)
    dist = lambda x: max(abs(a - b) for a, b in zip(G.centered_crt(x, Q), want))
    assert dist(hoisted) < 2 ** 20
    assert dist(written) > 2 ** 100


def test_mkf_rotate_decrypts_to_the_rotated_message():
    logN = 9
    Q, P = O.gen_primes(50, 15, 3), O.gen_primes(60, 15, 2)
    ctx = O.Ctx(logN, Q, P, gamma=1)
    k = 3
    keys = G.MKFKeys(ctx, k, seed=60)
    g = O.Prng(61)
    N = ctx.N
    Dl = 1 << 30
    xs = [((np.array([int(v) for v in g.uniform([15], N)[0]]) - 7) * Dl).astype(np.int64) for _ in range(k)]
    parts = [keys.encrypt(p, O.small_to_rns(x, Q), g) for p, x in enumerate(xs)]
    c0 = parts[0][0]
    for p in parts[1:]:
        c0 = G.addmod(c0, p[0], Q)
    ct = np.stack([c0] + [p[1] for p in parts])
    m = np.sum(np.stack(xs), axis=0)
    for r in (1, 6):
        gal = O.galois_element(logN, r)
        rk = keys.rotation_keys(gal, g)
        o0, o1 = ctx.mkf_ckks_rotate(ct, rk, gal)
        got = G.centered_crt(keys.decrypt(o0, o1), Q)
        want = G.galois_small(m, gal)
        assert max(abs(a - int(b)) for a, b in zip(got, want)) < 2 ** 24
    # planted mistake: keys of another rotation
    o0, o1 = ctx.mkf_ckks_rotate(ct, keys.rotation_keys(O.galois_element(logN, 2), g), O.galois_element(logN, 1))
    got = G.centered_crt(keys.decrypt(o0, o1), Q)
    want = G.galois_small(m, O.galois_element(logN, 1))
    assert max(abs(a - int(b)) for a, b in zip(got, want)) > 2 ** 100
