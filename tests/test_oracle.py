"""Pin the CPU oracle against the reference's own vectors and against mathematics.

* RTL known answers: 64 MRed vectors (hardware/u55c/rtl/tb_mred64_pipeline.sv)
  and 128 butterfly vectors (tb_ntt_butterfly_x8.sv), extracted as data by
  tests/golden/extract_reference_vectors.py.
* HLS constant tables (hardware/u55c/hls/include/par1_ckks_hls_constants.hpp):
  every basis-extension / ModDown / rescale / N^-1 table the oracle derives.
* Closed forms that a plausible mistake (dropped term, wrong sign or index,
  transposed operand) breaks: NTT product = negacyclic schoolbook product,
  ModUp = CRT integer mod p, ModDown = floor(X / P), rescale = floor((X + h)/q_L),
  gadget digit = NTT of the digit integer, and decrypt-and-compare of the whole
  KKLSS MK-CKKS mul + relin (tolerance of mkckks/mkckks_test.go:352-358).
"""
import json
import math
import os

import numpy as np
import pytest

import mkgen as G
import oracle as O
from params import config

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


# ---------------------------------------------------------------- known answers

def test_mred_known_answers():
    kats = _load("mred64_kats.json")
    assert len(kats) == 64
    for v in kats:
        assert O.mred_params(v["q"]) == v["q_inv"]
        assert (v["q"] * v["q_inv"]) % (1 << 64) == 1
        assert O.mred(v["x"], v["y"], v["q"], v["q_inv"]) == v["expected"], v["index"]
        # the definition: x y 2^-64 mod q
        assert v["expected"] == v["x"] * v["y"] * pow(1 << 64, -1, v["q"]) % v["q"]


def test_butterfly_known_answers():
    b = _load("butterfly_kats.json")
    assert len(b["vectors"]) == 128
    for v in b["vectors"]:
        assert O.butterfly(v["u"], v["v"], v["w"], b["q"], b["q_inv"]) == (v["sum"], v["diff"])


def test_planted_mred_mistake_is_caught():
    """A Montgomery product missing the 2^-64 factor fails the known answers."""
    kats = _load("mred64_kats.json")
    wrong = [v for v in kats if (v["x"] * v["y"]) % v["q"] != v["expected"]]
    assert len(wrong) > 32


# ---------------------------------------------------------------- HLS constant tables

def test_hls_ckks_tables():
    c = _load("hls_ckks_constants.json")
    Q, P = c["Q"], c["P"]
    QP = Q + P
    a, b, v = O.modup_tables(P, Q)  # ModUp P -> Q of ModDown
    assert [int(x) for x in a] == c["P2Q_INV"]
    assert [[int(b[t][i]) for i in range(len(P))] for t in range(len(Q))] == c["P2Q_QSTAR"]
    assert [int(v[t][1]) for t in range(len(Q))] == c["P2Q_MINUS_P"]
    Pp = math.prod(P)
    assert [O.mform(pow(Pp, -1, q), q) for q in Q] == c["P_INV_Q"]
    for beta in range(c["BETA"]):  # MKF alpha = 2 digits
        a, b, v = O.modup_tables(Q[2 * beta:2 * beta + 2], QP)
        assert [int(x) for x in a] == c["DECOMP_INV"][beta]
        assert [[int(b[t][i]) for i in range(2)] for t in range(len(QP))] == c["DECOMP_QSTAR"][beta]
        assert [int(v[t][1]) for t in range(len(QP))] == c["DECOMP_MINUS_Q"][beta]
    assert [pow(2, 128, q) for q in Q] == c["R2_Q"]
    assert [O.mform(pow(c["N"], -1, q), q) for q in Q] == c["N_INV_Q"]
    assert [O.mred_params(q) for q in QP] == c["MRED_QP"]
    qL = Q[-1]
    assert [O.mform((-pow(qL, -1, q)) % q, q) for q in Q[:-1]] == c["RESCALE_5_TO_4"]
    # the ring built from the same moduli uses the same N^-1 (MForm)
    ctx = O.Ctx(c["LOG_N"], Q, P, gamma=1)
    for i, q in enumerate(Q):
        assert ctx.tables(i)[3] == c["N_INV_Q"][i]


# ---------------------------------------------------------------- NTT

def _small_ctx():
    logN, Q, P = config("small")
    return O.Ctx(logN, Q, P, gamma=2)


def test_psi_is_primitive_2n_root():
    ctx = _small_ctx()
    for i, q in enumerate(ctx.QP):
        psi, psiinv, g, ninv = ctx.tables(i)
        # psi[brv(1)] = psi^1 in Montgomery form; psi^N = -1 and psi^(2N) = 1
        p1 = int(psi[1 << (ctx.logN - 1)]) * pow(1 << 64, -1, q) % q
        assert pow(p1, ctx.N, q) == q - 1
        assert p1 == pow(g, (q - 1) // (2 * ctx.N), q)
        # rule: smallest generator >= 3
        assert all(any(pow(h, (q - 1) // f, q) == 1 for f in _factors(q - 1)) for h in range(3, g))


def _factors(n):
    f, p = [], 2
    while p * p <= n:
        if n % p == 0:
            f.append(p)
            while n % p == 0:
                n //= p
        p += 1 if p == 2 else 2
    if n > 1:
        f.append(n)
    return f


def test_ntt_roundtrip_and_negacyclic_product():
    ctx = _small_ctx()
    g = O.Prng(5)
    x = g.uniform(ctx.QP, ctx.N)
    for i in range(len(ctx.QP)):
        assert np.array_equal(ctx.ntt(ctx.ntt(x[i], i), i, inverse=True), x[i])
    # product in the NTT domain == negacyclic convolution (sparse operands keep the
    # schoolbook product cheap)
    q = ctx.QP[0]
    a = np.zeros(ctx.N, np.uint64)
    b = np.zeros(ctx.N, np.uint64)
    ia = [0, 3, 1000, ctx.N - 1]
    ib = [1, 7, ctx.N - 2]
    for j, idx in enumerate(ia):
        a[idx] = (j + 2) * 1234567
    for j, idx in enumerate(ib):
        b[idx] = q - (j + 3) * 7654321
    A, B = ctx.ntt(a, 0), ctx.ntt(b, 0)
    C = np.array([(int(u) * int(v)) % q for u, v in zip(A, B)], np.uint64)
    c = ctx.ntt(C, 0, inverse=True)
    exp = [0] * ctx.N
    for i in ia:
        for j in ib:
            k = i + j
            s = int(a[i]) * int(b[j])
            if k >= ctx.N:
                exp[k - ctx.N] -= s
            else:
                exp[k] += s
    assert [int(v) for v in c] == [e % q for e in exp]


def test_ntt_is_the_bitreversed_evaluation():
    """NTT output index j holds a(psi^(2 brv(j) + 1)) (Lattigo's CT ordering)."""
    ctx = _small_ctx()
    g = O.Prng(6)
    q = ctx.QP[1]
    a = g.uniform([q], ctx.N)[0]
    A = ctx.ntt(a, 1)
    psi_m = ctx.tables(1)[0][1 << (ctx.logN - 1)]
    psi = int(psi_m) * pow(1 << 64, -1, q) % q
    for j in (0, 1, 5, ctx.N - 1):
        e = 2 * int(format(j, "0%db" % ctx.logN)[::-1], 2) + 1
        x = pow(psi, e, q)
        val = 0
        for c in reversed([int(v) for v in a]):
            val = (val * x + c) % q
        assert int(A[j]) == val


# ---------------------------------------------------------------- basis extension

def _crt(rows, moduli):
    M = math.prod(moduli)
    return [sum(int(rows[i][c]) * (M // q) * pow(M // q, -1, q) for i, q in enumerate(moduli)) % M
            for c in range(rows.shape[1])]


def test_modup_is_crt_integer():
    ctx = _small_ctx()
    g = O.Prng(7)
    Q, P = ctx.Q[:3], ctx.P
    x = g.uniform(Q, 64)
    out = O.modup_exact(x, Q, P)
    X = _crt(x, Q)
    for j, p in enumerate(P):
        assert [int(v) for v in out[j]] == [v % p for v in X]


def test_moddown_is_floor_division_by_P():
    ctx = _small_ctx()
    g = O.Prng(8)
    Q, P = ctx.Q, ctx.P
    x = g.uniform(Q + P, 64)
    out = O.moddown(x[:len(Q)], x[len(Q):], Q, P)
    X = _crt(x, Q + P)
    Pp = math.prod(P)
    for i, q in enumerate(Q):
        assert [int(v) for v in out[i]] == [(v // Pp) % q for v in X]


def test_modup_brute_force_small_moduli():
    """Tiny primes: every residue pair; the float64 estimate is exact here."""
    Q, P = [17, 97], [193, 257]
    xs = np.array([[a for a in range(17) for _ in range(97)], [b for _ in range(17) for b in range(97)]],
                  np.uint64)
    out = O.modup_exact(xs, Q, P)
    X = _crt(xs, Q)
    for j, p in enumerate(P):
        assert [int(v) for v in out[j]] == [v % p for v in X]


def test_rescale_is_rounded_division():
    ctx = _small_ctx()
    g = O.Prng(9)
    x = g.uniform(ctx.Q, 64 if ctx.N < 64 else ctx.N)
    out = ctx.div_round_last(x, 1)
    X = _crt(x, ctx.Q)
    qL = ctx.Q[-1]
    for i, q in enumerate(ctx.Q[:-1]):
        assert [int(v) for v in out[i][:64]] == [((v + qL // 2) // qL) % q for v in X[:64]]
    # the NTT-domain variant gives the same residues under the NTT
    xn = ctx.ntt_poly(x)
    outn = ctx.div_round_last(xn, 1, ntt=True)
    assert np.array_equal(ctx.ntt_poly(outn, inverse=True), out)


def test_rescale_count_follows_the_scale_loop():
    ctx = _small_ctx()
    q = ctx.Q
    assert ctx.rescale_count(3, float(q[3]) ** 2, float(q[3]))[0] == 1
    assert ctx.rescale_count(3, 2.0 ** 80, 2.0 ** 40)[0] == 0  # 2^80 / q_3 < 2^39
    nb, s = ctx.rescale_count(3, float(q[3]) * float(q[2]) * 2.0 ** 50, 2.0 ** 50)
    assert nb == 2 and abs(s - 2.0 ** 50) / 2.0 ** 50 < 1e-9


# ---------------------------------------------------------------- gadget decomposition

@pytest.mark.parametrize("gamma", [2, 1])
def test_decompose_digit_is_ntt_of_digit_integer(gamma):
    """alpha = 1 (KKLSS, gamma=2): digit i = residue of limb i; alpha = 2 (gamma=1):
    digit = CRT integer of the digit's two limbs; every row is its NTT mod that row."""
    logN, Q, P = config("small")
    ctx = O.Ctx(logN, Q, P, gamma=gamma)
    g = O.Prng(10)
    a = g.uniform(Q, ctx.N)
    D = ctx.decompose(a)
    alpha = ctx.alpha
    for i in range(ctx.beta()):
        dq = Q[i * alpha:(i + 1) * alpha]
        X = _crt(a[i * alpha:(i + 1) * alpha], dq)
        for r, m in enumerate(Q + P):
            row = np.array([v % m for v in X], np.uint64)
            assert np.array_equal(D[i][r], ctx.ntt(row, r)), (i, r)


# ---------------------------------------------------------------- end to end (decrypt)

def test_kklss_mulrelin_decrypts_to_the_product():
    """Dec(MulRelin(ct0, ct1)) = m0 m1 + small noise; after rescale within the
    reference test's bound log2|err| <= -log2(scale) + logSlots + 12, here in the
    coefficient domain: |err| < 2^12 relative to the rescaled message unit."""
    logN = 10
    Q, P = O.gen_primes(50, 15, 4), O.gen_primes(60, 15, 2)
    ctx = O.Ctx(logN, Q, P, gamma=2)
    keys = G.KKLSSKeys(ctx, 3, seed=0)
    g = O.Prng(1)
    N = ctx.N
    D = 1 << 40

    def small():
        return (np.array([int(v) for v in g.uniform([15], N)[0]]) - 7).astype(np.int64)

    xs, ys = [small() for _ in range(2)], [small() for _ in range(2)]
    parts0 = [keys.encrypt(p, O.small_to_rns(x * D, Q), g) for p, x in zip([0, 1], xs)]
    parts1 = [keys.encrypt(p, O.small_to_rns(y * D, Q), g) for p, y in zip([1, 2], ys)]
    op0 = np.stack([G.addmod(parts0[0][0], parts0[1][0], Q), parts0[0][1], parts0[1][1]])
    op1 = np.stack([G.addmod(parts1[0][0], parts1[1][0], Q), parts1[0][1], parts1[1][1]])
    out, ids = ctx.kklss_mulrelin(op0, [0, 1], op1, [1, 2], keys.b, keys.d, keys.v, keys.u)
    assert ids == [0, 1, 2]
    got = G.centered_crt(keys.decrypt(out, ids), Q)
    X = [int(v) * D for v in xs[0] + xs[1]]
    Y = [int(v) * D for v in ys[0] + ys[1]]
    exp = G.negacyclic_mul(X, Y)
    err = max(abs(a - b) for a, b in zip(got, exp))
    assert err < 2 ** 60  # m0 m1 ~ 2^99; fresh-noise cross terms ~ 2^57
    # a relinearization that drops a term fails by ~Q
    resc = np.stack([ctx.div_round_last(out[i], 1) for i in range(out.shape[0])])
    got2 = G.centered_crt(keys.decrypt(resc, ids), Q[:-1])
    exp2 = [(e + Q[-1] // 2) // Q[-1] for e in exp]
    assert max(abs(a - b) for a, b in zip(got2, exp2)) < 2 ** 12


def test_decomposition_at_the_operands_own_level():
    """An operand above the product's level is decomposed at its own level
    (mkckks/evaluator.go:433-439): with alpha = 2 the last digit read at level 2
    is the CRT integer of limbs 2 and 3 of the level-4 operand, not limb 2 alone,
    and every row 0..2 and P holds its NTT; digits entirely below the level are
    those of the ordinary decomposition."""
    logN = 9
    Q, P = O.gen_primes(50, 15, 5), O.gen_primes(60, 15, 4)
    ctx = O.Ctx(logN, Q, P, gamma=2)
    assert ctx.alpha == 2
    a = O.Prng(31).uniform(Q, ctx.N)
    level = 2
    D = ctx.decompose_src(a, level)
    plain = ctx.decompose(a[:level + 1], level)
    assert D.shape == plain.shape == (2, level + 1 + len(P), ctx.N)
    assert np.array_equal(D[0], plain[0])
    assert not np.array_equal(D[1], plain[1])
    X = _crt(a[2:4], Q[2:4])
    for r, m in enumerate(Q[:level + 1] + P):
        ri = r if r <= level else len(Q) + r - level - 1
        row = np.array([v % m for v in X], np.uint64)
        assert np.array_equal(D[1][r], ctx.ntt(row, ri)), r
    # src level = level: the ordinary decomposition
    assert np.array_equal(ctx.decompose_src(a[:level + 1], level), plain)


def test_kklss_mulrelin_operands_at_their_own_levels():
    """MulRelinNew of operands at levels 4 and 2 with alpha = 2: op0's last digit
    at level 2 spans limbs 2 and 3, so the words differ from dropping op0 to level 2
    first, and both decrypt to m0 m1 (each is a valid key switch mod Q_2); with
    alpha = 1, or when level + 1 is a multiple of alpha, the two agree."""
    logN = 9
    Q, P = O.gen_primes(50, 15, 5), O.gen_primes(60, 15, 4)
    ctx = O.Ctx(logN, Q, P, gamma=2)
    keys = G.KKLSSKeys(ctx, 2, seed=33)
    g = O.Prng(34)
    N, D = ctx.N, 1 << 30
    x = (np.array([int(v) for v in g.uniform([15], N)[0]]) - 7).astype(np.int64)
    y = (np.array([int(v) for v in g.uniform([15], N)[0]]) - 7).astype(np.int64)
    op0 = np.stack(keys.encrypt(0, O.small_to_rns(x * D, Q), g))
    op1 = np.stack(keys.encrypt(1, O.small_to_rns(y * D, Q), g))
    exp = G.negacyclic_mul([int(v) * D for v in x], [int(v) * D for v in y])
    args = (keys.b, keys.d, keys.v, keys.u)
    for lv1, differs in ((2, True), (3, False)):
        own, ids = ctx.kklss_mulrelin_levels(op0, [0], op1[:, :lv1 + 1], [1], *args)
        drop, ids2 = ctx.kklss_mulrelin(op0[:, :lv1 + 1], [0], op1[:, :lv1 + 1], [1], *args, level=lv1)
        assert ids == ids2 == [0, 1]
        assert (not np.array_equal(own, drop)) == differs, lv1
        for out in (own, drop):
            got = G.centered_crt(keys.decrypt(out, ids), Q[:lv1 + 1])
            assert max(abs(a - b) for a, b in zip(got, exp)) < 2 ** 60, lv1
    # alpha = 1: the digit of one limb does not depend on the level
    ctx1 = O.Ctx(logN, Q, P[:2], gamma=2)
    k1 = G.KKLSSKeys(ctx1, 2, seed=35)
    own, _ = ctx1.kklss_mulrelin_levels(op0, [0], op1[:, :3], [1], k1.b, k1.d, k1.v, k1.u)
    drop, _ = ctx1.kklss_mulrelin(op0[:, :3], [0], op1[:, :3], [1], k1.b, k1.d, k1.v, k1.u, level=2)
    assert np.array_equal(own, drop)


def test_planted_relin_mistake_is_caught():
    """Swapping the roles of d and b keys (a transposed operand) breaks decryption."""
    logN = 9
    Q, P = O.gen_primes(50, 15, 3), O.gen_primes(60, 15, 2)
    ctx = O.Ctx(logN, Q, P, gamma=2)
    keys = G.KKLSSKeys(ctx, 2, seed=3)
    g = O.Prng(4)
    D = 1 << 30
    x = np.ones(ctx.N, np.int64)
    c0a, c1a = keys.encrypt(0, O.small_to_rns(x * D, Q), g)
    c0b, c1b = keys.encrypt(1, O.small_to_rns(x * D, Q), g)
    op0 = np.stack([c0a, c1a])
    op1 = np.stack([c0b, c1b])
    out, ids = ctx.kklss_mulrelin(op0, [0], op1, [1], keys.d, keys.b, keys.v, keys.u)  # swapped
    got = G.centered_crt(keys.decrypt(out, ids), Q)
    exp = G.negacyclic_mul([D] * ctx.N, [D] * ctx.N)
    assert max(abs(a - b) for a, b in zip(got, exp)) > 2 ** 80


def test_mkf_ckks_mulrelin_decrypts_to_the_product():
    """MKF CKKSCase.Run restated: fold + Karatsuba + relin + rescale decrypts under
    the aggregate secret to round(m_left m_right / q_L) (ckks.go:155-165 checks the
    decoded product within 1e-6; here the coefficient-domain bound)."""
    logN = 10
    Q, P = O.gen_primes(50, 15, 4), O.gen_primes(60, 15, 2)
    ctx = O.Ctx(logN, Q, P, gamma=1)  # Lattigo rlwe decomposition: alpha = K = 2
    k = 3
    keys = G.MKFKeys(ctx, k, seed=5)
    g = O.Prng(6)
    N = ctx.N
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
    o0, o1 = ctx.mkf_ckks_mulrelin(left, right, keys.fold, keys.relin)
    got = G.centered_crt(keys.decrypt(o0, o1), Q)
    X = [int(v) * D for v in np.sum(np.stack(xs), axis=0)]
    Y = [int(v) * D for v in np.sum(np.stack(ys), axis=0)]
    exp = G.negacyclic_mul(X, Y)
    assert max(abs(a - b) for a, b in zip(got, exp)) < 2 ** 62
    r0, r1 = ctx.div_round_last(o0, 1, ntt=True), ctx.div_round_last(o1, 1, ntt=True)
    got2 = G.centered_crt(keys.decrypt(r0, r1), Q[:-1])
    exp2 = [(e + Q[-1] // 2) // Q[-1] for e in exp]
    assert max(abs(a - b) for a, b in zip(got2, exp2)) < 2 ** 14


def test_mkf_switch_is_a_key_switch():
    """switchOne(c, fold_i) = (p0, p1) with p0 + p1 s = c s_i + small (common.go:79-107)."""
    logN = 9
    Q, P = O.gen_primes(50, 15, 3), O.gen_primes(60, 15, 1)
    ctx = O.Ctx(logN, Q, P, gamma=1)  # alpha = K = 1: the copy decomposition
    keys = G.MKFKeys(ctx, 2, seed=7)
    g = O.Prng(8)
    c = g.uniform(Q, ctx.N)
    p0, p1 = ctx.mkf_switch_one(c, keys.fold[1], in_ntt=True, ntt_output=True)
    got = G.centered_crt(keys.decrypt(p0, p1), Q)
    si = ctx.ntt_poly(O.small_to_rns(keys.small[1], Q))
    want = G.centered_crt(ctx.ntt_poly(G.mulmod(c, si, Q), inverse=True), Q)
    assert max(abs(a - b) for a, b in zip(got, want)) < 2 ** 30


# ---------------------------------------------------------------- KKLSS MK-BFV

def _bfv_small(logN=10, L=3, seed=0, parties=3):
    Q = O.gen_primes(50, 15, L)
    QM = O.gen_primes(50, 15, L, exclude=Q)
    P = O.gen_primes(60, 15, 2)
    bctx = O.BfvCtx(logN, Q, QM, P, 65537)
    return bctx, G.BFVKeys(bctx, parties, seed=seed)


def test_bfv_modup_and_rescale_closed_forms():
    """ModUpQtoR keeps [x]_Q in R; Rescale gives floor(QMul [x]_Q / Q) in R
    (mkbfv/basis_extension.go:50-66, 88-104); raw limbs are congruent."""
    bctx, _ = _bfv_small(logN=4, parties=0)
    g = O.Prng(9)
    x = g.uniform(bctx.Q, bctx.N)
    X = G.centered_crt(x, bctx.Q)
    Qp = math.prod(bctx.Q)
    QMp = math.prod(bctx.QMul)
    up = bctx.modup_q_to_r(x)
    rs = bctx.rescale_q_to_r(x)
    for c in range(bctx.N):
        xv = X[c] % Qp
        for l, q in enumerate(bctx.R):
            assert int(up[l][c]) % q == xv % q
            assert int(rs[l][c]) % q == (QMp * xv // Qp) % q
            assert int(up[l][c]) < 3 * q and int(rs[l][c]) < 3 * q


def test_bfv_quantize_is_scaled_floor():
    """Quantize(x) = floor(t [x]_{Q QMul} / QMul) mod Q up to the float64 estimate's +-1
    (exact here: |error| of the basis extension is 0 or QMul)."""
    bctx, _ = _bfv_small(logN=4, parties=0)
    g = O.Prng(10)
    x = g.uniform(bctx.R, bctx.N)                  # coefficient domain over R
    xn = bctx.ntt_r(x)
    got = bctx.quantize(xn)
    Rp = math.prod(bctx.R)
    QMp = math.prod(bctx.QMul)
    Qp = math.prod(bctx.Q)
    Xv = G.centered_crt(x, bctx.R)
    for c in range(bctx.N):
        xv = Xv[c] % Rp
        tx = (bctx.t * xv) % Rp
        exp = tx // QMp
        v = G.centered_crt(got[:, c:c + 1], bctx.Q)[0] % Qp
        assert v == exp % Qp


def test_bfv_decompose_copies_raw_limbs():
    """Digit i of an R poly is R limb i (its integer value) in every QP row, NTT'd."""
    bctx, _ = _bfv_small(logN=4, parties=0)
    g = O.Prng(11)
    aR = bctx.modup_q_to_r(g.uniform(bctx.Q, bctx.N))
    D = bctx.decompose(aR)
    for d in range(2 * bctx.L):
        for r, q in enumerate(bctx.QP):
            row = np.array([int(v) % q for v in aR[d]], np.uint64)
            assert np.array_equal(D[d][r], bctx.qp.ntt(row, r)), (d, r)


def test_kklss_bfv_mulrelin_decrypts_exactly():
    """Dec(MulRelin(ct0, ct1)) = m0 m1 mod t exactly, as mkbfv_test.go:407-413
    (coefficient encoding: the negacyclic product mod t)."""
    bctx, keys = _bfv_small()
    g = O.Prng(12)
    N, t = bctx.N, bctx.t
    m = [[int(v) for v in g.uniform([t], N)[0]] for _ in range(4)]
    p0 = [keys.encrypt(p, keys.encode(x), g) for p, x in zip([0, 1], m[:2])]
    p1 = [keys.encrypt(p, keys.encode(x), g) for p, x in zip([1, 2], m[2:])]
    Q = bctx.Q
    op0 = np.stack([G.addmod(p0[0][0], p0[1][0], Q), p0[0][1], p0[1][1]])
    op1 = np.stack([G.addmod(p1[0][0], p1[1][0], Q), p1[0][1], p1[1][1]])
    out, ids = bctx.mulrelin(op0, [0, 1], op1, [1, 2], keys.bkey, keys.dkey, keys.vkey, keys.u)
    assert ids == [0, 1, 2]
    a = [(x + y) % t for x, y in zip(m[0], m[1])]
    b = [(x + y) % t for x, y in zip(m[2], m[3])]
    exp = [v % t for v in G.negacyclic_mul(a, b)]
    assert keys.decode(out, ids) == exp


def test_bfv_planted_branch_swap_is_caught():
    """Swapping the two gadget branches of d (a transposed operand) breaks decryption."""
    bctx, keys = _bfv_small(logN=9, L=3, seed=2, parties=2)
    g = O.Prng(13)
    N, t, L = bctx.N, bctx.t, bctx.L
    m0 = [int(v) for v in g.uniform([t], N)[0]]
    m1 = [int(v) for v in g.uniform([t], N)[0]]
    c0 = keys.encrypt(0, keys.encode(m0), g)
    c1 = keys.encrypt(1, keys.encode(m1), g)
    swapped = [np.concatenate([d[L:], d[:L]]) for d in keys.dkey]
    out, ids = bctx.mulrelin(np.stack(c0), [0], np.stack(c1), [1], keys.bkey, swapped, keys.vkey, keys.u)
    exp = [v % t for v in G.negacyclic_mul(m0, m1)]
    assert keys.decode(out, ids) != exp


def test_kklss_mulrelin_below_top_level_decrypts():
    """At level L-2 the x / y accumulators stay full-level switching keys (the
    reference's swkPool layout, keyswitch_hoisted.go:60-70): decryption still
    gives m0 m1 + small noise (a row-layout mix-up fails by ~Q)."""
    logN = 9
    Q, P = O.gen_primes(50, 15, 4), O.gen_primes(60, 15, 2)
    ctx = O.Ctx(logN, Q, P, gamma=2)
    keys = G.KKLSSKeys(ctx, 2, seed=21)
    g = O.Prng(22)
    N, D, level = ctx.N, 1 << 30, 2
    x = (np.array([int(v) for v in g.uniform([15], N)[0]]) - 7).astype(np.int64)
    y = (np.array([int(v) for v in g.uniform([15], N)[0]]) - 7).astype(np.int64)
    c0 = keys.encrypt(0, O.small_to_rns(x * D, Q), g)
    c1 = keys.encrypt(1, O.small_to_rns(y * D, Q), g)
    op0 = np.stack(c0)[:, :level + 1]
    op1 = np.stack(c1)[:, :level + 1]
    out, ids = ctx.kklss_mulrelin(op0, [0], op1, [1], keys.b, keys.d, keys.v, keys.u, level=level)
    got = G.centered_crt(keys.decrypt(out, ids), Q[:level + 1])
    exp = G.negacyclic_mul([int(v) * D for v in x], [int(v) * D for v in y])
    assert max(abs(a - b) for a, b in zip(got, exp)) < 2 ** 60


# ---------------------------------------------------------------- MKF MK-BFV

def test_mkf_bfv_mulrelin_decrypts_exactly():
    """MKF BFVCase.Run restated (mkf_cpu/bfv.go:212-252): fold, lift / rescale into
    R, Karatsuba, quantize, relinearize; Dec_s(c0, Gamma) = (sum m_left)(sum m_right)
    mod t exactly, as the reference's validation (:242-250)."""
    logN, L, k = 10, 3, 3
    Q = O.gen_primes(50, 15, L)
    QM = O.gen_primes(50, 15, L, exclude=Q)
    P = O.gen_primes(60, 15, 2)
    bctx = O.BfvCtx(logN, Q, QM, P, 65537, gamma=1)
    ctx = bctx.qp
    keys = G.MKFKeys(ctx, k, seed=41)
    g = O.Prng(42)
    N, t = ctx.N, bctx.t
    Qp = math.prod(Q)
    delta = Qp // t

    def enc(ms):
        c0 = np.zeros((L, N), np.uint64)
        comps = []
        for p, m in enumerate(ms):
            pt = np.stack([np.array([(int(x) * delta) % q for x in m], np.uint64) for q in Q])
            a, b = keys.encrypt(p, pt, g)
            c0 = G.addmod(c0, ctx.ntt_poly(a, inverse=True), Q)
            comps.append(ctx.ntt_poly(b, inverse=True))
        return np.stack([c0] + comps)

    ml = [[int(x) for x in g.uniform([t], N)[0]] for _ in range(k)]
    mr = [[int(x) for x in g.uniform([t], N)[0]] for _ in range(k)]
    left, right = enc(ml), enc(mr)
    c0, gamma = bctx.mkf_mulrelin(left, right, keys.fold, keys.relin)
    x = G.centered_crt(keys.decrypt(ctx.ntt_poly(c0), ctx.ntt_poly(gamma)), Q)
    got = [((t * v + Qp // 2) // Qp) % t for v in x]
    a = [sum(col) % t for col in zip(*ml)]
    b = [sum(col) % t for col in zip(*mr)]
    assert got == [v % t for v in G.negacyclic_mul(a, b)]
