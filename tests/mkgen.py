"""Test-side key generation, encryption and decryption over the CPU oracle.

The reference samples keys with a crypto PRNG (mkrlwe/keygen.go:26), so "keys
from seed s" is the tests' convention: the oracle's xoshiro256** stream.
The key equations follow mkrlwe/keygen.go:
  secret  s  : ternary, P(0)=1/2 (:58-75), stored as MForm(NTT(s)) over QP (:44-55)
  b_i = MForm(e_i - a_i s)                              (:161-169)
  d_i = MForm(e_i + g_i P s - a_i r)                    (:171-176, GenSwitchingKey :270-327)
  v_i = MForm(-(e_i + g_i P r) - u_i s)                 (:178-185)
where a = CRS[0], u = CRS[-1] are uniform MForm PolyQP vectors (params.go:16-61)
and g_i P s adds P*s on the Q limbs of digit i only.
"""
import numpy as np

import oracle as O


def _mulmod_rows(a, b, moduli):
    out = np.empty_like(a)
    for i, q in enumerate(moduli):
        out[i] = np.array([(int(x) * int(y)) % q for x, y in zip(a[i], b[i])], np.uint64) \
            if a.shape[-1] <= 64 else _mulmod_vec(a[i], b[i], q)
    return out


def _mulmod_vec(x, y, q):
    # exact 64x64 mod q via Python ints on object arrays
    xo = x.astype(object)
    yo = y.astype(object)
    return np.array((xo * yo) % q, dtype=np.uint64)


def mulmod(a, b, moduli):
    a = np.asarray(a, np.uint64)
    b = np.asarray(b, np.uint64)
    return np.stack([_mulmod_vec(a[i], b[i], q) for i, q in enumerate(moduli)])


def addmod(a, b, moduli):
    return np.stack([np.array((a[i].astype(object) + b[i].astype(object)) % q, np.uint64)
                     for i, q in enumerate(moduli)])


def submod(a, b, moduli):
    return np.stack([np.array((a[i].astype(object) - b[i].astype(object)) % q, np.uint64)
                     for i, q in enumerate(moduli)])


def mform(a, moduli):
    return np.stack([np.array((a[i].astype(object) * (1 << 64)) % q, np.uint64) for i, q in enumerate(moduli)])


def inv_mform(a, moduli):
    return np.stack([np.array((a[i].astype(object) * pow(1 << 64, -1, q)) % q, np.uint64)
                     for i, q in enumerate(moduli)])


class KKLSSKeys:
    """CRS + per-party (b, d, v) relinearization keys + secrets for a KKLSS context."""

    def __init__(self, ctx, nparties, seed=0):
        self.ctx = ctx
        g = O.Prng(seed)
        N, QP = ctx.N, ctx.QP
        beta = ctx.beta()
        self.beta = beta
        # CRS[0] = a, CRS[-1] = u (uniform, MForm)
        self.a = np.stack([mform(g.uniform(QP, N), QP) for _ in range(beta)])
        self.u = np.stack([mform(g.uniform(QP, N), QP) for _ in range(beta)])
        self.sk, self.sk_small, self.b, self.d, self.v, self.pk = [], [], [], [], [], []
        for _ in range(nparties):
            s_small = g.ternary(N)
            r_small = g.ternary(N)
            s = self._secret(s_small)
            r = self._secret(r_small)
            self.sk.append(s)
            self.sk_small.append(s_small)
            self.b.append(self._gen_b(g, s))
            self.d.append(self._gen_d(g, s, r))
            self.v.append(self._gen_v(g, s, r))
            self.pk.append(self._gen_pk(g, s))

    def _secret(self, small):
        ctx = self.ctx
        return mform(ctx.ntt_poly(O.small_to_rns(small, ctx.QP)), ctx.QP)

    def _err_ntt(self, g):
        ctx = self.ctx
        return ctx.ntt_poly(O.small_to_rns(g.gaussian(ctx.N), ctx.QP))

    def _plain(self, x_mform):
        return inv_mform(x_mform, self.ctx.QP)

    def _gen_b(self, g, s):
        QP = self.ctx.QP
        out = []
        for i in range(self.beta):
            as_ = mulmod(self._plain(self.a[i]), self._plain(s), QP)
            out.append(mform(submod(self._err_ntt(g), as_, QP), QP))
        return np.stack(out)

    def _gadget(self, g, s, i):
        """MForm(e + g_i P s): P*s added on the Q limbs of digit i (keygen.go:270-327)."""
        ctx = self.ctx
        QP = ctx.QP
        e = self._err_ntt(g)
        Pprod = 1
        for p in ctx.P:
            Pprod *= p
        splain = self._plain(s)
        alpha = ctx.alpha
        for j in range(alpha):
            idx = i * alpha + j
            if idx >= ctx.L:
                break
            q = QP[idx]
            e[idx] = np.array((e[idx].astype(object) + splain[idx].astype(object) * (Pprod % q)) % q, np.uint64)
        return e

    def _gen_d(self, g, s, r):
        QP = self.ctx.QP
        out = []
        for i in range(self.beta):
            x = self._gadget(g, s, i)
            ar = mulmod(self._plain(self.a[i]), self._plain(r), QP)
            out.append(mform(submod(x, ar, QP), QP))
        return np.stack(out)

    def _gen_v(self, g, s, r):
        QP = self.ctx.QP
        out = []
        for i in range(self.beta):
            x = self._gadget(g, r, i)
            us = mulmod(self._plain(self.u[i]), self._plain(s), QP)
            neg = submod(np.zeros_like(x), addmod(x, us, QP), QP)
            out.append(mform(neg, QP))
        return np.stack(out)

    def _gen_pk(self, g, s):
        """pk = (-s a_0 + e, a_0) over Q (keygen.go:88-109), NTT form, not MForm."""
        ctx = self.ctx
        Q = ctx.Q
        L = ctx.L
        a0 = self._plain(self.a[0])[:L]
        e = self._err_ntt(g)[:L]
        pk0 = submod(e, mulmod(a0, self._plain(s)[:L], Q), Q)
        return pk0, a0

    def encrypt(self, party, m_coeff, g):
        """Public-key encryption (mkrlwe/encryptor.go:55-119), coefficient-domain output."""
        ctx = self.ctx
        Q, N = ctx.Q, ctx.N
        pk0, pk1 = self.pk[party]
        vv = ctx.ntt_poly(O.small_to_rns(g.ternary(N), Q))
        c0 = ctx.ntt_poly(mulmod(vv, pk0, Q), inverse=True)
        c1 = ctx.ntt_poly(mulmod(vv, pk1, Q), inverse=True)
        c0 = addmod(c0, O.small_to_rns(g.gaussian(N), Q), Q)
        c1 = addmod(c1, O.small_to_rns(g.gaussian(N), Q), Q)
        c0 = addmod(c0, m_coeff, Q)
        return c0, c1

    def decrypt(self, ct, ids, nlimbs=None):
        """c0 + sum_i c_i s_i (mkrlwe/decryptor.go:34-66) on the first nlimbs limbs."""
        ctx = self.ctx
        nl = ct.shape[1] if nlimbs is None else nlimbs
        Q = ctx.Q[:nl]
        acc = ct[0][:nl].copy()
        for j, pid in enumerate(ids):
            s = self._plain(self.sk[pid])[:nl]
            cj = ctx.ntt_poly(ct[1 + j][:nl])
            acc = addmod(acc, ctx.ntt_poly(mulmod(cj, s, Q), inverse=True), Q)
        return acc


def galois_small(x, gal):
    """sigma_gal(x) for an integer polynomial x (length N) in Z[X]/(X^N+1):
    X^i -> X^(i gal), folded with X^N = -1."""
    N = len(x)
    out = [0] * N
    for i, v in enumerate(x):
        e = (i * gal) % (2 * N)
        if e < N:
            out[e] += int(v)
        else:
            out[e - N] -= int(v)
    return np.array(out, np.int64)


def kklss_galois_keys(keys, gal_of, g, rotations=(), conj_crs=None, crs=None):
    """Rotation and conjugation keys for a KKLSSKeys/BFVKeys instance
    (mkrlwe/keygen.go:190-268), returned as dicts.

    rotation key of party i for rotidx r (GenRotationKey :190-230):
        rk_j = MForm(-a_j sigma_{g^-1}(s) + e_j + g_j P s),   a = CRS[r], g = galEl(r)
    conjugation key (GenConjugationKey :240-268):
        ck_j = MForm(-a_j s + e_j + g_j P sigma_{2N-1}(s)),   a = CRS[-2]
    crs maps r -> MForm uniform CRS (generated when absent)."""
    ctx = keys.ctx
    QP, N = ctx.QP, ctx.N
    crs = {} if crs is None else crs
    rot = {}
    for r in rotations:
        if r not in crs:
            crs[r] = np.stack([mform(g.uniform(QP, N), QP) for _ in range(keys.beta)])
        ginv = pow(gal_of(r), -1, 2 * N)
        for pid, s_small in enumerate(keys.sk_small):
            s_rot = ctx.ntt_poly(O.small_to_rns(galois_small(s_small, ginv), QP))
            out = []
            for j in range(keys.beta):
                x = keys._gadget(g, keys.sk[pid], j)
                ar = mulmod(keys._plain(crs[r][j]), s_rot, QP)
                out.append(mform(submod(x, ar, QP), QP))
            rot[(pid, r)] = np.stack(out)
    conj = {}
    if conj_crs is not None:
        for pid, s_small in enumerate(keys.sk_small):
            s = keys._plain(keys.sk[pid])
            s_c = mform(ctx.ntt_poly(O.small_to_rns(galois_small(s_small, 2 * N - 1), QP)), QP)
            out = []
            for j in range(keys.beta):
                x = keys._gadget(g, s_c, j)
                as_ = mulmod(keys._plain(conj_crs[j]), s, QP)
                out.append(mform(submod(x, as_, QP), QP))
            conj[pid] = np.stack(out)
    return rot, conj, crs


class MKFKeys:
    """MKF aggregate-secret keys (software/mkf/mkf_cpu/common.go:32-51) over a
    gamma = 1 context: secrets s_i, s = sum s_i, fold keys s_i -> s and the
    relinearization key s^2 -> s, each a Lattigo switching key
    [beta][2][L+K][N] in MForm: ([-a s_out + e + P s_in g_b], [a])."""

    def __init__(self, ctx, k, seed=0):
        self.ctx = ctx
        g = O.Prng(seed)
        N, QP = ctx.N, ctx.QP
        self.small = [g.ternary(N) for _ in range(k)]
        agg = np.sum(np.stack(self.small), axis=0)
        self.agg_small = agg
        self.s = self._ntt_small(agg)                 # plain NTT(s) over QP
        s_in = [self._ntt_small(x) for x in self.small]
        self.fold = [self._swk(g, si, self.s) for si in s_in]
        s2 = mulmod(self.s, self.s, QP)
        self.relin = self._swk(g, s2, self.s)
        self.pk = [self._pk(g, self._ntt_small(x)) for x in self.small]

    def _ntt_small(self, x):
        return self.ctx.ntt_poly(O.small_to_rns(x, self.ctx.QP))

    def _swk(self, g, s_in, s_out):
        ctx = self.ctx
        QP, N = ctx.QP, ctx.N
        Pprod = 1
        for p in ctx.P:
            Pprod *= p
        out = []
        for b in range(ctx.beta()):
            a = g.uniform(QP, N)
            e = ctx.ntt_poly(O.small_to_rns(g.gaussian(N), QP))
            v0 = submod(e, mulmod(a, s_out, QP), QP)
            for j in range(ctx.alpha):
                idx = b * ctx.alpha + j
                if idx >= ctx.L:
                    break
                q = QP[idx]
                v0[idx] = np.array((v0[idx].astype(object) + s_in[idx].astype(object) * (Pprod % q)) % q, np.uint64)
            out.append(np.stack([mform(v0, QP), mform(a, QP)]))
        return np.stack(out)

    def rotation_keys(self, gal, g):
        """Party-specific Galois keys sigma_gal(s_i) -> s (true_mkf_ckks.go:379-390)."""
        return [self._swk(g, self._ntt_small(galois_small(x, gal)), self.s) for x in self.small]

    def _pk(self, g, s):
        ctx = self.ctx
        Q, N = ctx.Q, ctx.N
        a = g.uniform(Q, N)
        e = ctx.ntt_poly(O.small_to_rns(g.gaussian(N), Q))
        return submod(e, mulmod(a, s[:ctx.L], Q), Q), a

    def encrypt(self, party, m_coeff, g):
        """NTT-domain public-key encryption (the multiCiphertext inputs of ckks.go:81-100)."""
        ctx = self.ctx
        Q, N = ctx.Q, ctx.N
        pk0, pk1 = self.pk[party]
        v = ctx.ntt_poly(O.small_to_rns(g.ternary(N), Q))
        e0 = ctx.ntt_poly(O.small_to_rns(g.gaussian(N), Q))
        e1 = ctx.ntt_poly(O.small_to_rns(g.gaussian(N), Q))
        m = ctx.ntt_poly(m_coeff)
        return addmod(addmod(mulmod(v, pk0, Q), e0, Q), m, Q), addmod(mulmod(v, pk1, Q), e1, Q)

    def decrypt(self, c0, c1):
        """c0 + c1 s (NTT domain in, coefficient domain out)."""
        ctx = self.ctx
        nl = c0.shape[0]
        Q = ctx.Q[:nl]
        return ctx.ntt_poly(addmod(c0, mulmod(c1, self.s[:nl], Q), Q), inverse=True)


def centered_crt(x, moduli):
    """Lift an RNS poly (nlimbs, N) to centered integers via CRT (Python ints)."""
    Qp = 1
    for q in moduli:
        Qp *= q
    out = []
    N = x.shape[1]
    coefs = [0] * N
    for i, q in enumerate(moduli):
        Mi = Qp // q
        inv = pow(Mi, -1, q)
        for c in range(N):
            coefs[c] += int(x[i][c]) * inv % q * Mi
    for c in range(N):
        v = coefs[c] % Qp
        if v > Qp // 2:
            v -= Qp
        out.append(v)
    return out


def negacyclic_mul(a, b):
    """Schoolbook product in Z[X]/(X^N+1) of Python-int lists."""
    N = len(a)
    ma = max((abs(int(x)) for x in a), default=0)
    mb = max((abs(int(x)) for x in b), default=0)
    if ma * mb * N < (1 << 62):  # exact in int64: numpy convolution
        c = np.convolve(np.asarray(a, np.int64), np.asarray(b, np.int64))
        r = c[:N].copy()
        r[:N - 1] -= c[N:]
        return [int(x) for x in r]
    out = [0] * N
    for i, ai in enumerate(a):
        if ai == 0:
            continue
        for j, bj in enumerate(b):
            k = i + j
            if k < N:
                out[k] += ai * bj
            else:
                out[k - N] -= ai * bj
    return out


class BFVKeys(KKLSSKeys):
    """KKLSS MK-BFV keys (software/kklss/mkbfv/keygen.go:24-162) over an
    ``oracle.BfvCtx``.  CRS a1 = CRS[0], a2 = CRS[-3], u = CRS[-1]; per party
      b_br = MForm(e - a_br s)                       (:46-63)
      d_br = MForm(G_br,i s + e - a_br r)             (:65-73, GenBFVSwitchingKey :91-162)
      v    = MForm(-(e + g_i P r) - u s)              (:75-85)
    with the BFV gadget G_1,i = floor((QQMul/q_i) t [(QQMul/q_i)^-1]_{q_i} P / QMul)
    over the Q primes and G_2,i likewise over the QMul primes.  The tensors the
    evaluator consumes: bkey = [b_1; b_2], dkey = [d_1; d_2] (2L digits), vkey = v."""

    def __init__(self, bctx, nparties, seed=0):
        self.bctx = bctx
        ctx = bctx.qp
        self.ctx = ctx
        g = O.Prng(seed)
        N, QP, L = ctx.N, ctx.QP, ctx.L
        self.beta = L
        self.a = np.stack([mform(g.uniform(QP, N), QP) for _ in range(L)])    # CRS[0]
        self.a2 = np.stack([mform(g.uniform(QP, N), QP) for _ in range(L)])   # CRS[-3]
        self.u = np.stack([mform(g.uniform(QP, N), QP) for _ in range(L)])    # CRS[-1]
        Qprod = 1
        for q in bctx.Q:
            Qprod *= q
        QMprod = 1
        for q in bctx.QMul:
            QMprod *= q
        Pprod = 1
        for p in bctx.P:
            Pprod *= p
        QQ = Qprod * QMprod
        self.G1 = [(QQ // qi) * bctx.t * pow(QQ // qi, -1, qi) * Pprod // QMprod for qi in bctx.Q]
        self.G2 = [(QQ // qi) * bctx.t * pow(QQ // qi, -1, qi) * Pprod // QMprod for qi in bctx.QMul]
        self.sk, self.sk_small, self.pk = [], [], []
        self.bkey, self.dkey, self.vkey = [], [], []
        for _ in range(nparties):
            s_small, r_small = g.ternary(N), g.ternary(N)
            s, r = self._secret(s_small), self._secret(r_small)
            self.sk.append(s)
            self.sk_small.append(s_small)
            b1 = self._gen_b_crs(g, s, self.a)
            b2 = self._gen_b_crs(g, s, self.a2)
            d1 = self._gen_d_bfv(g, s, r, self.a, self.G1)
            d2 = self._gen_d_bfv(g, s, r, self.a2, self.G2)
            v = self._gen_v(g, s, r)
            self.bkey.append(np.concatenate([b1, b2]))
            self.dkey.append(np.concatenate([d1, d2]))
            self.vkey.append(v)
            self.pk.append(self._gen_pk(g, s))

    def _gen_b_crs(self, g, s, crs):
        QP = self.ctx.QP
        return np.stack([mform(submod(self._err_ntt(g), mulmod(self._plain(crs[i]), self._plain(s), QP), QP), QP)
                         for i in range(self.beta)])

    def _gen_d_bfv(self, g, s, r, crs, G):
        QP = self.ctx.QP
        splain = self._plain(s)
        out = []
        for i in range(self.beta):
            gs = np.stack([np.array((splain[j].astype(object) * (G[i] % q)) % q, np.uint64)
                           for j, q in enumerate(QP)])
            x = addmod(gs, self._err_ntt(g), QP)
            out.append(mform(submod(x, mulmod(self._plain(crs[i]), self._plain(r), QP), QP), QP))
        return np.stack(out)

    def encode(self, m):
        """Plaintext Delta m, Delta = floor(Q / t) (coefficient encoding of m mod t)."""
        Qprod = 1
        for q in self.bctx.Q:
            Qprod *= q
        delta = Qprod // self.bctx.t
        vals = [(int(x) % self.bctx.t) * delta for x in m]
        return np.stack([np.array([v % q for v in vals], np.uint64) for q in self.bctx.Q])

    def decode(self, ct, ids):
        """round(t (c0 + sum c_i s_i) / Q) mod t (bfv Decryptor + DecodeInt, coefficient form)."""
        Q = self.bctx.Q
        Qprod = 1
        for q in Q:
            Qprod *= q
        x = centered_crt(self.decrypt(ct, ids), Q)
        t = self.bctx.t
        return [((t * v + Qprod // 2) // Qprod) % t for v in x]
