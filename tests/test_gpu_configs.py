"""GPU parity at the full shapes of the uneven-batch and serving workloads.

Uneven batch: one hipGraph launch over three rings -- MKF MK-CKKS at N=2^13 with
4 Q + 1 P (parties k = 2, 4, 8, 16), KKLSS MK-CKKS at N=2^14 with 8 Q + 2 P
(k = 8 and 16, also with different party sets on the two sides) and at N=2^15
with 16 Q + 3 P (k = 16 and 2) -- every output equal, bit for bit, to the
oracle's MulRelinNew + Rescale (mkckks/evaluator.go:419-446, 360-401;
mkf_cpu/ckks.go:102-151), on two replays.

Serving: 32 registered parties per class with their keys resident in HBM --
KKLSS MK-CKKS at N=2^14 (8 Q + 2 P) and N=2^15 (16 Q + 3 P), KKLSS MK-BFV at
N=2^14 (the reference's PN14QP439) and N=2^15 (8 Q + 8 QMul + 2 P) --, requests
with k from 2 to 16: every request's words are the same alone and in a mix of
all of them, and a request of each class equals the oracle (KKLSS MK-BFV:
mkbfv/evaluator.go:85-152).

Keys and operands are uniform residues: bit-exact parity does not need valid
encryptions.  A party's keys are a function of (seed, party), so the oracle
regenerates only the parties a checked request uses.
"""
import numpy as np
import pytest

import oracle as O
from params import PN14QP439, primes

pytestmark = pytest.mark.gpu


def _kklss_party_keys(ref, beta, seed, pid):
    g = O.Prng(seed * 1000 + pid)
    return [np.stack([g.uniform(ref.QP, ref.N) for _ in range(beta)]) for _ in range(3)]


def _crs(ref, beta, seed):
    g = O.Prng(seed * 1000 + 999)
    return np.stack([g.uniform(ref.QP, ref.N) for _ in range(beta)])


def _kklss_expected(ref, beta, seed, op0, ids0, op1, ids1, nres):
    pids = sorted(set(ids0) | set(ids1))
    idx = {p: i for i, p in enumerate(pids)}
    keys = [_kklss_party_keys(ref, beta, seed, p) for p in pids]
    out, oids = ref.kklss_mulrelin(op0, [idx[i] for i in ids0], op1, [idx[i] for i in ids1],
                                   [k[0] for k in keys], [k[1] for k in keys], [k[2] for k in keys],
                                   _crs(ref, beta, seed))
    assert [pids[i] for i in oids] == pids
    return np.stack([ref.div_round_last(out[i], nres) for i in range(out.shape[0])])


def _bfv_party_keys(ref, seed, pid):
    g = O.Prng(seed * 1000 + pid)
    L, QP, N = ref.L, ref.QP, ref.N
    b = np.stack([g.uniform(QP, N) for _ in range(2 * L)])
    d = np.stack([g.uniform(QP, N) for _ in range(2 * L)])
    v = np.stack([g.uniform(QP, N) for _ in range(L)])
    return b, d, v


def _bfv_expected(ref, seed, op0, ids0, op1, ids1):
    pids = sorted(set(ids0) | set(ids1))
    tiny = np.zeros((1, 1), np.uint64)  # never read: only the ids in the request are
    b, d, v = [tiny] * (max(pids) + 1), [tiny] * (max(pids) + 1), [tiny] * (max(pids) + 1)
    for p in pids:
        b[p], d[p], v[p] = _bfv_party_keys(ref, seed, p)
    g = O.Prng(seed * 1000 + 999)
    u = np.stack([g.uniform(ref.QP, ref.N) for _ in range(ref.L)])
    out, oids = ref.mulrelin(op0, ids0, op1, ids1, b, d, v, u)
    assert oids == pids
    return out


# ---------------------------------------------------------------------------
# uneven batch (256-op workload's rings and party counts)

UNEVEN_KKLSS = ((14, 8, 2, 81, [(list(range(8)), list(range(8))), (list(range(16)), list(range(16))),
                                (list(range(0, 16, 2)), list(range(16)))]),
                (15, 16, 3, 83, [(list(range(16)), list(range(16))), ([3, 9], [3, 9])]))


@pytest.fixture(scope="module")
def uneven():
    import mkf_amd

    branches = []
    # MKF MK-CKKS at N = 2^13, 4 Q + 1 P: one group per party count
    Q, P = primes(50, 4), primes(60, 1)
    ctx, ref = mkf_amd.Context(13, Q, P, gamma=1), O.Ctx(13, Q, P, gamma=1)
    g = O.Prng(85)
    beta, nl, N = ref.beta(), ref.L, ref.N

    def swk():
        return np.stack([np.stack([g.uniform(ref.QP, N), g.uniform(ref.QP, N)]) for _ in range(beta)])

    ops, exps, bufs = [], [], []
    for k in (2, 4, 8, 16):
        fold, relin = [swk() for _ in range(k)], swk()
        ctx.register_mkf_group(k, fold, relin)
        left = np.stack([g.uniform(ref.Q, N) for _ in range(1 + k)])
        right = np.stack([g.uniform(ref.Q, N) for _ in range(1 + k)])
        dl, dr, do = ctx.alloc(left.nbytes), ctx.alloc(right.nbytes), ctx.alloc((1 + k) * (nl - 1) * N * 8)
        ctx.h2d(dl, left)
        ctx.h2d(dr, right)
        o0, o1 = ref.mkf_ckks_mulrelin(left, right, fold, relin)
        r0, r1 = ref.div_round_last(o0, 1, ntt=True), ref.div_round_last(o1, 1, ntt=True)
        ops.append(dict(left=dl, right=dr, out=do, group=k, level=nl - 1, nres=1))
        exps.append(np.stack([r0] + [r1] * k))
        bufs += [dl, dr, do]
    branches.append((ctx, "mkf", ops, exps, bufs))
    # KKLSS MK-CKKS at N = 2^14 (8 Q + 2 P) and N = 2^15 (16 Q + 3 P), 16 parties
    for logN, L, K, seed, shapes in UNEVEN_KKLSS:
        Q, P = primes(50, L), primes(60, K)
        ctx, ref = mkf_amd.Context(logN, Q, P, gamma=2), O.Ctx(logN, Q, P, gamma=2)
        beta, nl, N = ctx.beta, ref.L, ref.N
        for pid in range(16):
            ctx.register_party(pid, *_kklss_party_keys(ref, beta, seed, pid))
        ctx.set_crs_u(_crs(ref, beta, seed))
        g = O.Prng(seed + 1)
        ops, exps, bufs = [], [], []
        for ids0, ids1 in shapes:
            op0 = np.stack([g.uniform(ref.Q, N) for _ in range(1 + len(ids0))])
            op1 = np.stack([g.uniform(ref.Q, N) for _ in range(1 + len(ids1))])
            exp = _kklss_expected(ref, beta, seed, op0, ids0, op1, ids1, 1)
            d0, d1, do = ctx.alloc(op0.nbytes), ctx.alloc(op1.nbytes), ctx.alloc(exp.nbytes)
            ctx.h2d(d0, op0)
            ctx.h2d(d1, op1)
            ops.append(dict(op0=d0, op1=d1, out=do, ids0=ids0, ids1=ids1, level=nl - 1, nres=1))
            exps.append(exp)
            bufs += [d0, d1, do]
        branches.append((ctx, "kklss", ops, exps, bufs))
    yield branches
    for ctx, _, _, _, bufs in branches:
        for p in bufs:
            ctx.free(p)
        ctx.close()


def test_uneven_batch_full_shapes_match_oracle(uneven):
    import mkf_amd

    batch = mkf_amd.Batch()
    for ctx, scheme, ops, _, _ in uneven:
        assert batch.add(ctx, ops, scheme) == 1
    batch.finalize()
    for rep in range(2):
        for ctx, _, ops, exps, _ in uneven:
            for op, e in zip(ops, exps):
                ctx.h2d(op["out"], np.zeros(e.shape, np.uint64))
        batch.submit()
        batch.wait()
        for ctx, scheme, ops, exps, _ in uneven:
            for i, (op, e) in enumerate(zip(ops, exps)):
                got = ctx.d2h(np.zeros(e.shape, np.uint64), op["out"])
                assert np.array_equal(got, e), (rep, scheme, ctx.N, i)
    batch.close()


# ---------------------------------------------------------------------------
# serving: 32 resident parties per class, k from 2 to 16

NPARTY = 32
SERVE_SEED = {("kklss", 14): 91, ("kklss", 15): 92, ("bfv", 14): 93, ("bfv", 15): 94}


def _serve_classes():
    import mkf_amd

    out = []
    for scheme, logN in SERVE_SEED:
        seed = SERVE_SEED[(scheme, logN)]
        if scheme == "kklss":
            L, K = (8, 2) if logN == 14 else (16, 3)
            Q, P = primes(50, L), primes(60, K)
            ctx, ref = mkf_amd.Context(logN, Q, P, gamma=2), O.Ctx(logN, Q, P, gamma=2)
            for pid in range(NPARTY):
                ctx.register_party(pid, *_kklss_party_keys(ref, ctx.beta, seed, pid))
            ctx.set_crs_u(_crs(ref, ctx.beta, seed))
        else:
            if logN == 14:
                Q, QM, P, t = PN14QP439["Q"], PN14QP439["QMul"], PN14QP439["P"], PN14QP439["t"]
            else:
                Q = primes(55, 8)
                Q, QM, P, t = Q, primes(55, 8, exclude=Q), primes(61, 2), 65537
            ctx, ref = mkf_amd.Context(logN, Q, P, gamma=2, t=t, qmul=QM), O.BfvCtx(logN, Q, QM, P, t)
            for pid in range(NPARTY):
                ctx.register_bfv_party(pid, *_bfv_party_keys(ref, seed, pid))
            g = O.Prng(seed * 1000 + 999)
            ctx.set_crs_u(np.stack([g.uniform(ref.QP, ref.N) for _ in range(ref.L)]))
        out.append(dict(scheme=scheme, logN=logN, seed=seed, ctx=ctx, ref=ref))
    return out


@pytest.fixture(scope="module")
def serve_classes():
    cls = _serve_classes()
    yield cls
    for c in cls:
        c["ctx"].close()


def _requests(classes, per_class, rng):
    g = O.Prng(95)
    reqs = []
    for ci, c in enumerate(classes):
        ref = c["ref"]
        for j in range(per_class):
            k = 16 if j == 0 else int(rng.integers(2, 17))
            ids = sorted(int(x) for x in rng.choice(NPARTY, k, replace=False))
            if j % 2:  # different party sets on the two sides (union of k parties)
                ids0 = ids[: max(1, k // 2)]
                ids1 = ids[k // 3:]
            else:
                ids0 = ids1 = ids
            op0 = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + len(ids0))])
            op1 = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + len(ids1))])
            nres = 1 if c["scheme"] == "kklss" else 0
            out = np.zeros((1 + k, ref.L - nres, ref.N), np.uint64)
            reqs.append(dict(ci=ci, in0=op0, in1=op1, ids0=ids0, ids1=ids1, nres=nres, out=out, k=k))
    return reqs


def test_serving_32_parties_alone_vs_mix_and_oracle(serve_classes):
    import mkf_amd

    classes = serve_classes
    srv = mkf_amd.Server(streams=4, max_batch=8, mem_bytes=48 << 30)
    cl = [srv.add_class(c["ctx"], c["scheme"]) for c in classes]
    srv.start()
    rng = np.random.default_rng(96)
    reqs = _requests(classes, 4, rng)

    def submit(r, out):
        return srv.enqueue(cl[r["ci"]], r["in0"], r["in1"], out, ids0=r["ids0"], ids1=r["ids1"], nres=r["nres"])

    alone = []
    for r in reqs:  # alone: the server is idle, the request is its own batch
        out = np.zeros_like(r["out"])
        st, lat, _ = srv.wait(submit(r, out))
        assert st == 0 and lat > 0, (classes[r["ci"]]["scheme"], r["k"])
        alone.append(out)
    # all at once: batched per class, batches of different classes overlap on the device
    order = list(rng.permutation(len(reqs)))
    tickets = [(i, submit(reqs[i], reqs[i]["out"])) for i in order]
    for i, t in tickets:
        assert srv.wait(t)[0] == 0, i
    for i, r in enumerate(reqs):
        assert np.array_equal(alone[i], r["out"]), (classes[r["ci"]]["scheme"], classes[r["ci"]]["logN"], i)
    stats = srv.stats()
    assert stats["completed"] == 2 * len(reqs) and stats["max_batch"] >= 2
    srv.close()
    # the oracle on one request of each class: the largest k at N = 2^14, the
    # smallest at N = 2^15 (bounds the oracle's time)
    for ci, c in enumerate(classes):
        mine = [r for r in reqs if r["ci"] == ci]
        r = max(mine, key=lambda x: x["k"]) if c["logN"] == 14 else min(mine, key=lambda x: x["k"])
        if c["scheme"] == "kklss":
            exp = _kklss_expected(c["ref"], c["ctx"].beta, c["seed"], r["in0"], r["ids0"], r["in1"], r["ids1"], 1)
        else:
            exp = _bfv_expected(c["ref"], c["seed"], r["in0"], r["ids0"], r["in1"], r["ids1"])
        assert np.array_equal(r["out"], exp), (c["scheme"], c["logN"], r["k"])
