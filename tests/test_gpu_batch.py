"""GPU parity of uneven batches (one hipGraph launch over several contexts):
ops of different party counts k and ring degrees N -- MKF MK-CKKS at N=2^13
(one P prime), KKLSS MK-CKKS at N=2^14 and N=2^15 -- each equal, bit for bit, to
the oracle's MulRelinNew + Rescale (mkckks/evaluator.go:419-446, 360-401;
mkf_cpu/ckks.go:102-151), on every replay, with and without workspace chunking."""
import numpy as np
import pytest

import oracle as O
from params import primes

pytestmark = pytest.mark.gpu


def _kklss_keys(ctx, ref, parties, seed):
    g = O.Prng(seed)
    keys = {}
    for pid in parties:
        keys[pid] = [np.stack([g.uniform(ref.QP, ref.N) for _ in range(ctx.beta)]) for _ in range(3)]
        ctx.register_party(pid, *keys[pid])
    u = np.stack([g.uniform(ref.QP, ref.N) for _ in range(ctx.beta)])
    ctx.set_crs_u(u)
    return keys, u


def _kklss_expected(ref, op0, ids0, op1, ids1, keys, u):
    pids = sorted(keys)
    idx = {p: i for i, p in enumerate(pids)}
    out, _ = ref.kklss_mulrelin(op0, [idx[i] for i in ids0], op1, [idx[i] for i in ids1],
                                [keys[p][0] for p in pids], [keys[p][1] for p in pids],
                                [keys[p][2] for p in pids], u)
    return np.stack([ref.div_round_last(out[i], 1) for i in range(out.shape[0])])


class _Branch:
    """Ops of one context with device buffers and oracle outputs."""

    def __init__(self, ctx, ref, scheme):
        self.ctx, self.ref, self.scheme = ctx, ref, scheme
        self.ops, self.exp, self.bufs = [], [], []

    def add(self, op, exp, bufs):
        self.ops.append(op)
        self.exp.append(exp)
        self.bufs += bufs

    def check(self, tag):
        for i, (op, e) in enumerate(zip(self.ops, self.exp)):
            got = self.ctx.d2h(np.zeros(e.shape, np.uint64), op["out"])
            assert np.array_equal(got, e), (tag, self.scheme, i)

    def clear_outputs(self):
        for op, e in zip(self.ops, self.exp):
            self.ctx.h2d(op["out"], np.zeros(e.shape, np.uint64))


@pytest.fixture(scope="module")
def branches():
    import mkf_amd

    out = []
    # MKF MK-CKKS, N = 2^13, 4 Q + 1 P (Lattigo decomposition alpha = K = 1)
    Q, P = primes(50, 4), primes(60, 1)
    ctx, ref = mkf_amd.Context(13, Q, P, gamma=1), O.Ctx(13, Q, P, gamma=1)
    br = _Branch(ctx, ref, "mkf")
    g = O.Prng(71)
    beta = ref.beta()

    def swk():
        return np.stack([np.stack([g.uniform(ref.QP, ref.N), g.uniform(ref.QP, ref.N)]) for _ in range(beta)])

    groups = {}
    for grp, k in ((1, 2), (2, 4)):
        groups[grp] = ([swk() for _ in range(k)], swk())
        ctx.register_mkf_group(grp, *groups[grp])
    nl, N = ref.L, ref.N
    for grp in (1, 2, 1):
        k = len(groups[grp][0])
        left = np.stack([g.uniform(ref.Q, N) for _ in range(1 + k)])
        right = np.stack([g.uniform(ref.Q, N) for _ in range(1 + k)])
        dl, dr, do = ctx.alloc(left.nbytes), ctx.alloc(right.nbytes), ctx.alloc((1 + k) * (nl - 1) * N * 8)
        ctx.h2d(dl, left)
        ctx.h2d(dr, right)
        o0, o1 = ref.mkf_ckks_mulrelin(left, right, *groups[grp])
        r0, r1 = ref.div_round_last(o0, 1, ntt=True), ref.div_round_last(o1, 1, ntt=True)
        br.add(dict(left=dl, right=dr, out=do, group=grp, level=nl - 1, nres=1), np.stack([r0] + [r1] * k),
               [dl, dr, do])
    out.append(br)
    # KKLSS MK-CKKS at N = 2^14 (4 Q + 2 P) and N = 2^15 (3 Q + 2 P)
    for logN, L, seed, shapes in ((14, 4, 73, [([0, 1], [0, 1]), ([0, 1, 2, 3], [0, 1, 2, 3]), ([2], [0, 3]),
                                              ([1, 3], [1, 2, 3])]),
                                  (15, 3, 75, [([0, 1, 2, 3], [0, 1, 2, 3]), ([0, 2], [1])])):
        Q, P = primes(50, L), primes(60, 2)
        ctx, ref = mkf_amd.Context(logN, Q, P, gamma=2), O.Ctx(logN, Q, P, gamma=2)
        keys, u = _kklss_keys(ctx, ref, [0, 1, 2, 3], seed)
        br = _Branch(ctx, ref, "kklss")
        g = O.Prng(seed + 1)
        nl, N = ref.L, ref.N
        for ids0, ids1 in shapes:
            op0 = np.stack([g.uniform(ref.Q, N) for _ in range(1 + len(ids0))])
            op1 = np.stack([g.uniform(ref.Q, N) for _ in range(1 + len(ids1))])
            exp = _kklss_expected(ref, op0, ids0, op1, ids1, keys, u)
            d0, d1, do = ctx.alloc(op0.nbytes), ctx.alloc(op1.nbytes), ctx.alloc(exp.nbytes)
            ctx.h2d(d0, op0)
            ctx.h2d(d1, op1)
            br.add(dict(op0=d0, op1=d1, out=do, ids0=ids0, ids1=ids1, level=nl - 1, nres=1), exp, [d0, d1, do])
        out.append(br)
    yield out
    for br in out:
        for p in br.bufs:
            br.ctx.free(p)
        br.ctx.close()


@pytest.mark.parametrize("budget", ["none", "one_op"])
def test_uneven_batch_one_graph_matches_oracle(branches, budget):
    import mkf_amd

    batch = mkf_amd.Batch()
    chunks = []
    for br in branches:
        # "one_op": a budget below any two-op workspace, so every chunk holds one op
        chunks.append(batch.add(br.ctx, br.ops, br.scheme, ws_budget=1 if budget == "one_op" else 0))
    if budget == "one_op":
        assert chunks == [len(br.ops) for br in branches]
    else:
        assert chunks == [1] * len(branches)
    for i, br in enumerate(branches):
        n, arena = batch.branch_info(i)
        assert n == chunks[i] and arena > 0
    batch.finalize()
    for rep in range(3):  # each replay recomputes every output
        for br in branches:
            br.clear_outputs()
        batch.submit()
        batch.wait()
        for br in branches:
            br.check((budget, rep))
        # eager work on the contexts between replays
        br0 = branches[1]
        br0.ctx.kklss_ckks_mulrelin(np.zeros((3, br0.ref.L, br0.ref.N), np.uint64), [0, 1],
                                    np.zeros((3, br0.ref.L, br0.ref.N), np.uint64), [0, 1], 2.0 ** 50, 2.0 ** 50,
                                    2.0 ** 50)
    assert batch.time(2) > 0
    batch.close()
