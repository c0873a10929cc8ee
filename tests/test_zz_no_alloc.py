"""The allocation-free KKLSS MK-CKKS entry point (mkhip_kklss_ckks_mulrelin_noalloc)
against the plan entry point it mirrors (mkhip_plan_kklss_ckks + launch) and the
oracle: two calls in a row on the same scratch, the second smaller (fewer
parties, a lower level), then a batch larger than the scratch, refused with
MKHIP_ERR_SCRATCH."""
import numpy as np
import pytest

import oracle as O
from params import config

pytestmark = pytest.mark.gpu


def _oracle(ref, op0, ids0, op1, ids1, keys, u, nres):
    pids = sorted(set(ids0) | set(ids1))
    idx = {p: i for i, p in enumerate(pids)}
    out, _ = ref.kklss_mulrelin(op0, [idx[i] for i in ids0], op1, [idx[i] for i in ids1],
                                [keys[p][0] for p in pids], [keys[p][1] for p in pids],
                                [keys[p][2] for p in pids], u)
    return np.stack([ref.div_round_last(out[i], nres) for i in range(out.shape[0])])


def test_noalloc_matches_plan_entry_point():
    import mkf_amd as mk

    logN, Q, P = config("c2")  # N = 2^14, 8 Q + 2 P
    ctx, ref = mk.Context(logN, Q, P, gamma=2), O.Ctx(logN, Q, P, gamma=2)
    g = O.Prng(201)
    N, beta = ref.N, ctx.beta
    keys = {}
    for pid in range(4):
        keys[pid] = [np.stack([g.uniform(ref.QP, N) for _ in range(beta)]) for _ in range(3)]
        ctx.register_party(pid, *keys[pid])
    u = np.stack([g.uniform(ref.QP, N) for _ in range(beta)])
    ctx.set_crs_u(u)
    bufs = []
    try:
        # call 1: k = 4 at the top level; call 2: parties {1, 3} x {0} one level lower
        for ids0, ids1, level in (([0, 1, 2, 3], [0, 1, 2, 3], ref.L - 1), ([1, 3], [0], ref.L - 2)):
            nl = level + 1
            op0 = np.stack([g.uniform(ref.Q[:nl], N) for _ in range(1 + len(ids0))])
            op1 = np.stack([g.uniform(ref.Q[:nl], N) for _ in range(1 + len(ids1))])
            ku = len(set(ids0) | set(ids1))
            shape = (1 + ku, nl - 1, N)
            d0, d1 = ctx.alloc(op0.nbytes), ctx.alloc(op1.nbytes)
            da, db = ctx.alloc(int(np.prod(shape)) * 8), ctx.alloc(int(np.prod(shape)) * 8)
            bufs += [d0, d1, da, db]
            ctx.h2d(d0, op0)
            ctx.h2d(d1, op1)
            op = dict(op0=d0, op1=d1, ids0=ids0, ids1=ids1, level=level, nres=1)
            ctx.kklss_ckks_mulrelin_noalloc([dict(op, out=da)])
            got = ctx.d2h(np.zeros(shape, np.uint64), da)
            pl = mk.Plan(ctx, [dict(op, out=db)], use_graph=False)
            pl.launch()
            ctx.sync()
            pl.close()
            ref_words = ctx.d2h(np.zeros(shape, np.uint64), db)
            assert np.array_equal(got, ref_words), (ids0, ids1, level)
            exp = _oracle(ref, op0, ids0, op1, ids1, keys, u, 1)
            assert np.array_equal(got, exp), (ids0, ids1, level)
        # more work than the scratch holds: two headline ops
        big = dict(op0=bufs[0], op1=bufs[1], out=bufs[2], ids0=[0, 1, 2, 3], ids1=[0, 1, 2, 3], level=ref.L - 1,
                   nres=1)
        with pytest.raises(mk.MkhipError) as e:
            ctx.kklss_ckks_mulrelin_noalloc([big, big])
        assert e.value.status == -12  # MKHIP_ERR_SCRATCH
    finally:
        for p in bufs:
            ctx.free(p)
        ctx.close()
