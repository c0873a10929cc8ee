"""GPU tests of the serving loop (continuous batching, admission by device
memory): every result is bit-identical alone and in a mix of other requests,
and equal to the oracle's MulRelinNew (KKLSS MK-CKKS with rescale,
mkckks/evaluator.go:419-446; KKLSS MK-BFV, mkbfv/evaluator.go:85-152)."""
import numpy as np
import pytest

import oracle as O
from params import bfv_config, config

pytestmark = pytest.mark.gpu

NPARTY = 6


@pytest.fixture(scope="module")
def classes():
    """A KKLSS MK-CKKS context and a KKLSS MK-BFV context with NPARTY parties each."""
    import mkf_amd

    out = {}
    logN, Q, P = config("small")
    ctx, ref = mkf_amd.Context(logN, Q, P, gamma=2), O.Ctx(logN, Q, P, gamma=2)
    g = O.Prng(91)
    keys = [[np.stack([g.uniform(ref.QP, ref.N) for _ in range(ctx.beta)]) for _ in range(3)]
            for _ in range(NPARTY)]
    for pid in range(NPARTY):
        ctx.register_party(pid, *keys[pid])
    u = np.stack([g.uniform(ref.QP, ref.N) for _ in range(ctx.beta)])
    ctx.set_crs_u(u)
    out["ckks"] = (ctx, ref, keys, u)
    logN, Q, QM, P, t = bfv_config("bsmall")
    ctx, ref = mkf_amd.Context(logN, Q, P, gamma=2, t=t, qmul=QM), O.BfvCtx(logN, Q, QM, P, t)
    g = O.Prng(92)
    L, QP, N = ref.L, ref.QP, ref.N
    b = [np.stack([g.uniform(QP, N) for _ in range(2 * L)]) for _ in range(NPARTY)]
    d = [np.stack([g.uniform(QP, N) for _ in range(2 * L)]) for _ in range(NPARTY)]
    v = [np.stack([g.uniform(QP, N) for _ in range(L)]) for _ in range(NPARTY)]
    u = np.stack([g.uniform(QP, N) for _ in range(L)])
    for pid in range(NPARTY):
        ctx.register_bfv_party(pid, b[pid], d[pid], v[pid])
    ctx.set_crs_u(u)
    out["bfv"] = (ctx, ref, (b, d, v), u)
    yield out
    for c in out.values():
        c[0].close()


def _request(kind, cl, ref, g, ids0, ids1):
    op0 = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + len(ids0))])
    op1 = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + len(ids1))])
    ku = len(set(ids0) | set(ids1))
    nres = 1 if kind == "ckks" else 0
    out = np.zeros((1 + ku, ref.L - nres, ref.N), np.uint64)
    return dict(kind=kind, cls=cl, in0=op0, in1=op1, out=out, ids0=ids0, ids1=ids1, nres=nres)


def _expected(classes, r):
    ctx, ref, keys, u = classes[r["kind"]]
    if r["kind"] == "ckks":
        o, _ = ref.kklss_mulrelin(r["in0"], r["ids0"], r["in1"], r["ids1"], [k[0] for k in keys],
                                  [k[1] for k in keys], [k[2] for k in keys], u)
        return np.stack([ref.div_round_last(o[i], 1) for i in range(o.shape[0])])
    b, d, v = keys
    o, _ = ref.mulrelin(r["in0"], r["ids0"], r["in1"], r["ids1"], b, d, v, u)
    return o


def _server(classes, **kw):
    import mkf_amd

    s = mkf_amd.Server(**kw)
    cl = {"ckks": s.add_class(classes["ckks"][0], "kklss"), "bfv": s.add_class(classes["bfv"][0], "bfv")}
    s.start()
    return s, cl


def _submit(s, r):
    return s.enqueue(r["cls"], r["in0"], r["in1"], r["out"], ids0=r["ids0"], ids1=r["ids1"], nres=r["nres"])


def _mix(cl, classes, g, rng, n):
    reqs = []
    for _ in range(n):
        kind = "ckks" if rng.random() < 0.6 else "bfv"
        ref = classes[kind][1]
        k0, k1 = int(rng.integers(1, 5)), int(rng.integers(1, 5))
        ids0 = sorted(int(x) for x in rng.choice(NPARTY, k0, replace=False))
        ids1 = sorted(int(x) for x in rng.choice(NPARTY, k1, replace=False))
        reqs.append(_request(kind, cl[kind], ref, g, ids0, ids1))
    return reqs


def test_result_same_alone_and_in_mix(classes):
    s, cl = _server(classes, streams=3, max_batch=8, mem_bytes=2 << 30)
    g = O.Prng(93)
    rng = np.random.default_rng(93)
    probes = [_request("ckks", cl["ckks"], classes["ckks"][1], g, [0, 1], [1, 2]),
              _request("bfv", cl["bfv"], classes["bfv"][1], g, [2, 3, 4], [0, 4])]
    alone = []
    for r in probes:  # alone: the server is idle, the request is its own batch
        st, lat, _ = s.wait(_submit(s, r))
        assert st == 0 and lat > 0
        alone.append(r["out"].copy())
        assert np.array_equal(r["out"], _expected(classes, r)), r["kind"]
    # in a mix: the same operands re-submitted among other requests, all enqueued at once
    mix = _mix(cl, classes, g, rng, 14)
    again = [dict(r, out=np.zeros_like(r["out"])) for r in probes]
    order = mix[:5] + [again[0]] + mix[5:10] + [again[1]] + mix[10:]
    tickets = [_submit(s, r) for r in order]
    for t in tickets:
        st, _, _ = s.wait(t)
        assert st == 0
    for a, r in zip(alone, again):
        assert np.array_equal(a, r["out"]), r["kind"]
    for r in mix[:4]:  # mix members themselves against the oracle
        assert np.array_equal(r["out"], _expected(classes, r)), r["kind"]
    stats = s.stats()
    assert stats["completed"] == len(order) + len(probes)
    assert stats["max_batch"] >= 2  # requests that waited were batched together
    s.close()


def _block_bytes(ctx, ref, ids0, ids1):
    """Device block of one KKLSS MK-CKKS request: plan workspace + operands + output."""
    import mkf_amd

    fake = 1 << 20  # plan building only records pointers; nothing is launched
    p = mkf_amd.Plan(ctx, [dict(op0=fake, op1=fake, out=fake, ids0=ids0, ids1=ids1, level=ref.L - 1, nres=1)],
                     use_graph=False)
    ws = p.workspace_bytes
    p.close()
    ku = len(set(ids0) | set(ids1))
    return ws + ((2 + len(ids0) + len(ids1)) * ref.L + (1 + ku) * (ref.L - 1)) * ref.N * 8


def test_admission_by_device_memory(classes):
    """A pool that holds one request's block but not two: every request still
    completes with the right result, one batch at a time, and a request larger
    than the whole pool fails with MKHIP_ERR_OOM."""
    ctx, ref = classes["ckks"][0], classes["ckks"][1]
    ids0, ids1 = [0, 1, 2], [3, 4]
    one = _block_bytes(ctx, ref, ids0, ids1)
    big = _block_bytes(ctx, ref, list(range(NPARTY)), list(range(NPARTY)))
    pool = int(1.6 * one)
    assert big > pool
    s, cl = _server(classes, streams=4, max_batch=16, mem_bytes=pool)
    g = O.Prng(94)
    reqs = [_request("ckks", cl["ckks"], ref, g, ids0, ids1) for _ in range(6)]
    tickets = [_submit(s, r) for r in reqs]
    for t in tickets:
        assert s.wait(t)[0] == 0
    for r in reqs[:3]:
        assert np.array_equal(r["out"], _expected(classes, r))
    st = s.stats()
    assert st["completed"] == len(reqs) and st["max_batch"] == 1 and st["max_inflight"] == 1
    r = _request("ckks", cl["ckks"], ref, g, list(range(NPARTY)), list(range(NPARTY)))
    status, _, _ = s.wait(_submit(s, r))
    assert status == -7  # MKHIP_ERR_OOM
    s.close()


def test_device_resident_requests_in_a_mix(classes):
    """Requests whose operands and result stay in HBM (served in place, no host
    copies) give the host requests' words, batched with host requests of both
    schemes; two are checked against the oracle."""
    s, cl = _server(classes, streams=2, max_batch=8, mem_bytes=2 << 30)
    g = O.Prng(95)
    rng = np.random.default_rng(95)
    probes = [_request("ckks", cl["ckks"], classes["ckks"][1], g, [0, 1, 3], [1, 2]),
              _request("bfv", cl["bfv"], classes["bfv"][1], g, [2, 4], [0, 4, 5])]
    dev = []
    for r in probes:
        ctx = classes[r["kind"]][0]
        d = [ctx.alloc(a.nbytes) for a in (r["in0"], r["in1"], r["out"])]
        ctx.h2d(d[0], r["in0"])
        ctx.h2d(d[1], r["in1"])
        dev.append((ctx, d))
    mix = _mix(cl, classes, g, rng, 10)
    tickets = [_submit(s, r) for r in mix[:5]]
    for r, (ctx, d) in zip(probes, dev):
        tickets.append(s.enqueue_device(r["cls"], d[0], d[1], d[2], ids0=r["ids0"], ids1=r["ids1"],
                                        nres=r["nres"]))
    tickets += [_submit(s, r) for r in mix[5:]]
    for t in tickets:
        assert s.wait(t)[0] == 0
    for r, (ctx, d) in zip(probes, dev):
        got = ctx.d2h(np.zeros_like(r["out"]), d[2])
        assert np.array_equal(got, _expected(classes, r)), r["kind"]
        for p in d:
            ctx.free(p)
    for r in mix[:2]:
        assert np.array_equal(r["out"], _expected(classes, r)), r["kind"]
    assert s.stats()["max_batch"] >= 2
    s.close()
