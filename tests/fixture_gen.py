"""Small fixture directories in the reference's export format (manifest.json +
uint64 payloads, software/kklss/mkckks/fpga_fixture_test.go, mkbfv/
fpga_fixture_test.go, mkf/mkf_cpu/fpga_fixture_test.go), with the expected
outputs computed by the oracle.  Keys and ciphertexts are uniform residues."""
import numpy as np

import oracle as O
from params import bfv_config, config


def kklss_ckks(path, n=2, seed=7, root_rule=0, name="small"):
    from mkf_amd import fixtures

    logN, Q, P = config(name)
    ref = O.Ctx(logN, Q, P, gamma=2, root_rule=root_rule)
    g = O.Prng(seed)
    beta = ref.beta()
    keys = np.stack([np.stack([np.stack([g.uniform(ref.QP, ref.N) for _ in range(beta)]) for _ in range(3)])
                     for _ in range(n)])
    u = np.stack([g.uniform(ref.QP, ref.N) for _ in range(beta)])
    left = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + n)])
    right = np.stack([g.uniform(ref.Q, ref.N) for _ in range(1 + n)])
    ids = list(range(n))
    out, _ = ref.kklss_mulrelin(left, ids, right, ids, [keys[p, 0] for p in ids], [keys[p, 1] for p in ids],
                                [keys[p, 2] for p in ids], u)
    exp = np.stack([ref.div_round_last(out[i], 1) for i in range(out.shape[0])])
    fixtures.save_kklss_ckks(path, logN, Q, P, left, right, keys, u, exp)
    return exp


def kklss_bfv(path, n=2, seed=8, name="bsmall"):
    from mkf_amd import fixtures

    logN, Q, QM, P, t = bfv_config(name)
    ref = O.BfvCtx(logN, Q, QM, P, t)
    g = O.Prng(seed)
    L, QP, N = ref.L, ref.QP, ref.N
    keys = np.stack([np.stack([np.stack([np.stack([g.uniform(QP, N) for _ in range(L)]) for _ in range(3)])
                               for _ in range(2)]) for _ in range(n)])
    u = np.stack([g.uniform(QP, N) for _ in range(L)])
    ct = np.stack([g.uniform(ref.Q, N) for _ in range(1 + n)])
    ids = list(range(n))
    b = [np.concatenate([keys[p, 0, 0], keys[p, 1, 0]]) for p in ids]
    d = [np.concatenate([keys[p, 0, 1], keys[p, 1, 1]]) for p in ids]
    v = [keys[p, 0, 2] for p in ids]
    exp, _ = ref.mulrelin(ct, ids, ct, ids, b, d, v, u)
    fixtures.save_kklss_bfv(path, logN, Q, QM, P, t, ct, keys, u, exp)
    return exp


def mkf(path, bfv=False, n=2, seed=9):
    from mkf_amd import fixtures

    if bfv:
        logN, Q, QM, P, t = bfv_config("bsmall")
        ref = O.BfvCtx(logN, Q, QM, P, t, gamma=1)
        qp = ref.qp
    else:
        logN, Q, P = config("c1")
        QM, t = None, 0
        ref = qp = O.Ctx(logN, Q, P, gamma=1)
    g = O.Prng(seed)
    beta = qp.beta()

    def swk():
        return np.stack([np.stack([g.uniform(qp.QP, qp.N), g.uniform(qp.QP, qp.N)]) for _ in range(beta)])

    fold, relin = [swk() for _ in range(n)], swk()
    left = np.stack([g.uniform(qp.Q, qp.N) for _ in range(1 + n)])
    right = np.stack([g.uniform(qp.Q, qp.N) for _ in range(1 + n)])
    if bfv:
        c0, gamma = ref.mkf_mulrelin(left, right, fold, relin)
    else:
        o0, o1 = ref.mkf_ckks_mulrelin(left, right, fold, relin)
        c0, gamma = ref.div_round_last(o0, 1, ntt=True), ref.div_round_last(o1, 1, ntt=True)
    exp = np.stack([c0, gamma] + [gamma] * n)
    fixtures.save_mkf(path, "MKF-BFV" if bfv else "MKF-CKKS", logN, Q, P, left, right, fold, relin, exp, QMul=QM,
                      t=t)
    return exp
