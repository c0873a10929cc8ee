"""Reader, writer and GPU checker for the reference's bit-exact fixture directories.

The reference's Go tests export one directory per scheme: a manifest.json plus
little-endian uint64 payloads, modulus-major then coefficient-major:

  KKLSS-CKKS  software/kklss/mkckks/fpga_fixture_test.go:27-45 (manifest), :46-160
              (writers), :198-240 (MulRelinNew(left, right) of 16 users)
  KKLSS-BFV   software/kklss/mkbfv/fpga_fixture_test.go:23-42, :55-118, :201-236
              (MulRelinNew(input, input))
  MKF-CKKS /  software/mkf/mkf_cpu/fpga_fixture_test.go:24-43, :45-129, :299-390
  MKF-BFV     (CKKSCase.Run / BFVCase.Run; output c0, Gamma, Gamma per party)

`load(dir)` reads (and SHA-256-checks) a directory, `check(dir)` runs it through
the HIP library and compares bit for bit -- trying the two primitive-root rules
the library supports, so a Go-generated fixture also tells which rule Lattigo
used -- and the `save_*` functions write directories in the same format (the
tests use them to build small fixtures from the oracle).

    python -m mkf_amd.fixtures <fixture_dir> [...]
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT_RULES = (0, 1)  # 0: smallest generator >= 3 (default), 1: >= 2 (MKHIP_ROOT_RULE_SMALLEST2)


class FixtureError(ValueError):
    pass


def _sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 22), b""):
            h.update(chunk)
    return h.hexdigest()


def _read(dirpath, name, shape, man, verify):
    path = os.path.join(dirpath, name)
    if verify and man.get("sha256", {}).get(name) not in (None, _sha256(path)):
        raise FixtureError("%s: SHA-256 does not match the manifest" % name)
    a = np.fromfile(path, dtype="<u8")
    n = int(np.prod(shape))
    if a.size != n:
        raise FixtureError("%s: %d words, expected %d for shape %s" % (name, a.size, n, shape))
    return a.reshape(shape).astype(np.uint64, copy=False)


def _digits(dirpath, name, per_digit_words):
    size = os.path.getsize(os.path.join(dirpath, name)) // 8
    if size % per_digit_words:
        raise FixtureError("%s: size is not a whole number of key digits" % name)
    return size // per_digit_words


def load(dirpath, verify=True):
    """The fixture as a dict: scheme, logN, Q, P (QMul, t), parties, levels and arrays."""
    with open(os.path.join(dirpath, "manifest.json")) as f:
        man = json.load(f)
    scheme = man["scheme"]
    N = int(man["polynomial_degree"])
    logN = N.bit_length() - 1
    Q, P = [int(x) for x in man["q"]], [int(x) for x in man["p"]]
    L, K, n = len(Q), len(P), int(man["parties"])
    lin, lout = int(man["input_level_q"]), int(man["output_level_q"])
    fx = dict(scheme=scheme, manifest=man, logN=logN, N=N, Q=Q, P=P, parties=n, in_level=lin, out_level=lout)
    row = (L + K) * N
    if scheme == "KKLSS-CKKS":
        fx["input"] = _read(dirpath, man["input_ciphertexts_file"], (2, 1 + n, lin + 1, N), man, verify)
        beta = _digits(dirpath, man["relinearization_key_file"], n * 3 * row)
        fx["keys"] = _read(dirpath, man["relinearization_key_file"], (n, 3, beta, L + K, N), man, verify)
        fx["crs_u"] = _read(dirpath, man["crs_u_file"], (beta, L + K, N), man, verify)
        fx["expected"] = _read(dirpath, man["output_ciphertext_file"], (1 + n, lout + 1, N), man, verify)
    elif scheme == "KKLSS-BFV":
        fx["QMul"], fx["t"] = [int(x) for x in man["q_mul"]], int(man["plaintext_modulus"])
        fx["input"] = _read(dirpath, man["input_ciphertext_file"], (1 + n, lin + 1, N), man, verify)
        beta = _digits(dirpath, man["relinearization_key_file"], n * 2 * 3 * row)
        fx["keys"] = _read(dirpath, man["relinearization_key_file"], (n, 2, 3, beta, L + K, N), man, verify)
        fx["crs_u"] = _read(dirpath, man["crs_u_file"], (beta, L + K, N), man, verify)
        fx["expected"] = _read(dirpath, man["output_ciphertext_file"], (1 + n, lout + 1, N), man, verify)
    elif scheme in ("MKF-CKKS", "MKF-BFV"):
        if scheme == "MKF-BFV":
            fx["QMul"], fx["t"] = [int(x) for x in man["q_mul"]], int(man["plaintext_modulus"])
        fx["input"] = _read(dirpath, man["input_file"], (2, 1 + n, lin + 1, N), man, verify)
        beta = _digits(dirpath, man["switching_key_file"], (n + 1) * 2 * row)
        keys = _read(dirpath, man["switching_key_file"], (n + 1, beta, 2, L + K, N), man, verify)
        fx["fold_keys"], fx["relin_key"] = keys[:n], keys[n]
        fx["expected"] = _read(dirpath, man["output_file"], (2 + n, lout + 1, N), man, verify)
    else:
        raise FixtureError("unknown scheme %r" % scheme)
    return fx


# ---------------------------------------------------------------------------
# writers (the reference's layout and manifest fields)

def _write(dirpath, name, arr, hashes):
    path = os.path.join(dirpath, name)
    np.ascontiguousarray(arr, dtype="<u8").tofile(path)
    hashes[name] = _sha256(path)


def _manifest(dirpath, man, hashes):
    man["sha256"] = hashes
    with open(os.path.join(dirpath, "manifest.json"), "w") as f:
        json.dump(man, f, indent=2)
        f.write("\n")


def save_kklss_ckks(dirpath, logN, Q, P, left, right, keys, crs_u, expected, parameter="synthetic"):
    """left/right [(1+n)][nl][N], keys [n][3][beta][L+K][N], crs_u [beta][L+K][N], expected [(1+n)][nl'][N]."""
    os.makedirs(dirpath, exist_ok=True)
    n = left.shape[0] - 1
    ids = ["user%d" % i for i in range(n)]
    h = {}
    _write(dirpath, "input_ciphertexts_q.bin", np.stack([left, right]), h)
    _write(dirpath, "relinearization_keys_qp.bin", keys, h)
    _write(dirpath, "crs_u_qp.bin", crs_u, h)
    _write(dirpath, "expected_ciphertext_q.bin", expected, h)
    _manifest(dirpath, {
        "scheme": "KKLSS-CKKS", "parameter": parameter, "parties": n, "polynomial_degree": 1 << logN,
        "q": [int(x) for x in Q], "p": [int(x) for x in P], "input_level_q": left.shape[1] - 1,
        "output_level_q": expected.shape[1] - 1, "beta": int(keys.shape[2]), "ciphertext_order": ["0"] + ids,
        "switching_key_order": "user, relin_component[0..2], decomposition[0..beta-1], Q_then_P, modulus_major, "
                               "coefficient_major",
        "input_ciphertexts_file": "input_ciphertexts_q.bin", "input_order": "side[left,right], component[0,user0..], "
                                                                           "modulus, coefficient",
        "relinearization_key_file": "relinearization_keys_qp.bin", "crs_u_file": "crs_u_qp.bin",
        "output_ciphertext_file": "expected_ciphertext_q.bin"}, h)


def save_kklss_bfv(dirpath, logN, Q, QMul, P, t, ct, keys, crs_u, expected, parameter="synthetic"):
    """ct [(1+n)][L][N] (multiplied by itself), keys [n][2][3][beta][L+K][N], expected [(1+n)][L][N]."""
    os.makedirs(dirpath, exist_ok=True)
    n = ct.shape[0] - 1
    h = {}
    _write(dirpath, "input_ciphertext_q.bin", ct, h)
    _write(dirpath, "relinearization_keys_qp.bin", keys, h)
    _write(dirpath, "crs_u_qp.bin", crs_u, h)
    _write(dirpath, "expected_ciphertext_q.bin", expected, h)
    _manifest(dirpath, {
        "scheme": "KKLSS-BFV", "parameter": parameter, "parties": n, "polynomial_degree": 1 << logN,
        "q": [int(x) for x in Q], "q_mul": [int(x) for x in QMul], "p": [int(x) for x in P],
        "plaintext_modulus": int(t), "input_level_q": ct.shape[1] - 1, "output_level_q": expected.shape[1] - 1,
        "beta": int(keys.shape[3]), "ciphertext_order": ["0"] + ["user%d" % i for i in range(n)],
        "switching_key_order": "user, BFV_branch[0..1], relin_component[0..2], decomposition[0..beta-1], Q_then_P, "
                               "modulus_major, coefficient_major",
        "input_ciphertext_file": "input_ciphertext_q.bin", "relinearization_key_file": "relinearization_keys_qp.bin",
        "crs_u_file": "crs_u_qp.bin", "output_ciphertext_file": "expected_ciphertext_q.bin"}, h)


def save_mkf(dirpath, scheme, logN, Q, P, left, right, fold_keys, relin_key, expected, QMul=None, t=0,
             parameter="synthetic"):
    """scheme MKF-CKKS / MKF-BFV; left/right [(1+n)][nl][N]; keys [beta][2][L+K][N];
    expected [(2+n)][nl'][N] = c0, Gamma, Gamma per party."""
    os.makedirs(dirpath, exist_ok=True)
    n = left.shape[0] - 1
    h = {}
    _write(dirpath, "input_ciphertexts_q.bin", np.stack([left, right]), h)
    _write(dirpath, "fold_and_relin_keys_qp.bin", np.concatenate([np.stack(fold_keys), relin_key[None]]), h)
    _write(dirpath, "expected_output_q.bin", expected, h)
    man = {"scheme": scheme, "parameter": parameter, "parties": n, "polynomial_degree": 1 << logN,
           "q": [int(x) for x in Q], "p": [int(x) for x in P], "input_level_q": left.shape[1] - 1,
           "output_level_q": expected.shape[1] - 1, "beta": int(relin_key.shape[0]),
           "input_order": "side[left,right], component[c0,party0..], modulus, coefficient",
           "switching_key_order": "fold_key[party0..], then aggregate_relin_key; decomposition, output_poly[0..1], "
                                  "Q_then_P, modulus, coefficient",
           "output_order": "c0, gamma, repeated_gamma[party0..], modulus, coefficient",
           "input_file": "input_ciphertexts_q.bin", "switching_key_file": "fold_and_relin_keys_qp.bin",
           "output_file": "expected_output_q.bin"}
    if QMul is not None:
        man["q_mul"], man["plaintext_modulus"] = [int(x) for x in QMul], int(t)
    _manifest(dirpath, man, h)


# ---------------------------------------------------------------------------
# running a fixture through the HIP library

def _dev(ctx, arr, bufs):
    d = ctx.alloc(arr.nbytes)
    ctx.h2d(d, np.ascontiguousarray(arr))
    bufs.append(d)
    return d


def run(fx, root_rule=0):
    """The library's output for the fixture's inputs, shaped like fx['expected']."""
    import mkf_amd

    s, n, N, L = fx["scheme"], fx["parties"], fx["N"], len(fx["Q"])
    lin, lout = fx["in_level"], fx["out_level"]
    flags = 1 if root_rule == 1 else 0
    ids = list(range(n))
    bufs = []
    if s == "KKLSS-CKKS":
        ctx = mkf_amd.Context(fx["logN"], fx["Q"], fx["P"], gamma=2, flags=flags)
        try:
            for p in ids:
                ctx.register_party(p, *[np.ascontiguousarray(fx["keys"][p, c]) for c in range(3)])
            ctx.set_crs_u(np.ascontiguousarray(fx["crs_u"]))
            out = np.zeros((1 + n, lout + 1, N), np.uint64)
            op = dict(op0=_dev(ctx, fx["input"][0], bufs), op1=_dev(ctx, fx["input"][1], bufs),
                      out=_dev(ctx, out, bufs), ids0=ids, ids1=ids, level=lin, nres=lin - lout)
            plan = mkf_amd.Plan(ctx, [op], use_graph=False, scheme="kklss")
            plan.launch()
            ctx.sync()
            ctx.d2h(out, op["out"])
            plan.close()
            return out
        finally:
            for b in bufs:
                ctx.free(b)
            ctx.close()
    if s == "KKLSS-BFV":
        ctx = mkf_amd.Context(fx["logN"], fx["Q"], fx["P"], gamma=2, t=fx["t"], qmul=fx["QMul"], flags=flags)
        try:
            k = fx["keys"]
            for p in ids:  # b, d of both branches (2 beta digits), v of branch 0
                ctx.register_bfv_party(p, np.concatenate([k[p, 0, 0], k[p, 1, 0]]),
                                       np.concatenate([k[p, 0, 1], k[p, 1, 1]]), np.ascontiguousarray(k[p, 0, 2]))
            ctx.set_crs_u(np.ascontiguousarray(fx["crs_u"]))
            out, _ = ctx.kklss_bfv_mulrelin(fx["input"], ids, fx["input"], ids)
            return out
        finally:
            ctx.close()
    if s in ("MKF-CKKS", "MKF-BFV"):
        bfv = s == "MKF-BFV"
        ctx = mkf_amd.Context(fx["logN"], fx["Q"], fx["P"], gamma=1, t=fx.get("t", 0), qmul=fx.get("QMul"),
                              flags=flags)
        try:
            ctx.register_mkf_group(0, [np.ascontiguousarray(x) for x in fx["fold_keys"]],
                                   np.ascontiguousarray(fx["relin_key"]))
            if bfv:
                res = ctx.mkf_bfv_mulrelin(0, fx["input"][0], fx["input"][1])
            else:
                res = np.zeros((1 + n, lout + 1, N), np.uint64)
                op = dict(left=_dev(ctx, fx["input"][0], bufs), right=_dev(ctx, fx["input"][1], bufs),
                          out=_dev(ctx, res, bufs), group=0, level=lin, nres=lin - lout)
                plan = mkf_amd.Plan(ctx, [op], use_graph=False, scheme="mkf")
                plan.launch()
                ctx.sync()
                ctx.d2h(res, op["out"])
                plan.close()
            # c0, Gamma, then Gamma once per party (the reference's output order)
            return np.concatenate([res[:2], res[1:]])
        finally:
            for b in bufs:
                ctx.free(b)
            ctx.close()
    raise FixtureError("unknown scheme %r" % s)


def check(dirpath, verify=True):
    """(matched root rule or None, number of mismatching words under each rule tried)."""
    fx = load(dirpath, verify)
    diffs = {}
    for rule in ROOT_RULES:
        got = run(fx, rule)
        if got.shape != fx["expected"].shape:
            raise FixtureError("output shape %s, fixture %s" % (got.shape, fx["expected"].shape))
        diffs[rule] = int(np.count_nonzero(got != fx["expected"]))
        if diffs[rule] == 0:
            return rule, diffs
    return None, diffs


def main(argv):
    if not argv:
        print(__doc__)
        return 2
    bad = 0
    for d in argv:
        rule, diffs = check(d)
        if rule is None:
            bad += 1
            print("MISMATCH %s: differing words per root rule %s" % (d, diffs))
        else:
            print("ok %s: bit-exact (primitive-root rule %d)" % (d, rule))
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
