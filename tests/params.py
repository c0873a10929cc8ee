"""Parameter sets used by the tests and the benchmark.

The benchmark workloads use synthetic NTT-friendly primes (~50-bit Q, ~60-bit P,
q = 1 mod 2^16 so that every N up to 2^15 is supported); they are generated
deterministically by scanning down from 2^bits + 1.  P1 is the reference's own
KKLSS/MKF parameter literal (software/mkf/mkf_cpu/params.go:15-28,
software/kklss/mkckks/mkckks_test.go:75-89).
"""
import oracle as O

_cache = {}


def primes(bits, count, exclude=()):
    key = (bits, count, tuple(exclude))
    if key not in _cache:
        _cache[key] = O.gen_primes(bits, 15, count, exclude)
    return _cache[key]


def config(name):
    """(logN, Q, P) of a named workload."""
    if name == "c1":  # plain MK-CKKS, k=2, N=2^13, 4 Q + 1 P (MKF only)
        return 13, primes(50, 4), primes(60, 1)
    if name == "c2":  # batched MK-CKKS, k=4, N=2^14, 8 Q + 2 P
        return 14, primes(50, 8), primes(60, 2)
    if name == "small":  # 4 Q + 2 P at N=2^13 (KKLSS-compatible variant of c1)
        return 13, primes(50, 4), primes(60, 2)
    if name == "n15":  # 3 Q + 2 P at N=2^15 (limbs split over two workgroups)
        return 15, primes(50, 3), primes(60, 2)
    if name == "P1":
        return 14, [0x7ffffffffe70001, 0xffffffff00001, 0xfffffffe40001, 0xfffffffe20001,
                    0xfffffffbe0001, 0xfffffffa60001], [0xffffffffffc0001, 0xfffffffff840001]
    raise KeyError(name)


# KKLSS MK-BFV parameter literal PN14QP439 (software/kklss/mkbfv/mkbfv_test.go:76-107):
# N = 2^14, 6 Q primes (5 x 53 + 1 x 54 bit), 6 QMul primes, 2 x 60-bit P, t = 65537.
PN14QP439 = dict(
    logN=14,
    Q=[0x1fffffffe30001, 0x1fffffffd10001, 0x1fffffffbf0001, 0x1fffffffb60001, 0x1fffffff920001,
       0x3fffffffd60001],
    QMul=[0x1fffffffd80001, 0x1fffffffc50001, 0x1fffffffb90001, 0x1fffffffa50001, 0x1fffffff900001,
          0x3fffffffca0001],
    P=[0xffffffffffc0001, 0xfffffffff840001],
    t=65537,
)


def bfv_config(name):
    """(logN, Q, QMul, P, t) of a named MK-BFV workload."""
    if name == "b3":  # the MK-BFV workload: k=4, N=2^14, 6 Q + 6 QMul + 2 P, t=65537
        c = PN14QP439
        return c["logN"], c["Q"], c["QMul"], c["P"], c["t"]
    if name == "bsmall":  # 3 Q + 3 QMul + 2 P at N=2^13
        Q = primes(50, 3)
        return 13, Q, primes(50, 3, exclude=Q), primes(60, 2), 65537
    if name == "b15":  # 2 Q + 2 QMul + 2 P at N=2^15
        Q = primes(50, 2)
        return 15, Q, primes(50, 2, exclude=Q), primes(60, 2), 65537
    raise KeyError(name)
