"""Benchmark: KKLSS MK-CKKS mul+relin per second at N=2^14, k=4 (batched vs plain).

Headline workload: N = 2^14, 8 Q primes (~50-bit) + 2 P primes (~60-bit, gadget
digits over the Q limbs, alpha = 1, beta = 8), k = 4 parties with evaluation
keys resident in HBM, a batch of 64 independent ciphertext pairs (uniform
residues, seeded) per step, one hipGraph launch per step.  A step is one
mul+relin+rescale of the whole batch; `value` is operations per second over the
timed steps (inputs already resident in HBM).

Also reported, each with its batched and plain (one op per launch) rates and
per-stage device times:
  mkf_ckks   MKF MK-CKKS at the same shape (gamma 1: alpha = 2, beta = 4), B = 64
  kklss_bfv  KKLSS MK-BFV, N = 2^14, k = 4, t = 65537, 6 Q + 6 QMul + 2 P
             (the reference's PN14QP439 primes), B = 32
  mkf_bfv    MKF MK-BFV at the MK-BFV shape (gamma 1), B = 32
plus the roofline of the headline's dominant kernel (HIP events on the library's
stream), the CPU oracle timed on a bounded sample of the headline workload, and
the Galois key switches (KKLSS hoisted rotations, conjugation, MKF rotation) at
the headline shape.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]
                       [--schemes kklss_ckks,mkf_ckks,kklss_bfv,mkf_bfv]
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

# Each server stream needs a hardware queue of its own: HIP's default of 4 queues
# per process makes the serving legs' streams share queues, and a batch then waits
# behind another stream's batch (p99 of 30-120 ms at loads where 16 queues give
# 5-8 ms).  Raised before the HIP runtime starts; a larger setting is kept.
try:
    _hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
except ValueError:
    _hwq = 4
if _hwq < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MK-CKKS mul+relin/s at N=2^14 k=4 (batched vs plain); p50/p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
ALL_SCHEMES = ["kklss_ckks", "mkf_ckks", "kklss_bfv", "mkf_bfv"]
TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic_kklss_ckks_mulrelin_n14_l8_k4_b64.json")
# the library's tuning knobs by the names the bench reports them under
TUNING_KNOBS = {"fused_decomp": 1, "xcd_remap": 2, "decomp_order": 3, "fp64": 4, "split_limb": 5, "split_ext": 6,
                "digit_pre": 7, "xcd_ext": 8, "ext_cs": 9, "inv_lazy": 10}


def parse_tuning(text):
    """'decomp_order=1,split_limb=1' -> {knob: value} (knob names of TUNING_KNOBS)."""
    out = {}
    for item in filter(None, text.split(",")):
        k, v = item.split("=")
        if k not in TUNING_KNOBS:
            raise argparse.ArgumentTypeError("unknown tuning knob %r" % k)
        out[k] = int(v)
    return out

# The reference's MK-BFV parameter literal PN14QP439 (software/kklss/mkbfv/mkbfv_test.go:76-107)
BFV_Q = [0x1fffffffe30001, 0x1fffffffd10001, 0x1fffffffbf0001, 0x1fffffffb60001, 0x1fffffff920001,
         0x3fffffffd60001]
BFV_QMUL = [0x1fffffffd80001, 0x1fffffffc50001, 0x1fffffffb90001, 0x1fffffffa50001, 0x1fffffff900001,
            0x3fffffffca0001]
BFV_P = [0xffffffffffc0001, 0xfffffffff840001]
BFV_T = 65537


def ntt_primes(bits, count):
    """NTT-friendly primes q = 1 mod 2^16 scanning down from 2^bits + 1 (deterministic)."""
    out, x, step = [], (1 << bits) + 1, 1 << 16
    while len(out) < count:
        x -= step
        if _is_prime(x):
            out.append(x)
    return out


def _is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def uniform(rng, moduli, N):
    return np.stack([rng.integers(0, q, size=N, dtype=np.uint64) for q in moduli])


def pct(xs, p):
    return float(np.percentile(np.asarray(xs), p))


def stage_alg_bytes(B, k, L, K, beta, N, W=8):
    """Algorithmic HBM bytes per launch of each KKLSS MK-CKKS plan stage, every
    datum counted once: an op's input limbs once per (op, side, party), the
    parties' keys once per launch (the batch shares them), the x / y keys once
    per op (written once, read once), every output once.  In limbs of N W bytes:

      decompose_mac   inputs 2kL B + keys (d, b) 2k beta rows + digits written
                      2k beta rows B + x / y written 2 beta rows B
      decompose       inputs + digits;  mac_xy: digits + keys + x / y
      tensor_ntt      the two c0s read and written (4 L B)
      tensor          c0s + the parties' diagonal digit rows (2 L + 2k L) B
                      read, (1 + k) L B written
      ext_xy          digits of both sides 2k beta rows B + x / y 2 beta rows B
                      read, 2k rows B accumulators written
      moddown_xy      2k rows B read + k L B accumulation targets read,
                      2k L B written
      decompose_t     t_i k L B read, digits k beta rows B written
      ext_vu          digits k beta rows B + keys (v_i, u) (k + 1) beta rows
                      read, 2k rows B written
      moddown_rescale 2k rows B + (1 + k) L B read, (1 + k)(L - 1) B written"""
    rows = L + K
    limb = N * W
    dig = 2 * k * beta * rows
    return {
        "decompose_mac": (2 * k * L * B + dig + dig * B + 2 * beta * rows * B) * limb,
        "decompose": (2 * k * L * B + dig * B) * limb,
        "mac_xy": (dig * B + dig + 2 * beta * rows * B) * limb,
        "tensor_ntt": 4 * L * B * limb,
        "tensor": ((2 + 2 * k) * L + (1 + k) * L) * B * limb,
        "ext_xy": (dig + 2 * beta * rows + 2 * k * rows) * B * limb,
        "moddown_xy": (2 * k * rows + k * L + 2 * k * L) * B * limb,
        "decompose_t": (k * L + k * beta * rows) * B * limb,
        "ext_vu": ((k * beta * rows + 2 * k * rows) * B + (k + 1) * beta * rows) * limb,
        "moddown_rescale": (2 * k * rows + (1 + k) * L + (1 + k) * (L - 1)) * B * limb,
    }


def stage_valu_work(B, k, L, K, beta, N):
    """Modular-arithmetic work per launch of the transform-bearing KKLSS stages, by
    the form the kernels use at the headline shape (Q primes below 2^57, P primes
    of 60 bits): (forward butterflies on Q rows -- unreduced --, butterflies of
    the reduced form -- forward and inverse on P rows --, inverse butterflies on
    Q rows -- unreduced --, Montgomery products -- the fused x / y MAC, the
    tensor products --, products summed in carry-save form -- the external
    products).  A limb transform is N/2 log2(N) butterflies."""
    rows = L + K
    nb = N // 2 * (N.bit_length() - 1)
    dig = 2 * k * beta
    return {
        "decompose_mac": (dig * L * nb * B, dig * K * nb * B, 0, dig * rows * N * B, 0),
        "decompose": (dig * L * nb * B, dig * K * nb * B, 0, 0, 0),
        "mac_xy": (0, 0, 0, dig * rows * N * B, 0),
        "tensor_ntt": (2 * L * nb * B, 0, 0, 0, 0),
        "tensor": (0, 0, (1 + k) * L * nb * B, (1 + 2 * k) * L * N * B, 0),
        "ext_xy": (0, 2 * k * K * nb * B, 2 * k * L * nb * B, 0, dig * rows * N * B),
        "decompose_t": (k * beta * L * nb * B, k * beta * K * nb * B, 0, 0, 0),
        "ext_vu": (0, 2 * k * K * nb * B, 2 * k * L * nb * B, 0, dig * rows * N * B),
    }


def measure_peaks(iters=5):
    """This GPU's peaks, measured: device copy bandwidth over 1 GiB buffers and the
    in-register issue rates of the integer primitives (mkhip_probe_peak)."""
    import mkf_amd

    ctx = mkf_amd.Context(13, ntt_primes(50, 2), ntt_primes(60, 1), gamma=1)
    try:
        lib_copy = ctx.probe_peak(mkf_amd.PROBE_COPY, 1 << 30, iters) / 1e9
        stream_copy = ctx.probe_peak(mkf_amd.PROBE_COPY_STREAM, 1 << 30, iters) / 1e9
        # the measured HBM peak: the faster of the library's copy and the streaming copy
        out = {"copy_gbs": round(max(lib_copy, stream_copy), 1), "copy_kernel_gbs": round(lib_copy, 1),
               "copy_stream_gbs": round(stream_copy, 1)}
        for name, kind in (("mul_lo_u32_per_s", mkf_amd.PROBE_MUL32), ("mad_u64_u32_per_s", mkf_amd.PROBE_MAD64),
                           ("bfly_lazy_per_s", mkf_amd.PROBE_BFLY_LAZY), ("bfly_reduced_per_s", mkf_amd.PROBE_BFLY_RED),
                           ("mred_add_per_s", mkf_amd.PROBE_MRED), ("acc3_mac_per_s", mkf_amd.PROBE_ACC3),
                           ("ibfly_lazy_per_s", mkf_amd.PROBE_IBFLY_LAZY)):
            out[name] = float("%.4g" % ctx.probe_peak(kind, 4096, iters))
    finally:
        ctx.close()
    return out


def op_alg_bytes(scheme, B, k, L, K, beta, N, W=8):
    """Algorithmic bytes per op: inputs + outputs + keys shared by the batch."""
    b_in = 2 * (k + 1) * L * N * W
    if scheme == "kklss_ckks":
        return b_in + (k + 1) * (L - 1) * N * W + (3 * k + 1) * beta * (L + K) * N * W / B
    if scheme == "mkf_ckks":
        return b_in + (k + 1) * (L - 1) * N * W + (k + 1) * beta * 2 * (L + K) * N * W / B
    if scheme == "kklss_bfv":
        return b_in + (k + 1) * L * N * W + (5 * k + 1) * beta * (L + K) * N * W / B
    return b_in + (k + 1) * L * N * W + (k + 1) * beta * 2 * (L + K) * N * W / B


class Workload:
    """One scheme at one shape: context, resident keys, batched + plain plans."""

    def __init__(self, scheme, rng, batch):
        import mkf_amd

        self.scheme, self.B = scheme, batch
        self.k = 4
        if scheme.endswith("ckks"):
            self.logN, self.L, self.K = 14, 8, 2
            self.Q, self.P, self.QM, self.t = ntt_primes(50, 8), ntt_primes(60, 2), None, 0
        else:
            self.logN, self.L, self.K = 14, len(BFV_Q), len(BFV_P)
            self.Q, self.P, self.QM, self.t = BFV_Q, BFV_P, BFV_QMUL, BFV_T
        self.N = 1 << self.logN
        gamma = 2 if scheme.startswith("kklss") else 1
        self.ctx = mkf_amd.Context(self.logN, self.Q, self.P, gamma=gamma, t=self.t, qmul=self.QM)
        self.beta = self.ctx.beta
        QP, N, k = self.Q + self.P, self.N, self.k
        self.keys, self.u = [], None
        if scheme == "kklss_ckks":
            for pid in range(k):
                kk = [uniform_stack(rng, QP, N, self.beta) for _ in range(3)]
                self.ctx.register_party(pid, *kk)
                self.keys.append(kk)
        elif scheme == "kklss_bfv":
            for pid in range(k):
                b = uniform_stack(rng, QP, N, 2 * self.beta)
                d = uniform_stack(rng, QP, N, 2 * self.beta)
                v = uniform_stack(rng, QP, N, self.beta)
                self.ctx.register_bfv_party(pid, b, d, v)
        else:
            fold = [np.stack([np.stack([uniform(rng, QP, N), uniform(rng, QP, N)]) for _ in range(self.beta)])
                    for _ in range(k)]
            relin = np.stack([np.stack([uniform(rng, QP, N), uniform(rng, QP, N)]) for _ in range(self.beta)])
            self.ctx.register_mkf_group(0, fold, relin)
        if scheme.startswith("kklss"):
            self.u = uniform_stack(rng, QP, N, self.beta)
            self.ctx.set_crs_u(self.u)
        self.nres = 0
        if scheme.endswith("ckks"):
            scale = float(self.Q[-1])
            self.nres, _ = self.ctx.rescale_count(self.L - 1, scale * scale, scale)
        self.bufs = []

    def make_ops(self, n, rng):
        L, N, k = self.L, self.N, self.k
        ids = list(range(k))
        in_words = (1 + k) * L * N
        out_words = (1 + k) * (L - self.nres) * N
        ops = []
        for _ in range(n):
            d0, d1, do = self.ctx.alloc(in_words * 8), self.ctx.alloc(in_words * 8), self.ctx.alloc(out_words * 8)
            self.ctx.h2d(d0, np.concatenate([uniform(rng, self.Q, N) for _ in range(1 + k)]))
            self.ctx.h2d(d1, np.concatenate([uniform(rng, self.Q, N) for _ in range(1 + k)]))
            self.bufs += [d0, d1, do]
            if self.scheme.startswith("kklss"):
                ops.append(dict(op0=d0, op1=d1, out=do, ids0=ids, ids1=ids, level=L - 1, nres=self.nres))
            else:
                ops.append(dict(left=d0, right=d1, out=do, group=0, level=L - 1, nres=self.nres))
        return ops

    def plan(self, ops):
        import mkf_amd

        sch = {"kklss_ckks": "kklss", "kklss_bfv": "bfv", "mkf_ckks": "mkf", "mkf_bfv": "mkf_bfv"}[self.scheme]
        return mkf_amd.Plan(self.ctx, ops, use_graph=True, scheme=sch)

    def close(self):
        for p in self.bufs:
            self.ctx.free(p)
        self.ctx.close()


def uniform_stack(rng, moduli, N, n):
    return np.stack([uniform(rng, moduli, N) for _ in range(n)])


def time_plan(ctx, plan, steps, warmup):
    for _ in range(warmup):
        plan.launch()
    ctx.sync()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        plan.launch()
        ctx.sync()
        lat.append(time.perf_counter() - s)
    return time.perf_counter() - t0, lat


def run_scheme(scheme, args, rng):
    B = args.batch if scheme.endswith("ckks") else args.bfv_batch
    w = Workload(scheme, rng, B)
    ops = w.make_ops(B, rng)
    sch = {"kklss_ckks": "kklss", "kklss_bfv": "bfv", "mkf_ckks": "mkf", "mkf_bfv": "mkf_bfv"}[scheme]
    if args.tuning and scheme == "kklss_ckks":
        for knob, val in args.tuning.items():
            w.ctx.set_tuning(TUNING_KNOBS[knob], val)
        tune_ms = []
    else:
        tune_ms = w.ctx.autotune(ops, scheme=sch, iters=args.autotune_iters) if args.autotune else []
    plan = w.plan(ops)
    elapsed, lat = time_plan(w.ctx, plan, args.steps, args.warmup)
    ops_s = B * args.steps / elapsed
    stage_ms = plan.stage_times(args.stage_iters)
    names = plan.stages
    per_run = [plan.stage_times(1) for _ in range(args.ref_runs)]
    res = {
        "ops_per_s": round(ops_s, 2), "ms_per_step": round(elapsed / args.steps * 1e3, 3), "batch": B,
        "p50_ms": round(pct(lat, 50) * 1e3, 3), "p99_ms": round(pct(lat, 99) * 1e3, 3),
        "stages_ms": {names[s]: round(stage_ms[s], 4) for s in range(len(names))},
        "workspace_gb": round(plan.workspace_bytes / 1e9, 3),
        "autotune_ms": [round(t, 3) for t in tune_ms],
        "tuning": {"fused_decomp": w.ctx.get_tuning(1), "xcd_remap": w.ctx.get_tuning(2),
                   "decomp_order": w.ctx.get_tuning(3), "split_limb": w.ctx.get_tuning(5),
                   "split_ext": w.ctx.get_tuning(6), "digit_pre": w.ctx.get_tuning(7),
                   "xcd_ext": w.ctx.get_tuning(8), "ext_cs": w.ctx.get_tuning(9),
                   "inv_lazy": w.ctx.get_tuning(10)},
        "shape": {"logN": w.logN, "L": w.L, "K": w.K, "k": w.k, "alpha": w.ctx.alpha, "beta": w.beta,
                  "nres": w.nres, "t": w.t},
    }
    res["reference_stages_ms"] = {g: round(t, 4) for g, t in ref_stage_split(names, stage_ms).items()}
    param = "N%dL%dK%d-B%d" % (w.logN, w.L, w.K, B)
    res["mkfbench_rows"] = mkfbench_rows(scheme, param, w.k, B, names, per_run, args.warmup)
    alg = op_alg_bytes(scheme, B, w.k, w.L, w.K, w.beta, w.N)
    res["alg_bytes_per_op"] = int(alg)
    res["op_gbs"] = round(alg * ops_s / 1e9, 1)
    if args.plain_ops > 0:
        pops = w.make_ops(1, rng)
        pplan = w.plan(pops)
        pel, plat = time_plan(w.ctx, pplan, args.plain_ops, 3)
        res["plain"] = {"ops_per_s": round(args.plain_ops / pel, 2), "p50_ms": round(pct(plat, 50) * 1e3, 3),
                        "p99_ms": round(pct(plat, 99) * 1e3, 3)}
        res["batched_over_plain"] = round(ops_s / res["plain"]["ops_per_s"], 3)
        pplan.close()
        if scheme == "kklss_ckks":
            res["plain_host"] = plain_host_rate(w, rng, args.plain_ops)
    if scheme == "kklss_ckks" and args.fp64_leg:
        # the same batch with the opt-in FP64 transforms for moduli below 2^50 (the
        # same words; the reported value is the integer form above)
        import mkf_amd

        w.ctx.set_tuning(mkf_amd.TUNE_FP64, 1)
        try:
            fplan = w.plan(ops)
            fel, flat = time_plan(w.ctx, fplan, args.steps, args.warmup)
            fst = fplan.stage_times(args.stage_iters)
            res["fp64_arithmetic"] = {
                "ops_per_s": round(B * args.steps / fel, 2), "ms_per_step": round(fel / args.steps * 1e3, 3),
                "p99_ms": round(pct(flat, 99) * 1e3, 3),
                "stages_ms": {fplan.stages[s_]: round(fst[s_], 4) for s_ in range(len(fst))}}
            fplan.close()
        finally:
            w.ctx.set_tuning(mkf_amd.TUNE_FP64, 0)
    if scheme == "kklss_ckks" and args.batch_sweep:
        res["batch_sweep"] = batch_sweep(w, ops, args, rng, res.get("plain"))
    extra = (w, plan, stage_ms, names)
    return res, extra


# The reference's stage names (software/mkf/stageprof/stageprof.go:17-25) and the
# plan stages whose device time each one covers.  The MKF profile fields of
# mkfbench (mkf_cpu/profile.go:8-19, attributed in common.go:79-139 and
# ckks.go:102-152) use the same split: Decomposition (DecomposeNTT, with its
# inverse transform of the NTT-domain input), External (KeyswitchHoistedNoModDown
# and the fold's additions), Modulus (ModDown, Rescale), Tensor, Memory (the k
# copies of Gamma).
REF_STAGES = {
    "tensor_product": ("tensor_ntt", "tensor", "karatsuba", "lift"),
    "gadget_decomposition": ("decompose_mac", "decompose", "decompose_t", "intt_in", "digit_pre", "intt_mu2",
                             "digit_pre_mu2", "decompose_mu2", "decompose_relin"),
    "external_product_accumulation": ("ext_xy", "ext_vu", "mac_xy", "mac_fold", "mac_relin"),
    "modup_moddown_rescale": ("moddown_xy", "moddown_rescale", "moddown_final", "moddown_fold", "moddown_relin",
                              "intt_last", "rescale", "quantize"),
    "memory_data_movement": ("materialize",),
}
REF_FIELD = {"tensor_product": "tensor_ns", "gadget_decomposition": "decomposition_ns",
             "external_product_accumulation": "external_product_ns",
             "modup_moddown_rescale": "modup_moddown_rescale_ns", "memory_data_movement": "memory_copy_ns"}
MKFBENCH_HEADER = ["scheme", "parameter", "parties", "run", "warmup", "gomaxprocs",
                   "total_ns", "tensor_ns", "decomposition_ns", "external_product_ns",
                   "modup_moddown_rescale_ns", "memory_copy_ns", "other_ns",
                   "decomposition_count", "external_product_poly_count",
                   "validation_passed", "expected", "observed", "max_error",
                   "go_version", "goos", "goarch", "timestamp_utc"]


def ref_stage_split(names, times):
    """Plan stage times grouped under the reference's stage names; stages no group
    names (the N = 2^15 finish passes) count as other_synchronization."""
    out = {k: 0.0 for k in list(REF_STAGES) + ["other_synchronization"]}
    for n_, t in zip(names, times):
        base = n_[:-len("_finish")] if n_.endswith("_finish") else None
        grp = next((g for g, members in REF_STAGES.items() if n_ in members), None)
        if grp is None or base is not None:
            grp = "other_synchronization"
        out[grp] += t
    return out


def switch_counts(scheme, k):
    """Decompositions and external-product polynomials per op, counted as the
    reference's profile counts them: MKF one decomposition and 2 output polys per
    switch, 2k fold switches + 1 relinearization (common.go:84-89); KKLSS the 2k
    operand decompositions + k for the t_i, and 4k external products."""
    if scheme.startswith("mkf"):
        return 2 * k + 1, 2 * (2 * k + 1)
    return 3 * k, 4 * k


def mkfbench_rows(scheme, param, k, B, names, runs, warmup):
    """Rows in the reference's mkfbench CSV format (cmd/mkfbench/main.go:17-24),
    one per timed launch; times are per op (launch time / B, device time of each
    stage from HIP events).  Validation columns are left empty: the bench does not
    decrypt (bit-exact parity is checked by the tests)."""
    import datetime

    ndec, next_ = switch_counts(scheme, k)
    rows = []
    label = {"kklss_ckks": "KKLSS-CKKS", "kklss_bfv": "KKLSS-BFV", "mkf_ckks": "MKF-CKKS", "mkf_bfv": "MKF-BFV"}[scheme]
    for i, times in enumerate(runs):
        g = ref_stage_split(names, times)
        per = {f: int(round(g[st] * 1e6 / B)) for st, f in REF_FIELD.items()}
        total = int(round(sum(times) * 1e6 / B))
        rows.append([label, param, k, i + 1, warmup, 1, total, per["tensor_ns"], per["decomposition_ns"],
                     per["external_product_ns"], per["modup_moddown_rescale_ns"], per["memory_copy_ns"],
                     int(round(g["other_synchronization"] * 1e6 / B)), ndec, next_, "", "", "", "",
                     "hip", "linux", "gfx950", datetime.datetime.utcnow().isoformat() + "Z"])
    return rows


def plain_host_rate(w, rng, n):
    """The host-pointer entry point (mkhip_kklss_ckks_mulrelin: operands copied in
    from host memory, the result copied back, one op per call) at the headline
    shape: ops/s and latency, operands in ordinary (pageable) host memory."""
    N, k, L = w.N, w.k, w.L
    ids = list(range(k))
    op0 = np.stack([uniform(rng, w.Q, N) for _ in range(1 + k)])
    op1 = np.stack([uniform(rng, w.Q, N) for _ in range(1 + k)])
    scale = float(w.Q[-1])
    for _ in range(3):
        w.ctx.kklss_ckks_mulrelin(op0, ids, op1, ids, scale, scale, scale)
    lat = []
    t0 = time.perf_counter()
    for _ in range(n):
        s = time.perf_counter()
        w.ctx.kklss_ckks_mulrelin(op0, ids, op1, ids, scale, scale, scale)
        lat.append(time.perf_counter() - s)
    el = time.perf_counter() - t0
    return {"ops_per_s": round(n / el, 2), "p50_ms": round(pct(lat, 50) * 1e3, 3), "p99_ms": round(pct(lat, 99) * 1e3, 3)}


def batch_sweep(w, ops, args, rng, plain):
    """The headline plan at smaller batches (the first b ops of the batch, same
    knobs): throughput and latency per launch, and the highest-throughput batch
    whose p99 launch latency stays under 10 ms, against the plain per-op path."""
    rows = []
    for b in args.batch_sweep:
        if b >= len(ops):
            continue
        pl = w.plan(ops[:b])
        el, lat = time_plan(w.ctx, pl, args.steps, args.warmup)
        pl.close()
        rows.append({"batch": b, "ops_per_s": round(b * args.steps / el, 2), "p50_ms": round(pct(lat, 50) * 1e3, 3),
                     "p99_ms": round(pct(lat, 99) * 1e3, 3)})
    ok = [r for r in rows if r["p99_ms"] < 10.0]
    best = max(ok, key=lambda r: r["ops_per_s"]) if ok else None
    out = {"sizes": rows, "best_batch_under_p99_10ms": best}
    if best and plain and plain.get("ops_per_s"):
        out["best_over_plain"] = round(best["ops_per_s"] / plain["ops_per_s"], 3)
    return out


ROT_STEPS = (1, 2, 4, 8)


def run_rotations(args, rng):
    """Galois key switches at the headline shape (N = 2^14, 8 Q + 2 P, k = 4):
    KKLSS RotateHoisted of every op by 1, 2, 4 and 8 from one decomposition,
    KKLSS rotation by 1 alone, conjugation, and the MKF party-specific rotation
    (gamma 1); B ops per launch, one hipGraph per step."""
    import mkf_amd

    logN, L, K, k, B = 14, 8, 2, 4, args.rot_batch
    N = 1 << logN
    Q, P = ntt_primes(50, L), ntt_primes(60, K)
    QP = Q + P
    out = {"shape": {"logN": logN, "L": L, "K": K, "k": k, "batch": B}}
    ctx = mkf_amd.Context(logN, Q, P, gamma=2)
    beta = ctx.beta
    for r in ROT_STEPS:
        ctx.set_crs(r, uniform_stack(rng, QP, N, beta))
        for pid in range(k):
            ctx.add_rotation_key(pid, r, uniform_stack(rng, QP, N, beta))
    ctx.set_crs(mkf_amd.CONJUGATE, uniform_stack(rng, QP, N, beta))
    for pid in range(k):
        ctx.add_conjugation_key(pid, uniform_stack(rng, QP, N, beta))
    words = (1 + k) * L * N
    bufs = []

    def leg(rots):
        ops = []
        for _ in range(B):
            din, dout = ctx.alloc(words * 8), ctx.alloc(len(rots) * words * 8)
            ctx.h2d(din, np.concatenate([uniform(rng, Q, N) for _ in range(1 + k)]))
            bufs.extend([din, dout])
            ops.append({"in": din, "out": dout, "ids": list(range(k)), "level": L - 1, "rot": list(rots)})
        plan = mkf_amd.Plan(ctx, ops, use_graph=True, scheme="rotate")
        el, lat = time_plan(ctx, plan, args.rot_steps, 2)
        st = plan.stage_times(args.stage_iters)
        names = plan.stages
        n = B * len(rots) * args.rot_steps
        # keys read once per launch: 2 (rotation key + CRS) per party per rotation
        key_bytes = len(rots) * (k + 1) * beta * (L + K) * N * 8
        io_bytes = B * (1 + len(rots)) * words * 8
        res = {"rotations_per_s": round(n / el, 1), "ms_per_step": round(el / args.rot_steps * 1e3, 3),
               "p99_ms": round(pct(lat, 99) * 1e3, 3), "rotations_per_launch": B * len(rots),
               "alg_gbs": round((key_bytes + io_bytes) * args.rot_steps / el / 1e9, 1),
               "stages_ms": {names[i]: round(st[i], 4) for i in range(len(names))}}
        plan.close()
        return res

    out["kklss_hoisted_1_2_4_8"] = leg(ROT_STEPS)
    out["kklss_rotate_1"] = leg((1,))
    out["kklss_conjugate"] = leg((mkf_amd.CONJUGATE,))
    for p_ in bufs:
        ctx.free(p_)
    ctx.close()
    # MKF party-specific rotation
    ctx = mkf_amd.Context(logN, Q, P, gamma=1)
    beta = ctx.beta

    def swk():
        return np.stack([np.stack([uniform(rng, QP, N), uniform(rng, QP, N)]) for _ in range(beta)])

    ctx.register_mkf_group(0, [swk() for _ in range(k)], swk())
    ctx.register_mkf_rotation(0, 1, [swk() for _ in range(k)])
    ops, bufs = [], []
    for _ in range(B):
        din, dout = ctx.alloc(words * 8), ctx.alloc(words * 8)
        ctx.h2d(din, np.concatenate([uniform(rng, Q, N) for _ in range(1 + k)]))
        bufs.extend([din, dout])
        ops.append({"in": din, "out": dout, "group": 0, "level": L - 1, "rotation": 1})
    plan = mkf_amd.Plan(ctx, ops, use_graph=True, scheme="mkf_rotate")
    el, lat = time_plan(ctx, plan, args.rot_steps, 2)
    st = plan.stage_times(args.stage_iters)
    out["mkf_rotate_1"] = {"rotations_per_s": round(B * args.rot_steps / el, 1),
                           "ms_per_step": round(el / args.rot_steps * 1e3, 3),
                           "p99_ms": round(pct(lat, 99) * 1e3, 3), "rotations_per_launch": B,
                           "stages_ms": {plan.stages[i]: round(st[i], 4) for i in range(len(st))}}
    plan.close()
    for p_ in bufs:
        ctx.free(p_)
    ctx.close()
    return out


def run_poly(args, rng):
    """MKF EvaluatePoly (true_mkf_ckks_polynomial.go:79-273) at the MKF headline shape
    (N = 2^14, 8 Q + 2 P, gamma 1, k = 4): the degree-59 Chebyshev ReLU activation of
    the reference's ResNet-20 tests (resnet20_activation_test.go:15,
    resnet20_real_block1_test.go:226-233) on B ciphertexts per plan launch (one
    hipGraph), scale = the last prime (one prime per rescale)."""
    import mkf_amd

    logN, L, K, k, B = 14, 8, 2, 4, args.poly_batch
    N = 1 << logN
    Q, P = ntt_primes(50, L), ntt_primes(60, K)
    QP = Q + P
    ctx = mkf_amd.Context(logN, Q, P, gamma=1)
    beta = ctx.beta

    def swk():
        return np.stack([np.stack([uniform(rng, QP, N), uniform(rng, QP, N)]) for _ in range(beta)])

    ctx.register_mkf_group(0, [swk() for _ in range(k)], swk())
    relu = np.polynomial.chebyshev.chebinterpolate(lambda t: np.maximum(t, 0.0), 59)
    poly = mkf_amd.Polynomial(list(relu), basis=mkf_amd.BASIS_CHEBYSHEV)
    D = float(Q[-1])
    level = L - 1
    lo, so, nm = ctx.mkf_poly_query(0, level, D, poly, D)
    words_in, words_out = (1 + k) * L * N, (1 + k) * (lo + 1) * N
    ins, outs = [], []
    for _ in range(B):
        d = ctx.alloc(words_in * 8)
        ctx.h2d(d, np.concatenate([uniform(rng, Q, N) for _ in range(1 + k)]))
        ins.append(d)
        outs.append(ctx.alloc(words_out * 8))
    t0 = time.perf_counter()
    plan = mkf_amd.PolyPlan(ctx, 0, ins, level, D, poly, D, outs, use_graph=True)
    build_s = time.perf_counter() - t0
    el, lat = time_plan(ctx, plan, args.poly_steps, 1)
    st = plan.stage_times(1)
    by = {}
    for name_, t in zip(plan.stages, st):
        by[name_] = by.get(name_, 0.0) + t
    res = {"polynomial": "Chebyshev ReLU, degree 59", "batch": B, "level_in": level, "level_out": lo,
           "mulrelins_per_eval": nm, "stages_per_launch": len(st),
           "evals_per_s": round(B * args.poly_steps / el, 1),
           "mulrelins_per_s": round(B * nm * args.poly_steps / el, 1),
           "ms_per_step": round(el / args.poly_steps * 1e3, 3), "p99_ms": round(pct(lat, 99) * 1e3, 3),
           "plan_build_s": round(build_s, 3), "workspace_gb": round(plan.workspace_bytes / 1e9, 3),
           "stages_ms": {n_: round(t, 3) for n_, t in sorted(by.items(), key=lambda x: -x[1])}}
    plan.close()
    for d in ins + outs:
        ctx.free(d)
    ctx.close()
    return res


# uneven batch: (logN, L, K, scheme) per ring degree; K = 1 only admits the MKF
# construction (KKLSS needs alpha = floor(K/2) >= 1)
UNEVEN_SHAPES = {13: (4, 1, "mkf"), 14: (8, 2, "kklss"), 15: (16, 3, "kklss")}
UNEVEN_K = (2, 4, 8, 16)


def run_uneven(args, rng):
    """256 MK-CKKS mul+relin ops, k uniform in {2,4,8,16}, N uniform in
    {2^13,2^14,2^15} (4/8/16 Q primes, 1/2/3 P primes), one graph launch per step."""
    import mkf_amd

    nops = args.uneven_ops
    draws = [(int(rng.choice(list(UNEVEN_SHAPES))), int(rng.choice(UNEVEN_K))) for _ in range(nops)]
    nparty = max(UNEVEN_K)
    ctxs, bufs, branches = {}, [], []
    t_setup = time.perf_counter()
    for logN, (L, K, scheme) in UNEVEN_SHAPES.items():
        sel = [k for (n, k) in draws if n == logN]
        if not sel:
            continue
        Q, P, N = ntt_primes(50, L), ntt_primes(60, K), 1 << logN
        QP = Q + P
        ctx = mkf_amd.Context(logN, Q, P, gamma=2 if scheme == "kklss" else 1)
        ctxs[logN] = ctx
        beta = ctx.beta
        if scheme == "kklss":
            for pid in range(nparty):
                ctx.register_party(pid, *[uniform_stack(rng, QP, N, beta) for _ in range(3)])
            ctx.set_crs_u(uniform_stack(rng, QP, N, beta))
        else:
            def swk():
                return np.stack([np.stack([uniform(rng, QP, N), uniform(rng, QP, N)]) for _ in range(beta)])
            for k in UNEVEN_K:
                ctx.register_mkf_group(k, [swk() for _ in range(k)], swk())
        # operand pool: 2 (1 + kmax) uniform polys; an op reads 1 + k consecutive ones per side
        npool = 2 * (1 + nparty)
        pool = ctx.alloc(npool * L * N * 8)
        ctx.h2d(pool, np.concatenate([uniform(rng, Q, N) for _ in range(npool)]))
        bufs.append((ctx, pool))
        scale = float(Q[-1])
        nres, _ = ctx.rescale_count(L - 1, scale * scale, scale)
        ops = []
        for k in sel:
            s0, s1 = (int(x) for x in rng.integers(0, npool - k, size=2))
            out = ctx.alloc((1 + k) * (L - nres) * N * 8)
            bufs.append((ctx, out))
            a, b = pool + s0 * L * N * 8, pool + s1 * L * N * 8
            if scheme == "kklss":
                ids = sorted(int(x) for x in rng.choice(nparty, size=k, replace=False))
                ops.append(dict(op0=a, op1=b, out=out, ids0=ids, ids1=ids, level=L - 1, nres=nres))
            else:
                ops.append(dict(left=a, right=b, out=out, group=k, level=L - 1, nres=nres))
        tune = []
        if args.autotune:
            tune = ctx.autotune(ops[:args.uneven_tune_ops], scheme=scheme, iters=3)
        branches.append((logN, scheme, ctx, ops, tune))
    setup_s = time.perf_counter() - t_setup
    batch = mkf_amd.Batch()
    info = {}
    budget = int(args.uneven_budget_gb * 1e9)
    for logN, scheme, ctx, ops, tune in branches:
        nch = batch.add(ctx, ops, scheme, ws_budget=budget)
        ks = [len(o["ids0"]) if scheme == "kklss" else o["group"] for o in ops]
        info["N=2^%d" % logN] = {"scheme": scheme + "_ckks", "L": UNEVEN_SHAPES[logN][0],
                                 "K": UNEVEN_SHAPES[logN][1], "ops": len(ops),
                                 "k_counts": {str(k): ks.count(k) for k in UNEVEN_K}, "chunks": nch,
                                 "autotune_ms": [round(t, 3) for t in tune],
                                 "tuning": {"fused_decomp": ctx.get_tuning(1), "xcd_remap": ctx.get_tuning(2),
                                            "decomp_order": ctx.get_tuning(3), "split_limb": ctx.get_tuning(5),
                                            "split_ext": ctx.get_tuning(6), "digit_pre": ctx.get_tuning(7),
                                            "xcd_ext": ctx.get_tuning(8), "ext_cs": ctx.get_tuning(9),
                                            "inv_lazy": ctx.get_tuning(10)}}
    batch.finalize()
    for _ in range(args.warmup):
        batch.submit()
    batch.wait()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.uneven_steps):
        s = time.perf_counter()
        batch.submit()
        batch.wait()
        lat.append(time.perf_counter() - s)
    el = time.perf_counter() - t0
    res = {"ops": nops, "ops_per_s": round(nops * args.uneven_steps / el, 2),
           "ms_per_launch": round(el / args.uneven_steps * 1e3, 3), "p50_ms": round(pct(lat, 50) * 1e3, 3),
           "p99_ms": round(pct(lat, 99) * 1e3, 3), "graph_launches_per_step": 1,
           "workspace_gb": round(batch.workspace_bytes / 1e9, 3), "ws_budget_gb_per_branch": args.uneven_budget_gb,
           "branches": info, "setup_s": round(setup_s, 1)}
    batch.close()
    for ctx, p in bufs:
        ctx.free(p)
    for ctx in ctxs.values():
        ctx.close()
    return res


# serving classes: (scheme, logN, L, K); BFV adds L QMul primes and t = 65537
SERVE_CLASSES = [("kklss", 14, 8, 2), ("kklss", 15, 16, 3), ("bfv", 14, 6, 2), ("bfv", 15, 8, 2)]


def _serve_setup(args, rng, specs, nparty, kmax):
    import mkf_amd

    classes = []
    for scheme, logN, L, K in specs:
        N = 1 << logN
        if scheme == "bfv" and logN == 14:
            Q, QM, P, t = BFV_Q, BFV_QMUL, BFV_P, BFV_T
        elif scheme == "bfv":
            allq = ntt_primes(55, 2 * L)
            Q, QM, P, t = allq[:L], allq[L:], ntt_primes(61, K), BFV_T
        else:
            Q, QM, P, t = ntt_primes(50, L), None, ntt_primes(60, K), 0
        ctx = mkf_amd.Context(logN, Q, P, gamma=2, t=t, qmul=QM)
        QP, beta = Q + P, ctx.beta
        for pid in range(nparty):
            if scheme == "kklss":
                ctx.register_party(pid, *[uniform_stack(rng, QP, N, beta) for _ in range(3)])
            else:
                ctx.register_bfv_party(pid, uniform_stack(rng, QP, N, 2 * beta), uniform_stack(rng, QP, N, 2 * beta),
                                       uniform_stack(rng, QP, N, beta))
        ctx.set_crs_u(uniform_stack(rng, QP, N, beta))
        nres = 0
        if scheme == "kklss":
            scale = float(Q[-1])
            nres, _ = ctx.rescale_count(L - 1, scale * scale, scale)
        # operand pool in pinned host memory: a request reads 1 + k consecutive polys per side
        npool = 2 * (1 + kmax)
        pool = mkf_amd.PinnedArray((npool, L, N))
        for i in range(npool):
            pool.a[i] = uniform(rng, Q, N)
        # results land in a ring of pinned output slots (the bench does not read them)
        outs = [mkf_amd.PinnedArray((1 + kmax, L - nres, N)) for _ in range(8)]
        # the same operand pool and an output ring in device memory, for requests
        # served in place (ciphertexts that stay in HBM between operations)
        dpool = ctx.alloc(pool.a.nbytes)
        ctx.h2d(dpool, pool.a)
        douts = [ctx.alloc((1 + kmax) * (L - nres) * N * 8) for _ in range(8)]
        classes.append(dict(scheme=scheme, logN=logN, L=L, K=K, ctx=ctx, nres=nres, pool=pool, npool=npool,
                            outs=outs, dpool=dpool, douts=douts))
    return classes


def _serve_run(classes, args, rng, rate, duration, streams, nparty, kmin, kmax, on_device=False):
    """Poisson arrivals at `rate` req/s for `duration` s, k uniform in kmin..kmax;
    returns latency stats.  on_device: operands and results in device memory
    (served in place, no PCIe copies)."""
    import mkf_amd

    srv = mkf_amd.Server(streams=streams, max_batch=args.serve_max_batch, mem_bytes=int(args.serve_mem_gb * 1e9))
    ids = [srv.add_class(c["ctx"], c["scheme"]) for c in classes]
    srv.start()
    # draw the whole arrival schedule first, then replay it against the wall clock
    n = max(1, int(rng.poisson(rate * duration)))
    gaps = rng.exponential(1.0 / rate, size=n)
    reqs = []
    for i in range(n):
        ci = int(rng.integers(0, len(classes)))
        c = classes[ci]
        k = int(rng.integers(kmin, kmax + 1))
        pids = sorted(int(x) for x in rng.choice(nparty, size=k, replace=False))
        s0, s1 = (int(x) for x in rng.integers(0, c["npool"] - k, size=2))
        out = c["outs"][i % len(c["outs"])].a[:1 + k]
        reqs.append((ids[ci], c, pids, s0, s1, out, k, ci))
    # operand views made before the clock starts; the collector is off while the
    # schedule is replayed, so that a collection pause does not turn into a burst
    views = [(c["pool"].a[s0:s0 + 1 + k], c["pool"].a[s1:s1 + 1 + k]) for (_, c, _, s0, s1, _, k, _) in reqs]
    if on_device:
        dviews = []
        for i, (_, c, _, s0, s1, _, k, _) in enumerate(reqs):
            poly = c["L"] * (1 << c["logN"]) * 8
            dviews.append((c["dpool"] + s0 * poly, c["dpool"] + s1 * poly, c["douts"][i % len(c["douts"])]))
    lag = []
    gc.collect()
    gc.disable()
    try:
        tickets, t0 = [], time.perf_counter()
        due = t0
        for (cl, c, pids, s0, s1, out, k, ci), (v0, v1), gap in zip(reqs, views, gaps):
            due += gap
            while True:
                now = time.perf_counter()
                if now >= due:
                    break
                if due - now > 0.0003:
                    time.sleep(min(due - now - 0.0002, 0.0005))
            lag.append(now - due)
            if on_device:
                d0, d1, dout = dviews[len(tickets)]
                tickets.append(srv.enqueue_device(cl, d0, d1, dout, ids0=pids, ids1=pids, nres=c["nres"]))
            else:
                tickets.append(srv.enqueue(cl, v0, v1, out, ids0=pids, ids1=pids, nres=c["nres"]))
    finally:
        gc.enable()
    lat, qd, bad = [], [], 0
    by_class = {}
    for t, r in zip(tickets, reqs):
        st, l, q = srv.wait(t)
        bad += st != 0
        lat.append(l)
        qd.append(q)
        by_class.setdefault(r[7], []).append(l)
    el = time.perf_counter() - t0
    stats = srv.stats()
    srv.close()
    res = {"offered_rps": rate, "requests": n, "errors": bad, "achieved_ops_per_s": round(n / el, 2),
           "p50_ms": round(pct(lat, 50), 3), "p99_ms": round(pct(lat, 99), 3),
           "queue_p50_ms": round(pct(qd, 50), 3), "arrival_lag_p99_ms": round(pct(lag, 99) * 1e3, 3),
           "arrival_lag_max_ms": round(max(lag) * 1e3, 3), "mean_batch": round(stats["mean_batch"], 2),
           "max_batch": stats["max_batch"], "max_inflight": stats["max_inflight"]}
    if len(classes) > 1:
        res["per_class"] = {"%s_N2^%d" % (classes[ci]["scheme"], classes[ci]["logN"]):
                            {"requests": len(v), "p50_ms": round(pct(v, 50), 3), "p99_ms": round(pct(v, 99), 3)}
                            for ci, v in sorted(by_class.items())}
    return res


def _serve_sweep(args, rng, specs, nparty, kmin, kmax, loads, burst, streams, on_device=False):
    t_setup = time.perf_counter()
    classes = _serve_setup(args, rng, specs, nparty, kmax)
    setup_s = time.perf_counter() - t_setup
    run = lambda rate, dur: _serve_run(classes, args, rng, rate, dur, streams, nparty, kmin, kmax, on_device)
    # capacity: a saturating burst (every request at once), after a warm-up burst
    run(1e6, min(burst, 64) / 1e6)
    sat = run(1e6, burst / 1e6)
    sweep = []
    for f in loads:
        sweep.append(run(sat["achieved_ops_per_s"] * f, args.serve_seconds))
        sweep[-1]["load"] = f
    res = {"classes": [{"scheme": c["scheme"] + ("_ckks" if c["scheme"] == "kklss" else ""), "logN": c["logN"],
                        "L": c["L"], "K": c["K"]} for c in classes],
           "parties": nparty, "k": [kmin, kmax], "streams": streams, "max_batch": args.serve_max_batch,
           "mem_gb": args.serve_mem_gb, "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
           "saturated": sat, "sweep": sweep, "setup_s": round(setup_s, 1),
           "operands": "device" if on_device else "host"}
    for c in classes:
        for o in c["outs"]:
            o.close()
        c["pool"].close()
        c["ctx"].free(c["dpool"])
        for d in c["douts"]:
            c["ctx"].free(d)
        c["ctx"].close()
    return res


def run_serving(args, rng):
    """Serving loop: Poisson arrivals of KKLSS MK-CKKS / MK-BFV mul+relin requests,
    k uniform in 2..16, N in {2^14, 2^15}, operands from registered parties whose
    keys stay resident; continuous batching, admission by device memory."""
    return _serve_sweep(args, rng, SERVE_CLASSES, args.serve_parties, 2, args.serve_kmax, args.serve_load,
                        args.serve_burst, args.serve_streams)


def run_serving_headline(args, rng, on_device=False):
    """The serving loop at the headline shape (KKLSS MK-CKKS, N=2^14, L=8, K=2,
    k=4): Poisson arrivals at fractions of the saturated rate; on_device: the
    requests' operands and results stay in HBM."""
    return _serve_sweep(args, rng, [("kklss", 14, 8, 2)], 4, 4, 4, args.serve_headline_load,
                        args.serve_headline_burst,
                        args.serve_device_streams if on_device else args.serve_headline_streams, on_device)


def _cpu_threads(requested):
    if requested > 0:
        return requested
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline(logN, Q, P, k, keys, u, threads, warm=5, timed=10):
    """The CPU oracle (C restatement of the reference path) under the reference's
    timing policy (profile_exact_test.go:45-61: 5 warm-up runs, then 10 timed
    runs): one thread, and `threads` threads running independent ops at once
    (each thread 5 warm-up + 10 timed ops; the timed region starts when every
    thread has finished its warm-up).  The C calls release the interpreter lock."""
    import threading

    import oracle as O

    N = 1 << logN
    ids = list(range(k))
    kb, kd, kv = [kk[0] for kk in keys], [kk[1] for kk in keys], [kk[2] for kk in keys]

    def leg(nthr):
        refs = [O.Ctx(logN, Q, P, gamma=2) for _ in range(nthr)]
        inputs = []
        for t in range(nthr):
            r = np.random.default_rng(1234 + t)
            inputs.append([(np.stack([uniform(r, Q, N) for _ in range(1 + k)]),
                            np.stack([uniform(r, Q, N) for _ in range(1 + k)])) for _ in range(2)])
        bar = threading.Barrier(nthr + 1)
        done = threading.Barrier(nthr + 1)

        def one(t, i):
            op0, op1 = inputs[t][i % 2]
            out, _ = refs[t].kklss_mulrelin(op0, ids, op1, ids, kb, kd, kv, u)
            for c in range(out.shape[0]):
                refs[t].div_round_last(out[c], 1)

        def worker(t):
            for i in range(warm):
                one(t, i)
            bar.wait()
            for i in range(timed):
                one(t, i)
            done.wait()

        th = [threading.Thread(target=worker, args=(t,)) for t in range(nthr)]
        for x in th:
            x.start()
        bar.wait()
        t0 = time.perf_counter()
        done.wait()
        el = time.perf_counter() - t0
        for x in th:
            x.join()
        return nthr * timed / el, el

    v1, el1 = leg(1)
    res = {"value": round(v1, 3), "unit": "ops/s", "cores": 1, "kind": "port",
           "sample": "KKLSS MK-CKKS mul+relin+rescale (N=2^14, L=8, K=2, k=4) with the C oracle, one thread: "
                     "%d warm-up + %d timed ops (%.1f s timed)" % (warm, timed, el1)}
    if threads > 1:
        vn, eln = leg(threads)
        res["multi_thread"] = {"value": round(vn, 3), "unit": "ops/s", "cores": threads,
                               "sample": "%d threads, each %d warm-up + %d timed independent ops (%.1f s timed)"
                                         % (threads, warm, timed, eln)}
    res["reference_published"] = {"value_ms_per_op": 536.0, "note": "the reference's Go KKLSS-CKKS mul+relin, "
                                  "P1 parameters, 16 parties, Apple arm64, one thread (different hardware and "
                                  "parameters)"}
    return res


def copy_calibration(mb):
    """A plain device-to-device copy kernel of known size (k_copy), run three
    times, for scaling the PMC byte counters of the other kernels."""
    import mkf_amd

    ctx = mkf_amd.Context(13, ntt_primes(50, 2), ntt_primes(60, 1), gamma=1)
    nbytes = int(mb * (1 << 20)) // 16 * 16
    src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
    for _ in range(3):
        ctx.d2d(dst, src, nbytes)
    ctx.sync()
    ctx.free(src)
    ctx.free(dst)
    ctx.close()
    return {"kernel": "k_copy", "bytes_read": nbytes, "bytes_written": nbytes, "launches": 3}


def load_pmc_summary(kernel):
    """The committed rocprofv3 PMC summary of `kernel` (HBM bytes per launch from
    FETCH_SIZE / WRITE_SIZE, VALUBusy), if it covers that kernel."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return {}
    return t if t.get("kernel") == kernel else {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--bfv-batch", type=int, default=32)
    ap.add_argument("--plain-ops", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-iters", type=int, default=3)
    ap.add_argument("--schemes", default=",".join(ALL_SCHEMES))
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="skip the library's autotuner (plans use the default knobs)")
    ap.add_argument("--autotune-iters", type=int, default=5,
                    help="timed launches per autotuner candidate (candidates a few %% apart need several)")
    ap.add_argument("--tuning", type=parse_tuning, default={},
                    help="fixed knobs for the headline scheme instead of the autotuner, e.g. "
                         "'decomp_order=1,split_limb=1,xcd_remap=0,xcd_ext=1' (reproducible profiles)")
    ap.add_argument("--uneven-ops", type=int, default=256,
                    help="ops of the uneven batch (k in {2,4,8,16}, N in {2^13,2^14,2^15}); 0 skips it")
    ap.add_argument("--uneven-steps", type=int, default=5)
    ap.add_argument("--uneven-budget-gb", type=float, default=64.0,
                    help="workspace budget of each ring degree's branch of the uneven batch (64: the "
                         "N=2^15 branch in about 3 chunks, 92 GB in all)")
    ap.add_argument("--uneven-tune-ops", type=int, default=16,
                    help="ops of each branch the autotuner times")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="length of each Poisson arrival run of the serving loop; 0 skips serving")
    ap.add_argument("--serve-load", type=lambda x: [float(v) for v in x.split(",")], default=[0.3, 0.6, 0.9],
                    help="offered load as fractions of the measured saturated rate")
    ap.add_argument("--serve-burst", type=int, default=256, help="requests of the saturating burst")
    ap.add_argument("--serve-headline-load", type=lambda x: [float(v) for v in x.split(",")],
                    default=[0.5, 0.6, 0.7, 0.8, 0.85, 0.9, 0.95])
    ap.add_argument("--serve-headline-burst", type=int, default=384,
                    help="requests of the saturating burst at the headline shape")
    ap.add_argument("--serve-parties", type=int, default=32)
    ap.add_argument("--serve-kmax", type=int, default=16)
    ap.add_argument("--serve-streams", type=int, default=4,
                    help="server streams (batches in flight) of the mixed workload")
    ap.add_argument("--serve-headline-streams", type=int, default=8,
                    help="server streams at the headline shape, host operands (8 measured the most ops/s "
                         "under p99 10 ms with 16 hardware queues)")
    ap.add_argument("--serve-device-streams", type=int, default=4,
                    help="server streams at the headline shape, device-resident operands")
    ap.add_argument("--serve-max-batch", type=int, default=32)
    ap.add_argument("--serve-mem-gb", type=float, default=64.0)
    ap.add_argument("--rot-batch", type=int, default=16, help="ops per launch of the rotation legs (0: skip)")
    ap.add_argument("--rot-steps", type=int, default=5)
    ap.add_argument("--poly-batch", type=int, default=32,
                    help="ciphertexts per launch of the polynomial-evaluation leg (0: skip)")
    ap.add_argument("--poly-steps", type=int, default=3)
    ap.add_argument("--no-fp64-leg", dest="fp64_leg", action="store_false",
                    help="skip the headline batch re-timed with the opt-in FP64 transforms")
    ap.add_argument("--batch-sweep", type=lambda x: [int(v) for v in x.split(",") if v], default=[8, 16, 24, 32, 48],
                    help="headline batch sizes timed for the best batch under p99 < 10 ms ('' skips)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the op-parallel CPU leg (0: the cores this process may use, at most 16)")
    ap.add_argument("--ref-runs", type=int, default=10,
                    help="timed launches written as mkfbench rows per scheme (the reference's -runs)")
    ap.add_argument("--mkfbench-csv", default="",
                    help="write the mkfbench-format rows (cmd/mkfbench/main.go:17-24) of every scheme to this file")
    ap.add_argument("--no-peaks", dest="peaks", action="store_false",
                    help="skip the measured copy bandwidth and integer issue rates")
    ap.add_argument("--copy-calib-mb", type=float, default=0.0,
                    help="run a plain D2D copy of this many MiB first (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (mul+relin ops shard "
              "across GPUs with zero per-op traffic). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    schemes = [s for s in args.schemes.split(",") if s]
    for s in schemes:
        if s not in ALL_SCHEMES:
            ap.error("unknown scheme %s" % s)

    calib = copy_calibration(args.copy_calib_mb) if args.copy_calib_mb > 0 else None
    peaks = measure_peaks() if args.peaks else None
    if peaks:
        print("# peaks: %s" % json.dumps(peaks), file=sys.stderr, flush=True)
    rng = np.random.default_rng(0)
    results, head = {}, None
    for s in schemes:
        res, extra = run_scheme(s, args, rng)
        results[s] = res
        if s == "kklss_ckks":
            head = (res, extra)
        else:
            extra[1].close()
            extra[0].close()
        print("# %s: %.1f ops/s batched" % (s, res["ops_per_s"]), file=sys.stderr, flush=True)

    line = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic"}
    if head is not None:
        import mkf_amd

        res, (w, plan, stage_ms, names) = head
        B, k, L, K, beta, N = w.B, w.k, w.L, w.K, w.beta, w.N
        alg = stage_alg_bytes(B, k, L, K, beta, N)
        work = stage_valu_work(B, k, L, K, beta, N)
        dom = max(range(len(stage_ms)), key=lambda s: stage_ms[s])
        dom_bytes = alg[names[dom]]
        achieved = dom_bytes / (stage_ms[dom] * 1e-3) / 1e9
        kern = mkf_amd.stage_kernel(names[dom])
        pmc = load_pmc_summary(kern)
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc.get("bytes_per_launch"),
                    "kernel": kern, "stage": names[dom], "kernel_ms": round(stage_ms[dom], 4),
                    "alg_bytes_per_launch": dom_bytes}
        if pmc.get("bytes_per_launch"):
            roofline["traffic_over_alg"] = round(pmc["bytes_per_launch"] / dom_bytes, 3)
            roofline["traffic_kernel_ms"] = pmc.get("kernel_ms")
            # the profiled instance (the counters are per template instance: whole or
            # half-limb workgroups) beside the one this run's tuning launched
            roofline["traffic_instance"] = pmc.get("instance")
            if kern == "k_decomp_mac":
                roofline["instance"] = "k_decomp_mac<%d, %s>" % (w.logN, "true" if res["tuning"]["split_limb"] else "false")
        if pmc.get("valu_busy_pct") is not None:
            roofline["valu_busy_pct"] = pmc["valu_busy_pct"]
        # every stage's algorithmic bytes over its measured time, as a fraction of the HBM peak
        roofline["stages_frac"] = {n_: round(alg[n_] / (stage_ms[i_] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                   for i_, n_ in enumerate(names) if n_ in alg and stage_ms[i_] > 0}
        if peaks:
            roofline["measured_copy_gbs"] = peaks["copy_gbs"]
            roofline["frac_of_measured_copy"] = round(achieved / peaks["copy_gbs"], 4)

            rates = ("bfly_lazy_per_s", "bfly_reduced_per_s", "ibfly_lazy_per_s", "mred_add_per_s",
                     "acc3_mac_per_s")

            def valu_frac(stage, ms):
                t_min = sum(n_ / peaks[r_] for n_, r_ in zip(work[stage], rates))
                return t_min / (ms * 1e-3)

            # the integer-VALU roofline: the stage's butterflies and products at the
            # measured in-register rates of the same primitives, over its time
            if names[dom] in work:
                lz, rd, il, mr, cs = work[names[dom]]
                roofline["valu"] = {
                    "bound": "int32-valu", "butterflies_unreduced": lz, "butterflies_reduced": rd,
                    "inv_butterflies_unreduced": il, "mred_products": mr, "acc3_products": cs,
                    "peak_bfly_unreduced_per_s": peaks["bfly_lazy_per_s"],
                    "peak_bfly_reduced_per_s": peaks["bfly_reduced_per_s"],
                    "peak_ibfly_unreduced_per_s": peaks["ibfly_lazy_per_s"], "peak_mred_per_s": peaks["mred_add_per_s"],
                    "peak_acc3_per_s": peaks["acc3_mac_per_s"],
                    "frac": round(valu_frac(names[dom], stage_ms[dom]), 4)}
            roofline["stages_valu_frac"] = {n_: round(valu_frac(n_, stage_ms[i_]), 4)
                                            for i_, n_ in enumerate(names) if n_ in work and stage_ms[i_] > 0}
        cpu = None
        if not args.no_cpu:
            cpu = cpu_baseline(w.logN, w.Q, w.P, k, w.keys, w.u, _cpu_threads(args.cpu_threads))
        line.update({
            "value": res["ops_per_s"], "unit": "ops/s", "ms_per_step": res["ms_per_step"],
            "config": {"workload": "kklss-mkckks-mulrelin-rescale batched", "logN": w.logN, "L": L, "K": K,
                       "k": k, "batch": B, "alpha": w.ctx.alpha, "beta": beta, "nres": w.nres, "graph": True},
            "batched": {"ops_per_s": res["ops_per_s"], "p50_ms": res["p50_ms"], "p99_ms": res["p99_ms"],
                        "ops_per_launch": B},
            "plain": res.get("plain"), "batched_over_plain": res.get("batched_over_plain"),
            "plain_host": res.get("plain_host"),
            "stages_ms": res["stages_ms"], "reference_stages_ms": res["reference_stages_ms"],
            "workspace_gb": res["workspace_gb"],
            "fp64_arithmetic": res.get("fp64_arithmetic"), "batch_sweep": res.get("batch_sweep"),
            "peaks": peaks, "roofline": roofline, "cpu_baseline": cpu,
            "tuning": res["tuning"], "autotune_ms": res["autotune_ms"],
        })
        plan.close()
        w.close()
    rows = []
    for r in results.values():
        rows += r.pop("mkfbench_rows", [])
    if args.mkfbench_csv:
        import csv

        with open(args.mkfbench_csv, "w", newline="") as f:
            cw = csv.writer(f)
            cw.writerow(MKFBENCH_HEADER)
            cw.writerows(rows)
    line["schemes"] = {s: r for s, r in results.items() if s != "kklss_ckks"}
    if args.serve_seconds > 0:
        line["serving_headline"] = run_serving_headline(args, rng)
        sh = line["serving_headline"]
        plain = line.get("plain") or {}
        ok = [r for r in sh["sweep"] if r["p99_ms"] < 10.0 and r["errors"] == 0]
        best = max(ok, key=lambda r: r["achieved_ops_per_s"]) if ok else None
        sh["best_under_p99_10ms"] = best
        if best and plain.get("ops_per_s"):
            # against the device-resident plain path (operands already in HBM) and the
            # host-pointer plain path (operands copied in and out per op, as served)
            sh["served_over_plain"] = round(best["achieved_ops_per_s"] / plain["ops_per_s"], 3)
            ph = line.get("plain_host") or {}
            if ph.get("ops_per_s"):
                sh["served_over_plain_host"] = round(best["achieved_ops_per_s"] / ph["ops_per_s"], 3)
        print("# serving (headline shape): saturated %.1f ops/s" % sh["saturated"]["achieved_ops_per_s"],
              file=sys.stderr, flush=True)
        # the same with device-resident operands: continuous batching without the PCIe copies,
        # against the device-resident plain path
        line["serving_headline_device"] = run_serving_headline(args, rng, on_device=True)
        sd = line["serving_headline_device"]
        ok = [r for r in sd["sweep"] if r["p99_ms"] < 10.0 and r["errors"] == 0]
        sd["best_under_p99_10ms"] = max(ok, key=lambda r: r["achieved_ops_per_s"]) if ok else None
        if sd["best_under_p99_10ms"] and plain.get("ops_per_s"):
            sd["served_over_plain"] = round(sd["best_under_p99_10ms"]["achieved_ops_per_s"] / plain["ops_per_s"], 3)
        print("# serving (headline shape, device-resident operands): saturated %.1f ops/s"
              % sd["saturated"]["achieved_ops_per_s"], file=sys.stderr, flush=True)
        line["serving"] = run_serving(args, rng)
        print("# serving: saturated %.1f ops/s" % line["serving"]["saturated"]["achieved_ops_per_s"],
              file=sys.stderr, flush=True)
    if args.uneven_ops > 0:
        line["uneven_batch"] = run_uneven(args, rng)
        print("# uneven batch: %.1f ops/s" % line["uneven_batch"]["ops_per_s"], file=sys.stderr, flush=True)
    if args.rot_batch > 0:
        line["rotations"] = run_rotations(args, rng)
        print("# rotations: %.1f hoisted rotations/s" % line["rotations"]["kklss_hoisted_1_2_4_8"]["rotations_per_s"],
              file=sys.stderr, flush=True)
    if args.poly_batch > 0:
        line["polynomial"] = run_poly(args, rng)
        print("# polynomial: %.1f degree-59 evaluations/s" % line["polynomial"]["evals_per_s"],
              file=sys.stderr, flush=True)
    if calib:
        line["copy_calibration"] = calib
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
