"""Benchmark: KKLSS MK-CKKS mul+relin per second at N=2^14, k=4 (batched vs plain).

Workload (the headline configuration): N = 2^14, 8 Q primes (~50-bit) + 2 P
primes (~60-bit, gadget digits over the Q limbs, alpha = 1, beta = 8), k = 4
parties with evaluation keys resident in HBM, a batch of 64 independent
ciphertext pairs (uniform residues, seeded) per step, one hipGraph launch per
step.  A step is one mul+relin+rescale of the whole batch; `value` is
operations per second over the timed steps (inputs already resident in HBM).

Also reported: the plain per-op GPU path (one op per launch, synchronised) with
p50/p99 latency, the per-stage device times (HIP events on the library's stream),
the roofline of the dominant kernel, and the CPU oracle timed on a bounded
sample of the same workload on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MK-CKKS mul+relin/s at N=2^14 k=4 (batched vs plain); p50/p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


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
    """Algorithmic bytes per launch of each plan stage (each datum counted once).

    Per-unit figures: a limb transform (NTT / INTT) moves 2 N W bytes (read +
    write); a key-switch MAC moves (digit words + key words + output words) W;
    a ModDown reads (L + K) rows and writes L rows.  Keys shared by the batch
    count once per launch."""
    rows = L + K
    limb = N * W
    return [
        B * 2 * k * beta * rows * 2 * limb,                                     # decompose: limb NTTs
        (B * 2 * k * beta * rows + 2 * k * beta * rows + B * 2 * beta * rows) * limb,  # x/y MAC
        B * 2 * (1 + k) * L * 2 * limb,                                         # tensor input NTTs
        B * (1 + k) * L * 2 * limb,                                             # tensor + INTT
        B * 2 * k * (2 * beta * rows + rows) * limb,                            # Ext(x|y): D + x|y + out
        B * 2 * k * (rows + L) * limb,                                          # ModDown
        B * k * beta * rows * 2 * limb,                                         # decompose t
        (B * k * (beta * rows + 2 * rows) + (k + 1) * beta * rows) * limb,      # Ext(v,u): D + out, keys once
        B * (2 * k * (rows + L) + (1 + k) * (L + L - 1)) * limb,                 # final ModDown + rescale
    ]


def cpu_baseline(logN, Q, P, k, keys, u, budget_s):
    """The CPU oracle (C restatement of the reference path, 1 thread) on a bounded sample."""
    import oracle as O

    ref = O.Ctx(logN, Q, P, gamma=2)
    rng = np.random.default_rng(1234)
    ids = list(range(k))
    n, t0 = 0, time.perf_counter()
    while True:
        op0 = np.stack([uniform(rng, Q, 1 << logN) for _ in range(1 + k)])
        op1 = np.stack([uniform(rng, Q, 1 << logN) for _ in range(1 + k)])
        out, _ = ref.kklss_mulrelin(op0, ids, op1, ids, [kk[0] for kk in keys], [kk[1] for kk in keys],
                                    [kk[2] for kk in keys], u)
        for i in range(out.shape[0]):
            ref.div_round_last(out[i], 1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 64:
            break
    return {"value": n / el, "unit": "ops/s", "cores": 1, "kind": "port",
            "sample": "%d KKLSS MK-CKKS mul+relin+rescale ops (N=2^14, L=8, K=2, k=4) with the C oracle, "
                      "single thread, %.1f s" % (n, el)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--plain-ops", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-iters", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (ops shard across GPUs "
              "with zero per-op traffic; see the design notes). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import mkf_amd

    logN, L, K, k, B = 14, 8, 2, 4, args.batch
    N = 1 << logN
    Q = ntt_primes(50, L)
    P = ntt_primes(60, K)
    QP = Q + P
    ctx = mkf_amd.Context(logN, Q, P, gamma=2)
    beta = ctx.beta
    rng = np.random.default_rng(0)
    keys = []
    for pid in range(k):
        kk = [np.stack([uniform(rng, QP, N) for _ in range(beta)]) for _ in range(3)]
        ctx.register_party(pid, *kk)
        keys.append(kk)
    u = np.stack([uniform(rng, QP, N) for _ in range(beta)])
    ctx.set_crs_u(u)
    ids = list(range(k))
    scale = float(Q[-1])
    nres, _ = ctx.rescale_count(L - 1, scale * scale, scale)

    in_words = (1 + k) * L * N
    out_words = (1 + k) * (L - nres) * N

    def make_ops(n):
        ops, bufs = [], []
        for _ in range(n):
            d0, d1 = ctx.alloc(in_words * 8), ctx.alloc(in_words * 8)
            do = ctx.alloc(out_words * 8)
            ctx.h2d(d0, np.concatenate([uniform(rng, Q, N) for _ in range(1 + k)]))
            ctx.h2d(d1, np.concatenate([uniform(rng, Q, N) for _ in range(1 + k)]))
            bufs += [d0, d1, do]
            ops.append(dict(op0=d0, op1=d1, out=do, ids0=ids, ids1=ids, level=L - 1, nres=nres))
        return ops, bufs

    # ---- batched: one graph launch per step --------------------------------
    ops, bufs = make_ops(B)
    plan = mkf_amd.Plan(ctx, ops, use_graph=True)
    for _ in range(args.warmup):
        plan.launch()
    ctx.sync()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        plan.launch()
        ctx.sync()
        lat.append(time.perf_counter() - s)
    elapsed = time.perf_counter() - t0
    batched_ops = B * args.steps / elapsed
    # per-stage device times (HIP events on the stream the kernels run on)
    stage_ms = plan.stage_times(args.stage_iters)
    stage_names = plan.stages
    ws_bytes = plan.workspace_bytes

    # ---- plain per-op path: one op per launch ------------------------------
    pbufs, pplan, plat, plain_ops = [], None, [float("nan")], float("nan")
    if args.plain_ops > 0:
        pops, pbufs = make_ops(1)
        pplan = mkf_amd.Plan(ctx, pops, use_graph=True)
        for _ in range(3):
            pplan.launch()
        ctx.sync()
        plat = []
        t1 = time.perf_counter()
        for _ in range(args.plain_ops):
            s = time.perf_counter()
            pplan.launch()
            ctx.sync()
            plat.append(time.perf_counter() - s)
        plain_ops = args.plain_ops / (time.perf_counter() - t1)

    # ---- roofline of the dominant kernel ------------------------------------
    rows = L + K
    alg = stage_alg_bytes(B, k, L, K, beta, N)
    dom = max(range(len(stage_ms)), key=lambda s: stage_ms[s])
    alg_bytes = alg[dom]
    achieved = alg_bytes / (stage_ms[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": mkf_amd.STAGE_KERNELS[stage_names[dom]], "stage": stage_names[dom],
                "kernel_ms": round(stage_ms[dom], 4), "alg_bytes_per_launch": alg_bytes}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(logN, Q, P, k, keys, u, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(batched_ops, 2),
        "unit": "ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "kklss-mkckks-mulrelin-rescale batched", "logN": logN, "L": L, "K": K,
                   "k": k, "batch": B, "alpha": ctx.alpha, "beta": beta, "nres": nres, "graph": True},
        "batched": {"ops_per_s": round(batched_ops, 2), "p50_ms": round(pct(lat, 50) * 1e3, 3),
                    "p99_ms": round(pct(lat, 99) * 1e3, 3), "ops_per_launch": B},
        "plain": {"ops_per_s": round(plain_ops, 2), "p50_ms": round(pct(plat, 50) * 1e3, 3),
                  "p99_ms": round(pct(plat, 99) * 1e3, 3)},
        "batched_over_plain": round(batched_ops / plain_ops, 3),
        "stages_ms": {stage_names[s]: round(stage_ms[s], 4) for s in range(len(stage_ms))},
        "workspace_gb": round(ws_bytes / 1e9, 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    plan.close()
    if pplan is not None:
        pplan.close()
    for p in bufs + pbufs:
        ctx.free(p)
    ctx.close()


if __name__ == "__main__":
    main()
