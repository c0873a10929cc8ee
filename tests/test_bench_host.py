"""Host-side logic of bench.py (no GPU): the algorithmic-byte accounting of the
headline plan, the reference's stage names, the mkfbench CSV rows and the
hardware-queue setting of the serving legs."""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402


def test_headline_alg_bytes_count_each_datum_once():
    # B = 64, k = 4, L = 8, K = 2, beta = 8, N = 2^14: inputs 4096 limbs, keys 640,
    # digits written 40960, x / y written 10240
    b = bench.stage_alg_bytes(64, 4, 8, 2, 8, 1 << 14)
    limb = (1 << 14) * 8
    assert b["decompose_mac"] == (4096 + 640 + 40960 + 10240) * limb
    # fused = unfused decomposition + MAC minus the digits the fused kernel never reads back
    assert b["decompose"] + b["mac_xy"] - 40960 * limb == b["decompose_mac"]


def test_reference_stage_split_covers_every_stage():
    names = ["decompose_mac", "tensor_ntt", "tensor", "tensor_finish", "ext_xy", "moddown_xy", "decompose_t",
             "ext_vu", "moddown_rescale"]
    times = [4.0, 0.1, 0.5, 0.05, 2.0, 0.3, 1.8, 1.3, 0.3]
    g = bench.ref_stage_split(names, times)
    assert abs(sum(g.values()) - sum(times)) < 1e-12
    assert abs(g["gadget_decomposition"] - 5.8) < 1e-12
    assert abs(g["other_synchronization"] - 0.05) < 1e-12  # split-limb finish pass
    assert set(g) == {"tensor_product", "gadget_decomposition", "external_product_accumulation",
                      "modup_moddown_rescale", "memory_data_movement", "other_synchronization"}


def test_mkfbench_rows_follow_the_reference_header():
    names = ["intt_in", "decompose", "mac_fold", "moddown_fold", "karatsuba", "rescale"]
    rows = bench.mkfbench_rows("mkf_ckks", "N14L8K2-B64", 4, 64, names, [[0.5, 3.7, 1.2, 1.4, 0.07, 0.17]] * 3, 5)
    assert len(rows) == 3 and all(len(r) == len(bench.MKFBENCH_HEADER) for r in rows)
    r = dict(zip(bench.MKFBENCH_HEADER, rows[0]))
    assert r["scheme"] == "MKF-CKKS" and r["parties"] == 4 and r["run"] == 1 and r["warmup"] == 5
    assert r["decomposition_count"] == 9 and r["external_product_poly_count"] == 18  # 2k+1 switches, 2 polys each
    parts = sum(r[c] for c in ("tensor_ns", "decomposition_ns", "external_product_ns", "modup_moddown_rescale_ns",
                               "memory_copy_ns", "other_ns"))
    assert abs(parts - r["total_ns"]) <= 6  # per-op rounding of each column


def test_hardware_queues_raised_for_the_server_streams():
    """bench.py gives each server stream a hardware queue of its own: a setting
    below 16 is raised to 16 before the HIP runtime starts, a larger one kept."""
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = "import os, bench; print(os.environ['GPU_MAX_HW_QUEUES'])"
    for given, expect in (("4", "16"), ("32", "32"), (None, "16")):
        env = dict(os.environ)
        env.pop("GPU_MAX_HW_QUEUES", None)
        if given is not None:
            env["GPU_MAX_HW_QUEUES"] = given
        out = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True,
                             check=True)
        assert out.stdout.strip() == expect
