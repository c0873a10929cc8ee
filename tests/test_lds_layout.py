"""The transforms' padded LDS layout (one pad word after every 16): for every
pass of every ring degree, each lane's element positions split exactly into a
lane base plus compile-time offsets that survive the padding, fit the
workgroup's allocation and never collide.  Host-only (tests/cpp/test_lds_layout)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_padded_lds_layout_is_exact_and_collision_free():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "test_lds_layout"], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    r = subprocess.run([os.path.join(HERE, "cpp", "test_lds_layout")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all LDS layout checks passed" in r.stdout
