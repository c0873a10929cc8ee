"""Extract the reference's own known-answer vectors and constant tables as data.

Writes JSON fixtures next to this script from files of the reference repository
(read as text; nothing there is compiled, imported or run):

* ``mred64_kats.json`` -- the 64 (x, y, q, q_inv, expected) vectors of
  hardware/u55c/rtl/tb_mred64_pipeline.sv:19-82 (expected = x*y*2^-64 mod q);
* ``butterfly_kats.json`` -- the 16x8 (u, v, w, sum, diff) vectors of
  hardware/u55c/rtl/tb_ntt_butterfly_x8.sv (q = 0x07ffffffffe70001);
* ``hls_ckks_constants.json`` / ``hls_bfv_constants.json`` -- every integer
  table of hardware/u55c/hls/include/par1_{ckks,bfv}_hls_constants.hpp.

Usage: python tests/golden/extract_reference_vectors.py [/path/to/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _hex(s):
    return int(s, 16)


def mred_kats(root):
    path = os.path.join(root, "hardware/u55c/rtl/tb_mred64_pipeline.sv")
    pat = re.compile(
        r"v\((\d+),\s*64'h([0-9a-f]+),\s*64'h([0-9a-f]+),\s*64'h([0-9a-f]+),"
        r"\s*64'h([0-9a-f]+),\s*64'h([0-9a-f]+)\);")
    out = []
    with open(path) as f:
        for m in pat.finditer(f.read()):
            out.append({"index": int(m.group(1)), "x": _hex(m.group(2)), "y": _hex(m.group(3)),
                        "q": _hex(m.group(4)), "q_inv": _hex(m.group(5)), "expected": _hex(m.group(6))})
    return out


def butterfly_kats(root):
    path = os.path.join(root, "hardware/u55c/rtl/tb_ntt_butterfly_x8.sv")
    text = open(path).read()
    q = _hex(re.search(r"logic \[63:0\] q=64'h([0-9a-f]+)", text).group(1))
    qinv = _hex(re.search(r"q_inv=64'h([0-9a-f]+)", text).group(1))
    pat = re.compile(
        r"setv\((\d+),(\d+),64'h([0-9a-f]+),64'h([0-9a-f]+),64'h([0-9a-f]+),"
        r"64'h([0-9a-f]+),64'h([0-9a-f]+)\);")
    vecs = []
    for m in pat.finditer(text):
        vecs.append({"t": int(m.group(1)), "lane": int(m.group(2)), "u": _hex(m.group(3)),
                     "v": _hex(m.group(4)), "w": _hex(m.group(5)), "sum": _hex(m.group(6)),
                     "diff": _hex(m.group(7))})
    return {"q": q, "q_inv": qinv, "vectors": vecs}


def hls_tables(path):
    """Parse `static const ap_uint<64> NAME[...] = {...};` and scalar constexprs."""
    text = open(path).read()
    out = {}
    for m in re.finditer(r"static constexpr unsigned(?: long long)? (\w+) = ([0-9A-Z_ +]+?)(?:ULL)?;", text):
        val = m.group(2).strip()
        try:
            out[m.group(1)] = int(val)
        except ValueError:
            parts = [p.strip() for p in val.split("+")]
            out[m.group(1)] = sum(out[p] if p in out else int(p) for p in parts)
    for m in re.finditer(r"static const ap_uint<64> (\w+)((?:\[[^\]]+\])+) = (\{.*?\});", text, re.S):
        name, body = m.group(1), m.group(3)
        body = re.sub(r"(\d+)ULL", r"\1", body).replace("{", "[").replace("}", "]")
        body = re.sub(r",\s*\]", "]", body)
        out[name] = json.loads(body)
    return out


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(HERE, "mred64_kats.json"), "w") as f:
        json.dump(mred_kats(root), f, indent=0)
    with open(os.path.join(HERE, "butterfly_kats.json"), "w") as f:
        json.dump(butterfly_kats(root), f, indent=0)
    inc = os.path.join(root, "hardware/u55c/hls/include")
    for scheme in ("ckks", "bfv"):
        with open(os.path.join(HERE, "hls_%s_constants.json" % scheme), "w") as f:
            json.dump(hls_tables(os.path.join(inc, "par1_%s_hls_constants.hpp" % scheme)), f, indent=0)


if __name__ == "__main__":
    main()
