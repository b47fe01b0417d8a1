"""Writes ggx_energy_ref.json: the 32x32 table of E_ss(mu, r) that the reference renderer compiles in
(src/gpu/ggx_energy.cuh:9-171, produced by its ggx-energy suite, tests/ggx_tests.cu:2109-2480), as a fixture for
tests/test_quality_tables.py. Only the numbers are kept.

python tests/golden/make_ggx_energy.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "src", "gpu", "ggx_energy.cuh")) as f:
        text = f.read()
    body = text[text.index("ggx_energy_table"):]
    body = body[body.index("{") + 1:body.index("}")]
    body = re.sub(r"/\*.*?\*/", " ", body, flags=re.S)
    values = [float(tok.strip().rstrip("f")) for tok in body.split(",") if tok.strip()]
    assert len(values) == 32 * 32, len(values)
    with open(os.path.join(HERE, "ggx_energy_ref.json"), "w") as f:
        json.dump({"size": 32, "layout": "[j*32+i]: r = j/31 (alpha = r^2), sqrt(cos theta_v) = i/31",
                   "values": values}, f)


if __name__ == "__main__":
    main()
