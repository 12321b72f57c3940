"""Extract the reference's min-sum normalisation tables into min_sum_norm.json.

Sources in the reference tree, both the same table indexed by the number of
parity nodes mb:
  cuPHY/src/cuphy/error_correction/ldpc.cpp:139-235   g_min_sum_norm_BG{1,2}_Z384
  5GModel/nr_matlab/pxsch/LDPC_decode.m:101-208       (the same values)
The script checks that the two agree before writing the fixture.

    python extract_norm_table.py [/root/reference]
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def cpp_table(text, name):
    m = re.search(name + r"\[\d+\]\s*=\s*\{(.*?)\};", text, re.S)
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return [float(x.strip().rstrip("f")) for x in body.split(",") if x.strip()]


def m_table(text, name):
    m = re.search(name + r"\s*=\s*\.\.\.\s*\[(.*?)\];", text, re.S)
    return [float(x) for x in m.group(1).split()]


cpp = open(os.path.join(ref, "cuPHY/src/cuphy/error_correction/ldpc.cpp")).read()
mat = open(os.path.join(ref, "5GModel/nr_matlab/pxsch/LDPC_decode.m")).read()
out = {}
for bg in (1, 2):
    name = "g_min_sum_norm_BG%d_Z384" % bg
    a, b = cpp_table(cpp, name), m_table(mat, name)
    assert a == b, (bg, a, b)
    assert len(a) == (47 if bg == 1 else 43)
    out["BG%d" % bg] = a
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "min_sum_norm.json")
json.dump(out, open(path, "w"), indent=0)
print("wrote", path)
