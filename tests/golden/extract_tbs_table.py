"""Extract TS 38.214 Table 5.1.3.2-1 (TBS for N_info <= 3824) from the
reference into tbs_table.json.

Sources in the reference tree (the script checks they agree):
  cuPHY/src/cuphy/utils.cuh:62                      TBS_table[93]
  pyaerial/src/aerial/phy5g/ldpc/util.py:210-303    tbs_select

    python extract_tbs_table.py [/root/reference]
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
cuh = open(os.path.join(ref, "cuPHY/src/cuphy/utils.cuh")).read()
a = [int(x) for x in re.search(r"TBS_table\[93\]\s*=\s*\{([^}]*)\}", cuh).group(1).split(",")]
py = open(os.path.join(ref, "pyaerial/src/aerial/phy5g/ldpc/util.py")).read()
b = [int(x) for x in re.findall(r"\d+", re.search(r"tbs_select = \[(.*?)\]", py, re.S).group(1))]
assert a == b and len(a) == 93, (a, b)
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tbs_table.json")
json.dump(a, open(path, "w"))
print("wrote", path)
