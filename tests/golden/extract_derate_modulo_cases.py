"""Extract the reference's de-rate-matching bit-selection cases into
derate_modulo_cases.json.

Source: cuPHY/test/derate_match_modulo/test_derate_match_modulo.cu:122-228,
the table `testcases` of (E, K, F, k0, N_cb, Zc) rows that the reference's
test runs through its textbook iterative bit selection (ref_calc_modhole,
:21-38) and its closed form.  K there is the end of the filler hole in
circular-buffer coordinates (TS 38.212 K - 2 Zc), so the hole is
[K - F, K).  Two rows are hand-crafted multi-loop cases, the others come
from the reference's PUSCH test vectors (named in the comment kept as
"tv").

    python extract_derate_modulo_cases.py [/root/reference]
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = open(os.path.join(ref, "cuPHY/test/derate_match_modulo/test_derate_match_modulo.cu")).read()
body = re.search(r"int testcases\[\]\[6\]\s*=\s*\{(.*?)\n\};", src, re.S).group(1)
rows = []
for m in re.finditer(r"\{\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\s*\},?\s*//\s*([^\n]*)", body):
    E, K, F, k0, Ncb, Zc = (int(x) for x in m.groups()[:6])
    rows.append({"E": E, "K": K, "F": F, "k0": k0, "Ncb": Ncb, "Zc": Zc, "tv": m.group(7).strip()})
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "derate_modulo_cases.json")
json.dump(rows, open(path, "w"), indent=0)
print("wrote %d rows to %s" % (len(rows), path))
