"""Extract the reference's BLER-threshold fixtures into bler_thresholds.json.

Source: cuPHY/util/ldpc/test/ldpc_decode_BG{1,2}_Z*_BLER0.1.txt in the
reference tree, the fixtures of its statistical regression test
(cuPHY/util/ldpc/ldpc_perf_collect.py --mode test, :772-856): each row
"BG Z num_parity num_iter SNR max_BER max_BLER" states that at that complex
Es/N0 (dB) the decoder's BLER over 800 QPSK code blocks must be <= max_BLER.

    python extract_bler_thresholds.py [/root/reference]
"""
import glob
import json
import os
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
rows = []
for path in sorted(glob.glob(os.path.join(ref, "cuPHY/util/ldpc/test/ldpc_decode_BG*_BLER0.1.txt"))):
    for line in open(path):
        f = line.split()
        if not f or f[0].startswith("#"):
            continue
        bg, Z, mb, it = (int(x) for x in f[:4])
        rows.append({"bg": bg, "Z": Z, "mb": mb, "iters": it, "snr_db": float(f[4]), "max_ber": float(f[5]),
                     "max_bler": float(f[6]), "file": os.path.basename(path)})
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bler_thresholds.json")
json.dump(rows, open(out, "w"), indent=0)
print("wrote %d rows to %s" % (len(rows), out))
