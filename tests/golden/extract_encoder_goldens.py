"""Extract the reference's LDPC-encoder golden vectors into encoder_goldens.npz.

Source: 5GModel/nr_matlab/standalone/standalone_LDPC_encoder_test/
mat_gen_data_1_1144_{1_0,1_1,752_0}.hdf5 in the reference tree (written by
that directory's save_TbCodedCbs.m / hdf5_write_nv.m from the MATLAB
encoder LDPC_encode_main.m): datasets `sourceData` (uint8 info bits) and
`encodedData` (uint8 codeword bits, all 68 columns of BG1 at Zc = 52).
h5py is not available, so this reads the two contiguous datasets directly:
each dataset's object header carries a dataspace message (type 1) with its
dimensions and a data-layout message (type 8, class 1 = contiguous) with the
file address and byte size of its raw data.  Only the first 4 code blocks of
the 752-CB file are kept, to keep the fixture small.

    python extract_encoder_goldens.py [/root/reference]
"""
import os
import struct
import sys

import numpy as np

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
D = os.path.join(ref, "5GModel/nr_matlab/standalone/standalone_LDPC_encoder_test")


def datasets(path):
    s = open(path, "rb").read()
    spaces, layouts = [], []
    for i in range(len(s) - 24):
        t, sz = struct.unpack_from("<HH", s, i)
        if s[i + 5:i + 8] != b"\0\0\0":
            continue
        if t == 1 and 8 <= sz <= 64 and s[i + 8] == 1 and 1 <= s[i + 9] <= 3:
            dims = struct.unpack_from("<%dQ" % s[i + 9], s, i + 16)
            if all(0 < d < 10 ** 8 for d in dims):
                spaces.append(dims)
        if t == 8 and 16 <= sz <= 40 and s[i + 8] == 3 and s[i + 9] == 1:
            addr, size = struct.unpack_from("<QQ", s, i + 10)
            if addr + size <= len(s):
                layouts.append((addr, size))
    assert len(spaces) == 2 and len(layouts) == 2, (spaces, layouts)
    out = []
    for dims, (addr, size) in zip(spaces, layouts):
        assert int(np.prod(dims)) == size
        # MATLAB writes column-major: HDF5 dims (n_cb, n_bits) -> rows are code blocks
        out.append(np.frombuffer(s[addr:addr + size], np.uint8).reshape(dims))
    return out  # [sourceData, encodedData]


arrs = {}
for tag, keep in (("1_0", None), ("1_1", None), ("752_0", 4)):
    src, enc = datasets(os.path.join(D, "mat_gen_data_1_1144_%s.hdf5" % tag))
    if keep:
        src, enc = src[:keep], enc[:keep]
    arrs["src_" + tag] = src
    arrs["enc_" + tag] = enc
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "encoder_goldens.npz")
np.savez_compressed(out, **arrs)
print("wrote", out, {k: v.shape for k, v in arrs.items()})
