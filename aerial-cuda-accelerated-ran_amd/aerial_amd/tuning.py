"""Per-(BG, Z) choice of code-block packing for the batched decoder.

For every base graph, lifting size and mb class (mb <= 8: the register-lean
kernel variant; larger mb: the full one) the batched decoder needs two
numbers: P, the code blocks of that configuration decoded side by side by
one workgroup (lanes [c*Z, (c+1)*Z) on code block c), and the workgroup
size, 64 * ceil(P*Z / 64).  A bigger P fills more lanes of each wave; a
smaller workgroup lets more workgroups share a CU.  The best trade depends
on registers, LDS and wave scheduling, so `autotune` measures it: for each
(BG, Z, mb class) it times one persistent launch over a few thousand
workgroups' worth of code blocks for every candidate P (the largest P for
each workgroup size of 1..6 waves, and P = 1; for the full class also the
768-thread row-pair workgroup, one code block decoded by two lane groups
on the independent block rows of each barrier segment) and keeps the
fastest.  The
result is a JSON table shipped with the package (tuning_gfx950.json);
without a table entry `model_choice` picks by a residency model.

Run on the GPU:  python -m aerial_amd.tuning --out <path.json>
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from .workloads import LIFTING_SIZES, nsys, nrows

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(HERE, "tuning_gfx950.json")
SMALL_MAXR = 8
LDS_PER_CU = 160 * 1024
# resident waves per SIMD of the persistent kernels (register-bound):
# ~113 VGPRs for the mb <= 8 variant, ~220 for the full one
WAVES_PER_SIMD = {"small": 4, "full": 2}


def mb_class(mb):
    return "small" if mb <= SMALL_MAXR else "full"


def candidates(Z):
    """P values worth measuring: for each workgroup size of w waves (w = 1..6)
    the largest P with P*Z <= 64w, and P = 1."""
    out = {1}
    for w in range(1, 7):
        p = (64 * w) // Z
        if p >= 1:
            out.add(p)
    return sorted(out)


def block_of(P, Z):
    return 64 * (-(-P * Z // 64))


def _stride(bg, Z, mb):
    nval = (nsys(bg) + mb) * Z
    return (nval + 7) & ~7


def model_choice(bg, Z, mb, elem=2):
    """Residency model: maximise code blocks in flight per CU."""
    cls = mb_class(mb)
    mb_rep = SMALL_MAXR if cls == "small" else nrows(bg)
    waves_cu = 4 * WAVES_PER_SIMD[cls]
    best = (0, 1, block_of(1, Z))
    for P in candidates(Z):
        w = -(-P * Z // 64)
        lds = P * _stride(bg, Z, mb_rep) * elem
        wgs = min(waves_cu // w, LDS_PER_CU // max(lds, 1))
        if wgs < 1:
            continue
        score = wgs * P
        if score > best[0]:
            best = (score, P, 64 * w)
    return best[1], best[2]


def load_table(path=TABLE_PATH):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("entries", {})


def choose(table, bg, Z, mb):
    """(P, workgroup size) for a code block configuration."""
    e = table.get("%d,%d,%s" % (bg, Z, mb_class(mb)))
    if e is not None:
        return int(e["P"]), int(e["block"])
    return model_choice(bg, Z, mb)


# ---------------------------------------------------------------------------
# measurement (GPU)
# ---------------------------------------------------------------------------
def measure(decoder, bg, Z, mb, P, n_groups=4096, reps=3, max_values=96 << 20, block=None):
    """Code blocks per second of one persistent launch over n_groups groups
    of P code blocks (BG, Z, mb; binary16 LLRs drawn at random, 10
    iterations, no early termination), workgroups of `block` threads
    (default 64 * ceil(P*Z / 64))."""
    import ctypes as C
    import torch
    from . import _lib as L
    from .batch import ITEM_DTYPE, group_values
    from .decoder import _stream_handle

    nval = (nsys(bg) + mb) * Z
    stride = (nval + 7) & ~7
    n_cbs = max(P, min(n_groups * P, max_values // stride))
    n_groups = -(-n_cbs // P)
    n_cbs = n_groups * P
    g = torch.Generator(device="cuda")
    g.manual_seed(1000 * bg + Z)
    llr = (torch.randn(n_cbs, stride, device="cuda", generator=g) * 4.0).to(torch.float16)
    words = (nsys(bg) * Z + 31) // 32
    out = torch.empty(n_cbs, words, dtype=torch.int32, device="cuda")
    arr = np.zeros(n_cbs, dtype=ITEM_DTYPE)
    idx = np.arange(n_cbs, dtype=np.uint64)
    arr["llr"] = llr.data_ptr() + idx * (stride * 2)
    arr["out"] = out.data_ptr() + idx * (words * 4)
    arr["BG"], arr["Z"], arr["num_parity_nodes"], arr["Kb"] = bg, Z, mb, nsys(bg)
    arr["max_iterations"] = 10
    h = int(np.array([0.8125], dtype=np.float16).view(np.uint16)[0])
    arr["norm"] = h | (h << 16)
    arr["clamp_value"] = 32.0
    items = torch.from_numpy(arr.view(np.uint8).copy()).cuda()
    first = torch.from_numpy((np.arange(n_groups + 1, dtype=np.uint32) * P).view(np.int32)).cuda()
    sz = C.c_size_t()
    lib = L.lib()
    lib.cuphyHipLDPCBatchDecodeGetWorkspaceSize(C.byref(sz))
    ws = torch.zeros(sz.value // 4, dtype=torch.int32, device="cuda")
    block = block or block_of(P, Z)
    mv = group_values(bg, Z, mb, P)
    stream = torch.cuda.current_stream()

    def launch():
        L.check("cuphyHipLDPCBatchDecodeGroups", lib.cuphyHipLDPCBatchDecodeGroups(
            decoder.dec, C.c_void_p(items.data_ptr()), C.c_void_p(first.data_ptr()), n_groups, L.CUPHY_R_16F, 1,
            mv, mb, block, C.c_void_p(ws.data_ptr()), _stream_handle(stream)))

    launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        launch()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    return n_cbs / (ms * 1e-3)


def autotune(decoder, zs=None, log=sys.stderr):
    entries = {}
    for bg in (1, 2):
        for Z in (zs or LIFTING_SIZES):
            for cls, mb in (("small", 4), ("full", nrows(bg))):
                res = {}
                for P in candidates(Z):
                    res[P] = measure(decoder, bg, Z, mb, P)
                best = max(res, key=lambda p: res[p])
                best_rate, block = res[best], block_of(best, Z)
                # one code block per 384-thread workgroup (the latency-regime
                # launch shape), for reference
                wide = measure(decoder, bg, Z, mb, 1, block=384)
                pair = measure(decoder, bg, Z, mb, 1, block=768) if cls == "full" else None
                if pair is not None and pair > best_rate:
                    best, best_rate, block = 1, pair, 768
                entries["%d,%d,%s" % (bg, Z, cls)] = {
                    "P": best, "block": block, "cbs_per_s": round(best_rate, 1),
                    "model_P": model_choice(bg, Z, mb)[0], "one_per_384_threads": round(wide, 1),
                    "row_pair_768_threads": round(pair, 1) if pair is not None else None,
                    "candidates": {str(p): round(v, 1) for p, v in res.items()}}
                print("BG%d Z=%d %s: P=%d block %d (%.3g CB/s; P=1 %.3g; one per 384 threads %.3g; row pair %s)"
                      % (bg, Z, cls, best, block, best_rate, res[1], wide,
                         "%.3g" % pair if pair is not None else "-"), file=log, flush=True)
    return entries


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--z", type=int, nargs="*", help="lifting sizes to tune (default: all 51)")
    a = ap.parse_args()
    import torch
    from .decoder import LdpcDecoder
    dec = LdpcDecoder()
    t0 = time.time()
    entries = autotune(dec, a.z)
    doc = {"device": torch.cuda.get_device_name(0), "seconds": round(time.time() - t0, 1),
           "method": "persistent launch, binary16 LLRs, 10 iterations, no early termination; "
                     "mb = 4 (small class) and 46/42 (full class); code blocks per second",
           "entries": entries}
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    print("wrote %s (%d entries, %.0f s)" % (a.out, len(entries), time.time() - t0), file=sys.stderr)


if __name__ == "__main__":
    main()
