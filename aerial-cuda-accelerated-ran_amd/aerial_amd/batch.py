"""Heterogeneous batched decode: persistent launches per slot.

Builds the device work-item array for cuphyHipLDPCBatchDecode /
cuphyHipLDPCBatchDecodeGroups from a list of transport blocks (any BG, Z, mb
per TB), orders work longest-first (a longest-processing-time list schedule
over the device work queue), and can capture the launches in a HIP graph
(torch.cuda.CUDAGraph) for replay per slot.

Two regimes:
  * latency (fewer code blocks than `split_min`): every code block gets a
    workgroup of its own, so each one runs at its own speed -- a 768-thread
    row-pair workgroup (two lane groups taking the independent block rows
    of each barrier segment at the same time) unless all code blocks are
    high-rate (mb <= 8: no two consecutive rows are independent), then 384;
  * throughput (at least `split_min` code blocks): code blocks of one
    (BG, Z, mb) are packed several to a workgroup and workgroup sizes follow
    Z, as the tuning table chooses per (BG, Z, mb class)
    (aerial_amd.tuning); one persistent launch per (mb class, workgroup
    size), run concurrently on separate streams.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from . import tuning
from .decoder import _stream_handle
from .workloads import num_edges, nsys

SMALL_MAXR = 8  # mb class boundary of the register-lean kernel variant

# numpy view of cuphyHipLDPCWorkItem_t (48 bytes)
ITEM_DTYPE = np.dtype([("llr", "<u8"), ("out", "<u8"), ("soft", "<u8"), ("BG", "<i2"), ("Z", "<i2"),
                       ("num_parity_nodes", "<i2"), ("Kb", "<i2"), ("max_iterations", "<i2"), ("flags", "<u2"),
                       ("norm", "<u4"), ("clamp_value", "<f4"), ("status", "<u4")])
assert ITEM_DTYPE.itemsize == C.sizeof(L.LDPCWorkItem)


def _half_bits(x):
    return int(np.array([x], dtype=np.float16).view(np.uint16)[0])


def group_values(bg, Z, mb, count):
    """LDS values of a group of `count` code blocks (regions of whole 16 B)."""
    nval = (nsys(bg) + mb) * Z
    return count * ((nval + 7) & ~7)


class _Launch:
    """Work of one persistent launch (one kernel variant and workgroup size)
    with its own device queue."""

    def __init__(self, ws_words, block, grouped):
        self.workspace = torch.zeros(ws_words, dtype=torch.int32, device="cuda")
        self.block = block
        self.grouped = grouped
        self.items = None
        self.group_first = None
        self.n_items = 0
        self.n_groups = 0
        self.max_values = 0
        self.max_mb = 4


class BatchDecoder:
    """Heterogeneous batch over one or more persistent launches.

    Latency regime (< split_min code blocks): code blocks with mb <= 8 go to
    the register-lean kernel variant when all of them are high-rate, else
    everything shares the full variant's launch (measured on the 16-UE
    heterogeneous slot: one launch is faster than two there).  Throughput
    regime: groups and workgroup sizes from the tuning table, one launch per
    (mb class, workgroup size) on separate streams (fork/join)."""

    def __init__(self, decoder, split_by_mb=True, split_min=256, pack=True, table=None, pair=True):
        self.dec = decoder
        self.pair = pair  # latency regime: row-pair workgroups (768 threads)
        self._lib = L.lib()
        self.split_by_mb = split_by_mb
        self.split_min = split_min
        self.pack = pack
        self.table = table if table is not None else tuning.load_table()
        sz = C.c_size_t()
        L.check("cuphyHipLDPCBatchDecodeGetWorkspaceSize",
                self._lib.cuphyHipLDPCBatchDecodeGetWorkspaceSize(C.byref(sz)))
        self._ws_words = sz.value // 4
        self.subs = []
        self.graph = None
        self.aux_streams = []

    @property
    def n_items(self):
        return sum(b.n_items for b in self.subs)

    @property
    def workspaces(self):
        return [b.workspace for b in self.subs]

    def _choice(self, bg, Z, mb, regime_throughput, pair_launch=False):
        """(code blocks per workgroup, workgroup size) of a configuration."""
        if not regime_throughput and pair_launch:
            return 1, 768
        if not (self.pack and regime_throughput):
            return 1, 384
        return tuning.choose(self.table, bg, Z, mb)

    def build(self, tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=10, alpha=0.8125,
              clamp_value=32.0, flags=0, soft=None, order="lpt"):
        """tbs: objects with bg, Z, mb, ncb, llr (ncb, stride) and out (ncb, words)
        device tensors (aerial_amd.workloads.TransportBlock).  soft: optional
        list of (ncb, >= Kb*Z) float16 tensors per TB."""
        n_small = sum(tb.ncb for tb in tbs if tb.mb <= SMALL_MAXR)
        n_all = sum(tb.ncb for tb in tbs)
        throughput = n_all >= self.split_min
        split = self.split_by_mb and (n_small >= self.split_min or n_small == n_all)
        f16 = fp16_math or llr_type == L.CUPHY_R_16F
        h = _half_bits(alpha)
        norm = (h | (h << 16)) if f16 else int(np.array([alpha], dtype=np.float32).view(np.uint32)[0])

        # code blocks -> groups of one configuration -> launches
        launches = {}
        for t, tb in enumerate(tbs):
            Kb = getattr(tb, "Kb", None) or nsys(tb.bg)
            small = (tb.mb <= SMALL_MAXR) if split else False
            P, block = self._choice(tb.bg, tb.Z, tb.mb, throughput, pair_launch=self.pair and not small)
            key = (small, block)
            # estimated group time: waves per SIMD x edges per row sweep
            waves = -(-P * tb.Z // 64)
            cost = -(-waves // 4) * num_edges(tb.bg, tb.mb)
            for c0 in range(0, tb.ncb, P):
                launches.setdefault(key, []).append((cost, t, c0, min(P, tb.ncb - c0), Kb))
        self.subs = []
        self.order = []
        for key in sorted(launches, key=lambda k: (not k[0], -k[1])):
            groups = launches[key]
            if order == "lpt":
                groups.sort(key=lambda g: -g[0])
            n_items = sum(g[3] for g in groups)
            arr = np.zeros(n_items, dtype=ITEM_DTYPE)
            first = np.zeros(len(groups) + 1, dtype=np.uint32)
            sb = _Launch(self._ws_words, key[1], True)
            i = 0
            for gi, (cost, t, c0, cnt, Kb) in enumerate(groups):
                tb = tbs[t]
                first[gi] = i
                cs = np.arange(c0, c0 + cnt, dtype=np.uint64)
                sl = slice(i, i + cnt)
                arr["llr"][sl] = tb.llr.data_ptr() + cs * (tb.llr.stride(0) * tb.llr.element_size())
                arr["out"][sl] = tb.out.data_ptr() + cs * (tb.out.stride(0) * 4)
                if soft is not None and soft[t] is not None:
                    arr["soft"][sl] = soft[t].data_ptr() + cs * (soft[t].stride(0) * 2)
                arr["BG"][sl] = tb.bg
                arr["Z"][sl] = tb.Z
                arr["num_parity_nodes"][sl] = tb.mb
                arr["Kb"][sl] = Kb
                arr["max_iterations"][sl] = max_iterations
                arr["flags"][sl] = flags
                arr["norm"][sl] = norm
                arr["clamp_value"][sl] = clamp_value
                sb.max_values = max(sb.max_values, group_values(tb.bg, tb.Z, tb.mb, cnt))
                sb.max_mb = max(sb.max_mb, tb.mb)
                self.order += [(t, c) for c in range(c0, c0 + cnt)]
                i += cnt
            first[len(groups)] = i
            sb.items = torch.from_numpy(arr.view(np.uint8).copy()).cuda()
            # no group list when every group is one code block: the library then
            # uses its lane-major one-code-block kernel
            if any(g[3] > 1 for g in groups) or sb.block == 768:
                sb.group_first = torch.from_numpy(first.view(np.int32).copy()).cuda()
            sb.n_items = n_items
            sb.n_groups = len(groups)
            self.subs.append(sb)
        while len(self.aux_streams) < len(self.subs) - 1:
            self.aux_streams.append(torch.cuda.Stream())
        self.llr_type = llr_type
        self.fp16_math = fp16_math
        return self

    def _launch_one(self, sb, stream):
        st = self._lib.cuphyHipLDPCBatchDecodeGroups(
            self.dec.dec, C.c_void_p(sb.items.data_ptr()),
            C.c_void_p(sb.group_first.data_ptr() if sb.group_first is not None else None), sb.n_groups,
            self.llr_type, int(self.fp16_math), sb.max_values, sb.max_mb, sb.block,
            C.c_void_p(sb.workspace.data_ptr()), _stream_handle(stream))
        return L.check("cuphyHipLDPCBatchDecodeGroups", st)

    def launch(self, stream=None):
        main = stream if stream is not None else torch.cuda.current_stream()
        if len(self.subs) == 0:
            return
        if len(self.subs) == 1:
            return self._launch_one(self.subs[0], main)
        # fork: the other launches run on auxiliary streams, then join
        aux = self.aux_streams[:len(self.subs) - 1]
        for a in aux:
            a.wait_stream(main)
        self._launch_one(self.subs[0], main)
        for sb, a in zip(self.subs[1:], aux):
            self._launch_one(sb, a)
        for a in aux:
            main.wait_stream(a)

    def capture(self):
        """Capture one launch into a HIP graph (after a warm-up launch)."""
        self.launch()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
