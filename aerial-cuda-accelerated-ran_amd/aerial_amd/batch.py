"""Heterogeneous batched decode: one persistent launch per slot.

Builds the device work-item array for cuphyHipLDPCBatchDecode from a list of
transport blocks (any BG, Z, mb per TB), orders items longest-first (a
longest-processing-time list schedule over the device work queue), and can
capture the launch in a HIP graph (torch.cuda.CUDAGraph) for replay per slot.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .decoder import _stream_handle
from .workloads import num_edges, nsys


def _half_bits(x):
    return int(torch.tensor([x], dtype=torch.float16).view(torch.int16).item()) & 0xFFFF


SMALL_MAXR = 8  # mb class boundary of the register-lean kernel variant


class _SubBatch:
    """Work items of one kernel variant (mb class) with their own queue."""

    def __init__(self, lib, ws_words):
        self.workspace = torch.zeros(ws_words, dtype=torch.int32, device="cuda")
        self.items = None
        self.n_items = 0
        self.max_values = 0
        self.max_mb = 4


class BatchDecoder:
    """Heterogeneous batch over one or two persistent launches.

    Code blocks with mb <= 8 (high code rates) go to the register-lean
    kernel variant when there are at least `split_min` of them (enough to
    fill the GPU on their own), the rest to the full one; the launches run
    on separate streams (fork/join) and share the GPU concurrently.  A
    smaller high-rate class joins the full variant's launch: measured on the
    16-UE heterogeneous slot, one launch is faster than two there."""

    def __init__(self, decoder, split_by_mb=True, split_min=256):
        self.dec = decoder
        self._lib = L.lib()
        self.split_by_mb = split_by_mb
        self.split_min = split_min
        sz = C.c_size_t()
        L.check("cuphyHipLDPCBatchDecodeGetWorkspaceSize",
                self._lib.cuphyHipLDPCBatchDecodeGetWorkspaceSize(C.byref(sz)))
        self._ws_words = sz.value // 4
        self.subs = []
        self.graph = None
        self.aux_streams = [torch.cuda.Stream()]

    @property
    def n_items(self):
        return sum(b.n_items for b in self.subs)

    @property
    def workspaces(self):
        return [b.workspace for b in self.subs]

    def build(self, tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=10, alpha=0.8125,
              clamp_value=32.0, flags=0, soft=None, order="lpt"):
        """tbs: objects with bg, Z, mb, ncb, llr (ncb, stride) and out (ncb, words)
        device tensors (aerial_amd.workloads.TransportBlock).  soft: optional
        list of (ncb, >= Kb*Z) float16 tensors per TB."""
        n_small = sum(tb.ncb for tb in tbs if tb.mb <= SMALL_MAXR)
        n_all = sum(tb.ncb for tb in tbs)
        split = self.split_by_mb and (n_small >= self.split_min or n_small == n_all)
        classes = {}
        for t, tb in enumerate(tbs):
            Kb = getattr(tb, "Kb", None) or nsys(tb.bg)
            cost = tb.Z * num_edges(tb.bg, tb.mb)
            key = (tb.mb <= SMALL_MAXR) if split else False
            for c in range(tb.ncb):
                classes.setdefault(key, []).append((cost, t, c, Kb))
        h = _half_bits(alpha)
        self.subs = []
        self.order = []
        for key in sorted(classes, key=lambda k: not k):
            rows = classes[key]
            if order == "lpt":
                rows.sort(key=lambda r: -r[0])
            sb = _SubBatch(self._lib, self._ws_words)
            arr = (L.LDPCWorkItem * len(rows))()
            for i, (cost, t, c, Kb) in enumerate(rows):
                tb = tbs[t]
                w = arr[i]
                w.llr = tb.llr.data_ptr() + c * tb.llr.stride(0) * tb.llr.element_size()
                w.out = tb.out.data_ptr() + c * tb.out.stride(0) * 4
                if soft is not None and soft[t] is not None:
                    w.soft = soft[t].data_ptr() + c * soft[t].stride(0) * 2
                w.BG, w.Z, w.num_parity_nodes, w.Kb = tb.bg, tb.Z, tb.mb, Kb
                w.max_iterations = max_iterations
                w.flags = flags
                if fp16_math or llr_type == L.CUPHY_R_16F:
                    w.norm.f16x2 = h | (h << 16)
                else:
                    w.norm.f32 = alpha
                w.clamp_value = clamp_value
                sb.max_values = max(sb.max_values, (nsys(tb.bg) + tb.mb) * tb.Z)
                sb.max_mb = max(sb.max_mb, tb.mb)
            raw = np.frombuffer(bytes(arr), dtype=np.uint8)
            sb.items = torch.from_numpy(raw.copy()).cuda()
            sb.n_items = len(rows)
            self.subs.append(sb)
            self.order += [(t, c) for _, t, c, _ in rows]
        self.llr_type = llr_type
        self.fp16_math = fp16_math
        return self

    def _launch_one(self, sb, stream):
        st = self._lib.cuphyHipLDPCBatchDecode(self.dec.dec, C.c_void_p(sb.items.data_ptr()), sb.n_items,
                                               self.llr_type, int(self.fp16_math), sb.max_values, sb.max_mb,
                                               C.c_void_p(sb.workspace.data_ptr()), _stream_handle(stream))
        return L.check("cuphyHipLDPCBatchDecode", st)

    def launch(self, stream=None):
        main = stream if stream is not None else torch.cuda.current_stream()
        if len(self.subs) == 0:
            return
        if len(self.subs) == 1:
            return self._launch_one(self.subs[0], main)
        # fork: the other variants run on auxiliary streams, then join
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
