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


class BatchDecoder:
    def __init__(self, decoder):
        self.dec = decoder
        self._lib = L.lib()
        sz = C.c_size_t()
        L.check("cuphyHipLDPCBatchDecodeGetWorkspaceSize",
                self._lib.cuphyHipLDPCBatchDecodeGetWorkspaceSize(C.byref(sz)))
        self.workspace = torch.zeros(sz.value // 4, dtype=torch.int32, device="cuda")
        self.items = None
        self.n_items = 0
        self.graph = None

    def build(self, tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=10, alpha=0.8125,
              clamp_value=32.0, flags=0, soft=None, order="lpt"):
        """tbs: objects with bg, Z, mb, ncb, llr (ncb, stride) and out (ncb, words)
        device tensors (aerial_amd.workloads.TransportBlock).  soft: optional
        list of (ncb, >= Kb*Z) float16 tensors per TB."""
        rows = []
        for t, tb in enumerate(tbs):
            Kb = getattr(tb, "Kb", None) or nsys(tb.bg)
            cost = tb.Z * num_edges(tb.bg, tb.mb)
            for c in range(tb.ncb):
                rows.append((cost, t, c, Kb))
        if order == "lpt":
            rows.sort(key=lambda r: -r[0])
        arr = (L.LDPCWorkItem * len(rows))()
        maxv = 0
        h = _half_bits(alpha)
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
            if fp16_math:
                w.norm.f16x2 = h | (h << 16)
            else:
                w.norm.f32 = alpha
            w.clamp_value = clamp_value
            maxv = max(maxv, (nsys(tb.bg) + tb.mb) * tb.Z)
        raw = np.frombuffer(bytes(arr), dtype=np.uint8)
        self.items = torch.from_numpy(raw.copy()).cuda()
        self.n_items = len(rows)
        self.llr_type = llr_type
        self.fp16_math = fp16_math
        self.max_values = maxv
        self.order = [(t, c) for _, t, c, _ in rows]
        return self

    def launch(self, stream=None):
        st = self._lib.cuphyHipLDPCBatchDecode(self.dec.dec, C.c_void_p(self.items.data_ptr()), self.n_items,
                                               self.llr_type, int(self.fp16_math), self.max_values,
                                               C.c_void_p(self.workspace.data_ptr()), _stream_handle(stream))
        return L.check("cuphyHipLDPCBatchDecode", st)

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
