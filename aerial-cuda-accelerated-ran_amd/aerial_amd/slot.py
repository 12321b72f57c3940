"""Fused PUSCH slot decode: de-rate-matching + HARQ combining + layered NMS
decoding + CB/TB CRC, code block by code block, in one persistent launch
per kernel variant (cuphyHipPuschSlotDecode).

The reference runs the same three stages as separate kernels per TB group
(pusch_rx.cpp:3007-3021); here every CB of the slot is one item of a device
work queue, ordered longest first, and the launch can be captured in a HIP
graph and replayed per slot.  HARQ buffers (binary16, [2Zc | Ncb] per CB,
stride Ncb_padded) stay on the device across transmissions: ndi = 1
overwrites them, ndi = 0 combines into them.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .decoder import _stream_handle
from .workloads import num_edges, nsys

SLOT_CB_WORDS = L.MAX_WORDS_PER_CODE_BLOCK  # decoded-word stride per CB
SMALL_MAXR = 8   # row capacities of the fused kernel's variants
MID_MAXR = 16

_LLR_TYPES = {torch.int8: L.CUPHY_R_8I, torch.float16: L.CUPHY_R_16F, torch.float32: L.CUPHY_R_32F}


def half_bits(x):
    return int(torch.tensor([x], dtype=torch.float16).view(torch.int16).item()) & 0xFFFF


FUSED_BLOCK = 384                    # threads of a fused-kernel workgroup
FUSED_MAX_GROUP = 16                 # CBs of one work item at most
SLOT_MAX_LDS_VALUES = 68 * 384 + 2 * 12288


def lds_values(p, mb, descramble=False, count=1):
    """LDS values a work item of `count` CBs of TB parameters p needs: per
    CB, the HARQ circular buffer and the decoder's APP array share one
    binary16 array (all but the last rounded up to whole 16-byte units); a
    descrambled TB adds Gold-sequence words after them (two values per word,
    for the largest per-CB E starting at bit 31 of a word, plus one word)."""
    arr1 = max(2 * p.Zc + p.Ncb, (nsys(p.bg) + mb) * p.Zc)
    arr = (count - 1) * ((arr1 + 7) & ~7) + arr1
    if not descramble or p.uciOnPuschFlag:
        return arr
    nq, C = p.Nl * p.Qm, p.num_CBs
    q = p.encodedSize // nq // C
    emax = nq * q + (nq if q * nq * C < p.encodedSize else 0)
    words = (31 + emax - 1) // 32 + 1
    return ((arr + 1) & ~1) + 2 * (words + 1)


def group_size(p, mb, descramble=False, lds_limit=SLOT_MAX_LDS_VALUES):
    """CBs per work item: consecutive CBs of the TB decoded side by side in
    one workgroup (count * Zc <= FUSED_BLOCK), within lds_limit."""
    n = min(FUSED_MAX_GROUP, FUSED_BLOCK // p.Zc, p.num_CBs)
    while n > 1 and lds_values(p, mb, descramble, n) > lds_limit:
        n -= 1
    return max(n, 1)


def make_item(t, cb, count=1):
    """Work item: (TB << 16) | ((count - 1) << 12) | first CB."""
    return (t << 16) | ((count - 1) << 12) | cb


class SlotTb:
    """One TB transmission of a slot with its device outputs.

    params: PerTbParams (ndi, rv/k0, G, ...); llr: this transmission's LLRs in
    the rate-match input layout (int8 / float16 / float32 device tensor);
    harq: float16 device tensor of >= (firstCodeBlockIndex + C) * Ncb_padded
    values; mb: decoder parity rows (cover every position received so far);
    descramble: negate LLRs where the Gold sequence of params.cinit is 1."""

    def __init__(self, params, llr, harq, mb, max_iterations=10, early_term=False, llr_scale=1.0, alpha=None,
                 clamp_value=32.0, outputs=None, descramble=False):
        self.params = params
        self.llr = llr
        self.harq = harq
        self.mb = int(mb)
        self.max_iterations = max_iterations
        self.early_term = early_term
        self.descramble = descramble
        self.llr_scale = llr_scale
        self.alpha = alpha
        self.clamp_value = clamp_value
        ncb = int(params.firstCodeBlockIndex + params.num_CBs)
        if outputs is None:
            outputs = SlotTb.alloc_outputs(params)
        self.cb_bits, self.cb_crc, self.tb_bytes, self.tb_crc, self.cb_iterations = outputs
        assert harq.dtype == torch.float16 and harq.numel() >= ncb * params.Ncb_padded
        assert self.cb_bits.numel() >= ncb * SLOT_CB_WORDS and self.cb_crc.numel() >= ncb

    @staticmethod
    def data_bytes(p):
        kp = p.K - p.F
        return (kp - (24 if p.num_CBs > 1 else 0) + 7) // 8

    @staticmethod
    def alloc_outputs(p, device="cuda"):
        ncb = int(p.firstCodeBlockIndex + p.num_CBs)
        nbytes = (ncb * SlotTb.data_bytes(p) + 3) // 4 * 4
        return (torch.zeros((ncb, SLOT_CB_WORDS), dtype=torch.int32, device=device),
                torch.zeros(ncb, dtype=torch.int32, device=device),
                torch.zeros(nbytes, dtype=torch.uint8, device=device),
                torch.zeros(1, dtype=torch.int32, device=device),
                torch.zeros(ncb, dtype=torch.uint8, device=device))


class _Launch:
    """CB items of one kernel variant (mb class) with their own workspace."""

    def __init__(self):
        self.items = None
        self.n_items = 0
        self.max_values = 0
        self.max_mb = 4
        self.max_group = 1  # CBs of the largest item (1: the one-CB-per-workgroup kernel)
        self.workspace = None


def row_class(mb):
    """Kernel variant of a TB's CBs: 0 (mb <= 8), 1 (mb <= 16), 2 (all rows)."""
    return 0 if mb <= SMALL_MAXR else 1 if mb <= MID_MAXR else 2


def launch_classes(counts, split_min):
    """Launch of each row class, given the CBs per class: a class joins the
    next larger one until the CBs gathered reach split_min (enough to fill
    the GPU) or no larger class has CBs.  Returns {row class: launch class}."""
    last = max((c for c in range(3) if counts[c] > 0), default=2)
    launch_class, pending = {}, []
    for c in range(3):
        pending.append(c)
        if sum(counts[k] for k in pending) >= split_min or c >= last:
            for k in pending:
                launch_class[k] = c
            pending = []
    return launch_class


class SlotDecoder:
    """Fused decode of a list of SlotTb; one persistent launch per kernel
    variant (row class), each on its own stream: a class with fewer than
    split_min CBs (too few to fill the GPU) joins the next larger one."""

    def __init__(self, decoder, split_min=256, concurrent=True, pack=True):
        self.dec = decoder
        self.split_min = split_min
        self.concurrent = concurrent
        self.pack = pack  # consecutive CBs of a TB side by side in one workgroup (small Zc)
        self._lib = L.lib()
        self.launches = []
        self.graph = None
        self.aux_streams = [torch.cuda.Stream(), torch.cuda.Stream()]
        self._norm_cache = {}

    def _alpha_bits(self, tb):
        if tb.alpha is not None:
            return half_bits(tb.alpha)
        key = (tb.params.bg, tb.params.Zc, tb.mb)
        if key not in self._norm_cache:
            cfg = L.LDPCDecodeConfigDesc()
            cfg.llr_type = L.CUPHY_R_16F
            cfg.num_parity_nodes = tb.mb
            cfg.Z = tb.params.Zc
            cfg.max_iterations = 1
            cfg.Kb = nsys(tb.params.bg)
            cfg.BG = tb.params.bg
            L.check("cuphyErrorCorrectionLDPCDecodeSetNormalization",
                    self._lib.cuphyErrorCorrectionLDPCDecodeSetNormalization(self.dec.dec, C.byref(cfg)))
            self._norm_cache[key] = cfg.norm.f16x2 & 0xFFFF
        return self._norm_cache[key]

    def build(self, tbs, order="lpt"):
        if not tbs:
            self.launches = []
            return self
        dtypes = {tb.llr.dtype for tb in tbs}
        assert len(dtypes) == 1, "one LLR type per slot"
        self.llr_type = _LLR_TYPES[dtypes.pop()]
        n = len(tbs)
        assert n <= 65536
        # work items hold a 12-bit CB index (make_item)
        assert all(tb.params.num_CBs <= 4096 for tb in tbs), "at most 4096 code blocks per TB"

        parr = (L.PerTbParams * n)(*[tb.params for tb in tbs])
        self._dev_params = torch.from_numpy(np.frombuffer(bytes(parr), dtype=np.uint8).copy()).cuda()
        psize = C.sizeof(L.PerTbParams)
        jobs = (L.PuschSlotTb * n)()
        share_offset = 0
        counts = [0, 0, 0]
        for tb in tbs:
            counts[row_class(tb.mb)] += tb.params.num_CBs
        launch_class = launch_classes(counts, self.split_min)
        classes = {}
        for t, tb in enumerate(tbs):
            p = tb.params
            j = jobs[t]
            j.prm = self._dev_params.data_ptr() + t * psize
            j.llr = tb.llr.data_ptr()
            j.harq = tb.harq.data_ptr()
            j.cb_bits = tb.cb_bits.data_ptr()
            j.cb_crc = tb.cb_crc.data_ptr()
            j.tb_bytes = tb.tb_bytes.data_ptr()
            j.tb_crc = tb.tb_crc.data_ptr()
            j.cb_iterations = tb.cb_iterations.data_ptr() if tb.cb_iterations is not None else None
            j.num_parity_nodes = tb.mb
            j.max_iterations = tb.max_iterations
            j.flags = ((L.CUPHY_LDPC_DECODE_EARLY_TERM if tb.early_term else 0) |
                       (L.CUPHY_HIP_SLOT_TB_DESCRAMBLE if tb.descramble else 0))
            j.llr_scale = tb.llr_scale
            h = self._alpha_bits(tb)
            j.norm.f16x2 = h | (h << 16)
            j.clamp_value = tb.clamp_value
            j.share_offset = share_offset
            share_offset += p.num_CBs
            cost = p.Zc * num_edges(p.bg, tb.mb) * max(1, tb.max_iterations)
            key = launch_class[row_class(tb.mb)]
            g = group_size(p, tb.mb, tb.descramble) if self.pack else 1
            for c in range(0, p.num_CBs, g):
                n_cb = min(g, p.num_CBs - c)
                classes.setdefault(key, []).append((cost * n_cb, t, c, n_cb))
        self._jobs = torch.from_numpy(np.frombuffer(bytes(jobs), dtype=np.uint8).copy()).cuda()
        self.n_tbs = n
        sz = C.c_size_t()
        L.check("cuphyHipPuschSlotDecodeGetWorkspaceSize",
                self._lib.cuphyHipPuschSlotDecodeGetWorkspaceSize(share_offset, C.byref(sz)))
        self.launches = []
        for key in sorted(classes):
            rows = classes[key]
            if order == "lpt":
                rows.sort(key=lambda r: -r[0])
            ln = _Launch()
            ids = np.array([make_item(t, c, n_cb) for _, t, c, n_cb in rows], dtype=np.uint32).view(np.int32)
            ln.items = torch.from_numpy(ids).cuda()
            ln.n_items = len(rows)
            for _, t, _, n_cb in rows:
                ln.max_values = max(ln.max_values, lds_values(tbs[t].params, tbs[t].mb, tbs[t].descramble, n_cb))
                ln.max_group = max(ln.max_group, n_cb)
                ln.max_mb = max(ln.max_mb, tbs[t].mb)
            ln.workspace = torch.zeros(sz.value // 4, dtype=torch.int32, device="cuda")
            self.launches.append(ln)
        return self

    @property
    def workspaces(self):
        return [ln.workspace for ln in self.launches]

    def _launch_one(self, ln, stream, n_items=None):
        n_items = ln.n_items if n_items is None else n_items
        st = self._lib.cuphyHipPuschSlotDecode(self.dec.dec, C.c_void_p(self._jobs.data_ptr()), self.n_tbs,
                                               C.c_void_p(ln.items.data_ptr()), n_items, self.llr_type,
                                               ln.max_values, ln.max_mb, ln.max_group,
                                               C.c_void_p(ln.workspace.data_ptr()),
                                               _stream_handle(stream))
        return L.check("cuphyHipPuschSlotDecode", st)

    def launch(self, stream=None):
        main = stream if stream is not None else torch.cuda.current_stream()
        if not self.launches:
            return
        if len(self.launches) == 1 or not self.concurrent:
            for ln in self.launches:
                self._launch_one(ln, main)
            return
        aux = self.aux_streams[:len(self.launches) - 1]
        for s in aux:
            s.wait_stream(main)
        self._launch_one(self.launches[0], main)
        for ln, s in zip(self.launches[1:], aux):
            self._launch_one(ln, s)
        for s in aux:
            main.wait_stream(s)

    def capture(self):
        """Capture one launch into a HIP graph.  The captured launch combines
        into the HARQ buffers when replayed, like an eager launch."""
        for ln in self.launches:  # kernel attributes set outside the capture
            self._launch_one(ln, torch.cuda.current_stream(), n_items=0)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch()
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
