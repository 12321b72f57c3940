"""PuschRx-style LDPC graph layer: transport blocks of a slot grouped into
decode descriptors by configuration, one graph kernel node per descriptor,
node parameters updated and unused nodes disabled per slot.

This is the way the reference's PUSCH pipeline drives the decoder
(pusch_rx.cpp:3007-3021 adds nMaxLdpcHetConfigs decoder nodes to its graph
with empty parameters; pusch_rx.cpp:6275-6415 (prepareLDPCStreamsTB)
collects the TBs of a slot with identical (BG, Z, mb, iterations) into
descriptors of at most 32 TBs; pusch_rx.cpp:4450-4468 sets each used node's
parameters from cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor with
cuGraphExecKernelNodeSetParams and enables it, and disables the rest).  The
same steps here use the HIP graph API on the launch descriptors this
library returns, so a PuschRx built on HIP drives this decoder unchanged.
"""
import ctypes as C

import torch

from . import _lib as L
from .decoder import LdpcDecodeConfig, _stream_handle
from .pusch import DeRateMatch
from .workloads import nsys

_api = None


def _hip_graph_api():
    global _api
    if _api is None:
        h = C.CDLL("libamdhip64.so")
        VP, P = C.c_void_p, C.POINTER
        h.hipGraphCreate.argtypes = [P(VP), C.c_uint]
        h.hipGraphAddKernelNode.argtypes = [P(VP), VP, P(VP), C.c_size_t, P(L.HipKernelNodeParams)]
        h.hipGraphInstantiate.argtypes = [P(VP), VP, P(VP), C.c_char_p, C.c_size_t]
        h.hipGraphExecKernelNodeSetParams.argtypes = [VP, VP, P(L.HipKernelNodeParams)]
        h.hipGraphNodeSetEnabled.argtypes = [VP, VP, C.c_uint]
        h.hipGraphLaunch.argtypes = [VP, VP]
        h.hipGraphExecDestroy.argtypes = [VP]
        h.hipGraphDestroy.argtypes = [VP]
        for f in ("hipGraphCreate", "hipGraphAddKernelNode", "hipGraphInstantiate", "hipGraphExecKernelNodeSetParams",
                  "hipGraphNodeSetEnabled", "hipGraphLaunch", "hipGraphExecDestroy", "hipGraphDestroy"):
            getattr(h, f).restype = C.c_int
        _api = h
    return _api


def _ok(name, rc):
    if rc != 0:
        raise RuntimeError("%s failed: hipError %d" % (name, rc))


class LdpcDecodeGraph:
    """A graph with `max_configs` decoder kernel nodes (the reference's
    nMaxLdpcHetConfigs), instantiated once and updated every slot."""

    def __init__(self, decoder, max_configs=8):
        self.dec = decoder
        self.max_configs = max_configs
        self._h = _hip_graph_api()
        # empty node parameters: the library's empty kernel, which the
        # rate-match API hands out for its reset/clamp slots
        self._rm = DeRateMatch(fp_config=3)
        empty = self._empty_params()
        self.graph = C.c_void_p()
        _ok("hipGraphCreate", self._h.hipGraphCreate(C.byref(self.graph), 0))
        self.nodes = []
        for _ in range(max_configs):
            n = C.c_void_p()
            _ok("hipGraphAddKernelNode", self._h.hipGraphAddKernelNode(C.byref(n), self.graph, None, 0,
                                                                      C.byref(empty)))
            self.nodes.append(n)
        self.exec = C.c_void_p()
        _ok("hipGraphInstantiate", self._h.hipGraphInstantiate(C.byref(self.exec), self.graph, None, None, 0))
        self.enabled = [True] * max_configs
        self._cfgs = []

    def _empty_params(self):
        from .pusch import make_tb_params
        p, _ = make_tb_params(2, 8, 1, 80, 2, 1, 200, 0, ndi=1)
        dummy_in = torch.zeros(1024, dtype=torch.float16, device="cuda")
        dummy_out = torch.zeros(1024, dtype=torch.float16, device="cuda")
        self._dummy = (dummy_in, dummy_out)
        self._rm.run([p], [dummy_in], [dummy_out])  # setup fills the launch configuration
        torch.cuda.synchronize()
        node = L.HipKernelNodeParams()
        C.memmove(C.byref(node), C.byref(self._rm.cfg.resetKernelNodeParamsDriver), C.sizeof(node))
        return node

    def close(self):
        if self.exec:
            self._h.hipGraphExecDestroy(self.exec)
            self.exec = C.c_void_p()
        if self.graph:
            self._h.hipGraphDestroy(self.graph)
            self.graph = C.c_void_p()
        self._rm.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def group(tbs, max_iterations):
        """prepareLDPCStreamsTB: TBs with identical (BG, Z, mb, Kb, llr type,
        iterations) share a descriptor, at most 32 TBs each."""
        sets = {}
        order = []
        for t, tb in enumerate(tbs):
            Kb = getattr(tb, "Kb", None) or nsys(tb.bg)
            key = (tb.bg, tb.Z, tb.mb, Kb, tb.llr.dtype, max_iterations)
            lst = sets.get(key)
            if lst is None or len(lst[-1]) == L.CUPHY_LDPC_DECODE_DESC_MAX_TB:
                sets.setdefault(key, []).append([])
                order.append((key, len(sets[key]) - 1))
            sets[key][-1].append(t)
        return [(key, sets[key][i]) for key, i in order]

    def run(self, tbs, max_iterations=10, alpha=0.8125, clamp_value=32.0, flags=0, stream=None):
        """Decode the slot's TBs (objects with bg, Z, mb, llr (ncb, stride)
        and out (ncb, words) device tensors) through the graph."""
        groups = self.group(tbs, max_iterations)
        if len(groups) > self.max_configs:
            raise ValueError("%d decoder configurations in the slot, graph has %d nodes"
                             % (len(groups), self.max_configs))
        lib = L.lib()
        self._cfgs = []
        for i, (key, members) in enumerate(groups):
            bg, Z, mb, Kb, dtype, iters = key
            llr_type = L.CUPHY_R_16F if dtype == torch.float16 else L.CUPHY_R_32F
            cfg = LdpcDecodeConfig(bg, Z, mb, Kb=Kb, llr_type=llr_type, max_iterations=iters, alpha=alpha,
                                   clamp_value=clamp_value, flags=flags)
            lc = L.LDPCDecodeLaunchConfig()
            lc.decode_desc = self.dec.build_desc(cfg, [tbs[t].llr for t in members], [tbs[t].out for t in members])
            L.check("cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor",
                    lib.cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor(self.dec.dec, C.byref(lc)))
            _ok("hipGraphExecKernelNodeSetParams",
                self._h.hipGraphExecKernelNodeSetParams(self.exec, self.nodes[i],
                                                        C.byref(lc.kernel_node_params_driver)))
            if not self.enabled[i]:
                _ok("hipGraphNodeSetEnabled", self._h.hipGraphNodeSetEnabled(self.exec, self.nodes[i], 1))
                self.enabled[i] = True
            self._cfgs.append(lc)
        for i in range(len(groups), self.max_configs):
            if self.enabled[i]:
                _ok("hipGraphNodeSetEnabled", self._h.hipGraphNodeSetEnabled(self.exec, self.nodes[i], 0))
                self.enabled[i] = False
        _ok("hipGraphLaunch", self._h.hipGraphLaunch(self.exec, _stream_handle(stream)))
        return len(groups)
