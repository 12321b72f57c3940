"""LDPC encoding and DL rate matching through the C ABI
(cuphyLDPCEncodeGetDescrInfo / cuphySetupLDPCEncode, cuphyHipDlRateMatch).

LdpcEncoder mirrors pyAerial's LdpcEncoder (pyaerial/src/aerial/phy5g/ldpc/
encoder.py, pybind11/pycuphy_ldpc_encoder.cpp): code blocks of K' bits per
TB in, coded bits (N = 66 Zc / 50 Zc, the first 2 Zc systematic bits
punctured) out, with the base graph and lifting size derived from the TB
size and code rate.  Arrays are torch device tensors of 0/1 bytes; the
encoder itself works on bit-packed 32-bit words (bit i of word w = bit
32 w + i).
"""
import ctypes as C

import numpy as np
import torch

from . import _hip
from . import _lib as L
from .decoder import _stream_handle
from .pusch import params_to_device, set_per_tb_params


def pack_bits(bits):
    """(n, C) 0/1 uint8 device tensor -> (C, ceil(n/32)) int32 words."""
    n, ncw = bits.shape
    nw = (n + 31) // 32
    padded = torch.zeros((ncw, nw * 32), dtype=torch.int64, device=bits.device)
    padded[:, :n] = bits.t().to(torch.int64)
    weights = torch.tensor([1 << i for i in range(32)], dtype=torch.int64, device=bits.device)
    words = (padded.view(ncw, nw, 32) * weights).sum(dim=2)
    return torch.where(words >= 1 << 31, words - (1 << 32), words).to(torch.int32)


def unpack_bits(words, n):
    """(C, nw) int32 words -> (n, C) uint8 bits."""
    w = words.to(torch.int64) & 0xFFFFFFFF
    shifts = torch.arange(32, device=words.device, dtype=torch.int64)
    bits = ((w.unsqueeze(-1) >> shifts) & 1).reshape(words.shape[0], -1)[:, :n]
    return bits.t().contiguous().to(torch.uint8)


class _BitTensor:
    """A cuphyTensorDescriptor_t of type CUPHY_BIT [nbits x ncw], column stride in bits."""

    def __init__(self, nbits, ncw, stride_bits):
        self._lib = L.lib()
        self.h = C.c_void_p()
        L.check("cuphyCreateTensorDescriptor", self._lib.cuphyCreateTensorDescriptor(C.byref(self.h)))
        dims = (C.c_int * 2)(nbits, ncw)
        strides = (C.c_int * 2)(1, stride_bits)
        L.check("cuphySetTensorDescriptor", self._lib.cuphySetTensorDescriptor(self.h, L.CUPHY_BIT, 2, dims, strides, 0))

    def __del__(self):
        try:
            self._lib.cuphyDestroyTensorDescriptor(self.h)
        except Exception:
            pass


class LdpcEncoder:
    """Encoder over the reference's encoder API: one setup and one kernel
    launch per call (all TBs of one (BG, Zc) batched in it)."""

    def __init__(self, puncturing=True, max_tbs=64):
        self._lib = L.lib()
        self.puncturing = puncturing
        self.max_tbs = max_tbs
        dsz, dal, wsz = C.c_size_t(), C.c_size_t(), C.c_size_t()
        L.check("cuphyLDPCEncodeGetDescrInfo",
                self._lib.cuphyLDPCEncodeGetDescrInfo(C.byref(dsz), C.byref(dal), max_tbs, C.byref(wsz)))
        self.cpu_desc = (C.c_uint8 * dsz.value)()
        self.gpu_desc = torch.zeros(dsz.value, dtype=torch.uint8, device="cuda")
        self.h_ws = (C.c_uint8 * wsz.value)()
        self.d_ws = torch.zeros(wsz.value, dtype=torch.uint8, device="cuda")
        self.cfg = L.LDPCEncodeLaunchConfig()

    def encode_words(self, bg, Z, info_words, num_rows=0, rv=0, stream=None):
        """info_words: list of (C, in_words) int32 device tensors (one per TB,
        equal shapes), Kb * Z information bits per code word (Kb = 22 / 10)
        -> list of (C, out_words) int32 tensors of coded bits."""
        n = len(info_words)
        assert 1 <= n <= self.max_tbs
        ncw, in_words = info_words[0].shape
        nsys = 22 if bg == 1 else 10
        rows = num_rows if (rv == 0 and 0 < num_rows < (46 if bg == 1 else 42)) else (46 if bg == 1 else 42)
        nbits = (nsys + rows) * Z - (2 * Z if self.puncturing else 0)
        out_words = (nbits + 31) // 32
        outs = [torch.empty((ncw, out_words), dtype=torch.int32, device="cuda") for _ in range(n)]
        tin = _BitTensor(in_words * 32, ncw, in_words * 32)
        tout = _BitTensor(out_words * 32, ncw, out_words * 32)
        in_ptrs = (C.c_void_p * n)(*[x.data_ptr() for x in info_words])
        out_ptrs = (C.c_void_p * n)(*[o.data_ptr() for o in outs])
        st = self._lib.cuphySetupLDPCEncode(
            C.byref(self.cfg), tin.h, None, tout.h, None, bg, Z, int(self.puncturing), num_rows, rv, 1, n,
            C.cast(in_ptrs, C.c_void_p), C.cast(out_ptrs, C.c_void_p), C.cast(self.h_ws, C.c_void_p),
            C.c_void_p(self.d_ws.data_ptr()), C.cast(self.cpu_desc, C.c_void_p), C.c_void_p(self.gpu_desc.data_ptr()),
            1, _stream_handle(stream))
        L.check("cuphySetupLDPCEncode", st)
        _hip.launch_node(self.cfg.m_kernelNodeParams, _stream_handle(stream))
        return outs

    def encode(self, code_blocks, tb_sizes, code_rates, redundancy_versions):
        """pyAerial's encode(): code_blocks[t] is a (K', C) 0/1 uint8 device
        tensor (CRCs attached, segmented); returns (N, C) coded bits per TB,
        N = 66 Zc / 50 Zc (68 Zc / 52 Zc without puncturing)."""
        out = []
        for cbs, tbs, R, rv in zip(code_blocks, tb_sizes, code_rates, redundancy_versions):
            kp, ncw = cbs.shape
            p, _, _ = set_per_tb_params(tbs, R, 2, 2 * ncw, rv)
            assert p.num_CBs == ncw and p.K - p.F == kp, "code blocks do not match the TB size and code rate"
            info = torch.zeros((p.K, ncw), dtype=torch.uint8, device=cbs.device)
            info[:kp] = cbs  # filler bits 0
            words = self.encode_words(p.bg, p.Zc, [pack_bits(info)], rv=rv)[0]
            n = (66 if p.bg == 1 else 50) * p.Zc + (0 if self.puncturing else 2 * p.Zc)
            out.append(unpack_bits(words, n))
        return out


def dl_rate_match(params, coded_words, emax, scrambling=False, stream=None):
    """cuphyHipDlRateMatch over TBs: params[t] a PerTbParams (N, Ncb, k0,
    G in encodedSize, Qm, Nl, num_CBs, cinit ...), coded_words[t] a (C,
    ceil(N/32)) int32 device tensor of the encoder's punctured output ->
    (num_tbs, max_cbs, emax/32) int32 words, CB r of TB t in its slot."""
    n = len(params)
    max_cbs = max(p.num_CBs for p in params)
    _, dev = params_to_device(params)
    ptrs = torch.tensor([w.data_ptr() for w in coded_words], dtype=torch.int64, device="cuda")
    out = torch.full((n, max_cbs, emax // 32), -1, dtype=torch.int32, device="cuda")
    st = L.lib().cuphyHipDlRateMatch(C.c_void_p(dev.data_ptr()), n, C.c_void_p(ptrs.data_ptr()), max_cbs, emax,
                                     int(scrambling), C.c_void_p(out.data_ptr()), _stream_handle(stream))
    L.check("cuphyHipDlRateMatch", st)
    out._keep = (dev, ptrs)  # the argument buffers live until the caller synchronizes
    return out
