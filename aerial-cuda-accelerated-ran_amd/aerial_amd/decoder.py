"""Host-side wrapper of the LDPC decoder C ABI over torch device tensors.

Mirrors the reference's C++ convenience layer cuphy::LDPC_decoder /
LDPC_decode_desc_set (cuPHY/src/cuphy/ldpc/ldpc_api.hpp:104-321): a decoder
object, a per-(BG, Z, mb) configuration, and transport-block descriptors of
up to 32 TBs per launch.  Tensors must live on the GPU; the call is
asynchronous on the given (or current) torch stream.
"""
import ctypes as C

import torch

from . import _lib as L


def _stream_handle(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class LdpcDecodeConfig:
    """cuphyLDPCDecodeConfigDesc_t builder (cuphy.h:3945-3960)."""

    def __init__(self, bg, Z, mb, Kb=None, llr_type=L.CUPHY_R_32F, max_iterations=10, alpha=None,
                 clamp_value=32.0, flags=0, algo=0):
        c = L.LDPCDecodeConfigDesc()
        c.llr_type = llr_type
        c.num_parity_nodes = mb
        c.Z = Z
        c.max_iterations = max_iterations
        c.Kb = Kb if Kb is not None else (22 if bg == 1 else 10)
        c.flags = flags
        c.BG = bg
        c.algo = algo
        c.clamp_value = clamp_value
        c.workspace = None
        self.desc = c
        self.alpha = alpha

    def nsys(self):
        return 22 if self.desc.BG == 1 else 10

    def num_llrs(self):
        return (self.nsys() + self.desc.num_parity_nodes) * self.desc.Z

    def out_words(self):
        return (self.desc.Kb * self.desc.Z + 31) // 32


class LdpcDecoder:
    """cuphyCreateLDPCDecoder / TransportBlockDecode / Destroy."""

    def __init__(self):
        self._lib = L.lib()
        self.ctx = C.c_void_p()
        L.check("cuphyCreateContext", self._lib.cuphyCreateContext(C.byref(self.ctx), 0))
        self.dec = C.c_void_p()
        L.check("cuphyCreateLDPCDecoder", self._lib.cuphyCreateLDPCDecoder(self.ctx, C.byref(self.dec), 0))

    def close(self):
        if self.dec:
            self._lib.cuphyDestroyLDPCDecoder(self.dec)
            self.dec = C.c_void_p()
        if self.ctx:
            self._lib.cuphyDestroyContext(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def finalize_config(self, cfg):
        """Fill the normalisation: the table value (SetNormalization) or cfg.alpha."""
        c = cfg.desc
        if cfg.alpha is None:
            L.check("cuphyErrorCorrectionLDPCDecodeSetNormalization",
                    self._lib.cuphyErrorCorrectionLDPCDecodeSetNormalization(self.dec, C.byref(c)))
        elif c.llr_type == L.CUPHY_R_16F or c.algo == 2:
            h = int(torch.tensor([cfg.alpha], dtype=torch.float16).view(torch.int16).item()) & 0xFFFF
            c.norm.f16x2 = h | (h << 16)
        else:
            c.norm.f32 = cfg.alpha
        return c

    def build_desc(self, cfg, llrs, outs, softs=None):
        """LDPCDecodeDesc for up to 32 TBs.

        llrs[i]: (num_cw, stride_elements) float32/float16 device tensor
        outs[i]: (num_cw, stride_words) int32 device tensor
        softs[i]: optional (num_cw, stride) float16 device tensor
        """
        assert 1 <= len(llrs) <= L.CUPHY_LDPC_DECODE_DESC_MAX_TB and len(outs) == len(llrs)
        d = L.LDPCDecodeDesc()
        d.config = self.finalize_config(cfg)
        d.num_tbs = len(llrs)
        want = torch.float32 if d.config.llr_type == L.CUPHY_R_32F else torch.float16
        for i, (x, o) in enumerate(zip(llrs, outs)):
            assert x.is_cuda and o.is_cuda and x.dtype == want and o.dtype == torch.int32
            assert x.dim() == 2 and o.dim() == 2 and x.stride(1) == 1 and o.stride(1) == 1
            assert x.shape[0] == o.shape[0] and x.shape[1] >= cfg.num_llrs() and o.shape[1] >= cfg.out_words()
            d.llr_input[i].addr = x.data_ptr()
            d.llr_input[i].stride_elements = x.stride(0)
            d.llr_input[i].num_codewords = x.shape[0]
            d.tb_output[i].addr = o.data_ptr()
            d.tb_output[i].stride_words = o.stride(0)
            d.tb_output[i].num_codewords = o.shape[0]
            if softs is not None and softs[i] is not None:
                s = softs[i]
                assert s.is_cuda and s.dtype == torch.float16 and s.shape[0] == x.shape[0]
                d.llr_output[i].addr = s.data_ptr()
                d.llr_output[i].stride_elements = s.stride(0)
                d.llr_output[i].num_codewords = s.shape[0]
        if softs is not None:
            d.config.flags |= L.CUPHY_LDPC_DECODE_WRITE_SOFT_OUTPUTS
        return d

    def decode_tb(self, cfg, llrs, outs, softs=None, stream=None):
        d = self.build_desc(cfg, llrs, outs, softs)
        return self.launch(d, stream)

    def launch(self, desc, stream=None):
        st = self._lib.cuphyErrorCorrectionLDPCTransportBlockDecode(self.dec, C.byref(desc), _stream_handle(stream))
        return L.check("cuphyErrorCorrectionLDPCTransportBlockDecode", st)
