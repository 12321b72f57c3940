"""PUSCH back end around the decoder: de-rate-matching/HARQ combining and
CB/TB CRC checking through the C ABI (cuphySetupPuschRxRateMatch,
cuphySetupPuschRxCrcDecode), plus PerTbParams construction.

Mirrors pyAerial's LdpcDeRateMatch / CrcChecker (pyaerial/src/aerial/phy5g/
ldpc/derate_match.py, crc_check.py) and the parameter derivation of
pycuphy_ldpc_util.cpp:75-171; arrays are torch device tensors.
"""
import ctypes as C

import numpy as np
import torch

from . import _hip
from . import _lib as L
from .decoder import _stream_handle

LLR_MAX_ABS = 10000.0


def k0_of(bg, rv, Ncb, Zc):
    """TS 38.212 Table 5.4.2.1-2 (pycuphy_ldpc_util.cpp:122-157)."""
    num = (0, 17, 33, 56) if bg == 1 else (0, 13, 25, 43)
    den = 66 if bg == 1 else 50
    return (num[rv] * Ncb // (den * Zc)) * Zc


def num_parity_nodes(bg, Zc, K, F, G, C, k0, Ncb):
    """mb covering the received parity (pycuphy_ldpc_util.cpp:117-157)."""
    Kd = K - F - 2 * Zc
    ncb_par = min(G // C + k0, Ncb)
    mb = (ncb_par - Kd + Zc - 1) // Zc
    return max(4, min(46 if bg == 1 else 42, mb))


def make_tb_params(bg, Zc, C, Kprime, Qm, Nl, G, rv, ndi=1, n_bbu_layers=None, layer_map=None, user_group=0,
                   first_cb=0, Nref=0):
    """PerTbParams of one TB (fields as setPerTbParams fills them)."""
    p = L.PerTbParams()
    p.ndi, p.rv, p.Qm, p.bg, p.Nl, p.num_CBs, p.Zc = ndi, rv, Qm, bg, Nl, C, Zc
    N = (66 if bg == 1 else 50) * Zc
    p.N = N
    p.Ncb = min(N, Nref) if Nref > 0 else N
    p.Ncb_padded = (N + 2 * Zc + 7) // 8 * 8
    p.G = G
    p.K = (22 if bg == 1 else 10) * Zc
    p.F = p.K - Kprime
    p.k0 = k0_of(bg, rv, p.Ncb, Zc)
    p.encodedSize = G
    p.firstCodeBlockIndex = first_cb
    p.nBBULayers = Nl if n_bbu_layers is None else n_bbu_layers
    lm = list(range(Nl)) if layer_map is None else layer_map
    for i in range(Nl):
        p.layer_map_array[i] = lm[i]
    p.userGroupIndex = user_group
    p.isDataPresent = 1
    mb = num_parity_nodes(bg, Zc, p.K, p.F, G, C, p.k0, p.Ncb)
    p.nZpBitsPerCb = mb * Zc + p.K
    return p, mb


def get_tb_size(num_symbols, num_prbs, num_layers, code_rate, mod_order, num_dmrs_cdm_grps_no_data1_symbols=0):
    """TS 38.214 5.1.3.2 transport block size and code-block count of an
    allocation (cuphyHipTransportBlockSize; pyAerial get_tb_size,
    util.py:171-333, with the code rate as a fraction)."""
    tbs, ncb = C.c_uint32(), C.c_uint32()
    L.check("cuphyHipTransportBlockSize",
            L.lib().cuphyHipTransportBlockSize(num_symbols, num_prbs, num_layers, code_rate, mod_order,
                                               num_dmrs_cdm_grps_no_data1_symbols, C.byref(tbs), C.byref(ncb)))
    return tbs.value, ncb.value


def set_per_tb_params(tb_size, code_rate, qam_mod, rate_match_len, rv, ndi=1, cinit=0, user_group=0, num_layers=1,
                      num_ue_grp_layers=None, layer_map=None, n_dmrs_cdm_grps_no_data=2, lbrm=None):
    """PerTbParams of a PUSCH TB from its size (bits, without CRC) and code
    rate, as pyAerial's setPerTbParams (pycuphy_ldpc_util.cpp:75-171) with
    the PUSCH pipeline's limited-buffer rate matching (lbrm = (n_PRB_LBRM,
    maxLayers, maxQm) or None).  Returns (params, parity nodes, info nodes Kb)."""
    p = L.PerTbParams()
    nl = num_layers
    ng = nl if num_ue_grp_layers is None else num_ue_grp_layers
    lm = (C.c_uint32 * L.MAX_N_LAYERS_PUSCH)(*(list(range(nl)) if layer_map is None else list(layer_map)))
    lb = None
    if lbrm is not None:
        lb = L.LbrmPrm(i_lbrm=1, maxLayers=lbrm[1], maxQm=lbrm[2], n_PRB_LBRM=lbrm[0])
    mb, kb = C.c_uint16(), C.c_uint16()
    L.check("cuphyHipSetPerTbParams",
            L.lib().cuphyHipSetPerTbParams(C.byref(p), tb_size, code_rate, qam_mod, ndi, rv, rate_match_len, cinit,
                                           user_group, nl, ng, lm, n_dmrs_cdm_grps_no_data,
                                           C.byref(lb) if lb is not None else None, C.byref(mb), C.byref(kb)))
    return p, mb.value, kb.value


def params_to_device(params):
    arr = (L.PerTbParams * len(params))(*params)
    host = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
    return arr, torch.from_numpy(host).cuda()


class DeRateMatch:
    """cuphyCreate/Setup/DestroyPuschRxRateMatch + launch of its kernel nodes."""

    def __init__(self, fp_config=3, descrambling=False):
        """descrambling: negate LLR n of each TB where its Gold sequence
        (PerTbParams.cinit) has c(n) = 1, as cuphyCreatePuschRxRateMatch's
        descramblingOn."""
        self._lib = L.lib()
        self.h = C.c_void_p()
        L.check("cuphyCreatePuschRxRateMatch",
                self._lib.cuphyCreatePuschRxRateMatch(C.byref(self.h), fp_config, int(descrambling)))
        sz, al = C.c_size_t(), C.c_size_t()
        L.check("cuphyPuschRxRateMatchGetDescrInfo",
                self._lib.cuphyPuschRxRateMatchGetDescrInfo(C.byref(sz), C.byref(al)))
        self.cpu_desc = (C.c_uint8 * sz.value)()
        self.gpu_desc = torch.zeros(sz.value, dtype=torch.uint8, device="cuda")
        self.cfg = L.RateMatchLaunchCfg()

    def close(self):
        if self.h:
            self._lib.cuphyDestroyPuschRxRateMatch(self.h)
            self.h = C.c_void_p()

    def run(self, params, llr_in, harq_out, stream=None):
        """params: list of PerTbParams (UE index = list index); llr_in: one
        device tensor per user group; harq_out: one device tensor per UE."""
        n = len(params)
        self._host_prms, self._dev_prms = params_to_device(params)
        idx = (C.c_uint16 * n)(*range(n))
        tin = (L.TensorPrm * len(llr_in))()
        for i, t in enumerate(llr_in):
            tin[i].pAddr = t.data_ptr()
        self._out_ptrs = torch.tensor([t.data_ptr() for t in harq_out], dtype=torch.int64).cuda()
        st = self._lib.cuphySetupPuschRxRateMatch(
            self.h, n, idx, self._host_prms, C.c_void_p(self._dev_prms.data_ptr()), tin, tin,
            C.cast(C.c_void_p(self._out_ptrs.data_ptr()), C.POINTER(C.c_void_p)), C.cast(self.cpu_desc, C.c_void_p),
            C.c_void_p(self.gpu_desc.data_ptr()), 1, C.c_void_p(C.addressof(self.cfg)), _stream_handle(stream))
        L.check("cuphySetupPuschRxRateMatch", st)
        sh = _stream_handle(stream)
        _hip.launch_node(self.cfg.resetKernelNodeParamsDriver, sh)
        _hip.launch_node(self.cfg.kernelNodeParamsDriver, sh)
        _hip.launch_node(self.cfg.clampKernelNodeParamsDriver, sh)


class CrcDecode:
    """cuphyCreate/Setup/DestroyPuschRxCrcDecode + launch of CB and TB kernels."""

    def __init__(self, reverse_bytes=True):
        self._lib = L.lib()
        self.h = C.c_void_p()
        L.check("cuphyCreatePuschRxCrcDecode", self._lib.cuphyCreatePuschRxCrcDecode(C.byref(self.h), int(reverse_bytes)))
        sz, al = C.c_size_t(), C.c_size_t()
        L.check("cuphyPuschRxCrcDecodeGetDescrInfo", self._lib.cuphyPuschRxCrcDecodeGetDescrInfo(C.byref(sz), C.byref(al)))
        self.cpu_desc = (C.c_uint8 * sz.value)()
        self.gpu_desc = torch.zeros(sz.value, dtype=torch.uint8, device="cuda")
        self.cb_cfg = L.CrcDecodeLaunchCfg()
        self.tb_cfg = L.CrcDecodeLaunchCfg()

    def close(self):
        if self.h:
            self._lib.cuphyDestroyPuschRxCrcDecode(self.h)
            self.h = C.c_void_p()

    def run(self, params, cb_words, cb_crc, tb_bytes, tb_crc, stream=None):
        n = len(params)
        self._host_prms, self._dev_prms = params_to_device(params)
        idx = (C.c_uint16 * n)(*range(n))
        st = self._lib.cuphySetupPuschRxCrcDecode(
            self.h, n, idx, C.c_void_p(cb_crc.data_ptr()), C.c_void_p(tb_bytes.data_ptr()),
            C.c_void_p(cb_words.data_ptr()), C.c_void_p(tb_crc.data_ptr()), self._host_prms,
            C.c_void_p(self._dev_prms.data_ptr()), C.cast(self.cpu_desc, C.c_void_p),
            C.c_void_p(self.gpu_desc.data_ptr()), 1, C.c_void_p(C.addressof(self.cb_cfg)), C.c_void_p(C.addressof(self.tb_cfg)),
            _stream_handle(stream))
        L.check("cuphySetupPuschRxCrcDecode", st)
        sh = _stream_handle(stream)
        _hip.launch_node(self.cb_cfg.kernelNodeParamsDriver, sh)
        _hip.launch_node(self.tb_cfg.kernelNodeParamsDriver, sh)
