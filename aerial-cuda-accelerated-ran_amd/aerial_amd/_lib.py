"""ctypes view of the C ABI in include/cuphy_hip.h (lib/libcuphy_hip.so).

The structures below mirror the header field by field.  Loading fails
loudly (ImportError) when the HIP library has not been built: there is no
CPU fallback anywhere in this package.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libcuphy_hip.so")

# cuphyStatus_t
CUPHY_STATUS_SUCCESS = 0
CUPHY_STATUS_INTERNAL_ERROR = 1
CUPHY_STATUS_NOT_SUPPORTED = 2
CUPHY_STATUS_INVALID_ARGUMENT = 3
CUPHY_STATUS_ALLOC_FAILED = 5
CUPHY_STATUS_UNSUPPORTED_TYPE = 9
CUPHY_STATUS_UNSUPPORTED_RANK = 11
CUPHY_STATUS_UNSUPPORTED_CONFIG = 12

# cuphyDataType_t
CUPHY_R_32F = 0
CUPHY_R_16F = 2
CUPHY_R_8I = 3
CUPHY_BIT = 20

# decoder flags
CUPHY_LDPC_DECODE_DEFAULT = 0
CUPHY_LDPC_DECODE_EARLY_TERM = 1
CUPHY_LDPC_DECODE_CHOOSE_THROUGHPUT = 2
CUPHY_LDPC_DECODE_WRITE_SOFT_OUTPUTS = 4

CUPHY_LDPC_DECODE_DESC_MAX_TB = 32
MAX_N_LAYERS_PUSCH = 8
CUPHY_MAX_N_CSI1_PRMS = 4
CUPHY_MAX_N_CSI2_REPORTS_PER_UE = 16
MAX_WORDS_PER_CODE_BLOCK = 264
SLOT_CB_ERROR = 0xFFFFFFFF  # CRC value of a CB/TB the fused slot kernel rejected
# reasons a batch launch did not decode a work item (cuphyHipLDPCWorkItem_t::status)
ITEM_BAD_BG, ITEM_BAD_Z, ITEM_BAD_MB, ITEM_BAD_KB = 0x01, 0x02, 0x04, 0x08
ITEM_BAD_ITER, ITEM_NULL_BUFFER, ITEM_NO_FIT, ITEM_GROUP = 0x10, 0x20, 0x40, 0x80
CUPHY_HIP_HARQ_LAST_TX = 0x100
CUPHY_HIP_SLOT_TB_NOTIFY = 0x200
CUPHY_HIP_SLOT_TB_DESCRAMBLE = 0x400
CUPHY_HIP_HARQ_SERVER_RESIDENT = 0x1


class Norm(C.Union):
    _fields_ = [("f32", C.c_float), ("f16x2", C.c_uint32)]


class LDPCDecodeConfigDesc(C.Structure):
    _fields_ = [("llr_type", C.c_int), ("num_parity_nodes", C.c_int16), ("Z", C.c_int16),
                ("max_iterations", C.c_int16), ("Kb", C.c_int16), ("norm", Norm), ("flags", C.c_uint32),
                ("BG", C.c_int16), ("algo", C.c_int16), ("clamp_value", C.c_float), ("workspace", C.c_void_p)]


class TransportBlockLLRDesc(C.Structure):
    _fields_ = [("addr", C.c_void_p), ("stride_elements", C.c_int32), ("num_codewords", C.c_int32)]


class TransportBlockDataDesc(C.Structure):
    _fields_ = [("addr", C.c_void_p), ("stride_words", C.c_int32), ("num_codewords", C.c_int32)]


class LDPCDecodeDesc(C.Structure):
    _fields_ = [("config", LDPCDecodeConfigDesc), ("num_tbs", C.c_int32),
                ("llr_input", TransportBlockLLRDesc * CUPHY_LDPC_DECODE_DESC_MAX_TB),
                ("tb_output", TransportBlockDataDesc * CUPHY_LDPC_DECODE_DESC_MAX_TB),
                ("llr_output", TransportBlockLLRDesc * CUPHY_LDPC_DECODE_DESC_MAX_TB)]


class Dim3(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32)]


class HipKernelNodeParams(C.Structure):
    _fields_ = [("blockDim", Dim3), ("extra", C.c_void_p), ("func", C.c_void_p), ("gridDim", Dim3),
                ("kernelParams", C.c_void_p), ("sharedMemBytes", C.c_uint32)]


class LDPCDecodeLaunchConfig(C.Structure):
    _fields_ = [("kernel_node_params_driver", HipKernelNodeParams), ("kernel_args", C.c_void_p * 2),
                ("decode_desc", LDPCDecodeDesc)]


class CalcCsi2SizePrm(C.Structure):
    _fields_ = [("nPart1Prms", C.c_uint8), ("prmSizes", C.c_uint8 * CUPHY_MAX_N_CSI1_PRMS),
                ("prmOffsets", C.c_uint16 * CUPHY_MAX_N_CSI1_PRMS), ("csi2sizeMapIdx", C.c_uint16)]


class PerTbParams(C.Structure):
    _fields_ = [("ndi", C.c_uint32), ("rv", C.c_uint32), ("Qm", C.c_uint32), ("bg", C.c_uint32),
                ("Nl", C.c_uint32), ("num_CBs", C.c_uint32), ("Zc", C.c_uint32), ("N", C.c_uint32),
                ("Ncb", C.c_uint32), ("Ncb_padded", C.c_uint32), ("G", C.c_uint32), ("K", C.c_uint32),
                ("F", C.c_uint32), ("cinit", C.c_uint32), ("nDataBytes", C.c_uint32),
                ("nZpBitsPerCb", C.c_uint32), ("k0", C.c_uint32), ("ldpcEarlyTerminationPerUe", C.c_uint8),
                ("ldpcMaxNumItrPerUe", C.c_uint8), ("firstCodeBlockIndex", C.c_uint32),
                ("encodedSize", C.c_uint32), ("layer_map_array", C.c_uint32 * MAX_N_LAYERS_PUSCH),
                ("userGroupIndex", C.c_uint32), ("nBBULayers", C.c_uint32), ("startLLR", C.c_uint32),
                ("mScUciSum", C.c_uint32), ("codedBitsSum", C.c_uint32), ("isDataPresent", C.c_uint8),
                ("betaOffsetCsi2", C.c_uint8), ("qPrimeAck", C.c_uint32), ("codeRate", C.c_float),
                ("alpha", C.c_float), ("qPrimeCsi1", C.c_uint32), ("isEarlyHarq", C.c_uint8),
                ("uciOnPuschFlag", C.c_uint8), ("csi2Flag", C.c_uint8), ("G_schAndCsi2", C.c_uint32),
                ("G_harq", C.c_uint32), ("G_csi1", C.c_uint32), ("G_csi2", C.c_uint32),
                ("G_harq_rvd", C.c_uint32), ("nBitsHarq", C.c_uint32), ("nBitsCsi2", C.c_uint16),
                ("nCsiReports", C.c_uint8), ("rankBitOffset", C.c_uint8), ("nRanksBits", C.c_uint8),
                ("d_schAndCsi2LLRs", C.c_void_p), ("d_csi1LLRs", C.c_void_p), ("d_harqLLrs", C.c_void_p),
                ("tbSize", C.c_uint32), ("nCsi2Reports", C.c_uint16),
                ("calcCsi2SizePrms", CalcCsi2SizePrm * CUPHY_MAX_N_CSI2_REPORTS_PER_UE),
                ("enableTfPrcd", C.c_uint8), ("nDmrsCdmGrpsNoData", C.c_uint8),
                ("debug_d_derateCbsIndices", C.c_void_p)]


class LDPCWorkItem(C.Structure):
    _fields_ = [("llr", C.c_void_p), ("out", C.c_void_p), ("soft", C.c_void_p), ("BG", C.c_int16),
                ("Z", C.c_int16), ("num_parity_nodes", C.c_int16), ("Kb", C.c_int16),
                ("max_iterations", C.c_int16), ("flags", C.c_uint16), ("norm", Norm),
                ("clamp_value", C.c_float), ("status", C.c_uint32)]


class PuschSlotTb(C.Structure):
    _fields_ = [("prm", C.c_void_p), ("llr", C.c_void_p), ("harq", C.c_void_p), ("cb_bits", C.c_void_p),
                ("cb_crc", C.c_void_p), ("tb_bytes", C.c_void_p), ("tb_crc", C.c_void_p),
                ("cb_iterations", C.c_void_p), ("num_parity_nodes", C.c_int16), ("max_iterations", C.c_int16),
                ("flags", C.c_uint16), ("reserved", C.c_uint16), ("llr_scale", C.c_float), ("norm", Norm),
                ("clamp_value", C.c_float), ("share_offset", C.c_uint32)]


class HarqServerConfig(C.Structure):
    _fields_ = [("harq_pool_bytes", C.c_size_t), ("max_inflight", C.c_int), ("max_tbs_per_launch", C.c_int),
                ("max_cbs_per_launch", C.c_int), ("llr_type", C.c_int), ("flags", C.c_uint32),
                ("harq_timeout_us", C.c_double), ("resident_idle_us", C.c_double)]


class HarqRequest(C.Structure):
    _fields_ = [("cell_id", C.c_uint32), ("rnti", C.c_uint16), ("harq_pid", C.c_uint8), ("reserved0", C.c_uint8),
                ("tag", C.c_uint64), ("prm", PerTbParams), ("llr", C.c_void_p),
                ("cb_bits", C.c_void_p), ("cb_crc", C.c_void_p), ("tb_bytes", C.c_void_p),
                ("cb_iterations", C.c_void_p), ("num_parity_nodes", C.c_int16), ("max_iterations", C.c_int16),
                ("flags", C.c_uint16), ("reserved", C.c_uint16), ("llr_scale", C.c_float), ("norm", Norm),
                ("clamp_value", C.c_float)]


class HarqCompletion(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("cell_id", C.c_uint32), ("rnti", C.c_uint16), ("harq_pid", C.c_uint8),
                ("reserved0", C.c_uint8), ("tb_crc", C.c_uint32), ("status", C.c_int32),
                ("t_submit_us", C.c_double), ("t_launch_us", C.c_double), ("t_done_us", C.c_double)]


class HarqServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("launches", C.c_uint64),
                ("pool_bytes_in_use", C.c_uint64), ("pool_bytes_peak", C.c_uint64), ("pending", C.c_uint64),
                ("admission_waits", C.c_uint64), ("evictions", C.c_uint64), ("pool_bytes_cached", C.c_uint64),
                ("pool_reclaims", C.c_uint64)]


class HarqTraceEntry(C.Structure):
    _fields_ = [("req", HarqRequest), ("arrival_us", C.c_double), ("next", C.c_int32), ("reserved", C.c_int32)]


class LbrmPrm(C.Structure):
    _fields_ = [("i_lbrm", C.c_uint8), ("maxLayers", C.c_uint8), ("maxQm", C.c_uint8), ("reserved", C.c_uint8),
                ("n_PRB_LBRM", C.c_uint16), ("reserved2", C.c_uint16)]


class LDPCEncodeLaunchConfig(C.Structure):
    _fields_ = [("m_kernelNodeParams", HipKernelNodeParams), ("m_desc", C.c_void_p), ("m_kernelArgs", C.c_void_p * 1)]


class TensorPrm(C.Structure):
    _fields_ = [("desc", C.c_void_p), ("pAddr", C.c_void_p)]


class RateMatchLaunchCfg(C.Structure):
    _fields_ = [("kernelNodeParamsDriver", HipKernelNodeParams), ("resetKernelNodeParamsDriver", HipKernelNodeParams),
                ("clampKernelNodeParamsDriver", HipKernelNodeParams), ("desc", C.c_void_p),
                ("kernelArgs", C.c_void_p * 1)]


class CrcDecodeLaunchCfg(C.Structure):
    _fields_ = [("kernelNodeParamsDriver", HipKernelNodeParams), ("desc", C.c_void_p), ("kernelArgs", C.c_void_p * 1)]


# Every entry point declared in include/cuphy_hip.h, with its ctypes signature.
_P = C.POINTER
_VP = C.c_void_p
_I = C.c_int
SIGNATURES = {
    "cuphyGetErrorString": ([_I], C.c_char_p),
    "cuphyGetErrorName": ([_I], C.c_char_p),
    "cuphyCreateContext": ([_P(_VP), C.c_uint], _I),
    "cuphyDestroyContext": ([_VP], _I),
    "cuphyCreateTensorDescriptor": ([_P(_VP)], _I),
    "cuphyDestroyTensorDescriptor": ([_VP], _I),
    "cuphySetTensorDescriptor": ([_VP, _I, _I, _P(_I), _P(_I), C.c_uint], _I),
    "cuphyGetTensorDescriptor": ([_VP, _I, _P(_I), _P(_I), _P(_I), _P(_I)], _I),
    "cuphyGetTensorSizeInBytes": ([_VP, _P(C.c_size_t)], _I),
    "cuphyCreateLDPCDecoder": ([_VP, _P(_VP), C.c_uint], _I),
    "cuphyDestroyLDPCDecoder": ([_VP], _I),
    "cuphyErrorCorrectionLDPCDecode": ([_VP, _VP, _VP, _VP, _VP, _VP, _VP, _P(LDPCDecodeConfigDesc), _VP], _I),
    "cuphyErrorCorrectionLDPCTransportBlockDecode": ([_VP, _P(LDPCDecodeDesc), _VP], _I),
    "cuphyErrorCorrectionLDPCDecodeGetWorkspaceSize": ([_VP, _P(LDPCDecodeConfigDesc), _I, _P(C.c_size_t)], _I),
    "cuphyErrorCorrectionLDPCDecodeSetNormalization": ([_VP, _P(LDPCDecodeConfigDesc)], _I),
    "cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor": ([_VP, _P(LDPCDecodeLaunchConfig)], _I),
    "cuphyPuschRxRateMatchGetDescrInfo": ([_P(C.c_size_t), _P(C.c_size_t)], _I),
    "cuphyCreatePuschRxRateMatch": ([_P(_VP), _I, _I], _I),
    "cuphySetupPuschRxRateMatch": ([_VP, C.c_uint16, _P(C.c_uint16), _P(PerTbParams), _VP, _P(TensorPrm),
                                    _P(TensorPrm), _P(_VP), _VP, _VP, C.c_uint8, _VP, _VP], _I),
    "cuphyDestroyPuschRxRateMatch": ([_VP], _I),
    "cuphyPuschRxCrcDecodeGetDescrInfo": ([_P(C.c_size_t), _P(C.c_size_t)], _I),
    "cuphyCreatePuschRxCrcDecode": ([_P(_VP), _I], _I),
    "cuphySetupPuschRxCrcDecode": ([_VP, C.c_uint16, _P(C.c_uint16), _VP, _VP, _VP, _VP, _P(PerTbParams), _VP,
                                    _VP, _VP, C.c_uint8, _VP, _VP, _VP], _I),
    "cuphyDestroyPuschRxCrcDecode": ([_VP], _I),
    "cuphyHipLDPCEncodeCodeBlocks": ([_I, _I, _VP, _I, _I, _VP, _I, _VP], _I),
    "cuphyHipLDPCBatchDecodeGetWorkspaceSize": ([_P(C.c_size_t)], _I),
    "cuphyHipLDPCBatchDecode": ([_VP, _VP, _I, _I, _I, _I, _I, _VP, _VP], _I),
    "cuphyHipLDPCBatchDecodeGroups": ([_VP, _VP, _VP, _I, _I, _I, _I, _I, _I, _VP, _VP], _I),
    "cuphyHipPuschSlotDecodeGetWorkspaceSize": ([_I, _P(C.c_size_t)], _I),
    "cuphyHipPuschSlotDecode": ([_VP, _VP, _I, _VP, _I, _I, _I, _I, _I, _VP, _VP], _I),
    "cuphyHipPuschSlotDecodeNoAlloc": ([_VP, _VP, _I, _VP, _I, _I, _I, _I, _I, _I, _VP], _I),
    "cuphyHipHarqServerCreate": ([_VP, _P(HarqServerConfig), _P(_VP)], _I),
    "cuphyHipHarqServerDestroy": ([_VP], _I),
    "cuphyHipHarqServerSubmit": ([_VP, _P(HarqRequest)], _I),
    "cuphyHipHarqServerPoll": ([_VP, _VP, _I, _P(_I)], _I),
    "cuphyHipHarqServerRelease": ([_VP, C.c_uint32, C.c_uint16, C.c_uint8], _I),
    "cuphyHipHarqServerGetStats": ([_VP, _P(HarqServerStats)], _I),
    "cuphyHipHarqServeTrace": ([_VP, _VP, _I, C.c_double, _VP, _P(_I)], _I),
    "cuphyHipLDPCWorkItemCheck": ([_P(LDPCWorkItem), _I, _P(C.c_uint32)], _I),
    "cuphyLDPCEncodeGetDescrInfo": ([_P(C.c_size_t), _P(C.c_size_t), C.c_uint16, _P(C.c_size_t)], _I),
    "cuphySetupLDPCEncode": ([_P(LDPCEncodeLaunchConfig), _VP, _VP, _VP, _VP, _I, _I, C.c_uint8, _I, _I, C.c_uint8,
                              _I, _VP, _VP, _VP, _VP, _VP, _VP, C.c_uint8, _VP], _I),
    "cuphyHipDlRateMatch": ([_VP, _I, _VP, _I, _I, C.c_uint8, _VP, _VP], _I),
    "cuphyHipTransportBlockSize": ([_I, _I, _I, C.c_float, C.c_uint32, _I, _P(C.c_uint32), _P(C.c_uint32)], _I),
    "cuphyHipSetPerTbParams": ([_P(PerTbParams), C.c_uint32, C.c_float, C.c_uint8, C.c_uint32, C.c_uint32,
                                C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint8, C.c_uint8, _P(C.c_uint32), C.c_uint8,
                                _P(LbrmPrm), _P(C.c_uint16), _P(C.c_uint16)], _I),
}

_lib = None


class CuphyError(RuntimeError):
    def __init__(self, fn, status):
        self.status = status
        super().__init__("%s failed: %s" % (fn, status_name(status)))


def lib():
    """The loaded HIP library; raises ImportError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("HIP library %s is not built (run __graft_entry__.build() or "
                              "aerial-cuda-accelerated-ran_amd/build.py)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def status_name(status):
    return lib().cuphyGetErrorName(status).decode()


def check(fn_name, status):
    if status != CUPHY_STATUS_SUCCESS:
        raise CuphyError(fn_name, status)
    return status
