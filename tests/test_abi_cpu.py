"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/cuphy_hip.h, and rejects bad arguments the way the
reference's cuphy_ldpc.cpp does (no compute call is made)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_entry_points():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        names += re.findall(r"CUPHYWINAPI\s+(\w+)\s*\(", text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def L():
    from aerial_amd import _lib
    _lib.lib()
    return _lib


def test_header_declares_the_reference_surface():
    names = header_entry_points()
    for required in ("cuphyCreateLDPCDecoder", "cuphyDestroyLDPCDecoder", "cuphyErrorCorrectionLDPCDecode",
                     "cuphyErrorCorrectionLDPCTransportBlockDecode", "cuphyErrorCorrectionLDPCDecodeGetWorkspaceSize",
                     "cuphyErrorCorrectionLDPCDecodeSetNormalization",
                     "cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor", "cuphyCreatePuschRxRateMatch",
                     "cuphySetupPuschRxRateMatch", "cuphyDestroyPuschRxRateMatch", "cuphyPuschRxRateMatchGetDescrInfo",
                     "cuphyCreatePuschRxCrcDecode", "cuphySetupPuschRxCrcDecode", "cuphyDestroyPuschRxCrcDecode",
                     "cuphyPuschRxCrcDecodeGetDescrInfo"):
        assert required in names


def test_library_exports_every_declared_symbol(L):
    lib = L.lib()
    missing = [n for n in header_entry_points() if not hasattr(lib, n)]
    assert not missing, missing


def test_api_phases_carry_roctx_ranges(L):
    """The library calls roctx (range push/pop around the API's phases, the
    counterpart of the reference's NVTX PUSH_RANGE/POP_RANGE) and links the
    roctx library, so that rocprofv3 --marker-trace shows the ranges."""
    import subprocess
    path = L.lib()._name
    undef = subprocess.run(["nm", "-D", "--undefined-only", path], capture_output=True, text=True, check=True).stdout
    assert "roctxRangePushA" in undef and "roctxRangePop" in undef
    needed = subprocess.run(["readelf", "-d", path], capture_output=True, text=True, check=True).stdout
    assert "librocprofiler-sdk-roctx" in needed


def test_every_declared_symbol_has_a_python_signature():
    from aerial_amd import _lib
    missing = [n for n in header_entry_points() if n not in _lib.SIGNATURES]
    assert not missing, missing


def test_struct_sizes_match_header(L):
    # sizes of the C structs (x86-64): checked against the header layout
    assert C.sizeof(L.LDPCDecodeConfigDesc) == 40
    assert C.sizeof(L.LDPCDecodeDesc) == 40 + 8 + 3 * 32 * 16
    assert C.sizeof(L.PerTbParams) == 488


def test_error_strings(L):
    lib = L.lib()
    assert lib.cuphyGetErrorName(0) == b"CUPHY_STATUS_SUCCESS"
    assert lib.cuphyGetErrorName(12) == b"CUPHY_STATUS_UNSUPPORTED_CONFIG"
    assert lib.cuphyGetErrorName(99) == b"CUPHY_UNKNOWN_STATUS"
    assert lib.cuphyGetErrorString(3).startswith(b"An invalid argument")


def test_null_arguments_rejected(L):
    lib = L.lib()
    INV = L.CUPHY_STATUS_INVALID_ARGUMENT
    assert lib.cuphyCreateContext(None, 0) == INV
    assert lib.cuphyDestroyContext(None) == INV
    assert lib.cuphyCreateLDPCDecoder(None, None, 0) == INV
    assert lib.cuphyDestroyLDPCDecoder(None) == INV
    assert lib.cuphyErrorCorrectionLDPCTransportBlockDecode(None, None, None) == INV
    assert lib.cuphyErrorCorrectionLDPCDecodeSetNormalization(None, None) == INV
    assert lib.cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor(None, None) == INV
    cfg = L.LDPCDecodeConfigDesc()
    sz = C.c_size_t()
    assert lib.cuphyErrorCorrectionLDPCDecodeGetWorkspaceSize(None, C.byref(cfg), 1, C.byref(sz)) == INV
    assert lib.cuphyErrorCorrectionLDPCDecode(None, None, None, None, None, None, None, None, None) == INV


def test_tensor_descriptor_semantics(L):
    lib = L.lib()
    t = C.c_void_p()
    assert lib.cuphyCreateTensorDescriptor(C.byref(t)) == 0
    dims = (C.c_int * 2)(25344 + 768, 8)
    assert lib.cuphySetTensorDescriptor(t, L.CUPHY_R_32F, 2, dims, None, 0) == 0
    dt, nd = C.c_int(), C.c_int()
    od, os_ = (C.c_int * 2)(), (C.c_int * 2)()
    assert lib.cuphyGetTensorDescriptor(t, 2, C.byref(dt), C.byref(nd), od, os_) == 0
    assert (dt.value, nd.value, list(od), list(os_)) == (L.CUPHY_R_32F, 2, [26112, 8], [1, 26112])
    sz = C.c_size_t()
    assert lib.cuphyGetTensorSizeInBytes(t, C.byref(sz)) == 0 and sz.value == 26112 * 8 * 4
    # bit tensor: first dimension padded to whole 32-bit words
    bdims = (C.c_int * 2)(8448 + 5, 3)
    assert lib.cuphySetTensorDescriptor(t, L.CUPHY_BIT, 2, bdims, None, 0) == 0
    assert lib.cuphyGetTensorDescriptor(t, 2, None, None, od, os_) == 0
    assert list(os_) == [1, 8480]
    # overlapping strides and bad bit strides are rejected
    bad = (C.c_int * 2)(1, 10)
    assert lib.cuphySetTensorDescriptor(t, L.CUPHY_R_16F, 2, (C.c_int * 2)(20, 4), bad, 0) == 3
    assert lib.cuphySetTensorDescriptor(t, L.CUPHY_BIT, 2, (C.c_int * 2)(40, 4), (C.c_int * 2)(1, 40), 0) == 3
    assert lib.cuphySetTensorDescriptor(t, L.CUPHY_R_32F, 0, dims, None, 0) == 3
    assert lib.cuphyDestroyTensorDescriptor(t) == 0


def test_descriptor_info_queries(L):
    lib = L.lib()
    a, b = C.c_size_t(), C.c_size_t()
    if hasattr(lib, "cuphyPuschRxRateMatchGetDescrInfo"):
        assert lib.cuphyPuschRxRateMatchGetDescrInfo(C.byref(a), C.byref(b)) == 0
        assert a.value > 0 and b.value > 0
        assert lib.cuphyPuschRxRateMatchGetDescrInfo(None, C.byref(b)) == 3
    if hasattr(lib, "cuphyPuschRxCrcDecodeGetDescrInfo"):
        assert lib.cuphyPuschRxCrcDecodeGetDescrInfo(C.byref(a), C.byref(b)) == 0
        assert a.value > 0 and b.value > 0


# ctypes mirror -> C type name in include/cuphy_hip.h
_MIRRORS = [("LDPCDecodeConfigDesc", "cuphyLDPCDecodeConfigDesc_t"),
            ("TransportBlockLLRDesc", "cuphyTransportBlockLLRDesc_t"),
            ("TransportBlockDataDesc", "cuphyTransportBlockDataDesc_t"),
            ("LDPCDecodeDesc", "cuphyLDPCDecodeDesc_t"),
            ("LDPCDecodeLaunchConfig", "cuphyLDPCDecodeLaunchConfig_t"),
            ("PerTbParams", "PerTbParams"),
            ("LDPCWorkItem", "cuphyHipLDPCWorkItem_t"),
            ("PuschSlotTb", "cuphyHipPuschSlotTb_t"),
            ("TensorPrm", "cuphyTensorPrm_t"),
            ("RateMatchLaunchCfg", "cuphyPuschRxRateMatchLaunchCfg_t"),
            ("CrcDecodeLaunchCfg", "cuphyPuschRxCrcDecodeLaunchCfg_t"),
            ("HarqServerConfig", "cuphyHipHarqServerConfig_t"),
            ("HarqRequest", "cuphyHipHarqRequest_t"),
            ("HarqCompletion", "cuphyHipHarqCompletion_t"),
            ("HarqServerStats", "cuphyHipHarqServerStats_t"),
            ("HarqTraceEntry", "cuphyHipHarqTraceEntry_t")]


def test_ctypes_layouts_match_the_c_header(tmp_path):
    """Size and every field offset of each ctypes mirror equal the C
    compiler's for the header struct (gcc, x86-64)."""
    import shutil
    import subprocess
    from aerial_amd import _lib
    if shutil.which("gcc") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("gcc or ROCm headers not available")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "cuphy_hip.h"', "int main(void) {"]
    expect = []
    for py, cname in _MIRRORS:
        cls = getattr(_lib, py)
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        expect.append(("%s size" % py, C.sizeof(cls)))
        for f in cls._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, f[0]))
            expect.append(("%s.%s" % (py, f[0]), getattr(cls, f[0]).offset))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
                    "-D__HIP_PLATFORM_AMD__", str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    bad = [(name, want, g) for (name, want), g in zip(expect, got) if want != g]
    assert len(got) == len(expect) and not bad, bad


def test_harq_server_argument_validation(L):
    """Creation arguments are checked before any device call."""
    lib = L.lib()
    INV = L.CUPHY_STATUS_INVALID_ARGUMENT
    cfg = L.HarqServerConfig(harq_pool_bytes=1 << 20, max_inflight=2, max_tbs_per_launch=16, max_cbs_per_launch=64,
                             llr_type=L.CUPHY_R_16F)
    h = C.c_void_p()
    assert lib.cuphyHipHarqServerCreate(None, C.byref(cfg), C.byref(h)) == INV
    assert lib.cuphyHipHarqServerDestroy(None) == INV
    assert lib.cuphyHipHarqServerSubmit(None, None) == INV
    n = C.c_int()
    assert lib.cuphyHipHarqServerPoll(None, None, 0, C.byref(n)) == INV
    assert lib.cuphyHipHarqServerRelease(None, 0, 1, 2) == INV
    assert lib.cuphyHipHarqServeTrace(None, None, 0, 0.0, None, C.byref(n)) == INV


def test_set_normalization_is_the_reference_table(L):
    """SetNormalization gives, for every (BG, parity nodes), the reference's
    g_min_sum_norm_BG{1,2}_Z384 entry (ldpc.cpp:139-235 = LDPC_decode.m:101-208,
    fixture extracted from both by tests/golden/extract_norm_table.py):
    binary32 as is, binary16 rounded to nearest even into both halves."""
    import json
    import numpy as np
    table = json.load(open(os.path.join(ROOT, "tests", "golden", "min_sum_norm.json")))
    lib = L.lib()
    # the decoder object reads only the context's device fields; SetNormalization is host only,
    # so a zero-filled context stands in for a device context here
    ctx = (C.c_uint8 * 64)()
    dec = C.c_void_p()
    assert lib.cuphyCreateLDPCDecoder(C.cast(ctx, C.c_void_p), C.byref(dec), 0) == L.CUPHY_STATUS_SUCCESS
    try:
        for bg in (1, 2):
            vals = table["BG%d" % bg]
            for mb in range(0, len(vals) + 2):
                for llr_type in (L.CUPHY_R_32F, L.CUPHY_R_16F):
                    cfg = L.LDPCDecodeConfigDesc()
                    cfg.llr_type, cfg.BG, cfg.num_parity_nodes, cfg.Z, cfg.Kb = llr_type, bg, mb, 384, 22 if bg == 1 else 10
                    st = lib.cuphyErrorCorrectionLDPCDecodeSetNormalization(dec, C.byref(cfg))
                    if mb < 4 or mb >= len(vals):
                        assert st == L.CUPHY_STATUS_INVALID_ARGUMENT, (bg, mb)
                        continue
                    assert st == L.CUPHY_STATUS_SUCCESS
                    if llr_type == L.CUPHY_R_32F:
                        assert cfg.norm.f32 == np.float32(vals[mb]), (bg, mb)
                    else:
                        h = int(np.array([vals[mb]], np.float32).astype(np.float16).view(np.uint16)[0])
                        assert cfg.norm.f16x2 == (h | (h << 16)), (bg, mb)
    finally:
        lib.cuphyDestroyLDPCDecoder(dec)


def test_work_item_check(L):
    """The work-item predicate the persistent kernels apply, on the host:
    every valid (BG, lifting Z, mb, Kb) passes; each kind of bad field is
    named, alone and combined."""
    lib = L.lib()
    lifting = sorted({a << j for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8) if (a << j) <= 384})

    def check(max_mb=46, **kw):
        w = L.LDPCWorkItem(llr=16, out=32, soft=None, BG=1, Z=384, num_parity_nodes=46, Kb=22, max_iterations=10,
                           flags=0, clamp_value=32.0, status=0)
        for k, v in kw.items():
            setattr(w, k, v)
        r = C.c_uint32(0xDEAD)
        assert lib.cuphyHipLDPCWorkItemCheck(C.byref(w), max_mb, C.byref(r)) == L.CUPHY_STATUS_SUCCESS
        return r.value

    for bg in (1, 2):
        for Z in lifting:
            for mb in range(4, (46 if bg == 1 else 42) + 1):
                for kb in ((22,) if bg == 1 else (6, 8, 9, 10)):
                    assert check(BG=bg, Z=Z, num_parity_nodes=mb, Kb=kb) == 0, (bg, Z, mb, kb)
    for Z in [z for z in range(-2, 400) if z not in lifting] + [1000, -32768]:
        assert check(Z=Z) == L.ITEM_BAD_Z, Z
    for bg in (0, 3, -1, 7):
        assert check(BG=bg) & L.ITEM_BAD_BG, bg
    for mb in (0, 3, 47, -5):
        assert check(num_parity_nodes=mb) == L.ITEM_BAD_MB, mb
    assert check(BG=2, Kb=10, num_parity_nodes=43) == L.ITEM_BAD_MB
    assert check(num_parity_nodes=9, max_mb=8) == L.ITEM_BAD_MB      # beyond the high-code-rate kernel's rows
    assert check(num_parity_nodes=8, max_mb=8) == 0
    assert check(num_parity_nodes=46, max_mb=9) == 0                 # max_mb > 8 runs the 46-row kernel
    for bg, kb in ((1, 10), (1, 21), (2, 22), (2, 7), (2, 5), (2, 11)):
        assert check(BG=bg, Kb=kb, num_parity_nodes=10) == L.ITEM_BAD_KB, (bg, kb)
    assert check(max_iterations=-1) == L.ITEM_BAD_ITER
    assert check(max_iterations=0) == 0
    assert check(llr=None) == L.ITEM_NULL_BUFFER and check(out=None) == L.ITEM_NULL_BUFFER
    assert check(BG=3, Z=5 * 7, num_parity_nodes=2, max_iterations=-2, out=None) == \
        L.ITEM_BAD_BG | L.ITEM_BAD_Z | L.ITEM_BAD_MB | L.ITEM_BAD_KB | L.ITEM_BAD_ITER | L.ITEM_NULL_BUFFER
    r = C.c_uint32()
    assert lib.cuphyHipLDPCWorkItemCheck(None, 46, C.byref(r)) == L.CUPHY_STATUS_INVALID_ARGUMENT


def test_encoder_setup_validation(L):
    """cuphyLDPCEncodeGetDescrInfo sizes and cuphySetupLDPCEncode's argument
    checks (ldpc_encode.cu:416-470); setup without the descriptor copy makes
    no device call."""
    lib = L.lib()
    dsz, dal, wsz = C.c_size_t(), C.c_size_t(), C.c_size_t()
    assert lib.cuphyLDPCEncodeGetDescrInfo(C.byref(dsz), C.byref(dal), 10, C.byref(wsz)) == 0
    assert dsz.value > 0 and dal.value in (4, 8, 16) and wsz.value == 10 * 2 * 16
    assert lib.cuphyLDPCEncodeGetDescrInfo(None, C.byref(dal), 10, C.byref(wsz)) == L.CUPHY_STATUS_INVALID_ARGUMENT

    def tensor(t, nbits, ncw):
        h = C.c_void_p()
        assert lib.cuphyCreateTensorDescriptor(C.byref(h)) == 0
        dims = (C.c_int * 2)(nbits, ncw)
        strides = (C.c_int * 2)(1, (nbits + 31) // 32 * 32)
        assert lib.cuphySetTensorDescriptor(h, t, 2, dims, strides, 0) == 0
        return h

    tin, tout = tensor(L.CUPHY_BIT, 22 * 384, 4), tensor(L.CUPHY_BIT, 66 * 384, 4)
    tf = tensor(L.CUPHY_R_32F, 22 * 384, 4)
    cfg = L.LDPCEncodeLaunchConfig()
    cpu = (C.c_uint8 * dsz.value)()
    hws = (C.c_uint8 * wsz.value)()
    dws = C.c_void_p(0x1000)  # not dereferenced without the copy
    buf = C.c_void_p(0x2000)

    def setup(i=tin, o=tout, bg=1, Z=384, punct=1, rows=0, rv=0, batching=0, n=1, h=hws, d=dws):
        return lib.cuphySetupLDPCEncode(C.byref(cfg), i, buf, o, buf, bg, Z, punct, rows, rv, batching, n, None, None,
                                        C.cast(h, C.c_void_p) if h is not None else None, d,
                                        C.cast(cpu, C.c_void_p), None, 0, None)

    assert setup() == L.CUPHY_STATUS_SUCCESS
    g = cfg.m_kernelNodeParams
    assert (g.gridDim.x, g.gridDim.y, g.blockDim.x) == (4, 1, 256) and cfg.m_kernelArgs[0] is not None
    assert setup(rv=4) == L.CUPHY_STATUS_INVALID_ARGUMENT and setup(rv=-1) == L.CUPHY_STATUS_INVALID_ARGUMENT
    assert setup(batching=1, n=2) == L.CUPHY_STATUS_INVALID_ARGUMENT        # no batched address arrays
    assert setup(h=None) == L.CUPHY_STATUS_INVALID_ARGUMENT
    assert setup(i=tf) == L.CUPHY_STATUS_UNSUPPORTED_TYPE
    assert setup(bg=3) == L.CUPHY_STATUS_UNSUPPORTED_CONFIG and setup(Z=17) == L.CUPHY_STATUS_UNSUPPORTED_CONFIG
    assert setup(punct=0, rows=3) == L.CUPHY_STATUS_INVALID_ARGUMENT        # fewer than 4 rows in LDS
    assert setup(punct=1, rows=2) == L.CUPHY_STATUS_SUCCESS
    assert setup(punct=0, rows=0) == L.CUPHY_STATUS_INVALID_ARGUMENT       # 68 Z bits do not fit the output
    for h in (tin, tout, tf):
        lib.cuphyDestroyTensorDescriptor(h)
