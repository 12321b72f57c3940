"""GPU parity of the HIP layered decoder against the oracle (bit-exact).

Every case prints its name, shape and seed before touching the GPU.
The binary32 path must equal the oracle's NR_ORC_FP32 mode bit for bit
(hard decisions and the binary16-rounded APP soft outputs); the binary16
paths must equal the oracle's NR_ORC_FP16 mode bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import vectors as V

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.fixture(scope="module")
def dec():
    _need_gpu()
    from aerial_amd.decoder import LdpcDecoder
    d = LdpcDecoder()
    yield d
    d.close()


def _log(msg):
    print("[case]", msg, flush=True)


def run_gpu(dec, bg, Z, mb, llr, n_iter=10, alpha=0.8125, dtype="fp32", flags=0, algo=0, clamp=32.0,
            soft=True, Kb=None):
    from aerial_amd import _lib as L
    from aerial_amd.decoder import LdpcDecodeConfig
    llr_type = L.CUPHY_R_32F if dtype in ("fp32", "fp32in_fp16math") else L.CUPHY_R_16F
    if dtype == "fp32in_fp16math":
        algo = 2
    cfg = LdpcDecodeConfig(bg, Z, mb, Kb=Kb, llr_type=llr_type, max_iterations=n_iter, alpha=alpha,
                           clamp_value=clamp, flags=flags, algo=algo)
    ncw = llr.shape[0]
    stride = (llr.shape[1] + 7) // 8 * 8
    host = np.zeros((ncw, stride), np.float32)
    host[:, :llr.shape[1]] = llr
    if llr_type == L.CUPHY_R_16F:
        x = torch.from_numpy(V.to_half_rn(host)).cuda()
    else:
        x = torch.from_numpy(host).cuda()
    nw = cfg.out_words()
    out = torch.full((ncw, nw + 3), -1, dtype=torch.int32, device="cuda")
    kbz = cfg.desc.Kb * Z
    sft = torch.zeros((ncw, (kbz + 1) // 2 * 2 + 2), dtype=torch.float16, device="cuda") if soft else None
    dec.decode_tb(cfg, [x], [out], [sft] if soft else None)
    torch.cuda.synchronize()
    o = out.cpu().numpy().view(np.uint32)
    assert (o[:, nw:] == 0xFFFFFFFF).all(), "decoder wrote past its output words"
    return o[:, :nw], (sft.cpu().numpy()[:, :kbz] if soft else None)


def run_oracle(bg, Z, mb, llr, n_iter=10, alpha=0.8125, mode=O.FP32, clamp=32.0, et=False, Kb=None):
    kb = Kb if Kb is not None else V.nsys(bg)
    cfg = O.dec_cfg(bg, Z, mb, n_iter=n_iter, alpha=alpha, clamp=clamp, mode=mode, kb_out=kb, early_term=et)
    hards, softs, iters = [], [], []
    for i in range(llr.shape[0]):
        h, it, app = O.decode_cb(cfg, llr[i], want_app=True)
        hards.append(h)
        softs.append(app[:kb * Z].astype(np.float16))
        iters.append(it)
    return np.stack(hards), np.stack(softs), iters


def assert_soft_equal(a, b):
    assert np.array_equal(a.view(np.uint16), b.view(np.uint16)), \
        "soft outputs differ at %d positions" % int((a.view(np.uint16) != b.view(np.uint16)).sum())


def test_config1_fp32_bit_exact(dec):
    _log("config1 BG1 Z=384 mb=46 8 CBs fp32 seed=1")
    llr, info = V.config1_vectors()
    h_gpu, s_gpu = run_gpu(dec, 1, 384, 46, llr)
    h_ref, s_ref, _ = run_oracle(1, 384, 46, llr)
    assert np.array_equal(h_gpu, h_ref)
    assert_soft_equal(s_gpu, s_ref)
    # and the decode is correct at this SNR
    bits = np.stack([O.unpack_bits(h, 22 * 384) for h in h_gpu])
    assert (bits == info).all()


def test_config1_fp16_bit_exact(dec):
    _log("config1 BG1 Z=384 mb=46 8 CBs fp16 seed=1")
    llr, _ = V.config1_vectors()
    llr16 = V.to_half_rn(llr).astype(np.float32)
    h_gpu, s_gpu = run_gpu(dec, 1, 384, 46, llr16, dtype="fp16")
    h_ref, s_ref, _ = run_oracle(1, 384, 46, llr16, mode=O.FP16)
    assert np.array_equal(h_gpu, h_ref)
    assert_soft_equal(s_gpu, s_ref)


def test_fp32_input_fp16_math_bit_exact(dec):
    _log("BG1 Z=384 mb=8 4 CBs fp32 in / fp16 math seed=3")
    llr, _ = V.make_cbs(1, 384, 8, 4, seed=3, snr_db=4.0)
    h_gpu, s_gpu = run_gpu(dec, 1, 384, 8, llr, dtype="fp32in_fp16math")
    h_ref, s_ref, _ = run_oracle(1, 384, 8, llr, mode=O.FP16)
    assert np.array_equal(h_gpu, h_ref)
    assert_soft_equal(s_gpu, s_ref)


SWEEP = [
    (1, 2, 4), (1, 3, 46), (1, 15, 10), (1, 36, 20), (1, 64, 4), (1, 104, 33), (1, 208, 46), (1, 352, 5),
    (2, 2, 42), (2, 7, 4), (2, 30, 12), (2, 64, 42), (2, 120, 8), (2, 240, 25), (2, 384, 4), (2, 384, 42),
    (1, 288, 46), (2, 288, 9), (1, 320, 13), (2, 320, 30), (1, 384, 7), (1, 384, 30), (2, 352, 42),
]


@pytest.mark.parametrize("bg,Z,mb", SWEEP)
@pytest.mark.parametrize("dtype", ["fp32", "fp16"])
def test_sweep_bit_exact(dec, bg, Z, mb, dtype):
    _log("sweep BG%d Z=%d mb=%d %s seed=7" % (bg, Z, mb, dtype))
    llr, _ = V.make_cbs(bg, Z, mb, 3, seed=7, snr_db=1.0)
    if dtype == "fp16":
        llr = V.to_half_rn(llr).astype(np.float32)
    h_gpu, s_gpu = run_gpu(dec, bg, Z, mb, llr, n_iter=6, alpha=0.75, dtype=dtype)
    h_ref, s_ref, _ = run_oracle(bg, Z, mb, llr, n_iter=6, alpha=0.75, mode=O.FP32 if dtype == "fp32" else O.FP16)
    assert np.array_equal(h_gpu, h_ref)
    assert_soft_equal(s_gpu, s_ref)


@pytest.mark.parametrize("bg,Z,mb,ncb", [(1, 2, 46, 70), (2, 3, 42, 45), (1, 16, 10, 9), (2, 30, 6, 5)])
def test_small_z_packed_tb(dec, bg, Z, mb, ncb):
    # 64 // Z code blocks share one wave (tb_group_size): full and partial groups
    _log("packed small-Z TB BG%d Z=%d mb=%d %d CBs fp16 seed=21" % (bg, Z, mb, ncb))
    llr, _ = V.make_cbs(bg, Z, mb, ncb, seed=21, snr_db=1.5)
    llr = V.to_half_rn(llr).astype(np.float32)
    h_gpu, s_gpu = run_gpu(dec, bg, Z, mb, llr, n_iter=8, dtype="fp16")
    h_ref, s_ref, _ = run_oracle(bg, Z, mb, llr, n_iter=8, mode=O.FP16)
    assert np.array_equal(h_gpu, h_ref)
    assert_soft_equal(s_gpu, s_ref)


def test_small_z_packed_tb_early_termination(dec):
    _log("packed small-Z TB BG2 Z=8 mb=20 40 CBs fp32 ET seed=23")
    from aerial_amd import _lib as L
    llr, _ = V.make_cbs(2, 8, 20, 40, seed=23, snr_db=2.0)
    h_gpu, s_gpu = run_gpu(dec, 2, 8, 20, llr, n_iter=25, flags=L.CUPHY_LDPC_DECODE_EARLY_TERM)
    h_ref, s_ref, iters = run_oracle(2, 8, 20, llr, n_iter=25, et=True)
    assert len(set(iters)) > 1, iters
    assert np.array_equal(h_gpu, h_ref)
    assert_soft_equal(s_gpu, s_ref)


@pytest.mark.parametrize("Kb", [6, 8, 9, 10])
def test_bg2_kb_output_width(dec, Kb):
    _log("BG2 Z=64 mb=10 Kb=%d" % Kb)
    llr, _ = V.make_cbs(2, 64, 10, 2, seed=11, snr_db=3.0, filler=(10 - Kb) * 64)
    h_gpu, s_gpu = run_gpu(dec, 2, 64, 10, llr, Kb=Kb)
    h_ref, s_ref, _ = run_oracle(2, 64, 10, llr, Kb=Kb)
    assert np.array_equal(h_gpu, h_ref)
    assert_soft_equal(s_gpu, s_ref)


def test_zero_iterations(dec):
    # pyaerial/tests/test_ldpc_decoder.py:142-190: 0 iterations -> soft = clip(LLR, +-32)
    _log("BG1 Z=128 mb=6 0 iterations")
    llr, _ = V.make_cbs(1, 128, 6, 2, seed=5, snr_db=-3.0)
    llr[:, 2 * 128:2 * 128 + 50] *= 100.0
    h_gpu, s_gpu = run_gpu(dec, 1, 128, 6, llr, n_iter=0)
    expect_soft = np.clip(llr[:, :22 * 128], -32, 32).astype(np.float16)
    assert_soft_equal(s_gpu, expect_soft + np.float16(0))
    expect_bits = (llr[:, :22 * 128] <= 0).astype(np.uint8)
    got = np.stack([O.unpack_bits(h, 22 * 128) for h in h_gpu])
    assert (got == expect_bits).all()


def test_early_termination_matches_oracle(dec):
    from aerial_amd import _lib as L
    _log("BG1 Z=256 mb=12 ET, 5 CBs high SNR")
    llr, info = V.make_cbs(1, 256, 12, 5, seed=9, snr_db=6.0)
    h_gpu, s_gpu = run_gpu(dec, 1, 256, 12, llr, n_iter=20, flags=L.CUPHY_LDPC_DECODE_EARLY_TERM)
    h_ref, s_ref, iters = run_oracle(1, 256, 12, llr, n_iter=20, et=True)
    assert max(iters) < 20, iters  # ET actually stopped early
    assert np.array_equal(h_gpu, h_ref)
    assert_soft_equal(s_gpu, s_ref)


def test_early_termination_fp16_large_z(dec):
    from aerial_amd import _lib as L
    _log("BG1 Z=352 mb=20 ET fp16, 4 CBs")
    llr, info = V.make_cbs(1, 352, 20, 4, seed=19, snr_db=3.0)
    llr = V.to_half_rn(llr).astype(np.float32)
    h_gpu, s_gpu = run_gpu(dec, 1, 352, 20, llr, n_iter=20, flags=L.CUPHY_LDPC_DECODE_EARLY_TERM, dtype="fp16")
    h_ref, s_ref, iters = run_oracle(1, 352, 20, llr, n_iter=20, et=True, mode=O.FP16)
    assert max(iters) < 20, iters
    assert np.array_equal(h_gpu, h_ref)
    assert_soft_equal(s_gpu, s_ref)


def test_multi_tb_descriptor(dec):
    from aerial_amd import _lib as L
    from aerial_amd.decoder import LdpcDecodeConfig
    _log("3 TBs BG1 Z=96 mb=9 ncw=(1,4,2) with distinct strides")
    cfg = LdpcDecodeConfig(1, 96, 9, llr_type=L.CUPHY_R_32F, max_iterations=8, alpha=0.7)
    n = cfg.num_llrs()
    xs, outs, refs = [], [], []
    for t, ncw in enumerate((1, 4, 2)):
        llr, _ = V.make_cbs(1, 96, 9, ncw, seed=20 + t, snr_db=2.0)
        stride = n + 8 * (t + 1)
        host = np.zeros((ncw, stride), np.float32)
        host[:, :n] = llr
        xs.append(torch.from_numpy(host).cuda())
        outs.append(torch.full((ncw, cfg.out_words() + 5 * t), -1, dtype=torch.int32, device="cuda"))
        refs.append(run_oracle(1, 96, 9, llr, n_iter=8, alpha=0.7)[0])
    dec.decode_tb(cfg, xs, outs)
    torch.cuda.synchronize()
    for t in range(3):
        o = outs[t].cpu().numpy().view(np.uint32)
        assert np.array_equal(o[:, :cfg.out_words()], refs[t])
        assert (o[:, cfg.out_words():] == 0xFFFFFFFF).all()


def test_tensor_interface_matches_tb_interface(dec):
    from aerial_amd import _lib as L
    _log("tensor API BG2 Z=160 mb=7 3 CBs")
    lib = L.lib()
    bg, Z, mb = 2, 160, 7
    llr, _ = V.make_cbs(bg, Z, mb, 3, seed=31, snr_db=2.5)
    h_ref, _ = run_gpu(dec, bg, Z, mb, llr, soft=False)
    nllr = llr.shape[1]
    x = torch.from_numpy(np.ascontiguousarray(llr)).cuda()
    kbits = 10 * Z
    wstride = (kbits + 31) // 32
    out = torch.zeros((3, wstride), dtype=torch.int32, device="cuda")
    tl, to = C.c_void_p(), C.c_void_p()
    lib.cuphyCreateTensorDescriptor(C.byref(tl))
    lib.cuphyCreateTensorDescriptor(C.byref(to))
    dims = (C.c_int * 2)(nllr, 3)
    L.check("set", lib.cuphySetTensorDescriptor(tl, L.CUPHY_R_32F, 2, dims, None, 0))
    dims_o = (C.c_int * 2)(kbits, 3)
    L.check("set", lib.cuphySetTensorDescriptor(to, L.CUPHY_BIT, 2, dims_o, None, 0))
    from aerial_amd.decoder import LdpcDecodeConfig
    cfg = LdpcDecodeConfig(bg, Z, mb, llr_type=L.CUPHY_R_32F, max_iterations=10, alpha=0.8125)
    c = dec.finalize_config(cfg)
    st = lib.cuphyErrorCorrectionLDPCDecode(dec.dec, to, C.c_void_p(out.data_ptr()), tl, C.c_void_p(x.data_ptr()),
                                            None, None, C.byref(c), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    L.check("cuphyErrorCorrectionLDPCDecode", st)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), h_ref)
    lib.cuphyDestroyTensorDescriptor(tl)
    lib.cuphyDestroyTensorDescriptor(to)


def test_launch_descriptor_runs(dec):
    """GetLaunchDescriptor fills a hipKernelNodeParams that hipLaunchKernel accepts."""
    from aerial_amd import _lib as L
    from aerial_amd.decoder import LdpcDecodeConfig
    _log("launch descriptor BG1 Z=192 mb=5 2 CBs")
    lib = L.lib()
    llr, _ = V.make_cbs(1, 192, 5, 2, seed=41, snr_db=4.0)
    h_ref, _ = run_gpu(dec, 1, 192, 5, llr, soft=False)
    cfg = LdpcDecodeConfig(1, 192, 5, llr_type=L.CUPHY_R_32F, max_iterations=10, alpha=0.8125)
    x = torch.from_numpy(np.ascontiguousarray(llr)).cuda()
    out = torch.zeros((2, cfg.out_words()), dtype=torch.int32, device="cuda")
    lc = L.LDPCDecodeLaunchConfig()
    lc.decode_desc = dec.build_desc(cfg, [x], [out])
    L.check("GetLaunchDescriptor", lib.cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor(dec.dec, C.byref(lc)))
    p = lc.kernel_node_params_driver
    assert p.gridDim.x == 2 and p.blockDim.x == 192 and p.func
    hip = C.CDLL("libamdhip64.so")
    hip.hipLaunchKernel.argtypes = [C.c_void_p, L.Dim3, L.Dim3, C.c_void_p, C.c_size_t, C.c_void_p]
    rc = hip.hipLaunchKernel(p.func, p.gridDim, p.blockDim, p.kernelParams, p.sharedMemBytes,
                             C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), h_ref)
    # mb > 8: the descriptor carries the row-pair kernel (two lane groups of
    # roundup64(Z) lanes), with the same results as the direct call
    _log("launch descriptor BG1 Z=192 mb=20 2 CBs (row-pair)")
    llr, _ = V.make_cbs(1, 192, 20, 2, seed=42, snr_db=1.0)
    h_ref, _ = run_gpu(dec, 1, 192, 20, llr, soft=False)
    cfg = LdpcDecodeConfig(1, 192, 20, llr_type=L.CUPHY_R_32F, max_iterations=10, alpha=0.8125)
    x = torch.from_numpy(np.ascontiguousarray(llr)).cuda()
    out = torch.zeros((2, cfg.out_words()), dtype=torch.int32, device="cuda")
    lc = L.LDPCDecodeLaunchConfig()
    lc.decode_desc = dec.build_desc(cfg, [x], [out])
    L.check("GetLaunchDescriptor", lib.cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor(dec.dec, C.byref(lc)))
    p = lc.kernel_node_params_driver
    assert p.gridDim.x == 2 and p.blockDim.x == 384 and p.func
    rc = hip.hipLaunchKernel(p.func, p.gridDim, p.blockDim, p.kernelParams, p.sharedMemBytes,
                             C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), h_ref)
