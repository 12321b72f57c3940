"""Statistical decoding-performance checks on the GPU (paired noise).

1. binary16 LLR storage + binary16 arithmetic against the binary32 path
   (bit-identical to the oracle's binary32 restatement of the reference CPU
   decoder, see test_decoder_gpu.py): the Es/N0 at BLER 1e-2 differs by at
   most 0.05 dB; the APP deviation of the binary16 soft outputs is stated.
2. the int8 fused slot (int8 LLRs, 4 steps per unit, +-127; de-rate-match,
   HARQ write, 8 iterations with early termination, CRC) against the
   binary32 path on the unquantised LLRs of the same transmissions, at the
   points where the binary32 BLER is about 1e-1 and 1e-2.
Every run prints its measured numbers."""
import numpy as np
import pytest

pytestmark = [pytest.mark.gpu]
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from aerial_amd.decoder import LdpcDecoder
    d = LdpcDecoder()
    yield d
    d.close()


def block_errors(tb, out):
    bits = out.view(torch.int32)
    got = ((bits.unsqueeze(-1) >> torch.arange(32, device="cuda", dtype=torch.int32)) & 1).reshape(tb.ncb, -1)
    got = got[:, :tb.info_bits_per_cb].to(torch.uint8)
    return (got != tb.info).any(dim=1)


def run_point(dec, bg, Z, mb, snr_db, n_cb, seed, chunk=4000, want_soft=False):
    """Block-error flags of the binary32 and binary16 paths on the same
    LLRs (BPSK-AWGN at snr_db), plus APP deviation statistics."""
    from aerial_amd import _lib as L
    from aerial_amd import workloads as W
    from aerial_amd.batch import BatchDecoder
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    K = W.nsys(bg) * Z
    n_tx = (W.nsys(bg) + mb - 2) * Z
    e32, e16, dev = [], [], []
    done = 0
    while done < n_cb:
        n = min(chunk, n_cb - done)
        tb = W.fill_tb(W.TransportBlock(bg, Z, mb, n, n_tx, snr_db), gen, llr_dtype=torch.float32)
        out32 = torch.zeros_like(tb.out)
        soft32 = torch.zeros((n, K), dtype=torch.float16, device="cuda") if want_soft else None
        tb.out = out32
        BatchDecoder(dec).build([tb], llr_type=L.CUPHY_R_32F, fp16_math=False, max_iterations=10,
                                soft=[soft32] if want_soft else None).launch()
        e32.append(block_errors(tb, out32))
        tb16 = W.TransportBlock(bg, Z, mb, n, n_tx, snr_db)
        tb16.llr = tb.llr.half()
        tb16.info = tb.info
        tb16.out = torch.zeros_like(out32)
        soft16 = torch.zeros((n, K), dtype=torch.float16, device="cuda") if want_soft else None
        BatchDecoder(dec).build([tb16], llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=10,
                                soft=[soft16] if want_soft else None).launch()
        e16.append(block_errors(tb16, tb16.out))
        if want_soft:
            d = (soft16.float() - soft32.float()).abs()
            dev.append((float(d.mean().item()), float(d.max().item()),
                        float((soft16.float() * soft32.float() < 0).float().mean().item())))
        done += n
    torch.cuda.synchronize()
    return torch.cat(e32).cpu().numpy(), torch.cat(e16).cpu().numpy(), dev


def snr_at(snrs, bler, target):
    """Es/N0 where log10(BLER) crosses log10(target), linear interpolation."""
    lb = np.log10(np.maximum(bler, 1e-6))
    lt = np.log10(target)
    for i in range(len(snrs) - 1):
        if lb[i] >= lt >= lb[i + 1]:
            f = (lb[i] - lt) / max(1e-12, lb[i] - lb[i + 1])
            return snrs[i] + f * (snrs[i + 1] - snrs[i])
    return None


def test_fp16_storage_within_0p05_db_at_bler_1e2(dec):
    bg, Z, mb = 1, 384, 10  # R = 8448 / 11520 = 0.73
    print("[case] BLER fp16 vs fp32, BG1 Z=384 mb=10", flush=True)
    coarse = np.arange(0.5, 3.01, 0.25)
    b = [run_point(dec, bg, Z, mb, s, 1000, 1000 + i)[0].mean() for i, s in enumerate(coarse)]
    s0 = snr_at(coarse, np.array(b), 1e-2)
    assert s0 is not None, list(zip(coarse, b))
    fine = np.round(np.arange(s0 - 0.2, s0 + 0.21, 0.1), 3)
    b32, b16, devs = [], [], []
    for i, s in enumerate(fine):
        e32, e16, dev = run_point(dec, bg, Z, mb, float(s), 20000, 2000 + i, want_soft=(i == 2))
        b32.append(e32.mean())
        b16.append(e16.mean())
        devs += dev
    s32 = snr_at(fine, np.array(b32), 1e-2)
    s16 = snr_at(fine, np.array(b16), 1e-2)
    app = np.array(devs).mean(axis=0) if devs else None
    print("BLER fp32 %s" % dict(zip(fine.tolist(), np.round(b32, 5).tolist())))
    print("BLER fp16 %s" % dict(zip(fine.tolist(), np.round(b16, 5).tolist())))
    print("Es/N0 at BLER 1e-2: fp32 %.4f dB, fp16 %.4f dB, difference %.4f dB" % (s32, s16, s16 - s32))
    if app is not None:
        print("APP (soft output) deviation fp16 vs fp32 at %.2f dB: mean |d| %.4f, max |d| %.2f, sign flips %.2e"
              % (fine[2], app[0], app[1], app[2]))
    assert s32 is not None and s16 is not None
    assert abs(s16 - s32) <= 0.05


def test_int8_fused_slot_bler_against_binary32(dec):
    """BG1 Zc=384 R=0.93 TBs of 20 CBs (E = 9088), consistent-Gaussian LLRs:
    CB errors of the int8 fused slot vs the binary32 path (de-rate-match
    kernel fp32 -> binary32 decode, same early termination, CB CRC)."""
    from aerial_amd import _lib as L
    from aerial_amd import pusch_workloads as PW
    from aerial_amd.decoder import LdpcDecodeConfig
    from aerial_amd.pusch import CrcDecode, DeRateMatch
    from aerial_amd.slot import SlotDecoder, SlotTb
    print("[case] BLER int8 fused vs fp32 path, BG1 Z=384 R=0.93", flush=True)
    n_tb, C = 250, 20

    def point(snr, seed):
        specs = [(1, 384, C, 8448, 8, 1, 9088, snr)] * n_tb
        tbs = PW.build_tbs(specs, seed)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        e8 = e32 = 0
        rm = DeRateMatch(fp_config=0)
        crc = CrcDecode()
        slot = []
        for tb in tbs:
            llr = PW.tx_llrs(tb, 0, gen, mode="cg", dtype=torch.float32)
            p, mb = tb.params(0, 1)
            q = torch.clamp(torch.round(llr * 4.0), -127, 127).to(torch.int8)
            harq8 = torch.zeros(C * p.Ncb_padded, dtype=torch.float16, device="cuda")
            slot.append(SlotTb(p, q, harq8, mb, max_iterations=8, early_term=True, llr_scale=0.25))
            # binary32 path on the unquantised LLRs
            harq32 = torch.zeros(C * p.Ncb_padded, dtype=torch.float32, device="cuda")
            rm.run([p], [llr], [harq32])
            words = torch.zeros((C, L.MAX_WORDS_PER_CODE_BLOCK), dtype=torch.int32, device="cuda")
            cfg = LdpcDecodeConfig(1, 384, mb, llr_type=L.CUPHY_R_32F, max_iterations=8,
                                   flags=L.CUPHY_LDPC_DECODE_EARLY_TERM)
            dec.decode_tb(cfg, [harq32.view(C, -1)], [words])
            cb_crc = torch.zeros(C, dtype=torch.int32, device="cuda")
            tb_bytes = torch.zeros(C * 1056 + 64, dtype=torch.uint8, device="cuda")
            tb_crc = torch.zeros(1, dtype=torch.int32, device="cuda")
            crc.run([p], words, cb_crc, tb_bytes, tb_crc)
            torch.cuda.synchronize()
            e32 += int((cb_crc != 0).sum().item())
        SlotDecoder(dec).build(slot).launch()
        torch.cuda.synchronize()
        e8 = sum(int((st.cb_crc != 0).sum().item()) for st in slot)
        rm.close()
        crc.close()
        return e32 / (n_tb * C), e8 / (n_tb * C)

    # Es/N0 (BPSK-equivalent) grid around the R = 0.93 waterfall
    grid = np.round(np.arange(3.0, 5.01, 0.25), 3)
    res = [point(float(s), 300 + i) for i, s in enumerate(grid)]
    b32 = np.array([r[0] for r in res])
    b8 = np.array([r[1] for r in res])
    print("CB BLER binary32 %s" % dict(zip(grid.tolist(), b32.round(5).tolist())))
    print("CB BLER int8 fused %s" % dict(zip(grid.tolist(), b8.round(5).tolist())))
    for target in (1e-1, 1e-2):
        s32 = snr_at(grid, b32, target)
        s8 = snr_at(grid, b8, target)
        print("Es/N0 at CB BLER %.0e: binary32 %s dB, int8 fused %s dB" % (target, s32, s8))
        if s32 is not None and s8 is not None:
            assert s8 - s32 <= 0.25, (target, s32, s8)  # int8 (0.25 LLR steps, +-31.75) costs little
