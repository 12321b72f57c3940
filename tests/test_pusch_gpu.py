"""GPU parity of de-rate-matching/HARQ combining and CB/TB CRC against the
oracle (bit-exact HARQ buffers, CRC remainders and TB bytes)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


# (bg, Zc, C, Kprime, Qm, Nl, G, rv) -- includes repetition (E > Ncb - F),
# filler bits, k0 inside/after the filler hole, all Qm.
RM_CASES = [
    (1, 384, 2, 8448, 8, 4, 2 * 9024, 0),
    (1, 384, 3, 8424, 2, 1, 3 * 4000, 2),
    (1, 64, 1, 1300, 4, 2, 6000, 3),            # E > Ncb - F: wrap-around repetition
    (1, 208, 4, 4400, 6, 2, 4 * 9600, 1),
    (2, 52, 2, 480, 2, 1, 2 * 3000, 0),          # BG2 + filler, repetition
    (2, 384, 3, 3840, 8, 2, 3 * 4800, 3),
    (2, 20, 1, 150, 1, 1, 2000, 2),              # Qm = 1, heavy repetition
    (2, 128, 5, 1232, 4, 4, 5 * 2400, 1),
]


def tb_llrs(G, seed):
    rng = np.random.RandomState(seed)
    v = rng.uniform(-40.0, 40.0, G).astype(np.float32)
    v[rng.randint(0, G, G // 50)] *= 400.0  # some beyond the +-10000 clamp
    return v


def qam_stride_layout(llr_demap, Qm, Nl, n_bbu, layer_map, fill=123.0):
    """Input buffer in the reference's equalizer-output layout: symbol m of the
    TB at RE m // Nl, layer layer_map[m % Nl], 8 LLR slots per symbol."""
    nsym = llr_demap.size // Qm
    nre = (nsym + Nl - 1) // Nl
    buf = np.full((nre, n_bbu, 8), fill, np.float32)
    for m in range(nsym):
        buf[m // Nl, layer_map[m % Nl], :Qm] = llr_demap[m * Qm:(m + 1) * Qm]
    return buf.reshape(-1)


# descrambling cases: the RM_CASES plus one TB of 50 CBs (G = 600,000: the
# Gold sequence far from its start)
DESCR_CASES = RM_CASES + [(1, 384, 50, 8448, 8, 1, 50 * 12000, 1)]


# (case, descrambling): every RM case without and with descrambling, the long TB with it only
RM_RUNS = [(c, d) for d in (False, True) for c in range(len(DESCR_CASES)) if d or c < len(RM_CASES)]


@pytest.mark.parametrize("fp", [0, 1, 2, 3])
@pytest.mark.parametrize("case,descr", RM_RUNS)
def test_derate_match_two_transmissions(case, fp, descr):
    _need_gpu()
    from aerial_amd.pusch import DeRateMatch, make_tb_params
    bg, Zc, C, Kp, Qm, Nl, G, rv = DESCR_CASES[case]
    in16, out16 = fp in (1, 3), fp in (2, 3)
    cinit = (0x1234567 * (case + 1) + 0x9E37) & 0x7FFFFFFF if descr else 0
    print("[case] derate BG%d Zc=%d C=%d K'=%d Qm=%d Nl=%d G=%d rv=%d fp=%d descr=%d cinit=%#x"
          % (bg, Zc, C, Kp, Qm, Nl, G, rv, fp, descr, cinit), flush=True)
    n_bbu = Nl + (1 if case % 2 else 0)
    lmap = list(range(n_bbu))[::-1][:Nl] if case % 2 else list(range(Nl))
    rm = DeRateMatch(fp_config=fp, descrambling=descr)
    p1, _ = make_tb_params(bg, Zc, C, Kp, Qm, Nl, G, rv, ndi=1, n_bbu_layers=n_bbu, layer_map=lmap)
    ocfg = O.rm_cfg(bg=bg, Zc=Zc, C=C, Qm=Qm, Nl=Nl, G=G, K=p1.K, F=p1.F, Ncb=p1.Ncb,
                    Ncb_padded=p1.Ncb_padded, k0=0, ndi=1, descramble=int(descr), cinit=cinit)
    n_cmp = 2 * Zc + p1.Ncb
    ref = np.zeros(C * p1.Ncb_padded, np.float32)
    out_t = torch.float16 if out16 else torch.float32
    harq = torch.full((C * p1.Ncb_padded,), 77.0, dtype=out_t, device="cuda")
    for tx, (ndi, rv_) in enumerate([(1, rv), (0, (rv + 2) % 4)]):
        p, _ = make_tb_params(bg, Zc, C, Kp, Qm, Nl, G, rv_, ndi=ndi, n_bbu_layers=n_bbu, layer_map=lmap)
        p.cinit = cinit
        llr = tb_llrs(G, 1000 * case + tx)
        if in16:
            llr = llr.astype(np.float16).astype(np.float32)
        ocfg.ndi, ocfg.k0 = ndi, O.k0(bg, rv_, p.Ncb, Zc)
        assert p.k0 == ocfg.k0
        O.derate_match(ocfg, llr, ref, fp16=out16)
        buf = qam_stride_layout(llr, Qm, Nl, n_bbu, lmap)
        x = torch.from_numpy(buf).cuda()
        if in16:
            x = x.half()
        rm.run([p], [x], [harq])
        torch.cuda.synchronize()
        got = harq.float().cpu().numpy().reshape(C, -1)[:, :n_cmp]
        exp = ref.reshape(C, -1)[:, :n_cmp]
        bad = np.argwhere(got.view(np.uint32) != exp.view(np.uint32))
        assert bad.size == 0, "transmission %d: %d mismatches, first %s got %r exp %r" % (
            tx, len(bad), bad[0], got[tuple(bad[0])], exp[tuple(bad[0])])
    rm.close()


def test_derate_match_multi_ue_first_cb_offset():
    """Several UEs in one setup, one of them starting at a CB offset."""
    _need_gpu()
    from aerial_amd.pusch import DeRateMatch, make_tb_params
    print("[case] derate 3 UEs, firstCodeBlockIndex=1", flush=True)
    specs = [(1, 96, 2, 2112, 4, 1, 2 * 3000, 0, 0), (2, 72, 3, 720, 2, 2, 3 * 1600, 2, 1), (1, 160, 1, 3520, 6, 1, 4800, 3, 0)]
    rm = DeRateMatch(fp_config=3)
    params, xs, outs, refs = [], [], [], []
    for u, (bg, Zc, C, Kp, Qm, Nl, G, rv, first) in enumerate(specs):
        p, _ = make_tb_params(bg, Zc, C, Kp, Qm, Nl, G, rv, ndi=1, user_group=u, first_cb=first)
        llr = tb_llrs(G, 500 + u).astype(np.float16).astype(np.float32)
        ref = np.zeros((C + first) * p.Ncb_padded, np.float32)
        ocfg = O.rm_cfg(bg=bg, Zc=Zc, C=C, Qm=Qm, Nl=Nl, G=G, K=p.K, F=p.F, Ncb=p.Ncb, Ncb_padded=p.Ncb_padded,
                        k0=p.k0, ndi=1)
        sub = np.zeros(C * p.Ncb_padded, np.float32)
        O.derate_match(ocfg, llr, sub, fp16=True)
        ref[first * p.Ncb_padded:] = sub
        params.append(p)
        xs.append(torch.from_numpy(qam_stride_layout(llr, Qm, Nl, Nl, list(range(Nl)))).cuda().half())
        outs.append(torch.zeros((C + first) * p.Ncb_padded, dtype=torch.float16, device="cuda"))
        refs.append((ref, first, C, p))
    rm.run(params, xs, outs)
    torch.cuda.synchronize()
    for (ref, first, C, p), o in zip(refs, outs):
        n = 2 * p.Zc + p.Ncb
        got = o.float().cpu().numpy().reshape(C + first, -1)[first:, :n]
        assert np.array_equal(got, ref.reshape(C + first, -1)[first:, :n])
    rm.close()


def _modulo_cases():
    """Distinct rows of the reference's de-rate-match table
    (test_derate_match_modulo.cu:122-228: E, K - 2Zc, F, k0, Ncb, Zc of its
    PUSCH test vectors and two multi-loop cases; fixture in tests/golden)."""
    import json
    import os
    rows = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "derate_modulo_cases.json")))
    seen, out = set(), []
    for r in rows:
        key = (r["E"], r["K"], r["F"], r["k0"], r["Ncb"], r["Zc"])
        if key not in seen:
            seen.add(key)
            out.append(r)
    return out


def test_derate_match_reference_cases():
    """Every distinct (E, K, F, k0, Ncb, Zc) of the reference's bit-selection
    table as a one-CB TB (incl. the limited-buffer Ncb = 17915 of TVnr_7321,
    k0 inside the filler hole and E up to 5 Ncb), two transmissions
    (overwrite, then combine at the same k0): the binary32 HARQ buffer equals
    the oracle's, whose bit selection the CPU tests pin to the reference's
    textbook selection on the same rows."""
    _need_gpu()
    from aerial_amd.pusch import DeRateMatch, make_tb_params
    cases = _modulo_cases()
    print("[case] derate: %d reference (E, K, F, k0, Ncb, Zc) rows, fp32" % len(cases), flush=True)
    rm = DeRateMatch(fp_config=0)
    for i, r in enumerate(cases):
        Zc, F, E = r["Zc"], r["F"], r["E"]
        K = r["K"] + 2 * Zc
        bg = 1 if K == 22 * Zc else 2
        Qm = next(q for q in (8, 6, 4, 2, 1) if E % q == 0)
        N = (66 if bg == 1 else 50) * Zc
        p, _ = make_tb_params(bg, Zc, 1, K - F, Qm, 1, E, 0, ndi=1, Nref=r["Ncb"] if r["Ncb"] < N else 0)
        assert p.Ncb == r["Ncb"] and p.K == K and p.F == F
        p.k0 = r["k0"]
        ocfg = O.rm_cfg(bg=bg, Zc=Zc, C=1, Qm=Qm, Nl=1, G=E, K=K, F=F, Ncb=r["Ncb"], Ncb_padded=p.Ncb_padded,
                        k0=r["k0"], ndi=1)
        ref = np.zeros(p.Ncb_padded, np.float32)
        harq = torch.full((p.Ncb_padded,), 9.0, dtype=torch.float32, device="cuda")
        for tx in range(2):
            llr = tb_llrs(E, 7000 + 10 * i + tx)
            p.ndi = ocfg.ndi = 1 - tx
            O.derate_match(ocfg, llr, ref)
            rm.run([p], [torch.from_numpy(qam_stride_layout(llr, Qm, 1, 1, [0])).cuda()], [harq])
            torch.cuda.synchronize()
            n = 2 * Zc + r["Ncb"]
            got = harq.cpu().numpy()[:n]
            assert np.array_equal(got.view(np.uint32), ref[:n].view(np.uint32)), (r, tx)
    rm.close()


def test_derate_match_uci_on_pusch():
    """uciOnPuschFlag: the UL-SCH LLRs are G values in transmission order
    (no QAM_STRIDE layout) and the E split follows G, with encodedSize (all
    bits of the allocation) larger than G; two transmissions against the
    oracle."""
    _need_gpu()
    from aerial_amd.pusch import DeRateMatch, make_tb_params
    print("[case] derate uci-on-pusch BG1 Zc=208 C=3 Qm=4 Nl=2", flush=True)
    bg, Zc, C, Kp, Qm, Nl = 1, 208, 3, 4400, 4, 2
    G = 8 * 1879  # 1879 symbol pairs over 3 CBs: E = 5008, 5008, 5016
    rm = DeRateMatch(fp_config=0)
    ref = None
    for tx, (ndi, rv) in enumerate([(1, 0), (0, 2)]):
        p, _ = make_tb_params(bg, Zc, C, Kp, Qm, Nl, G, rv, ndi=ndi)
        p.uciOnPuschFlag = 1
        p.encodedSize = G + 8 * 137  # UCI bits of the allocation beyond the UL-SCH's G
        if ref is None:
            ref = np.zeros(C * p.Ncb_padded, np.float32)
            harq = torch.full((C * p.Ncb_padded,), 5.0, dtype=torch.float32, device="cuda")
        ocfg = O.rm_cfg(bg=bg, Zc=Zc, C=C, Qm=Qm, Nl=Nl, G=G, K=p.K, F=p.F, Ncb=p.Ncb, Ncb_padded=p.Ncb_padded,
                        k0=O.k0(bg, rv, p.Ncb, Zc), ndi=ndi)
        llr = tb_llrs(G, 4100 + tx)
        O.derate_match(ocfg, llr, ref)
        x = torch.from_numpy(llr).cuda()
        p.d_schAndCsi2LLRs = x.data_ptr()  # the UL-SCH LLRs of a UCI-on-PUSCH TB come from PerTbParams
        rm.run([p], [x], [harq])
        torch.cuda.synchronize()
        n = 2 * Zc + p.Ncb
        got = harq.cpu().numpy().reshape(C, -1)[:, :n]
        assert np.array_equal(got.view(np.uint32), ref.reshape(C, -1)[:, :n].view(np.uint32)), tx
    rm.close()


# ---------------------------------------------------------------------------
# CRC
# ---------------------------------------------------------------------------
def build_tb(A, C, seed):
    """TB of A bits + CRC24A (or CRC16 / CRC24A for C = 1) segmented into C
    code blocks with CRC24B (C > 1); returns (cb words (C, 264) uint32, K')."""
    rng = np.random.RandomState(seed)
    a = rng.randint(0, 2, A).astype(np.uint8)
    if C == 1:
        L = 16 if A <= 3824 else 24
        crc = O.crc(a, O.CRC16 if L == 16 else O.CRC24A)
        tb = np.concatenate([a, np.array([(crc >> (L - 1 - i)) & 1 for i in range(L)], np.uint8)])
        cbs = [tb]
    else:
        crc = O.crc(a, O.CRC24A)
        tb = np.concatenate([a, np.array([(crc >> (23 - i)) & 1 for i in range(24)], np.uint8)])
        d = tb.size // C
        assert d * C == tb.size
        cbs = []
        for r in range(C):
            seg = tb[r * d:(r + 1) * d]
            c = O.crc(seg, O.CRC24B)
            cbs.append(np.concatenate([seg, np.array([(c >> (23 - i)) & 1 for i in range(24)], np.uint8)]))
    Kp = cbs[0].size
    words = np.zeros((C, O.WORDS_PER_CB), np.uint32)
    for r, cb in enumerate(cbs):
        w = O.pack_bits(cb)
        words[r, :w.size] = w
    return words, Kp


CRC_CASES = [(8424 * 3 - 24 - 48, 3), (1000, 1), (3824, 1), (4000, 1), (8448 * 2 - 48 - 24 - 800, 2), (6000 - 24, 6)]


@pytest.mark.parametrize("reverse", [True])
def test_crc_decode_matches_oracle(reverse):
    _need_gpu()
    from aerial_amd.pusch import CrcDecode, make_tb_params
    print("[case] crc %d TBs" % len(CRC_CASES), flush=True)
    params, words, refs = [], [], []
    for t, (A, C) in enumerate(CRC_CASES):
        if C > 1:
            A = (A + 24) // (8 * C) * (8 * C) - 24  # C*(K'-24) = A + 24, K'-24 a multiple of 8
        w, Kp = build_tb(A, C, 40 + t)
        if t == 1:
            w[0, 3] ^= 1 << 5  # corrupt a single-CB TB
        if t == 4:
            w[1, 10] ^= 1      # corrupt CB 1 of a 2-CB TB
        bg = 1 if Kp > 3840 else 2
        Zc = next(z for z in sorted({a << j for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8)})
                  if (22 if bg == 1 else 10) * z >= Kp)
        p, _ = make_tb_params(bg, Zc, C, Kp, 2, 1, C * 4000, 0)
        params.append(p)
        words.append(w)
        refs.append(O.crc_desegment(w.reshape(-1), C, Kp))
    allw = torch.from_numpy(np.concatenate(words).reshape(-1).view(np.int32).copy()).cuda()
    ncb = sum(C for _, C in CRC_CASES)
    cb_crc = torch.full((ncb,), -1, dtype=torch.int32, device="cuda")
    tb_crc = torch.full((len(params),), -1, dtype=torch.int32, device="cuda")
    tb_bytes = torch.full((sum(len(r[1]) for r in refs) + 64,), 0xAB, dtype=torch.uint8, device="cuda")
    crc = CrcDecode(reverse_bytes=reverse)
    crc.run(params, allw, cb_crc, tb_bytes, tb_crc)
    torch.cuda.synchronize()
    got_cb = cb_crc.cpu().numpy().view(np.uint32)
    got_tb = tb_crc.cpu().numpy().view(np.uint32)
    got_bytes = tb_bytes.cpu().numpy()
    off_cb, off_tb = 0, 0
    for t, (ref_cb, ref_bytes, ref_tb) in enumerate(refs):
        C = CRC_CASES[t][1]
        assert np.array_equal(got_cb[off_cb:off_cb + C], ref_cb), t
        assert got_tb[t] == ref_tb, t
        assert np.array_equal(got_bytes[off_tb:off_tb + len(ref_bytes)], ref_bytes), t
        off_cb += C
        off_tb += len(ref_bytes)
    # pass/fail pattern: only the corrupted TBs fail
    assert [int(got_tb[t] != 0) for t in range(len(refs))] == [0, 1, 0, 0, 1, 0]
    crc.close()
