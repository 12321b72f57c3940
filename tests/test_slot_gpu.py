"""Fused slot decode (de-rate-matching + HARQ combining + binary16 layered
NMS decoding + CB/TB CRC in one persistent launch, cuphyHipPuschSlotDecode):
bit-exact against the oracle's three stages (HARQ buffers, decoded bits,
iterations run, CB CRCs, TB bytes, TB CRCs) over two HARQ transmissions,
for int8 / binary16 / binary32 input LLRs, with and without descrambling;
equal to the unfused GPU pipeline; replayable from a graph."""
import numpy as np
import pytest

import oracle as O
import pusch_vectors as P

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# (bg, Zc, C, K', Qm, Nl, nbbu, E per CB, Es/N0 dB)
SLOT = [
    (1, 384, 3, 8448, 8, 2, 2, 9024, 2.5),     # R ~ 0.93: mb 4..8 class on tx 1
    (2, 208, 2, 1800, 4, 1, 2, 4000, -1.0),    # BG2 with filler, layer map over 2 BBU layers
    (1, 104, 4, 2200, 2, 1, 1, 3000, 0.0),
    (1, 64, 1, 1296, 6, 3, 3, 4500, 1.0),      # C = 1: CRC16 TB; E > Ncb - F: repetition
    (2, 52, 2, 480, 2, 1, 1, 3000, -2.0),      # repetition, BG2 filler
    (2, 384, 2, 3840, 8, 4, 4, 4096, 3.0),     # high rate BG2
    (2, 32, 12, 280, 2, 1, 1, 600, 0.5),       # 12 CBs of Zc=32: one work item of 12 side by side
]
TX = [(1, 0), (0, 2)]  # (ndi, rv) of the two transmissions
ITERS = 6


@pytest.fixture(scope="module")
def dec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from aerial_amd.decoder import LdpcDecoder
    d = LdpcDecoder()
    yield d
    d.close()


def half_to_float(h):
    return float(np.array([h], np.uint16).view(np.float16)[0])


class Case:
    """Device state and oracle state of one TB across transmissions."""

    def __init__(self, k, spec, dtype, descramble=False):
        from aerial_amd.pusch import make_tb_params
        bg, Zc, C, Kp, Qm, Nl, nbbu, E, snr = spec
        self.spec, self.dtype, self.k = spec, dtype, k
        self.descramble = descramble
        self.cinit = (0x2468ACE * (k + 1) + 0x1357) & 0x7FFFFFFF if descramble else 0
        self.tb = P.TbVector(bg, Zc, C, Kp, seed=700 + k)
        self.G = C * E
        assert E % (Qm * Nl) == 0
        self.lmap = list(range(nbbu))[::-1][:Nl] if nbbu > Nl else list(range(Nl))
        p0, _ = make_tb_params(bg, Zc, C, Kp, Qm, Nl, self.G, 0, ndi=1, n_bbu_layers=nbbu, layer_map=self.lmap)
        self.Ncb_padded = p0.Ncb_padded
        self.harq = torch.full((C * p0.Ncb_padded,), 55.0, dtype=torch.float16, device="cuda")
        self.ref_harq = np.zeros(C * p0.Ncb_padded, np.float32)
        self.mb = 4

    def transmission(self, ndi, rv, tx):
        from aerial_amd.pusch import make_tb_params
        bg, Zc, C, Kp, Qm, Nl, nbbu, E, snr = self.spec
        p, mb = make_tb_params(bg, Zc, C, Kp, Qm, Nl, self.G, rv, ndi=ndi, n_bbu_layers=nbbu, layer_map=self.lmap)
        # k0 and parity nodes as the oracle derives them (derate_match.m:86-129); the host helper must agree
        k0 = O.k0(bg, rv, p.Ncb, Zc)
        mb_o = O.num_parity_nodes(bg, Zc, p.K, p.F, self.G, C, k0)
        assert (p.k0, mb) == (k0, mb_o), (self.spec, rv)
        self.mb = mb_o if ndi else max(self.mb, mb_o)
        llr = P.bpsk_llrs(self.tb.tx_bits(self.G, Qm, Nl, rv), snr, seed=9000 + 10 * self.k + tx)
        if self.descramble:  # the channel carries the scrambled bits b(n) ^ c(n)
            llr = np.where(O.gold(self.cinit, 0, self.G) != 0, -llr, llr).astype(np.float32)
        if self.dtype == "int8":
            q = P.quantize_int8(llr)
            seen = q.astype(np.float32) * 0.25
            buf = P.qam_stride_layout(q.astype(np.float32), Qm, Nl, nbbu, self.lmap)
            dev = torch.from_numpy(buf.astype(np.int8)).cuda()
            scale = 0.25
        else:
            seen = llr
            buf = P.qam_stride_layout(llr, Qm, Nl, nbbu, self.lmap)
            dev = torch.from_numpy(buf).cuda()
            if self.dtype == "fp16":
                dev = dev.half()
                seen = llr.astype(np.float16).astype(np.float32)
            scale = 1.0
        ocfg = O.rm_cfg(bg=bg, Zc=Zc, C=C, Qm=Qm, Nl=Nl, G=self.G, K=p.K, F=p.F, Ncb=p.Ncb,
                        Ncb_padded=p.Ncb_padded, k0=k0, ndi=ndi, descramble=int(self.descramble),
                        cinit=self.cinit)
        O.derate_match(ocfg, seen, self.ref_harq, fp16=True)
        p.cinit = self.cinit
        self.p = p
        return p, dev, scale

    def oracle_decode(self, alpha, early_term):
        bg, Zc, C, Kp = self.spec[:4]
        cfg = O.dec_cfg(bg, Zc, self.mb, n_iter=ITERS, alpha=alpha, clamp=32.0, mode=O.FP16, early_term=early_term)
        words = np.zeros((C, O.WORDS_PER_CB), np.uint32)
        iters = np.zeros(C, np.int64)
        h = self.ref_harq.reshape(C, -1)
        for r in range(C):
            w, it, _ = O.decode_cb(cfg, h[r, :(P.nsys(bg) + self.mb) * Zc])
            words[r, :w.size] = w
            iters[r] = it
        cb_crc, tb_bytes, tb_crc = O.crc_desegment(words.reshape(-1), C, Kp)
        return words, iters, cb_crc, tb_bytes, tb_crc


def run_slot(dec, dtype, early_term, specs=SLOT, descramble=False):
    from aerial_amd.slot import SlotDecoder, SlotTb
    cases = [Case(k, s, dtype, descramble) for k, s in enumerate(specs)]
    # int8: one launch of the full variant; otherwise one launch per row
    # class (mb <= 8, mb <= 16 -- tx 1's BG1 Zc=104 and BG2 Zc=32 TBs -- and
    # the full variant), on separate streams
    sd = SlotDecoder(dec, split_min=256 if dtype == "int8" else 1)
    for tx, (ndi, rv) in enumerate(TX):
        print("[case] slot %s tx %d ndi %d rv %d ET %d descr %d" % (dtype, tx, ndi, rv, early_term, descramble),
              flush=True)
        tbs = []
        for c in cases:
            p, dev, scale = c.transmission(ndi, rv, tx)
            tbs.append(SlotTb(p, dev, c.harq, c.mb, max_iterations=ITERS, early_term=early_term, llr_scale=scale,
                              descramble=descramble))
        sd.build(tbs).launch()
        torch.cuda.synchronize()
        assert all(int(w[:2].abs().sum().item()) == 0 for w in sd.workspaces), "queue not returned to zero"
        assert all(int(w.abs().sum().item()) == 0 for w in sd.workspaces), "TB words not returned to zero"
        for c, tb in zip(cases, tbs):
            bg, Zc, C, Kp = c.spec[:4]
            alpha = P.table_alpha(bg, c.mb)  # the reference's table, as SetNormalization must give it
            assert half_to_float(sd._alpha_bits(tb)) == alpha
            words, iters, cb_crc, tb_bytes, tb_crc = c.oracle_decode(alpha, early_term)
            n = 2 * Zc + c.p.Ncb
            got_h = c.harq.float().cpu().numpy().reshape(C, -1)[:, :n]
            assert np.array_equal(got_h.view(np.uint32), c.ref_harq.reshape(C, -1)[:, :n].view(np.uint32)), c.spec
            nw = (P.nsys(bg) * Zc + 31) // 32
            got_w = tb.cb_bits.cpu().numpy().view(np.uint32)
            assert np.array_equal(got_w[:, :nw], words[:, :nw]), (c.spec, tx)
            assert np.array_equal(tb.cb_iterations.cpu().numpy().astype(np.int64), iters), (c.spec, tx)
            assert np.array_equal(tb.cb_crc.cpu().numpy().view(np.uint32), cb_crc), (c.spec, tx)
            assert int(tb.tb_crc.cpu().numpy().view(np.uint32)[0]) == tb_crc, (c.spec, tx)
            assert np.array_equal(tb.tb_bytes.cpu().numpy()[:len(tb_bytes)], tb_bytes), (c.spec, tx)
    return cases


@pytest.mark.parametrize("dtype", ["fp16", "int8", "fp32"])
def test_slot_matches_oracle(dec, dtype):
    cases = run_slot(dec, dtype, early_term=False)
    assert len(cases) == len(SLOT)


@pytest.mark.parametrize("dtype", ["fp16", "int8"])
def test_slot_early_termination_matches_oracle(dec, dtype):
    """Syndrome early termination with identical stopping: iterations per CB,
    bits, CRCs and HARQ buffers equal the oracle's, for binary16 and for the
    int8 LLRs of the fused receive-side slot."""
    run_slot(dec, dtype, early_term=True)


@pytest.mark.parametrize("dtype", ["fp16", "int8", "fp32"])
def test_slot_descrambled_matches_oracle(dec, dtype):
    """Descrambling (per-TB c_init) before combining, both combining paths
    (whole symbols; repetition and layer-mapped LLRs), two transmissions."""
    run_slot(dec, dtype, early_term=False, descramble=True)


def test_slot_packed_equals_one_cb_per_workgroup(dec):
    """Consecutive CBs of a TB decoded side by side in one workgroup (each
    stopping on its own syndrome) give the same outputs, iterations
    included, as one CB per workgroup."""
    from aerial_amd.slot import SlotDecoder, SlotTb
    print("[case] slot packed vs unpacked, fp16, early termination", flush=True)
    outs = []
    for pack in (True, False):
        cases = [Case(k, s, "fp16") for k, s in enumerate(SLOT)]
        sd = SlotDecoder(dec, pack=pack)
        res = []
        for tx, (ndi, rv) in enumerate(TX):
            tbs = []
            for c in cases:
                p, dev, scale = c.transmission(ndi, rv, tx)
                tbs.append(SlotTb(p, dev, c.harq, c.mb, max_iterations=ITERS, early_term=True, llr_scale=scale))
            sd.build(tbs).launch()
            torch.cuda.synchronize()
            if pack:
                assert sum(ln.n_items for ln in sd.launches) < sum(c.spec[2] for c in cases)
            res.append([(tb.cb_bits.clone(), tb.cb_crc.clone(), tb.tb_bytes.clone(), tb.tb_crc.clone(),
                         tb.cb_iterations.clone(), c.harq.clone()) for c, tb in zip(cases, tbs)])
        outs.append(res)
    for tx in range(len(TX)):
        for k, (a, b) in enumerate(zip(outs[0][tx], outs[1][tx])):
            for j, (x, y) in enumerate(zip(a, b)):
                assert torch.equal(x, y), (tx, SLOT[k], j)


@pytest.mark.parametrize("descramble", [False, True])
def test_slot_decodes_payload_and_passes_crc(dec, descramble):
    """At high SNR every TB passes its CRC and the TB bytes are the payload
    (descramble: the channel carries scrambled bits)."""
    from aerial_amd.slot import SlotDecoder, SlotTb
    specs = [(b, z, c, k, q, l, n, e, 8.0) for (b, z, c, k, q, l, n, e, _) in SLOT]
    print("[case] slot fp16 high SNR payload check descr %d" % descramble, flush=True)
    cases = [Case(k, s, "fp16", descramble) for k, s in enumerate(specs)]
    tbs = []
    for c in cases:
        p, dev, scale = c.transmission(1, 0, 0)
        tbs.append(SlotTb(p, dev, c.harq, c.mb, max_iterations=ITERS, early_term=True, llr_scale=scale,
                          descramble=descramble))
    SlotDecoder(dec).build(tbs).launch()
    torch.cuda.synchronize()
    for c, tb in zip(cases, tbs):
        assert not tb.cb_crc.cpu().numpy().any() and int(tb.tb_crc.item()) == 0, c.spec
        payload = np.unpackbits(tb.tb_bytes.cpu().numpy())[:c.tb.A]
        assert np.array_equal(payload, c.tb.a), c.spec


@pytest.mark.parametrize("descramble", [False, True])
def test_slot_equals_unfused_pipeline(dec, descramble):
    """Fused result == de-rate-match kernel + TB decode + CRC kernels."""
    from aerial_amd import _lib as L
    from aerial_amd.decoder import LdpcDecodeConfig
    from aerial_amd.pusch import CrcDecode, DeRateMatch
    from aerial_amd.slot import SlotDecoder, SlotTb
    print("[case] slot fused vs unfused fp16 descr %d" % descramble, flush=True)
    cases = [Case(k, s, "fp16", descramble) for k, s in enumerate(SLOT)]
    harq2 = [c.harq.clone() for c in cases]
    rm = DeRateMatch(fp_config=3, descrambling=descramble)
    crc = CrcDecode(reverse_bytes=True)
    sd = SlotDecoder(dec)
    for tx, (ndi, rv) in enumerate(TX):
        tbs, params, inputs = [], [], []
        for c in cases:
            p, dev, scale = c.transmission(ndi, rv, tx)
            p.userGroupIndex = len(params)  # the RM API takes TB t's LLRs from input tensor t
            tbs.append(SlotTb(p, dev, c.harq, c.mb, max_iterations=ITERS, llr_scale=scale, descramble=descramble))
            params.append(p)
            inputs.append(dev)
        sd.build(tbs).launch()
        rm.run(params, inputs, harq2)
        ncb = sum(c.spec[2] for c in cases)
        words = torch.zeros((ncb, L.MAX_WORDS_PER_CODE_BLOCK), dtype=torch.int32, device="cuda")
        off = 0
        for c, tb, h in zip(cases, tbs, harq2):
            bg, Zc, C = c.spec[:3]
            cfg = LdpcDecodeConfig(bg, Zc, c.mb, llr_type=L.CUPHY_R_16F, max_iterations=ITERS)
            dec.decode_tb(cfg, [h.view(C, -1)], [words[off:off + C]])
            off += C
        cb_crc = torch.zeros(ncb, dtype=torch.int32, device="cuda")
        tb_crc = torch.zeros(len(cases), dtype=torch.int32, device="cuda")
        tb_bytes = torch.zeros(ncb * 1056 + 64, dtype=torch.uint8, device="cuda")
        crc.run(params, words, cb_crc, tb_bytes, tb_crc)
        torch.cuda.synchronize()
        off = boff = 0
        for t, (c, tb, h) in enumerate(zip(cases, tbs, harq2)):
            C = c.spec[2]
            assert torch.equal(c.harq, h), c.spec
            assert torch.equal(tb.cb_bits, words[off:off + C]), c.spec
            assert torch.equal(tb.cb_crc, cb_crc[off:off + C]), c.spec
            assert int(tb.tb_crc.item()) == int(tb_crc[t].item()), c.spec
            nb = tb.tb_bytes.numel()
            assert torch.equal(tb.tb_bytes, tb_bytes[boff:boff + nb]), c.spec
            off += C
            boff += nb
    rm.close()
    crc.close()


def test_slot_graph_replay(dec):
    """A captured slot replays with identical results (fresh transmission
    each replay: ndi = 1), with eager launches in between."""
    from aerial_amd.slot import SlotDecoder, SlotTb
    print("[case] slot graph capture + 3 replays", flush=True)
    cases = [Case(k, s, "fp16") for k, s in enumerate(SLOT)]
    tbs = []
    for c in cases:
        p, dev, scale = c.transmission(1, 0, 0)
        tbs.append(SlotTb(p, dev, c.harq, c.mb, max_iterations=ITERS, llr_scale=scale))
    sd = SlotDecoder(dec).build(tbs)
    sd.launch()
    torch.cuda.synchronize()
    def harq_region(c):
        C, Zc = c.spec[2], c.spec[1]
        return c.harq.view(C, -1)[:, :2 * Zc + c.p.Ncb]

    def bits_region(c, tb):  # the decoded words of each CB's 264-word row
        return tb.cb_bits[:, :(P.nsys(c.spec[0]) * c.spec[1] + 31) // 32]

    ref = [(bits_region(c, tb).clone(), tb.cb_crc.clone(), tb.tb_crc.clone(), tb.tb_bytes.clone(),
            harq_region(c).clone()) for c, tb in zip(cases, tbs)]
    sd.capture()
    for rep in range(3):
        for tb, c in zip(tbs, cases):
            tb.cb_bits.fill_(-1)
            tb.cb_crc.fill_(-1)
            tb.tb_crc.fill_(-1)
            c.harq.fill_(3.0)
        torch.cuda.synchronize()
        sd.replay()
        torch.cuda.synchronize()
        for (b, cc, tc, by, h), tb, c in zip(ref, tbs, cases):
            assert torch.equal(bits_region(c, tb), b) and torch.equal(tb.cb_crc, cc) and torch.equal(tb.tb_crc, tc)
            assert torch.equal(tb.tb_bytes, by) and torch.equal(harq_region(c), h)
        assert all(int(w[:2].abs().sum().item()) == 0 for w in sd.workspaces)
        sd.launch()
        torch.cuda.synchronize()


def test_slot_rejects_cb_beyond_variant(dec):
    """A CB whose mb exceeds the launched variant's row capacity is not
    decoded: its CRC and its TB CRC carry the rejection marker; so are the
    CBs of a grouped item launched on the one-CB kernel."""
    import ctypes as C
    from aerial_amd import _lib as L
    from aerial_amd.slot import SlotDecoder, SlotTb
    print("[case] slot rejection marker", flush=True)
    c = Case(0, (1, 104, 2, 2200, 2, 1, 1, 3000, 0.0), "fp16")
    p, dev, scale = c.transmission(1, 0, 0)
    assert c.mb > 8
    tb = SlotTb(p, dev, c.harq, c.mb, max_iterations=2, llr_scale=scale)
    sd = SlotDecoder(dec).build([tb])
    ln = sd.launches[0]
    st = L.lib().cuphyHipPuschSlotDecode(dec.dec, C.c_void_p(sd._jobs.data_ptr()), 1, C.c_void_p(ln.items.data_ptr()),
                                         ln.n_items, L.CUPHY_R_16F, ln.max_values, 8, ln.max_group,
                                         C.c_void_p(ln.workspace.data_ptr()), None)
    L.check("cuphyHipPuschSlotDecode", st)
    torch.cuda.synchronize()
    assert (tb.cb_crc.cpu().numpy().view(np.uint32) == L.SLOT_CB_ERROR).all()
    assert int(tb.tb_crc.cpu().numpy().view(np.uint32)[0]) == L.SLOT_CB_ERROR
    assert int(ln.workspace[:2].abs().sum().item()) == 0
    # an item of two CBs given to the one-CB-per-workgroup kernel
    # (max_cbs_per_item = 1) is rejected as a whole, not half decoded
    assert ln.n_items == 1 and ln.max_group == 2
    tb.cb_crc.zero_()
    tb.tb_crc.zero_()
    st = L.lib().cuphyHipPuschSlotDecode(dec.dec, C.c_void_p(sd._jobs.data_ptr()), 1, C.c_void_p(ln.items.data_ptr()),
                                         ln.n_items, L.CUPHY_R_16F, ln.max_values, 46, 1,
                                         C.c_void_p(ln.workspace.data_ptr()), None)
    L.check("cuphyHipPuschSlotDecode", st)
    torch.cuda.synchronize()
    assert (tb.cb_crc.cpu().numpy().view(np.uint32) == L.SLOT_CB_ERROR).all()
    assert int(tb.tb_crc.cpu().numpy().view(np.uint32)[0]) == L.SLOT_CB_ERROR
    assert int(ln.workspace.abs().sum().item()) == 0
