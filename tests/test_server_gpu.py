"""HARQ serving engine: continuous batching with device-resident HARQ
buffers.  Each TB's results in a mix of overlapping arrivals (concurrent
launches, retransmissions combined across launches) are bit-identical to
the oracle's chain over the same transmissions (binary16 de-rate-matching
with HARQ combining at rv 0, 2, 3, 1 -> layered decoding with syndrome early
termination -> CB/TB CRC) and to decoding that TB alone, transmission by
transmission, with the fused slot decoder; admission by HBM budget holds the
pool to its size."""
import numpy as np
import pytest

import oracle as O
import pusch_vectors as P

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from aerial_amd.decoder import LdpcDecoder
    d = LdpcDecoder()
    yield d
    d.close()


@pytest.fixture(scope="module")
def templates(dec):
    from aerial_amd.server import TraceTemplates, serving_specs
    return TraceTemplates(serving_specs(24, seed=31), seed=31)


def alone(dec, templates, ti, n_tx, max_iterations):
    """Decode template ti's first n_tx transmissions alone, in order."""
    from aerial_amd.server import RV_SEQUENCE
    from aerial_amd.slot import SlotDecoder, SlotTb
    tb = templates.tbs[ti]
    p0, _ = tb.params(0, 1)
    harq = torch.zeros(tb.C * p0.Ncb_padded, dtype=torch.float16, device="cuda")
    outs = SlotTb.alloc_outputs(p0)
    sd = SlotDecoder(dec)
    for k in range(n_tx):
        rv = RV_SEQUENCE[k]
        p, _ = tb.params(rv, 1 if k == 0 else 0)
        st = SlotTb(p, tb.llrs[rv], harq, templates.mbs[ti][k], max_iterations=max_iterations, early_term=True,
                    outputs=outs)
        sd.build([st]).launch()
    torch.cuda.synchronize()
    return outs


def check_against_alone(dec, templates, done, meta, outs, arrivals, max_iterations):
    from aerial_amd import _lib as L
    n_tx = {}
    for i, (a, k, ti) in enumerate(meta):
        if done[i].status == L.CUPHY_STATUS_SUCCESS:
            n_tx[a] = max(n_tx.get(a, 0), k + 1)
    assert len(n_tx) == len(arrivals), "every TB got at least its first transmission"
    for a, (_, _, ti) in enumerate(arrivals):
        ref = alone(dec, templates, ti, n_tx[a], max_iterations)
        tb = templates.tbs[ti]
        nw = (tb.Kp + 31) // 32
        got = outs[a]
        assert torch.equal(got[0][:, :nw], ref[0][:, :nw]), (a, ti)
        for j in (1, 2, 4):  # CB CRCs, TB bytes, iterations per CB
            assert torch.equal(got[j], ref[j]), (a, ti, j)
        # the completion's TB CRC is the one the alone run produced last
        last = 4 * a + n_tx[a] - 1
        assert done[last].tb_crc == int(ref[3].cpu().numpy().view(np.uint32)[0]), (a, ti)


def oracle_chain(tb, n_tx, max_iterations):
    """The oracle over a TB's first n_tx transmissions: binary16 HARQ
    combining (derate_match.m) of each transmission's LLRs at rv 0, 2, 3, 1,
    then every CB decoded in binary16 with syndrome early termination at the
    parity nodes covering what was received so far and the reference's
    normalisation for them, then CB/TB CRC and TB bytes.  k0 and the parity
    nodes come from the oracle.  Returns (words (C, 264), iterations,
    cb_crc, tb_bytes, tb_crc) of the last transmission."""
    from aerial_amd.server import RV_SEQUENCE
    bg, Zc, C, Kp, Qm, Nl = tb.bg, tb.Zc, tb.C, tb.Kp, tb.Qm, tb.Nl
    K = P.nsys(bg) * Zc
    N = (66 if bg == 1 else 50) * Zc
    npad = (N + 2 * Zc + 7) // 8 * 8
    harq = np.zeros(C * npad, np.float32)
    mb = 4
    for k in range(n_tx):
        rv = RV_SEQUENCE[k]
        k0 = O.k0(bg, rv, N, Zc)
        m = O.num_parity_nodes(bg, Zc, K, K - Kp, tb.G, C, k0)
        mb = m if k == 0 else max(mb, m)
        cfg = O.rm_cfg(bg=bg, Zc=Zc, C=C, Qm=Qm, Nl=Nl, G=tb.G, K=K, F=K - Kp, Ncb=N, Ncb_padded=npad, k0=k0,
                       ndi=int(k == 0))
        llr = tb.llrs[rv].view(-1, P.QAM_STRIDE)[:, :Qm].reshape(-1).float().cpu().numpy()
        O.derate_match(cfg, llr, harq, fp16=True)
    dcfg = O.dec_cfg(bg, Zc, mb, n_iter=max_iterations, alpha=P.table_alpha(bg, mb), clamp=32.0, mode=O.FP16,
                     early_term=True)
    h = harq.reshape(C, npad)[:, :(P.nsys(bg) + mb) * Zc]
    words, iters = O.decode_cbs_it(dcfg, h, nthreads=8, hard_stride=O.WORDS_PER_CB)
    cb_crc, tb_bytes, tb_crc = O.crc_desegment(words.reshape(-1), C, Kp)
    return words, iters, cb_crc, tb_bytes, tb_crc


def check_against_oracle(templates, done, meta, outs, arrivals, max_iterations):
    """Every TB's final outputs (decoded words, iterations per CB, CB CRCs,
    TB bytes, the completion's TB CRC) equal the oracle's chain over the
    transmissions the TB received."""
    from aerial_amd import _lib as L
    n_tx = {}
    for i, (a, k, ti) in enumerate(meta):
        if done[i].status == L.CUPHY_STATUS_SUCCESS:
            n_tx[a] = max(n_tx.get(a, 0), k + 1)
    assert len(n_tx) == len(arrivals)
    for a, (_, _, ti) in enumerate(arrivals):
        tb = templates.tbs[ti]
        words, iters, cb_crc, tb_bytes, tb_crc = oracle_chain(tb, n_tx[a], max_iterations)
        got = outs[a]
        nw = (tb.Kp + 31) // 32
        assert np.array_equal(got[0].cpu().numpy().view(np.uint32)[:, :nw], words[:, :nw]), (a, ti)
        assert np.array_equal(got[4].cpu().numpy().astype(np.int64), iters.astype(np.int64)), (a, ti)
        assert np.array_equal(got[1].cpu().numpy().view(np.uint32), cb_crc), (a, ti)
        assert np.array_equal(got[2].cpu().numpy()[:len(tb_bytes)], tb_bytes), (a, ti)
        assert done[4 * a + n_tx[a] - 1].tb_crc == tb_crc, (a, ti)


@pytest.mark.parametrize("resident", [False, True])
def test_results_equal_alone_and_in_the_mix(dec, templates, resident):
    from aerial_amd.server import HarqServer, build_trace, poisson_arrivals
    print("[case] server: 40 overlapping TBs, rtt 300us, resident=%s" % resident, flush=True)
    srv = HarqServer(dec, pool_bytes=1 << 30, max_inflight=3, max_tbs=64, max_cbs=1024, resident=resident)
    arrivals = poisson_arrivals(n_cells=8, rate_per_s=200000.0, duration_us=600.0, n_templates=len(templates.tbs),
                                seed=5)
    arrivals = arrivals[:40]
    assert len(arrivals) == 40
    entries, outs, meta = build_trace(srv, templates, arrivals, max_iterations=8)
    done = srv.serve_trace(entries, harq_rtt_us=300.0)
    st = srv.stats()
    assert st["pending"] == 0 and st["pool_bytes_in_use"] == 0, st
    assert st["launches"] >= (1 if resident else 2)
    check_against_oracle(templates, done, meta, outs, arrivals, 8)
    check_against_alone(dec, templates, done, meta, outs, arrivals, 8)
    # some TBs needed retransmissions, and overlapping launches happened
    ks = [k for i, (a, k, ti) in enumerate(meta) if done[i].status == 0]
    assert max(ks) >= 1
    srv.close()


def test_resident_frequent_idle_exits(dec, templates):
    """A resident launch that leaves 100 us after its last work, fed TBs
    with gaps of that order: the launch leaves and restarts many times, its
    workgroups handing their unfilled tickets to the next launch, while TBs
    keep arriving.  Every transmission completes exactly once and every TB
    equals the oracle's chain and its decode alone."""
    from aerial_amd.server import HarqServer, build_trace, poisson_arrivals
    print("[case] server: resident launch with 100 us idle exit, 40 TBs over 4 ms", flush=True)
    srv = HarqServer(dec, pool_bytes=1 << 30, max_tbs=64, max_cbs=1024, resident=True, resident_idle_us=100.0)
    arrivals = poisson_arrivals(n_cells=4, rate_per_s=10000.0, duration_us=6000.0, n_templates=len(templates.tbs),
                                seed=9)[:40]
    assert len(arrivals) == 40
    entries, outs, meta = build_trace(srv, templates, arrivals, max_iterations=8)
    done = srv.serve_trace(entries, harq_rtt_us=250.0)
    st = srv.stats()
    srv.close()
    assert st["launches"] >= 4, st
    assert st["completed"] == st["submitted"] and st["pending"] == 0 and st["pool_bytes_in_use"] == 0, st
    released = [i for i in range(len(meta)) if done[i].status == 0]
    assert len(set(done[i].tag for i in released)) == len(released)
    check_against_oracle(templates, done, meta, outs, arrivals, 8)
    check_against_alone(dec, templates, done, meta, outs, arrivals, 8)


@pytest.mark.parametrize("resident", [False, True])
def test_tb_completes_before_its_launch_ends(dec, resident):
    # a one-CB TB and a 24-CB low-SNR TB in the same launch: the small TB's
    # CRC is written by the kernel as soon as its own CB is decoded, and the
    # engine completes it while the large TB is still decoding
    from aerial_amd.server import HarqServer, TraceTemplates
    from aerial_amd.slot import SlotTb
    print("[case] server: per-TB completion, BG2 Z=64 1 CB + BG1 Z=384 24 CBs in one launch, resident=%s"
          % resident, flush=True)
    specs = [(2, 64, 1, 640, 2, 1, 1280, 8.0), (1, 384, 24, 8448, 2, 1, 42240, -1.0)]
    tt = TraceTemplates(specs, seed=3)
    srv = HarqServer(dec, pool_bytes=1 << 30, max_inflight=2, max_tbs=16, max_cbs=256, resident=resident)
    outs = []
    for i, tb in enumerate(tt.tbs):
        p, _ = tb.params(0, 1)
        o = SlotTb.alloc_outputs(p)
        outs.append(o)
        srv.submit(srv.request(100 + i, p, tb.llrs[0], o, tt.mbs[i][0], max_iterations=10, early_term=True,
                               last=True, tag=i))
    done = {}
    while len(done) < 2:
        for c in srv.poll():
            done[c.tag] = c
    st = srv.stats()
    srv.close()
    assert st["launches"] == 1, st
    small, big = done[0], done[1]
    assert small.status == 0 and big.status == 0
    assert small.t_launch_us == big.t_launch_us
    assert small.t_done_us < big.t_done_us, (small.t_done_us, big.t_done_us)
    # the one-CB TB's CB CRC is its TB CRC
    assert int(outs[0][1].cpu()[0]) == small.tb_crc == 0


@pytest.mark.parametrize("resident", [False, True])
def test_admission_by_hbm_budget(dec, templates, resident):
    """A pool that holds about two TBs still serves a burst of ten; its peak
    use never exceeds its size and new TBs waited for space."""
    from aerial_amd.server import HarqServer, build_trace
    print("[case] server: burst of 10 TBs into a small pool, resident=%s" % resident, flush=True)
    sizes = []
    for tb in templates.tbs[:10]:
        p, _ = tb.params(0, 1)
        sizes.append((p.firstCodeBlockIndex + p.num_CBs) * p.Ncb_padded * 2)
    pool = (max(sizes) * 2 + 255) // 256 * 256 + 512
    srv = HarqServer(dec, pool_bytes=pool, max_inflight=2, max_tbs=64, max_cbs=1024, resident=resident)
    arrivals = [(0.0, 0, i) for i in range(10)]
    entries, outs, meta = build_trace(srv, templates, arrivals, max_iterations=8)
    done = srv.serve_trace(entries, harq_rtt_us=50.0)
    st = srv.stats()
    assert st["pool_bytes_peak"] <= pool, st
    assert st["admission_waits"] > 0, st
    assert st["pool_bytes_in_use"] == 0
    check_against_oracle(templates, done, meta, outs, arrivals, 8)
    check_against_alone(dec, templates, done, meta, outs, arrivals, 8)
    srv.close()


@pytest.mark.parametrize("resident", [False, True])
def test_retransmission_waits_for_its_predecessor(dec, templates, resident):
    """Two transmissions of one HARQ process submitted back to back run in
    order (the second combines into the first's buffer)."""
    from aerial_amd.server import HarqServer, RV_SEQUENCE
    from aerial_amd.slot import SlotTb
    print("[case] server: back-to-back submit of tx 0 and tx 1, resident=%s" % resident, flush=True)
    # a TB whose first transmission fails (a passing TB's buffer is freed
    # and its retransmission rejected, as the next test checks)
    ti = next((i for i in range(len(templates.tbs))
               if int(alone(dec, templates, i, 1, 8)[3].cpu().numpy().view(np.uint32)[0]) != 0), None)
    if ti is None:
        pytest.skip("no template fails its first transmission")
    tb = templates.tbs[ti]
    srv = HarqServer(dec, pool_bytes=1 << 28, max_inflight=3, max_tbs=8, max_cbs=256, resident=resident)
    p0, _ = tb.params(0, 1)
    outs = SlotTb.alloc_outputs(p0)
    for k in range(2):
        p, _ = tb.params(RV_SEQUENCE[k], 1 if k == 0 else 0)
        srv.submit(srv.request(77, p, tb.llrs[RV_SEQUENCE[k]], outs, templates.mbs[ti][k], max_iterations=8,
                               tag=k, last=(k == 1)))
    got = []
    while len(got) < 2:
        got += srv.poll()
    assert [c.tag for c in got] == [0, 1]
    torch.cuda.synchronize()
    ref = alone(dec, templates, ti, 2, 8)
    nw = (tb.Kp + 31) // 32
    assert torch.equal(outs[0][:, :nw], ref[0][:, :nw])
    assert got[1].tb_crc == int(ref[3].cpu().numpy().view(np.uint32)[0])
    assert srv.poll() == []  # launches ended (synchronized): their buffers are released
    assert srv.stats()["pool_bytes_in_use"] == 0  # LAST_TX freed the buffer
    srv.close()


def test_resident_launch_restarts_after_idle(dec, templates):
    """The resident launch leaves after 20 ms without work, its workgroups
    leaving their tickets behind; work submitted afterwards starts a new
    launch that takes those tickets over, and every TB still decodes exactly
    as alone."""
    import time
    from aerial_amd.server import HarqServer
    from aerial_amd.slot import SlotTb
    print("[case] server: resident launch idle exit and restart", flush=True)
    srv = HarqServer(dec, pool_bytes=1 << 28, max_tbs=16, max_cbs=512, resident=True)
    results, crcs = [], {}
    for wave_, tis in enumerate(([0, 1, 2], [3, 4], [5])):
        for ti in tis:
            tb = templates.tbs[ti]
            p0, _ = tb.params(0, 1)
            outs = SlotTb.alloc_outputs(p0)
            srv.submit(srv.request(200 + ti, p0, tb.llrs[0], outs, templates.mbs[ti][0], max_iterations=8,
                                   last=True, tag=ti))
            results.append((ti, outs))
        got = []
        while len(got) < len(tis):
            got += srv.poll()
        assert all(c.status == 0 for c in got)
        crcs.update({c.tag: c.tb_crc for c in got})
        t_end = time.time() + 0.06  # idle: the launch leaves after 20 ms
        while time.time() < t_end:
            srv.poll()
            time.sleep(0.001)
    st = srv.stats()
    srv.close()
    assert st["launches"] == 3, st
    assert st["pool_bytes_in_use"] == 0
    for ti, outs in results:
        ref = alone(dec, templates, ti, 1, 8)
        nw = (templates.tbs[ti].Kp + 31) // 32
        assert torch.equal(outs[0][:, :nw], ref[0][:, :nw]), ti
        assert torch.equal(outs[1], ref[1]) and torch.equal(outs[2], ref[2]), ti
        assert crcs[ti] == int(ref[3].cpu().numpy().view(np.uint32)[0]), ti


@pytest.mark.parametrize("resident", [False, True])
def test_harq_buffer_timeout_eviction(dec, templates, resident):
    """A HARQ buffer left unused for harq_timeout_us is freed; a later
    retransmission of that process is rejected."""
    import time
    from aerial_amd import _lib as L
    from aerial_amd.server import HarqServer, RV_SEQUENCE
    from aerial_amd.slot import SlotTb
    print("[case] server: HARQ buffer timeout 2 ms, resident=%s" % resident, flush=True)
    ti = next((i for i in range(len(templates.tbs))
               if int(alone(dec, templates, i, 1, 8)[3].cpu().numpy().view(np.uint32)[0]) != 0), None)
    if ti is None:
        pytest.skip("no template fails its first transmission")
    tb = templates.tbs[ti]
    srv = HarqServer(dec, pool_bytes=1 << 28, max_inflight=2, max_tbs=8, max_cbs=256, resident=resident,
                     harq_timeout_us=2000.0)
    p0, _ = tb.params(0, 1)
    outs = SlotTb.alloc_outputs(p0)
    srv.submit(srv.request(55, p0, tb.llrs[0], outs, templates.mbs[ti][0], max_iterations=8, tag=0))
    got = []
    while not got:
        got += srv.poll()
    assert got[0].tb_crc != 0
    if not resident:  # batched: the launch's buffers are released when it has ended
        torch.cuda.synchronize()
        srv.poll()
    assert srv.stats()["pool_bytes_in_use"] > 0  # kept for the retransmission
    t_end = time.time() + 0.02
    while time.time() < t_end:
        srv.poll()
        time.sleep(0.0005)
    st = srv.stats()
    assert st["evictions"] == 1 and st["pool_bytes_in_use"] == 0, st
    p1, _ = tb.params(RV_SEQUENCE[1], 0)
    srv.submit(srv.request(55, p1, tb.llrs[RV_SEQUENCE[1]], outs, templates.mbs[ti][1], max_iterations=8, tag=1))
    got = []
    while not got:
        got += srv.poll()
    assert got[0].tag == 1 and got[0].status == L.CUPHY_STATUS_INVALID_ARGUMENT
    srv.close()


@pytest.mark.parametrize("resident", [False, True])
def test_descrambled_requests(dec, resident):
    """Requests with CUPHY_HIP_SLOT_TB_DESCRAMBLE: the channel carries
    scrambled bits (per-TB c_init); every TB passes its CRC and its outputs
    equal the fused slot decoder's alone.  In resident mode a CB whose
    sequence words exceed the launch's LDS budget is rejected at submit."""
    import oracle as O
    from aerial_amd import _lib as L
    from aerial_amd.server import HarqServer, TraceTemplates
    from aerial_amd.slot import SlotDecoder, SlotTb
    print("[case] server: descrambled requests, resident=%s" % resident, flush=True)
    specs = [(1, 384, 4, 8448, 8, 1, 9088, 12.0), (2, 104, 1, 1040, 4, 1, 2080, 8.0),
             (1, 208, 3, 4576, 2, 1, 9152, 8.0), (2, 64, 4, 640, 2, 1, 1280, 8.0)]  # last: one item of 4 CBs
    tt = TraceTemplates(specs, seed=11)
    srv = HarqServer(dec, pool_bytes=1 << 30, max_inflight=2, max_tbs=16, max_cbs=256, resident=resident)
    reqs = []
    for i, tb in enumerate(tt.tbs):
        p, _ = tb.params(0, 1)
        p.cinit = (0x5A5A5A5 + 7919 * i) & 0x7FFFFFFF
        c = O.gold(p.cinit, 0, tb.G).reshape(-1, tb.Qm)
        sign = torch.ones((tb.G // tb.Qm, 8), dtype=torch.float16, device="cuda")
        sign[:, :tb.Qm] = torch.from_numpy(1.0 - 2.0 * c.astype(np.float32)).to(sign)
        llr = (tb.llrs[0].view(-1, 8) * sign).reshape(-1).contiguous()
        reqs.append((p, llr))
        o = SlotTb.alloc_outputs(p)
        srv.submit(srv.request(300 + i, p, llr, o, tt.mbs[i][0], max_iterations=10, early_term=True, last=True,
                               tag=i, descramble=True))
        reqs[-1] += (o,)
    if resident:
        # a server that is never polled (nothing it accepted would launch)
        from aerial_amd.slot import lds_values
        chk = HarqServer(dec, pool_bytes=1 << 26, max_inflight=1, max_tbs=4, max_cbs=16, resident=True)
        p, llr, o = reqs[1]
        r = chk.request(399, p, llr, o, tt.mbs[1][0], descramble=True)
        r.prm.encodedSize = r.prm.G = 1000000  # 31,252 sequence words for its one CB
        # beyond the resident launch's LDS: one 68 x 384 array and 2048 words
        assert lds_values(r.prm, tt.mbs[1][0], True) > 68 * 384 + 2 * 2048
        rc = L.lib().cuphyHipHarqServerSubmit(chk.h, r)
        assert chk.stats()["submitted"] == 0
        chk.close()
        assert rc == L.CUPHY_STATUS_INVALID_ARGUMENT
    done = {}
    while len(done) < len(reqs):
        for c in srv.poll():
            done[c.tag] = c
    srv.close()
    for i, (p, llr, o) in enumerate(reqs):
        assert done[i].status == 0 and done[i].tb_crc == 0, (i, done[i].status, done[i].tb_crc)
        harq = torch.zeros(tt.tbs[i].C * p.Ncb_padded, dtype=torch.float16, device="cuda")
        ref = SlotTb.alloc_outputs(p)
        SlotDecoder(dec).build([SlotTb(p, llr, harq, tt.mbs[i][0], max_iterations=10, early_term=True,
                                       outputs=ref, descramble=True)]).launch()
        torch.cuda.synchronize()
        nw = (tt.tbs[i].Kp + 31) // 32
        assert torch.equal(o[0][:, :nw], ref[0][:, :nw]), i
        for j in (1, 2, 4):
            assert torch.equal(o[j], ref[j]), (i, j)
        assert int(ref[3].item()) == 0, i


def test_invalid_requests_rejected(dec, templates):
    from aerial_amd import _lib as L
    from aerial_amd.server import HarqServer
    from aerial_amd.slot import SlotTb
    print("[case] server: request validation", flush=True)
    srv = HarqServer(dec, pool_bytes=1 << 24, max_inflight=1, max_tbs=4, max_cbs=64)
    tb = templates.tbs[0]
    p, _ = tb.params(0, 1)
    outs = SlotTb.alloc_outputs(p)
    r = srv.request(1, p, tb.llrs[0], outs, 4)
    r.num_parity_nodes = 3  # below the base graph's 4 core rows
    assert L.lib().cuphyHipHarqServerSubmit(srv.h, r) == L.CUPHY_STATUS_INVALID_ARGUMENT
    with pytest.raises(L.CuphyError):
        srv.release(12345)
    # a retransmission of an unknown (or already decoded) HARQ process
    # completes with an error
    p1, _ = tb.params(2, 0)
    srv.submit(srv.request(999, p1, tb.llrs[2], outs, templates.mbs[0][1], tag=5))
    got = []
    while not got:
        got = srv.poll()
    assert got[0].tag == 5 and got[0].status == L.CUPHY_STATUS_INVALID_ARGUMENT
    srv.close()
