"""Persistent heterogeneous batch decoder (one launch, device work queue):
bit-exact against the oracle, independent of item order, replayable from a
captured HIP graph, and leaves its queue state at zero."""
from types import SimpleNamespace

import numpy as np
import pytest

import oracle as O
import vectors as V

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MIX = [(1, 384, 46, 3), (2, 64, 42, 2), (1, 72, 4, 5), (2, 208, 9, 1), (1, 2, 20, 2), (2, 15, 4, 3),
       (1, 352, 12, 2), (2, 384, 4, 4), (1, 160, 33, 1), (2, 96, 25, 2)]


@pytest.fixture(scope="module")
def dec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from aerial_amd.decoder import LdpcDecoder
    d = LdpcDecoder()
    yield d
    d.close()


def make_tbs(dtype, seed=50, mix=MIX):
    tbs, refs = [], []
    mode = O.FP32 if dtype == "fp32" else O.FP16
    for k, (bg, Z, mb, ncb) in enumerate(mix):
        llr, _ = V.make_cbs(bg, Z, mb, ncb, seed=seed + k, snr_db=1.5)
        if dtype != "fp32":
            llr = V.to_half_rn(llr).astype(np.float32)
        stride = (llr.shape[1] + 7) // 8 * 8
        host = np.zeros((ncb, stride), np.float32)
        host[:, :llr.shape[1]] = llr
        t = torch.from_numpy(host).cuda()
        if dtype == "fp16":
            t = t.half()
        nw = (V.nsys(bg) * Z + 31) // 32
        out = torch.full((ncb, nw), -1, dtype=torch.int32, device="cuda")
        tbs.append(SimpleNamespace(bg=bg, Z=Z, mb=mb, ncb=ncb, llr=t, out=out))
        cfg = O.dec_cfg(bg, Z, mb, n_iter=7, alpha=0.8125, clamp=32.0, mode=mode)
        refs.append(np.stack([O.decode_cb(cfg, llr[i])[0] for i in range(ncb)]))
    return tbs, refs


def check(tbs, refs):
    for tb, ref in zip(tbs, refs):
        assert np.array_equal(tb.out.cpu().numpy().view(np.uint32), ref), (tb.bg, tb.Z, tb.mb)


@pytest.mark.parametrize("dtype", ["fp32", "fp16", "fp32in_fp16math"])
def test_batch_bit_exact(dec, dtype):
    from aerial_amd import _lib as L
    from aerial_amd.batch import BatchDecoder
    print("[case] batch of %d TBs %s" % (len(MIX), dtype), flush=True)
    tbs, refs = make_tbs(dtype)
    llr_type = L.CUPHY_R_16F if dtype == "fp16" else L.CUPHY_R_32F
    b = BatchDecoder(dec).build(tbs, llr_type=llr_type, fp16_math=(dtype != "fp32"), max_iterations=7)
    b.launch()
    torch.cuda.synchronize()
    check(tbs, refs)
    assert all(int(w.abs().sum().item()) == 0 for w in b.workspaces), "queue state not returned to zero"


@pytest.mark.parametrize("dtype", ["fp32", "fp16"])
def test_batch_row_pair_equals_one_lane_group(dec, dtype):
    """Row-pair workgroups (768 threads, two lane groups on the independent
    rows of each barrier segment) give the hard and soft outputs of the
    384-thread decode, with early termination (per-CB stopping) and without."""
    from aerial_amd import _lib as L
    from aerial_amd.batch import BatchDecoder
    llr_type = L.CUPHY_R_16F if dtype == "fp16" else L.CUPHY_R_32F
    for et in (0, L.CUPHY_LDPC_DECODE_EARLY_TERM):
        print("[case] batch row-pair vs one lane group %s et=%d" % (dtype, et), flush=True)
        outs = []
        for pair in (True, False):
            tbs, _ = make_tbs(dtype, seed=70)
            soft = [torch.zeros((tb.ncb, V.nsys(tb.bg) * tb.Z), dtype=torch.float16, device="cuda") for tb in tbs]
            b = BatchDecoder(dec, pair=pair).build(tbs, llr_type=llr_type, fp16_math=(dtype != "fp32"),
                                                   max_iterations=7, flags=et, soft=soft)
            assert (max(sb.block for sb in b.subs) == 768) == pair
            b.launch()
            torch.cuda.synchronize()
            outs.append([(tb.out.clone(), s_.clone()) for tb, s_ in zip(tbs, soft)])
            assert all(int(w.abs().sum().item()) == 0 for w in b.workspaces)
        for k, ((h1, s1), (h2, s2)) in enumerate(zip(*outs)):
            assert torch.equal(h1, h2), (MIX[k], et)
            assert torch.equal(s1.view(torch.int16), s2.view(torch.int16)), (MIX[k], et)


def test_batch_order_independent(dec):
    from aerial_amd import _lib as L
    from aerial_amd.batch import BatchDecoder
    print("[case] batch order fifo vs lpt fp16", flush=True)
    tbs, refs = make_tbs("fp16", seed=70)
    BatchDecoder(dec).build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=7, order="fifo").launch()
    torch.cuda.synchronize()
    check(tbs, refs)
    for tb in tbs:
        tb.out.fill_(-1)
    BatchDecoder(dec).build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=7, order="lpt").launch()
    torch.cuda.synchronize()
    check(tbs, refs)
    for tb in tbs:
        tb.out.fill_(-1)
    BatchDecoder(dec, split_by_mb=False).build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=7).launch()
    torch.cuda.synchronize()
    check(tbs, refs)
    for tb in tbs:
        tb.out.fill_(-1)
    b = BatchDecoder(dec, split_min=1, pack=False).build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True,
                                                          max_iterations=7)
    assert len(b.subs) == 2  # high-rate CBs on the register-lean variant, concurrently
    b.launch()
    torch.cuda.synchronize()
    check(tbs, refs)


def test_batch_graph_replay(dec):
    from aerial_amd import _lib as L
    from aerial_amd.batch import BatchDecoder
    print("[case] batch graph capture + 3 replays fp16", flush=True)
    tbs, refs = make_tbs("fp16", seed=90)
    b = BatchDecoder(dec).build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=7)
    b.capture()
    for rep in range(3):
        for tb in tbs:
            tb.out.fill_(-1)
        torch.cuda.synchronize()
        b.replay()
        torch.cuda.synchronize()
        check(tbs, refs)
        assert all(int(w.abs().sum().item()) == 0 for w in b.workspaces)
        # eager work between replays
        b.launch()
        torch.cuda.synchronize()
        check(tbs, refs)


def test_batch_single_and_empty(dec):
    from aerial_amd import _lib as L
    from aerial_amd.batch import BatchDecoder
    print("[case] batch single item and empty batch", flush=True)
    tbs, refs = make_tbs("fp32", seed=95, mix=[(1, 384, 46, 1)])
    b = BatchDecoder(dec).build(tbs, llr_type=L.CUPHY_R_32F, fp16_math=False, max_iterations=7)
    b.launch()
    torch.cuda.synchronize()
    check(tbs, refs)
    b.subs = []
    b.launch()  # no-op
    torch.cuda.synchronize()


# Packed groups: several code blocks of one (BG, Z, mb) per workgroup.
PACK_MIX = [(1, 2, 20, 70), (2, 15, 4, 30), (1, 36, 8, 12), (2, 64, 42, 9), (1, 72, 4, 7), (2, 104, 10, 5),
            (1, 176, 12, 4), (1, 384, 6, 2), (2, 7, 42, 25), (1, 320, 30, 2)]  # last: a row-pair launch


def _max_pack_table(mix):
    t = {}
    for bg, Z, mb, _ in mix:
        P = 384 // Z
        t["%d,%d,%s" % (bg, Z, "small" if mb <= 8 else "full")] = {"P": P, "block": 64 * (-(-P * Z // 64))}
    return t


def make_packed(seed, et, n_iter):
    tbs, refs = [], []
    for k, (bg, Z, mb, ncb) in enumerate(PACK_MIX):
        # SNR near the waterfall so that code blocks stop at different iterations
        llr, _ = V.make_cbs(bg, Z, mb, ncb, seed=seed + k, snr_db=2.0)
        llr = V.to_half_rn(llr).astype(np.float32)
        stride = (llr.shape[1] + 7) // 8 * 8
        host = np.zeros((ncb, stride), np.float32)
        host[:, :llr.shape[1]] = llr
        kbz = V.nsys(bg) * Z
        nw = (kbz + 31) // 32
        tb = SimpleNamespace(bg=bg, Z=Z, mb=mb, ncb=ncb, llr=torch.from_numpy(host).cuda().half(),
                             out=torch.full((ncb, nw), -1, dtype=torch.int32, device="cuda"),
                             soft=torch.zeros((ncb, kbz), dtype=torch.float16, device="cuda"))
        cfg = O.dec_cfg(bg, Z, mb, n_iter=n_iter, alpha=0.8125, clamp=32.0, mode=O.FP16, early_term=et)
        hs, ss, its = [], [], []
        for i in range(ncb):
            h, it, app = O.decode_cb(cfg, llr[i], want_app=True)
            hs.append(h)
            ss.append(app[:kbz].astype(np.float16))
            its.append(it)
        tbs.append(tb)
        refs.append((np.stack(hs), np.stack(ss), its))
    return tbs, refs


@pytest.mark.parametrize("et", [False, True])
@pytest.mark.parametrize("table", ["max_pack", "tuned"])
def test_batch_packed_bit_exact(dec, table, et):
    from aerial_amd import _lib as L
    from aerial_amd.batch import BatchDecoder
    print("[case] packed batch table=%s et=%s fp16 seed=300" % (table, et), flush=True)
    n_iter = 12
    tbs, refs = make_packed(300, et, n_iter)
    if et:
        spread = [len(set(r[2])) for r in refs]
        assert max(spread) > 1, "early termination did not vary within a TB"
    b = BatchDecoder(dec, split_min=1, table=_max_pack_table(PACK_MIX) if table == "max_pack" else None)
    b.build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=n_iter,
            flags=L.CUPHY_LDPC_DECODE_EARLY_TERM if et else 0, soft=[tb.soft for tb in tbs])
    if table == "max_pack":
        assert any(sb.n_items > sb.n_groups for sb in b.subs), "no code blocks were packed"
    b.launch()
    torch.cuda.synchronize()
    for tb, (h, sft, _) in zip(tbs, refs):
        assert np.array_equal(tb.out.cpu().numpy().view(np.uint32), h), (tb.bg, tb.Z, tb.mb)
        got = tb.soft.cpu().numpy()
        assert np.array_equal(got.view(np.uint16), sft.view(np.uint16)), (tb.bg, tb.Z, tb.mb)
    assert all(int(w.abs().sum().item()) == 0 for w in b.workspaces)


@pytest.mark.parametrize("et", [False, True])
def test_batch_lane_major_equals_column_major(dec, et):
    """One code block per group: a launch without a group list (the
    lane-major one-code-block kernel) and the same launch with a one-item
    group list (the column-major kernel) give identical hard and soft
    outputs, both equal to the oracle."""
    from aerial_amd import _lib as L
    from aerial_amd.batch import BatchDecoder
    print("[case] lane-major vs column-major one-CB launches et=%s fp16 seed=310" % et, flush=True)
    n_iter = 12
    outs = []
    for listed in (False, True):
        tbs, refs = make_packed(310, et, n_iter)
        b = BatchDecoder(dec, split_min=1, pack=False, pair=False)
        b.build(tbs, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=n_iter,
                flags=L.CUPHY_LDPC_DECODE_EARLY_TERM if et else 0, soft=[tb.soft for tb in tbs])
        for sb in b.subs:
            assert sb.block <= 384 and sb.n_items == sb.n_groups
            if listed:
                sb.group_first = torch.arange(sb.n_groups + 1, dtype=torch.int32, device="cuda")
            else:
                assert sb.group_first is None
        b.launch()
        torch.cuda.synchronize()
        for tb, (h, sft, _) in zip(tbs, refs):
            assert np.array_equal(tb.out.cpu().numpy().view(np.uint32), h), (listed, tb.bg, tb.Z, tb.mb)
            assert np.array_equal(tb.soft.cpu().numpy().view(np.uint16), sft.view(np.uint16)), (listed, tb.bg, tb.Z)
        assert all(int(w.abs().sum().item()) == 0 for w in b.workspaces)
        outs.append([(tb.out.cpu().numpy(), tb.soft.cpu().numpy()) for tb in tbs])
    for (h0, s0), (h1, s1) in zip(*outs):
        assert np.array_equal(h0, h1) and np.array_equal(s0.view(np.uint16), s1.view(np.uint16))


def test_full_load_slot_slice_matches_oracle(dec):
    """Two cells of the 20-cell full-load slot (280 CBs, BG1 Zc=384, R=0.93
    -> mb=4, consistent-Gaussian LLRs with mu from the cell's SNR, binary16,
    10 iterations), generated and batch-decoded exactly as the benchmark's
    full-load section does (same seed, same first cells, graph replay of the
    tuned throughput-regime launch): every code block equals the oracle's
    binary16 decode bit for bit, and the plain per-TB form gives the same
    words."""
    from aerial_amd import _lib as L
    from aerial_amd import workloads as W
    from aerial_amd.batch import BatchDecoder
    from aerial_amd.decoder import LdpcDecodeConfig
    print("[case] full-load slot slice: 2 cells x 140 CBs BG1 Zc=384 mb=4 fp16 10 iterations seed=3", flush=True)
    c3 = W.full_load_slot(seed=3, n_cells=2, cbs_per_cell=140)
    assert all((tb.bg, tb.Z, tb.mb, tb.ncb) == (1, 384, 4, 140) for tb in c3)
    b = BatchDecoder(dec).build(c3, llr_type=L.CUPHY_R_16F, fp16_math=True, max_iterations=10, alpha=0.8125)
    assert sum(tb.ncb for tb in c3) >= b.split_min  # throughput regime, as the 2,800-CB slot
    b.capture()
    for tb in c3:
        tb.out.fill_(-1)
    torch.cuda.synchronize()
    b.replay()
    torch.cuda.synchronize()
    batched = [tb.out.clone() for tb in c3]
    assert all(int(w.abs().sum().item()) == 0 for w in b.workspaces)
    for tb in c3:
        tb.out.fill_(-1)
        cfg = LdpcDecodeConfig(1, 384, 4, llr_type=L.CUPHY_R_16F, max_iterations=10, alpha=0.8125, clamp_value=32.0)
        dec.decode_tb(cfg, [tb.llr], [tb.out])
    torch.cuda.synchronize()
    ocfg = O.dec_cfg(1, 384, 4, n_iter=10, alpha=0.8125, clamp=32.0, mode=O.FP16)
    for k, (tb, got) in enumerate(zip(c3, batched)):
        assert torch.equal(tb.out, got), "plain and batched words differ in cell %d" % k
        llr = tb.llr.float().cpu().numpy()[:, :tb.num_llrs]
        ref = O.decode_cbs(ocfg, llr, nthreads=8)
        assert np.array_equal(got.cpu().numpy().view(np.uint32), ref), "cell %d differs from the oracle" % k


def _launch_items(dec, arr, first, max_values, max_mb, block, llr_type):
    """One cuphyHipLDPCBatchDecodeGroups launch over a host item array;
    returns the items as the device left them (status words)."""
    import ctypes as C
    from aerial_amd import _lib as L
    from aerial_amd.batch import ITEM_DTYPE
    items = torch.from_numpy(arr.view(np.uint8).copy()).cuda()
    gf = torch.from_numpy(np.asarray(first, np.uint32).view(np.int32).copy()).cuda() if first is not None else None
    ws = torch.zeros(64, dtype=torch.int32, device="cuda")
    n = (len(first) - 1) if first is not None else len(arr)
    L.check("cuphyHipLDPCBatchDecodeGroups", L.lib().cuphyHipLDPCBatchDecodeGroups(
        dec.dec, C.c_void_p(items.data_ptr()), C.c_void_p(gf.data_ptr() if gf is not None else None), n, llr_type,
        1, max_values, max_mb, block, C.c_void_p(ws.data_ptr()), None))
    torch.cuda.synchronize()
    assert int(ws.abs().sum().item()) == 0
    return items.cpu().numpy().view(ITEM_DTYPE)


@pytest.mark.parametrize("mode", ["grouped", "lane_major", "row_pair"])
def test_batch_rejects_invalid_items(dec, mode):
    """Invalid work items (BG not 1/2, Z not a lifting size, mb or Kb out of
    range, negative iterations, null buffers, another configuration inside a
    group, more rows than the launch's kernel holds) are not decoded: their
    outputs stay untouched and their status word names the reason, while the
    valid items of the same launch equal the oracle bit for bit."""
    from aerial_amd import _lib as L
    from aerial_amd.batch import ITEM_DTYPE, _half_bits, group_values
    print("[case] batch with invalid items, %s" % mode, flush=True)
    mix = [(1, 64, 10, 4), (2, 104, 6, 3), (1, 36, 20, 4)]
    tbs, refs = make_tbs("fp16", seed=61, mix=mix)
    h = _half_bits(0.8125)
    sink = torch.full((64, 512), -1, dtype=torch.int32, device="cuda")  # outputs of the invalid items

    def item(tb, c, out=None, **kw):
        a = np.zeros(1, ITEM_DTYPE)
        a["llr"] = tb.llr.data_ptr() + c * tb.llr.stride(0) * 2
        a["out"] = (tb.out.data_ptr() + c * tb.out.stride(0) * 4) if out is None else out
        a["BG"], a["Z"], a["num_parity_nodes"], a["Kb"] = tb.bg, tb.Z, tb.mb, 22 if tb.bg == 1 else 10
        a["max_iterations"], a["norm"], a["clamp_value"], a["status"] = 7, h | (h << 16), 32.0, 0xABCD
        for k, v in kw.items():
            a[k] = v
        return a

    sink_row = [0]

    def bad(tb, **kw):
        r = sink_row[0]
        sink_row[0] += 1
        out = kw.pop("out", sink.data_ptr() + r * 512 * 4)
        return item(tb, 0, out=out, **kw)

    A, B, Cc = tbs
    BAD = [(dict(BG=3), L.ITEM_BAD_BG | L.ITEM_BAD_KB), (dict(Z=17), L.ITEM_BAD_Z), (dict(Z=0), L.ITEM_BAD_Z),
           (dict(Z=500), L.ITEM_BAD_Z), (dict(num_parity_nodes=47), L.ITEM_BAD_MB),
           (dict(num_parity_nodes=3), L.ITEM_BAD_MB), (dict(Kb=21), L.ITEM_BAD_KB),
           (dict(max_iterations=-1), L.ITEM_BAD_ITER), (dict(llr=0), L.ITEM_NULL_BUFFER)]
    rows, expect, first = [], [], [0]
    if mode == "grouped":
        # group 1: A's four code blocks with a different mb third: decoded up to it
        rows += [item(A, 0), item(A, 1), bad(A, num_parity_nodes=11), item(A, 3)]
        expect += [0, 0, L.ITEM_GROUP, L.ITEM_GROUP]
        first.append(len(rows))
        # group 2: an invalid first item rejects the whole group
        rows += [bad(B, Kb=7), bad(B), bad(B)]
        expect += [L.ITEM_BAD_KB, L.ITEM_GROUP, L.ITEM_GROUP]
        first.append(len(rows))
        # group 3: a null output buffer cuts the group
        rows += [item(Cc, 0), item(Cc, 1), bad(Cc, out=0), bad(Cc)]
        expect += [0, 0, L.ITEM_NULL_BUFFER, L.ITEM_GROUP]
        first.append(len(rows))
        decoded = {(0, 0), (0, 1), (2, 0), (2, 1)}
        block, grp = 384, first
    else:
        for t, tb in enumerate(tbs):
            for c in range(tb.ncb):
                rows.append(item(tb, c))
                expect.append(0)
        decoded = {(t, c) for t, tb in enumerate(tbs) for c in range(tb.ncb)}
        block, grp = (768, list(range(len(rows) + len(BAD) + 1))) if mode == "row_pair" else (384, None)
    for kw, why in BAD:
        rows.append(bad(A, **kw))
        expect.append(why)
    if mode == "grouped":
        for kw, why in BAD:
            first.append(first[-1] + 1)
        grp = first
    elif mode == "row_pair":
        grp = list(range(len(rows) + 1))
    arr = np.concatenate(rows)
    max_values = max(group_values(tb.bg, tb.Z, tb.mb, tb.ncb) for tb in tbs)
    got = _launch_items(dec, arr, grp, max_values, max(tb.mb for tb in tbs), block, L.CUPHY_R_16F)
    assert list(got["status"]) == expect, (list(got["status"]), expect)
    assert int((sink != -1).sum().item()) == 0, "an invalid item's output was written"
    for t, (tb, ref) in enumerate(zip(tbs, refs)):
        out = tb.out.cpu().numpy().view(np.uint32)
        for c in range(tb.ncb):
            if (t, c) in decoded:
                assert np.array_equal(out[c], ref[c]), (mode, t, c)
            else:
                assert (out[c] == 0xFFFFFFFF).all(), (mode, t, c)


def test_batch_rejects_rows_beyond_high_rate_kernel(dec):
    """A launch whose max_parity_nodes (<= 8) selects the high-code-rate
    kernel rejects an item with more parity rows instead of decoding part of
    them; its other items decode as the oracle."""
    from aerial_amd import _lib as L
    from aerial_amd.batch import ITEM_DTYPE, _half_bits, group_values
    print("[case] batch: mb 10 item in an mb <= 8 launch", flush=True)
    tbs, refs = make_tbs("fp16", seed=62, mix=[(1, 96, 6, 2), (2, 64, 10, 1)])
    h = _half_bits(0.8125)
    arr = np.zeros(3, ITEM_DTYPE)
    for i, (t, c) in enumerate([(0, 0), (0, 1), (1, 0)]):
        tb = tbs[t]
        arr["llr"][i] = tb.llr.data_ptr() + c * tb.llr.stride(0) * 2
        arr["out"][i] = tb.out.data_ptr() + c * tb.out.stride(0) * 4
        arr["BG"][i], arr["Z"][i], arr["num_parity_nodes"][i], arr["Kb"][i] = tb.bg, tb.Z, tb.mb, 22 if tb.bg == 1 else 10
        arr["max_iterations"][i], arr["norm"][i], arr["clamp_value"][i] = 7, h | (h << 16), 32.0
    for grp in (None, [0, 1, 2, 3]):
        for tb in tbs:
            tb.out.fill_(-1)
        got = _launch_items(dec, arr, grp, max(group_values(tb.bg, tb.Z, tb.mb, 1) for tb in tbs), 8, 384,
                            L.CUPHY_R_16F)
        assert list(got["status"]) == [0, 0, L.ITEM_BAD_MB]
        assert np.array_equal(tbs[0].out.cpu().numpy().view(np.uint32), refs[0])
        assert (tbs[1].out.cpu().numpy().view(np.uint32) == 0xFFFFFFFF).all()
