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
