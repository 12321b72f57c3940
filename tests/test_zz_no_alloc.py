"""Allocation-free fused slot decode (cuphyHipPuschSlotDecodeNoAlloc): its
outputs equal cuphyHipPuschSlotDecode's (workspace argument) bit for bit on
two calls in a row, the second with a smaller slot, and it rejects a slot
beyond its scratch array before launching anything."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

NOALLOC_MAX_CBS = 16384


@pytest.fixture(scope="module")
def dec():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from aerial_amd.decoder import LdpcDecoder
    d = LdpcDecoder()
    yield d
    d.close()


def _slot(dec, templates, idx):
    """First transmissions (ndi = 1, so a rerun writes the same HARQ
    buffers) of the templates idx, as one launch of the full kernel variant."""
    from aerial_amd.slot import SlotDecoder, SlotTb
    tbs = []
    for ti in idx:
        tb = templates.tbs[ti]
        p, _ = tb.params(0, 1)
        harq = torch.zeros(tb.C * p.Ncb_padded, dtype=torch.float16, device="cuda")
        tbs.append(SlotTb(p, tb.llrs[0], harq, templates.mbs[ti][0], max_iterations=8, early_term=True))
    sd = SlotDecoder(dec, split_min=1 << 30).build(tbs)
    assert len(sd.launches) == 1
    return sd, tbs


def _outputs(tbs):
    return [(t.cb_bits.clone(), t.cb_crc.clone(), t.tb_bytes.clone(), t.tb_crc.clone(), t.cb_iterations.clone(),
             t.harq.clone()) for t in tbs]


def _clear(tbs):
    for t in tbs:
        for x in (t.cb_bits, t.cb_crc, t.tb_crc):
            x.fill_(-1)
        t.tb_bytes.fill_(0xAB)
        t.cb_iterations.fill_(0xEE)
        t.harq.fill_(7.0)


def _no_alloc(dec, sd, n_cbs):
    from aerial_amd import _lib as L
    ln = sd.launches[0]
    return L.lib().cuphyHipPuschSlotDecodeNoAlloc(dec.dec, C.c_void_p(sd._jobs.data_ptr()), sd.n_tbs,
                                                  C.c_void_p(ln.items.data_ptr()), ln.n_items, n_cbs, sd.llr_type,
                                                  ln.max_values, ln.max_mb, ln.max_group, None)


def test_no_alloc_equals_workspace_entry_point(dec):
    from aerial_amd import _lib as L
    from aerial_amd.server import TraceTemplates, serving_specs
    templates = TraceTemplates(serving_specs(12, seed=77), seed=77)
    for call, idx in enumerate((list(range(12)), [3, 7, 1])):  # the second slot is smaller
        print("[case] no-alloc slot call %d: %d TBs" % (call, len(idx)), flush=True)
        sd, tbs = _slot(dec, templates, idx)
        n_cbs = sum(t.params.num_CBs for t in tbs)
        _clear(tbs)
        sd.launch()
        torch.cuda.synchronize()
        ref = _outputs(tbs)
        assert all(int(w.abs().sum().item()) == 0 for w in sd.workspaces)
        _clear(tbs)
        L.check("cuphyHipPuschSlotDecodeNoAlloc", _no_alloc(dec, sd, n_cbs))
        torch.cuda.synchronize()
        got = _outputs(tbs)
        for k, (a, b) in enumerate(zip(ref, got)):
            for j, (x, y) in enumerate(zip(a, b)):
                assert torch.equal(x, y), (call, k, j)


def test_no_alloc_rejects_slot_beyond_scratch(dec):
    from aerial_amd import _lib as L
    from aerial_amd.server import TraceTemplates, serving_specs
    templates = TraceTemplates(serving_specs(2, seed=78), seed=78)
    sd, tbs = _slot(dec, templates, [0, 1])
    _clear(tbs)
    torch.cuda.synchronize()
    assert _no_alloc(dec, sd, NOALLOC_MAX_CBS + 1) == L.CUPHY_STATUS_INVALID_ARGUMENT
    assert _no_alloc(dec, sd, -1) == L.CUPHY_STATUS_INVALID_ARGUMENT
    torch.cuda.synchronize()
    assert all(int((t.cb_crc == -1).all().item()) for t in tbs)  # nothing ran
