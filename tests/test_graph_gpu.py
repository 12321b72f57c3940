"""PuschRx-style graph layer: TBs grouped into decode descriptors by
configuration (at most 32 TBs each), one graph node per descriptor, node
parameters updated and unused nodes disabled per slot
(pusch_rx.cpp:3007-3021, 4450-4468, 6275-6415).  Every slot's hard bits equal
the oracle's."""
from types import SimpleNamespace

import numpy as np
import pytest

import oracle as O
import vectors as V

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


def make_slot(spec, seed, n_iter):
    """spec: (bg, Z, mb, ncb, n_tbs) entries -> TBs (binary16 LLRs) and oracle bits."""
    tbs, refs = [], []
    k = 0
    for bg, Z, mb, ncb, n_tbs in spec:
        for _ in range(n_tbs):
            llr, _ = V.make_cbs(bg, Z, mb, ncb, seed=seed + k, snr_db=1.5)
            k += 1
            llr = V.to_half_rn(llr).astype(np.float32)
            stride = (llr.shape[1] + 7) // 8 * 8
            host = np.zeros((ncb, stride), np.float32)
            host[:, :llr.shape[1]] = llr
            nw = (V.nsys(bg) * Z + 31) // 32
            tbs.append(SimpleNamespace(bg=bg, Z=Z, mb=mb, ncb=ncb, llr=torch.from_numpy(host).cuda().half(),
                                       out=torch.full((ncb, nw), -1, dtype=torch.int32, device="cuda")))
            cfg = O.dec_cfg(bg, Z, mb, n_iter=n_iter, alpha=0.8125, clamp=32.0, mode=O.FP16)
            refs.append(np.stack([O.decode_cb(cfg, llr[i])[0] for i in range(ncb)]))
    return tbs, refs


SLOTS = [
    [(1, 384, 46, 2, 1), (2, 64, 42, 3, 2), (1, 72, 4, 2, 3), (2, 208, 9, 1, 1), (1, 2, 20, 5, 1), (2, 15, 4, 2, 2)],
    [(2, 16, 4, 1, 40), (1, 96, 8, 2, 2), (2, 384, 42, 1, 1)],   # 40 TBs of one configuration: 2 descriptors
    [(1, 160, 33, 2, 2)],                                        # one node used, the rest disabled
]


def test_graph_slots_bit_exact(dec):
    from aerial_amd.graph import LdpcDecodeGraph
    g = LdpcDecodeGraph(dec, max_configs=8)
    expected_nodes = [6, 4, 1]
    for s, spec in enumerate(SLOTS):
        print("[case] graph slot %d: %s" % (s, spec), flush=True)
        tbs, refs = make_slot(spec, seed=600 + 100 * s, n_iter=7)
        n = g.run(tbs, max_iterations=7)
        torch.cuda.synchronize()
        assert n == expected_nodes[s]
        assert sum(g.enabled) == n
        for tb, ref in zip(tbs, refs):
            assert np.array_equal(tb.out.cpu().numpy().view(np.uint32), ref), (s, tb.bg, tb.Z, tb.mb)
    # the first slot again on the updated graph (nodes re-enabled)
    tbs, refs = make_slot(SLOTS[0], seed=600, n_iter=7)
    g.run(tbs, max_iterations=7)
    torch.cuda.synchronize()
    for tb, ref in zip(tbs, refs):
        assert np.array_equal(tb.out.cpu().numpy().view(np.uint32), ref)
    g.close()


def test_graph_rejects_more_configurations_than_nodes(dec):
    from aerial_amd.graph import LdpcDecodeGraph
    print("[case] graph: 3 configurations, 2 nodes", flush=True)
    g = LdpcDecodeGraph(dec, max_configs=2)
    tbs, _ = make_slot([(1, 64, 4, 1, 1), (1, 72, 4, 1, 1), (1, 80, 4, 1, 1)], seed=900, n_iter=2)
    with pytest.raises(ValueError):
        g.run(tbs, max_iterations=2)
    g.close()
