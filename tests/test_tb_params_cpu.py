"""Per-TB parameter derivation (TB size -> base graph, segmentation, lifting
size, filler bits, limited-buffer Ncb, k0, parity nodes): the oracle's
restatement of derive_coding_main.m / derate_match.m is pinned by the
reference's own data (TBS table, the (E, K, F, k0, Ncb, Zc) tuples of its
PUSCH test vectors in test_derate_match_modulo.cu, an LBRM test vector) and
by invariants of the standard; the library's C-ABI derivation
(cuphyHipTransportBlockSize, cuphyHipSetPerTbParams; host code, no GPU) is
then compared with the oracle field by field."""
import ctypes as C
import json
import os
from collections import defaultdict

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    return json.load(open(os.path.join(GOLD, name)))


def tv_groups():
    """Test-vector rows of the de-rate-match cases grouped by vector: name ->
    (rows, largest CB number listed)."""
    g = defaultdict(list)
    for r in golden("derate_modulo_cases.json"):
        if not r["tv"].startswith("TVnr"):
            continue
        name, _, c = r["tv"].partition(" c=")
        g[name].append((r, int(c)))
    return {k: (v, max(c for _, c in v)) for k, v in g.items()}


def bg_of(row):
    K = row["K"] + 2 * row["Zc"]
    return 1 if K == 22 * row["Zc"] else 2


# ---------------------------------------------------------------------------
# oracle pins
# ---------------------------------------------------------------------------
def test_oracle_tbs_table_is_the_references():
    assert O.tbs_table() == golden("tbs_table.json")


def test_oracle_tbs_segments_exactly():
    """TS 38.214 chooses TBS so that code-block segmentation is exact
    (B + 24 C divisible by C, K' integral) and the TB CRC length follows the
    3824-bit boundary: every TBS the oracle derives over a grid of
    allocations segments exactly, small TBS come from the table, and TBS is
    non-decreasing in the allocation size."""
    table = set(golden("tbs_table.json"))
    for Qm in (2, 4, 6, 8):
        for nl in (1, 2, 4):
            for R in (120 / 1024, 251 / 1024, 449 / 1024, 658 / 1024, 873 / 1024, 948 / 1024):
                prev = 0
                for nprb in list(range(1, 40)) + list(range(40, 274, 7)):
                    tbs = O.tb_size(144, nprb, R, Qm, nl)
                    assert tbs >= prev, (Qm, nl, R, nprb)
                    prev = tbs
                    if tbs <= 3824:
                        assert tbs in table
                    c = O.derive_coding(tbs, R)  # asserts K' integral
                    assert c["C"] * c["Kp"] == c["B"] + (24 * c["C"] if c["C"] > 1 else 0)
                    assert c["Kp"] <= (8448 if c["bg"] == 1 else 3840)
                    assert 0 <= c["F"] < c["K"]


def test_oracle_lifting_size_is_smallest_that_fits():
    lifting = sorted({a << j for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8) if (a << j) <= 384})
    for R in (0.2, 0.5, 0.9):
        for tbs in golden("tbs_table.json") + [O.tb_size(144, n, R, 6, 2) for n in range(10, 274, 3)]:
            c = O.derive_coding(tbs, R)
            assert c["Zc"] == min(z for z in lifting if z * c["Kb"] >= c["Kp"]), (tbs, R)
            assert c["K"] == c["Zc"] * (22 if c["bg"] == 1 else 10)


def test_oracle_segmentation_matches_reference_test_vectors():
    """Each multi-CB PUSCH test vector of the reference's de-rate-match table
    lists its first and last CB with (K - 2Zc, F, Ncb, Zc): the TB of C CBs of
    K' = K - F bits has C K' - 24 C - 24 payload bits, and the oracle's
    derivation from that payload must give back C, Zc and F."""
    n = 0
    for name, (rows, C_) in sorted(tv_groups().items()):
        if C_ < 2:
            continue
        r = rows[0][0]
        K = r["K"] + 2 * r["Zc"]
        Kp = K - r["F"]
        tbs = C_ * Kp - 24 * C_ - 24
        bg = bg_of(r)
        c = O.derive_coding(tbs, 0.9 if bg == 1 else 0.2)
        assert (c["bg"], c["C"], c["Zc"], c["F"], c["K"]) == (bg, C_, r["Zc"], r["F"], K), (name, c)
        n += 1
    assert n >= 60


def test_oracle_limited_buffer_matches_reference_test_vector():
    """TVnr_7321 / 7323: 107 CBs of BG1 Zc=384 with Ncb = 17915 < N = 25344,
    the limited buffer of a UE with n_PRB_LBRM = 273, 4 layers, 256QAM."""
    g = tv_groups()
    for name in ("TVnr_7321", "TVnr_7323"):
        rows, C_ = g[name]
        r = rows[0][0]
        assert C_ == 107 and r["Ncb"] == 17915
        tbs = C_ * (r["K"] + 2 * r["Zc"] - r["F"]) - 24 * C_ - 24
        c = O.derive_coding(tbs, 0.66, i_lbrm=1, n_prb_lbrm=273, max_layers=4, max_qm=8)
        assert c["C"] == 107 and min(c["N"], c["Nref"]) == 17915


def test_oracle_k0_matches_reference_test_vectors():
    """Nonzero k0 of the reference's vectors (rv 1..3, incl. the limited
    buffer and BG2 HARQ cases) are the oracle's k0 of some rv."""
    seen = 0
    for r in golden("derate_modulo_cases.json"):
        if r["k0"] == 0 or not r["tv"].startswith("TV"):
            continue
        bg = bg_of(r)
        assert r["k0"] in [O.k0(bg, rv, r["Ncb"], r["Zc"]) for rv in (1, 2, 3)], r
        seen += 1
    assert seen >= 10


def test_oracle_bit_selection_matches_textbook_on_reference_cases():
    """The oracle's bit selection equals the textbook iterative selection
    the reference's test uses (ref_calc_modhole,
    test_derate_match_modulo.cu:21-38: walk the circular buffer from k0,
    skip [K - F, K)) on every row of its table, incl. the multi-loop and
    k0-in-the-hole cases."""
    for r in golden("derate_modulo_cases.json"):
        E, K, F, k0, Ncb = r["E"], r["K"], r["F"], r["k0"], r["Ncb"]
        Kd = K - F
        walk = (k0 + np.arange(E + (E // max(1, Ncb - F) + 2) * Ncb)) % Ncb
        ref = walk[(walk < Kd) | (walk >= K)][:E]
        assert np.array_equal(O.bit_select(E, Kd, F, k0, Ncb), ref), r


def test_oracle_parity_nodes_cover_the_received_llrs():
    """nV_parity is the smallest count (4 at least, 46/42 at most) of parity
    columns whose end reaches the last received LLR, min(G/C + k0, N)."""
    rng = np.random.RandomState(3)
    for _ in range(3000):
        bg = int(rng.randint(1, 3))
        Zc = int(rng.choice([2, 15, 36, 64, 104, 208, 384]))
        K = Zc * (22 if bg == 1 else 10)
        F = int(rng.randint(0, K // 4))
        C_ = int(rng.randint(1, 30))
        G = int(rng.randint(1, 40000)) * C_
        rv = int(rng.randint(0, 4))
        N = Zc * (66 if bg == 1 else 50)
        k0 = O.k0(bg, rv, N, Zc)
        last = min(G // C_ + k0, N)
        Kd = K - F - 2 * Zc
        mb = next(m for m in range(1, 100) if Kd + m * Zc >= last)
        mb = max(4, min(46 if bg == 1 else 42, mb))
        assert O.num_parity_nodes(bg, Zc, K, F, G, C_, k0) == mb


# ---------------------------------------------------------------------------
# the library's derivation (C ABI, host only) against the oracle
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def L():
    from aerial_amd import _lib
    _lib.lib()
    return _lib


MCS_RATES = [120, 193, 308, 449, 602, 378, 434, 490, 553, 616, 658, 466, 517, 567, 616, 666, 719, 772, 822, 873,
             682.5, 711, 754, 797, 841, 885, 916.5, 948]


def test_tb_size_matches_oracle(L):
    from aerial_amd.pusch import get_tb_size
    n, diff = 0, []
    for nsym in (2, 7, 11, 12, 13):
        for nprb in (1, 3, 8, 24, 51, 106, 217, 273):
            for nl in (1, 2, 4):
                for Qm in (2, 4, 6, 8):
                    for x in MCS_RATES:
                        R = float(np.float32(x / 1024))
                        tbs, ncb = get_tb_size(nsym, nprb, nl, R, Qm)
                        ref = O.tb_size(12 * nsym, nprb, R, Qm, nl)
                        n += 1
                        if tbs != ref:
                            diff.append((nsym, nprb, nl, Qm, x, tbs, ref))
                        c = O.derive_coding(tbs, R)
                        assert ncb == c["C"] or tbs <= 3824
    assert not diff, diff[:10]
    assert n > 4000


def _oracle_params(tbs, R, Qm, G, rv, lbrm):
    c = O.derive_coding(tbs, R, *((1,) + lbrm if lbrm else (0, 273, 4, 8)))
    Ncb = min(c["N"], c["Nref"]) if lbrm else c["N"]
    k0 = O.k0(c["bg"], rv, Ncb, c["Zc"])
    Kd = c["Kp"] - 2 * c["Zc"]
    if lbrm:  # the reference rounds down within the limited buffer (pusch_rx.cpp:8190,8220)
        last = min(G // c["C"] + k0, Ncb)
        mb = max(m for m in range(0, 100) if Kd + m * c["Zc"] <= last) if last > Kd else 4
        mb = max(4, min(46 if c["bg"] == 1 else 42, mb))
    else:
        mb = O.num_parity_nodes(c["bg"], c["Zc"], c["K"], c["F"], G, c["C"], k0)
    data_bytes = (c["Kp"] - (24 if c["C"] > 1 else c["crc_len"]) + 7) // 8 * c["C"]
    return c, Ncb, k0, mb, data_bytes


@pytest.mark.parametrize("lbrm", [None, (273, 4, 8), (51, 2, 6), (106, 1, 8)])
def test_set_per_tb_params_matches_oracle(L, lbrm):
    from aerial_amd.pusch import set_per_tb_params
    rng = np.random.RandomState(17 if lbrm is None else lbrm[0])
    checked = 0
    for R in (0.2, 0.5, 0.7, 0.93):
        # sizes TS 38.214 gives at this rate (the CB count of a large TBS depends on R <= 1/4)
        for tbs in golden("tbs_table.json") + [O.tb_size(144, int(n), R, 8, 2) for n in range(20, 274, 5)]:
            Qm = int(rng.choice([2, 4, 6, 8]))
            nl = int(rng.choice([1, 2]))
            G = int(np.ceil(tbs / R / (Qm * nl))) * Qm * nl
            for rv in (0, 2, 3, 1):
                p, mb, kb = set_per_tb_params(tbs, R, Qm, G, rv, ndi=int(rv == 0), num_layers=nl, lbrm=lbrm)
                c, Ncb, k0, mb_ref, data_bytes = _oracle_params(tbs, float(np.float32(R)), Qm, G, rv, lbrm)
                got = (p.bg, p.num_CBs, p.Zc, p.K, p.F, p.N, p.Ncb, p.k0, kb, mb, p.nDataBytes)
                exp = (c["bg"], c["C"], c["Zc"], c["K"], c["F"], c["N"], Ncb, k0, c["Kb"], mb_ref, data_bytes)
                assert got == exp, (tbs, R, Qm, nl, rv, lbrm, got, exp)
                assert p.Ncb_padded % 8 == 0 and p.Ncb + 2 * p.Zc <= p.Ncb_padded < p.Ncb + 2 * p.Zc + 8
                assert p.nZpBitsPerCb == mb * p.Zc + p.K and p.G == p.encodedSize == G
                assert (p.tbSize, p.Qm, p.rv, p.Nl, p.isDataPresent) == (tbs, Qm, rv, nl, 1)
                checked += 1
    assert checked > 1000


def test_set_per_tb_params_rejects_invalid(L):
    lib = L.lib()
    p = L.PerTbParams()
    mb, kb = C.c_uint16(), C.c_uint16()
    lm = (C.c_uint32 * 8)(0, 1, 2, 3, 4, 5, 6, 7)

    def call(tbs=20000, R=0.5, Qm=2, rv=0, G=40000, nl=1, ng=1, lmap=lm, lbrm=None):
        return lib.cuphyHipSetPerTbParams(C.byref(p), tbs, R, Qm, 1, rv, G, 0, 0, nl, ng, lmap, 2,
                                          C.byref(lbrm) if lbrm is not None else None, C.byref(mb), C.byref(kb))

    assert call() == L.CUPHY_STATUS_SUCCESS
    bad = L.CUPHY_STATUS_INVALID_ARGUMENT
    assert call(rv=4) == bad
    assert call(Qm=3) == bad
    assert call(tbs=0) == bad
    assert call(R=0.0) == bad and call(R=1.5) == bad
    assert call(G=40001) == bad                       # not a multiple of Qm * layers
    assert call(nl=2, ng=1) == bad                    # more layers than the UE group has
    assert call(nl=1, ng=2, lmap=(C.c_uint32 * 8)(2, 0, 0, 0, 0, 0, 0, 0)) == bad
    assert call(lbrm=L.LbrmPrm(i_lbrm=1, maxLayers=4, maxQm=4, n_PRB_LBRM=100)) == bad
    assert call(lbrm=L.LbrmPrm(i_lbrm=1, maxLayers=4, maxQm=8, n_PRB_LBRM=20)) == bad
    assert call(lbrm=L.LbrmPrm(i_lbrm=0, maxLayers=0, maxQm=0, n_PRB_LBRM=0)) == L.CUPHY_STATUS_SUCCESS
    tb, ncb = C.c_uint32(), C.c_uint32()
    assert lib.cuphyHipTransportBlockSize(12, 0, 1, 0.5, 2, 0, C.byref(tb), C.byref(ncb)) == bad
    assert lib.cuphyHipTransportBlockSize(12, 10, 1, 0.0, 2, 0, C.byref(tb), C.byref(ncb)) == bad


def test_make_tb_params_matches_oracle():
    """The explicit-segmentation helper used by the fused-slot and serving
    workloads (BG, C, K', Zc given) agrees with the oracle's k0 and parity
    nodes."""
    from aerial_amd.pusch import make_tb_params
    rng = np.random.RandomState(5)
    lifting = sorted({a << j for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8) if (a << j) <= 384})
    for _ in range(2000):
        bg = int(rng.randint(1, 3))
        Zc = int(rng.choice(lifting))
        K = Zc * (22 if bg == 1 else 10)
        Kp = K - 8 * int(rng.randint(0, K // 32 + 1))
        C_ = int(rng.randint(1, 20))
        Qm = int(rng.choice([1, 2, 4, 6, 8]))
        G = C_ * Qm * int(rng.randint(max(1, Kp // (3 * Qm)), 3 * K // Qm + 2))
        rv = int(rng.randint(0, 4))
        p, mb = make_tb_params(bg, Zc, C_, Kp, Qm, 1, G, rv)
        assert p.k0 == O.k0(bg, rv, p.Ncb, Zc)
        assert mb == O.num_parity_nodes(bg, Zc, K, K - Kp, G, C_, p.k0)
