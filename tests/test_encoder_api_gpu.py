"""LDPC encoder behind the reference's encoder API (cuphySetupLDPCEncode:
bit-packed tensors, puncturing, a subset of parity rows for rv 0, batched
TBs) and DL rate matching (cuphyHipDlRateMatch: bit selection, bit
interleaving, scrambling), bit-exact against the oracle's encoder (pinned
by the reference's MATLAB goldens and H c = 0) and the oracle's rate-matching
steps (TS 38.212 5.4.2.1 E split, k0, bit selection pinned by the
reference's textbook selection; TS 38.211 Gold sequence)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def enc():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from aerial_amd.encoder import LdpcEncoder
    return LdpcEncoder(puncturing=True, max_tbs=8)


def nsys(bg):
    return 22 if bg == 1 else 10


def info_bits(bg, Z, ncw, seed, filler=0):
    rng = np.random.RandomState(seed)
    x = rng.randint(0, 2, (ncw, nsys(bg) * Z)).astype(np.uint8)
    if filler:
        x[:, nsys(bg) * Z - filler:] = 0
    return x


def words_to_bits(words, n):
    w = np.ascontiguousarray(words, dtype=np.int32).view(np.uint32)
    return np.unpackbits(w.view(np.uint8), axis=-1, bitorder="little")[..., :n]


# (bg, Z, puncture, num_rows (0: all), rv, TBs, code words, filler bits)
ENC_CASES = [(1, 384, 1, 0, 0, 1, 3, 0), (1, 384, 0, 0, 2, 2, 2, 224), (1, 2, 1, 0, 0, 1, 5, 0),
             (1, 15, 0, 0, 1, 1, 4, 0), (2, 384, 1, 0, 0, 3, 2, 96), (2, 7, 1, 0, 3, 1, 6, 0),
             (1, 208, 1, 8, 0, 2, 3, 0), (2, 104, 1, 6, 0, 1, 4, 40), (1, 64, 0, 20, 2, 1, 2, 0),
             (2, 36, 1, 2, 0, 1, 3, 0)]


@pytest.mark.parametrize("case", ENC_CASES)
def test_encoder_api_matches_oracle(enc, case):
    from aerial_amd.encoder import pack_bits
    bg, Z, punct, rows, rv, n_tbs, ncw, filler = case
    print("[case] encoder API BG%d Z=%d puncture=%d rows=%d rv=%d TBs=%d CWs=%d" % case[:7], flush=True)
    enc.puncturing = bool(punct)
    infos = [info_bits(bg, Z, ncw, 100 * sum(case) + t, filler) for t in range(n_tbs)]
    words = enc.encode_words(bg, Z, [pack_bits(torch.from_numpy(x.T.copy()).cuda()) for x in infos],
                             num_rows=rows, rv=rv)
    torch.cuda.synchronize()
    nrows = rows if (rv == 0 and 0 < rows < (46 if bg == 1 else 42)) else (46 if bg == 1 else 42)
    first = 2 * Z if punct else 0
    nbits = (nsys(bg) + nrows) * Z - first
    for t in range(n_tbs):
        got = words[t].cpu().numpy()
        bits = words_to_bits(got, got.shape[1] * 32)
        for c in range(ncw):
            cw = O.encode(bg, Z, infos[t][c])
            assert np.array_equal(bits[c, :nbits], cw[first:first + nbits]), (case, t, c)
            assert not bits[c, nbits:].any(), "bits after the code word are not zero"
    enc.puncturing = True


def dl_reference(params, cw_bits, cinit, scrambling):
    """Rate-matched bits of each CB from the oracle's steps: E_r, bit
    selection from k0 skipping the filler, interleaving f[j Qm + k] =
    e[k E/Qm + j], XOR with c(start_r + i)."""
    p = params
    cfg = O.rm_cfg(bg=p.bg, Zc=p.Zc, C=p.num_CBs, Qm=p.Qm, Nl=p.Nl, G=p.encodedSize, K=p.K, F=p.F, Ncb=p.Ncb,
                   Ncb_padded=0, k0=p.k0, ndi=1)
    Kd = p.K - p.F - 2 * p.Zc
    gold = O.gold(cinit, 0, p.encodedSize) if scrambling else None
    out, start = [], 0
    for r in range(p.num_CBs):
        E = O.rm_E(cfg, r)
        e = cw_bits[r][O.bit_select(E, Kd, p.F, p.k0, p.Ncb)]
        f = e.reshape(p.Qm, E // p.Qm).T.reshape(-1)
        if scrambling:
            f = f ^ gold[start:start + E]
        out.append(f)
        start += E
    return out


# (tbs, code rate, Qm, layers, rv, G, LBRM (n_PRB_LBRM, maxLayers, maxQm) or None, scrambling)
DL_CASES = [(8448 * 3 - 96, 0.9, 8, 2, 0, 3 * 9216, None, True),       # BG1 Zc=384, 3 CBs
            (7608, 0.3, 4, 1, 2, 2 * 9000, None, False),               # BG1 R < 0.67, one CB, k0 at rv 2
            (1200, 0.5, 2, 1, 3, 7000, None, True),                    # BG2, repetition (E > Ncb - F)
            (256, 0.2, 6, 1, 1, 1800, None, True),                     # BG2, small Zc, filler
            (120936, 0.66, 8, 4, 1, 184000, (51, 2, 6), True),         # limited buffer: Ncb < N, E > Ncb - F
            (40976, 0.8, 6, 3, 0, 3 * 17766, None, True)]              # uneven E split


@pytest.mark.parametrize("case", DL_CASES)
def test_dl_rate_match_matches_oracle(case):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from aerial_amd.encoder import dl_rate_match
    from aerial_amd.pusch import set_per_tb_params
    tbs, R, Qm, nl, rv, G, lbrm, scr = case
    print("[case] DL rate matching TBS=%d R=%.2f Qm=%d Nl=%d rv=%d G=%d lbrm=%s scr=%d" % case, flush=True)
    cinit = (0x51ED27 * (tbs % 977) + 3) & 0x7FFFFFFF
    p, _, _ = set_per_tb_params(tbs, R, Qm, G, rv, num_layers=nl, cinit=cinit, lbrm=lbrm)
    assert p.Ncb <= p.N and (lbrm is None) == (p.Ncb == p.N)
    rng = np.random.RandomState(tbs)
    cw = []
    for r in range(p.num_CBs):
        info = rng.randint(0, 2, p.K).astype(np.uint8)
        info[p.K - p.F:] = 0
        cw.append(O.encode(p.bg, p.Zc, info)[2 * p.Zc:2 * p.Zc + p.N])
    nw = (p.N + 31) // 32
    words = np.zeros((p.num_CBs, nw), np.uint32)
    for r in range(p.num_CBs):
        words[r] = O.pack_bits(cw[r])
    dev = torch.from_numpy(words.view(np.int32)).cuda()
    emax = ((max(O.rm_E(O.rm_cfg(bg=p.bg, Zc=p.Zc, C=p.num_CBs, Qm=Qm, Nl=nl, G=G, K=p.K, F=p.F, Ncb=p.Ncb,
                                         Ncb_padded=0, k0=0, ndi=1), r) for r in range(p.num_CBs)) + 31) // 32 + 1) * 32
    out = dl_rate_match([p], [dev], emax, scrambling=scr)
    torch.cuda.synchronize()
    ref = dl_reference(p, cw, cinit, scr)
    got = out.cpu().numpy()[0]
    for r in range(p.num_CBs):
        bits = words_to_bits(got[r], emax)
        E = ref[r].size
        assert np.array_equal(bits[:E], ref[r]), (case, r)
        assert not bits[E:].any(), (case, r)


def test_encode_then_rate_match_chain(enc):
    """Encoder API -> DL rate matching on the device equals the oracle's
    encoder -> rate-matching steps; the pyAerial-style encode() gives the
    same coded bits."""
    from aerial_amd.encoder import dl_rate_match, pack_bits
    from aerial_amd.pusch import set_per_tb_params
    print("[case] encoder API -> DL rate matching chain, TBS 60000", flush=True)
    tbs, R, Qm, G = 60000, 0.75, 6, 6 * 13500
    p, _, _ = set_per_tb_params(tbs, R, Qm, G, 2, cinit=12345)
    rng = np.random.RandomState(8)
    kp = p.K - p.F
    blocks = rng.randint(0, 2, (kp, p.num_CBs)).astype(np.uint8)
    cbs_dev = torch.from_numpy(blocks).cuda()
    coded = enc.encode([cbs_dev], [tbs], [R], [2])[0]                       # (N, C) bits
    info = torch.zeros((p.K, p.num_CBs), dtype=torch.uint8, device="cuda")
    info[:kp] = cbs_dev
    words = enc.encode_words(p.bg, p.Zc, [pack_bits(info)], rv=2)[0]
    out = dl_rate_match([p], [words], 32 * ((G // p.num_CBs + 40) // 32 + 1), scrambling=True)
    torch.cuda.synchronize()
    cw = []
    for r in range(p.num_CBs):
        full = np.zeros(p.K, np.uint8)
        full[:kp] = blocks[:, r]
        cw.append(O.encode(p.bg, p.Zc, full)[2 * p.Zc:2 * p.Zc + p.N])
    assert np.array_equal(coded.cpu().numpy().T, np.stack(cw))
    ref = dl_reference(p, cw, 12345, True)
    got = out.cpu().numpy()[0]
    for r in range(p.num_CBs):
        assert np.array_equal(words_to_bits(got[r], got.shape[1] * 32)[:ref[r].size], ref[r]), r
