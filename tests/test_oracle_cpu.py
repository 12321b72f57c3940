"""Pins of the CPU oracle (no GPU): every oracle function is checked against
the reference's own fixtures and known answers, against independent
restatements built differently (dense parity-check matrices, brute-force
extrinsic minima, forward rate matching), or against closed forms and
published catalogue values -- chosen so that a plausible mistake (a wrong
shift direction, a dropped term, a wrong sign/tie rule, a wrong index)
fails at least one of them."""
import json
import os

import numpy as np
import pytest

import oracle as O
import vectors as V

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
LIFTING = sorted({a << j for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8) if (a << j) <= 384})


# ---------------------------------------------------------------------------
# base-graph tables
# ---------------------------------------------------------------------------
def test_lifting_sets():
    assert len(LIFTING) == 51
    for Z in LIFTING:
        ils = O.ils(Z)
        a = (2, 3, 5, 7, 9, 11, 13, 15)[ils]
        assert Z % a == 0 and (Z // a) & (Z // a - 1) == 0
    assert O.ils(17) == -1 and O.ils(400) == -1


def test_row_degrees_and_structure():
    deg1 = [O.lib().nr_orc_row_degree(1, r) for r in range(46)]
    deg2 = [O.lib().nr_orc_row_degree(2, r) for r in range(42)]
    assert sum(deg1) == 316 and sum(deg2) == 197
    assert deg1[:5] == [19, 19, 19, 19, 3] and deg2[:4] == [8, 10, 8, 10]
    # every extension row r >= 4 owns column nsys + r with shift 0 (degree-1 column)
    for bg, nr in ((1, 46), (2, 42)):
        ns = V.nsys(bg)
        for r in range(4, nr):
            cols = dict(O.edges(bg, 384, r))
            assert cols.get(ns + r) == 0
            assert all(c < ns + 4 or c == ns + r for c in cols)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_committed_tables_regenerate_identically(tmp_path):
    """The committed header equals a fresh generation from the reference's
    38.212 text tables, which the generator cross-checks against the
    second copy in nrLDPC_templates.cuh."""
    import subprocess
    import sys
    root = os.path.dirname(HERE)
    gen = os.path.join(root, "aerial-cuda-accelerated-ran_amd", "gen_bg_tables.py")
    out = tmp_path / "t.h"
    subprocess.run([sys.executable, gen, "--ref", REF, "--out", str(out)], check=True, capture_output=True)
    committed = os.path.join(root, "aerial-cuda-accelerated-ran_amd", "csrc", "nr_bg_tables.h")
    assert out.read_text() == open(committed).read()


# ---------------------------------------------------------------------------
# encoder
# ---------------------------------------------------------------------------
def dense_H(bg, Z, mb):
    """Parity-check matrix from TS 38.212 5.3.2: block (r, col) with shift s
    is the identity cyclically shifted right s times, H[r*Z+i, col*Z+(i+s)%Z] = 1."""
    ns = V.nsys(bg)
    H = np.zeros((mb * Z, (ns + mb) * Z), np.uint8)
    for r in range(mb):
        for col, s in O.edges(bg, Z, r):
            for i in range(Z):
                H[r * Z + i, col * Z + (i + s) % Z] = 1
    return H


@pytest.mark.parametrize("bg", [1, 2])
def test_encoder_all_lifting_sizes_satisfy_h(bg):
    for Z in LIFTING:
        info = O.random_bits(Z * 7 + bg, V.nsys(bg) * Z)
        cw = O.encode(bg, Z, info)
        assert (cw[:V.nsys(bg) * Z] == info).all()
        assert O.syndrome(bg, Z, O.nrows(bg), cw) == 0, (bg, Z)


@pytest.mark.parametrize("bg,Z", [(1, 4), (1, 13), (2, 6), (2, 15)])
def test_encoder_against_dense_h(bg, Z):
    H = dense_H(bg, Z, O.nrows(bg))
    for seed in range(3):
        cw = O.encode(bg, Z, O.random_bits(seed, V.nsys(bg) * Z))
        assert not (H.astype(np.int64) @ cw % 2).any()
    # and the code is not trivial: H has full row rank over the parity part
    assert H.shape[0] == O.nrows(bg) * Z


# ---------------------------------------------------------------------------
# check node (compute_cC2V.m + update_C2V.m) vs the textbook definition on
# the input distribution of the reference KAT (test_ldpc_internal_rc.cu:
# 174-215: std::mt19937 default seed, uniform(-61440, 61440) -> binary16)
# ---------------------------------------------------------------------------
def mt19937_kat_values(n):
    rs = np.random.RandomState(5489)            # == std::mt19937 default seed
    raw = rs.randint(0, 2 ** 32, size=2 * n, dtype=np.uint64).astype(np.float64)
    r = (raw[0::2] + raw[1::2] * 2.0 ** 32) / 2.0 ** 64  # generate_canonical<double, 53>
    return (-61440.0 + 122880.0 * r).astype(np.float16)


def test_mt19937_kat_stream():
    v = mt19937_kat_values(2)
    assert v[0] == np.float16(-44792.585712011612) and v[1] == np.float16(41165.855538533928)


@pytest.mark.parametrize("mode", [O.FP32, O.FP16])
def test_check_node_is_extrinsic_min_sum(mode):
    vals = mt19937_kat_values(19 * 1024).astype(np.float32)
    rng = np.random.RandomState(1)
    pos = 0
    for d in list(range(2, 20)) * 40:
        t = vals[pos:pos + d].copy()
        pos = (pos + d) % (vals.size - 20)
        if rng.rand() < 0.3:                       # ties of the minimum
            t[rng.randint(d)] = np.float32(np.sign(t[0]) * abs(t[0])) if d > 1 else t[0]
            t[rng.randint(d)] = -abs(t[0])
        if rng.rand() < 0.1:
            t[rng.randint(d)] = 0.0
        got = O.check_node(t, mode)
        for j in range(d):
            others = np.delete(t, j)
            sgn = np.prod(np.where(others < 0, -1.0, 1.0))
            exp = np.float32(sgn * np.min(np.abs(others)))
            assert got[j] == exp and np.signbit(got[j]) == np.signbit(exp), (d, j, t, got)


# ---------------------------------------------------------------------------
# layered decoder vs an independent dense-matrix restatement (numpy,
# brute-force extrinsic minima, same per-op rounding)
# ---------------------------------------------------------------------------
def dense_layered(bg, Z, mb, llr, iters, alpha, clamp, dtype):
    ns = V.nsys(bg)
    n = (ns + mb) * Z
    a = np.clip(llr[:n].astype(np.float32), -clamp, clamp).astype(dtype)
    a[a == 0] = 0
    al = dtype(alpha)
    c2v = {}
    for _ in range(iters):
        for r in range(mb):
            e = O.edges(bg, Z, r)
            idx = [np.array([c * Z + (i + s) % Z for i in range(Z)]) for c, s in e]
            old = [c2v.get((r, j), np.zeros(Z, dtype)) for j in range(len(e))]
            t = np.stack([(a[idx[j]] - (al * old[j]).astype(dtype)).astype(dtype) for j in range(len(e))])
            for j in range(len(e)):
                others = np.delete(t, j, axis=0)
                mag = np.min(np.abs(others), axis=0).astype(dtype)
                neg = (np.sum(others < 0, axis=0) % 2) == 1
                c2v[(r, j)] = np.where(neg, -mag, mag).astype(dtype)
            for j in range(len(e)):
                v2c = np.where(t[j] < 0, t[j], np.abs(t[j])).astype(dtype)
                a[idx[j]] = (v2c + (al * c2v[(r, j)]).astype(dtype)).astype(dtype)
    return a


@pytest.mark.parametrize("bg,Z,mb,mode", [(1, 5, 8, O.FP32), (1, 3, 46, O.FP32), (2, 7, 12, O.FP32),
                                          (2, 4, 42, O.FP32), (1, 6, 10, O.FP16), (2, 9, 6, O.FP16)])
def test_decoder_equals_dense_restatement(bg, Z, mb, mode):
    llr, _ = V.make_cbs(bg, Z, mb, 1, seed=3, snr_db=0.5)
    llr = llr[0]
    if mode == O.FP16:
        llr = V.to_half_rn(llr).astype(np.float32)
    cfg = O.dec_cfg(bg, Z, mb, n_iter=4, alpha=0.75, clamp=20.0, mode=mode)
    _, _, app = O.decode_cb(cfg, llr, want_app=True)
    ref = dense_layered(bg, Z, mb, llr, 4, 0.75, 20.0, np.float32 if mode == O.FP32 else np.float16)
    assert np.array_equal(app.view(np.uint32), ref.astype(np.float32).view(np.uint32))


@pytest.mark.parametrize("mode", [O.FP32, O.FP64, O.FP16])
@pytest.mark.parametrize("bg,Z,mb", [(1, 384, 46), (2, 64, 20), (1, 13, 4)])
def test_noiseless_codewords_decode_exactly(bg, Z, mb, mode):
    # pyaerial/tests/test_ldpc_endtoend.py:152-184: noiseless +-L LLRs decode exactly
    llr, info = V.make_cbs(bg, Z, mb, 2, seed=4)
    cfg = O.dec_cfg(bg, Z, mb, n_iter=3, mode=mode)
    for i in range(2):
        hard, _, _ = O.decode_cb(cfg, llr[i])
        assert (O.unpack_bits(hard, V.nsys(bg) * Z) == info[i]).all()


def test_zero_iterations_is_clamped_input():
    # pyaerial/tests/test_ldpc_decoder.py:142-190
    llr, _ = V.make_cbs(1, 64, 5, 1, seed=6, snr_db=-2.0)
    llr = llr[0] * 3.0
    cfg = O.dec_cfg(1, 64, 5, n_iter=0, clamp=32.0)
    hard, it, app = O.decode_cb(cfg, llr, want_app=True)
    assert it == 0
    assert np.array_equal(app, np.clip(llr[:27 * 64], -32, 32) + np.float32(0))
    assert (O.unpack_bits(hard, 22 * 64) == (llr[:22 * 64] <= 0)).all()


def test_fp64_and_fp32_agree_at_high_snr():
    llr, info = V.make_cbs(1, 96, 12, 6, seed=8, snr_db=5.0)
    for i in range(6):
        h32, _, _ = O.decode_cb(O.dec_cfg(1, 96, 12, mode=O.FP32), llr[i])
        h64, _, _ = O.decode_cb(O.dec_cfg(1, 96, 12, mode=O.FP64), llr[i])
        assert np.array_equal(h32, h64)
        assert (O.unpack_bits(h32, 22 * 96) == info[i]).all()


def test_early_termination_stops_on_valid_codeword():
    llr, info = V.make_cbs(2, 128, 10, 3, seed=12, snr_db=6.0)
    for i in range(3):
        hard, it, app = O.decode_cb(O.dec_cfg(2, 128, 10, n_iter=30, early_term=True), llr[i], want_app=True)
        assert it < 30
        x = (app < 0).astype(np.uint8)
        assert O.syndrome(2, 128, 10, np.concatenate([x, np.zeros(32 * 128, np.uint8)])) == 0
        h_full, it_full, _ = O.decode_cb(O.dec_cfg(2, 128, 10, n_iter=it), llr[i])
        assert np.array_equal(hard, h_full)  # stopping == running exactly `it` iterations


def test_timed_decodes_equal_single_block_decodes():
    """The CPU baseline's timing entry points (pinned threads, one queue over
    the code blocks of several TBs) return what one-block decodes return."""
    a, _ = V.make_cbs(1, 64, 6, 3, seed=21, snr_db=1.0)
    b, _ = V.make_cbs(2, 96, 10, 5, seed=22, snr_db=0.5)
    ca, cb = O.dec_cfg(1, 64, 6), O.dec_cfg(2, 96, 10)
    want_a = np.stack([O.decode_cb(ca, x)[0] for x in a])
    want_b = np.stack([O.decode_cb(cb, x)[0] for x in b])
    times, hard = O.time_decode(cb, b, nthreads=3, warmup=1, reps=2)
    assert times.shape == (2,) and (times > 0).all()
    assert np.array_equal(hard, want_b)
    times, hards = O.time_slot([ca, cb], [a, b], nthreads=4, warmup=0, reps=2)
    assert (times > 0).all()
    assert np.array_equal(hards[0], want_a) and np.array_equal(hards[1], want_b)


def test_binary16_rounding_matches_ieee():
    rng = np.random.RandomState(0)
    xs = np.concatenate([rng.standard_normal(4000) * 10 ** rng.uniform(-8, 5, 4000),
                         [65504.0, 65519.99, 65520.0, 1e9, 6.1e-5, 5.96e-8, 2.98e-8, 2.99e-8, 0.0, -0.0, 1.00048828125,
                          1.000244140625, 1.000732421875]]).astype(np.float32)
    for x in xs:
        with np.errstate(over="ignore"):
            exp = np.float32(np.float16(x))
        got = O.round_half(x)
        assert got == exp and np.signbit(got) == np.signbit(exp), (x, got, exp)


# ---------------------------------------------------------------------------
# statistical pin: the reference's BLER-threshold fixtures
# (cuPHY/util/ldpc/test/*_BLER0.1.txt via tests/golden/bler_thresholds.json)
# ---------------------------------------------------------------------------
def kb_for(bg, Z):
    # cuPHY/examples/error_correction/ldpc_decode_test_vec.cpp:77-100
    if bg == 1:
        return 22
    return 10 if Z > 64 else (9 if Z > 56 else (8 if Z > 24 else 6))


def norm_table(bg, mb):
    # LDPC_decode.m:101-208 (== error_correction/ldpc.cpp:139-235)
    t1 = [0, 0, 0, 0, .79, .77, .75, .73, .75, .70, .67, .68, .67, .67, .68, .66, .65, .66, .64, .65, .65, .65, .65,
          .66, .66, .66, .66, .66, .66, .67, .66, .65, .64, .63, .63, .63, .63, .63, .62, .63, .63, .64, .63, .63, .63,
          .62, .63]
    t2 = [0, 0, 0, 0, .86, .84, .80, .77, .75, .75, .74, .74, .74, .73, .73, .73, .73, .72, .70, .71, .71, .71, .71,
          .70, .69, .70, .70, .70, .70, .70, .70, .70, .70, .68, .67, .67, .68, .69, .69, .69, .69, .69, .69]
    return (t1 if bg == 1 else t2)[mb]


def bler_at(bg, Z, mb, snr_db, n_cb, seed, iters=10):
    ns = V.nsys(bg)
    Kb = kb_for(bg, Z)
    F = (ns - Kb) * Z
    llr, info = V.make_cbs(bg, Z, mb, n_cb, seed=seed, snr_db=snr_db, filler=F)
    cfg = O.dec_cfg(bg, Z, mb, n_iter=iters, alpha=norm_table(bg, mb), clamp=32.0, kb_out=ns)
    hard = O.decode_cbs(cfg, llr, nthreads=os.cpu_count() or 1)
    errs = 0
    for i in range(n_cb):
        errs += int((O.unpack_bits(hard[i], ns * Z)[:Kb * Z] != info[i][:Kb * Z]).any())
    return errs / n_cb


FIX = json.load(open(os.path.join(HERE, "golden", "bler_thresholds.json")))
BLER_POINTS = [(1, 64, 4), (1, 64, 8), (1, 64, 46), (2, 64, 4), (2, 64, 42), (1, 128, 16), (2, 112, 24)]


@pytest.mark.parametrize("bg,Z,mb", BLER_POINTS)
def test_bler_threshold_fixture(bg, Z, mb):
    """At the fixture's SNR the reference decoder measured BLER <= 0.1 on 800
    code blocks (its SNR search targets 0.09).  Here: 1600 code blocks, and
    the bound is the fixture's plus three binomial standard deviations, so
    the test is a statistical check rather than a coin toss at the edge."""
    row = next(r for r in FIX if (r["bg"], r["Z"], r["mb"]) == (bg, Z, mb))
    n = 1600
    bler = bler_at(bg, Z, mb, row["snr_db"], n, seed=1000 + mb)
    p = row["max_bler"]
    assert bler <= p + 3.0 * np.sqrt(p * (1 - p) / n), (row, bler)
    # sensitivity: 1.5 dB below the threshold the BLER is well above it
    assert bler_at(bg, Z, mb, row["snr_db"] - 1.5, 200, seed=2000 + mb) > row["max_bler"]


# ---------------------------------------------------------------------------
# CRC: published catalogue check values (CRC of ASCII "123456789") and the
# self-check property
# ---------------------------------------------------------------------------
def ascii_bits(s):
    return np.unpackbits(np.frombuffer(s.encode(), np.uint8))


def test_crc_catalogue_check_values():
    b = ascii_bits("123456789")
    assert O.crc(b, O.CRC24A) == 0xCDE703  # CRC-24/LTE-A
    assert O.crc(b, O.CRC24B) == 0x23EF52  # CRC-24/LTE-B
    assert O.crc(b, O.CRC16) == 0x31C3     # CRC-16/XMODEM (TS 38.212 gCRC16, zero init)


@pytest.mark.parametrize("kind,L", [(O.CRC24A, 24), (O.CRC24B, 24), (O.CRC16, 16)])
def test_crc_appended_gives_zero_remainder(kind, L):
    rng = np.random.RandomState(kind)
    for n in (1, 7, 100, 3000):
        a = rng.randint(0, 2, n).astype(np.uint8)
        c = O.crc(a, kind)
        full = np.concatenate([a, [(c >> (L - 1 - i)) & 1 for i in range(L)]]).astype(np.uint8)
        assert O.crc(full, kind) == 0
        full[rng.randint(full.size)] ^= 1
        assert O.crc(full, kind) != 0


def test_crc_desegment_layout():
    # 2 CBs of 1000 data bits + CRC24B, TB CRC24A inside the last CB's data
    rng = np.random.RandomState(3)
    A = 2 * 1000 - 24
    a = rng.randint(0, 2, A).astype(np.uint8)
    c = O.crc(a, O.CRC24A)
    tb = np.concatenate([a, [(c >> (23 - i)) & 1 for i in range(24)]]).astype(np.uint8)
    words = np.zeros((2, 264), np.uint32)
    for r in range(2):
        seg = tb[r * 1000:(r + 1) * 1000]
        cc = O.crc(seg, O.CRC24B)
        cb = np.concatenate([seg, [(cc >> (23 - i)) & 1 for i in range(24)]]).astype(np.uint8)
        words[r, :32] = O.pack_bits(cb)
    cb_crc, tb_bytes, tb_crc = O.crc_desegment(words.reshape(-1), 2, 1024)
    assert list(cb_crc) == [0, 0] and tb_crc == 0
    assert np.array_equal(np.unpackbits(tb_bytes)[:tb.size], tb)  # MSB-first bytes, CBs back to back
    assert len(tb_bytes) % 4 == 0


# ---------------------------------------------------------------------------
# de-rate-matching: forward TS 38.212 5.4.2 rate matching (written here in
# the opposite direction) followed by the oracle's de-rate-matching
# ---------------------------------------------------------------------------
def rate_match_forward(w, E, Qm, k0, Ncb, fill_lo, fill_hi):
    """Bit selection (skipping filler <NULL>) then bit interleaving."""
    e, j = [], 0
    while len(e) < E:
        idx = (k0 + j) % Ncb
        if not (fill_lo <= idx < fill_hi):
            e.append(w[idx])
        j += 1
    e = np.array(e)
    f = np.empty(E, w.dtype)
    for jj in range(E // Qm):
        for k in range(Qm):
            f[jj * Qm + k] = e[k * (E // Qm) + jj]
    return f


# cases from cuPHY/test/derate_match_modulo/test_derate_match_modulo.cu:122-135
# style: (bg, Zc, Kprime, Qm, E, rv) with filler, k0 before/inside/after it
DRM_CASES = [(1, 384, 8448, 2, 12000, 0), (1, 104, 2200, 4, 3000, 2), (2, 52, 480, 6, 1200, 1),
             (2, 208, 1800, 8, 4096, 3), (1, 16, 300, 2, 500, 3)]


@pytest.mark.parametrize("case", DRM_CASES)
def test_derate_inverts_forward_rate_matching(case):
    bg, Zc, Kp, Qm, E, rv = case
    K = V.nsys(bg) * Zc
    N = (66 if bg == 1 else 50) * Zc
    F = K - Kp
    k0 = O.k0(bg, rv, N, Zc)
    w = (np.arange(N, dtype=np.float32) % 97) + 1.0   # distinct nonzero "LLRs" per circular-buffer position
    fill_lo, fill_hi = Kp - 2 * Zc, K - 2 * Zc
    f = rate_match_forward(w, E, Qm, k0, N, fill_lo, fill_hi)
    Ncb_padded = (N + 2 * Zc + 7) // 8 * 8
    cfg = O.rm_cfg(bg=bg, Zc=Zc, C=1, Qm=Qm, Nl=1, G=E, K=K, F=F, Ncb=N, Ncb_padded=Ncb_padded, k0=k0, ndi=1)
    assert O.rm_E(cfg, 0) == E
    harq = np.zeros(Ncb_padded, np.float32)
    O.derate_match(cfg, f, harq)
    out = harq[2 * Zc:2 * Zc + N]
    # expected: every selected position holds (number of times selected) * w
    count = np.zeros(N)
    j = 0
    sel = 0
    while sel < E:
        idx = (k0 + j) % N
        if not (fill_lo <= idx < fill_hi):
            count[idx] += 1
            sel += 1
        j += 1
    exp = np.minimum(count * w, 10000.0).astype(np.float32)
    exp[fill_lo:fill_hi] = 10000.0
    assert np.array_equal(out, exp)
    assert not harq[:2 * Zc].any()


def test_harq_combining_is_addition():
    bg, Zc, Kp, Qm, rv = 1, 64, 1400, 4, 0
    K, N = 22 * Zc, 66 * Zc
    Ncb_padded = (N + 2 * Zc + 7) // 8 * 8
    G = 3000
    rng = np.random.RandomState(9)
    a = rng.randint(-20, 20, G).astype(np.float32)  # integers: sums are exact
    b = rng.randint(-20, 20, G).astype(np.float32)
    k0 = O.k0(bg, rv, N, Zc)
    mk = lambda ndi: O.rm_cfg(bg=bg, Zc=Zc, C=1, Qm=Qm, Nl=1, G=G, K=K, F=K - Kp, Ncb=N, Ncb_padded=Ncb_padded,
                              k0=k0, ndi=ndi)
    h1 = np.zeros(Ncb_padded, np.float32)
    O.derate_match(mk(1), a, h1)
    O.derate_match(mk(0), b, h1)
    h2 = np.zeros(Ncb_padded, np.float32)
    O.derate_match(mk(1), a + b, h2)
    assert np.array_equal(h1, h2)


def test_rm_E_split_sums_to_G():
    for C, G, Qm, Nl in [(3, 3 * 4000 + 8, 2, 1), (7, 77777 // 24 * 24, 8, 3), (1, 640, 4, 2)]:
        cfg = O.rm_cfg(bg=1, Zc=384, C=C, Qm=Qm, Nl=Nl, G=G, K=8448, F=0, Ncb=25344, Ncb_padded=26112, k0=0, ndi=1)
        Es = [O.rm_E(cfg, r) for r in range(C)]
        assert sum(Es) == G and max(Es) - min(Es) <= Qm * Nl and all(e % (Qm * Nl) == 0 for e in Es)


def test_k0_table():
    # TS 38.212 Table 5.4.2.1-2
    assert [O.k0(1, rv, 66 * 384, 384) for rv in range(4)] == [0, 17 * 384, 33 * 384, 56 * 384]
    assert [O.k0(2, rv, 50 * 52, 52) for rv in range(4)] == [0, 13 * 52, 25 * 52, 43 * 52]
    # limited buffer: floor(num * Ncb / (den * Zc)) * Zc
    assert O.k0(1, 2, 20000, 384) == (33 * 20000 // (66 * 384)) * 384


# ---------------------------------------------------------------------------
# the reference's own encoder goldens (MATLAB LDPC_encode_main.m output,
# 5GModel/.../standalone_LDPC_encoder_test/mat_gen_data_1_1144_*.hdf5,
# extracted by tests/golden/extract_encoder_goldens.py): BG1, Zc = 52
# ---------------------------------------------------------------------------
def test_encoder_matches_reference_goldens():
    f = np.load(os.path.join(HERE, "golden", "encoder_goldens.npz"))
    n = 0
    for tag in ("1_0", "1_1", "752_0"):
        src, enc = f["src_" + tag], f["enc_" + tag]
        for i in range(src.shape[0]):
            cw = O.encode(1, 52, src[i])
            if enc.shape[1] == 66 * 52:  # first 2Zc columns punctured
                cw = cw[2 * 52:]
            assert np.array_equal(cw, enc[i]), (tag, i)
            n += 1
    assert n == 6


# ---------------------------------------------------------------------------
# Gold sequence (TS 38.211 5.2.1) and descrambling
# ---------------------------------------------------------------------------
def _gold_by_matrix_power(cinit, n):
    """c(n) from 31x31 GF(2) matrix powers of the two LFSRs (a method
    independent of the oracle's bit-serial recurrence)."""
    def step(taps):
        A = np.zeros((31, 31), np.int64)
        for i in range(30):
            A[i, i + 1] = 1
        for t in taps:
            A[30, t] = 1
        return A

    def power(A, e):
        R = np.eye(31, dtype=np.int64)
        while e:
            if e & 1:
                R = (R @ A) % 2
            A = (A @ A) % 2
            e >>= 1
        return R

    m = n + 1600
    s1 = np.zeros(31, np.int64)
    s1[0] = 1
    s2 = np.array([(cinit >> i) & 1 for i in range(31)], np.int64)
    x1 = (power(step([0, 3]), m) @ s1) % 2
    x2 = (power(step([0, 1, 2, 3]), m) @ s2) % 2
    return int(x1[0] ^ x2[0])


def test_gold_sequence_matches_matrix_power():
    rng = np.random.RandomState(7)
    for cinit in (0, 1, 0x12345, 1507348, 0x7FFFFFFF):
        ns = [0, 1, 30, 31, 32, 1000] + list(rng.randint(0, 2_000_000, 6))
        for n in ns:
            assert O.gold(cinit, int(n), 1)[0] == _gold_by_matrix_power(cinit, int(n)), (cinit, n)


def test_gold_sequence_windows_and_linearity():
    a, b = 0x1A2B3C, 0x0F0F0F
    full = O.gold(a, 0, 5000)
    assert np.array_equal(O.gold(a, 1234, 3000), full[1234:4234])  # any window of the same sequence
    # x2 is linear in c_init and x1 does not depend on it
    lhs = O.gold(a ^ b, 0, 4000) ^ O.gold(0, 0, 4000)
    rhs = O.gold(a, 0, 4000) ^ O.gold(b, 0, 4000)
    assert np.array_equal(lhs, rhs)
    assert 0.45 < full.mean() < 0.55


def test_descrambled_derate_match_is_derate_match_of_descrambled_llrs():
    # descramble_bits.m (LLR x (1 - 2c) over the TB) before derate_match.m
    bg, Zc, C, Qm, Nl, G = 1, 64, 3, 4, 2, 3 * 2400
    K = 22 * Zc
    cfg = O.rm_cfg(bg=bg, Zc=Zc, C=C, Qm=Qm, Nl=Nl, G=G, K=K, F=40, Ncb=66 * Zc, Ncb_padded=(68 * Zc + 7) // 8 * 8,
                   k0=O.lib().nr_orc_k0(bg, 2, 66 * Zc, Zc), ndi=1, descramble=1, cinit=0x4321)
    llr = np.random.RandomState(3).uniform(-30, 30, G).astype(np.float32)
    a = np.zeros(C * cfg.Ncb_padded, np.float32)
    O.derate_match(cfg, llr, a)
    c = O.gold(0x4321, 0, G)
    plain = O.rm_cfg(**{f: getattr(cfg, f) for f, _ in cfg._fields_})
    plain.descramble = 0
    b = np.zeros_like(a)
    O.derate_match(plain, llr * (1.0 - 2.0 * c.astype(np.float32)), b)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
