"""ctypes bindings of the CPU oracle (oracle/build/libnr_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libnr_oracle.so")

FP32, FP64, FP16 = 0, 1, 2
CRC24A, CRC24B, CRC16 = 0, 1, 2
WORDS_PER_CB = 264


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class DecCfg(C.Structure):
    _fields_ = [("bg", C.c_int), ("Z", C.c_int), ("mb", C.c_int), ("kb_out", C.c_int),
                ("max_iter", C.c_int), ("early_term", C.c_int), ("mode", C.c_int),
                ("alpha", C.c_float), ("clamp", C.c_float)]


class Coding(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("tbs", "bg", "crc_len", "B", "C", "Kp", "Kb", "Zc", "ils", "K", "F", "N",
                                       "tbs_lbrm", "Nref")]


class RmCfg(C.Structure):
    _fields_ = [("bg", C.c_int), ("Zc", C.c_int), ("C", C.c_int), ("Qm", C.c_int), ("Nl", C.c_int),
                ("G", C.c_int), ("K", C.c_int), ("F", C.c_int), ("Ncb", C.c_int),
                ("Ncb_padded", C.c_int), ("k0", C.c_int), ("ndi", C.c_int), ("descramble", C.c_int),
                ("cinit", C.c_uint)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.nr_orc_ils.argtypes = [C.c_int]
        L.nr_orc_num_rows.argtypes = [C.c_int]
        L.nr_orc_row_degree.argtypes = [C.c_int, C.c_int]
        L.nr_orc_edge.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_int), P(C.c_int)]
        L.nr_orc_random_bits.argtypes = [C.c_uint64, C.c_void_p, C.c_int]
        L.nr_orc_bpsk_awgn_llr.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_void_p]
        L.nr_orc_qpsk_awgn_llr.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_void_p]
        L.nr_orc_consistent_gaussian_llr.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.c_void_p]
        L.nr_orc_encode.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.nr_orc_syndrome.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.nr_orc_decode_cb.argtypes = [P(DecCfg), C.c_void_p, C.c_void_p, C.c_void_p]
        L.nr_orc_decode_cbs.argtypes = [P(DecCfg), C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                                        C.c_size_t, C.c_int]
        L.nr_orc_decode_cbs_it.argtypes = [P(DecCfg), C.c_void_p, C.c_size_t, C.c_int, C.c_void_p,
                                           C.c_size_t, C.c_void_p, C.c_int, C.c_int]
        L.nr_orc_num_parity_nodes.argtypes = [C.c_int] * 7
        L.nr_orc_bit_select.argtypes = [C.c_int] * 5 + [C.c_void_p]
        L.nr_orc_bit_select.restype = None
        L.nr_orc_tbs_table.argtypes = [C.c_int]
        L.nr_orc_tb_size.argtypes = [C.c_double, C.c_int, C.c_double, C.c_int, C.c_int]
        L.nr_orc_derive_coding.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, P(Coding)]
        L.nr_orc_check_node.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.nr_orc_round_half.argtypes = [C.c_float]
        L.nr_orc_round_half.restype = C.c_float
        L.nr_orc_crc.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.nr_orc_crc.restype = C.c_uint32
        L.nr_orc_rm_E.argtypes = [P(RmCfg), C.c_int]
        L.nr_orc_derate_match.argtypes = [P(RmCfg), C.c_void_p, C.c_void_p]
        L.nr_orc_gold.argtypes = [C.c_uint, C.c_int, C.c_int, C.c_void_p]
        L.nr_orc_gold.restype = None
        L.nr_orc_derate_match_fp16.argtypes = [P(RmCfg), C.c_void_p, C.c_void_p]
        L.nr_orc_k0.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.nr_orc_crc_desegment.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.nr_orc_time_decode.argtypes = [P(DecCfg), C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_size_t,
                                         C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.nr_orc_time_slot.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                       C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        _lib = L
    return _lib


def use_library(path):
    """Load the oracle from another build of the same sources (the CPU
    baseline's -march=native build); later calls go to it."""
    global _lib, LIB_PATH
    _lib = None
    LIB_PATH = path
    return lib()


def time_decode(cfg, llr2d, nthreads, warmup=2, reps=20):
    """Seconds of each of reps pinned decodes of the rows of llr2d (after
    warmup untimed ones) and the last decode's hard words."""
    llr2d = np.ascontiguousarray(llr2d, np.float32)
    ncb = llr2d.shape[0]
    nw = (cfg.kb_out * cfg.Z + 31) // 32
    hard = np.zeros((ncb, nw), np.uint32)
    times = np.zeros(reps, np.float64)
    rc = lib().nr_orc_time_decode(C.byref(cfg), _p(llr2d), llr2d.shape[1], ncb, _p(hard), nw, nthreads, warmup,
                                  reps, _p(times))
    assert rc == 0
    return times, hard


def time_slot(cfgs, llrs, nthreads, warmup=2, reps=20):
    """Seconds of each of reps pinned passes over a slot: TB t's code blocks
    are the rows of llrs[t], decoded with cfgs[t]; all code blocks share one
    queue over nthreads threads.  Also returns each TB's hard words."""
    n = len(cfgs)
    llrs = [np.ascontiguousarray(x, np.float32) for x in llrs]
    hards = [np.zeros((x.shape[0], (c.kb_out * c.Z + 31) // 32), np.uint32) for c, x in zip(cfgs, llrs)]
    cfg_arr = (DecCfg * n)(*cfgs)
    llr_p = (C.c_void_p * n)(*[x.ctypes.data for x in llrs])
    llr_s = (C.c_size_t * n)(*[x.shape[1] for x in llrs])
    ncbs = (C.c_int * n)(*[x.shape[0] for x in llrs])
    hard_p = (C.c_void_p * n)(*[h.ctypes.data for h in hards])
    hard_s = (C.c_size_t * n)(*[h.shape[1] for h in hards])
    times = np.zeros(reps, np.float64)
    rc = lib().nr_orc_time_slot(C.cast(cfg_arr, C.c_void_p), C.cast(llr_p, C.c_void_p), C.cast(llr_s, C.c_void_p),
                                C.cast(ncbs, C.c_void_p), n, C.cast(hard_p, C.c_void_p),
                                C.cast(hard_s, C.c_void_p), nthreads, warmup, reps, _p(times))
    assert rc == 0
    return times, hards


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def nsys(bg):
    return 22 if bg == 1 else 10


def nrows(bg):
    return 46 if bg == 1 else 42


def ils(Z):
    return lib().nr_orc_ils(Z)


def edges(bg, Z, r):
    """[(col, shift mod Z)] of row r."""
    L = lib()
    out = []
    c, s = C.c_int(), C.c_int()
    for j in range(L.nr_orc_row_degree(bg, r)):
        L.nr_orc_edge(bg, Z, r, j, C.byref(c), C.byref(s))
        out.append((c.value, s.value))
    return out


def random_bits(seed, n):
    b = np.empty(n, np.uint8)
    lib().nr_orc_random_bits(seed, _p(b), n)
    return b


def encode(bg, Z, info):
    info = np.ascontiguousarray(info, np.uint8)
    assert info.size == nsys(bg) * Z
    cw = np.empty((nsys(bg) + nrows(bg)) * Z, np.uint8)
    rc = lib().nr_orc_encode(bg, Z, _p(info), _p(cw))
    assert rc == 0, rc
    return cw


def syndrome(bg, Z, mb, x):
    x = np.ascontiguousarray(x, np.uint8)
    return lib().nr_orc_syndrome(bg, Z, mb, _p(x))


def bpsk_awgn_llr(bits, sigma2, seed):
    bits = np.ascontiguousarray(bits, np.uint8)
    out = np.empty(bits.size, np.float32)
    lib().nr_orc_bpsk_awgn_llr(_p(bits), bits.size, sigma2, seed, _p(out))
    return out


def qpsk_awgn_llr(bits, snr_db, seed):
    bits = np.ascontiguousarray(bits, np.uint8)
    out = np.empty(bits.size, np.float32)
    lib().nr_orc_qpsk_awgn_llr(_p(bits), bits.size, snr_db, seed, _p(out))
    return out


def consistent_gaussian_llr(bits, mu, seed):
    bits = np.ascontiguousarray(bits, np.uint8)
    out = np.empty(bits.size, np.float32)
    lib().nr_orc_consistent_gaussian_llr(_p(bits), bits.size, mu, seed, _p(out))
    return out


def dec_cfg(bg, Z, mb, n_iter=10, alpha=0.8125, clamp=32.0, mode=FP32, kb_out=None, early_term=False):
    return DecCfg(bg, Z, mb, nsys(bg) if kb_out is None else kb_out, n_iter, int(early_term), mode,
                  alpha, clamp)


def decode_cb(cfg, llr, want_app=False):
    """-> (hard words uint32, iterations run, app or None)"""
    llr = np.ascontiguousarray(llr, np.float32)
    n = (nsys(cfg.bg) + cfg.mb) * cfg.Z
    assert llr.size >= n
    nw = (cfg.kb_out * cfg.Z + 31) // 32
    hard = np.zeros(nw, np.uint32)
    app = np.empty(n, np.float32) if want_app else None
    it = lib().nr_orc_decode_cb(C.byref(cfg), _p(llr), _p(hard), _p(app) if want_app else None)
    assert it >= 0, it
    return hard, it, app


def decode_cbs(cfg, llr2d, nthreads=1, hard_stride=None):
    """Decode every row of llr2d; returns hard words (ncb, hard_stride)."""
    llr2d = np.ascontiguousarray(llr2d, np.float32)
    ncb = llr2d.shape[0]
    nw = (cfg.kb_out * cfg.Z + 31) // 32
    hs = nw if hard_stride is None else hard_stride
    hard = np.zeros((ncb, hs), np.uint32)
    rc = lib().nr_orc_decode_cbs(C.byref(cfg), _p(llr2d), llr2d.shape[1], ncb, _p(hard), hs, nthreads)
    assert rc == 0
    return hard


def decode_cbs_it(cfg, llr2d, nthreads=1, hard_stride=None, pin=False):
    """Decode every row of llr2d on nthreads host threads (pinned to CPUs
    when pin); returns (hard words (ncb, hard_stride), iterations (ncb,))."""
    llr2d = np.ascontiguousarray(llr2d, np.float32)
    ncb = llr2d.shape[0]
    nw = (cfg.kb_out * cfg.Z + 31) // 32
    hs = nw if hard_stride is None else hard_stride
    hard = np.zeros((ncb, hs), np.uint32)
    iters = np.zeros(ncb, np.int32)
    rc = lib().nr_orc_decode_cbs_it(C.byref(cfg), _p(llr2d), llr2d.shape[1], ncb, _p(hard), hs, _p(iters),
                                    nthreads, int(pin))
    assert rc == 0
    return hard, iters


def tbs_table():
    return [lib().nr_orc_tbs_table(i) for i in range(93)]


def tb_size(re_per_prb, n_prb, R, Qm, nl):
    """TS 38.214 5.1.3.2 TBS (derive_coding_main.m:78-125)."""
    return lib().nr_orc_tb_size(float(re_per_prb), n_prb, float(R), Qm, nl)


def derive_coding(tbs, R, i_lbrm=0, n_prb_lbrm=273, max_layers=4, max_qm=8):
    """Coding parameters of a TB (derive_coding_main.m:127-301) as a dict."""
    o = Coding()
    rc = lib().nr_orc_derive_coding(tbs, float(R), i_lbrm, n_prb_lbrm, max_layers, max_qm, C.byref(o))
    assert rc == 0, (tbs, R)
    return {n: getattr(o, n) for n, _ in o._fields_}


def bit_select(E, Kd, F, k0, Ncb):
    """Circular-buffer positions of the E selected bits (filler [Kd, Kd+F) skipped)."""
    pos = np.zeros(E, np.int32)
    lib().nr_orc_bit_select(E, Kd, F, k0, Ncb, _p(pos))
    return pos


def num_parity_nodes(bg, Zc, K, F, G, C_, k0):
    """Parity nodes of a transmission (derate_match.m:122-129)."""
    return lib().nr_orc_num_parity_nodes(bg, Zc, K, F, G, C_, k0)


def check_node(t, mode=FP32):
    t = np.ascontiguousarray(t, np.float32)
    out = np.empty_like(t)
    lib().nr_orc_check_node(t.size, _p(t), _p(out), mode)
    return out


def round_half(x):
    return np.float32(lib().nr_orc_round_half(float(x)))


def crc(bits, kind):
    bits = np.ascontiguousarray(bits, np.uint8)
    return int(lib().nr_orc_crc(_p(bits), bits.size, kind))


def k0(bg, rv, Ncb, Zc):
    return lib().nr_orc_k0(bg, rv, Ncb, Zc)


def rm_cfg(**kw):
    return RmCfg(**kw)


def gold(cinit, n0, nbits):
    """TS 38.211 5.2.1 Gold sequence bits c(n0 .. n0+nbits-1), uint8."""
    out = np.zeros(nbits, np.uint8)
    lib().nr_orc_gold(cinit, n0, nbits, out.ctypes.data)
    return out


def rm_E(cfg, r):
    return lib().nr_orc_rm_E(C.byref(cfg), r)


def derate_match(cfg, llr_demap, harq, fp16=False):
    """Update harq (C*Ncb_padded float32, in place) with the TB's G LLRs."""
    llr_demap = np.ascontiguousarray(llr_demap, np.float32)
    assert harq.dtype == np.float32 and harq.flags.c_contiguous
    assert harq.size >= cfg.C * cfg.Ncb_padded
    f = lib().nr_orc_derate_match_fp16 if fp16 else lib().nr_orc_derate_match
    f(C.byref(cfg), _p(llr_demap), _p(harq))
    return harq


def crc_desegment(cb_words, C_, Kp):
    cb_words = np.ascontiguousarray(cb_words, np.uint32)
    assert cb_words.size >= C_ * WORDS_PER_CB
    cb_crc = np.zeros(C_, np.uint32)
    tb = np.zeros(C_ * WORDS_PER_CB * 4 + 8, np.uint8)
    tb_crc = np.zeros(1, np.uint32)
    n = lib().nr_orc_crc_desegment(_p(cb_words), C_, Kp, _p(cb_crc), _p(tb), _p(tb_crc))
    return cb_crc, tb[:n].copy(), int(tb_crc[0])


def pack_bits(bits):
    """0/1 array -> uint32 words, bit i of word w = bits[32w + i]."""
    bits = np.asarray(bits, np.uint8)
    nw = (bits.size + 31) // 32
    padded = np.zeros(nw * 32, np.uint8)
    padded[:bits.size] = bits
    return np.packbits(padded, bitorder="little").view("<u4").copy()


def unpack_bits(words, n):
    words = np.ascontiguousarray(words, "<u4")
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].copy()
