"""Synthetic PUSCH transport blocks for the fused slot and the serving loop,
generated as a transmitter would produce them:

  payload -> TB CRC (CRC24A, or CRC16 for a one-CB TB of <= 3824 bits)
  -> segmentation with CRC24B per CB (TS 38.212 5.2.2), filler 0s
  -> LDPC encoding (the library's HIP encoder)
  -> rate matching per transmission: bit selection from k0(rv) around the
     circular buffer skipping filler bits, bit interleaving (TS 38.212 5.4.2)
  -> per-bit channel LLRs (BPSK/AWGN 2y/sigma^2, or consistent Gaussian
     N(+-mu, 2mu)), optionally int8 (saturated, `steps` per unit LLR)
  -> the de-rate-matcher's input layout (QAM_STRIDE = 8 slots per symbol).

CRC attachment uses a byte-table CRC vectorised over many messages
(numpy); everything per bit runs on the device (torch).
"""
from dataclasses import dataclass, field

import numpy as np
import torch

from .pusch import make_tb_params, k0_of
from .workloads import encode, nsys

QAM_STRIDE = 8
CRC24A_POLY, CRC24B_POLY, CRC16_POLY = 0x864CFB, 0x800063, 0x1021


def _crc_table(poly, L):
    t = np.zeros(256, np.uint32)
    mask = (1 << L) - 1
    for v in range(256):
        r = v << (L - 8)
        for _ in range(8):
            r = ((r << 1) & mask) ^ (poly if (r >> (L - 1)) & 1 else 0)
        t[v] = r
    return t


_TABLES = {}


def crc_many(bits, poly, L):
    """CRC remainders (zero initial value) of each row of a 0/1 matrix whose
    width is a multiple of 8 (MSB-first bytes)."""
    key = (poly, L)
    if key not in _TABLES:
        _TABLES[key] = _crc_table(poly, L)
    T = _TABLES[key]
    by = np.packbits(np.asarray(bits, np.uint8), axis=1)
    reg = np.zeros(by.shape[0], np.uint32)
    mask = np.uint32((1 << L) - 1)
    for j in range(by.shape[1]):
        reg = ((reg << np.uint32(8)) & mask) ^ T[((reg >> np.uint32(L - 8)) ^ by[:, j]) & 0xFF]
    return reg


def _append_crc(msgs, poly, L):
    crc = crc_many(msgs, poly, L)
    tail = ((crc[:, None] >> np.arange(L - 1, -1, -1, dtype=np.uint32)[None, :]) & 1).astype(np.uint8)
    return np.concatenate([msgs, tail], axis=1)


@dataclass
class PuschTb:
    """One TB with its code words on the device and per-transmission LLRs."""
    bg: int
    Zc: int
    C: int
    Kp: int
    Qm: int
    Nl: int
    E: int                       # rate-matched bits per CB (all CBs equal)
    snr_db: float
    payload: np.ndarray = None   # A payload bits (host)
    cw: torch.Tensor = None      # (C, (nsys + nrows) * Zc) uint8 code words (device)
    llrs: dict = field(default_factory=dict)  # rv -> input-layout LLR tensor

    @property
    def G(self):
        return self.C * self.E

    @property
    def A(self):
        return self.payload.size

    @property
    def info_bits(self):
        return self.C * self.Kp

    def params(self, rv, ndi, user_group=0):
        return make_tb_params(self.bg, self.Zc, self.C, self.Kp, self.Qm, self.Nl, self.G, rv, ndi=ndi,
                              user_group=user_group)


def build_tbs(specs, seed):
    """specs: (bg, Zc, C, Kp, Qm, Nl, E, snr_db) per TB -> PuschTb list with
    code words on the device (payloads and CRCs drawn from `seed`)."""
    rng = np.random.RandomState(seed)
    tbs = []
    groups = {}
    for i, s in enumerate(specs):
        tb = PuschTb(*s)
        bg, Zc, C, Kp = tb.bg, tb.Zc, tb.C, tb.Kp
        K = nsys(bg) * Zc
        assert Kp <= K and Kp % 8 == 0 and (C == 1 or (Kp - 24) % 8 == 0)
        A = C * (Kp - 24) - 24 if C > 1 else (Kp - 16 if Kp - 16 <= 3824 else Kp - 24)
        tb.payload = rng.randint(0, 2, A).astype(np.uint8)
        tbs.append(tb)
        groups.setdefault((bg, Zc, Kp, C > 1), []).append(i)
    # TB CRC per TB (vectorised over TBs of equal payload size), CB CRC per CB
    by_len = {}
    for i, tb in enumerate(tbs):
        by_len.setdefault(tb.A, []).append(i)
    tb_bits = [None] * len(tbs)
    for A, idx in by_len.items():
        first = tbs[idx[0]]
        L = 24 if (first.C > 1 or A > 3824) else 16
        poly = CRC24A_POLY if L == 24 else CRC16_POLY
        pad = (-A) % 8  # leading zeros do not change a zero-initial CRC
        m = np.stack([np.concatenate([np.zeros(pad, np.uint8), tbs[i].payload]) for i in idx])
        withcrc = _append_crc(m, poly, L)[:, pad:]
        for k, i in enumerate(idx):
            tb_bits[i] = withcrc[k]
    for key, idx in groups.items():
        bg, Zc, Kp, seg = key
        K = nsys(bg) * Zc
        rows = []
        for i in idx:
            tb = tbs[i]
            if seg:
                d = Kp - 24
                rows.append(tb_bits[i][:tb.C * d].reshape(tb.C, d))
            else:
                rows.append(tb_bits[i][None, :])
        segs = np.concatenate(rows)
        if seg:
            segs = _append_crc(segs, CRC24B_POLY, 24)
        info = np.zeros((segs.shape[0], K), np.uint8)
        info[:, :Kp] = segs
        cw = encode(bg, Zc, torch.from_numpy(info).cuda())
        off = 0
        for i in idx:
            tbs[i].cw = cw[off:off + tbs[i].C]
            off += tbs[i].C
    return tbs


_SEL_CACHE = {}


def _selection(bg, Zc, Kp, E, rv):
    """Circular-buffer indices of the E selected bits (filler skipped)."""
    key = (bg, Zc, Kp, E, rv)
    if key not in _SEL_CACHE:
        K = nsys(bg) * Zc
        N = (66 if bg == 1 else 50) * Zc
        F = K - Kp
        k0 = k0_of(bg, rv, N, Zc)
        lo, hi = Kp - 2 * Zc, K - 2 * Zc
        j = (k0 + np.arange((E // (N - F) + 2) * N)) % N
        j = j[(j < lo) | (j >= hi)][:E]
        _SEL_CACHE[key] = torch.from_numpy(j.astype(np.int64)).cuda()
    return _SEL_CACHE[key]


def tx_llrs(tb, rv, gen, mode="bpsk", dtype=torch.float16, steps=4.0):
    """This transmission's LLRs in the de-rate-matcher's input layout.

    mode "bpsk": LLR = 2y/sigma^2, y = +-1 + N(0, sigma^2), sigma^2 =
    10^(-snr/10)/2; mode "cg": consistent Gaussian N(+-mu, 2mu) with
    mu = 4 * 10^(snr/10) (the BPSK LLR statistics at that Es/N0).
    dtype int8: round(steps * LLR) saturated at +-127."""
    Zc = tb.Zc
    sel = _selection(tb.bg, Zc, tb.Kp, tb.E, rv)
    w = tb.cw[:, 2 * Zc:]
    e = w[:, sel]                                               # (C, E)
    f = e.view(tb.C, tb.Qm, tb.E // tb.Qm).transpose(1, 2).reshape(-1)  # interleaved, TB order
    x = 1.0 - 2.0 * f.to(torch.float32)
    n = torch.randn(x.shape, generator=gen, device="cuda", dtype=torch.float32)
    if mode == "bpsk":
        sigma2 = 10.0 ** (-tb.snr_db / 10.0) / 2.0
        llr = (x + n * float(np.sqrt(sigma2))) * (2.0 / sigma2)
    else:
        mu = 4.0 * 10.0 ** (tb.snr_db / 10.0)
        llr = mu * x + float(np.sqrt(2.0 * mu)) * n
    nsym = tb.G // tb.Qm
    buf = torch.zeros((nsym, QAM_STRIDE), dtype=torch.float32, device="cuda")
    buf[:, :tb.Qm] = llr.view(nsym, tb.Qm)
    buf = buf.reshape(-1)
    if dtype == torch.int8:
        return torch.clamp(torch.round(buf * steps), -127, 127).to(torch.int8)
    return buf.to(dtype)


def fused_slot_specs(seed=5, n_cells=64, bg1_cbs=140, bg2_tbs_per_cell=28, snr_range=(18.0, 24.0)):
    """Fused receive-side slot: per cell one BG1 Zc=384 TB of 140 CBs at
    R = 0.93 (256QAM, E = 9088) plus 28 one-CB BG2 TBs (Zc 104 / 208
    alternating, R = 0.5; 20% on top of the BG1 CBs), with a
    per-cell SNR drawn uniformly from snr_range (dB).  LLRs are consistent
    Gaussian with mu = SNR_linear / 5 (the 20-cell full-load slot's model),
    i.e. a BPSK-equivalent Es/N0 of SNR - 10*log10(20) dB."""
    rng = np.random.RandomState(seed)
    specs = []
    for _ in range(n_cells):
        snr = float(rng.uniform(*snr_range)) - 10.0 * np.log10(20.0)
        specs.append((1, 384, bg1_cbs, 8448, 8, 1, 9088, snr))
        for k in range(bg2_tbs_per_cell):
            if k % 2 == 0:
                specs.append((2, 104, 1, 1040, 8, 1, 2080, snr))
            else:
                specs.append((2, 208, 1, 2080, 8, 1, 4160, snr))
    return specs
