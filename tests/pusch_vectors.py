"""Seeded PUSCH transport-block vectors for the fused-slot and serving tests
(CPU, via the oracle): TB payload + CRC24A/CRC16, segmentation with CRC24B
(TS 38.212 5.2.2), LDPC encoding, forward rate matching (bit selection from
k0 skipping the filler bits, then bit interleaving, TS 38.212 5.4.2) and
BPSK-AWGN channel LLRs per transmitted bit, laid out like the equalizer
output the de-rate-matcher reads (QAM_STRIDE = 8 slots per symbol)."""
import json
import os

import numpy as np

import oracle as O

QAM_STRIDE = 8


def nsys(bg):
    return 22 if bg == 1 else 10


def tb_payload_bits(C, Kp):
    """Payload size A of a TB whose C CBs carry K' bits each (C > 1: K' - 24
    data bits per CB, TB CRC24A inside; C = 1: the TB CRC fills the rest)."""
    if C > 1:
        return C * (Kp - 24) - 24
    return Kp - 16 if Kp - 16 <= 3824 else Kp - 24


class TbVector:
    """One TB: payload, segmented/CRC'd/encoded code blocks."""

    def __init__(self, bg, Zc, C, Kp, seed):
        self.bg, self.Zc, self.C, self.Kp = bg, Zc, C, Kp
        K = nsys(bg) * Zc
        assert Kp <= K and Kp % 8 == 0
        self.K, self.F = K, K - Kp
        self.N = (66 if bg == 1 else 50) * Zc
        A = tb_payload_bits(C, Kp)
        self.A = A
        self.a = O.random_bits(seed, A)
        L = 24 if (C > 1 or A > 3824) else 16
        crc = O.crc(self.a, O.CRC24A if L == 24 else O.CRC16)
        tb = np.concatenate([self.a, np.array([(crc >> (L - 1 - i)) & 1 for i in range(L)], np.uint8)])
        self.cb_bits = []
        if C == 1:
            segs = [tb]
        else:
            d = Kp - 24
            segs = []
            for r in range(C):
                s = tb[r * d:(r + 1) * d]
                c = O.crc(s, O.CRC24B)
                segs.append(np.concatenate([s, np.array([(c >> (23 - i)) & 1 for i in range(24)], np.uint8)]))
        self.cws = np.zeros((C, (nsys(bg) + (46 if bg == 1 else 42)) * Zc), np.uint8)
        for r, s in enumerate(segs):
            assert s.size == Kp
            info = np.zeros(K, np.uint8)
            info[:Kp] = s
            self.cb_bits.append(s)
            self.cws[r] = O.encode(bg, Zc, info)

    def circular_buffer(self, r):
        """w: the N coded bits after the 2Zc punctured ones (filler bits at
        [K' - 2Zc, K - 2Zc) are <NULL>: never selected)."""
        return self.cws[r, 2 * self.Zc:2 * self.Zc + self.N]

    def tx_bits(self, G, Qm, Nl, rv):
        """Rate-matched, interleaved bits of the whole TB for one
        transmission: CB r contributes E_r bits (TS 38.212 5.4.2.1)."""
        cfg = O.rm_cfg(bg=self.bg, Zc=self.Zc, C=self.C, Qm=Qm, Nl=Nl, G=G, K=self.K, F=self.F, Ncb=self.N,
                       Ncb_padded=0, k0=0, ndi=1)
        k0 = O.k0(self.bg, rv, self.N, self.Zc)
        lo, hi = self.Kp - 2 * self.Zc, self.K - 2 * self.Zc
        out = []
        for r in range(self.C):
            E = O.rm_E(cfg, r)
            w = self.circular_buffer(r)
            idx = (k0 + np.arange((E // (self.N - self.F) + 2) * self.N)) % self.N
            idx = idx[(idx < lo) | (idx >= hi)][:E]
            assert idx.size == E
            e = w[idx]
            f = e.reshape(Qm, E // Qm).T.reshape(-1)  # f[j*Qm + k] = e[k*E/Qm + j]
            out.append(f)
        return np.concatenate(out)


def bpsk_llrs(bits, snr_db, seed):
    """Per-bit BPSK over real AWGN at Es/N0 = snr_db: LLR = 2y / sigma^2."""
    sigma2 = 10.0 ** (-snr_db / 10.0) / 2.0
    return O.bpsk_awgn_llr(bits, sigma2, seed)


def qam_stride_layout(llr_demap, Qm, Nl, n_bbu, layer_map, fill=0.0):
    """Equalizer-output layout: symbol m of the TB at RE m // Nl, layer
    layer_map[m % Nl], QAM_STRIDE LLR slots per symbol."""
    nsym = llr_demap.size // Qm
    nre = (nsym + Nl - 1) // Nl
    buf = np.full((nre, n_bbu, QAM_STRIDE), fill, np.float32)
    m = np.arange(nsym)
    buf[m // Nl, np.asarray(layer_map)[m % Nl], :Qm] = llr_demap[:nsym * Qm].reshape(nsym, Qm)
    return buf.reshape(-1)


def quantize_int8(llr, steps_per_unit=4):
    """int8 LLRs saturated at +-127 with `steps_per_unit` steps per unit LLR
    (numpy rounding, half to even); the value seen downstream is q / steps."""
    q = np.clip(np.round(np.asarray(llr, np.float64) * steps_per_unit), -127, 127).astype(np.int8)
    return q


_NORM = None


def table_alpha(bg, mb):
    """The reference's min-sum normalisation for (BG, parity nodes), rounded
    to binary16 (fixture golden/min_sum_norm.json, extracted from
    ldpc.cpp:139-235 and LDPC_decode.m:101-208)."""
    global _NORM
    if _NORM is None:
        _NORM = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                            "min_sum_norm.json")))
    return float(np.float16(_NORM["BG%d" % bg][mb]))
