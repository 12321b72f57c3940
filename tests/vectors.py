"""Seeded test-vector builders shared by the tests (CPU, via the oracle).

Code block i of a vector set uses info-bit seed `seed * 1000 + 2*i + 1` and
noise seed `seed * 1000 + 2*i + 2` (xoshiro256** seeded by splitmix64, in
the oracle library), so every test input is reproducible bit for bit.
"""
import numpy as np

import oracle as O

# BPSK over real AWGN at Es/N0 = 1.0 dB: sigma^2 = 10^(-0.1) / 2
C1_SIGMA2 = 10.0 ** (-0.1) / 2.0


def nsys(bg):
    return 22 if bg == 1 else 10


def cb_seeds(seed, i):
    return seed * 1000 + 2 * i + 1, seed * 1000 + 2 * i + 2


def make_cbs(bg, Z, mb, n_cb, seed, sigma2=None, snr_db=None, mu=None, filler=0):
    """-> (llr (n_cb, (nsys+mb)*Z) float32 with 2Z punctured zeros first,
           info bits (n_cb, nsys*Z) uint8)

    filler: number of filler bits at the end of the systematic part (zero
    bits, LLR = +inf before clamping, as cuphy_ex_ldpc's generator does,
    ldpc_decode_test_vec_gen.cpp:218-227)."""
    ns = nsys(bg)
    n_tx = (ns + mb - 2) * Z
    llr = np.zeros((n_cb, (ns + mb) * Z), np.float32)
    info = np.zeros((n_cb, ns * Z), np.uint8)
    for i in range(n_cb):
        sb, sn = cb_seeds(seed, i)
        bits = O.random_bits(sb, ns * Z)
        if filler:
            bits[ns * Z - filler:] = 0
        cw = O.encode(bg, Z, bits)
        tx = cw[2 * Z:2 * Z + n_tx]
        if sigma2 is not None:
            y = O.bpsk_awgn_llr(tx, sigma2, sn)
        elif snr_db is not None:
            y = O.qpsk_awgn_llr(tx, snr_db, sn)
        elif mu is not None:
            y = O.consistent_gaussian_llr(tx, mu, sn)
        else:  # noiseless +-4
            y = (4.0 * (1.0 - 2.0 * tx.astype(np.float32))).astype(np.float32)
        llr[i, 2 * Z:] = y
        if filler:
            llr[i, ns * Z - filler:ns * Z] = np.inf
        info[i] = bits
    return llr, info


def config1_vectors():
    """BG1, Z=384, mb=46, 8 CBs, BPSK-AWGN at Es/N0 = 1 dB, FP32 LLRs."""
    return make_cbs(1, 384, 46, 8, seed=1, sigma2=C1_SIGMA2)


def to_half_rn(x):
    """float32 -> float16 round-to-nearest-even (numpy's conversion)."""
    return np.asarray(x, np.float32).astype(np.float16)
