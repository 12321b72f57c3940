"""Synthetic uplink slots for benchmarking, generated on the device.

Info bits come from a seeded torch generator, code blocks are encoded with
the library's HIP encoder (cuphyHipLDPCEncodeCodeBlocks), transmitted over
BPSK/AWGN with torch's device RNG, and turned into LLRs 2y/sigma^2 in the
decoder input layout (2Z punctured zeros first, untransmitted parity 0).
Parameters of each transport block are drawn with numpy from a seed.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L

# TS 38.212 lifting sizes (Table 5.3.2-1)
LIFTING_SIZES = sorted({a << j for a in (2, 3, 5, 7, 9, 11, 13, 15) for j in range(8) if (a << j) <= 384})

# Edges of the first mb rows (row degrees of TS 38.212 Tables 5.3.2-2/3)
_BG1_DEG = [19, 19, 19, 19, 3, 8, 9, 7, 10, 9, 7, 8, 7, 6, 7, 7, 6, 6, 6, 6, 6, 6, 5, 5, 6, 5, 5, 4, 5, 5, 5, 5, 5,
            5, 5, 5, 5, 4, 5, 5, 4, 5, 4, 5, 5, 4]
_BG2_DEG = [8, 10, 8, 10, 4, 6, 6, 6, 4, 5, 5, 5, 4, 5, 5, 4, 5, 5, 4, 4, 4, 4, 3, 4, 4, 3, 5, 3, 4, 3, 5, 3, 4, 4, 4,
            4, 4, 3, 4, 4, 4, 4]


def num_edges(bg, mb):
    return sum((_BG1_DEG if bg == 1 else _BG2_DEG)[:mb])


def nsys(bg):
    return 22 if bg == 1 else 10


def nrows(bg):
    return 46 if bg == 1 else 42


def mb_for_rate(bg, Z, K, R):
    """Parity rows needed for N = ceil(K/R) transmitted bits (first 2Z punctured):
    mb = clamp(ceil((N - K)/Z) + 2, 4, 46/42) (cuphy_ex_ldpc, ldpc_decode_test_vec_gen.cpp:145)."""
    N = int(np.ceil(K / R))
    mb = int(np.ceil((N - K) / Z)) + 2
    return max(4, min(mb, nrows(bg))), N


@dataclass
class TransportBlock:
    bg: int
    Z: int
    mb: int
    ncb: int
    n_tx: int                 # transmitted LLRs per code block
    esn0_db: float
    llr: torch.Tensor = None  # (ncb, stride) float16/float32, device
    info: torch.Tensor = None  # (ncb, nsys*Z) uint8, device
    out: torch.Tensor = None  # (ncb, words) int32, device
    extra: dict = field(default_factory=dict)

    @property
    def info_bits_per_cb(self):
        return nsys(self.bg) * self.Z

    @property
    def num_llrs(self):
        return (nsys(self.bg) + self.mb) * self.Z

    @property
    def out_words(self):
        return (nsys(self.bg) * self.Z + 31) // 32

    def edge_updates(self, iters):
        return self.ncb * self.Z * num_edges(self.bg, self.mb) * iters


def encode(bg, Z, info, stream=None):
    """info (ncb, nsys*Z) uint8 device -> codewords (ncb, (nsys+nrows)*Z) uint8."""
    ncb = info.shape[0]
    cw = torch.empty((ncb, (nsys(bg) + nrows(bg)) * Z), dtype=torch.uint8, device=info.device)
    s = stream if stream is not None else torch.cuda.current_stream()
    st = L.lib().cuphyHipLDPCEncodeCodeBlocks(bg, Z, C.c_void_p(info.data_ptr()), info.stride(0), ncb,
                                              C.c_void_p(cw.data_ptr()), cw.stride(0), C.c_void_p(s.cuda_stream))
    L.check("cuphyHipLDPCEncodeCodeBlocks", st)
    return cw


def fill_tb(tb, gen, llr_dtype=torch.float16, sigma2=None):
    """Generate info bits, encode, BPSK/AWGN, LLRs for one transport block."""
    Z, bg = tb.Z, tb.bg
    info = torch.randint(0, 2, (tb.ncb, nsys(bg) * Z), generator=gen, device="cuda", dtype=torch.uint8)
    cw = encode(bg, Z, info)
    tx = cw[:, 2 * Z:2 * Z + tb.n_tx].to(torch.float32)
    if sigma2 is None:
        sigma2 = 10.0 ** (-tb.esn0_db / 10.0) / 2.0
    noise = torch.randn(tx.shape, generator=gen, device="cuda", dtype=torch.float32)
    y = (1.0 - 2.0 * tx) + noise * float(np.sqrt(sigma2))
    stride = (tb.num_llrs + 7) // 8 * 8
    llr = torch.zeros((tb.ncb, stride), dtype=torch.float32, device="cuda")
    llr[:, 2 * Z:2 * Z + tb.n_tx] = y * (2.0 / sigma2)
    tb.llr = llr.to(llr_dtype).contiguous()
    tb.info = info
    tb.out = torch.zeros((tb.ncb, tb.out_words), dtype=torch.int32, device="cuda")
    return tb


def heterogeneous_slot(seed=2, n_ue=16, llr_dtype=torch.float16, z_min=64):
    """16 UEs: BG in {1,2}, Zc uniform over lifting sizes in [z_min, 384],
    code rate U[0.2, 0.93], 1-24 CBs per TB, Es/N0 U[-1, 8] dB, BPSK-AWGN
    LLRs."""
    rng = np.random.RandomState(seed)
    zs = [z for z in LIFTING_SIZES if z_min <= z <= 384]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    tbs = []
    for _ in range(n_ue):
        bg = int(rng.randint(1, 3))
        Z = int(zs[rng.randint(len(zs))])
        R = float(rng.uniform(0.2, 0.93))
        ncb = int(rng.randint(1, 25))
        esn0 = float(rng.uniform(-1.0, 8.0))
        K = nsys(bg) * Z
        mb, N = mb_for_rate(bg, Z, K, R)
        n_tx = min(N - 2 * Z, (nsys(bg) + mb - 2) * Z)
        tb = TransportBlock(bg, Z, mb, ncb, n_tx, esn0, extra={"R": R})
        tbs.append(fill_tb(tb, gen, llr_dtype))
    return tbs


def config1_tb(seed=1, llr_dtype=torch.float32):
    """1 TB x 8 CBs, BG1 Z=384 mb=46, BPSK-AWGN Es/N0 = 1 dB (sigma^2 = 0.3972)."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    tb = TransportBlock(1, 384, 46, 8, 66 * 384, 1.0)
    return fill_tb(tb, gen, llr_dtype, sigma2=10.0 ** (-0.1) / 2.0)


def full_load_slot(seed=3, n_cells=20, cbs_per_cell=140, llr_dtype=torch.float16):
    """20 cells x ~140 CBs BG1 Z=384 at R = 0.93 (mb = 4), consistent-Gaussian
    LLRs N(+-mu, 2 mu) with mu from per-cell SNR U[18, 24] dB."""
    rng = np.random.RandomState(seed)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    tbs = []
    Z, bg = 384, 1
    K = 22 * Z
    mb, N = mb_for_rate(bg, Z, K, 0.93)
    for _ in range(n_cells):
        snr_db = float(rng.uniform(18.0, 24.0))
        mu = 2.0 * 10.0 ** (snr_db / 10.0) / 10.0  # mean LLR magnitude for this SNR (scaled to the clamp range)
        tb = TransportBlock(bg, Z, mb, cbs_per_cell, min(N - 2 * Z, (22 + mb - 2) * Z), snr_db)
        info = torch.randint(0, 2, (tb.ncb, K), generator=gen, device="cuda", dtype=torch.uint8)
        cw = encode(bg, Z, info)
        x = 1.0 - 2.0 * cw[:, 2 * Z:2 * Z + tb.n_tx].to(torch.float32)
        noise = torch.randn(x.shape, generator=gen, device="cuda", dtype=torch.float32)
        stride = (tb.num_llrs + 7) // 8 * 8
        llr = torch.zeros((tb.ncb, stride), dtype=torch.float32, device="cuda")
        llr[:, 2 * Z:2 * Z + tb.n_tx] = mu * x + float(np.sqrt(2.0 * mu)) * noise
        tb.llr = llr.to(llr_dtype).contiguous()
        tb.info = info
        tb.out = torch.zeros((tb.ncb, tb.out_words), dtype=torch.int32, device="cuda")
        tbs.append(tb)
    return tbs
