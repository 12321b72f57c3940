// Device core of the CRC checks (TS 38.212 5.1, zero initial value), shared
// by the PUSCH CRC-decode kernels and the fused slot kernel.  A thread
// computes the remainder of its contiguous chunk of words byte-wise with a
// 256-entry table, then multiplies it by x^(32 * words after the chunk)
// mod g with a table of x^(32*2^j) mod g; chunk results are XOR-combined
// (CRC linearity).  All tables are built at compile time.
#pragma once

#include "nr_ldpc_common.hpp"

namespace nrldpc {

struct crc_poly {
    uint32_t g;     // polynomial without the leading term
    int L;          // degree
};

constexpr crc_poly CRC24A_P{0x864CFBu, 24};
constexpr crc_poly CRC24B_P{0x800063u, 24};
constexpr crc_poly CRC16_P{0x1021u, 16};

// a*b mod g (both < 2^L), carry-less
constexpr uint32_t gf2_mulmod(uint32_t a, uint32_t b, crc_poly p)
{
    const uint32_t mask = (p.L == 32) ? 0xFFFFFFFFu : ((1u << p.L) - 1u);
    uint32_t r = 0;
    for (int i = p.L - 1; i >= 0; --i) {
        const bool top = (r >> (p.L - 1)) & 1u;
        r = (r << 1) & mask;
        if (top) r ^= p.g;
        if ((b >> i) & 1u) r ^= a;
    }
    return r;
}

struct crc_tables {
    uint32_t byte[3][256];   // byte-wise step: reg' = (reg << 8) ^ T[(reg >> (L-8)) ^ b]
    uint32_t pow32[3][24];   // x^(32 * 2^j) mod g
};

constexpr uint32_t byte_entry(uint32_t v, crc_poly p)
{
    const uint32_t mask = (1u << p.L) - 1u;
    uint32_t r = v << (p.L - 8);
    for (int i = 0; i < 8; ++i) r = ((r << 1) & mask) ^ (((r >> (p.L - 1)) & 1u) ? p.g : 0u);
    return r;
}

constexpr crc_tables make_crc_tables()
{
    crc_tables t{};
    const crc_poly ps[3] = {CRC24A_P, CRC24B_P, CRC16_P};
    for (int k = 0; k < 3; ++k) {
        for (uint32_t v = 0; v < 256; ++v) t.byte[k][v] = byte_entry(v, ps[k]);
        // x^32 mod g, then repeated squaring
        uint32_t x32 = 1;
        for (int i = 0; i < 32; ++i) x32 = gf2_mulmod(x32, 2u, ps[k]);
        t.pow32[k][0] = x32;
        for (int j = 1; j < 24; ++j) t.pow32[k][j] = gf2_mulmod(t.pow32[k][j - 1], t.pow32[k][j - 1], ps[k]);
    }
    return t;
}

static __device__ const crc_tables g_crc_tables = make_crc_tables();

enum { K_CRC24A = 0, K_CRC24B = 1, K_CRC16 = 2 };

__device__ __forceinline__ uint32_t dmulmod(uint32_t a, uint32_t b, int L, uint32_t g)
{
    const uint32_t mask = (1u << L) - 1u;
    uint32_t r = 0;
    for (int i = L - 1; i >= 0; --i) {
        const uint32_t top = (r >> (L - 1)) & 1u;
        r = ((r << 1) & mask) ^ (top ? g : 0u);
        if ((b >> i) & 1u) r ^= a;
    }
    return r;
}

// a*b mod g for a compile-time polynomial (fully unrolled).
template <int L, uint32_t G>
__device__ __forceinline__ uint32_t cmulmod(uint32_t a, uint32_t b)
{
    constexpr uint32_t mask = (1u << L) - 1u;
    // bit i of b and bit L-1 of r as all-ones / all-zero masks (sign-extending
    // field extracts); bits of r above L-1 are left to fall out of the mask
    // at the end, so that each step is a shift and two masked XORs
    uint32_t r = 0;
#pragma unroll
    for (int i = L - 1; i >= 0; --i) {
        const uint32_t top = (uint32_t)((int32_t)(r << (31 - (L - 1))) >> 31);
        const uint32_t bit = (uint32_t)((int32_t)(b << (31 - i)) >> 31);
        r = (r << 1) ^ (top & G) ^ (bit & a);
    }
    return r & mask;
}

__device__ __forceinline__ uint32_t rbit8x4(uint32_t w)
{
    // reverse the bits inside each byte (byte order kept)
    w = ((w >> 1) & 0x55555555u) | ((w & 0x55555555u) << 1);
    w = ((w >> 2) & 0x33333333u) | ((w & 0x33333333u) << 2);
    w = ((w >> 4) & 0x0F0F0F0Fu) | ((w & 0x0F0F0F0Fu) << 4);
    return w;
}

// Remainder of `nw` message words (bytes in memory order, each byte MSB-first
// in the message) held by the workgroup; `word(i)` yields message word i.
template <typename WordFn>
__device__ uint32_t block_crc(int nw, int kind, WordFn word, uint32_t* lds_red)
{
    const crc_poly p = kind == K_CRC24A ? CRC24A_P : (kind == K_CRC24B ? CRC24B_P : CRC16_P);
    const uint32_t* T = g_crc_tables.byte[kind];
    const int nt = blockDim.x, tid = threadIdx.x;
    const int per = (nw + nt - 1) / nt;
    const int w0 = min(nw, tid * per), w1 = min(nw, w0 + per);
    uint32_t reg = 0;
    for (int w = w0; w < w1; ++w) {
        const uint32_t v = word(w);
        for (int b = 0; b < 4; ++b) {
            const uint32_t byte = (v >> (8 * b)) & 0xFFu;
            reg = ((reg << 8) & ((1u << p.L) - 1u)) ^ T[((reg >> (p.L - 8)) ^ byte) & 0xFFu];
        }
    }
    // shift by the words that follow this chunk: x^(32 * (nw - w1))
    uint32_t m = (uint32_t)(nw - w1);
    for (int j = 0; m && reg; ++j, m >>= 1)
        if (m & 1u) reg = dmulmod(reg, g_crc_tables.pow32[kind][j], p.L, p.g);
    // XOR-reduce over the workgroup
    for (int off = 32; off > 0; off >>= 1) reg ^= __shfl_xor(reg, off, 64);
    if ((tid & 63) == 0) lds_red[tid >> 6] = reg;
    __syncthreads();
    uint32_t r = 0;
    for (int i = 0; i < (nt + 63) / 64; ++i) r ^= lds_red[i];
    __syncthreads();
    return r;
}

// Byte-granular shift factors: x^(8m) mod g for m < CRC_SHIFT_LO (shifts
// inside one code block, up to 1056 bytes), and for CRC24A x^(8 * 1024 * m)
// mod g for m < 256 (shifts inside a TB of up to 256 KiB: x^(8n) =
// x^(8 (n mod 1024)) * x^(8 * 1024 * (n / 1024))).  Built at compile time:
// each entry is the previous one advanced by one byte-wise CRC step.
constexpr int CRC_SHIFT_LO = 1088;
struct crc_shift_tables {
    uint32_t lo[3][CRC_SHIFT_LO];
    uint32_t tb_hi[256];
};

constexpr crc_shift_tables make_crc_shift_tables()
{
    crc_shift_tables t{};
    const crc_poly ps[3] = {CRC24A_P, CRC24B_P, CRC16_P};
    for (int k = 0; k < 3; ++k) {
        const uint32_t mask = (1u << ps[k].L) - 1u;
        uint32_t v = 1;  // x^0
        for (int m = 0; m < CRC_SHIFT_LO; ++m) {
            t.lo[k][m] = v;
            // v * x^8 mod g: one byte-wise step with a zero input byte
            v = ((v << 8) & mask) ^ byte_entry((v >> (ps[k].L - 8)) & 0xFFu, ps[k]);
        }
    }
    const uint32_t step = gf2_mulmod(t.lo[0][1023], t.lo[0][1], CRC24A_P);  // x^(8*1024)
    uint32_t v = 1;
    for (int m = 0; m < 256; ++m) {
        t.tb_hi[m] = v;
        v = gf2_mulmod(v, step, CRC24A_P);
    }
    return t;
}

static __device__ const crc_shift_tables g_crc_shift = make_crc_shift_tables();

__device__ __forceinline__ uint32_t crc_poly_of(int kind) { return kind == K_CRC24A ? CRC24A_P.g : (kind == K_CRC24B ? CRC24B_P.g : CRC16_P.g); }
__device__ __forceinline__ int crc_len_of(int kind) { return kind == K_CRC16 ? 16 : 24; }

// Remainder of the nbytes message bytes byte_at(0 .. nbytes-1) (MSB-first
// bytes), nbytes <= CRC_SHIFT_LO, computed by the workgroup with the byte
// table T of the kind (in LDS): every thread a contiguous chunk, shifted by
// x^(8 * bytes after it) with one multiplication, XOR-reduced.
template <typename ByteFn>
__device__ uint32_t block_crc_bytes(int nbytes, int kind, ByteFn byte_at, const uint32_t* T, uint32_t* lds_red)
{
    const int L = crc_len_of(kind);
    const uint32_t g = crc_poly_of(kind);
    const uint32_t mask = (1u << L) - 1u;
    const int nt = blockDim.x, tid = threadIdx.x;
    const int per = (nbytes + nt - 1) / nt;
    const int b0 = min(nbytes, tid * per), b1 = min(nbytes, b0 + per);
    uint32_t reg = 0;
    for (int b = b0; b < b1; ++b) reg = ((reg << 8) & mask) ^ T[((reg >> (L - 8)) ^ byte_at(b)) & 0xFFu];
    if (reg && nbytes > b1) reg = dmulmod(reg, g_crc_shift.lo[kind][nbytes - b1], L, g);
    for (int off = 32; off > 0; off >>= 1) reg ^= __shfl_xor(reg, off, 64);
    if ((tid & 63) == 0) lds_red[tid >> 6] = reg;
    __syncthreads();
    uint32_t r = 0;
    for (int i = 0; i < (nt + 63) / 64; ++i) r ^= lds_red[i];
    __syncthreads();
    return r;
}

// x^(8n) mod g24A for n < 256 KiB.
__device__ __forceinline__ uint32_t crc24a_shift_bytes(uint32_t n)
{
    return dmulmod(g_crc_shift.lo[K_CRC24A][n & 1023u], g_crc_shift.tb_hi[(n >> 10) & 255u], 24, CRC24A_P.g);
}

// CRC bits attached to each CB: 24 (CRC24B) when the TB is segmented.
__device__ __forceinline__ int cb_crc_bits(const PerTbParams& p)
{
    return p.num_CBs > 1 ? 24 : 0;
}

// Payload bytes one CB contributes to the TB byte buffer.
__device__ __forceinline__ int cb_data_bytes(const PerTbParams& p)
{
    return ((int)(p.K - p.F) - cb_crc_bits(p) + 7) / 8;
}

// CRC checked over a CB's K' bits: CRC24B when the TB is segmented, else
// the TB CRC itself (CRC16 up to 480 bytes incl. the CRC, i.e. A <= 3824).
__device__ __forceinline__ int cb_crc_kind(const PerTbParams& p)
{
    return p.num_CBs > 1 ? K_CRC24B : (cb_data_bytes(p) <= 480 ? K_CRC16 : K_CRC24A);
}

}  // namespace nrldpc
