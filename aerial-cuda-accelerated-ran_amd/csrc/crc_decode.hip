// PUSCH code-block and transport-block CRC check with TB assembly (gfx950).
//
// Drop-in for the reference's puschRxCrcDecode (cuPHY/src/cuphy/crc/crc.cu:
// 55-237, setup :1174-1254):
//   CB kernel -- per code block: CRC over the ceil(K'/32) decoded words
//     (K' = K - F bits incl. the CB CRC; with C > 1 CRC24B, with C = 1 the
//     TB CRC itself: CRC16 when ceil(K'/8) <= 480 bytes else CRC24A);
//     remainder 0 = pass, stored at [sum of previous TBs' CBs + CB index];
//     the CB's data bytes (K' - 24 bits when C > 1, K' bits when C = 1) are
//     copied into the TB byte buffer, MSB-first per byte when reverseBytes
//     (bit 32w + i of the decoder output = bit i of word w);
//   TB kernel -- CRC24A over the assembled TB bytes (zero-padded to whole
//     words), one workgroup per TB, remainder 0 = pass (C = 1: the CB
//     kernel's remainder).
// Layouts as in the reference: CBs of TB y start at word sum_{i<y} C_i*264;
// TB y's bytes start at sum_{i<y} roundup4(C_i * data_bytes_i).
// CRC arithmetic (TS 38.212 5.1, zero initial value): a thread computes the
// remainder of its contiguous chunk byte-wise with a 256-entry table, then
// multiplies it by x^(32 * words after the chunk) mod g with a table of
// x^(32*2^j) mod g; the chunk results are XOR-combined (CRC linearity).
// Pass/fail equals the reference's; a failing remainder's value is this
// definition's (the reference's depends on its LUT frame size).
#include "nr_ldpc_common.hpp"
#include "crc_decode.hpp"

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

__device__ const crc_tables g_crc_tables = make_crc_tables();

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

struct tb_layout {
    int cb_word_base, cb_count_base, tb_byte_base;
};

__device__ __forceinline__ int cb_crc_bits(const PerTbParams& p)
{
    return p.num_CBs > 1 ? 24 : 0;
}

__device__ tb_layout layout_of(const crc_descriptor& d, int slot)
{
    tb_layout l{0, 0, 0};
    for (int i = 0; i < slot; ++i) {
        const PerTbParams& q = d.tb_prms[d.sch_user_idxs[i]];
        const int Kp = (int)(q.K - q.F);
        const int data_bytes = (Kp - cb_crc_bits(q) + 7) / 8;
        l.cb_word_base += (int)q.num_CBs * MAX_WORDS_PER_CB;
        l.cb_count_base += (int)q.num_CBs;
        int bytes = (int)q.num_CBs * data_bytes;
        l.tb_byte_base += (bytes + 3) / 4 * 4;
    }
    return l;
}

__global__ __launch_bounds__(64) void crc_cb_kernel(const crc_descriptor* __restrict__ desc)
{
    __shared__ uint32_t red[1];
    const crc_descriptor& d = *desc;
    const int slot = blockIdx.y;
    const PerTbParams& p = d.tb_prms[d.sch_user_idxs[slot]];
    if ((int)blockIdx.x >= (int)p.num_CBs) return;
    const tb_layout l = layout_of(d, slot);
    const int cb = (int)blockIdx.x + (int)p.firstCodeBlockIndex;
    const int C = (int)p.num_CBs;
    const int Kp = (int)(p.K - p.F);
    const int nw = (Kp + 31) / 32;
    const uint32_t* words = d.cb_in + l.cb_word_base + (size_t)cb * MAX_WORDS_PER_CB;
    const bool rev = d.reverse_bytes;
    const int data_bytes = (Kp - cb_crc_bits(p) + 7) / 8;
    const int kind = (C > 1) ? K_CRC24B : (data_bytes <= 480 ? K_CRC16 : K_CRC24A);
    const uint32_t crc = block_crc(nw, kind, [&](int w) { return rev ? rbit8x4(words[w]) : words[w]; }, red);
    uint8_t* tb = d.tb_out + l.tb_byte_base;
    for (int b = threadIdx.x; b < data_bytes; b += blockDim.x) {
        const uint32_t v = rev ? rbit8x4(words[b >> 2]) : words[b >> 2];
        tb[(size_t)cb * data_bytes + b] = (uint8_t)(v >> (8 * (b & 3)));
    }
    if (threadIdx.x == 0) {
        d.cb_crc_out[l.cb_count_base + cb] = crc;
        if (blockIdx.x == 0) {
            if (C == 1) d.tb_crc_out[slot] = crc;
            const int bytes = C * data_bytes;
            for (int b = bytes; b < (bytes + 3) / 4 * 4; ++b) tb[b] = 0;
        }
    }
}

__global__ __launch_bounds__(256) void crc_tb_kernel(const crc_descriptor* __restrict__ desc)
{
    __shared__ uint32_t red[4];
    const crc_descriptor& d = *desc;
    const int slot = blockIdx.x;
    const PerTbParams& p = d.tb_prms[d.sch_user_idxs[slot]];
    if (p.num_CBs <= 1) return;
    const tb_layout l = layout_of(d, slot);
    const int Kp = (int)(p.K - p.F);
    const int data_bytes = (Kp - 24 + 7) / 8;
    const int nw = ((int)p.num_CBs * data_bytes + 3) / 4;
    const uint32_t* tbw = reinterpret_cast<const uint32_t*>(d.tb_out + l.tb_byte_base);
    const uint32_t crc = block_crc(nw, K_CRC24A, [&](int w) { return tbw[w]; }, red);
    if (threadIdx.x == 0) d.tb_crc_out[slot] = crc;
}

void* crc_cb_kernel_ptr() { return (void*)&crc_cb_kernel; }
void* crc_tb_kernel_ptr() { return (void*)&crc_tb_kernel; }

}  // namespace nrldpc
