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
#include "crc_core.hpp"
#include "crc_decode.hpp"

namespace nrldpc {

struct tb_layout {
    int cb_word_base, cb_count_base, tb_byte_base;
};

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
    const int data_bytes = cb_data_bytes(p);
    const int kind = cb_crc_kind(p);
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
