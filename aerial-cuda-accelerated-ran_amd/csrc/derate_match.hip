// PUSCH de-rate-matching with HARQ soft combining (gfx950).
//
// Drop-in for the reference's puschRxRateMatch (cuPHY/src/cuphy/rate_matching/
// rate_matching.cu:239-1097); the per-CB arithmetic is in derate_core.hpp.
//   * descrambling (descramblingOn, not for UCI on PUSCH, as
//     rate_matching.cu:276): LLR n of the TB's rate-matched sequence is
//     negated where the Gold sequence c(n) of the TB's c_init is 1
//     (rate_matching.cu:378-387, 530-535); the CB's words of c are generated
//     into LDS first (gold_core.hpp);
//   * ndi = 1: write; ndi = 0: add to the HARQ buffer; clamp to +-10000;
//     untouched positions 0, filler +10000, the 2Zc punctured positions 0
//     (rate_matching.cu:622-683), HARQ layout [2Zc | Ncb] with stride
//     Ncb_padded per CB.
// Design difference (deterministic, atomics-free): one workgroup per CB
// walks the selected bits in passes of L = Ncb - F with a barrier between
// passes; the reference uses atomicAdd in the wrap-around window plus a
// separate clamp pass.
#include "derate_core.hpp"
#include "derate_match.hpp"
#include "gold_core.hpp"

namespace nrldpc {

template <typename TO, typename TI> __device__ __forceinline__ TO cvt_llr(TI v) { return (TO)v; }

template <typename TI, typename TO, bool DESCR>
__global__ __launch_bounds__(256) void derate_match_kernel(const rm_descriptor* __restrict__ desc)
{
    const int cb_in_tb = blockIdx.x;
    const int tb_slot = blockIdx.y;
    const uint16_t ue = desc->sch_user_idxs[tb_slot];
    const PerTbParams& p = desc->tb_prms[ue];
    if (cb_in_tb >= (int)p.num_CBs) return;
    const int r = cb_in_tb + (int)p.firstCodeBlockIndex;
    TO* out = static_cast<TO*>(desc->out[ue]) + (size_t)r * p.Ncb_padded;
    const TI* in = static_cast<const TI*>(desc->llr_in[tb_slot]);
    const rm_geometry g = rm_geometry_of(p, cb_in_tb);
    TO* cbuf = out + 2 * g.Zc;

    // descrambling: words of c covering the CB's bits [start, start + E)
    __shared__ uint32_t seq[DESCR ? RM_SEQ_WORDS : 1];
    const bool descr = DESCR && !g.uci;
    const uint32_t w_base = (uint32_t)g.start >> 5;
    if (descr) {
        const uint32_t nw = min(((uint32_t)(g.start + g.E - 1) >> 5) - w_base + 1u, (uint32_t)RM_SEQ_WORDS);
        const uint32_t nwave = blockDim.x >> 6, wave = threadIdx.x >> 6;
        const uint32_t per = (nw + nwave - 1) / nwave;
        const uint32_t a = min(nw, wave * per), b = min(nw, a + per);
        gold_words_wave(p.cinit, w_base + a, w_base + b, seq + a);
        __syncthreads();
    }

    const int npass = rm_num_passes(g);
    for (int pass = 0; pass < npass; ++pass) {
        const int i0 = pass * g.L;
        const int span = g.ndi && pass == 0 ? g.L : min(g.L, g.E - i0);
        for (int d = threadIdx.x; d < span; d += blockDim.x) {
            const int i = i0 + d;  // index in the de-interleaved sequence
            const int o = rm_position(g, d);
            if (i >= g.E) {  // ndi = 1, first pass: selected-bit positions not transmitted
                cbuf[o] = (TO)0.0f;
                continue;
            }
            TO v = rm_num<TO>::clamp(cvt_llr<TO, TI>(in[rm_input_index(g, p, i)]));
            if (descr) {
                const uint32_t n = (uint32_t)rm_tx_position(g, i) - (w_base << 5);
                if ((seq[min(n >> 5, (uint32_t)RM_SEQ_WORDS - 1u)] >> (n & 31u)) & 1u) v = -v;
            }
            const TO prev = (g.ndi && pass == 0) ? (TO)0.0f : cbuf[o];
            cbuf[o] = rm_num<TO>::clamp(prev + v);
        }
        __syncthreads();
    }
    if (g.ndi) {
        for (int n = threadIdx.x; n < g.F; n += blockDim.x) cbuf[g.Kd + n] = (TO)LLR_MAX_ABS;
        for (int n = threadIdx.x; n < 2 * g.Zc; n += blockDim.x) out[n] = (TO)0.0f;
    }
}

// The reference launches separate reset and clamp kernels around the main
// one; this design needs neither, so their launch-configuration slots point
// at this empty kernel (grid 1) for callers that launch all three nodes.
__global__ void derate_match_noop_kernel(const rm_descriptor* desc) { (void)desc; }

template <bool DESCR>
void* rm_kernel_of(int fpconfig)
{
    switch (fpconfig) {
    case 0: return (void*)&derate_match_kernel<float, float, DESCR>;
    case 1: return (void*)&derate_match_kernel<_Float16, float, DESCR>;
    case 2: return (void*)&derate_match_kernel<float, _Float16, DESCR>;
    case 3: return (void*)&derate_match_kernel<_Float16, _Float16, DESCR>;
    default: return nullptr;
    }
}

void* rm_kernel(int fpconfig, bool descrambling)
{
    return descrambling ? rm_kernel_of<true>(fpconfig) : rm_kernel_of<false>(fpconfig);
}

void* rm_noop_kernel() { return (void*)&derate_match_noop_kernel; }

}  // namespace nrldpc
