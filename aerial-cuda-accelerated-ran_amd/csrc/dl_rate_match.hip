// DL rate matching of LDPC-encoded code blocks (TS 38.212 5.4.2) for gfx950:
// bit selection from k0 around the circular buffer of Ncb bits, skipping
// the filler bits, then bit interleaving, then (optionally) scrambling with
// the TS 38.211 7.3.1.1 Gold sequence of the TB's c_init.  The reference's
// PDSCH kernel (cuPHY/src/cuphy/dl_rate_matching/dl_rate_matching.cu:1951-2260)
// does the same with enable_layer_mapping = 0 and enable_modulation = 0 and
// writes CB r of TB t at bit (t * cmax + r) * emax of the output; so does
// this kernel.  Here every thread builds whole output words (32 output bits,
// each gathered from the CB's encoded bits staged in LDS) instead of
// OR-ing single bits into LDS words with atomics.
//
// Output bit i of CB r (i < E_r) is bit (i mod Qm) of symbol i / Qm, i.e.
// selected bit m = (i mod Qm) * E_r / Qm + i / Qm; selected bit m sits at
// circle position c = (c0 + m) mod (Ncb - F) of the filler-free circular
// buffer, buffer position c < Kd ? c : c + F (the same closed form as the
// uplink de-rate-matcher).  Scrambling XORs c(start_r + i), start_r the
// CB's first bit in the TB.
#include <hip/hip_runtime.h>

#include "derate_core.hpp"
#include "dl_rate_match.hpp"
#include "gold_core.hpp"

namespace nrldpc {

constexpr int DL_RM_MAX_N = 66 * 384;  // encoded bits of a BG1 Zc = 384 code block

__global__ __launch_bounds__(256) void dl_rate_match_kernel(const PerTbParams* __restrict__ prms,
                                                            const uint32_t* const* __restrict__ cw, int max_cbs,
                                                            int emax_words, int scrambling,
                                                            uint32_t* __restrict__ out)
{
    extern __shared__ uint32_t dl_lds[];
    const int t = blockIdx.y, r = blockIdx.x;
    const PerTbParams& p = prms[t];
    const int C = __builtin_amdgcn_readfirstlane((int)p.num_CBs);
    if (r >= C || r >= max_cbs) return;
    const int Qm = __builtin_amdgcn_readfirstlane((int)p.Qm), Nl = __builtin_amdgcn_readfirstlane((int)p.Nl);
    const bool prm_ok = (Qm == 1 || Qm == 2 || Qm == 4 || Qm == 6 || Qm == 8) && Nl >= 1 && Nl <= 8;
    rm_geometry g{};
    if (prm_ok) {
        g = rm_geometry_of(p, r);
        rm_make_uniform(g);
    }
    const int N = __builtin_amdgcn_readfirstlane((int)p.N);
    const int tid = threadIdx.x, nthr = blockDim.x;
    uint32_t* dst = out + ((size_t)t * max_cbs + r) * emax_words;
    // a CB that does not fit its slot or the kernel's buffers stays zero
    const bool ok = prm_ok && g.E > 0 && g.E <= 32 * emax_words && g.Ncb <= N && N <= DL_RM_MAX_N && g.L > 0 &&
                    g.Kd >= 0 && g.c0 >= 0 && g.c0 < g.L;
    const int words_in = (N + 31) >> 5;
    uint32_t* cb_in = dl_lds;
    uint32_t* seq = dl_lds + words_in;
    if (ok) {
        const uint32_t* src = cw[t] + (size_t)r * words_in;
        for (int w = tid; w < words_in; w += nthr) cb_in[w] = src[w];
        if (scrambling) {  // the CB's words of c, each wave generating a share
            const uint32_t w_base = (uint32_t)g.start >> 5;
            const int nw = (int)(((uint32_t)(g.start + g.E - 1) >> 5) - w_base + 1u);
            const int nwave = nthr >> 6, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
            const int per = (nw + nwave - 1) / nwave;
            const int a = min(nw, wave * per), b = min(nw, a + per);
            gold_words_wave(p.cinit, w_base + (uint32_t)a, w_base + (uint32_t)b, seq + a);
            if (tid == 0) seq[nw] = 0u;
        }
    }
    __syncthreads();
    const int e_words = ok ? (g.E + 31) >> 5 : 0;
    const int soff = g.start & 31;
    for (int w = tid; w < emax_words; w += nthr) {
        uint32_t v = 0;
        if (w < e_words) {
            const int nb = min(32, g.E - 32 * w);
            for (int b = 0; b < nb; ++b) {
                const int i = 32 * w + b;
                const int j = i / g.Qm, k = i - j * g.Qm;
                const int m = k * g.EQ + j;
                const int c = (g.c0 + m) % g.L;
                const int pos = (c < g.Kd) ? c : c + g.F;
                uint32_t bit = (cb_in[pos >> 5] >> (pos & 31)) & 1u;
                if (scrambling) {
                    const int n = soff + i;
                    bit ^= (seq[n >> 5] >> (n & 31)) & 1u;
                }
                v |= bit << b;
            }
        }
        dst[w] = v;
    }
}

cuphyStatus_t dl_rate_match(const PerTbParams* d_prms, int num_tbs, const uint32_t* const* d_cw, int max_cbs,
                            int emax, bool scrambling, uint32_t* d_out, hipStream_t strm)
{
    if (num_tbs < 0 || num_tbs > 65535 || max_cbs < 1 || max_cbs > 65535 || emax <= 0 || emax % 32 ||
        emax > DL_RM_MAX_EMAX)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    if (num_tbs == 0) return CUPHY_STATUS_SUCCESS;
    if (!d_prms || !d_cw || !d_out) return CUPHY_STATUS_INVALID_ARGUMENT;
    const int emax_words = emax / 32;
    const size_t lds = ((size_t)(DL_RM_MAX_N + 31) / 32 + (scrambling ? (size_t)emax_words + 2 : 0)) * sizeof(uint32_t);
    if (lds > 64u * 1024u && hipFuncSetAttribute((const void*)&dl_rate_match_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return CUPHY_STATUS_INTERNAL_ERROR;
    hipLaunchKernelGGL(dl_rate_match_kernel, dim3((unsigned)max_cbs, (unsigned)num_tbs), dim3(256), lds, strm, d_prms,
                       d_cw, max_cbs, emax_words, scrambling ? 1 : 0, d_out);
    return hipGetLastError() == hipSuccess ? CUPHY_STATUS_SUCCESS : CUPHY_STATUS_INTERNAL_ERROR;
}

}  // namespace nrldpc
