// PUSCH de-rate-matching with HARQ soft combining (gfx950).
//
// Drop-in for the reference's puschRxRateMatch (cuPHY/src/cuphy/rate_matching/
// rate_matching.cu:239-1097): per code block r of each scheduled TB,
//   * E_r and the CB's first LLR from G, C, Nl, Qm (TS 38.212 5.4.2.1,
//     rate_matching.cu:325-357),
//   * input LLR of transmitted bit (symbol j, bit k) at
//     (start/Qm)*QAM_STRIDE + k + j*QAM_STRIDE, or spread over nBBULayers
//     through layer_map_array (compute_llr_index, rate_matching.cu:81-110),
//   * de-interleaving: in-sequence index kIdx*E/Qm + jIdx (:112-131),
//   * bit selection from k0 around the Ncb circular buffer, skipping the
//     filler bits [Kd, Kd+F) (TS 38.212 5.4.2.1; derate_match.m:152-170),
//   * descrambling (descramblingOn) is not implemented: creation with it
//     enabled returns CUPHY_STATUS_NOT_SUPPORTED;
//   * ndi = 1: write; ndi = 0: add to the HARQ buffer; clamp to +-10000;
//     untouched positions 0, filler +10000, the 2Zc punctured positions 0
//     (rate_matching.cu:622-683), HARQ layout [2Zc | Ncb] with stride
//     Ncb_padded per CB.
// Design differences (deterministic, atomics-free): one workgroup per CB
// walks the E selected bits in passes of L = Ncb - F (within a pass every
// output position is distinct), with a barrier between passes, so repeated
// positions are combined in transmission order and clamped after every
// addition exactly like derate_match.m:156-163.  The reference instead
// uses atomicAdd in the wrap-around window plus a separate clamp pass.
// Selected position i maps to the filler-free circle as c = (c0 + i) mod L
// and back as c < Kd ? c : c + F, with c0 the circle position of k0.
#include "nr_ldpc_common.hpp"
#include "derate_match.hpp"

namespace nrldpc {

constexpr float LLR_MAX_ABS = 10000.0f;

template <typename T> struct rm_num;
template <> struct rm_num<float> {
    static __device__ __forceinline__ float clamp(float v)
    {
        return __builtin_fmaxf(__builtin_fminf(v, LLR_MAX_ABS), -LLR_MAX_ABS);
    }
};
template <> struct rm_num<_Float16> {
    static __device__ __forceinline__ _Float16 clamp(_Float16 v)
    {
        return __builtin_fmaxf16(__builtin_fminf16(v, (_Float16)LLR_MAX_ABS), (_Float16)-LLR_MAX_ABS);
    }
};

template <typename TO, typename TI> __device__ __forceinline__ TO cvt_llr(TI v) { return (TO)v; }

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void derate_match_kernel(const rm_descriptor* __restrict__ desc)
{
    const int cb_in_tb = blockIdx.x;
    const int tb_slot = blockIdx.y;
    const uint16_t ue = desc->sch_user_idxs[tb_slot];
    const PerTbParams& p = desc->tb_prms[ue];
    if (cb_in_tb >= (int)p.num_CBs) return;
    const int C = (int)p.num_CBs;
    const int Qm = (int)p.Qm, Nl = (int)p.Nl;
    const int r = cb_in_tb + (int)p.firstCodeBlockIndex;
    TO* out = static_cast<TO*>(desc->out[ue]) + (size_t)r * p.Ncb_padded;
    const TI* in = static_cast<const TI*>(desc->llr_in[tb_slot]);

    // ---- E and first LLR of this CB (rate_matching.cu:325-357) ----------
    const int nq = Nl * Qm;
    const int total = p.uciOnPuschFlag ? (int)p.G : (int)p.encodedSize;
    const int q1 = total / nq;
    const int q = q1 / C;
    const int rr = C - (q1 - q * C) - 1;
    const int El = nq * q;
    const int Eh = El + nq * ((long long)q * nq * C < (long long)total ? 1 : 0);
    int E, start;
    if (cb_in_tb <= rr) {
        E = El;
        start = cb_in_tb * El;
    } else {
        E = Eh;
        start = (rr + 1) * El + (cb_in_tb - rr - 1) * Eh;
    }
    const int EQ = E / Qm;
    const int Zc = (int)p.Zc, Ncb = (int)p.Ncb, F = (int)p.F, K = (int)p.K;
    const int Kd = K - F - 2 * Zc;
    const int L = Ncb - F;
    const int k0 = (int)p.k0;
    const int c0 = (k0 < Kd) ? k0 : ((k0 <= Kd + F) ? Kd : k0 - F);
    const bool ndi = p.ndi != 0;
    const int adj = (start / Qm) * QAM_STRIDE;
    const int nbbu = (int)p.nBBULayers;
    const bool uci = p.uciOnPuschFlag != 0;
    TO* cbuf = out + 2 * Zc;

    // ---- passes over the selected bits -----------------------------------
    const int npass = (ndi ? ((E > L) ? (E + L - 1) / L : 1) : (E + L - 1) / L);
    for (int pass = 0; pass < npass; ++pass) {
        const int i0 = pass * L;
        const int span = ndi && pass == 0 ? L : min(L, E - i0);
        for (int d = threadIdx.x; d < span; d += blockDim.x) {
            const int i = i0 + d;  // index in the de-interleaved sequence
            int c = c0 + d;
            if (c >= L) c -= L;
            const int o = (c < Kd) ? c : c + F;
            if (i >= E) {  // ndi = 1, first pass: selected-bit positions not transmitted
                cbuf[o] = (TO)0.0f;
                continue;
            }
            const int kIdx = i / EQ, jIdx = i - kIdx * EQ;
            int idx;
            if (uci) {
                idx = start + jIdx * Qm + kIdx;
            } else if (Nl < nbbu) {
                const int jl = jIdx / Nl;
                idx = (adj * nbbu / Nl) + kIdx + (jl * nbbu + (int)p.layer_map_array[jIdx - jl * Nl]) * QAM_STRIDE;
            } else {
                idx = adj + kIdx + jIdx * QAM_STRIDE;
            }
            TO v = rm_num<TO>::clamp(cvt_llr<TO, TI>(in[idx]));
            const TO prev = (ndi && pass == 0) ? (TO)0.0f : cbuf[o];
            cbuf[o] = rm_num<TO>::clamp(prev + v);
        }
        __syncthreads();
    }
    if (ndi) {
        for (int n = threadIdx.x; n < F; n += blockDim.x) cbuf[Kd + n] = (TO)LLR_MAX_ABS;
        for (int n = threadIdx.x; n < 2 * Zc; n += blockDim.x) out[n] = (TO)0.0f;
    }
}

// The reference launches separate reset and clamp kernels around the main
// one; this design needs neither, so their launch-configuration slots point
// at this empty kernel (grid 1) for callers that launch all three nodes.
__global__ void derate_match_noop_kernel(const rm_descriptor* desc) { (void)desc; }

void* rm_kernel(int fpconfig)
{
    switch (fpconfig) {
    case 0: return (void*)&derate_match_kernel<float, float>;
    case 1: return (void*)&derate_match_kernel<_Float16, float>;
    case 2: return (void*)&derate_match_kernel<float, _Float16>;
    case 3: return (void*)&derate_match_kernel<_Float16, _Float16>;
    default: return nullptr;
    }
}

void* rm_noop_kernel() { return (void*)&derate_match_noop_kernel; }

}  // namespace nrldpc
