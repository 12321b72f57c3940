// Device core of PUSCH de-rate-matching with HARQ combining (gfx950), shared
// by the standalone de-rate-match kernel and the fused slot kernel.
//
// Per code block r of a TB (reference: rate_matching.cu:239-683):
//   * E_r and the CB's first LLR from G, C, Nl, Qm (TS 38.212 5.4.2.1,
//     rate_matching.cu:325-357),
//   * input LLR of transmitted bit (symbol j, bit k) at
//     (start/Qm)*QAM_STRIDE + k + j*QAM_STRIDE, or spread over nBBULayers
//     through layer_map_array (compute_llr_index, rate_matching.cu:81-110),
//   * de-interleaving: in-sequence index kIdx*E/Qm + jIdx (:112-131),
//   * bit selection from k0 around the Ncb circular buffer, skipping the
//     filler bits [Kd, Kd+F) (TS 38.212 5.4.2.1; derate_match.m:152-170).
// Selected position i maps to the filler-free circle as c = (c0 + i) mod L,
// L = Ncb - F, and back as c < Kd ? c : c + F, with c0 the circle position
// of k0.  Within a pass of L consecutive selected bits every position is
// distinct, so a workgroup combines a pass without atomics; passes are
// separated by a barrier, which combines repeated positions in transmission
// order with a clamp after every addition (derate_match.m:156-163).
#pragma once

#include "nr_ldpc_common.hpp"

namespace nrldpc {

constexpr float LLR_MAX_ABS = 10000.0f;

struct rm_geometry {
    int E;       // rate-matched bits of this CB
    int start;   // index of its first bit in the TB's bit sequence
    int EQ;      // E / Qm
    int Zc, Ncb, F, Kd, L, c0;
    int adj;     // QAM_STRIDE-layout offset of the CB's first symbol
    int Qm, Nl, nbbu;
    bool ndi, uci;
};

__device__ __forceinline__ rm_geometry rm_geometry_of(const PerTbParams& p, int cb_in_tb)
{
    rm_geometry g;
    const int C = (int)p.num_CBs;
    g.Qm = (int)p.Qm;
    g.Nl = (int)p.Nl;
    const int nq = g.Nl * g.Qm;
    const int total = p.uciOnPuschFlag ? (int)p.G : (int)p.encodedSize;
    const int q1 = total / nq;
    const int q = q1 / C;
    const int rr = C - (q1 - q * C) - 1;
    const int El = nq * q;
    // the larger size compares with encodedSize as rate_matching.cu:345; it is used only for r > rr,
    // i.e. when q1 is not a multiple of C, where q*nq*C < total <= encodedSize holds either way
    const int Eh = El + nq * ((long long)q * nq * C < (long long)p.encodedSize ? 1 : 0);
    if (cb_in_tb <= rr) {
        g.E = El;
        g.start = cb_in_tb * El;
    } else {
        g.E = Eh;
        g.start = (rr + 1) * El + (cb_in_tb - rr - 1) * Eh;
    }
    g.EQ = g.E / g.Qm;
    g.Zc = (int)p.Zc;
    g.Ncb = (int)p.Ncb;
    g.F = (int)p.F;
    g.Kd = (int)p.K - g.F - 2 * g.Zc;
    g.L = g.Ncb - g.F;
    const int k0 = (int)p.k0;
    g.c0 = (k0 < g.Kd) ? k0 : ((k0 <= g.Kd + g.F) ? g.Kd : k0 - g.F);
    g.ndi = p.ndi != 0;
    g.adj = (g.start / g.Qm) * QAM_STRIDE;
    g.nbbu = (int)p.nBBULayers;
    g.uci = p.uciOnPuschFlag != 0;
    return g;
}

// The geometry as wave-uniform values (scalar registers), for callers that
// read PerTbParams through a generic pointer, whose loads count as divergent.
__device__ __forceinline__ void rm_make_uniform(rm_geometry& g)
{
    g.E = __builtin_amdgcn_readfirstlane(g.E);
    g.start = __builtin_amdgcn_readfirstlane(g.start);
    g.EQ = __builtin_amdgcn_readfirstlane(g.EQ);
    g.Zc = __builtin_amdgcn_readfirstlane(g.Zc);
    g.Ncb = __builtin_amdgcn_readfirstlane(g.Ncb);
    g.F = __builtin_amdgcn_readfirstlane(g.F);
    g.Kd = __builtin_amdgcn_readfirstlane(g.Kd);
    g.L = __builtin_amdgcn_readfirstlane(g.L);
    g.c0 = __builtin_amdgcn_readfirstlane(g.c0);
    g.adj = __builtin_amdgcn_readfirstlane(g.adj);
    g.Qm = __builtin_amdgcn_readfirstlane(g.Qm);
    g.Nl = __builtin_amdgcn_readfirstlane(g.Nl);
    g.nbbu = __builtin_amdgcn_readfirstlane(g.nbbu);
    g.ndi = __builtin_amdgcn_readfirstlane((int)g.ndi) != 0;
    g.uci = __builtin_amdgcn_readfirstlane((int)g.uci) != 0;
}

// Circular-buffer position (relative to the end of the 2Zc punctured
// values) of the d-th selected bit of a pass.
__device__ __forceinline__ int rm_position(const rm_geometry& g, int d)
{
    int c = g.c0 + d;
    if (c >= g.L) c -= g.L;
    return (c < g.Kd) ? c : c + g.F;
}

// Index in the input LLR buffer of de-interleaved bit i of the CB.
__device__ __forceinline__ int rm_input_index(const rm_geometry& g, const PerTbParams& p, int i)
{
    const int kIdx = i / g.EQ, jIdx = i - kIdx * g.EQ;
    if (g.uci) return g.start + jIdx * g.Qm + kIdx;
    if (g.Nl < g.nbbu) {
        const int jl = jIdx / g.Nl;
        return (g.adj * g.nbbu / g.Nl) + kIdx + (jl * g.nbbu + (int)p.layer_map_array[jIdx - jl * g.Nl]) * QAM_STRIDE;
    }
    return g.adj + kIdx + jIdx * QAM_STRIDE;
}

// Position in the TB's rate-matched (transmission-order) bit sequence of
// de-interleaved bit i of the CB: symbol jIdx, bit kIdx -> start + jIdx*Qm
// + kIdx (the descrambling index, rate_matching.cu:366-386).
__device__ __forceinline__ int rm_tx_position(const rm_geometry& g, int i)
{
    const int kIdx = i / g.EQ, jIdx = i - kIdx * g.EQ;
    return g.start + jIdx * g.Qm + kIdx;
}

// Number of passes over the selected bits; with ndi the first pass covers
// all L positions (untransmitted ones are written 0).
__device__ __forceinline__ int rm_num_passes(const rm_geometry& g)
{
    if (g.ndi) return (g.E > g.L) ? (g.E + g.L - 1) / g.L : 1;
    return (g.E + g.L - 1) / g.L;
}

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

}  // namespace nrldpc
