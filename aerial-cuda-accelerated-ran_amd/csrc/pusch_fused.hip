// Fused PUSCH receive-side decode of a slot (gfx950): for every code block,
// in one persistent launch,
//   de-rate-matching + HARQ combining  (derate_core.hpp; the reference's
//       puschRxRateMatch, rate_matching.cu:239-683)
//   -> layered normalized min-sum decode (ldpc_core.hpp; the reference's
//       ldpc2 transport-block kernels)
//   -> CB CRC and TB byte assembly, and the TB CRC by the last CB of the TB
//       to finish (crc_core.hpp; the reference's puschRxCrcDecode,
//       crc.cu:55-237).
// The reference runs these as separate kernels with HBM round trips in
// between (pusch_rx.cpp:3007-3021 builds them as graph nodes).  Here a
// workgroup takes the next CB from a device queue, builds the CB's HARQ
// circular buffer in LDS (binary16, [2Zc punctured | Ncb]), writes it back
// to its HBM HARQ buffer (the persistent soft state across
// retransmissions), converts the first (NSYS + mb) * Zc values in place to
// the decoder's APP array, decodes, and checks the CRC from the hard bits
// it just produced.  Results are bit-identical to the three separate
// stages (de-rate-match with binary16 output, binary16 decode, CRC).
//
// Input LLRs (int8 / binary16 / binary32) are scaled by llr_scale (int8
// LLRs saturated at +-127 with 4 steps per unit LLR: scale 0.25, exact)
// and rounded to binary16 before combining.
//
// The TB CRC24A is assembled from per-CB shares (CRC linearity): each CB's
// workgroup computes the remainder of its payload bytes shifted to their
// place in the TB and XORs it into a per-TB accumulator with an atomic; the
// workgroup that completes the TB reads the result.  Counters and
// accumulators return to 0, and so does the queue (reset by the last
// workgroup to leave), so the launch is replayable from a graph.
#include "crc_core.hpp"
#include "derate_core.hpp"
#include "ldpc_core.hpp"
#include "pusch_fused.hpp"

#include <algorithm>

namespace nrldpc {

constexpr uint32_t FUSED_CB_ERROR = 0xFFFFFFFFu;  // CRC value of a CB/TB whose parameters were rejected
constexpr uint32_t TB_ACC_ERROR = 0x80000000u;    // accumulator flag: a CB of the TB was rejected
constexpr int SLOT_CB_WORDS = MAX_DECODED_CODE_BLOCK_BIT_SIZE / 32;  // 264: decoded-word stride per CB

template <typename TI> __device__ __forceinline__ float llr_as_float(TI v) { return (float)v; }

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

// LDS state of a workgroup besides the HARQ/APP array.
struct fused_lds {
    uint32_t words[SLOT_CB_WORDS];  // hard bits of the current CB
    uint32_t crc_t[3][256];         // byte tables of the three CRCs
    uint32_t red[8];
    int item, last;
};

// Combine this transmission's E LLRs of the CB into the LDS circular buffer
// `cbuf` (passes of distinct positions, barrier after each).  Indices are
// advanced incrementally: one division per thread and pass.
template <typename TI>
__device__ __forceinline__ void combine_tx(_Float16* __restrict__ cbuf, const TI* __restrict__ in,
                                           const rm_geometry& g, const PerTbParams& p, float scale)
{
    const int nthr = blockDim.x;
    const int npass = rm_num_passes(g);
    for (int pass = 0; pass < npass; ++pass) {
        const int i0 = pass * g.L;
        const int span = min(g.L, g.E - i0);
        int d = threadIdx.x;
        if (d < span) {
            int c = g.c0 + d;  // circle position, kept in [0, L)
            while (c >= g.L) c -= g.L;
            int k = (i0 + d) / g.EQ;  // de-interleaving: i = k * EQ + j
            int j = (i0 + d) - k * g.EQ;
            for (; d < span; d += nthr) {
                int idx;
                if (g.uci) {
                    idx = g.start + j * g.Qm + k;
                } else if (g.Nl < g.nbbu) {
                    const int jl = j / g.Nl;
                    idx = (g.adj * g.nbbu / g.Nl) + k + (jl * g.nbbu + (int)p.layer_map_array[j - jl * g.Nl]) * QAM_STRIDE;
                } else {
                    idx = g.adj + k + j * QAM_STRIDE;
                }
                const int o = (c < g.Kd) ? c : c + g.F;
                const float x = llr_as_float(in[idx]) * scale;
                const _Float16 v = rm_num<_Float16>::clamp((_Float16)x);
                cbuf[o] = rm_num<_Float16>::clamp(cbuf[o] + v);
                c += nthr;
                while (c >= g.L) c -= g.L;
                j += nthr;
                while (j >= g.EQ) {
                    j -= g.EQ;
                    ++k;
                }
            }
        }
        __syncthreads();
    }
}

// One CB.  Returns this CB's share of the TB CRC24A (the remainder of its
// payload bytes shifted to their place in the TB), or TB_ACC_ERROR when
// the CB's parameters do not fit the launch (then only the error marker is
// written, into its CRC and into the TB CRC of a one-CB TB).
template <typename TI, int MAXR>
__device__ __forceinline__ uint32_t fused_cb(_Float16* __restrict__ hb, fused_lds& sh,
                                             const cuphyHipPuschSlotTb_t& job, const PerTbParams& p, int cb_in_tb,
                                             int lds_values)
{
    rm_geometry g = rm_geometry_of(p, cb_in_tb);
    rm_make_uniform(g);
    const int Zc = g.Zc;
    const int bg = __builtin_amdgcn_readfirstlane((int)p.bg);
    const int nsys = bg == 1 ? 22 : 10;
    const int mb = job.num_parity_nodes;
    const int nh = 2 * Zc + g.Ncb;
    const int nval = (nsys + mb) * Zc;
    const int r = cb_in_tb + __builtin_amdgcn_readfirstlane((int)p.firstCodeBlockIndex);
    const int Kb = __builtin_amdgcn_readfirstlane((int)p.K) / Zc;
    const int C = __builtin_amdgcn_readfirstlane((int)p.num_CBs);
    const size_t ncb_padded = (size_t)__builtin_amdgcn_readfirstlane((int)p.Ncb_padded);
    const bool ok = (bg == 1 || bg == 2) && z_index(Zc) >= 0 && mb >= 4 &&
                    mb <= (bg == 1 ? 46 : 42) && mb <= MAXR && nh <= lds_values && nval <= lds_values &&
                    Kb * Zc <= 32 * (int)SLOT_CB_WORDS && g.L > 0 && g.E > 0;
    if (!ok) {
        if (threadIdx.x == 0) {
            job.cb_crc[r] = FUSED_CB_ERROR;
            if (C == 1) *job.tb_crc = FUSED_CB_ERROR;
        }
        return TB_ACC_ERROR;
    }
    _Float16* harq = static_cast<_Float16*>(job.harq) + (size_t)r * ncb_padded;
    const TI* in = static_cast<const TI*>(job.llr);
    _Float16* cbuf = hb + 2 * Zc;
    const int tid = threadIdx.x, nthr = blockDim.x;
    // 16-byte vector path when the HBM buffer is 16-byte aligned (LDS is)
    const bool vec = ((uintptr_t)harq & 15u) == 0;
    const int nv = vec ? nh / 8 : 0;

    // ---- 1. HARQ circular buffer -> LDS ---------------------------------
    if (g.ndi) {
        const half8_t z = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int n = tid; n < nv; n += nthr) reinterpret_cast<half8_t*>(hb)[n] = z;
        for (int n = nv * 8 + tid; n < nh; n += nthr) hb[n] = (_Float16)0.0f;
    } else {
        for (int n = tid; n < nv; n += nthr) reinterpret_cast<half8_t*>(hb)[n] = reinterpret_cast<const half8_t*>(harq)[n];
        for (int n = nv * 8 + tid; n < nh; n += nthr) hb[n] = harq[n];
    }
    __syncthreads();
    if (g.ndi)
        for (int n = tid; n < g.F; n += nthr) cbuf[g.Kd + n] = (_Float16)LLR_MAX_ABS;

    // ---- 2. combine this transmission -------------------------------------
    combine_tx<TI>(cbuf, in, g, p, job.llr_scale);

    // ---- 3. write back the HARQ buffer; APP = clamp(HARQ), +0 for -0 ----
    {
        const _Float16 cl = (_Float16)job.clamp_value;
        const bool do_clamp = job.clamp_value > 0.0f;
        for (int n = tid; n < nv; n += nthr) {
            half8_t v = reinterpret_cast<const half8_t*>(hb)[n];
            reinterpret_cast<half8_t*>(harq)[n] = v;
            if (n * 8 < nval) {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    _Float16 a = v[e];
                    if (do_clamp) a = __builtin_fmaxf16(__builtin_fminf16(a, cl), -cl);
                    v[e] = (a == (_Float16)0) ? (_Float16)0 : a;
                }
                reinterpret_cast<half8_t*>(hb)[n] = v;  // values past nval are not read by the decoder
            }
        }
        for (int n = nv * 8 + tid; n < nh; n += nthr) {
            const _Float16 v = hb[n];
            harq[n] = v;
            if (n < nval) {
                _Float16 a = v;
                if (do_clamp) a = __builtin_fmaxf16(__builtin_fminf16(a, cl), -cl);
                hb[n] = (a == (_Float16)0) ? (_Float16)0 : a;
            }
        }
        for (int n = nh + tid; n < nval; n += nthr) hb[n] = (_Float16)0.0f;
    }
    __syncthreads();

    // ---- 4. decode -------------------------------------------------------
    cb_job dj;
    dj.llr = nullptr;
    dj.out = job.cb_bits + (size_t)r * SLOT_CB_WORDS;
    dj.soft = nullptr;
    dj.Z = Zc;
    dj.mb = mb;
    dj.Kb = Kb;
    dj.max_iter = job.max_iterations;
    dj.early_term = (job.flags & CUPHY_LDPC_DECODE_EARLY_TERM) != 0;
    dj.clamp = job.clamp_value;
    dj.alpha_f32 = job.norm.f32;
    dj.alpha_f16 = (uint16_t)(job.norm.f16x2 & 0xFFFFu);
    const int iters = (bg == 1) ? run_iterations<1, _Float16, MAXR>(hb, dj) : run_iterations<2, _Float16, MAXR>(hb, dj);
    store_hard<_Float16>(hb, Kb * Zc, dj.out, sh.words);
    if (job.cb_iterations && tid == 0) job.cb_iterations[r] = (uint8_t)iters;

    // ---- 5. CB CRC over ceil(K'/32) words, TB bytes (MSB-first) ---------
    const int Kp = (int)(p.K - p.F);
    const uint32_t* words = sh.words;
    auto byte_of = [&](int b) { return (rbit8x4(words[b >> 2]) >> (8 * (b & 3))) & 0xFFu; };
    const int kind = cb_crc_kind(p);
    const uint32_t crc = block_crc_bytes(4 * ((Kp + 31) / 32), kind, byte_of, sh.crc_t[kind], sh.red);
    const int data_bytes = cb_data_bytes(p);
    for (int b = tid; b < data_bytes; b += nthr) job.tb_bytes[(size_t)r * data_bytes + b] = (uint8_t)byte_of(b);
    if (tid == 0) {
        job.cb_crc[r] = crc;
        if (cb_in_tb == 0) {
            if (C == 1) *job.tb_crc = crc;
            const int bytes = C * data_bytes;
            for (int b = bytes; b < (bytes + 3) / 4 * 4; ++b) job.tb_bytes[b] = 0;
        }
    }
    if (C == 1) return 0;
    // ---- 6. this CB's share of the TB CRC24A ------------------------------
    // TB CRC = XOR over CBs of crc(payload bytes of CB c) * x^(8 * TB bytes
    // after them), the TB zero-padded to whole words (CRC linearity)
    uint32_t part = block_crc_bytes(data_bytes, K_CRC24A, byte_of, sh.crc_t[K_CRC24A], sh.red);
    const uint32_t total = (uint32_t)((C * data_bytes + 3) / 4 * 4);
    const uint32_t after = total - (uint32_t)(cb_in_tb + 1) * (uint32_t)data_bytes;
    if (part && after) part = dmulmod(part, crc24a_shift_bytes(after), 24, CRC24A_P.g);
    return part;
}

// Persistent fused slot kernel.  Workspace: queue state (2 words used of
// FUSED_WS_QUEUE_WORDS), then per TB a completion counter and a CRC
// accumulator.  A CB's share of its TB CRC is XORed into the accumulator by
// a device-scope atomic that has completed (its old value returned) before
// the same thread increments the TB's counter; the thread that brings the
// counter to C reads and clears the accumulator with another atomic.  All
// cross-workgroup data thus moves through atomics at the device coherence
// point: no cache write-back fences per code block.
template <typename TI, int MAXR>
__global__ __launch_bounds__(384) void pusch_slot_fused(const cuphyHipPuschSlotTb_t* __restrict__ tbs, int n_tbs,
                                                        const uint32_t* __restrict__ items, int n_items,
                                                        int lds_values, uint32_t* __restrict__ ws)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ fused_lds sh;
    _Float16* hb = reinterpret_cast<_Float16*>(lds_raw);
    uint32_t* queue = ws;
    uint32_t* tb_count = ws + FUSED_WS_QUEUE_WORDS;
    uint32_t* tb_acc = tb_count + n_tbs;
    for (int i = threadIdx.x; i < 3 * 256; i += blockDim.x) (&sh.crc_t[0][0])[i] = (&g_crc_tables.byte[0][0])[i];
    for (;;) {
        if (threadIdx.x == 0) sh.item = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        const int i = sh.item;
        __syncthreads();
        if (i >= n_items) break;
        const uint32_t it = items[i];
        const int t = (int)(it >> 16), cb = (int)(it & 0xFFFFu);
        if (t >= n_tbs) continue;  // malformed item: skipped
        const cuphyHipPuschSlotTb_t& job = tbs[t];
        const PerTbParams& p = *job.prm;
        const uint32_t C = (uint32_t)__builtin_amdgcn_readfirstlane((int)p.num_CBs);
        if ((uint32_t)cb >= C) continue;
        const uint32_t part = fused_cb<TI, MAXR>(hb, sh, job, p, cb, lds_values);
        if (C > 1 && threadIdx.x == 0) {
            // shares XOR; the rejection flag (bit 31, above every CRC value) ORs
            const uint32_t prev = (part == TB_ACC_ERROR) ? atomicOr(&tb_acc[t], part) : atomicXor(&tb_acc[t], part);
            // the XOR has been performed (its value returned) before the count
            asm volatile("s_waitcnt vmcnt(0)" ::"v"(prev) : "memory");
            if (atomicAdd(&tb_count[t], 1u) == C - 1) {
                const uint32_t acc = atomicExch(&tb_acc[t], 0u);
                atomicExch(&tb_count[t], 0u);
                *job.tb_crc = (acc & TB_ACC_ERROR) ? FUSED_CB_ERROR : acc;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&queue[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

void* fused_kernel(cuphyDataType_t llr_type, int max_parity_nodes)
{
    const bool small = max_parity_nodes <= SMALL_MAXR;
    switch (llr_type) {
    case CUPHY_R_8I: return small ? (void*)&pusch_slot_fused<int8_t, SMALL_MAXR> : (void*)&pusch_slot_fused<int8_t, 46>;
    case CUPHY_R_16F: return small ? (void*)&pusch_slot_fused<_Float16, SMALL_MAXR> : (void*)&pusch_slot_fused<_Float16, 46>;
    case CUPHY_R_32F: return small ? (void*)&pusch_slot_fused<float, SMALL_MAXR> : (void*)&pusch_slot_fused<float, 46>;
    default: return nullptr;
    }
}

}  // namespace nrldpc

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace nrldpc {

cuphyStatus_t ensure_lds_attr(void* func, size_t bytes);  // ldpc_decode.hip

cuphyStatus_t slot_decode(const cuphyHipPuschSlotTb_t* d_tbs, int n_tbs, const uint32_t* d_items, int n_items,
                          cuphyDataType_t llr_type, int max_lds_values, int max_parity_nodes, void* d_ws,
                          int cu_count, hipStream_t strm)
{
    if (n_items < 0 || n_tbs < 0 || n_tbs > 65536 || !d_ws) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (n_items > 0 && (!d_tbs || !d_items || n_tbs == 0)) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (max_lds_values <= 0 || max_lds_values > 68 * 384) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (max_parity_nodes < 4 || max_parity_nodes > 46) return CUPHY_STATUS_INVALID_ARGUMENT;
    void* func = fused_kernel(llr_type, max_parity_nodes);
    if (!func) return CUPHY_STATUS_UNSUPPORTED_TYPE;
    const size_t lds = ((size_t)max_lds_values * 2 + 15) / 16 * 16;
    cuphyStatus_t s = ensure_lds_attr(func, lds);  // also for an empty launch: prepares graph capture
    if (s != CUPHY_STATUS_SUCCESS) return s;
    if (n_items == 0) return CUPHY_STATUS_SUCCESS;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, func, 384, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    const int grid = std::min(n_items, std::max(1, cu_count) * per_cu);
    uint32_t* ws = static_cast<uint32_t*>(d_ws);
    void* args[6] = {const_cast<cuphyHipPuschSlotTb_t**>(&d_tbs), &n_tbs, const_cast<uint32_t**>(&d_items),
                     &n_items, &max_lds_values, &ws};
    if (hipLaunchKernel(func, dim3((unsigned)grid), dim3(384), args, lds, strm) != hipSuccess)
        return CUPHY_STATUS_INTERNAL_ERROR;
    return CUPHY_STATUS_SUCCESS;
}

}  // namespace nrldpc
