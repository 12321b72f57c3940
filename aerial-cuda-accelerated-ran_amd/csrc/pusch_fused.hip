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
// place in the TB, XORs it into the TB's accumulator word with a device
// atomic and then counts the CB done on the TB's counter word; the
// workgroup that finishes the TB's last CB takes (and zeroes) the
// accumulator, zeroes the counter and writes the TB CRC.  No workgroup ever
// waits for another.  With CUPHY_HIP_SLOT_TB_NOTIFY each CB's outputs are
// first written back from the XCD's L2 (agent-scope release), so the TB CRC
// appears only after every output of the TB is in memory: a host polling
// the TB CRC in host-mapped memory can take each TB as soon as its own code
// blocks are decoded.  The queue is returned to 0 by the last workgroup to
// leave, so the launch replays from a graph.
#include "crc_core.hpp"
#include "derate_core.hpp"
#include "gold_core.hpp"
#include "ldpc_core.hpp"
#include "pusch_fused.hpp"

#include <algorithm>
#include <mutex>

namespace nrldpc {

constexpr uint32_t FUSED_CB_ERROR = 0xFFFFFFFFu;  // CRC value of a CB/TB whose parameters were rejected
constexpr uint32_t TB_ACC_ERROR = 0x80000000u;    // accumulator flag: a CB of the TB was rejected
constexpr int SLOT_CB_WORDS = MAX_DECODED_CODE_BLOCK_BIT_SIZE / 32;  // 264: decoded-word stride per CB

template <typename TI> __device__ __forceinline__ float llr_as_float(TI v) { return (float)v; }

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

// LDS state of a workgroup besides the HARQ/APP array.
struct fused_lds {
    uint32_t words[SLOT_CB_WORDS];       // hard bits of the current CB
    uint32_t crc_t[3][256];              // byte tables of the three CRCs
    uint32_t shift[3][CRC_SHIFT_LO];     // x^(8m) mod g of the three CRCs
    uint32_t red[2][8];
    uint32_t cb_flags[FUSED_MAX_GROUP];  // per-CB early-termination flags of a group
    uint32_t cb_iters[FUSED_MAX_GROUP];  // iterations each CB of a group ran
    int item, next;
};

// Remainder of one word's 4 message bytes (MSB-first bytes) with byte table T.
template <int L>
__device__ __forceinline__ uint32_t crc_word(uint32_t w, const uint32_t* T)
{
    constexpr uint32_t mask = (1u << L) - 1u;
    uint32_t reg = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) reg = ((reg << 8) & mask) ^ T[((reg >> (L - 8)) ^ (w >> (8 * b))) & 0xFFu];
    return reg;
}

// XOR of x over the wave, returned in every lane: within each row of 16
// lanes by DPP (lane pairs, quads, half rows, rows: register-to-register,
// no LDS round trip), then across the four rows by readlane.  Every lane
// of the wave must be active (workgroups of whole waves, called in uniform
// control flow).
__device__ __forceinline__ uint32_t wave_xor(uint32_t x)
{
    x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, false);  // row_half_mirror
    x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x140, 0xF, 0xF, false);  // row_mirror
    return (uint32_t)(__builtin_amdgcn_readlane((int)x, 0) ^ __builtin_amdgcn_readlane((int)x, 16) ^
                      __builtin_amdgcn_readlane((int)x, 32) ^ __builtin_amdgcn_readlane((int)x, 48));
}

// The CB CRC (kind K of L bits over nb_cb = 4 * words bytes) and the CB's
// TB-CRC24A share (over its db payload bytes, before the shift to its
// place in the TB) in one pass: thread t takes word t, shifts its two
// remainders by x^(8 * bytes after it) with factors from LDS, and both are
// XOR-reduced together.
template <int L, uint32_t G>
__device__ __forceinline__ void cb_and_share_crc(const fused_lds& sh, const uint32_t* T, const uint32_t* S, int nb_cb,
                                                 int db, uint32_t* red0, uint32_t* red1, uint32_t& crc,
                                                 uint32_t& share)
{
    const int t = threadIdx.x;
    uint32_t rc = 0, rs = 0;
    if (4 * t < nb_cb) {
        // the two remainders as independent straight-line chains (a shift
        // of 0 bytes multiplies by x^0 = 1), so that their table lookups
        // and products interleave
        const uint32_t w = rbit8x4(sh.words[t]);
        const uint32_t* TA = sh.crc_t[K_CRC24A];
        rc = crc_word<L>(w, T);
        if (4 * t + 4 <= db) {
            rs = crc_word<24>(w, TA);
        } else if (4 * t < db) {  // the payload's last, partial word: bytes [4t, db)
            for (int b = 0; b < db - 4 * t; ++b)
                rs = ((rs << 8) & 0xFFFFFFu) ^ TA[((rs >> 16) ^ (w >> (8 * b))) & 0xFFu];
        }
        const int nb = max(0, min(4, db - 4 * t));
        rc = cmulmod<L, G>(rc, S[nb_cb - 4 * t - 4]);
        rs = cmulmod<24, CRC24A_P.g>(rs, sh.shift[K_CRC24A][max(0, db - 4 * t - nb)]);
    }
    rc = wave_xor(rc);
    rs = wave_xor(rs);
    if ((t & 63) == 0) {
        red0[t >> 6] = rc;
        red1[t >> 6] = rs;
    }
    __syncthreads();
    crc = share = 0;
    for (int i = 0; i < (int)((blockDim.x + 63) / 64); ++i) {
        crc ^= red0[i];
        share ^= red1[i];
    }
    __syncthreads();
}

// Combine this transmission's E LLRs of the CB into the LDS circular buffer
// `cbuf` (passes of distinct positions, barrier after each).  Indices are
// advanced incrementally: one division per thread and pass.  DESCR: LLR i
// (= k * EQ + j) is bit n = start + j * Qm + k of the TB's rate-matched
// sequence and is negated where c(n) = 1; seq holds the words of c from
// word start / 32 on, soff = start % 32.
template <typename TI, bool DESCR>
__device__ __forceinline__ void combine_tx(_Float16* __restrict__ cbuf, const TI* __restrict__ in,
                                           const rm_geometry& g, const PerTbParams& p, float scale,
                                           const uint32_t* __restrict__ seq, int soff)
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
                _Float16 v = rm_num<_Float16>::clamp((_Float16)x);
                if (DESCR) {
                    const int n = soff + j * g.Qm + k;
                    if ((seq[n >> 5] >> (n & 31)) & 1u) v = -v;
                }
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

template <typename TI> struct slot_vec;  // the QAM_STRIDE LLR slots of one symbol
template <> struct slot_vec<int8_t> {
    typedef int8_t type __attribute__((ext_vector_type(8)));
};
template <> struct slot_vec<_Float16> {
    typedef _Float16 type __attribute__((ext_vector_type(8)));
};
template <> struct slot_vec<float> {
    typedef float type __attribute__((ext_vector_type(8)));
};

// The QM bits of symbol j combined into the buffer: their positions are
// distinct (E <= L), so all QM reads are issued before the QM writes.
template <int QM, typename TI, bool DESCR, typename V>
__device__ __forceinline__ void combine_symbol(_Float16* __restrict__ cbuf, const V& v, const rm_geometry& g, int j,
                                               float scale, uint32_t bits)
{
    int o[QM];
    _Float16 old[QM];
#pragma unroll
    for (int k = 0; k < QM; ++k) {
        int c = g.c0 + k * g.EQ + j;
        if (c >= g.L) c -= g.L;
        o[k] = (c < g.Kd) ? c : c + g.F;
        old[k] = cbuf[o[k]];
    }
#pragma unroll
    for (int k = 0; k < QM; ++k) {
        _Float16 x = rm_num<_Float16>::clamp((_Float16)(llr_as_float(v[k]) * scale));
        if (DESCR && ((bits >> k) & 1u)) x = -x;
        cbuf[o[k]] = rm_num<_Float16>::clamp(old[k] + x);
    }
}

// Combining when no circular-buffer position repeats within this
// transmission (E <= L), the layers map one to one onto the buffer's
// layers and the LLRs are not UCI-multiplexed: thread t takes whole symbols
// j = t, t + nthr, ... (one vector load of the symbol's QAM_STRIDE slots,
// all loads of a thread issued before its LDS updates); bit k of symbol j
// is de-interleaved bit i = k * E/Qm + j, at circle position c0 + i, and
// bit n = start + j * Qm + k of the TB's rate-matched sequence (DESCR: the
// symbol's Qm sequence bits are taken from one 64-bit window of seq).
template <typename TI, bool DESCR>
__device__ __forceinline__ void combine_symbols(_Float16* __restrict__ cbuf, const TI* __restrict__ in,
                                                const rm_geometry& g, float scale, const uint32_t* __restrict__ seq,
                                                int soff)
{
    typedef typename slot_vec<TI>::type vec_t;
    constexpr int NB = 4;  // symbols per thread in flight
    const int nthr = blockDim.x;
    const vec_t* sym = reinterpret_cast<const vec_t*>(in + g.adj);
    for (int j0 = threadIdx.x; j0 < g.EQ; j0 += NB * nthr) {
        vec_t v[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int j = j0 + u * nthr;
            if (j < g.EQ) v[u] = sym[j];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int j = j0 + u * nthr;
            if (j >= g.EQ) break;
            uint32_t bits = 0;
            if (DESCR) {  // bits n0 .. n0 + Qm - 1, n0 = soff + j * Qm (word n0/32 + 1 exists)
                const int n0 = soff + j * g.Qm;
                const uint64_t w2 = ((uint64_t)seq[(n0 >> 5) + 1] << 32) | seq[n0 >> 5];
                bits = (uint32_t)(w2 >> (n0 & 31));
            }
            switch (g.Qm) {  // (uniform) modulation order as a constant
            case 8: combine_symbol<8, TI, DESCR>(cbuf, v[u], g, j, scale, bits); break;
            case 6: combine_symbol<6, TI, DESCR>(cbuf, v[u], g, j, scale, bits); break;
            case 4: combine_symbol<4, TI, DESCR>(cbuf, v[u], g, j, scale, bits); break;
            case 2: combine_symbol<2, TI, DESCR>(cbuf, v[u], g, j, scale, bits); break;
            default:
                for (int k = 0; k < g.Qm; ++k) {
                    int c = g.c0 + k * g.EQ + j;
                    if (c >= g.L) c -= g.L;
                    const int o = (c < g.Kd) ? c : c + g.F;
                    _Float16 x = rm_num<_Float16>::clamp((_Float16)(llr_as_float(v[u][k]) * scale));
                    if (DESCR && ((bits >> k) & 1u)) x = -x;
                    cbuf[o] = rm_num<_Float16>::clamp(cbuf[o] + x);
                }
            }
        }
    }
    __syncthreads();
}

// Words of the descrambling sequence covering a CB's rate-matched bits
// [start, start + E): from word start / 32 to word (start + E - 1) / 32.
__device__ __forceinline__ int rm_seq_words(const rm_geometry& g)
{
    return (int)(((uint32_t)(g.start + g.E - 1) >> 5) - ((uint32_t)g.start >> 5) + 1u);
}

// Steps 1-3 of one CB: its HARQ circular buffer into LDS `hb`, this
// transmission combined (descrambled with `seq` when descr), the buffer
// written back to HBM and converted in place to the decoder's APP array of
// nval values.  Ends with a barrier.
template <typename TI>
__device__ __forceinline__ void fused_combine(_Float16* __restrict__ hb, uint32_t* __restrict__ seq,
                                              const cuphyHipPuschSlotTb_t& job, const PerTbParams& p,
                                              const rm_geometry& g, int r, int nh, int nval, bool descr)
{
    const int Zc = g.Zc;
    const size_t ncb_padded = (size_t)__builtin_amdgcn_readfirstlane((int)p.Ncb_padded);
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
        for (int n0 = tid; n0 < nv; n0 += 4 * nthr) {  // four 16-byte loads in flight per thread
            half8_t v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (n0 + u * nthr < nv) v[u] = reinterpret_cast<const half8_t*>(harq)[n0 + u * nthr];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (n0 + u * nthr < nv) reinterpret_cast<half8_t*>(hb)[n0 + u * nthr] = v[u];
        }
        for (int n = nv * 8 + tid; n < nh; n += nthr) hb[n] = harq[n];
    }
    const uint32_t w_base = (uint32_t)g.start >> 5;
    if (descr) {  // each wave generates its share of the words while the loads land
        const int nw = rm_seq_words(g);
        const int nwave = nthr >> 6, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
        const int per = (nw + nwave - 1) / nwave;
        const int a = min(nw, wave * per), b = min(nw, a + per);
        gold_words_wave(p.cinit, w_base + (uint32_t)a, w_base + (uint32_t)b, seq + a);
        if (tid == 0) seq[nw] = 0u;
    }
    __syncthreads();
    if (g.ndi)
        for (int n = tid; n < g.F; n += nthr) cbuf[g.Kd + n] = (_Float16)LLR_MAX_ABS;

    // ---- 2. combine this transmission -------------------------------------
    const bool sym_path = !g.uci && g.Nl == g.nbbu && g.E <= g.L &&
                          ((uintptr_t)(in + g.adj) % sizeof(typename slot_vec<TI>::type)) == 0;
    const int soff = g.start & 31;
    if (descr) {
        if (sym_path)
            combine_symbols<TI, true>(cbuf, in, g, job.llr_scale, seq, soff);
        else
            combine_tx<TI, true>(cbuf, in, g, p, job.llr_scale, seq, soff);
    } else {
        if (sym_path)
            combine_symbols<TI, false>(cbuf, in, g, job.llr_scale, seq, soff);
        else
            combine_tx<TI, false>(cbuf, in, g, p, job.llr_scale, seq, soff);
    }

    // ---- 3. write back the HARQ buffer; APP = clamp(HARQ), +0 for -0 ----
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
    __syncthreads();
}

// A group of `count` consecutive CBs cb0 .. cb0 + count - 1 of one TB
// (count * Zc <= blockDim.x).  CB c's HARQ/APP array is at hb + c * S
// (S = max(2Zc + Ncb, (nsys + mb) * Zc), rounded up to whole 16-byte units
// when count > 1); descrambling words follow the last array.  Steps 1-3
// run CB by CB with the whole workgroup; the decode runs the group side by
// side (lanes [c * Zc, (c + 1) * Zc) on CB c, each CB stopping on its own
// syndrome: results as decoded alone); hard decisions and CRCs again CB by
// CB.  Returns, in every thread, the XOR of the CBs' shares of the TB
// CRC24A (each the remainder of the CB's payload bytes shifted to their
// place in the TB; for a one-CB TB its CRC, which is the TB CRC), or
// TB_ACC_ERROR when the group's parameters do not fit the launch (then only
// the error marker is written, into the CBs' CRCs).  The TB CRC itself is
// written by complete_cb.  GROUPED = false: the one-CB form (count = 1 at
// compile time, no group state; an item of several CBs is rejected).
template <typename TI, int MAXR, bool GROUPED>
__device__ __forceinline__ uint32_t fused_group(_Float16* __restrict__ hb, fused_lds& sh,
                                                const cuphyHipPuschSlotTb_t& job, const PerTbParams& p, int cb0,
                                                int count_in, int lds_values, uint32_t* __restrict__ tbw,
                                                uint32_t& xor_prev)
{
    const int count = GROUPED ? count_in : 1;
    rm_geometry g0 = rm_geometry_of(p, cb0);
    rm_make_uniform(g0);
    const int Zc = g0.Zc;
    const int bg = __builtin_amdgcn_readfirstlane((int)p.bg);
    const int nsys = bg == 1 ? 22 : 10;
    const int mb = job.num_parity_nodes;
    const int nh = 2 * Zc + g0.Ncb;
    const int nval = (nsys + mb) * Zc;
    const int arr = max(nh, nval);
    const int S = (count == 1) ? arr : group_stride(arr);
    const int r0 = cb0 + __builtin_amdgcn_readfirstlane((int)p.firstCodeBlockIndex);
    const int Kb = __builtin_amdgcn_readfirstlane((int)p.K) / Zc;
    const int C = __builtin_amdgcn_readfirstlane((int)p.num_CBs);
    // descrambling: the CB's words of c (and one more word) after the arrays
    const bool descr = (job.flags & CUPHY_HIP_SLOT_TB_DESCRAMBLE) != 0 && !g0.uci;
    const int seq_at = ((count - 1) * S + arr + 1) & ~1;
    // every CB of the group: E > 0 and (descrambled) its words within LDS
    int nw_max = 0;
    bool e_ok = g0.L > 0;
    for (int c = 0; c < count; ++c) {
        const rm_geometry g = (c == 0) ? g0 : rm_geometry_of(p, cb0 + c);
        e_ok = e_ok && g.E > 0;
        if (descr && g.E > 0) nw_max = max(nw_max, rm_seq_words(g));
    }
    const bool ok = (bg == 1 || bg == 2) && z_index(Zc) >= 0 && mb >= 4 && mb <= (bg == 1 ? 46 : 42) &&
                    mb <= MAXR && count == count_in && count >= 1 && count <= FUSED_MAX_GROUP &&
                    count * Zc <= (int)blockDim.x &&
                    (count - 1) * S + arr <= lds_values && Kb * Zc <= 32 * (int)SLOT_CB_WORDS && e_ok &&
                    (!descr || seq_at + 2 * (nw_max + 1) <= lds_values);
    const int tid = threadIdx.x, nthr = blockDim.x;
    if (!ok) {
        if (tid < count_in) job.cb_crc[r0 + tid] = FUSED_CB_ERROR;
        return TB_ACC_ERROR;
    }
    uint32_t* seq = reinterpret_cast<uint32_t*>(hb + seq_at);

    if (tid < count) sh.cb_iters[tid] = 0u;  // (read after the barriers below)

    // ---- 1-3. per CB: HARQ buffer, combining, write-back, APP --------------
    for (int c = 0; c < count; ++c) {
        rm_geometry g = g0;  // (the one-CB form keeps the geometry computed above)
        if (GROUPED) {
            g = rm_geometry_of(p, cb0 + c);
            rm_make_uniform(g);
        }
        fused_combine<TI>(hb + c * S, seq, job, p, g, r0 + c, nh, nval, descr);
    }

    // ---- 4. decode -------------------------------------------------------
    cb_job dj;
    dj.llr = nullptr;
    dj.out = nullptr;
    dj.soft = nullptr;
    dj.Z = Zc;
    dj.mb = mb;
    dj.Kb = Kb;
    dj.max_iter = job.max_iterations;
    dj.early_term = (job.flags & CUPHY_LDPC_DECODE_EARLY_TERM) != 0;
    dj.clamp = job.clamp_value;
    dj.alpha_f32 = job.norm.f32;
    dj.alpha_f16 = (uint16_t)(job.norm.f16x2 & 0xFFFFu);
    uint32_t* const flags = (count > 1) ? sh.cb_flags : nullptr;
    uint32_t* const cb_it = (count > 1) ? sh.cb_iters : nullptr;
    const int iters = (bg == 1) ? run_iterations<1, _Float16, MAXR>(hb, dj, count, S, flags, cb_it)
                                : run_iterations<2, _Float16, MAXR>(hb, dj, count, S, flags, cb_it);

    // ---- 5-6. per CB: hard decisions, CB CRC, TB bytes, TB-CRC share --------
    const int Kp = (int)(p.K - p.F);
    const int kind = cb_crc_kind(p);
    const int data_bytes = cb_data_bytes(p);
    const uint32_t* words = sh.words;
    auto byte_of = [&](int b) { return (rbit8x4(words[b >> 2]) >> (8 * (b & 3))) & 0xFFu; };
    uint32_t acc = 0;
    for (int c = 0; c < count; ++c) {
        const int r = r0 + c, cb_in_tb = cb0 + c;
        store_hard_words<_Float16>(hb + c * S, Kb * Zc, job.cb_bits + (size_t)r * SLOT_CB_WORDS, sh.words);
        if (job.cb_iterations && tid == 0) job.cb_iterations[r] = (uint8_t)(count == 1 ? iters : (int)sh.cb_iters[c]);
        uint32_t crc, part;
        if (kind == K_CRC16)
            cb_and_share_crc<16, CRC16_P.g>(sh, sh.crc_t[kind], sh.shift[kind], 4 * ((Kp + 31) / 32), data_bytes,
                                            sh.red[0], sh.red[1], crc, part);
        else if (kind == K_CRC24B)
            cb_and_share_crc<24, CRC24B_P.g>(sh, sh.crc_t[kind], sh.shift[kind], 4 * ((Kp + 31) / 32), data_bytes,
                                             sh.red[0], sh.red[1], crc, part);
        else
            cb_and_share_crc<24, CRC24A_P.g>(sh, sh.crc_t[kind], sh.shift[kind], 4 * ((Kp + 31) / 32), data_bytes,
                                             sh.red[0], sh.red[1], crc, part);
        if (tid == 0) {
            job.cb_crc[r] = crc;
            if (cb_in_tb == 0) {
                const int bytes = C * data_bytes;
                for (int b = bytes; b < (bytes + 3) / 4 * 4; ++b) job.tb_bytes[b] = 0;
            }
        }
        if (C > 1) {
            // TB CRC = XOR over CBs of crc(payload bytes of CB c) * x^(8 * TB
            // bytes after them), the TB zero-padded to whole words (CRC linearity)
            const uint32_t total = (uint32_t)((C * data_bytes + 3) / 4 * 4);
            const uint32_t after = total - (uint32_t)(cb_in_tb + 1) * (uint32_t)data_bytes;
            // uniform operands: the two products run on the scalar unit
            part = (uint32_t)__builtin_amdgcn_readfirstlane((int)part);
            if (part && after) {
                const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)sh.shift[K_CRC24A][after & 1023u]);
                part = cmulmod<24, CRC24A_P.g>(part,
                                               cmulmod<24, CRC24A_P.g>(lo, g_crc_shift.tb_hi[(after >> 10) & 255u]));
            }
            acc ^= part;
            // the group's share goes into the TB's accumulator now, so that
            // the atomic's round trip overlaps the byte stores below
            // (complete_cb raises the TB's count once it has returned)
            if (c + 1 == count && tid == 0)
                xor_prev = __hip_atomic_fetch_xor(tbw + 2u * job.share_offset, acc, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
        }
        for (int b = tid; b < data_bytes; b += nthr) job.tb_bytes[(size_t)r * data_bytes + b] = (uint8_t)byte_of(b);
        if (C == 1) return crc;
        if (c + 1 < count) __syncthreads();  // the words in LDS are the next CB's
    }
    return acc;
}

// Workspace words of a TB (after the queue words): TB-CRC accumulator at
// 2 * share_offset, count of finished CBs at 2 * share_offset + 1.  Called
// by one thread once the whole workgroup has finished a group of n CBs of
// the TB (all of their outputs issued).  v: the XOR of the group's shares
// (TB_ACC_ERROR: rejected group); for C > 1 and an accepted group, the
// shares are already in the accumulator and xor_prev is its old value.
__device__ __forceinline__ void complete_cb(const cuphyHipPuschSlotTb_t& job, uint32_t C, uint32_t v, uint32_t n,
                                            uint32_t* __restrict__ tbw, uint32_t xor_prev)
{
    const bool notify = (job.flags & CUPHY_HIP_SLOT_TB_NOTIFY) != 0;
    // every output of this CB out of the XCD's L2 before it is counted
    if (notify) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    uint32_t tb_crc;
    if (C == 1) {
        tb_crc = (v == TB_ACC_ERROR) ? FUSED_CB_ERROR : v;
    } else {
        uint32_t* acc = tbw + 2u * job.share_offset;
        // an accepted group's share was XORed in by fused_group (its old
        // value is xor_prev); a rejected group flags the accumulator here
        uint32_t prev = (v == TB_ACC_ERROR)
                            ? __hip_atomic_fetch_or(acc, 0x80000000u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                            : xor_prev;
        // the share is in the accumulator (its old value returned) before
        // the count is raised
        asm volatile("" : : "v"(prev) : "memory");
        if (__hip_atomic_fetch_add(acc + 1, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != C - n) return;
        const uint32_t a = __hip_atomic_exchange(acc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(acc + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        tb_crc = (a & 0x80000000u) ? FUSED_CB_ERROR : a;
    }
    if (notify)
        __hip_atomic_store(job.tb_crc, tb_crc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    else
        *job.tb_crc = tb_crc;
}

// Persistent fused slot kernel.  Workspace: queue state (2 words used of
// FUSED_WS_QUEUE_WORDS), then two words per CB slot of the launch, of which
// each TB uses the pair at its first slot (TB-CRC accumulator and count of
// finished CBs; see complete_cb).  The next item index is fetched while the current CB is
// worked on.  The queue is returned to 0 by the last workgroup to leave.
// The high-rate variant (MAXR <= 8) and the middle one (MAXR <= 16) are
// held to 4 waves per SIMD (at most 128 VGPRs): a 384-thread workgroup puts
// its 6 waves 2/2/1/1 on the CU's SIMDs, so a second workgroup on the CU
// needs 4 wave slots on two SIMDs; at its natural 157 VGPRs (3 slots) the
// high-rate variant ran one workgroup per CU.  (The middle variant's BG1
// rows beyond 8 spill a few registers; BG2 rows, half as long, do not.)
template <typename TI, int MAXR, bool GROUPED>
__global__ __launch_bounds__(384) __attribute__((amdgpu_waves_per_eu(MAXR <= MID_MAXR ? 4 : 1)))
void pusch_slot_fused(const cuphyHipPuschSlotTb_t* __restrict__ tbs, int n_tbs,
                                                        const uint32_t* __restrict__ items, int n_items,
                                                        int lds_values, uint32_t* __restrict__ ws)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ fused_lds sh;
    _Float16* hb = reinterpret_cast<_Float16*>(lds_raw);
    uint32_t* queue = ws;
    uint32_t* tbw = ws + FUSED_WS_QUEUE_WORDS;
    for (int i = threadIdx.x; i < 3 * 256; i += blockDim.x) (&sh.crc_t[0][0])[i] = (&g_crc_tables.byte[0][0])[i];
    for (int i = threadIdx.x; i < 3 * CRC_SHIFT_LO; i += blockDim.x) (&sh.shift[0][0])[i] = (&g_crc_shift.lo[0][0])[i];
    if (threadIdx.x == 0) sh.item = (int)atomicAdd(&queue[0], 1u);
    __syncthreads();
    for (;;) {
        const int i = sh.item;
        if (i >= n_items) break;
        if (threadIdx.x == 0) sh.next = (int)atomicAdd(&queue[0], 1u);  // consumed after this CB
        const uint32_t it = (uint32_t)__builtin_amdgcn_readfirstlane((int)items[i]);
        const int t = (int)(it >> 16), cb = item_cb(it), n = item_count(it);
        if (t < n_tbs) {
            const cuphyHipPuschSlotTb_t& job = tbs[t];
            const PerTbParams& p = *job.prm;
            const uint32_t C = (uint32_t)__builtin_amdgcn_readfirstlane((int)p.num_CBs);
            if ((uint32_t)(cb + n) <= C) {
                uint32_t xprev = 0;
                const uint32_t part = fused_group<TI, MAXR, GROUPED>(hb, sh, job, p, cb, n, lds_values, tbw, xprev);
                __syncthreads();  // every output of the group issued by every thread
                if (threadIdx.x == 0) complete_cb(job, C, part, (uint32_t)n, tbw, xprev);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) sh.item = sh.next;
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

// ---------------------------------------------------------------------------
// Resident serving kernel: one persistent launch that takes code blocks from
// a ring in host-mapped memory as the host publishes them (no launch per
// batch).  Work item i is ring items[i & mask] = (TB slot << 16) | CB; the
// TB's job descriptor (and, through it, its PerTbParams) is in host-mapped
// memory too and is copied to LDS per CB.  A workgroup takes a ticket (the
// next item index: one device atomic add, no retries) and waits until the
// host has published that item; one waiting workgroup at a time refreshes
// the device copy of the host's tail.  Every wait is bounded: a workgroup
// leaves after idle_ticks of the wall clock without work, when the host
// sets `stop`, or after life_ticks in all, and then leaves its ticket in
// abandoned[workgroup]; the same workgroup of the next launch (same grid)
// starts with it, so no item is lost or taken twice.  Leaving is collective:
// the first workgroup to leave marks the launch (st->leave = epoch, the
// launch's number), and every other workgroup then leaves at its next wait,
// parking its ticket the same way, without taking a new one -- so no
// published item waits behind a parked ticket while the launch goes on
// serving later ones; the launch ends within one code block's time.  The
// host launches the kernel again whenever work is open and none is
// running.  Each CB starts
// with an agent-scope acquire (inputs and HARQ buffers written since may
// sit behind stale lines of this XCD's L2) and ends with the TB-CRC
// protocol of complete_cb in notify mode.
// ---------------------------------------------------------------------------
constexpr uint32_t NO_TICKET = 0xFFFFFFFFu;

__device__ __forceinline__ int serve_next(const serve_ring_hdr* ring, serve_state* st, uint32_t* abandoned,
                                          bool first, uint64_t t_start, uint64_t idle_ticks, uint64_t life_ticks,
                                          uint32_t epoch)
{
    uint32_t i = NO_TICKET;
    if (first) {
        i = __hip_atomic_exchange(&abandoned[blockIdx.x], NO_TICKET, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (i == NO_TICKET) {
        // the launch is leaving: take no new ticket
        if (__hip_atomic_load(&st->leave, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) return -1;
        i = __hip_atomic_fetch_add(&st->head, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint64_t t_wait = wall_clock64();
    for (;;) {
        const uint32_t t = __hip_atomic_load(&st->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((int32_t)(t - i) > 0) return (int)(i & 0x7FFFFFFFu);
        // not published yet: one workgroup at a time reads the host's tail
        if (__hip_atomic_exchange(&st->token, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
            const uint32_t ht = __hip_atomic_load(&ring->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const uint32_t hs = __hip_atomic_load(&ring->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const uint32_t cur = __hip_atomic_load(&st->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((int32_t)(ht - cur) > 0) __hip_atomic_store(&st->tail, ht, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&st->stop, hs ? 1u : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&st->token, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((int32_t)(ht - i) > 0) return (int)(i & 0x7FFFFFFFu);
        }
        const uint64_t now = wall_clock64();
        if (__hip_atomic_load(&st->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
            __hip_atomic_load(&st->leave, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch ||
            now - t_wait > idle_ticks || now - t_start > life_ticks) {
            __hip_atomic_store(&abandoned[blockIdx.x], i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&st->leave, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return -1;
        }
        __builtin_amdgcn_s_sleep(32);
    }
}

template <typename TI>
__global__ __launch_bounds__(384) void pusch_serve_resident(const serve_ring_hdr* ring,
                                                            const uint32_t* __restrict__ ring_items, uint32_t ring_mask,
                                                            const cuphyHipPuschSlotTb_t* jobs, int n_slots,
                                                            serve_state* st, uint32_t* abandoned,
                                                            uint32_t* __restrict__ tbw, int lds_values,
                                                            uint64_t idle_ticks, uint64_t life_ticks, uint32_t epoch)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ fused_lds sh;
    __shared__ cuphyHipPuschSlotTb_t s_job;
    __shared__ PerTbParams s_prm;
    _Float16* hb = reinterpret_cast<_Float16*>(lds_raw);
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int i = tid; i < 3 * 256; i += nthr) (&sh.crc_t[0][0])[i] = (&g_crc_tables.byte[0][0])[i];
    for (int i = tid; i < 3 * CRC_SHIFT_LO; i += nthr) (&sh.shift[0][0])[i] = (&g_crc_shift.lo[0][0])[i];
    const uint64_t t_start = wall_clock64();
    bool first = true;
    for (;;) {
        if (tid == 0) sh.item = serve_next(ring, st, abandoned, first, t_start, idle_ticks, life_ticks, epoch);
        first = false;
        __syncthreads();
        const int i = sh.item;
        if (i < 0) break;
        const uint32_t it = (uint32_t)__builtin_amdgcn_readfirstlane((int)__hip_atomic_load(
            &ring_items[(uint32_t)i & ring_mask], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
        const int t = (int)(it >> 16), cb = item_cb(it), n = item_count(it);
        if (t < n_slots) {
            // the TB's descriptor and parameters from host memory into LDS
            const uint32_t* js = reinterpret_cast<const uint32_t*>(&jobs[t]);
            for (int k = tid; k < (int)(sizeof(cuphyHipPuschSlotTb_t) / 4); k += nthr)
                reinterpret_cast<uint32_t*>(&s_job)[k] =
                    __hip_atomic_load(js + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __syncthreads();
            const uint32_t* ps = reinterpret_cast<const uint32_t*>(s_job.prm);
            for (int k = tid; k < (int)(sizeof(PerTbParams) / 4); k += nthr)
                reinterpret_cast<uint32_t*>(&s_prm)[k] =
                    __hip_atomic_load(ps + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            if (tid == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __syncthreads();
            const uint32_t C = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_prm.num_CBs);
            if ((uint32_t)(cb + n) <= C) {
                uint32_t xprev = 0;
                const uint32_t part = fused_group<TI, 46, true>(hb, sh, s_job, s_prm, cb, n, lds_values, tbw, xprev);
                __syncthreads();  // every output of the group issued by every thread
                if (tid == 0) complete_cb(s_job, C, part, (uint32_t)n, tbw, xprev);
            }
        }
        __syncthreads();
    }
}

// Scratch of cuphyHipPuschSlotDecodeNoAlloc: the queue state and two words
// per CB of the launch (the workspace of cuphyHipPuschSlotDecode), for
// callers that cannot allocate device memory.  Each GPU the library runs
// on holds one copy until the process ends.
static __device__ __align__(256) uint32_t scratch[FUSED_WS_QUEUE_WORDS + 2 * NOALLOC_MAX_CBS];

// Zero n words (a kernel rather than a memset, so that a caller may capture
// the call in a graph).
__global__ __launch_bounds__(256) void zero_words(uint32_t* __restrict__ p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

template <typename TI, bool GROUPED>
void* fused_kernel_rows(int max_parity_nodes)
{
    return max_parity_nodes <= SMALL_MAXR ? (void*)&pusch_slot_fused<TI, SMALL_MAXR, GROUPED>
           : max_parity_nodes <= MID_MAXR ? (void*)&pusch_slot_fused<TI, MID_MAXR, GROUPED>
                                          : (void*)&pusch_slot_fused<TI, 46, GROUPED>;
}

template <bool GROUPED>
void* fused_kernel_of(cuphyDataType_t llr_type, int max_parity_nodes)
{
    switch (llr_type) {
    case CUPHY_R_8I: return fused_kernel_rows<int8_t, GROUPED>(max_parity_nodes);
    case CUPHY_R_16F: return fused_kernel_rows<_Float16, GROUPED>(max_parity_nodes);
    case CUPHY_R_32F: return fused_kernel_rows<float, GROUPED>(max_parity_nodes);
    default: return nullptr;
    }
}

// Kernel of a launch: LLR type, mb class (rows <= 8, <= 16, all), and
// whether items hold groups (the one-CB form keeps no group state:
// measurably faster per CB).
void* fused_kernel(cuphyDataType_t llr_type, int max_parity_nodes, bool grouped)
{
    return grouped ? fused_kernel_of<true>(llr_type, max_parity_nodes)
                   : fused_kernel_of<false>(llr_type, max_parity_nodes);
}

}  // namespace nrldpc

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace nrldpc {

cuphyStatus_t ensure_lds_attr(void* func, size_t bytes);  // ldpc_decode.hip

int serve_resident_grid(int cu_count)
{
    void* func = (void*)&pusch_serve_resident<_Float16>;
    const size_t lds = ((size_t)SERVE_LDS_VALUES * 2 + 15) / 16 * 16 + 64u;
    if (ensure_lds_attr(func, lds) != CUPHY_STATUS_SUCCESS) return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, func, FUSED_BLOCK, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return std::max(1, cu_count) * per_cu;
}

cuphyStatus_t serve_resident_launch(const serve_ring_hdr* ring, const uint32_t* ring_items, uint32_t ring_mask,
                                    const cuphyHipPuschSlotTb_t* jobs, int n_slots, serve_state* st, uint32_t* tbw,
                                    uint32_t* abandoned, int grid, cuphyDataType_t llr_type, uint64_t idle_ticks,
                                    uint64_t life_ticks, uint32_t epoch, hipStream_t strm)
{
    void* func = nullptr;
    switch (llr_type) {
    case CUPHY_R_8I: func = (void*)&pusch_serve_resident<int8_t>; break;
    case CUPHY_R_16F: func = (void*)&pusch_serve_resident<_Float16>; break;
    case CUPHY_R_32F: func = (void*)&pusch_serve_resident<float>; break;
    default: return CUPHY_STATUS_UNSUPPORTED_TYPE;
    }
    int lds_values = SERVE_LDS_VALUES;  // any TB (descrambled: E within the word budget)
    const size_t lds = ((size_t)lds_values * 2 + 15) / 16 * 16 + 64u;
    cuphyStatus_t s = ensure_lds_attr(func, lds);
    if (s != CUPHY_STATUS_SUCCESS) return s;
    void* args[12] = {const_cast<serve_ring_hdr**>(&ring), const_cast<uint32_t**>(&ring_items), &ring_mask,
                      const_cast<cuphyHipPuschSlotTb_t**>(&jobs), &n_slots, &st, &abandoned, &tbw, &lds_values,
                      &idle_ticks, &life_ticks, &epoch};
    if (hipLaunchKernel(func, dim3((unsigned)grid), dim3(FUSED_BLOCK), args, lds, strm) != hipSuccess)
        return CUPHY_STATUS_INTERNAL_ERROR;
    return CUPHY_STATUS_SUCCESS;
}

namespace {
constexpr int MAX_SCRATCH_DEVICES = 64;
std::mutex g_scratch_mutex;
uint32_t* g_scratch_ptr[MAX_SCRATCH_DEVICES] = {};  // device address of `scratch` on each GPU
}  // namespace

cuphyStatus_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= MAX_SCRATCH_DEVICES) return CUPHY_STATUS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return CUPHY_STATUS_SUCCESS;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return CUPHY_STATUS_INTERNAL_ERROR;
    if (cur != device && hipSetDevice(device) != hipSuccess) return CUPHY_STATUS_INTERNAL_ERROR;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);  // restoring the caller's device; the lookup's result decides
    if (e != hipSuccess || !ptr) return CUPHY_STATUS_INTERNAL_ERROR;
    g_scratch_ptr[device] = static_cast<uint32_t*>(ptr);
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t slot_decode_noalloc(const cuphyHipPuschSlotTb_t* d_tbs, int n_tbs, const uint32_t* d_items, int n_items,
                                  int n_cbs, cuphyDataType_t llr_type, int max_lds_values, int max_parity_nodes,
                                  int max_cbs_per_item, int cu_count, hipStream_t strm)
{
    if (n_cbs < 0 || n_cbs > NOALLOC_MAX_CBS) return CUPHY_STATUS_INVALID_ARGUMENT;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_SCRATCH_DEVICES) return CUPHY_STATUS_INTERNAL_ERROR;
    uint32_t* ws = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        ws = g_scratch_ptr[dev];
    }
    if (!ws) {  // first call on a GPU the decoder was not created on
        const cuphyStatus_t s = noalloc_scratch_init(dev);
        if (s != CUPHY_STATUS_SUCCESS) return s;
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        ws = g_scratch_ptr[dev];
    }
    // the queue state and the words of every TB start at zero
    const int words = FUSED_WS_QUEUE_WORDS + 2 * n_cbs;
    void* zargs[2] = {&ws, const_cast<int*>(&words)};
    if (hipLaunchKernel((void*)&zero_words, dim3((unsigned)((words + 255) / 256)), dim3(256), zargs, 0, strm) !=
        hipSuccess)
        return CUPHY_STATUS_INTERNAL_ERROR;
    return slot_decode(d_tbs, n_tbs, d_items, n_items, llr_type, max_lds_values, max_parity_nodes, max_cbs_per_item,
                       ws, cu_count, strm);
}

cuphyStatus_t slot_decode(const cuphyHipPuschSlotTb_t* d_tbs, int n_tbs, const uint32_t* d_items, int n_items,
                          cuphyDataType_t llr_type, int max_lds_values, int max_parity_nodes, int max_cbs_per_item,
                          void* d_ws, int cu_count, hipStream_t strm)
{
    if (n_items < 0 || n_tbs < 0 || n_tbs > 65536 || !d_ws) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (n_items > 0 && (!d_tbs || !d_items || n_tbs == 0)) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (max_lds_values <= 0 || max_lds_values > SLOT_MAX_LDS_VALUES) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (max_parity_nodes < 4 || max_parity_nodes > 46) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (max_cbs_per_item < 1 || max_cbs_per_item > FUSED_MAX_GROUP) return CUPHY_STATUS_INVALID_ARGUMENT;
    void* func = fused_kernel(llr_type, max_parity_nodes, max_cbs_per_item > 1);
    if (!func) return CUPHY_STATUS_UNSUPPORTED_TYPE;
    const size_t lds = ((size_t)max_lds_values * 2 + 15) / 16 * 16 + 64u;  // + whole 32-value groups of the hard decisions
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
