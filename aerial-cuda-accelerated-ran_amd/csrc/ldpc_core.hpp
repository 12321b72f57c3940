// Device core of the layered normalized min-sum decoder (gfx950), shared by
// the transport-block, persistent-batch and fused-slot kernels.
#pragma once

#include "nr_ldpc_common.hpp"

namespace nrldpc {

// Row capacity of the high-code-rate kernel variants (mb <= 8) and of the
// fused slot kernel's middle variant (mb <= 16).
constexpr int SMALL_MAXR = 8;
constexpr int MID_MAXR = 16;
// global loads a thread keeps in flight while loading a code block's LLRs
constexpr int LLR_LOADS = 8;
#ifndef TWO_SET_PAIRS
#define TWO_SET_PAIRS 6
#endif
#ifndef SDWA_PAIR
#define SDWA_PAIR 1
#endif

// Reduced shifts of every (BG, Z) in device memory, constant-initialised in
// the code object.  Read through an address-space-4 pointer so that the
// uniform loads become scalar (s_load) loads.
static __device__ const reduced_shift_tables g_shift_tables = make_reduced_shift_tables();
typedef const __attribute__((address_space(4))) uint32_t* shift_ptr;

// Position of every lifting size among the 51 (index Z, -1 for other
// values), as a constant table: a runtime search of a host constexpr array
// may be placed in per-lane memory, whose loads count as divergent.
struct lifting_index_table {
    int32_t v[385];
};
constexpr lifting_index_table make_lifting_index_table()
{
    lifting_index_table t{};
    for (int z = 0; z <= 384; ++z) t.v[z] = lifting_size_index(z);
    return t;
}
static __device__ const lifting_index_table g_lifting_index = make_lifting_index_table();

// Index of Z among the lifting sizes, -1 if Z is not one (uniform load).
__device__ __forceinline__ int z_index(int Z)
{
    typedef const __attribute__((address_space(4))) int32_t* cptr;
    return ((unsigned)Z <= 384u) ? ((cptr)(const void*)g_lifting_index.v)[Z] : -1;
}

template <typename T> struct num;
template <> struct num<float> {
    static __device__ __forceinline__ float vmin(float a, float b) { return __builtin_fminf(a, b); }
    static __device__ __forceinline__ float vmax(float a, float b) { return __builtin_fmaxf(a, b); }
    static __device__ __forceinline__ float vabs(float a) { return __builtin_fabsf(a); }
    static __device__ __forceinline__ float inf() { return __builtin_huge_valf(); }
};
template <> struct num<_Float16> {
    static __device__ __forceinline__ _Float16 vmin(_Float16 a, _Float16 b) { return __builtin_fminf16(a, b); }
    static __device__ __forceinline__ _Float16 vmax(_Float16 a, _Float16 b) { return __builtin_fmaxf16(a, b); }
    static __device__ __forceinline__ _Float16 vabs(_Float16 a) { return __builtin_fabsf16(a); }
    static __device__ __forceinline__ _Float16 inf() { return (_Float16)__builtin_huge_valf(); }
};

template <typename TIn, typename T> struct llr_in;
template <> struct llr_in<float, float> {
    static __device__ __forceinline__ float cvt(float v) { return v; }
};
template <> struct llr_in<float, _Float16> {
    static __device__ __forceinline__ _Float16 cvt(float v) { return (_Float16)v; }
};
template <> struct llr_in<_Float16, _Float16> {
    static __device__ __forceinline__ _Float16 cvt(_Float16 v) { return v; }
};

// Locate code word `cw` of the launch in the TB list (the reference's
// decode_desc_output_addr, ldpc2_dec_output.cuh:58-81).
struct cw_location {
    int tb, idx;
};
__device__ __forceinline__ cw_location locate_cw(const cuphyLDPCDecodeDesc_t& d, int cw)
{
    int tb = 0;
    for (; tb < d.num_tbs - 1; ++tb) {
        if (cw < d.llr_input[tb].num_codewords) break;
        cw -= d.llr_input[tb].num_codewords;
    }
    return {tb, cw};
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef uint16_t u16x2_t __attribute__((ext_vector_type(2)));

template <int BG> constexpr uint64_t kBarrierMask = barrier_mask<bg_traits<BG>>();

// Row-pair plan (two lane groups per code block): the rows between two
// barriers share no column, so two lane groups may process them at the same
// time.  Each row goes to one group: the rows of a barrier segment
// alternate between the groups, and a segment's first row goes to the
// group holding fewer edges so far (the per-lane message state is then
// about half the graph's).  Within its group a row's binary16 messages take
// consecutive slots from pair_slot[r] (rows padded to even degree) and its
// compressed binary32 row state the slot row_slot[r].  Rows r < NR only.
template <int BG, int NR> struct pair_plan {
    int group[64];
    int row_slot[64];
    int pair_slot[64];  // binary16: first message slot of the row (even; rows padded to even degree)
    int edges[2];
    int rows[2];
    int padded_edges[2];
};
template <int BG, int NR> constexpr pair_plan<BG, NR> make_pair_plan()
{
    using BGT = bg_traits<BG>;
    pair_plan<BG, NR> p{};
    int prev = 1;
    for (int r = 0; r < NR; ++r) {
        const bool first = (r == 0) || ((kBarrierMask<BG> >> r) & 1ull);
        const int g = first ? (p.edges[1] < p.edges[0] ? 1 : 0) : 1 - prev;
        p.group[r] = g;
        p.row_slot[r] = p.rows[g];
        p.pair_slot[r] = p.padded_edges[g];
        p.edges[g] += BGT::row_start(r + 1) - BGT::row_start(r);
        p.padded_edges[g] += (BGT::row_start(r + 1) - BGT::row_start(r) + 1) & ~1;
        p.rows[g] += 1;
        prev = g;
    }
    return p;
}
template <int BG, int NR> constexpr pair_plan<BG, NR> kPairPlan = make_pair_plan<BG, NR>();

// binary16 message slots: row r's edges from padded_row_start(r) on, every
// row's degree rounded up to even, so that each pair of a row's edges is one
// 32-bit register of packed halves.
template <int BG> constexpr int padded_row_start(int r)
{
    int s = 0;
    for (int i = 0; i < r; ++i) s += (bg_traits<BG>::row_start(i + 1) - bg_traits<BG>::row_start(i) + 1) & ~1;
    return s;
}

__device__ __forceinline__ uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }
__device__ __forceinline__ uint16_t h2u(_Float16 x) { return __builtin_bit_cast(uint16_t, x); }
__device__ __forceinline__ _Float16 u2h(uint16_t x) { return __builtin_bit_cast(_Float16, x); }

// (z + s) mod Z for z, s < Z: the smaller of p and p - Z as unsigned.
__device__ __forceinline__ uint32_t rot_index(uint32_t z, uint32_t s, uint32_t Z)
{
    const uint32_t p = z + s;
    const uint32_t q = p - Z;
    return q < p ? q : p;
}

// Value at byte offset `off` of a lane's LDS region.
template <typename T>
__device__ __forceinline__ T& lds_at(unsigned char* base, uint32_t off)
{
    return *reinterpret_cast<T*>(base + off);
}

// LDS addressed from LDS address 0 (row-pair kernels, whose APP values start
// there: they hold no static LDS, which the host checks before a launch).  A
// known base lets an edge's constant column offset go into the LDS
// instruction's offset field.
typedef __attribute__((address_space(3))) unsigned char lds_byte;
__device__ __forceinline__ lds_byte* lds_origin() { return (lds_byte*)(uintptr_t)0u; }
template <typename T> __device__ __forceinline__ T lds_ld(unsigned char* b, uint32_t off) { return lds_at<T>(b, off); }
template <typename T> __device__ __forceinline__ void lds_st(unsigned char* b, uint32_t off, T v) { lds_at<T>(b, off) = v; }
template <typename T> __device__ __forceinline__ T lds_ld(lds_byte* b, uint32_t off)
{
    return *reinterpret_cast<const __attribute__((address_space(3))) T*>(b + off);
}
template <typename T> __device__ __forceinline__ void lds_st(lds_byte* b, uint32_t off, T v)
{
    *reinterpret_cast<__attribute__((address_space(3))) T*>(b + off) = v;
}

// Lane-major APP layout (row-pair kernels): lane z's values of every column
// at z * pitch + column * sizeof(T), the pitch an odd number of 4-byte words
// so that 64 consecutive lanes touch 64 distinct LDS banks.  The column term
// is then a constant, which the LDS instruction takes as its offset.
__host__ __device__ constexpr uint32_t lm_pitch_bytes(int ncol, int elem)
{
    return ((((uint32_t)(ncol * elem) + 3u) >> 2) | 1u) << 2;
}

// Split lane-major layout (binary16 kernels holding more than SMALL_MAXR
// block rows): columns [0, S) in a lane-major array at LDS address 0 and
// columns [S, ncol) in a second one at LM_HALF, both with the pitch of the
// larger part, S = 34 (BG1) / 26 (BG2).  Z * pitch then stays at most 26,112
// bytes, so every rotated offset fits 16 bits (packed rotation), and an
// edge's column term -- 2c, or LM_HALF + 2(c - S) -- is still a constant
// that the LDS instruction takes as its offset.
constexpr uint32_t LM_HALF = 32768;
__host__ __device__ constexpr int lm_split_col(int BG) { return BG == 1 ? 34 : 26; }
__host__ __device__ constexpr bool lm_split_layout(int elem, int maxr) { return elem == 2 && maxr > SMALL_MAXR; }
__host__ __device__ constexpr uint32_t lm_pitch(int BG, int ncol, int elem, bool split)
{
    if (!split) return lm_pitch_bytes(ncol, elem);
    const int S = lm_split_col(BG);
    const int n0 = ncol < S ? ncol : S, n1 = ncol - S;
    return lm_pitch_bytes(n0 > n1 ? n0 : n1, elem);
}
// Byte offset of column c in a lane's row (add z * pitch for lane z).
__host__ __device__ constexpr uint32_t lm_col_offset(int BG, int c, int elem, bool split)
{
    return (split && c >= lm_split_col(BG)) ? LM_HALF + (uint32_t)(c - lm_split_col(BG)) * (uint32_t)elem
                                            : (uint32_t)c * (uint32_t)elem;
}
// LDS bytes of a code block's lane-major APP values.
__host__ __device__ constexpr uint32_t lm_app_bytes(int BG, int Z, int ncol, int elem, bool split)
{
    return (split && ncol > lm_split_col(BG)) ? LM_HALF + (uint32_t)Z * lm_pitch(BG, ncol, elem, true)
                                              : (uint32_t)Z * lm_pitch(BG, ncol, elem, split);
}
// Largest of those over Z <= 384 and every mb (launch sizing).
__host__ __device__ constexpr uint32_t lm_app_bytes_max(int elem, bool split)
{
    return split ? LM_HALF + 384u * lm_pitch_bytes(34, elem) : 384u * lm_pitch_bytes(68, elem);
}

// ---------------------------------------------------------------------------
// Row addressing: lane z of a code block whose APP values start at byte
// `base` reads APP[col][(z + s) mod Z] at base + col*Zb + rot_index(zb, sb, Zb),
// with zb, sb, Zb the lane, shift and lifting size in bytes (x sizeof(T)):
// add, subtract, min and one three-operand add per edge.
// ---------------------------------------------------------------------------

// Packed rotation (PK): when every rotated offset is below 2^16 (2 * Zb <=
// 65536), the lane keeps zb and zb - Zb as the two 16-bit halves of one
// register (lane_rot_base); adding the shift to both halves and one 16-bit
// unsigned min of the halves give rot_index: two instructions instead of
// three.  (zb - Zb wraps to at least 65536 - Zb > any offset.)  The add is
// a plain 32-bit add of sb * 0x10001 (two-operand encoding, twice the issue
// rate of a packed add): the low half zb + sb < 2 * Zb <= 65536 never
// carries into the high half, and the high half's carry leaves bit 31.
__device__ __forceinline__ uint32_t lane_rot_base(uint32_t zb, uint32_t Zb)
{
    return (zb & 0xFFFFu) | (((zb - Zb) & 0xFFFFu) << 16);
}
__device__ __forceinline__ uint32_t rot_index_packed(uint32_t zz, uint32_t sbb)
{
    const uint32_t pq = zz + sbb;  // sbb = sb * 0x10001
    return (uint32_t)__builtin_elementwise_min((uint16_t)pq, (uint16_t)(pq >> 16));
}

// Byte offset of edge E's APP value for the lane.  Column-major (LM false):
// zb, Zb = lane and lifting size x sizeof(T).  Lane-major (LM true): zb, Zb
// = lane and lifting size x pitch Pb.  PK: zb is lane_rot_base(zp, Zb), zp
// the plain lane offset (otherwise zp == zb).  An edge whose shift is 0 in
// every lifting set (the identity circulants of the parity columns) needs
// no rotation.
template <int BG, int E, typename T, bool LM, bool PK = false, bool SPL = false>
__device__ __forceinline__ uint32_t edge_off(shift_ptr sh, uint32_t zb, uint32_t zp, uint32_t Zb, uint32_t Pb)
{
    using BGT = bg_traits<BG>;
    constexpr bool ID = BGT::identity_edge(E);
    const uint32_t sb = ID ? 0u : sh[E] * ((LM ? Pb : (uint32_t)sizeof(T)) * (PK ? 0x10001u : 1u));
    const uint32_t rot = ID ? zp : PK ? rot_index_packed(zb, sb) : rot_index(zb, sb, Zb);
    if constexpr (LM)
        return rot + lm_col_offset(BG, BGT::col(E), (int)sizeof(T), SPL);
    else
        return (uint32_t)BGT::col(E) * Zb + rot;
}

// ---------------------------------------------------------------------------
// binary32 rows: compressed C2V per (row, lane) -- normalised min1/min2 and a
// word holding the argmin (bits 27..31) and the C2V sign of every edge
// (edge j at bit D-1-j).  Per edge: address 3, previous-message rebuild 4,
// subtract 1, min/min2/parity 3, new message + write-back 6 VALU ops.
// ---------------------------------------------------------------------------
// RS: the row's state slot (R, or its slot within a row-pair lane group).
template <int BG, int R, int NR, int RS = R, bool LM = false, typename BP = unsigned char*, bool PK = false,
          bool SPL = false>
__device__ __forceinline__ void process_row(float*, BP __restrict__ base, shift_ptr sh, uint32_t zb,
                                            uint32_t zp, uint32_t Zb, float alpha, float (&st_n1)[NR], float (&st_n2)[NR],
                                            uint32_t (&st_meta)[NR], uint32_t Pb = 0)
{
    using BGT = bg_traits<BG>;
    constexpr int E0 = BGT::row_start(R);
    constexpr int D = BGT::row_start(R + 1) - E0;
    float t[D];
    uint32_t off[D];
    const float on1 = st_n1[RS], on2 = st_n2[RS];
    const uint32_t om = st_meta[RS];
    const uint32_t oidx = om >> 27;
    static_for<D>([&](auto J) {
        constexpr int j = decltype(J)::value;
        off[j] = edge_off<BG, E0 + j, float, LM, PK, SPL>(sh, zb, zp, Zb, Pb);
        const float mag = (oidx == (uint32_t)j) ? on2 : on1;
        const float old = u2f(f2u(mag) | ((om << (32 - D + j)) & 0x80000000u));
        t[j] = lds_ld<float>(base, off[j]) - old;
    });
    float m1 = __builtin_huge_valf(), m2 = __builtin_huge_valf();
    uint32_t par = 0;
    static_for<D>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const float a = __builtin_fabsf(t[j]);
        // min(m2, max(m1, a)) (= med3(m1, a, m2) as m1 <= m2, but as two
        // full-rate operations: v_med3 issues at a quarter of v_min's rate)
        m2 = __builtin_fminf(m2, __builtin_fmaxf(m1, a));
        m1 = __builtin_fminf(m1, a);
        par ^= f2u(t[j]);
    });
    par &= 0x80000000u;
    const float a1 = alpha * m1, a2 = alpha * m2;
    uint32_t idx = 0, fs = 0;
    static_for<D>([&](auto J) {
        constexpr int j = decltype(J)::value;
        // |t| == min1 marks the argmin; on a tie min2 == min1, so every tied
        // edge gets the same value as the first-index rule gives it
        const bool is_min = __builtin_fabsf(t[j]) == m1;
        const float mag = is_min ? a2 : a1;
        idx = is_min ? (uint32_t)j : idx;
        const uint32_t cb = f2u(mag) | ((f2u(t[j]) ^ par) & 0x80000000u);
        lds_st<float>(base, off[j], t[j] + u2f(cb));
        fs = __builtin_amdgcn_alignbit(fs, cb, 31);  // (fs << 1) | sign(c_j)
    });
    st_n1[RS] = a1;
    st_n2[RS] = a2;
    st_meta[RS] = fs | (idx << 27);
}

// ---------------------------------------------------------------------------
// binary16 rows, two edges per 32-bit register (packed halves): the
// normalised C2V message of every edge kept explicitly (edges 2k, 2k+1 of a
// row in one register; a row of odd degree ends with half a pair whose
// upper half is a dummy edge at +inf).  Per pair of edges:
//   t   = APP - C2V                                     (packed subtract)
//   min1/min2 running over |t|: min2 = min(min2, max(min1, |t|)),
//   min1 = min(min1, |t|), two independent sets of accumulators on long rows
//   parity ^= t (sign bits)
// then per row the two halves (and accumulator sets) are merged, both halves
// of min1/min2 holding the row's values, and per pair
//   g     = min(|t|, min2)            (= min1 exactly for the argmin edge,
//                                      min2 for every other edge)
//   other = (min1 + min2) - g         (16-bit integer arithmetic on the bit
//                                      patterns: the magnitude not taken)
//   C2V   = other * (+-alpha)         (alpha's sign = sign(t) ^ parity)
//   APP   = t + C2V
// which is the extrinsic min-sum update with the first-index rule on ties
// (tied minima give min1 == min2) and the same binary16 roundings as
// alpha * min1 / alpha * min2 with the sign attached: the product's
// magnitude does not depend on the sign under round-to-nearest.
// No argmin tracking, no compare/select per edge.
// ---------------------------------------------------------------------------
__device__ __forceinline__ half2_t u2h2(uint32_t x) { return __builtin_bit_cast(half2_t, x); }
__device__ __forceinline__ uint32_t h22u(half2_t x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ u16x2_t u2m(uint32_t x) { return __builtin_bit_cast(u16x2_t, x); }
// Magnitudes (|t|, min1, min2: never NaN, never negative) compared as their
// 16-bit patterns, which order like the values: integer min/max need no
// NaN canonicalisation of their operands and take swapped halves directly.
__device__ __forceinline__ u16x2_t umin2(u16x2_t a, u16x2_t b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ u16x2_t umax2(u16x2_t a, u16x2_t b) { return __builtin_elementwise_max(a, b); }

// (v0 - m.lo, v1 - m.hi) in binary16, v0 and v1 two zero-extended 16-bit
// LDS loads, without packing them first: one two-operand subtract into the
// low half, one into the high half (SDWA destination select, low half
// preserved) -- two full-rate instructions in place of a three-operand
// pack and a packed subtract.  x - y is x + (-y) in IEEE arithmetic, so the
// halves equal those of the packed subtract.  The wait state: a VALU result
// written through a destination select must not be read by the next VALU
// instruction.
__device__ __forceinline__ half2_t sub_pair(uint32_t v0, uint32_t v1, half2_t m)
{
    uint32_t r;
    asm("v_sub_f16_e32 %0, %1, %2\n\t"
        "v_sub_f16_sdwa %0, %3, %2 dst_sel:WORD_1 dst_unused:UNUSED_PRESERVE src0_sel:WORD_0 src1_sel:WORD_1\n\t"
        "s_nop 0"
        : "=&v"(r)
        : "v"(v0), "v"(h22u(m)), "v"(v1));
    return u2h2(r);
}

// S0: message slot of the row's first edge (even: padded_row_start(R), or
// the row's pair_slot within a row-pair lane group).
template <int BG, int R, int NW, int S0 = padded_row_start<BG>(R), bool LM = false, typename BP = unsigned char*,
          bool PK = false, bool SPL = false>
__device__ __forceinline__ void process_row(_Float16*, BP __restrict__ base, shift_ptr sh, uint32_t zb,
                                            uint32_t zp, uint32_t Zb, _Float16 alpha, half2_t (&st)[NW], uint32_t Pb = 0)
{
    using BGT = bg_traits<BG>;
    constexpr int E0 = BGT::row_start(R);
    constexpr int D = BGT::row_start(R + 1) - E0;
    constexpr int P = (D + 1) / 2;
    static_assert(S0 % 2 == 0, "binary16 message slots of a row start at an even slot");
    constexpr int W0 = S0 / 2;
    constexpr bool TWO = P >= TWO_SET_PAIRS;  // two accumulator sets on long rows: shorter dependency chains
    constexpr int NSET = TWO ? 2 : 1;
    constexpr uint32_t SIGNS = 0x80008000u, MAGS = 0x7FFF7FFFu;
    uint32_t off[D];
    half2_t t[P];
    u16x2_t mag[P];
    u16x2_t m1a, m2a, m1b, m2b;
    uint32_t par = 0;
    static_for<P>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int j0 = 2 * k, j1 = 2 * k + 1;
        off[j0] = edge_off<BG, E0 + j0, _Float16, LM, PK, SPL>(sh, zb, zp, Zb, Pb);
        const uint32_t v0 = lds_ld<uint16_t>(base, off[j0]);
        if constexpr (j1 < D) {
            off[j1] = edge_off<BG, E0 + j1, _Float16, LM, PK, SPL>(sh, zb, zp, Zb, Pb);
            const uint32_t v1 = lds_ld<uint16_t>(base, off[j1]);
            if constexpr (SDWA_PAIR)
                t[k] = sub_pair(v0, v1, st[W0 + k]);
            else
                t[k] = u2h2(v0 | (v1 << 16)) - st[W0 + k];
        } else {
            t[k] = u2h2(v0 | 0x7C000000u) - st[W0 + k];  // the dummy edge: +inf, sign bit clear
        }
        const u16x2_t a = u2m(h22u(t[k]) & MAGS);
        mag[k] = a;
        if (k < NSET) {  // a set's first pair: min1 = |t|
            if (TWO && (k & 1))
                m1b = a;
            else
                m1a = a;
        } else if (k < 2 * NSET) {  // its second: min2 = max(min1, |t|)
            if (TWO && (k & 1)) {
                m2b = umax2(m1b, a);
                m1b = umin2(m1b, a);
            } else {
                m2a = umax2(m1a, a);
                m1a = umin2(m1a, a);
            }
        } else if (TWO && (k & 1)) {
            m2b = umin2(m2b, umax2(m1b, a));
            m1b = umin2(m1b, a);
        } else {
            m2a = umin2(m2a, umax2(m1a, a));
            m1a = umin2(m1a, a);
        }
        par ^= h22u(t[k]);
    });
    if constexpr (TWO) {
        const u16x2_t m1 = umin2(m1a, m1b);
        m2a = umin2(umin2(m2a, m2b), umax2(m1a, m1b));
        m1a = m1;
    }
    // both halves: the row's min1 / min2, and the parity in bits 15 and 31
    const u16x2_t m1s = m1a.yx, m2s = m2a.yx;
    const u16x2_t M1 = umin2(m1a, m1s);
    const u16x2_t M2 = umin2(umin2(m2a, m2s), umax2(m1a, m1s));
    const uint32_t px = par ^ __builtin_amdgcn_alignbit(par, par, 16);
    const uint32_t Q = h22u(half2_t{alpha, alpha}) ^ (px & SIGNS);
    const u16x2_t S = M1 + M2;
    static_for<P>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int j0 = 2 * k, j1 = 2 * k + 1;
        const u16x2_t g = umin2(mag[k], M2);
        const half2_t other = __builtin_bit_cast(half2_t, (u16x2_t)(S - g));
        // Q ^ (t & SIGNS) as one three-input bitwise instruction (table 0x6C: B ^ (A & C))
        const half2_t sa = u2h2(__builtin_amdgcn_bitop3_b32(h22u(t[k]), Q, SIGNS, (unsigned char)0x6C));
        const half2_t c = other * sa;
        const uint32_t nv = h22u(t[k] + c);
        lds_st<uint16_t>(base, off[j0], (uint16_t)nv);
        if constexpr (j1 < D) lds_st<uint16_t>(base, off[j1], (uint16_t)(nv >> 16));
        st[W0 + k] = c;
    });
}

// Parity of check row R, lane z: hard bit = APP < 0 (syndrome_check.m:38-40).
template <int BG, int R, typename T, bool LM = false, typename BP = unsigned char*, bool PK = false,
          bool SPL = false>
__device__ __forceinline__ uint32_t row_syndrome(BP __restrict__ base, shift_ptr sh, uint32_t zb, uint32_t zp,
                                                 uint32_t Zb, uint32_t Pb = 0)
{
    using BGT = bg_traits<BG>;
    constexpr int E0 = BGT::row_start(R);
    constexpr int D = BGT::row_start(R + 1) - E0;
    uint32_t par = 0;
    static_for<D>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const uint32_t off = edge_off<BG, E0 + j, T, LM, PK, SPL>(sh, zb, zp, Zb, Pb);
        if constexpr (std::is_same<T, _Float16>::value)
            par ^= (uint32_t)lds_ld<uint16_t>(base, off);  // sign bits (APP values are never -0)
        else
            par ^= (uint32_t)(lds_ld<T>(base, off) < (T)0);
    });
    if constexpr (std::is_same<T, _Float16>::value) par >>= 15;
    return par;
}

// Everything one code block's decode needs, resolved from a TB descriptor
// (transport-block kernel) or from a batch work item (persistent kernel).
struct cb_job {
    const void* llr;
    uint32_t* out;
    _Float16* soft;  // nullptr: no soft output
    int Z, mb, Kb, max_iter;
    bool early_term;
    float clamp;
    float alpha_f32;
    uint16_t alpha_f16;
};

// Phase 1 of a code block's decode: LLRs (global, type TIn) -> LDS `app`
// (type T), converted, clamped to +-clamp (clamp > 0) and with a signed zero
// made +0.  Every thread of the workgroup calls it; ends with a barrier.
template <typename T, typename TIn>
__device__ __forceinline__ void load_llrs_nosync(T* __restrict__ app, const TIn* __restrict__ src, int nval,
                                                 float clamp)
{
    const T clampv = (T)clamp;
    const bool do_clamp = clamp > 0.0f;
    const int step = (int)blockDim.x;
    // LLR_LOADS loads in flight per thread (an index past the end reloads
    // the thread's first value, whose store is skipped)
    for (int n0 = threadIdx.x; n0 < nval; n0 += LLR_LOADS * step) {
        TIn raw[LLR_LOADS];
#pragma unroll
        for (int u = 0; u < LLR_LOADS; ++u) {
            const int n = n0 + u * step;
            raw[u] = src[n < nval ? n : n0];
        }
#pragma unroll
        for (int u = 0; u < LLR_LOADS; ++u) {
            const int n = n0 + u * step;
            if (n >= nval) break;
            T v = llr_in<TIn, T>::cvt(raw[u]);
            if (do_clamp) v = num<T>::vmax(num<T>::vmin(v, clampv), -clampv);
            app[n] = (v == (T)0) ? (T)0 : v;
        }
    }
}
template <typename T, typename TIn>
__device__ __forceinline__ void load_llrs(T* __restrict__ app, const TIn* __restrict__ src, int nval, float clamp)
{
    load_llrs_nosync<T, TIn>(app, src, nval, clamp);
    __syncthreads();
}

// Values between the APP regions of consecutive code blocks of a group
// (whole 16-byte units).
__device__ __forceinline__ int group_stride(int nval) { return (nval + 7) & ~7; }

template <typename T> __device__ __forceinline__ T job_alpha(const cb_job& job);
template <> __device__ __forceinline__ float job_alpha<float>(const cb_job& job) { return job.alpha_f32; }
template <> __device__ __forceinline__ _Float16 job_alpha<_Float16>(const cb_job& job)
{
    return __builtin_bit_cast(_Float16, job.alpha_f16);
}

// Phase 2: the layered iterations over the APP values in LDS (already
// loaded).  Returns the number of iterations run (< max_iter only with
// early termination; the largest over the group).  Every thread of the
// workgroup calls it.
// MAXR: rows the register state is sized for (mb <= MAXR); a small MAXR
// (high code rates) needs far fewer VGPRs and so allows more resident waves.
// Groups: `count` code blocks of the same (BG, Z, mb, iterations) decoded
// together, lanes [c*Z, (c+1)*Z) on code block c, whose APP values start at
// app + c*stride.  With early termination every code block stops on its own
// syndrome (its lanes freeze, exactly as if decoded alone); `cb_flags` is
// LDS scratch of `count` words for that (unused when count == 1).
// cb_iters (optional, LDS, `count` words): iterations each code block ran.
// PAIR = 2: row-pair mode for one code block (count must be 1, blockDim.x
// >= 2 * roundup64(Z)): lanes [0, Z) and [Zw, Zw + Z) (Zw = roundup64(Z),
// wave-aligned) form two lane groups that process the rows of kPairPlan's
// group 0 and group 1; rows of one barrier segment run at the same time,
// and each lane holds the message state of its group's rows only.
// LM: lane-major APP values from LDS address 0 (lm_pitch_bytes; count must
// be 1, cb_flags two LDS words for the early-termination votes).
template <int BG, typename T, int MAXR, int PAIR = 1, bool LM = (PAIR == 2)>
__device__ __forceinline__ int run_iterations(T* __restrict__ app, const cb_job& job, int count = 1, int stride = 0,
                                              uint32_t* cb_flags = nullptr, uint32_t* cb_iters = nullptr)
{
    using BGT = bg_traits<BG>;
    constexpr int NR = (MAXR < BGT::NROWS) ? MAXR : BGT::NROWS;
    constexpr pair_plan<BG, NR> PP = kPairPlan<BG, NR>;
    const uint32_t Z = (uint32_t)job.Z;
    const int mb = job.mb;
    const uint32_t tid = threadIdx.x;
    const T alpha = job_alpha<T>(job);

    // C2V state: compressed per row (binary32) or explicit per edge (binary16)
    constexpr int NRS = (PAIR == 2) ? (PP.rows[0] > PP.rows[1] ? PP.rows[0] : PP.rows[1]) : NR;
    // binary16 message slots: every row padded to an even number of edges
    constexpr int NEDGE = (PAIR == 2) ? (PP.padded_edges[0] > PP.padded_edges[1] ? PP.padded_edges[0]
                                                                                  : PP.padded_edges[1])
                                      : padded_row_start<BG>(NR);
    constexpr int NW = NEDGE / 2;
    float st_n1[std::is_same<T, float>::value ? NRS : 1], st_n2[std::is_same<T, float>::value ? NRS : 1];
    uint32_t st_meta[std::is_same<T, float>::value ? NRS : 1];
    half2_t st_h[std::is_same<T, float>::value ? 1 : NW];
    if constexpr (std::is_same<T, float>::value) {
        static_for<NRS>([&](auto R) {
            st_n1[R] = 0.0f;
            st_n2[R] = 0.0f;
            st_meta[R] = 0;
        });
    } else {
        static_for<NW>([&](auto W) { st_h[W] = half2_t{(_Float16)0, (_Float16)0}; });
    }

    // lane -> (code block, check row); row-pair mode: lane -> (lane group, check row)
    uint32_t cb, zlane, lgrp = 0;
    bool in_group;
    if constexpr (PAIR == 2) {
        const uint32_t Zw = (Z + 63u) & ~63u;
        lgrp = tid >= Zw ? 1u : 0u;
        zlane = tid - lgrp * Zw;
        cb = 0;
        in_group = zlane < Z && tid < 2u * Zw;
        count = 1;
    } else {
        cb = (count == 1) ? (tid < Z ? 0u : 1u) : tid / Z;
        zlane = tid - (cb < (uint32_t)count ? cb : 0u) * Z;
        in_group = cb < (uint32_t)count;
    }
    bool active = in_group;
    unsigned char* const cm_base =
        reinterpret_cast<unsigned char*>(app) + (size_t)(in_group ? cb : 0u) * (size_t)stride * sizeof(T);
    // lane-major: the APP values start at LDS address 0 (app == LDS origin)
    const auto lane_base = [&]() {
        if constexpr (LM)
            return lds_origin();
        else
            return cm_base;
    }();
    // lane-major binary16 with more than SMALL_MAXR rows: split layout
    constexpr bool SPL = LM && lm_split_layout((int)sizeof(T), NR);
    const uint32_t Pb = LM ? lm_pitch(BG, BGT::NSYS + mb, (int)sizeof(T), SPL) : (uint32_t)sizeof(T);
    const uint32_t Zb = Z * Pb;
    // packed 16-bit rotations where 2 * Zb <= 65536 for every Z <= 384:
    // column-major (Zb = Z * sizeof(T)), lane-major binary16 with at most 8
    // block rows (pitch <= 60 bytes) and the split layout (pitch <= 68 bytes)
    constexpr bool PK = !LM || sizeof(T) == 2;

    shift_ptr sh = (shift_ptr)(const void*)&g_shift_tables.v[BG - 1][z_index((int)Z)][0];
    int it = 0;
    while (it < job.max_iter) {
        ++it;
        if (cb_iters && active && zlane == 0u && lgrp == 0u) cb_iters[cb] = (uint32_t)it;
        // Opaque copies of the lane offset and of the shift pointer, renewed
        // per row: the 316 (BG1) per-edge shifts and addresses are reloaded
        // and recomputed where used instead of being hoisted out of the
        // loops into (spilled) registers.
        uint32_t zp = zlane * Pb;
        uint32_t zb = PK ? lane_rot_base(zp, Zb) : zp;
        asm volatile("" : "+v"(zb), "+v"(zp));
        static_for<NR>([&](auto Rc) {
            constexpr int R = decltype(Rc)::value;
            // opaque row count: the R < mb tests stay scalar compares at
            // each row instead of 46 hoisted (and spilled) lane masks
            int mbr = mb;
            asm volatile("" : "+s"(mbr));
            if (R < mbr) {
                shift_ptr shr = sh;
                uint32_t Zr = Zb;
                asm volatile("" : "+s"(shr), "+s"(Zr));
                if (active && (PAIR == 1 || lgrp == (uint32_t)PP.group[R])) {
                    constexpr int RS = (PAIR == 2) ? PP.row_slot[R] : R;
                    constexpr int S0 = (PAIR == 2) ? PP.pair_slot[R] : padded_row_start<BG>(R);
                    if constexpr (std::is_same<T, float>::value)
                        process_row<BG, R, NRS, RS, LM, decltype(lane_base), PK, SPL>(app, lane_base, shr, zb, zp, Zr,
                                                                                      alpha, st_n1, st_n2, st_meta,
                                                                                      Pb);
                    else
                        process_row<BG, R, NW, S0, LM, decltype(lane_base), PK, SPL>(app, lane_base, shr, zb, zp, Zr,
                                                                                     alpha, st_h, Pb);
                }
                // barrier before the next row only where it shares a column
                // with the rows since the last barrier, and at the iteration end
                constexpr bool next_needs = (R + 1 >= NR) || ((kBarrierMask<BG> >> (R + 1)) & 1ull);
                if (next_needs || R + 1 == mbr) __syncthreads();
            }
        });
        if (job.early_term) {
            // parity of block rows [r0, r1) over this lane's check rows
            const auto syndrome = [&](auto r0c, auto r1c) {
                constexpr int r0 = decltype(r0c)::value, r1 = decltype(r1c)::value;
                uint32_t fail = 0;
                static_for<NR>([&](auto Rc) {
                    constexpr int R = decltype(Rc)::value;
                    if constexpr (R >= r0 && R < r1) {
                        int mbr = mb;
                        asm volatile("" : "+s"(mbr));
                        if (R < mbr) {
                            shift_ptr shr = sh;
                            uint32_t Zr = Zb;
                            asm volatile("" : "+s"(shr), "+s"(Zr));
                            if (active && (PAIR == 1 || lgrp == (uint32_t)PP.group[R]))
                                fail |= row_syndrome<BG, R, T, LM, decltype(lane_base), PK, SPL>(lane_base, shr, zb,
                                                                                                 zp, Zr, Pb);
                        }
                    }
                });
                return fail;
            };
            // OR of x over the workgroup (lane-major kernels: through the LDS
            // word cb_flags[w], w = 0 / 1 for the two votes of an iteration;
            // they hold no static LDS, so that their APP values start at LDS
            // address 0)
            const auto vote = [&](uint32_t x, int w) -> bool {
                if constexpr (LM) {
                    if (tid == 0u) cb_flags[w] = 0u;
                    __syncthreads();
                    if (x) cb_flags[w] = 1u;
                    __syncthreads();
                    return __builtin_amdgcn_readfirstlane(cb_flags[w]) != 0u;
                } else {
                    return __syncthreads_or((int)x) != 0;
                }
            };
            if (LM || count == 1) {
                // block row 0 first: a code block that has not converged
                // almost always fails it, so the other rows are checked only
                // when it passes (the stopping rule is unchanged: stop when
                // every row passes)
                if (vote(syndrome(std::integral_constant<int, 0>{}, std::integral_constant<int, 1>{}), 0)) continue;
                if (!vote(syndrome(std::integral_constant<int, 1>{}, std::integral_constant<int, NR>{}), 1)) break;
            } else {
                const uint32_t fail = syndrome(std::integral_constant<int, 0>{}, std::integral_constant<int, NR>{});
                // per code block: any failing check row keeps it iterating
                if (tid < (uint32_t)count) cb_flags[tid] = 0u;
                __syncthreads();
                if (active && fail) cb_flags[cb] = 1u;
                __syncthreads();
                if (active && cb_flags[cb] == 0u) active = false;
                if (!__syncthreads_or((int)active)) break;
            }
        }
    }
    return it;
}

// Phase 3: hard decisions of the first nbits APP values (APP <= 0 -> 1),
// 64 per wave per step by ballot, into `out` (global, LSB-first words).
// Used by the decoder kernels: it needs no per-bit constants, so a
// persistent kernel keeps its register count.
template <typename T>
__device__ __forceinline__ void store_hard(const T* __restrict__ app, int nbits, uint32_t* __restrict__ out)
{
    const int nwords = (nbits + 31) >> 5;
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int nwaves = (int)((blockDim.x + 63u) >> 6);
    for (int c = wave; c * 64 < nbits; c += nwaves) {
        const int n = c * 64 + lane;
        const bool bit = (n < nbits) && (app[n] <= (T)0);
        const uint64_t m = __ballot(bit);
        if (lane == 0) {
            out[2 * c] = (uint32_t)m;
            if (2 * c + 1 < nwords) out[2 * c + 1] = (uint32_t)(m >> 32);
        }
    }
}

// Hard decisions as LSB-first 32-bit words into `out` (global) and into
// `lds_words` (LDS), ending with a barrier.  Thread w builds word w from its
// 32 APP values with 16-byte LDS reads (the buffer must hold whole groups of
// 32 values); bits past nbits are 0.  Used by the fused slot kernel, whose
// CRC pass reads the words back from LDS.
template <typename T>
__device__ __forceinline__ void store_hard_words(const T* __restrict__ app, int nbits, uint32_t* __restrict__ out,
                                                 uint32_t* __restrict__ lds_words)
{
    typedef T vec_t __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr int PER = 16 / (int)sizeof(T);  // values per 16-byte read
    const int nwords = (nbits + 31) >> 5;
    // opaque copy of the thread index: keeps the per-thread addresses from
    // being hoisted out of a persistent kernel's item loop into registers
    // held across the decode
    int w0 = (int)threadIdx.x;
    asm volatile("" : "+v"(w0));
    for (int w = w0; w < nwords; w += blockDim.x) {
        const vec_t* src = reinterpret_cast<const vec_t*>(app + 32 * w);
        uint32_t word = 0;
#pragma unroll
        for (int q = 0; q < 32 / PER; ++q) {
            const vec_t v = src[q];
#pragma unroll
            for (int e = 0; e < PER; ++e) word |= (uint32_t)(v[e] <= (T)0) << (q * PER + e);
        }
        const int tail = nbits - 32 * w;  // valid bits of this word
        if (tail < 32) word &= (1u << tail) - 1u;
        out[w] = word;
        lds_words[w] = word;
    }
    __syncthreads();
}

// Buffers of one code block of a group.
struct cb_buffers {
    const void* llr;
    uint32_t* out;
    _Float16* soft;  // nullptr: no soft output
};

// Decode `count` code blocks of one (BG, Z, mb) configuration with the whole
// workgroup (count * Z <= blockDim.x); LDS `app` holds count regions of
// group_stride((NSYS + mb) * Z) values of type T.  buf(c) gives code block
// c's buffers.  Every thread of the workgroup calls it.
template <int BG, typename T, typename TIn, int MAXR, typename Buf>
__device__ __forceinline__ int decode_group(T* __restrict__ app, const cb_job& job, int count, const Buf& buf,
                                            uint32_t* cb_flags)
{
    const int nval = (bg_traits<BG>::NSYS + job.mb) * job.Z;
    const int stride = group_stride(nval);
    for (int c = 0; c < count; ++c)
        load_llrs_nosync<T, TIn>(app + c * stride, static_cast<const TIn*>(buf(c).llr), nval, job.clamp);
    __syncthreads();
    const int iters = run_iterations<BG, T, MAXR>(app, job, count, stride, cb_flags);
    const int nbits = job.Kb * job.Z;
    for (int c = 0; c < count; ++c) {
        const cb_buffers b = buf(c);
        store_hard<T>(app + c * stride, nbits, b.out);
        // soft outputs: Kb*Z binary16 APP values
        if (b.soft)
            for (int n = threadIdx.x; n < nbits; n += blockDim.x) b.soft[n] = (_Float16)app[c * stride + n];
    }
    return iters;
}

// One code block in row-pair mode (run_iterations PAIR = 2; blockDim.x >=
// 2 * roundup64(Z)): LLRs into LDS, iterations, hard (and soft) outputs.
// One code block with lane-major APP values (lm_pitch_bytes; LDS of Z *
// pitch bytes from LDS address 0): value n = column * Z + z of the code word
// at z * pitch + column * sizeof(T).  PAIR = 2: row-pair mode (blockDim.x >=
// 2 * roundup64(Z)); PAIR = 1: one lane group (blockDim.x >= Z).  flag: one
// LDS word outside the values (the early-termination vote).
template <int BG, typename T, typename TIn, int MAXR, int PAIR>
__device__ __forceinline__ int decode_lm(T* __restrict__ app, const cb_job& job, const cb_buffers& b, uint32_t* flag)
{
    const int Z = job.Z;
    const int ncol = bg_traits<BG>::NSYS + job.mb;
    constexpr bool SPL = lm_split_layout((int)sizeof(T), MAXR);
    const uint32_t pitch = lm_pitch(BG, ncol, (int)sizeof(T), SPL);
    unsigned char* const lds = reinterpret_cast<unsigned char*>(app);
    // LLRs: thread t takes lane z = t mod Z of columns t / Z, t / Z + per, ...
    // (per = blockDim.x / Z >= 1 column streams; coalesced over z)
    {
        const TIn* src = static_cast<const TIn*>(b.llr);
        const int per = (int)blockDim.x / Z;
        const int t = (int)threadIdx.x;
        const T clampv = (T)job.clamp;
        const bool do_clamp = job.clamp > 0.0f;
        if (t < per * Z) {
            const int c0 = t / Z;
            const int z = t - c0 * Z;
            for (int col0 = c0; col0 < ncol; col0 += LLR_LOADS * per) {
                TIn raw[LLR_LOADS];  // loads in flight (a column past the end reloads column c0)
#pragma unroll
                for (int u = 0; u < LLR_LOADS; ++u) {
                    const int col = col0 + u * per;
                    raw[u] = src[(col < ncol ? col : c0) * Z + z];
                }
#pragma unroll
                for (int u = 0; u < LLR_LOADS; ++u) {
                    const int col = col0 + u * per;
                    if (col >= ncol) break;
                    T v = llr_in<TIn, T>::cvt(raw[u]);
                    if (do_clamp) v = num<T>::vmax(num<T>::vmin(v, clampv), -clampv);
                    lds_at<T>(lds, (uint32_t)z * pitch + lm_col_offset(BG, col, (int)sizeof(T), SPL)) =
                        (v == (T)0) ? (T)0 : v;
                }
            }
        }
        __syncthreads();
    }
    const int iters = run_iterations<BG, T, MAXR, PAIR, true>(app, job, 1, 0, flag);
    const int nbits = job.Kb * Z;
    // hard decisions, 64 per wave per step by ballot (as store_hard); each
    // lane walks its (column, z) position instead of dividing per bit
    {
        const int nwords = (nbits + 31) >> 5;
        const int tid = (int)threadIdx.x;
        const int lane = tid & 63;
        const int nwaves = (int)((blockDim.x + 63u) >> 6);
        const int step = nwaves * 64;
        int n = (tid >> 6) * 64 + lane;
        int col = n / Z, z = n - col * Z;
        for (int c = tid >> 6; c * 64 < nbits; c += nwaves) {
            const bool bit =
                (n < nbits) && (lds_at<T>(lds, (uint32_t)z * pitch + lm_col_offset(BG, col, (int)sizeof(T), SPL)) <= (T)0);
            const uint64_t m = __ballot(bit);
            if (lane == 0) {
                b.out[2 * c] = (uint32_t)m;
                if (2 * c + 1 < nwords) b.out[2 * c + 1] = (uint32_t)(m >> 32);
            }
            n += step;
            z += step;
            while (z >= Z) {
                z -= Z;
                ++col;
            }
        }
    }
    if (b.soft) {
        const int step = (int)blockDim.x;
        int n = (int)threadIdx.x;
        int col = n / Z, z = n - col * Z;
        for (; n < nbits; n += step) {
            b.soft[n] = (_Float16)lds_at<T>(lds, (uint32_t)z * pitch + lm_col_offset(BG, col, (int)sizeof(T), SPL));
            z += step;
            while (z >= Z) {
                z -= Z;
                ++col;
            }
        }
    }
    return iters;
}

// Code blocks per workgroup of a transport-block launch: small lifting
// sizes share a wave (64 / Z code blocks for Z <= 32), larger ones get a
// workgroup each.  Host and device use the same rule, so a launch
// descriptor needs nothing beyond cuphyLDPCDecodeDesc_t.
__host__ __device__ constexpr int tb_group_size(int Z) { return (Z >= 2 && Z <= 32) ? 64 / Z : 1; }
constexpr int TB_MAX_GROUP = 32;

// Number of code words of a transport-block launch.
__host__ __device__ inline int desc_codewords(const cuphyLDPCDecodeDesc_t& d)
{
    int n = 0;
    for (int i = 0; i < d.num_tbs; ++i) n += d.llr_input[i].num_codewords;
    return n;
}

// Shared configuration of a transport-block launch.
template <typename TIn>
__device__ __forceinline__ cb_job tb_job(const cuphyLDPCDecodeDesc_t& desc)
{
    const cuphyLDPCDecodeConfigDesc_t& cfg = desc.config;
    cb_job job;
    job.llr = nullptr;
    job.out = nullptr;
    job.soft = nullptr;
    job.Z = cfg.Z;
    job.mb = cfg.num_parity_nodes;
    job.Kb = cfg.Kb;
    job.max_iter = cfg.max_iterations;
    job.early_term = (cfg.flags & CUPHY_LDPC_DECODE_EARLY_TERM) != 0;
    job.clamp = cfg.clamp_value;
    job.alpha_f32 = cfg.norm.f32;
    job.alpha_f16 = (uint16_t)(cfg.norm.f16x2 & 0xFFFFu);
    return job;
}

// Buffers of code word `cw` of a transport-block launch.
template <typename TIn>
__device__ __forceinline__ cb_buffers tb_buffers(const cuphyLDPCDecodeDesc_t& desc, int cw)
{
    const cw_location loc = locate_cw(desc, cw);
    cb_buffers b;
    b.llr = static_cast<const TIn*>(desc.llr_input[loc.tb].addr) +
            (size_t)loc.idx * (size_t)desc.llr_input[loc.tb].stride_elements;
    b.out = desc.tb_output[loc.tb].addr + (size_t)loc.idx * (size_t)desc.tb_output[loc.tb].stride_words;
    b.soft = ((desc.config.flags & CUPHY_LDPC_DECODE_WRITE_SOFT_OUTPUTS) && desc.llr_output[loc.tb].addr)
                 ? static_cast<_Float16*>(desc.llr_output[loc.tb].addr) +
                       (size_t)loc.idx * (size_t)desc.llr_output[loc.tb].stride_elements
                 : nullptr;
    return b;
}

// Reasons (CUPHY_HIP_ITEM_*) a batch launch whose kernel holds max_mb block
// rows does not decode a work item; zi is Z's index among the lifting sizes
// (-1: not one).  The same predicate runs on the host
// (cuphyHipLDPCWorkItemCheck) and in the persistent kernels.
__host__ __device__ inline uint32_t work_item_reasons(int BG, int zi, int mb, int Kb, int max_iter, bool has_llr,
                                                      bool has_out, int max_mb)
{
    uint32_t r = 0;
    if (BG != 1 && BG != 2) r |= CUPHY_HIP_ITEM_BAD_BG;
    if (zi < 0) r |= CUPHY_HIP_ITEM_BAD_Z;
    if (mb < 4 || mb > (BG == 2 ? 42 : 46) || mb > max_mb) r |= CUPHY_HIP_ITEM_BAD_MB;
    if (BG == 1 ? Kb != 22 : !(Kb == 6 || Kb == 8 || Kb == 9 || Kb == 10)) r |= CUPHY_HIP_ITEM_BAD_KB;
    if (max_iter < 0) r |= CUPHY_HIP_ITEM_BAD_ITER;
    if (!has_llr || !has_out) r |= CUPHY_HIP_ITEM_NULL_BUFFER;
    return r;
}

__device__ __forceinline__ uint32_t work_item_reasons(const cuphyHipLDPCWorkItem_t& w, int max_mb)
{
    return work_item_reasons(w.BG, z_index(w.Z), w.num_parity_nodes, w.Kb, w.max_iterations, w.llr != nullptr,
                             w.out != nullptr, max_mb);
}

// Work items a and b share the configuration a group decodes with.
__device__ __forceinline__ bool same_config(const cuphyHipLDPCWorkItem_t& a, const cuphyHipLDPCWorkItem_t& b)
{
    return a.BG == b.BG && a.Z == b.Z && a.num_parity_nodes == b.num_parity_nodes && a.Kb == b.Kb &&
           a.max_iterations == b.max_iterations && a.flags == b.flags && a.norm.f16x2 == b.norm.f16x2 &&
           __builtin_bit_cast(uint32_t, a.clamp_value) == __builtin_bit_cast(uint32_t, b.clamp_value);
}

// Shared configuration of a batch work item (the first of its group).
__device__ __forceinline__ cb_job item_job(const cuphyHipLDPCWorkItem_t& w)
{
    cb_job job;
    job.llr = w.llr;
    job.out = w.out;
    job.soft = static_cast<_Float16*>(w.soft);
    job.Z = w.Z;
    job.mb = w.num_parity_nodes;
    job.Kb = w.Kb;
    job.max_iter = w.max_iterations;
    job.early_term = (w.flags & CUPHY_LDPC_DECODE_EARLY_TERM) != 0;
    job.clamp = w.clamp_value;
    job.alpha_f32 = w.norm.f32;
    job.alpha_f16 = (uint16_t)(w.norm.f16x2 & 0xFFFFu);
    return job;
}

}  // namespace nrldpc
