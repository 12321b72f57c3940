// Shared definitions of the HIP LDPC decode path (device + host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>
#include <utility>

#include "cuphy_hip.h"
#include "nr_bg_tables.h"

namespace nrldpc {

// Compile-time view of a TS 38.212 base graph (tables in nr_bg_tables.h).
template <int BG> struct bg_traits;
// Rows are processed in order; a workgroup barrier is needed before row r
// only if r shares a column with a row processed since the last barrier
// (the layered schedule's read-after-write and write-after-read hazards
// across waves).  Bit r of the mask: barrier before row r.
template <typename Tr> constexpr uint64_t barrier_mask()
{
    uint64_t need = 0;
    uint64_t acc_lo = 0, acc_hi = 0;
    for (int r = 0; r < Tr::NROWS; ++r) {
        uint64_t lo = 0, hi = 0;
        for (int e = Tr::row_start(r); e < Tr::row_start(r + 1); ++e) {
            const int c = Tr::col(e);
            if (c < 64) lo |= 1ull << c; else hi |= 1ull << (c - 64);
        }
        if ((lo & acc_lo) || (hi & acc_hi)) {
            need |= 1ull << r;
            acc_lo = lo;
            acc_hi = hi;
        } else {
            acc_lo |= lo;
            acc_hi |= hi;
        }
    }
    return need;
}

template <> struct bg_traits<1> {
    static constexpr int NROWS = NR_BG1_NUM_ROWS;  // 46
    static constexpr int NSYS = 22;                // systematic columns incl. 2 punctured
    static constexpr int NEDGES = NR_BG1_NUM_EDGES;
    static constexpr int MAXDEG = 19;
    static constexpr int row_start(int r) { return nr_bg1_row_start[r]; }
    static constexpr int col(int e) { return nr_bg1_col[e]; }
    // shift 0 in every lifting set: an identity circulant for every Z
    static constexpr bool identity_edge(int e)
    {
        for (int i = 0; i < 8; ++i)
            if (nr_bg1_shift[i][e] != 0) return false;
        return true;
    }
};
template <> struct bg_traits<2> {
    static constexpr int NROWS = NR_BG2_NUM_ROWS;  // 42
    static constexpr int NSYS = 10;
    static constexpr int NEDGES = NR_BG2_NUM_EDGES;
    static constexpr int MAXDEG = 10;
    static constexpr int row_start(int r) { return nr_bg2_row_start[r]; }
    static constexpr int col(int e) { return nr_bg2_col[e]; }
    static constexpr bool identity_edge(int e)
    {
        for (int i = 0; i < 8; ++i)
            if (nr_bg2_shift[i][e] != 0) return false;
        return true;
    }
};

// TS 38.212 Table 5.3.2-1: lifting-set index of Z, -1 if Z is not a lifting size.
__host__ __device__ constexpr int lifting_set_index(int Z)
{
    const int a[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    for (int i = 0; i < 8; ++i)
        for (int v = a[i]; v <= 384; v <<= 1)
            if (v == Z) return i;
    return -1;
}

// The 51 lifting sizes in increasing order, and the position of Z among them.
constexpr int NUM_LIFTING_SIZES = 51;
constexpr int lifting_sizes[NUM_LIFTING_SIZES] = {
    2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12,  13,  14,  15,  16,  18,  20,  22,  24,  26,  28,  30,  32,  36,  40, 44,
    48, 52, 56, 60, 64, 72, 80, 88, 96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384};
__host__ __device__ constexpr int lifting_size_index(int Z)
{
    for (int i = 0; i < NUM_LIFTING_SIZES; ++i)
        if (lifting_sizes[i] == Z) return i;
    return -1;
}

// Circulant shifts reduced mod Z for every (BG, Z), edge order of
// nr_bg_tables.h; row stride SHIFT_ROW_STRIDE.  Built at compile time.
constexpr int SHIFT_ROW_STRIDE = 320;
struct reduced_shift_tables {
    uint32_t v[2][NUM_LIFTING_SIZES][SHIFT_ROW_STRIDE];
};
constexpr reduced_shift_tables make_reduced_shift_tables()
{
    reduced_shift_tables t{};
    for (int zi = 0; zi < NUM_LIFTING_SIZES; ++zi) {
        const int Z = lifting_sizes[zi];
        const int ils = lifting_set_index(Z);
        for (int e = 0; e < NR_BG1_NUM_EDGES; ++e) t.v[0][zi][e] = (uint32_t)(nr_bg1_shift[ils][e] % Z);
        for (int e = 0; e < NR_BG2_NUM_EDGES; ++e) t.v[1][zi][e] = (uint32_t)(nr_bg2_shift[ils][e] % Z);
    }
    return t;
}

// Row starts and columns of both base graphs as one constant-initialised
// table (runtime-indexed device copies: g_graph_tables in the .hip files).
struct graph_tables {
    uint16_t row_start[2][48];
    uint8_t col[2][320];
};
constexpr graph_tables make_graph_tables()
{
    graph_tables t{};
    for (int r = 0; r <= NR_BG1_NUM_ROWS; ++r) t.row_start[0][r] = nr_bg1_row_start[r];
    for (int r = 0; r <= NR_BG2_NUM_ROWS; ++r) t.row_start[1][r] = nr_bg2_row_start[r];
    for (int e = 0; e < NR_BG1_NUM_EDGES; ++e) t.col[0][e] = nr_bg1_col[e];
    for (int e = 0; e < NR_BG2_NUM_EDGES; ++e) t.col[1][e] = nr_bg2_col[e];
    return t;
}

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1.
template <typename F, int... I>
__host__ __device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__host__ __device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

}  // namespace nrldpc
