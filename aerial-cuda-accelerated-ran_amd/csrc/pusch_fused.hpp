// Host/device interface of the fused slot decode (pusch_fused.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>

#include "cuphy_hip.h"

namespace nrldpc {

// Workspace: the queue state (first FUSED_WS_QUEUE_WORDS words, of which 2
// are used) followed by two words per CB slot, of which each TB uses the
// pair at its first slot: TB-CRC accumulator and count of finished CBs.
constexpr int FUSED_WS_QUEUE_WORDS = 64;
// Largest num_cbs of cuphyHipPuschSlotDecodeNoAlloc (its scratch array
// holds FUSED_WS_QUEUE_WORDS + 2 * NOALLOC_MAX_CBS words, 128.25 KB).
constexpr int NOALLOC_MAX_CBS = 16384;

// Work items of the fused kernels: (TB << 16) | ((count - 1) << 12) | CB,
// the group of `count` consecutive CBs CB .. CB + count - 1 of the TB,
// decoded side by side in one workgroup (count * Zc <= 384).
constexpr int FUSED_MAX_GROUP = 16;
constexpr int FUSED_BLOCK = 384;  // threads of a fused-kernel workgroup
__host__ __device__ inline int item_cb(uint32_t it) { return (int)(it & 0xFFFu); }
__host__ __device__ inline int item_count(uint32_t it) { return (int)((it >> 12) & 0xFu) + 1; }
__host__ __device__ inline uint32_t make_item(int tb, int cb, int count)
{
    return ((uint32_t)tb << 16) | ((uint32_t)(count - 1) << 12) | (uint32_t)cb;
}

// LDS values (binary16) of the fused kernels: a slot launch takes up to
// SLOT_MAX_LDS_VALUES; a resident launch always has SERVE_LDS_VALUES (any
// HARQ/APP array, 68 * 384, plus 2048 descrambling words: E up to 65,473
// bits per descrambled CB of the largest array, two workgroups per CU).
constexpr int SLOT_MAX_LDS_VALUES = 68 * 384 + 2 * 12288;
constexpr int SERVE_LDS_VALUES = 68 * 384 + 2 * 2048;

// LDS values a CB of TB parameters p needs in the fused kernels: the HARQ
// circular buffer [2Zc | Ncb] and the decoder's APP array ((nsys + mb) * Zc)
// share one array; descrambled CBs add their sequence words after it (two
// values per word; words = (start + E - 1) / 32 - start / 32 + 1 of the
// largest E, start % 32 at most 31, plus one).  The same per-CB split of G
// into E as the de-rate-match kernel (TS 38.212 5.4.2.1).
// A group of `count` CBs holds count arrays, each but the last rounded up to
// whole 16-byte units.
inline int slot_lds_values(const PerTbParams& p, int mb, bool descramble, int count = 1)
{
    const int nsys = (p.bg == 1) ? 22 : 10;
    const int arr1 = std::max<int>((int)(2 * p.Zc + p.Ncb), (nsys + mb) * (int)p.Zc);
    const int arr = (count - 1) * ((arr1 + 7) & ~7) + arr1;
    if (!descramble || p.uciOnPuschFlag) return arr;
    const long long nq = (long long)p.Nl * p.Qm, C = p.num_CBs;
    const long long total = p.encodedSize;
    if (nq <= 0 || C <= 0) return arr;
    const long long q = total / nq / C;
    const long long emax = nq * q + (q * nq * C < total ? nq : 0);
    const long long words = (31 + emax - 1) / 32 + 1;
    return (int)std::min<long long>(((arr + 1) & ~1) + 2 * (words + 1), 1LL << 30);
}

// CBs per work item for a TB: as many as the workgroup's lanes hold (count *
// Zc <= FUSED_BLOCK, at most FUSED_MAX_GROUP and num_CBs) whose LDS need
// stays within lds_limit.
inline int fused_group_size(const PerTbParams& p, int mb, bool descramble, int lds_limit)
{
    int n = std::min<int>({FUSED_MAX_GROUP, FUSED_BLOCK / std::max<int>(1, (int)p.Zc), (int)p.num_CBs});
    while (n > 1 && slot_lds_values(p, mb, descramble, n) > lds_limit) --n;
    return std::max(n, 1);
}

// Resident serving launch (pusch_serve_resident): the host-written ring
// header (host-mapped pinned memory; items follow in their own array) and
// the device-side claim state (device memory, zero before the first launch
// and kept across launches).
struct serve_ring_hdr {
    uint32_t tail;  // items published (host: release store)
    uint32_t stop;  // host: 1 = leave once no work is left
    uint32_t pad[62];
};
struct serve_state {       // one 256-byte line per word: tickets do not contend with polls
    uint32_t head;         // next ticket (item index)
    uint32_t pad0[63];
    uint32_t tail;         // last host tail seen
    uint32_t pad1[63];
    uint32_t token;        // a workgroup is reading the host tail
    uint32_t stop;         // host stop seen
    uint32_t leave;        // = the launch's epoch once one of its workgroups left: the others follow
    uint32_t pad2[61];
};

// Workgroups of a resident launch (the same for every launch of a server:
// abandoned tickets are handed over by workgroup index); 0 on failure.
int serve_resident_grid(int cu_count);
// abandoned: `grid` words, all 0xFFFFFFFF before the first launch.
cuphyStatus_t serve_resident_launch(const serve_ring_hdr* ring, const uint32_t* ring_items, uint32_t ring_mask,
                                    const cuphyHipPuschSlotTb_t* jobs, int n_slots, serve_state* st, uint32_t* tbw,
                                    uint32_t* abandoned, int grid, cuphyDataType_t llr_type, uint64_t idle_ticks,
                                    uint64_t life_ticks, uint32_t epoch, hipStream_t strm);

// Look up the device address of the allocation-free entry point's scratch
// array on `device` (once per device; the device is made current for the
// lookup).
cuphyStatus_t noalloc_scratch_init(int device);

// slot_decode with the library's scratch array of the current device as the
// workspace, cleared first (n_cbs words pairs plus the queue state).
cuphyStatus_t slot_decode_noalloc(const cuphyHipPuschSlotTb_t* d_tbs, int n_tbs, const uint32_t* d_items, int n_items,
                                  int n_cbs, cuphyDataType_t llr_type, int max_lds_values, int max_parity_nodes,
                                  int max_cbs_per_item, int cu_count, hipStream_t strm);

// Enqueue one persistent fused launch (de-rate-match + HARQ combine +
// decode + CB/TB CRC) over the CB items of a slot.
cuphyStatus_t slot_decode(const cuphyHipPuschSlotTb_t* d_tbs, int n_tbs, const uint32_t* d_items, int n_items,
                          cuphyDataType_t llr_type, int max_lds_values, int max_parity_nodes, int max_cbs_per_item,
                          void* d_ws, int cu_count, hipStream_t strm);

}  // namespace nrldpc
