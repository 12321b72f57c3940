// Host-side interface of the HIP layered LDPC decoder (ldpc_decode.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "cuphy_hip.h"

namespace nrldpc {

struct launch_geometry {
    void* func = nullptr;
    dim3 grid, block;
    size_t lds_bytes = 0;
};

// Z is one of the 51 TS 38.212 lifting sizes.
bool is_lifting_size(int Z);

// (BG, Z, mb, Kb, LLR type) is decodable by the kernels of this library.
bool config_supported(const cuphyLDPCDecodeConfigDesc_t& c);

// Kernel and launch shape for a transport-block descriptor.
cuphyStatus_t fill_launch(const cuphyLDPCDecodeDesc_t& desc, launch_geometry& g);

// Raise a kernel's dynamic-LDS limit above the 64 KB default, once per
// (kernel, size) high-water mark.
cuphyStatus_t ensure_lds_attr(void* func, size_t bytes);

// Enqueue the decode of every code word of `desc` on `strm`.
cuphyStatus_t decode_tb(const cuphyLDPCDecodeDesc_t& desc, hipStream_t strm);

// Populate a kernel-node launch configuration (graph capture / driver launch).
cuphyStatus_t get_launch_config(cuphyLDPCDecodeLaunchConfig_t& lc);

// Reasons (CUPHY_HIP_ITEM_*) a launch holding max_mb block rows does not
// decode the item, except those that depend on the launch's LDS and lanes.
uint32_t work_item_check(const cuphyHipLDPCWorkItem_t& w, int max_mb);

// Enqueue one persistent launch over a heterogeneous batch of work-item
// groups (d_group_first: n_groups + 1 item offsets, or null for one item per
// group).  max_values_per_group: LDS values of the largest group; max_mb:
// the largest num_parity_nodes of the batch (<= 8 selects the
// high-code-rate variant with the smaller register footprint);
// block_threads: workgroup size (a multiple of 64, <= 384).
cuphyStatus_t decode_batch(cuphyHipLDPCWorkItem_t* d_items, const uint32_t* d_group_first, int n_groups,
                           cuphyDataType_t llr_type, bool fp16_math, int max_values_per_group, int max_mb,
                           int block_threads, uint32_t* d_queue, int cu_count, hipStream_t strm);

}  // namespace nrldpc
