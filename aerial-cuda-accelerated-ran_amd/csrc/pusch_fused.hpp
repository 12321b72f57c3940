// Host/device interface of the fused slot decode (pusch_fused.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "cuphy_hip.h"

namespace nrldpc {

// Workspace: the queue state (first FUSED_WS_QUEUE_WORDS words, of which 2
// are used) followed by the TB-CRC share slots (one per CB).
constexpr int FUSED_WS_QUEUE_WORDS = 64;

// Enqueue one persistent fused launch (de-rate-match + HARQ combine +
// decode + CB/TB CRC) over the CB items of a slot.
cuphyStatus_t slot_decode(const cuphyHipPuschSlotTb_t* d_tbs, int n_tbs, const uint32_t* d_items, int n_items,
                          cuphyDataType_t llr_type, int max_lds_values, int max_parity_nodes, void* d_ws,
                          int cu_count, hipStream_t strm);

}  // namespace nrldpc
