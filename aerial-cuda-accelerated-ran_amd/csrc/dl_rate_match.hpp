// Host-side interface of the DL rate-matching kernel (dl_rate_match.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "cuphy_hip.h"

namespace nrldpc {

// Largest emax (bits per rate-matched CB slot) of a launch.
constexpr int DL_RM_MAX_EMAX = 32 * 12288;

// Rate-match (and scramble) every CB of num_tbs TBs (cuphyHipDlRateMatch).
cuphyStatus_t dl_rate_match(const PerTbParams* d_prms, int num_tbs, const uint32_t* const* d_cw, int max_cbs,
                            int emax, bool scrambling, uint32_t* d_out, hipStream_t strm);

}  // namespace nrldpc
