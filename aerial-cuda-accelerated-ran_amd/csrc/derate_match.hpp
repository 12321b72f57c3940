// Host/device interface of the PUSCH de-rate-match kernel (derate_match.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "cuphy_hip.h"

namespace nrldpc {

// Device descriptor built by cuphySetupPuschRxRateMatch (the reference's
// puschRxRateMatchDescr, rate_matching.hpp:23-31).
struct rm_descriptor {
    const void* llr_in[MAX_N_TBS_SUPPORTED];  // input LLRs of scheduled TB slot i
    uint16_t sch_user_idxs[MAX_N_TBS_SUPPORTED];
    void** out;                               // HARQ buffer of each UE index
    const PerTbParams* tb_prms;               // device copy, indexed by UE index
    int descrambling_on;
};

// Kernel of a FP configuration (0: fp32->fp32, 1: fp16->fp32, 2: fp32->fp16,
// 3: fp16->fp16), with or without descrambling; nullptr otherwise.
void* rm_kernel(int fpconfig, bool descrambling);
void* rm_noop_kernel();

constexpr unsigned RM_BLOCK = 256;
// Descrambling sequence words per CB held in LDS (48 KB): E up to
// 32 * (RM_SEQ_WORDS - 1) = 393,184 rate-matched bits per CB (a CB's bits
// may start anywhere in a word).
constexpr int RM_SEQ_WORDS = 12288;
constexpr int RM_MAX_DESCR_E = 32 * (RM_SEQ_WORDS - 1);

}  // namespace nrldpc
