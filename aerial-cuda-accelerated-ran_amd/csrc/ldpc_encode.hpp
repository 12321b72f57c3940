// Host-side interface of the HIP LDPC encoder (ldpc_encode.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "cuphy_hip.h"

namespace nrldpc {

// Encode ncb code blocks: info bytes (nsys*Z bits, 0/1) -> codeword bytes
// ((nsys + nrows) * Z bits).
cuphyStatus_t encode_cbs(int bg, int Z, const uint8_t* info, int info_stride, int ncb, uint8_t* cw, int cw_stride,
                         hipStream_t strm);

}  // namespace nrldpc
