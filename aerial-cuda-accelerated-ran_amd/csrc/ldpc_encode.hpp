// Host-side interface of the HIP LDPC encoder (ldpc_encode.hip).
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "cuphy_hip.h"

namespace nrldpc {

// A bit tensor of the encoder API: code word c at addr + c * stride_words.
struct enc_tensor {
    uint32_t* addr;
    int32_t stride_words;
    int32_t num_cw;
};

// Descriptor of a cuphySetupLDPCEncode launch (kernel argument by value);
// input / output: num_TBs tensors each, in device memory (the workspace).
struct enc_descr {
    const enc_tensor* input;
    const enc_tensor* output;
    int32_t num_TBs;
    int16_t BG, Kb, Z, num_rows;
    int32_t puncture;
};

// The bit-packed encoder kernel (argument: enc_descr by value; grid
// (code words, TBs), 256 threads, (nsys + num_rows + 4) * Z bytes of LDS).
void* encode_bits_kernel();

// Encode ncb code blocks: info bytes (nsys*Z bits, 0/1) -> codeword bytes
// ((nsys + nrows) * Z bits).
cuphyStatus_t encode_cbs(int bg, int Z, const uint8_t* info, int info_stride, int ncb, uint8_t* cw, int cw_stride,
                         hipStream_t strm);

}  // namespace nrldpc
