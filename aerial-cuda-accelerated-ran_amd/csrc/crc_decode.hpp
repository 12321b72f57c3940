// Host/device interface of the PUSCH CB/TB CRC kernels (crc_decode.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include "cuphy_hip.h"

namespace nrldpc {

constexpr int MAX_WORDS_PER_CB = MAX_DECODED_CODE_BLOCK_BIT_SIZE / 32;  // 264

// Device descriptor built by cuphySetupPuschRxCrcDecode (the reference's
// puschRxCrcDecodeDescr).
struct crc_descriptor {
    uint16_t sch_user_idxs[MAX_N_TBS_SUPPORTED];
    uint32_t* cb_crc_out;        // one remainder per code block
    uint8_t* tb_out;             // assembled transport blocks
    const uint32_t* cb_in;       // decoder output words, 264 per code block
    uint32_t* tb_crc_out;        // one remainder per transport block
    const PerTbParams* tb_prms;  // device copy, indexed by UE index
    int reverse_bytes;
};

void* crc_cb_kernel_ptr();
void* crc_tb_kernel_ptr();

constexpr unsigned CRC_CB_BLOCK = 64;
constexpr unsigned CRC_TB_BLOCK = 256;

}  // namespace nrldpc
