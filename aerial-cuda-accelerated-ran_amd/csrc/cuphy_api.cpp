// C ABI of the LDPC decode path: context, tensor descriptors, error strings
// and the LDPC decoder entry points.  Argument validation and return codes
// follow the reference's cuPHY/src/cuphy/cuphy_ldpc.cpp:26-222 and
// error_correction/ldpc.cpp:709-913.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <new>

#include "trace.hpp"
#include "cuphy_hip.h"
#include "api_objects.hpp"
#include "ldpc_decode.hpp"
#include "ldpc_encode.hpp"
#include "derate_match.hpp"
#include "dl_rate_match.hpp"
#include "crc_decode.hpp"
#include "pusch_fused.hpp"

// ---------------------------------------------------------------------------
// context (cuphy.h:1156-1208): records the device it was created on
// ---------------------------------------------------------------------------
struct cuphyContext {
    int device = 0;
    int cu_count = 0;
    size_t lds_per_block = 0;
};

// ---------------------------------------------------------------------------
// tensor descriptor (cuphy.h:1225-1470): type, rank, dims, strides
// ---------------------------------------------------------------------------
struct cuphyTensorDescriptor {
    cuphyDataType_t type = CUPHY_VOID;
    int rank = 0;
    int dims[CUPHY_DIM_MAX] = {0};
    int strides[CUPHY_DIM_MAX] = {0};
};

// ---------------------------------------------------------------------------
// PUSCH de-rate-match and CRC-decode objects (cuphy.h:3349-3490, 3154-3293)
// ---------------------------------------------------------------------------
struct rm_object : cuphyPuschRxRateMatch {
    int fpconfig = 0;
    bool descrambling = false;
};

// Largest E_r of a TB (TS 38.212 5.4.2.1, as the kernel computes it).
static uint32_t rm_max_E(const PerTbParams& p)
{
    const uint64_t nq = (uint64_t)p.Nl * p.Qm, C = p.num_CBs;
    const uint64_t total = p.uciOnPuschFlag ? p.G : p.encodedSize;
    if (nq == 0 || C == 0) return 0;
    const uint64_t q = (total / nq) / C;
    return (uint32_t)(nq * q + nq * ((q * nq * C < (uint64_t)p.encodedSize) ? 1u : 0u));
}
struct crc_object : cuphyPuschRxCrcDecode {
    int reverse_bytes = 1;
};

namespace {

void set_node(hipKernelNodeParams& n, void* func, dim3 grid, dim3 block, void** args)
{
    n.func = func;
    n.gridDim = grid;
    n.blockDim = block;
    n.sharedMemBytes = 0;
    n.kernelParams = args;
    n.extra = nullptr;
}

size_t type_bytes(cuphyDataType_t t)
{
    switch (t) {
    case CUPHY_R_32F: case CUPHY_R_32I: case CUPHY_R_32U: return 4;
    case CUPHY_R_64F: return 8;
    case CUPHY_R_16F: case CUPHY_R_16I: case CUPHY_R_16U: return 2;
    case CUPHY_C_16F: return 4;
    case CUPHY_R_8I: case CUPHY_R_8U: return 1;
    case CUPHY_BIT: return 0;  // sub-byte: stored in 32-bit words
    default: return 0;
    }
}

bool valid_type(cuphyDataType_t t)
{
    return t == CUPHY_BIT || type_bytes(t) > 0;
}

// Normalisation tables g_min_sum_norm_BG{1,2}_Z384 indexed by mb
// (error_correction/ldpc.cpp:139-235; identical to LDPC_decode.m:101-208).
const float kNormBG1[47] = {0.0f,  0.0f,  0.0f,  0.0f,  0.79f, 0.77f, 0.75f, 0.73f, 0.75f, 0.70f, 0.67f, 0.68f,
                            0.67f, 0.67f, 0.68f, 0.66f, 0.65f, 0.66f, 0.64f, 0.65f, 0.65f, 0.65f, 0.65f, 0.66f,
                            0.66f, 0.66f, 0.66f, 0.66f, 0.66f, 0.67f, 0.66f, 0.65f, 0.64f, 0.63f, 0.63f, 0.63f,
                            0.63f, 0.63f, 0.62f, 0.63f, 0.63f, 0.64f, 0.63f, 0.63f, 0.63f, 0.62f, 0.63f};
const float kNormBG2[43] = {0.0f,  0.0f,  0.0f,  0.0f,  0.86f, 0.84f, 0.80f, 0.77f, 0.75f, 0.75f, 0.74f,
                            0.74f, 0.74f, 0.73f, 0.73f, 0.73f, 0.73f, 0.72f, 0.70f, 0.71f, 0.71f, 0.71f,
                            0.71f, 0.70f, 0.69f, 0.70f, 0.70f, 0.70f, 0.70f, 0.70f, 0.70f, 0.70f, 0.70f,
                            0.68f, 0.67f, 0.67f, 0.68f, 0.69f, 0.69f, 0.69f, 0.69f, 0.69f, 0.69f};

// IEEE binary16 bits of a float, round to nearest even.
uint16_t float_to_half_bits(float f)
{
    const _Float16 h = (_Float16)f;
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
const char* CUPHYWINAPI cuphyGetErrorName(cuphyStatus_t status)
{
    switch (status) {
    case CUPHY_STATUS_SUCCESS: return "CUPHY_STATUS_SUCCESS";
    case CUPHY_STATUS_INTERNAL_ERROR: return "CUPHY_STATUS_INTERNAL_ERROR";
    case CUPHY_STATUS_NOT_SUPPORTED: return "CUPHY_STATUS_NOT_SUPPORTED";
    case CUPHY_STATUS_INVALID_ARGUMENT: return "CUPHY_STATUS_INVALID_ARGUMENT";
    case CUPHY_STATUS_ARCH_MISMATCH: return "CUPHY_STATUS_ARCH_MISMATCH";
    case CUPHY_STATUS_ALLOC_FAILED: return "CUPHY_STATUS_ALLOC_FAILED";
    case CUPHY_STATUS_SIZE_MISMATCH: return "CUPHY_STATUS_SIZE_MISMATCH";
    case CUPHY_STATUS_MEMCPY_ERROR: return "CUPHY_STATUS_MEMCPY_ERROR";
    case CUPHY_STATUS_INVALID_CONVERSION: return "CUPHY_STATUS_INVALID_CONVERSION";
    case CUPHY_STATUS_UNSUPPORTED_TYPE: return "CUPHY_STATUS_UNSUPPORTED_TYPE";
    case CUPHY_STATUS_UNSUPPORTED_LAYOUT: return "CUPHY_STATUS_UNSUPPORTED_LAYOUT";
    case CUPHY_STATUS_UNSUPPORTED_RANK: return "CUPHY_STATUS_UNSUPPORTED_RANK";
    case CUPHY_STATUS_UNSUPPORTED_CONFIG: return "CUPHY_STATUS_UNSUPPORTED_CONFIG";
    case CUPHY_STATUS_UNSUPPORTED_ALIGNMENT: return "CUPHY_STATUS_UNSUPPORTED_ALIGNMENT";
    case CUPHY_STATUS_VALUE_OUT_OF_RANGE: return "CUPHY_STATUS_VALUE_OUT_OF_RANGE";
    case CUPHY_STATUS_REF_MISMATCH: return "CUPHY_STATUS_REF_MISMATCH";
    default: return "CUPHY_UNKNOWN_STATUS";
    }
}

const char* CUPHYWINAPI cuphyGetErrorString(cuphyStatus_t status)
{
    switch (status) {
    case CUPHY_STATUS_SUCCESS: return "Success";
    case CUPHY_STATUS_INTERNAL_ERROR: return "Internal error";
    case CUPHY_STATUS_NOT_SUPPORTED: return "An operation was requested that is not currently supported";
    case CUPHY_STATUS_INVALID_ARGUMENT: return "An invalid argument was provided";
    case CUPHY_STATUS_ARCH_MISMATCH: return "Requested operation not supported on current architecture";
    case CUPHY_STATUS_ALLOC_FAILED: return "Memory allocation failed";
    case CUPHY_STATUS_SIZE_MISMATCH: return "The size of the operands provided to the function do not match";
    case CUPHY_STATUS_MEMCPY_ERROR: return "An error occurred during a memcpy operation";
    case CUPHY_STATUS_INVALID_CONVERSION: return "An invalid conversion operation was requested";
    case CUPHY_STATUS_UNSUPPORTED_TYPE: return "An operation was requested on an unsupported type";
    case CUPHY_STATUS_UNSUPPORTED_LAYOUT: return "An operation was requested on an unsupported layout";
    case CUPHY_STATUS_UNSUPPORTED_RANK: return "An operation was requested on an unsupported rank";
    case CUPHY_STATUS_UNSUPPORTED_CONFIG: return "An operation was requested on an unsupported configuration";
    case CUPHY_STATUS_UNSUPPORTED_ALIGNMENT: return "One or more API arguments don't have the required alignment";
    case CUPHY_STATUS_VALUE_OUT_OF_RANGE: return "Data conversion could not occur because an input value was out of range";
    case CUPHY_STATUS_REF_MISMATCH: return "Mismatch found when comparing to reference";
    default: return "Unknown status value";
    }
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
cuphyStatus_t CUPHYWINAPI cuphyCreateContext(cuphyContext_t* pcontext, unsigned int flags)
{
    (void)flags;
    if (!pcontext) return CUPHY_STATUS_INVALID_ARGUMENT;
    *pcontext = nullptr;
    cuphyContext* c = new (std::nothrow) cuphyContext;
    if (!c) return CUPHY_STATUS_ALLOC_FAILED;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        delete c;
        return CUPHY_STATUS_INTERNAL_ERROR;
    }
    c->device = dev;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess) c->cu_count = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess) c->lds_per_block = (size_t)v;
    *pcontext = c;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyDestroyContext(cuphyContext_t ctx)
{
    if (!ctx) return CUPHY_STATUS_INVALID_ARGUMENT;
    delete ctx;
    return CUPHY_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// tensor descriptors
// ---------------------------------------------------------------------------
cuphyStatus_t CUPHYWINAPI cuphyCreateTensorDescriptor(cuphyTensorDescriptor_t* ptensorDesc)
{
    if (!ptensorDesc) return CUPHY_STATUS_INVALID_ARGUMENT;
    *ptensorDesc = new (std::nothrow) cuphyTensorDescriptor;
    return *ptensorDesc ? CUPHY_STATUS_SUCCESS : CUPHY_STATUS_ALLOC_FAILED;
}

cuphyStatus_t CUPHYWINAPI cuphyDestroyTensorDescriptor(cuphyTensorDescriptor_t tensorDesc)
{
    if (!tensorDesc) return CUPHY_STATUS_INVALID_ARGUMENT;
    delete tensorDesc;
    return CUPHY_STATUS_SUCCESS;
}

// Validation as tensor_layout_any::validate (tensor_desc.cpp:150-213);
// default strides tight, CUPHY_BIT first dimension padded to whole words,
// COALESCE pads the second-dimension stride to a multiple of 128 bytes.
cuphyStatus_t CUPHYWINAPI cuphySetTensorDescriptor(cuphyTensorDescriptor_t tensorDesc, cuphyDataType_t type,
                                                   int numDimensions, const int dimensions[], const int strides[],
                                                   unsigned int flags)
{
    if (!tensorDesc || !dimensions || !valid_type(type)) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (numDimensions <= 0 || numDimensions > CUPHY_DIM_MAX) return CUPHY_STATUS_INVALID_ARGUMENT;
    for (int i = 0; i < numDimensions; ++i)
        if (dimensions[i] < 1) return CUPHY_STATUS_INVALID_ARGUMENT;
    int st[CUPHY_DIM_MAX] = {0};
    const bool use_given = strides && !(flags & (CUPHY_TENSOR_ALIGN_TIGHT | CUPHY_TENSOR_ALIGN_COALESCE));
    if (use_given) {
        for (int i = 0; i < numDimensions; ++i) st[i] = strides[i];
        if (type == CUPHY_BIT && (st[0] != 1 || (numDimensions > 1 && st[1] % 32))) return CUPHY_STATUS_INVALID_ARGUMENT;
        // non-overlap: sorted by stride, each stride covers the previous extent
        std::array<int, CUPHY_DIM_MAX> idx{};
        for (int i = 0; i < numDimensions; ++i) idx[i] = i;
        std::sort(idx.begin(), idx.begin() + numDimensions, [&](int a, int b) {
            return st[a] != st[b] ? st[a] < st[b] : dimensions[a] < dimensions[b];
        });
        for (int i = 1; i < numDimensions; ++i)
            if (dimensions[idx[i]] > 1 && (long long)st[idx[i]] < (long long)st[idx[i - 1]] * dimensions[idx[i - 1]])
                return CUPHY_STATUS_INVALID_ARGUMENT;
    } else {
        st[0] = 1;
        for (int i = 1; i < numDimensions; ++i) {
            int prev = (i == 1 && type == CUPHY_BIT) ? ((dimensions[0] + 31) / 32) * 32 : dimensions[i - 1];
            if (i == 1 && (flags & CUPHY_TENSOR_ALIGN_COALESCE) && type != CUPHY_BIT) {
                const int mult = (int)(128 / type_bytes(type));
                prev = (prev + mult - 1) / mult * mult;
            }
            st[i] = st[i - 1] * prev;
            if (i > 1) st[i] = st[i - 1] * dimensions[i - 1];
        }
    }
    tensorDesc->type = type;
    tensorDesc->rank = numDimensions;
    for (int i = 0; i < CUPHY_DIM_MAX; ++i) {
        tensorDesc->dims[i] = i < numDimensions ? dimensions[i] : 1;
        tensorDesc->strides[i] = i < numDimensions ? st[i] : 0;
    }
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyGetTensorDescriptor(cuphyTensorDescriptor_t tensorDesc, int numDimsRequested,
                                                   cuphyDataType_t* dataType, int* numDims, int dimensions[],
                                                   int strides[])
{
    if (!tensorDesc || (numDimsRequested > 0 && !dimensions)) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (dataType) *dataType = tensorDesc->type;
    if (numDims) *numDims = tensorDesc->rank;
    for (int i = 0; i < numDimsRequested && i < CUPHY_DIM_MAX; ++i) {
        dimensions[i] = tensorDesc->dims[i];
        if (strides) strides[i] = tensorDesc->strides[i];
    }
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyGetTensorSizeInBytes(cuphyTensorDescriptor_t tensorDesc, size_t* psz)
{
    if (!tensorDesc || !psz) return CUPHY_STATUS_INVALID_ARGUMENT;
    size_t last = 0;
    for (int i = 0; i < tensorDesc->rank; ++i)
        last = std::max(last, (size_t)tensorDesc->strides[i] * (size_t)tensorDesc->dims[i]);
    *psz = (tensorDesc->type == CUPHY_BIT) ? (last + 31) / 32 * 4 : last * type_bytes(tensorDesc->type);
    return CUPHY_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// LDPC decoder
// ---------------------------------------------------------------------------
cuphyStatus_t CUPHYWINAPI cuphyCreateLDPCDecoder(cuphyContext_t context, cuphyLDPCDecoder_t* pdecoder,
                                                 unsigned int flags)
{
    (void)flags;
    if (!pdecoder || !context) return CUPHY_STATUS_INVALID_ARGUMENT;
    *pdecoder = nullptr;
    cuphyLDPCDecoder* d = new (std::nothrow) cuphyLDPCDecoder;
    if (!d) return CUPHY_STATUS_ALLOC_FAILED;
    d->device = context->device;
    d->cu_count = context->cu_count;
    // set-up of the allocation-free slot entry point on this decoder's GPU; a
    // failure here is retried by that entry point's first call
    (void)nrldpc::noalloc_scratch_init(d->device);
    *pdecoder = d;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyDestroyLDPCDecoder(cuphyLDPCDecoder_t decoder)
{
    if (!decoder) return CUPHY_STATUS_INVALID_ARGUMENT;
    delete decoder;
    return CUPHY_STATUS_SUCCESS;
}

// Tensor interface: LLR tensor [nLLR x nCW] (stride[1] = per-codeword
// stride in elements), output CUPHY_BIT tensor [K x nCW] (stride[1] in
// bits, multiple of 32), optional binary16 soft-output tensor.  Forwarded
// to the transport-block kernels as a one-TB descriptor.
cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCDecode(cuphyLDPCDecoder_t decoder, cuphyTensorDescriptor_t tensorDescDst,
                                                         void* dstAddr, cuphyTensorDescriptor_t tensorDescLLR,
                                                         const void* LLRAddr, cuphyTensorDescriptor_t tensorDescSoftOutputs,
                                                         void* softOutputsAddr, const cuphyLDPCDecodeConfigDesc_t* config,
                                                         hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphyErrorCorrectionLDPCDecode");
    // cuphy_ldpc.cpp:36-55
    if (!decoder || !tensorDescDst || !dstAddr || !tensorDescLLR || !LLRAddr || !config ||
        config->max_iterations < 0 || config->BG < 1 || config->BG > 2 ||
        (config->BG == 1 ? config->Kb != 22
                         : (config->Kb != 6 && config->Kb != 8 && config->Kb != 9 && config->Kb != 10)) ||
        config->Z < 2 || config->Z > 384 || config->num_parity_nodes < 4 ||
        (config->BG == 1 ? config->num_parity_nodes > 46 : config->num_parity_nodes > 42) ||
        ((tensorDescSoftOutputs == nullptr) != (softOutputsAddr == nullptr)))
        return CUPHY_STATUS_INVALID_ARGUMENT;
    if (config->llr_type != tensorDescLLR->type) return CUPHY_STATUS_INVALID_ARGUMENT;  // :63-66
    if (tensorDescSoftOutputs) {                                                        // :69-91
        if (tensorDescSoftOutputs->type != CUPHY_R_16F) return CUPHY_STATUS_INVALID_ARGUMENT;
        if (reinterpret_cast<uintptr_t>(softOutputsAddr) % 4) return CUPHY_STATUS_INVALID_ARGUMENT;
        if (tensorDescSoftOutputs->rank > 1 && tensorDescSoftOutputs->dims[1] > 1 &&
            tensorDescSoftOutputs->strides[1] % 2)
            return CUPHY_STATUS_INVALID_ARGUMENT;
    }
    // ldpc.cpp:733-745
    if (tensorDescLLR->rank > 2) return CUPHY_STATUS_UNSUPPORTED_RANK;
    if (tensorDescDst->type != CUPHY_BIT || tensorDescDst->rank > 2) return CUPHY_STATUS_UNSUPPORTED_TYPE;
    if (!nrldpc::config_supported(*config)) return CUPHY_STATUS_UNSUPPORTED_CONFIG;

    cuphyLDPCDecodeDesc_t d;
    std::memset(&d, 0, sizeof(d));
    d.config = *config;
    d.num_tbs = 1;
    const int ncw = tensorDescLLR->rank > 1 ? tensorDescLLR->dims[1] : 1;
    d.llr_input[0].addr = const_cast<void*>(LLRAddr);
    d.llr_input[0].stride_elements = tensorDescLLR->rank > 1 ? tensorDescLLR->strides[1] : tensorDescLLR->dims[0];
    d.llr_input[0].num_codewords = ncw;
    d.tb_output[0].addr = static_cast<uint32_t*>(dstAddr);
    d.tb_output[0].stride_words =
        tensorDescDst->rank > 1 ? tensorDescDst->strides[1] / 32 : (tensorDescDst->dims[0] + 31) / 32;
    d.tb_output[0].num_codewords = ncw;
    if (softOutputsAddr) {
        d.config.flags |= CUPHY_LDPC_DECODE_WRITE_SOFT_OUTPUTS;
        d.llr_output[0].addr = softOutputsAddr;
        d.llr_output[0].stride_elements =
            tensorDescSoftOutputs->rank > 1 ? tensorDescSoftOutputs->strides[1] : tensorDescSoftOutputs->dims[0];
        d.llr_output[0].num_codewords = ncw;
    } else {
        d.config.flags &= ~(uint32_t)CUPHY_LDPC_DECODE_WRITE_SOFT_OUTPUTS;
    }
    return nrldpc::decode_tb(d, strm);
}

cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCTransportBlockDecode(cuphyLDPCDecoder_t decoder,
                                                                       const cuphyLDPCDecodeDesc_t* decodeDesc,
                                                                       hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphyErrorCorrectionLDPCTransportBlockDecode");
    if (!decoder || !decodeDesc) return CUPHY_STATUS_INVALID_ARGUMENT;  // cuphy_ldpc.cpp:105-109
    const cuphyLDPCDecodeConfigDesc_t& c = decodeDesc->config;
    if (c.max_iterations < 0 || c.BG < 1 || c.BG > 2 || c.Z < 2 || c.Z > 384 || c.num_parity_nodes < 4 ||
        (c.BG == 1 ? c.num_parity_nodes > 46 : c.num_parity_nodes > 42) || decodeDesc->num_tbs < 0 ||
        decodeDesc->num_tbs > CUPHY_LDPC_DECODE_DESC_MAX_TB)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    for (int i = 0; i < decodeDesc->num_tbs; ++i)
        if (decodeDesc->llr_input[i].num_codewords < 0 ||
            (decodeDesc->llr_input[i].num_codewords > 0 &&
             (!decodeDesc->llr_input[i].addr || !decodeDesc->tb_output[i].addr)))
            return CUPHY_STATUS_INVALID_ARGUMENT;
    if (decodeDesc->num_tbs == 0) return CUPHY_STATUS_SUCCESS;
    if (!nrldpc::config_supported(c)) return CUPHY_STATUS_UNSUPPORTED_CONFIG;
    return nrldpc::decode_tb(*decodeDesc, strm);
}

// ---------------------------------------------------------------------------
// LDPC encoder API (cuphy.h:3781-3850; setup as error_correction/
// ldpc_encode.cu:394-620): bit tensors [K bits x C] in, [N bits x C] out.
// ---------------------------------------------------------------------------
cuphyStatus_t CUPHYWINAPI cuphyLDPCEncodeGetDescrInfo(size_t* pDescrSizeBytes, size_t* pDescrAlignBytes,
                                                      uint16_t maxUes, size_t* pWorkspaceBytes)
{
    if (!pDescrSizeBytes || !pDescrAlignBytes || !pWorkspaceBytes) return CUPHY_STATUS_INVALID_ARGUMENT;
    *pDescrSizeBytes = sizeof(nrldpc::enc_descr);
    *pDescrAlignBytes = alignof(nrldpc::enc_descr);
    *pWorkspaceBytes = 2u * (size_t)maxUes * sizeof(nrldpc::enc_tensor);  // inputs, then outputs
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphySetupLDPCEncode(cuphyLDPCEncodeLaunchConfig_t cfg, cuphyTensorDescriptor_t inDesc,
                                               void* inAddr, cuphyTensorDescriptor_t outDesc, void* outAddr, int BG,
                                               int Z, uint8_t puncture, int maxParityNodes, int max_rv,
                                               uint8_t batching, int batched_TBs, void** inBatchedAddr,
                                               void** outBatchedAddr, void* h_workspace, void* d_workspace,
                                               void* cpu_desc, void* gpu_desc, uint8_t enable_desc_async_copy,
                                               hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphySetupLDPCEncode");
    // ldpc_encode.cu:416-470
    if (!cfg || !inDesc || !outDesc || !cpu_desc || max_rv < 0 || max_rv > 3) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (batching && (batched_TBs < 1 || batched_TBs > 65535 || !inBatchedAddr || !outBatchedAddr))
        return CUPHY_STATUS_INVALID_ARGUMENT;
    if (!h_workspace || !d_workspace || (enable_desc_async_copy && !gpu_desc)) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (inDesc->rank > 2 || outDesc->rank > 2) return CUPHY_STATUS_UNSUPPORTED_RANK;
    if (inDesc->type != CUPHY_BIT || outDesc->type != CUPHY_BIT) return CUPHY_STATUS_UNSUPPORTED_TYPE;
    if (BG < 1 || BG > 2 || !nrldpc::is_lifting_size(Z)) return CUPHY_STATUS_UNSUPPORTED_CONFIG;
    // a subset of the parity rows only for rv 0 (ldpc_encode.cu:455-458)
    const int all_rows = BG == 1 ? 46 : 42;
    const bool fewer = max_rv == 0 && maxParityNodes != 0 && maxParityNodes < all_rows;
    const int num_rows = fewer ? maxParityNodes : all_rows;
    if (num_rows + (puncture ? 2 : 0) < 4) return CUPHY_STATUS_INVALID_ARGUMENT;  // :466-469
    const int nsys = BG == 1 ? 22 : 10;
    const int out_bits = (nsys + num_rows) * Z - (puncture ? 2 * Z : 0);
    auto cols = [](const cuphyTensorDescriptor& t) { return t.rank > 1 ? t.dims[1] : 1; };
    auto stride_words = [](const cuphyTensorDescriptor& t) {
        return t.rank > 1 ? t.strides[1] / 32 : (t.dims[0] + 31) / 32;
    };
    if (inDesc->dims[0] < nsys * Z || outDesc->dims[0] < out_bits) return CUPHY_STATUS_INVALID_ARGUMENT;
    if ((inDesc->rank > 1 && inDesc->strides[1] % 32) || (outDesc->rank > 1 && outDesc->strides[1] % 32))
        return CUPHY_STATUS_UNSUPPORTED_ALIGNMENT;
    const int n_tbs = batching ? batched_TBs : 1;
    nrldpc::enc_tensor* h_in = static_cast<nrldpc::enc_tensor*>(h_workspace);
    nrldpc::enc_tensor* h_out = h_in + n_tbs;
    for (int t = 0; t < n_tbs; ++t) {
        void* ia = batching ? inBatchedAddr[t] : inAddr;
        void* oa = batching ? outBatchedAddr[t] : outAddr;
        if (!ia || !oa) return CUPHY_STATUS_INVALID_ARGUMENT;
        h_in[t] = nrldpc::enc_tensor{static_cast<uint32_t*>(ia), stride_words(*inDesc), cols(*inDesc)};
        h_out[t] = nrldpc::enc_tensor{static_cast<uint32_t*>(oa), stride_words(*outDesc), cols(*outDesc)};
    }
    nrldpc::enc_descr& d = *static_cast<nrldpc::enc_descr*>(cpu_desc);
    d.input = static_cast<const nrldpc::enc_tensor*>(d_workspace);
    d.output = static_cast<const nrldpc::enc_tensor*>(d_workspace) + n_tbs;
    d.num_TBs = n_tbs;
    d.BG = (int16_t)BG;
    d.Kb = (int16_t)nsys;  // the information columns the kernel reads (ldpc_encode.cu:483)
    d.Z = (int16_t)Z;
    d.num_rows = (int16_t)num_rows;
    d.puncture = puncture ? 1 : 0;
    if (enable_desc_async_copy &&
        (hipMemcpyAsync(d_workspace, h_workspace, 2u * n_tbs * sizeof(nrldpc::enc_tensor), hipMemcpyHostToDevice,
                        strm) != hipSuccess ||
         hipMemcpyAsync(gpu_desc, cpu_desc, sizeof(nrldpc::enc_descr), hipMemcpyHostToDevice, strm) != hipSuccess))
        return CUPHY_STATUS_MEMCPY_ERROR;
    cfg->m_desc = gpu_desc;
    // the descriptor is the kernel's by-value argument (ldpc_encode.cu:544-546)
    cfg->m_kernelArgs[0] = cpu_desc;
    const int max_cw = cols(*inDesc);
    set_node(cfg->m_kernelNodeParams, nrldpc::encode_bits_kernel(), dim3((unsigned)std::max(1, max_cw), (unsigned)n_tbs, 1),
             dim3(256), &cfg->m_kernelArgs[0]);
    cfg->m_kernelNodeParams.sharedMemBytes = (unsigned)((nsys + std::max(num_rows, 4) + 4) * Z);
    return CUPHY_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// DL rate matching of encoded code blocks (library extension)
// ---------------------------------------------------------------------------
cuphyStatus_t CUPHYWINAPI cuphyHipDlRateMatch(const PerTbParams* d_tb_prms, int num_tbs, const uint32_t* const* d_cw,
                                              int max_cbs, int emax, uint8_t scrambling, uint32_t* d_out,
                                              hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphyHipDlRateMatch");
    return nrldpc::dl_rate_match(d_tb_prms, num_tbs, d_cw, max_cbs, emax, scrambling != 0, d_out, strm);
}

// cuphy_ldpc.cpp:116-155: the kernels need no workspace.
cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCDecodeGetWorkspaceSize(cuphyLDPCDecoder_t decoder,
                                                                         const cuphyLDPCDecodeConfigDesc_t* config,
                                                                         int numCodeWords, size_t* sizeInBytes)
{
    if (!decoder || !config || (config->BG != 1 && config->BG != 2) ||
        (config->Kb != 22 && config->Kb != 10 && config->Kb != 9 && config->Kb != 8 && config->Kb != 6) ||
        !nrldpc::is_lifting_size(config->Z) || numCodeWords <= 0 || !sizeInBytes ||
        config->num_parity_nodes < 4 || (config->BG == 1 ? config->num_parity_nodes > 46 : config->num_parity_nodes > 42))
        return CUPHY_STATUS_INVALID_ARGUMENT;
    if (config->llr_type != CUPHY_R_32F && config->llr_type != CUPHY_R_16F) return CUPHY_STATUS_UNSUPPORTED_CONFIG;
    *sizeInBytes = 0;
    return CUPHY_STATUS_SUCCESS;
}

// ldpc.cpp:851-879
cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCDecodeSetNormalization(cuphyLDPCDecoder_t decoder,
                                                                         cuphyLDPCDecodeConfigDesc_t* decodeDesc)
{
    if (!decoder || !decodeDesc) return CUPHY_STATUS_INVALID_ARGUMENT;
    cuphyLDPCDecodeConfigDesc_t& c = *decodeDesc;
    if ((c.llr_type != CUPHY_R_16F && c.llr_type != CUPHY_R_32F) || c.num_parity_nodes < 4 ||
        (c.BG == 1 && c.num_parity_nodes > 46) || (c.BG == 2 && c.num_parity_nodes > 42))
        return CUPHY_STATUS_INVALID_ARGUMENT;
    const float norm = (c.BG == 1) ? kNormBG1[c.num_parity_nodes] : kNormBG2[c.num_parity_nodes];
    if (c.llr_type == CUPHY_R_16F) {
        const uint32_t h = float_to_half_bits(norm);
        c.norm.f16x2 = h | (h << 16);
    } else {
        c.norm.f32 = norm;
    }
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor(cuphyLDPCDecoder_t decoder,
                                                                            cuphyLDPCDecodeLaunchConfig_t* launchConfig)
{
    CUPHY_HIP_RANGE("cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor");
    if (!decoder || !launchConfig) return CUPHY_STATUS_INVALID_ARGUMENT;
    return nrldpc::get_launch_config(*launchConfig);
}

// ---------------------------------------------------------------------------
// PUSCH de-rate-matching (cuphy_ldpc.cpp:228-307, rate_matching.cu:982-1097)
// ---------------------------------------------------------------------------
cuphyStatus_t CUPHYWINAPI cuphyPuschRxRateMatchGetDescrInfo(size_t* pDescrSizeBytes, size_t* pDescrAlignBytes)
{
    if (!pDescrSizeBytes || !pDescrAlignBytes) return CUPHY_STATUS_INVALID_ARGUMENT;
    *pDescrSizeBytes = sizeof(nrldpc::rm_descriptor);
    *pDescrAlignBytes = alignof(nrldpc::rm_descriptor);
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyCreatePuschRxRateMatch(cuphyPuschRxRateMatchHndl_t* pPuschRxRateMatchHndl, int FPconfig,
                                                      int descramblingOn)
{
    if (!pPuschRxRateMatchHndl) return CUPHY_STATUS_INVALID_ARGUMENT;
    *pPuschRxRateMatchHndl = nullptr;
    if (!nrldpc::rm_kernel(FPconfig, false)) return CUPHY_STATUS_INVALID_ARGUMENT;
    rm_object* o = new (std::nothrow) rm_object;
    if (!o) return CUPHY_STATUS_ALLOC_FAILED;
    o->fpconfig = FPconfig;
    o->descrambling = descramblingOn != 0;
    *pPuschRxRateMatchHndl = o;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphySetupPuschRxRateMatch(cuphyPuschRxRateMatchHndl_t puschRxRateMatchHndl, uint16_t nSchUes,
                                                     uint16_t* pSchUserIdxsCpu, const PerTbParams* pTbPrmsCpu,
                                                     const PerTbParams* pTbPrmsGpu, cuphyTensorPrm_t* pTPrmRmIn,
                                                     cuphyTensorPrm_t* pTPrmCdm1RmIn, void** ppRmOut, void* pCpuDesc,
                                                     void* pGpuDesc, uint8_t enableCpuToGpuDescrAsyncCpy,
                                                     cuphyPuschRxRateMatchLaunchCfg_t* pLaunchCfg, hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphySetupPuschRxRateMatch");
    if (!puschRxRateMatchHndl || !pTbPrmsCpu || !pTbPrmsGpu || !pTPrmRmIn || !pTPrmCdm1RmIn || !ppRmOut || !pCpuDesc ||
        !pGpuDesc || !pLaunchCfg || !pSchUserIdxsCpu || nSchUes > MAX_N_TBS_SUPPORTED)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    const rm_object& o = static_cast<const rm_object&>(*puschRxRateMatchHndl);
    nrldpc::rm_descriptor& d = *static_cast<nrldpc::rm_descriptor*>(pCpuDesc);
    uint32_t cmax = 0;
    for (uint32_t i = 0; i < nSchUes; ++i) {
        const uint16_t ue = pSchUserIdxsCpu[i];
        const PerTbParams& p = pTbPrmsCpu[ue];
        if (p.Qm != 1 && p.Qm != 2 && p.Qm != 4 && p.Qm != 6 && p.Qm != 8) return CUPHY_STATUS_INVALID_ARGUMENT;
        d.sch_user_idxs[i] = ue;
        if (p.uciOnPuschFlag)
            d.llr_in[i] = p.d_schAndCsi2LLRs;
        else
            d.llr_in[i] = (p.nDmrsCdmGrpsNoData == 1 ? pTPrmCdm1RmIn : pTPrmRmIn)[p.userGroupIndex].pAddr;
        cmax = std::max(cmax, p.num_CBs);
        if (o.descrambling && !p.uciOnPuschFlag && rm_max_E(p) > nrldpc::RM_MAX_DESCR_E)
            return CUPHY_STATUS_UNSUPPORTED_CONFIG;  // more descrambling bits per CB than the kernel holds
    }
    d.out = ppRmOut;
    d.tb_prms = pTbPrmsGpu;
    d.descrambling_on = o.descrambling ? 1 : 0;
    if (enableCpuToGpuDescrAsyncCpy &&
        hipMemcpyAsync(pGpuDesc, pCpuDesc, sizeof(nrldpc::rm_descriptor), hipMemcpyHostToDevice, strm) != hipSuccess)
        return CUPHY_STATUS_MEMCPY_ERROR;
    pLaunchCfg->desc = pGpuDesc;
    pLaunchCfg->kernelArgs[0] = &pLaunchCfg->desc;
    set_node(pLaunchCfg->kernelNodeParamsDriver, nrldpc::rm_kernel(o.fpconfig, o.descrambling), dim3(std::max(1u, cmax), std::max<unsigned>(1, nSchUes), 1),
             dim3(nrldpc::RM_BLOCK), pLaunchCfg->kernelArgs);
    set_node(pLaunchCfg->resetKernelNodeParamsDriver, nrldpc::rm_noop_kernel(), dim3(1), dim3(64), pLaunchCfg->kernelArgs);
    set_node(pLaunchCfg->clampKernelNodeParamsDriver, nrldpc::rm_noop_kernel(), dim3(1), dim3(64), pLaunchCfg->kernelArgs);
    if (nSchUes == 0) pLaunchCfg->kernelNodeParamsDriver.func = nrldpc::rm_noop_kernel();
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyDestroyPuschRxRateMatch(cuphyPuschRxRateMatchHndl_t puschRxRateMatchHndl)
{
    if (!puschRxRateMatchHndl) return CUPHY_STATUS_INVALID_ARGUMENT;
    delete static_cast<rm_object*>(puschRxRateMatchHndl);
    return CUPHY_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// PUSCH CB/TB CRC decode (cuphy_ldpc.cpp:312-392, crc.cu:1174-1254)
// ---------------------------------------------------------------------------
cuphyStatus_t CUPHYWINAPI cuphyPuschRxCrcDecodeGetDescrInfo(size_t* pDescrSizeBytes, size_t* pDescrAlignBytes)
{
    if (!pDescrSizeBytes || !pDescrAlignBytes) return CUPHY_STATUS_INVALID_ARGUMENT;
    *pDescrSizeBytes = sizeof(nrldpc::crc_descriptor);
    *pDescrAlignBytes = alignof(nrldpc::crc_descriptor);
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyCreatePuschRxCrcDecode(cuphyPuschRxCrcDecodeHndl_t* puschRxCrcDecodeHndl, int reverseBytes)
{
    if (!puschRxCrcDecodeHndl) return CUPHY_STATUS_INVALID_ARGUMENT;
    *puschRxCrcDecodeHndl = nullptr;
    crc_object* o = new (std::nothrow) crc_object;
    if (!o) return CUPHY_STATUS_ALLOC_FAILED;
    o->reverse_bytes = reverseBytes ? 1 : 0;
    *puschRxCrcDecodeHndl = o;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphySetupPuschRxCrcDecode(cuphyPuschRxCrcDecodeHndl_t puschRxCrcDecodeHndl, uint16_t nSchUes,
                                                     uint16_t* pSchUserIdxsCpu, uint32_t* pOutputCBCRCs,
                                                     uint8_t* pOutputTBs, const uint32_t* pInputCodeBlocks,
                                                     uint32_t* pOutputTBCRCs, const PerTbParams* pTbPrmsCpu,
                                                     const PerTbParams* pTbPrmsGpu, void* pCpuDesc, void* pGpuDesc,
                                                     uint8_t enableCpuToGpuDescrAsyncCpy,
                                                     cuphyPuschRxCrcDecodeLaunchCfg_t* pCbCrcLaunchCfg,
                                                     cuphyPuschRxCrcDecodeLaunchCfg_t* pTbCrcLaunchCfg, hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphySetupPuschRxCrcDecode");
    if (!puschRxCrcDecodeHndl || !pOutputCBCRCs || !pOutputTBs || !pInputCodeBlocks || !pOutputTBCRCs || !pTbPrmsCpu ||
        !pTbPrmsGpu || !pCpuDesc || !pGpuDesc || !pCbCrcLaunchCfg || !pTbCrcLaunchCfg || !pSchUserIdxsCpu ||
        nSchUes > MAX_N_TBS_SUPPORTED)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    const crc_object& o = static_cast<const crc_object&>(*puschRxCrcDecodeHndl);
    nrldpc::crc_descriptor& d = *static_cast<nrldpc::crc_descriptor*>(pCpuDesc);
    uint32_t cmax = 0;
    for (uint32_t i = 0; i < nSchUes; ++i) {
        d.sch_user_idxs[i] = pSchUserIdxsCpu[i];
        const PerTbParams& p = pTbPrmsCpu[pSchUserIdxsCpu[i]];
        if (p.num_CBs > MAX_N_CBS_PER_TB_SUPPORTED || p.K < p.F || p.K - p.F > MAX_DECODED_CODE_BLOCK_BIT_SIZE)
            return CUPHY_STATUS_INVALID_ARGUMENT;
        cmax = std::max(cmax, p.num_CBs);
    }
    d.cb_crc_out = pOutputCBCRCs;
    d.tb_out = pOutputTBs;
    d.cb_in = pInputCodeBlocks;
    d.tb_crc_out = pOutputTBCRCs;
    d.tb_prms = pTbPrmsGpu;
    d.reverse_bytes = o.reverse_bytes;
    if (enableCpuToGpuDescrAsyncCpy &&
        hipMemcpyAsync(pGpuDesc, pCpuDesc, sizeof(nrldpc::crc_descriptor), hipMemcpyHostToDevice, strm) != hipSuccess)
        return CUPHY_STATUS_MEMCPY_ERROR;
    pCbCrcLaunchCfg->desc = pGpuDesc;
    pCbCrcLaunchCfg->kernelArgs[0] = &pCbCrcLaunchCfg->desc;
    set_node(pCbCrcLaunchCfg->kernelNodeParamsDriver, nrldpc::crc_cb_kernel_ptr(),
             dim3(std::max(1u, cmax), std::max<unsigned>(1, nSchUes), 1), dim3(nrldpc::CRC_CB_BLOCK),
             pCbCrcLaunchCfg->kernelArgs);
    pTbCrcLaunchCfg->desc = pGpuDesc;
    pTbCrcLaunchCfg->kernelArgs[0] = &pTbCrcLaunchCfg->desc;
    set_node(pTbCrcLaunchCfg->kernelNodeParamsDriver, nrldpc::crc_tb_kernel_ptr(), dim3(std::max<unsigned>(1, nSchUes)),
             dim3(nrldpc::CRC_TB_BLOCK), pTbCrcLaunchCfg->kernelArgs);
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyDestroyPuschRxCrcDecode(cuphyPuschRxCrcDecodeHndl_t puschRxCrcDecodeHndl)
{
    if (!puschRxCrcDecodeHndl) return CUPHY_STATUS_INVALID_ARGUMENT;
    delete static_cast<crc_object*>(puschRxCrcDecodeHndl);
    return CUPHY_STATUS_SUCCESS;
}

// ---------------------------------------------------------------------------
// extensions
// ---------------------------------------------------------------------------
cuphyStatus_t CUPHYWINAPI cuphyHipLDPCEncodeCodeBlocks(int BG, int Z, const uint8_t* info, int info_stride,
                                                       int num_code_blocks, uint8_t* cw, int cw_stride,
                                                       hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphyHipLDPCEncodeCodeBlocks");
    return nrldpc::encode_cbs(BG, Z, info, info_stride, num_code_blocks, cw, cw_stride, strm);
}

cuphyStatus_t CUPHYWINAPI cuphyHipLDPCBatchDecodeGetWorkspaceSize(size_t* sizeInBytes)
{
    if (!sizeInBytes) return CUPHY_STATUS_INVALID_ARGUMENT;
    *sizeInBytes = 256;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyHipLDPCWorkItemCheck(const cuphyHipLDPCWorkItem_t* item, int max_parity_nodes,
                                                    uint32_t* reasons)
{
    if (!item || !reasons) return CUPHY_STATUS_INVALID_ARGUMENT;
    *reasons = nrldpc::work_item_check(*item, max_parity_nodes);
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyHipLDPCBatchDecode(cuphyLDPCDecoder_t decoder, cuphyHipLDPCWorkItem_t* d_items,
                                                  int num_items, cuphyDataType_t llr_type, int fp16_math,
                                                  int max_values_per_cb, int max_parity_nodes, void* d_workspace,
                                                  hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphyHipLDPCBatchDecode");
    if (!decoder || !d_workspace || max_parity_nodes < 4 || max_parity_nodes > 46) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (max_values_per_cb <= 0 || max_values_per_cb > 68 * 384) return CUPHY_STATUS_INVALID_ARGUMENT;
    return nrldpc::decode_batch(d_items, nullptr, num_items, llr_type, fp16_math != 0, max_values_per_cb,
                                max_parity_nodes, 384, static_cast<uint32_t*>(d_workspace), decoder->cu_count, strm);
}

cuphyStatus_t CUPHYWINAPI cuphyHipLDPCBatchDecodeGroups(cuphyLDPCDecoder_t decoder,
                                                        cuphyHipLDPCWorkItem_t* d_items,
                                                        const uint32_t* d_group_first, int num_groups,
                                                        cuphyDataType_t llr_type, int fp16_math,
                                                        int max_values_per_group, int max_parity_nodes,
                                                        int block_threads, void* d_workspace, hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphyHipLDPCBatchDecodeGroups");
    if (!decoder || !d_workspace || max_parity_nodes < 4 || max_parity_nodes > 46)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    return nrldpc::decode_batch(d_items, d_group_first, num_groups, llr_type, fp16_math != 0, max_values_per_group,
                                max_parity_nodes, block_threads, static_cast<uint32_t*>(d_workspace),
                                decoder->cu_count, strm);
}

cuphyStatus_t CUPHYWINAPI cuphyHipPuschSlotDecodeGetWorkspaceSize(int num_cbs, size_t* sizeInBytes)
{
    if (!sizeInBytes || num_cbs < 0) return CUPHY_STATUS_INVALID_ARGUMENT;
    *sizeInBytes = (size_t)(nrldpc::FUSED_WS_QUEUE_WORDS + 2 * (size_t)num_cbs) * sizeof(uint32_t);
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyHipPuschSlotDecode(cuphyLDPCDecoder_t decoder, const cuphyHipPuschSlotTb_t* d_tbs,
                                                  int num_tbs, const uint32_t* d_cb_items, int num_cb_items,
                                                  cuphyDataType_t llr_type, int max_lds_values, int max_parity_nodes,
                                                  int max_cbs_per_item, void* d_workspace, hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphyHipPuschSlotDecode");
    if (!decoder) return CUPHY_STATUS_INVALID_ARGUMENT;
    return nrldpc::slot_decode(d_tbs, num_tbs, d_cb_items, num_cb_items, llr_type, max_lds_values, max_parity_nodes,
                               max_cbs_per_item, d_workspace, decoder->cu_count, strm);
}

cuphyStatus_t CUPHYWINAPI cuphyHipPuschSlotDecodeNoAlloc(cuphyLDPCDecoder_t decoder, const cuphyHipPuschSlotTb_t* d_tbs,
                                                         int num_tbs, const uint32_t* d_cb_items, int num_cb_items,
                                                         int num_cbs, cuphyDataType_t llr_type, int max_lds_values,
                                                         int max_parity_nodes, int max_cbs_per_item, hipStream_t strm)
{
    CUPHY_HIP_RANGE("cuphyHipPuschSlotDecodeNoAlloc");
    if (!decoder) return CUPHY_STATUS_INVALID_ARGUMENT;
    return nrldpc::slot_decode_noalloc(d_tbs, num_tbs, d_cb_items, num_cb_items, num_cbs, llr_type, max_lds_values,
                                       max_parity_nodes, max_cbs_per_item, decoder->cu_count, strm);
}

}  // extern "C"
