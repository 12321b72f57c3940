/*
 * cuphy_hip.h -- C ABI of the MI355X (gfx950) uplink LDPC decode path.
 *
 * Drop-in for the LDPC decoder, PUSCH de-rate-match and PUSCH CRC-decode
 * sections of the reference's cuPHY C API.  Type names, field order, enum
 * values and entry-point names are those of cuPHY/src/cuphy/cuphy.h so that
 * a caller written against that header (PuschRx, the pyAerial bindings,
 * cuphy_ex_ldpc) compiles against this one; the only change a caller sees is
 * the stream type (hipStream_t instead of cudaStream_t) and the launch
 * descriptor (hipKernelNodeParams instead of CUDA_KERNEL_NODE_PARAMS).
 *
 * Every entry point is asynchronous on the given stream, never synchronises
 * the host, and returns a cuphyStatus_t.
 */
#ifndef CUPHY_HIP_H
#define CUPHY_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUPHYWINAPI

/* ------------------------------------------------------------------------
 * Status codes -- cuphy.h:346-363 (values identical)
 * ---------------------------------------------------------------------- */
typedef enum {
    CUPHY_STATUS_SUCCESS = 0,
    CUPHY_STATUS_INTERNAL_ERROR = 1,
    CUPHY_STATUS_NOT_SUPPORTED = 2,
    CUPHY_STATUS_INVALID_ARGUMENT = 3,
    CUPHY_STATUS_ARCH_MISMATCH = 4,
    CUPHY_STATUS_ALLOC_FAILED = 5,
    CUPHY_STATUS_SIZE_MISMATCH = 6,
    CUPHY_STATUS_MEMCPY_ERROR = 7,
    CUPHY_STATUS_INVALID_CONVERSION = 8,
    CUPHY_STATUS_UNSUPPORTED_TYPE = 9,
    CUPHY_STATUS_UNSUPPORTED_LAYOUT = 10,
    CUPHY_STATUS_UNSUPPORTED_RANK = 11,
    CUPHY_STATUS_UNSUPPORTED_CONFIG = 12,
    CUPHY_STATUS_UNSUPPORTED_ALIGNMENT = 13,
    CUPHY_STATUS_VALUE_OUT_OF_RANGE = 14,
    CUPHY_STATUS_REF_MISMATCH = 15,
    CUPHY_N_STATUS_CONFIGS = 16
} cuphyStatus_t;

/* ------------------------------------------------------------------------
 * Data types -- cuphy.h:960-983 (values of the CUDA/HIP R types kept)
 * ---------------------------------------------------------------------- */
typedef enum {
    CUPHY_VOID = -1,
    CUPHY_R_32F = 0,
    CUPHY_R_64F = 1,
    CUPHY_R_16F = 2,
    CUPHY_R_8I = 3,
    CUPHY_C_16F = 6,
    CUPHY_R_8U = 8,
    CUPHY_R_32I = 10,
    CUPHY_R_32U = 12,
    CUPHY_BIT = 20,
    CUPHY_R_16I = 21,
    CUPHY_R_16U = 23
} cuphyDataType_t;

/* cuphy.h:43-62,118,143-145 */
#define MAX_DECODED_CODE_BLOCK_BIT_SIZE 8448
#define MAX_ENCODED_CODE_BLOCK_BIT_SIZE 25344
#define MAX_N_LAYERS_PUSCH 8
#define MAX_N_TBS_PER_CELL_GROUP_SUPPORTED 256
#define MAX_N_TBS_SUPPORTED (MAX_N_TBS_PER_CELL_GROUP_SUPPORTED)
#define MAX_N_CBS_PER_TB_SUPPORTED 152
#define CUPHY_LDPC_DECODE_DESC_MAX_TB (32)
#define QAM_STRIDE 8
#define CUPHY_MAX_N_CSI1_PRMS (4)
#define CUPHY_MAX_N_CSI2_REPORTS_PER_UE (16)

/* LDPC decoder flags -- cuphy.h:1128-1144.  EARLY_TERM is implemented
 * here (syndrome check over all mb*Z parity equations after each
 * iteration); in the reference it is documented as unsupported. */
#define CUPHY_LDPC_DECODE_DEFAULT (0)
#define CUPHY_LDPC_DECODE_EARLY_TERM (0x01)
#define CUPHY_LDPC_DECODE_CHOOSE_THROUGHPUT (0x2)
#define CUPHY_LDPC_DECODE_WRITE_SOFT_OUTPUTS (0x04)

/* cuphySetTensorDescriptor() flags -- cuphy.h:1091-1097 */
#define CUPHY_TENSOR_N_DIM_1 (1)
#define CUPHY_TENSOR_N_DIM_2 (2)
#define CUPHY_TENSOR_N_DIM_3 (3)
#define CUPHY_TENSOR_N_DIM_4 (4)
#define CUPHY_TENSOR_N_DIM_5 (5)
#define CUPHY_DIM_MAX CUPHY_TENSOR_N_DIM_5
#define CUPHY_TENSOR_ALIGN_DEFAULT 0x00
#define CUPHY_TENSOR_ALIGN_TIGHT 0x01
#define CUPHY_TENSOR_ALIGN_COALESCE 0x02

/* ------------------------------------------------------------------------
 * Context and errors -- cuphy.h:1056-1075, 1156-1208
 * ---------------------------------------------------------------------- */
struct cuphyContext;
typedef struct cuphyContext* cuphyContext_t;

const char* CUPHYWINAPI cuphyGetErrorString(cuphyStatus_t status);   /* cuphy.h:1060 */
const char* CUPHYWINAPI cuphyGetErrorName(cuphyStatus_t status);     /* cuphy.h:1075 */
cuphyStatus_t CUPHYWINAPI cuphyCreateContext(cuphyContext_t* pcontext, unsigned int flags); /* cuphy.h:1186 */
cuphyStatus_t CUPHYWINAPI cuphyDestroyContext(cuphyContext_t ctx);   /* cuphy.h:1208 */

/* ------------------------------------------------------------------------
 * Tensor descriptors -- cuphy.h:1225-1470 (only the 2-D semantics that the
 * LDPC tensor interface uses are meaningful)
 * ---------------------------------------------------------------------- */
struct cuphyTensorDescriptor;
typedef struct cuphyTensorDescriptor* cuphyTensorDescriptor_t;

typedef struct _cuphyTensorPrm {
    cuphyTensorDescriptor_t desc;
    void* pAddr;
} cuphyTensorPrm_t;

cuphyStatus_t CUPHYWINAPI cuphyCreateTensorDescriptor(cuphyTensorDescriptor_t* ptensorDesc);      /* cuphy.h:1318 */
cuphyStatus_t CUPHYWINAPI cuphyDestroyTensorDescriptor(cuphyTensorDescriptor_t tensorDesc);       /* cuphy.h:1338 */
cuphyStatus_t CUPHYWINAPI cuphySetTensorDescriptor(cuphyTensorDescriptor_t tensorDesc,             /* cuphy.h:1403 */
                                                   cuphyDataType_t type, int numDimensions,
                                                   const int dimensions[], const int strides[],
                                                   unsigned int flags);
cuphyStatus_t CUPHYWINAPI cuphyGetTensorDescriptor(cuphyTensorDescriptor_t tensorDesc,             /* cuphy.h:1432 */
                                                   int numDimsRequested, cuphyDataType_t* dataType,
                                                   int* numDims, int dimensions[], int strides[]);
cuphyStatus_t CUPHYWINAPI cuphyGetTensorSizeInBytes(cuphyTensorDescriptor_t tensorDesc, size_t* psz); /* cuphy.h:1457 */

/* ------------------------------------------------------------------------
 * LDPC decoder -- cuphy.h:3858-4258
 * ---------------------------------------------------------------------- */
struct cuphyLDPCDecoder;
typedef struct cuphyLDPCDecoder* cuphyLDPCDecoder_t;

typedef struct {                 /* cuphy.h:3916-3923 */
    void* addr;
    int32_t stride_elements;
    int32_t num_codewords;
} cuphyTransportBlockLLRDesc_t;

typedef struct {                 /* cuphy.h:3926-3933 */
    uint32_t* addr;
    int32_t stride_words;
    int32_t num_codewords;
} cuphyTransportBlockDataDesc_t;

typedef union {                  /* cuphy.h:3936-3942; f16x2 = two IEEE binary16 bit patterns */
    float f32;
    uint32_t f16x2;
} cuphyLDPCNormalization_t;

typedef struct {                 /* cuphy.h:3945-3960 */
    cuphyDataType_t llr_type;    /* CUPHY_R_16F or CUPHY_R_32F (CUPHY_R_8I: int8 LLRs, scale in clamp_value) */
    int16_t num_parity_nodes;    /* mb */
    int16_t Z;
    int16_t max_iterations;
    int16_t Kb;
    cuphyLDPCNormalization_t norm;
    uint32_t flags;
    int16_t BG;
    int16_t algo;                /* 0 = automatic */
    float clamp_value;
    void* workspace;
} cuphyLDPCDecodeConfigDesc_t;

typedef struct {                 /* cuphy.h:3963-3972 */
    cuphyLDPCDecodeConfigDesc_t config;
    int32_t num_tbs;
    cuphyTransportBlockLLRDesc_t llr_input[CUPHY_LDPC_DECODE_DESC_MAX_TB];
    cuphyTransportBlockDataDesc_t tb_output[CUPHY_LDPC_DECODE_DESC_MAX_TB];
    cuphyTransportBlockLLRDesc_t llr_output[CUPHY_LDPC_DECODE_DESC_MAX_TB];
} cuphyLDPCDecodeDesc_t;

typedef struct {                 /* cuphy.h:3975-3982; hipKernelNodeParams replaces CUDA_KERNEL_NODE_PARAMS */
    hipKernelNodeParams kernel_node_params_driver;
    void* kernel_args[2];
    cuphyLDPCDecodeDesc_t decode_desc;
} cuphyLDPCDecodeLaunchConfig_t;

cuphyStatus_t CUPHYWINAPI cuphyCreateLDPCDecoder(cuphyContext_t context, cuphyLDPCDecoder_t* pdecoder, /* cuphy.h:3889 */
                                                 unsigned int flags);
cuphyStatus_t CUPHYWINAPI cuphyDestroyLDPCDecoder(cuphyLDPCDecoder_t decoder);                         /* cuphy.h:3913 */

cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCDecode(cuphyLDPCDecoder_t decoder,                   /* cuphy.h:4077 */
                                                         cuphyTensorDescriptor_t tensorDescDst, void* dstAddr,
                                                         cuphyTensorDescriptor_t tensorDescLLR, const void* LLRAddr,
                                                         cuphyTensorDescriptor_t tensorDescSoftOutputs,
                                                         void* softOutputsAddr,
                                                         const cuphyLDPCDecodeConfigDesc_t* config,
                                                         hipStream_t strm);

cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCTransportBlockDecode(cuphyLDPCDecoder_t decoder,     /* cuphy.h:4147 */
                                                                       const cuphyLDPCDecodeDesc_t* decodeDesc,
                                                                       hipStream_t strm);

cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCDecodeGetWorkspaceSize(cuphyLDPCDecoder_t decoder,  /* cuphy.h:4193 */
                                                                         const cuphyLDPCDecodeConfigDesc_t* config,
                                                                         int numCodeWords, size_t* sizeInBytes);

cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCDecodeSetNormalization(cuphyLDPCDecoder_t decoder,  /* cuphy.h:4228 */
                                                                         cuphyLDPCDecodeConfigDesc_t* decodeDesc);

cuphyStatus_t CUPHYWINAPI cuphyErrorCorrectionLDPCDecodeGetLaunchDescriptor(cuphyLDPCDecoder_t decoder, /* cuphy.h:4257 */
                                                                            cuphyLDPCDecodeLaunchConfig_t* launchConfig);

/* ------------------------------------------------------------------------
 * Per-TB parameters -- cuphy.h:3506-3575 (layout identical; the three
 * __half* members are declared void*)
 * ---------------------------------------------------------------------- */
typedef struct _cuphyCalcCsi2SizePrm {   /* cuphy.h:811-818 */
    uint8_t nPart1Prms;
    uint8_t prmSizes[CUPHY_MAX_N_CSI1_PRMS];
    uint16_t prmOffsets[CUPHY_MAX_N_CSI1_PRMS];
    uint16_t csi2sizeMapIdx;
} cuphyCalcCsi2SizePrm_t;

typedef struct PerTbParams {
    uint32_t ndi;
    uint32_t rv;
    uint32_t Qm;
    uint32_t bg;
    uint32_t Nl;
    uint32_t num_CBs;
    uint32_t Zc;
    uint32_t N;
    uint32_t Ncb;
    uint32_t Ncb_padded;
    uint32_t G;
    uint32_t K;
    uint32_t F;
    uint32_t cinit;
    uint32_t nDataBytes;
    uint32_t nZpBitsPerCb;
    uint32_t k0;
    uint8_t ldpcEarlyTerminationPerUe;
    uint8_t ldpcMaxNumItrPerUe;
    uint32_t firstCodeBlockIndex;
    uint32_t encodedSize;
    uint32_t layer_map_array[MAX_N_LAYERS_PUSCH];
    uint32_t userGroupIndex;
    uint32_t nBBULayers;
    uint32_t startLLR;
    uint32_t mScUciSum;
    uint32_t codedBitsSum;
    uint8_t isDataPresent;
    uint8_t betaOffsetCsi2;
    uint32_t qPrimeAck;
    float codeRate;
    float alpha;
    uint32_t qPrimeCsi1;
    uint8_t isEarlyHarq;
    uint8_t uciOnPuschFlag;
    uint8_t csi2Flag;
    uint32_t G_schAndCsi2;
    uint32_t G_harq;
    uint32_t G_csi1;
    uint32_t G_csi2;
    uint32_t G_harq_rvd;
    uint32_t nBitsHarq;
    uint16_t nBitsCsi2;
    uint8_t nCsiReports;
    uint8_t rankBitOffset;
    uint8_t nRanksBits;
    void* d_schAndCsi2LLRs;
    void* d_csi1LLRs;
    void* d_harqLLrs;
    uint32_t tbSize;
    uint16_t nCsi2Reports;
    cuphyCalcCsi2SizePrm_t calcCsi2SizePrms[CUPHY_MAX_N_CSI2_REPORTS_PER_UE];
    uint8_t enableTfPrcd;
    uint8_t nDmrsCdmGrpsNoData;
    uint32_t* debug_d_derateCbsIndices;
} PerTbParams;

/* ------------------------------------------------------------------------
 * PUSCH de-rate-matching + HARQ combining -- cuphy.h:3349-3490
 * ---------------------------------------------------------------------- */
struct cuphyPuschRxRateMatch { uint8_t reserved; };
typedef struct cuphyPuschRxRateMatch* cuphyPuschRxRateMatchHndl_t;

typedef struct {                 /* cuphy.h:3359-3366 */
    hipKernelNodeParams kernelNodeParamsDriver;       /* main de-rate-match kernel      */
    hipKernelNodeParams resetKernelNodeParamsDriver;  /* unused here (no atomics): gridDim 0 */
    hipKernelNodeParams clampKernelNodeParamsDriver;  /* unused here (no atomics): gridDim 0 */
    void* desc;
    void* kernelArgs[1];
} cuphyPuschRxRateMatchLaunchCfg_t;

cuphyStatus_t CUPHYWINAPI cuphyPuschRxRateMatchGetDescrInfo(size_t* pDescrSizeBytes, size_t* pDescrAlignBytes); /* cuphy.h:3392 */
cuphyStatus_t CUPHYWINAPI cuphyCreatePuschRxRateMatch(cuphyPuschRxRateMatchHndl_t* pPuschRxRateMatchHndl,      /* cuphy.h:3419 */
                                                      int FPconfig, int descramblingOn);
cuphyStatus_t CUPHYWINAPI cuphySetupPuschRxRateMatch(cuphyPuschRxRateMatchHndl_t puschRxRateMatchHndl,         /* cuphy.h:3460 */
                                                     uint16_t nSchUes, uint16_t* pSchUserIdxsCpu,
                                                     const PerTbParams* pTbPrmsCpu, const PerTbParams* pTbPrmsGpu,
                                                     cuphyTensorPrm_t* pTPrmRmIn, cuphyTensorPrm_t* pTPrmCdm1RmIn,
                                                     void** ppRmOut, void* pCpuDesc, void* pGpuDesc,
                                                     uint8_t enableCpuToGpuDescrAsyncCpy,
                                                     cuphyPuschRxRateMatchLaunchCfg_t* pLaunchCfg, hipStream_t strm);
cuphyStatus_t CUPHYWINAPI cuphyDestroyPuschRxRateMatch(cuphyPuschRxRateMatchHndl_t puschRxRateMatchHndl);     /* cuphy.h:3490 */

/* ------------------------------------------------------------------------
 * PUSCH CB/TB CRC decode -- cuphy.h:3154-3293
 * ---------------------------------------------------------------------- */
struct cuphyPuschRxCrcDecode { uint8_t reserved; };
typedef struct cuphyPuschRxCrcDecode* cuphyPuschRxCrcDecodeHndl_t;

typedef struct {                 /* cuphy.h:3168-3173 */
    hipKernelNodeParams kernelNodeParamsDriver;
    void* desc;
    void* kernelArgs[1];
} cuphyPuschRxCrcDecodeLaunchCfg_t;

cuphyStatus_t CUPHYWINAPI cuphyPuschRxCrcDecodeGetDescrInfo(size_t* pDescrSizeBytes, size_t* pDescrAlignBytes); /* cuphy.h:3195 */
cuphyStatus_t CUPHYWINAPI cuphyCreatePuschRxCrcDecode(cuphyPuschRxCrcDecodeHndl_t* puschRxCrcDecodeHndl,        /* cuphy.h:3217 */
                                                      int reverseBytes);
cuphyStatus_t CUPHYWINAPI cuphySetupPuschRxCrcDecode(cuphyPuschRxCrcDecodeHndl_t puschRxCrcDecodeHndl,          /* cuphy.h:3254 */
                                                     uint16_t nSchUes, uint16_t* pSchUserIdxsCpu,
                                                     uint32_t* pOutputCBCRCs, uint8_t* pOutputTBs,
                                                     const uint32_t* pInputCodeBlocks, uint32_t* pOutputTBCRCs,
                                                     const PerTbParams* pTbPrmsCpu, const PerTbParams* pTbPrmsGpu,
                                                     void* pCpuDesc, void* pGpuDesc,
                                                     uint8_t enableCpuToGpuDescrAsyncCpy,
                                                     cuphyPuschRxCrcDecodeLaunchCfg_t* pCbCrcLaunchCfg,
                                                     cuphyPuschRxCrcDecodeLaunchCfg_t* pTbCrcLaunchCfg,
                                                     hipStream_t strm);
cuphyStatus_t CUPHYWINAPI cuphyDestroyPuschRxCrcDecode(cuphyPuschRxCrcDecodeHndl_t puschRxCrcDecodeHndl);        /* cuphy.h:3293 */

/* ------------------------------------------------------------------------
 * Extensions of this library (no cuphy.h counterpart with this signature)
 * ---------------------------------------------------------------------- */

/* Systematic TS 38.212 LDPC encoding of ncb code blocks on the device: the
 * test-signal source of the decode path (the reference's DL encoder API,
 * cuphy.h:3781-3850, is a different, tensor-based interface).  info: ncb rows
 * of nsys*Z bytes (0/1, filler bits 0), row stride info_stride bytes;
 * cw: ncb rows of (nsys + nrows)*Z bytes, row stride cw_stride bytes, first
 * 2Z (punctured) columns included. */
cuphyStatus_t CUPHYWINAPI cuphyHipLDPCEncodeCodeBlocks(int BG, int Z, const uint8_t* info, int info_stride,
                                                       int num_code_blocks, uint8_t* cw, int cw_stride,
                                                       hipStream_t strm);

/* One code block of a heterogeneous batch (any BG, Z, mb per item).  The
 * batch form has no cuphy.h counterpart: the reference groups TBs of equal
 * (BG, Z, mb) into descriptors of <= 32 TBs (pusch_rx.cpp:6275-6388) and
 * launches one kernel per group. */
typedef struct {
    const void* llr;                 /* (nsys + mb) * Z LLRs of the batch's llr_type, 2Z punctured first */
    uint32_t* out;                   /* Kb * Z hard bits, LSB-first 32-bit words */
    void* soft;                      /* optional Kb * Z binary16 soft outputs, NULL for none */
    int16_t BG;
    int16_t Z;
    int16_t num_parity_nodes;        /* mb */
    int16_t Kb;
    int16_t max_iterations;
    uint16_t flags;                  /* CUPHY_LDPC_DECODE_EARLY_TERM */
    cuphyLDPCNormalization_t norm;   /* f32, or f16x2 when the batch computes in binary16 */
    float clamp_value;
    uint32_t status;                 /* written by the launch: 0 = decoded; otherwise CUPHY_HIP_ITEM_* reasons,
                                        and the item's outputs are left untouched */
} cuphyHipLDPCWorkItem_t;

/* Reasons a batch launch does not decode a work item (cuphyHipLDPCWorkItem_t::status),
 * the device-side counterpart of the argument checks of cuphy_ldpc.cpp:36-91. */
#define CUPHY_HIP_ITEM_BAD_BG        0x01u  /* BG not 1 or 2 */
#define CUPHY_HIP_ITEM_BAD_Z         0x02u  /* Z not one of the 51 TS 38.212 lifting sizes */
#define CUPHY_HIP_ITEM_BAD_MB        0x04u  /* num_parity_nodes outside 4..46 (BG2: 4..42), or above 8 in a launch
                                               whose max_parity_nodes <= 8 selects the high-code-rate kernel */
#define CUPHY_HIP_ITEM_BAD_KB        0x08u  /* Kb not 22 (BG1) or 6, 8, 9, 10 (BG2) */
#define CUPHY_HIP_ITEM_BAD_ITER      0x10u  /* max_iterations < 0 */
#define CUPHY_HIP_ITEM_NULL_BUFFER   0x20u  /* llr or out is NULL */
#define CUPHY_HIP_ITEM_NO_FIT        0x40u  /* the item's values or lanes exceed what the launch provides */
#define CUPHY_HIP_ITEM_GROUP         0x80u  /* not decoded with its group: an earlier item of the group is invalid
                                               or has another configuration (BG, Z, mb, Kb, iterations, flags,
                                               normalisation, clamp) */

/* Host-side check of one work item with the same predicate the batch
 * kernels apply (every reason except NO_FIT and GROUP, which depend on the
 * launch): *reasons = 0 for an item a launch with max_parity_nodes decodes. */
cuphyStatus_t CUPHYWINAPI cuphyHipLDPCWorkItemCheck(const cuphyHipLDPCWorkItem_t* item, int max_parity_nodes,
                                                    uint32_t* reasons);

/* Workspace of cuphyHipLDPCBatchDecode: the device work-queue state.  The
 * caller zero-fills it once; every completed launch leaves it zeroed, so
 * launches can be replayed from a captured graph.  One workspace per
 * concurrently running batch. */
cuphyStatus_t CUPHYWINAPI cuphyHipLDPCBatchDecodeGetWorkspaceSize(size_t* sizeInBytes);

/* Decode num_items code blocks (device array d_items) with one persistent
 * launch: each workgroup takes the next item from the device queue until
 * the batch is drained.  llr_type: CUPHY_R_32F or CUPHY_R_16F for every
 * item; fp16_math: binary16 arithmetic (always for CUPHY_R_16F);
 * max_values_per_cb: max over items of (nsys + mb) * Z (sizes LDS);
 * max_parity_nodes: max over items of mb (<= 8 selects the register-lean
 * high-code-rate variant).  Order the items longest first for the shortest
 * makespan.  Each item's status word receives 0 or the reasons it was not
 * decoded (CUPHY_HIP_ITEM_*); the items are otherwise only read. */
cuphyStatus_t CUPHYWINAPI cuphyHipLDPCBatchDecode(cuphyLDPCDecoder_t decoder, cuphyHipLDPCWorkItem_t* d_items,
                                                  int num_items, cuphyDataType_t llr_type, int fp16_math,
                                                  int max_values_per_cb, int max_parity_nodes, void* d_workspace,
                                                  hipStream_t strm);

/* Grouped form of cuphyHipLDPCBatchDecode: the device queue hands out groups
 * of items, group g being items [d_group_first[g], d_group_first[g + 1])
 * (d_group_first: num_groups + 1 device words; null: one item per group,
 * num_groups items, decoded with block_threads >= roundup64(Z) threads).
 * The items of a group share
 * BG, Z, num_parity_nodes, Kb, max_iterations, flags, norm and clamp_value
 * and are decoded side by side by one workgroup of block_threads threads (a
 * multiple of 64, <= 384, >= count * Z), lanes [c*Z, (c+1)*Z) on item c, each
 * with its own early termination; results equal the items decoded one by
 * one.  max_values_per_group: max over groups of count * roundup8((nsys + mb)
 * * Z) (sizes LDS).  The reference packs small-Z code blocks of one TB per
 * CTA (ldpc2_reg_index_fp_desc_dyn_small.cu:115-618); here any configuration
 * can be packed, with the group size chosen per (BG, Z) by the tuning tables
 * of the host layer.  block_threads = 768 selects row-pair mode instead:
 * every item is decoded alone (a group's items one after another) by two
 * lane groups of a 768-thread workgroup that take the column-disjoint block
 * rows of each barrier segment at the same time (12 waves per code block);
 * for batches of fewer code blocks than CUs, whose time is that of the
 * slowest code block.  Same workspace contract as cuphyHipLDPCBatchDecode. */
cuphyStatus_t CUPHYWINAPI cuphyHipLDPCBatchDecodeGroups(cuphyLDPCDecoder_t decoder,
                                                        cuphyHipLDPCWorkItem_t* d_items,
                                                        const uint32_t* d_group_first, int num_groups,
                                                        cuphyDataType_t llr_type, int fp16_math,
                                                        int max_values_per_group, int max_parity_nodes,
                                                        int block_threads, void* d_workspace, hipStream_t strm);

/* One transport block (one transmission) of a fused slot decode.  The
 * reference runs de-rate-matching, LDPC decoding and CRC checking as three
 * kernels per TB group (pusch_rx.cpp:3007-3021); cuphyHipPuschSlotDecode
 * runs them fused, code block by code block, in one persistent launch. */
typedef struct {
    const PerTbParams* prm;          /* device pointer: the TB's parameters, as for cuphySetupPuschRxRateMatch */
    const void* llr;                 /* this transmission's LLRs in the rate-match input layout (QAM_STRIDE
                                        slots per symbol), element type = llr_type of the launch */
    void* harq;                      /* binary16 HARQ buffers: CB r at r * Ncb_padded values, [2Zc | Ncb];
                                        ndi = 1 overwrites, ndi = 0 combines */
    uint32_t* cb_bits;               /* decoded bits: CB r at r * 264 words, bit i of word w = bit 32w + i */
    uint32_t* cb_crc;                /* CRC remainder of CB r (0 = pass; 0xFFFFFFFF = parameters rejected) */
    uint8_t* tb_bytes;               /* TB payload bytes, MSB-first per byte, num_CBs * data bytes, padded to 4 */
    uint32_t* tb_crc;                /* TB CRC remainder (0 = pass; 0xFFFFFFFF = parameters rejected) */
    uint8_t* cb_iterations;          /* optional: iterations run per CB r, NULL for none */
    int16_t num_parity_nodes;        /* mb of the decoder, 4..46 (BG2: ..42) */
    int16_t max_iterations;
    uint16_t flags;                  /* CUPHY_LDPC_DECODE_EARLY_TERM | CUPHY_HIP_SLOT_TB_NOTIFY |
                                        CUPHY_HIP_SLOT_TB_DESCRAMBLE */
    uint16_t reserved;
    float llr_scale;                 /* combined LLR = stored value * llr_scale, rounded to binary16
                                        (int8 LLRs with 4 steps per unit LLR: 0.25) */
    cuphyLDPCNormalization_t norm;   /* f16x2: binary16 normalisation */
    float clamp_value;               /* decoder input clamp */
    uint32_t share_offset;           /* first of the TB's num_CBs TB-CRC share slots in the workspace */
} cuphyHipPuschSlotTb_t;

/* cuphyHipPuschSlotTb_t.flags: publish the TB CRC as soon as the TB is decoded
 * (see cuphyHipPuschSlotDecode) */
#define CUPHY_HIP_SLOT_TB_NOTIFY 0x200u
/* cuphyHipPuschSlotTb_t.flags: descramble the LLRs with the TS 38.211 6.3.1.1
 * Gold sequence of c_init = PerTbParams.cinit before combining (the
 * reference's descramblingOn of cuphySetupPuschRxRateMatch,
 * rate_matching.cu:276, 378-387; not for UCI-multiplexed TBs).  The CB's
 * sequence words are held in LDS after its HARQ/APP values: max_lds_values
 * must cover, per descrambled CB, align2(max(2Zc + Ncb, (nsys + mb) * Zc)) +
 * 2 * (words + 1) with words = (start + E - 1) / 32 - start / 32 + 1 (start:
 * index of the CB's first rate-matched bit in the TB); a CB that does not
 * fit is rejected. */
#define CUPHY_HIP_SLOT_TB_DESCRAMBLE 0x400u

/* Workspace of cuphyHipPuschSlotDecode: queue state and num_cbs CB slots of
 * two words (TB t owns slots share_offset .. share_offset + num_CBs - 1 and
 * uses the pair at its first slot as TB-CRC accumulator and completion
 * count; the TBs of one launch must not overlap).  Zero-fill once; all of
 * it returns to zero after every launch (graph-replayable).  One workspace
 * per running launch. */
cuphyStatus_t CUPHYWINAPI cuphyHipPuschSlotDecodeGetWorkspaceSize(int num_cbs, size_t* sizeInBytes);

/* Fused de-rate-match + HARQ combine + layered NMS decode (binary16) + CB
 * CRC + TB assembly + TB CRC for num_cb_items work items in one persistent
 * launch.  d_cb_items[i] = (TB index << 16) | ((count - 1) << 12) | CB
 * index within the TB: the `count` (1..16) consecutive CBs from that index
 * on, decoded side by side in one 384-thread workgroup (count * Zc <= 384;
 * each CB stops on its own syndrome, results as decoded alone); count = 1
 * gives one CB per item, (TB index << 16) | CB index (CB index < 4096).
 * Every CB of a TB must be in the launch exactly once.  A TB's CRC
 * (*tb_crc) is written as soon as its last CB is decoded.  With
 * CUPHY_HIP_SLOT_TB_NOTIFY in the TB's flags it is written only after every
 * other output of the TB has reached device memory: a caller that points
 * tb_crc into host-mapped pinned memory, pre-set to a value no CRC takes
 * (e.g. 0xFFFFFFFE), can take each TB as soon as that value changes, before
 * the launch ends (the TB's HARQ buffer must still not be reused by another
 * launch before this one ends).  llr_type: CUPHY_R_8I,
 * CUPHY_R_16F or CUPHY_R_32F.  max_lds_values: max over items of
 * (count - 1) * roundup8(A) + A with A = max(2Zc + Ncb, (nsys + mb) * Zc),
 * plus the sequence words of descrambled TBs (see
 * CUPHY_HIP_SLOT_TB_DESCRAMBLE), at most 68 * 384 + 2 * 12288 (sizes LDS;
 * an item that does not fit is rejected); max_cbs_per_item: the largest
 * count of the items (1 runs the one-CB-per-workgroup kernel, which keeps
 * no group state; an item of several CBs is then rejected); max_parity_nodes: max mb
 * (<= 8 selects the register-lean high-code-rate variant, <= 16 the 16-row
 * variant, otherwise the full one; a CB with a larger mb is rejected). */
cuphyStatus_t CUPHYWINAPI cuphyHipPuschSlotDecode(cuphyLDPCDecoder_t decoder, const cuphyHipPuschSlotTb_t* d_tbs,
                                                  int num_tbs, const uint32_t* d_cb_items, int num_cb_items,
                                                  cuphyDataType_t llr_type, int max_lds_values, int max_parity_nodes,
                                                  int max_cbs_per_item, void* d_workspace, hipStream_t strm);

/* Allocation-free form of cuphyHipPuschSlotDecode.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * queue state and TB-CRC words live in a scratch array of the library of
 * 64 + 2 * 16384 words on each GPU (held until the process ends), cleared by
 * every call; num_cbs is what cuphyHipPuschSlotDecodeGetWorkspaceSize would
 * be given (share_offset + num_CBs of every TB <= num_cbs), at most 16384,
 * else CUPHY_STATUS_INVALID_ARGUMENT.  Calls on one GPU share the array:
 * calls that overlap in time on one GPU race (serialize them on one stream).
 * No hipMalloc, hipFree or allocating library call.  Other arguments and
 * results as cuphyHipPuschSlotDecode. */
cuphyStatus_t CUPHYWINAPI cuphyHipPuschSlotDecodeNoAlloc(cuphyLDPCDecoder_t decoder, const cuphyHipPuschSlotTb_t* d_tbs,
                                                         int num_tbs, const uint32_t* d_cb_items, int num_cb_items,
                                                         int num_cbs, cuphyDataType_t llr_type, int max_lds_values,
                                                         int max_parity_nodes, int max_cbs_per_item, hipStream_t strm);

/* ------------------------------------------------------------------------
 * HARQ serving engine (no cuphy.h counterpart: cuPHY is driven slot by slot
 * by its caller).  Transmissions of transport blocks are submitted as they
 * arrive; the engine keeps each HARQ process's soft buffer (binary16,
 * [2Zc | Ncb] per CB) in a device pool of a fixed byte budget, admits a new
 * TB only when its buffer fits (admission by HBM budget), and continuously
 * batches admitted transmissions into fused slot launches
 * (cuphyHipPuschSlotDecode), up to max_inflight launches running at once on
 * their own streams.  Single-threaded: every call from one host thread.
 * ---------------------------------------------------------------------- */
struct cuphyHipHarqServer;
typedef struct cuphyHipHarqServer* cuphyHipHarqServer_t;

#define CUPHY_HIP_HARQ_LAST_TX 0x100u   /* request flag: free the HARQ buffer after this transmission */

/* cuphyHipHarqServerConfig_t.flags: run one resident launch that takes code
 * blocks as they are admitted (see cuphyHipHarqServerCreate) */
#define CUPHY_HIP_HARQ_SERVER_RESIDENT 0x1u

typedef struct {
    size_t harq_pool_bytes;          /* device bytes for HARQ soft buffers */
    int max_inflight;                /* launches in flight at once, 1..8 (batched mode) */
    int max_tbs_per_launch;          /* 1..65536; resident mode: TBs in flight at once (<= 65535) */
    int max_cbs_per_launch;          /* >= 1; resident mode: code blocks in flight at once */
    cuphyDataType_t llr_type;        /* element type of every request's LLRs */
    uint32_t flags;                  /* 0 (batched launches) or CUPHY_HIP_HARQ_SERVER_RESIDENT */
    double harq_timeout_us;          /* > 0: free a HARQ buffer unused this long (the reference's
                                        HarqPoolManager timeout, harq_pool.hpp); 0: never */
    double resident_idle_us;         /* resident mode: the launch leaves this long after its last work
                                        (0: 20000) and is started again when work arrives */
} cuphyHipHarqServerConfig_t;

typedef struct {
    uint32_t cell_id;                /* HARQ process (cell, RNTI, process id); one TB at a time, as the */
    uint16_t rnti;                   /* reference's HarqPoolManager keys its buffers (harq_pool.hpp:389-545) */
    uint8_t harq_pid;
    uint8_t reserved0;
    uint64_t tag;                    /* returned in the completion */
    PerTbParams prm;                 /* this transmission's parameters (ndi, rv/k0, G, ...) */
    const void* llr;                 /* device: this transmission's LLRs, rate-match input layout */
    uint32_t* cb_bits;               /* device outputs as in cuphyHipPuschSlotTb_t */
    uint32_t* cb_crc;
    uint8_t* tb_bytes;
    uint8_t* cb_iterations;          /* optional, NULL for none */
    int16_t num_parity_nodes;        /* decoder mb: cover every position received so far */
    int16_t max_iterations;
    uint16_t flags;                  /* CUPHY_LDPC_DECODE_EARLY_TERM | CUPHY_HIP_HARQ_LAST_TX |
                                        CUPHY_HIP_SLOT_TB_DESCRAMBLE (c_init = prm.cinit; rejected when
                                        the CB's sequence words exceed the launch's LDS: resident mode
                                        E up to 65,473 bits per CB) */
    uint16_t reserved;
    float llr_scale;
    cuphyLDPCNormalization_t norm;   /* f16x2 */
    float clamp_value;
} cuphyHipHarqRequest_t;

typedef struct {
    uint64_t tag;
    uint32_t cell_id;                /* the request's HARQ process */
    uint16_t rnti;
    uint8_t harq_pid;
    uint8_t reserved0;
    uint32_t tb_crc;                 /* TB CRC remainder, 0 = pass */
    int32_t status;                  /* CUPHY_STATUS_SUCCESS, or why the request was rejected */
    double t_submit_us;              /* host steady clock, microseconds since server creation */
    double t_launch_us;
    double t_done_us;                /* when the engine observed the launch complete */
} cuphyHipHarqCompletion_t;

typedef struct {
    uint64_t submitted, completed, launches;
    uint64_t pool_bytes_in_use, pool_bytes_peak;
    uint64_t pending;                /* submitted, not yet launched */
    uint64_t admission_waits;        /* polls at which a new TB waited for pool space */
    uint64_t evictions;              /* HARQ buffers freed by harq_timeout_us */
    uint64_t pool_bytes_cached;      /* freed buffers kept in their size class for reuse */
    uint64_t pool_reclaims;          /* times the cached buffers went back to the pool to make room */
} cuphyHipHarqServerStats_t;

cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerCreate(cuphyLDPCDecoder_t decoder, const cuphyHipHarqServerConfig_t* cfg,
                                                   cuphyHipHarqServer_t* server);
cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerDestroy(cuphyHipHarqServer_t server);
/* Queue one transmission (copied); returns at once.  CUPHY_STATUS_INVALID_ARGUMENT
 * for parameters the engine cannot run (num_CBs above 4096 or max_cbs_per_launch,
 * a HARQ buffer larger than the pool, mb outside the base graph, ...). */
cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerSubmit(cuphyHipHarqServer_t server, const cuphyHipHarqRequest_t* req);
/* Non-blocking progress: retire finished launches (their completions go to
 * out[0 .. *n_out-1]; the rest wait for the next call), admit pending
 * requests, start launches on free streams. */
cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerPoll(cuphyHipHarqServer_t server, cuphyHipHarqCompletion_t* out,
                                                 int max_out, int* n_out);
/* Free a HARQ process's buffer (e.g. the MAC gave up on the TB). */
cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerRelease(cuphyHipHarqServer_t server, uint32_t cell_id, uint16_t rnti,
                                                    uint8_t harq_pid);
cuphyStatus_t CUPHYWINAPI cuphyHipHarqServerGetStats(cuphyHipHarqServer_t server, cuphyHipHarqServerStats_t* stats);

/* Load-test driver: run a trace of transmissions in real time through the
 * engine.  trace[i].arrival_us >= 0: a first transmission, released at that
 * time (microseconds from the call); arrival_us < 0: a retransmission,
 * released harq_rtt_us after its predecessor (the entry whose next == i)
 * completed with a failing TB CRC.  done[i] receives entry i's completion
 * (tag = i, status CUPHY_STATUS_NOT_SUPPORTED if never released);
 * *n_done = completions.  Returns when every released entry completed. */
typedef struct {
    cuphyHipHarqRequest_t req;
    double arrival_us;
    int32_t next;                    /* index of the TB's next transmission, -1: none */
    int32_t reserved;
} cuphyHipHarqTraceEntry_t;

cuphyStatus_t CUPHYWINAPI cuphyHipHarqServeTrace(cuphyHipHarqServer_t server, const cuphyHipHarqTraceEntry_t* trace,
                                                 int n, double harq_rtt_us, cuphyHipHarqCompletion_t* done,
                                                 int* n_done);

/* ------------------------------------------------------------------------
 * LDPC encoder (cuphy.h:3781-3850): bit tensors in, bit tensors out
 * ---------------------------------------------------------------------- */

/* Launch configuration of the encoder (cuphy.h:3781-3790; the kernel node
 * parameters are hipKernelNodeParams here). */
struct cuphyLDPCEncodeLaunchConfig {
    hipKernelNodeParams m_kernelNodeParams;
    void* m_desc;
    void* m_kernelArgs[1];
};
typedef struct cuphyLDPCEncodeLaunchConfig* cuphyLDPCEncodeLaunchConfig_t;

/* Descriptor size / alignment and workspace bytes for maxUes TBs
 * (cuphy.h:3799-3802). */
cuphyStatus_t CUPHYWINAPI cuphyLDPCEncodeGetDescrInfo(size_t* pDescrSizeBytes, size_t* pDescrAlignBytes,
                                                      uint16_t maxUes, size_t* pWorkspaceBytes);

/* Set up an encode launch (cuphy.h:3828-3847): inDesc CUPHY_BIT [K bits x C]
 * (K >= Kb * Z, Kb = 22 / 10, filler bits as the caller set them),
 * outDesc CUPHY_BIT [N bits x C], N >= (Kb + num_rows) * Z less 2 Z when
 * puncture; column strides multiples of 32 bits.  num_rows = maxParityNodes
 * when max_rv == 0 and 0 < maxParityNodes < 46 / 42, else every parity
 * row.  batching: inBatchedAddr / outBatchedAddr give batched_TBs TBs of
 * the descriptors' shape.  h_workspace / d_workspace hold the TBs' tensor
 * addresses (copied with the descriptor when enable_desc_async_copy);
 * launch m_kernelNodeParams on a stream or put it in a graph.  Errors as
 * the reference (ldpc_encode.cu:416-470). */
cuphyStatus_t CUPHYWINAPI cuphySetupLDPCEncode(cuphyLDPCEncodeLaunchConfig_t ldpcEncodeLaunchConfig,
                                               cuphyTensorDescriptor_t inDesc, void* inAddr,
                                               cuphyTensorDescriptor_t outDesc, void* outAddr, int BG, int Z,
                                               uint8_t puncture, int maxParityNodes, int max_rv, uint8_t batching,
                                               int batched_TBs, void** inBatchedAddr, void** outBatchedAddr,
                                               void* h_workspace, void* d_workspace, void* cpu_desc, void* gpu_desc,
                                               uint8_t enable_desc_async_copy, hipStream_t strm);

/* DL rate matching of encoded code blocks: for each TB t (d_tb_prms[t]:
 * bg, Zc, num_CBs, N = encoded bits per CB, Ncb, K, F, k0, Qm, Nl,
 * encodedSize = G, cinit) and CB r, bit selection from k0 around Ncb
 * skipping the filler bits, bit interleaving, and, when scrambling, XOR
 * with the TB's Gold sequence; CB r's E_r bits at bit (t * max_cbs + r) *
 * emax of d_out, the rest of its emax-bit slot zero.  d_cw: device array of
 * num_tbs device pointers, CB r of TB t at word r * ceil(N / 32) (the
 * encoder's punctured output).  The reference's cuphySetupDlRateMatching
 * (cuphy.h:3726-3752) with layer mapping and modulation disabled produces
 * this layout; that entry point also takes the PDSCH DMRS / UE-group
 * structures and is not reproduced.  emax: a multiple of 32, at most
 * 393,216; a CB whose E_r exceeds it (or N > 25,344) is left zero. */
cuphyStatus_t CUPHYWINAPI cuphyHipDlRateMatch(const PerTbParams* d_tb_prms, int num_tbs, const uint32_t* const* d_cw,
                                              int max_cbs, int emax, uint8_t scrambling, uint32_t* d_out,
                                              hipStream_t strm);

/* ------------------------------------------------------------------------
 * Per-TB parameter derivation (host only, no device call)
 * ---------------------------------------------------------------------- */

/* TS 38.214 5.1.3.2 transport block size and code-block count of an
 * allocation: num_symbols data symbols, num_prbs PRBs, num_layers layers,
 * code rate R (0 < R <= 1), modulation order Qm, num_dmrs_cdm_grps_no_data1_symbols
 * DMRS symbols counted as half a data symbol.  Replaces
 * get_TB_size_and_num_CBs (cuPHY/src/cuphy/utils.cuh:60-105) and pyAerial's
 * get_tb_size (pyaerial/src/aerial/phy5g/ldpc/util.py:171-333, code rate as a
 * fraction here).  INVALID_ARGUMENT on a non-positive size or rate. */
cuphyStatus_t CUPHYWINAPI cuphyHipTransportBlockSize(int num_symbols, int num_prbs, int num_layers, float code_rate,
                                                     uint32_t Qm, int num_dmrs_cdm_grps_no_data1_symbols,
                                                     uint32_t* tb_size, uint32_t* num_CBs);

/* Limited-buffer rate matching of a UE (TS 38.212 5.4.2.1, I_LBRM = 1):
 * the PUSCH UE parameters i_lbrm, n_PRB_LBRM, maxLayers, maxQm
 * (cuPHY/src/cuphy_channels/pusch_rx.cpp:8146-8164). */
typedef struct {
    uint8_t  i_lbrm;       /* 0: Ncb = N; 1: Ncb = min(N, Nref)            */
    uint8_t  maxLayers;    /* 1..4                                          */
    uint8_t  maxQm;        /* 6 or 8                                        */
    uint8_t  reserved;
    uint16_t n_PRB_LBRM;   /* 32..273                                       */
    uint16_t reserved2;
} cuphyHipLbrmPrm_t;

/* PerTbParams of one PUSCH TB from its size and code rate, with the decoder's
 * parity-node count and info-node count: base graph, CB segmentation (C, K',
 * Zc, K, F, TS 38.212 5.2.2 / 6.2.2), N, Ncb (limited buffer when
 * lbrm->i_lbrm), Ncb_padded, k0 of rv (Table 5.4.2.1-2), parity nodes
 * covering min(G/C + k0, Ncb) (rounded up; down with LBRM, as the
 * reference), nZpBitsPerCb, nDataBytes.  Same arguments and meaning as
 * pyAerial's setPerTbParams (pyaerial/pybind11/pycuphy_ldpc_util.cpp:75-171:
 * tbSize bits without CRC, codeRate as a fraction, rateMatchLen = G) with the
 * PUSCH pipeline's LBRM handling (pusch_rx.cpp:8003-8236 expandBackEndParameters);
 * lbrm may be NULL (no limited buffer).  Fields the derivation does not set
 * are zeroed.  INVALID_ARGUMENT where the reference throws or rejects: rv > 3,
 * Qm not in {1,2,4,6,8}, zero TB size or rate, G not a multiple of Qm*numLayers,
 * numLayers or numUeGrpLayers outside 1..8 or below numLayers, layer map
 * entries >= numUeGrpLayers, LBRM parameters out of range. */
cuphyStatus_t CUPHYWINAPI cuphyHipSetPerTbParams(PerTbParams* tbParams, uint32_t tbSize, float codeRate,
                                                 uint8_t qamMod, uint32_t ndi, uint32_t rv, uint32_t rateMatchLen,
                                                 uint32_t cinit, uint32_t userGroupIdx, uint8_t numLayers,
                                                 uint8_t numUeGrpLayers, const uint32_t* layerMapArray,
                                                 uint8_t nDmrsCdmGrpsNoData, const cuphyHipLbrmPrm_t* lbrm,
                                                 uint16_t* numParityNodes, uint16_t* Kb);

#ifdef __cplusplus
}
#endif

#endif /* CUPHY_HIP_H */
