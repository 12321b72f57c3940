// Host-side PUSCH per-TB parameter derivation: transport block size
// (TS 38.214 5.1.3.2), code-block segmentation (TS 38.212 5.2.2, 6.2.2),
// limited-buffer rate matching and the decoder's parity-node count.
//
// The arithmetic follows the reference's PUSCH host code step by step, so
// that every derived field equals what the reference hands its kernels:
//   cuPHY/src/cuphy/utils.cuh:60-105            get_TB_size_and_num_CBs (single-precision Ninfo)
//   cuPHY/src/cuphy_channels/pusch_rx.cpp:8003-8236  expandBackEndParameters
//   pyaerial/pybind11/pycuphy_ldpc_util.cpp:75-171   setPerTbParams
// Three deliberate differences: N_ref is computed exactly (see below); a payload of (Kcb - 24, Kcb] bits is segmented
// by the size with its TB CRC, as TS 38.212 5.2.2 (pusch_rx.cpp:8039 tests
// the payload alone and derives K' > Kcb there); and when the received LLRs of a CB do not reach
// the end of its systematic part (G/C + k0 < K' - 2Zc), the reference's
// unsigned subtraction wraps and selects the maximum number of parity nodes;
// here the count is the minimum, 4, as the reference's CPU path computes it
// (5GModel/nr_matlab/pxsch/derate_match.m:122-129).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cuphy_hip.h"

namespace {

// TS 38.214 Table 5.1.3.2-1: TBS for N_info <= 3824.
const uint32_t kTbsTable[93] = {
    24,   32,   40,   48,   56,   64,   72,   80,   88,   96,   104,  112,  120,  128,  136,  144,
    152,  160,  168,  176,  184,  192,  208,  224,  240,  256,  272,  288,  304,  320,  336,  352,
    368,  384,  408,  432,  456,  480,  504,  528,  552,  576,  608,  640,  672,  704,  736,  768,
    808,  848,  888,  928,  984,  1032, 1064, 1128, 1160, 1192, 1224, 1256, 1288, 1320, 1352, 1416,
    1480, 1544, 1608, 1672, 1736, 1800, 1864, 1928, 2024, 2088, 2152, 2216, 2280, 2408, 2472, 2536,
    2600, 2664, 2728, 2792, 2856, 2976, 3104, 3240, 3368, 3496, 3624, 3752, 3824};

// TS 38.212 Table 5.3.2-1 lifting sizes, ascending.
const uint32_t kLifting[51] = {2,  3,  4,  5,  6,  7,   8,   9,   10,  11,  12,  13,  14,  15,  16,  18,  20,
                               22, 24, 26, 28, 30, 32,  36,  40,  44,  48,  52,  56,  60,  64,  72,  80,  88,
                               96, 104, 112, 120, 128, 144, 160, 176, 192, 208, 224, 240, 256, 288, 320, 352, 384};

// Same steps and precisions as utils.cuh:60-105: N_info in binary32, the
// rounding of N_info to N'_info in binary64.
void tb_size_and_cbs(int num_symbols, int num_prbs, int num_layers, float code_rate, uint32_t Qm, int ncdm1,
                     uint32_t& num_CBs, uint32_t& tb_size)
{
    const int nre_prime = 12 * num_symbols + 6 * ncdm1;
    const uint32_t nre = (uint32_t)std::min(156, nre_prime) * (uint32_t)num_prbs;
    const float ninfo = (float)nre * code_rate * (float)Qm * (float)num_layers;
    tb_size = 0;
    if (ninfo <= 3824.0f) {
        const int n = std::max(3, (int)std::floor(std::log2(ninfo)) - 6);
        const double p = std::pow(2.0, n);
        const uint32_t ninfo_prime = (uint32_t)std::max(24, (int)(p * std::floor((double)ninfo / p)));
        for (uint32_t t : kTbsTable)
            if (t >= ninfo_prime) {
                tb_size = t;
                break;
            }
        num_CBs = 1;
    } else {
        const uint32_t n = (uint32_t)(std::floor(std::log2(ninfo - 24.0f)) - 5.0f);
        const double p = std::pow(2.0, (double)n);
        const uint32_t ninfo_prime = (uint32_t)std::max(3840, (int)(p * std::round(((double)ninfo - 24.0) / p)));
        const float num = (float)(ninfo_prime + 24u);
        if ((double)code_rate <= 0.25)
            num_CBs = (uint32_t)std::ceil((double)num / 3816.0);
        else
            num_CBs = (ninfo_prime > 8424u) ? (uint32_t)std::ceil((double)num / 8424.0) : 1u;
        tb_size = (uint32_t)(8.0 * num_CBs * std::ceil((double)num / (8.0 * num_CBs)) - 24.0);
    }
}

}  // namespace

extern "C" {

cuphyStatus_t CUPHYWINAPI cuphyHipTransportBlockSize(int num_symbols, int num_prbs, int num_layers, float code_rate,
                                                     uint32_t Qm, int num_dmrs_cdm_grps_no_data1_symbols,
                                                     uint32_t* tb_size, uint32_t* num_CBs)
{
    if (!tb_size || !num_CBs || num_symbols < 0 || num_prbs <= 0 || num_layers <= 0 || Qm == 0 || Qm > 8 ||
        num_dmrs_cdm_grps_no_data1_symbols < 0 || !(code_rate > 0.0f) || code_rate > 1.0f ||
        12 * num_symbols + 6 * num_dmrs_cdm_grps_no_data1_symbols <= 0)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    tb_size_and_cbs(num_symbols, num_prbs, num_layers, code_rate, Qm, num_dmrs_cdm_grps_no_data1_symbols, *num_CBs,
                    *tb_size);
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t CUPHYWINAPI cuphyHipSetPerTbParams(PerTbParams* p, uint32_t tbSize, float codeRate, uint8_t qamMod,
                                                 uint32_t ndi, uint32_t rv, uint32_t rateMatchLen, uint32_t cinit,
                                                 uint32_t userGroupIdx, uint8_t numLayers, uint8_t numUeGrpLayers,
                                                 const uint32_t* layerMapArray, uint8_t nDmrsCdmGrpsNoData,
                                                 const cuphyHipLbrmPrm_t* lbrm, uint16_t* numParityNodes,
                                                 uint16_t* Kb_out)
{
    if (!p || !numParityNodes || !Kb_out || tbSize == 0 || !(codeRate > 0.0f) || codeRate > 1.0f || rv > 3)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    if (qamMod != 1 && qamMod != 2 && qamMod != 4 && qamMod != 6 && qamMod != 8) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (numLayers < 1 || numLayers > MAX_N_LAYERS_PUSCH || numUeGrpLayers < numLayers ||
        numUeGrpLayers > MAX_N_LAYERS_PUSCH)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    const uint32_t nq = (uint32_t)qamMod * numLayers;
    if (rateMatchLen == 0 || rateMatchLen % nq != 0) return CUPHY_STATUS_INVALID_ARGUMENT;
    for (int l = 0; l < numLayers; ++l) {
        const uint32_t m = layerMapArray ? layerMapArray[l] : (uint32_t)l;
        if (m >= numUeGrpLayers) return CUPHY_STATUS_INVALID_ARGUMENT;
    }
    const bool lb = lbrm && lbrm->i_lbrm;
    if (lb && (lbrm->maxLayers < 1 || lbrm->maxLayers > 4 || (lbrm->maxQm != 6 && lbrm->maxQm != 8) ||
               lbrm->n_PRB_LBRM < 32 || lbrm->n_PRB_LBRM > 273))  // utils.cuh:313-323
        return CUPHY_STATUS_INVALID_ARGUMENT;

    std::memset(p, 0, sizeof(*p));
    // base graph (TS 38.212 6.2.2; pusch_rx.cpp:8020-8023)
    const double R = (double)codeRate;
    const uint32_t bg = (tbSize <= 292 || (tbSize <= 3824 && R <= 0.67) || R <= 0.25) ? 2u : 1u;
    // segmentation (TS 38.212 5.2.2; pusch_rx.cpp:8025-8063)
    const uint32_t Kcb = bg == 1 ? 8448u : 3840u;
    // B = payload + TB CRC (CRC16 up to 3824 bits); more than one CB when B > Kcb, as TS 38.212 5.2.2,
    // utils.cuh:141-161 (get_K_prime) and derive_coding_main.m:163-172 (pusch_rx.cpp:8039 compares the
    // payload alone, which leaves K' > Kcb for payloads in (Kcb - 24, Kcb])
    uint32_t C = 1, crc_bytes = tbSize <= 3824 ? 2u : 3u;
    const uint32_t B = tbSize + 8 * crc_bytes;
    uint32_t Bp = B;
    if (B > Kcb) {
        C = (B + (Kcb - 24) - 1) / (Kcb - 24);
        Bp = B + 24 * C;
    }
    const uint32_t Kp = Bp / C;
    // info nodes and lifting size (pusch_rx.cpp:8080-8105)
    uint32_t Kb = 22;
    if (bg == 2) Kb = B > 640 ? 10u : (B > 560 ? 9u : (B > 192 ? 8u : 6u));
    uint32_t Zc = 0;
    for (uint32_t z : kLifting)
        if (z * Kb >= Kp) {
            Zc = z;
            break;
        }
    if (Zc == 0) return CUPHY_STATUS_INVALID_ARGUMENT;  // K' beyond 384 * Kb
    const uint32_t K = Zc * (bg == 1 ? 22u : 10u);
    const uint32_t F = K - Kp;
    const uint32_t N = Zc * (bg == 1 ? 66u : 50u);
    // limited buffer (pusch_rx.cpp:8145-8164): N_ref = floor(TBS_LBRM / (C * R_LBRM)), R_LBRM = 2/3, in
    // integers as utils.cuh:330-333 (pusch_rx.cpp:8162 divides in binary32, one less where 3 TBS_LBRM / (2C)
    // is an integer)
    uint32_t Ncb = N;
    if (lb) {
        uint32_t tbs_lbrm = 0, c_unused = 0;
        tb_size_and_cbs(13, lbrm->n_PRB_LBRM, lbrm->maxLayers, 948.0f / 1024.0f, lbrm->maxQm, 0, c_unused, tbs_lbrm);
        const uint32_t nref = (uint32_t)((uint64_t)tbs_lbrm * 3u / (2u * (uint64_t)C));
        Ncb = std::min(Ncb, nref);
    }
    // k0 (TS 38.212 Table 5.4.2.1-2)
    static const uint32_t num1[4] = {0, 17, 33, 56}, num2[4] = {0, 13, 25, 43};
    const uint32_t k0 = bg == 1 ? (num1[rv] * Ncb / (66 * Zc)) * Zc : (num2[rv] * Ncb / (50 * Zc)) * Zc;
    // parity nodes covering the received LLRs (pusch_rx.cpp:8189-8224)
    const long long Kd = (long long)K - F - 2LL * Zc;
    const long long n_par = (long long)std::min<uint32_t>(rateMatchLen / C + k0, Ncb) - Kd;
    long long mb = 4;
    if (n_par > 0) mb = lb ? n_par / Zc : (n_par + Zc - 1) / Zc;
    mb = std::max(4LL, std::min<long long>(bg == 1 ? 46 : 42, mb));

    p->tbSize = tbSize;
    p->codeRate = codeRate;
    p->Qm = qamMod;
    p->ndi = ndi;
    p->rv = rv;
    p->bg = bg;
    p->Nl = numLayers;
    p->num_CBs = C;
    p->Zc = Zc;
    p->N = N;
    p->Ncb = Ncb;
    p->Ncb_padded = (Ncb + 2 * Zc + 7) / 8 * 8;
    p->G = rateMatchLen;
    p->encodedSize = rateMatchLen;
    p->K = K;
    p->F = F;
    p->k0 = k0;
    p->cinit = cinit;
    p->nDataBytes = ((K - crc_bytes * 8 - F + 7) / 8) * C;  // pusch_rx.cpp:8066-8071
    p->firstCodeBlockIndex = 0;
    for (int l = 0; l < numLayers; ++l) p->layer_map_array[l] = layerMapArray ? layerMapArray[l] : (uint32_t)l;
    p->userGroupIndex = userGroupIdx;
    p->nBBULayers = numUeGrpLayers;
    p->startLLR = 0;
    p->nZpBitsPerCb = (uint32_t)mb * Zc + K;
    p->isDataPresent = 1;
    p->uciOnPuschFlag = 0;
    p->nDmrsCdmGrpsNoData = nDmrsCdmGrpsNoData;
    *numParityNodes = (uint16_t)mb;
    *Kb_out = (uint16_t)Kb;
    return CUPHY_STATUS_SUCCESS;
}

}  // extern "C"
