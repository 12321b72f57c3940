// Native driver of the decode path's C ABI (no Python): times the
// transport-block decode and the persistent batch decode on device-resident
// LLRs and prints one line per workload.  Used for hardware-counter runs
// (rocprofv3 --pmc), where the program under the profiler should be small.
//
//   decode_bench [tb|batch|wide|widepair|all|valupeak] [repetitions]
//
//   tb    : 1 TB x 8 CBs, BG1 Z=384 mb=46, binary32 LLRs, 10 iterations
//           (the plain per-TB launch of cuphyErrorCorrectionLDPCTransportBlockDecode)
//   batch : 2800 CBs BG1 Z=384 mb=4, binary16 LLRs, 10 iterations, one
//           persistent launch (cuphyHipLDPCBatchDecode)
//   wide  : 256 CBs BG1 Z=384 mb=46, binary16 LLRs, 10 iterations, one
//           persistent launch (one code block per CU: the latency case)
//   widepair : the same in row-pair mode (cuphyHipLDPCBatchDecodeGroups,
//           768-thread workgroups, one code block per group)
//
//   valupeak : VALU issue rate of the device in lane-ops per second for the
//           instruction types of the decoder's row body (each in 8 independent
//           chains per lane, 8 workgroups of 256 threads per CU); v_add_u32
//           gives the denominator of the decoder's VALU roofline
//
// LLRs are pseudo-random (+-0.5..8.5, xorshift from a fixed seed): decode
// time does not depend on them without early termination.
#include "cuphy_hip.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

#define CHECK_HIP(x)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) {                                                              \
            std::fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            std::exit(1);                                                                    \
        }                                                                                    \
    } while (0)
#define CHECK_CUPHY(x)                                                                       \
    do {                                                                                     \
        cuphyStatus_t s_ = (x);                                                              \
        if (s_ != CUPHY_STATUS_SUCCESS) {                                                    \
            std::fprintf(stderr, "%s:%d %s: status %d\n", __FILE__, __LINE__, #x, (int)s_);  \
            std::exit(1);                                                                    \
        }                                                                                    \
    } while (0)

uint32_t g_rng = 0x9E3779B9u;
float next_llr()
{
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 17;
    g_rng ^= g_rng << 5;
    const float mag = 0.5f + (float)(g_rng & 0xFFFF) * (8.0f / 65536.0f);
    return (g_rng >> 31) ? -mag : mag;
}

uint16_t to_half_bits(float f)
{
    _Float16 h = (_Float16)f;
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}

struct timer {
    hipEvent_t a, b;
    timer()
    {
        CHECK_HIP(hipEventCreate(&a));
        CHECK_HIP(hipEventCreate(&b));
    }
    ~timer()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

void run_tb(cuphyLDPCDecoder_t dec, hipStream_t strm, int reps)
{
    const int Z = 384, mb = 46, ncb = 8, nval = (22 + mb) * Z;
    std::vector<float> h((size_t)ncb * nval);
    for (int c = 0; c < ncb; ++c)
        for (int n = 0; n < nval; ++n) h[(size_t)c * nval + n] = (n < 2 * Z) ? 0.0f : next_llr();
    float* d_llr = nullptr;
    uint32_t* d_out = nullptr;
    CHECK_HIP(hipMalloc(&d_llr, h.size() * sizeof(float)));
    CHECK_HIP(hipMalloc(&d_out, (size_t)ncb * 264 * sizeof(uint32_t)));
    CHECK_HIP(hipMemcpy(d_llr, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));

    cuphyLDPCDecodeDesc_t d;
    std::memset(&d, 0, sizeof(d));
    d.config.llr_type = CUPHY_R_32F;
    d.config.num_parity_nodes = mb;
    d.config.Z = Z;
    d.config.max_iterations = 10;
    d.config.Kb = 22;
    d.config.BG = 1;
    d.config.clamp_value = 32.0f;
    d.config.norm.f32 = 0.8125f;
    d.num_tbs = 1;
    d.llr_input[0].addr = d_llr;
    d.llr_input[0].stride_elements = nval;
    d.llr_input[0].num_codewords = ncb;
    d.tb_output[0].addr = d_out;
    d.tb_output[0].stride_words = 264;
    d.tb_output[0].num_codewords = ncb;
    CHECK_CUPHY(cuphyErrorCorrectionLDPCTransportBlockDecode(dec, &d, strm));
    CHECK_HIP(hipStreamSynchronize(strm));
    timer t;
    CHECK_HIP(hipEventRecord(t.a, strm));
    for (int r = 0; r < reps; ++r) CHECK_CUPHY(cuphyErrorCorrectionLDPCTransportBlockDecode(dec, &d, strm));
    CHECK_HIP(hipEventRecord(t.b, strm));
    CHECK_HIP(hipEventSynchronize(t.b));
    float ms = 0;
    CHECK_HIP(hipEventElapsedTime(&ms, t.a, t.b));
    std::printf("tb: BG1 Z=384 mb=46 8 CBs fp32 10 it: %.2f us per TB launch (%d launches)\n", 1e3 * ms / reps,
                reps);
    (void)hipFree(d_llr);
    (void)hipFree(d_out);
}

// block 0: cuphyHipLDPCBatchDecode; otherwise cuphyHipLDPCBatchDecodeGroups
// with one code block per group and `block` threads per workgroup
void run_batch(cuphyLDPCDecoder_t dec, hipStream_t strm, int reps, int ncb, int mb, const char* name, int block = 0)
{
    const int Z = 384, nval = (22 + mb) * Z, stride = (nval + 7) & ~7;
    std::vector<uint16_t> h((size_t)ncb * stride, 0);
    for (int c = 0; c < ncb; ++c)
        for (int n = 2 * Z; n < nval; ++n) h[(size_t)c * stride + n] = to_half_bits(next_llr());
    uint16_t* d_llr = nullptr;
    uint32_t* d_out = nullptr;
    CHECK_HIP(hipMalloc(&d_llr, h.size() * sizeof(uint16_t)));
    CHECK_HIP(hipMalloc(&d_out, (size_t)ncb * 264 * sizeof(uint32_t)));
    CHECK_HIP(hipMemcpy(d_llr, h.data(), h.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    std::vector<cuphyHipLDPCWorkItem_t> items(ncb);
    const uint16_t a = to_half_bits(0.8125f);
    for (int c = 0; c < ncb; ++c) {
        cuphyHipLDPCWorkItem_t& w = items[c];
        std::memset(&w, 0, sizeof(w));
        w.llr = d_llr + (size_t)c * stride;
        w.out = d_out + (size_t)c * 264;
        w.BG = 1;
        w.Z = Z;
        w.num_parity_nodes = (int16_t)mb;
        w.Kb = 22;
        w.max_iterations = 10;
        w.norm.f16x2 = (uint32_t)a | ((uint32_t)a << 16);
        w.clamp_value = 32.0f;
    }
    cuphyHipLDPCWorkItem_t* d_items = nullptr;
    CHECK_HIP(hipMalloc(&d_items, items.size() * sizeof(items[0])));
    CHECK_HIP(hipMemcpy(d_items, items.data(), items.size() * sizeof(items[0]), hipMemcpyHostToDevice));
    size_t ws_bytes = 0;
    CHECK_CUPHY(cuphyHipLDPCBatchDecodeGetWorkspaceSize(&ws_bytes));
    void* d_ws = nullptr;
    CHECK_HIP(hipMalloc(&d_ws, ws_bytes));
    CHECK_HIP(hipMemset(d_ws, 0, ws_bytes));
    std::vector<uint32_t> first(ncb + 1);
    for (int c = 0; c <= ncb; ++c) first[c] = (uint32_t)c;
    uint32_t* d_first = nullptr;
    CHECK_HIP(hipMalloc(&d_first, first.size() * sizeof(uint32_t)));
    CHECK_HIP(hipMemcpy(d_first, first.data(), first.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    auto launch = [&]() {
        if (block == 0)
            CHECK_CUPHY(cuphyHipLDPCBatchDecode(dec, d_items, ncb, CUPHY_R_16F, 1, nval, mb, d_ws, strm));
        else
            CHECK_CUPHY(cuphyHipLDPCBatchDecodeGroups(dec, d_items, d_first, ncb, CUPHY_R_16F, 1, stride, mb, block,
                                                      d_ws, strm));
    };
    launch();
    CHECK_HIP(hipStreamSynchronize(strm));
    timer t;
    CHECK_HIP(hipEventRecord(t.a, strm));
    for (int r = 0; r < reps; ++r) launch();
    CHECK_HIP(hipEventRecord(t.b, strm));
    CHECK_HIP(hipEventSynchronize(t.b));
    float ms = 0;
    CHECK_HIP(hipEventElapsedTime(&ms, t.a, t.b));
    const double us = 1e3 * ms / reps;
    const double gbps = (double)ncb * 22 * Z / (us * 1e-6) / 1e9;
    std::printf("%s: %d CBs BG1 Z=384 mb=%d fp16 10 it: %.2f us per launch, %.3f Gbps (%d launches)\n", name, ncb,
                mb, us, gbps, reps);
    (void)hipFree(d_items);
    (void)hipFree(d_first);
    (void)hipFree(d_ws);
    (void)hipFree(d_llr);
    (void)hipFree(d_out);
}

constexpr int PEAK_UNROLL = 4;

// One VALU instruction of the probe (the operand registers hold 32-bit
// values; for the 16-bit forms only the low halves are used): K selects it.
template <int K>
__device__ __forceinline__ void peak_op(uint32_t& a, uint32_t b)
{
    if constexpr (K == 0) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a) : "v"(b));
    if constexpr (K == 1) asm volatile("v_add_f16 %0, %0, %1" : "+v"(a) : "v"(b));
    if constexpr (K == 2) asm volatile("v_pk_add_f16 %0, %0, %1" : "+v"(a) : "v"(b));
    if constexpr (K == 3) asm volatile("v_min_f16 %0, %0, %1" : "+v"(a) : "v"(b));
    if constexpr (K == 4) asm volatile("v_med3_f16 %0, %0, %1, %1" : "+v"(a) : "v"(b));
    if constexpr (K == 5) asm volatile("v_lshl_add_u32 %0, %1, 1, %0" : "+v"(a) : "v"(b));
    if constexpr (K == 6) asm volatile("v_add3_u32 %0, %0, %1, %1" : "+v"(a) : "v"(b));
    if constexpr (K == 7) asm volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(a) : "v"(b) : "vcc");
}
constexpr const char* PEAK_NAMES[8] = {"v_add_u32", "v_add_f16", "v_pk_add_f16", "v_min_f16",
                                       "v_med3_f16", "v_lshl_add_u32", "v_add3_u32", "v_cndmask_b32"};

// 8 chains x PEAK_UNROLL instructions per loop trip; the instructions are
// inline asm so that the count of VALU instructions is exactly what is timed.
template <int K>
__global__ __launch_bounds__(256) void valu_peak_kernel(uint32_t* out, int iters)
{
    uint32_t a[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) a[c] = threadIdx.x + (uint32_t)c;
    const uint32_t b = blockIdx.x | 1u;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int u = 0; u < PEAK_UNROLL; ++u) {
#pragma unroll
            for (int c = 0; c < 8; ++c) peak_op<K>(a[c], b);
        }
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = a[0] ^ a[1] ^ a[2] ^ a[3] ^ a[4] ^ a[5] ^ a[6] ^ a[7];
}

template <int K>
void run_valu_peak_one(hipStream_t strm, int reps, int cus, int clk_khz, uint32_t* d_out)
{
    const int blocks = cus * 8, threads = 256, iters = 8192;
    hipLaunchKernelGGL(valu_peak_kernel<K>, dim3(blocks), dim3(threads), 0, strm, d_out, iters);
    CHECK_HIP(hipStreamSynchronize(strm));
    timer t;
    CHECK_HIP(hipEventRecord(t.a, strm));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(valu_peak_kernel<K>, dim3(blocks), dim3(threads), 0, strm, d_out, iters);
    CHECK_HIP(hipEventRecord(t.b, strm));
    CHECK_HIP(hipEventSynchronize(t.b));
    float ms = 0;
    CHECK_HIP(hipEventElapsedTime(&ms, t.a, t.b));
    const double lane_ops = (double)blocks * threads * iters * 8 * PEAK_UNROLL * reps;
    const double rate = lane_ops / (ms * 1e-3);
    std::printf("valupeak %-15s: %6.2f T lane-ops/s (%5.1f lane-ops per CU per clock at the reported %.0f MHz)\n",
                PEAK_NAMES[K], rate / 1e12, rate / cus / (clk_khz * 1e3), clk_khz / 1e3);
}

void run_valu_peak(hipStream_t strm, int reps)
{
    int dev = 0, cus = 0, clk_khz = 0;
    CHECK_HIP(hipGetDevice(&dev));
    CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CHECK_HIP(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, dev));
    uint32_t* d_out = nullptr;
    CHECK_HIP(hipMalloc(&d_out, (size_t)cus * 8 * 256 * sizeof(uint32_t)));
    run_valu_peak_one<0>(strm, reps, cus, clk_khz, d_out);
    run_valu_peak_one<1>(strm, reps, cus, clk_khz, d_out);
    run_valu_peak_one<2>(strm, reps, cus, clk_khz, d_out);
    run_valu_peak_one<3>(strm, reps, cus, clk_khz, d_out);
    run_valu_peak_one<4>(strm, reps, cus, clk_khz, d_out);
    run_valu_peak_one<5>(strm, reps, cus, clk_khz, d_out);
    run_valu_peak_one<6>(strm, reps, cus, clk_khz, d_out);
    run_valu_peak_one<7>(strm, reps, cus, clk_khz, d_out);
    (void)hipFree(d_out);
}

}  // namespace

int main(int argc, char** argv)
{
    const std::string what = argc > 1 ? argv[1] : "all";
    const int reps = argc > 2 ? std::atoi(argv[2]) : 5;
    if (reps < 1 || (what != "tb" && what != "batch" && what != "wide" && what != "widepair" && what != "all" &&
                     what != "valupeak")) {
        std::fprintf(stderr, "usage: decode_bench [tb|batch|wide|widepair|all|valupeak] [repetitions >= 1]\n");
        return 2;
    }
    cuphyContext_t ctx = nullptr;
    cuphyLDPCDecoder_t dec = nullptr;
    CHECK_CUPHY(cuphyCreateContext(&ctx, 0));
    CHECK_CUPHY(cuphyCreateLDPCDecoder(ctx, &dec, 0));
    hipStream_t strm;
    CHECK_HIP(hipStreamCreate(&strm));
    if (what == "tb" || what == "all") run_tb(dec, strm, reps);
    if (what == "batch" || what == "all") run_batch(dec, strm, reps, 2800, 4, "batch");
    if (what == "wide" || what == "all") run_batch(dec, strm, reps, 256, 46, "wide");
    if (what == "widepair" || what == "all") run_batch(dec, strm, reps, 256, 46, "widepair", 768);
    if (what == "valupeak") run_valu_peak(strm, reps);
    CHECK_HIP(hipStreamDestroy(strm));
    (void)cuphyDestroyLDPCDecoder(dec);
    (void)cuphyDestroyContext(ctx);
    return 0;
}
