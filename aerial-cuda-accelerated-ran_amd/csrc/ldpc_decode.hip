// Layered normalized min-sum LDPC decoder for gfx950 (MI355X).
//
// One workgroup decodes one code block (CB).  The whole APP vector of the
// CB, (NSYS + mb) * Z values, lives in LDS (BG1, Z = 384: 26,112 values =
// 104,448 B in binary32, 52,224 B in binary16, of the 160 KB per CU).
// Lane z owns check row z of every block row: a circulant shift s is an LDS
// index rotation (z + s) mod Z, no data moves.  Each lane keeps the
// compressed check-to-variable (C2V) message of each of its rows in VGPRs
// (normalised min1, min2, argmin, final signs); the block rows are fully
// unrolled at compile time, so the register arrays are indexed by constants.
// A workgroup barrier separates consecutive rows.  No MFMA: nothing here is
// a dense contraction.
//
// Arithmetic follows the oracle operation by operation (the reference CPU
// path, 5GModel/nr_matlab/pxsch/MSA_layering.m:64-74 with compute_V2C.m:46,
// compute_cC2V.m:42-54, update_C2V.m:39-45, update_APP.m:40), so for a given
// storage type the hard decisions and APP values are bit-identical to the
// oracle's in that type:
//   v2c_j = APP[col_j][(z+s_j) mod Z] - a*c2v_j          (one RN op each)
//   min1/argmin (first index on ties), min2 over the others, sign product
//   c2v_j = (j == argmin ? min2 : min1) * sign_prod * sign_j
//   APP   = v2c_j + a*c2v_j    for every neighbour (extension columns too)
// a*c2v is held already multiplied, which is exact: RN is sign-symmetric.
// Hard decision APP <= 0 -> 1 (cuPHY ldpc2.cuh:100-113 and MSA_layering.m:138-139).
//
// Boundary: the kernel consumes cuphyLDPCDecodeDesc_t exactly as the
// reference's transport-block kernels do (ldpc2_reg_index_fp_desc_dyn.cu:
// 233-272): code word i of the launch is the i-th code word of the TB list,
// its LLRs start with the 2Z punctured positions (ldpc2.cuh:1484-1487), its
// hard output is Kb*Z bits, LSB-first in 32-bit words (ldpc2_dec_output.cuh:
// 568-625), its optional soft output Kb*Z binary16 values (:983-1016).

#include "nr_ldpc_common.hpp"
#include "ldpc_decode.hpp"

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <utility>
#include <vector>

namespace nrldpc {

// Reduced shifts of every (BG, Z) in device memory, constant-initialised in
// the code object.  Read through an address-space-4 pointer so that the
// uniform loads become scalar (s_load) loads.
__device__ const reduced_shift_tables g_shift_tables = make_reduced_shift_tables();
typedef const __attribute__((address_space(4))) uint32_t* shift_ptr;

template <typename T> struct num;
template <> struct num<float> {
    static __device__ __forceinline__ float vmin(float a, float b) { return __builtin_fminf(a, b); }
    static __device__ __forceinline__ float vmax(float a, float b) { return __builtin_fmaxf(a, b); }
    static __device__ __forceinline__ float vabs(float a) { return __builtin_fabsf(a); }
    static __device__ __forceinline__ float inf() { return __builtin_huge_valf(); }
};
template <> struct num<_Float16> {
    static __device__ __forceinline__ _Float16 vmin(_Float16 a, _Float16 b) { return __builtin_fminf16(a, b); }
    static __device__ __forceinline__ _Float16 vmax(_Float16 a, _Float16 b) { return __builtin_fmaxf16(a, b); }
    static __device__ __forceinline__ _Float16 vabs(_Float16 a) { return __builtin_fabsf16(a); }
    static __device__ __forceinline__ _Float16 inf() { return (_Float16)__builtin_huge_valf(); }
};

template <typename TIn, typename T> struct llr_in;
template <> struct llr_in<float, float> {
    static __device__ __forceinline__ float cvt(float v) { return v; }
};
template <> struct llr_in<float, _Float16> {
    static __device__ __forceinline__ _Float16 cvt(float v) { return (_Float16)v; }
};
template <> struct llr_in<_Float16, _Float16> {
    static __device__ __forceinline__ _Float16 cvt(_Float16 v) { return v; }
};

// Locate code word `cw` of the launch in the TB list (the reference's
// decode_desc_output_addr, ldpc2_dec_output.cuh:58-81).
struct cw_location {
    int tb, idx;
};
__device__ __forceinline__ cw_location locate_cw(const cuphyLDPCDecodeDesc_t& d, int cw)
{
    int tb = 0;
    for (; tb < d.num_tbs - 1; ++tb) {
        if (cw < d.llr_input[tb].num_codewords) break;
        cw -= d.llr_input[tb].num_codewords;
    }
    return {tb, cw};
}

// One block row R for lane z.  st_* hold the row's compressed C2V state.
template <int BG, int R, typename T, int NR>
__device__ __forceinline__ void process_row(T* __restrict__ app, shift_ptr sh, uint32_t z, uint32_t Z,
                                            T alpha, T (&st_n1)[NR], T (&st_n2)[NR], uint32_t (&st_meta)[NR])
{
    using BGT = bg_traits<BG>;
    using N = num<T>;
    constexpr int E0 = BGT::row_start(R);
    constexpr int D = BGT::row_start(R + 1) - E0;
    T t[D];
    uint32_t off[D];
    const T on1 = st_n1[R], on2 = st_n2[R];
    const uint32_t om = st_meta[R];
    const uint32_t oidx = om & 31u;
    // gather APP and subtract the previous C2V message
    static_for<D>([&](auto J) {
        constexpr int j = decltype(J)::value;
        constexpr int col = BGT::col(E0 + j);
        uint32_t p = z + sh[E0 + j];
        const uint32_t q = p - Z;
        p = (q < p) ? q : p;  // (z + s) mod Z, both < 2Z
        off[j] = (uint32_t)col * Z + p;
        const T a = app[off[j]];
        const T mag = (oidx == (uint32_t)j) ? on2 : on1;
        const T old = ((om >> (8 + j)) & 1u) ? -mag : mag;
        t[j] = a - old;
    });
    // min1 / argmin / min2 / sign bits
    T m1 = N::inf(), m2 = N::inf();
    uint32_t idx = 0, sg = 0;
    static_for<D>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const T a = N::vabs(t[j]);
        idx = (a < m1) ? (uint32_t)j : idx;
        m2 = N::vmin(m2, N::vmax(m1, a));
        m1 = N::vmin(m1, a);
        sg |= (uint32_t)(t[j] < (T)0) << j;
    });
    const uint32_t fs = (__builtin_popcount(sg) & 1) ? (sg ^ ((1u << D) - 1u)) : sg;
    const T a1 = alpha * m1, a2 = alpha * m2;
    // new C2V and APP write-back
    static_for<D>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const T mag = (idx == (uint32_t)j) ? a2 : a1;
        const T c = ((fs >> j) & 1u) ? -mag : mag;
        app[off[j]] = t[j] + c;
    });
    st_n1[R] = a1;
    st_n2[R] = a2;
    st_meta[R] = idx | (fs << 8);
}

// Parity of check row R, lane z: hard bit = APP < 0 (syndrome_check.m:38-40).
template <int BG, int R, typename T>
__device__ __forceinline__ uint32_t row_syndrome(const T* __restrict__ app, shift_ptr sh, uint32_t z, uint32_t Z)
{
    using BGT = bg_traits<BG>;
    constexpr int E0 = BGT::row_start(R);
    constexpr int D = BGT::row_start(R + 1) - E0;
    uint32_t par = 0;
    static_for<D>([&](auto J) {
        constexpr int j = decltype(J)::value;
        uint32_t p = z + sh[E0 + j];
        const uint32_t q = p - Z;
        p = (q < p) ? q : p;
        par ^= (uint32_t)(app[(uint32_t)BGT::col(E0 + j) * Z + p] < (T)0);
    });
    return par;
}

// Everything one code block's decode needs, resolved from a TB descriptor
// (transport-block kernel) or from a batch work item (persistent kernel).
struct cb_job {
    const void* llr;
    uint32_t* out;
    _Float16* soft;  // nullptr: no soft output
    int Z, mb, Kb, max_iter;
    bool early_term;
    float clamp;
    float alpha_f32;
    uint16_t alpha_f16;
};

// Decode one code block with the whole workgroup; LDS `app` holds
// (NSYS + mb) * Z values of type T.  Every thread of the workgroup calls it.
template <int BG, typename T, typename TIn>
__device__ __forceinline__ void decode_one(T* __restrict__ app, const cb_job& job)
{
    using BGT = bg_traits<BG>;
    constexpr int NR = BGT::NROWS;
    const uint32_t Z = (uint32_t)job.Z;
    const int mb = job.mb;
    const int nval = (BGT::NSYS + mb) * (int)Z;
    const uint32_t tid = threadIdx.x;
    const uint32_t nthr = blockDim.x;

    // ---- LLRs -> LDS: convert, clamp, signed zero -> +0 ----------------
    {
        const TIn* src = static_cast<const TIn*>(job.llr);
        const T clampv = (T)job.clamp;
        const bool do_clamp = job.clamp > 0.0f;
        for (int n = tid; n < nval; n += nthr) {
            T v = llr_in<TIn, T>::cvt(src[n]);
            if (do_clamp) v = num<T>::vmax(num<T>::vmin(v, clampv), -clampv);
            app[n] = (v == (T)0) ? (T)0 : v;
        }
    }
    __syncthreads();

    T alpha;
    if constexpr (std::is_same<T, float>::value)
        alpha = job.alpha_f32;
    else
        alpha = __builtin_bit_cast(_Float16, job.alpha_f16);

    T st_n1[NR], st_n2[NR];
    uint32_t st_meta[NR];
    static_for<NR>([&](auto R) {
        st_n1[R] = (T)0;
        st_n2[R] = (T)0;
        st_meta[R] = 0;
    });

    const bool active = tid < Z;
    shift_ptr sh = (shift_ptr)(const void*)&g_shift_tables.v[BG - 1][lifting_size_index((int)Z)][0];
    for (int it = 0; it < job.max_iter; ++it) {
        // Opaque copies of the lane index and of the shift pointer, renewed
        // per row: the 316 (BG1) per-edge shifts and addresses are reloaded
        // and recomputed where used instead of being hoisted out of the
        // loops into (spilled) registers.
        uint32_t zl = tid;
        asm volatile("" : "+v"(zl));
        static_for<NR>([&](auto Rc) {
            constexpr int R = decltype(Rc)::value;
            if (R < mb) {
                shift_ptr shr = sh;
                uint32_t Zr = Z;
                asm volatile("" : "+s"(shr), "+s"(Zr));
                if (active) process_row<BG, R, T, NR>(app, shr, zl, Zr, alpha, st_n1, st_n2, st_meta);
                __syncthreads();
            }
        });
        if (job.early_term) {
            uint32_t fail = 0;
            static_for<NR>([&](auto Rc) {
                constexpr int R = decltype(Rc)::value;
                if (R < mb) {
                    shift_ptr shr = sh;
                    uint32_t Zr = Z;
                    asm volatile("" : "+s"(shr), "+s"(Zr));
                    if (active) fail |= row_syndrome<BG, R, T>(app, shr, zl, Zr);
                }
            });
            if (!__syncthreads_or((int)fail)) break;
        }
    }

    // ---- hard decisions: Kb*Z bits, 64 per wave per step (ballot) --------
    const int nbits = job.Kb * (int)Z;
    const int nwords = (nbits + 31) >> 5;
    {
        const int lane = (int)(tid & 63u);
        const int wave = (int)(tid >> 6);
        const int nwaves = (int)((nthr + 63u) >> 6);
        for (int c = wave; c * 64 < nbits; c += nwaves) {
            const int n = c * 64 + lane;
            const bool bit = (n < nbits) && (app[n] <= (T)0);
            const uint64_t m = __ballot(bit);
            if (lane == 0) {
                job.out[2 * c] = (uint32_t)m;
                if (2 * c + 1 < nwords) job.out[2 * c + 1] = (uint32_t)(m >> 32);
            }
        }
    }
    // ---- soft outputs: Kb*Z binary16 APP values ---------------------------
    if (job.soft)
        for (int n = tid; n < nbits; n += nthr) job.soft[n] = (_Float16)app[n];
}

template <int BG, typename T, typename TIn>
__global__ __launch_bounds__(384) void ldpc_layered_tb(cuphyLDPCDecodeDesc_t desc)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const cuphyLDPCDecodeConfigDesc_t& cfg = desc.config;
    const cw_location loc = locate_cw(desc, blockIdx.x);
    cb_job job;
    job.llr = static_cast<const TIn*>(desc.llr_input[loc.tb].addr) +
              (size_t)loc.idx * (size_t)desc.llr_input[loc.tb].stride_elements;
    job.out = desc.tb_output[loc.tb].addr + (size_t)loc.idx * (size_t)desc.tb_output[loc.tb].stride_words;
    job.soft = ((cfg.flags & CUPHY_LDPC_DECODE_WRITE_SOFT_OUTPUTS) && desc.llr_output[loc.tb].addr)
                   ? static_cast<_Float16*>(desc.llr_output[loc.tb].addr) +
                         (size_t)loc.idx * (size_t)desc.llr_output[loc.tb].stride_elements
                   : nullptr;
    job.Z = cfg.Z;
    job.mb = cfg.num_parity_nodes;
    job.Kb = cfg.Kb;
    job.max_iter = cfg.max_iterations;
    job.early_term = (cfg.flags & CUPHY_LDPC_DECODE_EARLY_TERM) != 0;
    job.clamp = cfg.clamp_value;
    job.alpha_f32 = cfg.norm.f32;
    job.alpha_f16 = (uint16_t)(cfg.norm.f16x2 & 0xFFFFu);
    decode_one<BG, T, TIn>(reinterpret_cast<T*>(lds_raw), job);
}

// Persistent heterogeneous decoder: each workgroup repeatedly takes the
// next work item from a device queue (atomic counter) and decodes it, so
// code blocks of any (BG, Z, mb) share one launch.  The queue state
// (2 words: next item, finished workgroups) is returned to zero by the
// last workgroup to finish, so the launch can be replayed from a graph.
template <typename T, typename TIn>
__global__ __launch_bounds__(384) void ldpc_layered_persistent(const cuphyHipLDPCWorkItem_t* __restrict__ items,
                                                               int n_items, uint32_t* __restrict__ queue)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int s_item;
    T* app = reinterpret_cast<T*>(lds_raw);
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        const int i = s_item;
        __syncthreads();
        if (i >= n_items) break;
        const cuphyHipLDPCWorkItem_t& w = items[i];
        cb_job job;
        job.llr = w.llr;
        job.out = w.out;
        job.soft = static_cast<_Float16*>(w.soft);
        job.Z = w.Z;
        job.mb = w.num_parity_nodes;
        job.Kb = w.Kb;
        job.max_iter = w.max_iterations;
        job.early_term = (w.flags & CUPHY_LDPC_DECODE_EARLY_TERM) != 0;
        job.clamp = w.clamp_value;
        job.alpha_f32 = w.norm.f32;
        job.alpha_f16 = (uint16_t)(w.norm.f16x2 & 0xFFFFu);
        if (w.BG == 1)
            decode_one<1, T, TIn>(app, job);
        else
            decode_one<2, T, TIn>(app, job);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&queue[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

void* select_kernel(int bg, cuphyDataType_t llr_type, bool fp16_math)
{
    if (bg == 1) {
        if (llr_type == CUPHY_R_32F) return fp16_math ? (void*)&ldpc_layered_tb<1, _Float16, float> : (void*)&ldpc_layered_tb<1, float, float>;
        if (llr_type == CUPHY_R_16F) return (void*)&ldpc_layered_tb<1, _Float16, _Float16>;
    } else if (bg == 2) {
        if (llr_type == CUPHY_R_32F) return fp16_math ? (void*)&ldpc_layered_tb<2, _Float16, float> : (void*)&ldpc_layered_tb<2, float, float>;
        if (llr_type == CUPHY_R_16F) return (void*)&ldpc_layered_tb<2, _Float16, _Float16>;
    }
    return nullptr;
}

int total_codewords(const cuphyLDPCDecodeDesc_t& d)
{
    int n = 0;
    for (int i = 0; i < d.num_tbs; ++i) n += d.llr_input[i].num_codewords;
    return n;
}

}  // namespace

bool is_lifting_size(int Z) { return lifting_set_index(Z) >= 0; }

bool config_supported(const cuphyLDPCDecodeConfigDesc_t& c)
{
    if (c.BG != 1 && c.BG != 2) return false;
    if (lifting_set_index(c.Z) < 0) return false;
    if (c.num_parity_nodes < 4 || c.num_parity_nodes > (c.BG == 1 ? 46 : 42)) return false;
    if (c.BG == 1 ? (c.Kb != 22) : (c.Kb != 6 && c.Kb != 8 && c.Kb != 9 && c.Kb != 10)) return false;
    if (c.llr_type != CUPHY_R_32F && c.llr_type != CUPHY_R_16F) return false;
    if (c.max_iterations < 0) return false;
    return true;
}

// algo 0: automatic (binary32 math for binary32 LLRs, binary16 for binary16);
// algo 2: binary16 math for binary32 LLRs (the reference's default behaviour).
static bool use_fp16_math(const cuphyLDPCDecodeConfigDesc_t& c)
{
    return c.llr_type == CUPHY_R_16F || c.algo == 2;
}

cuphyStatus_t fill_launch(const cuphyLDPCDecodeDesc_t& desc, launch_geometry& g)
{
    const cuphyLDPCDecodeConfigDesc_t& c = desc.config;
    if (!config_supported(c)) return CUPHY_STATUS_UNSUPPORTED_CONFIG;
    if (desc.num_tbs < 1 || desc.num_tbs > CUPHY_LDPC_DECODE_DESC_MAX_TB) return CUPHY_STATUS_INVALID_ARGUMENT;
    const bool fp16 = use_fp16_math(c);
    g.func = select_kernel(c.BG, c.llr_type, fp16);
    if (!g.func) return CUPHY_STATUS_UNSUPPORTED_CONFIG;
    const int nsys = (c.BG == 1) ? 22 : 10;
    g.lds_bytes = (size_t)(nsys + c.num_parity_nodes) * (size_t)c.Z * (fp16 ? 2u : 4u);
    g.block = dim3((unsigned)((c.Z + 63) / 64 * 64));
    g.grid = dim3((unsigned)total_codewords(desc));
    return CUPHY_STATUS_SUCCESS;
}

// Raise a kernel's dynamic-LDS limit above the 64 KB default, once per
// (kernel, size) high-water mark.
static cuphyStatus_t ensure_lds_attr(void* func, size_t bytes)
{
    static std::mutex mu;
    static std::vector<std::pair<void*, size_t>> done;
    if (bytes <= 65536) return CUPHY_STATUS_SUCCESS;
    std::lock_guard<std::mutex> g(mu);
    for (auto& d : done)
        if (d.first == func && d.second >= bytes) return CUPHY_STATUS_SUCCESS;
    if (hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
        return CUPHY_STATUS_INTERNAL_ERROR;
    for (auto& d : done)
        if (d.first == func) {
            d.second = bytes;
            return CUPHY_STATUS_SUCCESS;
        }
    done.emplace_back(func, bytes);
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t decode_tb(const cuphyLDPCDecodeDesc_t& desc, hipStream_t strm)
{
    launch_geometry g;
    cuphyStatus_t s = fill_launch(desc, g);
    if (s != CUPHY_STATUS_SUCCESS) return s;
    if (g.grid.x == 0) return CUPHY_STATUS_SUCCESS;
    if ((s = ensure_lds_attr(g.func, g.lds_bytes)) != CUPHY_STATUS_SUCCESS) return s;
    void* args[1] = {const_cast<cuphyLDPCDecodeDesc_t*>(&desc)};
    if (hipLaunchKernel(g.func, g.grid, g.block, args, g.lds_bytes, strm) != hipSuccess)
        return CUPHY_STATUS_INTERNAL_ERROR;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t decode_batch(const cuphyHipLDPCWorkItem_t* d_items, int n_items, cuphyDataType_t llr_type,
                           bool fp16_math, int max_values_per_cb, uint32_t* d_queue, int cu_count, hipStream_t strm)
{
    if (n_items < 0 || !d_queue || (n_items > 0 && !d_items)) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (max_values_per_cb <= 0 || max_values_per_cb > 68 * 384) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (n_items == 0) return CUPHY_STATUS_SUCCESS;
    void* func = nullptr;
    size_t elem = 0;
    if (llr_type == CUPHY_R_32F && !fp16_math) {
        func = (void*)&ldpc_layered_persistent<float, float>;
        elem = 4;
    } else if (llr_type == CUPHY_R_32F) {
        func = (void*)&ldpc_layered_persistent<_Float16, float>;
        elem = 2;
    } else if (llr_type == CUPHY_R_16F) {
        func = (void*)&ldpc_layered_persistent<_Float16, _Float16>;
        elem = 2;
    } else {
        return CUPHY_STATUS_UNSUPPORTED_TYPE;
    }
    const size_t lds = (size_t)max_values_per_cb * elem;
    cuphyStatus_t s = ensure_lds_attr(func, lds);
    if (s != CUPHY_STATUS_SUCCESS) return s;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, func, 384, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    const int grid = std::min(n_items, std::max(1, cu_count) * per_cu);
    void* args[3] = {const_cast<cuphyHipLDPCWorkItem_t**>(&d_items), &n_items, &d_queue};
    if (hipLaunchKernel(func, dim3((unsigned)grid), dim3(384), args, lds, strm) != hipSuccess)
        return CUPHY_STATUS_INTERNAL_ERROR;
    return CUPHY_STATUS_SUCCESS;
}

cuphyStatus_t get_launch_config(cuphyLDPCDecodeLaunchConfig_t& lc)
{
    launch_geometry g;
    cuphyStatus_t s = fill_launch(lc.decode_desc, g);
    if (s != CUPHY_STATUS_SUCCESS) return s;
    if ((s = ensure_lds_attr(g.func, g.lds_bytes)) != CUPHY_STATUS_SUCCESS) return s;
    lc.kernel_args[0] = &lc.decode_desc;
    lc.kernel_args[1] = nullptr;
    hipKernelNodeParams& p = lc.kernel_node_params_driver;
    p.blockDim = g.block;
    p.gridDim = g.grid;
    p.func = g.func;
    p.sharedMemBytes = (unsigned int)g.lds_bytes;
    p.kernelParams = lc.kernel_args;
    p.extra = nullptr;
    return CUPHY_STATUS_SUCCESS;
}

}  // namespace nrldpc
