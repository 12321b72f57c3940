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
#include "ldpc_core.hpp"
#include "ldpc_decode.hpp"

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <utility>
#include <vector>

namespace nrldpc {


// Workgroup b decodes code words [b*P, b*P + P) of the launch, P =
// tb_group_size(Z) (several small-Z code blocks share a wave).
template <int BG, typename T, typename TIn, int MAXR>
__global__ __launch_bounds__(384) void ldpc_layered_tb(cuphyLDPCDecodeDesc_t desc)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ uint32_t s_cb_flags[TB_MAX_GROUP];
    const int P = tb_group_size(desc.config.Z);
    const int first = (int)blockIdx.x * P;
    const int count = min(P, desc_codewords(desc) - first);
    decode_group<BG, T, TIn, MAXR>(
        reinterpret_cast<T*>(lds_raw), tb_job<TIn>(desc), count,
        [&](int c) { return tb_buffers<TIn>(desc, first + c); }, s_cb_flags);
}

// Row-pair form of the transport-block kernel (one code block per
// workgroup, Z > 32, mb > 8): workgroup b decodes code word b with two lane
// groups of roundup64(Z) lanes that take the column-disjoint block rows of
// each barrier segment at the same time (run_iterations PAIR = 2).
template <int BG, typename T, typename TIn>
__global__ __launch_bounds__(768) void ldpc_layered_tb_pair(cuphyLDPCDecodeDesc_t desc)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const cb_job job = tb_job<TIn>(desc);
    // the vote word after the lane-major APP values (a multiple of 4 bytes)
    uint32_t* flag = reinterpret_cast<uint32_t*>(
        lds_raw + lm_app_bytes(BG, job.Z, bg_traits<BG>::NSYS + job.mb, (int)sizeof(T),
                               lm_split_layout((int)sizeof(T), 46)));
    decode_lm<BG, T, TIn, 46, 2>(reinterpret_cast<T*>(lds_raw), job, tb_buffers<TIn>(desc, (int)blockIdx.x), flag);
}

// Persistent heterogeneous decoder: each workgroup repeatedly takes the
// next group of work items from a device queue (atomic counter) and decodes
// it, so code blocks of any (BG, Z, mb) share one launch.  Group g is items
// [group_first[g], group_first[g + 1]) (one item per group when group_first
// is null): code blocks of one configuration decoded side by side, lanes
// [c*Z, (c+1)*Z) on item c.  Items are checked before use: an invalid first
// item rejects its group, and the group is cut before any item that is
// invalid, has another configuration, or does not fit the workgroup (count *
// Z > blockDim.x) or the LDS (count regions > lds_values); every item's
// status word says whether it was decoded (the host never builds such
// groups).  The queue state (2 words: next group, finished workgroups) is
// returned to zero by the last workgroup to finish, so the launch can be
// replayed from a graph.
constexpr int PERSISTENT_MAX_GROUP = 192;  // 384 lanes / Z = 2

template <typename T, typename TIn, int MAXR>
__global__ __launch_bounds__(384) void ldpc_layered_persistent(cuphyHipLDPCWorkItem_t* items,
                                                               const uint32_t* __restrict__ group_first,
                                                               int n_groups, int lds_values,
                                                               uint32_t* __restrict__ queue)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ int s_item;
    __shared__ uint32_t s_cb_flags[PERSISTENT_MAX_GROUP];
    T* app = reinterpret_cast<T*>(lds_raw);
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        // uniform: the item's fields then come in through scalar loads and
        // the per-row bounds (R < mb) are scalar branches
        const int g = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (g >= n_groups) break;
        const int first = group_first ? (int)group_first[g] : g;
        const int n_in_group = group_first ? (int)group_first[g + 1] - first : 1;
        cuphyHipLDPCWorkItem_t* grp = items + first;
        const cb_job job = item_job(grp[0]);
        const uint32_t bad0 = work_item_reasons(grp[0], MAXR);
        int cap = 0, count = 0;  // items that fit; items decoded (a prefix of the group)
        if (bad0 == 0) {
            const int nval = ((grp[0].BG == 1) ? 22 : 10) * job.Z + job.mb * job.Z;
            cap = min(n_in_group, min((int)blockDim.x / job.Z, lds_values / group_stride(nval)));
            cap = min(cap, PERSISTENT_MAX_GROUP);
            count = cap;
            for (int c = 1; c < cap; ++c)
                if (!same_config(grp[c], grp[0]) || grp[c].llr == nullptr || grp[c].out == nullptr) {
                    count = c;
                    break;
                }
        }
        for (int c = threadIdx.x; c < n_in_group; c += blockDim.x) {
            uint32_t st = 0;
            if (c >= count) {
                st = (c == 0) ? bad0 : work_item_reasons(grp[c], MAXR) | (same_config(grp[c], grp[0]) ? 0u : CUPHY_HIP_ITEM_GROUP);
                if (st == 0) st = (bad0 != 0 || c < cap) ? CUPHY_HIP_ITEM_GROUP : CUPHY_HIP_ITEM_NO_FIT;
            }
            grp[c].status = st;
        }
        if (count >= 1) {
            auto buf = [&](int c) {
                return cb_buffers{grp[c].llr, grp[c].out, static_cast<_Float16*>(grp[c].soft)};
            };
            if (grp[0].BG == 1)
                decode_group<1, T, TIn, MAXR>(app, job, count, buf, s_cb_flags);
            else
                decode_group<2, T, TIn, MAXR>(app, job, count, buf, s_cb_flags);
        }
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

// Row-pair persistent kernel (768 threads): every code block of the queue
// decoded alone by two lane groups that take the independent block rows of
// each barrier segment at the same time (run_iterations PAIR = 2), with 12
// waves on the CU's 4 SIMDs.  For batches with fewer code blocks than CUs,
// where the slot time is the slowest code block's decode time.  Same queue
// protocol as ldpc_layered_persistent; a group's code blocks are decoded
// one after another.
// Body of the lane-major persistent kernels (PAIR = 2: row-pair, 768
// threads; PAIR = 1: one lane group, one code block per workgroup of at
// least roundup64(Z) threads).  No static LDS: the APP values start at LDS
// address 0; the queue slot and the vote word are the last 16 bytes of the
// dynamic allocation.
template <typename T, typename TIn, int MAXR, int PAIR>
__device__ __forceinline__ void persistent_lm(cuphyHipLDPCWorkItem_t* items, const uint32_t* __restrict__ group_first,
                                              int n_groups, int lds_values, uint32_t* __restrict__ queue)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t lds_bytes = (uint32_t)lds_values * (uint32_t)sizeof(T);
    int* const s_item = reinterpret_cast<int*>(lds_raw + lds_bytes - 16u);
    uint32_t* const flag = reinterpret_cast<uint32_t*>(lds_raw + lds_bytes - 8u);
    T* app = reinterpret_cast<T*>(lds_raw);
    for (;;) {
        if (threadIdx.x == 0) *s_item = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        const int g = __builtin_amdgcn_readfirstlane(*s_item);
        __syncthreads();
        if (g >= n_groups) break;
        const int first = group_first ? (int)group_first[g] : g;
        const int count = group_first ? (int)group_first[g + 1] - first : 1;
        for (int c = 0; c < count; ++c) {
            cuphyHipLDPCWorkItem_t& w = items[first + c];
            const cb_job job = item_job(w);
            uint32_t st = work_item_reasons(w, MAXR);
            if (st == 0) {
                // lane-major APP values, below the last 16 bytes
                const uint32_t need = lm_app_bytes(w.BG, job.Z, ((w.BG == 1) ? 22 : 10) + job.mb, (int)sizeof(T),
                                                   lm_split_layout((int)sizeof(T), MAXR));
                if (need + 16u > lds_bytes || PAIR * ((job.Z + 63) & ~63) > (int)blockDim.x) st = CUPHY_HIP_ITEM_NO_FIT;
            }
            if (threadIdx.x == 0) w.status = st;
            if (st == 0) {
                const cb_buffers b{w.llr, w.out, static_cast<_Float16*>(w.soft)};
                if (w.BG == 1)
                    decode_lm<1, T, TIn, MAXR, PAIR>(app, job, b, flag);
                else
                    decode_lm<2, T, TIn, MAXR, PAIR>(app, job, b, flag);
            }
            __syncthreads();
        }
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

// Row-pair persistent kernel (768 threads): every code block of the queue
// decoded alone by two lane groups that take the independent block rows of
// each barrier segment at the same time (run_iterations PAIR = 2), with 12
// waves on the CU's 4 SIMDs.  For batches with fewer code blocks than CUs,
// where the slot time is the slowest code block's decode time.  Same queue
// protocol as ldpc_layered_persistent; a group's code blocks are decoded
// one after another.
template <typename T, typename TIn, int MAXR>
__global__ __launch_bounds__(768) void ldpc_layered_persistent_pair(cuphyHipLDPCWorkItem_t* items,
                                                                    const uint32_t* __restrict__ group_first,
                                                                    int n_groups, int lds_values,
                                                                    uint32_t* __restrict__ queue)
{
    persistent_lm<T, TIn, MAXR, 2>(items, group_first, n_groups, lds_values, queue);
}

// One code block per workgroup (one lane group), lane-major APP values:
// for launches without code-block groups.  Same queue protocol.
template <typename T, typename TIn, int MAXR>
__global__ __launch_bounds__(384) void ldpc_layered_persistent_lm(cuphyHipLDPCWorkItem_t* items,
                                                                  const uint32_t* __restrict__ group_first,
                                                                  int n_groups, int lds_values,
                                                                  uint32_t* __restrict__ queue)
{
    persistent_lm<T, TIn, MAXR, 1>(items, group_first, n_groups, lds_values, queue);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

void* select_persistent_pair_kernel(cuphyDataType_t llr_type, bool fp16_math)
{
    if (llr_type == CUPHY_R_32F && !fp16_math) return (void*)&ldpc_layered_persistent_pair<float, float, 46>;
    if (llr_type == CUPHY_R_32F) return (void*)&ldpc_layered_persistent_pair<_Float16, float, 46>;
    if (llr_type == CUPHY_R_16F) return (void*)&ldpc_layered_persistent_pair<_Float16, _Float16, 46>;
    return nullptr;
}

template <int MAXR>
void* select_persistent_lm_kernel(cuphyDataType_t llr_type, bool fp16_math)
{
    if (llr_type == CUPHY_R_32F && !fp16_math) return (void*)&ldpc_layered_persistent_lm<float, float, MAXR>;
    if (llr_type == CUPHY_R_32F) return (void*)&ldpc_layered_persistent_lm<_Float16, float, MAXR>;
    if (llr_type == CUPHY_R_16F) return (void*)&ldpc_layered_persistent_lm<_Float16, _Float16, MAXR>;
    return nullptr;
}

template <int MAXR>
void* select_tb_kernel(int bg, cuphyDataType_t llr_type, bool fp16_math)
{
    if (bg == 1) {
        if (llr_type == CUPHY_R_32F)
            return fp16_math ? (void*)&ldpc_layered_tb<1, _Float16, float, MAXR> : (void*)&ldpc_layered_tb<1, float, float, MAXR>;
        if (llr_type == CUPHY_R_16F) return (void*)&ldpc_layered_tb<1, _Float16, _Float16, MAXR>;
    } else if (bg == 2) {
        if (llr_type == CUPHY_R_32F)
            return fp16_math ? (void*)&ldpc_layered_tb<2, _Float16, float, MAXR> : (void*)&ldpc_layered_tb<2, float, float, MAXR>;
        if (llr_type == CUPHY_R_16F) return (void*)&ldpc_layered_tb<2, _Float16, _Float16, MAXR>;
    }
    return nullptr;
}

void* select_tb_pair_kernel(int bg, cuphyDataType_t llr_type, bool fp16_math)
{
    if (bg == 1) {
        if (llr_type == CUPHY_R_32F)
            return fp16_math ? (void*)&ldpc_layered_tb_pair<1, _Float16, float> : (void*)&ldpc_layered_tb_pair<1, float, float>;
        if (llr_type == CUPHY_R_16F) return (void*)&ldpc_layered_tb_pair<1, _Float16, _Float16>;
    } else if (bg == 2) {
        if (llr_type == CUPHY_R_32F)
            return fp16_math ? (void*)&ldpc_layered_tb_pair<2, _Float16, float> : (void*)&ldpc_layered_tb_pair<2, float, float>;
        if (llr_type == CUPHY_R_16F) return (void*)&ldpc_layered_tb_pair<2, _Float16, _Float16>;
    }
    return nullptr;
}

void* select_kernel(int bg, cuphyDataType_t llr_type, bool fp16_math, int mb)
{
    return mb <= SMALL_MAXR ? select_tb_kernel<SMALL_MAXR>(bg, llr_type, fp16_math)
                            : select_tb_kernel<46>(bg, llr_type, fp16_math);
}

template <int MAXR>
void* select_persistent_kernel(cuphyDataType_t llr_type, bool fp16_math)
{
    if (llr_type == CUPHY_R_32F && !fp16_math) return (void*)&ldpc_layered_persistent<float, float, MAXR>;
    if (llr_type == CUPHY_R_32F) return (void*)&ldpc_layered_persistent<_Float16, float, MAXR>;
    if (llr_type == CUPHY_R_16F) return (void*)&ldpc_layered_persistent<_Float16, _Float16, MAXR>;
    return nullptr;
}


}  // namespace

bool is_lifting_size(int Z) { return lifting_set_index(Z) >= 0; }

uint32_t work_item_check(const cuphyHipLDPCWorkItem_t& w, int max_mb)
{
    // the kernel variant a launch runs holds 8 block rows when max_mb <= 8, else 46
    return work_item_reasons(w.BG, lifting_size_index(w.Z), w.num_parity_nodes, w.Kb, w.max_iterations,
                             w.llr != nullptr, w.out != nullptr, max_mb <= SMALL_MAXR ? SMALL_MAXR : 46);
}

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

cuphyStatus_t check_lm_kernel(void* func);

cuphyStatus_t fill_launch(const cuphyLDPCDecodeDesc_t& desc, launch_geometry& g)
{
    const cuphyLDPCDecodeConfigDesc_t& c = desc.config;
    if (!config_supported(c)) return CUPHY_STATUS_UNSUPPORTED_CONFIG;
    if (desc.num_tbs < 1 || desc.num_tbs > CUPHY_LDPC_DECODE_DESC_MAX_TB) return CUPHY_STATUS_INVALID_ARGUMENT;
    const bool fp16 = use_fp16_math(c);
    const int P = tb_group_size(c.Z);
    // one code block per workgroup and rows past the high-rate core: row-pair
    // form (two lane groups; the 46-row variant holds one 6-wave workgroup
    // per CU anyway, the row-pair one 12 waves)
    const bool pair = P == 1 && c.num_parity_nodes > SMALL_MAXR;
    g.func = pair ? select_tb_pair_kernel(c.BG, c.llr_type, fp16) : select_kernel(c.BG, c.llr_type, fp16, c.num_parity_nodes);
    if (!g.func) return CUPHY_STATUS_UNSUPPORTED_CONFIG;
    if (pair) {
        const cuphyStatus_t s = check_lm_kernel(g.func);
        if (s != CUPHY_STATUS_SUCCESS) return s;
    }
    const int nsys = (c.BG == 1) ? 22 : 10;
    const size_t stride = (size_t)(((nsys + c.num_parity_nodes) * c.Z + 7) & ~7);
    // row-pair kernels keep the APP values lane-major (binary16: split layout)
    g.lds_bytes = pair ? (size_t)lm_app_bytes(c.BG, c.Z, nsys + c.num_parity_nodes, fp16 ? 2 : 4,
                                              lm_split_layout(fp16 ? 2 : 4, 46)) + 16u
                       : (size_t)P * stride * (fp16 ? 2u : 4u);
    g.block = pair ? dim3((unsigned)(2 * ((c.Z + 63) / 64 * 64))) : dim3((unsigned)((P * c.Z + 63) / 64 * 64));
    g.grid = dim3((unsigned)((desc_codewords(desc) + P - 1) / P));
    return CUPHY_STATUS_SUCCESS;
}

// Raise a kernel's dynamic-LDS limit above the 64 KB default, once per
// (kernel, size) high-water mark.
cuphyStatus_t ensure_lds_attr(void* func, size_t bytes)
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

// Lane-major kernels address their APP values from LDS address 0, so they
// must hold no static LDS (which would come first); checked once per kernel.
cuphyStatus_t check_lm_kernel(void* func)
{
    static std::mutex mu;
    static std::vector<void*> ok;
    std::lock_guard<std::mutex> g(mu);
    for (void* f : ok)
        if (f == func) return CUPHY_STATUS_SUCCESS;
    hipFuncAttributes attr{};
    if (hipFuncGetAttributes(&attr, func) != hipSuccess || attr.sharedSizeBytes != 0) return CUPHY_STATUS_INTERNAL_ERROR;
    ok.push_back(func);
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

cuphyStatus_t decode_batch(cuphyHipLDPCWorkItem_t* d_items, const uint32_t* d_group_first, int n_groups,
                           cuphyDataType_t llr_type, bool fp16_math, int max_values_per_group, int max_mb,
                           int block_threads, uint32_t* d_queue, int cu_count, hipStream_t strm)
{
    if (n_groups < 0 || !d_queue || (n_groups > 0 && !d_items)) return CUPHY_STATUS_INVALID_ARGUMENT;
    const bool pair = block_threads == 768;  // row-pair mode
    if (!pair && (block_threads < 64 || block_threads > 384 || block_threads % 64)) return CUPHY_STATUS_INVALID_ARGUMENT;
    const bool f16 = fp16_math || llr_type == CUPHY_R_16F;
    const size_t elem = f16 ? 2 : 4;
    // a group's LDS regions are at most 384 lanes' worth of code blocks
    if (max_values_per_group <= 0 || (size_t)max_values_per_group * elem + (pair || !d_group_first ? 8u * 384u : 0u) > 160u * 1024u)
        return CUPHY_STATUS_INVALID_ARGUMENT;
    if (n_groups == 0) return CUPHY_STATUS_SUCCESS;
    if (llr_type != CUPHY_R_32F && llr_type != CUPHY_R_16F) return CUPHY_STATUS_UNSUPPORTED_TYPE;
    // without code-block groups, one code block per workgroup: lane-major kernel
    const bool lm = pair || !d_group_first;
    void* func = pair                   ? select_persistent_pair_kernel(llr_type, f16)
                 : lm                   ? (max_mb <= SMALL_MAXR ? select_persistent_lm_kernel<SMALL_MAXR>(llr_type, f16)
                                                                : select_persistent_lm_kernel<46>(llr_type, f16))
                 : max_mb <= SMALL_MAXR ? select_persistent_kernel<SMALL_MAXR>(llr_type, f16)
                                        : select_persistent_kernel<46>(llr_type, f16);
    const unsigned block = (unsigned)block_threads;
    // lane-major: APP rows padded to an odd number of words, at most 7 bytes
    // per lane beyond the values (Z <= 384), and 16 bytes of queue slot and
    // vote; the split layout of the binary16 46-row kernels needs its
    // second array at LM_HALF
    size_t lds = (size_t)max_values_per_group * elem + (lm ? 8u * 384u : 0u);
    if (lm && lm_split_layout((int)elem, (pair || max_mb > SMALL_MAXR) ? 46 : SMALL_MAXR))
        lds = std::max(lds, (size_t)lm_app_bytes_max((int)elem, true) + 16u);
    cuphyStatus_t s = lm ? check_lm_kernel(func) : CUPHY_STATUS_SUCCESS;
    if (s == CUPHY_STATUS_SUCCESS) s = ensure_lds_attr(func, lds);
    if (s != CUPHY_STATUS_SUCCESS) return s;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, func, (int)block, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const int grid = std::min(n_groups, std::max(1, cu_count) * per_cu);
    int lds_values = (int)(lds / elem);
    void* args[5] = {&d_items, const_cast<uint32_t**>(&d_group_first),
                     &n_groups, &lds_values, &d_queue};
    if (hipLaunchKernel(func, dim3((unsigned)grid), dim3(block), args, lds, strm) != hipSuccess)
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
