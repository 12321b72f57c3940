// Systematic LDPC encoder for TS 38.212 BG1/BG2 (gfx950).
//
// Two entry kernels over one core: the decode path's test-signal source
// (bits as bytes, cuphyHipLDPCEncodeCodeBlocks: bench and serving workloads
// build their code blocks with it on the device) and the bit-packed encoder
// behind the reference's encoder API (cuphySetupLDPCEncode).  One workgroup
// encodes one code block in LDS, one byte per bit.  Check row z of block row r involves bit
// (z + s) mod Z of each connected column (circulant shifted right s times,
// TS 38.212 5.3.2).  Encoding uses the double-diagonal core of both base
// graphs (rows 0..3, parity columns p0..p3 = columns nsys..nsys+3):
//   lambda_i = sum of the systematic terms of core row i
//   BG1: row0 P^a p0 + p1,  row1 P^b p0 + p1 + p2,  row2 p2 + p3,  row3 P^a p0 + p3
//   BG2: row0 P^a p0 + p1,  row1 p1 + p2,  row2 P^b p0 + p2 + p3,  row3 P^a p0 + p3
// Summing the four rows leaves P^b p0 = sum_i lambda_i, so
// p0[y] = sum_i lambda_i[(y - b) mod Z]; p1, p3 follow from rows 0 and 3 and
// p2 from row 1.  Every extension column (nsys + r, r >= 4) has degree 1 and
// is the parity of its own row.  The result is checked against H*c = 0 and
// against the oracle's (independent, Gaussian-elimination) encoder.
#include "nr_ldpc_common.hpp"
#include "ldpc_encode.hpp"

namespace nrldpc {

__device__ const reduced_shift_tables g_enc_shift_tables = make_reduced_shift_tables();
__device__ const graph_tables g_enc_graph = make_graph_tables();

__device__ __forceinline__ int rot(int z, int s, int Z)
{
    const int p = z + s;
    return p >= Z ? p - Z : p;
}

// Encode one code block in LDS: x holds its nsys * Z systematic bits (one
// per byte) on entry and (nsys + nrows_out) * Z codeword bits on return
// (parity rows beyond nrows_out are not computed); lam: 4 * Z bytes of
// scratch.  Every thread of the workgroup calls it; ends with a barrier.
__device__ void encode_core(int bg, int Z, int nrows_out, uint8_t* __restrict__ x, uint8_t* __restrict__ lam)
{
    const int nsys = (bg == 1) ? 22 : 10;
    const int zi = lifting_size_index(Z);
    const uint32_t* sh = g_enc_shift_tables.v[bg - 1][zi];
    const uint16_t* rs = g_enc_graph.row_start[bg - 1];
    const uint8_t* col = g_enc_graph.col[bg - 1];
    const int tid = threadIdx.x, nt = blockDim.x;
    // core-row systematic sums and the core shifts a, b
    int a = 0, b = 0;
    for (int e = rs[0]; e < rs[1]; ++e)
        if (col[e] == nsys) a = (int)sh[e];
    const int brow = (bg == 1) ? 1 : 2;
    for (int e = rs[brow]; e < rs[brow + 1]; ++e)
        if (col[e] == nsys) b = (int)sh[e];
    for (int n = tid; n < 4 * Z; n += nt) {
        const int r = n / Z, z = n - r * Z;
        uint32_t acc = 0;
        for (int e = rs[r]; e < rs[r + 1]; ++e)
            if (col[e] < nsys) acc ^= x[col[e] * Z + rot(z, (int)sh[e], Z)];
        lam[n] = (uint8_t)acc;
    }
    __syncthreads();
    uint8_t* p0 = x + nsys * Z;
    uint8_t* p1 = p0 + Z;
    uint8_t* p2 = p1 + Z;
    uint8_t* p3 = p2 + Z;
    for (int y = tid; y < Z; y += nt) {
        const int z = rot(y, Z - b, Z);  // (y - b) mod Z
        p0[y] = lam[z] ^ lam[Z + z] ^ lam[2 * Z + z] ^ lam[3 * Z + z];
    }
    __syncthreads();
    for (int z = tid; z < Z; z += nt) {
        const uint8_t pa = p0[rot(z, a, Z)];
        p1[z] = lam[z] ^ pa;
        p3[z] = lam[3 * Z + z] ^ pa;
    }
    __syncthreads();
    for (int z = tid; z < Z; z += nt) {
        if (bg == 1)
            p2[z] = lam[Z + z] ^ p0[rot(z, b, Z)] ^ p1[z];
        else
            p2[z] = lam[Z + z] ^ p1[z];
    }
    __syncthreads();
    // extension rows: column nsys + r is the parity of row r
    for (int n = tid; n < (nrows_out - 4) * Z; n += nt) {
        const int r = 4 + n / Z, z = n % Z;
        uint32_t acc = 0;
        int ext_s = 0;
        for (int e = rs[r]; e < rs[r + 1]; ++e) {
            if (col[e] == nsys + r)
                ext_s = (int)sh[e];
            else
                acc ^= x[col[e] * Z + rot(z, (int)sh[e], Z)];
        }
        x[(nsys + r) * Z + rot(z, ext_s, Z)] = (uint8_t)acc;
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void ldpc_encode_kernel(int bg, int Z, const uint8_t* __restrict__ info, int info_stride,
                                                          uint8_t* __restrict__ cw, int cw_stride)
{
    extern __shared__ uint8_t enc_lds[];
    const int nsys = (bg == 1) ? 22 : 10;
    const int nrows = (bg == 1) ? 46 : 42;
    uint8_t* x = enc_lds;                         // (nsys + nrows) * Z codeword bits
    uint8_t* lam = enc_lds + (nsys + nrows) * Z;  // 4 * Z
    const uint8_t* src = info + (size_t)blockIdx.x * info_stride;
    uint8_t* dst = cw + (size_t)blockIdx.x * cw_stride;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int n = tid; n < nsys * Z; n += nt) x[n] = src[n] & 1u;
    __syncthreads();
    encode_core(bg, Z, nrows, x, lam);
    for (int n = tid; n < (nsys + nrows) * Z; n += nt) dst[n] = x[n];
}

// Bit-packed encoder of cuphySetupLDPCEncode: block (c, t) encodes code word
// c of TB t.  Input: Kb * Z information bits (bit i of word w = bit 32 w + i,
// filler bits as the caller set them); output: codeword bits from column 2
// (puncture) or 0 up to column nsys + num_rows, packed the same way, the
// rest of the last word 0.
__global__ __launch_bounds__(256) void ldpc_encode_bits_kernel(enc_descr d)
{
    extern __shared__ uint8_t enc_lds[];
    const int t = blockIdx.y, c = blockIdx.x;
    const enc_tensor in = d.input[t], out = d.output[t];
    if (c >= in.num_cw || c >= out.num_cw) return;
    const int Z = d.Z, bg = d.BG;
    const int nsys = (bg == 1) ? 22 : 10;
    uint8_t* x = enc_lds;
    uint8_t* lam = enc_lds + (nsys + max((int)d.num_rows, 4)) * Z;  // the four core parity columns always
    const uint32_t* src = in.addr + (size_t)c * in.stride_words;
    uint32_t* dst = out.addr + (size_t)c * out.stride_words;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int kbits = d.Kb * Z;
    for (int n = tid; n < nsys * Z; n += nt) x[n] = (n < kbits) ? (uint8_t)((src[n >> 5] >> (n & 31)) & 1u) : 0u;
    __syncthreads();
    encode_core(bg, Z, d.num_rows, x, lam);
    const int first = d.puncture ? 2 * Z : 0;
    const int nbits = (nsys + d.num_rows) * Z - first;
    for (int w = tid; w < (nbits + 31) / 32; w += nt) {
        uint32_t v = 0;
        const int b0 = 32 * w;
        const int nb = min(32, nbits - b0);
        for (int i = 0; i < nb; ++i) v |= (uint32_t)x[first + b0 + i] << i;
        dst[w] = v;
    }
}

cuphyStatus_t encode_cbs(int bg, int Z, const uint8_t* info, int info_stride, int ncb, uint8_t* cw, int cw_stride,
                         hipStream_t strm)
{
    if ((bg != 1 && bg != 2) || lifting_size_index(Z) < 0 || !info || !cw || ncb < 0) return CUPHY_STATUS_INVALID_ARGUMENT;
    const int nsys = (bg == 1) ? 22 : 10, nrows = (bg == 1) ? 46 : 42;
    if (info_stride < nsys * Z || cw_stride < (nsys + nrows) * Z) return CUPHY_STATUS_INVALID_ARGUMENT;
    if (ncb == 0) return CUPHY_STATUS_SUCCESS;
    const size_t lds = (size_t)(nsys + nrows + 4) * Z;
    hipLaunchKernelGGL(ldpc_encode_kernel, dim3(ncb), dim3(256), lds, strm, bg, Z, info, info_stride, cw, cw_stride);
    return hipGetLastError() == hipSuccess ? CUPHY_STATUS_SUCCESS : CUPHY_STATUS_INTERNAL_ERROR;
}

void* encode_bits_kernel() { return (void*)&ldpc_encode_bits_kernel; }

}  // namespace nrldpc
