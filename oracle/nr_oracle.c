/*
 * CPU oracle for the 5G NR uplink LDPC decode path -- TEST INFRASTRUCTURE.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * this library; the product never does.  See nr_oracle.h for the reference
 * files each function restates.
 *
 * Signed zero: the decoder canonicalises -0 inputs to +0.  In MATLAB
 * (compute_V2C.m:46-49, update_APP.m:40) the first update of a column
 * computes |v2c|*sgn, which maps a -0 input to +0 as well, so no result of
 * MSA_layering.m changes; doing it at load keeps the GPU kernel free of a
 * per-edge fix-up.
 *
 * Build: see Makefile (gcc -O2 -ffp-contract=off, no fast-math: every
 * float operation is one IEEE-754 RN operation).
 */
#define _GNU_SOURCE /* sched_getaffinity, pthread_setaffinity_np (pinned CPU-baseline threads) */
#include "nr_oracle.h"

#include <math.h>
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../aerial-cuda-accelerated-ran_amd/csrc/nr_bg_tables.h"

/* ======================================================================
 * lifting sizes, TS 38.212 Table 5.3.2-1
 * ==================================================================== */
int nr_orc_ils(int Z)
{
    static const int a[8] = {2, 3, 5, 7, 9, 11, 13, 15};
    for (int i = 0; i < 8; ++i)
        for (int j = 0; a[i] << j <= 384; ++j)
            if ((a[i] << j) == Z) return i;
    return -1;
}

int nr_orc_num_rows(int bg) { return bg == 1 ? NR_BG1_NUM_ROWS : NR_BG2_NUM_ROWS; }
int nr_orc_num_sys_cols(int bg) { return bg == 1 ? 22 : 10; }

int nr_orc_row_degree(int bg, int r)
{
    const uint16_t* rs = (bg == 1) ? nr_bg1_row_start : nr_bg2_row_start;
    return rs[r + 1] - rs[r];
}

/* load_Tanner.m:93-98: shift reduced mod Zc */
int nr_orc_edge(int bg, int Z, int r, int j, int* col, int* shift)
{
    const int ils = nr_orc_ils(Z);
    if (ils < 0) return -1;
    const uint16_t* rs = (bg == 1) ? nr_bg1_row_start : nr_bg2_row_start;
    const int e = rs[r] + j;
    if (bg == 1) {
        *col = nr_bg1_col[e];
        *shift = nr_bg1_shift[ils][e] % Z;
    } else {
        *col = nr_bg2_col[e];
        *shift = nr_bg2_shift[ils][e] % Z;
    }
    return 0;
}

/* ======================================================================
 * random numbers: splitmix64 seeding + xoshiro256** (public-domain
 * algorithms by Vigna / Blackman); Box-Muller for normals.
 * ==================================================================== */
static uint64_t splitmix64(uint64_t* x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void nr_orc_rng_seed(nr_orc_rng* g, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) g->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t nr_orc_rng_next(nr_orc_rng* g)
{
    uint64_t* s = g->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

double nr_orc_rng_uniform(nr_orc_rng* g)
{
    return (double)((nr_orc_rng_next(g) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

double nr_orc_rng_normal(nr_orc_rng* g)
{
    const double u1 = nr_orc_rng_uniform(g);
    const double u2 = nr_orc_rng_uniform(g);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void nr_orc_random_bits(uint64_t seed, uint8_t* bits, int n)
{
    nr_orc_rng g;
    nr_orc_rng_seed(&g, seed);
    for (int i = 0; i < n; ++i) bits[i] = (uint8_t)(nr_orc_rng_next(&g) >> 63);
}

/* ======================================================================
 * channel models
 * ==================================================================== */
void nr_orc_bpsk_awgn_llr(const uint8_t* bits, int n, double sigma2, uint64_t seed, float* llr)
{
    nr_orc_rng g;
    nr_orc_rng_seed(&g, seed);
    const double sigma = sqrt(sigma2);
    for (int i = 0; i < n; ++i) {
        const double x = bits[i] ? -1.0 : 1.0;
        const double y = x + sigma * nr_orc_rng_normal(&g);
        llr[i] = (float)(2.0 * y / sigma2);
    }
}

/* noise convention of cuPHY/examples/error_correction/ldpc_decode_test_vec_gen.cpp:183-188:
 * total noise variance 10^(-SNR/10), half per real component */
void nr_orc_qpsk_awgn_llr(const uint8_t* bits, int n, double snr_db, uint64_t seed, float* llr)
{
    nr_orc_rng g;
    nr_orc_rng_seed(&g, seed);
    const double n0 = pow(10.0, -snr_db / 10.0);
    const double var = n0 / 2.0;
    const double a = sqrt(0.5);
    const double sd = sqrt(var);
    for (int i = 0; i < n; ++i) {
        const double x = bits[i] ? -1.0 : 1.0;
        const double y = a * x + sd * nr_orc_rng_normal(&g);
        llr[i] = (float)(2.0 * a * y / var);
    }
}

void nr_orc_consistent_gaussian_llr(const uint8_t* bits, int n, double mu, uint64_t seed, float* llr)
{
    nr_orc_rng g;
    nr_orc_rng_seed(&g, seed);
    const double sd = sqrt(2.0 * mu);
    for (int i = 0; i < n; ++i) {
        const double x = bits[i] ? -1.0 : 1.0;
        llr[i] = (float)(mu * x + sd * nr_orc_rng_normal(&g));
    }
}

/* ======================================================================
 * IEEE binary16 rounding (round to nearest, ties to even)
 * ==================================================================== */
float nr_orc_round_half(float x)
{
    if (isnan(x)) return x;
    const float ax = fabsf(x);
    if (ax >= 65520.0f) return copysignf(INFINITY, x);
    if (ax < 6.103515625e-05f) { /* subnormal: quantum 2^-24 */
        const float q = nearbyintf(ax * 16777216.0f);
        return copysignf(q / 16777216.0f, x);
    }
    int e;
    frexpf(ax, &e);                              /* ax in [2^(e-1), 2^e) */
    const float quantum = ldexpf(1.0f, e - 11);  /* 10 fraction bits    */
    return copysignf(nearbyintf(ax / quantum) * quantum, x);
}

/* ======================================================================
 * layered normalized min-sum, instantiated per arithmetic mode
 * ==================================================================== */
#define NMS_NAME nms_fp32
#define NMS_T float
#define NMS_OP(x) (x)
#define NMS_ABS fabsf
#include "nr_nms_body.inc"
#undef NMS_NAME
#undef NMS_T
#undef NMS_OP
#undef NMS_ABS

#define NMS_NAME nms_fp64
#define NMS_T double
#define NMS_OP(x) (x)
#define NMS_ABS fabs
#include "nr_nms_body.inc"
#undef NMS_NAME
#undef NMS_T
#undef NMS_OP
#undef NMS_ABS

/* binary16: every product and sum of two binary16 values is formed in
 * binary32 and rounded once to binary16.  Since 24 >= 2*11 + 2 the double
 * rounding is innocuous, so each op equals the correctly rounded binary16
 * result. */
#define NMS_NAME nms_fp16
#define NMS_T float
#define NMS_OP(x) nr_orc_round_half(x)
#define NMS_ABS fabsf
#include "nr_nms_body.inc"
#undef NMS_NAME
#undef NMS_T
#undef NMS_OP
#undef NMS_ABS

int nr_orc_decode_cb(const nr_orc_dec_cfg* cfg, const float* llr, uint32_t* hard, float* app_out)
{
    if (!cfg || !llr || nr_orc_ils(cfg->Z) < 0 || (cfg->bg != 1 && cfg->bg != 2)) return -1;
    if (cfg->mb < 4 || cfg->mb > nr_orc_num_rows(cfg->bg)) return -1;
    switch (cfg->mode) {
    case NR_ORC_FP32: return nms_fp32(cfg, llr, hard, app_out);
    case NR_ORC_FP64: return nms_fp64(cfg, llr, hard, app_out);
    case NR_ORC_FP16: return nms_fp16(cfg, llr, hard, app_out);
    default: return -1;
    }
}

/* compute_cC2V.m:42-54 + update_C2V.m:39-45 for one lane */
void nr_orc_check_node(int d, const float* t, float* c2v, int mode)
{
    double min1 = INFINITY, min2 = INFINITY;
    int idx = -1, par = 0;
    for (int j = 0; j < d; ++j) {
        const double v = (mode == NR_ORC_FP16) ? nr_orc_round_half(t[j]) : t[j];
        const double a = fabs(v);
        if (a < min1) {
            min2 = min1;
            min1 = a;
            idx = j;
        } else if (a < min2) {
            min2 = a;
        }
        par ^= (v < 0);
    }
    for (int j = 0; j < d; ++j) {
        const double v = (mode == NR_ORC_FP16) ? nr_orc_round_half(t[j]) : t[j];
        const double mag = (j == idx) ? min2 : min1;
        c2v[j] = (float)((par ^ (v < 0)) ? -mag : mag);
    }
}

/* ======================================================================
 * encoder: H*c = 0 with H from the base graph (TS 38.212 5.3.2):
 * block (r, col) with shift s is the Z x Z identity cyclically shifted
 * right s times, i.e. check z of row r involves bit (z + s) mod Z of col.
 * Core parity (4 columns after the systematic ones) is solved with a
 * cached GF(2) inverse of the 4Z x 4Z core block; each extension parity
 * column is determined by its single row.
 * ==================================================================== */
typedef struct {
    int n;          /* 4Z */
    int w;          /* words per row */
    uint64_t* inv;  /* n rows x w words */
} core_inv_t;

static core_inv_t* g_core_inv[3][385];
static pthread_mutex_t g_core_mutex = PTHREAD_MUTEX_INITIALIZER;

static core_inv_t* core_inverse(int bg, int Z)
{
    pthread_mutex_lock(&g_core_mutex);
    core_inv_t* ci = g_core_inv[bg][Z];
    if (ci) {
        pthread_mutex_unlock(&g_core_mutex);
        return ci;
    }
    const int nsys = nr_orc_num_sys_cols(bg);
    const int n = 4 * Z, w = (n + 63) / 64;
    uint64_t* a = (uint64_t*)calloc((size_t)n * 2 * w, sizeof(uint64_t)); /* [A | I] */
    for (int r = 0; r < 4; ++r) {
        const int d = nr_orc_row_degree(bg, r);
        for (int j = 0; j < d; ++j) {
            int col, s;
            nr_orc_edge(bg, Z, r, j, &col, &s);
            if (col < nsys || col >= nsys + 4) continue;
            for (int z = 0; z < Z; ++z) {
                const int row = r * Z + z;
                const int c = (col - nsys) * Z + (z + s) % Z;
                a[(size_t)row * 2 * w + c / 64] ^= 1ull << (c % 64);
            }
        }
    }
    for (int i = 0; i < n; ++i) a[(size_t)i * 2 * w + w + i / 64] |= 1ull << (i % 64);
    int ok = 1;
    for (int c = 0; c < n && ok; ++c) {
        int p = -1;
        for (int i = c; i < n; ++i)
            if ((a[(size_t)i * 2 * w + c / 64] >> (c % 64)) & 1) {
                p = i;
                break;
            }
        if (p < 0) {
            ok = 0;
            break;
        }
        if (p != c)
            for (int k = 0; k < 2 * w; ++k) {
                uint64_t t = a[(size_t)p * 2 * w + k];
                a[(size_t)p * 2 * w + k] = a[(size_t)c * 2 * w + k];
                a[(size_t)c * 2 * w + k] = t;
            }
        for (int i = 0; i < n; ++i)
            if (i != c && ((a[(size_t)i * 2 * w + c / 64] >> (c % 64)) & 1))
                for (int k = 0; k < 2 * w; ++k) a[(size_t)i * 2 * w + k] ^= a[(size_t)c * 2 * w + k];
    }
    if (!ok) {
        free(a);
        pthread_mutex_unlock(&g_core_mutex);
        return NULL;
    }
    ci = (core_inv_t*)malloc(sizeof(core_inv_t));
    ci->n = n;
    ci->w = w;
    ci->inv = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n * w);
    for (int i = 0; i < n; ++i) memcpy(ci->inv + (size_t)i * w, a + (size_t)i * 2 * w + w, sizeof(uint64_t) * w);
    free(a);
    g_core_inv[bg][Z] = ci;
    pthread_mutex_unlock(&g_core_mutex);
    return ci;
}

int nr_orc_encode(int bg, int Z, const uint8_t* info, uint8_t* cw)
{
    if ((bg != 1 && bg != 2) || nr_orc_ils(Z) < 0) return -1;
    const int nsys = nr_orc_num_sys_cols(bg), nrows = nr_orc_num_rows(bg);
    const int ncol = nsys + nrows;
    core_inv_t* ci = core_inverse(bg, Z);
    if (!ci) return -2;
    memset(cw, 0, (size_t)ncol * Z);
    memcpy(cw, info, (size_t)nsys * Z);
    /* right-hand side of the 4 core rows from the systematic columns */
    uint64_t* b = (uint64_t*)calloc((size_t)ci->w, sizeof(uint64_t));
    for (int r = 0; r < 4; ++r) {
        const int d = nr_orc_row_degree(bg, r);
        for (int z = 0; z < Z; ++z) {
            int acc = 0;
            for (int j = 0; j < d; ++j) {
                int col, s;
                nr_orc_edge(bg, Z, r, j, &col, &s);
                if (col < nsys) acc ^= cw[col * Z + (z + s) % Z];
            }
            if (acc) b[(r * Z + z) / 64] |= 1ull << ((r * Z + z) % 64);
        }
    }
    for (int i = 0; i < ci->n; ++i) {
        int acc = 0;
        for (int k = 0; k < ci->w; ++k) acc ^= __builtin_parityll(ci->inv[(size_t)i * ci->w + k] & b[k]);
        cw[nsys * Z + i] = (uint8_t)acc;
    }
    free(b);
    /* extension rows: row r >= 4 owns column nsys + r (degree-1 column) */
    for (int r = 4; r < nrows; ++r) {
        const int d = nr_orc_row_degree(bg, r);
        int ext_s = -1;
        for (int j = 0; j < d; ++j) {
            int col, s;
            nr_orc_edge(bg, Z, r, j, &col, &s);
            if (col == nsys + r) ext_s = s;
        }
        if (ext_s < 0) return -3;
        for (int z = 0; z < Z; ++z) {
            int acc = 0;
            for (int j = 0; j < d; ++j) {
                int col, s;
                nr_orc_edge(bg, Z, r, j, &col, &s);
                if (col != nsys + r) acc ^= cw[col * Z + (z + s) % Z];
            }
            cw[(nsys + r) * Z + (z + ext_s) % Z] = (uint8_t)acc;
        }
    }
    return 0;
}

int nr_orc_syndrome(int bg, int Z, int mb, const uint8_t* x)
{
    int fails = 0;
    for (int r = 0; r < mb; ++r) {
        const int d = nr_orc_row_degree(bg, r);
        for (int z = 0; z < Z; ++z) {
            int acc = 0;
            for (int j = 0; j < d; ++j) {
                int col, s;
                nr_orc_edge(bg, Z, r, j, &col, &s);
                acc ^= x[col * Z + (z + s) % Z];
            }
            fails += acc;
        }
    }
    return fails;
}

/* ======================================================================
 * CRC, TS 38.212 5.1 (bit-serial long division, zero initial value)
 * ==================================================================== */
uint32_t nr_orc_crc(const uint8_t* bits, int n, int type)
{
    uint32_t poly, L;
    switch (type) {
    case NR_CRC24A: poly = 0x864CFBu; L = 24; break;
    case NR_CRC24B: poly = 0x800063u; L = 24; break;
    case NR_CRC16: poly = 0x1021u; L = 16; break;
    default: return 0xFFFFFFFFu;
    }
    const uint32_t top = 1u << (L - 1), mask = (L == 32) ? 0xFFFFFFFFu : ((1u << L) - 1u);
    uint32_t reg = 0;
    for (int i = 0; i < n; ++i) {
        const uint32_t fb = ((reg & top) ? 1u : 0u) ^ (bits[i] & 1u);
        reg = (reg << 1) & mask;
        if (fb) reg ^= poly;
    }
    return reg;
}

/* ======================================================================
 * de-rate-matching with HARQ combining
 * ==================================================================== */
int nr_orc_k0(int bg, int rv, int Ncb, int Zc)
{
    static const int n1[4] = {0, 17, 33, 56}, n2[4] = {0, 13, 25, 43};
    if (rv < 0 || rv > 3) return -1;
    if (bg == 1) return (int)(((long long)n1[rv] * Ncb) / (66LL * Zc)) * Zc;
    return (int)(((long long)n2[rv] * Ncb) / (50LL * Zc)) * Zc;
}

/* ======================================================================
 * transport block size and coding parameters
 * (5GModel/nr_matlab/pxsch/derive_coding_main.m:78-301)
 * ==================================================================== */
/* TS 38.214 Table 5.1.3.2-1 (the TBS_table of derive_coding_main.m:84) */
static const int kOrcTbs[93] = {
    24, 32, 40, 48, 56, 64, 72, 80, 88, 96, 104, 112, 120, 128, 136, 144, 152, 160, 168, 176, 184, 192, 208, 224,
    240, 256, 272, 288, 304, 320, 336, 352, 368, 384, 408, 432, 456, 480, 504, 528, 552, 576, 608, 640, 672, 704,
    736, 768, 808, 848, 888, 928, 984, 1032, 1064, 1128, 1160, 1192, 1224, 1256, 1288, 1320, 1352, 1416, 1480,
    1544, 1608, 1672, 1736, 1800, 1864, 1928, 2024, 2088, 2152, 2216, 2280, 2408, 2472, 2536, 2600, 2664, 2728,
    2792, 2856, 2976, 3104, 3240, 3368, 3496, 3624, 3752, 3824};

int nr_orc_tbs_table(int i) { return (i >= 0 && i < 93) ? kOrcTbs[i] : -1; }

/* derive_coding_main.m:86-125 in binary64: N_info = N_RE R Qm nl with
 * N_RE = min(156, re_per_prb) n_prb.  For R <= 1/4 the code-block count
 * uses N'_info (TS 38.214 5.1.3.2; MATLAB's line :111 writes N_info and a
 * strict <, which the standard does not). */
int nr_orc_tb_size(double re_per_prb, int n_prb, double R, int Qm, int nl)
{
    const double Nre = fmin(156.0, re_per_prb) * n_prb;
    const double Ninfo = Nre * R * Qm * nl;
    if (Ninfo <= 3824.0) {
        const double n = fmax(3.0, floor(log2(Ninfo)) - 6.0);
        const double Np = fmax(24.0, pow(2.0, n) * floor(Ninfo / pow(2.0, n)));
        for (int i = 0; i < 93; ++i)
            if (kOrcTbs[i] >= Np) return kOrcTbs[i];
        return -1;
    }
    const double n = floor(log2(Ninfo - 24.0)) - 5.0;
    const double Np = fmax(3840.0, pow(2.0, n) * round((Ninfo - 24.0) / pow(2.0, n)));
    double C = 1.0;
    if (R <= 0.25) C = ceil((Np + 24.0) / 3816.0);
    else if (Np > 8424.0) C = ceil((Np + 24.0) / 8424.0);
    return (int)(8.0 * C * ceil((Np + 24.0) / (8.0 * C)) - 24.0);
}

/* derive_coding_main.m:127-301: base graph, TB CRC, segmentation (C, K'),
 * K_b, lifting size Zc (smallest with Zc K_b >= K'), K, F, N and the
 * limited-buffer N_ref (0 without LBRM). */
int nr_orc_derive_coding(int tbs, double R, int i_lbrm, int n_prb_lbrm, int max_layers, int max_qm,
                         nr_orc_coding* o)
{
    static const int Zs[51] = {2, 4, 8, 16, 32, 64, 128, 256, 3, 6, 12, 24, 48, 96, 192, 384, 5, 10,
                               20, 40, 80, 160, 320, 7, 14, 28, 56, 112, 224, 9, 18, 36, 72, 144, 288, 11,
                               22, 44, 88, 176, 352, 13, 26, 52, 104, 208, 15, 30, 60, 120, 240};
    memset(o, 0, sizeof(*o));
    o->tbs = tbs;
    o->bg = (tbs <= 292 || (tbs <= 3824 && R <= 0.67) || R <= 0.25) ? 2 : 1;       /* :138-143 */
    o->crc_len = tbs > 3824 ? 24 : 16;                                              /* :145-154 */
    o->B = tbs + o->crc_len;
    const int Kcb = o->bg == 1 ? 8448 : 3840;                                       /* :163-167 */
    if (o->B <= Kcb) {
        o->C = 1;
        o->Kp = o->B;
    } else {
        o->C = (int)ceil((double)o->B / (Kcb - 24));
        if ((o->B + 24 * o->C) % o->C) return -1;                                  /* K' = B'/C exact */
        o->Kp = (o->B + 24 * o->C) / o->C;
    }
    if (o->bg == 1) o->Kb = 22;                                                     /* :182-198 */
    else o->Kb = o->B > 640 ? 10 : (o->B > 560 ? 9 : (o->B > 192 ? 8 : 6));
    int best = -1;                                                                  /* :207-211 */
    for (int i = 0; i < 51; ++i) {
        const int d = Zs[i] * o->Kb - o->Kp;
        if (d >= 0 && (best < 0 || d < Zs[best] * o->Kb - o->Kp)) best = i;
    }
    if (best < 0) return -1;
    o->Zc = Zs[best];
    o->ils = best < 8 ? 0 : best < 16 ? 1 : best < 23 ? 2 : best < 29 ? 3 : best < 35 ? 4 : best < 41 ? 5
                                                                                        : best < 46 ? 6 : 7;
    o->K = o->Zc * (o->bg == 1 ? 22 : 10);                                          /* :263-269 */
    o->F = o->K - o->Kp;
    o->N = o->Zc * (o->bg == 1 ? 66 : 50);
    if (i_lbrm) {                                                                   /* :276-289 */
        const double ninfo = 156.0 * n_prb_lbrm * max_qm * (948.0 / 1024.0) * (max_layers < 4 ? max_layers : 4);
        const double n = floor(log2(ninfo - 24.0)) - 5.0;
        const double Np = fmax(3840.0, pow(2.0, n) * round((ninfo - 24.0) / pow(2.0, n)));
        const double Cl = ceil((Np + 24.0) / 8424.0);
        const double tbs_lbrm = 8.0 * Cl * ceil((Np + 24.0) / (8.0 * Cl)) - 24.0;
        o->tbs_lbrm = (int)tbs_lbrm;
        o->Nref = (int)floor(tbs_lbrm / (o->C * (2.0 / 3.0)));
    }
    return 0;
}

/* derate_match.m:122-129: parity nodes covering the received LLRs,
 * nV_parity = clamp(ceil((min(floor(G/C) + k0, N) - (K - F - 2Zc)) / Zc), 4, 46/42),
 * N = 66 Zc (BG1) or 50 Zc (BG2). */
int nr_orc_num_parity_nodes(int bg, int Zc, int K, int F, int G, int C, int k0)
{
    const long long N = (bg == 1 ? 66LL : 50LL) * Zc;
    long long n = (long long)(G / C) + k0;
    if (n > N) n = N;
    const long long par = n - (K - F - 2LL * Zc);
    long long mb = (par + Zc - 1) / Zc; /* ceil for par > 0 */
    if (par <= 0) mb = 0;
    const int maxr = bg == 1 ? 46 : 42;
    if (mb < 4) mb = 4;
    if (mb > maxr) mb = maxr;
    return (int)mb;
}

/* derate_match.m:54-63 */
int nr_orc_rm_E(const nr_orc_rm_cfg* c, int r)
{
    const int nq = c->Nl * c->Qm;
    const int q = c->G / nq; /* G / (Nl*Qm), exact */
    if (r <= c->C - (q % c->C) - 1) return nq * (q / c->C);
    return nq * ((q + c->C - 1) / c->C);
}

#define LLR_MAX_ABS 10000.0f

/* build_Gold_sequence.m:16-48: x1(0) = 1, x2 = bits of c_init (LSB first),
 * x1(n+31) = x1(n+3) + x1(n), x2(n+31) = x2(n+3) + x2(n+2) + x2(n+1) + x2(n)
 * (mod 2), c(n) = x1(n+1600) + x2(n+1600). */
void nr_orc_gold(unsigned int cinit, int n0, int nbits, unsigned char* bits)
{
    const int NC = 1600, N = n0 + nbits;
    unsigned char* x1 = (unsigned char*)calloc((size_t)N + NC + 31, 1);
    unsigned char* x2 = (unsigned char*)calloc((size_t)N + NC + 31, 1);
    x1[0] = 1;
    for (int i = 0; i < 31; ++i) x2[i] = (unsigned char)((cinit >> i) & 1u);
    for (int n = 0; n < N + NC; ++n) {
        x1[n + 31] = x1[n + 3] ^ x1[n];
        x2[n + 31] = x2[n + 3] ^ x2[n + 2] ^ x2[n + 1] ^ x2[n];
    }
    for (int n = 0; n < nbits; ++n) bits[n] = x1[n0 + n + NC] ^ x2[n0 + n + NC];
    free(x1);
    free(x2);
}

/* Bit selection (derate_match.m:152-170, TS 38.212 5.4.2.1): circular-buffer
 * positions of the E selected bits, starting at k0 and skipping the filler
 * bits [Kd, Kd + F). */
void nr_orc_bit_select(int E, int Kd, int F, int k0, int Ncb, int* pos)
{
    int k = 0, j = 0;
    while (k < E) {
        const int idx = (k0 + j) % Ncb;
        if (idx < Kd || idx >= Kd + F) pos[k++] = idx;
        ++j;
    }
}

static void derate_match_impl(const nr_orc_rm_cfg* c, const float* llr_demap, float* harq, int fp16)
{
    const int twoZ = 2 * c->Zc;
    const int Kp = c->K - c->F;
    const int fill_lo = Kp - twoZ, fill_hi = c->K - twoZ; /* filler [lo, hi) in circular-buffer coords */
    size_t start = 0;
    /* descramble_bits.m:31: LLR x (1 - 2c) over the TB's G LLRs */
    unsigned char* cseq = NULL;
    if (c->descramble) {
        cseq = (unsigned char*)malloc((size_t)c->G);
        nr_orc_gold(c->cinit, 0, c->G, cseq);
    }
    for (int r = 0; r < c->C; ++r) {
        float* out = harq + (size_t)r * c->Ncb_padded;
        const int E = nr_orc_rm_E(c, r);
        const int EQ = E / c->Qm;
        if (c->ndi)
            for (int n = 0; n < c->Ncb_padded; ++n) out[n] = 0.0f;
        /* bit selection after de-interleaving (derate_match.m:138-170):
         * the k-th selected LLR is e_k = LLR_c[(k % EQ)*Qm + k / EQ] */
        int* pos = (int*)malloc((size_t)E * sizeof(int));
        nr_orc_bit_select(E, fill_lo, fill_hi - fill_lo, c->k0, c->Ncb, pos);
        for (int k = 0; k < E; ++k) {
            const int idx = pos[k];
            {
                const int kIdx = k / EQ, jIdx = k % EQ;
                const size_t n = start + (size_t)jIdx * c->Qm + kIdx;
                float v = llr_demap[n];
                if (cseq && cseq[n]) v = -v;
                float* o = &out[twoZ + idx];
                if (fp16) {
                    v = nr_orc_round_half(v);
                    if (v > LLR_MAX_ABS) v = LLR_MAX_ABS;
                    if (v < -LLR_MAX_ABS) v = -LLR_MAX_ABS;
                    float s = nr_orc_round_half(*o + v);
                    if (s > LLR_MAX_ABS) s = LLR_MAX_ABS;
                    if (s < -LLR_MAX_ABS) s = -LLR_MAX_ABS;
                    *o = s;
                } else {
                    if (v > LLR_MAX_ABS) v = LLR_MAX_ABS;
                    if (v < -LLR_MAX_ABS) v = -LLR_MAX_ABS;
                    float s = *o + v;
                    if (s > LLR_MAX_ABS) s = LLR_MAX_ABS;
                    if (s < -LLR_MAX_ABS) s = -LLR_MAX_ABS;
                    *o = s;
                }
            }
        }
        free(pos);
        for (int n = fill_lo; n < fill_hi; ++n) out[twoZ + n] = LLR_MAX_ABS; /* :179 */
        for (int n = 0; n < twoZ; ++n) out[n] = 0.0f;                        /* punctured */
        start += (size_t)E;
    }
    free(cseq);
}

void nr_orc_derate_match(const nr_orc_rm_cfg* c, const float* llr_demap, float* harq)
{
    derate_match_impl(c, llr_demap, harq, 0);
}

void nr_orc_derate_match_fp16(const nr_orc_rm_cfg* c, const float* llr_demap, float* harq)
{
    derate_match_impl(c, llr_demap, harq, 1);
}

/* ======================================================================
 * CB CRC / TB assembly / TB CRC (cuPHY crc.cu:55-237 conventions)
 * ==================================================================== */
#define WORDS_PER_CB 264

int nr_orc_crc_desegment(const uint32_t* cb_words, int C, int Kp,
                         uint32_t* cb_crc, uint8_t* tb_bytes, uint32_t* tb_crc)
{
    /* the CB CRC covers the decoded bits up to the next 32-bit boundary */
    const int nbits = ((Kp + 31) / 32) * 32;
    uint8_t* bits = (uint8_t*)malloc((size_t)nbits);
    int crc_bits;
    int type;
    if (C > 1) {
        crc_bits = 24;
        type = NR_CRC24B;
    } else {
        crc_bits = 0;
        type = ((Kp + 7) / 8 <= 480) ? NR_CRC16 : NR_CRC24A;
    }
    const int data_bytes = (Kp - crc_bits + 7) / 8;
    for (int r = 0; r < C; ++r) {
        const uint32_t* w = cb_words + (size_t)r * WORDS_PER_CB;
        for (int n = 0; n < nbits; ++n) bits[n] = (uint8_t)((w[n >> 5] >> (n & 31)) & 1u);
        cb_crc[r] = nr_orc_crc(bits, nbits, type);
        for (int b = 0; b < data_bytes; ++b) {
            uint8_t v = 0;
            for (int i = 0; i < 8; ++i) v = (uint8_t)((v << 1) | bits[8 * b + i]);
            tb_bytes[(size_t)r * data_bytes + b] = v;
        }
    }
    int total = C * data_bytes;
    while (total % 4) tb_bytes[total++] = 0;
    if (C == 1) {
        *tb_crc = cb_crc[0];
    } else {
        uint8_t* tbits = (uint8_t*)malloc((size_t)total * 8);
        for (int b = 0; b < total; ++b)
            for (int i = 0; i < 8; ++i) tbits[8 * b + i] = (uint8_t)((tb_bytes[b] >> (7 - i)) & 1u);
        *tb_crc = nr_orc_crc(tbits, total * 8, NR_CRC24A);
        free(tbits);
    }
    free(bits);
    return total;
}

/* ======================================================================
 * many code blocks on several host threads (CPU baseline): code block i
 * reads llr + i*llr_stride and writes hard + i*hard_stride; block i goes to
 * thread i % nthreads.
 * ==================================================================== */
/* the calling thread on the tid-th logical CPU this process may use */
static void pin_to(int tid)
{
    cpu_set_t allowed, one;
    if (sched_getaffinity(0, sizeof(allowed), &allowed) != 0) return;
    int n = 0;
    for (int c = 0; c < CPU_SETSIZE; ++c) {
        if (!CPU_ISSET(c, &allowed)) continue;
        if (n++ == tid) {
            CPU_ZERO(&one);
            CPU_SET(c, &one);
            pthread_setaffinity_np(pthread_self(), sizeof(one), &one);
            return;
        }
    }
}

typedef struct {
    const nr_orc_dec_cfg* cfg;
    const float* llr;
    size_t llr_stride;
    uint32_t* hard;
    size_t hard_stride;
    int* iters;
    int ncb, tid, nthreads, pin;
} par_job_t;

static void* par_worker(void* p)
{
    par_job_t* j = (par_job_t*)p;
    if (j->pin) pin_to(j->tid);
    for (int i = j->tid; i < j->ncb; i += j->nthreads) {
        const int it = nr_orc_decode_cb(j->cfg, j->llr + (size_t)i * j->llr_stride,
                                        j->hard ? j->hard + (size_t)i * j->hard_stride : NULL, NULL);
        if (j->iters) j->iters[i] = it;
    }
    return NULL;
}

int nr_orc_decode_cbs_it(const nr_orc_dec_cfg* cfg, const float* llr, size_t llr_stride, int ncb,
                         uint32_t* hard, size_t hard_stride, int* iters, int nthreads, int pin)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    par_job_t jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = (par_job_t){cfg, llr, llr_stride, hard, hard_stride, iters, ncb, t, nthreads, pin};
        if (pthread_create(&th[t], NULL, par_worker, &jobs[t]) != 0) return -1;
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    return 0;
}

int nr_orc_decode_cbs(const nr_orc_dec_cfg* cfg, const float* llr, size_t llr_stride, int ncb,
                      uint32_t* hard, size_t hard_stride, int nthreads)
{
    return nr_orc_decode_cbs_it(cfg, llr, llr_stride, ncb, hard, hard_stride, NULL, nthreads, 0);
}

/* ======================================================================
 * a slot of transport blocks (one decoder configuration each) on several
 * host threads: the code blocks of every TB form one queue, handed out one
 * at a time (an atomic counter), so that threads stay busy across TBs of
 * different sizes.
 * ==================================================================== */
typedef struct {
    const nr_orc_dec_cfg* cfgs;
    const float* const* llrs;
    const size_t* llr_strides;
    const int* first; /* ntb + 1 prefix of code-block counts */
    uint32_t* const* hards;
    const size_t* hard_strides;
    int ntb;
    int next; /* atomic */
} slot_queue_t;

typedef struct {
    slot_queue_t* q;
    int tid, pin;
} slot_job_t;

static void* slot_worker(void* p)
{
    slot_job_t* j = (slot_job_t*)p;
    slot_queue_t* q = j->q;
    if (j->pin) pin_to(j->tid);
    const int total = q->first[q->ntb];
    int t = 0;
    for (;;) {
        const int i = __atomic_fetch_add(&q->next, 1, __ATOMIC_RELAXED);
        if (i >= total) break;
        while (q->first[t + 1] <= i) ++t; /* i only grows within a thread */
        const int r = i - q->first[t];
        nr_orc_decode_cb(&q->cfgs[t], q->llrs[t] + (size_t)r * q->llr_strides[t],
                         q->hards[t] ? q->hards[t] + (size_t)r * q->hard_strides[t] : NULL, NULL);
    }
    return NULL;
}

int nr_orc_time_slot(const nr_orc_dec_cfg* cfgs, const float* const* llrs, const size_t* llr_strides,
                     const int* ncbs, int ntb, uint32_t* const* hards, const size_t* hard_strides, int nthreads,
                     int warmup, int reps, double* times)
{
    if (ntb < 1) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    int* first = (int*)malloc(sizeof(int) * (size_t)(ntb + 1));
    if (!first) return -1;
    first[0] = 0;
    for (int t = 0; t < ntb; ++t) first[t + 1] = first[t] + ncbs[t];
    slot_queue_t q = {cfgs, llrs, llr_strides, first, hards, hard_strides, ntb, 0};
    pthread_t th[256];
    slot_job_t jobs[256];
    int rc = 0;
    for (int i = 0; i < warmup + reps && rc == 0; ++i) {
        struct timespec a, b;
        q.next = 0;
        clock_gettime(CLOCK_MONOTONIC, &a);
        int started = 0;
        for (; started < nthreads; ++started) {
            jobs[started] = (slot_job_t){&q, started, 1};
            if (pthread_create(&th[started], NULL, slot_worker, &jobs[started]) != 0) {
                rc = -1;
                break;
            }
        }
        for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
        clock_gettime(CLOCK_MONOTONIC, &b);
        if (i >= warmup) times[i - warmup] = (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
    }
    free(first);
    return rc;
}

/* CPU-baseline timing: warmup + reps decodes of the ncb code blocks on
 * nthreads pinned threads; seconds of each timed decode (monotonic clock)
 * into times[reps].  Returns 0, or -1 if a thread could not start. */
int nr_orc_time_decode(const nr_orc_dec_cfg* cfg, const float* llr, size_t llr_stride, int ncb, uint32_t* hard,
                       size_t hard_stride, int nthreads, int warmup, int reps, double* times)
{
    for (int i = 0; i < warmup + reps; ++i) {
        struct timespec a, b;
        clock_gettime(CLOCK_MONOTONIC, &a);
        if (nr_orc_decode_cbs_it(cfg, llr, llr_stride, ncb, hard, hard_stride, NULL, nthreads, 1) != 0) return -1;
        clock_gettime(CLOCK_MONOTONIC, &b);
        if (i >= warmup) times[i - warmup] = (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
    }
    return 0;
}
