/*
 * CPU oracle for the 5G NR uplink LDPC decode path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against; it is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg, and never by
 * the product library.
 *
 * It restates, in plain C, the reference's CPU decode path (the 5GModel
 * MATLAB files) plus the cuPHY boundary conventions:
 *   - layered normalized min-sum:  5GModel/nr_matlab/pxsch/MSA_layering.m:16-140
 *       compute_V2C.m:42-51, compute_cC2V.m:37-55, update_C2V.m:36-47,
 *       update_APP.m:37-42, LDPC_decode.m:16-96 (punctured-column insertion,
 *       alpha table :101-208)
 *   - syndrome (PCE) check:        syndrome_check.m:16-54, compute_check.m:16-52
 *   - descrambling (Gold sequence): build_Gold_sequence.m:16-48 and
 *                                  descramble_bits.m:16-33 (LLR x (1 - 2c))
 *   - de-rate-matching + HARQ:     derate_match.m:16-196; output layout and
 *       input clamp from cuPHY/src/cuphy/rate_matching/rate_matching.cu:112-177,
 *       622-683 (HARQ buffer = [2Zc punctured zeros | Ncb circular buffer])
 *   - CRC:                         TS 38.212 5.1 polynomials as in
 *       cuPHY/src/cuphy/crc/crc.hpp:29-32 (bit-serial computeCRC :52-76);
 *       CB/TB desegmentation block_desegment.m:16-61, CRC_decode.m:16-60,
 *       TB byte assembly cuPHY/src/cuphy/crc/crc.cu:55-181 (reverseBytes=1)
 *   - LDPC encoding (test inputs): TS 38.212 5.3.2 (systematic encoding of
 *       H*c = 0), cf. 5GModel/.../standalone_LDPC_encoder_test/LDPC_encode_main.m
 *
 * Parity status of each function is stated next to it and summarised in the
 * repository's design notes.
 */
#pragma once
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic modes of the layered decoder. */
enum {
    NR_ORC_FP32 = 0, /* IEEE binary32, every op rounded RN (the exact mode) */
    NR_ORC_FP64 = 1, /* IEEE binary64 (MATLAB's native precision)           */
    NR_ORC_FP16 = 2  /* IEEE binary16 storage and arithmetic, RN per op      */
};

/* ---- lifting sizes / base graphs ------------------------------------- */
int  nr_orc_ils(int Z);                         /* lifting-set index 0..7 or -1 */
int  nr_orc_num_rows(int bg);                   /* 46 / 42                      */
int  nr_orc_num_sys_cols(int bg);               /* 22 / 10                      */
int  nr_orc_row_degree(int bg, int r);
/* column and shift (already reduced mod Z) of edge j of row r */
int  nr_orc_edge(int bg, int Z, int r, int j, int* col, int* shift);

/* ---- random numbers (xoshiro256** seeded by splitmix64) --------------- */
typedef struct { uint64_t s[4]; } nr_orc_rng;
void     nr_orc_rng_seed(nr_orc_rng* g, uint64_t seed);
uint64_t nr_orc_rng_next(nr_orc_rng* g);
double   nr_orc_rng_uniform(nr_orc_rng* g);      /* (0, 1]                       */
double   nr_orc_rng_normal(nr_orc_rng* g);       /* Box-Muller, one value/call   */
void     nr_orc_random_bits(uint64_t seed, uint8_t* bits, int n);

/* ---- channel models producing decoder LLRs ----------------------------- */
/* BPSK x = 1-2b over real AWGN of variance sigma2; llr = 2y/sigma2 */
void nr_orc_bpsk_awgn_llr(const uint8_t* bits, int n, double sigma2, uint64_t seed, float* llr);
/* QPSK (unit energy, Gray) at complex Es/N0 = snr_db: per bit y = a*x + n,
 * a = 1/sqrt(2), var(n) = N0/2, N0 = 10^(-snr/10); llr = 2*a*y/(N0/2) */
void nr_orc_qpsk_awgn_llr(const uint8_t* bits, int n, double snr_db, uint64_t seed, float* llr);
/* "consistent Gaussian" LLRs: llr = mu*x + sqrt(2 mu)*n */
void nr_orc_consistent_gaussian_llr(const uint8_t* bits, int n, double mu, uint64_t seed, float* llr);

/* ---- LDPC encoder ------------------------------------------------------ */
/* info: nsys*Z bits (0/1, filler bits as 0); cw: (nsys+nrows)*Z bits.
 * Returns 0 on success.  Parity: pinned by syndrome == 0 (nr_orc_syndrome). */
int nr_orc_encode(int bg, int Z, const uint8_t* info, uint8_t* cw);
/* number of unsatisfied checks of rows [0, mb) for a hard word x (0/1) */
int nr_orc_syndrome(int bg, int Z, int mb, const uint8_t* x);

/* ---- layered normalized min-sum decoder (MSA_layering.m) --------------- */
typedef struct {
    int   bg, Z, mb;       /* base graph, lifting size, parity rows used      */
    int   kb_out;          /* info columns written to hard/soft output        */
    int   max_iter;        /* iteration budget                                */
    int   early_term;      /* 1: full-PCE syndrome check after each iteration */
    int   mode;            /* NR_ORC_FP32 / FP64 / FP16                       */
    float alpha;           /* normalisation                                   */
    float clamp;           /* input clamp (|llr| <= clamp); <= 0 disables     */
} nr_orc_dec_cfg;

/* llr: (nsys + mb)*Z values, first 2Z = punctured positions (cuPHY layout).
 * hard: ceil(kb_out*Z/32) words, bit i of word w = (APP[32w+i] <= 0).
 * app_out (optional): (nsys+mb)*Z final APP values (as float).
 * Returns the number of iterations run. */
int nr_orc_decode_cb(const nr_orc_dec_cfg* cfg, const float* llr, uint32_t* hard, float* app_out);

/* ncb code blocks, block i at llr + i*llr_stride -> hard + i*hard_stride,
 * spread over nthreads host threads (CPU baseline) */
int nr_orc_decode_cbs(const nr_orc_dec_cfg* cfg, const float* llr, size_t llr_stride, int ncb,
                      uint32_t* hard, size_t hard_stride, int nthreads);

/* the same with the iterations run per code block (iters may be NULL);
 * pin != 0: thread t runs on the t-th CPU this process may use */
int nr_orc_decode_cbs_it(const nr_orc_dec_cfg* cfg, const float* llr, size_t llr_stride, int ncb,
                         uint32_t* hard, size_t hard_stride, int* iters, int nthreads, int pin);

/* CPU-baseline timing of a slot: TB t (configuration cfgs[t], ncbs[t] code
 * blocks at llrs[t] + r*llr_strides[t], hard decisions to hards[t] (or
 * NULL)); all code blocks on nthreads pinned threads from one queue;
 * warmup + reps passes, seconds of each timed pass into times[reps] */
int nr_orc_time_slot(const nr_orc_dec_cfg* cfgs, const float* const* llrs, const size_t* llr_strides,
                     const int* ncbs, int ntb, uint32_t* const* hards, const size_t* hard_strides, int nthreads,
                     int warmup, int reps, double* times);

/* CPU-baseline timing: warmup + reps decodes of the ncb code blocks on
 * nthreads pinned threads, seconds of each timed decode into times[reps] */
int nr_orc_time_decode(const nr_orc_dec_cfg* cfg, const float* llr, size_t llr_stride, int ncb, uint32_t* hard,
                       size_t hard_stride, int nthreads, int warmup, int reps, double* times);

/* one check-node update (compute_cC2V.m + update_C2V.m) on d inputs t[] in
 * the given mode; c2v[j] = sign * (j == argmin ? min2 : min1), unnormalised */
void nr_orc_check_node(int d, const float* t, float* c2v, int mode);

/* IEEE binary16 round-to-nearest-even of a float (returned as float) */
float nr_orc_round_half(float x);

/* ---- CRC --------------------------------------------------------------- */
enum { NR_CRC24A = 0, NR_CRC24B = 1, NR_CRC16 = 2 };
/* remainder of bits[0..n) (first bit = highest power) divided by g(D) */
uint32_t nr_orc_crc(const uint8_t* bits, int n, int type);

/* ---- de-rate-matching with HARQ combining (derate_match.m) ------------- */
typedef struct {
    int bg, Zc, C, Qm, Nl;   /* base graph, lifting, CBs, bits/symbol, layers */
    int G;                   /* rate-matched bits of the TB                   */
    int K, F;                /* systematic bits incl. punctured 2Zc, filler   */
    int Ncb, Ncb_padded;     /* circular buffer length, HARQ stride per CB    */
    int k0;                  /* starting position for the rv                  */
    int ndi;                 /* 1 = new data (overwrite), 0 = combine         */
    int descramble;          /* 1 = LLR n of the TB times (1 - 2 c(n))        */
    unsigned int cinit;      /* Gold sequence initialisation c_init            */
} nr_orc_rm_cfg;
/* Gold sequence bits c(n0 .. n0+nbits-1) of TS 38.211 5.2.1 for c_init, bit
 * per byte (build_Gold_sequence.m:16-48, bit-serial). */
void nr_orc_gold(unsigned int cinit, int n0, int nbits, unsigned char* bits);
/* E of code block r (TS 38.212 5.4.2.1) */
int  nr_orc_rm_E(const nr_orc_rm_cfg* c, int r);
/* llr_demap: G LLRs of the TB (layer-demapped, CB after CB, Qm-interleaved).
 * harq: C*Ncb_padded values updated in place. */
void nr_orc_derate_match(const nr_orc_rm_cfg* c, const float* llr_demap, float* harq);
/* same, HARQ buffer held in binary16 (every stored value rounded RN) */
void nr_orc_derate_match_fp16(const nr_orc_rm_cfg* c, const float* llr_demap, float* harq);
/* ---- TB size and coding parameters (derive_coding_main.m:78-301) ------- */
typedef struct {
    int tbs, bg, crc_len, B, C, Kp, Kb, Zc, ils, K, F, N;
    int tbs_lbrm, Nref;      /* limited buffer: TBS_LBRM and N_ref (0 without) */
} nr_orc_coding;
/* entry i of TS 38.214 Table 5.1.3.2-1 (-1 outside 0..92) */
int nr_orc_tbs_table(int i);
/* TBS of an allocation (N_RE = min(156, re_per_prb) * n_prb, rate R) */
int nr_orc_tb_size(double re_per_prb, int n_prb, double R, int Qm, int nl);
/* coding parameters of a TB of tbs payload bits; returns 0, or -1 if no
 * lifting size fits */
int nr_orc_derive_coding(int tbs, double R, int i_lbrm, int n_prb_lbrm, int max_layers, int max_qm,
                         nr_orc_coding* out);

/* circular-buffer positions of the E selected bits from k0, filler bits
 * [Kd, Kd + F) skipped (derate_match.m:152-170) */
void nr_orc_bit_select(int E, int Kd, int F, int k0, int Ncb, int* pos);
/* parity nodes the decoder uses for a transmission (derate_match.m:122-129) */
int  nr_orc_num_parity_nodes(int bg, int Zc, int K, int F, int G, int C, int k0);
/* TS 38.212 5.4.2.1 k0 */
int  nr_orc_k0(int bg, int rv, int Ncb, int Zc);

/* ---- CB/TB CRC check + TB assembly (block_desegment.m, CRC_decode.m) --- */
/* cb_words: C code blocks, stride 264 words, bit i of word w = bit 32w+i.
 * Kp = K - F (bits incl. CB CRC).  Writes per-CB remainder (0 = pass),
 * TB bytes (MSB-first per byte, CB data bytes back to back, zero padded to a
 * multiple of 4) and the TB remainder.  Returns the TB byte count written. */
int nr_orc_crc_desegment(const uint32_t* cb_words, int C, int Kp,
                         uint32_t* cb_crc, uint8_t* tb_bytes, uint32_t* tb_crc);

#ifdef __cplusplus
}
#endif
