/*
 * wv_oracle.h -- CPU restatement of the reference (oshkosher/Wavelets) CPU
 * compression path, used ONLY as a test oracle and as the reported CPU
 * baseline.  Nothing in the product (libwv_hip.so, the drop-in driver) links
 * or calls this code; tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg are the only callers.
 *
 * Every function names the reference file:line it restates.  The code is
 * written from the reference's semantics, in plain C, compiled with
 * -O2 -ffp-contract=off (the reference is plain g++ -O2 on x86-64, which
 * emits no FMA: Makefile:88,122), so float results follow the reference's
 * operation order exactly.
 *
 * Pinning status of each function (how a plausible mistake would be caught)
 * is listed in oracle/README.md.
 */
#ifndef WV_ORACLE_H
#define WV_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* enums: internal ids as in wavelet.h:25-46 */
enum { OR_QUANT_UNIFORM = 1, OR_QUANT_LOG = 2, OR_QUANT_LLOYD = 3 };
enum { OR_WAVELET_HAAR = 1, OR_WAVELET_CDF97 = 2 };
enum { OR_DTYPE_UINT8 = 1, OR_DTYPE_INT32 = 2, OR_DTYPE_FLOAT32 = 3 };

/* ---- a1: steps and padding (dwt_cpu.cc:61-84,212-219,
        test_compress_common.cc:362-382) ---- */
unsigned or_count_leading_zeros(unsigned x);
unsigned or_ceil_log2(unsigned x);
int or_padded_length(int length, int steps);
void or_set_steps(int steps[3], const int size[3]);

/* REPEAT padding + u8/i32/f32 -> float (test_compress_common.cc:386-488).
   in: original cube (size orig[3], x fastest); out: padded[3] floats. */
void or_pad_convert(const void *in, int dtype, const int orig[3],
                    float *out, const int padded[3]);

/* ---- a4/a5/a6: 1-D transforms (dwt_cpu.cc:139-203,683-799) ---- */
void or_cdf97_fwd(int length, float *data, int steps, float *tmp);
void or_cdf97_inv(int length, float *data, int steps, float *tmp);
void or_haar_fwd(int length, float *data, int steps, float *tmp);
void or_haar_inv(int length, float *data, int steps, float *tmp);

/* ---- a2/a3: rotations and the 3-D driver (wavelet.h:864-912,
        dwt_cpu.cc:624-648,943-985, test_compress_common.cc:492-500) ----
   data has size[0]*size[1]*size[2] floats; size is rotated in place exactly
   like Cube::size.  tmp must hold as many floats as data. */
void or_transpose3d_fwd(const float *src, float *dst, const int size[3]);
void or_transpose3d_back(const float *src, float *dst, const int size[3]);
void or_dwt3d(float *data, float *tmp, int size[3], int wavelet,
              const int steps[3], int inverse);
/* FP64 lifting filter bank over a padded [z][y][x] volume, no rotation
   (or_dwt64.c): the yardstick for the float transforms' error */
void or_dwt3d_f64(const float *in, const int size[3], int wavelet, const int steps[3],
                  double *out);

/* ---- a8: threshold (thresh_cpu.cc:80-109) ----
   sorted_abs (optional, count floats) receives the ascending |x|.
   Returns the threshold value; *thresh_idx = (int)(frac*count). */
float or_thresh(const float *data, int count, float frac, int *nonzero_count,
                float *max_val, float *min_val, float *max_abs,
                int *thresh_idx, float *sorted_abs);
/* ascending sort of non-negative floats (any correct sort gives the unique
   result std::sort gives at thresh_cpu.cc:94) */
void or_sort_abs(const float *data, int count, float *sorted_abs);

/* ---- a9/a10/a11: quantizers (quant.h:189-558) ---- */
typedef struct {
  int alg, bin_count;
  float threshold, max_val;
  /* uniform / log derived parameters (quant.h:196-242, 337-372) */
  float neg_scale, neg_inv_scale, neg_offset;
  float pos_scale, pos_inv_scale, pos_offset;
  /* codebook (quant.h:467-558) */
  int n_boundaries;
  const float *boundaries;
  const float *values;
} or_quantizer;

void or_quant_init(or_quantizer *q, int alg, int bin_count, float threshold,
                   float max_val, const float *boundaries,
                   const float *values);
int or_quant(const or_quantizer *q, float x);
float or_dequant(const or_quantizer *q, int bin);
void or_quantize(const or_quantizer *q, const float *in, int32_t *out,
                 int count);
void or_dequantize(const or_quantizer *q, const int32_t *in, float *out,
                   int count);

/* ---- a12: Lloyd (lloyds.cpp:30-141, test_compress_cpu.cc:404-459,
        test_compress_common.cc:749-846) ----
   max_iters < 0 means no cap (the reference has none). */
void or_lloyd(const float *points, unsigned psize, float *codebook,
              unsigned csize, float *partition, float *dist, float *reldist,
              unsigned *groups, float stop_criteria, int *iterations,
              int max_iters);
void or_initial_lloyd_codebook(float *codebook, int cs, float min_abs,
                               float max_abs);
/* bin_values (bin_count), bin_boundaries (bin_count-1) */
int or_set_bins_from_codebook(float *bin_values, float *bin_boundaries,
                              int bin_count, const float *codebook, int cs,
                              float threshold, float min_val, float max_val);
/* whole computeLloydQuantization; nonzero = sorted + idx + 1 */
int or_compute_lloyd_quantization(const float *nonzero, int count,
                                  int bin_count, float min_val, float max_val,
                                  float threshold, float *bin_boundaries,
                                  float *bin_values, int max_iters,
                                  int *iters_out);

/* ---- a14/a15: Huffman + dup strings + bit stream (huffman.h/.cc,
        bit_stream.h, test_compress_common.cc:897-994) ---- */
typedef struct or_huffman or_huffman;
/* counts: bin_count entries; zero_bin < 0 disables the dup encoding.
   The histogram is built from data exactly as huffScan() does. */
or_huffman *or_huff_create_scan(const int32_t *data, int count,
                                int bin_count, int zero_bin);
/* histogram given (GPU path: initHuffman(freqCounts, zeroBin) +
   huffScanDupsOnly) */
or_huffman *or_huff_create_counts(const int32_t *counts, const int32_t *data,
                                  int count, int bin_count, int zero_bin);
void or_huff_free(or_huffman *h);
/* counts after the scan: bin_count entries + dup count at [bin_count] */
void or_huff_get_counts(const or_huffman *h, int64_t *counts_out);
int or_huff_code_length(const or_huffman *h, int value);
int or_huff_longest(const or_huffman *h);
/* Huffman::totalEncodedLengthBytes (huffman.cc:241-252) */
int64_t or_huff_total_encoded_bytes(const or_huffman *h);
/* decode table (BFS order pairs); returns its length in ints */
int or_huff_decode_table(const or_huffman *h, int32_t *table, int cap);
/* encodes data; returns number of 32-bit words written (cap in words) or
   -1 when cap is too small.  *bits_out = bits written. */
int64_t or_huff_encode(const or_huffman *h, const int32_t *data, int count,
                       uint32_t *words, int64_t cap, int64_t *bits_out);
/* decode with a decode table (HuffmanDecoder::decodeFromStream) */
int or_huff_decode(const int32_t *table, int table_len, int dup_key,
                   int dup_value, const uint32_t *words, int64_t n_words,
                   int32_t *out, int count);
/* BitStreamWriter::write(unsigned,int) sequence for the bit-stream KATs */
int or_bitstream_pack(const uint32_t *values, const int32_t *nbits, int n,
                      uint32_t *words, int cap);

/* ---- a17: error metrics (test_compress_common.h:90-160,
        test_compress_common.cc:616-694) ---- */
typedef struct {
  double sum_value, sum_value_sq, sum_diff, sum_diff_sq;
  float max_value;
  int count;
} or_error_acc;
void or_error_compare(or_error_acc *acc, const void *orig, int dtype,
                      const int orig_size[3], const float *restored,
                      const int padded[3], int max_possible);
float or_error_mse(const or_error_acc *acc);
float or_error_psnr(const or_error_acc *acc);

/* ---- a16: .cube writer (cubelet_file.cc:18-84,259-327,
        wavelet.cc:283-343, wavelet_compress.proto) ---- */
typedef struct {
  int width, height, depth;             /* rotated, padded size */
  int orig_width, orig_height, orig_depth;
  int orig_dtype;                       /* OR_DTYPE_* or -1 */
  int steps[3];
  int wavelet;                          /* OR_WAVELET_* */
  float threshold_fraction, threshold_value;
  int bin_count, quant_alg;             /* OR_QUANT_* */
  float max_value;
  int n_boundaries, n_values;
  const float *boundaries, *values;
  int n_decode;
  const int32_t *decode_table;
  int dup_key, dup_value;               /* -1 = unset */
  int64_t byte_count;
  int max_possible;                     /* maximum_value, 0 = unset */
} or_cube_meta;
/* Writes header + one compressed cubelet + optional footer into buf.
   Returns the byte length or -1 if cap is too small. */
int64_t or_cube_write(const or_cube_meta *m, const uint32_t *payload,
                      int with_footer, uint8_t *buf, int64_t cap);
/* serialize only the CubeletBuffer metadata */
int or_cube_meta_bytes(const or_cube_meta *m, int64_t data_file_offset,
                       uint8_t *buf, int cap);

/* ---- whole pipeline (test_compress_cpu.cc:68-184 / the GPU driver
        test_compress_gpu.cu:150-436 with its footer) ---- */
typedef struct {
  int wavelet, quant_alg, bin_count;
  int steps[3];                         /* -1 = max */
  float thresh_frac;
  int compress_zeros;                   /* 0 = -noz */
  int lloyd_max_iters;                  /* <0 = no cap */
  int with_footer;                      /* GPU driver writes it */
} or_options;

typedef struct {
  int size[3];                          /* rotated padded size */
  int steps[3];                         /* clamped */
  float threshold, min_val, max_val, max_abs;
  int thresh_idx, nonzero_count, zero_bin;
  int lloyd_iters;
  int64_t compressed_bytes;             /* Huffman payload bytes */
  int64_t file_bytes;
  double t_dwt, t_thresh, t_quant, t_huff, t_total;   /* seconds */
} or_result;

/* Full compress of one cubelet.  Optional outputs: coef (padded count
   floats, transformed), bins (padded count ints), file (cap bytes). */
int or_compress(const void *in, int dtype, const int orig[3], int max_possible,
                const or_options *opt, or_result *res, float *coef,
                int32_t *bins, uint8_t *file, int64_t file_cap);

/* seconds, CLOCK_MONOTONIC (nixtimer.h:27-35) */
double or_now(void);

#ifdef __cplusplus
}
#endif
/* cube split of one large cubelet (cube_split_typed cubelet_convert.cc:424-553):
   boxes as 6 ints (offset x,y,z, size x,y,z) in write order; optional
   packed copies of the boxes' samples */
int or_split(const int size[3], const int cub[3], int overlap, int *boxes, int cap,
             const void *vol, int esize, void *out);

/* one test point of the parameter optimisation (testParameters,
   optimize_gpu.cu:52-202, with the CPU path's double error sums): forward
   transform of the original, then bin_count <= 0: |x| <= threshold -> 0;
   else quantize (uniform / log, max = max |x|; Lloyd: computeLloydQuantization
   over the sorted |x| from findSorted(threshold), no iteration cap, its
   iterations to *lloyd_iters when non-null), Huffman size
   (totalEncodedLengthBytes), dequantize; inverse; error over the original
   extent.  Returns the size in bytes (0 without quantization). */
int64_t or_opt_test(const void *in, int dtype, const int orig[3], const int steps[3],
                    int wavelet, int max_possible, int compress_zeros, float threshold,
                    int bin_count, int quant_alg, or_error_acc *acc, int *lloyd_iters);
/* the sorted |x| of the transformed data (getSorted's array), n entries */
int or_opt_sorted(const void *in, int dtype, const int orig[3], const int steps[3],
                  int wavelet, float *sorted_out, int64_t cap);

#endif
