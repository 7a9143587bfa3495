/*
 * wv_hip.h -- C ABI of libwv_hip.so, the MI355X (gfx950) implementation of
 * the Wavelets compression hot path: multi-level CDF 9/7 / Haar DWT over a
 * cubelet, threshold selection, uniform / log / Lloyd quantization with a
 * fused bin histogram, and the host Huffman + .cube writer that consume the
 * bins and counts.
 *
 * Conventions
 *  - plain pointers and sizes, no C++ or torch types; every call returns an
 *    int status (WV_OK or a negative WV_ERR_*) and never exits; the message
 *    of the last failure on the calling thread is wv_last_error().
 *  - device buffers are owned by the caller; every device call takes an
 *    explicit hipStream_t (passed as void*, 0 = the null stream) and is
 *    asynchronous unless its comment says it synchronises.
 *  - "geometry" sizes are (x, y, z) with x fastest, as Cube::size
 *    (wavelet.h:305-435).
 *
 * Each entry point cites the reference interface it replaces.
 */
#ifndef WV_HIP_H
#define WV_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WV_OK 0
#define WV_ERR_ARG -1          /* invalid argument / shape */
#define WV_ERR_HIP -2          /* HIP runtime error */
#define WV_ERR_UNSUPPORTED -3  /* shape or option outside what is built */
#define WV_ERR_NOMEM -4        /* workspace too small / budget exceeded */
#define WV_ERR_STATE -5        /* call order / busy */

/* ids follow wavelet.h:25-46 */
#define WV_QUANT_UNIFORM 1
#define WV_QUANT_LOG 2
#define WV_QUANT_LLOYD 3
#define WV_WAVELET_HAAR 1
#define WV_WAVELET_CDF97 2
#define WV_DTYPE_UINT8 1
#define WV_DTYPE_INT32 2
#define WV_DTYPE_FLOAT32 3
/* 16-bit signed input, read as such by the first pass and stored as an INT32
   cubelet (the format has no int16; shortImagesToCubelets.cc:9-27 widens
   16-bit data to INT32 the same way).  No counterpart in WaveletDataType. */
#define WV_DTYPE_INT16 4

const char *wv_last_error(void);
/* library version string, e.g. "wv_hip 1.0 gfx950" */
const char *wv_version(void);

/* ------------------------------------------------------------------ *
 * Geometry of one cubelet.  Replaces readData's size handling:
 * setWaveletSteps + padDataSize (test_compress_common.cc:362-382).
 * ------------------------------------------------------------------ */
typedef struct {
  int orig[3];       /* input size before padding */
  int steps[3];      /* transform steps per axis after clamping */
  int padded[3];     /* padded size (multiple of 2^steps per axis) */
  int out_size[3];   /* size of the transformed cube (rotated: z,x,y) */
  int wavelet;       /* WV_WAVELET_* */
  int dtype;         /* WV_DTYPE_* of the input */
  int64_t count;     /* padded element count */
  int64_t ws_floats; /* workspace floats wv_dwt3d needs */
} wv_geom;

/* steps < 0 or > ceil(log2(len)) become ceil(log2(len)). */
int wv_geom_init(wv_geom *g, const int orig[3], const int steps[3],
                 int wavelet, int dtype);

/* ------------------------------------------------------------------ *
 * Device statistics of a coefficient array (wv_stats_compute, or
 * wv_dwt3d_forward with a stats block): a 2048-bin histogram of bits
 * 30..20 of |x|, the signed extremes and max |x|.  One WV_STATS_BYTES
 * device block.  The threshold select does not need it.
 * ------------------------------------------------------------------ */
#define WV_STATS_BYTES 8256

int wv_stats_clear(void *stats_d, void *stream);
/* standalone statistics pass for coefficients that did not come out of
   wv_dwt3d_forward */
int wv_stats_compute(const float *coef_d, int64_t n, void *stats_d,
                     void *stream);

/* ------------------------------------------------------------------ *
 * Forward 3-D transform.  Replaces waveletTransformGPU(..., inverse=false)
 * (test_compress_gpu.cu:45-49,504-593; the math of cdf97_3d / haar_3d,
 * dwt_cpu.cc:624-648,943-965).  Reads the unpadded input in its own dtype
 * (u8 / i32 / f32, converted and REPEAT-padded on load as
 * translateCubeDataToFloat + padData do, test_compress_common.cc:386-488),
 * runs every level of every axis, and writes out_d in the rotated layout
 * (size g->out_size).  ws_d must hold g->ws_floats floats.  stats_d (may
 * be NULL) receives the coefficient statistics (an extra read pass).
 * Coefficients are bit-identical to the reference CPU transform.
 * ------------------------------------------------------------------ */
int wv_dwt3d_forward(const wv_geom *g, const void *in_d, float *out_d,
                     float *ws_d, void *stats_d, void *stream);

/* Same, recording a hipEvent_t (passed as void*) before the first pass and
   after every pass: events must hold wv_dwt3d_num_passes(g)+1 events.
   Lets a caller time each fused axis pass on its own stream. */
int wv_dwt3d_forward_ev(const wv_geom *g, const void *in_d, float *out_d,
                        float *ws_d, void *stats_d, void **events,
                        void *stream);
int wv_dwt3d_num_passes(const wv_geom *g);

/* thin wrappers over hipEvent_t for callers that time kernels on the
   library's streams (no reference counterpart; CudaTimer, cuda_timer.h:22-117,
   plays this role in the reference) */
int wv_event_create(void **ev);
int wv_event_destroy(void *ev);
int wv_event_record(void *ev, void *stream);
int wv_event_elapsed_ms(void *a, void *b, float *ms);

/* Inverse 3-D transform (waveletTransformGPU(..., inverse=true); the math
   of cdf97_inverse / haar_inv, dwt_cpu.cc:174-203,757-799,966-984).
   coef_d is in g->out_size layout; out_d receives the padded cube in x,y,z
   layout.  ws_d holds g->ws_floats floats.  coef_d is not modified. */
int wv_dwt3d_inverse(const wv_geom *g, const float *coef_d, float *out_d,
                     float *ws_d, void *stream);

/* ------------------------------------------------------------------ *
 * Threshold.  Replaces sortAbsAndFindLimitsGPU + the selection at
 * test_compress_gpu.cu:295-308,632-662 (thrust::sort of |x|) with a radix
 * select; semantics of thresh_cpu (thresh_cpu.cc:80-109): index
 * (int)(frac*(float)n), threshold = the idx-th smallest |x| (0 -> 1e-16),
 * nonzero = n-idx-1, min/max with FLT_MAX / FLT_MIN starts.
 * ------------------------------------------------------------------ */
typedef struct {
  float threshold;
  float min_val, max_val, max_abs;
  int32_t thresh_idx, nonzero_count;
  uint32_t threshold_key;   /* bit pattern of the threshold before 0->1e-16 */
  int32_t n_candidates;     /* internal: size of the last candidate set */
} wv_thresh_result;

/* workspace bytes for wv_threshold on n coefficients */
int64_t wv_threshold_ws_bytes(int64_t n);
/* stats_d is unused (may be NULL); result_d is a device wv_thresh_result.
   Does not synchronise. */
int wv_threshold(const float *coef_d, int64_t n, float frac,
                 const void *stats_d, void *ws_d, wv_thresh_result *result_d,
                 void *stream);

/* wv_threshold without a workspace argument and without allocating: no
   counterpart in the reference.  For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops).  The select's
   state and candidate list live in a static device array of the library
   (one per GPU, about 16 MiB, held until the process ends), so n is at most
   WV_NOALLOC_MAX_N (WV_ERR_ARG beyond).  Every call on a GPU shares that
   GPU's array: calls that overlap on one GPU (other streams, other host
   threads) race and must be serialised by the caller.  Same results as
   wv_threshold; does not synchronise. */
#define WV_NOALLOC_MAX_N 4194304
int wv_threshold_noalloc(const float *coef_d, int64_t n, float frac,
                         wv_thresh_result *result_d, void *stream);

/* ------------------------------------------------------------------ *
 * Quantizer parameters, computed on the host with the host libm exactly
 * as QuantUniform::init / QuantLog::init / QuantCodebook::init
 * (quant.h:196-242,337-372,487-492).  For the log quantizer the host also
 * tabulates the exact bin edges (the smallest float of every bin), so the
 * device assigns bins by comparison and reproduces the host logf bit for
 * bit without evaluating a logarithm.
 * ------------------------------------------------------------------ */
#define WV_MAX_BINS 65536
typedef struct {
  int alg, bin_count, zero_bin;
  float threshold, max_val;
  float neg_scale, neg_inv_scale, neg_offset;
  float pos_scale, pos_inv_scale, pos_offset;
  int n_pos_edges, n_neg_edges;  /* log: edge counts */
  float pos_clamp;               /* log: first x whose bin is clamped */
  int n_boundaries;              /* codebook: bin_count-1 */
  const float *pos_edges;        /* host arrays owned by the params */
  const float *neg_edges;
  const float *boundaries;       /* codebook boundaries (bin_count-1) */
  const float *values;           /* codebook values (bin_count) */
  const float *table_d;          /* device copy of the tables, set by
                                    wv_quant_params_upload */
  void *priv;
} wv_quant_params;

int wv_quant_params_init(wv_quant_params *p, int alg, int bin_count,
                         float threshold, float max_val,
                         const float *boundaries, const float *values);
void wv_quant_params_free(wv_quant_params *p);
/* device bytes of the parameter tables (edges or boundaries, then the
   bin_count dequantization values) */
int64_t wv_quant_table_bytes(const wv_quant_params *p);
/* copies the tables into table_d (wv_quant_table_bytes bytes, caller-owned
   device memory) on the stream and records it in p->table_d */
int wv_quant_params_upload(wv_quant_params *p, void *table_d, void *stream);
/* host quantization of one value (Quantizer::quant, quant.h:33-37) */
int wv_quant_host(const wv_quant_params *p, float x);
float wv_dequant_host(const wv_quant_params *p, int bin);

/* ------------------------------------------------------------------ *
 * Quantize + histogram.  Replaces quantizeGPU (quant_gpu.cu:122-182, the
 * thrust::transform with Quant*Functor) fused with computeFrequenciesGPU
 * (histogram_gpu.cu:117-179): bins_d[i] = quant(coef_d[i]) and
 * counts_d[b] = #{i : bins_d[i] == b}.  counts_d (bin_count + 1 uint32;
 * the extra entry counts bins outside [0, bin_count), which the reference
 * log quantizer can produce for |x| above its max) is cleared by the call.
 * bins_d may alias coef_d.
 * ------------------------------------------------------------------ */
int wv_quantize_hist(const float *coef_d, int64_t n, const wv_quant_params *p,
                     int32_t *bins_d, uint32_t *counts_d, void *stream);

/* Device-parameter uniform quantization: the threshold and max |x| are
   read from result_d on the device (no host round trip).  The parameter
   arithmetic is the same float expressions as QuantUniform::init. */
int wv_quantize_hist_uniform_dev(const float *coef_d, int64_t n,
                                 int bin_count,
                                 const wv_thresh_result *result_d,
                                 int32_t *bins_d, uint32_t *counts_d,
                                 void *stream);

/* Dequantize (dequantizeGPU, quant_gpu.cu:237-300) through the uploaded
   dequantization table. */
int wv_dequantize(const int32_t *bins_d, int64_t n, const wv_quant_params *p,
                  float *out_d, void *stream);

/* ------------------------------------------------------------------ *
 * Lloyd-Max codebook on the device.  Replaces computeLloydQuantizationGPU
 * / cudaLloyd (quant_gpu.cu:185-233, cudalloyds.cu:85-194) with the CPU
 * algorithm's semantics (computeLloydQuantization test_compress_cpu.cc:
 * 404-459, lloyd lloyds.cpp:30-141): points are the sorted |x| above rank
 * thresh_idx without the largest, sums in sequential float order, stop at
 * reldist <= 1e-6 or max_iters (<0: no cap).  Synchronises the stream.
 * bin_boundaries / bin_values are host arrays of bin_count-1 / bin_count.
 * ------------------------------------------------------------------ */
int64_t wv_lloyd_ws_bytes(int64_t n, int bin_count);
int wv_lloyd(const float *coef_d, int64_t n, const wv_thresh_result *res_host,
             int bin_count, int max_iters, void *ws_d, float *bin_boundaries,
             float *bin_values, int *iters_out, void *stream);

/* ------------------------------------------------------------------ *
 * Host Huffman coder + .cube writer (writeQuantData,
 * test_compress_common.cc:198-237,897-1022; huffman.h/.cc; bit_stream.h;
 * cubelet_file.cc:18-84,153-327).  counts are the device histogram; the
 * zero-bin run statistics are recomputed from bins like huffScanDupsOnly.
 * ------------------------------------------------------------------ */
typedef struct {
  int orig[3], steps[3], out_size[3];
  int wavelet, quant_alg, bin_count, orig_dtype;
  float thresh_frac, threshold, max_val;
  const float *boundaries, *values;   /* Lloyd only */
  int zero_bin;                       /* -1: dup encoding off (-noz) */
  int with_footer;                    /* test_compress_gpu writes it */
} wv_cube_desc;

/* Encodes bins (host array, n values) and writes a complete .cube file
   image into file_buf (cap bytes).  Returns bytes written, or a negative
   status; *payload_bytes (optional) receives the Huffman payload size. */
int64_t wv_write_cube(const wv_cube_desc *d, const int32_t *bins_host,
                      int64_t n, const uint32_t *counts_host,
                      uint8_t *file_buf, int64_t cap, int64_t *payload_bytes);

/* ------------------------------------------------------------------ *
 * Huffman coding on the device.  Replaces the host half of writeQuantData
 * after the int32 bins D2H (test_compress_gpu.cu:402): initHuffman +
 * huffScanDupsOnly (test_compress_common.cc:927-973), computeHuffmanCoding
 * (huffman.cc:42-97) and encodeToStream / dupEncodeString
 * (huffman.h:241-296) into a BitStreamWriter (bit_stream.h:87-162).
 * The zero-bin run statistics are computed on the device; the code book
 * (a tree over bin_count + 1 leaves) is built on the host from them; the
 * payload words are written on the device, bit for bit what the host
 * coder writes.  Only the payload needs to cross PCIe.
 * ------------------------------------------------------------------ */
typedef struct wv_huff wv_huff;   /* code book + decode table of one payload */

/* device workspace of wv_encode_device */
int64_t wv_encode_ws_bytes(int64_t n, int bin_count);
/* payload words that always suffice for n bins (the exact count is known
   inside wv_encode_device) */
int64_t wv_encode_words_bound(int64_t n, int bin_count);
/* bins_d: n bins; counts_d: their histogram (bin_count + 1 entries, as
   wv_quantize_hist writes it); dup_value: the zero bin, or -1 for no
   duplicate-string code (-noz).  Writes the payload (LSB-first uint32
   words) into words_d and its length into *nwords_out; when words_cap is
   too small nothing is written, *nwords_out holds the size needed and
   WV_ERR_NOMEM returns.  Synchronises the stream twice (run statistics,
   code book); the payload kernels are left running on it.  *book_out
   (optional; free with wv_huff_free) is needed to write the container. */
int wv_encode_device(const int32_t *bins_d, int64_t n, const uint32_t *counts_d,
                     int bin_count, int dup_value, void *ws_d, uint32_t *words_d,
                     int64_t words_cap, int64_t *nwords_out, wv_huff **book_out,
                     void *stream);
int64_t wv_huff_payload_bits(const wv_huff *book);
/* One leaf of a code book: value 0..bin_count-1, or bin_count (the dup
   key, when the book has one).  *count: the value's count in the tree (the
   zero bin's singles left over from dup strings; the dup key's: the dup
   strings); *bit_length: its code length (0 when it does not occur);
   bits: the code in stream order, bit i of bits[i / 32] written i-th
   (first = the root's branch; at most 64 bits).  What
   Huffman::printEncoding prints (huffman.cc:100-107, the -enc option). */
int wv_huff_code(const wv_huff *book, int value, int64_t *count, int *bit_length,
                 uint32_t bits[2]);
/* the .cube image around a device-coded payload (host copy of its words):
   the same bytes wv_write_cube writes for the same bins */
int64_t wv_write_cube_payload(const wv_cube_desc *d, const wv_huff *book,
                              const uint32_t *words_host, int64_t nwords,
                              uint8_t *file_buf, int64_t cap);
/* the same, with the payload still on the device (words_d, as
   wv_encode_device left it): it is copied straight into file_buf.
   Synchronises the stream. */
int64_t wv_write_cube_device(const wv_cube_desc *d, const wv_huff *book,
                             const uint32_t *words_d, int64_t nwords,
                             uint8_t *file_buf, int64_t cap, void *stream);
void wv_huff_free(wv_huff *book);

/* ------------------------------------------------------------------ *
 * Decompression.  Replaces decompressFile (test_compress_cpu.cc:187-236;
 * the reference GPU decompress is a stub, test_compress_gpu.cu:438-501):
 * readQuantData (test_compress_common.cc:850-894) -> dequantize (:567-584)
 * -> inverse waveletTransform -> translateCubeDataToOriginal (:699-747)
 * -> an uncompressed cubelet file (cubelet_file.cc:259-327).
 * ------------------------------------------------------------------ */
typedef struct {
  int size[3];             /* stored (rotated, padded) width, height, depth */
  int orig[3];             /* original width, height, depth */
  int orig_dtype;          /* WV_DTYPE_UINT8 / INT32 / FLOAT32; 0: not recorded */
  int steps[3];
  int wavelet;             /* WV_WAVELET_* */
  int quant_alg;           /* WV_QUANT_*; -1: unknown */
  int bin_count;
  float thresh_frac, threshold, max_val;
  int dup_key, dup_value;  /* -1 when absent */
  int maximum_value;       /* the cubelet's maximum_value, 0 if unset */
  int n_decode, n_boundaries, n_values;   /* table / codebook lengths */
  int64_t payload_offset, payload_bytes;  /* Huffman payload in the file */
} wv_cube_info;

/* Finds the first wavelet-compressed INT32 cubelet of a .cube file image
   (readQuantData's rule) and fills *info; decode_table / boundaries /
   values (may be NULL) receive the arrays when their capacities suffice. */
int wv_cube_parse(const uint8_t *file, int64_t len, wv_cube_info *info,
                  int32_t *decode_table, int decode_cap, float *boundaries,
                  float *values, int codebook_cap);
/* HuffmanDecoder::decodeFromStream (huffman.h:96-131) on the host: decodes
   up to count values; returns the number decoded (count on success) or a
   negative status for a malformed table. */
int64_t wv_huff_decode_host(const int32_t *table, int n_table, int dup_key,
                            int dup_value, const uint32_t *words, int64_t nwords,
                            int32_t *out, int64_t count);
/* The same decoding on the device (words_d: the payload words).  The
   stream is cut into 512-bit pieces decoded in parallel; pieces whose
   assumed start is not a code boundary decode again from where their
   predecessor ended, until every start agrees (*rounds_out: how many
   rounds that took).  out_d receives count values.  Synchronises.
   Returns WV_OK, or an error when the stream holds fewer values. */
int64_t wv_huff_decode_ws_bytes(int64_t nwords, int n_table);
int wv_huff_decode_device(const int32_t *table, int n_table, int dup_key,
                          int dup_value, const uint32_t *words_d, int64_t nwords,
                          int32_t *out_d, int64_t count, void *ws_d, int *rounds_out,
                          void *stream);
/* the device half: dequantize the bins (stored layout), invert the
   transform and restore the original type over the original size into
   out_d (x fastest; FLOAT32 when the original type was not recorded).
   boundaries / values: the Lloyd codebook (host), else NULL. */
int64_t wv_decompress_work_bytes(const wv_cube_info *info);
int wv_decompress_device(const wv_cube_info *info, const int32_t *bins_d,
                         const float *boundaries, const float *values,
                         void *work_d, int64_t work_bytes, void *out_d, void *stream);
/* one call on a file image: parse, payload H2D, device decode, device
   half, D2H into out_host.  Returns the output bytes; *ms_decode
   (optional) receives the time of the Huffman decode.  Synchronises. */
int64_t wv_decompress(const uint8_t *file, int64_t len, void *work_d,
                      int64_t work_bytes, uint8_t *out_host, int64_t out_cap,
                      wv_cube_info *info_out, double *ms_decode, void *stream);
/* an uncompressed cubelet file (addCubelet + close with the footer) */
int64_t wv_write_cube_raw(const int size[3], int dtype, const void *data,
                          uint8_t *file_buf, int64_t cap);

/* ------------------------------------------------------------------ *
 * One-call drop-in: the whole of compressFile (test_compress_gpu.cu:
 * 150-436) on a device-resident input.  Synchronises.  bins_host (may be
 * NULL) receives the bins; file_buf the .cube image.
 * ------------------------------------------------------------------ */
typedef struct {
  int wavelet, quant_alg, bin_count;
  int steps[3];
  float thresh_frac;
  int compress_zeros;     /* 0 = -noz */
  int lloyd_max_iters;    /* <0: no cap */
  int with_footer;
} wv_options;

typedef struct {
  wv_thresh_result thresh;
  int out_size[3], steps[3];
  int zero_bin, lloyd_iters;
  int64_t payload_bytes, file_bytes;
  double ms_device;       /* device pipeline time (events) */
  double ms_total;        /* including D2H and host Huffman */
} wv_compress_info;

int wv_compress(const void *in_d, int dtype, const int orig[3],
                const wv_options *opt, void *work_d, int64_t work_bytes,
                int32_t *bins_host, uint8_t *file_buf, int64_t file_cap,
                wv_compress_info *info, void *stream);
/* device bytes wv_compress needs in work_d */
int64_t wv_compress_work_bytes(const int orig[3], const wv_options *opt);
/* The device-resident part of wv_compress alone (compressFile's GPU
   stages, test_compress_gpu.cu:180-402, without the bins D2H and the host
   coder): on return the quantize + histogram may still run on stream; bins
   (int32, padded count) are at offset 0 of work_d and the bin_count + 1
   counts at wv_compress_counts_offset.  Synchronises once, for the
   threshold result (and, with Lloyd, per iteration).  codebook_out
   (optional host, 2*bin_count - 1 floats) receives the Lloyd boundaries
   then values. */
int wv_compress_device(const void *in_d, int dtype, const int orig[3],
                       const wv_options *opt, void *work_d, int64_t work_bytes,
                       float *codebook_out, wv_compress_info *info, void *stream);
int64_t wv_compress_counts_offset(const int orig[3], const wv_options *opt);

/* ------------------------------------------------------------------ *
 * Batched compression of many cubelets of uneven size.  No counterpart in
 * the reference, which compresses one cubelet per run of
 * test_compress_gpu (compressFile, test_compress_gpu.cu:150-436); every
 * item gets the results that path gives it alone.  Items are grouped by
 * shape / dtype / wavelet; each transform pass of a group is one launch of
 * the persistent pass kernel over all (cubelet, tile) items; select and
 * quantize run a few launches for the whole batch; both sequences are
 * replayed as hipGraphs.  Log quantizer tables are built on the host
 * (host logf); a batch with log jobs runs in two parts so that the host
 * builds one part's tables while the device transforms or quantizes the
 * other (one event wait per part).
 * ------------------------------------------------------------------ */
typedef struct {
  const void *in_d;        /* device input: orig[0]*orig[1]*orig[2] samples */
  int dtype;               /* WV_DTYPE_* */
  int orig[3];
  int wavelet;             /* WV_WAVELET_* */
  int steps[3];            /* as wv_geom_init */
  int quant_alg;           /* WV_QUANT_UNIFORM or WV_QUANT_LOG */
  int bin_count;
  float thresh_frac;       /* -thresh: fraction of |coef| set to zero */
  float *coef_d;           /* out: padded count coefficients (z,x,y layout) */
  int32_t *bins_d;         /* out: padded count bins; also the intermediate */
  uint32_t *counts_d;      /* out: bin_count + 1 counts */
  /* samples between rows (y) and slices (z) of in_d; {0, 0}: dense
     (orig[0], orig[0] * orig[1]).  A box inside a larger volume is read in
     place by the first transform pass (no extraction copy). */
  int64_t in_stride[2];
} wv_batch_item;

#define WV_BATCH_AUTOTUNE 1   /* time tile / occupancy candidates per group */
#define WV_BATCH_NO_GRAPH 2   /* launch eagerly instead of replaying graphs */

typedef struct wv_batch wv_batch;
/* validates and groups the items, allocates the batch's device tables,
   autotunes when asked (runs kernels; synchronises) */
int wv_batch_plan(const wv_batch_item *items, int n, int flags, wv_batch **out);
/* enqueues the batch on stream; returns after the quantize stage is
   enqueued (with log quantizers, after waiting for each part's threshold
   records, from which the host builds that part's tables) */
int wv_batch_submit(wv_batch *b, void *stream);
int wv_batch_poll(wv_batch *b);          /* 1 done, 0 running, <0 error */
int wv_batch_wait(wv_batch *b);
/* threshold result and zero bin of item i after completion */
int wv_batch_result(const wv_batch *b, int i, wv_thresh_result *res, int *zero_bin);
int wv_batch_size(const wv_batch *b);
int wv_batch_groups(const wv_batch *b);
int wv_batch_group_tuning(const wv_batch *b, int g, int *tile_cap, int *ept0, int *bpc,
                          int *ncub);
int wv_batch_destroy(wv_batch *b);
/* measurement: one eager run of the batch serialised on stream (no side
   streams, no graphs) with events between stages; stage_ms[0..3] = the
   transform passes (sum of each group's span), the selects, the quantize
   passes, and the number of transform pass launches.  Same outputs as a
   submit; synchronises.  No counterpart in the reference. */
int wv_batch_stage_times(wv_batch *b, void *stream, double *stage_ms);
/* after a submit: every item Huffman-coded on the device (items grouped by
   bin count, each group in one launch per phase), written as one .cube
   stream of compressed cubelets in item order, at offsets[3 * i] (NULL:
   0, 0, 0).  Waits for the batch; allocates its coding buffers;
   synchronises.  Returns the bytes (WV_ERR_NOMEM with *need set when cap is
   short).  No counterpart in the reference (one file per run there). */
int64_t wv_batch_write_stream(wv_batch *b, const int *offsets, int compress_zeros,
                              int with_footer, uint8_t *file_buf, int64_t cap, int64_t *need,
                              void *stream);

/* ------------------------------------------------------------------ *
 * Serving loop.  No counterpart in the reference (one file per run of
 * test_compress_gpu).  Requests of mixed shapes and dtypes are submitted
 * as they arrive; device workers (one stream each) admit the head of the
 * FIFO when a worker is free and the device bytes of admitted requests
 * plus its footprint fit hbm_budget_bytes, run the device-resident stages
 * (wv_compress_device), return bins and counts, and release the device
 * memory; encoder threads write the .cube images.  Every result equals the
 * single-cubelet path's for that request.
 * ------------------------------------------------------------------ */
typedef struct {
  int slots;                 /* device workers (concurrent requests) */
  int64_t hbm_budget_bytes;  /* device bytes admitted requests may hold */
  int encode;                /* 1: host Huffman + .cube image per request */
  int host_threads;          /* encoder threads (0: = slots) */
  int lloyd_batch;           /* 1: the Lloyd windows of the requests in flight share
                                launches: runs of a power-of-two size class join one
                                thread and stream at window boundaries, and each
                                window is one set of launches for all of them.  Pays
                                for streams of similar sizes (a window lasts as long
                                as its longest run); 0 (default): each worker runs
                                its own windows */
} wv_serve_config;

typedef struct {
  const void *in_host;       /* input samples, valid until the result is polled */
  int dtype;
  int orig[3];
  wv_options opt;
  int32_t *bins_out;         /* optional host: padded count bins */
  uint32_t *counts_out;      /* optional host: bin_count + 1 counts */
  uint8_t *file_out;         /* optional host .cube image (encode) */
  int64_t file_cap;
} wv_serve_request;

typedef struct {
  int64_t id;
  int status;                /* WV_OK or a WV_ERR_* */
  double t_submit, t_admit, t_done;   /* seconds since the server started */
  int64_t in_bytes, device_bytes, file_bytes;
  wv_thresh_result thresh;
  int zero_bin, lloyd_iters;
} wv_serve_result;

typedef struct wv_server wv_server;
int wv_serve_create(const wv_serve_config *cfg, wv_server **out);
int wv_serve_submit(wv_server *s, const wv_serve_request *r, int64_t *id);
/* moves up to max completed results into out; returns how many */
int wv_serve_poll(wv_server *s, wv_serve_result *out, int max);
int wv_serve_drain(wv_server *s);   /* until every submitted request completed */
int wv_serve_stats(wv_server *s, int64_t *peak_bytes, int *peak_inflight, double *now_s);
/* with lloyd_batch: the window launch sets so far and the Lloyd runs they
 * served (runs / windows = runs per launch set); zeros without it */
int wv_serve_lloyd_stats(wv_server *s, int64_t *windows, int64_t *run_windows);
int wv_serve_destroy(wv_server *s);  /* drains, then stops the threads */

/* ------------------------------------------------------------------ *
 * Cube split with overlap (cube split, cube_split_typed
 * cubelet_convert.cc:368-553): the cubelets of one large volume of
 * size[0] x size[1] x size[2] at cubelet size cub with `overlap` samples
 * shared wherever neighbours meet.  x and y: ceil-style counts
 * (size - overlap - 1) / (cub - overlap) + 1 at offsets k * (cub - overlap),
 * trimmed at the volume edge; z: the reference's frame loop -- a chunk of
 * cub[2] frames, then back up `overlap` frames, the frames left at the end
 * one shorter chunk.  Boxes come in the reference's write order (z chunk,
 * then y, then x).  Returns the number of boxes (those past cap are not
 * written; boxes may be NULL to count), or a negative WV_ERR_*:
 * 0 <= overlap < cub[a] on every axis.
 * ------------------------------------------------------------------ */
typedef struct {
  int offset[3];   /* x, y, z of the box's first sample in the volume */
  int size[3];
} wv_split_box;
int wv_split_plan(const int size[3], const int cub[3], int overlap, wv_split_box *boxes,
                  int cap);

/* ------------------------------------------------------------------ *
 * Many cubelets in one .cube stream (CubeletStreamWriter::addCubelet per
 * cubelet, close with the index footer: cubelet_file.cc:259-327).  The
 * stream is kept in host memory until finished.
 * ------------------------------------------------------------------ */
typedef struct wv_cube_stream wv_cube_stream;
wv_cube_stream *wv_cube_stream_open(void);
/* a compressed cubelet: d and book as for wv_write_cube_payload, the
   payload's nwords words on the host, the box offset in the volume */
int wv_cube_stream_add(wv_cube_stream *w, const wv_cube_desc *d, const int offset[3],
                       const wv_huff *book, const uint32_t *words, int64_t nwords);
/* an uncompressed cubelet of size[0]*size[1]*size[2] samples (host data,
   x fastest; WV_DTYPE_UINT8 / INT32 / FLOAT32) */
int wv_cube_stream_add_raw(wv_cube_stream *w, const int size[3], int dtype,
                           const void *data, const int offset[3]);
/* bytes of the finished stream, and the stream written to file_buf */
int64_t wv_cube_stream_bytes(const wv_cube_stream *w, int with_footer);
int64_t wv_cube_stream_finish(wv_cube_stream *w, int with_footer, uint8_t *file_buf,
                              int64_t cap);
void wv_cube_stream_free(wv_cube_stream *w);

/* ------------------------------------------------------------------ *
 * Split + batched compression of one device-resident volume: the boxes of
 * wv_split_plan are compressed by the batched path straight from the
 * volume (no extraction copies), each coded on the device, and written as
 * one .cube stream of compressed cubelets with their offsets.  The
 * reference splits with `cube split` and then compresses one cubelet per
 * run of test_compress_gpu.  Uniform or log quantizer (opt->quant_alg).
 * Allocates its device buffers; synchronises.  Returns the file bytes
 * (file_buf NULL or too small: WV_ERR_NOMEM with *need set).
 * ------------------------------------------------------------------ */
typedef struct {
  int cubelets;
  int64_t need;            /* bytes of the whole .cube stream */
  double ms_plan, ms_batch, ms_encode, ms_total;
} wv_split_info;
int64_t wv_split_compress(const void *vol_d, int dtype, const int size[3], const int cub[3],
                          int overlap, const wv_options *opt, int batch_flags,
                          uint8_t *file_buf, int64_t cap, wv_split_info *info,
                          void *stream);

/* ------------------------------------------------------------------ *
 * Parameter optimisation (-opt): OptimizationData + testParameters
 * (optimize.h:6-145, optimize_gpu.cu:15-202), the interface the
 * reference's error-control sweeps (optimize.cc:4-133) use.  create runs the
 * forward transform once, takes min / max / max |x| and sorts |x| on the
 * device; in_d (the original samples) must stay valid until destroy.  A
 * test point quantizes with threshold T and bin_count bins (uniform or
 * log, max = max |x|; Lloyd: the code book trained on the sorted |x| from
 * findSorted(T), all but the largest, with no iteration cap, as
 * computeLloydQuantization does (test_compress_cpu.cc:404-459, the
 * reference's fixedPSNR sweep, optimize.cc:53-90); bin_count <= 0: only
 * |x| <= T -> 0), inverts, and compares with the original over its extent
 * (restore rules of computeErrorRatesAfterDequant).  size_bytes is
 * Huffman::totalEncodedLengthBytes of that quantization (huffman.cc:
 * 241-252), 0 without quantization.  Error sums are double, deterministic.
 * ------------------------------------------------------------------ */
typedef struct wv_opt wv_opt;
typedef struct {
  int64_t size_bytes;
  float l1, l2, mse, psnr, rel;   /* getL1Error, getL2Error, getMeanSquaredError,
                                     getPeakSignalToNoiseRatio, getRelativeError */
  int zero_bin;                   /* quant(0); -1 without quantization */
  int lloyd_iters;                /* Lloyd iterations of the test point's code book */
} wv_opt_result;
int wv_opt_create(const void *in_d, int dtype, const int orig[3], const int steps[3],
                  int wavelet, int max_possible, int compress_zeros, wv_opt **out,
                  void *stream);
int64_t wv_opt_count(const wv_opt *o);
int wv_opt_stats(const wv_opt *o, float *min_val, float *max_val, float *max_abs);
float wv_opt_sorted(const wv_opt *o, int64_t index);        /* getSorted */
int64_t wv_opt_find_sorted(const wv_opt *o, float value);  /* findSorted */
int wv_opt_test(wv_opt *o, float threshold, int bin_count, int quant_alg, wv_opt_result *r,
                void *stream);
int wv_opt_destroy(wv_opt *o);

/* measurement: device-to-device streaming copy with 128-bit accesses
   (vec_per_thread = 4, 8 or 16 loads in flight per thread; optional
   non-temporal hints), blocks_per_cu blocks per CU; bytes and both
   pointers 16-byte aligned.  The measured HBM ceiling beside the published
   peak.  No counterpart in the reference. */
int wv_copy_stream(const void *src_d, void *dst_d, int64_t bytes, int nontemporal,
                   int blocks_per_cu, int vec_per_thread, void *stream);

#ifdef __cplusplus
}
#endif
#endif
