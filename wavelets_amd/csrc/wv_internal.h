// wv_internal.h -- declarations shared between the gfx950 kernel files and
// the C ABI implementation.  Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include "../../include/wv_hip.h"

#define WV_MAX_LINE 16384   // longest axis one LDS tile holds
#define WV_MAXT 256         // most lines in one LDS tile

namespace wv {

// Where a pass reads or writes element (u, v, p): u*su + v*sv + p*sp, each
// coordinate first clamped to its cap (REPEAT padding of the unpadded
// source; lines past the volume edge).  p runs along the pass's axis, u
// over neighbouring lines of a tile, v over tile rows.
struct LineMap {
  int64_t su, sv, sp;
  int ucap, vcap, pcap;
};

// One fused axis pass over the whole (padded) volume.
struct DwtPass {
  const void *src;
  float *dst;
  struct WvStats *stats;
  int X, Y, Z;        // padded volume dims (x, y, z)
  int ox, oy, oz;     // dims of the clamped source (the unpadded input)
  int levels;
  int L;              // line length (the pass axis)
  int Nu, Nv;         // lines: Nu neighbouring lines (u) x Nv rows (v)
  int ntu;            // tiles per row: ceil(Nu / T)
  LineMap ld, st;     // load / store addressing
  int ld_line, st_line;   // 1: lines are contiguous (sp == 1); 0: lines are
                          // neighbours (su == 1), positions strided
  int T, log2T;       // lines per tile (a power of two <= WV_MAXT)
  int log2dblock_T;   // log2(threads per block / T)
  unsigned magic_L;   // ceil(2^32 / L) for exact division of i < 2^16
  int ept;            // tile elements per thread (16 or 32)
  int nt;             // threads per block (WV_DBLOCK)
  int exact;          // L*T fills a tile exactly, no padding, 32-bit lane offsets
  // batches: ncub cubelets of this same geometry, cubelet c reading
  // src_tab[c] and writing dst_tab[c] (src / dst unused then); ntiles1 tiles
  // per cubelet
  const void *const *src_tab;
  float *const *dst_tab;
  int ncub, ntiles1;
  // launch overrides (0: defaults): tile width cap, elements per thread of
  // line-contiguous passes, resident blocks per CU
  int tcap, ept0, bpc;
  // first forward pass of a box inside a larger volume: source samples
  // between rows and slices (0: the dense unpadded source)
  int64_t src_sy, src_sz;
  // every dst_tab pointer is 16-byte aligned (batches; P.dst is checked
  // directly); vec_st: exact tiles store their segments as 16-byte vectors
  int dst_tab_aligned, vec_st;
  // tiles of the runtime-T kernels (any L): rt_exact: a full tile holds
  // exactly nt*ept elements and needs no clamping; st_block: such a tile's
  // lines are adjacent in the output (line stride L), so it stores one
  // contiguous block of T*L floats as 16-byte vectors
  int rt_exact, st_block;
  // the plane pass (x and y levels of T planes per tile, x,y,z in place):
  // plane set, levels2 = y levels (levels = x levels), magic_XY = exact
  // reciprocal of X*Y (magic_L then holds X's)
  int plane, levels2;
  unsigned magic_XY;
};

// Memory layouts of a cubelet, fastest axis first: x,y,z (the input and
// intermediates), y,x,z, and the reference's final z,x,y.
enum { WV_LAYOUT_XYZ = 0, WV_LAYOUT_YXZ = 1, WV_LAYOUT_ZXY = 2 };

// One axis pass: levels along axis, reading layout lin, writing lout.
// axis WV_PASS_PLANE: x levels (levels) and y levels (levels2) of whole
// xy-planes in one pass, x,y,z in place.
enum { WV_PASS_PLANE = 3 };
struct PassSpec {
  int axis, levels, lin, lout;
  int levels2;
};

// The forward pass sequence; returns the number of passes.  The reference
// visits x, rotates, y, rotates, z (cdf97_3d dwt_cpu.cc:943-965); here only
// the data layouts change between passes.  With x and y levels on planes
// of at most 32 x 32 (WV_DWT_PLANE_MAX), one plane pass runs both in x,y,z
// and the z pass writes z,x,y.  Otherwise, with x and y levels and depth
// >= 32 the x pass writes y,x,z and the y pass z,x,y, so every pass reads
// whole lines; else x and y stay in x,y,z and the z pass writes z,x,y (with
// depth 1 the two orders coincide and no z pass runs unless it has levels).
// No levels at all: one conversion pass in x,y,z.
int forward_specs(const int padded[3], const int steps[3], PassSpec s[3]);
// inverse: z (from z,x,y into x,y,z), then y, then x
int inverse_specs(const int padded[3], const int steps[3], PassSpec s[3]);
// Which buffer each pass writes: 1 the output, 0 the workspace (the last
// pass writes the output; a pass whose input and output layouts are equal
// runs in place).
void pass_buffers(const PassSpec *s, int n, int dst_out[3]);
// Fills the geometry of P (X/Y/Z, ox/oy/oz and the batch fields set by the
// caller) for spec s; clamp: the pass reads the unpadded source.  Returns
// the tiles of one launch (all cubelets) or -1.
int plan_pass(DwtPass &P, const PassSpec &s, bool clamp);

hipError_t launch_pass(const DwtPass &P, int ntiles, int wavelet, bool inverse,
                       int src_dtype, bool clamp, bool stats, hipStream_t s);

// statistics / threshold select
hipError_t stats_clear(void *stats, hipStream_t s);
hipError_t stats_compute(const float *coef, int64_t n, void *stats,
                         hipStream_t s);
// LSD radix sort of *count_d (at most max_n) uint32 keys, result in keys
// (wv_lloyd.hip); tmp of max_n keys, hist of radix_sort_hist_bytes(max_n)
int64_t radix_sort_hist_bytes(int64_t max_n);
hipError_t radix_sort_u32(uint32_t *keys, uint32_t *tmp, uint32_t *hist,
                          const uint32_t *count_d, int64_t max_n, hipStream_t s);

// Lloyd runs (wv_lloyd.hip).  lloyd_run is wv_lloyd with an optional
// batcher: with one, the run's windows share launches with the other runs
// in flight (the serving loop), on the batcher's own thread and stream.
struct LloydBatcher;
LloydBatcher *lloyd_batcher_create(int device);
void lloyd_batcher_destroy(LloydBatcher *b);
// window launch sets so far, and the runs they served (their ratio: runs per launch set)
void lloyd_batcher_stats(LloydBatcher *b, int64_t *windows, int64_t *run_windows);
int lloyd_run(const float *coef_d, int64_t n, const wv_thresh_result *res, int bin_count,
              int max_iters, void *ws_d, float *bin_boundaries, float *bin_values,
              int *iters_out, hipStream_t s, LloydBatcher *batcher);
// Lloyd over points already sorted on the device (the optimiser's test
// points): nonzero[0 .. count-1] ascending, ws of wv_lloyd_ws_bytes(n_ws, bins)
int lloyd_run_sorted(const float *nonzero, int64_t count, float threshold, int bin_count,
                     int max_iters, void *ws_d, int64_t n_ws, float min_val, float max_val,
                     float *bin_boundaries, float *bin_values, int *iters_out, hipStream_t s);
// wv_compress_device with an optional Lloyd batcher (wv_compress.cpp)
int compress_device(const void *in_d, int dtype, const int orig[3], const wv_options *opt,
                    void *work_d, int64_t work_bytes, float *codebook_out,
                    wv_compress_info *info, hipStream_t stream, LloydBatcher *lb);
hipError_t threshold_select(const float *coef, int64_t n, float frac,
                            const void *stats, void *ws,
                            wv_thresh_result *res, hipStream_t s);
int64_t threshold_ws_bytes(int64_t n);

// batched threshold select: one job per cubelet; the collect launch gives
// job j blocks [blk0, blk0 + nblk) (nblk <= sel_max_collect_blocks())
struct BSelJob {
  const float *coef;
  int64_t n;
  float frac;
  void *state;            // sel_state_bytes() of device memory
  uint32_t *cand;         // candidate keys, capacity cap
  uint32_t cap;
  wv_thresh_result *res;  // device
  int blk0, nblk;
};
int64_t sel_state_bytes();
int sel_max_collect_blocks();
int64_t sel_collect_chunk();
hipError_t threshold_select_batch(const BSelJob *jobs_d, const int *blk0_d, int njobs,
                                  int total_blocks, hipStream_t s);

// quantize + histogram
struct QuantDev {
  int alg, bin_count, zero_bin;
  float threshold;
  float neg_scale, neg_offset, pos_scale, pos_offset;
  int n_pos, n_neg;   // log edges / codebook boundaries
  float pos_clamp;    // log: |x| from which the top bin is clamped to B-1
  const float *table; // device: log: pos edges then neg edges;
                      //         codebook: boundaries
};
hipError_t quantize_hist(const float *coef, int64_t n, const QuantDev &q,
                         int32_t *bins, uint32_t *counts, hipStream_t s);
hipError_t quantize_hist_uniform_dev(const float *coef, int64_t n,
                                     int bin_count,
                                     const wv_thresh_result *res,
                                     int32_t *bins, uint32_t *counts,
                                     hipStream_t s);
// batched quantize + histogram: job j gets blocks [blk0, blk0 + nblk) of
// one launch; q points at the job's parameter block in device memory
// (filled by the host for log / codebook, by bquant_prepare for uniform)
struct BQuantJob {
  const float *coef;
  int64_t n;
  int32_t *bins;
  uint32_t *counts;       // bin_count + 1 entries
  QuantDev *q;
  const wv_thresh_result *res;
  int alg, bin_count;
  int blk0, nblk;
};
int quant_blocks_for(int64_t n);
QuantDev quant_dev_of(const wv_quant_params *p);   // table = p->table_d
hipError_t quantize_hist_batch(const BQuantJob *jobs_d, const int *blk0_d, int njobs,
                               int total_blocks, int max_bins, hipStream_t s);
hipError_t dequantize_table(const int32_t *bins, int64_t n, const float *table,
                            int bin_count, float *out, hipStream_t s);

// device Huffman coding (wv_encode.hip)
int64_t enc_ws_bytes(int64_t n, int bin_count);
uint64_t *enc_book_ptr(void *ws, int64_t n, int bin_count);
// run statistics of the duplicate value zb (-1: none): sd_host[0] single
// codes, sd_host[1] dup codes, plus per-range tables in ws that
// enc_write_payload uses; synchronises the stream
hipError_t enc_count_runs(const int32_t *bins, int64_t n, int zb, int bin_count, void *ws,
                          unsigned long long *sd_host, hipStream_t s);
// payload words from the code book at enc_book_ptr (nb1 packed codes)
// short_runs: the codes of every run shorter than 1024 fit 120 bits
hipError_t enc_write_payload(const int32_t *bins, int64_t n, int zb, int bin_count, int nb1,
                             bool short_runs, void *ws, uint32_t *words, int64_t nwords,
                             hipStream_t s);

// batches of jobs with one bin count (every job with a dup key or every
// job without): plan (job tables and workspace layout; ws may be null to
// size it), histograms + run statistics of all jobs (one launch per phase,
// one synchronisation), then all payloads from the packed host books
// (nb1 codes per job) into words + woff[j]
// one coding job: its bins, ranges and workspace tables
struct EncArgs {
  const int32_t *bins;
  int64_t n;
  int64_t nchunks;
  int zb;                    // duplicate value; -1: no duplicate coding
  int nb1;                   // code-book entries (the dup key is the last)
  int vec;                   // bins are 16-byte aligned
  int nranges;               // blocks; block b owns chunks [b per, (b+1) per)
  int64_t per;
  int64_t *range_last;       // per range: its last non-duplicate index, -1
  uint64_t *range_bits;      // per range: output bits
  unsigned long long *range_sd;   // per range: single codes, dup codes
  uint32_t *range_hist;      // per range: histogram of the coded values
  const uint64_t *book;      // codes: bits | length << 56
  uint32_t *words;
  uint64_t nbits_cap;        // bits the payload buffer holds
  unsigned long long *sd;    // [0] single codes of zb, [1] dup codes
  int blk0;                  // first block of this job in its launch (batches)
};

struct EncBatch {
  int nj = 0, bin_count = 0, nb1 = 0, blocks = 0;
  bool hist = false, ok = false;
  int64_t o_jobs = 0, o_blk0 = 0, o_sd = 0, o_ptrs = 0, o_counts = 0, o_books = 0, total = 0;
  std::vector<EncArgs> jobs;
  std::vector<int> blk0;
};
EncBatch enc_batch_plan(const int32_t *const *bins, const int64_t *n, const int *zb, int nj,
                        int bin_count, void *ws);
hipError_t enc_batch_counts(EncBatch &P, const uint32_t *const *counts_d, void *ws,
                            uint32_t *counts_host, unsigned long long *sd_host, hipStream_t s);
hipError_t enc_batch_payload(EncBatch &P, void *ws, const uint64_t *books_host, bool short_runs,
                             uint32_t *words, const int64_t *woff, const int64_t *nwords,
                             int64_t total_words, hipStream_t s);

// the items a batch was planned with (wv_batch.cpp)
const wv_batch_item *batch_items(const wv_batch *b);

// the whole batch coding (wv_huffman.cpp): books per job and all payload
// words on the host (job j's nwords[j] words at all + woff[j]); ws of
// encode_batch_ws_bytes; words_d of the summed payloads, or null to have
// one of the exact size allocated and freed inside
int64_t encode_batch_ws_bytes(const int64_t *n, int nj, int bin_count, int dup);
int encode_batch(const int32_t *const *bins_d, const int64_t *n, const uint32_t *const *counts_d,
                 int nj, int bin_count, int dup, void *ws_d, int64_t ws_bytes, uint32_t *words_d,
                 int64_t words_cap, std::vector<uint32_t> &all, std::vector<int64_t> &woff,
                 std::vector<int64_t> &nwords, std::vector<wv_huff *> &books, hipStream_t s,
                 uint32_t **words_dev_out = nullptr);
// (words_dev_out: the payloads stay on the device, all stays empty, and
// *words_dev_out is their buffer -- allocated here when words_d is null,
// then the caller's to free)
// the index footer of a stream whose records' indexed messages are given
std::string cube_index_tail(const std::vector<std::string> &index);
// grows a stream's buffer ahead of many records
void cube_stream_reserve(wv_cube_stream *w, size_t bytes);

// after wv_compress_device on work_d: device Huffman coding of its bins,
// payload copied to words_host (synchronises); and the container fields
int encode_work(const wv_geom &g, const wv_options *opt, int zero_bin, void *work_d,
                std::vector<uint32_t> &words_host, wv_huff **book, hipStream_t s);
void cube_desc_of(const wv_geom &g, const wv_options *opt, int dtype, const wv_thresh_result &res,
                  const std::vector<float> &bounds, const std::vector<float> &values,
                  int zero_bin, wv_cube_desc &d);

void set_error(const char *fmt, ...);

// Huffman::totalEncodedLengthBytes of a histogram (counts: bin_count
// entries; dup < 0: no dup strings)
int64_t huff_total_encoded_bytes(const uint32_t *counts, int bin_count, int dup,
                                 int64_t singles, int64_t dups);

// the CubeletBuffer messages of a compressed cubelet at a box offset: the
// record's (no data_file_offset) and the index's (data at
// data_offset_after_meta + the record message's length)
// (the parameters message, with the decode table, built first and
// separately: it does not depend on the record's place in the stream)
std::string compressed_params_msg(const wv_cube_desc *d, const wv_huff *book);
void compressed_cubelet_msgs(const wv_cube_desc *d, const std::string &params, uint32_t nbytes,
                             uint64_t data_offset_after_meta, const int offset[3],
                             std::string &meta, std::string &indexed);
// appends a compressed record whose parameters message is built
int cube_stream_add_params(wv_cube_stream *w, const wv_cube_desc *d, const std::string &params,
                           const int offset[3], const uint32_t *words, int64_t nwords);

// compute units of the current device (cached per device)
int cu_count();

}  // namespace wv
