// wv_split.cpp -- cube split with overlap (cube split, cube_split_typed
// cubelet_convert.cc:368-553) feeding the batched compression path.
//
// The reference cuts a large volume into cubelet files on the host and
// then compresses each cubelet with its own run of test_compress_gpu.  Here
// the volume stays in HBM: every box becomes a batch item whose first
// transform pass reads the box in place through the volume's row and slice
// strides (no extraction copy), the whole set runs as one batch (grouped
// persistent pass launches, batched select and quantize), each cubelet is
// Huffman-coded on the device, and the payloads are written as one .cube
// stream with the boxes' offsets (wv_batch_write_stream).  All boxes are
// coded together: their run statistics in one launch per phase, the code
// books on host threads, all payloads in one launch per phase (a box coded
// alone is a chain of small launches and two synchronisations, which left
// the device idle), and the payloads go from the device straight into the
// file image.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <map>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>
#include "wv_internal.h"

using namespace wv;

// cube_split_typed: x / y counts and offsets (cubelet_convert.cc:451-483),
// z from the frame loop (:490-553): a chunk is flushed when it holds cub[2]
// frames, then the loop backs up `overlap` frames unless it is at the last
// frame; what is left at the end is one shorter chunk.  Write order: z
// chunk, then the y-major, x-minor list of the x / y boxes.
extern "C" int wv_split_plan(const int size[3], const int cub[3], int overlap,
                             wv_split_box *boxes, int cap) {
  if (!size || !cub || overlap < 0) {
    set_error("wv_split_plan: bad argument");
    return WV_ERR_ARG;
  }
  for (int a = 0; a < 3; a++)
    if (size[a] <= 0 || cub[a] <= 0 || overlap >= cub[a]) {
      set_error("wv_split_plan: volume %dx%dx%d, cubelets %dx%dx%d, overlap %d (needs "
                "0 <= overlap < cubelet size)",
                size[0], size[1], size[2], cub[0], cub[1], cub[2], overlap);
      return WV_ERR_ARG;
    }
  // boxes per row / column: (size - overlap - 1) / (cub - overlap) + 1
  // with C's truncating division
  int count[2];
  for (int a = 0; a < 2; a++) count[a] = (size[a] - overlap - 1) / (cub[a] - overlap) + 1;
  std::vector<int> zoff, zlen;
  int start = 0, depth = 0;
  for (int f = 0; f < size[2]; f++) {
    depth++;
    if (depth == cub[2]) {
      zoff.push_back(start);
      zlen.push_back(depth);
      start = f + 1 - overlap;
      depth = 0;
      if (f < size[2] - 1) f -= overlap;
    }
  }
  if (depth > 0) {
    zoff.push_back(start);
    zlen.push_back(depth);
  }
  const int64_t total = (int64_t)zoff.size() * count[0] * count[1];
  if (total > 0x7fffffff) {
    set_error("wv_split_plan: %lld boxes", (long long)total);
    return WV_ERR_ARG;
  }
  int k = 0;
  for (size_t zc = 0; zc < zoff.size(); zc++)
    for (int y = 0; y < count[1]; y++)
      for (int x = 0; x < count[0]; x++, k++) {
        if (!boxes || k >= cap) continue;
        wv_split_box &b = boxes[k];
        b.offset[0] = x * (cub[0] - overlap);
        b.offset[1] = y * (cub[1] - overlap);
        b.offset[2] = zoff[zc];
        b.size[0] = std::min(cub[0], size[0] - b.offset[0]);
        b.size[1] = std::min(cub[1], size[1] - b.offset[1]);
        b.size[2] = zlen[zc];
      }
  return k;
}

namespace {

int esize(int dtype) {
  return dtype == WV_DTYPE_UINT8 ? 1 : dtype == WV_DTYPE_INT16 ? 2 : 4;
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct BatchHandle {
  wv_batch *b = nullptr;
  ~BatchHandle() {
    if (b) (void)wv_batch_destroy(b);
  }
};

#define SCHK(call)                  \
  do {                              \
    int _rc = (call);               \
    if (_rc != WV_OK) return _rc;   \
  } while (0)
#define SHIP(expr, what)                                         \
  do {                                                           \
    hipError_t _e = (expr);                                      \
    if (_e != hipSuccess) {                                      \
      set_error("%s: %s", what, hipGetErrorString(_e));          \
      return WV_ERR_HIP;                                         \
    }                                                            \
  } while (0)

double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

extern "C" int64_t wv_split_compress(const void *vol_d, int dtype, const int size[3],
                                     const int cub[3], int overlap, const wv_options *opt,
                                     int batch_flags, uint8_t *file_buf, int64_t cap,
                                     wv_split_info *info, void *stream) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!vol_d || !opt || !size || !cub) {
    set_error("wv_split_compress: null argument");
    return WV_ERR_ARG;
  }
  if (opt->quant_alg != WV_QUANT_UNIFORM && opt->quant_alg != WV_QUANT_LOG) {
    set_error("wv_split_compress: quantizer %d (uniform or log)", opt->quant_alg);
    return WV_ERR_UNSUPPORTED;
  }
  if (dtype != WV_DTYPE_UINT8 && dtype != WV_DTYPE_INT16 && dtype != WV_DTYPE_INT32 &&
      dtype != WV_DTYPE_FLOAT32) {
    set_error("wv_split_compress: dtype %d", dtype);
    return WV_ERR_ARG;
  }
  const int nbox = wv_split_plan(size, cub, overlap, nullptr, 0);
  if (nbox < 0) return nbox;
  std::vector<wv_split_box> boxes((size_t)nbox);
  wv_split_plan(size, cub, overlap, boxes.data(), nbox);
  // per box: geometry, then coefficient / bins / counts regions of one
  // device arena
  std::vector<wv_geom> geoms((size_t)nbox);
  std::vector<int64_t> off_coef(nbox), off_bins(nbox), off_counts(nbox);
  auto al = [](int64_t v) { return (v + 255) & ~(int64_t)255; };
  int64_t total = 0, max_count = 0;
  for (int i = 0; i < nbox; i++) {
    SCHK(wv_geom_init(&geoms[i], boxes[i].size, opt->steps, opt->wavelet, dtype));
    const int64_t n = geoms[i].count;
    max_count = std::max(max_count, n);
    off_coef[i] = total; total += al(n * 4);
    off_bins[i] = total; total += al(n * 4);
    off_counts[i] = total; total += al(((int64_t)opt->bin_count + 1) * 4);
  }
  DevBuf arena;
  SHIP(hipMalloc(&arena.p, (size_t)total), "wv_split_compress: device buffers");
  char *ar = (char *)arena.p;
  const int es = esize(dtype);
  std::vector<wv_batch_item> items((size_t)nbox);
  for (int i = 0; i < nbox; i++) {
    const wv_split_box &bx = boxes[i];
    wv_batch_item &it = items[i];
    it = wv_batch_item{};
    const int64_t first = ((int64_t)bx.offset[2] * size[1] + bx.offset[1]) * size[0] + bx.offset[0];
    it.in_d = (const char *)vol_d + first * es;
    it.dtype = dtype;
    std::memcpy(it.orig, bx.size, sizeof it.orig);
    it.wavelet = opt->wavelet;
    std::memcpy(it.steps, opt->steps, sizeof it.steps);
    it.quant_alg = opt->quant_alg;
    it.bin_count = opt->bin_count;
    it.thresh_frac = opt->thresh_frac;
    it.coef_d = (float *)(ar + off_coef[i]);
    it.bins_d = (int32_t *)(ar + off_bins[i]);
    it.counts_d = (uint32_t *)(ar + off_counts[i]);
    it.in_stride[0] = size[0];
    it.in_stride[1] = (int64_t)size[0] * size[1];
  }
  const double ms_plan = ms_since(t0);
  const auto t1 = std::chrono::steady_clock::now();
  BatchHandle bh;
  // one submit: launched directly (a graph would first need an eager run)
  SCHK(wv_batch_plan(items.data(), nbox, batch_flags | WV_BATCH_NO_GRAPH, &bh.b));
  SCHK(wv_batch_submit(bh.b, stream));
  SCHK(wv_batch_wait(bh.b));
  const double ms_batch = ms_since(t1);
  // every box coded on the device together and written with its offset,
  // the payloads copied from the device straight into file_buf
  const auto t2 = std::chrono::steady_clock::now();
  std::vector<int> offs((size_t)nbox * 3);
  for (int i = 0; i < nbox; i++) std::memcpy(&offs[(size_t)i * 3], boxes[i].offset, 12);
  int64_t need = 0;
  const int64_t nb = wv_batch_write_stream(bh.b, offs.data(), opt->compress_zeros,
                                           opt->with_footer, file_buf, cap, &need, stream);
  if (info) {
    info->cubelets = nbox;
    info->need = need;
    info->ms_plan = ms_plan;
    info->ms_batch = ms_batch;
    info->ms_encode = ms_since(t2);
    info->ms_total = ms_since(t0);
  }
  return nb;
}

// A finished batch as one .cube stream: items grouped by (bin count, dup
// value), each group coded together (encode_batch), records in item order
// with the given box offsets (NULL: all 0).
extern "C" int64_t wv_batch_write_stream(wv_batch *b, const int *offsets, int compress_zeros,
                                         int with_footer, uint8_t *file_buf, int64_t cap,
                                         int64_t *need_out, void *stream) {
  if (!b) {
    set_error("wv_batch_write_stream: null batch");
    return WV_ERR_ARG;
  }
  SCHK(wv_batch_wait(b));
  const int n = wv_batch_size(b);
  const wv_batch_item *items = batch_items(b);
  std::vector<wv_geom> geoms((size_t)n);
  std::vector<wv_thresh_result> res((size_t)n);
  std::vector<int> zbs((size_t)n);
  std::map<std::pair<int, int>, std::vector<int>> groups;
  for (int i = 0; i < n; i++) {
    SCHK(wv_geom_init(&geoms[i], items[i].orig, items[i].steps, items[i].wavelet, items[i].dtype));
    SCHK(wv_batch_result(b, i, &res[i], &zbs[i]));
    groups[{items[i].bin_count, compress_zeros ? zbs[i] : -1}].push_back(i);
  }
  // per group: its payloads (left on the device); per item: group, offset
  // and length
  std::vector<DevBuf> gdev(groups.size());
  std::vector<int> gof((size_t)n);
  std::vector<int64_t> iwoff((size_t)n), inw((size_t)n);
  std::vector<wv_huff *> books((size_t)n, nullptr);
  auto free_books = [&] {
    for (auto *bk : books) wv_huff_free(bk);
  };
  // one coding workspace, sized for the largest group (encode_batch
  // synchronises its stream before it returns, so the groups can share it)
  int64_t wsb = 0;
  for (auto &kv : groups) {
    std::vector<int64_t> cnt;
    for (int i : kv.second) cnt.push_back(geoms[i].count);
    wsb = std::max(wsb, encode_batch_ws_bytes(cnt.data(), (int)cnt.size(), kv.first.first,
                                              kv.first.second));
  }
  DevBuf ws;
  if (wsb > 0 && hipMalloc(&ws.p, (size_t)wsb) != hipSuccess) {
    set_error("wv_batch_write_stream: coding workspace of %lld bytes", (long long)wsb);
    return WV_ERR_NOMEM;
  }
  int gi = 0;
  for (auto &kv : groups) {
    const std::vector<int> &ix = kv.second;
    const int nj = (int)ix.size(), bins = kv.first.first, dup = kv.first.second;
    std::vector<const int32_t *> bp(nj);
    std::vector<const uint32_t *> cp(nj);
    std::vector<int64_t> cnt(nj);
    for (int k = 0; k < nj; k++) {
      bp[k] = items[ix[k]].bins_d;
      cp[k] = items[ix[k]].counts_d;
      cnt[k] = geoms[ix[k]].count;
    }
    std::vector<uint32_t> none;
    std::vector<int64_t> gwoff, gnw;
    std::vector<wv_huff *> gb;
    uint32_t *dev = nullptr;
    const int rc = encode_batch(bp.data(), cnt.data(), cp.data(), nj, bins, dup, ws.p, wsb,
                                nullptr, 0, none, gwoff, gnw, gb, (hipStream_t)stream, &dev);
    if (rc != WV_OK) {
      free_books();
      return rc;
    }
    gdev[gi].p = dev;
    for (int k = 0; k < nj; k++) {
      gof[ix[k]] = gi;
      iwoff[ix[k]] = gwoff[k];
      inw[ix[k]] = gnw[k];
      books[ix[k]] = gb[k];
    }
    gi++;
  }
  for (int i = 0; i < n; i++)
    if (inw[i] > 0x3fffffff) {   // the record's byte_count is a uint32
      free_books();
      set_error("wv_batch_write_stream: item %d: payload of %lld words exceeds a record's 4 GiB",
                i, (long long)inw[i]);
      return WV_ERR_ARG;
    }
  std::vector<wv_cube_desc> descs((size_t)n);
  std::vector<std::string> params((size_t)n);
  {
    const int nt = std::max(1, std::min(16, n / 16));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&, t] {
        for (int i = t; i < n; i += nt) {
          wv_options opt{};
          opt.wavelet = items[i].wavelet;
          opt.quant_alg = items[i].quant_alg;
          opt.bin_count = items[i].bin_count;
          std::memcpy(opt.steps, items[i].steps, sizeof opt.steps);
          opt.thresh_frac = items[i].thresh_frac;
          opt.compress_zeros = compress_zeros;
          opt.with_footer = with_footer;
          cube_desc_of(geoms[i], &opt, items[i].dtype, res[i], {}, {}, zbs[i], descs[i]);
          params[i] = compressed_params_msg(&descs[i], books[i]);
          wv_huff_free(books[i]);
          books[i] = nullptr;
        }
      });
    for (auto &th : pool) th.join();
  }
  // the layout of the stream: header, then per record its length, message
  // and payload, then the index footer
  static const int zero3[3] = {0, 0, 0};
  std::vector<std::string> metas((size_t)n), indexed((size_t)n);
  std::vector<int64_t> data_at((size_t)n);
  int64_t pos = 64;
  for (int i = 0; i < n; i++) {
    compressed_cubelet_msgs(&descs[i], params[i], (uint32_t)(inw[i] * 4), (uint64_t)pos + 4,
                            offsets ? offsets + 3 * i : zero3, metas[i], indexed[i]);
    data_at[i] = pos + 4 + (int64_t)metas[i].size();
    pos = data_at[i] + inw[i] * 4;
  }
  const std::string tail = with_footer ? cube_index_tail(indexed) : std::string();
  const int64_t need = pos + (int64_t)tail.size();
  if (need_out) *need_out = need;
  if (!file_buf || cap < need) {
    set_error("wv_batch_write_stream: file buffer %lld < %lld", (long long)cap, (long long)need);
    return WV_ERR_NOMEM;
  }
  // records averaging 128 KB or more: each payload goes from the device
  // straight to its place in the file image; smaller ones (where a copy's
  // fixed cost dominates): one copy of each group's payloads, then host
  // copies.  The choice is made for the whole stream from the average: the
  // items of a group are interleaved, so staging only the small records of
  // a group would copy the large ones between them twice.
  hipStream_t s = (hipStream_t)stream;
  std::memset(file_buf, 0, 64);
  std::memcpy(file_buf, "SCU cubelets 1.0\n", 17);
  int64_t payload_words = 0;
  for (int i = 0; i < n; i++) payload_words += inw[i];
  const bool direct_all = payload_words * 4 >= (int64_t)n * (128 << 10);
  auto is_direct = [&](int) { return direct_all; };
  const size_t ng = gdev.size();
  std::vector<int64_t> glo(ng, INT64_MAX), ghi(ng, 0);
  for (int i = 0; i < n; i++)
    if (inw[i] > 0 && !is_direct(i)) {
      glo[gof[i]] = std::min(glo[gof[i]], iwoff[i]);
      ghi[gof[i]] = std::max(ghi[gof[i]], iwoff[i] + inw[i]);
    }
  std::vector<std::vector<uint32_t>> staged(ng);
  for (size_t g = 0; g < ng; g++) {
    if (ghi[g] <= 0) continue;
    staged[g].resize((size_t)(ghi[g] - glo[g]));
    if (hipMemcpyAsync(staged[g].data(), (const uint32_t *)gdev[g].p + glo[g],
                       (size_t)(ghi[g] - glo[g]) * 4, hipMemcpyDeviceToHost, s) != hipSuccess) {
      set_error("wv_batch_write_stream: payload D2H");
      return WV_ERR_HIP;
    }
  }
  for (int i = 0; i < n; i++) {
    if (inw[i] <= 0 || !is_direct(i)) continue;
    if (hipMemcpyAsync(file_buf + data_at[i], (const uint32_t *)gdev[gof[i]].p + iwoff[i],
                       (size_t)inw[i] * 4, hipMemcpyDeviceToHost, s) != hipSuccess) {
      set_error("wv_batch_write_stream: payload D2H");
      return WV_ERR_HIP;
    }
  }
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error("wv_batch_write_stream: sync");
    return WV_ERR_HIP;
  }
  for (int i = 0; i < n; i++) {
    const uint32_t mlen = (uint32_t)metas[i].size();
    uint8_t *r = file_buf + data_at[i] - 4 - mlen;
    std::memcpy(r, &mlen, 4);
    std::memcpy(r + 4, metas[i].data(), mlen);
    if (inw[i] > 0 && !is_direct(i))
      std::memcpy(file_buf + data_at[i], staged[gof[i]].data() + (iwoff[i] - glo[gof[i]]),
                  (size_t)inw[i] * 4);
  }
  if (!tail.empty()) std::memcpy(file_buf + pos, tail.data(), tail.size());
  return need;
}
