// wv_serve.cpp -- a serving loop: compress requests of mixed shapes and
// dtypes as they arrive, on one GPU, with continuous batching.
//
// The reference compresses one file per process run (test_compress_gpu,
// compressFile test_compress_gpu.cu:150-436); this has no counterpart there.
// Every request gets exactly the bins, counts and .cube bytes the
// single-cubelet path gives it.
//
//   submit     returns at once; the request joins a FIFO with its device
//              footprint (input copy + wv_compress_work_bytes).
//   admission  a device worker takes the head of the FIFO when the bytes
//              held by admitted requests plus the head's footprint fit the
//              HBM budget.  Requests that arrive while others run are thus
//              admitted as soon as a worker and memory free up, and their
//              kernels run concurrently with the others' on the worker
//              streams.
//   device     H2D of the input, then wv_compress_device (transform,
//              threshold, [Lloyd], quantize + histogram; the cubelet stays
//              in HBM throughout) and the device Huffman coder; only the
//              payload (and bins / counts when asked for) comes back.
//              With lloyd_batch, the Lloyd windows of all requests in
//              flight share launches: a run joins a batcher thread at a
//              window boundary and its worker waits for it there.
//              The device memory returns to the pool right after.
//   encode     host threads write the .cube container around the payload,
//              so the device worker moves on to the next request at once.
//
// Device memory comes from a caching pool (power-of-two size classes) so
// steady-state requests do not call hipMalloc / hipFree (hipFree would
// synchronise the device under the other workers' streams).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>
#include "wv_device.h"
#include "wv_internal.h"

using namespace wv;

namespace {

typedef std::chrono::steady_clock Clock;

struct Pool {
  std::mutex mu;
  std::multimap<size_t, void *> free_blocks;
  std::vector<void *> all;
  static size_t class_of(size_t b) {
    size_t c = 1 << 20;
    while (c < b) c <<= 1;
    return c;
  }
  void *get(size_t bytes) {
    const size_t c = class_of(bytes);
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = free_blocks.find(c);
      if (it != free_blocks.end()) {
        void *p = it->second;
        free_blocks.erase(it);
        return p;
      }
    }
    void *p = nullptr;
    if (hipMalloc(&p, c) != hipSuccess) {
      // drop the cached blocks and try once more
      std::vector<void *> drop;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (auto &kv : free_blocks) drop.push_back(kv.second);
        free_blocks.clear();
      }
      for (void *q : drop) {
        (void)hipFree(q);
        forget(q);
      }
      if (hipMalloc(&p, c) != hipSuccess) return nullptr;
    }
    std::lock_guard<std::mutex> lk(mu);
    all.push_back(p);
    return p;
  }
  void put(void *p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(mu);
    free_blocks.emplace(class_of(bytes), p);
  }
  void forget(void *p) {
    std::lock_guard<std::mutex> lk(mu);
    all.erase(std::remove(all.begin(), all.end(), p), all.end());
  }
  ~Pool() {
    for (void *p : all) (void)hipFree(p);
  }
};

struct Req {
  int64_t id;
  wv_serve_request r;
  int64_t in_bytes = 0, work_bytes = 0, footprint = 0;
  double t_submit = 0, t_admit = 0;
  // device results kept for the container writer
  std::vector<uint32_t> words;   // Huffman payload
  wv_huff *book = nullptr;
  std::vector<float> codebook;   // Lloyd: boundaries then values
  wv_compress_info info{};
  int status = WV_OK;
};

int64_t dtype_bytes(int dtype) {
  switch (dtype) {
    case WV_DTYPE_UINT8: return 1;
    case WV_DTYPE_INT16: return 2;
    default: return 4;
  }
}

}  // namespace

struct wv_server {
  wv_serve_config cfg;
  Clock::time_point t0;
  Pool pool;
  std::mutex mu;
  std::condition_variable cv_work;    // pending requests / stop
  std::condition_variable cv_done;    // completions
  std::condition_variable cv_enc;     // encode queue
  std::deque<Req *> pending;
  std::deque<Req *> to_encode;
  std::vector<wv_serve_result> done;
  int64_t next_id = 0, submitted = 0, completed = 0;
  int64_t bytes_admitted = 0;
  int64_t peak_bytes = 0;
  int inflight = 0, peak_inflight = 0;
  bool stop = false;
  std::vector<std::thread> workers, encoders;
  int device = 0;
  LloydBatcher *lloyd = nullptr;   // with cfg.lloyd_batch

  double now() const {
    return std::chrono::duration<double>(Clock::now() - t0).count();
  }

  void finish(Req *q) {
    wv_serve_result res{};
    res.id = q->id;
    res.status = q->status;
    res.t_submit = q->t_submit;
    res.t_admit = q->t_admit;
    res.t_done = now();
    res.in_bytes = q->in_bytes;
    res.device_bytes = q->footprint;
    res.file_bytes = q->info.file_bytes;
    res.thresh = q->info.thresh;
    res.zero_bin = q->info.zero_bin;
    res.lloyd_iters = q->info.lloyd_iters;
    delete q;
    std::lock_guard<std::mutex> lk(mu);
    done.push_back(res);
    completed++;
    cv_done.notify_all();
  }

  void device_worker() {
    (void)hipSetDevice(device);
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return;
    for (;;) {
      Req *q = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        // admission: the head of the FIFO, when its footprint fits the budget
        cv_work.wait(lk, [&] {
          if (stop) return true;
          if (pending.empty()) return false;
          const Req *h = pending.front();
          return bytes_admitted + h->footprint <= cfg.hbm_budget_bytes || inflight == 0;
        });
        if (stop && pending.empty()) break;
        q = pending.front();
        pending.pop_front();
        bytes_admitted += q->footprint;
        inflight++;
        peak_bytes = std::max(peak_bytes, bytes_admitted);
        peak_inflight = std::max(peak_inflight, inflight);
      }
      q->t_admit = now();
      run_device(q, s);
      {
        std::lock_guard<std::mutex> lk(mu);
        bytes_admitted -= q->footprint;
        inflight--;
        cv_work.notify_all();
      }
      if (q->status == WV_OK && cfg.encode && q->r.file_out) {
        std::lock_guard<std::mutex> lk(mu);
        to_encode.push_back(q);
        cv_enc.notify_one();
      } else {
        finish(q);
      }
    }
    (void)hipStreamDestroy(s);
  }

  void run_device(Req *q, hipStream_t s) {
    const wv_serve_request &r = q->r;
    void *din = pool.get((size_t)q->in_bytes);
    void *work = pool.get((size_t)q->work_bytes);
    if (!din || !work) {
      q->status = WV_ERR_NOMEM;
      pool.put(din, (size_t)q->in_bytes);
      pool.put(work, (size_t)q->work_bytes);
      return;
    }
    int rc = hipMemcpyAsync(din, r.in_host, (size_t)q->in_bytes, hipMemcpyHostToDevice, s) ==
                     hipSuccess ? WV_OK : WV_ERR_HIP;
    if (r.opt.quant_alg == WV_QUANT_LLOYD) q->codebook.resize(2 * r.opt.bin_count - 1);
    if (rc == WV_OK)
      rc = compress_device(din, r.dtype, r.orig, &r.opt, work, q->work_bytes,
                           q->codebook.empty() ? nullptr : q->codebook.data(), &q->info, s,
                           lloyd);
    if (rc == WV_OK) {
      wv_geom g;
      rc = wv_geom_init(&g, r.orig, r.opt.steps, r.opt.wavelet, r.dtype);
      const int nb = r.opt.bin_count + 1;
      const int64_t coff = wv_compress_counts_offset(r.orig, &r.opt);
      if (rc == WV_OK && r.counts_out &&
          hipMemcpyAsync(r.counts_out, (char *)work + coff, nb * 4, hipMemcpyDeviceToHost, s) !=
              hipSuccess)
        rc = WV_ERR_HIP;
      if (rc == WV_OK && r.bins_out &&
          hipMemcpyAsync(r.bins_out, work, g.count * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = WV_ERR_HIP;
      // encode_work synchronises the stream after the payload copy
      if (rc == WV_OK && cfg.encode && r.file_out)
        rc = encode_work(g, &r.opt, q->info.zero_bin, work, q->words, &q->book, s);
      else if (rc == WV_OK && hipStreamSynchronize(s) != hipSuccess)
        rc = WV_ERR_HIP;
    }
    if (rc != WV_OK && q->book) {
      wv_huff_free(q->book);
      q->book = nullptr;
    }
    q->status = rc;
    pool.put(din, (size_t)q->in_bytes);
    pool.put(work, (size_t)q->work_bytes);
  }

  void encoder() {
    for (;;) {
      Req *q = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_enc.wait(lk, [&] { return stop || !to_encode.empty(); });
        if (to_encode.empty()) break;
        q = to_encode.front();
        to_encode.pop_front();
      }
      const wv_serve_request &r = q->r;
      wv_geom g;
      int rc = wv_geom_init(&g, r.orig, r.opt.steps, r.opt.wavelet, r.dtype);
      if (rc == WV_OK) {
        wv_cube_desc d{};
        std::memcpy(d.orig, g.orig, sizeof d.orig);
        std::memcpy(d.steps, g.steps, sizeof d.steps);
        std::memcpy(d.out_size, g.out_size, sizeof d.out_size);
        d.wavelet = r.opt.wavelet;
        d.quant_alg = r.opt.quant_alg;
        d.bin_count = r.opt.bin_count;
        d.orig_dtype = r.dtype == WV_DTYPE_INT16 ? WV_DTYPE_INT32 : r.dtype;
        d.thresh_frac = r.opt.thresh_frac;
        d.threshold = q->info.thresh.threshold;
        d.max_val = q->info.thresh.max_abs;
        d.zero_bin = r.opt.compress_zeros ? q->info.zero_bin : -1;
        d.with_footer = r.opt.with_footer;
        if (r.opt.quant_alg == WV_QUANT_LLOYD) {
          d.boundaries = q->codebook.data();
          d.values = q->codebook.data() + (r.opt.bin_count - 1);
        }
        const int64_t nb = wv_write_cube_payload(&d, q->book, q->words.data(),
                                                 (int64_t)q->words.size(), r.file_out,
                                                 r.file_cap);
        if (nb < 0) rc = (int)nb;
        else q->info.file_bytes = nb;
      }
      wv_huff_free(q->book);
      q->book = nullptr;
      std::vector<uint32_t>().swap(q->words);
      q->status = rc;
      finish(q);
    }
  }
};

extern "C" {

int wv_serve_create(const wv_serve_config *cfg, wv_server **out) {
  if (!cfg || !out || cfg->slots < 1 || cfg->slots > 64 || cfg->hbm_budget_bytes <= 0) {
    set_error("bad serve config");
    return WV_ERR_ARG;
  }
  wv_server *s = new wv_server;
  s->cfg = *cfg;
  s->t0 = Clock::now();
  (void)hipGetDevice(&s->device);
  if (cfg->lloyd_batch) {
    s->lloyd = lloyd_batcher_create(s->device);
    if (!s->lloyd) {
      delete s;
      return WV_ERR_HIP;
    }
  }
  for (int i = 0; i < cfg->slots; i++) s->workers.emplace_back(&wv_server::device_worker, s);
  const int ne = cfg->encode ? (cfg->host_threads > 0 ? cfg->host_threads : cfg->slots) : 0;
  for (int i = 0; i < ne; i++) s->encoders.emplace_back(&wv_server::encoder, s);
  *out = s;
  return WV_OK;
}

int wv_serve_submit(wv_server *s, const wv_serve_request *r, int64_t *id) {
  if (!s || !r || !r->in_host) { set_error("bad arguments"); return WV_ERR_ARG; }
  wv_geom g;
  int rc = wv_geom_init(&g, r->orig, r->opt.steps, r->opt.wavelet, r->dtype);
  if (rc != WV_OK) return rc;
  if (r->opt.bin_count < 2 || r->opt.bin_count > WV_MAX_BINS) {
    set_error("bin count");
    return WV_ERR_ARG;
  }
  Req *q = new Req;
  q->r = *r;
  q->in_bytes = (int64_t)r->orig[0] * r->orig[1] * r->orig[2] * dtype_bytes(r->dtype);
  q->work_bytes = wv_compress_work_bytes(r->orig, &r->opt);
  if (q->work_bytes < 0) { delete q; set_error("work size"); return WV_ERR_ARG; }
  q->footprint = (int64_t)Pool::class_of((size_t)q->in_bytes) +
                 (int64_t)Pool::class_of((size_t)q->work_bytes);
  if (q->footprint > s->cfg.hbm_budget_bytes) {
    delete q;
    set_error("request needs more device memory than the budget");
    return WV_ERR_NOMEM;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  q->id = s->next_id++;
  q->t_submit = s->now();
  s->pending.push_back(q);
  s->submitted++;
  if (id) *id = q->id;
  s->cv_work.notify_all();
  return WV_OK;
}

int wv_serve_poll(wv_server *s, wv_serve_result *out, int max) {
  if (!s || (!out && max > 0)) { set_error("bad arguments"); return WV_ERR_ARG; }
  std::lock_guard<std::mutex> lk(s->mu);
  const int n = std::min<int>(max, (int)s->done.size());
  for (int i = 0; i < n; i++) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + n);
  return n;
}

int wv_serve_drain(wv_server *s) {
  if (!s) { set_error("bad arguments"); return WV_ERR_ARG; }
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->completed == s->submitted; });
  return WV_OK;
}

int wv_serve_stats(wv_server *s, int64_t *peak_bytes, int *peak_inflight, double *now_s) {
  if (!s) { set_error("bad arguments"); return WV_ERR_ARG; }
  std::lock_guard<std::mutex> lk(s->mu);
  if (peak_bytes) *peak_bytes = s->peak_bytes;
  if (peak_inflight) *peak_inflight = s->peak_inflight;
  if (now_s) *now_s = s->now();
  return WV_OK;
}

int wv_serve_destroy(wv_server *s) {
  if (!s) return WV_OK;
  (void)wv_serve_drain(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
    s->cv_enc.notify_all();
  }
  for (auto &t : s->workers) t.join();
  for (auto &t : s->encoders) t.join();
  if (s->lloyd) lloyd_batcher_destroy(s->lloyd);
  delete s;
  return WV_OK;
}

int wv_serve_lloyd_stats(wv_server *s, int64_t *windows, int64_t *run_windows) {
  if (!s || !windows || !run_windows) { set_error("bad arguments"); return WV_ERR_ARG; }
  *windows = *run_windows = 0;
  if (s->lloyd) lloyd_batcher_stats(s->lloyd, windows, run_windows);
  return WV_OK;
}

}  // extern "C"
