// wv_batch.cpp -- many cubelets of uneven size compressed as one batch.
//
// The reference compresses one cubelet per run of test_compress_gpu
// (compressFile, test_compress_gpu.cu:150-436).  A batch runs the same
// stages for all of its cubelets in a handful of launches:
//
//   transform  cubelets are grouped by (padded shape, steps, dtype,
//              wavelet); every group runs each of its axis passes as ONE
//              launch of the persistent pass kernel over the (cubelet,
//              tile) items of the whole group (cubelet pointer tables), with
//              the group's compile-time tile variant.  Groups are spread
//              over several streams so small groups overlap.
//   select     three launches for all cubelets: per-cubelet sample, one
//              flattened collect (each cubelet owns a block range), and a
//              per-cubelet finish (window check, LDS radix select, result).
//   quantize   two launches: per-cubelet counts / uniform parameters, then
//              one flattened fused quantize + histogram.
//
// The transform + select sequence and the quantize sequence are captured as
// hipGraphs on the first submit and replayed afterwards (the kernels read
// their work from device tables, so replays need no re-capture).  Between
// them the host computes the log quantizer's bin-edge tables with its own
// logf, exactly as for a single cubelet, on a pool of threads.  To hide that
// host stage, a batch with log jobs is cut into two parts of about equal
// work (whole groups): the device runs transform + select of part 0, then of
// part 1, then the quantize of part 0 and of part 1, while the host builds
// part 0's tables during part 1's transform and part 1's during part 0's
// quantize.
//
// An autotuner (WV_BATCH_AUTOTUNE) times tile-width caps, elements per
// thread of x-line passes, and resident blocks per CU for every group on a
// sample of its cubelets and keeps the fastest, cached per process by group
// key.
//
// Each cubelet's bins buffer doubles as the transform's intermediate (it
// has the same size as the coefficients), so a batch needs no per-cubelet
// scratch beyond the select state and candidate list.
#include <hip/hip_runtime.h>
#include <array>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>
#include "wv_device.h"
#include "wv_internal.h"

namespace wv {

#define BATCH_STREAMS 4

struct Tuning {
  int tcap = 0, ept0 = 0, bpc = 0;   // 0: defaults
};

struct Group {
  int padded[3], orig[3], steps[3];
  int dtype, wavelet;
  int64_t src_sy, src_sz;      // first pass: source strides (0: dense)
  std::vector<int> items;
  PassSpec spec[3];
  int npass;
  int dst_out[3];              // per pass: 1 writes the coefficients, 0 the bins buffer
  int dst_aligned = 1;         // every output pointer of the group 16-byte aligned
  const void **src_tab_d[3];   // per pass: device pointer tables
  float **dst_tab_d[3];
  Tuning tune;
};

typedef std::tuple<int, int, int, int, int, int, int, int, int, int, int, int64_t, int64_t>
    GroupKey;

static GroupKey key_of(const wv_geom &g, int dtype, int64_t sy, int64_t sz) {
  return GroupKey(g.padded[0], g.padded[1], g.padded[2], g.orig[0], g.orig[1], g.orig[2],
                  g.steps[0], g.steps[1], g.steps[2], dtype, g.wavelet, sy, sz);
}

// A fixed set of host threads for the edge tables (threads are started
// once per batch, not per submit).  run(n, f) calls f(0 .. n-1) on them and
// returns when all calls have returned.
class HostPool {
 public:
  explicit HostPool(unsigned n) {
    for (unsigned t = 0; t < n; t++) th_.emplace_back([this] { loop(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void run(int n, const std::function<void(int)> &f) {
    if (th_.empty()) {
      for (int i = 0; i < n; i++) f(i);
      return;
    }
    std::unique_lock<std::mutex> lk(mu_);
    fn_ = &f;
    next_ = 0;
    total_ = n;
    busy_ = 0;
    gen_++;
    cv_.notify_all();
    done_.wait(lk, [&] { return next_ >= total_ && busy_ == 0; });
    fn_ = nullptr;
  }

 private:
  void loop() {
    unsigned seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || (gen_ != seen && next_ < total_); });
      if (stop_) return;
      seen = gen_;
      while (next_ < total_) {
        const int i = next_++;
        busy_++;
        const std::function<void(int)> *f = fn_;
        lk.unlock();
        (*f)(i);
        lk.lock();
        busy_--;
      }
      if (busy_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)> *fn_ = nullptr;
  int next_ = 0, total_ = 0, busy_ = 0;
  unsigned gen_ = 0;
  bool stop_ = false;
};

static std::mutex g_tune_mu;
static std::map<GroupKey, Tuning> g_tune_cache;

static int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace wv

using namespace wv;

struct wv_batch {
  std::vector<wv_batch_item> items;
  std::vector<wv_geom> geoms;
  std::vector<Group> groups;
  int flags = 0;
  // device memory: one arena
  char *arena = nullptr;
  BSelJob *sel_d = nullptr;
  int *selblk_d = nullptr;
  int sel_total = 0;
  BQuantJob *q_d = nullptr;
  int *qblk_d = nullptr;
  int q_total = 0;
  QuantDev *qdev_d = nullptr;
  float *qtab_d = nullptr;
  int qtab_stride = 0;
  wv_thresh_result *res_d = nullptr;
  // pinned host mirrors
  wv_thresh_result *res_h = nullptr;
  QuantDev *qdev_h = nullptr;
  float *qtab_h = nullptr;
  wv_thresh_result *res_h2 = nullptr;   // after the last transform part: final
  bool host_params = false;    // any log job
  // parts: groups -> part, items -> slots of the job arrays (part 0 first)
  int nparts = 1;
  std::vector<int> group_part, slot, slot_item;
  int part_begin[3] = {0, 0, 0};
  int sel_total_p[2] = {0, 0}, q_total_p[2] = {0, 0};
  HostPool *pool = nullptr;
  int max_bins = 2;
  hipStream_t side[BATCH_STREAMS] = {};
  hipStream_t cap = nullptr;   // graphs are captured here (the caller's
                               // stream may be the null stream)
  hipEvent_t fork = nullptr, join[BATCH_STREAMS] = {};
  hipEvent_t part_done[2] = {};
  hipGraphExec_t gA[2] = {}, gB[2] = {};
  hipStream_t last_stream = nullptr;   // may be the null stream
  bool submitted = false;
};

namespace wv {

// Enqueues one group's transform passes on stream s.
static int enqueue_group(wv_batch *b, Group &G, hipStream_t s, int ncub_limit) {
  const int ncub = ncub_limit > 0 ? std::min<int>(ncub_limit, (int)G.items.size())
                                  : (int)G.items.size();
  for (int i = 0; i < G.npass; i++) {
    DwtPass P{};
    P.X = G.padded[0]; P.Y = G.padded[1]; P.Z = G.padded[2];
    P.ox = G.orig[0]; P.oy = G.orig[1]; P.oz = G.orig[2];
    P.src_tab = G.src_tab_d[i];
    P.dst_tab = G.dst_tab_d[i];
    P.ncub = ncub;
    if (i == 0) {
      P.src_sy = G.src_sy;
      P.src_sz = G.src_sz;
    }
    P.tcap = G.tune.tcap;
    P.ept0 = G.tune.ept0;
    P.bpc = G.tune.bpc;
    P.dst_tab_aligned = G.dst_aligned;
    const int ntiles = plan_pass(P, G.spec[i], i == 0);
    if (ntiles < 0) { set_error("batch: unsupported pass geometry"); return WV_ERR_UNSUPPORTED; }
    const bool clamp = i == 0;
    hipError_t e = launch_pass(P, ntiles, G.wavelet, false,
                               clamp ? G.dtype : WV_DTYPE_FLOAT32, clamp, false, s);
    if (e != hipSuccess) {
      set_error("batch pass: %s", hipGetErrorString(e));
      return WV_ERR_HIP;
    }
  }
  return WV_OK;
}

// transform of every group of a part (spread over the side streams) +
// select of the part's cubelets
static int enqueue_A(wv_batch *b, hipStream_t s, int part) {
  hipError_t e = hipEventRecord(b->fork, s);
  if (e != hipSuccess) { set_error("batch: %s", hipGetErrorString(e)); return WV_ERR_HIP; }
  for (int k = 0; k < BATCH_STREAMS; k++) {
    e = hipStreamWaitEvent(b->side[k], b->fork, 0);
    if (e != hipSuccess) { set_error("batch: %s", hipGetErrorString(e)); return WV_ERR_HIP; }
  }
  // largest groups first, round-robin over the streams
  std::vector<int> order;
  for (size_t i = 0; i < b->groups.size(); i++)
    if (b->group_part[i] == part) order.push_back((int)i);
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    const Group &A = b->groups[x], &B = b->groups[y];
    const int64_t ca = (int64_t)A.padded[0] * A.padded[1] * A.padded[2] * A.items.size();
    const int64_t cb = (int64_t)B.padded[0] * B.padded[1] * B.padded[2] * B.items.size();
    return ca > cb;
  });
  for (size_t i = 0; i < order.size(); i++) {
    const int rc = enqueue_group(b, b->groups[order[i]], b->side[i % BATCH_STREAMS], 0);
    if (rc != WV_OK) return rc;
  }
  for (int k = 0; k < BATCH_STREAMS; k++) {
    e = hipEventRecord(b->join[k], b->side[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, b->join[k], 0);
    if (e != hipSuccess) { set_error("batch: %s", hipGetErrorString(e)); return WV_ERR_HIP; }
  }
  const int pb = b->part_begin[part], pe = b->part_begin[part + 1];
  e = threshold_select_batch(b->sel_d + pb, b->selblk_d + pb, pe - pb, b->sel_total_p[part], s);
  if (e != hipSuccess) { set_error("batch select: %s", hipGetErrorString(e)); return WV_ERR_HIP; }
  return WV_OK;
}

static int enqueue_B(wv_batch *b, hipStream_t s, int part) {
  const int pb = b->part_begin[part], pe = b->part_begin[part + 1];
  hipError_t e = quantize_hist_batch(b->q_d + pb, b->qblk_d + pb, pe - pb, b->q_total_p[part],
                                     b->max_bins, s);
  if (e != hipSuccess) { set_error("batch quantize: %s", hipGetErrorString(e)); return WV_ERR_HIP; }
  return WV_OK;
}

typedef std::function<int(wv_batch *, hipStream_t)> EnqueueFn;

static int capture(wv_batch *b, const EnqueueFn &fn, hipGraphExec_t *out) {
  hipStream_t s = b->cap;
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) { set_error("capture: %s", hipGetErrorString(e)); return WV_ERR_HIP; }
  const int rc = fn(b, s);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(s, &g);
  if (rc != WV_OK) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) { set_error("capture end: %s", hipGetErrorString(e)); return WV_ERR_HIP; }
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) { set_error("graph instantiate: %s", hipGetErrorString(e)); return WV_ERR_HIP; }
  return WV_OK;
}

// Host parameters of a part's log jobs from the select results res (by
// item): edge tables into qtab_h, parameter blocks into qdev_h (by slot).
static int host_params(wv_batch *b, int part, const wv_thresh_result *res) {
  std::vector<int> jobs;
  for (int k = b->part_begin[part]; k < b->part_begin[part + 1]; k++)
    if (b->items[b->slot_item[k]].quant_alg == WV_QUANT_LOG) jobs.push_back(k);
  if (jobs.empty()) return WV_OK;
  std::vector<int> rcs(jobs.size(), WV_OK);
  b->pool->run((int)jobs.size(), [&](int j) {
    const int k = jobs[j];
    const int i = b->slot_item[k];
    const wv_thresh_result &r = res[i];
    wv_quant_params p;
    // QuantLog with maxValue = max |x| (test_compress_cpu.cc:262-265)
    int rc = wv_quant_params_init(&p, WV_QUANT_LOG, b->items[i].bin_count, r.threshold,
                                  r.max_abs, nullptr, nullptr);
    if (rc != WV_OK) { rcs[j] = rc; return; }
    float *tab = b->qtab_h + (size_t)k * b->qtab_stride;
    std::memcpy(tab, p.pos_edges, sizeof(float) * p.n_pos_edges);
    std::memcpy(tab + p.n_pos_edges, p.neg_edges, sizeof(float) * p.n_neg_edges);
    p.table_d = b->qtab_d + (size_t)k * b->qtab_stride;
    b->qdev_h[k] = quant_dev_of(&p);
    wv_quant_params_free(&p);
  });
  for (int rc : rcs)
    if (rc != WV_OK) return rc;
  return WV_OK;
}

// Autotuning of one group: time its passes on a sample of its cubelets for
// each candidate and keep the fastest.
static int autotune_group(wv_batch *b, Group &G, hipStream_t s) {
  const GroupKey key = key_of(b->geoms[G.items[0]], G.dtype, G.src_sy, G.src_sz);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) { G.tune = it->second; return WV_OK; }
  }
  const int64_t per = (int64_t)G.padded[0] * G.padded[1] * G.padded[2];
  // enough cubelets for about 64 M elements (several waves of tiles)
  int sample = (int)std::min<int64_t>((int64_t)G.items.size(),
                                      std::max<int64_t>(1, (64ll << 20) / per));
  std::vector<Tuning> cands;
  for (int tcap : {0, 128, 64, 32, 16})
    for (int ept0 : {16, 32})
      for (int bpc : {0, 1, 2}) {
        Tuning t;
        t.tcap = tcap; t.ept0 = ept0; t.bpc = bpc;
        cands.push_back(t);
      }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    set_error("autotune: event");
    return WV_ERR_HIP;
  }
  float best = 1e30f;
  Tuning bt;
  for (const Tuning &t : cands) {
    G.tune = t;
    int rc = enqueue_group(b, G, s, sample);   // warm (sets attributes, caches)
    if (rc != WV_OK) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < 3 && rc == WV_OK; r++) rc = enqueue_group(b, G, s, sample);
    (void)hipEventRecord(e1, s);
    if (rc != WV_OK || hipStreamSynchronize(s) != hipSuccess) {
      (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
      return rc != WV_OK ? rc : WV_ERR_HIP;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best) { best = ms; bt = t; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  G.tune = bt;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = bt;
  return WV_OK;
}

static void free_batch(wv_batch *b) {
  if (!b) return;
  for (int p = 0; p < 2; p++) {
    if (b->gA[p]) (void)hipGraphExecDestroy(b->gA[p]);
    if (b->gB[p]) (void)hipGraphExecDestroy(b->gB[p]);
    if (b->part_done[p]) (void)hipEventDestroy(b->part_done[p]);
  }
  delete b->pool;
  for (int k = 0; k < BATCH_STREAMS; k++) {
    if (b->side[k]) (void)hipStreamDestroy(b->side[k]);
    if (b->join[k]) (void)hipEventDestroy(b->join[k]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->arena) (void)hipFree(b->arena);
  if (b->res_h) (void)hipHostFree(b->res_h);
  if (b->res_h2) (void)hipHostFree(b->res_h2);
  if (b->qdev_h) (void)hipHostFree(b->qdev_h);
  if (b->qtab_h) (void)hipHostFree(b->qtab_h);
  delete b;
}

}  // namespace wv

#define BCHK(expr, what)                                            \
  do {                                                              \
    hipError_t _e = (expr);                                         \
    if (_e != hipSuccess) {                                         \
      set_error("%s: %s", what, hipGetErrorString(_e));             \
      free_batch(b);                                                \
      return WV_ERR_HIP;                                            \
    }                                                               \
  } while (0)

extern "C" {

int wv_batch_plan(const wv_batch_item *items, int n, int flags, wv_batch **out) {
  if (!items || n <= 0 || !out) { set_error("bad arguments"); return WV_ERR_ARG; }
  *out = nullptr;
  wv_batch *b = new wv_batch;
  b->flags = flags;
  b->items.assign(items, items + n);
  b->geoms.resize(n);
  std::map<GroupKey, int> gidx;
  int max_bins = 2;
  for (int i = 0; i < n; i++) {
    const wv_batch_item &it = items[i];
    if (!it.in_d || !it.coef_d || !it.bins_d || !it.counts_d) {
      set_error("batch item %d: null buffer", i);
      free_batch(b);
      return WV_ERR_ARG;
    }
    if (it.quant_alg != WV_QUANT_UNIFORM && it.quant_alg != WV_QUANT_LOG) {
      set_error("batch item %d: quantizer %d not batched (uniform or log)", i, it.quant_alg);
      free_batch(b);
      return WV_ERR_UNSUPPORTED;
    }
    if (it.bin_count < 2 || it.bin_count > WV_MAX_BINS) {
      set_error("batch item %d: bin count %d", i, it.bin_count);
      free_batch(b);
      return WV_ERR_ARG;
    }
    int rc = wv_geom_init(&b->geoms[i], it.orig, it.steps, it.wavelet, it.dtype);
    if (rc != WV_OK) { free_batch(b); return rc; }
    max_bins = std::max(max_bins, it.bin_count);
    if (it.quant_alg == WV_QUANT_LOG) b->host_params = true;
    // source strides: a box inside a larger volume (dense strides count
    // as dense, so such boxes group with dense cubelets)
    int64_t sy = it.in_stride[0], sz = it.in_stride[1];
    if (sy || sz) {
      if (sy < it.orig[0] || sz < sy * it.orig[1]) {
        set_error("batch item %d: strides %lld, %lld for %dx%dx%d", i, (long long)sy,
                  (long long)sz, it.orig[0], it.orig[1], it.orig[2]);
        free_batch(b);
        return WV_ERR_ARG;
      }
      if (sy == it.orig[0] && sz == (int64_t)it.orig[0] * it.orig[1]) sy = sz = 0;
    }
    const GroupKey k = key_of(b->geoms[i], it.dtype, sy, sz);
    auto f = gidx.find(k);
    if (f == gidx.end()) {
      Group G{};
      const wv_geom &g = b->geoms[i];
      std::memcpy(G.padded, g.padded, sizeof G.padded);
      std::memcpy(G.orig, g.orig, sizeof G.orig);
      std::memcpy(G.steps, g.steps, sizeof G.steps);
      G.dtype = it.dtype;
      G.wavelet = it.wavelet;
      G.src_sy = sy;
      G.src_sz = sz;
      G.npass = forward_specs(g.padded, g.steps, G.spec);
      pass_buffers(G.spec, G.npass, G.dst_out);
      gidx[k] = (int)b->groups.size();
      b->groups.push_back(G);
      f = gidx.find(k);
    }
    b->groups[f->second].items.push_back(i);
  }

  // parts: with log jobs, two parts of about equal work (largest groups
  // first, each to the lighter part); the job arrays list part 0's items
  // first (slots)
  const int ng = (int)b->groups.size();
  b->nparts = b->host_params && ng >= 2 ? 2 : 1;
  b->group_part.assign(ng, 0);
  if (b->nparts == 2) {
    std::vector<int> ord(ng);
    for (int g = 0; g < ng; g++) ord[g] = g;
    auto work = [&](int g) {
      const Group &G = b->groups[g];
      return (int64_t)G.padded[0] * G.padded[1] * G.padded[2] * (int64_t)G.items.size();
    };
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return work(x) > work(y); });
    int64_t load[2] = {0, 0};
    for (int g : ord) {
      const int pt = load[0] <= load[1] ? 0 : 1;
      b->group_part[g] = pt;
      load[pt] += work(g);
    }
  }
  b->slot.assign(n, 0);
  b->slot_item.clear();
  for (int pt = 0; pt < b->nparts; pt++) {
    b->part_begin[pt] = (int)b->slot_item.size();
    for (int g = 0; g < ng; g++)
      if (b->group_part[g] == pt)
        for (int i : b->groups[g].items) {
          b->slot[i] = (int)b->slot_item.size();
          b->slot_item.push_back(i);
        }
  }
  b->part_begin[b->nparts] = n;

  // arena layout
  const size_t A = 256;
  auto up = [&](size_t x) { return (x + A - 1) & ~(A - 1); };
  size_t off = 0;
  const size_t o_sel = off; off = up(off + sizeof(BSelJob) * n);
  const size_t o_selblk = off; off = up(off + sizeof(int) * n);
  const size_t o_q = off; off = up(off + sizeof(BQuantJob) * n);
  const size_t o_qblk = off; off = up(off + sizeof(int) * n);
  const size_t o_qdev = off; off = up(off + sizeof(QuantDev) * n);
  b->qtab_stride = (max_bins + 63) & ~63;
  b->max_bins = max_bins;
  const size_t o_qtab = off; off = up(off + sizeof(float) * (size_t)b->qtab_stride * n);
  const size_t o_res = off; off = up(off + sizeof(wv_thresh_result) * n);
  const size_t o_tabs = off;
  size_t tab_bytes = 0;
  for (auto &G : b->groups) tab_bytes += up(sizeof(void *) * G.items.size()) * 2 * G.npass;
  off = up(off + tab_bytes);
  const size_t state_bytes = (size_t)sel_state_bytes();
  const size_t o_state = off; off = up(off + state_bytes * n);
  std::vector<size_t> cand_off(n);
  std::vector<uint32_t> cand_cap(n);
  for (int i = 0; i < n; i++) {
    const int64_t cnt = b->geoms[i].count;
    // the window holds about 2% of the ranks; 1/8 + 16 K leaves ample room
    // (an overflow falls back to the exact select of that cubelet)
    const int64_t cap = std::min<int64_t>(cnt, cnt / 8 + 16384);
    cand_cap[i] = (uint32_t)cap;
    cand_off[i] = off;
    off = up(off + (size_t)cap * 4);
  }
  BCHK(hipMalloc((void **)&b->arena, off), "batch arena");
  char *ar = b->arena;
  b->sel_d = (BSelJob *)(ar + o_sel);
  b->selblk_d = (int *)(ar + o_selblk);
  b->q_d = (BQuantJob *)(ar + o_q);
  b->qblk_d = (int *)(ar + o_qblk);
  b->qdev_d = (QuantDev *)(ar + o_qdev);
  b->qtab_d = (float *)(ar + o_qtab);
  b->res_d = (wv_thresh_result *)(ar + o_res);
  BCHK(hipHostMalloc((void **)&b->res_h, sizeof(wv_thresh_result) * n, 0), "pinned");
  BCHK(hipHostMalloc((void **)&b->res_h2, sizeof(wv_thresh_result) * n, 0), "pinned");
  BCHK(hipHostMalloc((void **)&b->qdev_h, sizeof(QuantDev) * n, 0), "pinned");
  BCHK(hipHostMalloc((void **)&b->qtab_h, sizeof(float) * (size_t)b->qtab_stride * n, 0),
       "pinned");
  std::memset(b->qdev_h, 0, sizeof(QuantDev) * n);

  // job tables
  std::vector<BSelJob> sj(n);
  std::vector<int> sblk(n);
  std::vector<BQuantJob> qj(n);
  std::vector<int> qblk(n);
  int stot = 0, qtot = 0;
  const int64_t chunk = sel_collect_chunk();
  for (int k = 0; k < n; k++) {
    const int i = b->slot_item[k];
    for (int pt = 0; pt < b->nparts; pt++)
      if (k == b->part_begin[pt]) stot = qtot = 0;   // block ranges restart per part
    const int64_t cnt = b->geoms[i].count;
    BSelJob &s = sj[k];
    s.coef = b->items[i].coef_d;
    s.n = cnt;
    s.frac = b->items[i].thresh_frac;
    s.state = ar + o_state + state_bytes * i;
    s.cand = (uint32_t *)(ar + cand_off[i]);
    s.cap = cand_cap[i];
    s.res = b->res_d + i;
    // at least 8 collect chunks per block, at most the partials capacity
    const int nch = ceil_div(cnt, chunk);
    s.nblk = std::max(1, std::min(sel_max_collect_blocks(), ceil_div(nch, 8)));
    s.blk0 = stot;
    sblk[k] = stot;
    stot += s.nblk;
    BQuantJob &q = qj[k];
    q.coef = b->items[i].coef_d;
    q.n = cnt;
    q.bins = b->items[i].bins_d;
    q.counts = b->items[i].counts_d;
    q.q = b->qdev_d + k;
    q.res = b->res_d + i;
    q.alg = b->items[i].quant_alg;
    q.bin_count = b->items[i].bin_count;
    q.nblk = quant_blocks_for(cnt);
    q.blk0 = qtot;
    qblk[k] = qtot;
    qtot += q.nblk;
    for (int pt = 0; pt < b->nparts; pt++)
      if (k + 1 == b->part_begin[pt + 1]) {
        b->sel_total_p[pt] = stot;
        b->q_total_p[pt] = qtot;
      }
  }
  b->sel_total = b->sel_total_p[0] + (b->nparts > 1 ? b->sel_total_p[1] : 0);
  b->q_total = b->q_total_p[0] + (b->nparts > 1 ? b->q_total_p[1] : 0);
  BCHK(hipMemcpy(b->sel_d, sj.data(), sizeof(BSelJob) * n, hipMemcpyHostToDevice), "H2D");
  BCHK(hipMemcpy(b->selblk_d, sblk.data(), sizeof(int) * n, hipMemcpyHostToDevice), "H2D");
  BCHK(hipMemcpy(b->q_d, qj.data(), sizeof(BQuantJob) * n, hipMemcpyHostToDevice), "H2D");
  BCHK(hipMemcpy(b->qblk_d, qblk.data(), sizeof(int) * n, hipMemcpyHostToDevice), "H2D");
  BCHK(hipMemcpy(b->qdev_d, b->qdev_h, sizeof(QuantDev) * n, hipMemcpyHostToDevice), "H2D");

  // pass pointer tables: pass 0 reads the input; a pass writes the
  // coefficients or, as intermediate, the cubelet's bins buffer
  char *tp = ar + o_tabs;
  for (auto &G : b->groups) {
    const int nc = (int)G.items.size();
    for (int p = 0; p < G.npass; p++) {
      std::vector<const void *> src(nc);
      std::vector<float *> dst(nc);
      for (int c = 0; c < nc; c++) {
        const wv_batch_item &it = b->items[G.items[c]];
        float *inter = (float *)it.bins_d;
        src[c] = p == 0 ? it.in_d
                        : (const void *)(G.dst_out[p - 1] ? it.coef_d : inter);
        dst[c] = G.dst_out[p] ? it.coef_d : inter;
        if (((uintptr_t)dst[c] & 15) != 0) G.dst_aligned = 0;
      }
      G.src_tab_d[p] = (const void **)tp;
      BCHK(hipMemcpy(tp, src.data(), sizeof(void *) * nc, hipMemcpyHostToDevice), "H2D");
      tp += up(sizeof(void *) * nc);
      G.dst_tab_d[p] = (float **)tp;
      BCHK(hipMemcpy(tp, dst.data(), sizeof(void *) * nc, hipMemcpyHostToDevice), "H2D");
      tp += up(sizeof(void *) * nc);
    }
  }
  for (int k = 0; k < BATCH_STREAMS; k++) {
    BCHK(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking), "stream");
    BCHK(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming), "event");
  }
  BCHK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "event");
  for (int pt = 0; pt < 2; pt++)
    BCHK(hipEventCreateWithFlags(&b->part_done[pt], hipEventDisableTiming), "event");
  if (b->host_params) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 16) nt = 16;
    b->pool = new HostPool(nt);
  }
  BCHK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking), "stream");
  if (flags & WV_BATCH_AUTOTUNE) {
    for (auto &G : b->groups) {
      const int rc = autotune_group(b, G, b->side[0]);
      if (rc != WV_OK) { free_batch(b); return rc; }
    }
  }
  *out = b;
  return WV_OK;
}

int wv_batch_submit(wv_batch *b, void *stream) {
  if (!b) { set_error("null batch"); return WV_ERR_ARG; }
  hipStream_t s = (hipStream_t)stream;
  b->last_stream = s;
  b->submitted = true;
  const bool graphs = !(b->flags & WV_BATCH_NO_GRAPH);
  const int n = (int)b->items.size();
  auto run = [&](hipGraphExec_t *g, const EnqueueFn &fn) -> int {
    if (!graphs) return fn(b, s);
    if (!*g) {
      // one eager run first: kernel attributes are set outside any capture
      int rc = fn(b, s);
      if (rc != WV_OK) return rc;
      if (hipStreamSynchronize(s) != hipSuccess) { set_error("batch sync"); return WV_ERR_HIP; }
      rc = capture(b, fn, g);
      if (rc != WV_OK) return rc;
    }
    if (hipGraphLaunch(*g, s) != hipSuccess) { set_error("graph launch"); return WV_ERR_HIP; }
    return WV_OK;
  };
  int rc;
  // transform + select of every part; with log jobs, each part's results
  // come back (the last copy holds every cubelet's final result)
  for (int pt = 0; pt < b->nparts; pt++) {
    rc = run(&b->gA[pt], [pt](wv_batch *bb, hipStream_t ss) { return enqueue_A(bb, ss, pt); });
    if (rc != WV_OK) return rc;
    if (b->host_params) {
      wv_thresh_result *dst = pt == b->nparts - 1 ? b->res_h2 : b->res_h;
      if (hipMemcpyAsync(dst, b->res_d, sizeof(wv_thresh_result) * n, hipMemcpyDeviceToHost,
                         s) != hipSuccess ||
          hipEventRecord(b->part_done[pt], s) != hipSuccess) {
        set_error("batch: threshold results D2H");
        return WV_ERR_HIP;
      }
    }
  }
  // quantize of every part, each after the host built its log tables
  for (int pt = 0; pt < b->nparts; pt++) {
    if (b->host_params) {
      if (hipEventSynchronize(b->part_done[pt]) != hipSuccess) {
        set_error("batch: threshold results");
        return WV_ERR_HIP;
      }
      rc = host_params(b, pt, pt == b->nparts - 1 ? b->res_h2 : b->res_h);
      if (rc != WV_OK) return rc;
      const int pb = b->part_begin[pt], pe = b->part_begin[pt + 1];
      if (hipMemcpyAsync(b->qtab_d + (size_t)pb * b->qtab_stride,
                         b->qtab_h + (size_t)pb * b->qtab_stride,
                         sizeof(float) * (size_t)b->qtab_stride * (pe - pb), hipMemcpyHostToDevice,
                         s) != hipSuccess ||
          hipMemcpyAsync(b->qdev_d + pb, b->qdev_h + pb, sizeof(QuantDev) * (pe - pb),
                         hipMemcpyHostToDevice, s) != hipSuccess) {
        set_error("batch: parameter upload");
        return WV_ERR_HIP;
      }
    }
    rc = run(&b->gB[pt], [pt](wv_batch *bb, hipStream_t ss) { return enqueue_B(bb, ss, pt); });
    if (rc != WV_OK) return rc;
  }
  if (!b->host_params)
    (void)hipMemcpyAsync(b->res_h2, b->res_d, sizeof(wv_thresh_result) * n,
                         hipMemcpyDeviceToHost, s);
  return WV_OK;
}

// One eager run of the whole batch serialised on stream (no side streams,
// no graphs) with events between its stages: stage_ms[0] the transform
// passes of every group (the sum of each group's span), [1] the selects,
// [2] the quantize passes; [3] the number of transform pass launches.
// Synchronises.  The outputs are those of a submit.
int wv_batch_stage_times(wv_batch *b, void *stream, double *stage_ms) {
  if (!b || !stage_ms) { set_error("wv_batch_stage_times: bad arguments"); return WV_ERR_ARG; }
  hipStream_t s = (hipStream_t)stream;
  const int n = (int)b->items.size();
  std::vector<hipEvent_t> ev;
  auto mark = [&]() -> int {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess || hipEventRecord(e, s) != hipSuccess) {
      set_error("wv_batch_stage_times: event");
      return WV_ERR_HIP;
    }
    ev.push_back(e);
    return WV_OK;
  };
  auto cleanup = [&] {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  };
  // segments: (first event, second event, stage)
  std::vector<std::array<int, 3>> seg;
  int launches = 0, rc = WV_OK;
  for (int pt = 0; pt < b->nparts && rc == WV_OK; pt++) {
    for (size_t g = 0; g < b->groups.size() && rc == WV_OK; g++) {
      if (b->group_part[g] != pt) continue;
      const int a = (int)ev.size();
      if ((rc = mark()) != WV_OK) break;
      if ((rc = enqueue_group(b, b->groups[g], s, 0)) != WV_OK) break;
      if ((rc = mark()) != WV_OK) break;
      seg.push_back({a, a + 1, 0});
      launches += b->groups[g].npass;
    }
    if (rc != WV_OK) break;
    const int a = (int)ev.size();
    const int pb = b->part_begin[pt], pe = b->part_begin[pt + 1];
    if ((rc = mark()) != WV_OK) break;
    if (threshold_select_batch(b->sel_d + pb, b->selblk_d + pb, pe - pb, b->sel_total_p[pt], s) !=
        hipSuccess) {
      set_error("wv_batch_stage_times: select");
      rc = WV_ERR_HIP;
      break;
    }
    if ((rc = mark()) != WV_OK) break;
    seg.push_back({a, a + 1, 1});
  }
  if (rc == WV_OK && b->host_params) {
    if (hipMemcpyAsync(b->res_h2, b->res_d, sizeof(wv_thresh_result) * n, hipMemcpyDeviceToHost,
                       s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("wv_batch_stage_times: threshold results");
      rc = WV_ERR_HIP;
    }
    for (int pt = 0; pt < b->nparts && rc == WV_OK; pt++) {
      rc = host_params(b, pt, b->res_h2);
      const int pb = b->part_begin[pt], pe = b->part_begin[pt + 1];
      if (rc == WV_OK &&
          (hipMemcpyAsync(b->qtab_d + (size_t)pb * b->qtab_stride,
                          b->qtab_h + (size_t)pb * b->qtab_stride,
                          sizeof(float) * (size_t)b->qtab_stride * (pe - pb),
                          hipMemcpyHostToDevice, s) != hipSuccess ||
           hipMemcpyAsync(b->qdev_d + pb, b->qdev_h + pb, sizeof(QuantDev) * (pe - pb),
                          hipMemcpyHostToDevice, s) != hipSuccess)) {
        set_error("wv_batch_stage_times: parameter upload");
        rc = WV_ERR_HIP;
      }
    }
  }
  for (int pt = 0; pt < b->nparts && rc == WV_OK; pt++) {
    const int a = (int)ev.size();
    if ((rc = mark()) != WV_OK) break;
    if ((rc = enqueue_B(b, s, pt)) != WV_OK) break;
    if ((rc = mark()) != WV_OK) break;
    seg.push_back({a, a + 1, 2});
  }
  if (rc == WV_OK && hipStreamSynchronize(s) != hipSuccess) {
    set_error("wv_batch_stage_times: sync");
    rc = WV_ERR_HIP;
  }
  if (rc == WV_OK) {
    stage_ms[0] = stage_ms[1] = stage_ms[2] = 0;
    for (auto &g : seg) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, ev[g[0]], ev[g[1]]);
      stage_ms[g[2]] += ms;
    }
    stage_ms[3] = launches;
    if (!b->host_params)
      (void)hipMemcpy(b->res_h2, b->res_d, sizeof(wv_thresh_result) * n, hipMemcpyDeviceToHost);
    b->last_stream = s;
    b->submitted = true;
  }
  cleanup();
  return rc;
}

int wv_batch_poll(wv_batch *b) {
  if (!b) { set_error("null batch"); return WV_ERR_ARG; }
  if (!b->submitted) return 1;
  const hipError_t e = hipStreamQuery(b->last_stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  set_error("batch poll: %s", hipGetErrorString(e));
  return WV_ERR_HIP;
}

int wv_batch_wait(wv_batch *b) {
  if (!b) { set_error("null batch"); return WV_ERR_ARG; }
  if (b->submitted && hipStreamSynchronize(b->last_stream) != hipSuccess) {
    set_error("batch wait");
    return WV_ERR_HIP;
  }
  return WV_OK;
}

int wv_batch_size(const wv_batch *b) { return b ? (int)b->items.size() : 0; }

int wv_batch_groups(const wv_batch *b) { return b ? (int)b->groups.size() : 0; }

int wv_batch_group_tuning(const wv_batch *b, int g, int *tile_cap, int *ept0, int *bpc,
                          int *ncub) {
  if (!b || g < 0 || g >= (int)b->groups.size()) { set_error("bad group"); return WV_ERR_ARG; }
  const Group &G = b->groups[g];
  if (tile_cap) *tile_cap = G.tune.tcap;
  if (ept0) *ept0 = G.tune.ept0;
  if (bpc) *bpc = G.tune.bpc;
  if (ncub) *ncub = (int)G.items.size();
  return WV_OK;
}

int wv_batch_result(const wv_batch *b, int i, wv_thresh_result *res, int *zero_bin) {
  if (!b || i < 0 || i >= (int)b->items.size() || !res) { set_error("bad arguments"); return WV_ERR_ARG; }
  *res = b->res_h2[i];
  if (zero_bin) *zero_bin = (b->items[i].bin_count - 1) >> 1;   // quant(0) for uniform / log
  return WV_OK;
}

}  // extern "C"

namespace wv {
const wv_batch_item *batch_items(const wv_batch *b) { return b ? b->items.data() : nullptr; }
}  // namespace wv

extern "C" {

int wv_batch_destroy(wv_batch *b) {
  free_batch(b);
  return WV_OK;
}

}  // extern "C"
