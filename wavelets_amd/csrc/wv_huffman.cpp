// wv_huffman.cpp -- the host half of the compress path: Huffman coding of
// the quantized bins with the zero-bin duplicate-string code, the LSB-first
// word bit stream, and the .cube container writer.
//
// Semantics follow the reference host code the GPU driver calls after the
// device stages: initHuffman(freqCounts, zeroBin) + huffScanDupsOnly
// (test_compress_common.cc:927-973), Huffman::computeHuffmanCoding
// (huffman.cc:42-97, ties broken toward the later node), encodeToStream /
// dupEncodeString (huffman.h:241-296), BitStreamWriter (bit_stream.h:60-162),
// writeQuantData (test_compress_common.cc:198-237), and the container of
// cubelet_file.cc:18-84,259-327 with the metadata of wavelet.cc:283-343.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <queue>
#include <string>
#include <thread>
#include <vector>
#include "wv_internal.h"

namespace wvh {

// run of n copies of the duplicate value: how many 3-bit "dup" codes (for
// 2^(e+3) values, greedy, e <= 7) and how many single codes it needs
static void run_codes(int64_t n, int64_t &dups, int64_t &singles) {
  dups = 0;
  while (n > 8) {
    int e = 7;
    while (e > 0 && (int64_t(1) << (e + 3)) > n) e--;
    n -= int64_t(1) << (e + 3);
    dups++;
  }
  singles = n;
}

struct Code {
  int len = 0;
  std::vector<uint32_t> words;   // LSB-first
};

class Coder {
 public:
  // counts: device histogram of bins (bin_count entries); dup_value < 0
  // disables the duplicate-string code
  Coder(const uint32_t *counts, int bin_count, int dup_value,
        const int32_t *bins, int64_t n)
      : nb_(bin_count), dup_(dup_value >= 0 && dup_value < bin_count ? dup_value : -1) {
    freq_.assign(counts, counts + bin_count);
    if (dup_ >= 0) {
      freq_[dup_] = 0;
      int64_t run = 0, d, s;
      for (int64_t i = 0; i < n; i++) {
        if (bins[i] == dup_) { run++; continue; }
        if (run) { run_codes(run, d, s); freq_[dup_] += s; dupcount_ += d; run = 0; }
      }
      if (run) { run_codes(run, d, s); freq_[dup_] += s; dupcount_ += d; }
      freq_.push_back(dupcount_);
    }
    build();
  }

  // the same code book from finished frequencies: counts with the
  // duplicate value's entry replaced by its single codes, plus the dup
  // key's count (the run statistics the device computes)
  Coder(const uint32_t *counts, int bin_count, int dup_value, int64_t dup_singles,
        int64_t dup_codes)
      : nb_(bin_count), dup_(dup_value >= 0 && dup_value < bin_count ? dup_value : -1) {
    freq_.assign(counts, counts + bin_count);
    if (dup_ >= 0) {
      freq_[dup_] = dup_singles;
      dupcount_ = dup_codes;
      freq_.push_back(dupcount_);
    }
    build();
  }

  int dup() const { return dup_; }
  int leaves() const { return (int)freq_.size(); }
  int64_t freq(int v) const { return freq_[v]; }

  int key() const { return nb_; }
  const std::vector<int32_t> &decode_table() const { return table_; }
  const Code &code(int v) const { return codes_[v]; }
  int longest() const { return longest_; }

  // encodes n bins; returns the number of bits
  int64_t encode(const int32_t *bins, int64_t n, std::vector<uint32_t> &out) const {
    Writer w(out);
    for (int64_t i = 0; i < n; i++) {
      const int v = bins[i];
      if (v == dup_) {
        int64_t len = 1;
        while (i + len < n && bins[i + len] == dup_) len++;
        i += len - 1;
        while (len > 8) {
          int e = 7;
          while (e > 0 && (int64_t(1) << (e + 3)) > len) e--;
          len -= int64_t(1) << (e + 3);
          w.put(codes_[nb_]);
          w.put_bits((uint32_t)e, 3);
        }
        while (len-- > 0) w.put(codes_[v]);
      } else {
        w.put(codes_[v]);
      }
    }
    return w.finish();
  }

 private:
  struct Writer {
    std::vector<uint32_t> &out;
    uint64_t acc = 0;
    int fill = 0;
    int64_t bits = 0;
    explicit Writer(std::vector<uint32_t> &o) : out(o) {}
    void put_bits(uint32_t v, int k) {
      if (k <= 0) return;
      if (k < 32) v &= (1u << k) - 1;
      acc |= (uint64_t)v << fill;
      fill += k;
      bits += k;
      if (fill >= 32) {
        out.push_back((uint32_t)acc);
        acc >>= 32;
        fill -= 32;
      }
    }
    void put(const Code &c) {
      int left = c.len;
      for (size_t w = 0; left > 0; w++) {
        const int k = left > 32 ? 32 : left;
        put_bits(c.words[w], k);
        left -= k;
      }
    }
    int64_t finish() {
      if (fill > 0) out.push_back((uint32_t)acc);
      acc = 0;
      fill = 0;
      return bits;
    }
  };

  struct Node {
    int64_t count;
    int left = -1, right = -1, value = -1;
  };

  void build() {
    const int leaves = (int)freq_.size();
    nodes_.clear();
    nodes_.reserve(2 * leaves);
    // the queue pops the smallest count; equal counts pop the node created
    // later first (the reference compares node addresses in one array)
    auto later_first = [this](int a, int b) {
      if (nodes_[a].count != nodes_[b].count) return nodes_[a].count > nodes_[b].count;
      return a < b;
    };
    std::priority_queue<int, std::vector<int>, decltype(later_first)> pq(later_first);
    for (int i = 0; i < leaves; i++) {
      Node nd;
      nd.count = freq_[i];
      nd.value = i;
      nodes_.push_back(nd);
      if (freq_[i] > 0) pq.push(i);
    }
    while (pq.size() > 1) {
      const int a = pq.top(); pq.pop();
      const int b = pq.top(); pq.pop();
      Node nd;
      nd.count = nodes_[a].count + nodes_[b].count;
      nd.left = a;
      nd.right = b;
      nodes_.push_back(nd);
      pq.push((int)nodes_.size() - 1);
    }
    const int root = pq.top();
    codes_.assign(leaves, Code());
    longest_ = 0;
    std::vector<int> path;   // 0/1 per depth
    assign(root, path);
    // decode table from a breadth-first numbering (huffman.cc:177-228)
    std::vector<int> order{root};
    std::vector<int> id(nodes_.size(), -1);
    for (size_t h = 0; h < order.size(); h++) {
      const int v = order[h];
      id[v] = (int)h;
      if (nodes_[v].value < 0) {
        order.push_back(nodes_[v].left);
        order.push_back(nodes_[v].right);
      }
    }
    table_.resize(2 * order.size());
    for (size_t h = 0; h < order.size(); h++) {
      const Node &nd = nodes_[order[h]];
      if (nd.value >= 0) { table_[2 * h] = 0; table_[2 * h + 1] = nd.value; }
      else { table_[2 * h] = id[nd.left]; table_[2 * h + 1] = id[nd.right]; }
    }
  }

  void assign(int v, std::vector<int> &path) {
    const Node &nd = nodes_[v];
    if (nd.value >= 0) {
      Code &c = codes_[nd.value];
      c.len = (int)path.size();
      c.words.assign(path.size() / 32 + 1, 0u);
      for (size_t i = 0; i < path.size(); i++)
        if (path[i]) c.words[i / 32] |= 1u << (i % 32);
      longest_ = std::max(longest_, c.len);
      return;
    }
    path.push_back(0);
    assign(nd.left, path);
    path.back() = 1;
    assign(nd.right, path);
    path.pop_back();
  }

  int nb_, dup_;
  int64_t dupcount_ = 0;
  std::vector<int64_t> freq_;
  std::vector<Node> nodes_;
  std::vector<Code> codes_;
  std::vector<int32_t> table_;
  int longest_ = 0;
};

// ---- protobuf 2 wire format (only what the .cube metadata uses) ----
class Proto {
 public:
  std::string b;
  void varint(uint64_t v) {
    while (v >= 0x80) { b.push_back((char)(v | 0x80)); v >>= 7; }
    b.push_back((char)v);
  }
  void tag(int field, int wire) { varint(((uint64_t)field << 3) | (uint64_t)wire); }
  void u32(int field, uint32_t v) { tag(field, 0); varint(v); }
  void i32(int field, int32_t v) { tag(field, 0); varint((uint64_t)(int64_t)v); }
  void f32(int field, float v) { tag(field, 5); raw(&v, 4); }
  void raw(const void *p, size_t n) { b.append((const char *)p, n); }
  void bytes(int field, const std::string &s) { tag(field, 2); varint(s.size()); b += s; }
};

static int proto_quant_id(int alg) {   // wavelet.cc:21-35 -> proto enum
  return alg == WV_QUANT_LOG ? 1 : alg == WV_QUANT_UNIFORM ? 2 : 3;
}

static std::string params_msg(const wv_cube_desc *d, const Coder &c,
                              int dup_key, int dup_value) {
  Proto p;
  p.i32(1, d->steps[0]);
  p.i32(2, d->steps[1]);
  p.i32(3, d->steps[2]);
  p.i32(4, d->wavelet);
  p.f32(6, d->thresh_frac);
  p.f32(7, d->threshold);
  p.i32(8, d->bin_count);
  p.i32(9, proto_quant_id(d->quant_alg));
  if (d->quant_alg == WV_QUANT_LLOYD) {
    Proto pb, pv;
    for (int i = 0; i < d->bin_count - 1; i++) pb.raw(&d->boundaries[i], 4);
    for (int i = 0; i < d->bin_count; i++) pv.raw(&d->values[i], 4);
    if (!pb.b.empty()) p.bytes(11, pb.b);
    if (!pv.b.empty()) p.bytes(12, pv.b);
  } else {
    p.f32(10, d->max_val);
  }
  Proto t;
  for (int32_t v : c.decode_table()) t.varint((uint64_t)(int64_t)v);
  if (!t.b.empty()) p.bytes(13, t.b);
  p.u32(14, (uint32_t)d->orig[0]);
  p.u32(15, (uint32_t)d->orig[1]);
  p.u32(16, (uint32_t)d->orig[2]);
  if (d->orig_dtype > 0) p.i32(17, d->orig_dtype);
  p.u32(18, 0);   // transform_2d = false
  if (dup_key >= 0) p.i32(19, dup_key);
  if (dup_value >= 0) p.i32(20, dup_value);
  return p.b;
}

static std::string cubelet_msg(const wv_cube_desc *d, const std::string &params,
                               uint32_t byte_count, uint64_t data_offset,
                               const int *offset = nullptr) {
  Proto m;
  m.u32(1, (uint32_t)d->out_size[0]);
  m.u32(2, (uint32_t)d->out_size[1]);
  m.u32(3, (uint32_t)d->out_size[2]);
  m.u32(4, offset ? (uint32_t)offset[0] : 0u);
  m.u32(5, offset ? (uint32_t)offset[1] : 0u);
  m.u32(6, offset ? (uint32_t)offset[2] : 0u);
  m.u32(7, 2);                  // INT32
  m.u32(8, byte_count);
  m.u32(9, 2);                  // WAVELET
  if (data_offset > 0) { m.tag(13, 0); m.varint(data_offset); }
  m.bytes(14, params);
  return m.b;
}

// The .cube image around a payload of nbytes (LSB-first words): the file
// header and the record's metadata (head), then the payload, then the
// GPU driver's footer (tail).
struct Container {
  std::string head, tail;
  Container(const wv_cube_desc *d, const Coder &coder, uint32_t nbytes) {
    const int dup = coder.dup();
    const int dup_key = dup >= 0 ? d->bin_count : -1;
    const std::string params = params_msg(d, coder, dup_key, dup);
    const std::string meta = cubelet_msg(d, params, nbytes, 0);
    char header[64] = {0};
    std::memcpy(header, "SCU cubelets 1.0\n", 17);
    head.append(header, 64);
    const uint32_t mlen = (uint32_t)meta.size();
    head.append((const char *)&mlen, 4);
    head += meta;
    if (d->with_footer) {
      const int32_t eof = -1;
      tail.append((const char *)&eof, 4);
      Proto index;
      index.bytes(5, cubelet_msg(d, params, nbytes, head.size()));
      tail += index.b;
      const int32_t ilen = (int32_t)index.b.size();
      tail.append((const char *)&ilen, 4);
      tail += "cube";
    }
  }
  int64_t size(uint32_t nbytes) const { return (int64_t)(head.size() + nbytes + tail.size()); }
};

static int64_t write_container(const wv_cube_desc *d, const Coder &coder, const void *payload,
                               uint32_t nbytes, uint8_t *file_buf, int64_t cap) {
  const Container ct(d, coder, nbytes);
  const int64_t total = ct.size(nbytes);
  if (total > cap) return WV_ERR_NOMEM;
  std::memcpy(file_buf, ct.head.data(), ct.head.size());
  if (nbytes) std::memcpy(file_buf + ct.head.size(), payload, nbytes);
  std::memcpy(file_buf + ct.head.size() + nbytes, ct.tail.data(), ct.tail.size());
  return total;
}

}  // namespace wvh

struct wv_huff {
  wvh::Coder coder;
  int bin_count;
  int64_t payload_bits;
  wv_huff(const uint32_t *counts, int b, int dup, int64_t singles, int64_t dups)
      : coder(counts, b, dup, singles, dups), bin_count(b), payload_bits(0) {}
};

extern "C" int wv_huff_code(const wv_huff *book, int value, int64_t *count, int *bit_length,
                            uint32_t bits[2]) {
  if (!book || value < 0 || value >= book->coder.leaves()) {
    wv::set_error("wv_huff_code: value %d outside the code book", value);
    return WV_ERR_ARG;
  }
  const wvh::Code &c = book->coder.code(value);
  if (count) *count = book->coder.freq(value);
  if (bit_length) *bit_length = c.len;
  if (bits) {
    bits[0] = c.words.size() > 0 ? c.words[0] : 0u;
    bits[1] = c.words.size() > 1 ? c.words[1] : 0u;
  }
  return WV_OK;
}

namespace wv {
// the CubeletBuffer message of a compressed cubelet (copyToCubeletBuffer,
// wavelet.cc:283-343) at box offset `offset`, without and with
// data_file_offset (the record's copy and the index's); the parameters
// message (with the decode table) is built once for both
std::string compressed_params_msg(const wv_cube_desc *d, const wv_huff *book) {
  const int dup = book->coder.dup();
  const int dup_key = dup >= 0 ? d->bin_count : -1;
  return wvh::params_msg(d, book->coder, dup_key, dup);
}

void compressed_cubelet_msgs(const wv_cube_desc *d, const std::string &params, uint32_t nbytes,
                             uint64_t data_offset_after_meta, const int offset[3],
                             std::string &meta, std::string &indexed) {
  meta = wvh::cubelet_msg(d, params, nbytes, 0, offset);
  indexed = wvh::cubelet_msg(d, params, nbytes, data_offset_after_meta + meta.size(), offset);
}

// Huffman::totalEncodedLengthBytes (huffman.cc:241-252): every value's code
// length times its count (the dup value's count being its singles), plus
// the dup key's code per dup string -- its 3 length bits are not counted
// -- rounded up to whole 32-bit words (summed in 64 bits)
int64_t huff_total_encoded_bytes(const uint32_t *counts, int bin_count, int dup,
                                 int64_t singles, int64_t dups) {
  const wvh::Coder c(counts, bin_count, dup >= 0 && dup < bin_count ? dup : -1, singles, dups);
  int64_t total = 0;
  for (int v = 0; v < bin_count; v++) total += (int64_t)c.code(v).len * c.freq(v);
  if (c.dup() >= 0) total += dups * (int64_t)c.code(bin_count).len;
  return (total + 31) / 32 * 4;
}
}  // namespace wv

extern "C" int64_t wv_write_cube(const wv_cube_desc *d, const int32_t *bins,
                                 int64_t n, const uint32_t *counts,
                                 uint8_t *file_buf, int64_t cap,
                                 int64_t *payload_bytes) {
  using namespace wvh;
  if (!d || !bins || !counts || n <= 0 || d->bin_count < 1) return WV_ERR_ARG;
  Coder coder(counts, d->bin_count, d->zero_bin, bins, n);
  std::vector<uint32_t> words;
  words.reserve((size_t)(n / 8 + 16));
  coder.encode(bins, n, words);
  const uint32_t nbytes = (uint32_t)(words.size() * 4);
  if (payload_bytes) *payload_bytes = nbytes;
  return write_container(d, coder, words.data(), nbytes, file_buf, cap);
}

extern "C" int64_t wv_encode_ws_bytes(int64_t n, int bin_count) {
  if (n <= 0 || bin_count < 1 || bin_count > WV_MAX_BINS) return -1;
  return wv::enc_ws_bytes(n, bin_count);
}

extern "C" int64_t wv_encode_words_bound(int64_t n, int bin_count) {
  if (n <= 0 || bin_count < 1 || bin_count > WV_MAX_BINS) return -1;
  // Huffman codes average at most log2(leaves) + 1 bits per coded symbol;
  // at most n symbols are coded and each dup code (>= 8 bins) adds 3 bits
  int lg = 0;
  while ((1ll << lg) < (int64_t)bin_count + 1) lg++;
  return (n * (lg + 1) + (n / 8) * 3) / 32 + 2;
}

extern "C" int wv_encode_device(const int32_t *bins_d, int64_t n, const uint32_t *counts_d,
                                int bin_count, int dup_value, void *ws_d, uint32_t *words_d,
                                int64_t words_cap, int64_t *nwords_out, wv_huff **book_out,
                                void *stream) {
  using namespace wv;
  if (book_out) *book_out = nullptr;
  if (!bins_d || !counts_d || !ws_d || !nwords_out || n <= 0 || bin_count < 1 ||
      bin_count > WV_MAX_BINS) {
    set_error("wv_encode_device: bad argument");
    return WV_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const int dup = dup_value >= 0 && dup_value < bin_count ? dup_value : -1;
  std::vector<uint32_t> counts((size_t)bin_count + 1);
  hipError_t e = hipMemcpyAsync(counts.data(), counts_d, counts.size() * 4,
                                hipMemcpyDeviceToHost, s);
  unsigned long long sd[2] = {0, 0};
  // zero-run statistics (and the per-range histograms the payload
  // kernels size their output from); synchronises
  if (e == hipSuccess) e = enc_count_runs(bins_d, n, dup, bin_count, ws_d, sd, s);
  if (e != hipSuccess) {
    set_error("wv_encode_device: %s", hipGetErrorString(e));
    return WV_ERR_HIP;
  }
  // bins outside [0, bin_count) have no code of their own: the host coder
  // gives the value bin_count the dup key's code and has none for others
  const uint64_t outside = counts[bin_count];
  if (outside && dup < 0) {
    set_error("wv_encode_device: %llu bins outside [0, %d) and no dup key",
              (unsigned long long)outside, bin_count);
    return WV_ERR_ARG;
  }
  wv_huff *book = new wv_huff(counts.data(), bin_count, dup, (int64_t)sd[0], (int64_t)sd[1]);
  const wvh::Coder &c = book->coder;
  const int nb1 = c.leaves();
  if (c.longest() > 56) {
    delete book;
    set_error("wv_encode_device: code of %d bits", c.longest());
    return WV_ERR_ARG;
  }
  std::vector<uint64_t> packed((size_t)nb1);
  uint64_t bits = 0;
  for (int v = 0; v < nb1; v++) {
    const wvh::Code &cd = c.code(v);
    const uint64_t cb = cd.len ? (cd.words[0] | (cd.len > 32 ? (uint64_t)cd.words[1] << 32 : 0)) : 0;
    packed[v] = cb | ((uint64_t)cd.len << 56);
    bits += (uint64_t)c.freq(v) * (uint64_t)cd.len;
  }
  if (dup >= 0) bits += 3ull * sd[1] + outside * (uint64_t)c.code(bin_count).len;
  const int64_t nwords = (int64_t)((bits + 31) / 32);
  *nwords_out = nwords;
  if (nwords > words_cap || (nwords > 0 && !words_d)) {
    delete book;
    set_error("wv_encode_device: payload needs %lld words, buffer holds %lld",
              (long long)nwords, (long long)words_cap);
    return WV_ERR_NOMEM;
  }
  e = hipMemcpyAsync(enc_book_ptr(ws_d, n, bin_count), packed.data(), packed.size() * 8,
                     hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);   // packed is a local
  // longest code string of a run shorter than 1024 (its dup codes for the
  // set bits above bit 2, then up to 8 single codes)
  bool short_runs = true;
  if (dup >= 0) {
    const int u = c.code(bin_count).len + 3, zl = c.code(dup).len;
    for (int r = 9; r < 1024 && short_runs; r++) {
      int hi = r >> 3, singles = r & 7;
      if ((hi & 1) && singles == 0) { hi &= ~1; singles = 8; }
      short_runs = __builtin_popcount(hi) * u + singles * zl <= 120;
    }
    short_runs = short_runs && 8 * zl <= 120;
  }
  if (e == hipSuccess && nwords > 0)
    e = enc_write_payload(bins_d, n, dup, bin_count, nb1, short_runs, ws_d, words_d, nwords, s);
  if (e != hipSuccess) {
    delete book;
    set_error("wv_encode_device: %s", hipGetErrorString(e));
    return WV_ERR_HIP;
  }
  book->payload_bits = (int64_t)bits;
  if (book_out) *book_out = book;
  else delete book;
  return WV_OK;
}

namespace wv {

int64_t encode_batch_ws_bytes(const int64_t *n, int nj, int bin_count, int dup) {
  std::vector<const int32_t *> bins((size_t)nj, nullptr);
  std::vector<int> zb((size_t)nj, dup);
  return enc_batch_plan(bins.data(), n, zb.data(), nj, bin_count, nullptr).total;
}

// Device Huffman coding of nj bins arrays with one bin count and one dup
// value: the histograms and run statistics of all of them in one launch
// per phase and one synchronisation, the code books on host threads, all
// payloads in one launch per phase, one D2H.  The same books and words as
// wv_encode_device gives each array alone.
int encode_batch(const int32_t *const *bins_d, const int64_t *n, const uint32_t *const *counts_d,
                 int nj, int bin_count, int dup, void *ws_d, int64_t ws_bytes, uint32_t *words_d,
                 int64_t words_cap, std::vector<uint32_t> &all, std::vector<int64_t> &woff,
                 std::vector<int64_t> &nwords, std::vector<wv_huff *> &books, hipStream_t s,
                 uint32_t **words_dev_out) {
  if (words_dev_out) *words_dev_out = nullptr;
  books.assign((size_t)nj, nullptr);
  woff.assign((size_t)nj, 0);
  nwords.assign((size_t)nj, 0);
  all.clear();
  if (nj <= 0) return WV_OK;
  if (bin_count < 1 || bin_count > WV_MAX_BINS) {
    set_error("encode_batch: %d bins", bin_count);
    return WV_ERR_ARG;
  }
  dup = dup >= 0 && dup < bin_count ? dup : -1;
  std::vector<int> zbs((size_t)nj, dup);
  EncBatch P = enc_batch_plan(bins_d, n, zbs.data(), nj, bin_count, ws_d);
  if (P.total > ws_bytes) {
    set_error("encode_batch: workspace %lld < %lld", (long long)ws_bytes, (long long)P.total);
    return WV_ERR_NOMEM;
  }
  const int m = bin_count + 1;
  std::vector<uint32_t> counts((size_t)nj * m);
  std::vector<unsigned long long> sd((size_t)nj * 2);
  hipError_t e = enc_batch_counts(P, counts_d, ws_d, counts.data(), sd.data(), s);
  if (e != hipSuccess) {
    set_error("encode_batch: run statistics: %s", hipGetErrorString(e));
    return WV_ERR_HIP;
  }
  // code books on host threads
  const int nb1 = P.nb1;
  std::vector<uint64_t> packed((size_t)nj * nb1);
  std::vector<int64_t> bits((size_t)nj);
  std::vector<char> shorts((size_t)nj, 1);
  std::vector<int> rcs((size_t)nj, WV_OK);
  std::vector<std::string> errs((size_t)nj);
  auto book_of = [&](int j) {
    const uint32_t *cj = counts.data() + (size_t)j * m;
    if (cj[bin_count] && dup < 0) {
      rcs[j] = WV_ERR_ARG;
      errs[j] = "bins outside [0, bin_count) and no dup key";
      return;
    }
    wv_huff *book = new wv_huff(cj, bin_count, dup, (int64_t)sd[2 * j], (int64_t)sd[2 * j + 1]);
    const wvh::Coder &c = book->coder;
    if (c.leaves() != nb1 || c.longest() > 56) {
      delete book;
      rcs[j] = WV_ERR_ARG;
      errs[j] = "code longer than 56 bits";
      return;
    }
    uint64_t b = 0;
    uint64_t *pk = packed.data() + (size_t)j * nb1;
    for (int v = 0; v < nb1; v++) {
      const wvh::Code &cd = c.code(v);
      const uint64_t cb =
          cd.len ? (cd.words[0] | (cd.len > 32 ? (uint64_t)cd.words[1] << 32 : 0)) : 0;
      pk[v] = cb | ((uint64_t)cd.len << 56);
      b += (uint64_t)c.freq(v) * (uint64_t)cd.len;
    }
    if (dup >= 0) {
      b += 3ull * sd[2 * j + 1] + (uint64_t)cj[bin_count] * (uint64_t)c.code(bin_count).len;
      const int u = c.code(bin_count).len + 3, zl = c.code(dup).len;
      bool sr = 8 * zl <= 120;
      for (int r = 9; r < 1024 && sr; r++) {
        int hi = r >> 3, singles = r & 7;
        if ((hi & 1) && singles == 0) { hi &= ~1; singles = 8; }
        sr = __builtin_popcount(hi) * u + singles * zl <= 120;
      }
      shorts[j] = sr;
    }
    bits[j] = (int64_t)b;
    nwords[j] = (int64_t)((b + 31) / 32);
    book->payload_bits = (int64_t)b;
    books[j] = book;
  };
  {
    const int nt = std::max(1, std::min(16, nj / 8));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++)
      pool.emplace_back([&, t] {
        for (int j = t; j < nj; j += nt) book_of(j);
      });
    for (auto &th : pool) th.join();
  }
  for (int j = 0; j < nj; j++)
    if (rcs[j] != WV_OK) {
      for (auto *b : books) delete b;
      books.assign((size_t)nj, nullptr);
      set_error("encode_batch: job %d: %s", j, errs[j].c_str());
      return rcs[j];
    }
  int64_t total = 0;
  bool all_short = true;
  for (int j = 0; j < nj; j++) {
    woff[j] = total;
    total += nwords[j];
    all_short = all_short && shorts[j];
  }
  // words_d null: a buffer of exactly the payloads' size, freed here
  void *own = nullptr;
  if (!words_d && total > 0) {
    if (hipMalloc(&own, (size_t)total * 4) != hipSuccess) {
      for (auto *b : books) delete b;
      books.assign((size_t)nj, nullptr);
      set_error("encode_batch: %lld payload words", (long long)total);
      return WV_ERR_NOMEM;
    }
    words_d = (uint32_t *)own;
    words_cap = total;
  }
  if (total > words_cap) {
    for (auto *b : books) delete b;
    books.assign((size_t)nj, nullptr);
    set_error("encode_batch: payloads need %lld words, buffer holds %lld", (long long)total,
              (long long)words_cap);
    return WV_ERR_NOMEM;
  }
  if (!words_dev_out) all.resize((size_t)total);
  e = hipSuccess;
  if (total > 0) {
    e = enc_batch_payload(P, ws_d, packed.data(), all_short, words_d, woff.data(), nwords.data(),
                          total, s);
    if (e == hipSuccess && !words_dev_out)
      e = hipMemcpyAsync(all.data(), words_d, (size_t)total * 4, hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);   // packed is a local
  if (words_dev_out && e == hipSuccess) {
    *words_dev_out = words_d;   // the payloads stay on the device (own: the caller frees)
    own = nullptr;
  }
  if (own) (void)hipFree(own);
  if (e != hipSuccess) {
    for (auto *b : books) delete b;
    books.assign((size_t)nj, nullptr);
    set_error("encode_batch: payloads: %s", hipGetErrorString(e));
    return WV_ERR_HIP;
  }
  return WV_OK;
}

}  // namespace wv

extern "C" int64_t wv_huff_payload_bits(const wv_huff *book) {
  return book ? book->payload_bits : -1;
}

extern "C" int64_t wv_write_cube_payload(const wv_cube_desc *d, const wv_huff *book,
                                         const uint32_t *words, int64_t nwords,
                                         uint8_t *file_buf, int64_t cap) {
  if (!d || !book || (nwords > 0 && !words) || nwords < 0 || d->bin_count != book->bin_count ||
      nwords != (book->payload_bits + 31) / 32)
    return WV_ERR_ARG;
  const int want_dup = d->zero_bin >= 0 && d->zero_bin < d->bin_count ? d->zero_bin : -1;
  if (want_dup != book->coder.dup()) return WV_ERR_ARG;
  return wvh::write_container(d, book->coder, words, (uint32_t)(nwords * 4), file_buf, cap);
}

extern "C" int64_t wv_write_cube_device(const wv_cube_desc *d, const wv_huff *book,
                                        const uint32_t *words_d, int64_t nwords,
                                        uint8_t *file_buf, int64_t cap, void *stream) {
  if (!d || !book || (nwords > 0 && !words_d) || nwords < 0 || d->bin_count != book->bin_count ||
      nwords != (book->payload_bits + 31) / 32)
    return WV_ERR_ARG;
  const int want_dup = d->zero_bin >= 0 && d->zero_bin < d->bin_count ? d->zero_bin : -1;
  if (want_dup != book->coder.dup()) return WV_ERR_ARG;
  const uint32_t nbytes = (uint32_t)(nwords * 4);
  const wvh::Container ct(d, book->coder, nbytes);
  const int64_t total = ct.size(nbytes);
  if (total > cap) return WV_ERR_NOMEM;
  std::memcpy(file_buf, ct.head.data(), ct.head.size());
  hipStream_t s = (hipStream_t)stream;
  if (nbytes) {
    hipError_t e = hipMemcpyAsync(file_buf + ct.head.size(), words_d, nbytes,
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      wv::set_error("wv_write_cube_device: %s", hipGetErrorString(e));
      return WV_ERR_HIP;
    }
  }
  std::memcpy(file_buf + ct.head.size() + nbytes, ct.tail.data(), ct.tail.size());
  return total;
}

extern "C" void wv_huff_free(wv_huff *book) { delete book; }
