// wv_cube.cpp -- the decompression side of the .cube path: reading the
// container, Huffman decoding on the host, and writing the reconstructed
// (uncompressed) cubelet file.
//
// decompressFile (test_compress_cpu.cc:187-236) reads the first compressed
// cubelet (readQuantData, test_compress_common.cc:850-894: the first record
// with data type INT32 and wavelet compression), decodes its payload with
// HuffmanDecoder::decodeFromStream (huffman.h:96-131), dequantizes, inverts
// the transform, restores the original data type (translateCubeDataToOriginal
// test_compress_common.cc:699-747) and writes an uncompressed cubelet
// (CubeletStreamWriter::addCubelet / close, cubelet_file.cc:259-327, with the
// metadata of Cube::copyToCubeletBuffer, wavelet.cc:283-299).  The device
// half (dequantize, inverse transform, type restore) is wv_decompress_device.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include "wv_internal.h"

namespace {

// ---- protobuf 2 wire format: reading ----
struct Reader {
  const uint8_t *p, *e;
  bool varint(uint64_t &v) {
    v = 0;
    for (int s = 0; s < 64 && p < e; s += 7) {
      const uint8_t b = *p++;
      v |= (uint64_t)(b & 0x7f) << s;
      if (!(b & 0x80)) return true;
    }
    return false;
  }
  bool fixed32(uint32_t &v) {
    if (e - p < 4) return false;
    std::memcpy(&v, p, 4);
    p += 4;
    return true;
  }
  bool bytes(const uint8_t *&b, size_t &n) {
    uint64_t l;
    if (!varint(l) || (uint64_t)(e - p) < l) return false;
    b = p;
    n = (size_t)l;
    p += l;
    return true;
  }
  bool skip(int wt) {
    uint64_t v;
    uint32_t f;
    const uint8_t *b;
    size_t n;
    switch (wt) {
      case 0: return varint(v);
      case 1: if (e - p < 8) return false; p += 8; return true;
      case 2: return bytes(b, n);
      case 5: return fixed32(f);
      default: return false;
    }
  }
};

inline float f32_of(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// repeated float (packed [packed=true] or one element per field)
bool read_floats(Reader &r, int wt, std::vector<float> &out) {
  if (wt == 5) {
    uint32_t u;
    if (!r.fixed32(u)) return false;
    out.push_back(f32_of(u));
    return true;
  }
  if (wt != 2) return false;
  const uint8_t *b;
  size_t n;
  if (!r.bytes(b, n) || (n & 3)) return false;
  for (size_t i = 0; i < n; i += 4) {
    uint32_t u;
    std::memcpy(&u, b + i, 4);
    out.push_back(f32_of(u));
  }
  return true;
}

bool read_int32s(Reader &r, int wt, std::vector<int32_t> &out) {
  uint64_t v;
  if (wt == 0) {
    if (!r.varint(v)) return false;
    out.push_back((int32_t)v);
    return true;
  }
  if (wt != 2) return false;
  const uint8_t *b;
  size_t n;
  if (!r.bytes(b, n)) return false;
  Reader s{b, b + n};
  while (s.p < s.e) {
    if (!s.varint(v)) return false;
    out.push_back((int32_t)v);
  }
  return true;
}

struct Params {
  wv_cube_info info;
  std::vector<int32_t> decode;
  std::vector<float> bounds, values;
};

// CompressionParametersBuffer (wavelet_compress.proto:11-77)
bool parse_params(const uint8_t *buf, size_t len, Params &P) {
  wv_cube_info &I = P.info;
  Reader r{buf, buf + len};
  I.quant_alg = WV_QUANT_LOG;   // [default = LOG]
  while (r.p < r.e) {
    uint64_t key;
    if (!r.varint(key)) return false;
    const int field = (int)(key >> 3), wt = (int)(key & 7);
    uint64_t v;
    uint32_t u;
    switch (field) {
      case 1: case 2: case 3:
        if (wt != 0 || !r.varint(v)) return false;
        I.steps[field - 1] = (int)(int32_t)v;
        break;
      case 4:
        if (wt != 0 || !r.varint(v)) return false;
        I.wavelet = (int)v;   // proto HAAR 1, CDF97 2 = WV_WAVELET_*
        break;
      case 6: case 7: case 10:
        if (wt != 5 || !r.fixed32(u)) return false;
        (field == 6 ? I.thresh_frac : field == 7 ? I.threshold : I.max_val) = f32_of(u);
        break;
      case 8:
        if (wt != 0 || !r.varint(v)) return false;
        I.bin_count = (int)(int32_t)v;
        break;
      case 9:   // proto LOG 1, UNIFORM 2, LLOYD 3 (wavelet.cc:21-35)
        if (wt != 0 || !r.varint(v)) return false;
        I.quant_alg = v == 1 ? WV_QUANT_LOG : v == 2 ? WV_QUANT_UNIFORM
                    : v == 3 ? WV_QUANT_LLOYD : -1;
        break;
      case 11:
        if (!read_floats(r, wt, P.bounds)) return false;
        break;
      case 12:
        if (!read_floats(r, wt, P.values)) return false;
        break;
      case 13:
        if (!read_int32s(r, wt, P.decode)) return false;
        break;
      case 14: case 15: case 16:
        if (wt != 0 || !r.varint(v)) return false;
        I.orig[field - 14] = (int)v;
        break;
      case 17:
        if (wt != 0 || !r.varint(v)) return false;
        I.orig_dtype = (int)v;   // proto UINT8 1, INT32 2, FLOAT32 3 = WV_DTYPE_*
        break;
      case 19:
        if (wt != 0 || !r.varint(v)) return false;
        I.dup_key = (int)(int32_t)v;
        break;
      case 20:
        if (wt != 0 || !r.varint(v)) return false;
        I.dup_value = (int)(int32_t)v;
        break;
      default:
        if (!r.skip(wt)) return false;
    }
  }
  return true;
}

// CubeletBuffer (wavelet_compress.proto:80-160)
bool parse_cubelet(const uint8_t *buf, size_t len, Params &P, int &dtype, int &compression,
                   uint64_t &byte_count, int &max_value, bool &has_params) {
  Reader r{buf, buf + len};
  dtype = compression = max_value = 0;
  byte_count = 0;
  has_params = false;
  while (r.p < r.e) {
    uint64_t key, v;
    if (!r.varint(key)) return false;
    const int field = (int)(key >> 3), wt = (int)(key & 7);
    if (wt == 0) {
      if (!r.varint(v)) return false;
      switch (field) {
        case 1: case 2: case 3: P.info.size[field - 1] = (int)v; break;
        case 7: dtype = (int)v; break;
        case 8: byte_count = v; break;
        case 9: compression = (int)v; break;
        case 15: max_value = (int)(int32_t)v; break;
        default: break;
      }
    } else if (field == 14 && wt == 2) {
      const uint8_t *b;
      size_t n;
      if (!r.bytes(b, n) || !parse_params(b, n, P)) return false;
      has_params = true;
    } else if (!r.skip(wt)) {
      return false;
    }
  }
  return true;
}

// ---- protobuf 2 wire format: writing (the uncompressed cubelet) ----
struct Proto {
  std::string b;
  void varint(uint64_t v) {
    while (v >= 0x80) { b.push_back((char)(v | 0x80)); v >>= 7; }
    b.push_back((char)v);
  }
  void u32(int field, uint64_t v) { varint((uint64_t)field << 3); varint(v); }
  void bytes(int field, const std::string &s) {
    varint(((uint64_t)field << 3) | 2);
    varint(s.size());
    b += s;
  }
};

// Cube::copyToCubeletBuffer for an uncompressed cube (offsets 0, type,
// compression NONE, byte count); data_file_offset only in the index copy
std::string raw_cubelet_msg(const int size[3], int dtype, uint64_t nbytes, uint64_t offset,
                            const int *box = nullptr) {
  Proto m;
  m.u32(1, (uint32_t)size[0]);
  m.u32(2, (uint32_t)size[1]);
  m.u32(3, (uint32_t)size[2]);
  m.u32(4, box ? (uint32_t)box[0] : 0u);
  m.u32(5, box ? (uint32_t)box[1] : 0u);
  m.u32(6, box ? (uint32_t)box[2] : 0u);
  m.u32(7, (uint32_t)dtype);
  m.u32(8, (uint32_t)nbytes);
  m.u32(9, 0);
  if (offset > 0) m.u32(13, offset);
  return m.b;
}

}  // namespace

extern "C" int wv_cube_parse(const uint8_t *file, int64_t len, wv_cube_info *info,
                             int32_t *decode_table, int decode_cap, float *boundaries,
                             float *values, int codebook_cap) {
  using wv::set_error;
  if (!file || !info || len < 64) {
    set_error("wv_cube_parse: bad argument");
    return WV_ERR_ARG;
  }
  if (std::memcmp(file, "SCU cubelets 1.0", 16) != 0) {
    set_error("not a cubelet file");
    return WV_ERR_ARG;
  }
  int64_t pos = 64;
  for (;;) {
    if (pos + 4 > len) break;
    uint32_t mlen;
    std::memcpy(&mlen, file + pos, 4);
    pos += 4;
    if (mlen == 0xffffffffu || pos + (int64_t)mlen > len) break;
    Params P;
    std::memset(&P.info, 0, sizeof P.info);
    P.info.dup_key = P.info.dup_value = -1;
    P.info.orig_dtype = 0;
    int dtype, comp, maxv;
    uint64_t nbytes;
    bool has_params;
    if (!parse_cubelet(file + pos, mlen, P, dtype, comp, nbytes, maxv, has_params)) {
      set_error("malformed cubelet metadata at offset %lld", (long long)pos);
      return WV_ERR_ARG;
    }
    pos += mlen;
    if (pos + (int64_t)nbytes > len) break;
    // readQuantData: the first INT32 wavelet-compressed cubelet
    if (dtype == WV_DTYPE_INT32 && comp == 2 && has_params) {
      wv_cube_info &I = P.info;
      I.maximum_value = maxv;
      I.payload_offset = pos;
      I.payload_bytes = (int64_t)nbytes;
      I.n_decode = (int)P.decode.size();
      // leaves (pairs (0, v)) hold bin values 0..bin_count-1 or the dup key
      // bin_count; anything else cannot come from the coder
      for (size_t h = 0; h + 1 < P.decode.size(); h += 2)
        if (P.decode[h] == 0 && (P.decode[h + 1] < 0 || P.decode[h + 1] > I.bin_count)) {
          set_error("decode table leaf %d outside [0, %d]", P.decode[h + 1], I.bin_count);
          return WV_ERR_ARG;
        }
      I.n_boundaries = (int)P.bounds.size();
      I.n_values = (int)P.values.size();
      *info = I;
      if (decode_table) {
        if (decode_cap < I.n_decode) { set_error("decode table buffer too small"); return WV_ERR_NOMEM; }
        if (!P.decode.empty()) std::memcpy(decode_table, P.decode.data(), P.decode.size() * 4);
      }
      if (boundaries || values) {
        if (codebook_cap < I.n_boundaries || codebook_cap < I.n_values) {
          set_error("codebook buffers too small");
          return WV_ERR_NOMEM;
        }
        if (boundaries && !P.bounds.empty())
          std::memcpy(boundaries, P.bounds.data(), P.bounds.size() * 4);
        if (values && !P.values.empty())
          std::memcpy(values, P.values.data(), P.values.size() * 4);
      }
      return WV_OK;
    }
    pos += (int64_t)nbytes;
  }
  set_error("No compressed cubelet found");
  return WV_ERR_ARG;
}

// HuffmanDecoder::decodeFromStream (huffman.h:96-131): for each value, walk
// the table from the root one bit at a time (LSB-first words; bits past the
// end read as 0) until a leaf; the dup key then reads 3 more bits e and
// stands for 2^(e+3) copies of the dup value.  Decoding stops when the
// stream is empty at the start of a value.  A 12-bit lookup table resolves
// most codes in one step.
extern "C" int64_t wv_huff_decode_host(const int32_t *table, int n_table, int dup_key,
                                       int dup_value, const uint32_t *words, int64_t nwords,
                                       int32_t *out, int64_t count) {
  using wv::set_error;
  if (!table || n_table < 2 || (nwords > 0 && !words) || (count > 0 && !out) || count < 0) {
    set_error("wv_huff_decode_host: bad argument");
    return WV_ERR_ARG;
  }
  const int64_t total = nwords * 32;
  auto bit_at = [&](int64_t p) -> unsigned {
    return p < total ? (words[p >> 5] >> (p & 31)) & 1u : 0u;
  };
  // 64 bits from position p (zeros past the end)
  auto peek64 = [&](int64_t p) -> uint64_t {
    const int64_t w = p >> 5;
    const int sh = (int)(p & 31);
    const uint64_t w0 = w < nwords ? words[w] : 0u;
    const uint64_t w1 = w + 1 < nwords ? words[w + 1] : 0u;
    const uint64_t lo = w0 | (w1 << 32);   // bits [32 w, 32 w + 64)
    if (!sh) return lo;
    const uint64_t w2 = w + 2 < nwords ? words[w + 2] : 0u;
    return (lo >> sh) | (w2 << (64 - sh));
  };
  // one step down the table from node position pos with bit b; -1 if the
  // table is malformed
  auto step = [&](int pos, unsigned b) -> int {
    if (pos + (int)b >= n_table) return -1;
    const int64_t next = (int64_t)table[pos + b] * 2;
    if (next < 0 || next + 1 >= n_table) return -1;
    return (int)next;
  };
  constexpr int K = 12;
  struct Ent { int32_t v; int16_t len; int16_t leaf; int32_t pos; };
  std::vector<Ent> lut((size_t)1 << K);
  for (uint32_t p = 0; p < (1u << K); p++) {
    int pos = 0;
    Ent e{0, K, 0, 0};
    for (int b = 0; b < K; b++) {
      pos = step(pos, (p >> b) & 1u);
      if (pos < 0) { e.leaf = -1; break; }
      if (table[pos] == 0) {
        e.v = table[pos + 1];
        e.len = (int16_t)(b + 1);
        e.leaf = 1;
        break;
      }
    }
    if (e.leaf == 0) e.pos = pos;
    lut[p] = e;
  }
  int64_t pb = 0, i = 0;
  while (i < count) {
    if (pb >= total) return i;
    const uint64_t win = peek64(pb);
    const Ent &e = lut[win & ((1u << K) - 1)];
    int v;
    if (e.leaf == 1) {
      v = e.v;
      pb += e.len;
    } else if (e.leaf == 0) {
      int pos = e.pos;
      pb += K;
      for (;;) {
        pos = step(pos, bit_at(pb++));
        if (pos < 0) { set_error("malformed Huffman decode table"); return WV_ERR_ARG; }
        if (table[pos] == 0) break;
        if (pb > total + 64) { set_error("Huffman code runs past the stream"); return WV_ERR_ARG; }
      }
      v = table[pos + 1];
    } else {
      set_error("malformed Huffman decode table");
      return WV_ERR_ARG;
    }
    if (v == dup_key) {
      unsigned ebits = 0;
      for (int k = 0; k < 3; k++) ebits |= bit_at(pb++) << k;
      int64_t n = (int64_t)1 << (ebits + 3);
      if (n > count - i) n = count - i;
      for (int64_t k = 0; k < n; k++) out[i + k] = dup_value;
      i += n;
    } else {
      out[i++] = v;
    }
  }
  return count;
}

extern "C" int64_t wv_write_cube_raw(const int size[3], int dtype, const void *data,
                                     uint8_t *file_buf, int64_t cap) {
  if (!size || !data || !file_buf || size[0] <= 0 || size[1] <= 0 || size[2] <= 0 ||
      (dtype != WV_DTYPE_UINT8 && dtype != WV_DTYPE_INT32 && dtype != WV_DTYPE_FLOAT32))
    return WV_ERR_ARG;
  const uint64_t nbytes = (uint64_t)size[0] * size[1] * size[2] * (dtype == WV_DTYPE_UINT8 ? 1 : 4);
  const std::string meta = raw_cubelet_msg(size, dtype, nbytes, 0);
  std::string head;
  char header[64] = {0};
  std::memcpy(header, "SCU cubelets 1.0\n", 17);
  head.append(header, 64);
  const uint32_t mlen = (uint32_t)meta.size();
  head.append((const char *)&mlen, 4);
  head += meta;
  const uint64_t data_offset = head.size();
  std::string tail;
  const int32_t eof = -1;
  tail.append((const char *)&eof, 4);
  Proto index;
  index.bytes(5, raw_cubelet_msg(size, dtype, nbytes, data_offset));
  tail += index.b;
  const int32_t ilen = (int32_t)index.b.size();
  tail.append((const char *)&ilen, 4);
  tail += "cube";
  const int64_t total = (int64_t)(head.size() + nbytes + tail.size());
  if (total > cap) return WV_ERR_NOMEM;
  std::memcpy(file_buf, head.data(), head.size());
  std::memcpy(file_buf + head.size(), data, nbytes);
  std::memcpy(file_buf + head.size() + nbytes, tail.data(), tail.size());
  return total;
}

// ---- many cubelets in one stream (CubeletStreamWriter::addCubelet /
// close, cubelet_file.cc:259-327): each record is its length, the
// CubeletBuffer message and the data; the index repeats every message with
// data_file_offset set
struct wv_cube_stream {
  std::string body;
  std::vector<std::string> index;
  wv_cube_stream() {
    char header[64] = {0};
    std::memcpy(header, "SCU cubelets 1.0\n", 17);
    body.append(header, 64);
  }
};

namespace {

int stream_add_record(wv_cube_stream *w, const std::string &meta,
                      const std::string &indexed, const void *data, uint64_t nbytes) {
  const uint32_t mlen = (uint32_t)meta.size();
  w->body.append((const char *)&mlen, 4);
  w->body += meta;
  if (nbytes) w->body.append((const char *)data, nbytes);
  w->index.push_back(indexed);
  return WV_OK;
}

std::string stream_tail(const wv_cube_stream *w) {
  std::string tail;
  const int32_t eof = -1;
  tail.append((const char *)&eof, 4);
  Proto index;
  for (const std::string &m : w->index) index.bytes(5, m);
  tail += index.b;
  const int32_t ilen = (int32_t)index.b.size();
  tail.append((const char *)&ilen, 4);
  tail += "cube";
  return tail;
}

}  // namespace

extern "C" wv_cube_stream *wv_cube_stream_open(void) { return new wv_cube_stream(); }

namespace wv {
std::string cube_index_tail(const std::vector<std::string> &index) {
  wv_cube_stream w;
  w.index = index;
  return stream_tail(&w);
}

void cube_stream_reserve(wv_cube_stream *w, size_t bytes) {
  if (w) w->body.reserve(w->body.size() + bytes);
}

int cube_stream_add_params(wv_cube_stream *w, const wv_cube_desc *d, const std::string &params,
                           const int offset[3], const uint32_t *words, int64_t nwords) {
  if (!w || !d || !offset || nwords < 0 || (nwords && !words) || nwords > 0x3fffffffll) {
    set_error("wv_cube_stream_add: bad argument");
    return WV_ERR_ARG;
  }
  const uint32_t nbytes = (uint32_t)(nwords * 4);
  std::string meta, indexed;
  // the data follows the length word and the record's message
  compressed_cubelet_msgs(d, params, nbytes, w->body.size() + 4, offset, meta, indexed);
  return stream_add_record(w, meta, indexed, words, nbytes);
}
}  // namespace wv

extern "C" int wv_cube_stream_add(wv_cube_stream *w, const wv_cube_desc *d, const int offset[3],
                                  const wv_huff *book, const uint32_t *words, int64_t nwords) {
  if (!d || !book) {
    wv::set_error("wv_cube_stream_add: bad argument");
    return WV_ERR_ARG;
  }
  return wv::cube_stream_add_params(w, d, wv::compressed_params_msg(d, book), offset, words,
                                    nwords);
}

extern "C" int wv_cube_stream_add_raw(wv_cube_stream *w, const int size[3], int dtype,
                                      const void *data, const int offset[3]) {
  if (!w || !size || !data || !offset || size[0] <= 0 || size[1] <= 0 || size[2] <= 0 ||
      (dtype != WV_DTYPE_UINT8 && dtype != WV_DTYPE_INT32 && dtype != WV_DTYPE_FLOAT32)) {
    wv::set_error("wv_cube_stream_add_raw: bad argument");
    return WV_ERR_ARG;
  }
  const uint64_t nbytes =
      (uint64_t)size[0] * size[1] * size[2] * (dtype == WV_DTYPE_UINT8 ? 1 : 4);
  const std::string meta = raw_cubelet_msg(size, dtype, nbytes, 0, offset);
  const uint64_t data_at = w->body.size() + 4 + meta.size();
  return stream_add_record(w, meta, raw_cubelet_msg(size, dtype, nbytes, data_at, offset), data,
                           nbytes);
}

extern "C" int64_t wv_cube_stream_bytes(const wv_cube_stream *w, int with_footer) {
  if (!w) return WV_ERR_ARG;
  return (int64_t)(w->body.size() + (with_footer ? stream_tail(w).size() : 0));
}

extern "C" int64_t wv_cube_stream_finish(wv_cube_stream *w, int with_footer, uint8_t *file_buf,
                                         int64_t cap) {
  if (!w) return WV_ERR_ARG;
  const std::string tail = with_footer ? stream_tail(w) : std::string();
  const int64_t total = (int64_t)(w->body.size() + tail.size());
  if (!file_buf || total > cap) return WV_ERR_NOMEM;
  std::memcpy(file_buf, w->body.data(), w->body.size());
  if (!tail.empty()) std::memcpy(file_buf + w->body.size(), tail.data(), tail.size());
  return total;
}

extern "C" void wv_cube_stream_free(wv_cube_stream *w) { delete w; }
