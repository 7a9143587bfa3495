// ref_harness.cc -- ORACLE ONLY.  A small extern "C" caller around the
// reference's own, protobuf-free CPU sources (huffman.cc, huffman.h,
// bit_stream.h, bit_stack.h, quant.h, lloyds.cpp), compiled in place from
// /root/reference into oracle/_ref/libref.so.  It is used by the CPU tests to
// pin the C restatement in liboracle.so against the reference's real code.
// Nothing here is shipped or called by the product path.
#include <cstring>
#include <vector>
#include "huffman.h"
#include "quant.h"
#include "lloyds.h"

extern "C" {

// QuantUniform / QuantLog (quant.h:189-439): quantize an array.
int ref_quantize(int alg, int bin_count, float threshold, float max_val,
                 const float *in, int *out, int n) {
  if (alg == 1) {
    QuantUniform q;
    q.init(bin_count, threshold, max_val);
    for (int i = 0; i < n; i++) out[i] = q.quant(in[i]);
  } else if (alg == 2) {
    QuantLog q;
    q.init(bin_count, threshold, max_val);
    for (int i = 0; i < n; i++) out[i] = q.quant(in[i]);
  } else {
    return -1;
  }
  return 0;
}

int ref_dequantize(int alg, int bin_count, float threshold, float max_val,
                   const int *in, float *out, int n) {
  if (alg == 1) {
    QuantUniform q;
    q.init(bin_count, threshold, max_val);
    for (int i = 0; i < n; i++) out[i] = q.dequant(in[i]);
  } else if (alg == 2) {
    QuantLog q;
    q.init(bin_count, threshold, max_val);
    for (int i = 0; i < n; i++) out[i] = q.dequant(in[i]);
  } else {
    return -1;
  }
  return 0;
}

// QuantCodebook host quant (quant.h:515-526)
void ref_quantize_codebook(const float *values, int nv, const float *bounds,
                           int nb, const float *in, int *out, int n) {
  std::vector<float> v(values, values + nv), b(bounds, bounds + nb);
  QuantCodebook q;
  q.init(v, b);
  for (int i = 0; i < n; i++) out[i] = q.quant(in[i]);
}

// lloyd() from lloyds.cpp, unmodified (no iteration cap).
void ref_lloyd(const float *points, unsigned psize, float *codebook,
               unsigned csize, float *partition, float *dist, float *reldist,
               float stop, int *iterations) {
  std::vector<unsigned> groups(psize);
  lloyd(points, psize, codebook, csize, partition, *dist, *reldist,
        groups.data(), stop, iterations);
}

// Huffman built the way the GPU driver does it: initHuffman(freqCounts,
// zeroBin) then the huffScanDupsOnly loop (test_compress_common.cc:927-973),
// which only calls public Huffman methods.  Returns the encoded word count;
// fills the decode table and per-value code lengths.
long long ref_huffman(const int *counts, int bin_count, int zero_bin,
                      const int *data, int n, unsigned *words,
                      long long cap_words, int *decode, int *decode_len,
                      int decode_cap, int *code_len, long long *bits) {
  Huffman h;
  std::vector<int> c(counts, counts + bin_count);
  h.init(c.data(), bin_count);
  if (zero_bin >= 0) {
    h.setDuplicateValue(zero_bin);
    int dupCount = h.getCount(zero_bin);
    h.update(zero_bin, -dupCount);
    int run = 0;
    for (int i = 0; i < n; i++) {
      if (data[i] == zero_bin) {
        run++;
      } else if (run > 0) {
        h.addDupString(run);
        run = 0;
      }
    }
    if (run > 0) h.addDupString(run);
  }
  h.computeHuffmanCoding();
  std::vector<uint32_t> out;
  {
    BitStreamMemorySink sink(&out);
    BitStreamWriter<BitStreamMemorySink> w(&sink);
    h.encodeToStream(&w, data, n);
    *bits = (long long)w.size();
  }
  std::vector<int> table;
  h.getDecoderTable(table);
  *decode_len = (int)table.size();
  for (int i = 0; i < (int)table.size() && i < decode_cap; i++)
    decode[i] = table[i];
  int nvals = bin_count + (zero_bin >= 0 ? 1 : 0);
  for (int v = 0; v < nvals; v++) code_len[v] = h.encodedLength(v);
  if ((long long)out.size() > cap_words) return -1;
  std::memcpy(words, out.data(), out.size() * 4);
  return (long long)out.size();
}

// BitStreamWriter::write(unsigned, int) sequence (bit_stream.h:87-93)
int ref_bitstream(const unsigned *values, const int *nbits, int n,
                  unsigned *words, int cap) {
  std::vector<uint32_t> out;
  {
    BitStreamMemorySink sink(&out);
    BitStreamWriter<BitStreamMemorySink> w(&sink);
    for (int i = 0; i < n; i++) w.write(values[i], nbits[i]);
    w.flush();
  }
  if ((int)out.size() > cap) return -1;
  std::memcpy(words, out.data(), out.size() * 4);
  return (int)out.size();
}

}  // extern "C"
