// wv_compress_gpu -- the drop-in driver for test_compress_gpu's compress
// path (test_compress_gpu.cu:150-436, options test_compress_common.cc:35-70,
// 180-290), over libwv_hip.so.
//
//   wv_compress_gpu [opts] <input.cube> <output.cube>
//     -steps <n | x,y | x,y,z>  wavelet steps (default -1: as many as fit)
//     -thresh <f>               fraction of |coef| set to zero (default .5)
//     -qcount <n>               quantization bins (default 256)
//     -wave <haar|cdf97>        wavelet (default cdf97)
//     -qalg <uniform|log|lloyd> quantizer (default log)
//     -noz                      no dup-run coding of the zero bin
//     -err                      round trip on the GPU (dequantize, inverse
//                               transform) and print MSE / PSNR
//     -lloyd-iters <n>          Lloyd iteration cap (default: none)
//     -enc                      print the Huffman encoding of every value
//                               (Huffman::printEncoding, huffman.cc:100-107)
//     -v                        verbose: a line per stage after it ran
//     -bq <f>, -aq <f>, -pb, -experiment
//                               accepted as the reference GPU driver does
//                               (parsed, no effect there; -experiment
//                               implies -q)
//     -2d                       rejected as the reference GPU driver rejects
//                               it (test_compress_gpu.cu:165-168)
//     -q                        quiet
//     -d                        decompress (test_compress_common.cc:80):
//                               the first compressed cubelet of the input
//                               back to an uncompressed cubelet of its
//                               original type and size (decompressFile,
//                               test_compress_cpu.cc:187-236)
//
// Reads the first uncompressed cubelet of the input .cube file (uint8,
// int32 or float32), compresses it on the GPU and writes one compressed
// cubelet with the index footer, byte for byte what the reference GPU
// driver writes for the same input.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/wv_hip.h"

namespace {

struct InputCube {
  int size[3] = {0, 0, 0};
  int dtype = 0;          // proto DataType: UINT8 1, INT32 2, FLOAT32 3
  int max_value = 0;      // maximum_value (0: unset)
  std::vector<uint8_t> data;
};

// --- minimal proto2 reader (varints, length-delimited, fixed32/64) ---------
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
  bool skip(int wt) {
    uint64_t v;
    switch (wt) {
      case 0: return varint(v);
      case 1: if (e - p < 8) return false; p += 8; return true;
      case 2: if (!varint(v) || (uint64_t)(e - p) < v) return false; p += v; return true;
      case 5: if (e - p < 4) return false; p += 4; return true;
      default: return false;
    }
  }
};

// CubeletBuffer (wavelet_compress.proto): width 1, height 2, depth 3,
// data_type 7, byte_count 8, compression_algorithm 9, maximum_value 15
bool parse_cubelet(const uint8_t *buf, size_t len, InputCube &c, uint64_t &byte_count,
                   int &compression) {
  Reader r{buf, buf + len};
  compression = 0;
  byte_count = 0;
  while (r.p < r.e) {
    uint64_t key;
    if (!r.varint(key)) return false;
    const int field = (int)(key >> 3), wt = (int)(key & 7);
    if (wt == 0) {
      uint64_t v;
      if (!r.varint(v)) return false;
      switch (field) {
        case 1: c.size[0] = (int)v; break;
        case 2: c.size[1] = (int)v; break;
        case 3: c.size[2] = (int)v; break;
        case 7: c.dtype = (int)v; break;
        case 8: byte_count = v; break;
        case 9: compression = (int)v; break;
        case 15: c.max_value = (int)(int32_t)v; break;
        default: break;
      }
    } else if (!r.skip(wt)) {
      return false;
    }
  }
  return true;
}

// CubeletStreamReader::open / next / getCubeData (cubelet_file.cc:153-257):
// a 64-byte header, then [u32 length][CubeletBuffer][data] records up to the
// 0xffffffff index marker
bool read_first_uncompressed(const char *path, InputCube &c) {
  FILE *f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); return false; }
  char header[64];
  if (fread(header, 1, 64, f) != 64 || strncmp(header, "SCU cubelets 1.0", 16) != 0) {
    fprintf(stderr, "%s is not a cubelet file\n", path);
    fclose(f);
    return false;
  }
  for (;;) {
    uint32_t len;
    if (fread(&len, 4, 1, f) != 1 || len == 0xffffffffu) break;
    std::vector<uint8_t> msg(len);
    if (fread(msg.data(), 1, len, f) != len) break;
    InputCube cand;
    uint64_t bytes = 0;
    int comp = 0;
    if (!parse_cubelet(msg.data(), len, cand, bytes, comp)) break;
    if (comp == 0) {
      cand.data.resize(bytes);
      if (fread(cand.data.data(), 1, bytes, f) != bytes) break;
      c = std::move(cand);
      fclose(f);
      return true;
    }
    if (fseek(f, (long)bytes, SEEK_CUR) != 0) break;
  }
  fclose(f);
  fprintf(stderr, "No uncompressed cubelets in the input data.\n");
  return false;
}

double now_ms();

// decompressFile: read, decode on the host, dequantize + inverse transform +
// original type on the GPU, write an uncompressed cubelet with the footer
int decompress_file(const char *in_path, const char *out_path, bool quiet) {
  const double t0 = now_ms();
  FILE *f = fopen(in_path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", in_path); return 1; }
  std::vector<uint8_t> file;
  {
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) file.insert(file.end(), buf, buf + n);
    fclose(f);
  }
  wv_cube_info info;
  if (wv_cube_parse(file.data(), (int64_t)file.size(), &info, nullptr, 0, nullptr, nullptr, 0) !=
      WV_OK) {
    fprintf(stderr, "%s\n", wv_last_error());
    return 1;
  }
  if (!quiet)
    printf("Read %dx%dx%d data file: %.2f ms\n", info.size[0], info.size[1], info.size[2],
           now_ms() - t0);
  const int64_t work_bytes = wv_decompress_work_bytes(&info);
  if (work_bytes < 0) { fprintf(stderr, "%s\n", wv_last_error()); return 1; }
  void *work_d = nullptr;
  if (hipMalloc(&work_d, work_bytes) != hipSuccess) { fprintf(stderr, "hipMalloc failed\n"); return 1; }
  const int dt = info.orig_dtype == WV_DTYPE_UINT8 || info.orig_dtype == WV_DTYPE_INT32
                     ? info.orig_dtype : WV_DTYPE_FLOAT32;
  const int64_t out_bytes =
      (int64_t)info.orig[0] * info.orig[1] * info.orig[2] * (dt == WV_DTYPE_UINT8 ? 1 : 4);
  std::vector<uint8_t> out((size_t)out_bytes);
  const double t1 = now_ms();
  double ms_decode = 0;
  const int64_t n = wv_decompress(file.data(), (int64_t)file.size(), work_d, work_bytes,
                                  out.data(), out_bytes, &info, &ms_decode, nullptr);
  (void)hipFree(work_d);
  if (n < 0) { fprintf(stderr, "decompress: %s\n", wv_last_error()); return 1; }
  const double t2 = now_ms();
  std::vector<uint8_t> cube((size_t)out_bytes + (1 << 16));
  const int64_t nb = wv_write_cube_raw(info.orig, dt, out.data(), cube.data(), (int64_t)cube.size());
  if (nb < 0) { fprintf(stderr, "write: %s\n", wv_last_error()); return 1; }
  FILE *g = fopen(out_path, "wb");
  if (!g || fwrite(cube.data(), 1, (size_t)nb, g) != (size_t)nb) {
    fprintf(stderr, "cannot write %s\n", out_path);
    if (g) fclose(g);
    return 1;
  }
  fclose(g);
  if (!quiet) {
    printf("Huffman decode (host): %.2f ms\n", ms_decode);
    printf("Decompress (decode, H2D, dequantize + inverse transform + type on the GPU, D2H): "
           "%.2f ms\n", t2 - t1);
    printf("Write file: %.2f ms\n", now_ms() - t2);
    printf("Total: %.2f ms\n", now_ms() - t0);
  }
  return 0;
}

void usage() {
  fprintf(stderr,
          "\n  wv_compress_gpu [opts] <input.cube> <output.cube>\n"
          "    -steps <n|x,y|x,y,z> -thresh <f> -qcount <n> -wave <haar|cdf97>\n"
          "    -qalg <uniform|log|lloyd> -noz -err -lloyd-iters <n> -q -v -enc\n"
          "    -bq <file> -aq <file> -pb -experiment (accepted, no effect), -2d (rejected)\n"
          "    -opt : parameter optimisation (the reference's default sweep,\n"
          "           threshold fraction x log bin count -> size, pSNR); no output file\n"
          "  wv_compress_gpu -d [-q] <compressed.cube> <output.cube>\n\n");
  exit(1);
}

#define HOK(expr)                                                         \
  do {                                                                    \
    hipError_t _e = (expr);                                               \
    if (_e != hipSuccess) {                                               \
      fprintf(stderr, "%s: %s\n", #expr, hipGetErrorString(_e));          \
      return 1;                                                           \
    }                                                                     \
  } while (0)
#define WOK(expr)                                                         \
  do {                                                                    \
    int _r = (expr);                                                      \
    if (_r != WV_OK) {                                                    \
      fprintf(stderr, "%s: %s\n", #expr, wv_last_error());                \
      return 1;                                                           \
    }                                                                     \
  } while (0)

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch()).count();
}

// -opt: the device OptimizationData of the input, then plotAll
// (optimize.cc:99-122, the reference's default optimizeParameters): for
// threshold fractions .5, .51, ... (float steps, below .999) and bin counts
// 100, 110, 121, ... ((int)(n * 1.1), up to 2000) the log quantizer's size
// and pSNR.  Like the reference, no compressed file is written.
int optimize(const InputCube &cube, const wv_options &opt, bool quiet) {
  const int *orig = cube.size;
  const int maxp = cube.max_value > 0 ? cube.max_value : 0;
  if (cube.dtype != WV_DTYPE_UINT8 && maxp <= 0) {
    fprintf(stderr, "Parameter optimisation needs maximum_value in the input cubelet.\n");
    return 1;
  }
  void *in_d = nullptr;
  HOK(hipMalloc(&in_d, cube.data.size()));
  HOK(hipMemcpy(in_d, cube.data.data(), cube.data.size(), hipMemcpyHostToDevice));
  const double t0 = now_ms();
  wv_opt *o = nullptr;
  WOK(wv_opt_create(in_d, cube.dtype, orig, opt.steps, opt.wavelet, maxp, opt.compress_zeros,
                    &o, nullptr));
  const double t1 = now_ms();
  const int64_t n = wv_opt_count(o);
  printf("Threshold\tBins\tsize\tpSNR\n");
  int points = 0;
  for (float frac = .5f; frac < .999f; frac += 0.01f) {
    const float t = wv_opt_sorted(o, (int64_t)(int)((float)n * frac));
    for (int bins = 100; bins <= 2000; bins = (int)(bins * 1.1)) {
      wv_opt_result r;
      WOK(wv_opt_test(o, t, bins, WV_QUANT_LOG, &r, nullptr));
      printf("%.3f\t%d\t%lld\t%.6g\n", frac, bins, (long long)r.size_bytes, r.psnr);
      points++;
    }
  }
  fflush(stdout);
  const double t2 = now_ms();
  if (!quiet)
    fprintf(stderr, "transform + sort %.2f ms; %d test points in %.1f ms (%.3f ms each)\n",
            t1 - t0, points, t2 - t1, (t2 - t1) / points);
  wv_opt_destroy(o);
  (void)hipFree(in_d);
  return 0;
}

// Huffman::printEncoding (huffman.cc:100-107): the values that occur
// (leaves with a count, the dup key included), by decreasing count, ties
// by increasing value, as "value (count): length bits"
void print_book(const wv_huff *book, int bin_count, bool with_dup) {
  struct Leaf { int v; int64_t count; int len; uint32_t bits[2]; };
  std::vector<Leaf> leaves;
  for (int v = 0; v < bin_count + (with_dup ? 1 : 0); v++) {
    Leaf l{v, 0, 0, {0, 0}};
    if (wv_huff_code(book, v, &l.count, &l.len, l.bits) != WV_OK || l.count <= 0) continue;
    leaves.push_back(l);
  }
  std::stable_sort(leaves.begin(), leaves.end(),
                   [](const Leaf &a, const Leaf &b) { return a.count > b.count; });
  for (const Leaf &l : leaves) {
    std::string s((size_t)l.len, '0');
    for (int i = 0; i < l.len; i++)
      if ((l.bits[i >> 5] >> (i & 31)) & 1u) s[(size_t)i] = '1';
    printf("%3d (%8d): %d %s\n", l.v, (int)l.count, l.len, s.c_str());
  }
}

}  // namespace

int main(int argc, char **argv) {
  wv_options opt;
  std::memset(&opt, 0, sizeof opt);
  opt.wavelet = WV_WAVELET_CDF97;          // DEFAULT_WAVELET (wavelet.h:51)
  opt.quant_alg = WV_QUANT_LOG;            // DEFAULT_QUANTIZE_ALGORITHM
  opt.bin_count = 256;                     // DEFAULT_QUANTIZE_BINS
  opt.steps[0] = opt.steps[1] = opt.steps[2] = -1;   // DEFAULT_WAVELET_STEPS
  opt.thresh_frac = .5f;                   // DEFAULT_THRESHOLD_FRACTION
  opt.compress_zeros = 1;
  opt.lloyd_max_iters = -1;
  opt.with_footer = 1;                     // the GPU driver writes the index
  bool quiet = false, do_err = false, do_decompress = false, do_opt = false;
  bool print_encoding = false, do_2d = false, verbose = false;
  int a = 1;
  for (; a < argc && argv[a][0] == '-'; a++) {
    const char *s = argv[a];
    if (!strcmp(s, "-steps")) {
      if (++a >= argc) usage();
      int x = 1, y = 1, z = 1;
      const int n = sscanf(argv[a], "%d,%d,%d", &x, &y, &z);
      if (n == 1) { y = x; z = x; }
      else if (n == 2) { z = 0; }
      else if (n != 3) { fprintf(stderr, "Invalid # of wavelet transform steps: \"%s\"\n", argv[a]); return 1; }
      opt.steps[0] = x; opt.steps[1] = y; opt.steps[2] = z;
    } else if (!strcmp(s, "-thresh")) {
      if (++a >= argc) usage();
      if (sscanf(argv[a], "%f", &opt.thresh_frac) != 1 || opt.thresh_frac < 0 ||
          opt.thresh_frac > 1) {
        fprintf(stderr, "Invalid threshold proportion: \"%s\"\n", argv[a]);
        return 1;
      }
    } else if (!strcmp(s, "-qcount")) {
      if (++a >= argc) usage();
      if (sscanf(argv[a], "%d", &opt.bin_count) != 1 || opt.bin_count < 1) {
        fprintf(stderr, "Invalid # quantize bins: \"%s\"\n", argv[a]);
        return 1;
      }
    } else if (!strcmp(s, "-wave")) {
      if (++a >= argc) usage();
      if (!strcmp(argv[a], "haar")) opt.wavelet = WV_WAVELET_HAAR;
      else if (!strcmp(argv[a], "cdf97")) opt.wavelet = WV_WAVELET_CDF97;
      else { fprintf(stderr, "Invalid wavelet: \"%s\"\n", argv[a]); return 1; }
    } else if (!strcmp(s, "-qalg")) {
      if (++a >= argc) usage();
      if (!strcmp(argv[a], "uniform")) opt.quant_alg = WV_QUANT_UNIFORM;
      else if (!strcmp(argv[a], "log")) opt.quant_alg = WV_QUANT_LOG;
      else if (!strcmp(argv[a], "lloyd")) opt.quant_alg = WV_QUANT_LLOYD;
      else { fprintf(stderr, "Invalid quantize algorithm: \"%s\"\n", argv[a]); return 1; }
    } else if (!strcmp(s, "-noz")) {
      opt.compress_zeros = 0;
    } else if (!strcmp(s, "-err")) {
      do_err = true;
    } else if (!strcmp(s, "-opt")) {
      do_opt = true;
    } else if (!strcmp(s, "-lloyd-iters")) {
      if (++a >= argc) usage();
      opt.lloyd_max_iters = atoi(argv[a]);
    } else if (!strcmp(s, "-bq") || !strcmp(s, "-aq")) {
      if (++a >= argc) usage();   // file names the GPU driver never writes
    } else if (!strcmp(s, "-pb")) {
      // doPrintQuantizationBins: parsed, unused by the GPU driver
    } else if (!strcmp(s, "-enc")) {
      print_encoding = true;
    } else if (!strcmp(s, "-2d")) {
      do_2d = true;
    } else if (!strcmp(s, "-v")) {
      verbose = true;
    } else if (!strcmp(s, "-experiment")) {
      quiet = true;   // runQuantizationExperiments: sets quiet only here
    } else if (!strcmp(s, "-q")) {
      quiet = true;
    } else if (!strcmp(s, "-d")) {
      do_decompress = true;
    } else {
      fprintf(stderr, "Unrecognized option: \"%s\"\n", s);
      return 1;
    }
  }
  if (argc - a != 2) usage();
  const char *in_path = argv[a], *out_path = argv[a + 1];
  if (do_decompress) return decompress_file(in_path, out_path, quiet);
  if (do_2d) {
    fprintf(stderr, "-2d option not implemented for test_compress_gpu yet.\n");
    return 1;
  }

  const double t0 = now_ms();
  InputCube cube;
  if (!read_first_uncompressed(in_path, cube)) return 1;
  if (cube.dtype < WV_DTYPE_UINT8 || cube.dtype > WV_DTYPE_FLOAT32) {
    fprintf(stderr, "unsupported cubelet data type %d\n", cube.dtype);
    return 1;
  }
  const int *orig = cube.size;
  wv_geom g;
  WOK(wv_geom_init(&g, orig, opt.steps, opt.wavelet, cube.dtype));
  if (!quiet)
    printf("Read %dx%dx%d data file: %.2f ms\n", orig[0], orig[1], orig[2], now_ms() - t0);

  if (do_opt) return optimize(cube, opt, quiet);

  void *in_d = nullptr, *work_d = nullptr;
  const int64_t work_bytes = wv_compress_work_bytes(orig, &opt);
  HOK(hipMalloc(&in_d, cube.data.size()));
  HOK(hipMalloc(&work_d, work_bytes));
  HOK(hipMemcpy(in_d, cube.data.data(), cube.data.size(), hipMemcpyHostToDevice));
  std::vector<float> codebook(opt.quant_alg == WV_QUANT_LLOYD ? 2 * opt.bin_count - 1 : 0);
  wv_compress_info info;
  const double t1 = now_ms();
  WOK(wv_compress_device(in_d, cube.dtype, orig, &opt, work_d, work_bytes,
                         codebook.empty() ? nullptr : codebook.data(), &info, nullptr));
  const double t2 = now_ms();
  // Huffman coding on the device: the bins stay in the work buffer (they
  // overwrite the coefficients); only the payload comes back
  void *enc_ws = nullptr, *words_d = nullptr;
  const int64_t words_cap = wv_encode_words_bound(g.count, opt.bin_count);
  HOK(hipMalloc(&enc_ws, wv_encode_ws_bytes(g.count, opt.bin_count)));
  HOK(hipMalloc(&words_d, words_cap * 4));
  int64_t nwords = 0;
  wv_huff *book = nullptr;
  WOK(wv_encode_device((const int32_t *)work_d, g.count,
                       (const uint32_t *)((char *)work_d + wv_compress_counts_offset(orig, &opt)),
                       opt.bin_count, opt.compress_zeros ? info.zero_bin : -1, enc_ws,
                       (uint32_t *)words_d, words_cap, &nwords, &book, nullptr));
  HOK(hipFree(enc_ws));
  const double t2e = now_ms();
  if (verbose) {
    printf("transform, threshold %g (index %d), quantize + histogram: %.2f ms\n",
           info.thresh.threshold, info.thresh.thresh_idx, t2 - t1);
    printf("Huffman coding on the device: %lld payload words, %.2f ms\n", (long long)nwords,
           t2e - t2);
  }
  if (print_encoding) print_book(book, opt.bin_count, opt.compress_zeros != 0);

  wv_cube_desc d;
  std::memset(&d, 0, sizeof d);
  std::memcpy(d.orig, g.orig, sizeof d.orig);
  std::memcpy(d.steps, g.steps, sizeof d.steps);
  std::memcpy(d.out_size, g.out_size, sizeof d.out_size);
  d.wavelet = opt.wavelet;
  d.quant_alg = opt.quant_alg;
  d.bin_count = opt.bin_count;
  d.orig_dtype = cube.dtype;
  d.thresh_frac = opt.thresh_frac;
  d.threshold = info.thresh.threshold;
  d.max_val = info.thresh.max_abs;
  if (!codebook.empty()) {
    d.boundaries = codebook.data();
    d.values = codebook.data() + (opt.bin_count - 1);
  }
  d.zero_bin = opt.compress_zeros ? info.zero_bin : -1;
  d.with_footer = opt.with_footer;
  std::vector<uint8_t> file((size_t)nwords * 4 + ((size_t)opt.bin_count << 5) + (1 << 20));
  const int64_t payload = nwords * 4;
  // the payload is copied from the device straight into the file image
  const int64_t nbytes = wv_write_cube_device(&d, book, (const uint32_t *)words_d, nwords,
                                              file.data(), (int64_t)file.size(), nullptr);
  wv_huff_free(book);
  HOK(hipFree(words_d));
  if (nbytes < 0) { fprintf(stderr, "encode: %s\n", wv_last_error()); return 1; }
  FILE *f = fopen(out_path, "wb");
  if (!f || fwrite(file.data(), 1, (size_t)nbytes, f) != (size_t)nbytes) {
    fprintf(stderr, "cannot write %s\n", out_path);
    if (f) fclose(f);
    return 1;
  }
  fclose(f);
  const double t3 = now_ms();
  if (!quiet) {
    printf("Threshold %g (index %d, %d nonzero), max |x| %g\n", info.thresh.threshold,
           info.thresh.thresh_idx, info.thresh.nonzero_count, info.thresh.max_abs);
    if (opt.quant_alg == WV_QUANT_LLOYD) printf("Lloyd iterations: %d\n", info.lloyd_iters);
    printf("GPU stages (transform, threshold, quantize + histogram): %.2f ms\n", t2 - t1);
    printf("GPU Huffman coding: %.2f ms\n", t2e - t2);
    printf("Payload D2H + container + write: %.2f ms, %lld bytes (%lld payload)\n", t3 - t2e,
           (long long)nbytes, (long long)payload);
  }

  if (do_err) {
    // dequantize (bins still on the device) and invert on the GPU; compare
    // with the original over its size (computeErrorRatesAfterDequant,
    // test_compress_common.cc:616-694)
    double max_possible = cube.dtype == WV_DTYPE_UINT8
                              ? (cube.max_value > 0 ? cube.max_value : 255)
                              : cube.max_value;
    if (max_possible <= 0) {
      printf("Error metrics need maximum_value in the input cubelet.\n");
    } else {
      wv_quant_params p;
      WOK(wv_quant_params_init(&p, opt.quant_alg, opt.bin_count, info.thresh.threshold,
                               opt.quant_alg == WV_QUANT_LLOYD ? 0 : info.thresh.max_abs,
                               codebook.empty() ? nullptr : codebook.data(),
                               codebook.empty() ? nullptr : codebook.data() + (opt.bin_count - 1)));
      void *tab_d = nullptr, *deq_d = nullptr, *rest_d = nullptr, *ws_d = nullptr;
      HOK(hipMalloc(&tab_d, wv_quant_table_bytes(&p)));
      HOK(hipMalloc(&deq_d, g.count * 4));
      HOK(hipMalloc(&rest_d, g.count * 4));
      HOK(hipMalloc(&ws_d, (g.ws_floats > 0 ? g.ws_floats : 1) * 4));
      WOK(wv_quant_params_upload(&p, tab_d, nullptr));
      WOK(wv_dequantize((const int32_t *)work_d, g.count, &p, (float *)deq_d, nullptr));
      WOK(wv_dwt3d_inverse(&g, (const float *)deq_d, (float *)rest_d, (float *)ws_d, nullptr));
      std::vector<float> rest(g.count);
      HOK(hipMemcpy(rest.data(), rest_d, g.count * 4, hipMemcpyDeviceToHost));
      double sum_sq = 0;
      int64_t cnt = 0;
      for (int z = 0; z < orig[2]; z++)
        for (int y = 0; y < orig[1]; y++) {
          const float *r = rest.data() + ((size_t)z * g.padded[1] + y) * g.padded[0];
          const size_t o = ((size_t)z * orig[1] + y) * orig[0];
          for (int x = 0; x < orig[0]; x++) {
            float a0, b0;
            if (cube.dtype == WV_DTYPE_UINT8) {   // clamp + truncate (wavelet.h:155-163)
              const float v = r[x];
              a0 = cube.data[o + x];
              b0 = v <= 0 ? 0.f : v >= 255 ? 255.f : (float)(unsigned char)v;
            } else if (cube.dtype == WV_DTYPE_INT32) {   // (int)(f+0.5f) (wavelet.h:210-212)
              a0 = (float)((const int32_t *)cube.data.data())[o + x];
              b0 = (float)(int)(r[x] + 0.5f);
            } else {   // compared against (int)restored (test_compress_common.cc:679-686)
              a0 = ((const float *)cube.data.data())[o + x];
              b0 = (float)(int)r[x];
            }
            const double dd = (double)a0 - (double)b0;
            sum_sq += dd * dd;
            cnt++;
          }
        }
      const float mse = cnt ? (float)(sum_sq / cnt) : 0.f;
      const float psnr = mse > 0 ? (float)(10 * log10(max_possible * max_possible / mse)) : 0.f;
      printf("MSE %.6f  PSNR %.4f dB\n", mse, psnr);
      wv_quant_params_free(&p);
      (void)hipFree(tab_d); (void)hipFree(deq_d); (void)hipFree(rest_d); (void)hipFree(ws_d);
    }
  }
  (void)hipFree(in_d);
  (void)hipFree(work_d);
  return 0;
}
