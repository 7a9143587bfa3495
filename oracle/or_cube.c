/*
 * or_cube.c -- ORACLE ONLY (test infrastructure).  Restatement of the .cube
 * container writer: 64-byte header, [u32 length][CubeletBuffer][payload],
 * optional index footer.  protobuf 2 wire encoding is written by hand
 * (fields in field-number order, only fields that the reference sets).
 */
#include <string.h>
#include "wv_oracle.h"

typedef struct { uint8_t *p; int64_t n, cap; } obuf;

static void put_byte(obuf *b, uint8_t v) { if (b->n < b->cap) b->p[b->n] = v; b->n++; }
static void put_raw(obuf *b, const void *src, int64_t n) {
  for (int64_t i = 0; i < n; i++) put_byte(b, ((const uint8_t *)src)[i]);
}
static void put_varint(obuf *b, uint64_t v) {
  while (v >= 0x80) { put_byte(b, (uint8_t)(v | 0x80)); v >>= 7; }
  put_byte(b, (uint8_t)v);
}
/* int32 fields: negative values are sign-extended to 64 bits */
static void put_int32(obuf *b, int32_t v) { put_varint(b, (uint64_t)(int64_t)v); }
static void put_tag(obuf *b, int field, int wt) { put_varint(b, ((uint64_t)field << 3) | (unsigned)wt); }
static void put_f32(obuf *b, float f) { put_raw(b, &f, 4); }
static int varint_len(uint64_t v) { int n = 1; while (v >= 0x80) { v >>= 7; n++; } return n; }

/* proto enum values (wavelet_compress.proto:23-27,45-49,95-99) and the
   internal->proto quantizer mapping (wavelet.cc:21-35) */
static int proto_quant(int alg) {
  return alg == OR_QUANT_UNIFORM ? 2 : alg == OR_QUANT_LOG ? 1 : 3;
}

/* CompressionParametersBuffer as set by Cube::copyToCubeletBuffer,
   wavelet.cc:304-341 */
static void put_params(obuf *b, const or_cube_meta *m) {
  put_tag(b, 1, 0); put_int32(b, m->steps[0]);
  put_tag(b, 2, 0); put_int32(b, m->steps[1]);
  put_tag(b, 3, 0); put_int32(b, m->steps[2]);
  put_tag(b, 4, 0); put_varint(b, (uint64_t)m->wavelet);
  put_tag(b, 6, 5); put_f32(b, m->threshold_fraction);
  put_tag(b, 7, 5); put_f32(b, m->threshold_value);
  put_tag(b, 8, 0); put_int32(b, m->bin_count);
  put_tag(b, 9, 0); put_varint(b, (uint64_t)proto_quant(m->quant_alg));
  if (m->quant_alg == OR_QUANT_UNIFORM || m->quant_alg == OR_QUANT_LOG) {
    put_tag(b, 10, 5); put_f32(b, m->max_value);
  } else {
    if (m->n_boundaries > 0) {
      put_tag(b, 11, 2); put_varint(b, (uint64_t)m->n_boundaries * 4);
      for (int i = 0; i < m->n_boundaries; i++) put_f32(b, m->boundaries[i]);
    }
    if (m->n_values > 0) {
      put_tag(b, 12, 2); put_varint(b, (uint64_t)m->n_values * 4);
      for (int i = 0; i < m->n_values; i++) put_f32(b, m->values[i]);
    }
  }
  if (m->n_decode > 0) {
    uint64_t len = 0;
    for (int i = 0; i < m->n_decode; i++)
      len += (uint64_t)varint_len((uint64_t)(int64_t)m->decode_table[i]);
    put_tag(b, 13, 2); put_varint(b, len);
    for (int i = 0; i < m->n_decode; i++) put_int32(b, m->decode_table[i]);
  }
  put_tag(b, 14, 0); put_varint(b, (uint32_t)m->orig_width);
  put_tag(b, 15, 0); put_varint(b, (uint32_t)m->orig_height);
  put_tag(b, 16, 0); put_varint(b, (uint32_t)m->orig_depth);
  if (m->orig_dtype > 0) { put_tag(b, 17, 0); put_varint(b, (uint64_t)m->orig_dtype); }
  put_tag(b, 18, 0); put_varint(b, 0);    /* transform_2d = false */
  if (m->dup_key >= 0) { put_tag(b, 19, 0); put_int32(b, m->dup_key); }
  if (m->dup_value >= 0) { put_tag(b, 20, 0); put_int32(b, m->dup_value); }
}

/* CubeletBuffer of a compressed cube, wavelet.cc:283-303 plus
   addCubelet's byte_count (cubelet_file.cc:271-276); data_file_offset
   (field 13) only in the footer copy (cubelet_file.cc:281). */
static void put_cubelet(obuf *b, const or_cube_meta *m, int64_t dfo) {
  put_tag(b, 1, 0); put_varint(b, (uint32_t)m->width);
  put_tag(b, 2, 0); put_varint(b, (uint32_t)m->height);
  put_tag(b, 3, 0); put_varint(b, (uint32_t)m->depth);
  put_tag(b, 4, 0); put_varint(b, 0);
  put_tag(b, 5, 0); put_varint(b, 0);
  put_tag(b, 6, 0); put_varint(b, 0);
  put_tag(b, 7, 0); put_varint(b, 2);      /* INT32 */
  put_tag(b, 8, 0); put_varint(b, (uint64_t)m->byte_count);
  put_tag(b, 9, 0); put_varint(b, 2);      /* WAVELET */
  if (dfo > 0) { put_tag(b, 13, 0); put_varint(b, (uint64_t)dfo); }
  obuf sizer = { NULL, 0, 0 };
  put_params(&sizer, m);
  put_tag(b, 14, 2); put_varint(b, (uint64_t)sizer.n);
  put_params(b, m);
  if (m->max_possible > 0) { put_tag(b, 15, 0); put_int32(b, m->max_possible); }
}

int or_cube_meta_bytes(const or_cube_meta *m, int64_t dfo, uint8_t *buf,
                       int cap) {
  obuf b = { buf, 0, cap };
  put_cubelet(&b, m, dfo);
  return b.n <= cap ? (int)b.n : -1;
}

/* CubeletStreamWriter::open/addCubelet/close, cubelet_file.cc:153-327 */
int64_t or_cube_write(const or_cube_meta *m, const uint32_t *payload,
                      int with_footer, uint8_t *buf, int64_t cap) {
  obuf b = { buf, 0, cap };
  uint8_t header[64];
  memset(header, 0, sizeof header);
  memcpy(header, "SCU cubelets 1.0\n", 17);
  put_raw(&b, header, 64);
  obuf sizer = { NULL, 0, 0 };
  put_cubelet(&sizer, m, 0);
  uint32_t len = (uint32_t)sizer.n;
  put_raw(&b, &len, 4);
  put_cubelet(&b, m, 0);
  int64_t dfo = b.n;
  put_raw(&b, payload, m->byte_count);
  if (with_footer) {
    int32_t eof_marker = -1;
    put_raw(&b, &eof_marker, 4);
    obuf inner = { NULL, 0, 0 };
    put_cubelet(&inner, m, dfo);
    int64_t start = b.n;
    put_tag(&b, 5, 2);             /* CubeletIndexBuffer.cubelets */
    put_varint(&b, (uint64_t)inner.n);
    put_cubelet(&b, m, dfo);
    int32_t index_len = (int32_t)(b.n - start);
    put_raw(&b, &index_len, 4);
    put_raw(&b, "cube", 4);
  }
  return b.n <= cap ? b.n : -1;
}
