/*
 * or_huff.c -- ORACLE ONLY (test infrastructure).  CPU restatement of the
 * reference Huffman coder with its duplicate-string extension for runs of
 * the zero bin, the LSB-first 32-bit-word bit stream, and the decoder.
 */
#include <stdlib.h>
#include <string.h>
#include "wv_oracle.h"

/* HuffmanDup, huffman.h:13-58: a run longer than 8 uses key + 3 bits e
   covering 2^(e+3) values, greedily longest-first; the rest are singles. */
#define DUP_BITS 3
#define DUP_OFFSET 3
static int dup_decode_len(unsigned v) { return 1 << (v + DUP_OFFSET); }
static int dup_min_len(void) { return dup_decode_len(0); }
static unsigned dup_find_longest(int length) {
  const unsigned maxv = (1u << DUP_BITS) - 1;
  if (length >= dup_decode_len(maxv)) return maxv;
  unsigned e = 0, len = (unsigned)dup_min_len();
  while (length >= (int)(len << 1)) { e++; len <<= 1; }
  return e;
}
static void dup_count(int length, int *singles, int *dups) {
  *dups = 0;
  while (length > dup_min_len()) {
    unsigned e = dup_find_longest(length);
    length -= dup_decode_len(e);
    (*dups)++;
  }
  *singles = length;
}

struct or_huffman {
  int size;            /* bin count; dup key is 'size' */
  int dup_value;       /* -1: none */
  int64_t *counts;     /* size entries */
  int64_t dup_value_count;
  /* tree */
  int n_leaves;        /* size (+1 with dup) */
  int n_nodes;
  int *left, *right, *value;
  int64_t *ncount;
  int root;
  int *code_len;       /* per leaf value */
  uint32_t **code;     /* per leaf value, LSB-first words */
  int longest;
  int32_t *decode;     /* BFS table */
  int decode_len;
};

/* Huffman::addDupString, huffman.h:210-217 */
static void add_dup_string(or_huffman *h, int length) {
  int s, d;
  dup_count(length, &s, &d);
  h->counts[h->dup_value] += s;
  h->dup_value_count += d;
}

/* priority order of CompareHuffmanNode, huffman.cc:9-18: the queue top is
   the smallest count; among equal counts the node stored later in the node
   array (higher address) comes first.  Returns 1 if a pops before b. */
static int pops_before(const or_huffman *h, int a, int b) {
  if (h->ncount[a] != h->ncount[b]) return h->ncount[a] < h->ncount[b];
  return a > b;
}
static void heap_push(const or_huffman *h, int *heap, int *n, int v) {
  int i = (*n)++;
  heap[i] = v;
  while (i > 0) {
    int p = (i - 1) / 2;
    if (!pops_before(h, heap[i], heap[p])) break;
    int t = heap[i]; heap[i] = heap[p]; heap[p] = t; i = p;
  }
}
static int heap_pop(const or_huffman *h, int *heap, int *n) {
  int top = heap[0];
  heap[0] = heap[--(*n)];
  int i = 0;
  for (;;) {
    int l = 2 * i + 1, r = l + 1, m = i;
    if (l < *n && pops_before(h, heap[l], heap[m])) m = l;
    if (r < *n && pops_before(h, heap[r], heap[m])) m = r;
    if (m == i) break;
    int t = heap[i]; heap[i] = heap[m]; heap[m] = t; i = m;
  }
  return top;
}

static void assign_codes(or_huffman *h, int node, uint32_t *bits, int depth) {
  if (h->value[node] >= 0) {
    int v = h->value[node];
    int words = depth / 32 + 1;   /* BitStack keeps a partial word */
    h->code_len[v] = depth;
    h->code[v] = (uint32_t *)calloc((size_t)words, 4);
    memcpy(h->code[v], bits, (size_t)words * 4);
    if (depth > h->longest) h->longest = depth;
    return;
  }
  bits[depth / 32] &= ~(1u << (depth % 32));
  assign_codes(h, h->left[node], bits, depth + 1);
  bits[depth / 32] |= (1u << (depth % 32));
  assign_codes(h, h->right[node], bits, depth + 1);
  bits[depth / 32] &= ~(1u << (depth % 32));
}

/* Huffman::computeHuffmanCoding huffman.cc:42-97, computeEncodeTable
   :110-144, computeEncodedBits :147-174, computeDecodeTable :177-211,
   orderNodes (BFS) :214-228. */
static void build(or_huffman *h) {
  const int nn = h->size + (h->dup_value >= 0 ? 1 : 0);
  h->n_leaves = nn;
  const int cap = 2 * nn - 1;
  h->left = (int *)malloc(sizeof(int) * cap);
  h->right = (int *)malloc(sizeof(int) * cap);
  h->value = (int *)malloc(sizeof(int) * cap);
  h->ncount = (int64_t *)malloc(sizeof(int64_t) * cap);
  int *heap = (int *)malloc(sizeof(int) * cap);
  int hn = 0;
  for (int i = 0; i < nn; i++) {
    h->value[i] = i;
    h->left[i] = h->right[i] = -1;
    h->ncount[i] = (i < h->size) ? h->counts[i] : h->dup_value_count;
    if (h->ncount[i] > 0) heap_push(h, heap, &hn, i);
  }
  int next = nn;
  while (hn > 1) {
    int a = heap_pop(h, heap, &hn);
    int b = heap_pop(h, heap, &hn);
    h->value[next] = -1;
    h->left[next] = a;
    h->right[next] = b;
    h->ncount[next] = h->ncount[a] + h->ncount[b];
    heap_push(h, heap, &hn, next);
    next++;
  }
  h->n_nodes = next;
  h->root = heap[0];
  free(heap);

  h->code_len = (int *)calloc((size_t)nn, sizeof(int));
  h->code = (uint32_t **)calloc((size_t)nn, sizeof(uint32_t *));
  h->longest = 0;
  uint32_t bits[64];
  memset(bits, 0, sizeof bits);
  assign_codes(h, h->root, bits, 0);

  /* decode table from a breadth-first ordering */
  int *order = (int *)malloc(sizeof(int) * (size_t)h->n_nodes);
  int *id = (int *)malloc(sizeof(int) * (size_t)h->n_nodes);
  int qh = 0, qt = 0;
  order[qt++] = h->root;
  while (qh < qt) {
    int n = order[qh];
    id[n] = qh;
    qh++;
    if (h->value[n] < 0) { order[qt++] = h->left[n]; order[qt++] = h->right[n]; }
  }
  h->decode_len = 2 * qt;
  h->decode = (int32_t *)malloc(sizeof(int32_t) * (size_t)h->decode_len);
  for (int i = 0; i < qt; i++) {
    int n = order[i];
    if (h->value[n] >= 0) { h->decode[2 * i] = 0; h->decode[2 * i + 1] = h->value[n]; }
    else { h->decode[2 * i] = id[h->left[n]]; h->decode[2 * i + 1] = id[h->right[n]]; }
  }
  free(order);
  free(id);
}

static or_huffman *alloc_h(int bin_count, int zero_bin) {
  or_huffman *h = (or_huffman *)calloc(1, sizeof *h);
  h->size = bin_count;
  h->counts = (int64_t *)calloc((size_t)bin_count, sizeof(int64_t));
  /* Huffman::setDuplicateValue huffman.h:178-181 */
  h->dup_value = (zero_bin < 0 || zero_bin >= bin_count) ? -1 : zero_bin;
  return h;
}

/* initHuffman without freqCounts + huffScan,
   test_compress_common.cc:901-921,975-981 */
or_huffman *or_huff_create_scan(const int32_t *data, int count, int bin_count,
                                int zero_bin) {
  or_huffman *h = alloc_h(bin_count, zero_bin);
  int run = 0;
  for (int i = 0; i < count; i++) {
    if (data[i] == zero_bin) {
      run++;
    } else {
      if (run > 0) { add_dup_string(h, run); run = 0; }
      h->counts[data[i]]++;
    }
  }
  if (run > 0) add_dup_string(h, run);
  build(h);
  return h;
}

/* initHuffman with freqCounts + huffScanDupsOnly,
   test_compress_common.cc:927-950,960-973 */
or_huffman *or_huff_create_counts(const int32_t *counts, const int32_t *data,
                                  int count, int bin_count, int zero_bin) {
  or_huffman *h = alloc_h(bin_count, zero_bin);
  for (int i = 0; i < bin_count; i++) h->counts[i] = counts[i];
  if (zero_bin >= 0) {
    h->counts[zero_bin] = 0;
    int run = 0;
    for (int i = 0; i < count; i++) {
      if (data[i] == zero_bin) run++;
      else if (run > 0) { add_dup_string(h, run); run = 0; }
    }
    if (run > 0) add_dup_string(h, run);
  }
  build(h);
  return h;
}

void or_huff_free(or_huffman *h) {
  if (!h) return;
  for (int i = 0; i < h->n_leaves; i++) free(h->code[i]);
  free(h->code); free(h->code_len); free(h->decode);
  free(h->left); free(h->right); free(h->value); free(h->ncount);
  free(h->counts); free(h);
}

void or_huff_get_counts(const or_huffman *h, int64_t *out) {
  for (int i = 0; i < h->size; i++) out[i] = h->counts[i];
  out[h->size] = h->dup_value >= 0 ? h->dup_value_count : 0;
}

int or_huff_code_length(const or_huffman *h, int v) { return h->code_len[v]; }

/* Huffman::totalEncodedLengthBytes, huffman.cc:241-252: sum over values of
   code length times count, plus the dup key's length per dup string (the
   3-bit length fields are not counted), rounded up to 32-bit words.  The
   reference sums in int; 64 bits here. */
int64_t or_huff_total_encoded_bytes(const or_huffman *h) {
  int64_t total = 0;
  for (int v = 0; v < h->size; v++) total += (int64_t)h->code_len[v] * h->counts[v];
  if (h->dup_value >= 0) total += h->dup_value_count * h->code_len[h->size];
  return (total + 31) / 32 * 4;
}
int or_huff_longest(const or_huffman *h) { return h->longest; }

int or_huff_decode_table(const or_huffman *h, int32_t *table, int cap) {
  if (table && cap >= h->decode_len)
    memcpy(table, h->decode, sizeof(int32_t) * (size_t)h->decode_len);
  return h->decode_len;
}

/* BitStreamWriter (bit_stream.h:60-162) is an LSB-first packer into 32-bit
   words with a final partial word flushed. */
typedef struct { uint32_t *w; int64_t cap, nbits; int overflow; } bitw;
static void put_bits(bitw *b, const uint32_t *src, int n) {
  for (int i = 0; i < n; i++) {
    int64_t pos = b->nbits + i;
    int64_t wi = pos >> 5;
    if (wi >= b->cap) { b->overflow = 1; return; }
    if ((pos & 31) == 0) b->w[wi] = 0;
    if ((src[i >> 5] >> (i & 31)) & 1u) b->w[wi] |= 1u << (pos & 31);
  }
  b->nbits += n;
}

/* Huffman::encodeToStream huffman.h:241-265, dupEncodeString :278-296 */
int64_t or_huff_encode(const or_huffman *h, const int32_t *data, int count,
                       uint32_t *words, int64_t cap, int64_t *bits_out) {
  bitw b = { words, cap, 0, 0 };
  const int key = h->size;
  for (int i = 0; i < count; i++) {
    int v = data[i];
    if (v == h->dup_value) {
      int len = 1;
      while (len + i < count && data[i + len] == h->dup_value) len++;
      i += len - 1;
      while (len > dup_min_len()) {
        unsigned e = dup_find_longest(len);
        len -= dup_decode_len(e);
        put_bits(&b, h->code[key], h->code_len[key]);
        put_bits(&b, &e, DUP_BITS);
      }
      while (len > 0) { put_bits(&b, h->code[v], h->code_len[v]); len--; }
    } else {
      put_bits(&b, h->code[v], h->code_len[v]);
    }
    if (b.overflow) return -1;
  }
  if (bits_out) *bits_out = b.nbits;
  return (b.nbits + 31) / 32;
}

/* HuffmanDecoder::decodeFromStream huffman.h:96-131 over a
   BitStreamReader (bit_stream.h:216-311): reading past the last word
   yields zero bits and marks the stream empty. */
int or_huff_decode(const int32_t *table, int table_len, int dup_key,
                   int dup_value, const uint32_t *words, int64_t n_words,
                   int32_t *out, int count) {
  int64_t pos_bit = 0, total = n_words * 32;
  (void)table_len;
  for (int i = 0; i < count; i++) {
    int pos = 0;
    if (pos_bit >= total) return i;
    for (;;) {
      unsigned bit = 0;
      if (pos_bit < total) bit = (words[pos_bit >> 5] >> (pos_bit & 31)) & 1u;
      pos_bit++;
      pos = table[pos + (int)bit] * 2;
      if (table[pos] == 0) {
        int v = table[pos + 1];
        if (v == dup_key) {
          unsigned e = 0;
          for (int k = 0; k < DUP_BITS; k++) {
            unsigned bb = 0;
            if (pos_bit < total) bb = (words[pos_bit >> 5] >> (pos_bit & 31)) & 1u;
            pos_bit++;
            e |= bb << k;
          }
          int n = dup_decode_len(e);
          while (n > 0 && i < count) { out[i++] = dup_value; n--; }
          i--;
        } else {
          out[i] = v;
        }
        break;
      }
    }
  }
  return count;
}

int or_bitstream_pack(const uint32_t *values, const int32_t *nbits, int n,
                      uint32_t *words, int cap) {
  bitw b = { words, cap, 0, 0 };
  for (int i = 0; i < n; i++) {
    if (nbits[i] <= 0) continue;
    put_bits(&b, &values[i], nbits[i]);
    if (b.overflow) return -1;
  }
  return (int)((b.nbits + 31) / 32);
}
