/*
 * or_split.c -- ORACLE ONLY (test infrastructure).  Restatement of
 * `cube split` on one large cubelet (cube_split_typed,
 * cubelet_convert.cc:424-553), following its frame loop: a set of
 * temporary cubelets, one per x / y box, receives frame after frame; when
 * they hold cub[2] frames they are flushed (written) and their z offset set
 * to frame + 1 - overlap, and the loop backs up `overlap` frames unless it
 * is at the last frame; leftover frames are flushed at the end.
 *
 * Writes each flushed box as 6 ints (x, y, z offset, then x, y, z size)
 * into boxes (up to cap boxes) and, when vol/out are given, the box's
 * samples (esize bytes each, x fastest) packed one box after the other
 * into out.  Returns the number of boxes.
 */
#include <string.h>
#include "wv_oracle.h"

int or_split(const int size[3], const int cub[3], int overlap, int *boxes, int cap,
             const void *vol, int esize, void *out) {
  /* number of cubelets in each direction (cubelet_convert.cc:452-455) */
  int cx = (size[0] - overlap - 1) / (cub[0] - overlap) + 1;
  int cy = (size[1] - overlap - 1) / (cub[1] - overlap) + 1;
  int ntemp = cx * cy;
  int tx[4096], ty[4096], tw[4096], th[4096];
  int tz = 0, depth = 0, nout = 0;
  int64_t out_pos = 0;
  if (ntemp > 4096) return -1;
  for (int y = 0; y < cy; y++)
    for (int x = 0; x < cx; x++) {
      int k = y * cx + x;
      tx[k] = x * (cub[0] - overlap);
      ty[k] = y * (cub[1] - overlap);
      tw[k] = cub[0];
      th[k] = cub[1];
      /* trim cubes that stick out over the end (:477-483) */
      if (tw[k] + tx[k] > size[0]) tw[k] = size[0] - tx[k];
      if (th[k] + ty[k] > size[1]) th[k] = size[1] - ty[k];
    }
  for (int frame = 0; frame < size[2]; frame++) {
    depth++;
    if (depth == cub[2]) {
      for (int k = 0; k < ntemp; k++) {
        if (nout < cap) {
          int *b = boxes + 6 * nout;
          b[0] = tx[k]; b[1] = ty[k]; b[2] = tz;
          b[3] = tw[k]; b[4] = th[k]; b[5] = depth;
          if (vol && out) {
            for (int z = 0; z < depth; z++)
              for (int yy = 0; yy < th[k]; yy++) {
                int64_t src = (((int64_t)(tz + z) * size[1] + ty[k] + yy) * size[0] + tx[k]);
                memcpy((char *)out + out_pos * esize, (const char *)vol + src * esize,
                       (size_t)tw[k] * esize);
                out_pos += tw[k];
              }
          }
        }
        nout++;
      }
      tz = frame + 1 - overlap;
      depth = 0;
      if (frame < size[2] - 1) frame -= overlap;
    }
  }
  if (depth > 0) {
    for (int k = 0; k < ntemp; k++) {
      if (nout < cap) {
        int *b = boxes + 6 * nout;
        b[0] = tx[k]; b[1] = ty[k]; b[2] = tz;
        b[3] = tw[k]; b[4] = th[k]; b[5] = depth;
        if (vol && out) {
          for (int z = 0; z < depth; z++)
            for (int yy = 0; yy < th[k]; yy++) {
              int64_t src = (((int64_t)(tz + z) * size[1] + ty[k] + yy) * size[0] + tx[k]);
              memcpy((char *)out + out_pos * esize, (const char *)vol + src * esize,
                     (size_t)tw[k] * esize);
              out_pos += tw[k];
            }
        }
      }
      nout++;
    }
  }
  return nout;
}
