/*
 * gst_oracle.c -- CPU restatement of the GST decode path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP decoder, never part of it.
 * See gst_oracle.h for what each function restates and how it is pinned.
 */
#include "gst_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

static uint32_t rd32(const uint8_t *p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
         ((uint32_t)p[3] << 24);
}

static uint16_t rd16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }

static uint32_t popc32(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

/* ------------------------------------------------------------------------- */
/* Header                                                                     */

/* codec/codec_base.cpp:18-24 (memcpy of seven uint32), README.md:67-71 (size
 * rules), codec/decoder.cpp:430-476 (stream sizes summed into the upload). */
int gsto_parse_header(const uint8_t *file, size_t len, gsto_header *hdr) {
  if (!file || !hdr) return GSTO_ERR_ARG;
  if (len < GSTO_HEADER_BYTES + GSTO_FREQ_BYTES) return GSTO_ERR_HEADER;
  hdr->width = rd32(file + 0);
  hdr->height = rd32(file + 4);
  hdr->palette_bytes = rd32(file + 8);
  hdr->y_cmp_sz = rd32(file + 12);
  hdr->chroma_cmp_sz = rd32(file + 16);
  hdr->palette_sz = rd32(file + 20);
  hdr->indices_sz = rd32(file + 24);

  const uint64_t w = hdr->width, h = hdr->height;
  if (w == 0 || h == 0 || (w % 128) || (h % 128) || ((w * h) % 131072)) return GSTO_ERR_HEADER;
  const uint64_t nv = w * h / 16;
  if (nv > (1ull << 24)) return GSTO_ERR_HEADER;
  const uint64_t P = hdr->palette_bytes;
  if (P == 0 || (P % GSTO_GROUP_SYMS) || P > 4 * nv) return GSTO_ERR_HEADER;
  const uint64_t syms[4] = {2 * nv, 4 * nv, P, nv};
  const uint64_t sz[4] = {hdr->y_cmp_sz, hdr->chroma_cmp_sz, hdr->palette_sz, hdr->indices_sz};
  uint64_t total = 0;
  for (int k = 0; k < 4; ++k) {
    const uint64_t groups = syms[k] / GSTO_GROUP_SYMS;
    if (sz[k] % 4) return GSTO_ERR_HEADER;
    if (sz[k] < groups * (4 + 4 * GSTO_LANES)) return GSTO_ERR_HEADER;
    total += sz[k];
  }
  if (total >= (1ull << 32)) return GSTO_ERR_HEADER;
  if (GSTO_HEADER_BYTES + GSTO_FREQ_BYTES + total > len) return GSTO_ERR_HEADER;
  return GSTO_OK;
}

/* ------------------------------------------------------------------------- */
/* build_table                                                                */

/* ans/build_table.cl:12-83.  The work-group Blelloch scan (:26-54) runs on a
 * ushort array, so it is the exclusive prefix sum modulo 2^16; addition modulo
 * 2^16 is associative, so the sequential sum below gives the same words.  The
 * branch-free search (:61-75) and the entry store (:78-82) are kept literally. */
void gsto_build_table(const uint16_t freqs[256], gsto_table_entry table[2048]) {
  uint16_t cum[256];
  uint16_t run = 0;
  for (int i = 0; i < 256; ++i) {
    cum[i] = run;
    run = (uint16_t)(run + freqs[i]);
  }
  for (uint32_t id = 0; id < GSTO_TABLE_SIZE; ++id) {
    uint32_t low = 0, high = 255, x = (high + low) / 2;
    for (int i = 0; i < GSTO_TABLE_LOG; ++i) {
      const uint32_t too_high = (uint32_t)(id < cum[x]);
      const uint32_t too_low = (uint32_t)(x < 255 && cum[x + 1] <= id);
      const uint32_t lo1 = low + 1 > x ? low + 1 : x;   /* max(low + 1, x) */
      const uint32_t hi1 = high - 1 < x ? high - 1 : x; /* min(high - 1, x), uint wrap kept */
      low = too_low * lo1 + (1 - too_low) * low;
      high = too_high * hi1 + (1 - too_high) * high;
      x = (high + low) / 2;
    }
    table[id].freq = freqs[x];
    table[id].cum_freq = cum[x];
    table[id].symbol = (uint8_t)x;
  }
}

/* ------------------------------------------------------------------------- */
/* ans_decode                                                                 */

/* ans/ans_decode.cl:25-74 for one 32-lane work-group.  The local-memory mask
 * protocol (:43-62) is a ballot: every lane sets its bit, the barrier makes all
 * bits visible, each lane reads the total and its prefix count, then clears.  The
 * sequential form below computes the mask first, then applies the reads. */
int gsto_ans_decode_group(const gsto_table_entry table[2048], const uint8_t *stream,
                          size_t stream_len, uint32_t num_groups, uint32_t group,
                          uint8_t *out, uint32_t *final_state, uint32_t *final_word) {
  if (group >= num_groups || (uint64_t)num_groups * 4 > stream_len) return GSTO_ERR_STREAM;
  const uint32_t offset = rd32(stream + 4 * group);                       /* :30 */
  const uint32_t lo = group == 0 ? 4 * num_groups : rd32(stream + 4 * (group - 1));
  if ((offset & 3) || (lo & 3) || offset > stream_len || lo > offset ||
      offset - lo < 4 * GSTO_LANES)
    return GSTO_ERR_STREAM;

  uint32_t state[GSTO_LANES];
  for (uint32_t j = 0; j < GSTO_LANES; ++j)                               /* :31 */
    state[j] = rd32(stream + offset - 4 * GSTO_LANES + 4 * j);
  uint32_t next = (offset - GSTO_LANES * 4) / 2;                          /* :32 */
  const uint32_t lo_word = lo / 2;

  for (uint32_t i = 0; i < GSTO_SYMS_PER_LANE; ++i) {
    uint32_t mask = 0;
    uint8_t sym[GSTO_LANES];
    for (uint32_t j = 0; j < GSTO_LANES; ++j) {
      const uint32_t slot = state[j] & (GSTO_TABLE_SIZE - 1);            /* :38 */
      const gsto_table_entry *e = &table[slot];
      state[j] = (state[j] >> GSTO_TABLE_LOG) * e->freq - e->cum_freq + slot;  /* :40-41 */
      if (state[j] < GSTO_STATE_L) mask |= 1u << j;                      /* :44-46 */
      sym[j] = e->symbol;
    }
    const uint32_t total = popc32(mask);                                  /* :51 */
    for (uint32_t j = 0; j < GSTO_LANES; ++j) {
      if (!(mask >> j & 1)) continue;
      const uint32_t skip = total - (popc32(mask & ((1u << j) - 1)) + 1);  /* :54-55 */
      const uint32_t w = next - skip - 1;                                  /* :56 */
      if (w < lo_word || w >= next) return GSTO_ERR_STREAM;
      state[j] = (state[j] << 16) | rd16(stream + 2 * (size_t)w);
    }
    next -= total;                                                        /* :65 */
    for (uint32_t j = 0; j < GSTO_LANES; ++j)                             /* :71-72 */
      out[j * GSTO_SYMS_PER_LANE + GSTO_SYMS_PER_LANE - 1 - i] = sym[j];
  }
  if (final_state) memcpy(final_state, state, sizeof(state));
  if (final_word) *final_word = next;
  return GSTO_OK;
}

/* ------------------------------------------------------------------------- */
/* decode_indices / collect_indices                                           */

void gsto_decode_indices(const uint8_t *diffs, uint32_t nv, int32_t *out) {
  uint32_t acc = 0; /* int32 wrap-around, as the kernels' int adds */
  for (uint32_t i = 0; i < nv; ++i) {
    acc += (uint32_t)((int32_t)diffs[i] - 128);
    out[i] = (int32_t)acc;
  }
}

/* codec/decoder.cpp:304-393 launch sequence; codec/decode_indices.cl:6-64 (scan
 * launch) and :66-84 (down-propagation).  Launches, work-groups and barrier phases
 * are simulated in order; within a phase the work-items touch disjoint words
 * (checked by hand: the scan's write index b and read index a never coincide), so
 * sequential emulation is exact except for the one unordered double store noted in
 * the header, which `race_policy` resolves. */
void gsto_decode_indices_staged(const uint8_t *diffs, uint32_t nv, int32_t *out,
                                int race_policy, int32_t garbage) {
  enum { S = 128, SLOG = 7 };
  int stage = -1;
  for (;;) {
    stage++;
    const uint32_t shift = (uint32_t)stage * SLOG;
    const uint32_t n = shift >= 32 ? 0 : nv >> shift;
    if (n == 0) break;
    const uint32_t global = n < S ? n : ((n + S - 1) / S) * S;
    const uint32_t local = n < S ? n : S;
    const uint32_t idx_offset = 1u << (shift & 31);
    for (uint32_t g0 = 0; g0 < global; g0 += local) {
      int32_t scratch[S];
      for (uint32_t l = 0; l < S; ++l) scratch[l] = garbage + (int32_t)l;
      for (uint32_t l = 0; l < local; ++l) {          /* :17-31 */
        const uint32_t gid = g0 + l;
        const uint32_t gidx = idx_offset * (gid + 1) - 1;
        const uint32_t pgidx = idx_offset * gid - 1;
        if (stage == 0) scratch[l] = (int32_t)diffs[gid] - 128;
        else if (gidx < nv) scratch[l] = out[gidx];
        else if (pgidx < nv) scratch[l] = out[nv - 1];
        else scratch[l] = 0;
      }
      uint32_t offset = 1;                             /* :33-44 up-sweep */
      for (uint32_t i = S >> 1; i > 0; i >>= 1) {
        for (uint32_t tid = 0; tid < local; ++tid) {
          if (tid < i) {
            const uint32_t a = offset * (2 * tid + 1) - 1, b = offset * (2 * tid + 2) - 1;
            scratch[b] = (int32_t)((uint32_t)scratch[b] + (uint32_t)scratch[a]);
          }
        }
        offset *= 2;
      }
      offset >>= 1;                                    /* :46-56 inclusive down-sweep */
      for (uint32_t d = 2; d < S; d *= 2) {
        offset >>= 1;
        for (uint32_t tid = 0; tid < local; ++tid) {
          if (offset > 0 && 0 < tid && tid < d) {
            const uint32_t a = offset * (2 * tid) - 1, b = offset * (2 * tid + 1) - 1;
            scratch[b] = (int32_t)((uint32_t)scratch[b] + (uint32_t)scratch[a]);
          }
        }
      }
      /* :58-63 stores; two work-items of one group can store to out[nv-1]. */
      int have_tail = 0;
      int32_t tail_lo = 0, tail_hi = 0;
      int tail_count = 0;
      for (uint32_t l = 0; l < local; ++l) {
        const uint32_t gid = g0 + l;
        const uint32_t gidx = idx_offset * (gid + 1) - 1;
        const uint32_t pgidx = idx_offset * gid - 1;
        if (stage == 0 || gidx < nv) {
          if (gidx == nv - 1) {
            if (tail_count == 0) tail_lo = scratch[l]; else tail_hi = scratch[l];
            tail_count++;
            have_tail = 1;
          } else {
            out[gidx] = scratch[l];
          }
        } else if (pgidx < nv && gidx >= nv) {
          if (tail_count == 0) tail_lo = scratch[l]; else tail_hi = scratch[l];
          tail_count++;
          have_tail = 1;
        }
      }
      if (have_tail) {
        if (tail_count == 1) out[nv - 1] = tail_lo;
        else out[nv - 1] = race_policy == 0 ? tail_lo : tail_hi;
      }
    }
  }
  while (stage > 0) {                                  /* codec/decoder.cpp:358-393 */
    const int sh = (stage - 1) * SLOG;
    const uint32_t n = nv >> (sh > 0 ? sh : 0);
    const uint32_t global = ((n + S - 1) / S) * S;
    const uint32_t offset = 1u << ((uint32_t)stage * SLOG & 31);
    const uint32_t next_offset = 1u << ((uint32_t)(stage - 1) * SLOG & 31);
    for (uint32_t gid = 0; gid < global; ++gid) {      /* :66-84 */
      const uint32_t grp = gid / S, lid = gid % S;
      const uint32_t gidx = offset * grp - 1;
      const uint32_t tidx = next_offset * (gid + 1) - 1;
      int run = tidx < nv - 1;
      run = run && gidx < nv - 1;
      run = run && grp > 0;
      run = run && lid != S - 1;
      if (run) out[tidx] = (int32_t)((uint32_t)out[tidx] + (uint32_t)out[gidx]);
    }
    stage--;
  }
}

/* ------------------------------------------------------------------------- */
/* inv_wavelet                                                                */

/* codec/inverse_wavelet.cl:14-16 */
static int normalize_index(int idx, int range) {
  const int v = idx - (int)(idx >= range) * (idx - range + 2);
  return v < 0 ? -v : v;
}

enum { WDIM = 32 };

/* codec/inverse_wavelet.cl:28-44: even output 2x of row y, stored transposed. */
static void wavelet_even(const int32_t *src, int32_t *dst, int x, int y, int len, int mid) {
  const int idx = 2 * x;
  const int prev = mid + normalize_index(idx - 1, len) / 2;
  const int next = mid + normalize_index(idx + 1, len) / 2;
  const int32_t sp = src[y * WDIM + prev], sn = src[y * WDIM + next];
  dst[idx * WDIM + y] = src[y * WDIM + x] - (sp + sn + 2) / 4;
}

/* codec/inverse_wavelet.cl:46-64: odd output 2x+1 of row y, stored transposed. */
static void wavelet_odd(const int32_t *src, int32_t *dst, int x, int y, int len, int mid) {
  const int idx = mid + x;
  const int prev = normalize_index(2 * x, len);
  const int next = normalize_index(2 * x + 2, len);
  const int32_t dp = dst[prev * WDIM + y], dn = dst[next * WDIM + y];
  dst[(2 * x + 1) * WDIM + y] = src[y * WDIM + idx] + (dp + dn) / 2;
}

/* One level of codec/inverse_wavelet.cl:116-172: rows y < len over x < len, then
 * columns, on a 32x32 int32 tile held in `src` with `scratch` as the transposed
 * buffer.  Each barrier phase writes words no other work-item of that phase reads,
 * so the phases are emulated as loops. */
static void wavelet_level(int32_t *src, int32_t *scratch, int len) {
  const int mid = len >> 1;
  for (int y = 0; y < len; ++y) for (int x = 0; x < mid; ++x) wavelet_even(src, scratch, x, y, len, mid);
  for (int y = 0; y < len; ++y) for (int x = 0; x < mid; ++x) wavelet_odd(src, scratch, x, y, len, mid);
  for (int y = 0; y < len; ++y) for (int x = 0; x < mid; ++x) wavelet_even(scratch, src, x, y, len, mid);
  for (int y = 0; y < len; ++y) for (int x = 0; x < mid; ++x) wavelet_odd(scratch, src, x, y, len, mid);
}

void gsto_inv_wavelet_level(int32_t tile[1024], int len) {
  int32_t scratch[WDIM * WDIM];
  memset(scratch, 0, sizeof(scratch));
  wavelet_level(tile, scratch, len);
}

/* codec/inverse_wavelet.cl:69-192, one 16x16 work-group over one tile: load with
 * the -128 bias (:93-104), levels len = 2, 4, 8, 16 (:116-150) and 32 (:152-172,
 * two rows per work-item, the same arithmetic), int8 store (:174-191). */
void gsto_inv_wavelet_tile(const uint8_t in[1024], int8_t out[1024]) {
  int32_t local_data[2 * WDIM * WDIM];
  int32_t *src = local_data, *scratch = local_data + WDIM * WDIM;
  for (int k = 0; k < WDIM * WDIM; ++k) src[k] = (int32_t)in[k] - 128;
  for (int len = 2; len <= WDIM; len *= 2) wavelet_level(src, scratch, len);
  for (int k = 0; k < WDIM * WDIM; ++k) out[k] = (int8_t)src[k];
}

/* ------------------------------------------------------------------------- */
/* assemble                                                                   */

/* codec/assemble.cl:39-62: YCoCg -> RGB565 with truncating division; the
 * shifts are done on uint32 so negative channels keep their two's-complement bits
 * as the OpenCL int shifts do, and the ushort store keeps the low 16 bits. */
uint16_t gsto_ycocg_to_rgb565(int y, int co, int cg) {
  const int t = y - (cg / 2);
  const int g = cg + t;
  const int b = (t - co) / 2;
  const int r = b + co;
  return (uint16_t)(((uint32_t)r << 11) | ((uint32_t)g << 5) | (uint32_t)b);
}

static uint32_t palette_word(const uint32_t *palette, uint32_t palette_words, int32_t idx,
                             uint32_t *bad) {
  if ((uint32_t)idx >= palette_words) { (*bad)++; return 0; }
  return palette[(uint32_t)idx];
}

/* codec/assemble.cl:64-81 over the (bx, by) grid of one image. */
static uint32_t assemble_dxt_rows(const int8_t *planes, const int32_t *indices,
                                  const uint32_t *palette, uint32_t palette_words,
                                  uint32_t nv, uint32_t t0, uint32_t t1, uint8_t *out) {
  uint32_t bad = 0;
  for (uint32_t t = t0; t < t1; ++t) {
    const uint16_t ep1 = gsto_ycocg_to_rgb565(planes[0 * nv + t], planes[2 * nv + t], planes[3 * nv + t]);
    const uint16_t ep2 = gsto_ycocg_to_rgb565(planes[1 * nv + t], planes[4 * nv + t], planes[5 * nv + t]);
    const uint32_t word = palette_word(palette, palette_words, indices[t], &bad);
    uint8_t *o = out + 8 * (size_t)t;
    o[0] = (uint8_t)ep1; o[1] = (uint8_t)(ep1 >> 8);
    o[2] = (uint8_t)ep2; o[3] = (uint8_t)(ep2 >> 8);
    o[4] = (uint8_t)word; o[5] = (uint8_t)(word >> 8);
    o[6] = (uint8_t)(word >> 16); o[7] = (uint8_t)(word >> 24);
  }
  return bad;
}

uint32_t gsto_assemble_dxt(const int8_t *planes, const int32_t *indices,
                           const uint32_t *palette, uint32_t palette_words,
                           uint32_t width, uint32_t height, uint8_t *out) {
  const uint32_t nv = width / 4 * (height / 4);
  return assemble_dxt_rows(planes, indices, palette, palette_words, nv, 0, nv, out);
}

/* codec/assemble.cl:83-130: 4-colour palette in 8-bit RGB, then 16 texels. */
static uint32_t assemble_rgb_rows(const int8_t *planes, const int32_t *indices,
                                  const uint32_t *palette, uint32_t palette_words,
                                  uint32_t width, uint32_t height, uint32_t t0, uint32_t t1,
                                  uint8_t *out) {
  const uint32_t bx = width / 4, nv = bx * (height / 4);
  uint32_t bad = 0;
  for (uint32_t t = t0; t < t1; ++t) {
    int p[4][3];
    for (int e = 0; e < 2; ++e) {
      const int y = planes[(uint32_t)e * nv + t];
      const int co = planes[(2 + 2 * (uint32_t)e) * nv + t];
      const int cg = planes[(3 + 2 * (uint32_t)e) * nv + t];
      const int tt = y - (cg / 2);
      const int g = cg + tt;
      const int b = (tt - co) / 2;
      const int r = b + co;
      p[e][0] = (int)(((uint32_t)r << 3) | (uint32_t)(r >> 2));
      p[e][1] = (int)(((uint32_t)g << 2) | (uint32_t)(g >> 4));
      p[e][2] = (int)(((uint32_t)b << 3) | (uint32_t)(b >> 2));
    }
    for (int c = 0; c < 3; ++c) {
      p[2][c] = (2 * p[0][c] + p[1][c]) / 3;
      p[3][c] = (p[0][c] + 2 * p[1][c]) / 3;
    }
    uint32_t idx = palette_word(palette, palette_words, indices[t], &bad);
    const uint32_t gx = t % bx, gy = t / bx;
    for (int i = 0; i < 16; ++i) {
      const int *rgb = p[idx & 3];
      const uint32_t x = 4 * gx + (uint32_t)(i % 4), y = 4 * gy + (uint32_t)(i / 4);
      uint8_t *o = out + 3 * ((size_t)width * y + x);
      o[0] = (uint8_t)rgb[0]; o[1] = (uint8_t)rgb[1]; o[2] = (uint8_t)rgb[2];
      idx >>= 2;
    }
  }
  return bad;
}

uint32_t gsto_assemble_rgb(const int8_t *planes, const int32_t *indices,
                           const uint32_t *palette, uint32_t palette_words,
                           uint32_t width, uint32_t height, uint8_t *out) {
  const uint32_t nv = width / 4 * (height / 4);
  return assemble_rgb_rows(planes, indices, palette, palette_words, width, height, 0, nv, out);
}

/* ------------------------------------------------------------------------- */
/* Whole-image pipeline                                                       */

typedef struct task_pool {
  atomic_size_t next;
  size_t count;
  size_t chunk;
  void (*fn)(void *ctx, size_t begin, size_t end);
  void *ctx;
} task_pool;

static void *task_worker(void *arg) {
  task_pool *tp = (task_pool *)arg;
  for (;;) {
    const size_t b = atomic_fetch_add(&tp->next, tp->chunk);
    if (b >= tp->count) break;
    const size_t e = b + tp->chunk < tp->count ? b + tp->chunk : tp->count;
    tp->fn(tp->ctx, b, e);
  }
  return NULL;
}

static void parallel_for(int nthreads, size_t count, size_t chunk,
                         void (*fn)(void *, size_t, size_t), void *ctx) {
  if (nthreads <= 1 || count <= chunk) {
    fn(ctx, 0, count);
    return;
  }
  task_pool tp;
  atomic_init(&tp.next, 0);
  tp.count = count;
  tp.chunk = chunk;
  tp.fn = fn;
  tp.ctx = ctx;
  pthread_t th[256];
  int n = nthreads > 256 ? 256 : nthreads;
  int started = 0;
  for (int i = 1; i < n; ++i)
    if (pthread_create(&th[started], NULL, task_worker, &tp) == 0) started++;
  task_worker(&tp);
  for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
}

typedef struct decode_ctx {
  const uint8_t *file;
  gsto_header hdr;
  uint32_t nv;
  uint8_t *symbols;
  int8_t *planes;
  int32_t *indices;
  uint8_t *out;
  gsto_table_entry tables[4][2048];
  const uint8_t *stream[4];
  uint32_t stream_len[4];
  uint32_t groups[4];
  uint32_t group_base[5];
  uint32_t out_off[4];
  atomic_int ans_err;
  atomic_uint bad_idx;
  int rgb;
} decode_ctx;

static void ans_task(void *vctx, size_t b, size_t e) {
  decode_ctx *c = (decode_ctx *)vctx;
  for (size_t g = b; g < e; ++g) {
    int k = 0;
    while (g >= c->group_base[k + 1]) k++;
    const uint32_t local = (uint32_t)g - c->group_base[k];
    const int rc = gsto_ans_decode_group(c->tables[k], c->stream[k], c->stream_len[k],
                                         c->groups[k], local,
                                         c->symbols + c->out_off[k] + (size_t)local * GSTO_GROUP_SYMS,
                                         NULL, NULL);
    if (rc != GSTO_OK) atomic_store(&c->ans_err, rc);
  }
}

/* Tile t of plane p: codec/inverse_wavelet.cl:79-83 (plane base and stride) and
 * :93-95, :174-191 (tile position). */
static void wavelet_task(void *vctx, size_t b, size_t e) {
  decode_ctx *c = (decode_ctx *)vctx;
  const uint32_t bx = c->hdr.width / 4, tiles_x = bx / WDIM;
  const uint32_t tiles = c->nv / (WDIM * WDIM);
  for (size_t it = b; it < e; ++it) {
    const uint32_t p = (uint32_t)(it / tiles), t = (uint32_t)(it % tiles);
    const uint8_t *in = c->symbols + (size_t)p * c->nv + (size_t)t * WDIM * WDIM;
    int8_t tile[WDIM * WDIM];
    gsto_inv_wavelet_tile(in, tile);
    const uint32_t tx = t % tiles_x, ty = t / tiles_x;
    int8_t *plane = c->planes + (size_t)p * c->nv;
    for (uint32_t y = 0; y < WDIM; ++y)
      memcpy(plane + (size_t)(ty * WDIM + y) * bx + tx * WDIM, tile + y * WDIM, WDIM);
  }
}

static void assemble_task(void *vctx, size_t b, size_t e) {
  decode_ctx *c = (decode_ctx *)vctx;
  const uint32_t *palette = (const uint32_t *)(c->symbols + c->out_off[2]);
  const uint32_t words = c->hdr.palette_bytes / 4;
  uint32_t bad;
  if (c->rgb)
    bad = assemble_rgb_rows(c->planes, c->indices, palette, words, c->hdr.width,
                            c->hdr.height, (uint32_t)b, (uint32_t)e, c->out);
  else
    bad = assemble_dxt_rows(c->planes, c->indices, palette, words, c->nv, (uint32_t)b,
                            (uint32_t)e, c->out);
  if (bad) atomic_fetch_add(&c->bad_idx, bad);
}

static int decode_common(const uint8_t *file, size_t len, uint8_t *out, gsto_stages *stages,
                         int nthreads, int rgb) {
  if (!file || !out) return GSTO_ERR_ARG;
  decode_ctx *c = (decode_ctx *)calloc(1, sizeof(decode_ctx));
  if (!c) return GSTO_ERR_ARG;
  int rc = gsto_parse_header(file, len, &c->hdr);
  if (rc != GSTO_OK) { free(c); return rc; }
  c->file = file;
  c->rgb = rgb;
  c->out = out;
  const uint32_t nv = c->hdr.width / 4 * (c->hdr.height / 4);
  const uint32_t P = c->hdr.palette_bytes;
  c->nv = nv;
  /* codec/decoder.cpp:441-461 offsets */
  const uint32_t syms[4] = {2 * nv, 4 * nv, P, nv};
  const uint32_t sizes[4] = {c->hdr.y_cmp_sz, c->hdr.chroma_cmp_sz, c->hdr.palette_sz,
                             c->hdr.indices_sz};
  const uint8_t *streams = file + GSTO_HEADER_BYTES + GSTO_FREQ_BYTES;
  uint32_t in_off = 0, out_off = 0;
  c->group_base[0] = 0;
  for (int k = 0; k < 4; ++k) {
    uint16_t freqs[256];
    uint32_t sum = 0;
    for (int i = 0; i < 256; ++i) {
      freqs[i] = rd16(file + GSTO_HEADER_BYTES + 512 * k + 2 * i);
      sum += freqs[i];
    }
    if (sum != GSTO_TABLE_SIZE) { free(c); return GSTO_ERR_FREQ; }
    gsto_build_table(freqs, c->tables[k]);
    c->stream[k] = streams + in_off;
    c->stream_len[k] = sizes[k];
    c->groups[k] = syms[k] / GSTO_GROUP_SYMS;
    c->group_base[k + 1] = c->group_base[k] + c->groups[k];
    c->out_off[k] = out_off;
    in_off += sizes[k];
    out_off += syms[k];
  }
  const size_t nsym = (size_t)7 * nv + P;
  uint8_t *symbols = stages && stages->symbols ? stages->symbols : (uint8_t *)malloc(nsym);
  int8_t *planes = stages && stages->planes ? stages->planes : (int8_t *)malloc((size_t)6 * nv);
  int32_t *indices = stages && stages->indices ? stages->indices
                                                : (int32_t *)malloc((size_t)4 * nv);
  c->symbols = symbols;
  c->planes = planes;
  c->indices = indices;
  atomic_init(&c->ans_err, 0);
  atomic_init(&c->bad_idx, 0);
  if (!symbols || !planes || !indices) { rc = GSTO_ERR_ARG; goto done; }

  parallel_for(nthreads, c->group_base[4], 1, ans_task, c);
  rc = atomic_load(&c->ans_err);
  if (rc != GSTO_OK) goto done;
  parallel_for(nthreads, (size_t)6 * (nv / (WDIM * WDIM)), 16, wavelet_task, c);
  gsto_decode_indices(symbols + c->out_off[3], nv, indices);
  parallel_for(nthreads, nv, 4096, assemble_task, c);
  rc = (int)atomic_load(&c->bad_idx);

done:
  if (!(stages && stages->symbols)) free(symbols);
  if (!(stages && stages->planes)) free(planes);
  if (!(stages && stages->indices)) free(indices);
  free(c);
  return rc;
}

int gsto_decode_dxt(const uint8_t *file, size_t len, uint8_t *out, gsto_stages *stages,
                    int nthreads) {
  return decode_common(file, len, out, stages, nthreads, 0);
}

int gsto_decode_rgb(const uint8_t *file, size_t len, uint8_t *out, int nthreads) {
  return decode_common(file, len, out, NULL, nthreads, 1);
}
