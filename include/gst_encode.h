/*
 * gst_encode.h -- seeded synthetic textures and a CPU GST encoder (libgst_encode.so).
 *
 * Input tooling for tests and benchmarks: it produces .gtc files in the reference's
 * on-disk layout (GammaUNC/GST codec/encoder.cpp:122-144, codec/entropy.cpp:174-265)
 * from seeded synthetic RGB8 textures.  It is not byte-identical to the reference
 * encoder `gentenc` (different DXT1 compressor and palette-reuse search); every file
 * it writes decodes with the reference decoder semantics, and the DXT1 blocks it
 * reports are what a correct decoder must return.
 */
#ifndef GST_ENCODE_H_
#define GST_ENCODE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gste_stats {
  uint32_t palette_entries;      /* unique index words kept in the palette */
  uint32_t reused_blocks;        /* blocks that took an earlier palette word */
  uint32_t wavelet_overflows;    /* coefficients that did not fit int8 (then the file
                                    decodes to different endpoints for that tile) */
  uint32_t reserved;
} gste_stats;

/* Seeded synthetic RGB8 texture: two linear gradients per channel, 4-octave value
 * noise of amplitude `noise_milli`/1000 of the 0..255 range, and 8 hard-edged
 * axis-aligned rectangles.  Integer arithmetic only: identical on every host. */
void gste_generate_rgb(uint64_t seed, uint32_t width, uint32_t height, uint32_t noise_milli,
                       uint8_t *rgb);

/* DXT1 (4-colour mode) compression of an RGB8 image, blocks in raster order. */
void gste_compress_dxt1(const uint8_t *rgb, uint32_t width, uint32_t height, uint8_t *dxt);

/* Upper bound of the .gtc size for a width x height image. */
size_t gste_max_encoded_size(uint32_t width, uint32_t height);

/* Encodes DXT1 blocks into a .gtc file.  Palette reuse: a block may take one of the
 * last 127 palette words when that raises its mean squared error (over the 48
 * channel samples of `rgb`) by less than `reuse_threshold` (0 disables reuse;
 * `rgb` may then be NULL).  On reuse the block's index word in `dxt` is replaced,
 * so after the call `dxt` holds the blocks a decoder must reproduce.
 * Returns the file size, or 0 on bad arguments / too small `cap`. */
size_t gste_encode(uint8_t *dxt, const uint8_t *rgb, uint32_t width, uint32_t height,
                   uint32_t reuse_threshold, uint8_t *out, size_t cap, gste_stats *stats);

/* generate + compress + encode in one call.  `dxt_out` (width*height/2 bytes) may be
 * NULL.  Returns the file size or 0. */
size_t gste_make_texture(uint64_t seed, uint32_t width, uint32_t height, uint32_t noise_milli,
                         uint32_t reuse_threshold, uint8_t *out, size_t cap, uint8_t *dxt_out,
                         gste_stats *stats);

/* gste_make_texture over `count` textures on `nthreads` threads.  Texture i goes to
 * out + out_offsets[i] (capacity gste_max_encoded_size) and, if dxt is not NULL, its
 * blocks to dxt + dxt_offsets[i]; its size is written to sizes[i] (0 on failure). */
void gste_make_textures(uint32_t count, const uint64_t *seeds, const uint32_t *widths,
                        const uint32_t *heights, uint32_t noise_milli, uint32_t reuse_threshold,
                        uint8_t *out, const uint64_t *out_offsets, uint8_t *dxt,
                        const uint64_t *dxt_offsets, uint64_t *sizes, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* GST_ENCODE_H_ */
