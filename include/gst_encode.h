/*
 * gst_encode.h -- the GST encoder and seeded synthetic textures (libgst_encode.so).
 *
 * gste_compress_dxt is the reference encoder `gentenc` (GenTC::CompressDXT) byte for
 * byte: same DXT1 blocks, same .gtc file.  The other entry points are input tooling
 * for tests and benchmarks: a fast encoder (own DXT1 compressor and palette-reuse
 * search, not byte-identical to gentenc) over seeded synthetic RGB8 textures.  Every
 * file either writes has the reference's on-disk layout (GammaUNC/GST
 * codec/encoder.cpp:122-144, codec/entropy.cpp:174-265), and the DXT1 blocks they
 * report are what a correct decoder must return.
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

/* Status codes of gste_compress_dxt. */
#define GSTE_OK 0
#define GSTE_ERR_ARGUMENT (-1) /* NULL rgb/out_size, or a size gentenc cannot code: width and
                                  height multiples of 128 and width*height of 131072 */
#define GSTE_ERR_CAPACITY (-2) /* out is NULL or out_cap is short; *out_size holds the size */
#define GSTE_ERR_RANGE (-3)    /* a palette-index step does not fit a byte (gentenc asserts;
                                  its 127-entry lookback does not produce one) */
#define GSTE_ERR_UNCODABLE (-4) /* a stream holds zero bytes only, which gentenc's rANS
                                  coder refuses on an assertion (ans/ans_utils.h:42-47) */

/* GenTC::CompressDXT(width, height, rgb, dxt) and GenTC::CompressDXT(DXTImage(width,
 * height, rgb)) (codec/encoder.cpp:176-187, codec/dxt_image.cpp:444-468): the .gtc
 * file gentenc writes for an RGB8 image (`rgb`, width*height*3 bytes, rows top down).
 * `dxt_in` (width*height/2 bytes) gives the DXT1 blocks to start from; when NULL the
 * blocks are compressed as gentenc does, with stb_dxt v1.04 in its high-quality mode
 * (codec/dxt_image.cpp:362-385).  The blocks then go through gentenc's palette reuse
 * (codec/dxt_image.cpp:512-624), and `dxt_out` (width*height/2 bytes, may be NULL)
 * receives them: the image the file decodes to.  *out_size receives the file size
 * (also when out_cap is short).  Returns a GSTE_* status.  As in gentenc, a wavelet
 * coefficient outside [-128, 127] wraps, and such a file decodes to other endpoints
 * than `dxt_out`.  Single-threaded, no state between calls: concurrent calls are safe. */
int gste_compress_dxt(const uint8_t *rgb, uint32_t width, uint32_t height, const uint8_t *dxt_in,
                      uint8_t *out, size_t out_cap, size_t *out_size, uint8_t *dxt_out);

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

/* The same textures encoded as gentenc does: texture i is gste_generate_rgb(seeds[i],
 * widths[i], heights[i], noise_milli) through gste_compress_dxt (stb_dxt blocks, then
 * gentenc's palette reuse), so each file is byte for byte the reference encoder's.
 * Textures are taken largest first by `nthreads` threads; out + out_offsets[i] holds
 * gste_max_encoded_size bytes; dxt (may be NULL) receives the blocks the file decodes
 * to.  sizes[i] = file size, or 0 when gste_compress_dxt refused the texture. */
void gste_make_textures_gentenc(uint32_t count, const uint64_t *seeds, const uint32_t *widths,
                                const uint32_t *heights, uint32_t noise_milli, uint8_t *out,
                                const uint64_t *out_offsets, uint8_t *dxt, const uint64_t *dxt_offsets,
                                uint64_t *sizes, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* GST_ENCODE_H_ */
