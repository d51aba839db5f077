/*
 * gst_oracle.h -- CPU restatement of the GST decode path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP decoder in
 * gst_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in the product path links or calls it.
 *
 * Every function restates one piece of the reference decoder (GammaUNC/GST) and
 * cites the reference file:line it follows.  All arithmetic is integer and
 * follows the OpenCL kernels' types (uint32 rANS state, uint16 cumulative
 * frequencies, int32 wavelet intermediates, int8 plane store, uint16 RGB565 pack).
 *
 * How each function is pinned (see tests/test_oracle.py):
 *   gsto_build_table          closed form (exclusive scan + slot->symbol) and an
 *                             exhaustive check of the 11-step search over all 256
 *                             target symbols; reference KATs (ans/ans_ocl_test.cpp:64-155)
 *   gsto_ans_decode_group     reference CPU decoder ans::DecodeInterleaved
 *                             (ans/decode.cpp:240-292) on reference-encoded groups;
 *                             end-state invariants (state == k*M, pointer == group start)
 *   gsto_decode_indices       closed-form inclusive prefix sum; the reference's staged
 *                             scan restated in gsto_decode_indices_staged and compared
 *                             for every legal size up to 4096x4096
 *   gsto_inv_wavelet_tile     reference InverseWavelet2D (codec/wavelet.cpp:133-155) on
 *                             all tiles of reference-encoded planes; wavelet KATs
 *                             (codec/test/wavelet_test.cpp:129-158)
 *   gsto_assemble_dxt         exhaustive inverse check of rgb565->ycocg667 over all
 *                             65536 RGB565 values; end-to-end fixtures
 *   gsto_decode_dxt           reference end-to-end invariant
 *                             decode(CompressDXT(img)) == PhysicalBlocks()
 *                             (codec/test/codec_test.cpp:36-48) on reference-encoded files
 *   gsto_assemble_rgb         parity unpinned beyond the closed form restated here and
 *                             the DXT1 4-colour palette rule (no reference fixture exists)
 */
#ifndef GST_ORACLE_H_
#define GST_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSTO_TABLE_LOG 11
#define GSTO_TABLE_SIZE 2048u
#define GSTO_NUM_SYMBOLS 256u
#define GSTO_SYMS_PER_LANE 256u
#define GSTO_LANES 32u
#define GSTO_GROUP_SYMS 8192u
#define GSTO_STATE_L 32768u
#define GSTO_HEADER_BYTES 28u
#define GSTO_FREQ_BYTES 2048u

/* Error codes (negative). */
#define GSTO_OK 0
#define GSTO_ERR_ARG (-1)
#define GSTO_ERR_HEADER (-2)
#define GSTO_ERR_STREAM (-3)
#define GSTO_ERR_FREQ (-4)

/* GenTCHeader (codec/codec_base.h:9-20): seven little-endian uint32. */
typedef struct gsto_header {
  uint32_t width;
  uint32_t height;
  uint32_t palette_bytes;
  uint32_t y_cmp_sz;
  uint32_t chroma_cmp_sz;
  uint32_t palette_sz;
  uint32_t indices_sz;
} gsto_header;

/* AnsTableEntry (ans/build_table.cl:6-10), widened for host use. */
typedef struct gsto_table_entry {
  uint16_t freq;
  uint16_t cum_freq;
  uint8_t symbol;
} gsto_table_entry;

/* Header parse (codec/codec_base.cpp:18-24) plus the size rules of README.md:67-71
 * and the stream-size consistency the decoder relies on (codec/decoder.cpp:430-476). */
int gsto_parse_header(const uint8_t *file, size_t len, gsto_header *hdr);

/* Slot table from 256 uint16 frequencies (ans/build_table.cl:12-83). */
void gsto_build_table(const uint16_t freqs[256], gsto_table_entry table[2048]);

/* Decode one 32-stream rANS group (ans/ans_decode.cl:25-74).  `stream` is one
 * ANS stream (u32 end offsets then groups), `stream_len` its size in bytes and
 * `num_groups` its group count.  Writes 8192 symbols in the reference's lane-major
 * order: lane j's 256 symbols at out[j*256 .. j*256+255].  Returns GSTO_OK, or
 * GSTO_ERR_STREAM if the group's end offset is malformed or a renormalisation read
 * would leave the group's byte range [end_off[g-1] (or 4*num_groups), end_off[g]).
 * `final_state` (32 entries) and `final_word` may be NULL; they report the lane
 * states and the word pointer after the last step. */
int gsto_ans_decode_group(const gsto_table_entry table[2048], const uint8_t *stream,
                          size_t stream_len, uint32_t num_groups, uint32_t group,
                          uint8_t *out, uint32_t *final_state, uint32_t *final_word);

/* Inclusive prefix sum of (byte - 128) as int32 (codec/decode_indices.cl:6-84
 * composed by codec/decoder.cpp:304-393). */
void gsto_decode_indices(const uint8_t *diffs, uint32_t nv, int32_t *out);

/* The reference's multi-launch scan restated launch by launch
 * (codec/decoder.cpp:304-393, codec/decode_indices.cl:6-84).  `race_policy` picks
 * which of two work-items that store to out[nv-1] in the same launch lands last
 * (0: lower global id last, 1: higher global id last); the reference leaves it
 * unordered.  `garbage` seeds the local-memory words no work-item initialises. */
void gsto_decode_indices_staged(const uint8_t *diffs, uint32_t nv, int32_t *out,
                                int race_policy, int32_t garbage);

/* Inverse 5-level LeGall 5/3 on one 32x32 tile (codec/inverse_wavelet.cl:14-192).
 * `in` is 1024 bytes (raster within the tile, +128 bias); `out` receives 1024 int8
 * in tile raster order. */
void gsto_inv_wavelet_tile(const uint8_t in[1024], int8_t out[1024]);

/* One level (len = 2..32) of the same transform on a 32x32 int32 tile, in place:
 * rows y < len over x < len, then columns x < len over y < len. */
void gsto_inv_wavelet_level(int32_t tile[1024], int len);

/* One DXT1 block from the six planes (codec/assemble.cl:39-81). */
uint16_t gsto_ycocg_to_rgb565(int y, int co, int cg);

/* Whole-image DXT1 assembly (codec/assemble.cl:64-81).  planes: 6*nv int8 raster
 * planes in order Y1,Y2,Co1,Cg1,Co2,Cg2.  Palette indices outside
 * [0, palette_words) yield a zero index word and are counted in the return value
 * (the reference reads out of bounds there). */
uint32_t gsto_assemble_dxt(const int8_t *planes, const int32_t *indices,
                           const uint32_t *palette, uint32_t palette_words,
                           uint32_t width, uint32_t height, uint8_t *out);

/* RGB8 assembly (codec/assemble.cl:83-130). out: width*height*3 bytes. */
uint32_t gsto_assemble_rgb(const int8_t *planes, const int32_t *indices,
                           const uint32_t *palette, uint32_t palette_words,
                           uint32_t width, uint32_t height, uint8_t *out);

/* Stage outputs of a whole decode, for per-stage parity checks.  All buffers are
 * caller-owned and sized by the header: symbols 7*nv+P, planes 6*nv, indices nv. */
typedef struct gsto_stages {
  uint8_t *symbols;   /* ANS output, reference layout [Y 2nv][chroma 4nv][palette P][idx nv] */
  int8_t *planes;     /* inverse wavelet output, 6 raster planes */
  int32_t *indices;   /* decoded palette indices */
} gsto_stages;

/* Full decode of a .gtc file to DXT1 blocks (codec/decoder.cpp:98-428,487-532).
 * `out` must hold width*height/2 bytes.  `stages` may be NULL.  `nthreads` <= 1
 * decodes on the calling thread; larger values split every stage over that many
 * POSIX threads (one task per ANS group, wavelet tile and block row).  Returns
 * GSTO_OK, or a negative error, or a positive count of out-of-range palette
 * indices. */
int gsto_decode_dxt(const uint8_t *file, size_t len, uint8_t *out, gsto_stages *stages,
                    int nthreads);

/* Same pipeline with the RGB8 assembly (codec/decoder.cpp:546-556). */
int gsto_decode_rgb(const uint8_t *file, size_t len, uint8_t *out, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* GST_ORACLE_H_ */
