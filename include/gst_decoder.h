/*
 * gst_decoder.h -- C ABI of the MI355X (gfx950) GST decoder, libgst_amd.so.
 *
 * Drop-in for the decode path of GammaUNC/GST (codec/decoder.h).  Each entry point
 * below names the reference interface it replaces.  The HIP handles are passed as
 * opaque pointers: gst_stream is a hipStream_t, gst_event a hipEvent_t, and every
 * `d_` pointer is device memory.  All functions return GST_OK (0) or a negative
 * GST_ERR_* code; none aborts the process.
 *
 * Input formats are the reference's:
 *   .gtc file   = GenTCHeader (28 B) | 4 x u16[256] frequencies | 4 rANS streams
 *                 (codec/encoder.cpp:122-144)
 *   staging buf = [u32 out_off[4n]][u32 in_off[4n]] zero-padded to ceil(32n/512)*512 B,
 *                 then n x 2048 B frequencies, then the n x 4 streams back to back
 *                 (codec/decoder.cpp:109-209, demo/photos_sf.cpp:749-795); for n = 1
 *                 this is the file minus its header at offset 512
 *   output      = n x (w*h/2) bytes of DXT1 blocks in raster block order
 *                 (codec/assemble.cl:64-81), or n x (w*h*3) RGB8 bytes
 *                 (codec/assemble.cl:83-130)
 */
#ifndef GST_DECODER_H_
#define GST_DECODER_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GST_OK 0
#define GST_ERR_INVALID_ARGUMENT (-1)
#define GST_ERR_BAD_HEADER (-2)
#define GST_ERR_CORRUPT_STREAM (-3)
#define GST_ERR_OUT_OF_MEMORY (-4)
#define GST_ERR_HIP (-5)
#define GST_ERR_NOT_INITIALIZED (-6)
#define GST_ERR_BUSY (-7)
#define GST_ERR_TOO_LARGE (-8)

/* Device error bits raised by the kernels (gst_read_error_bits). */
#define GST_DEVERR_STREAM_OFFSETS 0x1u  /* group end offset malformed, or a renormalisation
                                           read left the group's byte range */
#define GST_DEVERR_FREQUENCIES 0x2u     /* a frequency table does not sum to 2048 */
#define GST_DEVERR_PALETTE_INDEX 0x4u   /* decoded palette index outside the palette */

typedef void *gst_stream; /* hipStream_t */
typedef void *gst_event;  /* hipEvent_t */

/* GenTCHeader (codec/codec_base.h:9-20): seven little-endian uint32, 28 bytes. */
typedef struct gst_header {
  uint32_t width;
  uint32_t height;
  uint32_t palette_bytes;
  uint32_t y_cmp_sz;
  uint32_t chroma_cmp_sz;
  uint32_t palette_sz;
  uint32_t indices_sz;
} gst_header;

/* ---- codec/decoder.h drop-in ------------------------------------------------- */

/* Replaces GenTC::InitializeDecoder (codec/decoder.h:15, codec/decoder.cpp:558-591).
 * Selects `device`, checks it is gfx950 and that every kernel's code object loads
 * with the work-group sizes it needs.  Returns GST_OK when the platform meets them. */
int gst_initialize_decoder(int device);

/* Replaces GenTCHeader::LoadFrom (codec/codec_base.cpp:18-24) and adds the size
 * rules the decoder relies on (README.md:67-71; codec/decoder.cpp:150): w and h
 * multiples of 128, w*h a multiple of 131072, palette_bytes a non-zero multiple
 * of 8192 no larger than w*h/4, stream sizes multiples of 4 that fit `len`. */
int gst_parse_header(const uint8_t *file, size_t len, gst_header *hdr);

/* Replaces GenTC::DecompressDXT (codec/decoder.h:16-17, codec/decoder.cpp:487-532):
 * host .gtc bytes in, host DXT1 blocks out (out_len >= w*h/2).  Blocking.  Returns
 * GST_ERR_CORRUPT_STREAM when the kernels raised an error bit. */
int gst_decompress_dxt(const uint8_t *file, size_t len, uint8_t *out, size_t out_len);

/* Replaces GenTC::LoadCompressedDXT (codec/decoder.h:19-21, codec/decoder.cpp:534-538).
 * d_cmp is a staging buffer (n = 1) in device memory, d_out receives w*h/2 bytes.
 * Work is enqueued on `stream` after the `num_init` events in `init`; `*done` (if
 * non-NULL) receives a new event the caller owns and must destroy. */
int gst_load_compressed_dxt(const gst_header *hdr, gst_stream stream, const void *d_cmp,
                            void *d_out, uint32_t num_init, const gst_event *init,
                            gst_event *done);

/* No counterpart in the reference: LoadCompressedDXT for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same single-image decode as gst_load_compressed_dxt (d_cmp in
 * the reference staging layout, d_out w*h/2 bytes), with no allocation of any kind: its
 * scratch is a static device array of GST_NOALLOC_SCRATCH_BYTES that each GPU the library
 * runs on holds until the process ends, enough for images of up to 2^20 blocks (e.g.
 * 4096x4096); larger images return GST_ERR_TOO_LARGE.  Device error bits (GST_DEVERR_*)
 * are OR-ed into *d_err_bits, a device word the caller owns and clears.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race and must be ordered by the caller.  No init events and no returned event:
 * record one on `stream` after the call. */
#define GST_NOALLOC_SCRATCH_BYTES (12u << 20)
int gst_load_compressed_dxt_noalloc(const gst_header *hdr, gst_stream stream, const void *d_cmp,
                                    void *d_out, uint32_t *d_err_bits);

/* Replaces GenTC::LoadCompressedDXTs (codec/decoder.h:23-25, codec/decoder.cpp:540-544):
 * n images in one staging buffer, outputs back to back.  Unlike the reference the
 * images may differ in size. */
int gst_load_compressed_dxts(const gst_header *hdrs, uint32_t n, gst_stream stream,
                             const void *d_cmp, void *d_out, uint32_t num_init,
                             const gst_event *init, gst_event *done);

/* Replace GenTC::LoadRGB / LoadRGBs (codec/decoder.h:27-33, codec/decoder.cpp:546-556):
 * the same pipeline with RGB8 output, w*h*3 bytes per image. */
int gst_load_rgb(const gst_header *hdr, gst_stream stream, const void *d_cmp, void *d_out,
                 uint32_t num_init, const gst_event *init, gst_event *done);
int gst_load_rgbs(const gst_header *hdrs, uint32_t n, gst_stream stream, const void *d_cmp,
                  void *d_out, uint32_t num_init, const gst_event *init, gst_event *done);

/* Replaces GenTC::RequiredScratchMem (codec/decoder.h:35, codec/decoder.cpp:41-47);
 * same formula, 4*2048*6 + 17*w*h/16 + palette_bytes. */
size_t gst_required_scratch_mem(const gst_header *hdr);

/* Replace GenTC::PreallocateDecompressor / FreeDecompressor (codec/decoder.h:36-37,
 * codec/decoder.cpp:478-485).  With a preallocation, decodes carve their scratch
 * from it in call order until FreeDecompressor (as the reference's bump allocator);
 * without one (or when it is used up) each decode takes stream-ordered scratch
 * that is released on its stream after the last kernel. */
int gst_preallocate_decompressor(size_t bytes);
int gst_free_decompressor(void);

/* Error bits (GST_DEVERR_*) raised on `stream` since the last call, read after the
 * stream's queued work completes (blocking).  Async entry points report corrupt
 * input only through this call. */
int gst_read_error_bits(gst_stream stream, uint32_t *bits);
/* Non-blocking form (no reference counterpart): enqueues on `stream` a copy of the
 * stream's error bits into `host_bits` (pinned host memory) and then their reset, so
 * each decode on the stream can be given its own bits; read `host_bits` once the
 * stream has passed that point. */
int gst_read_error_bits_async(gst_stream stream, uint32_t *host_bits);

/* ---- staging layout helpers (host) ------------------------------------------- */

/* Size of the staging buffer for n images and the byte offset of the stream section
 * (codec/decoder.cpp:109-111,172-209). */
int gst_staging_size(const gst_header *hdrs, uint32_t n, size_t *total, size_t *streams_offset);

/* Fills a host staging buffer from n .gtc files (demo/photos_sf.cpp:749-795,
 * demo/viewer.cpp:160-178 build the same layout). */
int gst_pack_staging(const uint8_t *const *files, const size_t *lens, uint32_t n,
                     uint8_t *out, size_t cap);

/* Upload of n ranges of pinned host memory (hipHostMalloc, 16-byte aligned) into device
 * memory (16-byte aligned) by a gather kernel on `stream`, one launch per 64 ranges: the
 * device reads the host memory over PCIe, about 50 GB/s whatever the range sizes, where a
 * DMA copy per range has a fixed cost of about 10 us (6 GB/s at 64 KiB).  The reference
 * uploads its pinned pages with one buffer copy each (demo/photos_sf.cpp:707-722).
 * Ranges of 0 bytes are skipped; GST_ERR_INVALID_ARGUMENT for a misaligned pointer or a
 * source the device cannot address.  Asynchronous: the host memory must stay unchanged
 * until `stream` has passed the copy. */
int gst_upload_gather(const void *const *host_src, void *const *dev_dst, const uint64_t *bytes, uint32_t n,
                      gst_stream stream);

/* ---- one launch per stage (per image; for parity tests and custom pipelines) ---
 * d_streams points at the image's stream section (Y stream first).  Scratch layout:
 * d_tables 4 x 2048 u32 plus 4 u32 status words; d_symbols 7*nv+P bytes in the
 * reference's ANS output layout [Y 2nv][chroma 4nv][palette P][indices nv]
 * (codec/decoder.cpp:441-461, ans/ans_decode.cl:71-72); d_group_sums nv/8192
 * int32; d_planes 6*nv int8 raster planes Y1,Y2,Co1,Cg1,Co2,Cg2; d_indices nv int32;
 * d_err one u32 of GST_DEVERR_* bits, which the stages only OR into: the caller
 * zeroes it before the first stage and reads it after the last. */

/* ans/build_table.cl:12-83 (codec/decoder.cpp:164-198): 4 tables from 4 x 256 u16. */
int gst_stage_build_table(const void *d_freqs, uint32_t *d_tables, uint32_t *d_err,
                          gst_stream stream);
/* ans/ans_decode.cl:97-143 (codec/decoder.cpp:214-236). */
int gst_stage_ans_decode(const gst_header *hdr, const void *d_streams, const uint32_t *d_tables,
                         uint8_t *d_symbols, int32_t *d_group_sums, uint32_t *d_err,
                         gst_stream stream);
/* codec/decode_indices.cl:6-84 (codec/decoder.cpp:302-393). */
int gst_stage_decode_indices(const gst_header *hdr, const uint8_t *d_symbols,
                             const int32_t *d_group_sums, int32_t *d_indices, gst_stream stream);
/* codec/inverse_wavelet.cl:69-192 (codec/decoder.cpp:242-300). */
int gst_stage_inv_wavelet(const gst_header *hdr, const uint8_t *d_symbols, int8_t *d_planes,
                          gst_stream stream);
/* codec/assemble.cl:64-81 (codec/decoder.cpp:395-417). */
int gst_stage_assemble_dxt(const gst_header *hdr, const uint8_t *d_symbols, const int8_t *d_planes,
                           const int32_t *d_indices, void *d_out, uint32_t *d_err,
                           gst_stream stream);
/* codec/assemble.cl:83-130. */
int gst_stage_assemble_rgb(const gst_header *hdr, const uint8_t *d_symbols, const int8_t *d_planes,
                           const int32_t *d_indices, void *d_out, uint32_t *d_err,
                           gst_stream stream);

/* ---- batched decode of textures of uneven size --------------------------------
 * No counterpart in the reference (its batched path needs equal sizes,
 * codec/decoder.cpp:105-121).  A batch is a plan over up to `capacity` images with
 * device-resident descriptors, tables and scratch.  Each image names its .gtc bytes
 * in device memory (the whole file, header included) and its output.  The five
 * stage kernels run as persistent grid-stride launches that read their work from
 * the plan, so one captured hipGraph replays any batch bound to the plan. */
typedef struct gst_batch gst_batch;

typedef struct gst_batch_config {
  uint32_t capacity_images;   /* max images per bind */
  uint64_t capacity_blocks;   /* max total w*h/16 over a bind */
  uint64_t capacity_palette;  /* max total palette_bytes over a bind */
  uint32_t ans_groups_per_wave;   /* 1 (32 streams per wave) or 2 (64); 0 = 2 */
  uint32_t ans_waves_per_group;   /* waves per work-group of the rANS kernel; 0 = 4 */
  uint32_t grid_ans;          /* persistent grid sizes; 0 = derived from the device */
  uint32_t grid_wavelet;
  uint32_t grid_indices;
  uint32_t grid_assemble;
  uint32_t rgb_output;        /* 1: RGB8 assembly instead of DXT1 */
  uint32_t wavelet_waves_per_group; /* waves per work-group of the wavelet kernel, each
                                       transforming 4 tiles at a time: 1, 2 or 4; 0 = 1 */
  uint32_t assembly_waves_per_group; /* waves per work-group of the DXT1 assembly with the
                                        index scan folded in: 1, 2 or 4; 0 = 4 */
  uint32_t separate_index_scan;     /* 1: index scan and assembly as two launches */
  uint32_t separate_table_stage;    /* 1: build_table as its own launch; 0: the rANS kernel
                                       builds each table in LDS (4 or 8 waves per group) */
  uint32_t assembly_wide;           /* fused assembly with 16 consecutive blocks per lane
                                       (one scan per 1024 blocks) instead of 4 x 4 */
  uint32_t separate_wavelet;        /* 1: the inverse wavelet as its own launch; 0: it runs
                                       inside the DXT1 assembly (planes stay in LDS) */
  uint32_t ans_form;                /* rANS kernel with 4 waves x 2 groups and tables built in
                                       LDS (the default shape): GST_ANS_FORM_* */
} gst_batch_config;

/* rANS work-group forms (gst_batch_config.ans_form, gst_tune_result.ans_form). */
#define GST_ANS_FORM_AUTO 0   /* chosen at each bind: work-groups of 8 groups of one stream
                                 while they fit the device in one pass at 6 per CU, else
                                 PAIRED (one captured graph replays both) */
#define GST_ANS_FORM_PAIRED 1 /* a stream's remainder of at most 4 groups shares a
                                 work-group with another stream's: fewer work-groups, at
                                 most 6 per CU */
#define GST_ANS_FORM_PLAIN 2  /* work-groups of 8 groups of one stream, 7 per CU */

int gst_batch_create(const gst_batch_config *cfg, gst_batch **out);
/* Binds n images: headers, device pointers to each .gtc file and each output.
 * Host-side validation; the descriptors are copied to the device on `stream`. */
int gst_batch_bind(gst_batch *b, const gst_header *hdrs, uint32_t n, const void *const *d_files,
                   void *const *d_outs, gst_stream stream);
/* Enqueues the five stages for the bound images. */
int gst_batch_decode(gst_batch *b, gst_stream stream);
/* Captures the five launches of the bound plan into a hipGraph (once); later
 * gst_batch_launch_graph replays it for whatever batch is bound. */
int gst_batch_capture_graph(gst_batch *b, gst_stream stream);
int gst_batch_launch_graph(gst_batch *b, gst_stream stream);
/* Per-stage kernel time (ms) of one decode, measured with hipEvents on `stream`:
 * build_table, ans_decode, decode_indices, inv_wavelet, assemble.  A stage that
 * launches nothing in the plan's shape (tables built inside the rANS kernel, an index
 * scan or wavelet folded into the assembly) reports NaN.  Blocking. */
int gst_batch_decode_timed(gst_batch *b, gst_stream stream, float ms[5]);
/* The same over `reps` (1..4096) decodes launched back to back, one host wait at the
 * end: mean per stage.  The events only time (no system-scope fence), so the kernels
 * run as they do without them, apart from the events' own packets. */
int gst_batch_decode_timed_reps(gst_batch *b, gst_stream stream, uint32_t reps, float ms[5]);
int gst_batch_error_bits(gst_batch *b, gst_stream stream, uint32_t *bits);
/* Non-blocking form: enqueues on `stream` a copy of the error bits into `host_bits`
 * (pinned host memory) followed by their reset; read it once `stream` has passed. */
int gst_batch_error_bits_async(gst_batch *b, gst_stream stream, uint32_t *host_bits);
/* Launch shapes chosen by gst_batch_autotune and the stage times they gave. */
typedef struct gst_tune_result {
  uint32_t ans_groups_per_wave;      /* 1: 32 streams per wavefront, 2: 64 */
  uint32_t ans_waves_per_group;
  uint32_t wavelet_waves_per_group;  /* 4 tiles per wave */
  uint32_t assembly_waves_per_group; /* 0: separate index scan and assembly */
  uint32_t ans_builds_tables;        /* 1: no build_table launch, tables built in LDS */
  uint32_t assembly_wide;            /* fused assembly: 1 = 16 consecutive blocks per lane */
  uint32_t grid_ans, grid_indices, grid_wavelet, grid_assemble;
  uint32_t candidates;               /* shapes timed */
  float stage_ms[5];                 /* median time of the chosen shape per stage; NaN for a
                                        stage the chosen shapes fold into another launch */
  float total_ms;
  uint32_t wavelet_fused;            /* 1: wavelet + index scan + DXT1 assembly in one kernel */
  uint32_t ans_form;                 /* GST_ANS_FORM_PAIRED, _PLAIN or _AUTO (8-group units in
                                        the kernel that also runs PAIRED) of the chosen rANS
                                        shape when it is 4 waves x 2 groups with tables built
                                        in LDS; else 0 */
} gst_tune_result;
/* No reference counterpart.  Times every launch shape of every stage on the bound
 * batch (median of `reps` launches each, 0 = 5) and keeps the fastest per stage:
 * rANS streams per wavefront {32, 64} x waves per work-group {1, 2, 4, 8} x grid,
 * wavelet waves per work-group {1, 2, 4} x grid, and the index scan + assembly either
 * fused (waves per work-group {1, 2, 4} x grid) or as two launches (grid each); then
 * the wavelet inside the DXT1 assembly (grid) replaces the last three stages when it
 * beats their sum.  The
 * bound images are decoded along the way.  Drops a captured graph
 * (capture again afterwards).  Blocking. */
int gst_batch_autotune(gst_batch *b, gst_stream stream, uint32_t reps, gst_tune_result *out);
/* Device bytes the plan holds (descriptors + tables + scratch). */
size_t gst_batch_device_bytes(const gst_batch *b);
int gst_batch_destroy(gst_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* GST_DECODER_H_ */
