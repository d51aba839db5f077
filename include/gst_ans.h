/* Stand-alone interleaved rANS decoder: the C ABI behind the counterpart of the
 * reference's ans::ocl::OpenCLDecoder (ans/ans_ocl.h:26-72, ans/ans_ocl.cpp).
 *
 * Parameters fixed by the reference (ans/ans.h:73-75, ans/ans_ocl_encode.cpp:10-18):
 * M = 2048 table slots, k = 2^4, b = 2^16 (16-bit renormalisation words, state
 * interval [2^15, 2^31)), 256 symbols per stream, alphabet of at most 256 symbols.
 * A decoder holds one table on the device; a decode uploads the words and states,
 * runs one wave per group of 1..32 interleaved streams and reads the symbols back,
 * blocking like the reference's methods.  C++ callers use include/gst/ans.hpp. */
#ifndef GST_ANS_H_
#define GST_ANS_H_

#include <stddef.h>
#include <stdint.h>

#include "gst_decoder.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GST_ANS_TABLE_SIZE 2048u       /* kANSTableSize (ans/ans.h:73) */
#define GST_ANS_SYMBOLS_PER_STREAM 256u /* kNumEncodedSymbols (ans/ans.h:74) */
#define GST_ANS_MAX_INTERLEAVED 32u     /* the reference kernel's 32-bit renormalisation
                                           mask (ans/ans_decode.cl:44-46) */

typedef struct gst_ans_decoder gst_ans_decoder;

/* ans::GenerateHistogram(counts, M) (ans/histogram.cpp:40-121): normalises n counts to
 * frequencies summing to M (zero counts stay zero, others at least 1).  The reference's
 * NormalizeFrequencies (ans/ans_ocl_encode.cpp:6-8) is this with M = 2048.
 * GST_ERR_INVALID_ARGUMENT for M <= 0 or all-zero counts. */
int gst_ans_normalize_frequencies(const uint32_t *counts, uint32_t n, int32_t M, uint32_t *out);

/* OpenCLDecoder::OpenCLDecoder(ctx, F, num_interleaved) (ans/ans_ocl.h:28-32,
 * ans/ans_ocl.cpp:35-48): normalises F (n <= 256 counts) and builds the table on the
 * device.  num_interleaved = streams per group for the grouped decode, 1..32. */
int gst_ans_decoder_create(const uint32_t *F, uint32_t n, uint32_t num_interleaved, gst_ans_decoder **out);

/* OpenCLDecoder::RebuildTable (ans/ans_ocl.h:47, ans/ans_ocl.cpp:112-148). */
int gst_ans_rebuild_table(gst_ans_decoder *d, const uint32_t *F, uint32_t n);

/* OpenCLDecoder::GetSymbols / GetFrequencies / GetCumulativeFrequencies
 * (ans/ans_ocl.h:49-51, ans/ans_ocl.cpp:54-101): the table's 2048 slots read back from
 * the device; any of the three outputs may be NULL. */
int gst_ans_get_table(const gst_ans_decoder *d, uint8_t *symbols, uint16_t *frequencies,
                      uint16_t *cumulative_frequencies);

/* The three OpenCLDecoder::Decode overloads (ans/ans_ocl.h:35-45,
 * ans/ans_ocl.cpp:155-425) in one call: n_data independent data streams, each holding
 * num_streams / n_data interleaved streams (1..32) encoded symbol by symbol in stream
 * order (16-bit words little-endian, in emission order); states[s] is the final
 * encoder state of stream s, streams of data stream i first.  out receives
 * num_streams x 256 bytes, stream s at s * 256.  Decode(state, data) is num_streams =
 * n_data = 1; Decode(states, data) is n_data = 1; Decode(states, vector data) is
 * num_streams = n_data x num_interleaved.  GST_ERR_CORRUPT_STREAM when a stream does
 * not decode back to the encoder's initial state (its symbols are then unspecified). */
int gst_ans_decode(const gst_ans_decoder *d, const uint32_t *states, uint32_t num_streams,
                   const uint8_t *const *data, const size_t *data_lens, uint32_t n_data, uint8_t *out);

/* OpenCLDecoder::~OpenCLDecoder (ans/ans_ocl.cpp:50-52). */
int gst_ans_decoder_destroy(gst_ans_decoder *d);

#ifdef __cplusplus
}
#endif

#endif /* GST_ANS_H_ */
