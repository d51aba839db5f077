/* oracle_sanitize_test.c -- the oracle (oracle/gst_oracle.c) decoding files under a
 * host sanitizer (AddressSanitizer + UndefinedBehaviorSanitizer, or ThreadSanitizer
 * for its thread pool), built by gst_amd/build.py; tests/test_sanitizers.py runs it.
 *
 * usage: oracle_sanitize_test <file.gtc> <expected.dxt> <expected.rgb or -> [threads]
 * Decodes the file to DXT1 blocks and RGB8 pixels on one thread and on `threads`
 * threads (default 4) and compares every result with the expected bytes.  A corrupt
 * copy of the file (garbled stream words) must decode without leaving its buffers. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gst_oracle.h"

static unsigned char *slurp(const char *path, size_t *len) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  unsigned char *p = (unsigned char *)malloc(n > 0 ? (size_t)n : 1);
  if (p && fread(p, 1, (size_t)n, f) != (size_t)n) {
    free(p);
    p = NULL;
  }
  fclose(f);
  *len = (size_t)n;
  return p;
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: oracle_sanitize_test <file.gtc> <expected.dxt> <expected.rgb|-> [threads]\n");
    return 2;
  }
  const int threads = argc > 4 ? atoi(argv[4]) : 4;
  size_t n_file = 0, n_dxt = 0, n_rgb = 0;
  unsigned char *file = slurp(argv[1], &n_file), *dxt = slurp(argv[2], &n_dxt);
  unsigned char *rgb = strcmp(argv[3], "-") ? slurp(argv[3], &n_rgb) : NULL;
  if (!file || !dxt) return 2;
  if (rgb && n_rgb != n_dxt * 6) {
    fprintf(stderr, "the RGB8 file is not 3 w h bytes\n");
    return 2;
  }
  int failures = 0;
  unsigned char *out = (unsigned char *)malloc(n_dxt);
  unsigned char *out_rgb = (unsigned char *)malloc(n_dxt * 6);  /* 3 w h bytes = 6 x the w h / 2 DXT1 bytes */
  const int counts[2] = {1, threads};
  for (int k = 0; k < 2; ++k) {
    memset(out, 0xA5, n_dxt);
    const int rc = gsto_decode_dxt(file, n_file, out, NULL, counts[k]);
    if (rc != GSTO_OK || memcmp(out, dxt, n_dxt)) {
      fprintf(stderr, "DXT1 decode on %d threads: rc %d, %s\n", counts[k], rc,
              rc == GSTO_OK ? "blocks differ" : "failed");
      ++failures;
    }
    if (rgb) {
      memset(out_rgb, 0xA5, n_rgb);
      const int rr = gsto_decode_rgb(file, n_file, out_rgb, counts[k]);
      if (rr != GSTO_OK || n_rgb != n_dxt * 6 || memcmp(out_rgb, rgb, n_rgb)) {
        fprintf(stderr, "RGB8 decode on %d threads: rc %d differs\n", counts[k], rr);
        ++failures;
      }
    }
  }
  /* Garbled words in the middle of the streams: any result but no stray access. */
  unsigned char *bad = (unsigned char *)malloc(n_file);
  memcpy(bad, file, n_file);
  for (size_t i = n_file / 2; i < n_file / 2 + 256 && i < n_file; ++i) bad[i] = (unsigned char)(i * 131u + 7u);
  (void)gsto_decode_dxt(bad, n_file, out, NULL, threads);
  /* A truncated file is refused. */
  if (gsto_decode_dxt(file, n_file / 2, out, NULL, threads) == GSTO_OK) {
    fprintf(stderr, "a truncated file decoded\n");
    ++failures;
  }
  free(bad);
  free(out);
  free(out_rgb);
  free(file);
  free(dxt);
  free(rgb);
  if (failures) return 1;
  printf("ok\n");
  return 0;
}
