/*
 * gst_serve.h -- continuous-batching decode server (part of libgst_amd.so).
 *
 * No counterpart in the reference, whose batched caller is a demo that decodes
 * fixed pages of equal-size images (demo/photos_sf.cpp:688-830).  Requests (host
 * .gtc bytes) are admitted against an HBM budget: each admitted request holds its
 * input and output in a device arena until it is released.  A scheduler thread
 * forms a batch from every admitted request as soon as one of `slots` decode
 * slots is free and launches it (one gst_batch plan per slot, replayed from a
 * captured hipGraph, each slot on its own stream), so requests that arrive while
 * other batches run are admitted as memory frees up and their work overlaps on the
 * device.  Outputs stay device-resident; an optional copy back to pinned host
 * memory is timed separately.  `plain` mode decodes every request alone through
 * the per-image LoadCompressedDXT path on one stream, for comparison.
 */
#ifndef GST_SERVE_H_
#define GST_SERVE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gst_server gst_server;

typedef struct gst_server_config {
  uint64_t hbm_budget_bytes;    /* device arena for request inputs + outputs */
  uint64_t host_staging_bytes;  /* pinned host staging for submitted inputs (0: 1 GiB) */
  uint32_t slots;               /* concurrent batches in flight (0: 3) */
  uint32_t max_batch_images;    /* per launch (0: 256) */
  uint64_t max_batch_blocks;    /* per launch, sum of w*h/16 (0: 16 Mi) */
  uint32_t copy_to_host;        /* 1: copy each output to pinned host memory after decode */
  uint32_t auto_release;        /* 1: free a request's device memory once it completes */
  uint32_t plain;               /* 1: decode each request alone via LoadCompressedDXT */
  uint32_t dry_run;             /* 1: no GPU work; decodes complete at once (scheduler tests) */
  uint32_t ans_groups_per_wave; /* batch plan launch shape (0: defaults) */
  uint32_t ans_waves_per_group;
  uint32_t direct_launch;       /* 0: replay each slot's captured hipGraph (one call per
                                   batch for the scheduler thread); 1: launch the stage kernels
                                   on the slot's stream (lower p50 at light load, more host
                                   time per batch) */
} gst_server_config;

typedef struct gst_completion {
  uint64_t tag;
  uint64_t ticket;
  int32_t status;               /* GST_OK or a GST_ERR_* code (also device error bits) */
  uint32_t batch_images;        /* size of the batch it was decoded in */
  double latency_device_us;     /* arrival -> output ready in HBM */
  double latency_host_us;       /* arrival -> output in host memory (copy_to_host) */
  double queue_us;              /* arrival -> batch launch */
  void *d_output;               /* valid until gst_server_release (or completion with auto_release) */
  uint64_t output_bytes;
} gst_completion;

typedef struct gst_server_stats {
  uint64_t submitted, completed, failed;
  uint64_t batches;
  uint64_t texels;              /* decoded */
  uint64_t algorithmic_bytes;   /* sum of per-image pipeline bytes of the decoded requests */
  double device_busy_ms;        /* union of batch intervals on the GPU timeline */
  double max_batch_images;
  double mean_batch_images;
  uint64_t peak_arena_bytes;    /* high-water mark of admitted requests */
  uint64_t admission_waits;     /* times the head request waited for memory */
} gst_server_stats;

int gst_server_create(const gst_server_config *cfg, gst_server **out);
/* Copies `file` into pinned staging and queues it; returns at once.  Blocks only
 * while host staging is full.  GST_ERR_BAD_HEADER for an invalid file,
 * GST_ERR_OUT_OF_MEMORY if the request can never fit the budget or a batch. */
int gst_server_submit(gst_server *s, const uint8_t *file, size_t len, uint64_t tag, uint64_t *ticket);
/* Up to `max` completions since the last poll. */
int gst_server_poll(gst_server *s, gst_completion *out, uint32_t max, uint32_t *n);
/* Blocks until every submitted request has completed. */
int gst_server_drain(gst_server *s);
int gst_server_release(gst_server *s, uint64_t ticket);
int gst_server_stats_get(gst_server *s, gst_server_stats *st);
int gst_server_destroy(gst_server *s);

/* Load generator (needs a server created with auto_release = 1): submits files
 * from `pool` (file i of pool_size chosen per
 * arrival by a seeded draw over `pool_weights`, or uniformly when NULL) either as
 * a Poisson process of rate `qps` for `duration_s` seconds (closed_loop = 0) or
 * keeping `closed_loop` requests outstanding for `duration_s` seconds.  Fills
 * per-request device and host latencies (microseconds, completion order) into
 * `lat_dev` / `lat_host` (capacity `lat_cap`, count in *n_lat) and the wall time
 * of the run (first arrival to last completion) into *wall_s. */
int gst_server_run_load(gst_server *s, const uint8_t *const *pool, const size_t *pool_lens,
                        const double *pool_weights, uint32_t pool_size, double qps, double duration_s,
                        uint32_t closed_loop, uint64_t seed, double *lat_dev, double *lat_host,
                        uint64_t lat_cap, uint64_t *n_lat, double *wall_s);

#ifdef __cplusplus
}
#endif

#endif /* GST_SERVE_H_ */
