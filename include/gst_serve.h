/*
 * gst_serve.h -- continuous-batching decode server (part of libgst_amd.so).
 *
 * No counterpart in the reference, whose batched caller is a demo that decodes
 * fixed pages of equal-size images (demo/photos_sf.cpp:688-830).  Requests are
 * admitted against an HBM budget: each admitted request holds its input (unless it
 * is already device-resident) and its output in a device arena until it is released.
 * Inputs submitted from the host go to pinned staging at submit.  A scheduler thread
 * forms a batch from every admitted request as soon as one of `slots` decode slots
 * is free and launches it (one gst_batch plan per slot, each slot on its own stream):
 * inputs admitted in that same round are copied to the arena on the slot's stream right
 * before their decode, and inputs that find no free slot go ahead on a copy stream of
 * the scheduler's (contiguous requests in one copy), overlapping the decodes in flight,
 * with the slot waiting for that copy when their batch launches;
 * so requests that arrive while other batches run are admitted as memory frees up and
 * their work overlaps on the device.  Outputs stay device-resident; an optional copy
 * back to pinned host memory is timed separately, and an optional expected output per
 * request is compared on the device after the decode.  `plain` mode decodes every
 * request alone through the per-image LoadCompressedDXT path on one stream, for
 * comparison.
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
  uint32_t slots;               /* concurrent batches in flight (0: 4; plain mode 1) */
  uint32_t max_batch_images;    /* per launch (0: 256) */
  uint64_t max_batch_blocks;    /* per launch, sum of w*h/16 (0: 16 Mi) */
  uint32_t copy_to_host;        /* 1: copy each output to pinned host memory after decode */
  uint32_t auto_release;        /* 1: free a request's device memory once it completes */
  uint32_t plain;               /* 1: decode each request alone via LoadCompressedDXT */
  uint32_t dry_run;             /* 1: no GPU work; decodes complete at once (scheduler tests) */
  uint32_t ans_groups_per_wave; /* batch plan launch shape (0: defaults) */
  uint32_t ans_waves_per_group;
  uint32_t graph_launch;        /* 1: replay each slot's captured hipGraph; 0: launch the
                                   stage kernels on the slot's stream (on MI355X both the
                                   scheduler's host time per batch and the decode are shorter
                                   this way: 34 against 53 us and about 3 us) */
  uint32_t poison_outputs;      /* 1: fill each request's output with 0xA5 before its decode, so
                                   a decode that writes nothing cannot pass a check */
  uint32_t fail_launch_every;   /* testing, dry_run only: every n-th batch launch fails (0: never) */
  uint32_t idle_spin_us;        /* an idle scheduler keeps polling for this long after its last
                                   work before it sleeps on the arrival condition (0: sleeps at
                                   once): trades a host core for the wake-up latency */
  uint32_t copy_kernel;         /* 1: the inputs admitted together reach the arena through one
                                   gather kernel reading the pinned staging over PCIe, instead of
                                   a DMA copy per run of adjacent requests */
} gst_server_config;

/* One request.  Either `file` holds the whole .gtc on the host (copied to pinned
 * staging at submit) and `d_file` is NULL, or `d_file` holds it in device memory
 * (pre-staged: no copy; the caller keeps it unchanged until completion) and `file`
 * holds at least its 28-byte header on the host. */
typedef struct gst_request {
  const uint8_t *file;
  const void *d_file;
  size_t len;                   /* file length */
  uint64_t tag;
  const void *d_expected;       /* optional: the expected output bytes in device memory (16-byte
                                   aligned); compared on the device after the decode, result in
                                   gst_completion.verified */
} gst_request;

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
  uint32_t verified;            /* 0: no expected output given; 1: output equal to it; 2: differs */
  uint32_t reserved;
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
  uint64_t verified;            /* completions compared with an expected output */
  uint64_t mismatches;          /* of those, outputs that differed */
  uint64_t h2d_bytes;           /* input bytes copied host -> device */
  uint64_t h2d_copies;          /* copies they took (contiguous requests share one) */
  uint64_t prestaged;           /* requests whose input was already device-resident */
  double launch_host_ms;        /* scheduler time spent enqueuing batches */
} gst_server_stats;

int gst_server_create(const gst_server_config *cfg, gst_server **out);
/* Copies `file` into pinned staging and queues it; returns at once.  Blocks only
 * while host staging is full.  GST_ERR_BAD_HEADER for an invalid file,
 * GST_ERR_OUT_OF_MEMORY if the request can never fit the budget or a batch.  Any number
 * of threads may submit at once (the copies run in parallel); a submit still running
 * when gst_server_destroy starts returns GST_ERR_BUSY, and none may start after. */
int gst_server_submit(gst_server *s, const uint8_t *file, size_t len, uint64_t tag, uint64_t *ticket);
/* General form: a host or pre-staged device input, with an optional expected output.
 * GST_ERR_INVALID_ARGUMENT for a misaligned d_expected. */
int gst_server_submit_request(gst_server *s, const gst_request *req, uint64_t *ticket);
/* Up to `max` completions since the last poll. */
int gst_server_poll(gst_server *s, gst_completion *out, uint32_t max, uint32_t *n);
/* Blocks until every submitted request has completed. */
int gst_server_drain(gst_server *s);
int gst_server_release(gst_server *s, uint64_t ticket);
int gst_server_stats_get(gst_server *s, gst_server_stats *st);
int gst_server_destroy(gst_server *s);

/* Load generator (needs a server created with auto_release = 1). */
typedef struct gst_load_spec {
  const uint8_t *const *pool;       /* host .gtc bytes of each pool file */
  const size_t *pool_lens;
  const double *pool_weights;       /* per-file draw weights, NULL: uniform */
  uint32_t pool_size;
  const void *const *pool_device;   /* NULL, or every pool file already in device memory (pre-staged:
                                       requests carry no copy) */
  const void *const *pool_expected; /* NULL, or every pool file's expected output in device memory:
                                       every completion is compared on the device */
  double qps;                       /* Poisson arrival rate when closed_loop = 0 */
  double duration_s;
  uint32_t closed_loop;             /* > 0: keep this many requests outstanding instead */
  uint64_t seed;
  uint32_t submit_threads;          /* closed loop: threads that submit, each copying its requests'
                                       inputs into the staging in parallel (0 or 1: the calling
                                       thread alone; one thread's memory copy is about 37 GB/s) */
} gst_load_spec;

typedef struct gst_load_result {
  uint64_t requests;                /* completed during the run */
  uint64_t failed;                  /* completions with a status other than GST_OK */
  uint64_t verified, mismatches;    /* completions compared / of those, differing */
  double wall_s;                    /* first arrival to last completion */
} gst_load_result;

/* Submits files from the pool (file i chosen per arrival by a seeded draw over the
 * weights) either as a Poisson process of rate `qps` for `duration_s` seconds or
 * keeping `closed_loop` requests outstanding for `duration_s` seconds, then drains.
 * Fills per-request device and host latencies (microseconds, completion order) into
 * `lat_dev` / `lat_host` (capacity `lat_cap`; either may be NULL). */
int gst_server_run_load(gst_server *s, const gst_load_spec *spec, double *lat_dev, double *lat_host,
                        uint64_t lat_cap, gst_load_result *res);

#ifdef __cplusplus
}
#endif

#endif /* GST_SERVE_H_ */
