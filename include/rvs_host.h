/*
 * rvs_host.h -- C ABI of the host-side library (librvs_host.so): synthetic scene
 * rendering and the reference-format front end (JSON configuration, camera
 * parameters, pose traces, raw YUV files) used by the drop-in Pipeline.
 */
#ifndef RVS_HOST_H
#define RVS_HOST_H

#include <stdint.h>
#include "rvs_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Render a seeded synthetic scene (kind 0: textured rectangles before a plane at
 * 10 m; kind 1: textured spheres inside a box) for one camera.  Colour planes
 * y (w*h), u, v ((w/2)*(h/2)) and the depth plane (w*h) are u8 when the bit depth
 * is <= 8, else u16; depth is quantised inverse depth in [znear, zfar].
 * threads <= 0 uses every hardware thread.  Returns 0 on success. */
int rvs_scene_render(int kind, uint64_t seed, const rvs_cam* cam, const float pos[3], const float rot[3],
                     int bitdepth_color, int bitdepth_depth, float znear, float zfar, void* y, void* u, void* v,
                     void* depth, int threads);

/* Message of the last failing call of this library on the calling thread. */
const char* rvs_host_last_error(void);

/* Parse `text` with the configuration JSON reader, walk `path` ("a/0/b"; numeric
 * segments index arrays, "?key" uses optional()), convert with `as` in
 * {double, float, int, string, bool, size, type, truthy}; overrides_text (may be
 * empty) is applied to every object reached.  Returns 0 and the value in `out`,
 * or 1 and the error text. */
int rvs_json_query(const char* text, const char* overrides_text, const char* path, const char* as, char* out,
                   int cap);
/* Parse a pose trace (PoseTraces.cpp:54-92): writes x,y,z,yaw,pitch,roll per row
 * (up to max_rows) and returns the row count, or -1. */
int rvs_pose_trace_parse(const char* text, float* out, int max_rows);
/* Load a view-synthesis configuration (and the files it names) and describe it
 * as JSON in `out`.  Returns 0, or 1 with the error text. */
int rvs_config_describe(const char* path, char* out, int cap);
/* Run the configuration-driven Application (Pipeline::execute): reads the input
 * YUV files, writes the configured outputs.  stagewise != 0 runs the
 * view-by-view plugin path instead of the fused one.  Returns the number of
 * frames or -1 (rvs_host_last_error). */
int rvs_app_run(const char* config_path, const char* source_path, int stagewise);

/* ------------------------------------------------------------------ serving
 * A request loop over resident source sets ("scenes"): target-view requests are
 * admitted by memory budget and batched as they arrive (continuous batching).
 * Each batch is one rvs_synthesize_batch_async launch set on one of
 * `compute_streams` streams shared by the `slots` in flight (in order on one
 * stream; overlapping when there are several), and its YUV planes are copied
 * back on the slot's own stream, overlapping later batches (with
 * `split_finish`, the batch's inpainting and encoding run there too).  Latency of a
 * request = arrival to the completion of its YUV copy to host memory. */
typedef struct {
    size_t memory_budget;       /* device bytes for in-flight batches (workspaces + outputs), one arena shared by the slots */
    int max_batch;              /* views per batch */
    int slots;                  /* batches in flight */
    uint32_t entry_cap;         /* binned entries per view of a batch (0 = library default); grows on overflow */
    int copy_to_host;           /* copy each request's YUV planes to pinned host memory */
    int keep_outputs;           /* keep a host copy per request (rvs_server_output) */
    size_t host_staging_bytes;  /* pinned staging per slot (0 = budget / (4 * slots), at most 1 GiB) */
    int compute_streams;        /* streams the slots' launch sets run on (0 or 1: one, in order) */
    int split_finish;           /* 1: inpainting + encoding of a batch on its slot's stream, overlapping the
                                   next batch's binning and raster (rvs_synthesize_batch_async_split) */
} rvs_serve_opts;

typedef struct {
    int scene;                  /* index of the resident source set */
    rvs_target target;
} rvs_request;

typedef struct {
    double arrival_ms, launch_ms, done_ms;   /* since the start of the run */
    int batch;                  /* batch id */
    int batch_views;            /* views in that batch */
    int rc;                     /* RVS_OK or an error code */
} rvs_request_stats;

/* scenes[k] points to n_src[k] resident sources (device pointers stay owned by
 * the caller and must outlive the server).  Returns NULL on failure. */
void* rvs_server_create(const rvs_source* const* scenes, const int* n_src, int n_scenes, const rvs_serve_opts* opts);
void rvs_server_destroy(void* server);
/* Open loop: request k arrives at the k-th event of a seeded Poisson process of
 * `rate` per second.  Closed loop: `concurrency` requests outstanding.  Both
 * block until every request is done, fill stats[n] and return 0, or the last
 * error code of a failed request (its stats say which). */
int rvs_server_run_open(void* server, const rvs_request* reqs, int n, double rate, uint64_t seed,
                        rvs_request_stats* stats);
int rvs_server_run_closed(void* server, const rvs_request* reqs, int n, int concurrency, rvs_request_stats* stats);
/* Host copy of request k's YUV planes (Y, U, V concatenated) from the last run
 * (keep_outputs); returns the byte count, or -1. */
long long rvs_server_output(void* server, int k, void* out, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
