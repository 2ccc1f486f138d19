/* voromesh_hip.h -- C ABI of libvoromesh_hip.so, the MI355X (gfx950) VoroMesh
 * optimizer.
 *
 * Plain pointers and sizes only.  Every function returns 0 on success and a
 * nonzero vm_status otherwise; vm_last_error() gives the message of the last
 * failure on the calling thread.  Host arrays are read or written during the
 * call and never retained (as in the reference, which copies synchronously).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   vm_create ... vm_extract     voromesh::VoroOptimizer, include/voromesh/voro_optimizer.h:70-138
 *                                 (implementation src/voromesh/voro_optimizer.cu:451-860)
 *   vm_sampler_*                 poisson::Sampler, include/poisson/poisson_disk.h:21-62
 *                                 (implementation src/poisson/poisson_disk.cu:117-201)
 *   vm_batch_*, vm_serve         no counterpart in the reference: many fits in one set of
 *                                 launches, and a continuous-batching request loop.
 */
#ifndef VOROMESH_HIP_H
#define VOROMESH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define VM_API __attribute__((visibility("default")))
#else
#define VM_API
#endif

typedef enum vm_status {
    VM_OK            = 0,
    VM_EINVAL        = 1, /* invalid argument */
    VM_EHIP          = 2, /* HIP runtime error */
    VM_ENOMEM        = 3, /* allocation failed or over budget */
    VM_ESTATE        = 4, /* call not valid in the current state */
    VM_EUNSUPPORTED  = 5
} vm_status;

/* Layout-identical to voromesh::Vec2 / poisson::Sample / float2
 * (include/voromesh/shape2d.h:11-14, include/poisson/poisson_disk.h:16-19). */
typedef struct vm_vec2 {
    float x, y;
} vm_vec2;

/* voromesh::OptimizerParams (include/voromesh/voro_optimizer.h:34-44) plus the
 * knobs the reference fixes at compile time. */
typedef struct vm_params {
    int      iterations;        /* 400: schedule length the milestones are given for */
    int      local_iterations;  /* 100 */
    float    lr_per_spacing;    /* 0.09: Adam lr in units of the spacing */
    float    sample_fraction;   /* 0.2: targets used per step */
    float    gamma;             /* 0.5: MultiStepLR decay */
    int      milestones[4];     /* {80, 150, 200, 250} */
    unsigned seed;              /* 1: subsampling hash seed */
    int      loss_k;            /* 12: neighbours per target (reference kLossK, voro_optimizer.cu:54) */
    float    grid_cell;         /* hash-grid cell in world units; 0: 2 x spacing */
    int      stage_rings;       /* cell rings the loss kernel stages in LDS first (1 or 2; with 2, chunks whose
                                   candidate disc two rings would likely miss start at 3); 0: auto */
    int      precision;         /* 0: FP32 sites/gradients/Adam; 1: all FP64 */
} vm_params;

typedef struct vm_handle vm_handle;

/* Clipped diagram and boundary mesh (voromesh::VoroMesh2D,
 * include/voromesh/voro_optimizer.h:46-68).  Arrays are owned by the handle
 * and stay valid until the next vm_extract or vm_destroy. */
typedef struct vm_mesh {
    int            num_slots;
    const vm_vec2 *generators;  /* [num_slots] (x < 0: empty slot) */
    const uint8_t *inside;      /* [num_slots] */
    const int32_t *cell_start;  /* [num_slots + 1] CSR offsets into cell_verts */
    const vm_vec2 *cell_verts;  /* [cell_start[num_slots]] */
    const int32_t *cell_labels; /* [cell_start[num_slots]]: label of edge v[k] -> v[k+1]; >= 0 neighbour slot, -1..-4 box */
    int            num_vertices;
    const vm_vec2 *vertices;
    int            num_edges;
    const int32_t *edges;       /* [num_edges * 2] */
    int            watertight;
    int            loops;
    int            unsafe_cells;
    int            overflow_cells;
    float          gpu_ms;      /* grid build + clip kernels (the boundary kernel runs after them) */
    float          host_ms;     /* host work after the device results arrived (CSR in slot order) */
} vm_mesh;

VM_API void        vm_default_params(vm_params *p);
VM_API const char *vm_last_error(void);
VM_API const char *vm_version(void);
/* Number of HIP devices visible (0 without a GPU); never fails. */
VM_API int         vm_device_count(void);
/* Device-to-device copy of `bytes` (multiple of 4) between two fresh buffers: the
 * known-size reference for calibrating HBM traffic counters.  No counterpart in the
 * reference. */
VM_API int         vm_copy_probe(int64_t bytes, float *gpu_ms);

/* ---- voromesh::VoroOptimizer --------------------------------------------- */
/* VoroOptimizer(width, height, spacing, params)  voro_optimizer.h:75, .cu:512-532 */
VM_API int vm_create(float width, float height, float spacing, const vm_params *params, vm_handle **out);
/* ~VoroOptimizer  .cu:534-547 */
VM_API int vm_destroy(vm_handle *h);
/* SetTargets(slots)  voro_optimizer.h:82, .cu:549-560; slots with x < 0 or NaN are empty */
VM_API int vm_set_targets(vm_handle *h, const vm_vec2 *slots, int n);
/* UpdateTargets(dirty, slots)  voro_optimizer.h:84, .cu:562-587 */
VM_API int vm_update_targets(vm_handle *h, const int32_t *dirty, int num_dirty, const vm_vec2 *slots, int num_slots);
/* SetOccupancy(masks, tilesX, tilesY, voxelSize)  voro_optimizer.h:87, .cu:589-599 */
VM_API int vm_set_occupancy(vm_handle *h, const uint64_t *masks, int tiles_x, int tiles_y, float voxel_size);
/* UpdateOccupancy(masks, tx0, ty0, tx1, ty1)  voro_optimizer.h:89, .cu:601-611 */
VM_API int vm_update_occupancy(vm_handle *h, const uint64_t *masks, int tx0, int ty0, int tx1, int ty1);
/* SetGenerators(slots, slotsW, slotsH, slotCellSize)  voro_optimizer.h:94, .cu:613-638 */
VM_API int vm_set_generators(vm_handle *h, const vm_vec2 *slots, int slots_w, int slots_h, float slot_cell_size);
/* UpdateGenerators(slots, cx0, cy0, cx1, cy1)  voro_optimizer.h:96, .cu:640-650 */
VM_API int vm_update_generators(vm_handle *h, const vm_vec2 *slots, int cx0, int cy0, int cx1, int cy1);
/* BeginSchedule(iterations, cx0, cy0, cx1, cy1)  voro_optimizer.h:102-103, .cu:652-687 */
VM_API int vm_begin_schedule(vm_handle *h, int iterations, int cx0, int cy0, int cx1, int cy1);
/* Step(iterations) -> GPU ms  voro_optimizer.h:106, .cu:724-744 */
VM_API int vm_step(vm_handle *h, int iterations, float *gpu_ms);
/* Iteration(), ScheduleLength(), LastBatchLoss(), NumSlots()  voro_optimizer.h:111,118-120 */
VM_API int vm_get_state(vm_handle *h, int *iteration, int *schedule_length, float *last_batch_loss, int *num_slots);
/* EvaluateLoss()  voro_optimizer.h:113, .cu:746-761 (mean over all targets, FP64 sum) */
VM_API int vm_evaluate_loss(vm_handle *h, double *mean_loss);
/* Extract()  voro_optimizer.h:116, .cu:763-858 */
VM_API int vm_extract(vm_handle *h, vm_mesh *out);
/* Current generator slots (x < 0: empty).  n must be NumSlots(). */
VM_API int vm_get_generators(vm_handle *h, vm_vec2 *slots, int n);

/* ---- diagnostics (parity tests) --------------------------------------------- */
/* k nearest live generator slots of each query by the device ring search; d2 in the
 * optimizer's arithmetic.  Missing entries: idx -1. */
VM_API int vm_knn(vm_handle *h, const vm_vec2 *queries, int num_queries, int k, int32_t *idx, float *d2);
/* One loss pass over all live targets at the current generators, as the loss kernel
 * runs it (no subsampling, no target box), without touching the schedule.  Outputs
 * (each optional): summed gradient per slot [num_slots*2], loss sum and count, and per
 * target slot: the k-NN list [n_target_slots*loss_k], nearest slot, bisector partner slot
 * and the loss term. */
VM_API int vm_loss_pass(vm_handle *h, double *grad_xy, double *loss_sum, int64_t *count, int32_t *knn,
                        int32_t *nearest, int32_t *partner, double *loss_terms);

/* The optimizer's iteration graphs so far: captures an existing exec took in place (hipGraphExecUpdate: a new
 * layout with the same launches) and execs instantiated (first use, or launches of a new shape).  No
 * counterpart in the reference (it re-instantiates after every capacity change, voro_optimizer.cu:554,618). */
VM_API int vm_graph_stats(vm_handle *h, int64_t *updates, int64_t *instantiations);
/* FP32 gradient sums are fixed point: a term c is clamped to +-2^(51 - g), g the instance's shift
 * (53 - ceil(log2 T) - ceil(log2 max(W, H)), within [3, 47]), before it is scaled by 2^g and added.  This
 * returns how many target terms (since the optimizer was created) had a clamped component: 0 unless two
 * neighbouring sites nearly coincide.  FP64 sums are never clamped.  No counterpart in the reference (its FP32
 * float atomics have no clamp). */
VM_API int vm_grad_saturation(vm_handle *h, int64_t *targets);

/* ---- poisson::Sampler ------------------------------------------------------ */
typedef struct vm_sampler vm_sampler;
/* Sampler(width, height, radius, trials, seed)  poisson_disk.h:24, .cu:119-132 (runs ResampleAll) */
VM_API int vm_sampler_create(float width, float height, float radius, int trials, unsigned seed, vm_sampler **out);
VM_API int vm_sampler_destroy(vm_sampler *s);
/* Resample(cx0, cy0, cx1, cy1) -> GPU ms  poisson_disk.h:33, .cu:141-177 */
VM_API int vm_sampler_resample(vm_sampler *s, int cx0, int cy0, int cx1, int cy1, float *gpu_ms);
/* Verify(cx0, cy0, cx1, cy1)  poisson_disk.h:38, .cu:179-199; *ok = 1 if valid */
VM_API int vm_sampler_verify(vm_sampler *s, int cx0, int cy0, int cx1, int cy1, int *ok);
/* Replaces the sampler's grid (host mirror and device copy) with `cells`, grid_w * grid_h entries (x < 0:
 * empty); later Resample calls continue from it.  No counterpart in the reference (its Cells() is read-only):
 * restores a saved grid, and lets the tests plant conflicts that Verify must report. */
VM_API int vm_sampler_load_cells(vm_sampler *s, const vm_vec2 *cells, int n);
/* GridW(), GridH(), CellSize()  poisson_disk.h:44-50 */
VM_API int vm_sampler_dims(vm_sampler *s, int *grid_w, int *grid_h, float *cell_size);
/* Cells(): host mirror, grid_w * grid_h entries  poisson_disk.h:42 */
VM_API int vm_sampler_cells(vm_sampler *s, vm_vec2 *cells, int n);
/* Cells() restricted to [cx0, cx1) x [cy0, cy1) (clamped): only those entries of `cells` (grid_w * grid_h) are
 * written.  No counterpart in the reference (its Resample refreshes the host mirror itself): lets a caller's
 * mirror follow a Resample in work proportional to the region. */
VM_API int vm_sampler_cells_region(vm_sampler *s, int cx0, int cy0, int cx1, int cy1, vm_vec2 *cells, int n);

/* Many independent samplers in one set of launches (batched initialisation).
 * Instance i samples [0,w[i]]x[0,h[i]] with radius r[i] and seed seed[i], `trials`
 * rounds; its grid (dims from vm_poisson_grid_dims) is written to cells + offsets[i]. */
VM_API int vm_poisson_grid_dims(float width, float height, float radius, int *grid_w, int *grid_h, float *cell_size);
VM_API int vm_poisson_batch(int n, const float *w, const float *h, const float *r, const unsigned *seed, int trials,
                            const int64_t *offsets, vm_vec2 *cells, float *gpu_ms);

/* vm_poisson_batch without any allocation: the instance table, the work tables and the grids live in a fixed
 * 16 MiB device array owned by the library (one per device, set up on the first call made on that device, kept
 * until the process ends); the call makes no hipMalloc / hipFree and no host heap allocation.  Same arguments,
 * results and error behaviour as vm_poisson_batch, with two limits: at most 16,384 instances, and the grids plus
 * the tables within the array (about 2 M grid cells in all; offsets[i] + grid_w * grid_h bounds the grids).
 * Larger requests fail with VM_EINVAL before any GPU work.  Calls in one process are serialized; the array keeps
 * the last call's contents (every grid is cleared before sampling, so results never depend on it).  No
 * counterpart in the reference: for callers that cannot allocate (tight serving loops, harnesses that forbid
 * allocation). */
VM_API int vm_poisson_batch_noalloc(int n, const float *w, const float *h, const float *r, const unsigned *seed,
                                    int trials, const int64_t *offsets, vm_vec2 *cells, float *gpu_ms);

/* ---- voxel canvas: target slots and occupancy from shapes and disc edits -------
 * voromesh::Canvas (include/voromesh/canvas.h:27-89, src/voromesh/canvas.cpp of
 * the reference).  Host-only: no GPU call.  Voxel (i, j) covers [i, i+1) x
 * [j, j+1) voxels; tiles are 8x8 voxels, voxel (i, j) is bit (j % 8) * 8 + i % 8
 * of its row-major tile word.  Free target slots hold (-1, -1). */
typedef struct vm_canvas vm_canvas;
typedef struct vm_canvas_edit {
    int32_t tx0, ty0, tx1, ty1;      /* occupancy tiles [tx0,tx1) x [ty0,ty1) rewritten (Canvas::Edit, canvas.h:38-45) */
    int32_t removed, added, flipped; /* surface samples removed / added, occupancy voxels flipped */
    int32_t num_dirty;               /* target slots whose value changed: vm_canvas_dirty */
} vm_canvas_edit;
/* Canvas(width, height, voxelSize) (canvas.h:47) */
VM_API int vm_canvas_create(float width, float height, float voxel_size, vm_canvas **out);
VM_API int vm_canvas_destroy(vm_canvas *c);
/* RasterizeShape(polygon) (canvas.h:51): closed polygon, last vertex not repeated */
VM_API int vm_canvas_rasterize(vm_canvas *c, const vm_vec2 *polygon, int n);
/* StampDisc(center, radius, inside) (canvas.h:57) */
VM_API int vm_canvas_stamp_disc(vm_canvas *c, vm_vec2 center, float radius, int inside, vm_canvas_edit *edit);
/* Edit::dirtySlots of the last stamp, in the reference's order (n >= num_dirty) */
VM_API int vm_canvas_dirty(const vm_canvas *c, int32_t *out, int n);
/* TilesX/TilesY/VoxelSize (canvas.h:67-69) and the voxel counts */
VM_API int vm_canvas_dims(const vm_canvas *c, int *voxels_x, int *voxels_y, int *tiles_x, int *tiles_y,
                          float *voxel_size);
/* TargetSlots().size(), SurfaceCount(), SurfaceTileCount() (canvas.h:63,70-71) */
VM_API int vm_canvas_counts(const vm_canvas *c, int64_t *num_slots, int64_t *surface_count, int64_t *surface_tiles);
/* TargetSlots() (canvas.h:63): n >= num_slots */
VM_API int vm_canvas_slots(const vm_canvas *c, vm_vec2 *out, int64_t n);
/* InsideMasks() (canvas.h:64): n >= tiles_x * tiles_y */
VM_API int vm_canvas_masks(const vm_canvas *c, uint64_t *out, int64_t n);
/* IsInside / IsSurface (canvas.h:59-60) */
VM_API int vm_canvas_voxel(const vm_canvas *c, int i, int j, int *inside, int *surface);

/* ---- batched fits: one set of launches per iteration for many instances ------ */
typedef struct vm_fit_desc {
    float          width, height;   /* domain [0,width] x [0,height] */
    const vm_vec2 *sites;           /* live sites (dense) */
    int            num_sites;
    const vm_vec2 *targets;         /* live targets */
    int            num_targets;
    float          lr;              /* Adam learning rate */
    float          gamma;
    int            milestones[4];   /* already scaled to this schedule */
    float          sample_fraction;
    unsigned       seed;
    int            loss_k;          /* 0: tuner choice for the size class */
    float          grid_cell;       /* 0: tuner choice */
    int            stage_rings;     /* 0: tuner choice */
    int            max_iterations;  /* 0: no limit */
    float          tol;             /* stop when |dL/L| < tol between steps (0: off) */
    float          chunk_extent;    /* target chunk diagonal in mean site spacings (0: tuner choice) */
    int            targets_per_lane; /* loss kernel targets per lane, 2-4 (0: tuner choice; FP64 batches use 3) */
} vm_fit_desc;

typedef struct vm_batch vm_batch;
VM_API int vm_batch_create(int precision, vm_batch **out);
VM_API int vm_batch_destroy(vm_batch *b);
/* Adds an instance; returns its id.  Valid until vm_batch_commit. */
VM_API int vm_batch_add(vm_batch *b, const vm_fit_desc *desc, int *id);
/* Lays out, uploads and bins every instance and captures one iteration as a graph. */
VM_API int vm_batch_commit(vm_batch *b);
/* Runs `iterations` optimizer iterations of every instance (graph replays). */
VM_API int vm_batch_step(vm_batch *b, int iterations, float *gpu_ms);
VM_API int vm_batch_get_sites(vm_batch *b, int id, vm_vec2 *out, int n);
/* Per instance: completed iterations, done flag and the mean loss of its last step. */
VM_API int vm_batch_status(vm_batch *b, int id, int *iteration, int *done, double *last_mean_loss);
/* Voronoi cells of fit `id` clipped to its box (FP64, as vm_extract) as a CSR mesh in the order of the fit's live
 * sites (cell_labels >= 0 are site indices), and with occupancy masks (8x8-voxel tile words, as vm_set_occupancy;
 * null: every cell outside) the watertight boundary of the inside cells.  The arrays stay valid until the next
 * vm_batch_extract on `b`.  No reference counterpart: the batched form of Extract (voro_optimizer.h:116). */
VM_API int vm_batch_extract(vm_batch *b, int id, const uint64_t *masks, int tiles_x, int tiles_y, float voxel_size,
                            vm_mesh *out);
/* Resolved per-instance choices: k, grid cell, staged rings, target chunk extent (mean spacings), targets per
 * lane of the loss kernel. */
VM_API int vm_batch_choice(vm_batch *b, int id, int *loss_k, float *grid_cell, int *stage_rings,
                           float *chunk_extent, int *targets_per_lane);
/* One loss pass over every instance at its current sites, as vm_loss_pass does for one optimizer (every live
 * target, no subsampling): outputs (each optional) concatenated instance by instance in the order they were added:
 * gradient sums [2 * sum of live sites], the loss sum and target count, and per live target (in input order) the
 * nearest site, bisector partner (instance-local site indices, -1 if unused) and loss term.  No counterpart in
 * the reference (diagnostics of batches: the FP32-vs-FP64 gradient gate). */
VM_API int vm_batch_loss_pass(vm_batch *b, double *grad_xy, double *loss_sum, int64_t *count, int32_t *nearest,
                              int32_t *partner, double *loss_terms);
/* One evaluation pass of the loss kernel over every instance, counting its wavefronts by
 * path: [0] bounding-box fast path, [1] ring search over the staged block, [2] ring
 * search in global memory (staging overflow), [3] no active target; [4] candidates
 * compacted by the fast path, [5] sites staged; [6] fast-path wavefronts in which at
 * least one target needed the exact k-NN selection after screening, [7] the number of
 * such targets. */
VM_API int vm_batch_loss_stats(vm_batch *b, uint64_t *counters);
/* Runs `iterations` optimizer iterations with HIP events around each phase on the batch's
 * stream and returns the mean ms per phase in phase_ms[5]: 0 scan, 1 scatter, 2 loss,
 * 3 exact k-NN selection for the targets the loss kernel could not settle, 4 adam. */
VM_API int vm_batch_profile(vm_batch *b, int iterations, float *phase_ms);
/* vm_grad_saturation summed over the batch's fits. */
VM_API int vm_batch_saturation(vm_batch *b, int64_t *targets);

/* ---- continuous-batching request loop ---------------------------------------- */
typedef struct vm_serve_config {
    double   hbm_budget_bytes;   /* device bytes available to resident fit state */
    int      max_resident;       /* instance slots (0: unlimited) */
    int      max_iterations;     /* per request (e.g. 300) */
    float    tol;                /* per request relative loss-change stop */
    int      poll_every;         /* iterations per launch between state readbacks */
    int      precision;
    double   max_resident_sites; /* admission stops while resident sites would pass this (0: no cap;
                                    a request is always admitted onto an empty device) */
    double   initial_pool_bytes; /* device pools at start (0: 256 MiB); they grow on demand */
} vm_serve_config;

typedef struct vm_serve_result {
    vm_vec2 *final_sites;     /* in: NULL or a buffer of num_sites entries for the fitted sites */
    double   latency_s;       /* completion - arrival */
    double   queue_s;         /* admission - arrival */
    int      iterations;
    double   final_mean_loss; /* mean loss of the last step */
} vm_serve_result;

typedef struct vm_serve_stats {
    double wall_s;
    double site_iters;        /* sum over completed requests of sites x iterations */
    double site_iters_run;    /* site_iters plus the iterations run so far by requests still resident */
    int    completed;
    int    max_resident_seen;
    double peak_bytes;        /* largest resident charge against the budget */
    int    arrived;           /* requests that arrived before the loop ended */
    int    admitted;
    int    pool_grows;        /* times a device pool was enlarged */
    double pool_bytes;        /* device pool capacity at the end */
    double gpu_busy_s;        /* device time of the iteration launches */
    double oldest_wait_s;     /* at the end: age of the oldest request not yet admitted (0: none) */
    double prep_busy_s;       /* host time of the preparation workers (sum over workers) */
    double sample_busy_s;     /* host time of the site-sampling calls (synthetic stream) */
    double loop_busy_s;       /* host time of the loop's retire / admit / upload work */
    double prep_wait_s;       /* part of prep_busy_s spent waiting for sampled sites (synthetic stream) */
    double upload_bytes;      /* bytes sent to the device for admissions and work tables */
    int    prep_workers;
} vm_serve_stats;

/* Serves requests[i] arriving at arrival_s[i] seconds after the call starts
 * (non-decreasing).  Continuous batching: a request is prepared once it has
 * arrived, admitted in arrival order as soon as an instance slot, room within
 * hbm_budget_bytes and room under max_resident_sites are free, advanced
 * together with every other resident request by one set of launches per
 * iteration, and retired when its relative loss change drops below tol or it
 * reaches max_iterations.  All requests share loss_k.  Returns when every
 * request has completed. */
VM_API int vm_serve(const vm_fit_desc *requests, const double *arrival_s, int n, const vm_serve_config *cfg,
                    vm_serve_result *results, vm_serve_stats *stats);

/* A synthetic request stream, generated lazily inside the library: Poisson
 * arrivals at `rate` per second during [0, duration_s); request i has a site
 * count drawn log-uniform in [min_sites, max_sites], Poisson-disk sites in the
 * unit square for it (the batched dart thrower, 30 rounds), and
 * targets_per_site targets per live site on a random simple polygon (5-40
 * vertices, sorted uniform angles, radii U[0.15, 0.45] around (0.5, 0.5)),
 * jittered by N(0, sigma^2).  Every draw comes from a counter-based generator
 * keyed by (seed, i), so a request is reproducible on its own
 * (vm_serve_synth_plan, vm_serve_synth_targets). */
typedef struct vm_serve_stream {
    double   rate;
    double   duration_s;
    uint64_t seed;
    int      min_sites, max_sites;
    int      targets_per_site;
    float    sigma;
    float    lr;               /* Adam learning rate (milestones 80/150/200/250, gamma 0.5) */
    int      loss_k;           /* shared by every request (0: 16) */
    int      drain;            /* 1: serve every arrival to completion; 0: stop at duration_s */
} vm_serve_stream;

typedef struct vm_serve_record {
    double arrival_s;
    double admit_s;           /* -1: not admitted when the loop ended */
    double done_s;            /* -1: not completed */
    int    num_sites, num_targets;
    int    iterations;
    double final_mean_loss;
} vm_serve_record;

/* Runs the loop of vm_serve over a synthetic stream.  A request is generated
 * (sites sampled, targets drawn, laid out) only once it has arrived and is
 * among the next requests to be admitted, so a backlog costs only its arrival
 * times.  records[i] describes arrival i; *num_records is the number of
 * arrivals.  Fails with VM_EINVAL if the stream brings more than max_records
 * arrivals.  No counterpart in the reference (the serving workload). */
VM_API int vm_serve_synthetic(const vm_serve_stream *stream, const vm_serve_config *cfg, vm_serve_record *records,
                              int max_records, int *num_records, vm_serve_stats *stats);
/* Request `index` of a stream: arrival time, drawn site count, Poisson-disk radius and sampler seed (its sites are
 * the live cells of vm_poisson_batch for that radius and seed in the unit square). */
VM_API int vm_serve_synth_plan(const vm_serve_stream *stream, int64_t index, double *arrival_s, int *num_sites,
                               float *radius, unsigned *sampler_seed);
/* Request `index`'s targets for its num_sites live sites (targets_per_site * num_sites points). */
VM_API int vm_serve_synth_targets(const vm_serve_stream *stream, int64_t index, int num_sites, vm_vec2 *targets,
                                  int capacity);

#ifdef __cplusplus
}
#endif

#endif /* VOROMESH_HIP_H */
