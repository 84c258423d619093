/*
 * onco.h — C ABI of liboncohip.so, the MI355X (gfx950) per-timestep hot path
 * of the Oncolytica agent-based tumour simulator.
 *
 * Every entry point returns an int status: ONC_OK (0) or a negative ONC_ERR_*
 * code; onc_last_error() returns a thread-local message for the last failure.
 * No torch or C++ types cross this boundary: plain pointers, sizes and PODs.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   - the backend plugin protocol ISimulationBackend
 *     (framework/src/oncolytica/core/runtime/_backend.py:6-36) is implemented in
 *     Python by oncolytica_amd.hip.HipBackend on top of the onc_sim_* calls;
 *   - the fixed WGSL template kernels (framework/src/oncolytica/gpu/compiler/
 *     template.wgsl:152-343) and their dispatch (gpu/runtime/_gpu_backend.py:
 *     113-171, 420-432) are replaced by onc_sim_step / onc_batch_step;
 *   - WGSL shader-module creation (gpu/runtime/_device_manager.py:81) is
 *     replaced by onc_module_build (hiprtc);
 *   - buffer upload/readback (gpu/runtime/_gpu_backend.py:353-405) is replaced
 *     by onc_sim_upload_* / onc_sim_download_* (SoA columns, no per-object pack).
 *
 * Threading: calls on one handle are not thread-safe; the caller serialises.
 * Steps are asynchronous on the handle's HIP stream; downloads synchronise.
 */
#ifndef ONCO_H
#define ONCO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ONC_ABI_VERSION 1

#define ONC_MAX_FIELDS 4
#define ONC_MAX_COLS 48
#define ONC_MAX_PARAMS 32
#define ONC_MAX_CHEM_COLS 16

/* status codes */
#define ONC_OK 0
#define ONC_ERR_INVALID (-1)  /* bad argument or handle */
#define ONC_ERR_HIP (-2)      /* a HIP runtime call failed */
#define ONC_ERR_NOMEM (-3)    /* device allocation failed or HBM budget exceeded */
#define ONC_ERR_COMPILE (-4)  /* hiprtc compilation or module load failed */
#define ONC_ERR_CAPACITY (-5) /* more cells than the sim's capacity */
#define ONC_ERR_STATE (-6)    /* call not valid in the handle's current state */
#define ONC_ERR_DEVICE (-7)   /* a kernel reported an internal error (bounded wait expired) */

/* Morton key formats (framework/src/oncolytica/core/runtime/_geometry.py:14-47).
 * ONC_KEY_MORTON30: 10 bits per axis, x -> bit 3i, y -> 3i+1, z -> 3i+2; equals
 *   the reference's (interleave_9bit(c/2) << 3) | sub-voxel closed form.
 * ONC_KEY_MORTON32: same for bits 0..29, plus x bit 10 -> key bit 30 and
 *   y bit 10 -> key bit 31 (x, y < 2048; z < 1024). Identical keys below 1024. */
#define ONC_KEY_MORTON30 30
#define ONC_KEY_MORTON32 32

/* field boundary conditions (engine/assets/shaders/tumor_sim.wgsl:414-446) */
#define ONC_BC_DIRICHLET 0    /* every boundary-layer voxel is held at bc_value */
#define ONC_BC_POINT_SOURCE 1 /* zero-flux faces; boundary-layer voxels of face z=0
                                 within src_radius of the face centre are held at bc_value */
#define ONC_BC_NONE 2         /* zero-flux faces, no fixed voxels (static or closed fields) */

/* cell column element types (all columns are 32-bit) */
#define ONC_COL_F32 0
#define ONC_COL_I32 1
#define ONC_COL_U32 2

/* fixed column slots: every cell schema starts with pos.x, pos.y, pos.z (f32), id (u32) */
#define ONC_COL_PX 0
#define ONC_COL_PY 1
#define ONC_COL_PZ 2
#define ONC_COL_ID 3

/* built-in tumour rule: parameter slots and the extra columns it uses */
#define ONC_TP_DIV_AGE 0     /* divide when age > div_age ...           */
#define ONC_TP_DIV_O2 1      /* ... and field 0 > div_o2                 */
#define ONC_TP_HYPOXIA 2     /* die when field 0 < hypoxia               */
#define ONC_TP_DRUG_KILL 3   /* die with probability drug_kill * field 1 */
#define ONC_TP_DAUGHTER_OFS 4 /* daughter offset length                  */
#define ONC_COL_AGE 4        /* f32 */
#define ONC_COL_GEN 5        /* u32 */

typedef struct onc_ctx onc_ctx;
typedef struct onc_module onc_module;
typedef struct onc_sim onc_sim;
typedef struct onc_batch onc_batch;

typedef struct {
    float rate;       /* D*dt/dx^2 of the explicit 7-point update */
    float decay;      /* first-order decay per step */
    float init;       /* initial value of every voxel */
    int32_t bc;       /* ONC_BC_* */
    float bc_value;   /* Dirichlet / source value */
    float src_radius; /* ONC_BC_POINT_SOURCE disc radius, in voxels */
    float uptake;     /* amount removed from a voxel per resident cell per step */
    float lo, hi;     /* clamp range of the updated value */
} onc_field_desc;

typedef struct {
    int32_t dims[3];  /* voxels per axis; bins are voxels */
    float dx;         /* voxel edge (= bin size = tissue_voxel_size) */
    int32_t key_format;
    uint32_t capacity; /* maximum live cells */
    uint64_t seed;     /* Philox4x32-10 key */
    /* mechanics */
    float cell_radius;        /* contact distance is 2*cell_radius */
    float interaction_radius; /* <= dx */
    float k_rep, k_adh;
    float mobility;           /* displacement = mobility * force */
    float max_disp;           /* displacement length clamp per step */
    int32_t n_fields;
    onc_field_desc fields[ONC_MAX_FIELDS];
    int32_t n_cols;                   /* >= 4 */
    int32_t col_types[ONC_MAX_COLS];  /* ONC_COL_* per column */
    int32_t n_params;
    float params[ONC_MAX_PARAMS];     /* rule parameters (Params dataclass values) */
    /* coarse chemistry grid: max(1, dims / 2) voxels per axis of edge 2 dx
       (core/utils/_types.py:331-346, _engine.py:75-80); n_chem_cols 32-bit
       columns of the Chemistry record, all zero at creation */
    int32_t n_chem_cols;
    int32_t chem_col_types[ONC_MAX_CHEM_COLS];  /* ONC_COL_* per column */
} onc_sim_desc;

typedef struct {
    char name[64];
    char arch[32];
    int32_t compute_units;
    int32_t xcds;
    uint64_t hbm_bytes;
    uint64_t free_bytes;
} onc_device_info;

/* ---- errors / version -------------------------------------------------- */
const char* onc_last_error(void);
int onc_abi_version(void);

/* ---- context (one per device) ------------------------------------------- */
int onc_ctx_create(int device, onc_ctx** out);
int onc_ctx_destroy(onc_ctx* ctx);
int onc_ctx_info(onc_ctx* ctx, onc_device_info* out);
/* raw device memory, for the pure-kernel entry points below */
int onc_dev_alloc(onc_ctx* ctx, size_t bytes, void** out);
int onc_dev_free(onc_ctx* ctx, void* p);
/* pinned (page-locked) host memory for large cell and field transfers
   (hipHostMalloc; the reference's buffers are mapped by wgpu,
   _device_manager.py:76-119) */
int onc_host_alloc(onc_ctx* ctx, size_t bytes, void** out);
int onc_host_free(onc_ctx* ctx, void* p);
int onc_h2d(onc_ctx* ctx, void* dst, const void* src, size_t bytes);
int onc_d2h(onc_ctx* ctx, void* dst, const void* src, size_t bytes);
int onc_ctx_sync(onc_ctx* ctx);

/* ---- rule modules (replaces WGSL create_shader_module,
 *      gpu/runtime/_device_manager.py:81; source comes from the Python
 *      AST -> HIP translator, cf. gpu/compiler/_compiler.py:233-352) ------ */
int onc_module_build(onc_ctx* ctx, const char* hip_src, const char* const* opts, int n_opts,
                     onc_module** out);
/* copies the hiprtc log (NUL-terminated, truncated to len) */
int onc_module_log(onc_module* mod, char* buf, size_t len);
int onc_module_destroy(onc_module* mod);

/* ---- one simulation (replaces GpuBackend.compile buffer setup,
 *      gpu/runtime/_gpu_backend.py:42-104) ------------------------------- */
int onc_sim_create(onc_ctx* ctx, const onc_sim_desc* desc, onc_sim** out);
int onc_sim_destroy(onc_sim* sim);
/* mod == NULL selects the built-in tumour rule; otherwise kernel `name` of mod.
   A module that also holds translated @tissue_rule / @chemistry_rule methods
   (kernels onc_tissue_pass / onc_chem_pass and the words onc_grid_tissue_passes
   / onc_grid_chem_passes) runs them every step after the diffusion sweep and
   before the cell rule (replaces the tissue and chemistry dispatches of
   GpuBackend.run_step, gpu/runtime/_gpu_backend.py:127-142) */
int onc_sim_set_rule(onc_sim* sim, onc_module* mod, const char* name);
int onc_sim_set_params(onc_sim* sim, const float* params, int n);
/* columns: n_cols host pointers, each n 32-bit elements (sync_to_device) */
int onc_sim_upload_cells(onc_sim* sim, const void* const* cols, uint32_t n, uint32_t next_id);
/* reads the live count into *n and up to cap elements per column (sync_to_host) */
int onc_sim_download_cells(onc_sim* sim, void* const* cols, uint32_t cap, uint32_t* n);
int onc_sim_upload_field(onc_sim* sim, int field, const float* host);
int onc_sim_download_field(onc_sim* sim, int field, float* host);
/* one column of the chemistry grid (32-bit words, x fastest over its voxels)
   (the chemistry part of sync_to_device / sync_to_host, _gpu_backend.py:369-371,
   399-405) */
int onc_sim_upload_chem(onc_sim* sim, int col, const void* host);
int onc_sim_download_chem(onc_sim* sim, int col, void* host);
/* run nsteps timesteps (asynchronous; replaces GpuBackend.run_step) */
int onc_sim_step(onc_sim* sim, int nsteps);
int onc_sim_sync(onc_sim* sim);
/* synchronises; returns live cells, next id and completed steps */
int onc_sim_counts(onc_sim* sim, uint32_t* n, uint32_t* next_id, uint32_t* steps);
/* The sim advances only while its completed steps are below `stop`; later
   steps of its batch leave it untouched (no reference counterpart: it lets a
   batch hold jobs with different horizons without host round trips).
   Default: no limit (0xFFFFFFFF). */
int onc_sim_set_stop(onc_sim* sim, uint32_t stop);
/* sets the completed-step counter (the Philox counter word of the next
   step's draws): restores a sim from a bundle so that it continues with the
   draws it would have made (no counterpart in the reference, which has no
   loader; its export writes state.bin, _gpu_backend.py:295-334) */
int onc_sim_set_step(onc_sim* sim, uint32_t step);
/* metric rules: runs the module's onc_metric_main / onc_metric_final kernels
   (translated @ol.metric_rule methods) over the sim's committed cells and
   writes the nvals reduced metric words (binary32 / int32 / uint32 bits, in
   the metrics class's field order, nvals <= 64; ONC_ERR_INVALID unless it
   equals the module's onc_metric_nvals word) to out (host memory). Sums
   in a fixed order set by the cell count (ONC_METRIC_CPT in onc_device.h),
   so binary32 metrics are reproducible; integer metrics wrap mod 2^32.
   (replaces the atomicAdd metric kernels of _shader_builder.py:193-204 and
   _stmt_translator.py:240-268, and the metric loop of CpuBackend.run_step,
   cpu/_cpu_backend.py:65-70) */
int onc_sim_metrics(onc_sim* sim, onc_module* module, int nvals, uint32_t* out);
/* sum over completed steps of the live cells at the start of each step */
int onc_sim_cell_updates(onc_sim* sim, uint64_t* cell_updates);
/* device bytes this sim holds (HBM admission accounting) */
int onc_sim_footprint(const onc_sim_desc* desc, uint64_t* bytes);

/* ---- batches: many sims of uneven sizes, one launch per phase over a
 *      packed segment table; the step can be captured as a hipGraph ----- */
int onc_batch_create(onc_ctx* ctx, onc_batch** out);
int onc_batch_destroy(onc_batch* b);
int onc_batch_add(onc_batch* b, onc_sim* sim);
int onc_batch_remove(onc_batch* b, onc_sim* sim);
int onc_batch_size(onc_batch* b, int* n);
/* use_graph: replay captured hipGraphs of the timestep (one step, and 16
   steps once the one-step graph has been replayed 64 times) nsteps times */
int onc_batch_step(onc_batch* b, int nsteps, int use_graph);
int onc_batch_sync(onc_batch* b);
/* tile shapes: cells per thread of a radix-sort tile and of a compaction
   tile (4, 8, 12 or 16), and the stencil block (x, y with x*y == 256, z-chunk).
   The tile sizes given here are kept whatever the population (it sets the
   tile policy's big_from to 0). */
int onc_batch_set_tuning(onc_batch* b, int sort_items, int compact_items, int stencil_bx,
                         int stencil_by, int stencil_bz);
/* Tile sizes by population (no counterpart in the reference, whose kernels
   have one fixed workgroup size): a batch's steps use the tuned sizes until
   its live cells reach big_from, then 16 / 8 cells per thread (sort /
   compaction, where larger) until they fall below small_below. The live count
   is read back asynchronously every 16 steps, so a change takes effect a few
   steps late; results are the same bits either way. Default 786432 / 524288;
   big_from = 0 keeps the tuned sizes. onc_batch_tile_items gives the sizes the
   batch's next steps use. */
int onc_batch_set_tile_policy(onc_batch* b, uint32_t big_from, uint32_t small_below);
int onc_batch_tile_items(onc_batch* b, int* sort_items, int* compact_items);
/* fused = 1 (the default): the neighbour forces and the 7-point sweep run as
   one launch, side by side on the same CUs (the force walk is bound by gather
   latency and ALU, the sweep by HBM); 0: two launches, forces then the sweep.
   Results are identical either way. (The reference runs them as separate
   compute passes, gpu/runtime/_gpu_backend.py:127-156.) */
int onc_batch_set_fused(onc_batch* b, int fused);
/* cap on the one-wave blocks of the neighbour-force walk: 0 (the default) =
   one per 64 cells of the batch's capacity; otherwise a positive multiple of
   8, and each wave strides through its XCD's 64-cell chunks. Results are
   identical for every value. */
int onc_batch_set_forces_waves(onc_batch* b, int waves);
/* fused = 1 (the default): when every sim of the batch runs the built-in
   tumour rule on the six-column tumour schema, the rule is evaluated inside
   the compaction kernel (one launch; the rule's output never goes through
   memory); 0: the rule kernel, then compaction. Results are identical. */
int onc_batch_set_fused_rule(onc_batch* b, int fused);
/* per-phase HIP-event timing of eager steps on the batch stream: bit k of
   phase_mask times phase k (0 plan, 1 Morton, 2 sort, 3 permute + bins,
   4 forces, 5 stencil, 6 rule, 7 compaction, 8 forces + stencil in one launch,
   9 rule + compaction in one launch); 0 disables */
int onc_batch_profile(onc_batch* b, int phase_mask);
/* the same, timing only every `every`-th eager step (the first, then every
   every-th after it): each event pair costs GPU time between launches, so a
   long run samples its steps instead (no counterpart in the reference) */
int onc_batch_profile_every(onc_batch* b, int phase_mask, int every);
#define ONC_PHASES 10
int onc_batch_phase_times(onc_batch* b, double* us_total, uint64_t* launches, int n);
/* how the last step's 7-point sweep treated the batch's field voxels (after
   committing the step): read and written (swept), written with a tile's one
   value without reading (filled), or left as they were (kept). A tile is
   filled or kept when its input neighbourhood holds one value that the update
   leaves unchanged; the fields are the same bits either way. No counterpart in
   the reference (its sweep touches every voxel); for measurement. */
int onc_batch_sweep_voxels(onc_batch* b, uint64_t* swept, uint64_t* filled, uint64_t* kept);

/* ---- pure kernels on device pointers (per-kernel parity and roofline) ---- */
/* keys[i] = Morton key of clamp(floor(p_i / dx), 0, dims-1) (template.wgsl:174-186) */
int onc_morton_keys(onc_ctx* ctx, const float* px, const float* py, const float* pz, uint32_t n,
                    const int32_t dims[3], float dx, int key_format, uint32_t* keys);
/* stable LSD radix sort of (key, value) pairs by the low key_bits bits
 * (replaces the atomic counting sort of template.wgsl:165-186, 335-343) */
int onc_sort_pairs(onc_ctx* ctx, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t n,
                   int key_bits, uint32_t* keys_out, uint32_t* vals_out);
/* raw neighbour forces of n cells whose positions (device pointers) are
 * sorted by Morton key (stable, as after a step's sort): for each cell the
 * linear repulsion / adhesion taper summed over the cells of its 27 Moore
 * voxels in visiting order, before integration; the sim descriptor gives the
 * grid, key format and force constants (its fields and capacity are not
 * used). (replaces the neighbour loop of the cell-rule kernel,
 * _stmt_translator.py:459-491, with the force law of tumor_sim.wgsl:505-628) */
int onc_forces(onc_ctx* ctx, const onc_sim_desc* desc, const float* px, const float* py, const float* pz,
               uint32_t n, float* fx, float* fy, float* fz);
/* exclusive prefix sum; *total (device pointer, may be NULL) receives the sum
 * (replaces template.wgsl:192-329, Blelloch 3-level scan) */
int onc_exscan_u32(onc_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total);
/* onc_exscan_u32 without any allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Its look-back status
 * and error words live in a fixed 2 MiB + 256 B device array per GPU, so n is
 * limited to 2^30 (ONC_ERR_INVALID beyond); calls that overlap on one GPU
 * share that array and race. Same results as onc_exscan_u32. */
int onc_exscan_u32_noalloc(onc_ctx* ctx, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total);
/* bin start/end for every occupied key of sorted keys; other entries untouched */
int onc_bins(onc_ctx* ctx, const uint32_t* sorted_keys, uint32_t n, uint32_t* start,
             uint32_t* end);
/* one explicit 7-point sweep of one field (tumor_sim.wgsl:414-446) */
int onc_stencil7(onc_ctx* ctx, const float* in, float* out, const int32_t dims[3],
                 const onc_field_desc* f);
/* plain copy of n 32-bit words (the calibration kernel of HBM byte counters) */
int onc_copy_u32(onc_ctx* ctx, const uint32_t* src, uint32_t* dst, uint64_t n);
/* GPU timing of the pure kernels above (no counterpart in the reference,
 * which has no timing code): with on = 1, each later call records HIP events
 * on the context stream around its kernel launches only (not its set-up
 * copies), and onc_ctx_last_kernel_us gives that span in microseconds. */
int onc_ctx_timing(onc_ctx* ctx, int on);
int onc_ctx_last_kernel_us(onc_ctx* ctx, double* us);
/* Philox4x32-10: out[4i..4i+3] = philox(ctr[4i..4i+3], key) */
int onc_philox(onc_ctx* ctx, const uint32_t* ctr, uint32_t n, uint64_t key, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* ONCO_H */
