// onc_internal.h — host-side declarations shared by the kernel launchers
// (onc_kernels.hip) and the C ABI runtime (onc_runtime.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "onc_device.h"

#define ONC_CTR_SORT0 0   // counters[0..3]: sort passes
#define ONC_CTR_COMPACT 4 // counters[4]: compaction scan
#define ONC_CTR_SCAN 5    // counters[5]: standalone exclusive scan
#define ONC_NCOUNTERS 16

#define ONC_BLOCK 256

// Sticky device error bits (onc_kargs::err): a bounded look-back wait expired
// in a sort pass (1), the compaction scan (2) or the standalone scan (4); a
// voxel held more than ONC_MAX_BIN_CELLS cells (8; the force walk keeps bin
// lengths in 16 bits).
#define ONC_DERR_SORT 1u
#define ONC_DERR_COMPACT 2u
#define ONC_DERR_SCAN 4u
#define ONC_DERR_BIG_BIN 8u
#define ONC_MAX_BIN_CELLS 65535u
#define ONC_MAX_METRICS 64  // metric fields of one metric module
#define ONC_SCAN_MIN_TILE 4096u  // fewest elements per tile of the exclusive scan (status words to allocate)

// Width in float4 columns of a float4 stencil tile: the smallest power of two
// >= X/4, capped at sbx (a power of two <= 256). Lanes of one row segment of
// min(width, 64) lanes hold consecutive float4 of one row, so x halos pass
// between neighbouring lanes; the tile is 256 / width rows tall.
__host__ __device__ inline int onc_stencil_tw(int X4, int sbx) {
    int p = 1;
    while (p < X4 && p < sbx) p <<= 1;
    return p;
}

// Phase ids for per-phase timing (onc_batch_phase_times).
enum onc_phase {
    ONC_PH_PLAN = 0,
    ONC_PH_MORTON = 1,
    ONC_PH_SORT = 2,
    ONC_PH_PERMUTE_BINS = 3,
    ONC_PH_FORCES = 4,
    ONC_PH_STENCIL = 5,
    ONC_PH_RULE = 6,
    ONC_PH_COMPACT = 7,
    ONC_PH_FORCES_STENCIL = 8,
    ONC_PH_RULE_COMPACT = 9,
};


// Launchers (all asynchronous on `s`). Grid sizes are chosen by the runtime;
// `items` (cells per thread of a tile kernel) is 4, 8, 12 or 16 for the sort
// passes and 4, 8 or 16 for compaction.
// Plan / commit modes of k_plan.
#define ONC_PLAN_STEP 0
#define ONC_PLAN_COMMIT 1
hipError_t onc_launch_plan(const onc_kargs& a, int mode, hipStream_t s);
// runs: count digits once per run of equal digits in a wave (large populations)
hipError_t onc_launch_morton(const onc_kargs& a, int grid, bool runs, hipStream_t s);
hipError_t onc_launch_hist_only(const onc_kargs& a, int items, int grid, hipStream_t s);
hipError_t onc_launch_sort_pass(const onc_kargs& a, int items, int pass, int grid, hipStream_t s);
// Blocks of the radix pass with `items` cells per thread that one CU holds at
// once (0 if the runtime cannot tell).
int onc_sort_pass_blocks_per_cu(int items);
// Tiles up to which a radix pass takes XCD-contiguous tile ranges: three
// quarters of the blocks the chip holds at once, so every tile's block is
// resident together even with some of the chip taken by other work (the
// library otherwise assumes the device to itself while a step runs).
inline unsigned onc_sort_xcd_tiles(int items, int cus) {
    return (unsigned)(onc_sort_pass_blocks_per_cu(items) * cus * 3 / 4);
}
// onc_forces: pack positions and keys, bin entries (epoch 1), then the raw
// force walk of one sim table entry S (device) with a.cell_prefix = {0, n};
// t0 / t1 (may be null) bracket the force kernel alone
hipError_t onc_launch_forces_raw(const onc_kargs& a, const float* px, const float* py, const float* pz, onc_u32 n,
                                 const onc_dsim* S, float* pos4, onc_u32* keys, onc_u32* bins, onc_u32* err,
                                 float* fx, float* fy, float* fz, hipEvent_t t0, hipEvent_t t1, hipStream_t s);
// the three-launch radix pass of large standalone sorts (tiles of 256 x ONC_SORT3_ITEMS keys)
#define ONC_SORT3_ITEMS 16
#define ONC_SORT3_MIN_N (1u << 20)
hipError_t onc_launch_sort_up(const onc_u32* keys, onc_u32 n, onc_u32 shift, onc_u32* counts, onc_u32 T,
                              hipStream_t s);
hipError_t onc_launch_sort_down(const onc_u32* kin, const onc_u32* vin, onc_u32* kout, onc_u32* vout, onc_u32 n,
                                onc_u32 shift, const onc_u32* offsets, onc_u32 T, hipStream_t s);
hipError_t onc_launch_permute_bins(const onc_kargs& a, int grid, hipStream_t s);
hipError_t onc_launch_forces(const onc_kargs& a, int grid, hipStream_t s);
hipError_t onc_launch_stencil(const onc_kargs& a, onc_u32 nblocks, hipStream_t s);
// forces + sweep in one launch of 8 * ngroups one-wave blocks (a.fs_sgroups of the groups sweep)
hipError_t onc_launch_forces_stencil(const onc_kargs& a, onc_u32 ngroups, hipStream_t s);
hipError_t onc_launch_rule_builtin(const onc_rule_args& r, int grid, hipStream_t s);
hipError_t onc_launch_compact(const onc_kargs& a, int items, int grid, hipStream_t s);
// built-in tumour rule fused into compaction (six-column tumour schema)
hipError_t onc_launch_rule_compact(const onc_kargs& a, int items, int grid, hipStream_t s);

// pure kernels
hipError_t onc_launch_morton_keys(const float* px, const float* py, const float* pz, onc_u32 n, int X,
                                  int Y, int Z, float dx, int fmt, onc_u32* keys, hipStream_t s);
hipError_t onc_launch_bins(const onc_u32* keys, onc_u32 n, onc_u32* start, onc_u32* end, hipStream_t s);
hipError_t onc_launch_philox(const onc_u32* ctr, onc_u32 n, onc_u32 k0, onc_u32 k1, onc_u32* out,
                             hipStream_t s);
// one block per 4096-element tile (static order, decoupled look-back)
hipError_t onc_launch_exscan(const onc_u32* in, onc_u32* out, onc_u32 n, onc_u32* total,
                             onc_u32* status, onc_u32* err, hipStream_t s);
hipError_t onc_launch_fill_u32(onc_u32* p, onc_u32 v, size_t n, hipStream_t s);
// fld[par] -> fld[par ^ 1] for every field of every unfrozen sim (after an odd
// number of tissue-rule passes); gx blocks per sim
hipError_t onc_launch_copy_fields(const onc_kargs& a, unsigned gx, hipStream_t s);
hipError_t onc_launch_field_init(float* u, int X, int Y, int Z, const onc_dfield& f, hipStream_t s);
hipError_t onc_launch_copy_u32(const onc_u32* src, onc_u32* dst, size_t n, hipStream_t s);
