// onc_device.h — device-side definitions shared by the fixed gfx950 kernels
// (onc_kernels.hip) and by rule modules that the Python translator emits and
// hiprtc compiles. Self-contained: no standard headers, so hiprtc can include
// it as a named header.
//
// Floating-point expressions here are evaluated in the exact order of the CPU
// oracle; the library and every rule module are compiled with
// -ffp-contract=off and IEEE division / square root, so binary32 results are
// bit-identical to the oracle's.
#pragma once

typedef unsigned int onc_u32;
typedef int onc_i32;
typedef unsigned long long onc_u64;

#define ONC_DEV_MAX_FIELDS 4
#define ONC_DEV_MAX_COLS 48
#define ONC_DEV_MAX_PARAMS 32
#define ONC_DEV_MAX_PASSES 4
#define ONC_DEV_MAX_CHEM 16
#define ONC_HIST_COPIES 8  // copies of the per-sim digit histograms (atomic spread)

// Device buffers reached through the segment table and the kernel arguments
// live in global memory. Typed so in device code, their loads and stores
// compile to global instructions; as generic (flat) instructions they would
// also count against the wait counter of LDS and scalar loads, so every LDS
// read or scalar load would first wait for all memory accesses in flight.
// The host sees plain pointers of the same size and layout (as does the device
// pass of a translation unit with host code only, which defines ONC_HOST_ONLY).
//
// The segment table itself is read through the constant address space: no
// kernel but the plan writes it (through a cast of its own), so uniform reads
// of a sim's constants and buffer pointers become scalar loads instead of
// vector loads whose completion every later memory wait would also count.
#if defined(__HIP_DEVICE_COMPILE__) && !defined(ONC_HOST_ONLY)
#define ONC_GLOBAL __attribute__((address_space(1)))
#define ONC_CONST __attribute__((address_space(4)))
#else
#define ONC_GLOBAL
#define ONC_CONST
#endif

typedef onc_u32 ONC_GLOBAL* onc_gcol;  // a state column in device memory

#define ONC_COL_ID_DEV 3  // the id column of every cell schema
#define ONC_FLAG_ALIVE 1u
#define ONC_FLAG_DIV 2u

// Mirrors onc_field_desc of include/onco.h (layout checked by static_assert).
struct onc_dfield {
    float rate, decay, init;
    int bc;
    float bc_value, src_radius, uptake, lo, hi;
};

// Dynamic per-sim counters, updated on the device only (graph replays need
// no host round trip). A step is committed by the next plan (or before the
// host reads the sim): the live cells are then st[parity][c][0, split)
// (survivors, in sorted order) followed by dau2[c][0, n - split) (the step's
// daughters in parent order).
struct onc_dstate {
    onc_u32 n;        // live cells
    onc_u32 next_id;  // id of the next daughter
    onc_u32 step;     // completed steps (Philox counter word 1)
    onc_u32 parity;   // which state / field buffer is current
    onc_u32 nsurv, ndiv;           // written by the compaction scan
    onc_u32 split;                 // live cells held in st[parity]; the rest are in dau2
    onc_u32 pending;               // a step ran and is not committed yet
    onc_u32 pad2;
    onc_u32 stop;                  // the sim advances only while step < stop
    onc_u64 cell_updates;          // sum over steps of live cells at step start
    onc_u32 pad[4];
};

// Bin table entries. The step with `completed` steps done writes its
// occupied bins with tag onc_bin_tag(completed) in [1, ONC_BIN_TAGS], never 0
// or 0xFFFF (what a cleared or filled entry holds); lengths are 16-bit (a
// voxel with more cells is reported as an error). A tag comes round again
// after ONC_BIN_TAGS steps; the Morton kernel of every step refills one of
// ONC_BIN_SLICES slices of the table, so each entry is refilled twice in that
// span and a stale entry can never carry the current tag.
#define ONC_BIN_TAGS 65534u
#define ONC_BIN_SLICES 32767u
__host__ __device__ inline onc_u32 onc_bin_tag(onc_u32 completed) { return completed % ONC_BIN_TAGS + 1u; }
// Length of entry b in the step with tag `tag` (0: empty, stale or unwritten).
__device__ __forceinline__ onc_u32 onc_bin_len(uint2 b, onc_u32 tag) {
    return (b.y & 0xFFFFu) == tag ? (b.y >> 16) : 0u;
}

// One entry of the packed segment table: a sim's geometry, model constants
// and device buffers.
struct onc_dsim {
    int dims[3];
    float dx;
    int key_format;
    int key_bits;
    onc_u32 cap;
    onc_u32 seed_lo, seed_hi;
    // This step's live count, buffer parity, completed steps and next daughter
    // id: copied from *state by the plan kernel at the start of every step, so
    // kernels read them together with the sim's other constants instead of
    // one dependent load later through the state pointer.
    onc_u32 n_live, par, step, next_id;
    onc_u32 frozen;  // step >= stop: this step leaves the sim untouched
    onc_u32 split;   // cells [0, split) are in st[par], [split, n_live) in dau2
    onc_u32 job0;    // the sweep job of this sim's field 0 in its batch (field f: job0 + f)
    onc_u32 pad_;
    float cell_radius, R, k_rep, k_adh, mobility, max_disp;
    int n_fields, n_cols;
    onc_dfield fields[ONC_DEV_MAX_FIELDS];
    float params[ONC_DEV_MAX_PARAMS];
    onc_dstate ONC_GLOBAL* state;
    onc_u32 ONC_GLOBAL* st[2][ONC_DEV_MAX_COLS];  // cell state columns, two buffers
    onc_u32 ONC_GLOBAL* dau[ONC_DEV_MAX_COLS];    // rule output: daughter of cell i at i
    onc_u32 ONC_GLOBAL* dau2[ONC_DEV_MAX_COLS];   // daughters compacted in parent order
    float ONC_GLOBAL* fld[2][ONC_DEV_MAX_FIELDS]; // fields, two buffers, x + X (y + Y z)
    onc_u32 ONC_GLOBAL* keys[2];
    onc_u32 ONC_GLOBAL* vals[2];
    onc_u32 ONC_GLOBAL* flags;
    // Bin table indexed by Morton key, 8 B per entry {start, len << 16 | tag}
    // (onc_bin_tag): an entry is valid in the step whose tag it carries; only
    // occupied bins are written, and the Morton kernel refills one slice of
    // the table per step, so stale entries never reach a repeated tag.
    onc_u32 ONC_GLOBAL* bins;
    onc_u32 table;  // entries of the bin table: Morton(dims - 1) + 1
    float ONC_GLOBAL* pos4;  // sorted positions as {x, y, z, 0}, written with the permute
    // Coarse chemistry grid: cdims voxels per axis, n_chem 32-bit columns, two
    // buffers. Every chemistry-rule pass reads one buffer and writes the other;
    // with chem_passes passes per step, the buffer holding the state after s
    // completed steps is (s * chem_passes) & 1.
    int cdims[3];
    int n_chem;
    onc_u32 chem_passes;
    onc_u32 pad_c;
    onc_u32 ONC_GLOBAL* chem[2][ONC_DEV_MAX_CHEM];
    // Uptake stamps of cubes of 2^ONC_QCUBE voxels per side, indexed by
    // Morton key >> (3 ONC_QCUBE): the step + 1 in which the permute kernel
    // last changed a field voxel of the cube (0: never). The sweep's quiet
    // tiles read them (onc_stencil_vec4).
    onc_u32 ONC_GLOBAL* qstamp;
};
#ifndef ONC_QCUBE
#define ONC_QCUBE 3
#endif

// Arguments of the grid-rule kernels of a translated module (onc_tissue_pass,
// onc_chem_pass): one launch per pass over every sim of the batch (blockIdx.y
// = sim). A tissue pass j reads fields fld[par ^ 1 ^ (j & 1)] and writes
// fld[par ^ (j & 1)]; a chemistry pass p of step s reads chem[(s P + p) & 1].
struct onc_grid_args {
    const onc_dsim ONC_CONST* sims;
    onc_u32 nsims;
    onc_u32 pass;
};

// Arguments of every batched kernel.
struct onc_kargs {
    const onc_dsim ONC_CONST* sims;
    onc_u32 nsims;
    onc_u32 passes;       // radix passes this step (max over sims)
    onc_u32 tile;         // cells per sort tile
    onc_u32 max_tiles;    // sort tiles the status array holds
    onc_u32 ONC_GLOBAL* tile_prefix; // [nsims + 1] sort tile prefix of live counts
    onc_u32 ctile;        // cells per compaction tile
    onc_u32 max_ctiles;   // compaction tiles the status array holds
    onc_u32 ONC_GLOBAL* ctile_prefix;// [nsims + 1] compaction tile prefix of live counts
    onc_u32 ONC_GLOBAL* cell_prefix; // [nsims + 1] prefix of live counts
    onc_u32 ONC_GLOBAL* hist;        // [ONC_HIST_COPIES][nsims][4][256] digit histograms
    onc_u32 ONC_GLOBAL* counters;    // [16] dynamic tile counters
    onc_u32 ONC_GLOBAL* sort_status; // [4][max_tiles][256] look-back words
    onc_u64 ONC_GLOBAL* comp_status; // [max_ctiles] look-back words of the compaction scan
    onc_u32 ONC_GLOBAL* jobs_prefix; // [njobs + 1] stencil blocks per (sim, field)
    onc_u32 ONC_GLOBAL* jobs;        // [njobs] sim << 8 | field
    onc_u32 njobs;
    int sbx, sby, sbz;    // stencil block shape
    onc_u32 fs_sgroups;   // fused forces + sweep launch: groups of 8 blocks that sweep
    onc_u32 ONC_GLOBAL* err;         // sticky error word (bounded waits that expired)
    onc_u32 xcd_tiles;    // sort passes: tiles whose blocks are all resident at once (0: none)
    // quiet sweep tiles: {state, value} per sweep block and field buffer,
    // [2][nqb][2] (null: every tile is swept)
    onc_u32 ONC_GLOBAL* quiet;
    onc_u32 nqb;
};

// ---- small helpers ---------------------------------------------------------
// Streaming (non-temporal) 16-byte store and load: the line is not kept for
// reuse in the caches (global_store / global_load with the nt bit).
typedef float onc_f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void onc_st_nt(float4* p, float4 v) {
    const onc_f4v t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, (onc_f4v*)p);
}
__device__ __forceinline__ float4 onc_ld_nt(const float4* p) {
    const onc_f4v t = __builtin_nontemporal_load((const onc_f4v*)p);
    return make_float4(t.x, t.y, t.z, t.w);
}

__device__ __forceinline__ float onc_clampf(float x, float lo, float hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}

// Voxel coordinate along one axis: floor of the IEEE quotient, clamped to
// [0, d-1] (floor and trunc agree on positions >= 0).
__device__ __forceinline__ int onc_vox1(float p, float dx, int d) {
    float q = p / dx;
    float f = __builtin_floorf(q);
    return f < 0.0f ? 0 : (f >= (float)d ? d - 1 : (int)f);
}

__device__ __forceinline__ onc_u32 onc_spread9(onc_u32 v) {
    v &= 0x1FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 3-D Morton key, 10 bits per axis (framework/src/oncolytica/core/runtime/
// _geometry.py:14-42 closed form); fmt 32 adds x bit 10 -> 30, y bit 10 -> 31.
__device__ __forceinline__ onc_u32 onc_morton(onc_u32 x, onc_u32 y, onc_u32 z, int fmt) {
    onc_u32 xl = x & 0x3FFu, yl = y & 0x3FFu, zl = z & 0x3FFu;
    onc_u32 j = onc_spread9(xl >> 1) | (onc_spread9(yl >> 1) << 1) | (onc_spread9(zl >> 1) << 2);
    onc_u32 k = (j << 3) | (xl & 1u) | ((yl & 1u) << 1) | ((zl & 1u) << 2);
    if (fmt == 32) k |= (((x >> 10) & 1u) << 30) | (((y >> 10) & 1u) << 31);
    return k;
}

// Morton keys of a voxel's 27 neighbours from its own key by dilated
// arithmetic: the x (y, z) bits of a key, kept under their mask M, form a
// dilated integer with x - 1 = (kx - 1) & M and x + 1 = ((kx | ~M) + 1) & M
// (the borrow or carry runs through the other axes' bits). The key of
// (x + dx, y + dy, z + dz) is x[dx + 1] | y[dy + 1] | z[dz + 1]: a few
// operations per neighbour instead of three bit interleavings each. Keys of
// coordinates outside [0, dims) are meaningless and must not be used.
struct onc_nkeys {
    onc_u32 x[3], y[3], z[3];
};
__device__ __forceinline__ onc_nkeys onc_neighbor_keys(onc_u32 k, int fmt) {
    const onc_u32 XM = fmt == 32 ? 0x49249249u : 0x09249249u;
    const onc_u32 YM = fmt == 32 ? 0x92492492u : 0x12492492u;
    const onc_u32 ZM = 0x24924924u;
    onc_nkeys n;
    const onc_u32 kx = k & XM, ky = k & YM, kz = k & ZM;
    n.x[0] = (kx - 1u) & XM;
    n.x[1] = kx;
    n.x[2] = ((kx | ~XM) + 1u) & XM;
    n.y[0] = (ky - 1u) & YM;
    n.y[1] = ky;
    n.y[2] = ((ky | ~YM) + 1u) & YM;
    n.z[0] = (kz - 1u) & ZM;
    n.z[1] = kz;
    n.z[2] = ((kz | ~ZM) + 1u) & ZM;
    return n;
}

__device__ __forceinline__ onc_u32 onc_cell_key(const onc_dsim& S, float px, float py, float pz) {
    return onc_morton((onc_u32)onc_vox1(px, S.dx, S.dims[0]), (onc_u32)onc_vox1(py, S.dx, S.dims[1]),
                      (onc_u32)onc_vox1(pz, S.dx, S.dims[2]), S.key_format);
}

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ void onc_philox(onc_u32 c0, onc_u32 c1, onc_u32 c2, onc_u32 c3, onc_u32 k0,
                                           onc_u32 k1, onc_u32 out[4]) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        onc_u32 hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
        onc_u32 hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
        onc_u32 n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Per-cell counter-based stream: call c reads word c & 3 of
// Philox(ctr = {id, step, c >> 2, tag}, key = seed). Cells draw with tag 0;
// voxels of grid-rule passes with a tag of their own (translated modules).
struct onc_rng {
    onc_u32 id, step, c, k0, k1, tag;
    onc_u32 buf[4];
};

__device__ __forceinline__ void onc_rng_init(onc_rng& r, const onc_dsim& S, onc_u32 id, onc_u32 step) {
    r.id = id;
    r.step = step;
    r.c = 0;
    r.tag = 0;
    r.k0 = S.seed_lo;
    r.k1 = S.seed_hi;
}

__device__ __forceinline__ float onc_rand(onc_rng& r) {
    if ((r.c & 3u) == 0) onc_philox(r.id, r.step, r.c >> 2, r.tag, r.k0, r.k1, r.buf);
    // a select chain, not r.buf[r.c & 3]: a runtime index would put the
    // buffer in scratch memory
    const onc_u32 q = r.c & 3u;
    const onc_u32 w = q == 0u ? r.buf[0] : (q == 1u ? r.buf[1] : (q == 2u ? r.buf[2] : r.buf[3]));
    r.c++;
    return (float)(w >> 8) * 5.9604644775390625e-08f;
}

// sin, cos on [0, pi/2] from the Taylor series (Horner in a^2), the same
// operation sequence as the oracle.
__device__ __forceinline__ void onc_sincos_q(float a, float* so, float* co) {
    float a2 = a * a;
    float s = (float)1.6059043836821613e-10;
    s = s * a2 + (float)-2.5052108385441720e-08;
    s = s * a2 + (float)2.7557319223985893e-06;
    s = s * a2 + (float)-1.9841269841269841e-04;
    s = s * a2 + (float)8.3333333333333333e-03;
    s = s * a2 + (float)-1.6666666666666667e-01;
    s = s * a2 + 1.0f;
    s = s * a;
    float c = (float)-1.1470745597729725e-11;
    c = c * a2 + (float)2.0876756987868099e-09;
    c = c * a2 + (float)-2.7557319223985891e-07;
    c = c * a2 + (float)2.4801587301587302e-05;
    c = c * a2 + (float)-1.3888888888888889e-03;
    c = c * a2 + (float)4.1666666666666667e-02;
    c = c * a2 + -0.5f;
    c = c * a2 + 1.0f;
    *so = s;
    *co = c;
}

// Unit vector: z = 2 u1 - 1, phi = 2 pi u2 with exact quadrant reduction
// (gpu/compiler/template.wgsl:84-98 z/phi form).
__device__ __forceinline__ void onc_rand_dir(onc_rng& r, float* x, float* y, float* z) {
    float u1 = onc_rand(r);
    float u2 = onc_rand(r);
    float zz = u1 * 2.0f - 1.0f;
    float rr = 1.0f - zz * zz;
    if (rr < 0.0f) rr = 0.0f;
    float rad = __builtin_sqrtf(rr);
    float t = u2 * 4.0f;
    int q = (int)t;
    float f = t - (float)q;
    float a = f * (float)1.5707963267948966;
    float s, c;
    onc_sincos_q(a, &s, &c);
    float cx, sy;
    switch (q & 3) {
        case 0: cx = c; sy = s; break;
        case 1: cx = -s; sy = c; break;
        case 2: cx = -c; sy = -s; break;
        default: cx = s; sy = -c; break;
    }
    *x = rad * cx;
    *y = rad * sy;
    *z = zz;
}

// Largest s with prefix[s] <= g (prefix is non-decreasing, prefix[0] = 0).
__device__ __forceinline__ int onc_find_seg(const onc_u32* prefix, onc_u32 nsegs, onc_u32 g) {
    int lo = 0, hi = (int)nsegs - 1;
    while (lo < hi) {
        int m = (lo + hi + 1) >> 1;
        if (prefix[m] <= g) lo = m; else hi = m - 1;
    }
    return lo;
}

// Value known to be equal on every lane of the wave: lets the compiler keep
// it (and what is loaded through it) in scalar registers.
__device__ __forceinline__ onc_u32 onc_uniform(onc_u32 x) { return __builtin_amdgcn_readfirstlane(x); }

// Largest s in [0, n) with prefix[s] <= g (prefix non-decreasing, prefix[0]
// = 0), for g equal on every lane; all 64 lanes of the wave must be active.
// Lane l loads prefix[base + l] and a ballot counts the entries <= g: one
// memory round trip per 64 entries instead of a binary search of dependent
// loads (a 64-sim batch: one load instead of six).
__device__ __forceinline__ int onc_find_seg_wave(const onc_u32* prefix, onc_u32 n, onc_u32 g) {
    const onc_u32 lane = threadIdx.x & 63u;
    int s = 0;
    for (onc_u32 base = 0; base < n; base += 64u) {
        const onc_u32 idx = base + lane;
        const bool le = idx < n && prefix[idx] <= g;
        const int c = __popcll(__ballot(le));
        if (c == 0) break;
        s = (int)base + c - 1;
        if (c < 64) break;
    }
    return (int)onc_uniform((onc_u32)s);
}

// XCD-aware block order. The 8 XCDs of an MI355X (each with its own L2) take
// workgroups round-robin by blockIdx; this bijection gives XCD k a contiguous
// run of logical block ids, so blocks that share halo data or neighbour bins
// run on the same XCD and hit the same L2.
__device__ __forceinline__ onc_u32 onc_xcd_block(onc_u32 bid, onc_u32 nblocks) {
    const onc_u32 q = nblocks >> 3, r = nblocks & 7u;
    const onc_u32 xcd = bid & 7u, idx = bid >> 3;
    return xcd * q + (xcd < r ? xcd : r) + idx;
}

// Interleaved block order for work of uneven cost: XCD k takes groups of G
// consecutive logical block ids, the groups round-robin over the 8 XCDs, so
// every XCD walks the whole id range at the same pace (a contiguous eighth per
// XCD would leave the XCDs holding the dense sims of a batch working alone at
// the end). Blocks past the last full round of 8 G keep their own id.
__device__ __forceinline__ onc_u32 onc_xcd_group(onc_u32 bid, onc_u32 nblocks, onc_u32 G) {
    const onc_u32 full = nblocks / (8u * G) * (8u * G);
    if (bid >= full) return bid;
    const onc_u32 xcd = bid & 7u, j = bid >> 3;
    return ((j / G) * 8u + xcd) * G + j % G;
}

// Persistent walk over the live cells of every sim of the segment table, one
// wave64 chunk of 64 consecutive cells at a time. Each wave takes a contiguous
// range of chunks (waves of an XCD take adjacent ranges), so a wave's Morton
// neighbourhood stays in one L2. When a chunk lies inside one sim (the common
// case) the sim index is wave-uniform, so the segment-table entry and the
// column pointers are read with scalar loads. `f(s, i)` gets the sim and the
// cell's index in it.
// With XCD_GROUP > 0 the blocks are spread by onc_xcd_group instead (for
// chunks of uneven cost).
// `wave` / `nwaves`: this wave's index among all waves taking part.
template <class F>
__device__ __forceinline__ void onc_for_cells_wave(const onc_u32* cell_prefix, onc_u32 nsims, onc_u32 wave,
                                                   onc_u32 nwaves, F f) {
    const onc_u32 total = cell_prefix[nsims];
    const onc_u32 lane = threadIdx.x & 63u;
    const onc_u32 nchunks = (total + 63u) >> 6;
    const onc_u32 c0 = (onc_u32)(((unsigned long long)nchunks * wave) / nwaves);
    const onc_u32 c1 = (onc_u32)(((unsigned long long)nchunks * (wave + 1u)) / nwaves);
    for (onc_u32 c = c0; c < c1; c++) {
        const onc_u32 base = c * 64u;
        const onc_u32 g = base + lane;
        const int s0 = onc_find_seg_wave(cell_prefix, nsims, base);
        const onc_u32 lo = cell_prefix[s0], hi = cell_prefix[s0 + 1];
        if (base + 63u < hi) {
            f(s0, g - lo);
        } else if (g < total) {
            const int s = onc_find_seg(cell_prefix, nsims, g);
            f(s, g - cell_prefix[s]);
        }
    }
}

// Walk for a grid of nblk one-wave blocks (a multiple of 8) that may hold
// fewer waves than chunks. Each sim's 64-cell chunks form groups of
// ONC_FORCE_GROUP consecutive chunks (2048 cells, a compact stretch of the
// sim's Morton order), and XCD x (blocks 8 i + x, the dispatcher's
// round-robin) takes groups x, x + 8, x + 16, ... of every sim: a group's bin
// entries and positions stay in one XCD's L2 (fetched once, plus the halo it
// shares with its neighbours), while dense and sparse stretches of every sim
// are dealt evenly over the XCDs. (One contiguous eighth of each sim per XCD
// fetched the least - the raw walk read 62 MB for 1M cells against 128 MB
// here - but left an XCD holding a dense stretch working alone: the walk of
// the 16M-cell clustered workload took 20 % longer, and the 500-step batched
// ensemble ran 5 % slower.)
// Wave i of XCD x takes its XCD's chunks i, i + W_x, i + 2 W_x, ... in the
// order (sim, chunk): with at least one wave per chunk each wave takes one
// chunk. `f(s, i)` gets the sim and the cell's index in it; lanes past a
// sim's last cell do not call it.
#define ONC_FORCE_GROUP 32u
template <class F>
__device__ __forceinline__ void onc_for_chunks_xcd(const onc_u32* cell_prefix, onc_u32 nsims, onc_u32 blk,
                                                   onc_u32 nblk, F f) {
    constexpr onc_u32 GG = ONC_FORCE_GROUP;
    const onc_u32 lane = threadIdx.x & 63u;
    const onc_u32 x = blk & 7u, i = blk >> 3;
    const onc_u32 wx = nblk >> 3;  // nblk is a multiple of 8 (>= 8): every XCD class has waves
    onc_u32 m = i;       // this wave's next chunk among XCD x's chunks
    onc_u32 before = 0;  // XCD x's chunks in the sims of earlier groups
    // An XCD takes fewer than n / 512 + GG + 1 chunks of a sim of n cells
    // (at most ceil(groups / 8) groups of GG chunks, each group but the
    // last full): a wave past that bound for the whole batch has no chunk
    // and leaves before the prefix scan.
    if (i >= (cell_prefix[nsims] >> 9) + (GG + 1u) * nsims + 1u) return;
    for (onc_u32 g0 = 0; g0 < nsims; g0 += 64u) {
        // lane l: sim g0 + l, its cell count, and how many of its chunks XCD x takes
        const onc_u32 s = g0 + lane;
        onc_u32 n = 0, cnt = 0;
        if (s < nsims) {
            n = cell_prefix[s + 1] - cell_prefix[s];
            const onc_u32 ch = (n + 63u) >> 6;
            const onc_u32 ngr = (ch + GG - 1u) / GG;
            const onc_u32 ng = x < ngr ? (ngr - x + 7u) / 8u : 0u;
            cnt = ng * GG;
            if (ng && ((ngr - 1u) & 7u) == x) cnt -= ngr * GG - ch;  // the sim's last group is short
        }
        onc_u32 incl = cnt;  // inclusive prefix over the lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const onc_u32 t = __shfl_up(incl, d);
            if ((int)lane >= d) incl += t;
        }
        const onc_u32 gtot = onc_uniform(__shfl(incl, 63));
        for (; m < before + gtot; m += wx) {
            const onc_u32 r = m - before;
            // the owning lane: the first whose inclusive prefix exceeds r
            const int sl = __popcll(__ballot(incl <= r));
            const onc_u32 sim = onc_uniform(g0 + (onc_u32)sl);
            const onc_u32 rr = r - (__shfl(incl, sl) - __shfl(cnt, sl));  // rank among the sim's chunks on XCD x
            const onc_u32 lc = onc_uniform((x + 8u * (rr / GG)) * GG + rr % GG);
            const onc_u32 ns = onc_uniform(__shfl(n, sl));
            const onc_u32 cell = lc * 64u + lane;
            if (cell < ns) f((int)sim, cell);
        }
        before += gtot;
    }
}

template <int XCD_GROUP = 0, class F>
__device__ __forceinline__ void onc_for_cells(const onc_u32* cell_prefix, onc_u32 nsims, F f) {
    const onc_u32 wpb = blockDim.x >> 6;
    const onc_u32 blk = XCD_GROUP > 0 ? onc_xcd_group(blockIdx.x, gridDim.x, (onc_u32)XCD_GROUP)
                                      : onc_xcd_block(blockIdx.x, gridDim.x);
    onc_for_cells_wave(cell_prefix, nsims, onc_uniform(blk * wpb + (threadIdx.x >> 6)), gridDim.x * wpb, f);
}

__device__ __forceinline__ onc_u32 onc_field_index(const onc_dsim& S, float px, float py, float pz) {
    onc_u32 X = (onc_u32)S.dims[0], Y = (onc_u32)S.dims[1];
    return (onc_u32)onc_vox1(px, S.dx, S.dims[0]) +
           X * ((onc_u32)onc_vox1(py, S.dx, S.dims[1]) + Y * (onc_u32)onc_vox1(pz, S.dx, S.dims[2]));
}

// Iterate the cells of the 27 Moore bins around (px, py, pz) in
// template.wgsl:11-21 order (x fastest), skipping out-of-range bins and the
// cell `self`. `body(j)` receives the sorted index of each neighbour. The
// nine bin entries of a z-plane are loaded together (independent 8-byte
// loads) before any of them is walked; `tag` is the step's bin tag.
template <class F>
__device__ __forceinline__ void onc_for_neighbors(const onc_dsim& S, onc_u32 tag, float px, float py, float pz,
                                                  onc_u32 self, F body) {
    const int vx = onc_vox1(px, S.dx, S.dims[0]), vy = onc_vox1(py, S.dx, S.dims[1]),
              vz = onc_vox1(pz, S.dx, S.dims[2]);
    const uint2* bins = (const uint2*)S.bins;
    const onc_nkeys nk = onc_neighbor_keys(onc_morton((onc_u32)vx, (onc_u32)vy, (onc_u32)vz, S.key_format),
                                           S.key_format);
    for (int oz = -1; oz <= 1; oz++) {
        const int nz = vz + oz;
        if (nz < 0 || nz >= S.dims[2]) continue;
        // a select, not nk.z[oz + 1]: a runtime index would put nk in scratch memory
        const onc_u32 zk = oz < 0 ? nk.z[0] : (oz == 0 ? nk.z[1] : nk.z[2]);
        uint2 b[9];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const int ny = vy + q / 3 - 1, nx = vx + q % 3 - 1;
            const bool in = nx >= 0 && ny >= 0 && nx < S.dims[0] && ny < S.dims[1];
            b[q] = make_uint2(0u, 0u);
            if (in) b[q] = bins[nk.x[q % 3] | nk.y[q / 3] | zk];
        }
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const onc_u32 len = onc_bin_len(b[q], tag);
            for (onc_u32 j = b[q].x; j < b[q].x + len; j++) {
                if (j == self) continue;
                body(j);
            }
        }
    }
}

// Same visiting order as onc_for_neighbors, but the candidates of a z-plane's
// nine bins are walked as one flat list, four at a time: the four position
// loads are issued together and the pairs are then accumulated in order.
// `pair(o, j)` receives the neighbour's packed position {x, y, z, _} and index.
template <class F>
__device__ __forceinline__ void onc_for_neighbor_pos4(const onc_dsim& S, onc_u32 tag, float px, float py, float pz,
                                                      onc_u32 self, const float4* __restrict__ pos4, F pair) {
    const int vx = onc_vox1(px, S.dx, S.dims[0]), vy = onc_vox1(py, S.dx, S.dims[1]),
              vz = onc_vox1(pz, S.dx, S.dims[2]);
    const uint2* __restrict__ bins = (const uint2*)S.bins;
    const onc_nkeys nk = onc_neighbor_keys(onc_morton((onc_u32)vx, (onc_u32)vy, (onc_u32)vz, S.key_format),
                                           S.key_format);
    for (int oz = -1; oz <= 1; oz++) {
        const int nz = vz + oz;
        if (nz < 0 || nz >= S.dims[2]) continue;
        const onc_u32 zk = oz < 0 ? nk.z[0] : (oz == 0 ? nk.z[1] : nk.z[2]);
        onc_u32 st[9], pre[10];
        pre[0] = 0u;
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const int ny = vy + q / 3 - 1, nx = vx + q % 3 - 1;
            const bool in = nx >= 0 && ny >= 0 && nx < S.dims[0] && ny < S.dims[1];
            uint2 b = make_uint2(0u, 0u);
            if (in) b = bins[nk.x[q % 3] | nk.y[q / 3] | zk];
            const onc_u32 c = onc_bin_len(b, tag);
            st[q] = b.x;
            pre[q + 1] = pre[q] + c;
        }
        const onc_u32 tot = pre[9];
        for (onc_u32 t0 = 0; t0 < tot; t0 += 4u) {
            onc_u32 jj[4];
            float4 p[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const onc_u32 t = t0 + (onc_u32)u;
                onc_u32 q = 0;
#pragma unroll
                for (int r = 1; r < 9; r++) q += (t >= pre[r]) ? 1u : 0u;
                onc_u32 sq = st[0], pq = pre[0];
#pragma unroll
                for (int r = 1; r < 9; r++)
                    if (q == (onc_u32)r) {
                        sq = st[r];
                        pq = pre[r];
                    }
                jj[u] = t < tot ? sq + (t - pq) : 0xFFFFFFFFu;
                p[u] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (jj[u] != 0xFFFFFFFFu) p[u] = pos4[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (jj[u] != 0xFFFFFFFFu && jj[u] != self) pair(p[u], jj[u]);
        }
    }
}

// Arguments of a rule kernel (built-in or hiprtc module): one thread per live
// cell of every sim in the segment table.
struct onc_rule_args {
    const onc_dsim ONC_CONST* sims;
    onc_u32 nsims;
    onc_u32 passes;
    const onc_u32 ONC_GLOBAL* cell_prefix;
};

// Arguments of the two kernels of a translated metric module (onc_sim_metrics).
// onc_metric_main: block b takes the committed cells [b, b+1) x 256 x
// ONC_METRIC_CPT; thread t adds the contributions of cells
// (b ONC_METRIC_CPT + j) 256 + t for j = 0, 1, ... in that order, then the
// block adds its 256 per-thread values in a fixed tree (stride 128, 64, ...,
// 1) and writes one partial per metric. onc_metric_final: thread t adds
// partials t, t + 256, ... in that order, then the same tree. The order is
// fixed by the cell count alone, so binary32 metrics are reproducible.
#define ONC_METRIC_CPT 8
struct onc_metric_args {
    const onc_dsim ONC_GLOBAL* S;    // the sim's segment-table entry
    onc_u32 n, split, parity;        // committed cells: st[parity][0, split), then dau2[0, n - split)
    onc_u32 ONC_GLOBAL* partials;    // [blocks][metrics]
};
struct onc_metric_fin_args {
    const onc_u32 ONC_GLOBAL* partials;
    onc_u32 nblocks;
    onc_u32 ONC_GLOBAL* out;         // [metrics], binary32 / int32 / uint32 bits
};
