// onc_kernels.hip — fixed gfx950 kernels of the per-timestep hot path.
//
// Every kernel serves a whole batch: it walks a packed segment table of sims
// (onc_dsim) and finds its sim by binary search over device-side prefixes of
// the live counts, so one launch per phase covers sims of uneven sizes and a
// captured hipGraph replays without host round trips. Kernels over cells are
// persistent: a grid sized to the machine loops over tiles.
//
// Replaces the WGSL template kernels of the reference
// (framework/src/oncolytica/gpu/compiler/template.wgsl:152-343): the atomic
// counting sort into a Morton-indexed voxel table (Kernel_ClearVoxels,
// Kernel_CountAndOffset, Kernel_Scan_L0/L1/L2, Kernel_AddBack_L1/L0,
// Kernel_Scatter, Kernel_UpdateState) becomes a stable onesweep LSD radix
// sort, bins built from the sorted keys, and flag-scan compaction.
#include "onc_internal.h"

#define ONC_FLAG_A 0x40000000u
#define ONC_FLAG_P 0x80000000u
#define ONC_CNT_MASK 0x3FFFFFFFu
#define ONC_SPIN_LIMIT (1u << 24)
// status words a sort tile's look-back reads per memory round trip
#ifndef ONC_SORT_LOOKBACK
#define ONC_SORT_LOOKBACK 8
#endif
#define ONC_F64_A (1ull << 62)
#define ONC_F64_P (2ull << 62)
#define ONC_M31 0x7FFFFFFFull

static_assert(sizeof(onc_dfield) == 36, "onc_dfield layout");

// ---------------------------------------------------------------------------
// wave64 / block helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ onc_u32 ld_relaxed(const onc_u32* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(onc_u32* p, onc_u32 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ onc_u64 ld_relaxed64(const onc_u64* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed64(onc_u64* p, onc_u64 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a 256-thread block (4 wave64).
// All threads must call it; *total receives the block sum.
__device__ __forceinline__ onc_u32 block_exscan256(onc_u32 v, onc_u32* s_tmp, onc_u32* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    onc_u32 x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        onc_u32 y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[w] = x;
    __syncthreads();
    onc_u32 before = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        onc_u32 t = s_tmp[i];
        if (i < w) before += t;
        tot += t;
    }
    __syncthreads();
    if (total) *total = tot;
    return before + x - v;
}

// Decoupled look-back for one digit of a sort tile: sums the aggregates of
// tiles g-1, g-2, ... down to the nearest inclusive prefix, reading Q status
// words per round so one memory round trip covers Q tiles. Tiles below
// `first` (another sim) read as an inclusive prefix of 0. When every tile
// publishes its aggregate at about the same time, the last of T tiles needs
// about sqrt(2T/Q) rounds (340 tiles: 9 rounds at Q = 8, 5 at 32, 3 at 64).
template <int Q>
__device__ __forceinline__ onc_u32 lookback_digit(const onc_u32* status, onc_u32 g, onc_u32 first, onc_u32 d,
                                                  onc_u32* err) {
    onc_u32 prefix = 0u, j = g - 1u, spins = 0u;
    for (;;) {
        onc_u32 v[Q];
#pragma unroll
        for (int q = 0; q < Q; q++)
            v[q] = (j >= first + (onc_u32)q) ? ld_relaxed(&status[(size_t)(j - (onc_u32)q) * 256u + d]) : ONC_FLAG_P;
        // q = words taken before the first one not yet published (or through
        // the first inclusive prefix); unrolled so v stays in registers
        int q = 0;
        bool done = false, stop = false;
#pragma unroll
        for (int u = 0; u < Q; u++) {
            if (!stop) {
                const onc_u32 x = v[u];
                if ((x & (ONC_FLAG_A | ONC_FLAG_P)) == 0u) {
                    stop = true;
                } else {
                    prefix += x & ONC_CNT_MASK;
                    q = u + 1;
                    if (x & ONC_FLAG_P) done = stop = true;
                }
            }
        }
        if (done) return prefix;
        j -= (onc_u32)q;
        if (q < Q) {
            if (++spins > ONC_SPIN_LIMIT) {
                atomicOr(err, 1u);
                return prefix;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
}

// Wave-wide decoupled look-back over 64-bit status words (2 flag bits on top):
// lane l reads tile j - l, so one round trip covers 64 predecessors. Returns
// the sums of the low (bits 0..30) and high (bits 31..61) counters, or of the
// low 32 bits when `split` is false. Must be called by all 64 lanes of a wave.
__device__ __forceinline__ void lookback_wave64(const onc_u64* status, onc_u32 g, onc_u32 first, bool split,
                                                onc_u32* err, onc_u32 code, onc_u32* sum_lo, onc_u32* sum_hi) {
    const int lane = threadIdx.x & 63;
    onc_u32 lo = 0u, hi = 0u, j = g - 1u, spins = 0u;
    for (;;) {
        const onc_u64 v = (j >= first + (onc_u32)lane) ? ld_relaxed64(&status[j - (onc_u32)lane]) : ONC_F64_P;
        const onc_u64 fl = v >> 62;
        const onc_u64 notready = __ballot(fl == 0ull);
        const onc_u64 isP = __ballot(fl == 2ull);
        onc_u64 win;
        if (isP) {
            const int p = __ffsll((unsigned long long)isP) - 1;
            win = (p == 63) ? ~0ull : ((2ull << p) - 1ull);
        } else {
            win = ~0ull;
        }
        if (notready & win) {
            if (++spins > ONC_SPIN_LIMIT) {
                if (lane == 0) atomicOr(err, code);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        const bool take = (win >> lane) & 1ull;
        onc_u32 a_ = take ? (split ? (onc_u32)(v & ONC_M31) : (onc_u32)v) : 0u;
        onc_u32 b_ = take && split ? (onc_u32)((v >> 31) & ONC_M31) : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            a_ += __shfl_xor(a_, o, 64);
            b_ += __shfl_xor(b_, o, 64);
        }
        lo += a_;
        hi += b_;
        if (isP) break;
        j -= 64u;
    }
    *sum_lo = lo;
    *sum_hi = hi;
}

// ---------------------------------------------------------------------------
// Step bookkeeping, one workgroup for the whole batch, at the start of every
// step (and, in commit mode, before the host reads a sim). It first closes the
// step that ran last for every sim that took part in it: the live count
// becomes survivors + the daughters that fit the capacity (the daughters stay
// in dau2 after the survivors, see onc_dstate), ids advance, the buffers flip.
// In step mode it then plans the step: which sims are frozen at their stop
// step, the prefixes of live cells and of sort / compaction tiles, the
// per-step scalars of the segment table, and cleared tile counters. (The
// digit histograms are cleared by the permute kernel, once the sort passes
// have read them.)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_plan(onc_kargs a, int mode, onc_dsim* table) {
    __shared__ onc_u32 s_tmp[8];
    __shared__ onc_u32 s_carry[3];
    const int tid = threadIdx.x;
    if (tid < 3) s_carry[tid] = 0u;
    __syncthreads();
    // The plan is the table's only device writer. It writes through `table`
    // (the same memory as a.sims): the other kernels read a.sims as constant
    // memory, and stores derived from a constant-memory pointer are undefined.
    onc_dsim* T = table;
    for (onc_u32 b0 = 0; b0 < a.nsims; b0 += 256) {
        const onc_u32 s = b0 + tid;
        onc_u32 nt = 0, nc = 0, nct = 0;
        if (s < a.nsims) {
            onc_dstate* st = T[s].state;
            if (st->pending) {
                const onc_u32 ns = st->nsurv, nd = st->ndiv, room = T[s].cap - ns;
                const onc_u32 avail = nd < room ? nd : room;
                st->cell_updates += st->n;
                st->n = ns + avail;
                st->split = ns;
                st->next_id += avail;
                st->parity ^= 1u;
                st->step += 1u;
                st->pending = 0u;
            }
            const bool frozen = st->step >= st->stop;
            T[s].n_live = frozen ? 0u : st->n;
            T[s].par = st->parity;
            T[s].step = st->step;
            T[s].next_id = st->next_id;
            T[s].split = st->split;
            T[s].frozen = frozen ? 1u : 0u;
            if (mode == ONC_PLAN_STEP && !frozen) {
                st->nsurv = 0u;
                st->ndiv = 0u;
                st->pending = 1u;
                nc = st->n;
                nt = (nc + a.tile - 1u) / a.tile;
                nct = (nc + a.ctile - 1u) / a.ctile;
            }
        }
        if (mode == ONC_PLAN_STEP) {
            onc_u32 tt, tc, tct;
            const onc_u32 et = block_exscan256(nt, s_tmp, &tt);
            const onc_u32 ec = block_exscan256(nc, s_tmp, &tc);
            const onc_u32 ect = block_exscan256(nct, s_tmp, &tct);
            if (s < a.nsims) {
                a.tile_prefix[s] = s_carry[0] + et;
                a.cell_prefix[s] = s_carry[1] + ec;
                a.ctile_prefix[s] = s_carry[2] + ect;
            }
            __syncthreads();
            if (tid == 0) {
                s_carry[0] += tt;
                s_carry[1] += tc;
                s_carry[2] += tct;
            }
            __syncthreads();
        }
    }
    if (mode == ONC_PLAN_STEP) {
        for (int i = tid; i < ONC_NCOUNTERS; i += 256) a.counters[i] = 0u;
        if (tid == 0) {
            a.tile_prefix[a.nsims] = s_carry[0];
            a.cell_prefix[a.nsims] = s_carry[1];
            a.ctile_prefix[a.nsims] = s_carry[2];
        }
    }
}

// ---------------------------------------------------------------------------
// Morton keys + digit histograms of every radix pass (one read of positions).
// Each block takes an equal contiguous share of the batch's live cells, in
// 256-cell chunks, so a small population still spreads over the whole chip;
// it counts digits in LDS for its current sim and adds the counts to that
// sim's histogram when its range moves on to the next sim (cells of a chunk
// that straddles two sims count straight into global memory). The kernel also
// clears the look-back words of this step's sort and compaction tiles.
// RUNS: a wave's cells are consecutive and nearly sorted, so its higher
// digits mostly agree; the first lane of each run of equal digits adds the
// run's length instead of every lane adding 1 to the same LDS word (at 1.9M
// cells 22.8 -> 20.3 us; at 80K cells 9.4 -> 11.5 us, so the runtime takes it
// for large populations only).
// ---------------------------------------------------------------------------
#define ONC_MORTON_CHUNKS 8
template <bool RUNS>
__global__ __launch_bounds__(256) void k_morton(onc_kargs a) {
    __shared__ onc_u32 s_hist[ONC_DEV_MAX_PASSES * 256];
    const int tid = threadIdx.x;
    const onc_u32 P = a.passes;
    const onc_u32 total = a.cell_prefix[a.nsims];
    const onc_u32 nch = (total + 255u) >> 8;
    const onc_u32 blk = onc_uniform(onc_xcd_block(blockIdx.x, gridDim.x));
    const onc_u32 c0 = (onc_u32)(((unsigned long long)nch * blk) / gridDim.x);
    const onc_u32 c1 = (onc_u32)(((unsigned long long)nch * (blk + 1u)) / gridDim.x);
    for (onc_u32 i = tid; i < P * 256u; i += 256) s_hist[i] = 0u;
    __syncthreads();
    // this block's copy of the digit histograms (blocks spread their device
    // atomics over ONC_HIST_COPIES copies; the sort passes add the copies)
    onc_u32* hist_copy = a.hist + (size_t)(blockIdx.x % ONC_HIST_COPIES) * a.nsims * 4u * 256u;
    auto flush = [&](int s) {
        __syncthreads();
        for (onc_u32 i = tid; i < P * 256u; i += 256) {
            const onc_u32 v = s_hist[i];
            if (v) atomicAdd(&hist_copy[(size_t)s * 4u * 256u + i], v);
            s_hist[i] = 0u;
        }
        __syncthreads();
    };
    auto count = [&](int s, int cur, onc_u32 key) {
        if (s == cur) {
            for (onc_u32 p = 0; p < P; p++) atomicAdd(&s_hist[p * 256u + ((key >> (8u * p)) & 255u)], 1u);
        } else {
            for (onc_u32 p = 0; p < P; p++)
                atomicAdd(&hist_copy[((size_t)s * 4u + p) * 256u + ((key >> (8u * p)) & 255u)], 1u);
        }
    };
    int cur = -1;
    for (onc_u32 c = c0; c < c1; c += ONC_MORTON_CHUNKS) {
        const onc_u32 ce = min(c + (onc_u32)ONC_MORTON_CHUNKS, c1);
        const onc_u32 base = c << 8;
        const onc_u32 last = min(ce << 8, total) - 1u;
        const int sa = onc_find_seg_wave(a.cell_prefix, a.nsims, base);
        const int sb = onc_find_seg_wave(a.cell_prefix, a.nsims, last);
        if (sa == sb) {
            // up to eight chunks of one sim: all position loads in flight at once
            if (sa != cur) {
                if (cur >= 0) flush(cur);
                cur = sa;
            }
            const onc_dsim& S = a.sims[sa];
            const onc_u32 lo = a.cell_prefix[sa];
            const onc_u32 par = S.par, split = S.split;
            float x[ONC_MORTON_CHUNKS], y[ONC_MORTON_CHUNKS], z[ONC_MORTON_CHUNKS];
#pragma unroll
            for (int q = 0; q < ONC_MORTON_CHUNKS; q++) {
                const onc_u32 g = base + (onc_u32)q * 256u + (onc_u32)tid;
                x[q] = y[q] = z[q] = 0.0f;
                if (g <= last) {
                    const onc_u32 i = g - lo;
                    const bool d = i >= split;
                    const onc_u32 j = d ? i - split : i;
                    x[q] = ((const float*)(d ? S.dau2[0] : S.st[par][0]))[j];
                    y[q] = ((const float*)(d ? S.dau2[1] : S.st[par][1]))[j];
                    z[q] = ((const float*)(d ? S.dau2[2] : S.st[par][2]))[j];
                }
            }
#pragma unroll
            for (int q = 0; q < ONC_MORTON_CHUNKS; q++) {
                const onc_u32 g = base + (onc_u32)q * 256u + (onc_u32)tid;
                if (!RUNS) {
                    if (g <= last) {
                        const onc_u32 i = g - lo;
                        const onc_u32 key = onc_cell_key(S, x[q], y[q], z[q]);
                        S.keys[0][i] = key;
                        S.vals[0][i] = i;
                        count(sa, cur, key);
                    }
                    continue;
                }
                const bool valid = g <= last;
                onc_u32 key = 0u;
                if (valid) {
                    const onc_u32 i = g - lo;
                    key = onc_cell_key(S, x[q], y[q], z[q]);
                    S.keys[0][i] = key;
                    S.vals[0][i] = i;
                }
                // digit counts once per run of equal digits in the wave (the
                // valid lanes are a prefix of the wave: g grows with the lane)
                const int L = (int)(threadIdx.x & 63u);
                const onc_u64 vmask = __ballot(valid);
                for (onc_u32 p = 0; p < P; p++) {
                    const onc_u32 d = (key >> (8u * p)) & 255u;
                    const onc_u32 dp = __shfl_up(d, 1, 64);
                    const bool start = valid && (L == 0 || dp != d);
                    const onc_u64 bounds = __ballot(start) | ~vmask;
                    if (start) {
                        const onc_u64 above = L == 63 ? 0ull : (bounds & ~((2ull << L) - 1ull));
                        const onc_u32 len = (above != 0ull ? (onc_u32)__builtin_ctzll(above) : 64u) - (onc_u32)L;
                        atomicAdd(&s_hist[p * 256u + d], len);
                    }
                }
            }
        } else {
            // chunks across a sim boundary: one chunk at a time
            for (onc_u32 cc = c; cc < ce; cc++) {
                const onc_u32 b0 = cc << 8;
                const int s0 = onc_find_seg_wave(a.cell_prefix, a.nsims, b0);
                if (s0 != cur) {
                    if (cur >= 0) flush(cur);
                    cur = s0;
                }
                const onc_u32 g = b0 + (onc_u32)tid;
                if (g < total) {
                    int s = s0;
                    if (g >= a.cell_prefix[s0 + 1]) s = onc_find_seg(a.cell_prefix, a.nsims, g);
                    const onc_dsim& S = a.sims[s];
                    const onc_u32 i = g - a.cell_prefix[s];
                    const bool d = i >= S.split;
                    const onc_u32 j = d ? i - S.split : i;
                    onc_gcol const* src = d ? S.dau2 : S.st[S.par];
                    const onc_u32 key = onc_cell_key(S, ((const float*)src[0])[j], ((const float*)src[1])[j],
                                                     ((const float*)src[2])[j]);
                    S.keys[0][i] = key;
                    S.vals[0][i] = i;
                    count(s, cur, key);
                }
            }
        }
    }
    if (cur >= 0) flush(cur);
    // look-back words of this step's tiles
    const onc_u32 nthreads = gridDim.x * 256u;
    const onc_u32 gid = blockIdx.x * 256u + (onc_u32)tid;
    const onc_u32 ts = a.tile_prefix[a.nsims] * 256u;
    for (onc_u32 p = 0; p < P; p++)
        for (onc_u32 i = gid; i < ts; i += nthreads) a.sort_status[(size_t)p * a.max_tiles * 256u + i] = 0u;
    const onc_u32 tc = a.ctile_prefix[a.nsims];
    for (onc_u32 i = gid; i < tc; i += nthreads) a.comp_status[i] = 0ull;
    // one slice of every sim's bin table back to the fill value (no kernel
    // reads the bins between the last step's rule and this step's permute)
    for (int s = 0; s < a.nsims; s++) {
        const onc_dsim& S = a.sims[s];
        const onc_u32 len = (S.table + ONC_BIN_SLICES - 1u) / ONC_BIN_SLICES;
        const onc_u32 b0 = (S.step % ONC_BIN_SLICES) * len;
        const onc_u32 b1 = min(b0 + len, S.table);
        for (onc_u32 i = b0 + gid; i < b1; i += nthreads)
            ((uint2*)S.bins)[i] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
    }
}

// Histograms only (keys already in keys[0]); used by onc_sort_pairs. A block
// accumulates the counts of all its tiles of one sim in LDS and adds them to
// the device histograms once, into copy blockIdx % ONC_HIST_COPIES: flushing
// per tile would put every tile's 256 x passes atomics on the same few
// hundred words, which serialise at their memory channel.
template <int ITEMS>
__global__ __launch_bounds__(256) void k_hist_only(onc_kargs a) {
    __shared__ onc_u32 s_hist[ONC_DEV_MAX_PASSES * 256];
    const int tid = threadIdx.x;
    const onc_u32 total = a.tile_prefix[a.nsims];
    const onc_u32 P = a.passes;
    const onc_u32 copy = blockIdx.x % ONC_HIST_COPIES;
    for (onc_u32 i = tid; i < P * 256u; i += 256) s_hist[i] = 0u;
    __syncthreads();
    int cur = -1;
    auto flush = [&](int s) {
        __syncthreads();
        for (onc_u32 p = 0; p < P; p++) {
            const onc_u32 v = s_hist[p * 256u + tid];
            if (v) atomicAdd(&a.hist[(((size_t)copy * a.nsims + (size_t)s) * 4u + p) * 256u + tid], v);
            s_hist[p * 256u + tid] = 0u;
        }
        __syncthreads();
    };
    for (onc_u32 g = blockIdx.x; g < total; g += gridDim.x) {
        const int s = onc_find_seg_wave(a.tile_prefix, a.nsims, g);
        if (s != cur) {
            if (cur >= 0) flush(cur);
            cur = s;
        }
        const onc_dsim& S = a.sims[s];
        const onc_u32 n = S.n_live;
        const onc_u32 base = (g - a.tile_prefix[s]) * (onc_u32)(256 * ITEMS);
        const onc_u32* __restrict__ kin = S.keys[0];
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const onc_u32 i = base + (onc_u32)(k * 256 + tid);
            if (i < n) {
                const onc_u32 key = __builtin_nontemporal_load(kin + i);
                for (onc_u32 p = 0; p < P; p++) atomicAdd(&s_hist[p * 256u + ((key >> (8u * p)) & 255u)], 1u);
            }
        }
        for (onc_u32 p = 0; p < P; p++) a.sort_status[((size_t)p * a.max_tiles + g) * 256u + tid] = 0u;
    }
    if (cur >= 0) flush(cur);
}

// ---------------------------------------------------------------------------
// One stable LSD radix pass (8-bit digit), onesweep style: wave64 ballot
// multi-split ranks each digit inside the tile in input order, a decoupled
// look-back over the sim's earlier tiles gives the tile's global offset per
// digit, and an LDS reorder makes the scatter write runs of equal digits.
// ---------------------------------------------------------------------------
template <int ITEMS>
__global__ __launch_bounds__(256) void k_sort_pass(onc_kargs a, int pass) {
    constexpr int TILE = 256 * ITEMS;
    __shared__ onc_u32 s_keys[TILE];
    __shared__ onc_u32 s_vals[TILE];
    __shared__ onc_u32 s_wc[4 * 256];
    __shared__ onc_u32 s_texcl[256];
    __shared__ onc_u32 s_gbase[256];
    __shared__ onc_u32 s_tmp[8];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const onc_u32 total = a.tile_prefix[a.nsims];
    const onc_u32 shift = 8u * (onc_u32)pass;
    const onc_u64 lt_mask = (1ull << lane) - 1ull;
    onc_u32* status = a.sort_status + (size_t)pass * a.max_tiles * 256u;
    // Tile order. The launch has one block per tile, so the loop below runs
    // at most once per block. When the step's tiles are no more than the
    // blocks the chip holds at once (a.xcd_tiles), every tile's block is
    // resident together and any order lets every look-back wait end: then
    // block b takes tile onc_xcd_block(b, total), so each XCD ranks and
    // scatters one contiguous range of tiles, and the neighbouring runs of a
    // digit that consecutive tiles write share lines in one L2 (tiles
    // round-robin over the XCDs write every line back as eight partial
    // copies). Otherwise block b takes tile b: a tile only waits on lower
    // tiles, whose blocks were dispatched before it, so the lowest unfinished
    // tile can always run. (A shared tile counter would also give that order,
    // but same-address device atomics serialise at one memory channel and
    // stall every other access queued behind them.)
    const bool xmap = total <= a.xcd_tiles;
    for (onc_u32 bb = blockIdx.x; bb < total; bb += gridDim.x) {
        const onc_u32 g = xmap ? onc_uniform(onc_xcd_block(bb, total)) : bb;
        for (int i = tid; i < 4 * 256; i += 256) s_wc[i] = 0u;
        __syncthreads();
        const int s = onc_find_seg_wave(a.tile_prefix, a.nsims, g);
        const onc_dsim& S = a.sims[s];
        const onc_u32 t = g - a.tile_prefix[s];
        const onc_u32 n = S.n_live;
        const onc_u32 base = t * (onc_u32)TILE;
        const onc_u32 cnt = min((onc_u32)TILE, n - base);
        const onc_u32* kin = S.keys[pass & 1];
        const onc_u32* vin = S.vals[pass & 1];
        onc_u32* kout = S.keys[(pass + 1) & 1];
        onc_u32* vout = S.vals[(pass + 1) & 1];

        // keys and values are read once: streaming loads
        onc_u32 key[ITEMS], val[ITEMS], rnk[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            onc_u32 idx = (onc_u32)(w * 64 * ITEMS + k * 64 + lane);
            key[k] = 0u;
            val[k] = 0u;
            if (idx < cnt) {
                key[k] = __builtin_nontemporal_load(kin + base + idx);
                val[k] = __builtin_nontemporal_load(vin + base + idx);
            }
        }
        // rank inside the wave, items in (k, lane) order = input order
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            onc_u32 idx = (onc_u32)(w * 64 * ITEMS + k * 64 + lane);
            bool valid = idx < cnt;
            onc_u32 d = (key[k] >> shift) & 255u;
            onc_u64 m = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; b++) {
                bool bit = (d >> b) & 1u;
                onc_u64 bb = __ballot(bit);
                m &= bit ? bb : ~bb;
            }
            onc_u32 pre = s_wc[w * 256 + d];
            onc_u32 r = (onc_u32)__popcll(m & lt_mask);
            rnk[k] = pre + r;
            if (valid && r == 0u) s_wc[w * 256 + d] = pre + (onc_u32)__popcll(m);
        }
        __syncthreads();
        // per digit (thread = digit): tile count, published at once so that
        // later tiles can look past this one, then the wave prefixes
        const onc_u32 c0 = s_wc[tid], c1 = s_wc[256 + tid], c2 = s_wc[512 + tid], c3 = s_wc[768 + tid];
        const onc_u32 tot = c0 + c1 + c2 + c3;
        onc_u32* mine = &status[(size_t)g * 256u + tid];
        st_relaxed(mine, (t == 0u ? ONC_FLAG_P : ONC_FLAG_A) | tot);
        s_wc[tid] = 0u;
        s_wc[256 + tid] = c0;
        s_wc[512 + tid] = c0 + c1;
        s_wc[768 + tid] = c0 + c1 + c2;
        const onc_u32 texcl = block_exscan256(tot, s_tmp, nullptr);
        s_texcl[tid] = texcl;
        onc_u32 hd = 0u;
#pragma unroll
        for (int c = 0; c < ONC_HIST_COPIES; c++)
            hd += a.hist[(((size_t)c * a.nsims + (size_t)s) * 4u + (onc_u32)pass) * 256u + tid];
        const onc_u32 gex = block_exscan256(hd, s_tmp, nullptr);  // ends with a barrier
        // reorder through LDS by digit before the look-back: the positions
        // inside the tile need no other tile, and the keys and values leave
        // the registers that the look-back's window of loads then uses
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            onc_u32 idx = (onc_u32)(w * 64 * ITEMS + k * 64 + lane);
            if (idx < cnt) {
                onc_u32 d = (key[k] >> shift) & 255u;
                onc_u32 pos = s_texcl[d] + s_wc[w * 256 + d] + rnk[k];
                s_keys[pos] = key[k];
                s_vals[pos] = val[k];
            }
        }
        // decoupled look-back over the sim's earlier tiles for digit tid
        onc_u32 prefix = 0u;
        if (t != 0u) {
            prefix = lookback_digit<ONC_SORT_LOOKBACK>(status, g, g - t, (onc_u32)tid, a.err);
            st_relaxed(mine, ONC_FLAG_P | (prefix + tot));
        }
        s_gbase[tid] = gex + prefix;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            onc_u32 j = (onc_u32)(k * 256 + tid);
            if (j < cnt) {
                onc_u32 kk = s_keys[j];
                onc_u32 d = (kk >> shift) & 255u;
                onc_u32 o = s_gbase[d] + (j - s_texcl[d]);
                kout[o] = kk;
                vout[o] = s_vals[j];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Radix pass in three launches, for large standalone sorts: per-tile digit
// counts (k_sort_up), an exclusive scan of the counts in digit-major order
// (k_exscan; entry d * T + t becomes the output position of tile t's first
// key of digit d), then rank and scatter (k_sort_down). No tile waits on
// another: on a chip whose L2 is split across 8 XCDs every look-back round
// trip goes through the memory-side cache, and a chain of them bounded the
// one-launch pass. Costs one more read of the keys (4 B/key) and 8 B per
// (tile, digit) of counts and offsets. With no waits between tiles, each XCD
// takes a contiguous range of tiles: the counts and offsets a digit row holds
// for neighbouring tiles then share cache lines in one L2 (instead of eight
// partial copies of each line), and so do the neighbouring runs of a digit
// in the scatter.
// ---------------------------------------------------------------------------
template <int ITEMS>
__global__ __launch_bounds__(256) void k_sort_up(const onc_u32* __restrict__ keys, onc_u32 n, onc_u32 shift,
                                                 onc_u32* __restrict__ counts, onc_u32 T) {
    constexpr int TILE = 256 * ITEMS;
    // one 256-bin LDS histogram per wave (fewer lanes on a bin's word); the
    // tile's loads are all issued before the first count (3 % faster passes
    // than a load and its four counts at a time)
    constexpr int NH = 4;
    __shared__ onc_u32 s_h[NH * 256];
    const int tid = threadIdx.x;
    onc_u32* h = s_h + (tid >> 6) * 256;
    const onc_u32 t = onc_xcd_block(blockIdx.x, gridDim.x);  // each XCD a contiguous range of tiles
#pragma unroll
    for (int c = 0; c < NH; c++) s_h[c * 256 + tid] = 0u;
    __syncthreads();
    const onc_u32 base = t * (onc_u32)TILE;
    if (base + (onc_u32)TILE <= n) {
        // whole tile: 16-byte loads (the key buffers are 256-byte aligned)
        typedef onc_u32 v4u __attribute__((ext_vector_type(4)));
        v4u q[ITEMS / 4];
#pragma unroll
        for (int k = 0; k < ITEMS / 4; k++) q[k] = __builtin_nontemporal_load((const v4u*)(keys + base) + k * 256 + tid);
#pragma unroll
        for (int k = 0; k < ITEMS / 4; k++) {
            atomicAdd(&h[(q[k].x >> shift) & 255u], 1u);
            atomicAdd(&h[(q[k].y >> shift) & 255u], 1u);
            atomicAdd(&h[(q[k].z >> shift) & 255u], 1u);
            atomicAdd(&h[(q[k].w >> shift) & 255u], 1u);
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const onc_u32 i = base + (onc_u32)(k * 256 + tid);
            if (i < n) atomicAdd(&h[(__builtin_nontemporal_load(keys + i) >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    onc_u32 c = 0u;
#pragma unroll
    for (int w = 0; w < NH; w++) c += s_h[w * 256 + tid];
    counts[(size_t)tid * T + t] = c;
}

template <int ITEMS>
__global__ __launch_bounds__(256) void k_sort_down(const onc_u32* __restrict__ kin, const onc_u32* __restrict__ vin,
                                                   onc_u32* __restrict__ kout, onc_u32* __restrict__ vout, onc_u32 n,
                                                   onc_u32 shift, const onc_u32* __restrict__ offsets, onc_u32 T) {
    constexpr int TILE = 256 * ITEMS;
    __shared__ __attribute__((aligned(16))) onc_u32 s_keys[TILE];
    __shared__ __attribute__((aligned(16))) onc_u32 s_vals[TILE];
    __shared__ onc_u32 s_wc[4 * 256];
    __shared__ onc_u32 s_texcl[256];
    __shared__ onc_u32 s_gbase[256];
    __shared__ onc_u32 s_tmp[8];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const onc_u64 lt_mask = (1ull << lane) - 1ull;
    const onc_u32 t = onc_xcd_block(blockIdx.x, gridDim.x);  // each XCD a contiguous range of tiles
    const onc_u32 base = t * (onc_u32)TILE;
    const onc_u32 cnt = min((onc_u32)TILE, n - base);
    // this tile's output base per digit (thread = digit), loaded first
    const onc_u32 gbase = offsets[(size_t)tid * T + t];
    for (int i = tid; i < 4 * 256; i += 256) s_wc[i] = 0u;
    onc_u32 key[ITEMS], val[ITEMS], rnk[ITEMS];
    if (cnt == (onc_u32)TILE) {
        // whole tile: 16-byte streaming loads into LDS in input order (the
        // pair buffers are 256-byte aligned), then each thread's items back
        typedef onc_u32 v4u __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int k = 0; k < ITEMS / 4; k++) {
            const int e = (k * 256 + tid) * 4;
            const v4u a = __builtin_nontemporal_load((const v4u*)(kin + base + e));
            const v4u b = __builtin_nontemporal_load((const v4u*)(vin + base + e));
            *(uint4*)&s_keys[e] = make_uint4(a.x, a.y, a.z, a.w);
            *(uint4*)&s_vals[e] = make_uint4(b.x, b.y, b.z, b.w);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const int idx = w * 64 * ITEMS + k * 64 + lane;
            key[k] = s_keys[idx];
            val[k] = s_vals[idx];
        }
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const onc_u32 idx = (onc_u32)(w * 64 * ITEMS + k * 64 + lane);
            key[k] = 0u;
            val[k] = 0u;
            if (idx < cnt) {
                key[k] = __builtin_nontemporal_load(kin + base + idx);
                val[k] = __builtin_nontemporal_load(vin + base + idx);
            }
        }
    }
    __syncthreads();
    // rank inside the wave, items in (k, lane) order = input order
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const onc_u32 idx = (onc_u32)(w * 64 * ITEMS + k * 64 + lane);
        const bool valid = idx < cnt;
        const onc_u32 d = (key[k] >> shift) & 255u;
        onc_u64 m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const onc_u64 bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        const onc_u32 pre = s_wc[w * 256 + d];
        const onc_u32 r = (onc_u32)__popcll(m & lt_mask);
        rnk[k] = pre + r;
        if (valid && r == 0u) s_wc[w * 256 + d] = pre + (onc_u32)__popcll(m);
    }
    __syncthreads();
    const onc_u32 c0 = s_wc[tid], c1 = s_wc[256 + tid], c2 = s_wc[512 + tid], c3 = s_wc[768 + tid];
    s_wc[tid] = 0u;
    s_wc[256 + tid] = c0;
    s_wc[512 + tid] = c0 + c1;
    s_wc[768 + tid] = c0 + c1 + c2;
    s_texcl[tid] = block_exscan256(c0 + c1 + c2 + c3, s_tmp, nullptr);
    s_gbase[tid] = gbase;
    __syncthreads();
    // reorder through LDS by digit, then scatter runs of equal digits
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const onc_u32 idx = (onc_u32)(w * 64 * ITEMS + k * 64 + lane);
        if (idx < cnt) {
            const onc_u32 d = (key[k] >> shift) & 255u;
            const onc_u32 pos = s_texcl[d] + s_wc[w * 256 + d] + rnk[k];
            s_keys[pos] = key[k];
            s_vals[pos] = val[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const onc_u32 j = (onc_u32)(k * 256 + tid);
        if (j < cnt) {
            const onc_u32 kk = s_keys[j];
            const onc_u32 d = (kk >> shift) & 255u;
            const onc_u32 o = s_gbase[d] + (j - s_texcl[d]);
            kout[o] = kk;
            vout[o] = s_vals[j];
        }
    }
}

// Fixed (Dirichlet) voxel predicate (engine/assets/shaders/tumor_sim.wgsl:420-423).
__device__ __forceinline__ bool onc_is_fixed(const onc_dfield& f, const int* d, int x, int y, int z) {
    if (f.bc == 0) return x == 0 || y == 0 || z == 0 || x == d[0] - 1 || y == d[1] - 1 || z == d[2] - 1;
    if (f.bc != 1 || z != 0) return false;
    float fx = ((float)x + 0.5f) - (float)d[0] * 0.5f;
    float fy = ((float)y + 0.5f) - (float)d[1] * 0.5f;
    float r = f.src_radius;
    return fx * fx + fy * fy <= r * r;
}

// Quiet-tile states (onc_kargs::quiet)
#define ONC_Q_NONE 0u      // not known to be uniform
#define ONC_Q_COMPUTED 1u  // computed, and every output came out equal
#define ONC_Q_FILLED 2u    // written as a uniform tile without reading the input
#define ONC_Q_KEPT 3u      // already held the value: neither read nor written

// ---------------------------------------------------------------------------
// Gather every state column into sorted order (current -> other buffer),
// write bin start/end of every occupied key, and apply the per-voxel uptake
// once per occupied voxel (tumor_sim.wgsl:387-412 reaction, before the sweep).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_permute_bins(onc_kargs a) {
    const onc_u32 P = a.passes;
    // the sort passes are done with this step's digit histograms: clear them
    // for the next step's Morton kernel
    for (onc_u32 i = blockIdx.x * 256u + threadIdx.x; i < ONC_HIST_COPIES * a.nsims * 4u * 256u;
         i += gridDim.x * 256u)
        a.hist[i] = 0u;
    onc_for_cells(a.cell_prefix, a.nsims, [&](int s, onc_u32 i) {
        const onc_dsim& S = a.sims[s];
        const onc_u32 n = S.n_live;
        const onc_u32 par = S.par;
        const onc_u32* __restrict__ keys = S.keys[P & 1u];
        const onc_u32 src0 = S.vals[P & 1u][i];
        // committed cells: survivors in st[par], then daughters in dau2
        const bool fromd = src0 >= S.split;
        const onc_u32 src = fromd ? src0 - S.split : src0;
        onc_gcol const* from = fromd ? S.dau2 : S.st[par];
        const onc_u32 k = keys[i];
        const onc_u32 kprev = i > 0u ? keys[i - 1u] : ~k;
        // gather in chunks of eight columns: loads first, then stores
        float px = 0.0f, py = 0.0f, pz = 0.0f;
        for (int c0 = 0; c0 < S.n_cols; c0 += 8) {
            onc_u32 v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = (c0 + q < S.n_cols) ? from[c0 + q][src] : 0u;
            if (c0 == 0) {
                px = __uint_as_float(v[0]);
                py = __uint_as_float(v[1]);
                pz = __uint_as_float(v[2]);
                ((float4*)S.pos4)[i] = make_float4(px, py, pz, 0.0f);
            }
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (c0 + q < S.n_cols) S.st[par ^ 1u][c0 + q][i] = v[q];
        }
        // Stamp the cube of each occupied voxel when the sim takes uptake (the
        // sweep's quiet tiles read the stamps; a voxel whose uptake the fixed
        // boundary skips is stamped too, which only keeps its tile swept):
        // the lanes of one call hold consecutive cells (lanes 0..last active),
        // so a cube's cells form a stretch of lanes and its first lane stores.
        if (a.quiet) {
            bool up = false;
            for (int f = 0; f < S.n_fields; f++) up = up || S.fields[f].uptake != 0.0f;
            if (up) {
                const onc_u32 cube = k >> (3 * ONC_QCUBE);
                const onc_u32 cprev = __shfl_up(cube, 1);
                if ((threadIdx.x & 63u) == 0u || cprev != cube) S.qstamp[cube] = S.step + 1u;
            }
        }
        // run lengths of equal keys from the wave's run starts: the lanes of
        // one call hold consecutive cells (lanes 0..last active), and a sim
        // boundary is itself a run start
        const onc_u64 act = __ballot(true);
        const onc_u64 starts = __ballot(kprev != k);
        if (kprev != k) {
            // the run ends at the next start in the wave, or continues past
            // the wave's last cell (then counted on from there)
            const int L = (int)(threadIdx.x & 63u);
            const onc_u64 above = L == 63 ? 0ull : (starts & ~((2ull << L) - 1ull));
            onc_u32 cnt;
            if (above != 0ull) {
                cnt = (onc_u32)(__builtin_ctzll(above) - L);
            } else {
                cnt = (onc_u32)(63 - __builtin_clzll(act) - L + 1);
                while (i + cnt < n && keys[i + cnt] == k) cnt++;
            }
            // the whole bin entry {start, len << 16 | tag} in one 8-byte store
            // (a longer bin is a device error; its length is clamped)
            ((uint2*)S.bins)[k] = make_uint2(i, (min(cnt, ONC_MAX_BIN_CELLS) << 16) | onc_bin_tag(S.step));
            if (cnt > ONC_MAX_BIN_CELLS) atomicOr(a.err, ONC_DERR_BIG_BIN);
            // uptake, once per occupied voxel, before the sweep
            const int vx = onc_vox1(px, S.dx, S.dims[0]), vy = onc_vox1(py, S.dx, S.dims[1]),
                      vz = onc_vox1(pz, S.dx, S.dims[2]);
            const size_t v = (size_t)vx + (size_t)S.dims[0] * ((size_t)vy + (size_t)S.dims[1] * (size_t)vz);
            for (int f = 0; f < S.n_fields; f++) {
                const onc_dfield& fd = S.fields[f];
                if (fd.uptake == 0.0f || onc_is_fixed(fd, S.dims, vx, vy, vz)) continue;
                float* u = S.fld[par][f];
                float w = u[v] - (float)cnt * fd.uptake;
                u[v] = w < fd.lo ? fd.lo : w;
            }
        }
    });
}


// ---------------------------------------------------------------------------
// Neighbour forces over the 27 Moore bins (engine/assets/shaders/
// tumor_sim.wgsl:505-628 structure; linear spring, linear adhesion taper)
// and explicit overdamped integration. Reads no field, so it runs beside the
// sweep.
// ---------------------------------------------------------------------------
#define ONC_FORCES_BLOCK 64  // one wave per block

// Candidates of the 27 Moore bins around (px, py, pz), z-plane by z-plane, in
// template.wgsl:11-21 order (x fastest) and bin order inside a bin;
// `pair(o, j)` gets each candidate's packed position, `self` included (and
// named again by spent slots): its d2 is exactly 0, which the pair filter
// rejects as it rejects any coincident cell. Per plane the
// entries of the plane's reachable bins are loaded together, the non-empty
// ranges are compacted into the thread's LDS slots, and a cursor walks them
// ONC_WALK_BATCH candidates at a time (their position loads are in flight
// together). `batch_end()` runs after every batch.
//
// A bin is reachable when the distance from (px, py, pz) to the nearest
// point of its voxel can be below R: face neighbours always, edge and corner
// neighbours only when the cell sits near that edge or corner (with dx = R,
// 20.6 of the 27 bins on average). A candidate in an unreachable bin is at
// least R from the cell, so the pair filter (d2 < R2) would reject it: the
// gaps are taken short by a margin (2^-20 of the grid's longest side, far
// above the rounding of the voxel assignment and of the gaps themselves) and
// their squared sum must reach R2 (1 + 2^-18), above the rounding of d2.
// Skipping such bins leaves every pair list, and so every force, unchanged.
//
// The cursor is branch-free: every slot reads the LDS range at the cursor's
// (clamped) next index and selects it when the current range ends, and a
// slot past the lane's last candidate names `self` (its load is a cached
// re-read of the cell's own position).
#ifndef ONC_WALK_BATCH
#define ONC_WALK_BATCH 4
#endif
#ifndef ONC_WALK_REACH
#define ONC_WALK_REACH 1
#endif
#ifndef ONC_WALK_SELECT
#define ONC_WALK_SELECT 1
#endif
#ifndef ONC_WALK_SENTINEL
#define ONC_WALK_SENTINEL 1
#endif
#define ONC_WALK_SLOTS (9 + ONC_WALK_SENTINEL)  // range slots per lane: 9 bins (+ the sentinel)
// A lane's pair list is flushed (wave-wide) once any lane holds more than
// ONC_PAIR_FLUSH pairs after a batch, so it never holds more than
// ONC_PAIR_FLUSH + ONC_WALK_BATCH.
#ifndef ONC_PAIR_FLUSH
#define ONC_PAIR_FLUSH 8
#endif
#define ONC_PAIR_LIST (ONC_PAIR_FLUSH + ONC_WALK_BATCH)
#ifndef ONC_FORCES_WPE
#define ONC_FORCES_WPE 5  // force-kernel waves per SIMD the register budget must allow
#endif
template <class F, class G>
__device__ __forceinline__ void forces_walk(const onc_dsim& S, onc_u32 tag, float px, float py, float pz,
                                            onc_u32 self, const float4* __restrict__ pos4, onc_u32 (*s_rs)[ONC_FORCES_BLOCK],
                                            unsigned short (*s_rl)[ONC_FORCES_BLOCK], F pair, G batch_end) {
    const int tid = threadIdx.x & (ONC_FORCES_BLOCK - 1);  // lane: LDS slots are per wave
    const int vx = onc_vox1(px, S.dx, S.dims[0]), vy = onc_vox1(py, S.dx, S.dims[1]),
              vz = onc_vox1(pz, S.dx, S.dims[2]);
    const uint2* __restrict__ bins = (const uint2*)S.bins;
    const onc_nkeys nk = onc_neighbor_keys(onc_morton((onc_u32)vx, (onc_u32)vy, (onc_u32)vz, S.key_format),
                                           S.key_format);
#if ONC_WALK_REACH
    // squared gaps to the lower (m) and upper (p) faces of the cell's voxel
    const float dx = S.dx;
    const float del = (float)max(S.dims[0], max(S.dims[1], S.dims[2])) * dx * 0x1p-20f;
    auto lo2 = [&](float p, int v) {
        const float g = (p - (float)v * dx) - del;
        return g > 0.0f ? g * g : 0.0f;
    };
    auto hi2 = [&](float p, int v) {
        const float g = ((float)(v + 1) * dx - p) - del;
        return g > 0.0f ? g * g : 0.0f;
    };
    const float gxm = lo2(px, vx), gxp = hi2(px, vx);
    const float gym = lo2(py, vy), gyp = hi2(py, vy);
    const float gzm = lo2(pz, vz), gzp = hi2(pz, vz);
    const float reach = (S.R * S.R) * (1.0f + 0x1p-18f);
#endif
    for (int oz = -1; oz <= 1; oz++) {
        const int nz = vz + oz;
        if (nz < 0 || nz >= S.dims[2]) continue;
#if ONC_WALK_REACH
        const float gz = oz < 0 ? gzm : (oz == 0 ? 0.0f : gzp);
        if (!(gz < reach)) continue;
#endif
        // a select, not nk.z[oz + 1]: a runtime index would put nk in scratch memory
        const onc_u32 zk = oz < 0 ? nk.z[0] : (oz == 0 ? nk.z[1] : nk.z[2]);
        uint2 b[9];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const int ny = vy + q / 3 - 1, nx = vx + q % 3 - 1;
            bool in = nx >= 0 && ny >= 0 && nx < S.dims[0] && ny < S.dims[1];
#if ONC_WALK_REACH
            const float gy = q / 3 == 0 ? gym : (q / 3 == 1 ? 0.0f : gyp);
            const float gx = q % 3 == 0 ? gxm : (q % 3 == 1 ? 0.0f : gxp);
            in = in && (gz + gy) + gx < reach;
#endif
            b[q] = make_uint2(0u, 0u);
            if (in) b[q] = bins[nk.x[q % 3] | nk.y[q / 3] | zk];
        }
        onc_u32 nr = 0;
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const onc_u32 len = onc_bin_len(b[q], tag);
            if (len > 0u) {
                s_rs[nr][tid] = b[q].x;
                s_rl[nr][tid] = (unsigned short)len;
                nr++;
            }
        }
#if ONC_WALK_SENTINEL
        // after the lane's last range, a range holding only the cell itself:
        // the cursor stays there, so a spent slot names `self` with no test
        s_rs[nr][tid] = self;
        s_rl[nr][tid] = 1;
        onc_u32 k = 0, j = s_rs[0][tid], e = j + s_rl[0][tid];
#else
        onc_u32 k = 0, j = 0, e = 0;
        if (nr > 0) {
            j = s_rs[0][tid];
            e = j + s_rl[0][tid];
        }
#endif
        while (k < nr) {
            onc_u32 jj[ONC_WALK_BATCH];
#pragma unroll
            for (int u = 0; u < ONC_WALK_BATCH; u++) {
#if ONC_WALK_SENTINEL
                jj[u] = j;
                const onc_u32 j1 = j + 1u;
                const bool end = j1 == e;
                k += end ? 1u : 0u;
                const onc_u32 kk = k < nr ? k : nr;
                const onc_u32 ns = s_rs[kk][tid], nl = s_rl[kk][tid];
                j = end ? ns : j1;
                e = end ? ns + nl : e;
#elif ONC_WALK_SELECT
                const bool v = k < nr;
                jj[u] = v ? j : self;
                const onc_u32 j1 = j + 1u;
                const bool end = v && j1 == e;
                k += end ? 1u : 0u;
                const onc_u32 kk = k < 8u ? k : 8u;
                const onc_u32 ns = s_rs[kk][tid], nl = s_rl[kk][tid];
                j = end ? ns : j1;
                e = end ? ns + nl : e;
#else
                jj[u] = k < nr ? j : self;
                if (k < nr) {
                    j++;
                    if (j == e) {
                        k++;
                        if (k < nr) {
                            j = s_rs[k][tid];
                            e = j + s_rl[k][tid];
                        }
                    }
                }
#endif
            }
            float4 p[ONC_WALK_BATCH];
#pragma unroll
            for (int u = 0; u < ONC_WALK_BATCH; u++) p[u] = pos4[jj[u]];
#pragma unroll
            for (int u = 0; u < ONC_WALK_BATCH; u++) pair(p[u], jj[u]);
            batch_end();
        }
    }
}

// LDS of one forces wave: per-thread list of interacting neighbours
// [slot][thread] (bank-conflict free) and the non-empty bin ranges of the
// current z-plane, followed by the sentinel range of the cell itself.
// Range lengths are 16-bit (a bin of more than 65535 cells is reported by the
// permute kernel as a device error). With 12-entry pair lists and ten range
// slots a wave takes 6.75 KB, so LDS holds 23 force waves per CU (their 78
// VGPRs would allow 24). The walk needs about 21: padding LDS to 18 and 15 waves per CU made
// the 1M-cell step 7 % and 11 % slower, while 24 ran as fast as 21, and 27
// or 28 (72 VGPRs with a small spill, earlier flushes) 2-5 % slower.
struct onc_forces_lds {
    onc_u32 pairs[ONC_PAIR_LIST][ONC_FORCES_BLOCK];
    onc_u32 rs[ONC_WALK_SLOTS][ONC_FORCES_BLOCK];
    unsigned short rl[ONC_WALK_SLOTS][ONC_FORCES_BLOCK];  // range lengths
};

// Forces of cell i of sim s through global memory (bins table, packed
// positions). Must be called by every lane of the wave that has a cell (the
// list flush is wave-wide). RAW: write the summed force to (rx, ry, rz)[i]
// instead of integrating the position (the onc_forces entry point).
template <bool RAW = false>
__device__ __forceinline__ void onc_forces_cell(const onc_kargs& a, int s, onc_u32 i, onc_forces_lds& L,
                                                float* rx = nullptr, float* ry = nullptr, float* rz = nullptr) {
    onc_u32 (*s_pairs)[ONC_FORCES_BLOCK] = L.pairs;
    onc_u32 (*s_rs)[ONC_FORCES_BLOCK] = L.rs;
    unsigned short (*s_rl)[ONC_FORCES_BLOCK] = L.rl;
    const int tid = threadIdx.x & (ONC_FORCES_BLOCK - 1);  // lane: LDS slots are per wave
    const onc_dsim& S = a.sims[s];
    const onc_u32 par = S.par;
    const onc_u32 tag = onc_bin_tag(S.step);
    const float4* __restrict__ p4 = (const float4*)S.pos4;
    const float4 me = p4[i];
    const float px = me.x, py = me.y, pz = me.z;
    const float R = S.R, Sc = 2.0f * S.cell_radius;
    const float R2 = R * R, RmS = R - Sc, krep = S.k_rep, kadh = S.k_adh;
    float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f;
    // Phase B: the pair law over the listed neighbours, in visiting order.
    onc_u32 npair = 0;
    auto law = [&](const float4& o) {
        const float ddx = px - o.x;
        const float ddy = py - o.y;
        const float ddz = pz - o.z;
        float d2 = ddx * ddx + ddy * ddy;
        d2 = d2 + ddz * ddz;
        const float dist = __builtin_sqrtf(d2);
        float f;
        if (dist < Sc) {
            f = krep * (Sc - dist);
        } else {
            if (kadh == 0.0f) return;
            f = -(kadh * ((R - dist) / RmS));
        }
        const float sc = f / dist;
        Fx = Fx + sc * ddx;
        Fy = Fy + sc * ddy;
        Fz = Fz + sc * ddz;
    };
    // the listed neighbours' positions are loaded four at a time, then
    // accumulated in list order
    auto flush = [&]() {
        onc_u32 k = 0;
        for (; k + 4u <= npair; k += 4u) {
            float4 o[4];
#pragma unroll
            for (int u = 0; u < 4; u++) o[u] = p4[s_pairs[k + u][tid]];
#pragma unroll
            for (int u = 0; u < 4; u++) law(o[u]);
        }
        for (; k < npair; k++) law(p4[s_pairs[k][tid]]);
        npair = 0;
    };
    // Phase A: cheap distance filter over the 27 bins; only pairs inside
    // the interaction radius reach the expensive (sqrt, divide) path, so
    // lanes diverge on short lists instead of on every candidate. A list
    // holds at most ONC_PAIR_FLUSH entries between batches (a batch adds at most ONC_WALK_BATCH);
    // once any lane of the wave holds more, every lane still walking
    // flushes together. A flush taken by one lane at a time would run
    // the wave through the pair law once per lane.
    forces_walk(S, tag, px, py, pz, i, p4, s_rs, s_rl,
                [&](const float4& o, onc_u32 j) {
                    const float ddx = px - o.x;
                    const float ddy = py - o.y;
                    const float ddz = pz - o.z;
                    float d2 = ddx * ddx + ddy * ddy;
                    d2 = d2 + ddz * ddz;
                    if (!(d2 < R2) || d2 == 0.0f) return;
                    s_pairs[npair][tid] = j;
                    npair++;
                },
                [&]() {
                    if (__ballot(npair > ONC_PAIR_FLUSH) != 0ull) flush();
                });
    flush();
    if (RAW) {
        rx[i] = Fx;
        ry[i] = Fy;
        rz[i] = Fz;
        return;
    }
    const float mob = S.mobility, md = S.max_disp;
    float dpx = mob * Fx, dpy = mob * Fy, dpz = mob * Fz;
    float l2 = dpx * dpx + dpy * dpy;
    l2 = l2 + dpz * dpz;
    if (l2 > md * md) {
        float sc = md / __builtin_sqrtf(l2);
        dpx = dpx * sc;
        dpy = dpy * sc;
        dpz = dpz * sc;
    }
    float* cx = (float*)S.st[par][0];
    float* cy = (float*)S.st[par][1];
    float* cz = (float*)S.st[par][2];
    cx[i] = onc_clampf(px + dpx, 0.0f, (float)S.dims[0] * S.dx);
    cy[i] = onc_clampf(py + dpy, 0.0f, (float)S.dims[1] * S.dx);
    cz[i] = onc_clampf(pz + dpz, 0.0f, (float)S.dims[2] * S.dx);
}

// Forces of the cells of the chunks of one-wave block `blk` of `nblk`
// (onc_for_chunks_xcd: blk % 8 is the block's XCD).
__device__ __forceinline__ void onc_forces_wave(const onc_kargs& a, onc_u32 blk, onc_u32 nblk,
                                                onc_forces_lds& L) {
    onc_for_chunks_xcd(a.cell_prefix, a.nsims, blk, nblk,
                       [&](int s, onc_u32 i) __attribute__((always_inline)) { onc_forces_cell(a, s, i, L); });
}

// The force walk is bound by the latency of its dependent gathers (bin
// entries, then candidate positions, then the listed neighbours): five waves
// per SIMD instead of four (96 VGPRs, a 20-byte spill) run it 6% faster on the
// dense 64-sim ensemble.
__global__ __launch_bounds__(ONC_FORCES_BLOCK) __attribute__((amdgpu_waves_per_eu(ONC_FORCES_WPE, 8))) void k_forces(onc_kargs a) {
    __shared__ onc_forces_lds L;
    onc_forces_wave(a, blockIdx.x, gridDim.x, L);
}

// The summed forces alone (onc_forces): the same walk and pair law, written
// out instead of integrated.
__global__ __launch_bounds__(ONC_FORCES_BLOCK) __attribute__((amdgpu_waves_per_eu(ONC_FORCES_WPE, 8)))
void k_forces_raw(onc_kargs a, float* fx, float* fy, float* fz) {
    __shared__ onc_forces_lds L;
    onc_for_chunks_xcd(a.cell_prefix, a.nsims, blockIdx.x, gridDim.x,
                       [&](int s, onc_u32 i) __attribute__((always_inline)) {
                           onc_forces_cell<true>(a, s, i, L, fx, fy, fz);
                       });
}

// Packed positions and keys of caller-supplied cells (onc_forces).
__global__ __launch_bounds__(256) void k_pack_keys(const float* __restrict__ px, const float* __restrict__ py,
                                                   const float* __restrict__ pz, onc_u32 n, const onc_dsim* S,
                                                   float4* __restrict__ pos4, onc_u32* __restrict__ keys) {
    const onc_dsim& D = *S;
    for (onc_u32 i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const float x = px[i], y = py[i], z = pz[i];
        pos4[i] = make_float4(x, y, z, 0.0f);
        keys[i] = onc_cell_key(D, x, y, z);
    }
}

// Bin entries {start, len << 16 | tag} of sorted keys: the first cell of every
// run of equal keys counts the run (bins longer than the force walk's limit
// set *err).
__global__ __launch_bounds__(256) void k_bins8(const onc_u32* __restrict__ keys, onc_u32 n, uint2* bins,
                                               onc_u32 tag, onc_u32* err) {
    for (onc_u32 i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const onc_u32 k = keys[i];
        if (i > 0u && keys[i - 1u] == k) continue;
        onc_u32 e = i + 1u;
        while (e < n && keys[e] == k) e++;
        bins[k] = make_uint2(i, (min(e - i, ONC_MAX_BIN_CELLS) << 16) | tag);
        if (e - i > ONC_MAX_BIN_CELLS) atomicOr(err, ONC_DERR_BIG_BIN);
    }
}

// ---------------------------------------------------------------------------
// 7-point explicit sweep, all (sim, field) jobs in one launch. A block of
// sbx x sby threads owns an xy tile and marches sbz planes along z, keeping
// the z-1 / z / z+1 values of its column in registers and loading plane z+1's
// x/y neighbours one iteration ahead (software pipelining); x/y neighbours
// come through L1/L2. u' = clamp(u + r (sum6 - 6u) - decay u) with the
// oracle's summation order; zero-flux faces mirror the centre.
// ---------------------------------------------------------------------------
// One output element of the sweep (the oracle's operation order).
__device__ __forceinline__ float onc_sweep1(const onc_dfield& fd, float c, float xm, float xp, float ym, float yp,
                                            float zm, float zp) {
    float sum = xm + xp;
    sum = sum + ym;
    sum = sum + yp;
    sum = sum + zm;
    sum = sum + zp;
    const float lap = sum - 6.0f * c;
    float w = c + fd.rate * lap;
    w = w - fd.decay * c;
    return onc_clampf(w, fd.lo, fd.hi);
}

// Waves of an NT-thread float4 sweep block with at least one voxel: xrem
// float4 columns and yrem rows of the tile lie in the grid; a wave holds a
// voxel iff its first lane does (lanes run along a row, rows of tw lanes).
__device__ __forceinline__ onc_u32 onc_stencil_live_waves(int tw, int th, int xrem, int yrem, int NT) {
    onc_u32 n = 0;
    for (int k = 0; k < NT / 64; k++)
        if ((64 * k) / tw < min(yrem, th) && (64 * k) % tw < min(xrem, tw)) n++;
    return n;
}

// Register-blocked sweep for rows whose length is a multiple of 4: a thread
// owns 4 consecutive voxels of a row (float4 loads / stores) and marches sbz
// planes along z; rows y-1, y+1 and planes z+1, z+2 come in as float4. A block of
// NT threads tiles tw = onc_stencil_tw(X/4, min(sbx, NT)) float4 columns by
// NT / tw rows,
// so each row segment of W = min(tw, 64) lanes holds consecutive float4 of
// one row: the x halo values pass between neighbouring lanes by shuffle, and
// only the segment's end lanes load them (one scalar load per segment
// instead of two per lane: the scalar halo loads, not HBM, limited the sweep).
//
// Quiet tiles (a.quiet set): the sweep keeps a state word and a value word
// per tile (the voxels of one block) and field buffer. State ONC_Q_NONE: not
// known to be uniform; any other state: every voxel of the tile holds the
// value in that buffer. A block whose tile and face-neighbour tiles all hold
// the same value v in the input buffer computes every output as v when v is a
// fixed point of the update (the update of a voxel whose seven stencil values
// are v gives v, bit for bit), a Dirichlet face holds v, and no source voxel
// lies in the tile: then it writes v (or nothing, when the output buffer's
// tile already holds v) without reading the input. Every other block computes
// its tile and records whether the outputs came out uniform. The permute
// kernel's uptake stamps the cubes it changes with the step (qstamp);
// host writes, step counter changes and batch rebuilds reset every tile to
// ONC_Q_NONE. The results are the bits of the full sweep.
template <int NT>
__device__ __forceinline__ void onc_stencil_vec4(const onc_kargs& a, const onc_dsim& S, const onc_dfield& fd, int f,
                                                 onc_u32 b, onc_u32 local, int sbx, int sbz) {
    const int X = S.dims[0], Y = S.dims[1], Z = S.dims[2];
    const int X4 = X >> 2;
    const int tw = onc_stencil_tw(X4, sbx < NT ? sbx : NT);
    const int th = NT / tw;
    const int W = tw < 64 ? tw : 64;
    const int ntx = (X4 + tw - 1) / tw, nty = (Y + th - 1) / th;
    const int tx = (int)(local % (onc_u32)ntx);
    const onc_u32 rest = local / (onc_u32)ntx;
    const int ty = (int)(rest % (onc_u32)nty);
    const int tz = (int)(rest / (onc_u32)nty);
    const int t = (int)threadIdx.x;
    const int x4 = tx * tw + t % tw;
    const int y = ty * th + t / tw;
    const int z0 = tz * sbz;
    const int z1 = min(Z, z0 + sbz);
    const onc_u32 par = S.par;
    const float* __restrict__ in = S.fld[par][f];
    float* __restrict__ out = S.fld[par ^ 1u][f];
    const size_t XY = (size_t)X * (size_t)Y;
    const bool dir = fd.bc == 0, psrc = fd.bc == 1;
    onc_u32 ONC_GLOBAL* qout = nullptr;
    if (a.quiet) {
        const onc_u32 ONC_GLOBAL* qin = a.quiet + 2u * ((size_t)par * a.nqb);
        qout = a.quiet + 2u * ((size_t)(par ^ 1u) * a.nqb);
        const onc_u32 vb = qin[2u * b + 1u];
        bool skip = qin[2u * b] != ONC_Q_NONE;
        const onc_u32 sx = 1u, sy = (onc_u32)ntx, sz = (onc_u32)ntx * (onc_u32)nty;
        const int ntz = (Z + sbz - 1) / sbz;
        auto same = [&](onc_u32 nb) { return qin[2u * nb] != ONC_Q_NONE && qin[2u * nb + 1u] == vb; };
        if (skip && tx > 0) skip = same(b - sx);
        if (skip && tx + 1 < ntx) skip = same(b + sx);
        if (skip && ty > 0) skip = same(b - sy);
        if (skip && ty + 1 < nty) skip = same(b + sy);
        if (skip && tz > 0) skip = same(b - sz);
        if (skip && tz + 1 < ntz) skip = same(b + sz);
        const float v = __uint_as_float(vb);
        if (skip) skip = __float_as_uint(onc_sweep1(fd, v, v, v, v, v, v, v)) == vb;
        if (skip && dir) skip = __float_as_uint(fd.bc_value) == vb;
        if (skip && psrc && z0 == 0) skip = false;
        bool kept = qout[2u * b] != ONC_Q_NONE && qout[2u * b + 1u] == vb;
        if (skip && fd.uptake != 0.0f) {
            // The flags describe the buffers as the sweeps wrote them. The
            // uptake of this step changed the input at the cubes stamped
            // step + 1, that of the previous step the output at those stamped
            // step: the tile and its one-voxel halo must hold none of the
            // first, and the output tile none of the second to be kept. (A
            // field without uptake is never changed between sweeps.)
            const int xa = max(tx * tw * 4 - 1, 0), xb = min((tx + 1) * tw * 4, X - 1);
            const int ya = max(ty * th - 1, 0), yb = min((ty + 1) * th, Y - 1);
            const int za = max(z0 - 1, 0), zb = min(z1, Z - 1);
            constexpr int QC = ONC_QCUBE;
            const int ncx = (xb >> QC) - (xa >> QC) + 1, ncy = (yb >> QC) - (ya >> QC) + 1, ncz = (zb >> QC) - (za >> QC) + 1;
            const onc_u32 now = S.step + 1u;
            bool din = false, dout = false;
            for (int c = t; c < ncx * ncy * ncz; c += NT) {
                const int cx = (xa >> QC) + c % ncx, cy = (ya >> QC) + (c / ncx) % ncy, cz = (za >> QC) + c / (ncx * ncy);
                const onc_u32 st = S.qstamp[onc_morton((onc_u32)cx, (onc_u32)cy, (onc_u32)cz, S.key_format)];
                din = din || st == now;
                dout = dout || st == S.step;
            }
            if (NT <= 64) {
                din = __ballot(din) != 0ull;
                dout = __ballot(dout) != 0ull;
            } else {
                din = __syncthreads_or(din);
                dout = __syncthreads_or(dout);
            }
            skip = !din;
            kept = kept && !dout;
        }
        if (skip) {
            const bool have = kept;
            if (!have && x4 < X4 && y < Y) {
                const float4 o = make_float4(v, v, v, v);
                size_t w = (size_t)x4 * 4u + (size_t)X * (size_t)y + XY * (size_t)z0;
                for (int z = z0; z < z1; z++, w += XY) onc_st_nt((float4*)(out + w), o);
            }
            if (t == 0) {
                qout[2u * b] = have ? ONC_Q_KEPT : ONC_Q_FILLED;
                qout[2u * b + 1u] = vb;
            }
            return;
        }
    }
    // Uniformity of the computed tile (NT = 256: four waves combine through
    // LDS atomics; set up before any lane leaves)
    __shared__ onc_u32 s_q[4];  // AND of the waves' verdicts, min and max first output, waves done
    if (NT > 64 && qout) {
        if (t == 0) {
            s_q[0] = 1u;
            s_q[1] = 0xFFFFFFFFu;
            s_q[2] = 0u;
            s_q[3] = 0u;
        }
        __syncthreads();
    }
    // lanes past the row end or the last row leave; a remaining lane only
    // takes a shuffled halo from a neighbour inside the row (hxm / hxp)
    if (x4 >= X4 || y >= Y) return;
    const int sl = t % W;  // lane inside its row segment
    const int x0 = x4 * 4;
    const size_t row = (size_t)x0 + (size_t)X * (size_t)y;
    const bool hxm = x0 > 0, hxp = x0 + 4 < X, hym = y > 0, hyp = y < Y - 1;
    const bool ldl = hxm && sl == 0, ldr = hxp && sl == W - 1;
    onc_u32 first = 0u;  // this lane's first output
    bool uni = true;     // every output of this lane equals it
    // The centre column runs two planes ahead of the plane being computed (its
    // z+1 neighbour must already be in registers), the y rows and x halos one
    // plane ahead: every load has a whole iteration to arrive.
    size_t v = row + XY * (size_t)z0;
    float4 c = *(const float4*)(in + v);
    float4 cm = z0 > 0 ? *(const float4*)(in + v - XY) : c;
    float4 cp = z0 + 1 < Z ? *(const float4*)(in + v + XY) : c;
    float4 ym = hym ? *(const float4*)(in + v - (size_t)X) : c;
    float4 yp = hyp ? *(const float4*)(in + v + (size_t)X) : c;
    float lx = ldl ? in[v - 1] : 0.0f;
    float rx = ldr ? in[v + 4] : 0.0f;
    for (int z = z0; z < z1; z++, v += XY) {
        const bool nxt = z + 1 < Z;
        const bool more = z + 1 < z1;
        const size_t vn = v + XY;
        float4 cpp = cp;
        float4 nym = cp, nyp = cp;
        float nlx = 0.0f, nrx = 0.0f;
        if (more) {
            if (z + 2 < Z) cpp = *(const float4*)(in + vn + XY);
            if (hym) nym = *(const float4*)(in + vn - (size_t)X);
            if (hyp) nyp = *(const float4*)(in + vn + (size_t)X);
            if (ldl) nlx = in[vn - 1];
            if (ldr) nrx = in[vn + 4];
        }
        const float sl_w = __shfl_up(c.w, 1, W);
        const float sr_x = __shfl_down(c.x, 1, W);
        const float hl = hxm ? (sl == 0 ? lx : sl_w) : c.x;
        const float hr = hxp ? (sl == W - 1 ? rx : sr_x) : c.w;
        const float4 zm = z > 0 ? cm : c;
        const float4 zp = nxt ? cp : c;
        float4 o;
        o.x = onc_sweep1(fd, c.x, hl, c.y, ym.x, yp.x, zm.x, zp.x);
        o.y = onc_sweep1(fd, c.y, c.x, c.z, ym.y, yp.y, zm.y, zp.y);
        o.z = onc_sweep1(fd, c.z, c.y, c.w, ym.z, yp.z, zm.z, zp.z);
        o.w = onc_sweep1(fd, c.w, c.z, hr, ym.w, yp.w, zm.w, zp.w);
        if (dir) {
            if (z == 0 || z == Z - 1 || y == 0 || y == Y - 1) {
                o = make_float4(fd.bc_value, fd.bc_value, fd.bc_value, fd.bc_value);
            } else {
                if (x0 == 0) o.x = fd.bc_value;
                if (x0 + 4 == X) o.w = fd.bc_value;
            }
        } else if (psrc && z == 0) {
            if (onc_is_fixed(fd, S.dims, x0, y, 0)) o.x = fd.bc_value;
            if (onc_is_fixed(fd, S.dims, x0 + 1, y, 0)) o.y = fd.bc_value;
            if (onc_is_fixed(fd, S.dims, x0 + 2, y, 0)) o.z = fd.bc_value;
            if (onc_is_fixed(fd, S.dims, x0 + 3, y, 0)) o.w = fd.bc_value;
        }
        // the next sweep reads this plane from HBM anyway: a streaming
        // store keeps it from evicting the planes this sweep still reads
        onc_st_nt((float4*)(out + v), o);
        if (qout) {
            const onc_u32 ox = __float_as_uint(o.x);
            if (z == z0) first = ox;
            uni = uni && ox == first && __float_as_uint(o.y) == first && __float_as_uint(o.z) == first &&
                  __float_as_uint(o.w) == first;
        }
        cm = c;
        c = cp;
        cp = cpp;
        ym = nym;
        yp = nyp;
        lx = nlx;
        rx = nrx;
    }
    if (!qout) return;
    // the block's verdict: every lane uniform, all with the same first output
    const onc_u32 f0 = onc_uniform(first);
    const bool wave_ok = __ballot(!uni || first != f0) == 0ull;
    if (NT <= 64) {
        if (t == 0) {
            qout[2u * b] = wave_ok ? ONC_Q_COMPUTED : ONC_Q_NONE;
            qout[2u * b + 1u] = wave_ok ? f0 : 0u;
        }
        return;
    }
    // the last live wave to finish writes the verdict
    if ((t & 63) == 0) {
        if (!wave_ok) atomicAnd(&s_q[0], 0u);
        atomicMin(&s_q[1], f0);
        atomicMax(&s_q[2], f0);
        const onc_u32 live = onc_stencil_live_waves(tw, th, X4 - tx * tw, Y - ty * th, NT);
        if (atomicAdd(&s_q[3], 1u) + 1u == live) {
            const bool ok = s_q[0] != 0u && s_q[1] == s_q[2];
            qout[2u * b] = ok ? ONC_Q_COMPUTED : ONC_Q_NONE;
            qout[2u * b + 1u] = ok ? s_q[1] : 0u;
        }
    }
}

// One block of the sweep: logical block `b` of the (sim, field) job table,
// NT threads. Rows that are not a multiple of 4 use bx x by scalar tiles with
// bx = min(sbx, NT), by = NT / bx.
template <int NT>
__device__ __forceinline__ void onc_stencil_block(const onc_kargs& a, onc_u32 b) {
    const int jb = onc_find_seg_wave(a.jobs_prefix, a.njobs, b);
    const onc_u32 job = onc_uniform(a.jobs[jb]);
    const onc_dsim& S = a.sims[job >> 8];
    if (S.frozen) return;
    const int f = (int)(job & 255u);
    const int X = S.dims[0], Y = S.dims[1], Z = S.dims[2];
    const int sbz = a.sbz;
    const onc_u32 local = b - a.jobs_prefix[jb];
    if ((X & 3) == 0) {
        const onc_dfield fdv = S.fields[f];
        onc_stencil_vec4<NT>(a, S, fdv, f, b, local, a.sbx, sbz);
        return;
    }
    const int sbx = a.sbx < NT ? a.sbx : NT, sby = NT / sbx;
    const onc_u32 nbx = (onc_u32)((X + sbx - 1) / sbx), nby = (onc_u32)((Y + sby - 1) / sby);
    const int bx = (int)(local % nbx);
    const onc_u32 rest = local / nbx;
    const int by = (int)(rest % nby);
    const int bz = (int)(rest / nby);
    const int x = bx * sbx + (int)(threadIdx.x % (unsigned)sbx);
    const int y = by * sby + (int)(threadIdx.x / (unsigned)sbx);
    if (x >= X || y >= Y) return;
    const int z0 = bz * sbz;
    const int z1 = min(Z, z0 + sbz);
    const onc_u32 par = S.par;
    const float* __restrict__ in = S.fld[par][f];
    float* __restrict__ out = S.fld[par ^ 1u][f];
    const onc_dfield fd = S.fields[f];
    const size_t XY = (size_t)X * (size_t)Y;
    const size_t col = (size_t)x + (size_t)X * (size_t)y;
    const bool hxm = x > 0, hxp = x < X - 1, hym = y > 0, hyp = y < Y - 1;
    // plane z0: centre, below, and in-plane neighbours
    size_t v = col + XY * (size_t)z0;
    float cc = in[v];
    float cm = z0 > 0 ? in[v - XY] : cc;
    float cp = z0 + 1 < Z ? in[v + XY] : cc;
    float xm = hxm ? in[v - 1] : cc;
    float xp = hxp ? in[v + 1] : cc;
    float ym = hym ? in[v - (size_t)X] : cc;
    float yp = hyp ? in[v + (size_t)X] : cc;
    for (int z = z0; z < z1; z++, v += XY) {
        // issue the loads of plane z+1 (its centre: plane z+2) before computing plane z
        const bool nxt = z + 1 < Z;
        const size_t vn = v + XY;
        float cpp = cp;
        float nxm = 0.0f, nxp = 0.0f, nym = 0.0f, nyp = 0.0f;
        if (nxt && z + 1 < z1) {
            if (z + 2 < Z) cpp = in[vn + XY];
            nxm = hxm ? in[vn - 1] : 0.0f;
            nxp = hxp ? in[vn + 1] : 0.0f;
            nym = hym ? in[vn - (size_t)X] : 0.0f;
            nyp = hyp ? in[vn + (size_t)X] : 0.0f;
        }
        if (onc_is_fixed(fd, S.dims, x, y, z)) {
            out[v] = fd.bc_value;
        } else {
            const float zm = z > 0 ? cm : cc;
            const float zp = nxt ? cp : cc;
            float sum = xm + xp;
            sum = sum + ym;
            sum = sum + yp;
            sum = sum + zm;
            sum = sum + zp;
            const float lap = sum - 6.0f * cc;
            float w = cc + fd.rate * lap;
            w = w - fd.decay * cc;
            out[v] = onc_clampf(w, fd.lo, fd.hi);
        }
        cm = cc;
        cc = cp;
        xm = hxm ? nxm : cp;
        xp = hxp ? nxp : cp;
        ym = hym ? nym : cp;
        yp = hyp ? nyp : cp;
        cp = cpp;
    }
}

__global__ __launch_bounds__(256) void k_stencil(onc_kargs a) {
    onc_stencil_block<256>(a, onc_uniform(onc_xcd_block(blockIdx.x, gridDim.x)));
}

// Forces and the sweep in one launch of one-wave blocks: neither reads what
// the other writes (forces: packed positions -> position columns; sweep:
// fields), and the force walk is bound by gather latency and ALU while the
// sweep streams HBM, so running them side by side on the same CUs hides much
// of one behind the other. Blocks go in groups of 8 (one per XCD), the sweep's
// groups spread evenly among the force groups (a_.fs_sgroups of the
// gridDim.x / 8 groups), so both parts keep their XCD-aware block orders and
// proceed together through the dispatch order.
__global__ __launch_bounds__(ONC_FORCES_BLOCK) __attribute__((amdgpu_waves_per_eu(ONC_FORCES_WPE, 8))) void k_forces_stencil(onc_kargs a) {
    __shared__ onc_forces_lds L;
    const onc_u32 G = gridDim.x >> 3, Gs = a.fs_sgroups;
    const onc_u32 pg = blockIdx.x >> 3, x = blockIdx.x & 7u;
    const onc_u32 s0 = (onc_u32)(((unsigned long long)pg * Gs) / G);
    const onc_u32 s1 = (onc_u32)(((unsigned long long)(pg + 1u) * Gs) / G);
    if (s1 > s0) {
        const onc_u32 ns = a.jobs_prefix[a.njobs];
        const onc_u32 b = onc_xcd_block(s0 * 8u + x, Gs * 8u);
        if (b < ns) onc_stencil_block<ONC_FORCES_BLOCK>(a, onc_uniform(b));
    } else {
        onc_forces_wave(a, (pg - s0) * 8u + x, (G - Gs) * 8u, L);
    }
}


// ---------------------------------------------------------------------------
// Built-in tumour rule (the benchmark workloads' model). Reads the permuted state (other
// buffer), the integrated position and the swept fields; writes the cell's
// columns to the current buffer, its alive/divide flags and its daughter.
// ---------------------------------------------------------------------------
// The rule for one cell i of sim S: reads its integrated position (current
// buffer), id / age / generation (permuted buffer) and the swept fields; gives
// its alive / divide decisions, its new age and generation, and where a
// daughter would go.
struct onc_tumor_out {
    float px, py, pz, age, dpx, dpy, dpz;
    onc_u32 id, gen, alive, div;
};

// The rule's inputs from the cell's columns: its integrated position
// (current buffer) and id / age / generation (permuted buffer).
struct onc_tumor_in {
    float px, py, pz, age;
    onc_u32 id, gen;
};
__device__ __forceinline__ onc_tumor_in onc_tumor_load(const onc_dsim& S, onc_u32 i) {
    const onc_u32 par = S.par;
    onc_gcol const* T = S.st[par ^ 1u];
    onc_gcol const* C = S.st[par];
    onc_tumor_in c;
    c.px = ((const float*)C[0])[i];
    c.py = ((const float*)C[1])[i];
    c.pz = ((const float*)C[2])[i];
    c.id = T[3][i];
    c.age = __uint_as_float(T[4][i]);
    c.gen = T[5][i];
    return c;
}

// Whether the rule reads the drug field (field 1) at the cell's voxel.
__device__ __forceinline__ bool onc_tumor_reads_drug(const onc_dsim& S) {
    return S.n_fields > 1 && S.params[3] > 0.0f;
}

// The rule given the cell's inputs and the swept O2 (field 0) and drug
// (field 1; read only when onc_tumor_reads_drug) at its voxel.
__device__ __forceinline__ onc_tumor_out onc_tumor_eval(const onc_dsim& S, const onc_tumor_in& c, float o2,
                                                        float drug) {
    onc_tumor_out o;
    o.px = c.px;
    o.py = c.py;
    o.pz = c.pz;
    o.id = c.id;
    o.gen = c.gen;
    float age = c.age;
    onc_rng rng;
    onc_rng_init(rng, S, o.id, S.step);
    age = age + 1.0f;
    o.alive = 1u;
    o.div = 0u;
    o.dpx = 0.0f;
    o.dpy = 0.0f;
    o.dpz = 0.0f;
    if (o2 < S.params[2]) {
        o.alive = 0u;
    } else {
        if (onc_tumor_reads_drug(S)) {
            const float u = onc_rand(rng);
            if (u < S.params[3] * drug) o.alive = 0u;
        }
        if (o.alive && age > S.params[0] && o2 > S.params[1]) {
            float ux, uy, uz;
            onc_rand_dir(rng, &ux, &uy, &uz);
            const float ofs = S.params[4];
            o.dpx = onc_clampf(o.px + ux * ofs, 0.0f, (float)S.dims[0] * S.dx);
            o.dpy = onc_clampf(o.py + uy * ofs, 0.0f, (float)S.dims[1] * S.dx);
            o.dpz = onc_clampf(o.pz + uz * ofs, 0.0f, (float)S.dims[2] * S.dx);
            age = 0.0f;
            o.gen = o.gen + 1u;
            o.div = 1u;
        }
    }
    o.age = age;
    return o;
}

__device__ __forceinline__ onc_tumor_out onc_tumor_rule(const onc_dsim& S, onc_u32 i) {
    const onc_tumor_in c = onc_tumor_load(S, i);
    const onc_u32 v = onc_field_index(S, c.px, c.py, c.pz);
    const float o2 = S.fld[S.par ^ 1u][0][v];
    const float drug = onc_tumor_reads_drug(S) ? S.fld[S.par ^ 1u][1][v] : 0.0f;
    return onc_tumor_eval(S, c, o2, drug);
}

__global__ __launch_bounds__(256) void k_rule_tumor(onc_rule_args r) {
    onc_for_cells(r.cell_prefix, r.nsims, [&](int s, onc_u32 i) {
        const onc_dsim& S = r.sims[s];
        const onc_u32 par = S.par;
        onc_gcol const* T = S.st[par ^ 1u];
        onc_gcol const* C = S.st[par];
        const onc_tumor_out o = onc_tumor_rule(S, i);
        C[3][i] = o.id;
        C[4][i] = __float_as_uint(o.age);
        C[5][i] = o.gen;
        for (int c = 6; c < S.n_cols; c++) C[c][i] = T[c][i];
        S.flags[i] = o.alive | (o.div << 1);
        if (o.div) {
            S.dau[0][i] = __float_as_uint(o.dpx);
            S.dau[1][i] = __float_as_uint(o.dpy);
            S.dau[2][i] = __float_as_uint(o.dpz);
            S.dau[3][i] = 0u;
            S.dau[4][i] = __float_as_uint(0.0f);
            S.dau[5][i] = o.gen;
            for (int c = 6; c < S.n_cols; c++) S.dau[c][i] = T[c][i];
        }
    });
}

// ---------------------------------------------------------------------------
// Compaction: scan of (alive, divide) flags with decoupled look-back;
// survivors are gathered in order into the other buffer, daughters into a
// dense staging area in parent order. The sim's last tile records totals.
// ---------------------------------------------------------------------------

// dst[c][slot[k]] = src[c][base + k * 256 + tid] for every column c and
// every item k whose slot is not NONE.
template <int ITEMS>
__device__ __forceinline__ void compact_copy(onc_gcol const* src, onc_gcol const* dst, int n_cols, onc_u32 base,
                                             const onc_u32* slot) {
    const int tid = threadIdx.x;
    for (int c0 = 0; c0 < n_cols; c0 += 4) {
        onc_u32 v[4][ITEMS];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (c0 + q >= n_cols) break;
            const onc_u32 ONC_GLOBAL* __restrict__ sp = src[c0 + q] + base;
#pragma unroll
            for (int k = 0; k < ITEMS; k++) v[q][k] = slot[k] != 0xFFFFFFFFu ? sp[k * 256 + tid] : 0u;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (c0 + q >= n_cols) break;
            onc_gcol __restrict__ dp = dst[c0 + q];
#pragma unroll
            for (int k = 0; k < ITEMS; k++)
                if (slot[k] != 0xFFFFFFFFu) dp[slot[k]] = v[q][k];
        }
    }
}

// Survivor slots and daughter ranks of one compaction tile: on entry
// s_aslot[j] holds the flags (bit 0 alive, bit 1 divides) of cell base + j of
// tile t (0 past the tile's cells); on exit s_aslot[j] / s_dslot[j] hold the
// cell's slot among the sim's survivors / its rank among the sim's daughters,
// or NONE. Scan of both counts with decoupled look-back; the sim's last tile
// records the totals. All threads of the block call it.
template <int ITEMS>
__device__ __forceinline__ void compact_slots(const onc_kargs& a, onc_u32 g, const onc_dsim& S, onc_u32 t,
                                              onc_u32 ntiles_s, onc_u32* s_aslot, onc_u32* s_dslot, onc_u32* s_tmp,
                                              onc_u32& s_pa, onc_u32& s_pd) {
    constexpr onc_u32 NONE = 0xFFFFFFFFu;
    const int tid = threadIdx.x;
    __syncthreads();
    onc_u32 la = 0u, ldv = 0u;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        onc_u32 f = s_aslot[tid * ITEMS + k];
        la += f & 1u;
        ldv += (f >> 1) & 1u;
    }
    onc_u32 tp;
    const onc_u32 ex = block_exscan256(la | (ldv << 16), s_tmp, &tp);
    const onc_u32 ta = tp & 0xFFFFu, td = tp >> 16;
    if (tid < 64) {
        onc_u32 pa = 0u, pd = 0u;
        onc_u64* mine = &a.comp_status[g];
        const onc_u64 agg = (onc_u64)ta | ((onc_u64)td << 31);
        if (t == 0u) {
            if (tid == 0) st_relaxed64(mine, ONC_F64_P | agg);
        } else {
            if (tid == 0) st_relaxed64(mine, ONC_F64_A | agg);
            lookback_wave64(a.comp_status, g, g - t, true, a.err, 2u, &pa, &pd);
            if (tid == 0) st_relaxed64(mine, ONC_F64_P | ((onc_u64)(pa + ta) | ((onc_u64)(pd + td) << 31)));
        }
        if (tid == 0) {
            s_pa = pa;
            s_pd = pd;
            if (t + 1u == ntiles_s) {
                S.state->nsurv = pa + ta;
                S.state->ndiv = pd + td;
            }
        }
    }
    __syncthreads();
    onc_u32 ra = s_pa + (ex & 0xFFFFu), rd = s_pd + (ex >> 16);
    onc_u32 fl[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) fl[k] = s_aslot[tid * ITEMS + k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
        const onc_u32 f = fl[k];
        s_aslot[tid * ITEMS + k] = (f & 1u) ? ra++ : NONE;
        s_dslot[tid * ITEMS + k] = (f & 2u) ? rd++ : NONE;
    }
    __syncthreads();
}

template <int ITEMS>
__global__ __launch_bounds__(256) void k_compact(onc_kargs a) {
    constexpr int TILE = 256 * ITEMS;
    constexpr onc_u32 NONE = 0xFFFFFFFFu;
    __shared__ onc_u32 s_aslot[TILE];
    __shared__ onc_u32 s_dslot[TILE];
    __shared__ onc_u32 s_tmp[8];
    __shared__ onc_u32 s_pa, s_pd;
    const int tid = threadIdx.x;
    const onc_u32 total = a.ctile_prefix[a.nsims];
    for (onc_u32 g = blockIdx.x; g < total; g += gridDim.x) {  // static tile order, as in k_sort_pass
        const int s = onc_find_seg_wave(a.ctile_prefix, a.nsims, g);
        const onc_dsim& S = a.sims[s];
        const onc_u32 t = g - a.ctile_prefix[s];
        const onc_u32 ntiles_s = a.ctile_prefix[s + 1] - a.ctile_prefix[s];
        const onc_u32 n = S.n_live;
        const onc_u32 base = t * (onc_u32)TILE;
        const onc_u32 cnt = min((onc_u32)TILE, n - base);
        const onc_u32 par = S.par;
        // flags, striped load into LDS
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            onc_u32 j = (onc_u32)(k * 256 + tid);
            s_aslot[j] = j < cnt ? S.flags[base + j] : 0u;
        }
        compact_slots<ITEMS>(a, g, S, t, ntiles_s, s_aslot, s_dslot, s_tmp, s_pa, s_pd);
        // Copies in groups of four columns: all loads of a group are issued
        // before any of its stores (a store to one column may alias a later
        // column's load, so the compiler keeps program order and each group
        // costs one memory round trip instead of one per column).
        onc_u32 sa[ITEMS], sd[ITEMS];
        bool anyd = false;
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const onc_u32 j = (onc_u32)(k * 256 + tid);
            sa[k] = j < cnt ? s_aslot[j] : NONE;
            sd[k] = j < cnt ? s_dslot[j] : NONE;
            anyd |= sd[k] != NONE;
        }
        compact_copy<ITEMS>(S.st[par], S.st[par ^ 1u], S.n_cols, base, sa);
        if (anyd) {
            compact_copy<ITEMS>(S.dau, S.dau2, S.n_cols, base, sd);
            // daughter ids: next_id + rank among the sim's daughters
#pragma unroll
            for (int k = 0; k < ITEMS; k++)
                if (sd[k] != NONE) S.dau2[ONC_COL_ID_DEV][sd[k]] = S.next_id + sd[k];
        }
        __syncthreads();
    }
}

// The built-in tumour rule fused into compaction, for the six-column tumour
// schema (pos.x, pos.y, pos.z, id, age, generation): each tile evaluates the
// rule for its cells into registers, scans the flags, and writes survivors
// and daughters straight to their slots, so neither the flags nor the rule's
// output columns nor the sparse daughters make a round trip through memory.
// Survivors go to the permuted buffer in place: a tile writes only after its
// look-back, i.e. after every earlier tile of its sim has read its cells, and
// its slots lie below the end of its own cells, which it holds in registers.
template <int ITEMS>
__global__ __launch_bounds__(256) void k_rule_compact(onc_kargs a) {
    constexpr int TILE = 256 * ITEMS;
    constexpr onc_u32 NONE = 0xFFFFFFFFu;
    __shared__ onc_u32 s_aslot[TILE];
    __shared__ onc_u32 s_dslot[TILE];
    __shared__ onc_u32 s_tmp[8];
    __shared__ onc_u32 s_pa, s_pd;
    const int tid = threadIdx.x;
    const onc_u32 total = a.ctile_prefix[a.nsims];
    for (onc_u32 g = blockIdx.x; g < total; g += gridDim.x) {  // static tile order, as in k_sort_pass
        const int s = onc_find_seg_wave(a.ctile_prefix, a.nsims, g);
        const onc_dsim& S = a.sims[s];
        const onc_u32 t = g - a.ctile_prefix[s];
        const onc_u32 ntiles_s = a.ctile_prefix[s + 1] - a.ctile_prefix[s];
        const onc_u32 n = S.n_live;
        const onc_u32 base = t * (onc_u32)TILE;
        const onc_u32 cnt = min((onc_u32)TILE, n - base);
        const onc_u32 par = S.par;
        // the rule for the tile's cells, PG items at a time with their loads
        // in two rounds (the columns, then the fields at the cells' voxels)
        // instead of two dependent rounds per item. Tiles of up to four cells
        // per thread (small populations, where the kernel's latency is its
        // cost) take all their items at once: the 1,000-step run of the
        // 1M-cell workload 1.5 % faster. Larger tiles take pairs: eight items
        // (or four) at once took 127 -> 164 (138) VGPRs, three waves per SIMD
        // instead of four, and the kernel from 35 to 37-38 us at 1.4M cells;
        // pairs keep 126 VGPRs and took it to 34.3 us.
        constexpr int PG = ITEMS <= 4 ? ITEMS : 2;
        const bool rd = onc_tumor_reads_drug(S);
        const float* __restrict__ f0 = S.fld[par ^ 1u][0];
        const float* __restrict__ f1 = rd ? S.fld[par ^ 1u][1] : f0;
        onc_tumor_out o[ITEMS];
#pragma unroll
        for (int k0 = 0; k0 < ITEMS; k0 += PG) {
            onc_tumor_in cin[PG];
#pragma unroll
            for (int q = 0; q < PG; q++) {
                const onc_u32 j = (onc_u32)((k0 + q) * 256 + tid);
                if (j < cnt) cin[q] = onc_tumor_load(S, base + j);
            }
            float o2[PG], drug[PG];
#pragma unroll
            for (int q = 0; q < PG; q++) {
                const onc_u32 j = (onc_u32)((k0 + q) * 256 + tid);
                o2[q] = 0.0f;
                drug[q] = 0.0f;
                if (j < cnt) {
                    const onc_u32 v = onc_field_index(S, cin[q].px, cin[q].py, cin[q].pz);
                    o2[q] = f0[v];
                    if (rd) drug[q] = f1[v];
                }
            }
#pragma unroll
            for (int q = 0; q < PG; q++) {
                const int k = k0 + q;
                const onc_u32 j = (onc_u32)(k * 256 + tid);
                o[k].alive = 0u;
                o[k].div = 0u;
                if (j < cnt) o[k] = onc_tumor_eval(S, cin[q], o2[q], drug[q]);
                s_aslot[j] = o[k].alive | (o[k].div << 1);
            }
        }
        compact_slots<ITEMS>(a, g, S, t, ntiles_s, s_aslot, s_dslot, s_tmp, s_pa, s_pd);
        onc_gcol const* D = S.st[par ^ 1u];
#pragma unroll
        for (int k = 0; k < ITEMS; k++) {
            const onc_u32 j = (onc_u32)(k * 256 + tid);
            if (j >= cnt) continue;
            const onc_u32 sa = s_aslot[j], sd = s_dslot[j];
            if (sa != NONE) {
                D[0][sa] = __float_as_uint(o[k].px);
                D[1][sa] = __float_as_uint(o[k].py);
                D[2][sa] = __float_as_uint(o[k].pz);
                D[3][sa] = o[k].id;
                D[4][sa] = __float_as_uint(o[k].age);
                D[5][sa] = o[k].gen;
            }
            if (sd != NONE) {
                S.dau2[0][sd] = __float_as_uint(o[k].dpx);
                S.dau2[1][sd] = __float_as_uint(o[k].dpy);
                S.dau2[2][sd] = __float_as_uint(o[k].dpz);
                S.dau2[3][sd] = S.next_id + sd;
                S.dau2[4][sd] = __float_as_uint(0.0f);
                S.dau2[5][sd] = o[k].gen;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// pure kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_morton_keys(const float* px, const float* py, const float* pz, onc_u32 n,
                                                    int X, int Y, int Z, float dx, int fmt, onc_u32* keys) {
    for (onc_u32 i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        keys[i] = onc_morton((onc_u32)onc_vox1(px[i], dx, X), (onc_u32)onc_vox1(py[i], dx, Y),
                             (onc_u32)onc_vox1(pz[i], dx, Z), fmt);
    }
}

__global__ __launch_bounds__(256) void k_bins(const onc_u32* keys, onc_u32 n, onc_u32* start, onc_u32* end) {
    for (onc_u32 i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const onc_u32 k = keys[i];
        if (i == 0u || keys[i - 1u] != k) start[k] = i;
        if (i + 1u == n || keys[i + 1u] != k) end[k] = i + 1u;
    }
}

__global__ __launch_bounds__(256) void k_philox(const onc_u32* ctr, onc_u32 n, onc_u32 k0, onc_u32 k1,
                                               onc_u32* out) {
    for (onc_u32 i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        onc_u32 o[4];
        onc_philox(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], k0, k1, o);
        out[4 * i] = o[0];
        out[4 * i + 1] = o[1];
        out[4 * i + 2] = o[2];
        out[4 * i + 3] = o[3];
    }
}

// Field initial state: init everywhere, bc_value on fixed voxels.
__global__ __launch_bounds__(256) void k_field_init(float* u, int X, int Y, int Z, onc_dfield f) {
    const size_t nv = (size_t)X * (size_t)Y * (size_t)Z;
    const int d[3] = {X, Y, Z};
    for (size_t v = (size_t)blockIdx.x * 256u + threadIdx.x; v < nv; v += (size_t)gridDim.x * 256u) {
        const int x = (int)(v % (size_t)X);
        const int y = (int)((v / (size_t)X) % (size_t)Y);
        const int z = (int)(v / ((size_t)X * (size_t)Y));
        u[v] = onc_is_fixed(f, d, x, y, z) ? f.bc_value : f.init;
    }
}

__global__ __launch_bounds__(256) void k_copy_u32(const onc_u32* __restrict__ src, onc_u32* __restrict__ dst,
                                                 size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (size_t)gridDim.x * 256u) dst[i] = src[i];
}

// The plain copy at the rate HBM gives a stream: one 16-byte streaming load
// and store per thread, one pass over the grid (16-byte aligned pointers;
// the last n % 4 words are copied by block 0).
__global__ __launch_bounds__(256) void k_copy_v4(const onc_u32* __restrict__ src, onc_u32* __restrict__ dst,
                                                size_t n) {
    typedef onc_u32 v4u __attribute__((ext_vector_type(4)));
    const size_t n4 = n >> 2;
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load((const v4u*)src + i), (v4u*)dst + i);
    if (blockIdx.x == 0 && threadIdx.x < (n & 3u)) dst[n4 * 4u + threadIdx.x] = src[n4 * 4u + threadIdx.x];
}

__global__ __launch_bounds__(256) void k_fill_u32(onc_u32* p, onc_u32 v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (size_t)gridDim.x * 256u) p[i] = v;
}

// Standalone exclusive scan: single pass, decoupled look-back, 64-bit status
// words (2 flag bits, 32-bit running sum). A tile is 256 x ITEMS elements
// held in registers: wave w owns the contiguous run [w, w+1) x 64 ITEMS of
// it, and lane l holds the 16-byte groups at k * 256 + 4 l (k < ITEMS / 4),
// so every load and store instruction of a wave covers 1 KB in one piece
// and no LDS transpose is needed. Per group the wave scans the lane sums
// with shuffles; four wave totals and one look-back give the tile's
// offsets. Large tiles (16K or 32K elements) keep the look-backs few: the
// scan is bound by the latency of the look-back chain, not by the bytes
// moved, once tiles are small. Streaming loads and stores: every element is read
// and written once. V4 = false reads and writes the same groups one element
// at a time (pointers that are not 16-byte aligned), as does a partial last
// tile.
template <int ITEMS, bool V4, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_exscan(const onc_u32* __restrict__ in, onc_u32* __restrict__ out,
                                                onc_u32 n, onc_u32* total, onc_u64* status, onc_u32* err) {
    // one block per tile, in static order: a tile's look-back waits only on
    // lower tiles, whose blocks were dispatched before it (a shared tile
    // counter would serialise every block on one device atomic)
    constexpr int K = ITEMS / 4;
    constexpr int TILE = 256 * ITEMS;
    __shared__ onc_u32 s_w[4];
    __shared__ onc_u32 s_pre;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const onc_u32 ntiles = (onc_u32)(((size_t)n + TILE - 1) / TILE);
    const onc_u32 g = blockIdx.x;
    if (g >= ntiles) return;
    const onc_u32 lim = n - g * (onc_u32)TILE;  // elements from the tile start to the end
    const onc_u32 wbase = (onc_u32)w * (onc_u32)(64 * ITEMS) + (onc_u32)lane * 4u;
    const onc_u32* __restrict__ src = in + (size_t)g * TILE;
    onc_u32* __restrict__ dst = out + (size_t)g * TILE;
    typedef onc_u32 v4u __attribute__((ext_vector_type(4)));
    const bool full = V4 && lim >= (onc_u32)TILE;  // every group whole: unguarded 16-byte accesses
    uint4 x[K];
    if (full) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const v4u t = __builtin_nontemporal_load((const v4u*)(src + wbase + k * 256));
            x[k] = make_uint4(t.x, t.y, t.z, t.w);
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const onc_u32 e = wbase + (onc_u32)(k * 256);
            x[k].x = e < lim ? __builtin_nontemporal_load(src + e) : 0u;
            x[k].y = e + 1u < lim ? __builtin_nontemporal_load(src + e + 1u) : 0u;
            x[k].z = e + 2u < lim ? __builtin_nontemporal_load(src + e + 2u) : 0u;
            x[k].w = e + 3u < lim ? __builtin_nontemporal_load(src + e + 3u) : 0u;
        }
    }
    // each group becomes the exclusive prefixes of its four elements inside
    // the wave's run (in place: no registers beyond the tile itself, so four
    // waves fit per SIMD)
    onc_u32 wtot = 0u;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const onc_u32 v = x[k].x + x[k].y + x[k].z + x[k].w;
        onc_u32 y = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const onc_u32 t = __shfl_up(y, o, 64);
            if (lane >= o) y += t;
        }
        const onc_u32 e0 = wtot + y - v, e1 = e0 + x[k].x, e2 = e1 + x[k].y, e3 = e2 + x[k].z;
        x[k] = make_uint4(e0, e1, e2, e3);
        wtot += __shfl(y, 63, 64);
    }
    if (lane == 0) s_w[w] = wtot;
    __syncthreads();
    onc_u32 before = 0u, tot = 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const onc_u32 t = s_w[i];
        if (i < w) before += t;
        tot += t;
    }
    if (tid < 64) {
        onc_u32 pre = 0u, unused = 0u;
        onc_u64* mine = &status[g];
        if (g == 0u) {
            if (tid == 0) st_relaxed64(mine, ONC_F64_P | (onc_u64)tot);
        } else {
            if (tid == 0) st_relaxed64(mine, ONC_F64_A | (onc_u64)tot);
            lookback_wave64(status, g, 0u, false, err, 4u, &pre, &unused);
            if (tid == 0) st_relaxed64(mine, ONC_F64_P | (onc_u64)(pre + tot));
        }
        if (tid == 0) {
            s_pre = pre;
            if (g + 1u == ntiles && total) *total = pre + tot;
        }
    }
    __syncthreads();
    const onc_u32 off = s_pre + before;
    if (full) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const v4u t = {off + x[k].x, off + x[k].y, off + x[k].z, off + x[k].w};
            __builtin_nontemporal_store(t, (v4u*)(dst + wbase + k * 256));
        }
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const onc_u32 e = wbase + (onc_u32)(k * 256);
            if (e < lim) __builtin_nontemporal_store(off + x[k].x, dst + e);
            if (e + 1u < lim) __builtin_nontemporal_store(off + x[k].y, dst + e + 1u);
            if (e + 2u < lim) __builtin_nontemporal_store(off + x[k].z, dst + e + 2u);
            if (e + 3u < lim) __builtin_nontemporal_store(off + x[k].w, dst + e + 3u);
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t onc_launch_plan(const onc_kargs& a, int mode, hipStream_t s) {
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(256), 0, s, a, mode, (onc_dsim*)a.sims);
    return hipGetLastError();
}

#define ONC_DISPATCH_ITEMS(items, KERNEL, grid, s, ...)                                      \
    do {                                                                                     \
        if ((items) == 4)                                                                    \
            hipLaunchKernelGGL(KERNEL<4>, dim3(grid), dim3(256), 0, s, __VA_ARGS__);        \
        else if ((items) == 8)                                                               \
            hipLaunchKernelGGL(KERNEL<8>, dim3(grid), dim3(256), 0, s, __VA_ARGS__);        \
        else if ((items) == 12)                                                              \
            hipLaunchKernelGGL(KERNEL<12>, dim3(grid), dim3(256), 0, s, __VA_ARGS__);       \
        else                                                                                 \
            hipLaunchKernelGGL(KERNEL<16>, dim3(grid), dim3(256), 0, s, __VA_ARGS__);       \
    } while (0)

hipError_t onc_launch_morton(const onc_kargs& a, int grid, bool runs, hipStream_t s) {
    if (runs)
        hipLaunchKernelGGL(k_morton<true>, dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(k_morton<false>, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t onc_launch_hist_only(const onc_kargs& a, int items, int grid, hipStream_t s) {
    ONC_DISPATCH_ITEMS(items, k_hist_only, grid, s, a);
    return hipGetLastError();
}
hipError_t onc_launch_sort_pass(const onc_kargs& a, int items, int pass, int grid, hipStream_t s) {
    ONC_DISPATCH_ITEMS(items, k_sort_pass, grid, s, a, pass);
    return hipGetLastError();
}
int onc_sort_pass_blocks_per_cu(int items) {
    int nb = 0;
    hipError_t e;
    if (items == 4)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sort_pass<4>, 256, 0);
    else if (items == 8)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sort_pass<8>, 256, 0);
    else if (items == 12)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sort_pass<12>, 256, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sort_pass<16>, 256, 0);
    return e == hipSuccess ? nb : 0;
}
hipError_t onc_launch_permute_bins(const onc_kargs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_permute_bins, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t onc_launch_forces(const onc_kargs& a, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_forces, dim3(grid), dim3(ONC_FORCES_BLOCK), 0, s, a);
    return hipGetLastError();
}
hipError_t onc_launch_forces_stencil(const onc_kargs& a, onc_u32 ngroups, hipStream_t s) {
    hipLaunchKernelGGL(k_forces_stencil, dim3(8u * ngroups), dim3(ONC_FORCES_BLOCK), 0, s, a);
    return hipGetLastError();
}
hipError_t onc_launch_stencil(const onc_kargs& a, onc_u32 nblocks, hipStream_t s) {
    if (nblocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_stencil, dim3(nblocks), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t onc_launch_rule_builtin(const onc_rule_args& r, int grid, hipStream_t s) {
    hipLaunchKernelGGL(k_rule_tumor, dim3(grid), dim3(256), 0, s, r);
    return hipGetLastError();
}
hipError_t onc_launch_compact(const onc_kargs& a, int items, int grid, hipStream_t s) {
    ONC_DISPATCH_ITEMS(items, k_compact, grid, s, a);
    return hipGetLastError();
}
hipError_t onc_launch_rule_compact(const onc_kargs& a, int items, int grid, hipStream_t s) {
    ONC_DISPATCH_ITEMS(items, k_rule_compact, grid, s, a);
    return hipGetLastError();
}

static int grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return (int)b;
}

hipError_t onc_launch_morton_keys(const float* px, const float* py, const float* pz, onc_u32 n, int X, int Y,
                                  int Z, float dx, int fmt, onc_u32* keys, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_morton_keys, dim3(grid_for(n)), dim3(256), 0, s, px, py, pz, n, X, Y, Z, dx, fmt, keys);
    return hipGetLastError();
}
hipError_t onc_launch_bins(const onc_u32* keys, onc_u32 n, onc_u32* start, onc_u32* end, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bins, dim3(grid_for(n)), dim3(256), 0, s, keys, n, start, end);
    return hipGetLastError();
}
hipError_t onc_launch_philox(const onc_u32* ctr, onc_u32 n, onc_u32 k0, onc_u32 k1, onc_u32* out,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_philox, dim3(grid_for(n)), dim3(256), 0, s, ctr, n, k0, k1, out);
    return hipGetLastError();
}
hipError_t onc_launch_sort_up(const onc_u32* keys, onc_u32 n, onc_u32 shift, onc_u32* counts, onc_u32 T,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_sort_up<ONC_SORT3_ITEMS>, dim3(T), dim3(256), 0, s, keys, n, shift, counts, T);
    return hipGetLastError();
}
hipError_t onc_launch_sort_down(const onc_u32* kin, const onc_u32* vin, onc_u32* kout, onc_u32* vout, onc_u32 n,
                                onc_u32 shift, const onc_u32* offsets, onc_u32 T, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_down<ONC_SORT3_ITEMS>, dim3(T), dim3(256), 0, s, kin, vin, kout, vout, n, shift,
                       offsets, T);
    return hipGetLastError();
}
hipError_t onc_launch_forces_raw(const onc_kargs& a, const float* px, const float* py, const float* pz, onc_u32 n,
                                 const onc_dsim* S, float* pos4, onc_u32* keys, onc_u32* bins, onc_u32* err,
                                 float* fx, float* fy, float* fz, hipEvent_t t0, hipEvent_t t1, hipStream_t s) {
    const unsigned g = (unsigned)std::min<size_t>((n + 255u) / 256u, 65535u);
    hipLaunchKernelGGL(k_pack_keys, dim3(g), dim3(256), 0, s, px, py, pz, n, S, (float4*)pos4, keys);
    hipLaunchKernelGGL(k_bins8, dim3(g), dim3(256), 0, s, keys, n, (uint2*)bins, onc_bin_tag(0u), err);
    // one wave per 64 cells, a multiple of 8 blocks (onc_for_chunks_xcd)
    const unsigned waves = (unsigned)((((size_t)n + 63u) / 64u + 7u) / 8u * 8u);
    if (t0) (void)hipEventRecord(t0, s);
    hipLaunchKernelGGL(k_forces_raw, dim3(waves ? waves : 8u), dim3(ONC_FORCES_BLOCK), 0, s, a, fx, fy, fz);
    if (t1) (void)hipEventRecord(t1, s);
    return hipGetLastError();
}
hipError_t onc_launch_exscan(const onc_u32* in, onc_u32* out, onc_u32 n, onc_u32* total, onc_u32* status,
                             onc_u32* err, hipStream_t s) {
    if (n == 0) return hipSuccess;
    // 32K-element tiles (235 VGPRs, two waves per SIMD) once there are at
    // least 512 of them, 16K-element tiles (124 VGPRs, four waves per SIMD)
    // below that; pointers that are not 16-byte aligned take 4K-element
    // tiles of one-element accesses
    if ((((uintptr_t)in) | ((uintptr_t)out)) & 15u) {
        hipLaunchKernelGGL((k_exscan<16, false, 1>), dim3((unsigned)(((size_t)n + 4095u) / 4096u)), dim3(256), 0, s, in, out, n, total,
                           (onc_u64*)status, err);
    } else if (n >= (1u << 24)) {
        hipLaunchKernelGGL((k_exscan<128, true, 2>), dim3((unsigned)(((size_t)n + 32767u) / 32768u)), dim3(256), 0, s, in, out, n,
                           total, (onc_u64*)status, err);
    } else {
        hipLaunchKernelGGL((k_exscan<64, true, 4>), dim3((unsigned)(((size_t)n + 16383u) / 16384u)), dim3(256), 0, s, in, out, n, total,
                           (onc_u64*)status, err);
    }
    return hipGetLastError();
}
// After an odd number of tissue-rule passes the fields sit in fld[par]; the
// cell rule and the commit expect them in fld[par ^ 1]. blockIdx.y = sim.
__global__ __launch_bounds__(256) void k_copy_fields(onc_kargs a) {
    const onc_dsim& S = a.sims[blockIdx.y];
    if (S.frozen) return;
    const size_t nvox = (size_t)S.dims[0] * (size_t)S.dims[1] * (size_t)S.dims[2];
    const onc_u32 par = S.par;
    for (int f = 0; f < S.n_fields; f++) {
        const float* __restrict__ src = S.fld[par][f];
        float* __restrict__ dst = S.fld[par ^ 1u][f];
        for (size_t v = (size_t)blockIdx.x * 256u + threadIdx.x; v < nvox; v += (size_t)gridDim.x * 256u)
            dst[v] = src[v];
    }
}
hipError_t onc_launch_copy_fields(const onc_kargs& a, unsigned gx, hipStream_t s) {
    if (a.nsims == 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_fields, dim3(gx ? gx : 1, a.nsims), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t onc_launch_fill_u32(onc_u32* p, onc_u32 v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}
hipError_t onc_launch_field_init(float* u, int X, int Y, int Z, const onc_dfield& f, hipStream_t s) {
    const size_t nv = (size_t)X * (size_t)Y * (size_t)Z;
    if (nv == 0) return hipSuccess;
    hipLaunchKernelGGL(k_field_init, dim3(grid_for(nv)), dim3(256), 0, s, u, X, Y, Z, f);
    return hipGetLastError();
}
hipError_t onc_launch_copy_u32(const onc_u32* src, onc_u32* dst, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (((((uintptr_t)src) | ((uintptr_t)dst)) & 15u) == 0 && n / 4u < ((size_t)1 << 36)) {
        const size_t blocks = (n / 4u + 255u) / 256u;
        hipLaunchKernelGGL(k_copy_v4, dim3((unsigned)(blocks ? blocks : 1)), dim3(256), 0, s, src, dst, n);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(n)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}
