// onc_runtime.hip — the C ABI of liboncohip.so (include/onco.h): device
// contexts, sims and their device buffers, batches over a packed segment
// table with an optional captured hipGraph per timestep, hiprtc rule modules,
// and the pure-kernel entry points.
//
// Host orchestration replaces GpuBackend.run_step / _dispatch_sort_pipeline
// (framework/src/oncolytica/gpu/runtime/_gpu_backend.py:113-171, 407-432):
// one HIP stream per batch, one launch per phase for all of its sims, live
// counts kept on the device.
// Host code only: the buffer pointers of the segment table stay plain
// pointers in this file's device pass too (it stores host-side values into
// them); the kernels' translation unit sees them as global-memory pointers.
#define ONC_HOST_ONLY 1

#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/onco.h"
#include "onc_device_src.h"  // generated: onc_device_h_src (text of onc_device.h)
#include "onc_internal.h"

static_assert(sizeof(onc_field_desc) == sizeof(onc_dfield), "field desc layout");

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define CK(call)                                                                                   \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(ONC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define CKR(expr)              \
    do {                       \
        int r_ = (expr);       \
        if (r_ != ONC_OK) return r_; \
    } while (0)

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct onc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 256;
    std::string arch;
    hipDeviceProp_t prop{};
    // scratch for pure kernels
    std::vector<void*> scratch_bufs;
    size_t scratch_bytes = 0;
    // this device's copy of the file-scope `scratch` array (onc_exscan_u32_noalloc)
    void* noalloc_scratch = nullptr;
    // optional GPU timing of the pure kernels (onc_ctx_timing)
    bool timing = false;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    double last_us = 0.0;
};

// Pure-kernel timing: events on the context stream around the launches of
// one call (not its set-up copies); read after the call's stream sync.
static void tmark(onc_ctx* c, int end) {
    if (c->timing) hipEventRecord(end ? c->t1 : c->t0, c->stream);
}
static int tread(onc_ctx* c) {
    if (!c->timing) return ONC_OK;
    float ms = 0.0f;
    CK(hipEventElapsedTime(&ms, c->t0, c->t1));
    c->last_us = 1000.0 * ms;
    return ONC_OK;
}

struct onc_module {
    onc_ctx* ctx = nullptr;
    hipModule_t mod = nullptr;
    std::string log;
};

struct onc_batch;

// Steps per replay of a batch's multi-step graph, and the one-step replays
// of the same batch after which it is captured.
#define ONC_GRAPH_STEPS 16
#define ONC_GRAPH_AFTER 64
// Default population thresholds of a batch's tile sizes (live cells), and
// the sizes above them. On the 1M-cell workload 16 / 8 cells per thread are
// 6-19 us per step faster than 8 / 4 from ~1M live cells up and 10-13 us
// slower below ~0.5M.
#define ONC_BIG_FROM 786432u
#define ONC_SMALL_BELOW 524288u
#define ONC_BIG_ITEMS 16
#define ONC_BIG_CITEMS 8

struct onc_sim {
    onc_ctx* ctx = nullptr;
    onc_sim_desc desc{};
    onc_dsim dsim{};        // host copy with device pointers
    onc_dstate* dstate = nullptr;
    std::vector<void*> allocs;
    uint64_t footprint = 0;
    size_t nvox = 0;
    uint32_t table = 0;     // bin table entries
    int key_bits = 0;
    hipFunction_t rule_fn = nullptr;  // nullptr: built-in tumour rule
    onc_module* rule_mod = nullptr;
    // translated grid rules of the rule module: passes per step and kernels
    onc_u32 tissue_passes = 0, chem_passes = 0;
    hipFunction_t tissue_fn = nullptr, chem_fn = nullptr;
    size_t nchem = 0;  // chemistry voxels
    onc_batch* owner = nullptr;
    onc_batch* solo = nullptr;
};

struct onc_batch {
    onc_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    std::vector<onc_sim*> sims;
    bool is_solo = false;
    bool dirty = true;
    // device tables
    onc_dsim* d_sims = nullptr;
    onc_u32* d_small = nullptr;  // prefixes, counters, err
    onc_u32* d_hist = nullptr;
    onc_u32* d_sort_status = nullptr;
    onc_u64* d_comp_status = nullptr;
    onc_u32* d_jobs = nullptr;
    onc_u32* d_quiet = nullptr;  // quiet sweep tiles: {state, value} per block and field buffer
    std::vector<onc_u32> jobs_h, jprefix_h;  // host copies of the sweep job table
    size_t cap_sims = 0, cap_tiles = 0, cap_ctiles = 0, cap_jobs = 0, cap_quiet = 0;
    onc_kargs ka{};
    onc_rule_args ra{};
    onc_u32 stencil_blocks = 0;
    uint64_t total_cap = 0;
    int passes = 1;
    // tuning
    int items = 8;    // cells per thread of a sort tile (this step's)
    int citems = 4;   // cells per thread of a compaction tile (this step's)
    // Tile sizes by population: items / citems are the tuned sizes (below)
    // until the batch's live cells reach big_from, then ONC_BIG_ITEMS /
    // ONC_BIG_CITEMS (if larger) until they fall below small_below. Fewer,
    // larger tiles shorten the look-back chains of a large population; small
    // populations need small tiles to spread over the chip. The same class
    // picks the Morton kernel's run-length digit counting (big only). The live count
    // comes back from the device asynchronously every ONC_GRAPH_STEPS steps
    // into one of two pinned slots; before a slot is reused the host waits for
    // its last copy, so the host runs at most ~2 x ONC_GRAPH_STEPS steps ahead
    // of the device and the count it acts on is that recent.
    // big_from = 0: the tuned sizes always (onc_batch_set_tuning sets this).
    int tuned_items = 8, tuned_citems = 4;
    uint32_t big_from = ONC_BIG_FROM, small_below = ONC_SMALL_BELOW;
    bool big = false;
    onc_u32* h_live = nullptr;  // pinned: two slots of 16 words
    hipEvent_t ev_live[2] = {nullptr, nullptr};
    bool live_used[2] = {false, false};
    uint64_t live_seq[2] = {0, 0};
    uint64_t live_next_seq = 1, live_seen_seq = 0, live_decided_seq = 0;
    uint32_t live_cells = 0;    // the newest count that came back
    int live_slot = 0;          // the slot the next readback uses
    int steps_since_readback = 0;
    int sbx = 64, sby = 4, sbz = 16;
    int fused = 1;    // forces and the sweep in one launch (k_forces_stencil)
    int force_waves = 0;  // cap on the one-wave blocks of the force walk (0: one per 64 cells of capacity)
    int fuse_rule = 1;    // built-in rule fused into compaction where the schema allows
    bool rule_compact = false;  // this batch runs k_rule_compact (set by batch_rebuild)
    onc_u32 fs_groups = 0;  // groups of 8 blocks of that launch
    // graphs: one step, and ONC_GRAPH_STEPS steps. A replay costs ~15 us of
    // GPU time before its first kernel starts (eager launches pipeline at ~3
    // us each), so a batch that keeps its members for a while replays the
    // longer graph; it is captured once the one-step graph of the same batch
    // has been replayed ONC_GRAPH_AFTER times (a capture of 16 steps costs
    // host time that a batch rebuilt every few ticks would not win back).
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph_n = nullptr;
    hipGraphExec_t exec_n = nullptr;
    uint64_t single_replays = 0;
    // profiling
    unsigned profile = 0;  // bit k: time phase k (onc_phase)
    bool prof_on = false;  // events are being recorded for the step being enqueued
    int prof_every = 1;    // time every prof_every-th eager step
    uint64_t prof_tick = 0;  // eager steps enqueued since profiling was set
    std::vector<unsigned> ev_mask;  // phases with both events recorded, per profiled step
    std::vector<hipEvent_t> ev;
    int ev_used = 0;
    double phase_us[ONC_PHASES] = {0};
    uint64_t phase_n[ONC_PHASES] = {0};
};

// Scratch of onc_exscan_u32_noalloc, for callers that cannot allocate: 256
// bytes holding the error word, then one 8-byte look-back status word per
// 4096-element tile of the largest input it accepts (2^30 elements), 2 MiB
// + 256 B in all. Every device that runs the library holds its own copy
// until the process ends. Host code never uses this name as an address: each
// context takes its device's address with hipGetSymbolAddress at creation
// (onc_ctx::noalloc_scratch) and passes that to the kernels.
#define ONC_NOALLOC_SCAN_MAX_N (1u << 30)
#define ONC_NOALLOC_ERR_WORDS 64u
#define ONC_NOALLOC_SCRATCH_BYTES (4u * ONC_NOALLOC_ERR_WORDS + 8u * (ONC_NOALLOC_SCAN_MAX_N / ONC_SCAN_MIN_TILE))
static __device__ __align__(256) unsigned char scratch[ONC_NOALLOC_SCRATCH_BYTES];

// ---------------------------------------------------------------------------
// errors / version
// ---------------------------------------------------------------------------
extern "C" const char* onc_last_error(void) { return g_err.c_str(); }
extern "C" int onc_abi_version(void) { return ONC_ABI_VERSION; }

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int onc_ctx_create(int device, onc_ctx** out) {
    if (!out) return fail(ONC_ERR_INVALID, "out is NULL");
    int n = 0;
    CK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(ONC_ERR_INVALID, "no such device");
    CK(hipSetDevice(device));
    auto* c = new onc_ctx();
    c->device = device;
    hipError_t e = hipGetDeviceProperties(&c->prop, device);
    if (e != hipSuccess) {
        delete c;
        return fail(ONC_ERR_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    }
    c->cus = c->prop.multiProcessorCount > 0 ? c->prop.multiProcessorCount : 256;
    c->arch = c->prop.gcnArchName;
    size_t colon = c->arch.find(':');
    if (colon != std::string::npos) c->arch = c->arch.substr(0, colon);
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(ONC_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    // this device's address of the no-alloc scratch (device current: set above)
    e = hipGetSymbolAddress(&c->noalloc_scratch, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !c->noalloc_scratch) {
        hipStreamDestroy(c->stream);
        delete c;
        return fail(ONC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    *out = c;
    return ONC_OK;
}

extern "C" int onc_ctx_destroy(onc_ctx* c) {
    if (!c) return ONC_OK;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (void* p : c->scratch_bufs) hipFree(p);
    if (c->t0) hipEventDestroy(c->t0);
    if (c->t1) hipEventDestroy(c->t1);
    hipStreamDestroy(c->stream);
    delete c;
    return ONC_OK;
}

extern "C" int onc_ctx_info(onc_ctx* c, onc_device_info* out) {
    if (!c || !out) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(c->device));
    memset(out, 0, sizeof(*out));
    snprintf(out->name, sizeof(out->name), "%s", c->prop.name);
    snprintf(out->arch, sizeof(out->arch), "%s", c->arch.c_str());
    out->compute_units = c->cus;
    out->xcds = c->cus >= 256 ? 8 : 1;
    size_t fr = 0, tot = 0;
    CK(hipMemGetInfo(&fr, &tot));
    out->hbm_bytes = tot;
    out->free_bytes = fr;
    return ONC_OK;
}

extern "C" int onc_dev_alloc(onc_ctx* c, size_t bytes, void** out) {
    if (!c || !out) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(c->device));
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 4);
    if (e != hipSuccess) return fail(ONC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    *out = p;
    return ONC_OK;
}
extern "C" int onc_dev_free(onc_ctx* c, void* p) {
    if (!c) return fail(ONC_ERR_INVALID, "NULL ctx");
    CK(hipSetDevice(c->device));
    CK(hipFree(p));
    return ONC_OK;
}
extern "C" int onc_host_alloc(onc_ctx* c, size_t bytes, void** out) {
    if (!c || !out) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(c->device));
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, bytes ? bytes : 4, hipHostMallocDefault);
    if (e != hipSuccess) return fail(ONC_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    *out = p;
    return ONC_OK;
}
extern "C" int onc_host_free(onc_ctx* c, void* p) {
    if (!c) return fail(ONC_ERR_INVALID, "NULL ctx");
    CK(hipSetDevice(c->device));
    if (p) CK(hipHostFree(p));
    return ONC_OK;
}
extern "C" int onc_h2d(onc_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return fail(ONC_ERR_INVALID, "NULL ctx");
    if (!bytes) return ONC_OK;
    CK(hipSetDevice(c->device));
    CK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return ONC_OK;
}
extern "C" int onc_d2h(onc_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c) return fail(ONC_ERR_INVALID, "NULL ctx");
    if (!bytes) return ONC_OK;
    CK(hipSetDevice(c->device));
    CK(hipStreamSynchronize(c->stream));
    CK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return ONC_OK;
}
extern "C" int onc_ctx_sync(onc_ctx* c) {
    if (!c) return fail(ONC_ERR_INVALID, "NULL ctx");
    CK(hipSetDevice(c->device));
    CK(hipStreamSynchronize(c->stream));
    return ONC_OK;
}

// ---------------------------------------------------------------------------
// rule modules (hiprtc)
// ---------------------------------------------------------------------------
extern "C" int onc_module_build(onc_ctx* c, const char* src, const char* const* opts, int n_opts,
                                onc_module** out) {
    if (!c || !src || !out) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(c->device));
    hiprtcProgram prog;
    const char* hdrs[1] = {onc_device_h_src};
    const char* names[1] = {"onc_device.h"};
    if (hiprtcCreateProgram(&prog, src, "onc_rule.hip", 1, hdrs, names) != HIPRTC_SUCCESS)
        return fail(ONC_ERR_COMPILE, "hiprtcCreateProgram failed");
    std::string arch = "--offload-arch=" + c->arch;
    std::vector<const char*> o = {arch.c_str(), "-O3", "-ffp-contract=off", "-std=c++17",
                                  "-fhip-fp32-correctly-rounded-divide-sqrt",
                                  "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature",
                                  "-Xclang", "-packed-fp32-ops"};
    for (int i = 0; i < n_opts; i++) o.push_back(opts[i]);
    hiprtcResult rc = hiprtcCompileProgram(prog, (int)o.size(), o.data());
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    if (rc != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(ONC_ERR_COMPILE, "hiprtc compilation failed:\n" + log);
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    std::vector<char> code(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    auto* m = new onc_module();
    m->ctx = c;
    m->log = log;
    hipError_t e = hipModuleLoadData(&m->mod, code.data());
    if (e != hipSuccess) {
        delete m;
        return fail(ONC_ERR_COMPILE, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    }
    *out = m;
    return ONC_OK;
}

extern "C" int onc_module_log(onc_module* m, char* buf, size_t len) {
    if (!m || !buf || !len) return fail(ONC_ERR_INVALID, "NULL argument");
    snprintf(buf, len, "%s", m->log.c_str());
    return ONC_OK;
}

extern "C" int onc_module_destroy(onc_module* m) {
    if (!m) return ONC_OK;
    hipSetDevice(m->ctx->device);
    hipModuleUnload(m->mod);
    delete m;
    return ONC_OK;
}

// ---------------------------------------------------------------------------
// sims
// ---------------------------------------------------------------------------
static uint32_t host_morton(uint32_t x, uint32_t y, uint32_t z, int fmt) {
    uint32_t k = 0;
    for (int b = 0; b < 10; b++) {
        k |= ((x >> b) & 1u) << (3 * b);
        k |= ((y >> b) & 1u) << (3 * b + 1);
        k |= ((z >> b) & 1u) << (3 * b + 2);
    }
    if (fmt == ONC_KEY_MORTON32) k |= (((x >> 10) & 1u) << 30) | (((y >> 10) & 1u) << 31);
    return k;
}

static int validate_desc(const onc_sim_desc* d) {
    if (!d) return fail(ONC_ERR_INVALID, "desc is NULL");
    if (d->key_format != ONC_KEY_MORTON30 && d->key_format != ONC_KEY_MORTON32)
        return fail(ONC_ERR_INVALID, "key_format must be 30 or 32");
    int lim_xy = d->key_format == ONC_KEY_MORTON32 ? 2048 : 1024;
    for (int i = 0; i < 3; i++) {
        int lim = i < 2 ? lim_xy : 1024;
        if (d->dims[i] < 1 || d->dims[i] > lim) return fail(ONC_ERR_INVALID, "dims out of range for key format");
    }
    if (host_morton((uint32_t)d->dims[0] - 1, (uint32_t)d->dims[1] - 1, (uint32_t)d->dims[2] - 1, d->key_format) >=
        0x80000000u)
        return fail(ONC_ERR_INVALID, "grid too large: the dense bin table needs Morton keys below 2^31");
    if (!(d->dx > 0.0f)) return fail(ONC_ERR_INVALID, "dx must be > 0");
    if (d->capacity < 1 || d->capacity > (1u << 30)) return fail(ONC_ERR_INVALID, "capacity must be in [1, 2^30]");
    if (d->n_cols < 4 || d->n_cols > ONC_MAX_COLS) return fail(ONC_ERR_INVALID, "n_cols must be in [4, 48]");
    for (int c = 0; c < 3; c++)
        if (d->col_types[c] != ONC_COL_F32) return fail(ONC_ERR_INVALID, "columns 0..2 (pos) must be f32");
    if (d->col_types[3] != ONC_COL_U32) return fail(ONC_ERR_INVALID, "column 3 (id) must be u32");
    if (d->n_fields < 0 || d->n_fields > ONC_MAX_FIELDS) return fail(ONC_ERR_INVALID, "n_fields must be in [0, 4]");
    if (d->n_params < 0 || d->n_params > ONC_MAX_PARAMS) return fail(ONC_ERR_INVALID, "n_params must be in [0, 32]");
    if (d->interaction_radius > d->dx) return fail(ONC_ERR_INVALID, "interaction_radius must be <= dx (bins are voxels)");
    if (d->interaction_radius < 0.0f || d->cell_radius < 0.0f || d->max_disp < 0.0f)
        return fail(ONC_ERR_INVALID, "negative mechanics parameter");
    for (int f = 0; f < d->n_fields; f++) {
        const onc_field_desc& fd = d->fields[f];
        if (fd.bc != ONC_BC_DIRICHLET && fd.bc != ONC_BC_POINT_SOURCE && fd.bc != ONC_BC_NONE)
            return fail(ONC_ERR_INVALID, "bad field bc");
        if (!(fd.lo <= fd.hi)) return fail(ONC_ERR_INVALID, "field lo > hi");
    }
    if (d->n_chem_cols < 0 || d->n_chem_cols > ONC_MAX_CHEM_COLS)
        return fail(ONC_ERR_INVALID, "n_chem_cols must be in [0, 16]");
    for (int c = 0; c < d->n_chem_cols; c++)
        if (d->chem_col_types[c] < ONC_COL_F32 || d->chem_col_types[c] > ONC_COL_U32)
            return fail(ONC_ERR_INVALID, "bad chemistry column type");
    return ONC_OK;
}

// Chemistry grid: max(1, dims / 2) voxels per axis.
static void chem_dims(const onc_sim_desc* d, int cd[3]) {
    for (int i = 0; i < 3; i++) cd[i] = std::max(1, d->dims[i] / 2);
}
static size_t chem_voxels(const onc_sim_desc* d) {
    int cd[3];
    chem_dims(d, cd);
    return (size_t)cd[0] * (size_t)cd[1] * (size_t)cd[2];
}

static void sim_layout(const onc_sim_desc* d, uint32_t* table, int* key_bits, size_t* nvox, uint64_t* bytes) {
    uint32_t maxkey = host_morton((uint32_t)d->dims[0] - 1, (uint32_t)d->dims[1] - 1, (uint32_t)d->dims[2] - 1,
                                  d->key_format);
    int kb = 1;
    while (kb < 32 && (maxkey >> kb) != 0u) kb++;
    *key_bits = kb;
    *table = maxkey + 1u;
    *nvox = (size_t)d->dims[0] * (size_t)d->dims[1] * (size_t)d->dims[2];
    uint64_t cap = d->capacity;
    uint64_t b = 0;
    b += 2ull * d->n_cols * cap * 4;  // state, two buffers
    b += 2ull * d->n_cols * cap * 4;  // daughters (sparse + compacted)
    b += 5ull * cap * 4;              // keys x2, vals x2, flags
    b += 16ull * cap;                 // packed sorted positions
    b += 2ull * d->n_fields * (*nvox) * 4;
    b += 8ull * (uint64_t)(*table);      // bins {start, len << 16 | tag}
    b += 2ull * d->n_chem_cols * chem_voxels(d) * 4;  // chemistry grid, two buffers
    b += 4ull * ((maxkey >> (3 * ONC_QCUBE)) + 1u);   // uptake stamps of the sweep's quiet tiles
    b += 256;
    *bytes = b;
}

extern "C" int onc_sim_footprint(const onc_sim_desc* d, uint64_t* bytes) {
    CKR(validate_desc(d));
    if (!bytes) return fail(ONC_ERR_INVALID, "bytes is NULL");
    uint32_t t;
    int kb;
    size_t nv;
    sim_layout(d, &t, &kb, &nv, bytes);
    return ONC_OK;
}

static int sim_alloc(onc_sim* s, size_t bytes, void** out) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 4);
    if (e != hipSuccess) return fail(ONC_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->allocs.push_back(p);
    *out = p;
    return ONC_OK;
}

static void fill_dsim_static(onc_sim* s) {
    const onc_sim_desc& d = s->desc;
    onc_dsim& D = s->dsim;
    for (int i = 0; i < 3; i++) D.dims[i] = d.dims[i];
    D.dx = d.dx;
    D.key_format = d.key_format;
    D.key_bits = s->key_bits;
    D.cap = d.capacity;
    D.seed_lo = (onc_u32)d.seed;
    D.seed_hi = (onc_u32)(d.seed >> 32);
    D.cell_radius = d.cell_radius;
    D.R = d.interaction_radius;
    D.k_rep = d.k_rep;
    D.k_adh = d.k_adh;
    D.mobility = d.mobility;
    D.max_disp = d.max_disp;
    D.n_fields = d.n_fields;
    D.n_cols = d.n_cols;
    memcpy(D.fields, d.fields, sizeof(D.fields));
    memcpy(D.params, d.params, sizeof(D.params));
    chem_dims(&d, D.cdims);
    D.n_chem = d.n_chem_cols;
    D.chem_passes = s->chem_passes;
}

static int batch_create_internal(onc_ctx* c, bool solo, onc_batch** out);
static int batch_commit(onc_batch* b);
static int sim_read_state(onc_sim* s, onc_dstate* st);

extern "C" int onc_sim_create(onc_ctx* c, const onc_sim_desc* d, onc_sim** out) {
    if (!c || !out) return fail(ONC_ERR_INVALID, "NULL argument");
    CKR(validate_desc(d));
    CK(hipSetDevice(c->device));
    auto* s = new onc_sim();
    s->ctx = c;
    s->desc = *d;
    sim_layout(d, &s->table, &s->key_bits, &s->nvox, &s->footprint);
    fill_dsim_static(s);
    auto cleanup = [&](int rc) {
        for (void* p : s->allocs) hipFree(p);
        delete s;
        return rc;
    };
    const size_t cap = d->capacity;
    void* p = nullptr;
    int rc;
    if ((rc = sim_alloc(s, 2ull * d->n_cols * cap * 4, &p)) != ONC_OK) return cleanup(rc);
    for (int b = 0; b < 2; b++)
        for (int col = 0; col < d->n_cols; col++)
            s->dsim.st[b][col] = (onc_u32*)p + ((size_t)b * d->n_cols + col) * cap;
    if ((rc = sim_alloc(s, 2ull * d->n_cols * cap * 4, &p)) != ONC_OK) return cleanup(rc);
    for (int col = 0; col < d->n_cols; col++) {
        s->dsim.dau[col] = (onc_u32*)p + (size_t)col * cap;
        s->dsim.dau2[col] = (onc_u32*)p + ((size_t)d->n_cols + col) * cap;
    }
    if ((rc = sim_alloc(s, 16ull * cap, &p)) != ONC_OK) return cleanup(rc);
    s->dsim.pos4 = (float*)p;
    if ((rc = sim_alloc(s, 5ull * cap * 4, &p)) != ONC_OK) return cleanup(rc);
    s->dsim.keys[0] = (onc_u32*)p;
    s->dsim.keys[1] = (onc_u32*)p + cap;
    s->dsim.vals[0] = (onc_u32*)p + 2 * cap;
    s->dsim.vals[1] = (onc_u32*)p + 3 * cap;
    s->dsim.flags = (onc_u32*)p + 4 * cap;
    if (d->n_fields > 0) {
        if ((rc = sim_alloc(s, 2ull * d->n_fields * s->nvox * 4, &p)) != ONC_OK) return cleanup(rc);
        for (int b = 0; b < 2; b++)
            for (int f = 0; f < d->n_fields; f++)
                s->dsim.fld[b][f] = (float*)p + ((size_t)b * d->n_fields + f) * s->nvox;
    }
    if ((rc = sim_alloc(s, 8ull * s->table, &p)) != ONC_OK) return cleanup(rc);
    s->dsim.bins = (onc_u32*)p;
    s->dsim.table = s->table;
    const size_t nstamps = (size_t)((s->table - 1u) >> (3 * ONC_QCUBE)) + 1u;
    if ((rc = sim_alloc(s, 4ull * nstamps, &p)) != ONC_OK) return cleanup(rc);
    s->dsim.qstamp = (onc_u32*)p;
    s->nchem = chem_voxels(d);
    void* chem_base = nullptr;
    if (d->n_chem_cols > 0) {
        if ((rc = sim_alloc(s, 2ull * d->n_chem_cols * s->nchem * 4, &chem_base)) != ONC_OK) return cleanup(rc);
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < d->n_chem_cols; c++)
                s->dsim.chem[b][c] = (onc_u32*)chem_base + ((size_t)b * d->n_chem_cols + c) * s->nchem;
    }
    if ((rc = sim_alloc(s, sizeof(onc_dstate), &p)) != ONC_OK) return cleanup(rc);
    s->dstate = (onc_dstate*)p;
    s->dsim.state = s->dstate;
    // initial device state
    hipStream_t st = c->stream;
    onc_dstate z{};
    z.stop = 0xFFFFFFFFu;
    hipError_t e = hipMemcpyAsync(s->dstate, &z, sizeof(z), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = onc_launch_fill_u32(s->dsim.bins, 0xFFFFFFFFu, 2ull * s->table, st);
    if (e == hipSuccess) e = onc_launch_fill_u32(s->dsim.qstamp, 0u, nstamps, st);
    if (e == hipSuccess && chem_base)
        e = onc_launch_fill_u32((onc_u32*)chem_base, 0u, 2ull * d->n_chem_cols * s->nchem, st);
    for (int f = 0; f < d->n_fields && e == hipSuccess; f++) {
        onc_dfield fd;
        memcpy(&fd, &d->fields[f], sizeof(fd));
        e = onc_launch_field_init(s->dsim.fld[0][f], d->dims[0], d->dims[1], d->dims[2], fd, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return cleanup(fail(ONC_ERR_HIP, std::string("sim init: ") + hipGetErrorString(e)));
    if ((rc = batch_create_internal(c, true, &s->solo)) != ONC_OK) return cleanup(rc);
    s->solo->sims.push_back(s);
    s->owner = s->solo;
    *out = s;
    return ONC_OK;
}

static int batch_destroy_internal(onc_batch* b);

extern "C" int onc_sim_destroy(onc_sim* s) {
    if (!s) return ONC_OK;
    hipSetDevice(s->ctx->device);
    if (s->owner && !s->owner->is_solo) {
        onc_batch* b = s->owner;
        hipStreamSynchronize(b->stream);
        b->sims.erase(std::remove(b->sims.begin(), b->sims.end(), s), b->sims.end());
        b->dirty = true;
    }
    if (s->solo) {
        hipStreamSynchronize(s->solo->stream);
        s->solo->sims.clear();
        batch_destroy_internal(s->solo);
    }
    hipDeviceSynchronize();
    for (void* p : s->allocs) hipFree(p);
    delete s;
    return ONC_OK;
}

// A u32 word of a module (0 when the module does not define it).
static int module_word(onc_module* m, const char* name, onc_u32* out) {
    hipDeviceptr_t gp = nullptr;
    size_t bytes = 0;
    *out = 0;
    if (hipModuleGetGlobal(&gp, &bytes, m->mod, name) != hipSuccess) {
        (void)hipGetLastError();
        return ONC_OK;
    }
    if (bytes != sizeof(onc_u32)) return fail(ONC_ERR_INVALID, std::string("module word ") + name + " is not 32-bit");
    CK(hipMemcpy(out, gp, sizeof(onc_u32), hipMemcpyDeviceToHost));
    return ONC_OK;
}

// Copy the chemistry columns from buffer `from` to the other one.
static int chem_copy(onc_sim* s, int from) {
    for (int c = 0; c < s->desc.n_chem_cols; c++)
        CK(hipMemcpy(s->dsim.chem[from ^ 1][c], s->dsim.chem[from][c], s->nchem * 4, hipMemcpyDeviceToDevice));
    return ONC_OK;
}

extern "C" int onc_sim_set_rule(onc_sim* s, onc_module* m, const char* name) {
    if (!s) return fail(ONC_ERR_INVALID, "NULL sim");
    CK(hipSetDevice(s->ctx->device));
    onc_u32 tp = 0, cp = 0;
    hipFunction_t f = nullptr, tf = nullptr, cf = nullptr;
    if (m) {
        if (!name) return fail(ONC_ERR_INVALID, "kernel name is NULL");
        hipError_t e = hipModuleGetFunction(&f, m->mod, name);
        if (e != hipSuccess) return fail(ONC_ERR_INVALID, std::string("no kernel named ") + name + " in module");
        CKR(module_word(m, "onc_grid_tissue_passes", &tp));
        CKR(module_word(m, "onc_grid_chem_passes", &cp));
        if (tp && hipModuleGetFunction(&tf, m->mod, "onc_tissue_pass") != hipSuccess)
            return fail(ONC_ERR_INVALID, "module declares tissue passes but has no onc_tissue_pass kernel");
        if (cp && hipModuleGetFunction(&cf, m->mod, "onc_chem_pass") != hipSuccess)
            return fail(ONC_ERR_INVALID, "module declares chemistry passes but has no onc_chem_pass kernel");
        if (tp && s->desc.n_fields == 0) return fail(ONC_ERR_INVALID, "tissue rules need tissue fields");
        if (cp && s->desc.n_chem_cols == 0) return fail(ONC_ERR_INVALID, "chemistry rules need chemistry columns");
    }
    if (cp != s->chem_passes && s->desc.n_chem_cols > 0) {
        // the committed chemistry state stays where the new pass count expects it
        onc_dstate st;
        CKR(sim_read_state(s, &st));
        const int cur_old = (int)((st.step * s->chem_passes) & 1u), cur_new = (int)((st.step * cp) & 1u);
        if (cur_old != cur_new) CKR(chem_copy(s, cur_old));
    }
    s->rule_fn = f;
    s->rule_mod = m;
    s->tissue_passes = tp;
    s->chem_passes = cp;
    s->tissue_fn = tf;
    s->chem_fn = cf;
    s->dsim.chem_passes = cp;
    if (s->owner) s->owner->dirty = true;
    return ONC_OK;
}

extern "C" int onc_sim_set_params(onc_sim* s, const float* params, int n) {
    if (!s || (n > 0 && !params)) return fail(ONC_ERR_INVALID, "NULL argument");
    if (n < 0 || n > ONC_MAX_PARAMS) return fail(ONC_ERR_INVALID, "n must be in [0, 32]");
    for (int i = 0; i < n; i++) s->desc.params[i] = params[i];
    s->desc.n_params = std::max(s->desc.n_params, n);
    memcpy(s->dsim.params, s->desc.params, sizeof(s->dsim.params));
    if (s->owner) s->owner->dirty = true;
    return ONC_OK;
}

// Device state of a sim with its last step committed.
static int sim_read_state(onc_sim* s, onc_dstate* st) {
    CKR(batch_commit(s->owner));
    CK(hipMemcpy(st, s->dstate, sizeof(*st), hipMemcpyDeviceToHost));
    return ONC_OK;
}

extern "C" int onc_sim_upload_cells(onc_sim* s, const void* const* cols, uint32_t n, uint32_t next_id) {
    if (!s || (n > 0 && !cols)) return fail(ONC_ERR_INVALID, "NULL argument");
    if (n > s->desc.capacity) return fail(ONC_ERR_CAPACITY, "more cells than capacity");
    CK(hipSetDevice(s->ctx->device));
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    for (int c = 0; c < s->desc.n_cols && n > 0; c++) {
        if (!cols[c]) return fail(ONC_ERR_INVALID, "column pointer is NULL");
        CK(hipMemcpy(s->dsim.st[st.parity][c], cols[c], (size_t)n * 4, hipMemcpyHostToDevice));
    }
    st.n = n;
    st.split = n;
    st.next_id = next_id;
    CK(hipMemcpy(s->dstate, &st, sizeof(st), hipMemcpyHostToDevice));
    return ONC_OK;
}

extern "C" int onc_sim_download_cells(onc_sim* s, void* const* cols, uint32_t cap, uint32_t* n) {
    if (!s) return fail(ONC_ERR_INVALID, "NULL sim");
    CK(hipSetDevice(s->ctx->device));
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    if (n) *n = st.n;
    // survivors of the last step, then its daughters (see onc_dstate)
    const uint32_t m = std::min(cap, st.n), m1 = std::min(m, st.split), m2 = m - m1;
    if (cols && m > 0)
        for (int c = 0; c < s->desc.n_cols; c++) {
            if (!cols[c]) continue;
            if (m1) CK(hipMemcpy(cols[c], s->dsim.st[st.parity][c], (size_t)m1 * 4, hipMemcpyDeviceToHost));
            if (m2) CK(hipMemcpy((char*)cols[c] + (size_t)m1 * 4, s->dsim.dau2[c], (size_t)m2 * 4, hipMemcpyDeviceToHost));
        }
    return ONC_OK;
}

static int ctx_scratch(onc_ctx* c, size_t bytes, void** out);

extern "C" int onc_sim_metrics(onc_sim* s, onc_module* m, int nvals, uint32_t* out) {
    if (!s || !m || !out) return fail(ONC_ERR_INVALID, "NULL argument");
    if (nvals < 1 || nvals > ONC_MAX_METRICS) return fail(ONC_ERR_INVALID, "nvals must be in [1, 64]");
    CK(hipSetDevice(s->ctx->device));
    hipFunction_t fm = nullptr, ff = nullptr;
    if (hipModuleGetFunction(&fm, m->mod, "onc_metric_main") != hipSuccess ||
        hipModuleGetFunction(&ff, m->mod, "onc_metric_final") != hipSuccess)
        return fail(ONC_ERR_INVALID, "module has no metric kernels (onc_metric_main, onc_metric_final)");
    // The kernels write as many words per block as the module's metric class
    // has fields; the scratch below is sized by nvals, so the two must agree.
    hipDeviceptr_t gp = nullptr;
    size_t gbytes = 0;
    uint32_t nv = 0;
    if (hipModuleGetGlobal(&gp, &gbytes, m->mod, "onc_metric_nvals") != hipSuccess || gbytes != sizeof(nv))
        return fail(ONC_ERR_INVALID, "module has no onc_metric_nvals word");
    CK(hipMemcpy(&nv, gp, sizeof(nv), hipMemcpyDeviceToHost));
    if ((uint32_t)nvals != nv)
        return fail(ONC_ERR_INVALID, "nvals (" + std::to_string(nvals) + ") differs from the module's metric count (" +
                                         std::to_string(nv) + ")");
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    onc_ctx* c = s->ctx;
    const uint32_t per = 256u * ONC_METRIC_CPT;
    const uint32_t nb = std::max<uint32_t>(1u, (uint32_t)(((uint64_t)st.n + per - 1) / per));
    const size_t o_part = (sizeof(onc_dsim) + 255) & ~(size_t)255;
    const size_t o_out = o_part + (((size_t)nb * nvals * 4 + 255) & ~(size_t)255);
    void* base;
    CKR(ctx_scratch(c, o_out + (size_t)nvals * 4, &base));
    char* B = (char*)base;
    CK(hipMemcpyAsync(B, &s->dsim, sizeof(onc_dsim), hipMemcpyHostToDevice, c->stream));
    onc_metric_args a{};
    a.S = (const onc_dsim*)B;
    a.n = st.n;
    a.split = st.split;
    a.parity = st.parity;
    a.partials = (onc_u32*)(B + o_part);
    void* pa[] = {&a};
    CK(hipModuleLaunchKernel(fm, nb, 1, 1, 256, 1, 1, 0, c->stream, pa, nullptr));
    onc_metric_fin_args f{};
    f.partials = a.partials;
    f.nblocks = nb;
    f.out = (onc_u32*)(B + o_out);
    void* pf[] = {&f};
    CK(hipModuleLaunchKernel(ff, 1, 1, 1, 256, 1, 1, 0, c->stream, pf, nullptr));
    CK(hipMemcpyAsync(out, f.out, (size_t)nvals * 4, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return ONC_OK;
}

extern "C" int onc_sim_upload_field(onc_sim* s, int f, const float* host) {
    if (!s || !host) return fail(ONC_ERR_INVALID, "NULL argument");
    if (f < 0 || f >= s->desc.n_fields) return fail(ONC_ERR_INVALID, "no such field");
    CK(hipSetDevice(s->ctx->device));
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    CK(hipMemcpy(s->dsim.fld[st.parity][f], host, s->nvox * 4, hipMemcpyHostToDevice));
    // the batch's record of uniform sweep tiles no longer holds
    if (s->owner && s->owner->d_quiet) {
        CK(hipMemsetAsync(s->owner->d_quiet, 0, s->owner->cap_quiet * 4, s->owner->stream));
        CK(hipStreamSynchronize(s->owner->stream));
    }
    return ONC_OK;
}

extern "C" int onc_sim_download_field(onc_sim* s, int f, float* host) {
    if (!s || !host) return fail(ONC_ERR_INVALID, "NULL argument");
    if (f < 0 || f >= s->desc.n_fields) return fail(ONC_ERR_INVALID, "no such field");
    CK(hipSetDevice(s->ctx->device));
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    CK(hipMemcpy(host, s->dsim.fld[st.parity][f], s->nvox * 4, hipMemcpyDeviceToHost));
    return ONC_OK;
}

static int chem_args(onc_sim* s, int col, const void* host, onc_u32** dev) {
    if (!s || !host) return fail(ONC_ERR_INVALID, "NULL argument");
    if (col < 0 || col >= s->desc.n_chem_cols) return fail(ONC_ERR_INVALID, "no such chemistry column");
    CK(hipSetDevice(s->ctx->device));
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    *dev = s->dsim.chem[(st.step * s->chem_passes) & 1u][col];
    return ONC_OK;
}

extern "C" int onc_sim_upload_chem(onc_sim* s, int col, const void* host) {
    onc_u32* dev = nullptr;
    CKR(chem_args(s, col, host, &dev));
    CK(hipMemcpy(dev, host, s->nchem * 4, hipMemcpyHostToDevice));
    return ONC_OK;
}

extern "C" int onc_sim_download_chem(onc_sim* s, int col, void* host) {
    onc_u32* dev = nullptr;
    CKR(chem_args(s, col, host, &dev));
    CK(hipMemcpy(host, dev, s->nchem * 4, hipMemcpyDeviceToHost));
    return ONC_OK;
}

extern "C" int onc_sim_set_stop(onc_sim* s, uint32_t stop) {
    if (!s) return fail(ONC_ERR_INVALID, "NULL sim");
    CK(hipSetDevice(s->ctx->device));
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    st.stop = stop;
    CK(hipMemcpy(s->dstate, &st, sizeof(st), hipMemcpyHostToDevice));
    return ONC_OK;
}

extern "C" int onc_sim_set_step(onc_sim* s, uint32_t step) {
    if (!s) return fail(ONC_ERR_INVALID, "NULL sim");
    CK(hipSetDevice(s->ctx->device));
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    if (step != st.step) {
        // A bin entry is valid while its tag is the running step's, and only
        // occupied keys are rewritten each step; the per-step refill of one
        // slice of the table assumes the steps go one by one. Entries left by
        // other steps can carry tags that the coming steps reuse, so they would
        // read as valid bins of voxels the cells do not occupy: clear the
        // table, as at creation.
        hipError_t e = onc_launch_fill_u32(s->dsim.bins, 0xFFFFFFFFu, 2ull * s->table, s->ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
        if (e != hipSuccess) return fail(ONC_ERR_HIP, std::string("set_step: ") + hipGetErrorString(e));
    }
    if (s->desc.n_chem_cols > 0) {
        // the committed chemistry state stays current under the new step count
        const int cur_old = (int)((st.step * s->chem_passes) & 1u), cur_new = (int)((step * s->chem_passes) & 1u);
        if (cur_old != cur_new) CKR(chem_copy(s, cur_old));
    }
    st.step = step;
    CK(hipMemcpy(s->dstate, &st, sizeof(st), hipMemcpyHostToDevice));
    // the uptake stamps count steps: the batch's record of uniform sweep
    // tiles no longer lines up with them
    if (s->owner && s->owner->d_quiet) {
        CK(hipMemsetAsync(s->owner->d_quiet, 0, s->owner->cap_quiet * 4, s->owner->stream));
        CK(hipStreamSynchronize(s->owner->stream));
    }
    return ONC_OK;
}

extern "C" int onc_sim_counts(onc_sim* s, uint32_t* n, uint32_t* next_id, uint32_t* steps) {
    if (!s) return fail(ONC_ERR_INVALID, "NULL sim");
    CK(hipSetDevice(s->ctx->device));
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    if (n) *n = st.n;
    if (next_id) *next_id = st.next_id;
    if (steps) *steps = st.step;
    return ONC_OK;
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
static int batch_create_internal(onc_ctx* c, bool solo, onc_batch** out) {
    auto* b = new onc_batch();
    b->ctx = c;
    b->is_solo = solo;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return fail(ONC_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = b;
    return ONC_OK;
}

static void batch_drop_graphs(onc_batch* b) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->exec_n) hipGraphExecDestroy(b->exec_n);
    if (b->graph_n) hipGraphDestroy(b->graph_n);
    b->exec = nullptr;
    b->graph = nullptr;
    b->exec_n = nullptr;
    b->graph_n = nullptr;
    b->single_replays = 0;
}

static void batch_free_tables(onc_batch* b) {
    batch_drop_graphs(b);
    hipFree(b->d_sims);
    hipFree(b->d_small);
    hipFree(b->d_hist);
    hipFree(b->d_sort_status);
    hipFree(b->d_comp_status);
    hipFree(b->d_jobs);
    hipFree(b->d_quiet);
    b->d_quiet = nullptr;
    b->cap_quiet = 0;
    b->d_sims = nullptr;
    b->d_small = nullptr;
    b->d_hist = nullptr;
    b->d_sort_status = nullptr;
    b->d_comp_status = nullptr;
    b->d_jobs = nullptr;
    b->cap_sims = b->cap_tiles = b->cap_ctiles = b->cap_jobs = 0;
}

static int batch_destroy_internal(onc_batch* b) {
    if (!b) return ONC_OK;
    hipStreamSynchronize(b->stream);
    batch_free_tables(b);
    for (hipEvent_t e : b->ev) hipEventDestroy(e);
    for (hipEvent_t e : b->ev_live)
        if (e) hipEventDestroy(e);
    if (b->h_live) hipHostFree(b->h_live);
    hipStreamDestroy(b->stream);
    delete b;
    return ONC_OK;
}

static onc_u32 force_waves(onc_batch* b);

// Blocks of one (sim, field) sweep with blocks of nt threads (256 for
// k_stencil, 64 for k_forces_stencil); must match onc_stencil_block: rows
// whose length is a multiple of 4 use the float4 tiling
// (onc_stencil_tw(X/4, min(sbx, nt)) float4 columns by nt / that many rows),
// others min(sbx, nt) x nt / min(sbx, nt) scalar tiles; both march sbz planes.
static onc_u32 stencil_blocks(int X, int Y, int Z, int sbx, int sbz, int nt) {
    const int nz = (Z + sbz - 1) / sbz;
    const int bx = sbx < nt ? sbx : nt;
    if ((X & 3) == 0) {
        const int X4 = X >> 2, tw = onc_stencil_tw(X4, bx), th = nt / tw;
        return (onc_u32)(((X4 + tw - 1) / tw) * ((Y + th - 1) / th) * nz);
    }
    const int by = nt / bx;
    return (onc_u32)(((X + bx - 1) / bx) * ((Y + by - 1) / by) * nz);
}

// (Re)build the device segment table, plan arrays and stencil job table.
static int batch_rebuild(onc_batch* b) {
    CK(hipSetDevice(b->ctx->device));
    CK(hipStreamSynchronize(b->stream));
    const size_t S = b->sims.size();
    // tables sized for the tuned (smallest) tiles; the tile sizes of the
    // coming steps follow the members' live cells
    const size_t tile = (size_t)256 * b->tuned_items, ctile = (size_t)256 * b->tuned_citems;
    // (read directly: the stream is idle, and sim_read_state would commit
    // through this batch, whose rebuild is under way)
    uint64_t live = 0;
    for (onc_sim* s : b->sims) {
        onc_dstate st;
        CK(hipMemcpy(&st, s->dstate, sizeof(st), hipMemcpyDeviceToHost));
        live += st.pending ? (uint64_t)st.nsurv + st.ndiv : st.n;
    }
    b->big = b->big_from != 0 && live >= b->big_from;
    b->items = b->big ? std::max(b->tuned_items, ONC_BIG_ITEMS) : b->tuned_items;
    b->citems = b->big ? std::max(b->tuned_citems, ONC_BIG_CITEMS) : b->tuned_citems;
    b->live_used[0] = b->live_used[1] = false;  // the stream is idle: any copy has landed
    b->live_seen_seq = b->live_decided_seq = b->live_next_seq - 1;
    b->steps_since_readback = 0;
    size_t max_tiles = 0, max_ctiles = 0;
    int passes = 1;
    uint64_t total_cap = 0;
    std::vector<onc_dsim> host(S);
    std::vector<onc_u32> jobs, jprefix;
    onc_u32 blocks = 0;
    // quiet sweep tiles, unless a sim's tissue rules also write the fields
    bool quiet = true;
    for (size_t i = 0; i < S; i++) {
        onc_sim* s = b->sims[i];
        host[i] = s->dsim;
        max_tiles += (s->desc.capacity + tile - 1) / tile;
        max_ctiles += (s->desc.capacity + ctile - 1) / ctile;
        passes = std::max(passes, (s->key_bits + 7) / 8);
        total_cap += s->desc.capacity;
        const int X = s->desc.dims[0], Y = s->desc.dims[1], Z = s->desc.dims[2];
        const onc_u32 nb = stencil_blocks(X, Y, Z, b->sbx, b->sbz, b->fused ? 64 : 256);
        host[i].job0 = (onc_u32)jobs.size();
        if (s->tissue_passes) quiet = false;
        for (int f = 0; f < s->desc.n_fields; f++) {
            jobs.push_back((onc_u32)(i << 8) | (onc_u32)f);
            jprefix.push_back(blocks);
            blocks += nb;
        }
    }
    if (max_tiles == 0) max_tiles = 1;
    if (max_ctiles == 0) max_ctiles = 1;
    jprefix.push_back(blocks);
    // (re)allocate
    batch_drop_graphs(b);
    if (S > b->cap_sims || max_tiles > b->cap_tiles || max_ctiles > b->cap_ctiles || jobs.size() + 1 > b->cap_jobs) {
        batch_free_tables(b);
        size_t cs = std::max<size_t>(S, 1), ct = max_tiles, cct = max_ctiles, cj = jobs.size() + 1;
        void* p;
        if (hipMalloc(&p, cs * sizeof(onc_dsim)) != hipSuccess) return fail(ONC_ERR_NOMEM, "batch table");
        b->d_sims = (onc_dsim*)p;
        if (hipMalloc(&p, (3 * (cs + 1) + ONC_NCOUNTERS + 16) * 4) != hipSuccess) return fail(ONC_ERR_NOMEM, "batch small");
        b->d_small = (onc_u32*)p;
        if (hipMalloc(&p, (size_t)ONC_HIST_COPIES * cs * 4 * 256 * 4) != hipSuccess)
            return fail(ONC_ERR_NOMEM, "batch hist");
        b->d_hist = (onc_u32*)p;
        CK(hipMemsetAsync(b->d_hist, 0, (size_t)ONC_HIST_COPIES * cs * 4 * 256 * 4, b->stream));
        if (hipMalloc(&p, (size_t)ONC_DEV_MAX_PASSES * ct * 256 * 4) != hipSuccess) return fail(ONC_ERR_NOMEM, "sort status");
        b->d_sort_status = (onc_u32*)p;
        if (hipMalloc(&p, cct * 8) != hipSuccess) return fail(ONC_ERR_NOMEM, "compaction status");
        b->d_comp_status = (onc_u64*)p;
        if (hipMalloc(&p, 2 * cj * 4) != hipSuccess) return fail(ONC_ERR_NOMEM, "stencil jobs");
        b->d_jobs = (onc_u32*)p;
        b->cap_sims = cs;
        b->cap_tiles = ct;
        b->cap_ctiles = cct;
        b->cap_jobs = cj;
        CK(hipMemsetAsync(b->d_small, 0, (3 * (cs + 1) + ONC_NCOUNTERS + 16) * 4, b->stream));
    }
    if ((size_t)blocks * 4 > b->cap_quiet) {
        hipFree(b->d_quiet);
        b->d_quiet = nullptr;
        b->cap_quiet = 0;
        void* p;
        if (hipMalloc(&p, (size_t)blocks * 16) != hipSuccess) return fail(ONC_ERR_NOMEM, "quiet tiles");
        b->d_quiet = (onc_u32*)p;
        b->cap_quiet = (size_t)blocks * 4;
    }
    // every tile starts as not known to be uniform (the tiling may have changed)
    if (b->d_quiet) CK(hipMemsetAsync(b->d_quiet, 0, b->cap_quiet * 4, b->stream));
    if (S) CK(hipMemcpyAsync(b->d_sims, host.data(), S * sizeof(onc_dsim), hipMemcpyHostToDevice, b->stream));
    b->jobs_h = jobs;
    b->jprefix_h = jprefix;
    std::vector<onc_u32> jj(jobs);
    jj.resize(b->cap_jobs, 0);
    std::vector<onc_u32> jp(jprefix);
    jp.resize(b->cap_jobs, blocks);
    CK(hipMemcpyAsync(b->d_jobs, jj.data(), b->cap_jobs * 4, hipMemcpyHostToDevice, b->stream));
    CK(hipMemcpyAsync(b->d_jobs + b->cap_jobs, jp.data(), b->cap_jobs * 4, hipMemcpyHostToDevice, b->stream));
    CK(hipStreamSynchronize(b->stream));
    onc_kargs& a = b->ka;
    a.sims = b->d_sims;
    a.nsims = (onc_u32)S;
    a.passes = (onc_u32)passes;
    a.tile = (onc_u32)(256 * b->items);
    a.max_tiles = (onc_u32)max_tiles;
    a.xcd_tiles = onc_sort_xcd_tiles(b->items, b->ctx->cus);
    a.quiet = quiet && blocks ? b->d_quiet : nullptr;
    a.nqb = blocks;
    a.ctile = (onc_u32)(256 * b->citems);
    a.max_ctiles = (onc_u32)max_ctiles;
    a.tile_prefix = b->d_small;
    a.cell_prefix = b->d_small + (b->cap_sims + 1);
    a.ctile_prefix = b->d_small + 2 * (b->cap_sims + 1);
    a.counters = b->d_small + 3 * (b->cap_sims + 1);
    a.err = a.counters + ONC_NCOUNTERS;
    a.hist = b->d_hist;
    a.sort_status = b->d_sort_status;
    a.comp_status = b->d_comp_status;
    a.jobs = b->d_jobs;
    a.jobs_prefix = b->d_jobs + b->cap_jobs;
    a.njobs = (onc_u32)jobs.size();
    a.sbx = b->sbx;
    a.sby = b->sby;
    a.sbz = b->sbz;
    b->ra.sims = b->d_sims;
    b->ra.nsims = (onc_u32)S;
    b->ra.passes = (onc_u32)passes;
    b->ra.cell_prefix = a.cell_prefix;
    b->stencil_blocks = blocks;
    b->total_cap = total_cap;
    // the fused rule + compaction serves the built-in rule on the six-column
    // tumour schema (every other column layout goes through k_rule_tumor,
    // which copies extra columns, and k_compact)
    b->rule_compact = b->fuse_rule != 0 && S > 0;
    for (size_t i = 0; i < S; i++)
        if (b->sims[i]->rule_fn || b->sims[i]->desc.n_cols != 6) b->rule_compact = false;
    a.fs_sgroups = (blocks + 7u) / 8u;
    b->fs_groups = a.fs_sgroups + force_waves(b) / 8u;  // groups of 8 one-wave blocks
    b->passes = passes;
    b->dirty = false;
    return ONC_OK;
}

// Drains the batch's stream and commits its sims' last step (the plan
// kernel in commit mode), so host reads see complete state.
static int batch_commit(onc_batch* b) {
    if (!b) return ONC_OK;
    CK(hipSetDevice(b->ctx->device));
    if (b->dirty) CKR(batch_rebuild(b));
    if (b->sims.empty()) return ONC_OK;
    CK(onc_launch_plan(b->ka, ONC_PLAN_COMMIT, b->stream));
    CK(hipStreamSynchronize(b->stream));
    return ONC_OK;
}

static int grid_cells(onc_batch* b) {
    uint64_t need = (b->total_cap + 255) / 256;
    uint64_t lim = (uint64_t)b->ctx->cus * 8;
    return (int)std::max<uint64_t>(1, std::min(need, lim));
}
// Forces: one-wave blocks, by default one per 64 cells of capacity (up to
// 65528), so that each wave walks about one chunk and a slow chunk (dense
// bins) holds a single wave slot while the dispatcher keeps the others busy;
// the force_waves tuning caps the count (each wave then strides through its
// XCD's chunks). Always a multiple of 8 (onc_for_chunks_xcd).
static onc_u32 force_waves(onc_batch* b) {
    uint64_t n = std::min<uint64_t>((b->total_cap + 63) / 64, 65528);
    if (b->force_waves > 0) n = std::min<uint64_t>(n, (uint64_t)b->force_waves);
    return (onc_u32)((std::max<uint64_t>(n, 8) + 7) / 8 * 8);
}
static int grid_forces(onc_batch* b) { return (int)force_waves(b); }
// Tile kernels (look-back scans): one block per tile the sims' capacity can
// need; blocks past this step's live tiles leave at once. A tile waits only
// on lower tiles, whose blocks were dispatched before it, so every wait ends;
// a smaller grid whose blocks loop over tiles could hold every resident block
// on a tile owned by a block that is not yet resident.
static int grid_tiles(onc_batch* b) { return (int)std::max<uint64_t>(1, b->ka.max_tiles); }
static int grid_ctiles(onc_batch* b) { return (int)std::max<uint64_t>(1, b->ka.max_ctiles); }
// Morton: up to four blocks per CU. Each holds an LDS histogram that it adds
// to global memory once per sim it touches; a block walks its share of the
// cells in serial rounds of dependent loads, so more blocks (shorter shares)
// finish sooner: four per CU instead of one made Morton 43 % faster on the
// 64-sim ensemble (eight were no faster), and the extra histogram atomics are
// spread over 8 copies.
static int grid_morton(onc_batch* b) {
    uint64_t need = (b->total_cap + 255) / 256;
    uint64_t lim = (uint64_t)b->ctx->cus * 4;
    return (int)std::max<uint64_t>(1, std::min(need, lim));
}

// Translated tissue and chemistry rules (all sims of a batch share one
// module): tissue passes, then a copy that leaves the fields in fld[par ^ 1]
// when their count is odd, then the chemistry passes. One launch per pass for
// the whole batch, blockIdx.y = sim.
static int launch_grid_rules(onc_batch* b) {
    if (b->sims.empty()) return ONC_OK;
    const onc_sim* s0 = b->sims[0];
    if (!s0->tissue_passes && !s0->chem_passes) return ONC_OK;
    size_t nvox = 0, nchem = 0;
    for (const onc_sim* s : b->sims) {
        nvox = std::max(nvox, s->nvox);
        nchem = std::max(nchem, s->nchem);
    }
    onc_grid_args g{};
    g.sims = b->ka.sims;
    g.nsims = (onc_u32)b->sims.size();
    void* args[] = {&g};
    const unsigned cap = (unsigned)b->ctx->cus * 4u;
    const unsigned gx_t = (unsigned)std::min<size_t>((nvox + 255) / 256, cap);
    const unsigned gx_c = (unsigned)std::min<size_t>((nchem + 255) / 256, cap);
    for (onc_u32 j = 0; j < s0->tissue_passes; j++) {
        g.pass = j;
        CK(hipModuleLaunchKernel(s0->tissue_fn, gx_t, g.nsims, 1, 256, 1, 1, 0, b->stream, args, nullptr));
    }
    if (s0->tissue_passes & 1u) CK(onc_launch_copy_fields(b->ka, gx_t, b->stream));
    for (onc_u32 p = 0; p < s0->chem_passes; p++) {
        g.pass = p;
        CK(hipModuleLaunchKernel(s0->chem_fn, gx_c, g.nsims, 1, 256, 1, 1, 0, b->stream, args, nullptr));
    }
    return ONC_OK;
}

static int launch_rule(onc_batch* b) {
    hipFunction_t fn = b->sims.empty() ? nullptr : b->sims[0]->rule_fn;
    if (!fn) {
        CK(onc_launch_rule_builtin(b->ra, grid_cells(b), b->stream));
        return ONC_OK;
    }
    void* args[] = {&b->ra};
    CK(hipModuleLaunchKernel(fn, (unsigned)grid_cells(b), 1, 1, 256, 1, 1, 0, b->stream, args, nullptr));
    return ONC_OK;
}

// Profiling: a begin and an end event per timed phase of every eager step,
// from a pool; durations are accumulated at the next sync, so the host never
// waits inside a run of steps. Only the phases in the mask get events (an
// event costs a few microseconds of GPU time, so timing one phase leaves the
// others nearly undisturbed).
#define ONC_EV_STEP (2 * ONC_PHASES)
static void prof_mark(onc_batch* b, int phase, int end) {
    if (!b->prof_on || !((b->profile >> phase) & 1u)) return;
    hipEventRecord(b->ev[(size_t)b->ev_used * ONC_EV_STEP + 2 * phase + end], b->stream);
    if (end) b->ev_mask.back() |= 1u << phase;
}

static int prof_reserve(onc_batch* b, int nsteps) {
    size_t need = (size_t)(b->ev_used + nsteps) * ONC_EV_STEP;
    while (b->ev.size() < need) {
        hipEvent_t e;
        CK(hipEventCreate(&e));
        b->ev.push_back(e);
    }
    return ONC_OK;
}

static int prof_collect(onc_batch* b) {
    for (int k = 0; k < b->ev_used; k++) {
        hipEvent_t* e = &b->ev[(size_t)k * ONC_EV_STEP];
        for (int ph = 0; ph < ONC_PHASES; ph++) {
            if (!((b->ev_mask[k] >> ph) & 1u)) continue;
            float ms = 0.0f;
            CK(hipEventElapsedTime(&ms, e[2 * ph], e[2 * ph + 1]));
            b->phase_us[ph] += 1000.0 * ms;
            b->phase_n[ph] += (ph == ONC_PH_SORT) ? (uint64_t)b->passes : 1u;
        }
    }
    b->ev_used = 0;
    b->ev_mask.clear();
    return ONC_OK;
}

// Enqueue one timestep for every sim of the batch. The plan commits the
// previous step; uptake is applied by the permute kernel; daughters stay in
// their compacted buffer until the next step's permute gathers them.
static int enqueue_step(onc_batch* b) {
    const onc_kargs& a = b->ka;
    hipStream_t st = b->stream;
    const int gt = grid_tiles(b), gc = grid_cells(b);
    prof_mark(b, ONC_PH_PLAN, 0);
    CK(onc_launch_plan(a, ONC_PLAN_STEP, st));
    prof_mark(b, ONC_PH_PLAN, 1);
    prof_mark(b, ONC_PH_MORTON, 0);
    CK(onc_launch_morton(a, grid_morton(b), b->big, st));
    prof_mark(b, ONC_PH_MORTON, 1);
    prof_mark(b, ONC_PH_SORT, 0);
    for (int p = 0; p < b->passes; p++) CK(onc_launch_sort_pass(a, b->items, p, gt, st));
    prof_mark(b, ONC_PH_SORT, 1);
    prof_mark(b, ONC_PH_PERMUTE_BINS, 0);
    CK(onc_launch_permute_bins(a, gc, st));
    prof_mark(b, ONC_PH_PERMUTE_BINS, 1);
    if (b->fused) {
        prof_mark(b, ONC_PH_FORCES_STENCIL, 0);
        CK(onc_launch_forces_stencil(a, b->fs_groups, st));
        prof_mark(b, ONC_PH_FORCES_STENCIL, 1);
    } else {
        prof_mark(b, ONC_PH_FORCES, 0);
        CK(onc_launch_forces(a, grid_forces(b), st));
        prof_mark(b, ONC_PH_FORCES, 1);
        prof_mark(b, ONC_PH_STENCIL, 0);
        CK(onc_launch_stencil(a, b->stencil_blocks, st));
        prof_mark(b, ONC_PH_STENCIL, 1);
    }
    CKR(launch_grid_rules(b));
    if (b->rule_compact) {
        prof_mark(b, ONC_PH_RULE_COMPACT, 0);
        CK(onc_launch_rule_compact(a, b->citems, grid_ctiles(b), st));
        prof_mark(b, ONC_PH_RULE_COMPACT, 1);
    } else {
        prof_mark(b, ONC_PH_RULE, 0);
        CKR(launch_rule(b));
        prof_mark(b, ONC_PH_RULE, 1);
        prof_mark(b, ONC_PH_COMPACT, 0);
        CK(onc_launch_compact(a, b->citems, grid_ctiles(b), st));
        prof_mark(b, ONC_PH_COMPACT, 1);
    }
    return ONC_OK;
}

static int check_rules(onc_batch* b) {
    for (size_t i = 1; i < b->sims.size(); i++)
        if (b->sims[i]->rule_fn != b->sims[0]->rule_fn)
            return fail(ONC_ERR_STATE, "all sims of a batch must use the same rule kernel");
    for (onc_sim* s : b->sims) {
        if (s->rule_fn) continue;
        const onc_sim_desc& d = s->desc;
        // the built-in tumour rule reads age (f32) / generation (u32) columns,
        // field 0 (O2) and five parameters
        if (d.n_cols < 6 || d.col_types[ONC_COL_AGE] != ONC_COL_F32 || d.col_types[ONC_COL_GEN] != ONC_COL_U32 ||
            d.n_fields < 1 || d.n_params < 5)
            return fail(ONC_ERR_STATE,
                        "built-in tumour rule needs columns age (f32) and generation (u32), one field and "
                        "five parameters; set a rule module for other schemas");
    }
    return ONC_OK;
}

// Tile sizes of the coming steps from the newest live count that has come
// back from the device (no wait here), and the next readback once
// ONC_GRAPH_STEPS steps are enqueued (waiting, if need be, for the copy that
// last used its slot). A change of sizes changes only launch arguments (the
// tables are sized for the smallest tiles) and drops the captured graphs;
// results do not depend on tile sizes.
static void batch_take_live(onc_batch* b, int k) {
    b->live_used[k] = false;
    if (b->live_seq[k] > b->live_seen_seq) {
        b->live_seen_seq = b->live_seq[k];
        b->live_cells = b->h_live[16 * k];
    }
}

static int batch_adapt_tiles(onc_batch* b) {
    if (b->big_from == 0) return ONC_OK;
    for (int k = 0; k < 2; k++) {
        if (!b->live_used[k]) continue;
        const hipError_t q = hipEventQuery(b->ev_live[k]);
        if (q == hipErrorNotReady) continue;
        CK(q);
        batch_take_live(b, k);
    }
    // a count taken here or while batch_count_steps waited for a slot
    if (b->live_seen_seq == b->live_decided_seq) return ONC_OK;
    b->live_decided_seq = b->live_seen_seq;
    const uint32_t live = b->live_cells;
    const bool big = b->big ? live >= b->small_below : live >= b->big_from;
    if (big == b->big) return ONC_OK;
    b->big = big;
    b->items = big ? std::max(b->tuned_items, ONC_BIG_ITEMS) : b->tuned_items;
    b->citems = big ? std::max(b->tuned_citems, ONC_BIG_CITEMS) : b->tuned_citems;
    b->ka.tile = (onc_u32)(256 * b->items);
    b->ka.ctile = (onc_u32)(256 * b->citems);
    b->ka.xcd_tiles = onc_sort_xcd_tiles(b->items, b->ctx->cus);
    batch_drop_graphs(b);
    return ONC_OK;
}

static int batch_count_steps(onc_batch* b, int steps) {
    b->steps_since_readback += steps;
    if (b->big_from == 0 || b->steps_since_readback < ONC_GRAPH_STEPS) return ONC_OK;
    if (!b->h_live) {
        void* p = nullptr;
        CK(hipHostMalloc(&p, 2 * 16 * 4, hipHostMallocDefault));
        b->h_live = (onc_u32*)p;
        for (hipEvent_t& e : b->ev_live) CK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    const int k = b->live_slot;
    if (b->live_used[k]) {
        CK(hipEventSynchronize(b->ev_live[k]));
        batch_take_live(b, k);
    }
    // the live cells of the last enqueued step (the plan's prefix total)
    CK(hipMemcpyAsync(b->h_live + 16 * k, b->ka.cell_prefix + b->ka.nsims, 4, hipMemcpyDeviceToHost, b->stream));
    CK(hipEventRecord(b->ev_live[k], b->stream));
    b->live_used[k] = true;
    b->live_seq[k] = b->live_next_seq++;
    b->live_slot = k ^ 1;
    b->steps_since_readback = 0;
    return ONC_OK;
}

static int batch_step_impl(onc_batch* b, int nsteps, int use_graph) {
    if (nsteps < 0) return fail(ONC_ERR_INVALID, "nsteps < 0");
    CK(hipSetDevice(b->ctx->device));
    if (b->dirty) CKR(batch_rebuild(b));
    if (b->sims.empty() || nsteps == 0) return ONC_OK;
    CKR(check_rules(b));
    const bool profiling = b->profile != 0u && !use_graph;
    if (profiling) CKR(prof_reserve(b, nsteps));
    if (!use_graph) {
        for (int k = 0; k < nsteps; k++) {
            if (k % ONC_GRAPH_STEPS == 0) CKR(batch_adapt_tiles(b));
            const bool timed = profiling && (b->prof_tick++ % (uint64_t)b->prof_every) == 0u;
            b->prof_on = timed;
            if (timed) b->ev_mask.push_back(0u);
            const int rc = enqueue_step(b);
            if (timed) b->ev_used++;
            b->prof_on = false;
            CKR(rc);
            CKR(batch_count_steps(b, 1));
        }
        return ONC_OK;
    }
    // Every step's work is decided on the device (the plan reads the live
    // counts), so k steps captured back to back replay exactly like k
    // replays of the one-step graph.
    auto capture = [&](int steps, hipGraph_t* graph, hipGraphExec_t* exec) -> int {
        CK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = ONC_OK;
        for (int k = 0; k < steps && rc == ONC_OK; k++) rc = enqueue_step(b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != ONC_OK) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(ONC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        *graph = g;
        CK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
        return ONC_OK;
    };
    for (int k = 0; k < nsteps;) {
        CKR(batch_adapt_tiles(b));
        if (nsteps - k >= ONC_GRAPH_STEPS) {
            if (!b->exec_n && b->single_replays >= ONC_GRAPH_AFTER)
                CKR(capture(ONC_GRAPH_STEPS, &b->graph_n, &b->exec_n));
            if (b->exec_n) {
                CK(hipGraphLaunch(b->exec_n, b->stream));
                k += ONC_GRAPH_STEPS;
                CKR(batch_count_steps(b, ONC_GRAPH_STEPS));
                continue;
            }
        }
        if (!b->exec) CKR(capture(1, &b->graph, &b->exec));
        CK(hipGraphLaunch(b->exec, b->stream));
        b->single_replays++;
        k++;
        CKR(batch_count_steps(b, 1));
    }
    return ONC_OK;
}

static int batch_sync_impl(onc_batch* b) {
    CK(hipSetDevice(b->ctx->device));
    CK(hipStreamSynchronize(b->stream));
    if (b->ev_used) CKR(prof_collect(b));
    if (b->dirty) return ONC_OK;
    onc_u32 err = 0;
    CK(hipMemcpy(&err, b->ka.err, 4, hipMemcpyDeviceToHost));
    if (err) {
        std::string m;
        if (err & (ONC_DERR_SORT | ONC_DERR_COMPACT | ONC_DERR_SCAN))
            m += "a kernel's bounded look-back wait expired (code " + std::to_string(err & 7u) + "); ";
        if (err & ONC_DERR_BIG_BIN)
            m += "a voxel held more than " + std::to_string(ONC_MAX_BIN_CELLS) + " cells (the force walk's limit); ";
        return fail(ONC_ERR_DEVICE, m);
    }
    return ONC_OK;
}

extern "C" int onc_sim_step(onc_sim* s, int nsteps) {
    if (!s) return fail(ONC_ERR_INVALID, "NULL sim");
    if (s->owner != s->solo) return fail(ONC_ERR_STATE, "sim belongs to a batch; step the batch");
    return batch_step_impl(s->solo, nsteps, 0);
}

extern "C" int onc_sim_sync(onc_sim* s) {
    if (!s) return fail(ONC_ERR_INVALID, "NULL sim");
    return batch_sync_impl(s->owner);
}

extern "C" int onc_batch_create(onc_ctx* c, onc_batch** out) {
    if (!c || !out) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(c->device));
    return batch_create_internal(c, false, out);
}

extern "C" int onc_batch_destroy(onc_batch* b) {
    if (!b) return ONC_OK;
    if (b->is_solo) return fail(ONC_ERR_INVALID, "cannot destroy a sim's private batch");
    hipSetDevice(b->ctx->device);
    // commit the last step, then hand every member back to its private batch
    int rc = batch_commit(b);
    for (onc_sim* s : b->sims) {
        s->owner = s->solo;
        s->solo->dirty = true;
    }
    if (rc != ONC_OK) {
        batch_destroy_internal(b);
        return rc;
    }
    batch_destroy_internal(b);
    return ONC_OK;
}

extern "C" int onc_batch_add(onc_batch* b, onc_sim* s) {
    if (!b || !s) return fail(ONC_ERR_INVALID, "NULL argument");
    if (b->is_solo) return fail(ONC_ERR_INVALID, "not a user batch");
    if (s->ctx != b->ctx) return fail(ONC_ERR_INVALID, "sim and batch belong to different contexts");
    if (s->owner != s->solo) return fail(ONC_ERR_STATE, "sim already belongs to a batch");
    if (!b->sims.empty() && b->sims[0]->rule_fn != s->rule_fn)
        return fail(ONC_ERR_STATE, "all sims of a batch must use the same rule kernel");
    CKR(batch_commit(s->solo));
    CK(hipStreamSynchronize(b->stream));
    b->sims.push_back(s);
    s->owner = b;
    b->dirty = true;
    return ONC_OK;
}

extern "C" int onc_batch_remove(onc_batch* b, onc_sim* s) {
    if (!b || !s) return fail(ONC_ERR_INVALID, "NULL argument");
    if (s->owner != b) return fail(ONC_ERR_STATE, "sim is not in this batch");
    CKR(batch_commit(b));
    b->sims.erase(std::remove(b->sims.begin(), b->sims.end(), s), b->sims.end());
    s->owner = s->solo;
    s->solo->dirty = true;
    b->dirty = true;
    return ONC_OK;
}

extern "C" int onc_batch_size(onc_batch* b, int* n) {
    if (!b || !n) return fail(ONC_ERR_INVALID, "NULL argument");
    *n = (int)b->sims.size();
    return ONC_OK;
}

extern "C" int onc_batch_step(onc_batch* b, int nsteps, int use_graph) {
    if (!b) return fail(ONC_ERR_INVALID, "NULL batch");
    if (b->is_solo) return fail(ONC_ERR_INVALID, "not a user batch");
    return batch_step_impl(b, nsteps, use_graph);
}

extern "C" int onc_batch_sync(onc_batch* b) {
    if (!b) return fail(ONC_ERR_INVALID, "NULL batch");
    return batch_sync_impl(b);
}

extern "C" int onc_batch_set_tuning(onc_batch* b, int items, int citems, int bx, int by, int bz) {
    if (!b) return fail(ONC_ERR_INVALID, "NULL batch");
    if (items != 4 && items != 8 && items != 12 && items != 16)
        return fail(ONC_ERR_INVALID, "sort items must be 4, 8, 12 or 16");
    if (citems != 4 && citems != 8 && citems != 12 && citems != 16)
        return fail(ONC_ERR_INVALID, "compaction items must be 4, 8, 12 or 16");
    if (bx < 1 || by < 1 || bx * by != 256 || bz < 1 || bz > 1024)
        return fail(ONC_ERR_INVALID, "stencil block must satisfy bx*by == 256, 1 <= bz <= 1024");
    CK(hipSetDevice(b->ctx->device));
    CK(hipStreamSynchronize(b->stream));
    b->tuned_items = b->items = items;
    b->tuned_citems = b->citems = citems;
    b->big_from = 0;  // tuned sizes are kept whatever the population (onc_batch_set_tile_policy)
    b->sbx = bx;
    b->sby = by;
    b->sbz = bz;
    b->dirty = true;
    return ONC_OK;
}

extern "C" int onc_batch_set_tile_policy(onc_batch* b, uint32_t big_from, uint32_t small_below) {
    if (!b) return fail(ONC_ERR_INVALID, "NULL batch");
    if (big_from != 0 && (small_below == 0 || small_below > big_from))
        return fail(ONC_ERR_INVALID, "tile policy needs 0 < small_below <= big_from (or big_from = 0)");
    CK(hipSetDevice(b->ctx->device));
    CK(hipStreamSynchronize(b->stream));
    b->big_from = big_from;
    b->small_below = small_below;
    b->dirty = true;
    return ONC_OK;
}

extern "C" int onc_batch_tile_items(onc_batch* b, int* items, int* citems) {
    if (!b || !items || !citems) return fail(ONC_ERR_INVALID, "NULL argument");
    *items = b->items;
    *citems = b->citems;
    return ONC_OK;
}

extern "C" int onc_batch_set_fused(onc_batch* b, int fused) {
    if (!b) return fail(ONC_ERR_INVALID, "NULL batch");
    if (fused != 0 && fused != 1) return fail(ONC_ERR_INVALID, "fused must be 0 or 1");
    CK(hipSetDevice(b->ctx->device));
    CK(hipStreamSynchronize(b->stream));
    b->fused = fused;
    b->dirty = true;
    return ONC_OK;
}

extern "C" int onc_batch_set_fused_rule(onc_batch* b, int fused) {
    if (!b) return fail(ONC_ERR_INVALID, "NULL batch");
    if (fused != 0 && fused != 1) return fail(ONC_ERR_INVALID, "fused must be 0 or 1");
    CK(hipSetDevice(b->ctx->device));
    CK(hipStreamSynchronize(b->stream));
    b->fuse_rule = fused;
    b->dirty = true;
    return ONC_OK;
}

extern "C" int onc_batch_set_forces_waves(onc_batch* b, int waves) {
    if (!b) return fail(ONC_ERR_INVALID, "NULL batch");
    if (waves < 0 || (waves > 0 && (waves < 8 || waves % 8 != 0)))
        return fail(ONC_ERR_INVALID, "force waves must be 0 or a positive multiple of 8");
    CK(hipSetDevice(b->ctx->device));
    CK(hipStreamSynchronize(b->stream));
    b->force_waves = waves;
    b->dirty = true;
    return ONC_OK;
}

extern "C" int onc_batch_profile(onc_batch* b, int phase_mask) {
    return onc_batch_profile_every(b, phase_mask, 1);
}

extern "C" int onc_batch_profile_every(onc_batch* b, int phase_mask, int every) {
    const int enable = phase_mask;
    if (!b) return fail(ONC_ERR_INVALID, "NULL batch");
    if (every < 1) return fail(ONC_ERR_INVALID, "every must be >= 1");
    CK(hipSetDevice(b->ctx->device));
    CK(hipStreamSynchronize(b->stream));
    b->ev_used = 0;
    b->ev_mask.clear();
    b->prof_every = every;
    b->prof_tick = 0;
    b->profile = (unsigned)enable & ((1u << ONC_PHASES) - 1u);
    for (int i = 0; i < ONC_PHASES; i++) {
        b->phase_us[i] = 0.0;
        b->phase_n[i] = 0;
    }
    return ONC_OK;
}

extern "C" int onc_batch_sweep_voxels(onc_batch* b, uint64_t* swept, uint64_t* filled, uint64_t* kept) {
    if (!b || !swept || !filled || !kept) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(b->ctx->device));
    CKR(batch_commit(b));
    *swept = *filled = *kept = 0;
    if (b->dirty) return ONC_OK;
    const onc_kargs& a = b->ka;
    std::vector<onc_u32> q;
    if (a.quiet) {
        q.resize(4ull * a.nqb);
        CK(hipMemcpy(q.data(), b->d_quiet, q.size() * 4, hipMemcpyDeviceToHost));
    }
    const int nt = b->fused ? 64 : 256;
    for (size_t j = 0; j < b->jobs_h.size(); j++) {
        onc_sim* s = b->sims[b->jobs_h[j] >> 8];
        onc_dstate st;
        CK(hipMemcpy(&st, s->dstate, sizeof(st), hipMemcpyDeviceToHost));
        if (st.step >= st.stop) continue;  // frozen: not swept
        const int X = s->desc.dims[0], Y = s->desc.dims[1], Z = s->desc.dims[2];
        if (!a.quiet || (X & 3) != 0) {
            *swept += (uint64_t)X * Y * Z;
            continue;
        }
        // the last sweep wrote the buffer that is current now
        const onc_u32* qo = q.data() + 2ull * st.parity * a.nqb;
        const int X4 = X >> 2, tw = onc_stencil_tw(X4, std::min(b->sbx, nt)), th = nt / tw;
        const int ntx = (X4 + tw - 1) / tw, nty = (Y + th - 1) / th;
        for (onc_u32 blk = b->jprefix_h[j]; blk < b->jprefix_h[j + 1]; blk++) {
            const onc_u32 l = blk - b->jprefix_h[j];
            const int tx = (int)(l % ntx), ty = (int)((l / ntx) % nty), tz = (int)(l / ((onc_u32)ntx * nty));
            const uint64_t vox = (uint64_t)(4 * std::min(tw, X4 - tx * tw)) * std::min(th, Y - ty * th) *
                                 std::min(b->sbz, Z - tz * b->sbz);
            const onc_u32 state = qo[2ull * blk];
            if (state == 2u) *filled += vox;
            else if (state == 3u) *kept += vox;
            else *swept += vox;
        }
    }
    return ONC_OK;
}

extern "C" int onc_batch_phase_times(onc_batch* b, double* us, uint64_t* launches, int n) {
    if (!b) return fail(ONC_ERR_INVALID, "NULL batch");
    for (int i = 0; i < n && i < ONC_PHASES; i++) {
        if (us) us[i] = b->phase_us[i];
        if (launches) launches[i] = b->phase_n[i];
    }
    return ONC_OK;
}

// ---------------------------------------------------------------------------
// pure kernels
// ---------------------------------------------------------------------------
static int ctx_scratch(onc_ctx* c, size_t bytes, void** out) {
    if (bytes > c->scratch_bytes) {
        CK(hipStreamSynchronize(c->stream));
        for (void* p : c->scratch_bufs) hipFree(p);
        c->scratch_bufs.clear();
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return fail(ONC_ERR_NOMEM, "scratch allocation");
        c->scratch_bufs.push_back(p);
        c->scratch_bytes = bytes;
    }
    *out = c->scratch_bufs[0];
    return ONC_OK;
}

extern "C" int onc_morton_keys(onc_ctx* c, const float* px, const float* py, const float* pz, uint32_t n,
                               const int32_t dims[3], float dx, int fmt, uint32_t* keys) {
    if (!c || !dims || (n && (!px || !py || !pz || !keys))) return fail(ONC_ERR_INVALID, "NULL argument");
    if (fmt != ONC_KEY_MORTON30 && fmt != ONC_KEY_MORTON32) return fail(ONC_ERR_INVALID, "bad key format");
    CK(hipSetDevice(c->device));
    tmark(c, 0);
    CK(onc_launch_morton_keys(px, py, pz, n, dims[0], dims[1], dims[2], dx, fmt, keys, c->stream));
    tmark(c, 1);
    CK(hipStreamSynchronize(c->stream));
    return tread(c);
}

extern "C" int onc_forces(onc_ctx* c, const onc_sim_desc* d, const float* px, const float* py, const float* pz,
                          uint32_t n, float* fx, float* fy, float* fz) {
    if (!c || !d || (n && (!px || !py || !pz || !fx || !fy || !fz))) return fail(ONC_ERR_INVALID, "NULL argument");
    CKR(validate_desc(d));
    if (n > (1u << 30)) return fail(ONC_ERR_INVALID, "n > 2^30");
    CK(hipSetDevice(c->device));
    if (n == 0) return ONC_OK;
    uint32_t table = 0;
    int key_bits = 0;
    size_t nvox = 0;
    uint64_t bytes = 0;
    sim_layout(d, &table, &key_bits, &nvox, &bytes);
    size_t off = 0;
    auto take = [&](size_t b) {
        size_t o = off;
        off += (b + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_sim = take(sizeof(onc_dsim)), o_small = take(256), o_pos = take((size_t)n * 16),
                 o_keys = take((size_t)n * 4), o_bins = take((size_t)table * 8);
    void* base;
    CKR(ctx_scratch(c, off, &base));
    char* B = (char*)base;
    onc_dsim ds{};
    for (int k = 0; k < 3; k++) ds.dims[k] = d->dims[k];
    ds.dx = d->dx;
    ds.key_format = d->key_format;
    ds.R = d->interaction_radius;
    ds.cell_radius = d->cell_radius;
    ds.k_rep = d->k_rep;
    ds.k_adh = d->k_adh;
    ds.mobility = d->mobility;
    ds.max_disp = d->max_disp;
    ds.step = 0;  // bin entries carry onc_bin_tag(0)
    ds.table = table;
    ds.n_live = n;
    ds.pos4 = (float*)(B + o_pos);
    ds.bins = (onc_u32*)(B + o_bins);
    onc_u32 small[4] = {0u, n, 0u, 0u};  // cell prefix {0, n}, error word
    hipStream_t st = c->stream;
    CK(hipMemcpyAsync(B + o_sim, &ds, sizeof(ds), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(B + o_small, small, sizeof(small), hipMemcpyHostToDevice, st));
    CK(hipMemsetAsync(B + o_bins, 0, (size_t)table * 8, st));  // tag 0: no entry valid
    onc_kargs a{};
    a.sims = (const onc_dsim*)(B + o_sim);
    a.nsims = 1;
    a.cell_prefix = (onc_u32*)(B + o_small);
    onc_u32* err = (onc_u32*)(B + o_small) + 2;
    CK(onc_launch_forces_raw(a, px, py, pz, n, (const onc_dsim*)(B + o_sim), (float*)(B + o_pos),
                             (onc_u32*)(B + o_keys), (onc_u32*)(B + o_bins), err, fx, fy, fz,
                             c->timing ? c->t0 : nullptr, c->timing ? c->t1 : nullptr, st));
    onc_u32 herr = 0;
    CK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (herr) return fail(ONC_ERR_DEVICE, "a voxel held more than " + std::to_string(ONC_MAX_BIN_CELLS) + " cells");
    return tread(c);
}

// Large standalone sorts: each pass as counts -> scan of the counts -> rank
// and scatter (three launches, no look-back chain between tiles).
static int sort_pairs_3launch(onc_ctx* c, const uint32_t* kin, const uint32_t* vin, uint32_t n, int passes,
                              uint32_t* kout, uint32_t* vout) {
    const size_t T = (n + 256u * ONC_SORT3_ITEMS - 1) / (256u * ONC_SORT3_ITEMS);
    const size_t nc = 256u * T;  // (digit, tile) counts
    const size_t stiles = (nc + ONC_SCAN_MIN_TILE - 1) / ONC_SCAN_MIN_TILE;
    if (nc > 0xFFFFFFFFull) return fail(ONC_ERR_INVALID, "too many sort tiles");
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_k0 = take(n * 4ull), o_k1 = take(n * 4ull), o_v0 = take(n * 4ull), o_v1 = take(n * 4ull);
    const size_t o_cnt = take(nc * 4), o_off = take(nc * 4);
    const size_t o_ss = take((size_t)passes * stiles * 8), o_err = take(256);
    void* base;
    CKR(ctx_scratch(c, off, &base));
    char* B = (char*)base;
    onc_u32* keys[2] = {(onc_u32*)(B + o_k0), (onc_u32*)(B + o_k1)};
    onc_u32* vals[2] = {(onc_u32*)(B + o_v0), (onc_u32*)(B + o_v1)};
    onc_u32* counts = (onc_u32*)(B + o_cnt);
    onc_u32* offsets = (onc_u32*)(B + o_off);
    onc_u32* err = (onc_u32*)(B + o_err);
    hipStream_t st = c->stream;
    CK(hipMemsetAsync(B + o_ss, 0, (size_t)passes * stiles * 8 + 256, st));
    CK(hipMemcpyAsync(keys[0], kin, n * 4ull, hipMemcpyDeviceToDevice, st));
    CK(hipMemcpyAsync(vals[0], vin, n * 4ull, hipMemcpyDeviceToDevice, st));
    tmark(c, 0);
    for (int p = 0; p < passes; p++) {
        const int i = p & 1;
        CK(onc_launch_sort_up(keys[i], n, 8u * (onc_u32)p, counts, (onc_u32)T, st));
        CK(onc_launch_exscan(counts, offsets, (onc_u32)nc, nullptr, (onc_u32*)(B + o_ss + (size_t)p * stiles * 8),
                             err, st));
        CK(onc_launch_sort_down(keys[i], vals[i], keys[i ^ 1], vals[i ^ 1], n, 8u * (onc_u32)p, offsets,
                                (onc_u32)T, st));
    }
    tmark(c, 1);
    CK(hipMemcpyAsync(kout, keys[passes & 1], n * 4ull, hipMemcpyDeviceToDevice, st));
    CK(hipMemcpyAsync(vout, vals[passes & 1], n * 4ull, hipMemcpyDeviceToDevice, st));
    onc_u32 herr = 0;
    CK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (herr) return fail(ONC_ERR_DEVICE, "scan look-back wait expired");
    return tread(c);
}

extern "C" int onc_sort_pairs(onc_ctx* c, const uint32_t* kin, const uint32_t* vin, uint32_t n, int key_bits,
                              uint32_t* kout, uint32_t* vout) {
    if (!c || (n && (!kin || !vin || !kout || !vout))) return fail(ONC_ERR_INVALID, "NULL argument");
    if (key_bits < 1 || key_bits > 32) return fail(ONC_ERR_INVALID, "key_bits must be in [1, 32]");
    if (n > (1u << 30)) return fail(ONC_ERR_INVALID, "n > 2^30");
    CK(hipSetDevice(c->device));
    if (n == 0) return ONC_OK;
    if (n >= ONC_SORT3_MIN_N) return sort_pairs_3launch(c, kin, vin, n, (key_bits + 7) / 8, kout, vout);
    const int items = 16;
    const size_t tile = 256 * items;
    const size_t tiles = (n + tile - 1) / tile;
    const int passes = (key_bits + 7) / 8;
    // layout: keys[2], vals[2], dsim, dstate, small, hist, sort status, comp status
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    size_t o_k0 = take(n * 4ull), o_k1 = take(n * 4ull), o_v0 = take(n * 4ull), o_v1 = take(n * 4ull);
    size_t o_sim = take(sizeof(onc_dsim)), o_st = take(sizeof(onc_dstate));
    size_t o_small = take((3 * 2 + ONC_NCOUNTERS + 16) * 4);
    size_t o_hist = take((size_t)ONC_HIST_COPIES * 4 * 256 * 4);
    size_t o_ss = take((size_t)ONC_DEV_MAX_PASSES * tiles * 256 * 4);
    size_t o_cs = take(tiles * 8);
    void* base;
    CKR(ctx_scratch(c, off, &base));
    char* B = (char*)base;
    onc_dsim ds{};
    ds.keys[0] = (onc_u32*)(B + o_k0);
    ds.keys[1] = (onc_u32*)(B + o_k1);
    ds.vals[0] = (onc_u32*)(B + o_v0);
    ds.vals[1] = (onc_u32*)(B + o_v1);
    ds.state = (onc_dstate*)(B + o_st);
    ds.cap = n;
    onc_dstate dst{};
    dst.stop = 0xFFFFFFFFu;
    dst.n = n;
    hipStream_t st = c->stream;
    CK(hipMemcpyAsync(B + o_sim, &ds, sizeof(ds), hipMemcpyHostToDevice, st));
    CK(hipMemcpyAsync(B + o_st, &dst, sizeof(dst), hipMemcpyHostToDevice, st));
    CK(hipMemsetAsync(B + o_small, 0, (3 * 2 + ONC_NCOUNTERS + 16) * 4, st));
    CK(hipMemsetAsync(B + o_hist, 0, (size_t)ONC_HIST_COPIES * 4 * 256 * 4, st));
    CK(hipMemcpyAsync(ds.keys[0], kin, n * 4ull, hipMemcpyDeviceToDevice, st));
    CK(hipMemcpyAsync(ds.vals[0], vin, n * 4ull, hipMemcpyDeviceToDevice, st));
    onc_kargs a{};
    a.sims = (const onc_dsim*)(B + o_sim);
    a.nsims = 1;
    a.passes = (onc_u32)passes;
    a.tile = (onc_u32)tile;
    a.max_tiles = (onc_u32)tiles;
    a.xcd_tiles = onc_sort_xcd_tiles(items, c->cus);
    onc_u32* small = (onc_u32*)(B + o_small);
    a.ctile = (onc_u32)tile;
    a.max_ctiles = (onc_u32)tiles;
    a.tile_prefix = small;
    a.cell_prefix = small + 2;
    a.ctile_prefix = small + 4;
    a.counters = small + 6;
    a.err = a.counters + ONC_NCOUNTERS;
    a.hist = (onc_u32*)(B + o_hist);
    a.sort_status = (onc_u32*)(B + o_ss);
    a.comp_status = (onc_u64*)(B + o_cs);
    // one block per tile: a tile's look-back waits only on lower tiles, whose
    // blocks were dispatched before it (a grid of fewer blocks looping over
    // tiles can hold every resident block on a tile of a block that never
    // becomes resident)
    const int grid = (int)tiles;
    CK(onc_launch_plan(a, ONC_PLAN_STEP, st));
    tmark(c, 0);
    CK(onc_launch_hist_only(a, items, (int)std::min<size_t>(tiles, (size_t)c->cus * 4), st));
    for (int p = 0; p < passes; p++) CK(onc_launch_sort_pass(a, items, p, grid, st));
    tmark(c, 1);
    CK(hipMemcpyAsync(kout, ds.keys[passes & 1], n * 4ull, hipMemcpyDeviceToDevice, st));
    CK(hipMemcpyAsync(vout, ds.vals[passes & 1], n * 4ull, hipMemcpyDeviceToDevice, st));
    onc_u32 err = 0;
    CK(hipMemcpyAsync(&err, a.err, 4, hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (err) return fail(ONC_ERR_DEVICE, "sort look-back wait expired");
    return tread(c);
}

extern "C" int onc_exscan_u32(onc_ctx* c, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* total) {
    if (!c || (n && (!in || !out))) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(c->device));
    if (n == 0) {
        if (total) CK(hipMemsetAsync(total, 0, 4, c->stream));
        CK(hipStreamSynchronize(c->stream));
        return ONC_OK;
    }
    const size_t tiles = (n + ONC_SCAN_MIN_TILE - 1) / ONC_SCAN_MIN_TILE;
    void* base;
    CKR(ctx_scratch(c, tiles * 8 + 256, &base));
    onc_u64* status = (onc_u64*)base;
    onc_u32* err = (onc_u32*)((char*)base + tiles * 8);
    CK(hipMemsetAsync(base, 0, tiles * 8 + 256, c->stream));
    tmark(c, 0);
    CK(onc_launch_exscan(in, out, n, total, (onc_u32*)status, err, c->stream));
    tmark(c, 1);
    onc_u32 herr = 0;
    CK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    if (herr) return fail(ONC_ERR_DEVICE, "scan look-back wait expired");
    return tread(c);
}

// The same scan with its look-back status and error words in the device's
// `scratch` array instead of the context's allocated scratch: no hipMalloc,
// hipFree or allocating library call. Every call on a device shares that
// device's array, so calls that overlap on one device (several threads or
// streams) race; the words are cleared by a fill kernel at each call.
extern "C" int onc_exscan_u32_noalloc(onc_ctx* c, const uint32_t* in, uint32_t* out, uint32_t n,
                                      uint32_t* total) {
    if (!c || (n && (!in || !out))) return fail(ONC_ERR_INVALID, "NULL argument");
    if (n > ONC_NOALLOC_SCAN_MAX_N)
        return fail(ONC_ERR_INVALID, "n exceeds the no-alloc scan's limit of 2^30 elements");
    if (!c->noalloc_scratch) return fail(ONC_ERR_STATE, "context has no scratch address");
    CK(hipSetDevice(c->device));
    if (n == 0) {
        if (total) CK(onc_launch_fill_u32(total, 0u, 1, c->stream));
        CK(hipStreamSynchronize(c->stream));
        return ONC_OK;
    }
    const size_t tiles = (n + ONC_SCAN_MIN_TILE - 1) / ONC_SCAN_MIN_TILE;
    onc_u32* err = (onc_u32*)c->noalloc_scratch;
    onc_u32* status = err + ONC_NOALLOC_ERR_WORDS;  // 8-byte words, 256-byte aligned
    CK(onc_launch_fill_u32(err, 0u, ONC_NOALLOC_ERR_WORDS + 2 * tiles, c->stream));
    CK(onc_launch_exscan(in, out, n, total, status, err, c->stream));
    onc_u32 herr = 0;
    CK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    if (herr) return fail(ONC_ERR_DEVICE, "scan look-back wait expired");
    return ONC_OK;
}

extern "C" int onc_bins(onc_ctx* c, const uint32_t* keys, uint32_t n, uint32_t* start, uint32_t* end) {
    if (!c || (n && (!keys || !start || !end))) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(c->device));
    CK(onc_launch_bins(keys, n, start, end, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return ONC_OK;
}

extern "C" int onc_stencil7(onc_ctx* c, const float* in, float* out, const int32_t dims[3],
                            const onc_field_desc* f) {
    if (!c || !in || !out || !dims || !f) return fail(ONC_ERR_INVALID, "NULL argument");
    for (int i = 0; i < 3; i++)
        if (dims[i] < 1 || dims[i] > 2048) return fail(ONC_ERR_INVALID, "dims out of range");
    CK(hipSetDevice(c->device));
    const int sbx = 64, sby = 4, sbz = 16;
    onc_dsim ds{};
    for (int i = 0; i < 3; i++) ds.dims[i] = dims[i];
    memcpy(&ds.fields[0], f, sizeof(onc_dfield));
    ds.n_fields = 1;
    ds.fld[0][0] = (float*)in;
    ds.fld[1][0] = out;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    size_t o_sim = take(sizeof(onc_dsim)), o_st = take(sizeof(onc_dstate)), o_j = take(16);
    void* base;
    CKR(ctx_scratch(c, off, &base));
    char* B = (char*)base;
    ds.state = (onc_dstate*)(B + o_st);
    onc_dstate dst{};
    dst.stop = 0xFFFFFFFFu;
    const onc_u32 nb = stencil_blocks(dims[0], dims[1], dims[2], sbx, sbz, 256);
    onc_u32 jobs[4] = {0u, 0u, nb, 0u};  // jobs[0] = job 0, jobs_prefix = {0, nb}
    CK(hipMemcpyAsync(B + o_sim, &ds, sizeof(ds), hipMemcpyHostToDevice, c->stream));
    CK(hipMemcpyAsync(B + o_st, &dst, sizeof(dst), hipMemcpyHostToDevice, c->stream));
    CK(hipMemcpyAsync(B + o_j, jobs, sizeof(jobs), hipMemcpyHostToDevice, c->stream));
    onc_kargs a{};
    a.sims = (const onc_dsim*)(B + o_sim);
    a.nsims = 1;
    a.jobs = (onc_u32*)(B + o_j);
    a.jobs_prefix = (onc_u32*)(B + o_j) + 1;
    a.njobs = 1;
    a.sbx = sbx;
    a.sby = sby;
    a.sbz = sbz;
    tmark(c, 0);
    CK(onc_launch_stencil(a, nb, c->stream));
    tmark(c, 1);
    CK(hipStreamSynchronize(c->stream));
    return tread(c);
}

extern "C" int onc_philox(onc_ctx* c, const uint32_t* ctr, uint32_t n, uint64_t key, uint32_t* out) {
    if (!c || (n && (!ctr || !out))) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(c->device));
    CK(onc_launch_philox(ctr, n, (onc_u32)key, (onc_u32)(key >> 32), out, c->stream));
    CK(hipStreamSynchronize(c->stream));
    return ONC_OK;
}

extern "C" int onc_sim_cell_updates(onc_sim* s, uint64_t* cu) {
    if (!s || !cu) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(s->ctx->device));
    onc_dstate st;
    CKR(sim_read_state(s, &st));
    *cu = st.cell_updates;
    return ONC_OK;
}

extern "C" int onc_ctx_timing(onc_ctx* c, int on) {
    if (!c) return fail(ONC_ERR_INVALID, "NULL context");
    CK(hipSetDevice(c->device));
    if (on && !c->t0) {
        CK(hipEventCreate(&c->t0));
        CK(hipEventCreate(&c->t1));
    }
    c->timing = on != 0;
    c->last_us = 0.0;
    return ONC_OK;
}

extern "C" int onc_ctx_last_kernel_us(onc_ctx* c, double* us) {
    if (!c || !us) return fail(ONC_ERR_INVALID, "NULL argument");
    if (!c->timing) return fail(ONC_ERR_STATE, "timing is off (onc_ctx_timing)");
    *us = c->last_us;
    return ONC_OK;
}

extern "C" int onc_copy_u32(onc_ctx* c, const uint32_t* src, uint32_t* dst, uint64_t n) {
    if (!c || (n && (!src || !dst))) return fail(ONC_ERR_INVALID, "NULL argument");
    CK(hipSetDevice(c->device));
    tmark(c, 0);
    CK(onc_launch_copy_u32(src, dst, (size_t)n, c->stream));
    tmark(c, 1);
    CK(hipStreamSynchronize(c->stream));
    return tread(c);
}
