/*
 * onco_oracle_real.c — CPU oracle, floating-point half: voxel keys of
 * positions, neighbour forces, the 7-point stencil, uptake, the built-in
 * tumour rule, stable compaction and the whole timestep.
 *
 * TEST INFRASTRUCTURE ONLY (see onco_oracle_int.c). Compiled twice: with
 * REAL=float (suffix _f32, the bit-exact twin of the HIP path when both are
 * built with -ffp-contract=off) and REAL=double (suffix _f64, the drift
 * oracle). Every floating-point expression below is written in the exact
 * operation order the HIP kernels use.
 *
 * Timestep (deterministic spec; the reference's CPU and GPU backends disagree
 * on order, cpu/_cpu_backend.py:36-70 vs gpu/runtime/_gpu_backend.py:113-171):
 *   1. key_i = Morton(clamp(floor(p_i / dx), 0, dims-1))
 *   2. stable radix sort of (key, index); permute every column
 *   3. bins = [start, end) of each occupied key
 *   4. uptake: every occupied, non-fixed voxel loses n_cells * uptake
 *   5. forces over the 27 Moore bins (template.wgsl:11-21 order, x fastest;
 *      out-of-range bins skipped as at template.wgsl:134-138), bin order
 *      within a bin, self skipped; p' = p + clamp(mobility * F)
 *   6. one explicit 7-point sweep per field
 *   7. rule on (permuted state, p', swept fields) -> alive / divide flags
 *   8. stable compaction: survivors in order, daughters appended in parent
 *      order (core/utils/_containers.py:107-117 order; replaces the atomic
 *      append of gpu/compiler/translator/_stmt_translator.py:442-453)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/onco.h"

#ifndef REAL
#define REAL float
#define SFX f32
#endif
#ifdef ORACLE_F32
#define RSQRT sqrtf
#define RFLOOR floorf
#else
#define RSQRT sqrt
#define RFLOOR floor
#endif
/* Planted mistakes, for test builds only: tests/test_oracle_mutants.py builds
 * this file with ORACLE_PLANT = 1..7 and checks that the closed-form step
 * tests fail on each. The default build (0) plants nothing.
 *   1 uptake run length off by one     2 uptake applied after the sweep
 *   3 no max_disp clamp                4 no domain clamp
 *   5 drug kill with <= instead of <   6 drug kill reads the second word
 *   7 drug kill draws with step + 1 */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif
#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

uint32_t oracle_morton(uint32_t x, uint32_t y, uint32_t z, int fmt);
int oracle_key_bits(const int32_t dims[3]);
void oracle_radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t n, int key_bits);
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

static REAL clampr(REAL x, REAL lo, REAL hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* Voxel coordinate along one axis: floor of the IEEE quotient, clamped.
 * The reference truncates (template.wgsl:180; core/utils/_types.py:557-559)
 * and clamps on sampling (core/utils/_types.py:325-327); positions are kept
 * >= 0, so floor and trunc agree. */
static int32_t vox1(REAL p, REAL dx, int32_t d) {
    REAL q = p / dx;
    REAL f = RFLOOR(q);
    int32_t c = f < (REAL)0 ? 0 : (f >= (REAL)d ? d - 1 : (int32_t)f);
    return c;
}

uint32_t FN(oracle_voxel_key)(REAL px, REAL py, REAL pz, REAL dx, const int32_t dims[3], int fmt) {
    return oracle_morton((uint32_t)vox1(px, dx, dims[0]), (uint32_t)vox1(py, dx, dims[1]),
                         (uint32_t)vox1(pz, dx, dims[2]), fmt);
}

void FN(oracle_keys)(const REAL* px, const REAL* py, const REAL* pz, uint32_t n, const int32_t dims[3],
                     REAL dx, int fmt, uint32_t* keys) {
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < n; i++) keys[i] = FN(oracle_voxel_key)(px[i], py[i], pz[i], dx, dims, fmt);
}

/* ---- Philox-driven draws -------------------------------------------------
 * rand(): call c of a cell in a step reads word (c & 3) of
 * Philox4x32-10(ctr = {id, step, c >> 2, 0}, key = seed) and maps it to
 * (w >> 8) * 2^-24, exact in binary32. */
typedef struct {
    uint32_t id, step, c;
    uint32_t key[2];
    uint32_t buf[4];
} rng_t;

static REAL rng_uniform(rng_t* r) {
    if ((r->c & 3u) == 0) {
        uint32_t ctr[4] = {r->id, r->step, r->c >> 2, 0u};
        oracle_philox4x32_10(ctr, r->key, r->buf);
    }
    uint32_t w = r->buf[r->c & 3u];
    r->c++;
    return (REAL)(w >> 8) * (REAL)5.9604644775390625e-08; /* 2^-24 */
}

/* sin and cos of a in [0, pi/2] from their Taylor series (Horner in a^2,
 * through a^13 and a^14), so the HIP kernels and this oracle evaluate the
 * same operations; no libm trig is involved. Pinned by: tests against numpy
 * sin/cos over [0, pi/2] (max error a few ulp in binary32). */
void FN(oracle_sincos_q)(REAL a, REAL* s_out, REAL* c_out) {
    REAL a2 = a * a;
    REAL s = (REAL)1.6059043836821613e-10;      /* 1/13! */
    s = s * a2 + (REAL)-2.5052108385441720e-08; /* -1/11! */
    s = s * a2 + (REAL)2.7557319223985893e-06;  /* 1/9! */
    s = s * a2 + (REAL)-1.9841269841269841e-04; /* -1/7! */
    s = s * a2 + (REAL)8.3333333333333333e-03;  /* 1/5! */
    s = s * a2 + (REAL)-1.6666666666666667e-01; /* -1/3! */
    s = s * a2 + (REAL)1;
    s = s * a;
    REAL c = (REAL)-1.1470745597729725e-11; /* -1/14! */
    c = c * a2 + (REAL)2.0876756987868099e-09;  /* 1/12! */
    c = c * a2 + (REAL)-2.7557319223985891e-07; /* -1/10! */
    c = c * a2 + (REAL)2.4801587301587302e-05;  /* 1/8! */
    c = c * a2 + (REAL)-1.3888888888888889e-03; /* -1/6! */
    c = c * a2 + (REAL)4.1666666666666667e-02;  /* 1/4! */
    c = c * a2 + (REAL)-0.5;
    c = c * a2 + (REAL)1;
    *s_out = s;
    *c_out = c;
}

/* Unit vector from two uniforms: z = 2 u1 - 1, phi = 2 pi u2, the z/phi form
 * of gpu/compiler/template.wgsl:84-98 (the CPU backend's acos form,
 * core/utils/_types.py:707-712, draws the same distribution). The angle is
 * reduced exactly by quadrant: t = 4 u2, q = floor(t), a = (t - q) pi/2. */
static void rng_dir(rng_t* r, REAL* x, REAL* y, REAL* z) {
    REAL u1 = rng_uniform(r);
    REAL u2 = rng_uniform(r);
    REAL zz = u1 * (REAL)2 - (REAL)1;
    REAL rr = (REAL)1 - zz * zz;
    if (rr < (REAL)0) rr = (REAL)0;
    REAL rad = RSQRT(rr);
    REAL t = u2 * (REAL)4;
    int32_t q = (int32_t)t;
    REAL f = t - (REAL)q;
    REAL a = f * (REAL)1.5707963267948966;
    REAL s, c;
    FN(oracle_sincos_q)(a, &s, &c);
    REAL cx, sy;
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

void FN(oracle_random_dir)(uint32_t id, uint32_t step, uint64_t seed, uint32_t c0, REAL out[3]) {
    rng_t r;
    r.id = id; r.step = step; r.c = 0;
    r.key[0] = (uint32_t)seed; r.key[1] = (uint32_t)(seed >> 32);
    /* advance to call c0 */
    while (r.c < c0) (void)rng_uniform(&r);
    rng_dir(&r, &out[0], &out[1], &out[2]);
}

/* ---- fields ---------------------------------------------------------------
 * Fixed (Dirichlet) voxels: every boundary-layer voxel for ONC_BC_DIRICHLET
 * (tumor_sim.wgsl:420-423); for ONC_BC_POINT_SOURCE the voxels of face z=0
 * whose centre lies within src_radius of the face centre. */
static int is_fixed(const onc_field_desc* f, const int32_t d[3], int32_t x, int32_t y, int32_t z) {
    if (f->bc == ONC_BC_DIRICHLET)
        return x == 0 || y == 0 || z == 0 || x == d[0] - 1 || y == d[1] - 1 || z == d[2] - 1;
    if (f->bc != ONC_BC_POINT_SOURCE || z != 0) return 0;
    REAL fx = ((REAL)x + (REAL)0.5) - (REAL)d[0] * (REAL)0.5;
    REAL fy = ((REAL)y + (REAL)0.5) - (REAL)d[1] * (REAL)0.5;
    REAL r = (REAL)f->src_radius;
    return fx * fx + fy * fy <= r * r;
}

void FN(oracle_field_init)(const onc_field_desc* f, const int32_t d[3], REAL* u) {
    size_t i = 0;
    for (int32_t z = 0; z < d[2]; z++)
        for (int32_t y = 0; y < d[1]; y++)
            for (int32_t x = 0; x < d[0]; x++, i++)
                u[i] = is_fixed(f, d, x, y, z) ? (REAL)f->bc_value : (REAL)f->init;
}

/* One explicit sweep (engine/assets/shaders/tumor_sim.wgsl:432-444, twin
 * engine/hlsl/Fields.hlsl:63-101): u' = clamp(u + r (sum6 - 6u) - decay u).
 * Neighbours outside the domain mirror the centre (zero flux); they only
 * occur on point-source fields since Dirichlet boundary voxels are fixed.
 * Layout: linear x + X (y + Y z). Pinned by: the constant-field fixed point,
 * the discrete sine-mode decay factor, and mass conservation under zero flux. */
void FN(oracle_stencil7)(const REAL* in, REAL* out, const int32_t d[3], const onc_field_desc* f) {
    const size_t X = (size_t)d[0], XY = (size_t)d[0] * (size_t)d[1];
    REAL rate = (REAL)f->rate, decay = (REAL)f->decay, lo = (REAL)f->lo, hi = (REAL)f->hi;
#pragma omp parallel for schedule(static)
    for (int32_t z = 0; z < d[2]; z++)
        for (int32_t y = 0; y < d[1]; y++)
            for (int32_t x = 0; x < d[0]; x++) {
                size_t v = (size_t)x + X * (size_t)y + XY * (size_t)z;
                if (is_fixed(f, d, x, y, z)) {
                    out[v] = (REAL)f->bc_value;
                    continue;
                }
                REAL c = in[v];
                REAL xm = x > 0 ? in[v - 1] : c;
                REAL xp = x < d[0] - 1 ? in[v + 1] : c;
                REAL ym = y > 0 ? in[v - X] : c;
                REAL yp = y < d[1] - 1 ? in[v + X] : c;
                REAL zm = z > 0 ? in[v - XY] : c;
                REAL zp = z < d[2] - 1 ? in[v + XY] : c;
                REAL s = xm + xp;
                s = s + ym;
                s = s + yp;
                s = s + zm;
                s = s + zp;
                REAL lap = s - (REAL)6 * c;
                REAL w = c + rate * lap;
                w = w - decay * c;
                out[v] = clampr(w, lo, hi);
            }
}

/* ---- neighbour forces -----------------------------------------------------
 * Pair law for delta = p_i - p_j, d = |delta| (tumor_sim.wgsl:541-570, with
 * the benchmark model's linear spring in place of the Hertz k o sqrt(o), and a linear
 * taper in place of the cosine taper):
 *   d < S = 2 r_cell:      f = k_rep (S - d)            (repulsion)
 *   S <= d < R:            f = -k_adh (R - d) / (R - S) (adhesion)
 *   F_i += (f / d) * delta; pairs with d >= R or d == 0 contribute nothing.
 * Visitation: 27 Moore bins in template.wgsl:11-21 order (x fastest), bin
 * order inside a bin, self skipped (_stmt_translator.py:459-491).
 * Bins are found by binary search in the sorted keys here; the HIP path uses
 * the start/end tables, which give the same ranges. */
static uint32_t lower_bound_u32(const uint32_t* a, uint32_t n, uint32_t k) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t m = lo + (hi - lo) / 2;
        if (a[m] < k) lo = m + 1; else hi = m;
    }
    return lo;
}

static void force_on(const onc_sim_desc* d, const uint32_t* keys, uint32_t n, const REAL* px,
                     const REAL* py, const REAL* pz, uint32_t i, REAL* fx, REAL* fy, REAL* fz) {
    REAL dx = (REAL)d->dx;
    REAL R = (REAL)d->interaction_radius;
    REAL S = (REAL)2 * (REAL)d->cell_radius;
    REAL R2 = R * R;
    REAL krep = (REAL)d->k_rep, kadh = (REAL)d->k_adh;
    REAL RmS = R - S;
    int32_t vx = vox1(px[i], dx, d->dims[0]), vy = vox1(py[i], dx, d->dims[1]),
            vz = vox1(pz[i], dx, d->dims[2]);
    REAL Fx = 0, Fy = 0, Fz = 0;
    for (int oz = -1; oz <= 1; oz++)
        for (int oy = -1; oy <= 1; oy++)
            for (int ox = -1; ox <= 1; ox++) {
                int32_t nx = vx + ox, ny = vy + oy, nz = vz + oz;
                if (nx < 0 || ny < 0 || nz < 0 || nx >= d->dims[0] || ny >= d->dims[1] || nz >= d->dims[2])
                    continue;
                uint32_t key = oracle_morton((uint32_t)nx, (uint32_t)ny, (uint32_t)nz, d->key_format);
                uint32_t s = lower_bound_u32(keys, n, key);
                for (uint32_t j = s; j < n && keys[j] == key; j++) {
                    if (j == i) continue;
                    REAL ddx = px[i] - px[j];
                    REAL ddy = py[i] - py[j];
                    REAL ddz = pz[i] - pz[j];
                    REAL d2 = ddx * ddx + ddy * ddy;
                    d2 = d2 + ddz * ddz;
                    if (!(d2 < R2) || d2 == (REAL)0) continue;
                    REAL dist = RSQRT(d2);
                    REAL f;
                    if (dist < S) {
                        f = krep * (S - dist);
                    } else {
                        if (kadh == (REAL)0) continue;
                        f = -(kadh * ((R - dist) / RmS));
                    }
                    REAL sc = f / dist; /* F += f * delta / d */
                    Fx = Fx + sc * ddx;
                    Fy = Fy + sc * ddy;
                    Fz = Fz + sc * ddz;
                }
            }
    *fx = Fx;
    *fy = Fy;
    *fz = Fz;
}

/* Raw forces on Morton-sorted positions (for the pair-law tests). */
void FN(oracle_forces_raw)(const onc_sim_desc* d, uint32_t n, const REAL* px, const REAL* py,
                           const REAL* pz, REAL* fx, REAL* fy, REAL* fz) {
    uint32_t* keys = (uint32_t*)malloc((size_t)(n ? n : 1) * 4);
    FN(oracle_keys)(px, py, pz, n, d->dims, (REAL)d->dx, d->key_format, keys);
    for (uint32_t i = 0; i < n; i++) force_on(d, keys, n, px, py, pz, i, &fx[i], &fy[i], &fz[i]);
    free(keys);
}

/* Integration of one cell: dp = mobility F, |dp| clamped to max_disp,
 * p' = clamp(p + dp, 0, dims * dx) per axis. */
static void integrate(const onc_sim_desc* d, REAL Fx, REAL Fy, REAL Fz, REAL* x, REAL* y, REAL* z) {
    REAL mob = (REAL)d->mobility, md = (REAL)d->max_disp;
    REAL dpx = mob * Fx, dpy = mob * Fy, dpz = mob * Fz;
    REAL l2 = dpx * dpx + dpy * dpy;
    l2 = l2 + dpz * dpz;
    if (l2 > md * md && ORACLE_PLANT != 3) {
        REAL sc = md / RSQRT(l2);
        dpx = dpx * sc;
        dpy = dpy * sc;
        dpz = dpz * sc;
    }
    REAL dx = (REAL)d->dx;
    if (ORACLE_PLANT == 4) {
        *x = *x + dpx;
        *y = *y + dpy;
        *z = *z + dpz;
        return;
    }
    *x = clampr(*x + dpx, (REAL)0, (REAL)d->dims[0] * dx);
    *y = clampr(*y + dpy, (REAL)0, (REAL)d->dims[1] * dx);
    *z = clampr(*z + dpz, (REAL)0, (REAL)d->dims[2] * dx);
}

/* ---- built-in tumour rule (the benchmark model; the DSL form is
 * oncolytica_amd/spheroid.py) ----------------------------------------
 * age += 1; die if O2 < hypoxia; else (two fields) die if rand() <
 * drug_kill * drug; else divide if age > div_age and O2 > div_o2: daughter
 * at p' + offset * random_dir(), age 0, generation + 1; parent age 0,
 * generation + 1. Death removes the cell at compaction (core/utils/_types.py:
 * 578-579, core/utils/_containers.py:114-117); division appends
 * (core/utils/_types.py:575-576). */
typedef struct {
    uint32_t alive, div;
    REAL dpx, dpy, dpz, dage;
    uint32_t dgen;
} rule_out;

static void tumor_rule(const onc_sim_desc* d, uint32_t step, REAL* px, REAL* py, REAL* pz, uint32_t id,
                       REAL* age, uint32_t* gen, REAL* const* fields, rule_out* o) {
    const REAL* P = NULL;
    REAL pr[ONC_MAX_PARAMS];
    for (int k = 0; k < ONC_MAX_PARAMS; k++) pr[k] = (REAL)d->params[k];
    P = pr;
    REAL dx = (REAL)d->dx;
    size_t X = (size_t)d->dims[0], XY = (size_t)d->dims[0] * (size_t)d->dims[1];
    size_t v = (size_t)vox1(*px, dx, d->dims[0]) + X * (size_t)vox1(*py, dx, d->dims[1]) +
               XY * (size_t)vox1(*pz, dx, d->dims[2]);
    rng_t r;
    r.id = id; r.step = ORACLE_PLANT == 7 ? step + 1u : step; r.c = 0;
    r.key[0] = (uint32_t)d->seed; r.key[1] = (uint32_t)(d->seed >> 32);
    REAL o2 = fields[0][v];
    *age = *age + (REAL)1;
    o->alive = 1;
    o->div = 0;
    if (o2 < P[ONC_TP_HYPOXIA]) {
        o->alive = 0;
        return;
    }
    if (d->n_fields > 1 && P[ONC_TP_DRUG_KILL] > (REAL)0) {
        REAL drug = fields[1][v];
        if (ORACLE_PLANT == 6) (void)rng_uniform(&r);
        REAL u = rng_uniform(&r);
        if (ORACLE_PLANT == 7) r.step = step;
        if (ORACLE_PLANT == 5 ? u <= P[ONC_TP_DRUG_KILL] * drug : u < P[ONC_TP_DRUG_KILL] * drug) {
            o->alive = 0;
            return;
        }
    }
    if (*age > P[ONC_TP_DIV_AGE] && o2 > P[ONC_TP_DIV_O2]) {
        REAL ux, uy, uz;
        rng_dir(&r, &ux, &uy, &uz);
        REAL ofs = P[ONC_TP_DAUGHTER_OFS];
        o->dpx = clampr(*px + ux * ofs, (REAL)0, (REAL)d->dims[0] * dx);
        o->dpy = clampr(*py + uy * ofs, (REAL)0, (REAL)d->dims[1] * dx);
        o->dpz = clampr(*pz + uz * ofs, (REAL)0, (REAL)d->dims[2] * dx);
        o->dage = (REAL)0;
        o->dgen = *gen + 1u;
        *age = (REAL)0;
        *gen = *gen + 1u;
        o->div = 1;
    }
}

/* Uptake (tumor_sim.wgsl:387-412, made linear): each occupied voxel that is
 * not fixed loses (cells in it) * uptake per field, clamped below at lo. The
 * cells are key-sorted, so a voxel's cells form one run of equal keys; t[0..2]
 * are their positions. Pinned by the closed-form step tests (a voxel holding
 * k cells, the lo clamp, a fixed voxel, and the sweep seeing the uptake). */
static void uptake(const onc_sim_desc* d, const uint32_t* keys, uint32_t n, REAL* const* t,
                   REAL* const* fields) {
    REAL dx = (REAL)d->dx;
    size_t X = (size_t)d->dims[0], XY = (size_t)d->dims[0] * (size_t)d->dims[1];
    for (uint32_t i = 0; i < n;) {
        uint32_t e = i + 1;
        while (e < n && keys[e] == keys[i]) e++;
        int32_t vx = vox1(t[0][i], dx, d->dims[0]), vy = vox1(t[1][i], dx, d->dims[1]),
                vz = vox1(t[2][i], dx, d->dims[2]);
        size_t v = (size_t)vx + X * (size_t)vy + XY * (size_t)vz;
        uint32_t run = ORACLE_PLANT == 1 ? e - i + 1u : e - i;
        for (int f = 0; f < d->n_fields; f++) {
            const onc_field_desc* fd = &d->fields[f];
            if (fd->uptake == 0.0f || is_fixed(fd, d->dims, vx, vy, vz)) continue;
            REAL u = fields[f][v] - (REAL)run * (REAL)fd->uptake;
            fields[f][v] = u < (REAL)fd->lo ? (REAL)fd->lo : u;
        }
        i = e;
    }
}

/* ---- one whole timestep of the tumour model, in place --------------------
 * Arrays have d->capacity elements; fields has n_fields pointers to
 * dims[0]*dims[1]*dims[2] values. *n, *next_id, *step are updated. */
int FN(oracle_tumor_run)(const onc_sim_desc* d, uint32_t* n_io, uint32_t* next_id_io, uint32_t* step_io,
                         REAL* px, REAL* py, REAL* pz, uint32_t* id, REAL* age, uint32_t* gen,
                         REAL* const* fields, int nsteps) {
    uint32_t cap = d->capacity;
    size_t nvox = (size_t)d->dims[0] * (size_t)d->dims[1] * (size_t)d->dims[2];
    int kb = oracle_key_bits(d->dims);
    uint32_t* keys = (uint32_t*)malloc((size_t)cap * 4 + 4);
    uint32_t* perm = (uint32_t*)malloc((size_t)cap * 4 + 4);
    REAL* t[4];
    for (int k = 0; k < 4; k++) t[k] = (REAL*)malloc((size_t)cap * sizeof(REAL) + sizeof(REAL));
    uint32_t* tid = (uint32_t*)malloc((size_t)cap * 4 + 4);
    uint32_t* tgen = (uint32_t*)malloc((size_t)cap * 4 + 4);
    rule_out* ro = (rule_out*)malloc((size_t)cap * sizeof(rule_out) + sizeof(rule_out));
    REAL* swept[ONC_MAX_FIELDS];
    for (int f = 0; f < d->n_fields; f++) swept[f] = (REAL*)malloc(nvox * sizeof(REAL));
    REAL dx = (REAL)d->dx;

    for (int s = 0; s < nsteps; s++) {
        uint32_t n = *n_io, step = *step_io;
        /* 1-2: keys, stable sort, permute */
        FN(oracle_keys)(px, py, pz, n, d->dims, dx, d->key_format, keys);
        for (uint32_t i = 0; i < n; i++) perm[i] = i;
        oracle_radix_sort_pairs(keys, perm, n, kb);
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < n; i++) {
            uint32_t j = perm[i];
            t[0][i] = px[j]; t[1][i] = py[j]; t[2][i] = pz[j]; t[3][i] = age[j];
            tid[i] = id[j]; tgen[i] = gen[j];
        }
        /* 4: uptake, one update per occupied voxel */
        if (ORACLE_PLANT != 2) uptake(d, keys, n, t, fields);
        /* 5: forces and integration into px/py/pz (old order no longer needed) */
#pragma omp parallel for schedule(dynamic, 4096)
        for (uint32_t i = 0; i < n; i++) {
            REAL Fx, Fy, Fz;
            force_on(d, keys, n, t[0], t[1], t[2], i, &Fx, &Fy, &Fz);
            REAL x = t[0][i], y = t[1][i], z = t[2][i];
            integrate(d, Fx, Fy, Fz, &x, &y, &z);
            px[i] = x; py[i] = y; pz[i] = z;
        }
        /* 6: fields */
        for (int f = 0; f < d->n_fields; f++) FN(oracle_stencil7)(fields[f], swept[f], d->dims, &d->fields[f]);
        for (int f = 0; f < d->n_fields; f++) memcpy(fields[f], swept[f], nvox * sizeof(REAL));
        if (ORACLE_PLANT == 2) uptake(d, keys, n, t, fields);
        /* 7: rule */
#pragma omp parallel for schedule(static)
        for (uint32_t i = 0; i < n; i++) {
            REAL a = t[3][i];
            uint32_t g = tgen[i];
            tumor_rule(d, step, &px[i], &py[i], &pz[i], tid[i], &a, &g, fields, &ro[i]);
            t[3][i] = a;
            tgen[i] = g;
        }
        /* 8: compaction (survivors then daughters, both in sorted order) */
        uint32_t ns = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (!ro[i].alive) continue;
            px[ns] = px[i]; py[ns] = py[i]; pz[ns] = pz[i];
            age[ns] = t[3][i]; id[ns] = tid[i]; gen[ns] = tgen[i];
            ns++;
        }
        uint32_t nid = *next_id_io, m = ns;
        for (uint32_t i = 0; i < n && m < cap; i++) {
            if (!ro[i].div) continue;
            px[m] = ro[i].dpx; py[m] = ro[i].dpy; pz[m] = ro[i].dpz;
            age[m] = ro[i].dage; gen[m] = ro[i].dgen; id[m] = nid++;
            m++;
        }
        *n_io = m;
        *next_id_io = nid;
        *step_io = step + 1;
    }
    free(keys); free(perm); free(tid); free(tgen); free(ro);
    for (int k = 0; k < 4; k++) free(t[k]);
    for (int f = 0; f < d->n_fields; f++) free(swept[f]);
    return 0;
}

/* ---- steps 1-6 for any rule ------------------------------------------------
 * The part of the timestep before the rules (the same code the tumour step
 * above runs): keys of the n cells and their stable sort (perm[i] = index of
 * the cell at sorted position i, keys[i] its key), uptake, forces with the
 * integration, and one sweep of every field. spx..spz receive the sorted
 * positions before the forces, ipx..ipz the integrated ones (sorted order);
 * fields are replaced by their swept values. Arrays hold n elements. */
int FN(oracle_pre_rule)(const onc_sim_desc* d, uint32_t n, const REAL* px, const REAL* py, const REAL* pz,
                        uint32_t* perm, uint32_t* keys, REAL* spx, REAL* spy, REAL* spz, REAL* ipx, REAL* ipy,
                        REAL* ipz, REAL* const* fields) {
    size_t nvox = (size_t)d->dims[0] * (size_t)d->dims[1] * (size_t)d->dims[2];
    REAL dx = (REAL)d->dx;
    FN(oracle_keys)(px, py, pz, n, d->dims, dx, d->key_format, keys);
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    oracle_radix_sort_pairs(keys, perm, n, oracle_key_bits(d->dims));
    for (uint32_t i = 0; i < n; i++) {
        spx[i] = px[perm[i]];
        spy[i] = py[perm[i]];
        spz[i] = pz[perm[i]];
    }
    REAL* t[3] = {spx, spy, spz};
    uptake(d, keys, n, t, fields);
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint32_t i = 0; i < n; i++) {
        REAL Fx, Fy, Fz;
        force_on(d, keys, n, spx, spy, spz, i, &Fx, &Fy, &Fz);
        REAL x = spx[i], y = spy[i], z = spz[i];
        integrate(d, Fx, Fy, Fz, &x, &y, &z);
        ipx[i] = x;
        ipy[i] = y;
        ipz[i] = z;
    }
    for (int f = 0; f < d->n_fields; f++) {
        REAL* sw = (REAL*)malloc((nvox ? nvox : 1) * sizeof(REAL));
        if (!sw) return -1;
        FN(oracle_stencil7)(fields[f], sw, d->dims, &d->fields[f]);
        memcpy(fields[f], sw, nvox * sizeof(REAL));
        free(sw);
    }
    return 0;
}

/* Unit vector from two uniforms (the operation sequence of rng_dir above),
 * for draws of streams with another counter layout. */
void FN(oracle_dir_from_uniforms)(REAL u1, REAL u2, REAL out[3]) {
    REAL zz = u1 * (REAL)2 - (REAL)1;
    REAL rr = (REAL)1 - zz * zz;
    if (rr < (REAL)0) rr = (REAL)0;
    REAL rad = RSQRT(rr);
    REAL t = u2 * (REAL)4;
    int32_t q = (int32_t)t;
    REAL f = t - (REAL)q;
    REAL a = f * (REAL)1.5707963267948966;
    REAL s, c;
    FN(oracle_sincos_q)(a, &s, &c);
    REAL cx, sy;
    switch (q & 3) {
        case 0: cx = c; sy = s; break;
        case 1: cx = -s; sy = c; break;
        case 2: cx = -c; sy = -s; break;
        default: cx = s; sy = -c; break;
    }
    out[0] = rad * cx;
    out[1] = rad * sy;
    out[2] = zz;
}
