/*
 * CPU oracle for the Plume simulation substep and volume ray march.
 *
 * TEST INFRASTRUCTURE ONLY. This file is a scalar restatement of the reference's Warp kernels, written
 * from their source, used as the checker for the HIP path (tests, the smoke check and the CPU baseline leg
 * of the benchmark). Nothing in the product links or calls it.
 *
 * Built twice: REAL = float (the reference's precision, same operation order, no FMA contraction) and
 * REAL = double with ORC_DOUBLE (the FP64 oracle for error statements). Every function cites the reference file:line it
 * follows (paths relative to the reference repository root, src/plume/...).
 *
 * Arrays are C-order [i][j][k], k contiguous. A 3-d field is passed as a pointer plus its shape n[3].
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_DOUBLE
typedef double REAL;
#define SQRT sqrt
#define EXP exp
#define POW pow
#define FLOOR floor
#define CEIL ceil
#define FABS fabs
#else
typedef float REAL;
#define SQRT sqrtf
#define EXP expf
#define POW powf
#define FLOOR floorf
#define CEIL ceilf
#define FABS fabsf
#endif

#define R(x) ((REAL)(x))

typedef struct {
    int32_t dims[3];
    REAL cell_m;
    REAL origin_m[3];
    int32_t open_lo[3];
    int32_t open_hi[3];
    REAL fill_hi[3];
    REAL ghost_gap;
} orc_info;

typedef struct {
    int32_t shape;
    REAL center_m[3];
    REAL radius_m;
    REAL half_size_m[3];
    REAL edge_softness;
    REAL start_s;
    REAL end_s;
    REAL smoke_per_s;
    REAL temperature_k;
    REAL velocity_mps[3];
    int32_t outward;
    REAL noise_scale_m;
    REAL noise_strength;
    int32_t stream;
} orc_emitter;

typedef struct {
    REAL right[3];
    REAL up[3];
    REAL toward[3];
    REAL left_m;
    REAL top_m;
    REAL pixel_m;
    int32_t perspective;
    REAL eye[3];
} orc_view;

typedef struct {
    REAL opacity_per_density;
    REAL glow_start_k;
    REAL smoke_hides_heat;
    REAL normal_fade_thickness;
    int32_t shade_layout;
    REAL shade_fill;
} orc_look;

static inline int64_t at(const int *n, int i, int j, int k) { return ((int64_t)i * n[1] + j) * n[2] + k; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline REAL rmin(REAL a, REAL b) { return a < b ? a : b; }
static inline REAL rmax(REAL a, REAL b) { return a > b ? a : b; }
static inline REAL rclamp(REAL x, REAL lo, REAL hi) { return rmin(rmax(x, lo), hi); }
/* wp.lerp(a, b, t) = a * (1 - t) + b * t */
static inline REAL lerp(REAL a, REAL b, REAL t) { return a * (R(1.0) - t) + b * t; }
static inline REAL len3(const REAL *a) { return SQRT(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* ---------------------------------------------------------------- sampling: sim/sample.py */

/* sample_index_space, sim/sample.py:10-29 (scalar field) */
static REAL sample_index(const REAL *f, const int *n, REAL gx, REAL gy, REAL gz) {
    REAL x = rclamp(gx, R(0.0), (REAL)(n[0] - 1));
    REAL y = rclamp(gy, R(0.0), (REAL)(n[1] - 1));
    REAL z = rclamp(gz, R(0.0), (REAL)(n[2] - 1));
    int i0 = (int)x, j0 = (int)y, k0 = (int)z;
    int i1 = imin(i0 + 1, n[0] - 1), j1 = imin(j0 + 1, n[1] - 1), k1 = imin(k0 + 1, n[2] - 1);
    REAL tx = x - (REAL)i0, ty = y - (REAL)j0, tz = z - (REAL)k0;
    REAL c00 = lerp(f[at(n, i0, j0, k0)], f[at(n, i1, j0, k0)], tx);
    REAL c10 = lerp(f[at(n, i0, j1, k0)], f[at(n, i1, j1, k0)], tx);
    REAL c01 = lerp(f[at(n, i0, j0, k1)], f[at(n, i1, j0, k1)], tx);
    REAL c11 = lerp(f[at(n, i0, j1, k1)], f[at(n, i1, j1, k1)], tx);
    return lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
}

/* sample_index_space on a vec3 array (AoS, 3 values per cell), sim/sample.py:10-29 */
static void sample_index_vec3(const REAL *f, const int *n, REAL gx, REAL gy, REAL gz, REAL *out) {
    REAL x = rclamp(gx, R(0.0), (REAL)(n[0] - 1));
    REAL y = rclamp(gy, R(0.0), (REAL)(n[1] - 1));
    REAL z = rclamp(gz, R(0.0), (REAL)(n[2] - 1));
    int i0 = (int)x, j0 = (int)y, k0 = (int)z;
    int i1 = imin(i0 + 1, n[0] - 1), j1 = imin(j0 + 1, n[1] - 1), k1 = imin(k0 + 1, n[2] - 1);
    REAL tx = x - (REAL)i0, ty = y - (REAL)j0, tz = z - (REAL)k0;
    for (int c = 0; c < 3; c++) {
        REAL c00 = lerp(f[3 * at(n, i0, j0, k0) + c], f[3 * at(n, i1, j0, k0) + c], tx);
        REAL c10 = lerp(f[3 * at(n, i0, j1, k0) + c], f[3 * at(n, i1, j1, k0) + c], tx);
        REAL c01 = lerp(f[3 * at(n, i0, j0, k1) + c], f[3 * at(n, i1, j0, k1) + c], tx);
        REAL c11 = lerp(f[3 * at(n, i0, j1, k1) + c], f[3 * at(n, i1, j1, k1) + c], tx);
        out[c] = lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
    }
}

/* range_index_space, sim/sample.py:32-50 */
static void range_index(const REAL *f, const int *n, REAL gx, REAL gy, REAL gz, REAL *lo, REAL *hi) {
    REAL x = rclamp(gx, R(0.0), (REAL)(n[0] - 1));
    REAL y = rclamp(gy, R(0.0), (REAL)(n[1] - 1));
    REAL z = rclamp(gz, R(0.0), (REAL)(n[2] - 1));
    int i0 = (int)x, j0 = (int)y, k0 = (int)z;
    REAL l = f[at(n, i0, j0, k0)], h = l;
    for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++)
            for (int dk = 0; dk < 2; dk++) {
                REAL value = f[at(n, imin(i0 + di, n[0] - 1), imin(j0 + dj, n[1] - 1), imin(k0 + dk, n[2] - 1))];
                l = rmin(l, value);
                h = rmax(h, value);
            }
    *lo = l;
    *hi = h;
}

/* is_outside_open_face, sim/sample.py:53-61 */
static int outside_open_face(const REAL *p, const orc_info *info) {
    for (int a = 0; a < 3; a++) {
        if (p[a] < R(0.0) && info->open_lo[a] == 1) return 1;
        if (p[a] > (REAL)info->dims[a] && info->open_hi[a] == 1) return 1;
    }
    return 0;
}

static void face_shape(const orc_info *info, int axis, int *n) {
    n[0] = info->dims[0];
    n[1] = info->dims[1];
    n[2] = info->dims[2];
    if (axis >= 0) n[axis] += 1;
}

/* velocity_at, sim/sample.py:74-82 */
static void velocity_at(const orc_info *info, const REAL *u, const REAL *v, const REAL *w, const REAL *p,
                        REAL *out) {
    int nu[3], nv[3], nw[3];
    face_shape(info, 0, nu);
    face_shape(info, 1, nv);
    face_shape(info, 2, nw);
    out[0] = sample_index(u, nu, p[0] - R(0.0), p[1] - R(0.5), p[2] - R(0.5));
    out[1] = sample_index(v, nv, p[0] - R(0.5), p[1] - R(0.0), p[2] - R(0.5));
    out[2] = sample_index(w, nw, p[0] - R(0.5), p[1] - R(0.5), p[2] - R(0.0));
}

/* trace, sim/advect.py:16-21 (midpoint rule) */
static void trace(const orc_info *info, const REAL *u, const REAL *v, const REAL *w, const REAL *p, REAL dto,
                  REAL *out) {
    REAL vel[3], mid[3], half = R(0.5) * dto;
    velocity_at(info, u, v, w, p, vel);
    for (int a = 0; a < 3; a++) mid[a] = p[a] - half * vel[a];
    velocity_at(info, u, v, w, mid, vel);
    for (int a = 0; a < 3; a++) out[a] = p[a] - dto * vel[a];
}

/* sample_field, sim/sample.py:64-71 */
static REAL sample_field(const REAL *f, const int *n, const REAL *p, const REAL *offset, const orc_info *info,
                         REAL outside, int uses_outside) {
    if (uses_outside == 1 && outside_open_face(p, info)) return outside;
    return sample_index(f, n, p[0] - offset[0], p[1] - offset[1], p[2] - offset[2]);
}

/* ---------------------------------------------------------------- advection: sim/advect.py */

/* semi_lagrangian_kernel, sim/advect.py:24-31 */
void orc_advect_sl(const orc_info *info, const REAL *src, REAL *dst, const int *n, const REAL *offset,
                   REAL outside, int uses_outside, const REAL *u, const REAL *v, const REAL *w, REAL dto) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++) {
                REAL p[3] = {(REAL)i + offset[0], (REAL)j + offset[1], (REAL)k + offset[2]}, s[3];
                trace(info, u, v, w, p, dto, s);
                dst[at(n, i, j, k)] = sample_field(src, n, s, offset, info, outside, uses_outside);
            }
}

/* maccormack_correct_kernel, sim/advect.py:34-50 */
void orc_maccormack_correct(const orc_info *info, const REAL *original, const REAL *hat, const REAL *back,
                            REAL *dst, const int *n, const REAL *offset, REAL outside, int uses_outside,
                            const REAL *u, const REAL *v, const REAL *w, REAL dto) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++) {
                int64_t c = at(n, i, j, k);
                REAL p[3] = {(REAL)i + offset[0], (REAL)j + offset[1], (REAL)k + offset[2]}, s[3];
                REAL corrected = hat[c] + R(0.5) * (original[c] - back[c]);
                trace(info, u, v, w, p, dto, s);
                REAL lo = outside, hi = outside;
                if (uses_outside == 0 || !outside_open_face(s, info))
                    range_index(original, n, s[0] - offset[0], s[1] - offset[1], s[2] - offset[2], &lo, &hi);
                dst[c] = rclamp(corrected, lo, hi);
            }
}

/* advect_maccormack, sim/advect.py:59-68 */
void orc_advect_maccormack(const orc_info *info, const REAL *data, REAL *hat, REAL *back, REAL *next,
                           const int *n, const REAL *offset, REAL outside, int uses_outside, const REAL *u,
                           const REAL *v, const REAL *w, REAL dto) {
    /* dt_over_cell is dt / cell_m, formed by the caller in double and rounded once, as the reference's
       Python does (sim/advect.py:62) */
    orc_advect_sl(info, data, hat, n, offset, outside, uses_outside, u, v, w, dto);
    orc_advect_sl(info, hat, back, n, offset, outside, uses_outside, u, v, w, -dto);
    orc_maccormack_correct(info, data, hat, back, next, n, offset, outside, uses_outside, u, v, w, dto);
}

/* ---------------------------------------------------------------- forces: sim/forces.py */

/* buoyancy_kernel, sim/forces.py:13-23 */
void orc_buoyancy(const orc_info *info, REAL *v, const REAL *smoke, const REAL *temperature, REAL gravity,
                  REAL heat_buoyancy, REAL smoke_weight, REAL ambient_k, REAL dt) {
    int nv[3], nc[3];
    face_shape(info, 1, nv);
    face_shape(info, -1, nc);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nv[0]; i++)
        for (int j = 0; j < nv[1]; j++)
            for (int k = 0; k < nv[2]; k++) {
                int below = imax(j - 1, 0), above = imin(j, info->dims[1] - 1);
                REAL heat = R(0.5) * (temperature[at(nc, i, below, k)] + temperature[at(nc, i, above, k)]);
                REAL density = R(0.5) * (smoke[at(nc, i, below, k)] + smoke[at(nc, i, above, k)]);
                REAL lift = heat_buoyancy * (heat - ambient_k) / ambient_k - smoke_weight * density;
                int64_t c = at(nv, i, j, k);
                v[c] = v[c] + gravity * lift * dt;
            }
}

/* drag_kernel, sim/forces.py:34-37 (one face array) */
void orc_drag(REAL *f, int64_t count, REAL wind, REAL keep) {
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < count; c++) f[c] = wind + (f[c] - wind) * keep;
}

/* ---------------------------------------------------------------- vorticity: sim/vorticity.py */

/* center_velocity_kernel, sim/vorticity.py:17-22 */
void orc_center_velocity(const orc_info *info, const REAL *u, const REAL *v, const REAL *w, REAL *center) {
    int nu[3], nv[3], nw[3], nc[3];
    face_shape(info, 0, nu);
    face_shape(info, 1, nv);
    face_shape(info, 2, nw);
    face_shape(info, -1, nc);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nc[0]; i++)
        for (int j = 0; j < nc[1]; j++)
            for (int k = 0; k < nc[2]; k++) {
                int64_t c = at(nc, i, j, k);
                center[3 * c + 0] = R(0.5) * (u[at(nu, i, j, k)] + u[at(nu, i + 1, j, k)]);
                center[3 * c + 1] = R(0.5) * (v[at(nv, i, j, k)] + v[at(nv, i, j + 1, k)]);
                center[3 * c + 2] = R(0.5) * (w[at(nw, i, j, k)] + w[at(nw, i, j, k + 1)]);
            }
}

/* central_difference, sim/vorticity.py:25-35 */
static void central_difference(const REAL *f, const int *n, const int *cell, int axis, REAL cell_m, REAL *out) {
    int lo[3] = {cell[0], cell[1], cell[2]}, hi[3] = {cell[0], cell[1], cell[2]};
    lo[axis] = imax(cell[axis] - 1, 0);
    hi[axis] = imin(cell[axis] + 1, n[axis] - 1);
    REAL span = (REAL)(hi[axis] - lo[axis]);
    if (span == R(0.0)) {
        out[0] = out[1] = out[2] = R(0.0);
        return;
    }
    REAL denom = span * cell_m;
    int64_t h = at(n, hi[0], hi[1], hi[2]), l = at(n, lo[0], lo[1], lo[2]);
    for (int c = 0; c < 3; c++) out[c] = (f[3 * h + c] - f[3 * l + c]) / denom;
}

/* curl_kernel, sim/vorticity.py:38-45 */
void orc_curl(const orc_info *info, const REAL *center, REAL *curl) {
    int nc[3];
    face_shape(info, -1, nc);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nc[0]; i++)
        for (int j = 0; j < nc[1]; j++)
            for (int k = 0; k < nc[2]; k++) {
                int cell[3] = {i, j, k};
                REAL dx[3], dy[3], dz[3];
                central_difference(center, nc, cell, 0, info->cell_m, dx);
                central_difference(center, nc, cell, 1, info->cell_m, dy);
                central_difference(center, nc, cell, 2, info->cell_m, dz);
                int64_t c = at(nc, i, j, k);
                curl[3 * c + 0] = dy[2] - dz[1];
                curl[3 * c + 1] = dz[0] - dx[2];
                curl[3 * c + 2] = dx[1] - dy[0];
            }
}

/* confinement_push_kernel, sim/vorticity.py:48-71 */
void orc_confinement_push(const orc_info *info, const REAL *curl, REAL *push, REAL strength) {
    int nc[3];
    face_shape(info, -1, nc);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nc[0]; i++)
        for (int j = 0; j < nc[1]; j++)
            for (int k = 0; k < nc[2]; k++) {
                int cell[3] = {i, j, k};
                REAL toward[3] = {R(0.0), R(0.0), R(0.0)};
                for (int a = 0; a < 3; a++) {
                    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
                    lo[a] = imax(cell[a] - 1, 0);
                    hi[a] = imin(cell[a] + 1, info->dims[a] - 1);
                    REAL span = (REAL)(hi[a] - lo[a]);
                    if (span > R(0.0)) {
                        REAL sh = len3(&curl[3 * at(nc, hi[0], hi[1], hi[2])]);
                        REAL sl = len3(&curl[3 * at(nc, lo[0], lo[1], lo[2])]);
                        toward[a] = (sh - sl) / (span * info->cell_m);
                    }
                }
                int64_t c = at(nc, i, j, k);
                REAL length = len3(toward);
                if (length < R(1.0e-12)) {
                    push[3 * c] = push[3 * c + 1] = push[3 * c + 2] = R(0.0);
                    continue;
                }
                REAL nrm[3] = {toward[0] / length, toward[1] / length, toward[2] / length};
                const REAL *w = &curl[3 * c];
                REAL cr[3] = {nrm[1] * w[2] - nrm[2] * w[1], nrm[2] * w[0] - nrm[0] * w[2],
                              nrm[0] * w[1] - nrm[1] * w[0]};
                REAL scale = strength * info->cell_m;
                for (int a = 0; a < 3; a++) push[3 * c + a] = scale * cr[a];
            }
}

/* add_face_push_kernel, sim/vorticity.py:74-84 */
void orc_add_face_push(const orc_info *info, REAL *f, const REAL *push, int axis, REAL dt) {
    int nf[3], nc[3];
    face_shape(info, axis, nf);
    face_shape(info, -1, nc);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nf[0]; i++)
        for (int j = 0; j < nf[1]; j++)
            for (int k = 0; k < nf[2]; k++) {
                int face[3] = {i, j, k}, below[3] = {i, j, k}, above[3] = {i, j, k};
                below[axis] = imax(face[axis] - 1, 0);
                above[axis] = imin(face[axis], info->dims[axis] - 1);
                REAL average = R(0.5) * (push[3 * at(nc, below[0], below[1], below[2]) + axis] +
                                         push[3 * at(nc, above[0], above[1], above[2]) + axis]);
                int64_t c = at(nf, i, j, k);
                f[c] = f[c] + average * dt;
            }
}

/* ---------------------------------------------------------------- noise: sim/noise.py */

/* hash_u32, sim/noise.py:13-21 (lowbias32) */
uint32_t orc_hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

#define LATTICE_OFFSET 1048576

/* lattice_value, sim/noise.py:24-32 */
REAL orc_lattice_value(int32_t seed, int32_t stream, int32_t i, int32_t j, int32_t k) {
    uint32_t h = orc_hash_u32((uint32_t)seed + (uint32_t)LATTICE_OFFSET);
    h = orc_hash_u32(h ^ (uint32_t)(stream + LATTICE_OFFSET));
    h = orc_hash_u32(h ^ (uint32_t)(i + LATTICE_OFFSET));
    h = orc_hash_u32(h ^ (uint32_t)(j + LATTICE_OFFSET));
    h = orc_hash_u32(h ^ (uint32_t)(k + LATTICE_OFFSET));
    return (REAL)(h >> 8) * (R(2.0) / R(16777215.0)) - R(1.0);
}

/* fade, sim/noise.py:35-37 */
static inline REAL fade(REAL t) { return t * t * (R(3.0) - R(2.0) * t); }

/* value_noise, sim/noise.py:40-57 */
REAL orc_value_noise(int32_t seed, int32_t stream, REAL px, REAL py, REAL pz) {
    REAL fx = FLOOR(px), fy = FLOOR(py), fz = FLOOR(pz);
    int i = (int)fx, j = (int)fy, k = (int)fz;
    REAL tx = fade(px - fx), ty = fade(py - fy), tz = fade(pz - fz);
    REAL c00 = lerp(orc_lattice_value(seed, stream, i, j, k), orc_lattice_value(seed, stream, i + 1, j, k), tx);
    REAL c10 = lerp(orc_lattice_value(seed, stream, i, j + 1, k), orc_lattice_value(seed, stream, i + 1, j + 1, k), tx);
    REAL c01 = lerp(orc_lattice_value(seed, stream, i, j, k + 1), orc_lattice_value(seed, stream, i + 1, j, k + 1), tx);
    REAL c11 = lerp(orc_lattice_value(seed, stream, i, j + 1, k + 1),
                    orc_lattice_value(seed, stream, i + 1, j + 1, k + 1), tx);
    return lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
}

/* ---------------------------------------------------------------- emitters: sim/emitters.py */

/* soft_edge, sim/emitters.py:41-50 */
static REAL soft_edge(REAL distance, REAL extent, REAL softness) {
    REAL inner = extent * (R(1.0) - softness);
    if (distance <= inner) return R(1.0);
    if (distance >= extent) return R(0.0);
    REAL t = (distance - inner) / (extent - inner);
    return R(1.0) - t * t * (R(3.0) - R(2.0) * t);
}

/* coverage, sim/emitters.py:53-67 */
static REAL coverage(const orc_emitter *e, const REAL *x, int32_t seed) {
    REAL off[3] = {x[0] - e->center_m[0], x[1] - e->center_m[1], x[2] - e->center_m[2]};
    REAL amount;
    if (e->shape == 0)
        amount = soft_edge(len3(off), e->radius_m, e->edge_softness);
    else
        amount = soft_edge(FABS(off[0]), e->half_size_m[0], e->edge_softness) *
                 soft_edge(FABS(off[1]), e->half_size_m[1], e->edge_softness) *
                 soft_edge(FABS(off[2]), e->half_size_m[2], e->edge_softness);
    if (amount > R(0.0) && e->noise_strength > R(0.0)) {
        REAL breakup = R(1.0) + e->noise_strength * orc_value_noise(seed, e->stream, x[0] / e->noise_scale_m,
                                                                    x[1] / e->noise_scale_m,
                                                                    x[2] / e->noise_scale_m);
        amount = amount * rmax(breakup, R(0.0));
    }
    return rmin(amount, R(1.0));
}

/* blend_factor, sim/emitters.py:70-77; BLEND_REFERENCE_S = 1/80 s (sim/emitters.py:20) */
static REAL blend_factor(REAL amount, REAL dt) {
    return R(1.0) - POW(R(1.0) - amount, dt / (REAL)(1.0 / 80.0));
}

/* is_active, sim/emitters.py:80-82 */
static int is_active(const orc_emitter *e, REAL time_s) { return time_s >= e->start_s && time_s < e->end_s; }

/* emit_scalars_kernel, sim/emitters.py:90-102 */
void orc_emit_scalars(const orc_info *info, REAL *smoke, REAL *temperature, const orc_emitter *emitters,
                      int n_emitters, int32_t seed, REAL time_s, REAL dt) {
    int nc[3];
    face_shape(info, -1, nc);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nc[0]; i++)
        for (int j = 0; j < nc[1]; j++)
            for (int k = 0; k < nc[2]; k++) {
                REAL x[3] = {info->origin_m[0] + ((REAL)i + R(0.5)) * info->cell_m,
                             info->origin_m[1] + ((REAL)j + R(0.5)) * info->cell_m,
                             info->origin_m[2] + ((REAL)k + R(0.5)) * info->cell_m};
                int64_t c = at(nc, i, j, k);
                for (int e = 0; e < n_emitters; e++) {
                    if (!is_active(&emitters[e], time_s)) continue;
                    REAL amount = coverage(&emitters[e], x, seed);
                    if (amount > R(0.0)) {
                        smoke[c] = smoke[c] + emitters[e].smoke_per_s * amount * dt;
                        temperature[c] = lerp(temperature[c], emitters[e].temperature_k, blend_factor(amount, dt));
                    }
                }
            }
}

/* emit_velocity_kernel, sim/emitters.py:105-121 (one face array) */
void orc_emit_velocity(const orc_info *info, REAL *component, int axis, const orc_emitter *emitters,
                       int n_emitters, int32_t seed, REAL time_s, REAL dt) {
    int nf[3];
    face_shape(info, axis, nf);
    REAL offset[3] = {R(0.5), R(0.5), R(0.5)};
    offset[axis] = R(0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nf[0]; i++)
        for (int j = 0; j < nf[1]; j++)
            for (int k = 0; k < nf[2]; k++) {
                REAL x[3] = {info->origin_m[0] + ((REAL)i + offset[0]) * info->cell_m,
                             info->origin_m[1] + ((REAL)j + offset[1]) * info->cell_m,
                             info->origin_m[2] + ((REAL)k + offset[2]) * info->cell_m};
                int64_t c = at(nf, i, j, k);
                for (int e = 0; e < n_emitters; e++) {
                    const orc_emitter *em = &emitters[e];
                    if (!is_active(em, time_s)) continue;
                    REAL amount = coverage(em, x, seed);
                    if (amount > R(0.0)) {
                        REAL target[3] = {em->velocity_mps[0], em->velocity_mps[1], em->velocity_mps[2]};
                        if (em->outward == 1) {
                            REAL away[3] = {x[0] - em->center_m[0], x[1] - em->center_m[1], x[2] - em->center_m[2]};
                            REAL la = len3(away);
                            if (la > R(0.0)) {
                                REAL speed = len3(em->velocity_mps);
                                for (int a = 0; a < 3; a++) target[a] = (away[a] / la) * speed;
                            }
                        }
                        component[c] = lerp(component[c], target[axis], blend_factor(amount, dt));
                    }
                }
            }
}

/* ---------------------------------------------------------------- turbulence: sim/turbulence.py */

/* potential, sim/turbulence.py:26-32 */
static REAL potential(int32_t seed, int component, REAL px, REAL py, REAL pz, REAL pattern_time) {
    REAL slice_lo = FLOOR(pattern_time);
    REAL blend = fade(pattern_time - slice_lo);
    int32_t first = (1 << 24) + 3 * (int32_t)slice_lo + component;
    return lerp(orc_value_noise(seed, first, px, py, pz), orc_value_noise(seed, first + 3, px, py, pz), blend);
}

/* turbulence_kernel, sim/turbulence.py:35-53 (one face array) */
void orc_turbulence(const orc_info *info, REAL *f, int axis, int32_t seed, REAL scale_m, REAL strength,
                    REAL pattern_time, REAL dt) {
    int nf[3];
    face_shape(info, axis, nf);
    REAL offset[3] = {R(0.5), R(0.5), R(0.5)};
    offset[axis] = R(0.0);
    int b = (axis + 1) % 3, cc = (axis + 2) % 3;
    REAL to_lattice = info->cell_m / scale_m;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nf[0]; i++)
        for (int j = 0; j < nf[1]; j++)
            for (int k = 0; k < nf[2]; k++) {
                REAL face[3] = {(REAL)i + offset[0], (REAL)j + offset[1], (REAL)k + offset[2]};
                REAL hb[3] = {R(0.0), R(0.0), R(0.0)}, hc[3] = {R(0.0), R(0.0), R(0.0)};
                hb[b] = R(0.5);
                hc[cc] = R(0.5);
                REAL pp[3], pm[3], qp[3], qm[3];
                for (int a = 0; a < 3; a++) {
                    pp[a] = (face[a] + hb[a]) * to_lattice;
                    pm[a] = (face[a] - hb[a]) * to_lattice;
                    qp[a] = (face[a] + hc[a]) * to_lattice;
                    qm[a] = (face[a] - hc[a]) * to_lattice;
                }
                REAL c_plus = potential(seed, cc, pp[0], pp[1], pp[2], pattern_time);
                REAL c_minus = potential(seed, cc, pm[0], pm[1], pm[2], pattern_time);
                REAL b_plus = potential(seed, b, qp[0], qp[1], qp[2], pattern_time);
                REAL b_minus = potential(seed, b, qm[0], qm[1], qm[2], pattern_time);
                REAL curl = ((c_plus - c_minus) - (b_plus - b_minus)) / to_lattice;
                int64_t c = at(nf, i, j, k);
                f[c] = f[c] + strength * curl * dt;
            }
}

/* ---------------------------------------------------------------- pressure: sim/pressure.py */

/* zero_closed_faces_kernel, sim/pressure.py:21-28 (one face array) */
void orc_zero_closed_faces(const orc_info *info, REAL *f, int axis) {
    int nf[3];
    face_shape(info, axis, nf);
    for (int i = 0; i < nf[0]; i++)
        for (int j = 0; j < nf[1]; j++)
            for (int k = 0; k < nf[2]; k++) {
                int idx[3] = {i, j, k};
                int index = idx[axis];
                if (index == 0 && info->open_lo[axis] == 0) f[at(nf, i, j, k)] = R(0.0);
                if (index == info->dims[axis] && info->open_hi[axis] == 0) f[at(nf, i, j, k)] = R(0.0);
            }
}

/* divergence_kernel, sim/pressure.py:31-36 */
void orc_divergence(const orc_info *info, const REAL *u, const REAL *v, const REAL *w, REAL *div) {
    int nu[3], nv[3], nw[3], nc[3];
    face_shape(info, 0, nu);
    face_shape(info, 1, nv);
    face_shape(info, 2, nw);
    face_shape(info, -1, nc);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nc[0]; i++)
        for (int j = 0; j < nc[1]; j++)
            for (int k = 0; k < nc[2]; k++) {
                REAL flux = (u[at(nu, i + 1, j, k)] - u[at(nu, i, j, k)]) + (v[at(nv, i, j + 1, k)] - v[at(nv, i, j, k)]) +
                            (w[at(nw, i, j, k + 1)] - w[at(nw, i, j, k)]);
                div[at(nc, i, j, k)] = flux / info->cell_m;
            }
}

/* cell_fill, sim/pressure.py:39-44 */
static inline REAL cell_fill(const orc_info *info, int axis, int index) {
    return index == info->dims[axis] - 1 ? info->fill_hi[axis] : R(1.0);
}

/* cell_volume, sim/pressure.py:47-49 */
static inline REAL cell_volume(const orc_info *info, const int *cell) {
    return cell_fill(info, 0, cell[0]) * cell_fill(info, 1, cell[1]) * cell_fill(info, 2, cell[2]);
}

/* neighbor_term, sim/pressure.py:52-69 */
static void neighbor_term(const REAL *q, const int *n, const int *cell, int axis, int step, const orc_info *info,
                          REAL *out) {
    REAL area = R(1.0);
    for (int other = 0; other < 3; other++)
        if (other != axis) area = area * cell_fill(info, other, cell[other]);
    REAL own = cell_fill(info, axis, cell[axis]);
    int nb[3] = {cell[0], cell[1], cell[2]};
    nb[axis] += step;
    if (nb[axis] >= 0 && nb[axis] < info->dims[axis]) {
        REAL coupling = area / (R(0.5) * (own + cell_fill(info, axis, nb[axis])));
        out[0] = coupling * q[at(n, nb[0], nb[1], nb[2])];
        out[1] = coupling;
        return;
    }
    if ((step < 0 && info->open_lo[axis] == 1) || (step > 0 && info->open_hi[axis] == 1)) {
        out[0] = R(0.0);
        out[1] = area / (R(0.5) * own + info->ghost_gap);
        return;
    }
    out[0] = R(0.0);
    out[1] = R(0.0);
}

/* neighbor_sum, sim/pressure.py:72-77 */
static void neighbor_sum(const REAL *q, const int *n, const int *cell, const orc_info *info, REAL *total) {
    total[0] = R(0.0);
    total[1] = R(0.0);
    for (int a = 0; a < 3; a++) {
        REAL lo[2], hi[2];
        neighbor_term(q, n, cell, a, -1, info, lo);
        neighbor_term(q, n, cell, a, 1, info, hi);
        total[0] = (total[0] + lo[0]) + hi[0];
        total[1] = (total[1] + lo[1]) + hi[1];
    }
}

/* Per-cell boundary flags: for each of the six neighbours (x-, x+, y-, y+, z-, z+) two bits, 0 an
   in-domain neighbour, 1 an open-face ghost at pressure 0, 2 a closed face (no coupling); bits 12-14 mark
   the last cell on x, y, z (whose fill is fill_hi). This is the integer content of neighbor_term's
   branches, sim/pressure.py:39-69. */
void orc_build_flags(const orc_info *info, uint16_t *flags) {
    const int *d = info->dims;
    for (int i = 0; i < d[0]; i++)
        for (int j = 0; j < d[1]; j++)
            for (int k = 0; k < d[2]; k++) {
                int cell[3] = {i, j, k};
                unsigned value = 0;
                for (int a = 0; a < 3; a++) {
                    for (int s = 0; s < 2; s++) {
                        int step = s == 0 ? -1 : 1;
                        int nb = cell[a] + step;
                        unsigned kind;
                        if (nb >= 0 && nb < d[a])
                            kind = 0;
                        else if ((step < 0 && info->open_lo[a] == 1) || (step > 0 && info->open_hi[a] == 1))
                            kind = 1;
                        else
                            kind = 2;
                        value |= kind << (2 * (2 * a + s));
                    }
                    if (cell[a] == d[a] - 1) value |= 1u << (12 + a);
                }
                flags[at(d, i, j, k)] = (uint16_t)value;
            }
}

/* red_black_sweep_kernel, sim/pressure.py:80-92 (one colour) */
void orc_rb_sweep(const orc_info *info, REAL *q, const REAL *rhs, int color, REAL relaxation) {
    const int *n = info->dims;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++) {
                if ((i + j + k) % 2 != color) continue;
                int cell[3] = {i, j, k};
                REAL total[2];
                neighbor_sum(q, n, cell, info, total);
                if (total[1] == R(0.0)) continue;
                int64_t c = at(n, i, j, k);
                REAL source = cell_volume(info, cell) * info->cell_m * info->cell_m * rhs[c];
                REAL target = (total[0] - source) / total[1];
                q[c] = q[c] + relaxation * (target - q[c]);
            }
}

/* red_black_sweeps, sim/pressure.py:135-141 */
void orc_rb_sweeps(const orc_info *info, REAL *q, const REAL *rhs, int sweeps, REAL relaxation) {
    for (int s = 0; s < sweeps; s++)
        for (int color = 0; color < 2; color++) orc_rb_sweep(info, q, rhs, color, relaxation);
}

/* subtract_gradient_kernel, sim/pressure.py:95-116 (one face array) */
void orc_subtract_gradient(const orc_info *info, REAL *f, const REAL *q, int axis) {
    int nf[3];
    face_shape(info, axis, nf);
    const int *n = info->dims;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nf[0]; i++)
        for (int j = 0; j < nf[1]; j++)
            for (int k = 0; k < nf[2]; k++) {
                int face[3] = {i, j, k}, below[3] = {i, j, k};
                int index = face[axis];
                below[axis] = index - 1;
                REAL q_above = R(0.0), q_below = R(0.0);
                if (index == 0) {
                    if (info->open_lo[axis] == 0) continue;
                    q_above = q[at(n, i, j, k)];
                } else if (index == info->dims[axis]) {
                    if (info->open_hi[axis] == 0) continue;
                    q_below = q[at(n, below[0], below[1], below[2])];
                } else {
                    q_above = q[at(n, i, j, k)];
                    q_below = q[at(n, below[0], below[1], below[2])];
                }
                int64_t c = at(nf, i, j, k);
                f[c] = f[c] - (q_above - q_below) / info->cell_m;
            }
}

/* ---------------------------------------------------------------- multigrid: sim/multigrid.py */

/* residual_kernel, sim/multigrid.py:36-43 */
void orc_residual(const orc_info *info, const REAL *q, const REAL *rhs, REAL *residual) {
    const int *n = info->dims;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++) {
                int cell[3] = {i, j, k};
                REAL total[2];
                neighbor_sum(q, n, cell, info, total);
                REAL scale = cell_volume(info, cell) * info->cell_m * info->cell_m;
                int64_t c = at(n, i, j, k);
                residual[c] = rhs[c] - (total[0] - total[1] * q[c]) / scale;
            }
}

/* restrict_kernel, sim/multigrid.py:46-60 */
void orc_restrict(const REAL *fine, const int *nf, REAL *coarse, const int *nc) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nc[0]; i++)
        for (int j = 0; j < nc[1]; j++)
            for (int k = 0; k < nc[2]; k++) {
                REAL total = R(0.0), count = R(0.0);
                for (int di = 0; di < 2; di++)
                    for (int dj = 0; dj < 2; dj++)
                        for (int dk = 0; dk < 2; dk++) {
                            int fi = 2 * i + di, fj = 2 * j + dj, fk = 2 * k + dk;
                            if (fi < nf[0] && fj < nf[1] && fk < nf[2]) {
                                total = total + fine[at(nf, fi, fj, fk)];
                                count = count + R(1.0);
                            }
                        }
                coarse[at(nc, i, j, k)] = total / count;
            }
}

/* prolong_add_kernel, sim/multigrid.py:63-68 */
void orc_prolong_add(const REAL *coarse, const int *nc, REAL *fine, const int *nf) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nf[0]; i++)
        for (int j = 0; j < nf[1]; j++)
            for (int k = 0; k < nf[2]; k++) {
                REAL gx = (REAL)i * R(0.5) - R(0.25), gy = (REAL)j * R(0.5) - R(0.25), gz = (REAL)k * R(0.5) - R(0.25);
                int64_t c = at(nf, i, j, k);
                fine[c] = fine[c] + sample_index(coarse, nc, gx, gy, gz);
            }
}

/* Multigrid._v_cycle, sim/multigrid.py:130-147, over caller-provided level arrays. COARSEST_SWEEPS = 20,
   COARSEST_SWEEPS_PER_CELL = 4, COARSEST_RELAXATION = 1.5, SMOOTHING_RELAXATION = 1.0
   (sim/multigrid.py:30-33). */
static void v_cycle(int index, int n_levels, const orc_info *infos, REAL **q, REAL **rhs, REAL **residual,
                    int sweeps) {
    const orc_info *info = &infos[index];
    if (index == n_levels - 1) {
        int longest = imax(info->dims[0], imax(info->dims[1], info->dims[2]));
        orc_rb_sweeps(info, q[index], rhs[index], imax(20, 4 * longest), R(1.5));
        return;
    }
    orc_rb_sweeps(info, q[index], rhs[index], sweeps, R(1.0));
    orc_residual(info, q[index], rhs[index], residual[index]);
    const orc_info *coarse = &infos[index + 1];
    orc_restrict(residual[index], info->dims, rhs[index + 1], coarse->dims);
    int64_t count = (int64_t)coarse->dims[0] * coarse->dims[1] * coarse->dims[2];
    memset(q[index + 1], 0, sizeof(REAL) * count);
    v_cycle(index + 1, n_levels, infos, q, rhs, residual, sweeps);
    orc_prolong_add(q[index + 1], coarse->dims, q[index], info->dims);
    orc_rb_sweeps(info, q[index], rhs[index], sweeps, R(1.0));
}

/* Multigrid.__call__, sim/multigrid.py:124-128 */
void orc_multigrid(int n_levels, const orc_info *infos, REAL **q, REAL **rhs, REAL **residual, int cycles,
                   int sweeps) {
    for (int c = 0; c < cycles; c++) v_cycle(0, n_levels, infos, q, rhs, residual, sweeps);
}

/* ---------------------------------------------------------------- budget and dissipation */

/* line_sums_kernel, sim/reduce.py:12-25 */
void orc_line_sums(const int *n, const REAL *a, const REAL *b, const REAL *c, REAL *lines) {
    int64_t plane = (int64_t)n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n[1]; j++)
        for (int k = 0; k < n[2]; k++) {
            REAL ta = R(0.0), tb = R(0.0), tc = R(0.0);
            for (int i = 0; i < n[0]; i++) {
                int64_t x = at(n, i, j, k);
                ta = ta + a[x];
                tb = tb + b[x];
                tc = tc + c[x];
            }
            int64_t l = (int64_t)j * n[2] + k;
            lines[l] = ta;
            lines[plane + l] = tb;
            lines[2 * plane + l] = tc;
        }
}

/* scale_kernel, sim/conserve.py:27-30 */
void orc_scale(REAL *f, int64_t count, REAL factor) {
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < count; c++) f[c] = f[c] * factor;
}

/* dissipate_kernel, sim/dissipate.py:10-15 */
void orc_dissipate(REAL *smoke, REAL *temperature, int64_t count, REAL smoke_keep, REAL heat_keep, REAL ambient_k) {
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < count; c++) {
        smoke[c] = smoke[c] * smoke_keep;
        temperature[c] = ambient_k + (temperature[c] - ambient_k) * heat_keep;
    }
}

/* ---------------------------------------------------------------- rendering: render/ */

/* slab_range, render/raymarch.py:55-69 */
static void slab_range(const REAL *origin, const REAL *dir, const REAL *bmin, const REAL *bmax, REAL *out) {
    REAL t_enter = R(-1.0e30), t_exit = R(1.0e30);
    for (int a = 0; a < 3; a++) {
        if (FABS(dir[a]) < R(1.0e-12)) {
            if (origin[a] < bmin[a] || origin[a] > bmax[a]) {
                out[0] = R(1.0);
                out[1] = R(0.0);
                return;
            }
        } else {
            REAL t0 = (bmin[a] - origin[a]) / dir[a];
            REAL t1 = (bmax[a] - origin[a]) / dir[a];
            t_enter = rmax(t_enter, rmin(t0, t1));
            t_exit = rmin(t_exit, rmax(t0, t1));
        }
    }
    out[0] = t_enter;
    out[1] = t_exit;
}

/* light_volume_kernel, render/light_volume.py:18-33 */
/* max_steps > 0: the benchmark's fixed-count mode (no reference counterpart), that many equal steps per
   cell over the same path instead of half-cell steps */
void orc_light_volume(const int *n, const REAL *smoke, const REAL *toward_light, REAL cell_m,
                      REAL opacity_per_density, int max_steps, REAL *light) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++) {
                REAL start[3] = {(REAL)i + R(0.5), (REAL)j + R(0.5), (REAL)k + R(0.5)};
                REAL zero[3] = {R(0.0), R(0.0), R(0.0)}, box_max[3] = {(REAL)n[0], (REAL)n[1], (REAL)n[2]};
                REAL span[2];
                slab_range(start, toward_light, zero, box_max, span);
                REAL length = span[1];
                int steps = (int)CEIL(length / R(0.5));
                if (max_steps > 0 && steps > 0) steps = max_steps;
                REAL column = R(0.0);
                if (steps > 0) {
                    REAL step = length / (REAL)steps;
                    for (int s = 0; s < steps; s++) {
                        REAL t = ((REAL)s + R(0.5)) * step;
                        REAL px = start[0] + toward_light[0] * t, py = start[1] + toward_light[1] * t,
                             pz = start[2] + toward_light[2] * t;
                        column = column + rmax(sample_index(smoke, n, px - R(0.5), py - R(0.5), pz - R(0.5)), R(0.0)) * step;
                    }
                }
                light[at(n, i, j, k)] = EXP(-opacity_per_density * column * cell_m);
            }
}

/* blur_axis_kernel, render/normal_field.py:12-22 */
void orc_blur_axis(const int *n, const REAL *src, REAL *dst, const REAL *weights, int radius, int axis) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++) {
                int cell[3] = {i, j, k};
                REAL total = R(0.0);
                for (int off = -radius; off <= radius; off++) {
                    int nb[3] = {i, j, k};
                    nb[axis] = imin(imax(cell[axis] + off, 0), n[axis] - 1);
                    total = total + weights[off + radius] * src[at(n, nb[0], nb[1], nb[2])];
                }
                dst[at(n, i, j, k)] = total;
            }
}

/* gradient_kernel, render/normal_field.py:25-38 */
void orc_gradient(const int *n, const REAL *src, REAL *gradient) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++) {
                int cell[3] = {i, j, k};
                REAL result[3] = {R(0.0), R(0.0), R(0.0)};
                for (int a = 0; a < 3; a++) {
                    int lo[3] = {i, j, k}, hi[3] = {i, j, k};
                    lo[a] = imax(cell[a] - 1, 0);
                    hi[a] = imin(cell[a] + 1, n[a] - 1);
                    REAL span = (REAL)(hi[a] - lo[a]);
                    if (span > R(0.0))
                        result[a] = (src[at(n, hi[0], hi[1], hi[2])] - src[at(n, lo[0], lo[1], lo[2])]) / span;
                }
                int64_t c = at(n, i, j, k);
                gradient[3 * c] = result[0];
                gradient[3 * c + 1] = result[1];
                gradient[3 * c + 2] = result[2];
            }
}

/* raymarch_kernel, render/raymarch.py:72-140 (pixel_ray at :72-83, FAR_M = 1e4 at :21). max_steps > 0 fixes
   every ray at that many samples (the fixed-count mode of the GPU march); 0 is the reference's half-cell
   stepping. */
void orc_raymarch(const orc_info *info, const REAL *smoke, const REAL *temperature, const REAL *gradient,
                  const int *ng, const REAL *light, const int *nl, const orc_view *view, const orc_look *look,
                  REAL *image, int height, int width, int max_steps) {
    const int *n = info->dims;
#pragma omp parallel for schedule(dynamic, 1)
    for (int py = 0; py < height; py++)
        for (int px = 0; px < width; px++) {
            REAL plane_x = view->left_m + ((REAL)px + R(0.5)) * view->pixel_m;
            REAL plane_y = view->top_m - ((REAL)py + R(0.5)) * view->pixel_m;
            REAL origin[3], dir[3];
            for (int a = 0; a < 3; a++) {
                /* perspective 2: the orthographic ray started on the camera plane (no FAR_M offset) */
                origin[a] = view->right[a] * plane_x + view->up[a] * plane_y +
                            view->toward[a] * (view->perspective == 2 ? R(0.0) : R(1.0e4));
                dir[a] = -view->toward[a];
            }
            if (view->perspective == 1) {
                REAL d[3];
                for (int a = 0; a < 3; a++) {
                    origin[a] = view->eye[a];
                    d[a] = view->right[a] * plane_x + view->up[a] * plane_y - view->toward[a];
                }
                REAL l = len3(d);
                for (int a = 0; a < 3; a++) dir[a] = l > R(0.0) ? d[a] / l : R(0.0);
            }
            REAL box_max[3], span[2];
            for (int a = 0; a < 3; a++) box_max[a] = info->origin_m[a] + (REAL)n[a] * info->cell_m;
            slab_range(origin, dir, info->origin_m, box_max, span);
            if (view->perspective == 1) span[0] = rmax(span[0], R(0.0));

            REAL tr = R(1.0), heat = R(0.0), shade_sum = R(0.0);
            REAL normal_sum[3] = {R(0.0), R(0.0), R(0.0)};
            if (span[1] > span[0]) {
                REAL length = span[1] - span[0];
                int steps = (int)CEIL(length / (R(0.5) * info->cell_m));
                if (max_steps > 0) steps = max_steps;
                REAL step_m = length / (REAL)steps;
                for (int s = 0; s < steps; s++) {
                    REAL t = span[0] + ((REAL)s + R(0.5)) * step_m;
                    REAL g[3];
                    for (int a = 0; a < 3; a++) {
                        REAL position = origin[a] + dir[a] * t;
                        g[a] = (position - info->origin_m[a]) / info->cell_m - R(0.5);
                    }
                    REAL density = rmax(sample_index(smoke, n, g[0], g[1], g[2]), R(0.0));
                    REAL alpha = R(1.0) - EXP(-look->opacity_per_density * density * step_m);
                    REAL glow = rmax(sample_index(temperature, n, g[0], g[1], g[2]) - look->glow_start_k, R(0.0));
                    heat = heat + glow * step_m * POW(tr, look->smoke_hides_heat);
                    if (look->shade_layout == 1) {
                        REAL lit = look->shade_fill + (R(1.0) - look->shade_fill) * sample_index(light, nl, g[0], g[1], g[2]);
                        shade_sum = shade_sum + tr * alpha * lit;
                    } else {
                        REAL slope[3];
                        sample_index_vec3(gradient, ng, g[0], g[1], g[2], slope);
                        REAL ls = len3(slope);
                        if (alpha > R(0.0) && ls > R(1.0e-9)) {
                            REAL wgt = tr * alpha;
                            for (int a = 0; a < 3; a++) normal_sum[a] = normal_sum[a] - wgt * (slope[a] / ls);
                        }
                    }
                    tr = tr * (R(1.0) - alpha);
                }
            }
            REAL thickness = R(1.0) - tr;
            REAL *pix = &image[((int64_t)py * width + px) * 5];
            pix[0] = heat;
            pix[1] = thickness;
            if (look->shade_layout == 1) {
                pix[2] = shade_sum;
                pix[3] = R(0.0);
                pix[4] = R(0.0);
                continue;
            }
            REAL normal[3] = {R(0.0), R(0.0), R(1.0)};
            REAL cam[3] = {
                normal_sum[0] * view->right[0] + normal_sum[1] * view->right[1] + normal_sum[2] * view->right[2],
                normal_sum[0] * view->up[0] + normal_sum[1] * view->up[1] + normal_sum[2] * view->up[2],
                normal_sum[0] * view->toward[0] + normal_sum[1] * view->toward[1] + normal_sum[2] * view->toward[2]};
            REAL lc = len3(cam);
            if (lc > R(1.0e-9)) {
                REAL nn[3] = {cam[0] / lc, cam[1] / lc, FABS(cam[2] / lc)};
                REAL weight = rclamp(thickness / look->normal_fade_thickness, R(0.0), R(1.0));
                REAL facing[3] = {R(0.0), R(0.0), R(1.0)}, mixed[3];
                for (int a = 0; a < 3; a++) mixed[a] = lerp(facing[a], nn[a], weight);
                REAL lm = len3(mixed);
                for (int a = 0; a < 3; a++) normal[a] = lm > R(0.0) ? mixed[a] / lm : R(0.0);
            }
            pix[2] = normal[0];
            pix[3] = normal[1];
            pix[4] = normal[2];
        }
}

/* OpenMP threads of the following calls (the benchmark's CPU baseline times the port at several counts) */
void orc_set_threads(int threads) { omp_set_num_threads(threads > 0 ? threads : 1); }
