// MacCormack advection shaped for gfx950 (sim/advect.py), bit-identical to the reference-shaped kernels in
// advect.hip for finite fields. Three savings, none of which changes a rounding:
//
//  1. The trace's first velocity sample sits on the grid: a field sample at index + offset reads each
//     velocity component at index + (offset - component offset), which is an exact integer on every axis
//     where the two offsets agree. There the trilinear weight is exactly 0 and lerp(a, b, 0) = a + b*0,
//     which is a for finite a and b except that a = -0 gives +0 when b >= +0. That sample is used only in
//     the midpoint p - (dt/2) v, and p - (+-0) = p for every p the trace starts from (p >= +0, and
//     +0 - (+-0) = +0), so those taps are skipped outright: the midpoint, and everything after it, is the
//     same bits. A cell centre reads 2 taps per component instead of 8, all issued together.
//  2. The forward pass samples `data` at the traced source, which is exactly where the correction takes
//     its clamp range, so it keeps the min and max of those 8 taps (`bounds`, 2 floats per sample) and the
//     correction skips its own forward trace.
//  3. Smoke and temperature sit at cell centres and move with the same velocity, so one trace serves both.
//
// Per sample: forward = trace (aligned + 24 taps) + 8 taps per field; correction = backward trace + 8 taps
// per field, instead of 56 + 112.
#include "common.h"

namespace plm {
namespace {

// Sample-point kinds: 0 cell centre, 1 u face, 2 v face, 3 w face. Offset of kind K on axis a.
__host__ __device__ constexpr float kind_offset(int kind, int axis) { return kind == axis + 1 ? 0.0f : 0.5f; }
// Axes (bit a) where velocity component c, sampled at a kind-K point, lands on an integer index.
__host__ __device__ constexpr int aligned_mask(int kind, int component) {
    return (kind_offset(kind, 0) == (component == 0 ? 0.0f : 0.5f) ? 1 : 0) |
           (kind_offset(kind, 1) == (component == 1 ? 0.0f : 0.5f) ? 2 : 0) |
           (kind_offset(kind, 2) == (component == 2 ? 0.0f : 0.5f) ? 4 : 0);
}

// sample_index_space (sim/sample.py:10-29) where the axes in MASK are known to sit exactly on an index in
// range, so their weight is 0 and their second taps are not read. The result may differ from the full sample
// only in the sign of a zero (see 1. above), so it serves the trace's midpoint alone.
template <int MASK>
__device__ __forceinline__ float sample_aligned(const float *__restrict__ f, Dims n, float gx, float gy, float gz) {
    const float hx = (float)(n.n0 - 1), hy = (float)(n.n1 - 1), hz = (float)(n.n2 - 1);
    const float x = (MASK & 1) ? gx : clamp_index(gx, hx);
    const float y = (MASK & 2) ? gy : clamp_index(gy, hy);
    const float z = (MASK & 4) ? gz : clamp_index(gz, hz);
    const int i0 = (int)x, j0 = (int)y, k0 = (int)z;
    const uint32_t o = offset_of(n, i0, j0, k0) * 4u;
    const uint32_t di = x < hx ? (uint32_t)(n.n1 * n.n2) * 4u : 0u, dj = y < hy ? (uint32_t)n.n2 * 4u : 0u;
    const uint32_t dk = z < hz ? 4u : 0u;
    const float tx = index_weight(x), ty = index_weight(y), tz = index_weight(z);
    auto xl = [&](uint32_t base) -> float {
        if (MASK & 1) return load_bytes(f, base);
        return lerp_(load_bytes(f, base), load_bytes(f, base + di), tx);
    };
    auto yl = [&](uint32_t base) -> float {
        if (MASK & 2) return xl(base);
        return lerp_(xl(base), xl(base + dj), ty);
    };
    if (MASK & 4) return yl(o);
    return lerp_(yl(o), yl(o + dk), tz);
}

// velocity_at (sim/sample.py:74-82) at the kind-K sample point of index (i, j, k)
template <int KIND>
__device__ __forceinline__ void velocity_on_grid(const plm_grid_info &g, const float *__restrict__ u,
                                                 const float *__restrict__ v, const float *__restrict__ w, float px,
                                                 float py, float pz, float &vx, float &vy, float &vz) {
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    vx = sample_aligned<aligned_mask(KIND, 0)>(u, Dims{nx + 1, ny, nz}, px - 0.0f, py - 0.5f, pz - 0.5f);
    vy = sample_aligned<aligned_mask(KIND, 1)>(v, Dims{nx, ny + 1, nz}, px - 0.5f, py - 0.0f, pz - 0.5f);
    vz = sample_aligned<aligned_mask(KIND, 2)>(w, Dims{nx, ny, nz + 1}, px - 0.5f, py - 0.5f, pz - 0.0f);
}

// trace (sim/advect.py:16-21) from a kind-K sample point
template <int KIND>
__device__ __forceinline__ void trace_from_grid(const plm_grid_info &g, const float *__restrict__ u,
                                                const float *__restrict__ v, const float *__restrict__ w, float px,
                                                float py, float pz, float dto, float &sx, float &sy, float &sz) {
    float vx, vy, vz;
    velocity_on_grid<KIND>(g, u, v, w, px, py, pz, vx, vy, vz);
    const float half = 0.5f * dto;
    const float mx = px - half * vx, my = py - half * vy, mz = pz - half * vz;
    velocity_at(g, u, v, w, mx, my, mz, vx, vy, vz);
    sx = px - dto * vx;
    sy = py - dto * vy;
    sz = pz - dto * vz;
}

// Trilinear sample and the min/max of its 8 taps at one continuous index (sample_index_space and
// range_index_space at the same point share their clamped corner).
__device__ __forceinline__ float sample_and_range(const float *__restrict__ f, Dims n, float gx, float gy, float gz,
                                                  float &lo, float &hi) {
    const Corner c = corner_of<4>(n, gx, gy, gz);
    const uint32_t o1 = c.o + c.dk;
    const float f000 = load_bytes(f, c.o), f100 = load_bytes(f, c.o + c.di), f010 = load_bytes(f, c.o + c.dj),
                f110 = load_bytes(f, c.o + c.di + c.dj), f001 = load_bytes(f, o1), f101 = load_bytes(f, o1 + c.di),
                f011 = load_bytes(f, o1 + c.dj), f111 = load_bytes(f, o1 + c.di + c.dj);
    lo = fmin_(fmin_(fmin_(f000, f100), fmin_(f010, f110)), fmin_(fmin_(f001, f101), fmin_(f011, f111)));
    hi = fmax_(fmax_(fmax_(f000, f100), fmax_(f010, f110)), fmax_(fmax_(f001, f101), fmax_(f011, f111)));
    const float c00 = lerp_(f000, f100, c.tx), c10 = lerp_(f010, f110, c.tx);
    const float c01 = lerp_(f001, f101, c.tx), c11 = lerp_(f011, f111, c.tx);
    return lerp_(lerp_(c00, c10, c.ty), lerp_(c01, c11, c.ty), c.tz);
}

struct Field {
    const float *data;
    float *hat;
    float *bounds;  // [0, count) lo, [count, 2 count) hi
    float *next;
    float outside;
    int uses_outside;
};

template <int KIND> __device__ __forceinline__ Dims kind_dims(const plm_grid_info &g) {
    return Dims{g.dims[0] + (KIND == 1), g.dims[1] + (KIND == 2), g.dims[2] + (KIND == 3)};
}

// Forward pass: hat = SL(data, +dto) and the clamp range at the same source, for NF fields of kind KIND.
template <int KIND, int NF>
__device__ __forceinline__ void forward_sample(const plm_grid_info &g, const Field *fields,
                                               const float *__restrict__ u, const float *__restrict__ v,
                                               const float *__restrict__ w, int i, int j, int k, float dto) {
    const Dims n = kind_dims<KIND>(g);
    const float ox = kind_offset(KIND, 0), oy = kind_offset(KIND, 1), oz = kind_offset(KIND, 2);
    const float px = (float)i + ox, py = (float)j + oy, pz = (float)k + oz;
    float sx, sy, sz;
    trace_from_grid<KIND>(g, u, v, w, px, py, pz, dto, sx, sy, sz);
    const int c = n.at(i, j, k), count = n.count();
    const bool outside = outside_open_face(g, sx, sy, sz);
    // every field's taps are read before anything is stored (the clamped taps lie in range wherever the
    // source is), so their loads are in flight together; a source beyond an open face then takes the
    // field's outside value, as the branch of sample_field does
    float value[NF], lo[NF], hi[NF];
#pragma unroll
    for (int f = 0; f < NF; f++)
        value[f] = sample_and_range(fields[f].data, n, sx - ox, sy - oy, sz - oz, lo[f], hi[f]);
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const Field &fd = fields[f];
        if (fd.uses_outside == 1 && outside) value[f] = lo[f] = hi[f] = fd.outside;
        fd.hat[c] = value[f];
        fd.bounds[c] = lo[f];
        fd.bounds[count + c] = hi[f];
    }
}

// Correction: back = SL(hat, -dto), next = clamp(hat + (data - back) / 2, lo, hi)
template <int KIND, int NF>
__device__ __forceinline__ void correct_sample(const plm_grid_info &g, const Field *fields,
                                               const float *__restrict__ u, const float *__restrict__ v,
                                               const float *__restrict__ w, int i, int j, int k, float dto) {
    const Dims n = kind_dims<KIND>(g);
    const float ox = kind_offset(KIND, 0), oy = kind_offset(KIND, 1), oz = kind_offset(KIND, 2);
    const float px = (float)i + ox, py = (float)j + oy, pz = (float)k + oz;
    float sx, sy, sz;
    trace_from_grid<KIND>(g, u, v, w, px, py, pz, -dto, sx, sy, sz);
    const int c = n.at(i, j, k), count = n.count();
    const bool outside = outside_open_face(g, sx, sy, sz);
    // every load before the first store, as in forward_sample
    float back[NF], hat[NF], data[NF], lo[NF], hi[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const Field &fd = fields[f];
        back[f] = sample_index(fd.hat, n, sx - ox, sy - oy, sz - oz);
        hat[f] = fd.hat[c];
        data[f] = fd.data[c];
        lo[f] = fd.bounds[c];
        hi[f] = fd.bounds[count + c];
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const Field &fd = fields[f];
        if (fd.uses_outside == 1 && outside) back[f] = fd.outside;
        const float corrected = hat[f] + 0.5f * (data[f] - back[f]);
        fd.next[c] = clamp_(corrected, lo[f], hi[f]);
    }
}

struct Faces3 {
    Field f[3];
};

__global__ __launch_bounds__(FACE_BLOCK) void k_velocity_forward(plm_grid_info g, FaceLaunch fl, Faces3 faces,
                                                                 float dto) {
    int axis, i, j, k;
    if (!packed_face(g, fl, axis, i, j, k)) return;
    const float *u = faces.f[0].data, *v = faces.f[1].data, *w = faces.f[2].data;
    if (axis == 0)
        forward_sample<1, 1>(g, &faces.f[0], u, v, w, i, j, k, dto);
    else if (axis == 1)
        forward_sample<2, 1>(g, &faces.f[1], u, v, w, i, j, k, dto);
    else
        forward_sample<3, 1>(g, &faces.f[2], u, v, w, i, j, k, dto);
}

__global__ __launch_bounds__(FACE_BLOCK) void k_velocity_correct2(plm_grid_info g, FaceLaunch fl, Faces3 faces,
                                                                  float dto) {
    int axis, i, j, k;
    if (!packed_face(g, fl, axis, i, j, k)) return;
    const float *u = faces.f[0].data, *v = faces.f[1].data, *w = faces.f[2].data;
    if (axis == 0)
        correct_sample<1, 1>(g, &faces.f[0], u, v, w, i, j, k, dto);
    else if (axis == 1)
        correct_sample<2, 1>(g, &faces.f[1], u, v, w, i, j, k, dto);
    else
        correct_sample<3, 1>(g, &faces.f[2], u, v, w, i, j, k, dto);
}

struct Scalars {
    Field f[2];
};

template <int NF>
__global__ __launch_bounds__(256) void k_scalars_forward(plm_grid_info g, Scalars s, const float *__restrict__ u,
                                                         const float *__restrict__ v, const float *__restrict__ w,
                                                         float dto) {
    int i, j, k;
    if (!thread_cell(cell_dims(g), i, j, k)) return;
    forward_sample<0, NF>(g, s.f, u, v, w, i, j, k, dto);
}

template <int NF>
__global__ __launch_bounds__(256) void k_scalars_correct(plm_grid_info g, Scalars s, const float *__restrict__ u,
                                                         const float *__restrict__ v, const float *__restrict__ w,
                                                         float dto) {
    int i, j, k;
    if (!thread_cell(cell_dims(g), i, j, k)) return;
    correct_sample<0, NF>(g, s.f, u, v, w, i, j, k, dto);
}

}  // namespace
}  // namespace plm

using namespace plm;

extern "C" {

int plm_advect_velocity_bounded(const plm_grid_info *info, const float *u, const float *v, const float *w,
                                float *u_hat, float *v_hat, float *w_hat, float *u_bounds, float *v_bounds,
                                float *w_bounds, float *u_next, float *v_next, float *w_next, float dt_over_cell,
                                void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && u_hat && v_hat && w_hat && u_bounds && v_bounds && w_bounds && u_next && v_next &&
                    w_next,
                "plm_advect_velocity_bounded: NULL array");
    hipStream_t s = as_stream(stream);
    Faces3 faces{{Field{u, u_hat, u_bounds, u_next, 0.0f, 0}, Field{v, v_hat, v_bounds, v_next, 0.0f, 0},
                  Field{w, w_hat, w_bounds, w_next, 0.0f, 0}}};
    TimingScope t("advect", s);
    const FaceLaunch fl = face_launch(*info);
    k_velocity_forward<<<fl.block_end[2], FACE_BLOCK, 0, s>>>(*info, fl, faces, dt_over_cell);
    PLM_TRY(check_launch("plm_advect_velocity_bounded forward"));
    k_velocity_correct2<<<fl.block_end[2], FACE_BLOCK, 0, s>>>(*info, fl, faces, dt_over_cell);
    return check_launch("plm_advect_velocity_bounded correct");
}

int plm_advect_scalars(const plm_grid_info *info, const float *a_data, float *a_hat, float *a_bounds, float *a_next,
                       float a_outside, const float *b_data, float *b_hat, float *b_bounds, float *b_next,
                       float b_outside, const float *u, const float *v, const float *w, float dt_over_cell,
                       void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(a_data && a_hat && a_bounds && a_next && u && v && w, "plm_advect_scalars: NULL array");
    const bool two = b_data != nullptr;
    PLM_REQUIRE(!two || (b_hat && b_bounds && b_next), "plm_advect_scalars: NULL array of the second field");
    hipStream_t s = as_stream(stream);
    Scalars sc{{Field{a_data, a_hat, a_bounds, a_next, a_outside, 1},
                Field{b_data, b_hat, b_bounds, b_next, b_outside, 1}}};
    const Dims n = cell_dims(*info);
    TimingScope t("advect", s);
    if (two)
        k_scalars_forward<2><<<grid3(n), block3(), 0, s>>>(*info, sc, u, v, w, dt_over_cell);
    else
        k_scalars_forward<1><<<grid3(n), block3(), 0, s>>>(*info, sc, u, v, w, dt_over_cell);
    PLM_TRY(check_launch("plm_advect_scalars forward"));
    if (two)
        k_scalars_correct<2><<<grid3(n), block3(), 0, s>>>(*info, sc, u, v, w, dt_over_cell);
    else
        k_scalars_correct<1><<<grid3(n), block3(), 0, s>>>(*info, sc, u, v, w, dt_over_cell);
    return check_launch("plm_advect_scalars correct");
}

}  // extern "C"
