// Body forces, vorticity confinement, emitters, noise, turbulence and dissipation:
// sim/forces.py, sim/vorticity.py, sim/emitters.py, sim/noise.py, sim/turbulence.py, sim/dissipate.py.
#include "common.h"

namespace plm {
namespace {

// ------------------------------------------------------------------ buoyancy, drag

// buoyancy_kernel, sim/forces.py:13-23, BUOYANCY_ROWS consecutive v rows (j) per thread: the rows' temperature
// and smoke (one more row than faces, the rows below and above shared between neighbouring faces) and their v
// are loaded first, then each face updated with the reference's operations. A quarter of the wavefronts of a
// face per thread, each with its loads in flight together.
constexpr int BUOYANCY_ROWS = 4;
__global__ __launch_bounds__(256) void k_buoyancy(plm_grid_info g, float *__restrict__ v,
                                                  const float *__restrict__ smoke,
                                                  const float *__restrict__ temperature, float gravity,
                                                  float heat_buoyancy, float smoke_weight, float ambient_k,
                                                  float dt) {
    constexpr int R = BUOYANCY_ROWS;
    const Dims nv = face_dims(g, 1), nc = cell_dims(g);
    const int k = blockIdx.x * BX + threadIdx.x, i = blockIdx.z;
    const int j0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.y * BY + threadIdx.y) * R));
    if (k >= nv.n2 || j0 >= nv.n1) return;
    // row m holds cell row clamp(j0 - 1 + m): face j0 + r averages rows r (below) and r + 1 (above)
    float heat_row[R + 1], smoke_row[R + 1], v_old[R];
#pragma unroll
    for (int m = 0; m <= R; m++) {
        const int row = min(max(j0 - 1 + m, 0), g.dims[1] - 1);
        heat_row[m] = temperature[nc.at(i, row, k)];
        smoke_row[m] = smoke[nc.at(i, row, k)];
    }
#pragma unroll
    for (int r = 0; r < R; r++) v_old[r] = v[nv.at(i, min(j0 + r, nv.n1 - 1), k)];
#pragma unroll
    for (int r = 0; r < R; r++) {
        if (j0 + r >= nv.n1) break;
        const float heat = 0.5f * (heat_row[r] + heat_row[r + 1]);
        const float density = 0.5f * (smoke_row[r] + smoke_row[r + 1]);
        const float lift = heat_buoyancy * (heat - ambient_k) / ambient_k - smoke_weight * density;
        v[nv.at(i, j0 + r, k)] = v_old[r] + gravity * lift * dt;
    }
}

inline dim3 buoyancy_grid(const plm_grid_info &g) {
    const Dims nv = face_dims(g, 1);
    return dim3((nv.n2 + BX - 1) / BX, (nv.n1 + BY * BUOYANCY_ROWS - 1) / (BY * BUOYANCY_ROWS), nv.n0);
}

// drag_kernel, sim/forces.py:34-37, on all three face arrays
__global__ __launch_bounds__(256) void k_drag(plm_grid_info g, FaceLaunch fl, Tri<float> f, float wx, float wy, float wz,
                                              float keep) {
    int axis, i, j, k;
    if (!packed_face(g, fl, axis, i, j, k)) return;
    const Dims d = face_dims(g, axis);
    const float wind = axis == 0 ? wx : (axis == 1 ? wy : wz);
    float *p = f.p[axis] + d.at(i, j, k);
    *p = wind + (*p - wind) * keep;
}

// ------------------------------------------------------------------ vorticity confinement

// center_velocity_kernel, sim/vorticity.py:17-22
__device__ __forceinline__ void center_velocity(const plm_grid_info &g, const float *__restrict__ u,
                                                const float *__restrict__ v, const float *__restrict__ w, int i,
                                                int j, int k, float &cx, float &cy, float &cz) {
    const Dims nu = face_dims(g, 0), nv = face_dims(g, 1), nw = face_dims(g, 2);
    cx = 0.5f * (u[nu.at(i, j, k)] + u[nu.at(i + 1, j, k)]);
    cy = 0.5f * (v[nv.at(i, j, k)] + v[nv.at(i, j + 1, k)]);
    cz = 0.5f * (w[nw.at(i, j, k)] + w[nw.at(i, j, k + 1)]);
}

__global__ __launch_bounds__(256) void k_center_velocity(plm_grid_info g, const float *__restrict__ u,
                                                         const float *__restrict__ v, const float *__restrict__ w,
                                                         float *__restrict__ center) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    float cx, cy, cz;
    center_velocity(g, u, v, w, i, j, k, cx, cy, cz);
    const int c = nc.at(i, j, k);
    center[3 * c] = cx;
    center[3 * c + 1] = cy;
    center[3 * c + 2] = cz;
}

// curl_kernel with central_difference, sim/vorticity.py:25-45. CENTER_FROM_FACES evaluates each centre
// velocity from the faces (same arithmetic as center_velocity_kernel) instead of reading `center`.
template <bool CENTER_FROM_FACES>
__global__ __launch_bounds__(256) void k_curl(plm_grid_info g, const float *__restrict__ center,
                                              const float *__restrict__ u, const float *__restrict__ v,
                                              const float *__restrict__ w, float *__restrict__ curl) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const int cell[3] = {i, j, k};
    float d[3][3];  // d[axis][component]
    // every axis's neighbours are read before any arithmetic (their clamped indices lie in the grid), so the
    // loads are in flight together; an axis of one cell (span 0) then takes the reference's 0
    float h[3][3], l[3][3], span[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int lo[3] = {i, j, k}, hi[3] = {i, j, k};
        lo[a] = max(cell[a] - 1, 0);
        hi[a] = min(cell[a] + 1, g.dims[a] - 1);
        span[a] = (float)(hi[a] - lo[a]);
        if (CENTER_FROM_FACES) {
            center_velocity(g, u, v, w, hi[0], hi[1], hi[2], h[a][0], h[a][1], h[a][2]);
            center_velocity(g, u, v, w, lo[0], lo[1], lo[2], l[a][0], l[a][1], l[a][2]);
        } else {
            const int ch = nc.at(hi[0], hi[1], hi[2]), cl = nc.at(lo[0], lo[1], lo[2]);
            for (int q = 0; q < 3; q++) {
                h[a][q] = center[3 * ch + q];
                l[a][q] = center[3 * cl + q];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float denom = span[a] * g.cell_m;
        for (int q = 0; q < 3; q++) d[a][q] = span[a] == 0.0f ? 0.0f : (h[a][q] - l[a][q]) / denom;
    }
    const int c = nc.at(i, j, k);
    curl[3 * c] = d[1][2] - d[2][1];
    curl[3 * c + 1] = d[2][0] - d[0][2];
    curl[3 * c + 2] = d[0][1] - d[1][0];
}

__device__ __forceinline__ float spin_at(const float *__restrict__ curl, int c) {
    return length3(curl[3 * c], curl[3 * c + 1], curl[3 * c + 2]);
}

// confinement_push_kernel, sim/vorticity.py:53-71
__global__ __launch_bounds__(256) void k_confinement_push(plm_grid_info g, const float *__restrict__ curl,
                                                          float *__restrict__ push, float strength) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const int cell[3] = {i, j, k};
    const int c = nc.at(i, j, k);
    // the six neighbours' and the cell's own curl are read first (clamped indices, in the grid), so the loads
    // are in flight together; an axis of one cell (span 0) keeps the reference's 0
    const float wx = curl[3 * c], wy = curl[3 * c + 1], wz = curl[3 * c + 2];
    float spin_hi[3], spin_lo[3], span[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int lo[3] = {i, j, k}, hi[3] = {i, j, k};
        lo[a] = max(cell[a] - 1, 0);
        hi[a] = min(cell[a] + 1, g.dims[a] - 1);
        span[a] = (float)(hi[a] - lo[a]);
        spin_hi[a] = spin_at(curl, nc.at(hi[0], hi[1], hi[2]));
        spin_lo[a] = spin_at(curl, nc.at(lo[0], lo[1], lo[2]));
    }
    float toward[3];
#pragma unroll
    for (int a = 0; a < 3; a++)
        toward[a] = span[a] > 0.0f ? (spin_hi[a] - spin_lo[a]) / (span[a] * g.cell_m) : 0.0f;
    const float length = length3(toward[0], toward[1], toward[2]);
    if (length < 1.0e-12f) {
        push[3 * c] = push[3 * c + 1] = push[3 * c + 2] = 0.0f;
        return;
    }
    const float nx = toward[0] / length, ny = toward[1] / length, nz = toward[2] / length;
    const float scale = strength * g.cell_m;
    push[3 * c] = scale * (ny * wz - nz * wy);
    push[3 * c + 1] = scale * (nz * wx - nx * wz);
    push[3 * c + 2] = scale * (nx * wy - ny * wx);
}

// add_face_push_kernel, sim/vorticity.py:74-84, one thread per FACE_PUSH_ROWS consecutive cells along j: the
// thread adds to the three faces on each cell's low side (u[i,j,k], v[i,j,k], w[i,j,k]) and, on the last cell
// along an axis, to that axis's face on the high boundary. Each face's value is the face kernel's,
// 0.5 (push[below] + push[above]) dt with the clamped neighbours; a cell's push vector is read once,
// contiguously (the j neighbour is the thread's own row below), the i and k neighbours' from L2. Every load is
// issued before the first store (clamped addresses; a high face read only by the boundary cell it belongs to).
constexpr int FACE_PUSH_ROWS = 4;
__global__ __launch_bounds__(256) void k_add_face_push(plm_grid_info g, Tri<float> f,
                                                       const float *__restrict__ push, float dt) {
    constexpr int R = FACE_PUSH_ROWS;
    const Dims nc = cell_dims(g), nu = face_dims(g, 0), nv = face_dims(g, 1), nw = face_dims(g, 2);
    const int k = blockIdx.x * BX + threadIdx.x, i = blockIdx.z;
    const int j0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.y * BY + threadIdx.y) * R));
    if (k >= nc.n2 || j0 >= nc.n1) return;
    // push y of rows j0 - 1 .. j0 + R - 1 (clamped); each row's own x and z and its i and k neighbours'
    float py_row[R + 1], px[R], pz[R], px_im[R], pz_km[R], u_lo[R], u_hi[R], v_lo[R], v_hi[R], w_lo[R], w_hi[R];
#pragma unroll
    for (int m = 0; m <= R; m++) py_row[m] = push[3 * nc.at(i, min(max(j0 - 1 + m, 0), nc.n1 - 1), k) + 1];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = min(j0 + r, nc.n1 - 1), c = nc.at(i, j, k);
        const bool hx = i == nc.n0 - 1, hy = j == nc.n1 - 1, hz = k == nc.n2 - 1;
        px[r] = push[3 * c];
        pz[r] = push[3 * c + 2];
        px_im[r] = push[3 * nc.at(max(i - 1, 0), j, k)];
        pz_km[r] = push[3 * nc.at(i, j, max(k - 1, 0)) + 2];
        u_lo[r] = f.p[0][nu.at(i, j, k)];
        u_hi[r] = f.p[0][nu.at(hx ? i + 1 : i, j, k)];
        v_lo[r] = f.p[1][nv.at(i, j, k)];
        v_hi[r] = f.p[1][nv.at(i, hy ? j + 1 : j, k)];
        w_lo[r] = f.p[2][nw.at(i, j, k)];
        w_hi[r] = f.p[2][nw.at(i, j, hz ? k + 1 : k)];
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = j0 + r;
        if (j >= nc.n1) break;
        const int cell[3] = {i, j, k};
        // below = push at the clamped lower neighbour (the cell itself on the low boundary)
        const float own[3] = {px[r], py_row[r + 1], pz[r]};
        const float below[3] = {i > 0 ? px_im[r] : px[r], j > 0 ? py_row[r] : py_row[r + 1], k > 0 ? pz_km[r] : pz[r]};
        const float low_face[3] = {u_lo[r], v_lo[r], w_lo[r]}, high_face[3] = {u_hi[r], v_hi[r], w_hi[r]};
        float *low_at[3] = {f.p[0] + nu.at(i, j, k), f.p[1] + nv.at(i, j, k), f.p[2] + nw.at(i, j, k)};
        float *high_at[3] = {f.p[0] + nu.at(i + 1, j, k), f.p[1] + nv.at(i, j + 1, k), f.p[2] + nw.at(i, j, k + 1)};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            *low_at[a] = low_face[a] + 0.5f * (below[a] + own[a]) * dt;
            if (cell[a] == g.dims[a] - 1) *high_at[a] = high_face[a] + 0.5f * (own[a] + own[a]) * dt;
        }
    }
}

inline dim3 face_push_grid(const plm_grid_info &g) {
    const Dims nc = cell_dims(g);
    return dim3((nc.n2 + BX - 1) / BX, (nc.n1 + BY * FACE_PUSH_ROWS - 1) / (BY * FACE_PUSH_ROWS), nc.n0);
}

// ------------------------------------------------------------------ noise, sim/noise.py

// hash_u32, sim/noise.py:13-21 (lowbias32)
__device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

constexpr int32_t LATTICE_OFFSET = 1048576;

// lattice_value, sim/noise.py:24-32
__device__ __forceinline__ float lattice_value(int32_t seed, int32_t stream, int32_t i, int32_t j, int32_t k) {
    uint32_t h = hash_u32((uint32_t)seed + (uint32_t)LATTICE_OFFSET);
    h = hash_u32(h ^ (uint32_t)(stream + LATTICE_OFFSET));
    h = hash_u32(h ^ (uint32_t)(i + LATTICE_OFFSET));
    h = hash_u32(h ^ (uint32_t)(j + LATTICE_OFFSET));
    h = hash_u32(h ^ (uint32_t)(k + LATTICE_OFFSET));
    return (float)(h >> 8) * (2.0f / 16777215.0f) - 1.0f;
}

__device__ __forceinline__ float fade(float t) { return t * t * (3.0f - 2.0f * t); }

// value_noise, sim/noise.py:40-57. The seed's and stream's hash prefix is shared by all 8 corners.
__device__ __forceinline__ float value_noise(int32_t seed, int32_t stream, float px, float py, float pz) {
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    const int i = (int)fx, j = (int)fy, k = (int)fz;
    const float tx = fade(px - fx), ty = fade(py - fy), tz = fade(pz - fz);
    const uint32_t hs = hash_u32(hash_u32((uint32_t)seed + (uint32_t)LATTICE_OFFSET) ^ (uint32_t)(stream + LATTICE_OFFSET));
    float corner[2][2][2];
#pragma unroll
    for (int di = 0; di < 2; di++) {
        const uint32_t hi = hash_u32(hs ^ (uint32_t)(i + di + LATTICE_OFFSET));
#pragma unroll
        for (int dj = 0; dj < 2; dj++) {
            const uint32_t hj = hash_u32(hi ^ (uint32_t)(j + dj + LATTICE_OFFSET));
#pragma unroll
            for (int dk = 0; dk < 2; dk++) {
                const uint32_t h = hash_u32(hj ^ (uint32_t)(k + dk + LATTICE_OFFSET));
                corner[di][dj][dk] = (float)(h >> 8) * (2.0f / 16777215.0f) - 1.0f;
            }
        }
    }
    const float c00 = lerp_(corner[0][0][0], corner[1][0][0], tx);
    const float c10 = lerp_(corner[0][1][0], corner[1][1][0], tx);
    const float c01 = lerp_(corner[0][0][1], corner[1][0][1], tx);
    const float c11 = lerp_(corner[0][1][1], corner[1][1][1], tx);
    return lerp_(lerp_(c00, c10, ty), lerp_(c01, c11, ty), tz);
}

__global__ void k_noise_probe(int32_t seed, int32_t stream, const float *__restrict__ points,
                              const int32_t *__restrict__ lattice, int count, float *__restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    if (points != nullptr)
        out[t] = value_noise(seed, stream, points[3 * t], points[3 * t + 1], points[3 * t + 2]);
    else
        out[t] = lattice_value(seed, stream, lattice[3 * t], lattice[3 * t + 1], lattice[3 * t + 2]);
}

// ------------------------------------------------------------------ emitters, sim/emitters.py

// soft_edge, sim/emitters.py:41-50
__device__ __forceinline__ float soft_edge(float distance, float extent, float softness) {
    const float inner = extent * (1.0f - softness);
    if (distance <= inner) return 1.0f;
    if (distance >= extent) return 0.0f;
    const float t = (distance - inner) / (extent - inner);
    return 1.0f - t * t * (3.0f - 2.0f * t);
}

// coverage, sim/emitters.py:53-67
__device__ __forceinline__ float coverage(const plm_emitter &e, float x, float y, float z, int32_t seed) {
    const float ox = x - e.center_m[0], oy = y - e.center_m[1], oz = z - e.center_m[2];
    float amount;
    if (e.shape == 0)
        amount = soft_edge(length3(ox, oy, oz), e.radius_m, e.edge_softness);
    else
        amount = soft_edge(fabsf(ox), e.half_size_m[0], e.edge_softness) *
                 soft_edge(fabsf(oy), e.half_size_m[1], e.edge_softness) *
                 soft_edge(fabsf(oz), e.half_size_m[2], e.edge_softness);
    if (amount > 0.0f && e.noise_strength > 0.0f) {
        const float breakup = 1.0f + e.noise_strength * value_noise(seed, e.stream, x / e.noise_scale_m,
                                                                    y / e.noise_scale_m, z / e.noise_scale_m);
        amount = amount * fmax_(breakup, 0.0f);
    }
    return fmin_(amount, 1.0f);
}

// blend_factor, sim/emitters.py:70-77 (BLEND_REFERENCE_S = 1/80 s, :20)
__device__ __forceinline__ float blend_factor(float amount, float dt) {
    return 1.0f - powf(1.0f - amount, dt / 0.0125f);
}

__device__ __forceinline__ bool is_active(const plm_emitter &e, float time_s) {
    return time_s >= e.start_s && time_s < e.end_s;
}

// emit_scalars_kernel, sim/emitters.py:90-102, at cell (i, j, k)
__device__ __forceinline__ void emit_cell(const plm_grid_info &g, float *__restrict__ smoke,
                                          float *__restrict__ temperature, const plm_emitter *__restrict__ emitters,
                                          int n_emitters, int32_t seed, float time_s, float dt, int i, int j, int k) {
    const Dims nc = cell_dims(g);
    const float x = g.origin_m[0] + ((float)i + 0.5f) * g.cell_m;
    const float y = g.origin_m[1] + ((float)j + 0.5f) * g.cell_m;
    const float z = g.origin_m[2] + ((float)k + 0.5f) * g.cell_m;
    const int c = nc.at(i, j, k);
    for (int n = 0; n < n_emitters; n++) {
        const plm_emitter e = emitters[n];
        if (!is_active(e, time_s)) continue;
        const float amount = coverage(e, x, y, z, seed);
        if (amount > 0.0f) {
            smoke[c] = smoke[c] + e.smoke_per_s * amount * dt;
            temperature[c] = lerp_(temperature[c], e.temperature_k, blend_factor(amount, dt));
        }
    }
}

__global__ __launch_bounds__(256) void k_emit_scalars(plm_grid_info g, float *__restrict__ smoke,
                                                      float *__restrict__ temperature,
                                                      const plm_emitter *__restrict__ emitters, int n_emitters,
                                                      int32_t seed, float time_s, float dt) {
    int i, j, k;
    if (!thread_cell(cell_dims(g), i, j, k)) return;
    emit_cell(g, smoke, temperature, emitters, n_emitters, seed, time_s, dt, i, j, k);
}

// emit_velocity_kernel, sim/emitters.py:105-121, at face (i, j, k) of face array `axis`
__device__ __forceinline__ void emit_face(const plm_grid_info &g, float *__restrict__ field, int axis,
                                          const plm_emitter *__restrict__ emitters, int n_emitters, int32_t seed,
                                          float time_s, float dt, int i, int j, int k) {
    const Dims d = face_dims(g, axis);
    const float x = g.origin_m[0] + ((float)i + (axis == 0 ? 0.0f : 0.5f)) * g.cell_m;
    const float y = g.origin_m[1] + ((float)j + (axis == 1 ? 0.0f : 0.5f)) * g.cell_m;
    const float z = g.origin_m[2] + ((float)k + (axis == 2 ? 0.0f : 0.5f)) * g.cell_m;
    float *p = field + d.at(i, j, k);
    for (int n = 0; n < n_emitters; n++) {
        const plm_emitter e = emitters[n];
        if (!is_active(e, time_s)) continue;
        const float amount = coverage(e, x, y, z, seed);
        if (amount > 0.0f) {
            float target[3] = {e.velocity_mps[0], e.velocity_mps[1], e.velocity_mps[2]};
            if (e.outward == 1) {
                const float ax = x - e.center_m[0], ay = y - e.center_m[1], az = z - e.center_m[2];
                const float la = length3(ax, ay, az);
                if (la > 0.0f) {
                    const float speed = length3(e.velocity_mps[0], e.velocity_mps[1], e.velocity_mps[2]);
                    target[0] = (ax / la) * speed;
                    target[1] = (ay / la) * speed;
                    target[2] = (az / la) * speed;
                }
            }
            *p = lerp_(*p, target[axis], blend_factor(amount, dt));
        }
    }
}

// emit_velocity_kernel on all three face arrays (the packed face launch)
__global__ __launch_bounds__(256) void k_emit_velocity(plm_grid_info g, FaceLaunch fl, Tri<float> f,
                                                       const plm_emitter *__restrict__ emitters, int n_emitters,
                                                       int32_t seed, float time_s, float dt) {
    int axis, i, j, k;
    if (!packed_face(g, fl, axis, i, j, k)) return;
    emit_face(g, f.p[axis], axis, emitters, n_emitters, seed, time_s, dt, i, j, k);
}

// Both emit kernels over a box of cells [lo, hi) and the faces of those cells (a face array's index runs to hi on
// its own axis): one thread per point of the box (hi - lo + 1 per axis), the cell and up to three faces.
struct EmitBox {
    int lo[3], hi[3];
};

__global__ __launch_bounds__(256) void k_emit_box(plm_grid_info g, EmitBox box, float *__restrict__ smoke,
                                                  float *__restrict__ temperature, Tri<float> f,
                                                  const plm_emitter *__restrict__ emitters, int n_emitters,
                                                  int32_t seed, float time_s, float dt) {
    const Dims span{box.hi[0] - box.lo[0] + 1, box.hi[1] - box.lo[1] + 1, box.hi[2] - box.lo[2] + 1};
    int a, b, c;
    if (!thread_cell(span, a, b, c)) return;
    const int i = box.lo[0] + a, j = box.lo[1] + b, k = box.lo[2] + c;
    const bool in_i = i < box.hi[0], in_j = j < box.hi[1], in_k = k < box.hi[2];
    if (in_i && in_j && in_k) emit_cell(g, smoke, temperature, emitters, n_emitters, seed, time_s, dt, i, j, k);
    if (in_j && in_k) emit_face(g, f.p[0], 0, emitters, n_emitters, seed, time_s, dt, i, j, k);
    if (in_i && in_k) emit_face(g, f.p[1], 1, emitters, n_emitters, seed, time_s, dt, i, j, k);
    if (in_i && in_j) emit_face(g, f.p[2], 2, emitters, n_emitters, seed, time_s, dt, i, j, k);
}

// ------------------------------------------------------------------ turbulence, sim/turbulence.py

// potential, sim/turbulence.py:26-32 (TURBULENCE_STREAM = 1 << 24, :23)
__device__ __forceinline__ float potential(int32_t seed, int component, float px, float py, float pz,
                                           float pattern_time) {
    const float slice_lo = floorf(pattern_time);
    const float blend = fade(pattern_time - slice_lo);
    const int32_t first = (1 << 24) + 3 * (int32_t)slice_lo + component;
    return lerp_(value_noise(seed, first, px, py, pz), value_noise(seed, first + 3, px, py, pz), blend);
}

// The lattice value of a full corner hash (lattice_value's last step)
__device__ __forceinline__ float corner_of(uint32_t h) { return (float)(h >> 8) * (2.0f / 16777215.0f) - 1.0f; }

// value_noise's trilinear blend of its 8 corner values, in its order
__device__ __forceinline__ float blend8(const float (&c)[2][2][2], float tx, float ty, float tz) {
    const float c00 = lerp_(c[0][0][0], c[1][0][0], tx);
    const float c10 = lerp_(c[0][1][0], c[1][1][0], tx);
    const float c01 = lerp_(c[0][0][1], c[1][0][1], tx);
    const float c11 = lerp_(c[0][1][1], c[1][1][1], tx);
    return lerp_(lerp_(c00, c10, ty), lerp_(c01, c11, ty), tz);
}

// value_noise of one stream (its seed-and-stream hash prefix `hs`) at two points lo and hi that are equal except
// along axis A, where 0 <= hi - lo < 1 lattice unit: both points' corners lie in one 3 x 2 x 2 block of lattice
// points (3 along A), which is hashed once (12 corner hashes instead of 16, and shared outer hashes). The
// corner values and the blend are value_noise's, so both results are bit-identical to two value_noise calls.
template <int A>
__device__ __forceinline__ void value_noise_pair(uint32_t hs, const float (&lo)[3], const float (&hi)[3],
                                                 float &v_lo, float &v_hi) {
    const float f[3] = {floorf(lo[0]), floorf(lo[1]), floorf(lo[2])};
    const float fa = floorf(hi[A]);
    const int base[3] = {(int)f[0], (int)f[1], (int)f[2]};
    const bool shift = (int)fa != base[A];  // hi's cell is lo's, or the next along A
    const float tl[3] = {fade(lo[0] - f[0]), fade(lo[1] - f[1]), fade(lo[2] - f[2])};
    float th[3] = {tl[0], tl[1], tl[2]};
    th[A] = fade(hi[A] - fa);
    constexpr int E0 = A == 0 ? 3 : 2, E1 = A == 1 ? 3 : 2, E2 = A == 2 ? 3 : 2;
    float lattice[E0][E1][E2];
#pragma unroll
    for (int di = 0; di < E0; di++) {
        const uint32_t h0 = hash_u32(hs ^ (uint32_t)(base[0] + di + LATTICE_OFFSET));
#pragma unroll
        for (int dj = 0; dj < E1; dj++) {
            const uint32_t h1 = hash_u32(h0 ^ (uint32_t)(base[1] + dj + LATTICE_OFFSET));
#pragma unroll
            for (int dk = 0; dk < E2; dk++)
                lattice[di][dj][dk] = corner_of(hash_u32(h1 ^ (uint32_t)(base[2] + dk + LATTICE_OFFSET)));
        }
    }
    float c_lo[2][2][2], c_hi[2][2][2];
#pragma unroll
    for (int di = 0; di < 2; di++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++)
#pragma unroll
            for (int dk = 0; dk < 2; dk++) {
                c_lo[di][dj][dk] = lattice[di][dj][dk];
                const float same = lattice[di][dj][dk];
                const float next = lattice[di + (A == 0)][dj + (A == 1)][dk + (A == 2)];
                c_hi[di][dj][dk] = shift ? next : same;
            }
    v_lo = blend8(c_lo, tl[0], tl[1], tl[2]);
    v_hi = blend8(c_hi, th[0], th[1], th[2]);
}

__device__ __forceinline__ uint32_t stream_prefix(int32_t seed, int32_t stream) {
    return hash_u32(hash_u32((uint32_t)seed + (uint32_t)LATTICE_OFFSET) ^ (uint32_t)(stream + LATTICE_OFFSET));
}

// potential() of component `comp` at the two points face - h and face + h (h = 0.5 cells along axis A), both
// time slices through value_noise_pair
template <int A>
__device__ __forceinline__ void potential_pair(int32_t seed, int comp, const float (&face)[3], float to_lattice,
                                               float pattern_time, float &minus, float &plus) {
    float lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float h = a == A ? 0.5f : 0.0f;
        lo[a] = (face[a] - h) * to_lattice;
        hi[a] = (face[a] + h) * to_lattice;
    }
    const float slice_lo = floorf(pattern_time);
    const float blend = fade(pattern_time - slice_lo);
    const int32_t first = (1 << 24) + 3 * (int32_t)slice_lo + comp;
    float lo0, hi0, lo1, hi1;
    value_noise_pair<A>(stream_prefix(seed, first), lo, hi, lo0, hi0);
    value_noise_pair<A>(stream_prefix(seed, first + 3), lo, hi, lo1, hi1);
    minus = lerp_(lo0, lo1, blend);
    plus = lerp_(hi0, hi1, blend);
}

// turbulence_kernel on one face array (AXIS) with the paired noise: b = AXIS + 1, c = AXIS + 2 (mod 3)
template <int AXIS>
__device__ __forceinline__ void turbulence_paired(const plm_grid_info &g, float *field, const Dims &d, int i, int j,
                                                  int k, int32_t seed, float to_lattice, float strength,
                                                  float pattern_time, float dt) {
    constexpr int B = (AXIS + 1) % 3, C = (AXIS + 2) % 3;
    const float face[3] = {(float)i + (AXIS == 0 ? 0.0f : 0.5f), (float)j + (AXIS == 1 ? 0.0f : 0.5f),
                           (float)k + (AXIS == 2 ? 0.0f : 0.5f)};
    float c_minus, c_plus, b_minus, b_plus;
    potential_pair<B>(seed, C, face, to_lattice, pattern_time, c_minus, c_plus);
    potential_pair<C>(seed, B, face, to_lattice, pattern_time, b_minus, b_plus);
    const float curl = ((c_plus - c_minus) - (b_plus - b_minus)) / to_lattice;
    float *p = field + d.at(i, j, k);
    *p = *p + strength * curl * dt;
}

// turbulence_kernel, sim/turbulence.py:35-53, on all three face arrays. PAIRED (the host's choice when half a
// cell is well under half a lattice unit, cell_m / scale_m < 0.99) evaluates each potential's two points with
// one shared lattice block; otherwise every point's noise is evaluated on its own. Both are bit-identical.
template <bool PAIRED>
__global__ __launch_bounds__(256) void k_turbulence(plm_grid_info g, FaceLaunch fl, Tri<float> f, int32_t seed, float scale_m,
                                                    float strength, float pattern_time, float dt) {
    int axis, i, j, k;
    if (!packed_face(g, fl, axis, i, j, k)) return;
    const Dims d = face_dims(g, axis);
    const float to_lattice = g.cell_m / scale_m;
    if constexpr (PAIRED) {
        if (axis == 0) turbulence_paired<0>(g, f.p[0], d, i, j, k, seed, to_lattice, strength, pattern_time, dt);
        else if (axis == 1) turbulence_paired<1>(g, f.p[1], d, i, j, k, seed, to_lattice, strength, pattern_time, dt);
        else turbulence_paired<2>(g, f.p[2], d, i, j, k, seed, to_lattice, strength, pattern_time, dt);
        return;
    }
    const float face[3] = {(float)i + (axis == 0 ? 0.0f : 0.5f), (float)j + (axis == 1 ? 0.0f : 0.5f),
                           (float)k + (axis == 2 ? 0.0f : 0.5f)};
    const int b = (axis + 1) % 3, c = (axis + 2) % 3;
    float hb[3] = {0.0f, 0.0f, 0.0f}, hc[3] = {0.0f, 0.0f, 0.0f};
    hb[b] = 0.5f;
    hc[c] = 0.5f;
    const float c_plus = potential(seed, c, (face[0] + hb[0]) * to_lattice, (face[1] + hb[1]) * to_lattice,
                                   (face[2] + hb[2]) * to_lattice, pattern_time);
    const float c_minus = potential(seed, c, (face[0] - hb[0]) * to_lattice, (face[1] - hb[1]) * to_lattice,
                                    (face[2] - hb[2]) * to_lattice, pattern_time);
    const float b_plus = potential(seed, b, (face[0] + hc[0]) * to_lattice, (face[1] + hc[1]) * to_lattice,
                                   (face[2] + hc[2]) * to_lattice, pattern_time);
    const float b_minus = potential(seed, b, (face[0] - hc[0]) * to_lattice, (face[1] - hc[1]) * to_lattice,
                                    (face[2] - hc[2]) * to_lattice, pattern_time);
    const float curl = ((c_plus - c_minus) - (b_plus - b_minus)) / to_lattice;
    float *p = f.p[axis] + d.at(i, j, k);
    *p = *p + strength * curl * dt;
}

// ------------------------------------------------------------------ elementwise

// dissipate_kernel, sim/dissipate.py:10-15
__global__ __launch_bounds__(256) void k_dissipate(int64_t count, float *__restrict__ smoke,
                                                   float *__restrict__ temperature, float smoke_keep,
                                                   float heat_keep, float ambient_k) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x) {
        smoke[c] = smoke[c] * smoke_keep;
        temperature[c] = ambient_k + (temperature[c] - ambient_k) * heat_keep;
    }
}

inline dim3 grid1(int64_t count) {
    int64_t blocks = (count + 255) / 256;
    return dim3((unsigned)(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536));
}

}  // namespace
}  // namespace plm

using namespace plm;

extern "C" {

int plm_buoyancy(const plm_grid_info *info, float *v, const float *smoke, const float *temperature, float gravity,
                 float heat_buoyancy, float smoke_weight, float ambient_k, float dt, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(v && smoke && temperature, "plm_buoyancy: NULL array");
    hipStream_t s = as_stream(stream);
    TimingScope t("forces", s);
    k_buoyancy<<<buoyancy_grid(*info), block3(), 0, s>>>(*info, v, smoke, temperature, gravity, heat_buoyancy,
                                                              smoke_weight, ambient_k, dt);
    return check_launch("plm_buoyancy");
}

int plm_drag(const plm_grid_info *info, float *u, float *v, float *w, const float wind[3], float keep,
             void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && wind, "plm_drag: NULL argument");
    hipStream_t s = as_stream(stream);
    const FaceLaunch fl = face_launch(*info);
    k_drag<<<fl.block_end[2], FACE_BLOCK, 0, s>>>(*info, fl, Tri<float>{{u, v, w}}, wind[0], wind[1], wind[2], keep);
    return check_launch("plm_drag");
}

int plm_center_velocity(const plm_grid_info *info, const float *u, const float *v, const float *w, float *center,
                        void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && center, "plm_center_velocity: NULL array");
    k_center_velocity<<<grid3(cell_dims(*info)), block3(), 0, as_stream(stream)>>>(*info, u, v, w, center);
    return check_launch("plm_center_velocity");
}

int plm_curl(const plm_grid_info *info, const float *center, float *curl, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(center && curl, "plm_curl: NULL array");
    k_curl<false><<<grid3(cell_dims(*info)), block3(), 0, as_stream(stream)>>>(*info, center, nullptr, nullptr,
                                                                              nullptr, curl);
    return check_launch("plm_curl");
}

int plm_confinement_push(const plm_grid_info *info, const float *curl, float *push, float strength,
                         void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(curl && push, "plm_confinement_push: NULL array");
    k_confinement_push<<<grid3(cell_dims(*info)), block3(), 0, as_stream(stream)>>>(*info, curl, push, strength);
    return check_launch("plm_confinement_push");
}

int plm_add_face_push(const plm_grid_info *info, float *u, float *v, float *w, const float *push, float dt,
                      void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && push, "plm_add_face_push: NULL array");
    k_add_face_push<<<face_push_grid(*info), block3(), 0, as_stream(stream)>>>(*info, Tri<float>{{u, v, w}}, push, dt);
    return check_launch("plm_add_face_push");
}

int plm_confinement(const plm_grid_info *info, float *u, float *v, float *w, float *center, float *curl,
                    float *push, float strength, float dt, int32_t fuse, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && curl && push, "plm_confinement: NULL array");
    PLM_REQUIRE(fuse || center, "plm_confinement: center is NULL");
    if (strength == 0.0f) return 0;
    hipStream_t s = as_stream(stream);
    const Dims nc = cell_dims(*info);
    TimingScope t("confinement", s);
    if (fuse) {
        k_curl<true><<<grid3(nc), block3(), 0, s>>>(*info, nullptr, u, v, w, curl);
    } else {
        k_center_velocity<<<grid3(nc), block3(), 0, s>>>(*info, u, v, w, center);
        PLM_TRY(check_launch("plm_confinement center"));
        k_curl<false><<<grid3(nc), block3(), 0, s>>>(*info, center, nullptr, nullptr, nullptr, curl);
    }
    PLM_TRY(check_launch("plm_confinement curl"));
    k_confinement_push<<<grid3(nc), block3(), 0, s>>>(*info, curl, push, strength);
    PLM_TRY(check_launch("plm_confinement push"));
    k_add_face_push<<<face_push_grid(*info), block3(), 0, s>>>(*info, Tri<float>{{u, v, w}}, push, dt);
    return check_launch("plm_confinement faces");
}

int plm_emit(const plm_grid_info *info, float *smoke, float *temperature, float *u, float *v, float *w,
             const plm_emitter *emitters, int32_t n_emitters, int32_t seed, float time_s, float dt, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(smoke && temperature && u && v && w, "plm_emit: NULL array");
    PLM_REQUIRE(n_emitters >= 0 && (n_emitters == 0 || emitters), "plm_emit: bad emitter array");
    if (n_emitters == 0) return 0;
    hipStream_t s = as_stream(stream);
    TimingScope t("emit", s);
    k_emit_scalars<<<grid3(cell_dims(*info)), block3(), 0, s>>>(*info, smoke, temperature, emitters, n_emitters,
                                                               seed, time_s, dt);
    PLM_TRY(check_launch("plm_emit scalars"));
    const FaceLaunch fl = face_launch(*info);
    k_emit_velocity<<<fl.block_end[2], FACE_BLOCK, 0, s>>>(*info, fl, Tri<float>{{u, v, w}}, emitters, n_emitters, seed,
                                                          time_s, dt);
    return check_launch("plm_emit velocity");
}

int plm_emit_box(const plm_grid_info *info, float *smoke, float *temperature, float *u, float *v, float *w,
                 const plm_emitter *emitters, int32_t n_emitters, int32_t seed, float time_s, float dt,
                 const int32_t lo[3], const int32_t hi[3], void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(smoke && temperature && u && v && w && lo && hi, "plm_emit_box: NULL array");
    PLM_REQUIRE(n_emitters >= 0 && (n_emitters == 0 || emitters), "plm_emit_box: bad emitter array");
    EmitBox box{};
    for (int a = 0; a < 3; a++) {
        PLM_REQUIRE(0 <= lo[a] && lo[a] <= hi[a] && hi[a] <= info->dims[a],
                    "plm_emit_box: box [%d, %d) outside the grid's %d cells on axis %d", lo[a], hi[a], info->dims[a], a);
        box.lo[a] = lo[a];
        box.hi[a] = hi[a];
    }
    if (n_emitters == 0) return 0;
    hipStream_t s = as_stream(stream);
    TimingScope t("emit", s);
    const Dims span{hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    k_emit_box<<<grid3(span), block3(), 0, s>>>(*info, box, smoke, temperature, Tri<float>{{u, v, w}}, emitters,
                                                 n_emitters, seed, time_s, dt);
    return check_launch("plm_emit_box");
}

int plm_noise_probe(int32_t seed, int32_t stream_id, const float *points, const int32_t *lattice, int32_t count,
                    float *out, void *stream) {
    PLM_REQUIRE(count >= 0 && out && (points || lattice), "plm_noise_probe: bad arguments");
    if (count == 0) return 0;
    k_noise_probe<<<(count + 255) / 256, 256, 0, as_stream(stream)>>>(seed, stream_id, points, lattice, count, out);
    return check_launch("plm_noise_probe");
}

int plm_turbulence(const plm_grid_info *info, float *u, float *v, float *w, int32_t seed, float scale_m,
                   float strength, float pattern_time, float dt, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w, "plm_turbulence: NULL array");
    PLM_REQUIRE(scale_m > 0.0f, "plm_turbulence: scale_m must be positive");
    if (strength == 0.0f) return 0;
    hipStream_t s = as_stream(stream);
    TimingScope t("turbulence", s);
    const FaceLaunch fl = face_launch(*info);
    if (info->cell_m / scale_m < 0.99f)
        k_turbulence<true><<<fl.block_end[2], FACE_BLOCK, 0, s>>>(*info, fl, Tri<float>{{u, v, w}}, seed, scale_m, strength,
                                                       pattern_time, dt);
    else
        k_turbulence<false><<<fl.block_end[2], FACE_BLOCK, 0, s>>>(*info, fl, Tri<float>{{u, v, w}}, seed, scale_m, strength,
                                                       pattern_time, dt);
    return check_launch("plm_turbulence");
}

int plm_dissipate(int64_t count, float *smoke, float *temperature, float smoke_keep, float heat_keep,
                  float ambient_k, void *stream) {
    PLM_REQUIRE(count >= 0 && smoke && temperature, "plm_dissipate: bad arguments");
    if (count == 0) return 0;
    k_dissipate<<<grid1(count), 256, 0, as_stream(stream)>>>(count, smoke, temperature, smoke_keep, heat_keep,
                                                            ambient_k);
    return check_launch("plm_dissipate");
}

}  // extern "C"
