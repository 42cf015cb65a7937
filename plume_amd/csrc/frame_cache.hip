// The viewer's use of the path: its frame cache (view/frame_cache.py) keeps each frame's smoke, temperature
// and cell-centred velocity as IEEE binary16, and its live run (view/live_run.py) reads the smoke's six
// boundary layers. All four kernels are streaming copies with a conversion, bound by HBM bandwidth.
//   - k_narrow_f16 / k_widen_f16: 8 values per thread in two contiguous passes of the block, each a 16-byte
//     float access and an 8-byte half access; rounding to binary16 is round-to-nearest-even (v_cvt_f16_f32
//     under the default mode), subnormal halves kept. A vec3h array widens as a flat array of 3N halves.
//   - k_narrow_center_velocity: center_velocity_kernel (sim/vorticity.py:17-22) and narrow_vector_kernel
//     (view/frame_cache.py:38-42) fused: the FP32 centre 0.5 (a + b) per axis, rounded once to binary16 and
//     stored as 3 halves per cell (the vec3h layout), with no FP32 vec3 array in between.
//   - k_boundary_layers: boundary_layer_kernel (view/live_run.py:41-57).
#include "common.h"

namespace plm {
namespace {

constexpr int CACHE_BLOCK = 256;
constexpr int PER_THREAD = 8;  // values per thread: two 16-byte float accesses

__device__ __forceinline__ uint16_t half_bits(float x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
__device__ __forceinline__ float half_value(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// A block converts PER_THREAD * CACHE_BLOCK values; its q-th pass takes the q-th run of 4 * CACHE_BLOCK of them,
// 4 per thread, so each load and store of a wavefront is contiguous, and a thread's passes are in flight together.
__global__ __launch_bounds__(CACHE_BLOCK) void k_narrow_f16(int64_t count, const float *__restrict__ src,
                                                            uint16_t *__restrict__ dst) {
    const int64_t base = (int64_t)blockIdx.x * CACHE_BLOCK * PER_THREAD + 4 * (int64_t)threadIdx.x;
    float4 v[PER_THREAD / 4];
#pragma unroll
    for (int q = 0; q < PER_THREAD / 4; q++) {
        const int64_t e = base + (int64_t)q * 4 * CACHE_BLOCK;
        v[q] = e + 4 <= count ? *reinterpret_cast<const float4 *>(src + e) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int q = 0; q < PER_THREAD / 4; q++) {
        const int64_t e = base + (int64_t)q * 4 * CACHE_BLOCK;
        if (e + 4 <= count) {
            const uint32_t lo = (uint32_t)half_bits(v[q].x) | ((uint32_t)half_bits(v[q].y) << 16);
            const uint32_t hi = (uint32_t)half_bits(v[q].z) | ((uint32_t)half_bits(v[q].w) << 16);
            *reinterpret_cast<uint2 *>(dst + e) = make_uint2(lo, hi);
        } else {
            for (int64_t x = e; x < count && x < e + 4; x++) dst[x] = half_bits(src[x]);
        }
    }
}

__global__ __launch_bounds__(CACHE_BLOCK) void k_widen_f16(int64_t count, const uint16_t *__restrict__ src,
                                                           float *__restrict__ dst) {
    const int64_t base = (int64_t)blockIdx.x * CACHE_BLOCK * PER_THREAD + 4 * (int64_t)threadIdx.x;
    uint2 h[PER_THREAD / 4];
#pragma unroll
    for (int q = 0; q < PER_THREAD / 4; q++) {
        const int64_t e = base + (int64_t)q * 4 * CACHE_BLOCK;
        h[q] = e + 4 <= count ? *reinterpret_cast<const uint2 *>(src + e) : make_uint2(0u, 0u);
    }
#pragma unroll
    for (int q = 0; q < PER_THREAD / 4; q++) {
        const int64_t e = base + (int64_t)q * 4 * CACHE_BLOCK;
        if (e + 4 <= count) {
            *reinterpret_cast<float4 *>(dst + e) =
                make_float4(half_value((uint16_t)(h[q].x & 0xffffu)), half_value((uint16_t)(h[q].x >> 16)),
                            half_value((uint16_t)(h[q].y & 0xffffu)), half_value((uint16_t)(h[q].y >> 16)));
        } else {
            for (int64_t x = e; x < count && x < e + 4; x++) dst[x] = half_value(src[x]);
        }
    }
}

__global__ __launch_bounds__(BX * BY) void k_narrow_center_velocity(plm_grid_info g, const float *__restrict__ u,
                                                                    const float *__restrict__ v,
                                                                    const float *__restrict__ w,
                                                                    uint16_t *__restrict__ center) {
    const Dims n = cell_dims(g);
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    const Dims du = face_dims(g, 0), dv = face_dims(g, 1), dw = face_dims(g, 2);
    const float cu = 0.5f * (u[du.at(i, j, k)] + u[du.at(i + 1, j, k)]);
    const float cv = 0.5f * (v[dv.at(i, j, k)] + v[dv.at(i, j + 1, k)]);
    const float cw = 0.5f * (w[dw.at(i, j, k)] + w[dw.at(i, j, k + 1)]);
    uint16_t *out = center + 3 * (int64_t)n.at(i, j, k);
    out[0] = half_bits(cu);
    out[1] = half_bits(cv);
    out[2] = half_bits(cw);
}

// layers[face, a, b] for face = 2 axis + side: the cell at index 0 (side 0) or n_axis - 1 (side 1) along
// `axis`, a and b along the other two axes in x, y, z order; 0 where (a, b) lies past that face.
__global__ __launch_bounds__(BX * BY) void k_boundary_layers(Dims n, const float *__restrict__ smoke,
                                                             float *__restrict__ layers, int extent_a,
                                                             int extent_b) {
    const int b = blockIdx.x * BX + threadIdx.x, a = blockIdx.y * BY + threadIdx.y, face = blockIdx.z;
    if (b >= extent_b || a >= extent_a) return;
    const int axis = face >> 1, side = face & 1;
    // (i, j, k) of (a, b) on the face, and the extents of a and b there
    const int i = axis == 0 ? side * (n.n0 - 1) : a;
    const int j = axis == 1 ? side * (n.n1 - 1) : axis == 0 ? a : b;
    const int k = axis == 2 ? side * (n.n2 - 1) : b;
    const int size_a = axis == 0 ? n.n1 : n.n0, size_b = axis == 2 ? n.n1 : n.n2;
    const float value = a < size_a && b < size_b ? smoke[n.at(i, j, k)] : 0.0f;
    layers[((int64_t)face * extent_a + a) * extent_b + b] = value;
}

inline dim3 grid_quads(int64_t count) {
    return dim3((unsigned)((count + PER_THREAD * CACHE_BLOCK - 1) / (PER_THREAD * CACHE_BLOCK)));
}

bool aligned(const void *p, uintptr_t bytes) { return (reinterpret_cast<uintptr_t>(p) & (bytes - 1)) == 0; }

}  // namespace
}  // namespace plm

using namespace plm;

extern "C" {

int plm_narrow_f16(int64_t count, const float *src, void *dst, void *stream) {
    PLM_REQUIRE(count >= 0 && (count == 0 || (src && dst)), "plm_narrow_f16: bad arguments");
    PLM_REQUIRE(aligned(src, 16) && aligned(dst, 8), "plm_narrow_f16: src must be 16-byte and dst 8-byte aligned");
    if (count == 0) return 0;
    const TimingScope t("narrow_f16", as_stream(stream), true);
    launch_timed(t, k_narrow_f16, grid_quads(count), dim3(CACHE_BLOCK), as_stream(stream), count, src,
                 static_cast<uint16_t *>(dst));
    return check_launch("plm_narrow_f16");
}

int plm_widen_f16(int64_t count, const void *src, float *dst, void *stream) {
    PLM_REQUIRE(count >= 0 && (count == 0 || (src && dst)), "plm_widen_f16: bad arguments");
    PLM_REQUIRE(aligned(src, 8) && aligned(dst, 16), "plm_widen_f16: src must be 8-byte and dst 16-byte aligned");
    if (count == 0) return 0;
    const TimingScope t("widen_f16", as_stream(stream), true);
    launch_timed(t, k_widen_f16, grid_quads(count), dim3(CACHE_BLOCK), as_stream(stream), count,
                 static_cast<const uint16_t *>(src), dst);
    return check_launch("plm_widen_f16");
}

int plm_narrow_center_velocity(const plm_grid_info *info, const float *u, const float *v, const float *w,
                               void *center, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && center, "plm_narrow_center_velocity: NULL array");
    const TimingScope t("narrow_center_velocity", as_stream(stream), true);
    launch_timed(t, k_narrow_center_velocity, grid3(cell_dims(*info)), block3(), as_stream(stream), *info, u, v, w,
                 static_cast<uint16_t *>(center));
    return check_launch("plm_narrow_center_velocity");
}

int plm_boundary_layers(const int32_t shape[3], const float *smoke, float *layers, int32_t extent_a,
                        int32_t extent_b, void *stream) {
    PLM_TRY(validate_dims(shape, "plm_boundary_layers"));
    PLM_REQUIRE(smoke && layers, "plm_boundary_layers: NULL array");
    PLM_REQUIRE(extent_a > 0 && extent_b > 0 && extent_a <= 65535 * BY, "plm_boundary_layers: bad layer extents");
    const dim3 grid((unsigned)((extent_b + BX - 1) / BX), (unsigned)((extent_a + BY - 1) / BY), 6);
    k_boundary_layers<<<grid, block3(), 0, as_stream(stream)>>>(dims_of(shape), smoke, layers, extent_a, extent_b);
    return check_launch("plm_boundary_layers");
}

}  // extern "C"
