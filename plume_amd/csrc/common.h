// Shared device helpers and C-ABI plumbing for libplume_hip (gfx950).
//
// Arithmetic follows the reference's Warp kernels operation by operation (the library is built with
// -ffp-contract=off and IEEE division/sqrt), so results match the reference's float32 CPU path bit for
// bit wherever no transcendental function is involved.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/plume_hip.h"

namespace plm {

// ------------------------------------------------------------------ errors and launch plumbing

int fail(int code, const char *fmt, ...);
int check_launch(const char *what);

// Event timing of one launch family, for benchmarks. A plain scope records an event pair around what it
// covers. A `bound` scope records nothing itself: the single kernel it covers is launched with
// hipExtLaunchKernelGGL(..., start(), stop(), ...), so the pair carries that dispatch's own begin and end
// timestamps and leaves out the gap before the dispatch starts. start() and stop() are null when the scope is
// inactive, and the launch then takes the ordinary path.
struct TimingScope {
    TimingScope(const char *family, hipStream_t stream, bool bound = false);
    ~TimingScope();
    hipEvent_t start() const;
    hipEvent_t stop() const;
    bool active;
    bool bound;
    hipStream_t stream;
    int slot;
};

// Launches `kernel` with the bound event pair of `timing` when it is active (hipExtLaunchKernelGGL), else as
// an ordinary launch.
template <typename Kernel, typename... Args>
void launch_timed(const TimingScope &timing, Kernel kernel, dim3 grid, dim3 block, hipStream_t stream, Args... args) {
    const hipEvent_t e0 = timing.start(), e1 = timing.stop();
    if (e0 != nullptr)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, e0, e1, 0u, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, stream, args...);
}

#define PLM_REQUIRE(cond, ...)                                                                     \
    do {                                                                                           \
        if (!(cond)) return ::plm::fail(PLM_EINVAL, __VA_ARGS__);                                  \
    } while (0)

#define PLM_TRY(expr)                                                                              \
    do {                                                                                           \
        int plm_rc_ = (expr);                                                                      \
        if (plm_rc_ != 0) return plm_rc_;                                                          \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

struct Dims {
    int n0, n1, n2;
    __host__ __device__ int count() const { return n0 * n1 * n2; }
    __host__ __device__ int at(int i, int j, int k) const { return (i * n1 + j) * n2 + k; }
};

__host__ __device__ inline Dims dims_of(const int32_t *s) { return Dims{s[0], s[1], s[2]}; }
__host__ __device__ inline Dims cell_dims(const plm_grid_info &g) { return Dims{g.dims[0], g.dims[1], g.dims[2]}; }
__host__ __device__ inline Dims face_dims(const plm_grid_info &g, int axis) {
    Dims d = cell_dims(g);
    if (axis == 0) d.n0 += 1;
    if (axis == 1) d.n1 += 1;
    if (axis == 2) d.n2 += 1;
    return d;
}

int validate_info(const plm_grid_info *info);
int validate_dims(const int32_t *shape, const char *what);

// 3-d launches: 64 threads along k (one wavefront on the contiguous axis), 4 rows of j, grid.z over i.
constexpr int BX = 64, BY = 4;
inline dim3 block3() { return dim3(BX, BY, 1); }
inline dim3 grid3(Dims d) { return dim3((d.n2 + BX - 1) / BX, (d.n1 + BY - 1) / BY, d.n0); }
// u, v and w stacked along grid.z: [0, nx] u, [nx+1, 2nx] v, [2nx+1, 3nx] w.
inline dim3 grid_faces(const plm_grid_info &g) {
    return dim3((g.dims[2] + 1 + BX - 1) / BX, (g.dims[1] + 1 + BY - 1) / BY, 3 * g.dims[0] + 1);
}

__device__ __forceinline__ bool thread_cell(Dims d, int &i, int &j, int &k) {
    k = blockIdx.x * BX + threadIdx.x;
    j = blockIdx.y * BY + threadIdx.y;
    i = blockIdx.z;
    return k < d.n2 && j < d.n1 && i < d.n0;
}

// Which face array this block works on (grid_faces), and the thread's face index in it.
__device__ __forceinline__ bool thread_face(const plm_grid_info &g, int &axis, Dims &d, int &i, int &j, int &k) {
    const int nx = g.dims[0];
    int z = blockIdx.z;
    if (z <= nx) {
        axis = 0;
        i = z;
    } else if (z <= 2 * nx) {
        axis = 1;
        i = z - (nx + 1);
    } else {
        axis = 2;
        i = z - (2 * nx + 1);
    }
    d = Dims{g.dims[0] + (axis == 0), g.dims[1] + (axis == 1), g.dims[2] + (axis == 2)};
    k = blockIdx.x * BX + threadIdx.x;
    j = blockIdx.y * BY + threadIdx.y;
    return k < d.n2 && j < d.n1;
}

// t / d for any 32-bit t: the multiply-high by m = ceil(2^32 / d) is floor(t / d) or one more (its error
// t (m d - 2^32) / (d 2^32) is below 1), corrected once
struct ExtentDiv {
    uint32_t d, m;
    __device__ __forceinline__ uint32_t operator()(uint32_t t) const {
        if (d == 1) return t;
        const uint32_t q = __umulhi(t, m);
        return q * d > t ? q - 1 : q;
    }
};

inline ExtentDiv extent_div(int d) {
    return ExtentDiv{(uint32_t)d, d == 1 ? 0u : (uint32_t)(((uint64_t(1) << 32) + (uint64_t)d - 1) / (uint64_t)d)};
}

// The three face arrays of a grid in one 1D launch, each enumerated in its own C order (one thread per face,
// 256 per block, a block never spanning two arrays): every wavefront holds 64 faces, where the (k, j) blocks
// of grid_faces leave most lanes of a 129-face row's third wavefront idle
constexpr int FACE_BLOCK = 256;
struct FaceLaunch {
    uint32_t block_end[3];  // cumulative blocks: u in [0, block_end[0]), v, then w
    uint32_t count[3];
    ExtentDiv div_n2[3], div_n1[3];
};

inline FaceLaunch face_launch(const plm_grid_info &g) {
    FaceLaunch f{};
    uint32_t blocks = 0;
    for (int a = 0; a < 3; a++) {
        const int n0 = g.dims[0] + (a == 0), n1 = g.dims[1] + (a == 1), n2 = g.dims[2] + (a == 2);
        f.count[a] = (uint32_t)n0 * (uint32_t)n1 * (uint32_t)n2;
        blocks += (f.count[a] + FACE_BLOCK - 1) / FACE_BLOCK;
        f.block_end[a] = blocks;
        f.div_n2[a] = extent_div(n2);
        f.div_n1[a] = extent_div(n1);
    }
    return f;
}

// The block's face array (uniform) and the thread's face in it; false past the array's end. The per-array
// values are picked with constant indices (a runtime index into the kernel arguments would copy them to
// scratch).
__device__ __forceinline__ bool packed_face(const plm_grid_info &g, const FaceLaunch &f, int &axis, int &i, int &j,
                                            int &k) {
    const uint32_t b = blockIdx.x;
    axis = b < f.block_end[0] ? 0 : b < f.block_end[1] ? 1 : 2;
    const uint32_t first = axis == 0 ? 0u : axis == 1 ? f.block_end[0] : f.block_end[1];
    const uint32_t count = axis == 0 ? f.count[0] : axis == 1 ? f.count[1] : f.count[2];
    const ExtentDiv dn2 = axis == 0 ? f.div_n2[0] : axis == 1 ? f.div_n2[1] : f.div_n2[2];
    const ExtentDiv dn1 = axis == 0 ? f.div_n1[0] : axis == 1 ? f.div_n1[1] : f.div_n1[2];
    const uint32_t t = (b - first) * FACE_BLOCK + threadIdx.x;
    if (t >= count) return false;
    const uint32_t r = dn2(t), c = dn1(r);
    k = (int)(t - r * dn2.d);
    i = (int)c;
    j = (int)(r - c * dn1.d);
    (void)g;
    return true;
}

template <typename T> struct Tri {
    T *p[3];
};

// ------------------------------------------------------------------ float helpers (Warp semantics)

__device__ __forceinline__ float fmin_(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmax_(float a, float b) { return a > b ? a : b; }
// wp.clamp(x, lo, hi) = min(max(x, lo), hi)
__device__ __forceinline__ float clamp_(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
// wp.lerp(a, b, t) = a * (1 - t) + b * t
__device__ __forceinline__ float lerp_(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float length3(float x, float y, float z) { return __fsqrt_rn(x * x + y * y + z * z); }

// Stored field values: FP32, or BF16 / FP16 storage of the mixed-precision solve (converted to FP32 on
// load, rounded to nearest even on store)
__device__ __forceinline__ float to_f32(float x) { return x; }
__device__ __forceinline__ float to_f32(__hip_bfloat16 x) { return __bfloat162float(x); }
__device__ __forceinline__ float to_f32(__half x) { return __half2float(x); }
template <typename T> __device__ __forceinline__ T from_f32(float x);
template <> __device__ __forceinline__ float from_f32<float>(float x) { return x; }
template <> __device__ __forceinline__ __hip_bfloat16 from_f32<__hip_bfloat16>(float x) { return __float2bfloat16(x); }
template <> __device__ __forceinline__ __half from_f32<__half>(float x) { return __float2half(x); }

// Element offset (i n1 + j) n2 + k with full-rate 24-bit multiplies: validate_dims keeps (nx + 1)(ny + 1) below
// 2^24 and every field below 2^31 / 3 elements, so the products and the byte offsets stay exact in 32 bits.
__device__ __forceinline__ uint32_t offset_of(Dims n, int i, int j, int k) {
    uint32_t row = __umul24((uint32_t)i, (uint32_t)n.n1);
    // kept a 24-bit product: fused with the add it becomes a quarter-rate 64-bit multiply-add
    asm volatile("" : "+v"(row));
    return __umul24(row + (uint32_t)j, (uint32_t)n.n2) + (uint32_t)k;
}

// f at a 32-bit byte offset from the field's base (the base stays scalar; one vector add per tap)
template <typename T> __device__ __forceinline__ float load_bytes(const T *__restrict__ f, uint32_t bytes) {
    return to_f32(*reinterpret_cast<const T *>(reinterpret_cast<const char *>(f) + bytes));
}
template <typename T> __device__ __forceinline__ float load_at(const T *__restrict__ f, uint32_t element) {
    return load_bytes(f, element * (uint32_t)sizeof(T));
}

// The clamped corner of a trilinear sample over elements of B bytes: the first tap's byte offset, the byte
// steps to the second tap on each axis (0 where the clamp pins both taps to the last index) and the weights.
struct Corner {
    uint32_t o, di, dj, dk;
    float tx, ty, tz;
};
// clamp_(g, 0, hi) for hi = n - 1 >= 0 in two instructions: fmax_(g, 0) is g or +0.0, never -0.0 or NaN, and
// non-negative floats order as their bit patterns do, so the unsigned minimum of the patterns is the select
// fmin_ makes, bit for bit (and needs no NaN quieting of its operands, as the float minimum would).
__device__ __forceinline__ float clamp_index(float g, float hi) {
    return __uint_as_float(min(__float_as_uint(fmax_(g, 0.0f)), __float_as_uint(hi)));
}

// x - (float)(int)x for a clamped coordinate 0 <= x < 2^24: x - floor(x), exact, which v_fract_f32 computes in
// one instruction.
__device__ __forceinline__ float index_weight(float x) { return __builtin_amdgcn_fractf(x); }

template <uint32_t B>
__device__ __forceinline__ Corner corner_of(Dims n, float gx, float gy, float gz) {
    const float hx = (float)(n.n0 - 1), hy = (float)(n.n1 - 1), hz = (float)(n.n2 - 1);
    const float x = clamp_index(gx, hx), y = clamp_index(gy, hy), z = clamp_index(gz, hz);
    const int i0 = (int)x, j0 = (int)y, k0 = (int)z;
    Corner c;
    c.o = offset_of(n, i0, j0, k0) * B;
    // i0 + 1 < n exactly when x < n - 1 (x <= n - 1, and n - 1 is a whole number)
    c.di = x < hx ? (uint32_t)(n.n1 * n.n2) * B : 0u;
    c.dj = y < hy ? (uint32_t)n.n2 * B : 0u;
    c.dk = z < hz ? B : 0u;
    c.tx = index_weight(x);
    c.ty = index_weight(y);
    c.tz = index_weight(z);
    return c;
}

// sample_index_space, sim/sample.py:10-29: i1 = min(i0 + 1, n - 1) etc., nested lerps over x, then y, then z
template <typename T>
__device__ __forceinline__ float sample_index(const T *__restrict__ f, Dims n, float gx, float gy, float gz) {
    const Corner c = corner_of<sizeof(T)>(n, gx, gy, gz);
    const uint32_t o1 = c.o + c.dk;
    const float c00 = lerp_(load_bytes(f, c.o), load_bytes(f, c.o + c.di), c.tx);
    const float c10 = lerp_(load_bytes(f, c.o + c.dj), load_bytes(f, c.o + c.di + c.dj), c.tx);
    const float c01 = lerp_(load_bytes(f, o1), load_bytes(f, o1 + c.di), c.tx);
    const float c11 = lerp_(load_bytes(f, o1 + c.dj), load_bytes(f, o1 + c.di + c.dj), c.tx);
    return lerp_(lerp_(c00, c10, c.ty), lerp_(c01, c11, c.ty), c.tz);
}

// range_index_space, sim/sample.py:32-50: min and max of the 8 taps (exact, so only the set matters)
__device__ __forceinline__ void range_index(const float *__restrict__ f, Dims n, float gx, float gy, float gz,
                                            float &lo, float &hi) {
    const Corner c = corner_of<4>(n, gx, gy, gz);
    const uint32_t o1 = c.o + c.dk;
    const float v[8] = {load_bytes(f, c.o),        load_bytes(f, o1),        load_bytes(f, c.o + c.dj),
                        load_bytes(f, o1 + c.dj), load_bytes(f, c.o + c.di), load_bytes(f, o1 + c.di),
                        load_bytes(f, c.o + c.di + c.dj), load_bytes(f, o1 + c.di + c.dj)};
    lo = v[0];
    hi = v[0];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        lo = fmin_(lo, v[t]);
        hi = fmax_(hi, v[t]);
    }
}

// is_outside_open_face, sim/sample.py:53-61
__device__ __forceinline__ bool outside_open_face(const plm_grid_info &g, float px, float py, float pz) {
    if (px < 0.0f && g.open_lo[0] == 1) return true;
    if (px > (float)g.dims[0] && g.open_hi[0] == 1) return true;
    if (py < 0.0f && g.open_lo[1] == 1) return true;
    if (py > (float)g.dims[1] && g.open_hi[1] == 1) return true;
    if (pz < 0.0f && g.open_lo[2] == 1) return true;
    if (pz > (float)g.dims[2] && g.open_hi[2] == 1) return true;
    return false;
}

// velocity_at, sim/sample.py:74-82
__device__ __forceinline__ void velocity_at(const plm_grid_info &g, const float *__restrict__ u,
                                            const float *__restrict__ v, const float *__restrict__ w, float px,
                                            float py, float pz, float &vx, float &vy, float &vz) {
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    vx = sample_index(u, Dims{nx + 1, ny, nz}, px - 0.0f, py - 0.5f, pz - 0.5f);
    vy = sample_index(v, Dims{nx, ny + 1, nz}, px - 0.5f, py - 0.0f, pz - 0.5f);
    vz = sample_index(w, Dims{nx, ny, nz + 1}, px - 0.5f, py - 0.5f, pz - 0.0f);
}

// trace, sim/advect.py:16-21
__device__ __forceinline__ void trace(const plm_grid_info &g, const float *__restrict__ u,
                                      const float *__restrict__ v, const float *__restrict__ w, float px, float py,
                                      float pz, float dto, float &sx, float &sy, float &sz) {
    float vx, vy, vz;
    velocity_at(g, u, v, w, px, py, pz, vx, vy, vz);
    const float half = 0.5f * dto;
    float mx = px - half * vx, my = py - half * vy, mz = pz - half * vz;
    velocity_at(g, u, v, w, mx, my, mz, vx, vy, vz);
    sx = px - dto * vx;
    sy = py - dto * vy;
    sz = pz - dto * vz;
}

// sample_field, sim/sample.py:64-71
__device__ __forceinline__ float sample_field(const float *__restrict__ f, Dims n, const plm_grid_info &g, float px,
                                              float py, float pz, float ox, float oy, float oz, float outside,
                                              int uses_outside) {
    if (uses_outside == 1 && outside_open_face(g, px, py, pz)) return outside;
    return sample_index(f, n, px - ox, py - oy, pz - oz);
}

constexpr float FAR_M = 1.0e4f;  // render/raymarch.py:21

// How far toward the camera an orthographic ray starts: FAR_M (the reference), or on the camera plane
// through the world origin (PLM_VIEW_ORTHO_NEAR; the same line, better conditioned in FP32)
__host__ __device__ inline float ortho_start_m(const plm_view &view) {
    return view.perspective == PLM_VIEW_ORTHO_NEAR ? 0.0f : FAR_M;
}

// slab_range, render/raymarch.py:55-69
__device__ __forceinline__ void slab_range(const float o[3], const float d[3], const float lo[3], const float hi[3],
                                           float &t_enter, float &t_exit) {
    t_enter = -1.0e30f;
    t_exit = 1.0e30f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (fabsf(d[a]) < 1.0e-12f) {
            if (o[a] < lo[a] || o[a] > hi[a]) {
                t_enter = 1.0f;
                t_exit = 0.0f;
                return;
            }
        } else {
            const float t0 = (lo[a] - o[a]) / d[a];
            const float t1 = (hi[a] - o[a]) / d[a];
            t_enter = fmax_(t_enter, fmin_(t0, t1));
            t_exit = fmin_(t_exit, fmax_(t0, t1));
        }
    }
}

// cell_fill, sim/pressure.py:39-44
__device__ __forceinline__ float cell_fill(const plm_grid_info &g, int axis, int index) {
    return index == g.dims[axis] - 1 ? g.fill_hi[axis] : 1.0f;
}

}  // namespace plm
