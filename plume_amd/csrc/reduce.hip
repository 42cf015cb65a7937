// Deterministic reductions: the smoke budget's line sums and totals (sim/reduce.py, sim/conserve.py) and the
// CFL cap's fastest face velocity (sim/step_cap.py). No float atomics: sums run in a fixed order, and the
// maximum (order-free) uses an integer atomic on the non-negative float's bits.
#include <mutex>

#include "common.h"

namespace plm {
namespace {

// line_sums_kernel, sim/reduce.py:12-25: one thread per (j, k) line, sequential over i in FP32. There are
// only n1 n2 lines, so each thread loads LINE_BURST values of i of all three fields at once (one memory
// round trip) and then adds them in order.
constexpr int LINE_THREADS = 64, LINE_BURST = 32;
__global__ __launch_bounds__(LINE_THREADS) void k_line_sums(Dims n, const float *__restrict__ a,
                                                            const float *__restrict__ b, const float *__restrict__ c,
                                                            float *__restrict__ lines) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int plane = n.n1 * n.n2;
    if (t >= plane) return;
    float ta = 0.0f, tb = 0.0f, tc = 0.0f;
    for (int i0 = 0; i0 < n.n0; i0 += LINE_BURST) {
        float va[LINE_BURST], vb[LINE_BURST], vc[LINE_BURST];
#pragma unroll
        for (int u = 0; u < LINE_BURST; u++) {
            const int i = min(i0 + u, n.n0 - 1);  // in range; the values past n0 are not added
            const uint32_t x = (uint32_t)(i * plane + t) * 4u;
            va[u] = load_bytes(a, x);
            vb[u] = load_bytes(b, x);
            vc[u] = load_bytes(c, x);
        }
#pragma unroll
        for (int u = 0; u < LINE_BURST; u++) {
            if (i0 + u < n.n0) {
                ta = ta + va[u];
                tb = tb + vb[u];
                tc = tc + vc[u];
            }
        }
    }
    lines[t] = ta;
    lines[plane + t] = tb;
    lines[2 * plane + t] = tc;
}

constexpr int BUDGET_THREADS = 1024;

// TripleSum's host float64 totals (sim/reduce.py:37) by a fixed-order tree, then SmokeBudget.hold's factor
// (sim/conserve.py:48-50): totals = before, semi-Lagrangian, MacCormack, factor. `copy`, when not null,
// receives the same four values.
__global__ __launch_bounds__(BUDGET_THREADS) void k_budget_totals(int plane, const float *__restrict__ lines,
                                                                  double *__restrict__ totals,
                                                                  double *__restrict__ copy) {
    __shared__ double s[3][BUDGET_THREADS];
    // each thread's lines in increasing order, BURST of them loaded at once (one memory round trip per burst,
    // where a load per step made every step wait for the last)
    constexpr int BURST = 8;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int t0 = threadIdx.x; t0 < plane; t0 += BURST * BUDGET_THREADS) {
        float v[3][BURST];
#pragma unroll
        for (int u = 0; u < BURST; u++) {
            const int t = min(t0 + u * BUDGET_THREADS, plane - 1);
            for (int q = 0; q < 3; q++) v[q][u] = lines[q * plane + t];
        }
#pragma unroll
        for (int u = 0; u < BURST; u++)
            if (t0 + u * BUDGET_THREADS < plane)
                for (int q = 0; q < 3; q++) acc[q] += (double)v[q][u];
    }
    for (int q = 0; q < 3; q++) s[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int half = BUDGET_THREADS / 2; half > 0; half /= 2) {
        if ((int)threadIdx.x < half)
            for (int q = 0; q < 3; q++) s[q][threadIdx.x] += s[q][threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double before = s[0][0], semi_lagrangian = s[1][0], maccormack = s[2][0];
    // held_total, sim/conserve.py:33-35
    const double target = fmax(semi_lagrangian, fmin(maccormack, before));
    totals[0] = before;
    totals[1] = semi_lagrangian;
    totals[2] = maccormack;
    totals[3] = maccormack > 0.0 ? target / maccormack : 1.0;
    if (copy != nullptr)
        for (int q = 0; q < 4; q++) copy[q] = totals[q];
}

// scale_kernel, sim/conserve.py:27-30, applied only when the factor is not 1 (sim/conserve.py:51-52)
__global__ __launch_bounds__(256) void k_scale_by_total(int64_t count, float *__restrict__ f,
                                                        const double *__restrict__ totals) {
    // four values per thread (one 16-byte access) over the aligned bulk, then the tail one by one; the first
    // values are read before the factor, so the two loads' latencies overlap
    const int64_t quads = (reinterpret_cast<uintptr_t>(f) & 15u) == 0 ? count / 4 : 0;
    const int64_t first = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
    float4 *f4 = reinterpret_cast<float4 *>(f);
    const float4 head = first < quads ? f4[first] : float4{0.0f, 0.0f, 0.0f, 0.0f};
    const double factor = totals[3];
    if (factor == 1.0) return;
    const float scale = (float)factor;
    for (int64_t c = first; c < quads; c += step) {
        float4 x = c == first ? head : f4[c];
        x.x = x.x * scale;
        x.y = x.y * scale;
        x.z = x.z * scale;
        x.w = x.w * scale;
        f4[c] = x;
    }
    for (int64_t c = 4 * quads + first; c < count; c += step) f[c] = f[c] * scale;
}

__global__ __launch_bounds__(256) void k_scale(int64_t count, float *__restrict__ f, float factor) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x)
        f[c] = f[c] * factor;
}

__global__ void k_clear_u32(unsigned *p) { *p = 0u; }

// max |f| over the three face arrays. |f| has its sign bit clear, so its bits order like unsigned integers,
// with +inf above every finite value and every NaN above +inf: the maximum is taken on the bits, so a NaN
// anywhere comes out as NaN (as NumPy's max propagates it, sim/step_cap.py:50-55) instead of being dropped
// the way fmaxf drops it.
__global__ __launch_bounds__(256) void k_max_abs3(Tri<const float> f, int64_t nu, int64_t nv, int64_t nw,
                                                  unsigned *__restrict__ out) {
    unsigned best = 0u;
    const int64_t total = nu + nv + nw;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < total; c += (int64_t)gridDim.x * blockDim.x) {
        float x;
        if (c < nu)
            x = f.p[0][c];
        else if (c < nu + nv)
            x = f.p[1][c - nu];
        else
            x = f.p[2][c - nu - nv];
        best = max(best, __float_as_uint(x) & 0x7fffffffu);
    }
    __shared__ unsigned s[256];
    s[threadIdx.x] = best;
    __syncthreads();
    for (int half = 128; half > 0; half /= 2) {
        if ((int)threadIdx.x < half) s[threadIdx.x] = max(s[threadIdx.x], s[threadIdx.x + half]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, s[0]);
}

// The scratch of plm_smoke_budget_noalloc: the four FP64 totals, then the line sums of up to
// PLM_BUDGET_NOALLOC_MAX_LINES (j, k) lines of the three fields. Each GPU holds its own copy until the process
// ends; calls on one GPU share it.
constexpr size_t NOALLOC_LINES_OFFSET = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_LINES_OFFSET + 3 * sizeof(float) * (size_t)PLM_BUDGET_NOALLOC_MAX_LINES;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_address[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device. In host code the array's
// name is only a stand-in for the symbol, so kernels get this address, never the name.
int scratch_address(void **address) {
    int device = -1;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return fail((int)e, "plm_smoke_budget_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (device < 0 || device >= MAX_DEVICES)
        return fail(PLM_EINVAL, "plm_smoke_budget_noalloc: device %d beyond the library's %d", device, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_address[device] == nullptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(e != hipSuccess ? (int)e : PLM_EINVAL, "plm_smoke_budget_noalloc: hipGetSymbolAddress: %s",
                        hipGetErrorString(e));
        g_scratch_address[device] = p;
    }
    *address = g_scratch_address[device];
    return 0;
}

inline dim3 grid1(int64_t count, int64_t cap = 16384) {
    int64_t blocks = (count + 255) / 256;
    return dim3((unsigned)(blocks < cap ? (blocks > 0 ? blocks : 1) : cap));
}

}  // namespace
}  // namespace plm

using namespace plm;

extern "C" {

int plm_line_sums(const int32_t shape[3], const float *a, const float *b, const float *c, float *lines,
                  void *stream) {
    PLM_TRY(validate_dims(shape, "plm_line_sums"));
    PLM_REQUIRE(a && b && c && lines, "plm_line_sums: NULL array");
    const Dims n = dims_of(shape);
    k_line_sums<<<(n.n1 * n.n2 + LINE_THREADS - 1) / LINE_THREADS, LINE_THREADS, 0, as_stream(stream)>>>(n, a, b, c,
                                                                                                     lines);
    return check_launch("plm_line_sums");
}

int plm_smoke_budget(const int32_t shape[3], const float *data, const float *hat, float *next, float *lines,
                     double *totals, void *stream) {
    PLM_TRY(validate_dims(shape, "plm_smoke_budget"));
    PLM_REQUIRE(data && hat && next && lines && totals, "plm_smoke_budget: NULL array");
    const Dims n = dims_of(shape);
    hipStream_t s = as_stream(stream);
    TimingScope t("budget", s);
    k_line_sums<<<(n.n1 * n.n2 + LINE_THREADS - 1) / LINE_THREADS, LINE_THREADS, 0, s>>>(n, data, hat, next, lines);
    PLM_TRY(check_launch("plm_smoke_budget line sums"));
    k_budget_totals<<<1, BUDGET_THREADS, 0, s>>>(n.n1 * n.n2, lines, totals, nullptr);
    PLM_TRY(check_launch("plm_smoke_budget totals"));
    k_scale_by_total<<<grid1((n.count() + 3) / 4), 256, 0, s>>>(n.count(), next, totals);
    return check_launch("plm_smoke_budget scale");
}

int plm_smoke_budget_noalloc(const int32_t shape[3], const float *data, const float *hat, float *next,
                             double *totals, void *stream) {
    PLM_TRY(validate_dims(shape, "plm_smoke_budget_noalloc"));
    PLM_REQUIRE(data && hat && next, "plm_smoke_budget_noalloc: NULL array");
    const Dims n = dims_of(shape);
    PLM_REQUIRE((int64_t)n.n1 * n.n2 <= PLM_BUDGET_NOALLOC_MAX_LINES,
                "plm_smoke_budget_noalloc: ny * nz = %d * %d lines exceeds the static scratch's %d", n.n1, n.n2,
                PLM_BUDGET_NOALLOC_MAX_LINES);
    void *base = nullptr;
    PLM_TRY(scratch_address(&base));
    double *held = static_cast<double *>(base);
    float *lines = reinterpret_cast<float *>(static_cast<unsigned char *>(base) + NOALLOC_LINES_OFFSET);
    hipStream_t s = as_stream(stream);
    TimingScope t("budget", s);
    // every word the kernels read is written first in this call: nothing to clear
    k_line_sums<<<(n.n1 * n.n2 + LINE_THREADS - 1) / LINE_THREADS, LINE_THREADS, 0, s>>>(n, data, hat, next, lines);
    PLM_TRY(check_launch("plm_smoke_budget_noalloc line sums"));
    k_budget_totals<<<1, BUDGET_THREADS, 0, s>>>(n.n1 * n.n2, lines, held, totals);
    PLM_TRY(check_launch("plm_smoke_budget_noalloc totals"));
    k_scale_by_total<<<grid1((n.count() + 3) / 4), 256, 0, s>>>(n.count(), next, held);
    return check_launch("plm_smoke_budget_noalloc scale");
}

int plm_scale(int64_t count, float *f, float factor, void *stream) {
    PLM_REQUIRE(count >= 0 && (count == 0 || f), "plm_scale: bad arguments");
    if (count == 0) return 0;
    k_scale<<<grid1(count), 256, 0, as_stream(stream)>>>(count, f, factor);
    return check_launch("plm_scale");
}

int plm_max_abs3(const plm_grid_info *info, const float *u, const float *v, const float *w, float *out,
                 void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && out, "plm_max_abs3: NULL array");
    hipStream_t s = as_stream(stream);
    const int64_t nu = face_dims(*info, 0).count(), nv = face_dims(*info, 1).count(), nw = face_dims(*info, 2).count();
    k_clear_u32<<<1, 1, 0, s>>>(reinterpret_cast<unsigned *>(out));
    PLM_TRY(check_launch("plm_max_abs3 clear"));
    k_max_abs3<<<grid1(nu + nv + nw, 2048), 256, 0, s>>>(Tri<const float>{{u, v, w}}, nu, nv, nw,
                                                         reinterpret_cast<unsigned *>(out));
    return check_launch("plm_max_abs3");
}

}  // extern "C"
