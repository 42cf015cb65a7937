// Pressure projection and geometric multigrid: sim/pressure.py and sim/multigrid.py.
//
// Kernels that restate the reference one launch per kernel (colour sweeps, residual, restrict, prolong) are
// kept for parity tests and small grids. The solver path uses three MI355X-shaped forms with the same
// arithmetic, so results are bit-identical:
//   - k_rb_sweep_fused: both colours of one sweep in one launch. A 64(k) x 8(j) tile streams along i with
//     a 2-cell halo in LDS; red is updated on the tile plus a 1-cell ring, then black on the tile. It
//     reads q_in and writes q_out (ping-pong), since a neighbouring tile's halo must see old values.
//   - k_residual_restrict: the coarse right-hand side straight from q and rhs (residual never stored).
//   - k_mg_tail: every level from the first with few enough cells down to the coarsest, solved by one
//     1024-thread workgroup in LDS, replacing ~90 launches per cycle at 128 x 192 x 128.
// Multigrid kernels take an optional device gate (tolerance mode): when gate[0] == 0 they return at once.
#include <vector>

#include "common.h"

namespace plm {
namespace {

__device__ __forceinline__ bool gated_off(const double *gate) { return gate != nullptr && gate[0] == 0.0; }

// ------------------------------------------------------------------ boundary faces, divergence, flags

// zero_closed_faces_kernel, sim/pressure.py:21-28. Only the six boundary planes can change, so the grid
// covers those: blockIdx.z = plane (axis * 2 + side), x/y over that plane's two other axes.
__global__ __launch_bounds__(256) void k_zero_closed(plm_grid_info g, Tri<float> f) {
    const int plane = blockIdx.z, axis = plane / 2, side = plane % 2;
    const bool closed = side == 0 ? g.open_lo[axis] == 0 : g.open_hi[axis] == 0;
    if (!closed) return;
    const Dims d{g.dims[0] + (axis == 0), g.dims[1] + (axis == 1), g.dims[2] + (axis == 2)};
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;  // the two other axes, ascending
    const int n1 = a1 == 0 ? d.n0 : d.n1, n2 = a2 == 1 ? d.n1 : d.n2;
    const int x = blockIdx.x * 64 + threadIdx.x % 64, y = blockIdx.y * 4 + threadIdx.x / 64;
    if (x >= n2 || y >= n1) return;
    int idx[3];
    idx[axis] = side == 0 ? 0 : g.dims[axis];
    idx[a1] = y;
    idx[a2] = x;
    f.p[axis][d.at(idx[0], idx[1], idx[2])] = 0.0f;
}

// divergence_kernel, sim/pressure.py:31-36
__global__ __launch_bounds__(256) void k_divergence(plm_grid_info g, const float *__restrict__ u,
                                                    const float *__restrict__ v, const float *__restrict__ w,
                                                    float *__restrict__ div) {
    const Dims nc = cell_dims(g), nu = face_dims(g, 0), nv = face_dims(g, 1), nw = face_dims(g, 2);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const float flux = (u[nu.at(i + 1, j, k)] - u[nu.at(i, j, k)]) + (v[nv.at(i, j + 1, k)] - v[nv.at(i, j, k)]) +
                       (w[nw.at(i, j, k + 1)] - w[nw.at(i, j, k)]);
    div[nc.at(i, j, k)] = flux / g.cell_m;
}

// The integer content of neighbor_term's branches (sim/pressure.py:52-69), per cell.
__device__ __forceinline__ int neighbor_kind(const plm_grid_info &g, int axis, int index, int step) {
    const int n = index + step;
    if (n >= 0 && n < g.dims[axis]) return 0;
    if ((step < 0 && g.open_lo[axis] == 1) || (step > 0 && g.open_hi[axis] == 1)) return 1;
    return 2;
}

__global__ __launch_bounds__(256) void k_build_flags(plm_grid_info g, uint16_t *__restrict__ flags) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const int cell[3] = {i, j, k};
    unsigned value = 0;
    for (int a = 0; a < 3; a++) {
        value |= (unsigned)neighbor_kind(g, a, cell[a], -1) << (2 * (2 * a));
        value |= (unsigned)neighbor_kind(g, a, cell[a], 1) << (2 * (2 * a + 1));
        if (cell[a] == g.dims[a] - 1) value |= 1u << (12 + a);
    }
    flags[nc.at(i, j, k)] = (uint16_t)value;
}

// ------------------------------------------------------------------ the finite-volume Poisson operator

// neighbor_term / neighbor_sum, sim/pressure.py:52-77. `q(i, j, k)` fetches a pressure. UNIT: the level
// has fill_hi = (1,1,1) and ghost_gap = 0.5 (the simulation grid), where every coupling is exactly 1, so
// coupling * q = q and the divisions are skipped; the sums are formed in the reference's order.
template <bool UNIT, typename Q>
__device__ __forceinline__ void neighbor_sum(const plm_grid_info &g, int i, int j, int k, Q q, float &t0,
                                             float &t1) {
    const int cell[3] = {i, j, k};
    t0 = 0.0f;
    t1 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int step = s == 0 ? -1 : 1;
            int nb[3] = {i, j, k};
            nb[a] += step;
            float v0 = 0.0f, v1 = 0.0f;
            if (UNIT) {
                if (nb[a] >= 0 && nb[a] < g.dims[a]) {
                    v0 = q(nb[0], nb[1], nb[2]);
                    v1 = 1.0f;
                } else if ((step < 0 && g.open_lo[a] == 1) || (step > 0 && g.open_hi[a] == 1)) {
                    v1 = 1.0f;
                }
            } else {
                float area = 1.0f;
#pragma unroll
                for (int other = 0; other < 3; other++)
                    if (other != a) area = area * cell_fill(g, other, cell[other]);
                const float own = cell_fill(g, a, cell[a]);
                if (nb[a] >= 0 && nb[a] < g.dims[a]) {
                    const float coupling = area / (0.5f * (own + cell_fill(g, a, nb[a])));
                    v0 = coupling * q(nb[0], nb[1], nb[2]);
                    v1 = coupling;
                } else if ((step < 0 && g.open_lo[a] == 1) || (step > 0 && g.open_hi[a] == 1)) {
                    v1 = area / (0.5f * own + g.ghost_gap);
                }
            }
            t0 = t0 + v0;
            t1 = t1 + v1;
        }
    }
}

// cell_volume, sim/pressure.py:47-49
__device__ __forceinline__ float cell_volume(const plm_grid_info &g, int i, int j, int k) {
    return cell_fill(g, 0, i) * cell_fill(g, 1, j) * cell_fill(g, 2, k);
}

// One red_black_sweep_kernel update (sim/pressure.py:86-92); returns q unchanged when no coupling.
__device__ __forceinline__ float gs_update(const plm_grid_info &g, int i, int j, int k, float q_c, float rhs_c,
                                          float t0, float t1, float relaxation) {
    if (t1 == 0.0f) return q_c;
    const float source = cell_volume(g, i, j, k) * g.cell_m * g.cell_m * rhs_c;
    const float target = (t0 - source) / t1;
    return q_c + relaxation * (target - q_c);
}

// residual_kernel's value, sim/multigrid.py:40-43
__device__ __forceinline__ float residual_value(const plm_grid_info &g, int i, int j, int k, float q_c, float rhs_c,
                                               float t0, float t1) {
    const float scale = cell_volume(g, i, j, k) * g.cell_m * g.cell_m;
    return rhs_c - (t0 - t1 * q_c) / scale;
}

struct GlobalQ {
    const float *q;
    Dims n;
    __device__ float operator()(int i, int j, int k) const { return q[n.at(i, j, k)]; }
};

inline bool unit_level(const plm_grid_info &g) {
    return g.fill_hi[0] == 1.0f && g.fill_hi[1] == 1.0f && g.fill_hi[2] == 1.0f && g.ghost_gap == 0.5f;
}

// red_black_sweep_kernel, sim/pressure.py:80-92 (one colour, in place)
template <bool UNIT>
__global__ __launch_bounds__(256) void k_rb_sweep(plm_grid_info g, float *__restrict__ q,
                                                  const float *__restrict__ rhs, int color, float relaxation,
                                                  const double *gate) {
    if (gated_off(gate)) return;
    const Dims n = cell_dims(g);
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    if ((i + j + k) % 2 != color) return;
    float t0, t1;
    neighbor_sum<UNIT>(g, i, j, k, GlobalQ{q, n}, t0, t1);
    const int c = n.at(i, j, k);
    if (t1 == 0.0f) return;
    q[c] = gs_update(g, i, j, k, q[c], rhs[c], t0, t1, relaxation);
}

// ------------------------------------------------------------------ both colours of a sweep in one launch

constexpr int FK = 64, FJ = 8, FH = 2;           // tile, halo
constexpr int PK = FK + 2 * FH, PJ = FJ + 2 * FH;  // LDS plane: 68 x 12
constexpr int FTHREADS = FK * FJ;                  // 512
constexpr int FCHUNK = 16;                         // i-planes per block

struct LdsPlane {
    const float *base;  // one plane of (PJ, PK) floats; (jj, kk) in tile-local coords, origin at -FH
    int j0, k0;         // global (j, k) of the tile's interior origin
    __device__ float operator()(int j, int k) const { return base[(j - j0 + FH) * PK + (k - k0 + FH)]; }
};

template <bool UNIT>
__global__ __launch_bounds__(FTHREADS) void k_rb_sweep_fused(plm_grid_info g, const float *__restrict__ q_in,
                                                             float *__restrict__ q_out,
                                                             const float *__restrict__ rhs, float relaxation,
                                                             const double *gate) {
    if (gated_off(gate)) return;
    __shared__ float old_ring[4][PJ * PK];  // old q, planes p..p+2 (+1 spare), tile + 2-cell halo
    __shared__ float red_ring[3][PJ * PK];  // red cells updated, planes p-1..p+1, tile + 1-cell ring
    const Dims n = cell_dims(g);
    const int k0 = blockIdx.x * FK, j0 = blockIdx.y * FJ;
    const int i_begin = blockIdx.z * FCHUNK, i_end = min(i_begin + FCHUNK, n.n0);
    const int tid = threadIdx.x;

    auto load_old = [&](int plane) {
        float *dst = old_ring[(plane + 4) & 3];
        for (int t = tid; t < PJ * PK; t += FTHREADS) {
            const int jj = t / PK, kk = t % PK;
            const int j = j0 - FH + jj, k = k0 - FH + kk;
            float value = 0.0f;
            if (plane >= 0 && plane < n.n0 && j >= 0 && j < n.n1 && k >= 0 && k < n.n2) value = q_in[n.at(plane, j, k)];
            dst[t] = value;
        }
    };
    // old q at (i, j, k) for a point inside the old ring
    auto old_at = [&](int i, int j, int k) {
        return old_ring[(i + 4) & 3][(j - j0 + FH) * PK + (k - k0 + FH)];
    };
    // red update of plane `plane` on the tile and its 1-cell ring
    auto red_plane = [&](int plane) {
        float *dst = red_ring[(plane + 3) % 3];
        if (plane < 0 || plane >= n.n0) return;
        for (int t = tid; t < (FJ + 2) * (FK + 2); t += FTHREADS) {
            const int jj = t / (FK + 2), kk = t % (FK + 2);
            const int j = j0 - 1 + jj, k = k0 - 1 + kk;
            if (j < 0 || j >= n.n1 || k < 0 || k >= n.n2) continue;
            if (((plane + j + k) & 1) != 0) continue;
            float t0, t1;
            neighbor_sum<UNIT>(g, plane, j, k, [&](int a, int b, int c) { return old_at(a, b, c); }, t0, t1);
            const float qc = old_at(plane, j, k);
            dst[(j - j0 + FH) * PK + (k - k0 + FH)] = gs_update(g, plane, j, k, qc, rhs[n.at(plane, j, k)], t0, t1,
                                                                relaxation);
        }
    };

    // prologue: old planes i_begin-2 .. i_begin+1, red planes i_begin-1 and i_begin
    for (int p = i_begin - 2; p <= i_begin + 1; p++) load_old(p);
    __syncthreads();
    red_plane(i_begin - 1);
    red_plane(i_begin);
    __syncthreads();

    const int jl = tid / FK, kl = tid % FK;
    const int j = j0 + jl, k = k0 + kl;
    const bool mine = j < n.n1 && k < n.n2;
    for (int p = i_begin; p < i_end; p++) {
        load_old(p + 2);  // overwrites plane p-2, whose last reader was red_plane(p-1)
        __syncthreads();
        red_plane(p + 1);
        __syncthreads();
        if (mine) {
            const int c = n.at(p, j, k);
            const int slot = (j - j0 + FH) * PK + (k - k0 + FH);
            float result;
            if (((p + j + k) & 1) == 0) {
                result = red_ring[(p + 3) % 3][slot];
            } else {
                float t0, t1;
                neighbor_sum<UNIT>(g, p, j, k,
                                   [&](int a, int b, int cc) {
                                       return red_ring[(a + 3) % 3][(b - j0 + FH) * PK + (cc - k0 + FH)];
                                   },
                                   t0, t1);
                result = gs_update(g, p, j, k, old_at(p, j, k), rhs[c], t0, t1, relaxation);
            }
            q_out[c] = result;
        }
        // the next iteration's load_old(p+3) overwrites plane p-1, read above by red_plane(p+1) and black(p)
        __syncthreads();
    }
}

// ------------------------------------------------------------------ gradient subtraction

// subtract_gradient_kernel, sim/pressure.py:95-116, on all three face arrays
__global__ __launch_bounds__(256) void k_subtract_gradient(plm_grid_info g, Tri<float> f,
                                                           const float *__restrict__ q) {
    int axis, i, j, k;
    Dims d;
    if (!thread_face(g, axis, d, i, j, k)) return;
    const Dims n = cell_dims(g);
    const int face[3] = {i, j, k};
    const int index = face[axis];
    int below[3] = {i, j, k};
    below[axis] = index - 1;
    float q_above = 0.0f, q_below = 0.0f;
    if (index == 0) {
        if (g.open_lo[axis] == 0) return;
        q_above = q[n.at(i, j, k)];
    } else if (index == g.dims[axis]) {
        if (g.open_hi[axis] == 0) return;
        q_below = q[n.at(below[0], below[1], below[2])];
    } else {
        q_above = q[n.at(i, j, k)];
        q_below = q[n.at(below[0], below[1], below[2])];
    }
    float *p = f.p[axis] + d.at(i, j, k);
    *p = *p - (q_above - q_below) / g.cell_m;
}

// ------------------------------------------------------------------ multigrid transfers

template <bool UNIT>
__global__ __launch_bounds__(256) void k_residual(plm_grid_info g, const float *__restrict__ q,
                                                  const float *__restrict__ rhs, float *__restrict__ r) {
    const Dims n = cell_dims(g);
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    float t0, t1;
    neighbor_sum<UNIT>(g, i, j, k, GlobalQ{q, n}, t0, t1);
    const int c = n.at(i, j, k);
    r[c] = residual_value(g, i, j, k, q[c], rhs[c], t0, t1);
}

// restrict_kernel, sim/multigrid.py:46-60
__global__ __launch_bounds__(256) void k_restrict(const float *__restrict__ fine, Dims nf, float *__restrict__ coarse,
                                                  Dims nc) {
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    float total = 0.0f, count = 0.0f;
    for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++)
            for (int dk = 0; dk < 2; dk++) {
                const int fi = 2 * i + di, fj = 2 * j + dj, fk = 2 * k + dk;
                if (fi < nf.n0 && fj < nf.n1 && fk < nf.n2) {
                    total = total + fine[nf.at(fi, fj, fk)];
                    count = count + 1.0f;
                }
            }
    coarse[nc.at(i, j, k)] = total / count;
}

// residual_kernel + restrict_kernel fused: each coarse cell forms its children's residuals in registers.
template <bool UNIT>
__global__ __launch_bounds__(256) void k_residual_restrict(plm_grid_info g, const float *__restrict__ q,
                                                           const float *__restrict__ rhs,
                                                           float *__restrict__ coarse, Dims nc,
                                                           const double *gate) {
    if (gated_off(gate)) return;
    const Dims nf = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    float total = 0.0f, count = 0.0f;
    for (int di = 0; di < 2; di++)
        for (int dj = 0; dj < 2; dj++)
            for (int dk = 0; dk < 2; dk++) {
                const int fi = 2 * i + di, fj = 2 * j + dj, fk = 2 * k + dk;
                if (fi < nf.n0 && fj < nf.n1 && fk < nf.n2) {
                    float t0, t1;
                    neighbor_sum<UNIT>(g, fi, fj, fk, GlobalQ{q, nf}, t0, t1);
                    const int c = nf.at(fi, fj, fk);
                    total = total + residual_value(g, fi, fj, fk, q[c], rhs[c], t0, t1);
                    count = count + 1.0f;
                }
            }
    coarse[nc.at(i, j, k)] = total / count;
}

// prolong_add_kernel, sim/multigrid.py:63-68
__global__ __launch_bounds__(256) void k_prolong_add(const float *__restrict__ coarse, Dims nc, float *__restrict__ fine,
                                                     Dims nf, const double *gate) {
    if (gated_off(gate)) return;
    int i, j, k;
    if (!thread_cell(nf, i, j, k)) return;
    const int c = nf.at(i, j, k);
    fine[c] = fine[c] + sample_index(coarse, nc, (float)i * 0.5f - 0.25f, (float)j * 0.5f - 0.25f,
                                     (float)k * 0.5f - 0.25f);
}

__global__ __launch_bounds__(256) void k_zero(int64_t count, float *__restrict__ f, const double *gate) {
    if (gated_off(gate)) return;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x)
        f[c] = 0.0f;
}

__global__ __launch_bounds__(256) void k_copy(int64_t count, const float *__restrict__ src, float *__restrict__ dst,
                                              const double *gate) {
    if (gated_off(gate)) return;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x)
        dst[c] = src[c];
}

inline dim3 grid1(int64_t count) {
    int64_t blocks = (count + 255) / 256;
    return dim3((unsigned)(blocks < 16384 ? (blocks > 0 ? blocks : 1) : 16384));
}

// ------------------------------------------------------------------ the coarse tail in one workgroup

constexpr int TAIL_THREADS = 1024;
constexpr int MAX_TAIL = 8;

struct TailLevels {
    int count;              // levels in the tail
    plm_grid_info info[MAX_TAIL];
    int offset[MAX_TAIL];   // float offset of the level's q in LDS; rhs follows at offset + cells
    int cells[MAX_TAIL];
};

struct LdsQ {
    const float *q;
    Dims n;
    __device__ float operator()(int i, int j, int k) const { return q[n.at(i, j, k)]; }
};

__device__ void tail_sweeps(const plm_grid_info &g, float *q, const float *rhs, int sweeps, float relaxation) {
    const Dims n = cell_dims(g);
    const int count = n.count();
    const bool unit = g.fill_hi[0] == 1.0f && g.fill_hi[1] == 1.0f && g.fill_hi[2] == 1.0f && g.ghost_gap == 0.5f;
    for (int s = 0; s < sweeps; s++) {
        for (int color = 0; color < 2; color++) {
            for (int c = threadIdx.x; c < count; c += TAIL_THREADS) {
                const int k = c % n.n2, j = (c / n.n2) % n.n1, i = c / (n.n1 * n.n2);
                if (((i + j + k) & 1) != color) continue;
                float t0, t1;
                if (unit)
                    neighbor_sum<true>(g, i, j, k, LdsQ{q, n}, t0, t1);
                else
                    neighbor_sum<false>(g, i, j, k, LdsQ{q, n}, t0, t1);
                q[c] = gs_update(g, i, j, k, q[c], rhs[c], t0, t1, relaxation);
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(TAIL_THREADS) void k_mg_tail(TailLevels t, float *__restrict__ q_top,
                                                          const float *__restrict__ rhs_top, int sweeps,
                                                          const double *gate) {
    if (gated_off(gate)) return;
    extern __shared__ float lds[];
    // top of the tail: q and rhs from global memory
    for (int c = threadIdx.x; c < t.cells[0]; c += TAIL_THREADS) {
        lds[t.offset[0] + c] = q_top[c];
        lds[t.offset[0] + t.cells[0] + c] = rhs_top[c];
    }
    __syncthreads();
    const int last = t.count - 1;
    // down: smooth, residual + restrict, zero the coarse q (sim/multigrid.py:137-143)
    for (int l = 0; l < last; l++) {
        const plm_grid_info &g = t.info[l];
        float *q = lds + t.offset[l];
        float *rhs = q + t.cells[l];
        tail_sweeps(g, q, rhs, sweeps, 1.0f);
        const Dims nf = cell_dims(g), nc = cell_dims(t.info[l + 1]);
        float *cq = lds + t.offset[l + 1];
        float *crhs = cq + t.cells[l + 1];
        const bool unit = g.fill_hi[0] == 1.0f && g.fill_hi[1] == 1.0f && g.fill_hi[2] == 1.0f && g.ghost_gap == 0.5f;
        for (int c = threadIdx.x; c < t.cells[l + 1]; c += TAIL_THREADS) {
            const int k = c % nc.n2, j = (c / nc.n2) % nc.n1, i = c / (nc.n1 * nc.n2);
            float total = 0.0f, count = 0.0f;
            for (int di = 0; di < 2; di++)
                for (int dj = 0; dj < 2; dj++)
                    for (int dk = 0; dk < 2; dk++) {
                        const int fi = 2 * i + di, fj = 2 * j + dj, fk = 2 * k + dk;
                        if (fi < nf.n0 && fj < nf.n1 && fk < nf.n2) {
                            float t0, t1;
                            if (unit)
                                neighbor_sum<true>(g, fi, fj, fk, LdsQ{q, nf}, t0, t1);
                            else
                                neighbor_sum<false>(g, fi, fj, fk, LdsQ{q, nf}, t0, t1);
                            const int f = nf.at(fi, fj, fk);
                            total = total + residual_value(g, fi, fj, fk, q[f], rhs[f], t0, t1);
                            count = count + 1.0f;
                        }
                    }
            crhs[c] = total / count;
            cq[c] = 0.0f;
        }
        __syncthreads();
    }
    // coarsest: max(COARSEST_SWEEPS, COARSEST_SWEEPS_PER_CELL * longest) sweeps at 1.5 (sim/multigrid.py:132-136)
    {
        const plm_grid_info &g = t.info[last];
        const int longest = max(g.dims[0], max(g.dims[1], g.dims[2]));
        float *q = lds + t.offset[last];
        tail_sweeps(g, q, q + t.cells[last], max(20, 4 * longest), 1.5f);
    }
    // up: prolong-add, smooth (sim/multigrid.py:145-147)
    for (int l = last - 1; l >= 0; l--) {
        const plm_grid_info &g = t.info[l];
        const Dims nf = cell_dims(g), nc = cell_dims(t.info[l + 1]);
        float *q = lds + t.offset[l];
        const float *cq = lds + t.offset[l + 1];
        for (int c = threadIdx.x; c < t.cells[l]; c += TAIL_THREADS) {
            const int k = c % nf.n2, j = (c / nf.n2) % nf.n1, i = c / (nf.n1 * nf.n2);
            q[c] = q[c] + sample_index(cq, nc, (float)i * 0.5f - 0.25f, (float)j * 0.5f - 0.25f, (float)k * 0.5f - 0.25f);
        }
        __syncthreads();
        tail_sweeps(g, q, q + t.cells[l], sweeps, 1.0f);
    }
    for (int c = threadIdx.x; c < t.cells[0]; c += TAIL_THREADS) q_top[c] = lds[t.offset[0] + c];
}

// ------------------------------------------------------------------ FP64 norms (tolerance mode)

constexpr int NORM_BLOCKS = 512, NORM_THREADS = 256;

template <bool UNIT>
__global__ __launch_bounds__(NORM_THREADS) void k_norm_partials(plm_grid_info g, const float *__restrict__ q,
                                                                const float *__restrict__ rhs,
                                                                double *__restrict__ partials, const double *gate) {
    if (gated_off(gate)) return;
    const Dims n = cell_dims(g);
    const int count = n.count();
    double rr = 0.0, bb = 0.0;
    for (int c = blockIdx.x * NORM_THREADS + threadIdx.x; c < count; c += NORM_BLOCKS * NORM_THREADS) {
        const int k = c % n.n2, j = (c / n.n2) % n.n1, i = c / (n.n1 * n.n2);
        float t0, t1;
        neighbor_sum<UNIT>(g, i, j, k, GlobalQ{q, n}, t0, t1);
        const double r = (double)residual_value(g, i, j, k, q[c], rhs[c], t0, t1);
        const double b = (double)rhs[c];
        rr += r * r;
        bb += b * b;
    }
    __shared__ double srr[NORM_THREADS], sbb[NORM_THREADS];
    srr[threadIdx.x] = rr;
    sbb[threadIdx.x] = bb;
    __syncthreads();
    for (int half = NORM_THREADS / 2; half > 0; half /= 2) {
        if ((int)threadIdx.x < half) {
            srr[threadIdx.x] += srr[threadIdx.x + half];
            sbb[threadIdx.x] += sbb[threadIdx.x + half];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = srr[0];
        partials[2 * blockIdx.x + 1] = sbb[0];
    }
}

// scratch layout: [0] relative residual, [1] gate (1 active / 0 converged), [2] |r|^2, [3] |rhs|^2,
// [4 ...] 2 * NORM_BLOCKS partials. `mode`: 0 plain norms (no gate change), 1 initial check, 2 after a cycle.
__global__ __launch_bounds__(NORM_THREADS) void k_norm_finish(double *scratch, double tolerance, int mode,
                                                              int32_t *cycles_run) {
    if (mode == 2 && scratch[1] == 0.0) return;
    const double *partials = scratch + 4;
    __shared__ double srr[NORM_THREADS], sbb[NORM_THREADS];
    double rr = 0.0, bb = 0.0;
    for (int b = threadIdx.x; b < NORM_BLOCKS; b += NORM_THREADS) {
        rr += partials[2 * b];
        bb += partials[2 * b + 1];
    }
    srr[threadIdx.x] = rr;
    sbb[threadIdx.x] = bb;
    __syncthreads();
    for (int half = NORM_THREADS / 2; half > 0; half /= 2) {
        if ((int)threadIdx.x < half) {
            srr[threadIdx.x] += srr[threadIdx.x + half];
            sbb[threadIdx.x] += sbb[threadIdx.x + half];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double r2 = srr[0], b2 = sbb[0];
    const double ratio = b2 > 0.0 ? sqrt(r2) / sqrt(b2) : (r2 > 0.0 ? INFINITY : 0.0);
    scratch[0] = ratio;
    scratch[2] = r2;
    scratch[3] = b2;
    if (mode == 0) return;
    if (mode == 1 && cycles_run != nullptr) cycles_run[0] = 0;
    if (mode == 2 && cycles_run != nullptr) {
        cycles_run[0] = cycles_run[0] + 1;
        cycles_run[1] = cycles_run[1] + 1;  // running total across solves, never reset here
    }
    scratch[1] = ratio > tolerance ? 1.0 : 0.0;
}

// ------------------------------------------------------------------ host orchestration

int launch_norms(const plm_grid_info &g, const float *q, const float *rhs, double *scratch, const double *gate,
                 hipStream_t s) {
    if (unit_level(g))
        k_norm_partials<true><<<NORM_BLOCKS, NORM_THREADS, 0, s>>>(g, q, rhs, scratch + 4, gate);
    else
        k_norm_partials<false><<<NORM_BLOCKS, NORM_THREADS, 0, s>>>(g, q, rhs, scratch + 4, gate);
    return check_launch("norm partials");
}

// `sweeps` sweeps on one level. q is updated in place; scratch (a level-sized buffer) enables the fused
// two-colour form, which ping-pongs q -> scratch -> q (an odd count starts with one colour-by-colour sweep).
int launch_sweeps(const plm_grid_info &g, float *q, const float *rhs, float *scratch, int sweeps, float relaxation,
                  const double *gate, hipStream_t s) {
    const Dims n = cell_dims(g);
    const bool unit = unit_level(g);
    int first_fused = 0;
    if (scratch == nullptr || n.count() < 4096) first_fused = sweeps;
    else if (sweeps % 2 == 1) first_fused = 1;
    TimingScope t("rb_sweeps", s);
    for (int sw = 0; sw < first_fused; sw++) {
        for (int color = 0; color < 2; color++) {
            if (unit)
                k_rb_sweep<true><<<grid3(n), block3(), 0, s>>>(g, q, rhs, color, relaxation, gate);
            else
                k_rb_sweep<false><<<grid3(n), block3(), 0, s>>>(g, q, rhs, color, relaxation, gate);
            PLM_TRY(check_launch("rb sweep"));
        }
    }
    const dim3 grid((n.n2 + FK - 1) / FK, (n.n1 + FJ - 1) / FJ, (n.n0 + FCHUNK - 1) / FCHUNK);
    for (int sw = first_fused; sw < sweeps; sw++) {
        const bool to_scratch = (sw - first_fused) % 2 == 0;
        const float *src = to_scratch ? q : scratch;
        float *dst = to_scratch ? scratch : q;
        TimingScope launch_timing(unit ? "rb_sweep_fused_top" : "rb_sweep_fused", s);
        if (unit)
            k_rb_sweep_fused<true><<<grid, FTHREADS, 0, s>>>(g, src, dst, rhs, relaxation, gate);
        else
            k_rb_sweep_fused<false><<<grid, FTHREADS, 0, s>>>(g, src, dst, rhs, relaxation, gate);
        PLM_TRY(check_launch("fused rb sweep"));
    }
    return 0;
}

struct Hierarchy {
    int n_levels;
    const plm_grid_info *infos;
    float *const *q;
    float *const *rhs;
    float *const *residual;
    int sweeps;
    int tail_start;
    TailLevels tail;
    size_t tail_bytes;
    const double *gate;
};

int v_cycle(const Hierarchy &h, int index, hipStream_t s) {
    if (index == h.tail_start) {
        k_mg_tail<<<1, TAIL_THREADS, h.tail_bytes, s>>>(h.tail, h.q[index], h.rhs[index], h.sweeps, h.gate);
        return check_launch("multigrid tail");
    }
    const plm_grid_info &g = h.infos[index];
    const Dims n = cell_dims(g);
    if (index == h.n_levels - 1) {  // coarsest level outside the tail (only when the tail is disabled)
        const int longest = max(g.dims[0], max(g.dims[1], g.dims[2]));
        return launch_sweeps(g, h.q[index], h.rhs[index], h.residual[index], max(20, 4 * longest), 1.5f, h.gate, s);
    }
    // residual[index] doubles as the ping-pong buffer of this level's fused sweeps
    PLM_TRY(launch_sweeps(g, h.q[index], h.rhs[index], h.residual[index], h.sweeps, 1.0f, h.gate, s));
    const plm_grid_info &cg = h.infos[index + 1];
    const Dims nc = cell_dims(cg);
    if (unit_level(g))
        k_residual_restrict<true><<<grid3(nc), block3(), 0, s>>>(g, h.q[index], h.rhs[index], h.rhs[index + 1], nc,
                                                                h.gate);
    else
        k_residual_restrict<false><<<grid3(nc), block3(), 0, s>>>(g, h.q[index], h.rhs[index], h.rhs[index + 1], nc,
                                                                 h.gate);
    PLM_TRY(check_launch("residual restrict"));
    k_zero<<<grid1(nc.count()), 256, 0, s>>>(nc.count(), h.q[index + 1], h.gate);
    PLM_TRY(check_launch("zero coarse"));
    PLM_TRY(v_cycle(h, index + 1, s));
    k_prolong_add<<<grid3(n), block3(), 0, s>>>(h.q[index + 1], nc, h.q[index], n, h.gate);
    PLM_TRY(check_launch("prolong add"));
    return launch_sweeps(g, h.q[index], h.rhs[index], h.residual[index], h.sweeps, 1.0f, h.gate, s);
}

}  // namespace
}  // namespace plm

using namespace plm;

extern "C" {

int plm_zero_closed_faces(const plm_grid_info *info, float *u, float *v, float *w, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w, "plm_zero_closed_faces: NULL array");
    const int longest = max(info->dims[0], max(info->dims[1], info->dims[2])) + 1;
    const dim3 grid((longest + 63) / 64, (longest + 3) / 4, 6);
    k_zero_closed<<<grid, 256, 0, as_stream(stream)>>>(*info, Tri<float>{{u, v, w}});
    return check_launch("plm_zero_closed_faces");
}

int plm_divergence(const plm_grid_info *info, const float *u, const float *v, const float *w, float *divergence,
                   void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && divergence, "plm_divergence: NULL array");
    k_divergence<<<grid3(cell_dims(*info)), block3(), 0, as_stream(stream)>>>(*info, u, v, w, divergence);
    return check_launch("plm_divergence");
}

int plm_build_flags(const plm_grid_info *info, uint16_t *flags, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(flags, "plm_build_flags: NULL array");
    k_build_flags<<<grid3(cell_dims(*info)), block3(), 0, as_stream(stream)>>>(*info, flags);
    return check_launch("plm_build_flags");
}

int plm_rb_sweep(const plm_grid_info *info, float *q, const float *rhs, int32_t color, float relaxation,
                 void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs, "plm_rb_sweep: NULL array");
    PLM_REQUIRE(color == 0 || color == 1, "plm_rb_sweep: color must be 0 or 1");
    const Dims n = cell_dims(*info);
    if (unit_level(*info))
        k_rb_sweep<true><<<grid3(n), block3(), 0, as_stream(stream)>>>(*info, q, rhs, color, relaxation, nullptr);
    else
        k_rb_sweep<false><<<grid3(n), block3(), 0, as_stream(stream)>>>(*info, q, rhs, color, relaxation, nullptr);
    return check_launch("plm_rb_sweep");
}

int plm_rb_sweeps(const plm_grid_info *info, float *q, const float *rhs, float *scratch, int32_t sweeps,
                  float relaxation, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs, "plm_rb_sweeps: NULL array");
    PLM_REQUIRE(sweeps >= 0, "plm_rb_sweeps: sweeps must be >= 0");
    return launch_sweeps(*info, q, rhs, scratch, sweeps, relaxation, nullptr, as_stream(stream));
}

int plm_subtract_gradient(const plm_grid_info *info, float *u, float *v, float *w, const float *q, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && q, "plm_subtract_gradient: NULL array");
    k_subtract_gradient<<<grid_faces(*info), block3(), 0, as_stream(stream)>>>(*info, Tri<float>{{u, v, w}}, q);
    return check_launch("plm_subtract_gradient");
}

int plm_residual(const plm_grid_info *info, const float *q, const float *rhs, float *residual, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs && residual, "plm_residual: NULL array");
    const Dims n = cell_dims(*info);
    if (unit_level(*info))
        k_residual<true><<<grid3(n), block3(), 0, as_stream(stream)>>>(*info, q, rhs, residual);
    else
        k_residual<false><<<grid3(n), block3(), 0, as_stream(stream)>>>(*info, q, rhs, residual);
    return check_launch("plm_residual");
}

int plm_restrict(const int32_t fine_shape[3], const float *fine, const int32_t coarse_shape[3], float *coarse,
                 void *stream) {
    PLM_TRY(validate_dims(fine_shape, "plm_restrict fine"));
    PLM_TRY(validate_dims(coarse_shape, "plm_restrict coarse"));
    PLM_REQUIRE(fine && coarse, "plm_restrict: NULL array");
    for (int a = 0; a < 3; a++)
        PLM_REQUIRE(coarse_shape[a] == (fine_shape[a] + 1) / 2, "plm_restrict: coarse shape must halve the fine");
    const Dims nc = dims_of(coarse_shape);
    k_restrict<<<grid3(nc), block3(), 0, as_stream(stream)>>>(fine, dims_of(fine_shape), coarse, nc);
    return check_launch("plm_restrict");
}

int plm_prolong_add(const int32_t coarse_shape[3], const float *coarse, const int32_t fine_shape[3], float *fine,
                    void *stream) {
    PLM_TRY(validate_dims(fine_shape, "plm_prolong_add fine"));
    PLM_TRY(validate_dims(coarse_shape, "plm_prolong_add coarse"));
    PLM_REQUIRE(fine && coarse, "plm_prolong_add: NULL array");
    const Dims nf = dims_of(fine_shape);
    k_prolong_add<<<grid3(nf), block3(), 0, as_stream(stream)>>>(coarse, dims_of(coarse_shape), fine, nf, nullptr);
    return check_launch("plm_prolong_add");
}

int plm_zero(int64_t count, float *f, void *stream) {
    PLM_REQUIRE(count >= 0 && (count == 0 || f), "plm_zero: bad arguments");
    if (count == 0) return 0;
    k_zero<<<grid1(count), 256, 0, as_stream(stream)>>>(count, f, nullptr);
    return check_launch("plm_zero");
}

int64_t plm_norm_scratch_doubles(const plm_grid_info *info) {
    (void)info;
    return 4 + 2 * NORM_BLOCKS;
}

int plm_residual_norms(const plm_grid_info *info, const float *q, const float *rhs, double *scratch, double *out,
                       void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs && scratch && out, "plm_residual_norms: NULL array");
    hipStream_t s = as_stream(stream);
    PLM_TRY(launch_norms(*info, q, rhs, scratch, nullptr, s));
    k_norm_finish<<<1, NORM_THREADS, 0, s>>>(scratch, 0.0, 0, nullptr);
    PLM_TRY(check_launch("norm finish"));
    PLM_REQUIRE(hipMemcpyAsync(out, scratch + 2, 2 * sizeof(double), hipMemcpyDeviceToDevice, s) == hipSuccess,
                "plm_residual_norms: copy failed");
    return 0;
}

int plm_vcycle(int32_t n_levels, const plm_grid_info *infos, float *const *q, float *const *rhs,
               float *const *residual, int32_t cycles, int32_t sweeps, double tolerance, double *norm_scratch,
               int32_t *cycles_run, void *stream) {
    PLM_REQUIRE(n_levels >= 1 && n_levels <= 16 && infos && q && rhs && residual, "plm_vcycle: bad level arrays");
    PLM_REQUIRE(cycles >= 0 && sweeps >= 0, "plm_vcycle: negative cycles or sweeps");
    for (int l = 0; l < n_levels; l++) {
        PLM_TRY(validate_info(&infos[l]));
        PLM_REQUIRE(q[l] && rhs[l] && residual[l], "plm_vcycle: NULL array at level %d", l);
        if (l > 0)
            for (int a = 0; a < 3; a++)
                PLM_REQUIRE(infos[l].dims[a] == (infos[l - 1].dims[a] + 1) / 2,
                            "plm_vcycle: level %d does not halve level %d", l, l - 1);
    }
    const bool tolerance_mode = tolerance > 0.0;
    PLM_REQUIRE(!tolerance_mode || norm_scratch, "plm_vcycle: tolerance mode needs norm_scratch");
    hipStream_t s = as_stream(stream);

    Hierarchy h{};
    h.n_levels = n_levels;
    h.infos = infos;
    h.q = q;
    h.rhs = rhs;
    h.residual = residual;
    h.sweeps = sweeps;
    h.gate = tolerance_mode ? norm_scratch + 1 : nullptr;
    // the tail starts at the first level from which every coarser level fits in LDS together
    int total = 0;
    h.tail_start = n_levels;
    for (int l = n_levels - 1; l >= 0; l--) {
        const int cells = infos[l].dims[0] * infos[l].dims[1] * infos[l].dims[2];
        if (total + cells > PLM_LDS_CELLS || n_levels - l > MAX_TAIL) break;
        total += cells;
        h.tail_start = l;
    }
    if (h.tail_start < n_levels) {
        h.tail.count = n_levels - h.tail_start;
        int offset = 0;
        for (int m = 0; m < h.tail.count; m++) {
            const plm_grid_info &g = infos[h.tail_start + m];
            h.tail.info[m] = g;
            h.tail.cells[m] = g.dims[0] * g.dims[1] * g.dims[2];
            h.tail.offset[m] = offset;
            offset += 2 * h.tail.cells[m];
        }
        h.tail_bytes = sizeof(float) * (size_t)offset;
    }

    TimingScope t("multigrid", s);
    if (tolerance_mode) {
        PLM_TRY(launch_norms(infos[0], q[0], rhs[0], norm_scratch, nullptr, s));
        k_norm_finish<<<1, NORM_THREADS, 0, s>>>(norm_scratch, tolerance, 1, cycles_run);
        PLM_TRY(check_launch("norm finish"));
    }
    for (int c = 0; c < cycles; c++) {
        PLM_TRY(v_cycle(h, 0, s));
        if (tolerance_mode) {
            PLM_TRY(launch_norms(infos[0], q[0], rhs[0], norm_scratch, h.gate, s));
            k_norm_finish<<<1, NORM_THREADS, 0, s>>>(norm_scratch, tolerance, 2, cycles_run);
            PLM_TRY(check_launch("norm finish"));
        }
    }
    return 0;
}

}  // extern "C"
