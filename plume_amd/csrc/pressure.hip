// Pressure projection and geometric multigrid: sim/pressure.py and sim/multigrid.py.
//
// Every form below has the reference's arithmetic, so results are bit-identical to the colour-by-colour,
// kernel-by-kernel restatement:
//   - Couplings. neighbor_term's coupling area / (0.5 (own + neighbour fill)) takes at most a few distinct
//     values per level (fills are 1 except the last cell per axis), so the host forms them once per level
//     in float32, in the same operations (`Coupling`), and the kernels index them instead of dividing.
//   - k_rb_sweep_tile: both colours of one sweep in one launch. A 64 (k) x 4 (j) x 4 (i) tile plus a 2-cell
//     halo sits in LDS; red is updated in place on the tile and a 1-cell ring (red reads only black), then
//     black on the tile. It reads q_in and writes q_out (ping-pong), since a neighbouring tile's halo must
//     see old values. Its first sweep of a level can start from q = 0 (the coarse q the reference zeroes)
//     or from q + prolongation of the coarser correction (prolong_add fused into the load).
//   - k_residual_restrict: the coarse right-hand side straight from q and rhs (the residual is never stored).
//   - k_level_smooth: the flat levels between the simulation grid and the tail, one or two sweeps a launch in
//     LDS tiles with a halo of two cells a sweep, fused with the prolongation before them or the residual and
//     restriction after them.
//   - k_mg_tail: every level from the first with few enough cells down to the coarsest, solved by one
//     1024-thread workgroup in LDS, instead of ~60 launches per cycle at 128 x 192 x 128.
// Multigrid kernels take an optional device gate (tolerance mode): when it reads 0 they do nothing (most read it
// once their first loads are in flight).
#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <type_traits>
#include <utility>
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

// zero_closed_faces_kernel then divergence_kernel (sim/pressure.py:21-36, 146-149) in one pass: each cell reads
// a closed boundary face as the 0 the zeroing writes, and the cell beside that face writes the 0, so the
// divergence and the faces end as the two kernels leave them. DIVERGENCE_ROWS consecutive rows (j) per thread,
// their v faces shared between neighbouring rows; every load comes before the stores.
constexpr int DIVERGENCE_ROWS = 4;
__global__ __launch_bounds__(256) void k_close_divergence(plm_grid_info g, float *__restrict__ u,
                                                          float *__restrict__ v, float *__restrict__ w,
                                                          float *__restrict__ div) {
    constexpr int R = DIVERGENCE_ROWS;
    const Dims nc = cell_dims(g), nu = face_dims(g, 0), nv = face_dims(g, 1), nw = face_dims(g, 2);
    const int k = blockIdx.x * BX + threadIdx.x, i = blockIdx.z;
    const int j0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.y * BY + threadIdx.y) * R));
    if (k >= nc.n2 || j0 >= nc.n1) return;
    const bool zx_lo = i == 0 && g.open_lo[0] == 0, zx_hi = i == g.dims[0] - 1 && g.open_hi[0] == 0;
    const bool zz_lo = k == 0 && g.open_lo[2] == 0, zz_hi = k == g.dims[2] - 1 && g.open_hi[2] == 0;
    // v faces j0 .. j0 + R (face m below row m, above row m - 1); rows past the grid read its last row
    float u_lo[R], u_hi[R], w_lo[R], w_hi[R], v_face[R + 1];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = min(j0 + r, nc.n1 - 1);
        u_lo[r] = u[nu.at(i, j, k)];
        u_hi[r] = u[nu.at(i + 1, j, k)];
        w_lo[r] = w[nw.at(i, j, k)];
        w_hi[r] = w[nw.at(i, j, k + 1)];
    }
#pragma unroll
    for (int m = 0; m <= R; m++) v_face[m] = v[nv.at(i, min(j0 + m, nv.n1 - 1), k)];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = j0 + r;
        if (j >= nc.n1) break;
        const bool zy_lo = j == 0 && g.open_lo[1] == 0, zy_hi = j == g.dims[1] - 1 && g.open_hi[1] == 0;
        const float a_lo = zx_lo ? 0.0f : u_lo[r], a_hi = zx_hi ? 0.0f : u_hi[r];
        const float b_lo = zy_lo ? 0.0f : v_face[r], b_hi = zy_hi ? 0.0f : v_face[r + 1];
        const float c_lo = zz_lo ? 0.0f : w_lo[r], c_hi = zz_hi ? 0.0f : w_hi[r];
        const float flux = (a_hi - a_lo) + (b_hi - b_lo) + (c_hi - c_lo);
        div[nc.at(i, j, k)] = flux / g.cell_m;
        if (zx_lo) u[nu.at(i, j, k)] = 0.0f;
        if (zx_hi) u[nu.at(i + 1, j, k)] = 0.0f;
        if (zy_lo) v[nv.at(i, j, k)] = 0.0f;
        if (zy_hi) v[nv.at(i, j + 1, k)] = 0.0f;
        if (zz_lo) w[nw.at(i, j, k)] = 0.0f;
        if (zz_hi) w[nw.at(i, j, k + 1)] = 0.0f;
    }
}

inline dim3 divergence_grid(const plm_grid_info &g) {
    const Dims nc = cell_dims(g);
    return dim3((nc.n2 + BX - 1) / BX, (nc.n1 + BY * DIVERGENCE_ROWS - 1) / (BY * DIVERGENCE_ROWS), nc.n0);
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

// The values neighbor_term (sim/pressure.py:52-69) and cell_volume (:47-49) can take on one level. With
// La = "cell is the last on axis a" and the other two axes b < c, area = (1 * F_b) * F_c, F = fill_hi if last
// else 1, indexed Lb + 2 Lc. An in-domain neighbour's coupling is area / (0.5 (own + its fill)), where
// own + fill is 1 + 1 or 1 + fill_hi (index 0 / 1); an open-face ghost's is area / (0.5 own + ghost_gap),
// own 1 or fill_hi (index 0 / 1). vol = (F_0 F_1) F_2, indexed L0 + 2 L1 + 4 L2.
struct Coupling {
    float in[3][4][2];
    float ghost[3][4][2];
    float vol[8];
};

struct LevelArgs {
    plm_grid_info g;
    Coupling cp;
    int unit;     // every coupling and volume is exactly 1 (the simulation grid)
    int flat;     // fill_hi (1, 1, 1): every coupling between two cells and every volume is exactly 1
    float ghost;  // the open-face ghost coupling of a flat level (1 on a unit level)
};

Coupling make_coupling(const plm_grid_info &g) {
    Coupling cp;
    for (int a = 0; a < 3; a++) {
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        for (int idx = 0; idx < 4; idx++) {
            const float Fb = (idx & 1) ? g.fill_hi[b] : 1.0f, Fc = (idx & 2) ? g.fill_hi[c] : 1.0f;
            volatile float area = 1.0f * Fb;  // volatile: keep each float operation as written
            area = area * Fc;
            const float fa = g.fill_hi[a];
            volatile float d0 = 1.0f + 1.0f, d1 = 1.0f + fa;
            volatile float h0 = 0.5f * d0, h1 = 0.5f * d1;
            cp.in[a][idx][0] = area / h0;
            cp.in[a][idx][1] = area / h1;
            volatile float g0 = 0.5f * 1.0f, g1 = 0.5f * fa;
            g0 = g0 + g.ghost_gap;
            g1 = g1 + g.ghost_gap;
            cp.ghost[a][idx][0] = area / g0;
            cp.ghost[a][idx][1] = area / g1;
        }
    }
    for (int idx = 0; idx < 8; idx++) {
        volatile float v = ((idx & 1) ? g.fill_hi[0] : 1.0f) * ((idx & 2) ? g.fill_hi[1] : 1.0f);
        v = v * ((idx & 4) ? g.fill_hi[2] : 1.0f);
        cp.vol[idx] = v;
    }
    return cp;
}

bool unit_level(const plm_grid_info &g) {
    return g.fill_hi[0] == 1.0f && g.fill_hi[1] == 1.0f && g.fill_hi[2] == 1.0f && g.ghost_gap == 0.5f;
}

LevelArgs level_args(const plm_grid_info &g) {
    LevelArgs L;
    L.g = g;
    L.cp = make_coupling(g);
    L.unit = unit_level(g) ? 1 : 0;
    L.flat = g.fill_hi[0] == 1.0f && g.fill_hi[1] == 1.0f && g.fill_hi[2] == 1.0f ? 1 : 0;
    L.ghost = L.cp.ghost[0][0][0];
    return L;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float uniform(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// One of four level constants by two lane predicates. Each value passes through readfirstlane so that the
// selects stay selects of scalar registers (as selects of loads they become one per-lane indexed load).
__device__ __forceinline__ float sel4(float v0, float v1, float v2, float v3, bool last_b, bool last_c) {
    v0 = uniform(v0);
    v1 = uniform(v1);
    v2 = uniform(v2);
    v3 = uniform(v3);
    return last_c ? (last_b ? v3 : v2) : (last_b ? v1 : v0);
}

// Coupling.in[a][area][pair_last] and Coupling.ghost[a][area][own_last], area = last_b + 2 last_c
__device__ __forceinline__ float coupling_in(const Coupling &cp, int a, bool last_b, bool last_c, bool pair_last) {
    const float x0 = sel4(cp.in[a][0][0], cp.in[a][1][0], cp.in[a][2][0], cp.in[a][3][0], last_b, last_c);
    const float x1 = sel4(cp.in[a][0][1], cp.in[a][1][1], cp.in[a][2][1], cp.in[a][3][1], last_b, last_c);
    return pair_last ? x1 : x0;
}
__device__ __forceinline__ float coupling_ghost(const Coupling &cp, int a, bool last_b, bool last_c, bool own_last) {
    const float x0 = sel4(cp.ghost[a][0][0], cp.ghost[a][1][0], cp.ghost[a][2][0], cp.ghost[a][3][0], last_b, last_c);
    const float x1 = sel4(cp.ghost[a][0][1], cp.ghost[a][1][1], cp.ghost[a][2][1], cp.ghost[a][3][1], last_b, last_c);
    return own_last ? x1 : x0;
}

// neighbor_sum, sim/pressure.py:72-77: (sum of coupling * q over the six neighbours, sum of couplings), added
// in the reference's order (x-, x+, y-, y+, z-, z+), ghosts and closed faces adding 0 to the first.
// UNIT levels skip the multiplications by an exact 1. q is read only inside the domain.
template <bool UNIT, typename Q>
__device__ __forceinline__ void neighbor_sum(const LevelArgs &L, int i, int j, int k, Q q, float &t0, float &t1) {
    const plm_grid_info &g = L.g;
    const int cell[3] = {i, j, k};
    const bool last[3] = {i == g.dims[0] - 1, j == g.dims[1] - 1, k == g.dims[2] - 1};
    t0 = 0.0f;
    t1 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        // step -1
        {
            int nb[3] = {i, j, k};
            nb[a] -= 1;
            float v0 = 0.0f, v1 = 0.0f;
            if (nb[a] >= 0) {
                const float cp = UNIT ? 1.0f : coupling_in(L.cp, a, last[b], last[c], last[a]);
                v0 = UNIT ? q(nb[0], nb[1], nb[2]) : cp * q(nb[0], nb[1], nb[2]);
                v1 = cp;
            } else if (g.open_lo[a] == 1) {
                v1 = UNIT ? 1.0f : coupling_ghost(L.cp, a, last[b], last[c], last[a]);
            }
            t0 = t0 + v0;
            t1 = t1 + v1;
        }
        // step +1
        {
            int nb[3] = {i, j, k};
            nb[a] += 1;
            float v0 = 0.0f, v1 = 0.0f;
            if (nb[a] < g.dims[a]) {
                const float cp = UNIT ? 1.0f : coupling_in(L.cp, a, last[b], last[c], cell[a] + 1 == g.dims[a] - 1);
                v0 = UNIT ? q(nb[0], nb[1], nb[2]) : cp * q(nb[0], nb[1], nb[2]);
                v1 = cp;
            } else if (g.open_hi[a] == 1) {
                v1 = UNIT ? 1.0f : coupling_ghost(L.cp, a, last[b], last[c], true);
            }
            t0 = t0 + v0;
            t1 = t1 + v1;
        }
    }
}

// neighbor_sum for a q that reads +0.0 outside the domain (an LDS tile whose halo holds 0 there): the same
// values in the same order, without branches. An outside neighbour's term is coupling * (+0.0) = +0.0 (or
// +0.0 itself on UNIT levels), what the reference adds for a ghost or a closed face.
template <bool UNIT, typename Q>
__device__ __forceinline__ void neighbor_sum_padded(const LevelArgs &L, int i, int j, int k, Q q, float &t0,
                                                    float &t1) {
    const plm_grid_info &g = L.g;
    const int cell[3] = {i, j, k};
    const bool last[3] = {i == g.dims[0] - 1, j == g.dims[1] - 1, k == g.dims[2] - 1};
    t0 = 0.0f;
    t1 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        const bool open_lo = uniform(g.open_lo[a]) == 1, open_hi = uniform(g.open_hi[a]) == 1;
        {
            int nb[3] = {i, j, k};
            nb[a] -= 1;
            const bool in = cell[a] > 0;
            const float qn = q(nb[0], nb[1], nb[2]);
            if (UNIT) {
                t0 = t0 + qn;
                t1 = t1 + ((in || open_lo) ? 1.0f : 0.0f);
            } else {
                const float cp = coupling_in(L.cp, a, last[b], last[c], last[a]);
                t0 = t0 + cp * qn;
                t1 = t1 + (in ? cp : open_lo ? coupling_ghost(L.cp, a, last[b], last[c], last[a]) : 0.0f);
            }
        }
        {
            int nb[3] = {i, j, k};
            nb[a] += 1;
            const bool in = !last[a];
            const float qn = q(nb[0], nb[1], nb[2]);
            if (UNIT) {
                t0 = t0 + qn;
                t1 = t1 + ((in || open_hi) ? 1.0f : 0.0f);
            } else {
                const float cp = coupling_in(L.cp, a, last[b], last[c], cell[a] + 1 == g.dims[a] - 1);
                t0 = t0 + cp * qn;
                t1 = t1 + (in ? cp : open_hi ? coupling_ghost(L.cp, a, last[b], last[c], true) : 0.0f);
            }
        }
    }
}

// The coupling table of a level with couplings other than 1 as an LDS array: Coupling.in (3 x 4 x 2 floats),
// Coupling.ghost (the same) and +0.0 at CP_ZERO. Selecting among uniform values costs a move and a select
// per value (one scalar operand per vector instruction); an indexed LDS read is one instruction.
constexpr int CP_GHOST = 24, CP_ZERO = 48, CP_WORDS = 49;

// neighbor_sum_padded for a level with couplings other than 1, every coupling read from the LDS table:
// per neighbour one read, of the in-domain coupling, the open-face ghost coupling or +0.0 as the neighbour
// is inside, beyond an open face or beyond a closed one. An outside neighbour's q is +0.0, so its t0 term
// is +0.0 whichever of these it is multiplied by, as in neighbor_sum. The same values in the same order.
template <typename Q>
__device__ __forceinline__ void neighbor_sum_lds(const LevelArgs &L, const float *cps, int i, int j, int k, Q q,
                                                 float &t0, float &t1) {
    const plm_grid_info &g = L.g;
    const int cell[3] = {i, j, k};
    const bool last[3] = {i == g.dims[0] - 1, j == g.dims[1] - 1, k == g.dims[2] - 1};
    t0 = 0.0f;
    t1 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        const bool open_lo = uniform(g.open_lo[a]) == 1, open_hi = uniform(g.open_hi[a]) == 1;
        const int row = a * 8 + (last[b] ? 2 : 0) + (last[c] ? 4 : 0);  // Coupling.in[a][last_b + 2 last_c]
        {
            int nb[3] = {i, j, k};
            nb[a] -= 1;
            const int pair_last = last[a] ? 1 : 0;
            const int index = cell[a] > 0 ? row + pair_last : open_lo ? CP_GHOST + row + pair_last : CP_ZERO;
            const float cp = cps[index];
            t0 = t0 + cp * q(nb[0], nb[1], nb[2]);
            t1 = t1 + cp;
        }
        {
            int nb[3] = {i, j, k};
            nb[a] += 1;
            const int pair_last = cell[a] + 1 == g.dims[a] - 1 ? 1 : 0;
            const int index = !last[a] ? row + pair_last : open_hi ? CP_GHOST + row + 1 : CP_ZERO;
            const float cp = cps[index];
            t0 = t0 + cp * q(nb[0], nb[1], nb[2]);
            t1 = t1 + cp;
        }
    }
}

// cell_volume, sim/pressure.py:47-49: (F_0 F_1) F_2 with F = fill_hi on the last cell of an axis, else 1
__device__ __forceinline__ float cell_volume(const LevelArgs &L, int i, int j, int k) {
    const plm_grid_info &g = L.g;
    float v = (i == g.dims[0] - 1 ? uniform(g.fill_hi[0]) : 1.0f) * (j == g.dims[1] - 1 ? uniform(g.fill_hi[1]) : 1.0f);
    v = v * (k == g.dims[2] - 1 ? uniform(g.fill_hi[2]) : 1.0f);
    return v;
}

// One red_black_sweep_kernel update (sim/pressure.py:86-92); q unchanged when nothing couples
__device__ __forceinline__ float gs_update(const LevelArgs &L, int i, int j, int k, float q_c, float rhs_c, float t0,
                                          float t1, float relaxation) {
    if (t1 == 0.0f) return q_c;
    const float source = cell_volume(L, i, j, k) * L.g.cell_m * L.g.cell_m * rhs_c;
    const float target = (t0 - source) / t1;
    return q_c + relaxation * (target - q_c);
}

// residual_kernel's value, sim/multigrid.py:40-43
__device__ __forceinline__ float residual_value(const LevelArgs &L, int i, int j, int k, float q_c, float rhs_c,
                                               float t0, float t1) {
    const float scale = cell_volume(L, i, j, k) * L.g.cell_m * L.g.cell_m;
    return rhs_c - (t0 - t1 * q_c) / scale;
}

// The residual of one cell of a unit-coupling level (every coupling and cell volume 1, so the scale is
// (1 cell_m) cell_m) from its six neighbours in the reference's order x-, x+, y-, y+, z-, z+, each passed as
// (value, inside): an outside neighbour adds +0.0, what neighbor_sum adds for a ghost or a closed face.
// `couples` is the sum of couplings, the faces with an inside neighbour or open: a sum of ones, exact in any
// order. The same values as residual_value(neighbor_sum).
__device__ __forceinline__ float unit_residual(float q_c, float rhs_c, float scale, const float (&qn)[6],
                                               const bool (&in)[6], int couples) {
    float t0 = 0.0f;
#pragma unroll
    for (int m = 0; m < 6; m++) t0 = t0 + (in[m] ? qn[m] : 0.0f);
    return rhs_c - (t0 - (float)couples * q_c) / scale;
}

// the faces of a cell along one axis that couple on a unit level: an inside neighbour, or an open face
__device__ __forceinline__ int unit_couples(const plm_grid_info &g, int a, bool lo_in, bool hi_in) {
    // bitwise, not ||: a short-circuit on the uniform flag would become a branch
    const int open_lo = uniform(g.open_lo[a]) == 1 ? 1 : 0, open_hi = uniform(g.open_hi[a]) == 1 ? 1 : 0;
    return ((int)lo_in | open_lo) + ((int)hi_in | open_hi);
}

struct GlobalQ {
    const float *q;
    Dims n;
    __device__ float operator()(int i, int j, int k) const { return q[n.at(i, j, k)]; }
};

template <typename T>
struct StoredQ {
    const T *q;
    Dims n;
    __device__ float operator()(int i, int j, int k) const { return to_f32(q[n.at(i, j, k)]); }
};

// red_black_sweep_kernel, sim/pressure.py:80-92 (one colour, in place)
template <bool UNIT>
__global__ __launch_bounds__(256) void k_rb_sweep(LevelArgs L, float *__restrict__ q, const float *__restrict__ rhs,
                                                  int color, float relaxation, const double *gate) {
    if (gated_off(gate)) return;
    const Dims n = cell_dims(L.g);
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    if ((i + j + k) % 2 != color) return;
    float t0, t1;
    neighbor_sum<UNIT>(L, i, j, k, GlobalQ{q, n}, t0, t1);
    if (t1 == 0.0f) return;
    const int c = n.at(i, j, k);
    q[c] = gs_update(L, i, j, k, q[c], rhs[c], t0, t1, relaxation);
}

// ------------------------------------------------------------------ both colours of a sweep in one launch

constexpr int TK = 64, TH = 2, LK = TK + 2 * TH;

enum SweepStart { FROM_Q = 0, FROM_ZERO = 1, FROM_PROLONG = 2 };

// sample_index_space of the coarse correction at a fine cell (prolong_add_kernel, sim/multigrid.py:63-68)
template <typename T>
__device__ __forceinline__ float prolongation(const T *__restrict__ coarse, Dims nc, int i, int j, int k) {
    return sample_index(coarse, nc, (float)i * 0.5f - 0.25f, (float)j * 0.5f - 0.25f, (float)k * 0.5f - 0.25f);
}

// The k part of prolongation() at fine k (the same for every row): byte offsets of the two coarse k taps
// and the weight
template <typename T>
__device__ __forceinline__ void prolong_k(Dims nc, int k, uint32_t &pk0, uint32_t &pk1, float &tz) {
    const float z = clamp_((float)k * 0.5f - 0.25f, 0.0f, (float)(nc.n2 - 1));
    const int kc0 = (int)z;
    pk0 = (uint32_t)kc0 * (uint32_t)sizeof(T);
    pk1 = (uint32_t)min(kc0 + 1, nc.n2 - 1) * (uint32_t)sizeof(T);
    tz = z - (float)kc0;
}

// prolongation() at (i, j, k) for a wavefront's row: i and j uniform, the k part from prolong_k
template <typename T>
__device__ __forceinline__ float prolong_at_row(const T *__restrict__ coarse, Dims nc, int i, int j, uint32_t pk0,
                                                uint32_t pk1, float tz) {
    const float x = clamp_((float)i * 0.5f - 0.25f, 0.0f, (float)(nc.n0 - 1));
    const float y = clamp_((float)j * 0.5f - 0.25f, 0.0f, (float)(nc.n1 - 1));
    const int ic0 = (int)x, jc0 = (int)y, ic1 = min(ic0 + 1, nc.n0 - 1), jc1 = min(jc0 + 1, nc.n1 - 1);
    const float tx = x - (float)ic0, ty = y - (float)jc0;
    constexpr uint32_t B = sizeof(T);
    const uint32_t r00 = (uint32_t)((ic0 * nc.n1 + jc0) * nc.n2) * B, r10 = (uint32_t)((ic1 * nc.n1 + jc0) * nc.n2) * B;
    const uint32_t r01 = (uint32_t)((ic0 * nc.n1 + jc1) * nc.n2) * B, r11 = (uint32_t)((ic1 * nc.n1 + jc1) * nc.n2) * B;
    const float c00 = lerp_(load_bytes(coarse, r00 + pk0), load_bytes(coarse, r10 + pk0), tx);
    const float c10 = lerp_(load_bytes(coarse, r01 + pk0), load_bytes(coarse, r11 + pk0), tx);
    const float c01 = lerp_(load_bytes(coarse, r00 + pk1), load_bytes(coarse, r10 + pk1), tx);
    const float c11 = lerp_(load_bytes(coarse, r01 + pk1), load_bytes(coarse, r11 + pk1), tx);
    return lerp_(lerp_(c00, c10, ty), lerp_(c01, c11, ty), tz);
}

// f(integral_constant<int, 0>) .. f(integral_constant<int, N - 1>), in order, each index a compile-time constant
template <typename F, int... U>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, U...>) {
    (f(std::integral_constant<int, U>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// t / d for 0 <= t < 2^32 / d: the multiply-high by m = ceil(2^32 / d) is exact there (its error t (m d - 2^32)
// / (d 2^32) stays below 1 / d). Tile counts and tail indices are far inside that range.
struct FastDiv {
    uint32_t d, m;
    __device__ __forceinline__ int operator()(int t) const { return d == 1 ? t : (int)__umulhi((uint32_t)t, m); }
};

FastDiv fast_div(int d) {
    FastDiv f;
    f.d = (uint32_t)d;
    f.m = d == 1 ? 0u : (uint32_t)(((uint64_t(1) << 32) + (uint64_t)d - 1) / (uint64_t)d);
    return f;
}

// Workgroups are handed to the 8 XCDs round-robin by linear id, each XCD with its own L2. Renumbering so
// that XCD x runs one contiguous range of tiles keeps a tile's halo neighbours in the same L2.
constexpr int XCDS = 8;
__device__ __forceinline__ int xcd_contiguous(int block, int blocks) {
    const int per = blocks / XCDS, rem = blocks % XCDS, x = block % XCDS;
    return x * per + min(x, rem) + block / XCDS;
}

// x / d, correctly rounded, where d is a cell's sum of couplings t1. For d = 4, 5 or 6, from r = RN(1 / d):
// q = RN(x r), then Markstein's correction q + RN(x - q d) r, both FMAs. Checked exhaustively against the
// IEEE quotient over every float x: equal for d = 4 and 5 except x = -0, and for d = 6 except |x| < 2^-125
// (subnormal quotients). Such x (and zeros), and lanes whose d is another value (`ieee`: a ghost coupling
// other than 1, a corner cell), take the IEEE division; lanes whose result is not used (`active` false) take
// neither. The test is one ballot per wavefront (a scalar value), so the common case saves no exec mask.
__device__ __forceinline__ float divide_flat(float x, float d, float r, bool ieee, bool active) {
    float q = x * r;
    const float e = __builtin_fmaf(-q, d, x);
    q = __builtin_fmaf(e, r, q);
    const bool slow = active && (ieee || __builtin_fabsf(x) < 0x1p-100f);
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(slow) != 0, 0)) {
        if (slow) q = x / d;
    }
    return q;
}

// The two-colour tile sweep on a flat level: fill_hi (1, 1, 1), so every coupling between two cells of the
// level and every cell volume is exactly 1, and a face on the level's boundary couples by `ghost` when open
// (1 on the simulation grid, 1 / (0.5 + ghost_gap) below it) or by 0 when closed. The same values as the
// general path (k_rb_sweep_tile below), with an update reduced to its arithmetic:
//   - neighbor_sum's t0 is the plain sum of the six neighbours in the reference's order x-, x+, y-, y+, z-,
//     z+ (1 q = q exactly; a neighbour outside the level is +0.0 in the tile, the +0.0 the reference adds
//     for a ghost or a closed face); its t1 = (((((0 + c_x-) + c_x+) + c_y-) + c_y+) + c_z-) + c_z+ is
//     (4 + c_z-) + c_z+ on a row whose four x and y neighbours lie in the level, a per-lane constant, and is
//     formed from the row's own couplings on the level's boundary rows;
//   - the source is cell_m² rhs, cell_m² rounded once, as (1 cell_m) cell_m; the division by t1 is
//     divide_flat;
//   - a wavefront updates two rows at once, one per half, of one class: the parity sigma of the cells' k - k0,
//     which fixes where a cell's k neighbours sit in the colour-split row. The pair is rows (2a, 2t + x) and
//     (2a + 1, 2t + 1 - x) of the colour's row block, so the second row is a fixed offset from the first that
//     depends on sigma alone: per-lane constants hold the half's offset and the cell's slot, and the seven
//     LDS reads of an update are one scalar row base plus one add, then constant offsets.
// EDGE false: the tile, its ring and its halo lie inside the level and the tile's 64 cells along k exist;
// loads and stores of a row go from the row's scalar base pointer. EDGE true: rows outside the level, and
// lanes past its last k, hold +0.0 in the tile and are never updated or stored. `edge` updates the ring
// rows' red cells in the k halo (k0 - 1 or k0 + 64) with the general update.
template <int TI, int TJ, int THREADS, int START, bool EDGE, typename T, typename EdgeUpdate>
__device__ __forceinline__ void sweep_tile_flat(const LevelArgs &L, float *__restrict__ s, const T *__restrict__ q_in,
                                                T *__restrict__ q_out, const T *__restrict__ rhs,
                                                const T *__restrict__ coarse, Dims nc, float relaxation, int i0,
                                                int j0, int k0, EdgeUpdate edge, const double *gate) {
    constexpr int WAVES = THREADS / 64;
    constexpr int LJ = TJ + 2 * TH, LI = TI + 2 * TH, HALF = LK / 2;
    constexpr uint32_t B = sizeof(T);
    // red: the ring (TI + 2) x (TJ + 2) rows; black: the tile's TI x TJ rows. Pairs per class = rows / 4.
    constexpr int RED_T = (TJ + 2) / 2, RED_CLASS = (TI + 2) / 2 * RED_T, RED_PAIRS = 2 * RED_CLASS;
    constexpr int BLACK_T = TJ / 2, BLACK_CLASS = TI / 2 * BLACK_T, BLACK_PAIRS = 2 * BLACK_CLASS;
    constexpr int RED_ITERS = (RED_PAIRS + WAVES - 1) / WAVES, BLACK_ITERS = (BLACK_PAIRS + WAVES - 1) / WAVES;
    static_assert(TI % 2 == 0 && TJ % 2 == 0 && WAVES <= LJ, "rows pair up within each class");
    const Dims n = cell_dims(L.g);
    const int n12 = n.n1 * n.n2;
    const int lane = threadIdx.x & 63, wave = uniform((int)(threadIdx.x >> 6));
    const int m = lane & 31;
    const bool upper = lane >= 32;
    const float cm2 = uniform(L.g.cell_m) * uniform(L.g.cell_m);
    const float ghost = uniform(L.ghost);
    bool open_lo[3], open_hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        open_lo[a] = uniform(L.g.open_lo[a]) == 1;
        open_hi[a] = uniform(L.g.open_hi[a]) == 1;
    }
    // per class sigma: the lane's cell k = k0 + 2 m + sigma, its z couplings, and t1, RN(1 / t1) and whether
    // t1 takes the IEEE division on a row whose four x and y neighbours lie in the level
    float cz_lo[2], cz_hi[2], t1[2], rcp[2];
    bool ieee[2], k_in[2];
    // per colour c (0 red, 1 black) and class sigma: the lane's LDS index relative to its pair's first row's
    // i-1 neighbour row, and its rhs byte offset relative to the first row's (i n1 + j) n2
    int lds_lane[2][2];
    uint32_t rhs_lane[2][2];
#pragma unroll
    for (int sigma = 0; sigma < 2; sigma++) {
        const int k = k0 + 2 * m + sigma;
        k_in[sigma] = !EDGE || k < n.n2;
        cz_lo[sigma] = k > 0 ? 1.0f : open_lo[2] ? ghost : 0.0f;
        cz_hi[sigma] = k < n.n2 - 1 ? 1.0f : open_hi[2] ? ghost : 0.0f;
        t1[sigma] = (4.0f + cz_lo[sigma]) + cz_hi[sigma];
        rcp[sigma] = 1.0f / t1[sigma];
        ieee[sigma] = !(t1[sigma] == 4.0f || t1[sigma] == 5.0f || t1[sigma] == 6.0f);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int step = c == 0 ? 1 - 2 * sigma : 2 * sigma - 1;  // the second row's j offset
            lds_lane[c][sigma] = (upper ? (LJ + step) * LK : 0) + sigma * HALF + m + 1;
            rhs_lane[c][sigma] = (uint32_t)((upper ? n12 + step * n.n2 : 0) + k) * B;
        }
    }
    const int g_tile = (i0 * n.n1 + j0) * n.n2;
    // pair p of colour c: its class, the LDS row index of its first row's i-1 neighbour row, and the first
    // row's tile-relative (ii, jj)
    auto decode = [&](int c, int pair, int &sigma, int &lds_row, int &ii, int &jj) {
        const int per_class = c == 0 ? RED_CLASS : BLACK_CLASS, per_a = c == 0 ? RED_T : BLACK_T;
        sigma = pair >= per_class;
        const int q = pair - sigma * per_class, a = q / per_a, t = q - a * per_a;
        // red rows (2a, 2t + sigma) of the ring (tile row - 1); black rows (2a, 2t + 1 - sigma) of the tile
        ii = c == 0 ? 2 * a - 1 : 2 * a;
        jj = c == 0 ? 2 * t + sigma - 1 : 2 * t + 1 - sigma;
        lds_row = (ii + TH - 1) * LJ + (jj + TH);
    };
    // the lane's row (i, j) in a pair whose first row is (ii, jj): the first row on the lower half, the second
    // on the upper
    auto lane_row = [&](int c, int sigma, int ii, int jj, int &i, int &j) {
        const int step = c == 0 ? 1 - 2 * sigma : 2 * sigma - 1;
        i = i0 + ii + (upper ? 1 : 0);
        j = j0 + jj + (upper ? step : 0);
    };
    auto row_inside = [&](int i, int j) { return (unsigned)i < (unsigned)n.n0 && (unsigned)j < (unsigned)n.n1; };
    auto ld = [](const T *base, uint32_t byte_off) {
        return to_f32(*reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off));
    };
    auto load_rhs = [&](int c, int pair, int &sigma, int &lds_row) {
        int ii, jj;
        decode(c, pair, sigma, lds_row, ii, jj);
        if (pair >= (c == 0 ? RED_PAIRS : BLACK_PAIRS)) return 0.0f;
        const uint32_t off = (uint32_t)(g_tile + ii * n12 + jj * n.n2) * B + (sigma ? rhs_lane[c][1] : rhs_lane[c][0]);
        if (!EDGE) return ld(rhs, off);
        // on an EDGE tile the first row's offset may lie outside the level; only an inside cell's is used
        int i, j;
        lane_row(c, sigma, ii, jj, i, j);
        float value = 0.0f;
        if (row_inside(i, j) && (sigma ? k_in[1] : k_in[0])) value = ld(rhs, off);
        return value;
    };
    float rhs_red[RED_ITERS], rhs_black[BLACK_ITERS];
    int red_sigma[RED_ITERS], red_row[RED_ITERS], black_sigma[BLACK_ITERS], black_row[BLACK_ITERS];
#pragma unroll
    for (int u = 0; u < RED_ITERS; u++) rhs_red[u] = load_rhs(0, wave + u * WAVES, red_sigma[u], red_row[u]);
#pragma unroll
    for (int u = 0; u < BLACK_ITERS; u++)
        rhs_black[u] = load_rhs(1, wave + u * WAVES, black_sigma[u], black_row[u]);
    // The tile's start values, row by row: row r = wave + u WAVES of the (TI + 4) x (TJ + 4) rows, its offset
    // carried from row to row in scalar registers; the k halo (4 cells a row) by a second loop. All loads of a
    // batch are in flight before the first LDS write (one batch of every row, unless the start adds the
    // prolongation).
    const int own_slot = ((lane + TH) & 1) * HALF + ((lane + TH) >> 1);
    const uint32_t lane_b = (uint32_t)(k0 + lane) * B;
    const bool k_load = !EDGE || k0 + lane < n.n2;
    // FROM_PROLONG: the prolongation's k part per lane (the same for every row), its i and j part per row; a
    // lane past the last k takes the last k's (its value is not used)
    uint32_t pk0 = 0, pk1 = 0;
    float tz = 0.0f;
    if (START == FROM_PROLONG) prolong_k<T>(nc, min(k0 + lane, n.n2 - 1), pk0, pk1, tz);
    auto prolong_row = [&](int i, int j) { return prolong_at_row(coarse, nc, i, j, pk0, pk1, tz); };
    int ii = 0, jj = wave, g = g_tile - TH * n12 + (wave - TH) * n.n2;  // row (ii, jj) at (i0 - 2 + ii, j0 - 2 + jj)
    constexpr int ROWS = LI * LJ / WAVES;
    constexpr int BATCH = START == FROM_PROLONG ? (ROWS % 3 == 0 ? 3 : 2) : ROWS;
    static_assert((LI * LJ) % WAVES == 0 && ROWS % BATCH == 0, "halo rows split evenly over wavefronts");
    constexpr int HALO = LI * LJ * 4, HALO_ITERS = (HALO + THREADS - 1) / THREADS;
    float halo[HALO_ITERS];
    auto load_halo = [&]() {
#pragma unroll
        for (int u = 0; u < HALO_ITERS; u++) {
            const int t = min((int)threadIdx.x + u * THREADS, HALO - 1), row = t >> 2, h = t & 3;
            const int i = i0 - TH + row / LJ, j = j0 - TH + row % LJ;
            const int k = h < 2 ? k0 - TH + h : k0 + TK - 2 + h;
            float value = 0.0f;
            if (START != FROM_ZERO && (!EDGE || row_inside(i, j)) && k >= 0 && k < n.n2) {
                value = ld(q_in, (uint32_t)((i * n.n1 + j) * n.n2 + k) * B);
                if (START == FROM_PROLONG) value = value + prolongation(coarse, nc, i, j, k);
            }
            halo[u] = value;
        }
    };
    if (START != FROM_PROLONG) load_halo();
    for (int b = 0; b < ROWS; b += BATCH) {
        float v[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; u++) {
            v[u] = 0.0f;
            const int i = i0 - TH + ii, j = j0 - TH + jj;
            if (START != FROM_ZERO && (!EDGE || (row_inside(i, j) && k_load))) {
                v[u] = ld(q_in, (uint32_t)g * B + lane_b);
                if (START == FROM_PROLONG) v[u] = v[u] + prolong_row(i, j);
            }
            // next row: WAVES rows on, wrapping into the next i row at most once (WAVES <= LJ)
            jj += WAVES;
            const bool wrap = jj >= LJ;
            jj -= wrap ? LJ : 0;
            ii += wrap ? 1 : 0;
            g += wrap ? n12 + (WAVES - LJ) * n.n2 : WAVES * n.n2;
        }
        // the device-side gate is read once the first loads are in flight (they are in range whatever it says),
        // so its latency overlaps theirs; a gated-off block leaves before any write or barrier
        if (b == 0 && gated_off(gate)) return;
#pragma unroll
        for (int u = 0; u < BATCH; u++) s[(wave + (b + u) * WAVES) * LK + own_slot] = v[u];
    }
    if (START == FROM_PROLONG) load_halo();
#pragma unroll
    for (int u = 0; u < HALO_ITERS; u++) {
        const int t = threadIdx.x + u * THREADS, row = t >> 2, h = t & 3;
        const int kk = h < 2 ? h : TK + h;
        if (t < HALO) s[row * LK + (kk & 1) * HALF + (kk >> 1)] = halo[u];
    }
    __syncthreads();
    // One update: red_black_sweep_kernel's arithmetic (sim/pressure.py:86-92), the neighbours added in the
    // reference's order x-, x+, y-, y+, z-, z+. `base` indexes the cell's i-1 neighbour; an empty asm keeps it
    // the one address register that every read offsets from. `write` false leaves the cell as it is.
    auto update = [&](auto sigma_c, int base, float rhs_c, float t1_c, float rcp_c, bool ieee_c, bool write) {
        constexpr int SIGMA = decltype(sigma_c)::value;
        asm volatile("" : "+v"(base));
        float *b = s + base;
        const float q_im = b[0], q_jm = b[(LJ - 1) * LK], q_c = b[LJ * LK];
        const float q_jp = b[(LJ + 1) * LK], q_ip = b[2 * LJ * LK];
        const float q_km = b[LJ * LK + (SIGMA == 0 ? HALF - 1 : -HALF)];
        const float q_kp = b[LJ * LK + (SIGMA == 0 ? HALF : 1 - HALF)];
        float t0 = 0.0f + q_im;
        t0 = t0 + q_ip;
        t0 = t0 + q_jm;
        t0 = t0 + q_jp;
        t0 = t0 + q_km;
        t0 = t0 + q_kp;
        const float target = divide_flat(t0 - cm2 * rhs_c, t1_c, rcp_c, ieee_c, !EDGE || write);
        const float next = q_c + relaxation * (target - q_c);
        if (!EDGE || write) b[LJ * LK] = next;
    };
    // One pair of colour c, class SIGMA: on an EDGE tile a lane whose row lies outside the level (or whose k
    // does) writes nothing, and a wavefront holding a row on the level's boundary forms that row's t1 from
    // its couplings (a t1 of 0, a cell coupled to nothing, leaves q as it is).
    auto pair_update = [&](auto sigma_c, int c, int pair, int lds_row, float rhs_c) {
        constexpr int SIGMA = decltype(sigma_c)::value;
        float t = t1[SIGMA], r = rcp[SIGMA];
        bool slow = ieee[SIGMA], write = true;
        if constexpr (EDGE) {
            int sg, lr, ii, jj, i, j;
            decode(c, pair, sg, lr, ii, jj);
            lane_row(c, SIGMA, ii, jj, i, j);
            write = row_inside(i, j) && k_in[SIGMA];
            const bool interior = i > 0 && i < n.n0 - 1 && j > 0 && j < n.n1 - 1;
            if (__builtin_amdgcn_ballot_w64(write && !interior) != 0) {
                if (!interior) {
                    const float cx_lo = i > 0 ? 1.0f : open_lo[0] ? ghost : 0.0f;
                    const float cx_hi = i < n.n0 - 1 ? 1.0f : open_hi[0] ? ghost : 0.0f;
                    const float cy_lo = j > 0 ? 1.0f : open_lo[1] ? ghost : 0.0f;
                    const float cy_hi = j < n.n1 - 1 ? 1.0f : open_hi[1] ? ghost : 0.0f;
                    float sum = 0.0f + cx_lo;
                    sum = sum + cx_hi;
                    sum = sum + cy_lo;
                    sum = sum + cy_hi;
                    sum = sum + cz_lo[SIGMA];
                    t = sum + cz_hi[SIGMA];
                    r = t == 4.0f ? 0.25f : t == 5.0f ? 0.2f : 1.0f / 6.0f;
                    slow = !(t == 4.0f || t == 5.0f || t == 6.0f);
                    write = write && t != 0.0f;
                }
            }
        }
        update(sigma_c, lds_row * LK + lds_lane[c][SIGMA], rhs_c, t, r, slow, write);
    };
    // Iteration u of a colour takes pairs u WAVES .. u WAVES + WAVES - 1 (one per wavefront), and pairs are
    // numbered class 0 first: an iteration whose pairs all lie in one class, or all exist, knows it at compile
    // time, so only the iterations that straddle a class or the end test at run time.
    static_for<RED_ITERS>([&](auto u_c) {
        constexpr int U = decltype(u_c)::value, FIRST = U * WAVES, LAST = FIRST + WAVES - 1;
        if (LAST >= RED_PAIRS && wave + FIRST >= RED_PAIRS) return;
        const int pair = wave + FIRST;
        if constexpr (LAST < RED_CLASS) pair_update(std::integral_constant<int, 0>{}, 0, pair, red_row[U], rhs_red[U]);
        else if constexpr (FIRST >= RED_CLASS)
            pair_update(std::integral_constant<int, 1>{}, 0, pair, red_row[U], rhs_red[U]);
        else if (red_sigma[U] == 0) pair_update(std::integral_constant<int, 0>{}, 0, pair, red_row[U], rhs_red[U]);
        else pair_update(std::integral_constant<int, 1>{}, 0, pair, red_row[U], rhs_red[U]);
    });
    edge();
    __syncthreads();
    static_for<BLACK_ITERS>([&](auto u_c) {
        constexpr int U = decltype(u_c)::value, FIRST = U * WAVES, LAST = FIRST + WAVES - 1;
        if (LAST >= BLACK_PAIRS && wave + FIRST >= BLACK_PAIRS) return;
        const int pair = wave + FIRST;
        if constexpr (LAST < BLACK_CLASS)
            pair_update(std::integral_constant<int, 0>{}, 1, pair, black_row[U], rhs_black[U]);
        else if constexpr (FIRST >= BLACK_CLASS)
            pair_update(std::integral_constant<int, 1>{}, 1, pair, black_row[U], rhs_black[U]);
        else if (black_sigma[U] == 0) pair_update(std::integral_constant<int, 0>{}, 1, pair, black_row[U], rhs_black[U]);
        else pair_update(std::integral_constant<int, 1>{}, 1, pair, black_row[U], rhs_black[U]);
    });
    __syncthreads();
    constexpr int OUT_ROWS = TI * TJ / WAVES;
    float v[OUT_ROWS];
#pragma unroll
    for (int u = 0; u < OUT_ROWS; u++) {
        const int row = wave + u * WAVES;
        v[u] = s[((row / TJ + TH) * LJ + (row % TJ + TH)) * LK + own_slot];
    }
#pragma unroll
    for (int u = 0; u < OUT_ROWS; u++) {
        const int row = wave + u * WAVES;
        if (EDGE && !(i0 + row / TJ < n.n0 && j0 + row % TJ < n.n1 && k_load)) continue;
        const uint32_t off = (uint32_t)(g_tile + (row / TJ) * n12 + (row % TJ) * n.n2) * B + lane_b;
        *reinterpret_cast<T *>(reinterpret_cast<char *>(q_out) + off) = from_f32<T>(v[u]);
    }
}

// One two-colour sweep over a TI x TJ x 64 tile (i, j, k) with THREADS threads, row by row: a row is the
// tile's 64 cells along k plus a 2-cell halo each side, 68 floats in LDS stored colour-split (even k - k0 + 2
// in the first 34 slots, odd in the last 34), so the cells of one colour and their k neighbours sit in
// consecutive LDS words. LDS holds (TI + 4) x (TJ + 4) rows. Loads: the 64 interior cells of a row are one
// aligned 256-byte access by one wavefront, the 4 halo cells a small second loop. Red runs on the tile and
// a 1-cell ring (the ring's new values are what black needs): 32 cells per row on half a wavefront, plus the
// one red cell of each ring row that lies in the k halo; black runs on the tile's rows; then the rows are
// written out. The rhs values both passes need are loaded into registers before the first barrier. T is the
// stored type of q, rhs and the coarse correction (FP32, or BF16 / FP16 in the mixed-precision solve); the
// tile and all arithmetic are FP32.
// The tile sweep of tile number `tile` (k fastest, then j, then i) with the block's LDS: s holds (TI + 4) x
// (TJ + 4) rows of LK floats, cps the coupling table (levels with couplings other than 1).
template <int TI, int TJ, int THREADS, bool UNIT, int START, typename T>
__device__ __forceinline__ void sweep_tile(const LevelArgs &L, const T *__restrict__ q_in, T *__restrict__ q_out,
                                           const T *__restrict__ rhs, const T *__restrict__ coarse, Dims nc,
                                           float relaxation, int tile, FastDiv tiles_k, FastDiv tiles_j, float *s,
                                           float *cps, const double *gate = nullptr) {
    constexpr int WAVES = THREADS / 64;
    constexpr int LJ = TJ + 2 * TH, LI = TI + 2 * TH, HALF = LK / 2;
    constexpr int RI = TI + 2, RJ = TJ + 2;
    constexpr int RED_PAIRS = (RI * RJ + 1) / 2, RED_ITERS = (RED_PAIRS + WAVES - 1) / WAVES;
    constexpr int BLACK_ITERS = TI * TJ / 2 / WAVES;
    static_assert(THREADS % 64 == 0 && (TI * TJ) % (2 * WAVES) == 0, "tile rows split evenly over wavefronts");
    static_assert(RI * RJ <= THREADS, "one thread per ring row for the k-halo red cells");
    if (!UNIT && threadIdx.x == 0) {
        // constant indices: read straight from the kernel arguments
        const float *table = &L.cp.in[0][0][0];
#pragma unroll
        for (int m = 0; m < 24; m++) cps[m] = table[m];
        const float *ghost = &L.cp.ghost[0][0][0];
#pragma unroll
        for (int m = 0; m < 24; m++) cps[CP_GHOST + m] = ghost[m];
        cps[CP_ZERO] = 0.0f;
    }
    const Dims n = cell_dims(L.g);
    const int tr = tiles_k(tile), tk = tile - tr * (int)tiles_k.d, ti = tiles_j(tr), tj = tr - ti * (int)tiles_j.d;
    const int k0 = tk * TK, j0 = tj * TJ, i0 = ti * TI;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int half_lane = lane & 31, upper = lane >> 5;
    // slot of k in a row: kk = k - k0 + 2 in [0, 68), even kk first
    auto slot = [&](int k) {
        const int kk = k - k0 + TH;
        return (kk & 1) * HALF + (kk >> 1);
    };
    auto at = [&](int i, int j, int k) -> float & {
        return s[((i - i0 + TH) * LJ + (j - j0 + TH)) * LK + slot(k)];
    };
    auto inside = [&](int i, int j, int k) {
        return i >= 0 && i < n.n0 && j >= 0 && j < n.n1 && k >= 0 && k < n.n2;
    };
    // red cell of ring row `row` (ii, jj over RI x RJ) at k0 + 2 m + ((i + j) & 1): (i + j + k) even
    auto red_main = [&](int row, int m, int &i, int &j, int &k) {
        i = i0 - 1 + row / RJ;
        j = j0 - 1 + row % RJ;
        k = k0 + 2 * m + ((i + j) & 1);
        return row < RI * RJ && inside(i, j, k);
    };
    // the red cell of a ring row in the k halo: k0 - 1 when i + j is odd, else k0 + 64
    auto red_edge = [&](int row, int &i, int &j, int &k) {
        i = i0 - 1 + row / RJ;
        j = j0 - 1 + row % RJ;
        k = ((i + j) & 1) ? k0 - 1 : k0 + TK;
        return row < RI * RJ && inside(i, j, k);
    };
    auto black_cell = [&](int row, int m, int &i, int &j, int &k) {
        i = i0 + row / TJ;
        j = j0 + row % TJ;
        k = k0 + 2 * m + ((i + j + 1) & 1);
        return i < n.n0 && j < n.n1 && k < n.n2;
    };
    auto update = [&](int i, int j, int k, float rhs_c) {
        float t0, t1;
        if (UNIT)
            neighbor_sum_padded<true>(L, i, j, k, [&](int a, int b, int c) { return at(a, b, c); }, t0, t1);
        else
            neighbor_sum_lds(L, cps, i, j, k, [&](int a, int b, int c) { return at(a, b, c); }, t0, t1);
        float &cell = at(i, j, k);
        cell = gs_update(L, i, j, k, cell, rhs_c, t0, t1, relaxation);
    };
    if (UNIT || L.flat == 1) {
        auto edge_cells = [&]() {
            int i, j, k;
            if (red_edge(threadIdx.x, i, j, k)) update(i, j, k, to_f32(rhs[n.at(i, j, k)]));
        };
        if (i0 >= TH && i0 + TI + TH <= n.n0 && j0 >= TH && j0 + TJ + TH <= n.n1 && k0 + TK <= n.n2)
            sweep_tile_flat<TI, TJ, THREADS, START, false, T>(L, s, q_in, q_out, rhs, coarse, nc, relaxation, i0, j0,
                                                              k0, edge_cells, gate);
        else
            sweep_tile_flat<TI, TJ, THREADS, START, true, T>(L, s, q_in, q_out, rhs, coarse, nc, relaxation, i0, j0,
                                                             k0, edge_cells, gate);
        return;
    }
    if (gated_off(gate)) return;
    float rhs_red[RED_ITERS], rhs_black[BLACK_ITERS], rhs_edge = 0.0f;
#pragma unroll
    for (int u = 0; u < RED_ITERS; u++) {
        int i, j, k;
        rhs_red[u] = red_main(2 * (wave + u * WAVES) + upper, half_lane, i, j, k) ? to_f32(rhs[n.at(i, j, k)]) : 0.0f;
    }
    {
        int i, j, k;
        if (red_edge(threadIdx.x, i, j, k)) rhs_edge = to_f32(rhs[n.at(i, j, k)]);
    }
#pragma unroll
    for (int u = 0; u < BLACK_ITERS; u++) {
        int i, j, k;
        rhs_black[u] = black_cell(2 * (wave + u * WAVES) + upper, half_lane, i, j, k) ? to_f32(rhs[n.at(i, j, k)]) : 0.0f;
    }
    // The tile's start values. Reads outside the level use a clamped address and are replaced by 0, so the
    // loads carry no branches; rows go in batches whose loads are all in flight before the LDS writes.
    auto start_value = [&](int i, int j, int k) {
        if (START == FROM_ZERO) return 0.0f;
        const bool in = inside(i, j, k);
        const int ci = min(max(i, 0), n.n0 - 1), cj = min(max(j, 0), n.n1 - 1), ck = min(max(k, 0), n.n2 - 1);
        float value = to_f32(q_in[n.at(ci, cj, ck)]);
        if (START == FROM_PROLONG) value = value + prolongation(coarse, nc, ci, cj, ck);
        return in ? value : 0.0f;
    };
    // rows (k = k0 + lane): a FROM_PROLONG start takes the prolongation's k part per lane at the clamped k
    uint32_t pk0 = 0, pk1 = 0;
    float tz = 0.0f;
    if (START == FROM_PROLONG) prolong_k<T>(nc, min(max(k0 + lane, 0), n.n2 - 1), pk0, pk1, tz);
    auto start_row = [&](int i, int j) {
        if (START != FROM_PROLONG) return start_value(i, j, k0 + lane);
        const int k = k0 + lane;
        const bool in = inside(i, j, k);
        const int ci = min(max(i, 0), n.n0 - 1), cj = min(max(j, 0), n.n1 - 1), ck = min(max(k, 0), n.n2 - 1);
        const float value = to_f32(q_in[n.at(ci, cj, ck)]) + prolong_at_row(coarse, nc, ci, cj, pk0, pk1, tz);
        return in ? value : 0.0f;
    };
    constexpr int ROWS = LI * LJ / WAVES, BATCH = ROWS % 6 == 0 ? 6 : START == FROM_PROLONG ? 4 : 8;
    static_assert((LI * LJ) % WAVES == 0 && ROWS % BATCH == 0, "halo rows split evenly over wavefronts");
    for (int b = 0; b < ROWS; b += BATCH) {
        float v[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; u++) {
            const int row = wave + (b + u) * WAVES;
            v[u] = start_row(i0 - TH + row / LJ, j0 - TH + row % LJ);
        }
#pragma unroll
        for (int u = 0; u < BATCH; u++) s[(wave + (b + u) * WAVES) * LK + slot(k0 + lane)] = v[u];
    }
    {
        constexpr int HALO = LI * LJ * 4, HALO_ITERS = (HALO + THREADS - 1) / THREADS;
        float v[HALO_ITERS];
#pragma unroll
        for (int u = 0; u < HALO_ITERS; u++) {
            const int t = min((int)threadIdx.x + u * THREADS, HALO - 1), row = t >> 2, h = t & 3;
            const int k = h < 2 ? k0 - TH + h : k0 + TK - 2 + h;
            v[u] = start_value(i0 - TH + row / LJ, j0 - TH + row % LJ, k);
        }
#pragma unroll
        for (int u = 0; u < HALO_ITERS; u++) {
            const int t = threadIdx.x + u * THREADS, row = t >> 2, h = t & 3;
            if (t < HALO) s[row * LK + slot(h < 2 ? k0 - TH + h : k0 + TK - 2 + h)] = v[u];
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < RED_ITERS; u++) {
        int i, j, k;
        if (red_main(2 * (wave + u * WAVES) + upper, half_lane, i, j, k)) update(i, j, k, rhs_red[u]);
    }
    {
        int i, j, k;
        if (red_edge(threadIdx.x, i, j, k)) update(i, j, k, rhs_edge);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < BLACK_ITERS; u++) {
        int i, j, k;
        if (black_cell(2 * (wave + u * WAVES) + upper, half_lane, i, j, k)) update(i, j, k, rhs_black[u]);
    }
    __syncthreads();
    constexpr int OUT_ROWS = TI * TJ / WAVES;
    float v[OUT_ROWS];
#pragma unroll
    for (int u = 0; u < OUT_ROWS; u++) {
        const int row = wave + u * WAVES;
        v[u] = s[(row / TJ + TH) * LJ * LK + (row % TJ + TH) * LK + slot(k0 + lane)];
    }
#pragma unroll
    for (int u = 0; u < OUT_ROWS; u++) {
        const int row = wave + u * WAVES;
        const int i = i0 + row / TJ, j = j0 + row % TJ, k = k0 + lane;
        if (i < n.n0 && j < n.n1 && k < n.n2) q_out[n.at(i, j, k)] = from_f32<T>(v[u]);
    }
}

// One two-colour sweep over a TI x TJ x 64 tile per block (sweep_tile), tiles renumbered so that each XCD runs
// one contiguous range of them.
template <int TI, int TJ, int THREADS, bool UNIT, int START, typename T = float>
__global__ __launch_bounds__(THREADS) void k_rb_sweep_tile(LevelArgs L, const T *__restrict__ q_in,
                                                           T *__restrict__ q_out, const T *__restrict__ rhs,
                                                           const T *__restrict__ coarse, Dims nc,
                                                           float relaxation, const double *gate, FastDiv tiles_k,
                                                           FastDiv tiles_j) {
    __shared__ float s[(TI + 2 * TH) * (TJ + 2 * TH) * LK];
    __shared__ float cps[UNIT ? 1 : CP_WORDS];
    sweep_tile<TI, TJ, THREADS, UNIT, START, T>(L, q_in, q_out, rhs, coarse, nc, relaxation,
                                                xcd_contiguous(blockIdx.x, gridDim.x), tiles_k, tiles_j, s, cps, gate);
}

// The tile shapes the autotuner chooses among: TI, TJ (k is 64 wide), threads per block.
struct TileShape {
    int ti, tj, threads;
};
constexpr TileShape TILE_SHAPES[] = {{4, 4, 256}, {2, 8, 256}, {8, 4, 256}, {4, 8, 512}, {8, 8, 512}, {2, 4, 256}};
constexpr int N_TILE_SHAPES = 6;
constexpr int DEFAULT_TILE_SHAPE = 0;

// One launch of a tile shape. With timing events (a bound TimingScope) the launch goes through
// hipExtLaunchKernelGGL so that the events take this dispatch's own begin and end timestamps.
template <int TI, int TJ, int THREADS, bool UNIT, int START, typename T>
void launch_tile_kernel(dim3 grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const LevelArgs &L, const T *src,
                        T *dst, const T *rhs, const T *coarse, Dims nc, float relaxation, const double *gate,
                        FastDiv tiles_k, FastDiv tiles_j) {
    if (e0 != nullptr)
        hipExtLaunchKernelGGL(k_rb_sweep_tile<TI, TJ, THREADS, UNIT, START, T>, grid, dim3(THREADS), 0, s, e0, e1, 0u,
                              L, src, dst, rhs, coarse, nc, relaxation, gate, tiles_k, tiles_j);
    else
        k_rb_sweep_tile<TI, TJ, THREADS, UNIT, START, T><<<grid, THREADS, 0, s>>>(L, src, dst, rhs, coarse, nc,
                                                                                relaxation, gate, tiles_k, tiles_j);
}

template <bool UNIT, int START, typename T>
void launch_tile_shape(int shape, dim3 grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const LevelArgs &L,
                       const T *src, T *dst, const T *rhs, const T *coarse, Dims nc, float relaxation,
                       const double *gate, FastDiv tiles_k, FastDiv tiles_j) {
#define PLM_TILE(TI, TJ, THREADS)                                                                                    \
    launch_tile_kernel<TI, TJ, THREADS, UNIT, START, T>(grid, s, e0, e1, L, src, dst, rhs, coarse, nc, relaxation,  \
                                                        gate, tiles_k, tiles_j)
    switch (shape) {
    case 1: PLM_TILE(2, 8, 256); break;
    case 2: PLM_TILE(8, 4, 256); break;
    case 3: PLM_TILE(4, 8, 512); break;
    case 4: PLM_TILE(8, 8, 512); break;
    case 5: PLM_TILE(2, 4, 256); break;
    default: PLM_TILE(4, 4, 256); break;
    }
#undef PLM_TILE
}

// The shape pinned for a level size (by plm_autotune_sweep or plm_set_sweep_tile), else the default.
struct TileChoices {
    std::mutex mutex;
    std::map<std::array<int, 3>, int> shape;
};
TileChoices &tile_choices() {
    static TileChoices c;
    return c;
}

int tile_shape_for(const Dims &n) {
    TileChoices &c = tile_choices();
    std::lock_guard<std::mutex> lock(c.mutex);
    auto it = c.shape.find({n.n0, n.n1, n.n2});
    return it == c.shape.end() ? DEFAULT_TILE_SHAPE : it->second;
}

// ------------------------------------------------------------------ gradient subtraction

// subtract_gradient_kernel, sim/pressure.py:95-116, one thread per GRADIENT_ROWS consecutive cells along j: the
// thread updates the three faces on each cell's low side (u[i,j,k], v[i,j,k], w[i,j,k]) and, on the last cell
// along an axis, that axis's face on the high boundary, so q is read once, coalesced (the lower neighbours from
// L1/L2, the j neighbour from the thread's own row below). Each face's update is the face kernel's:
// f -= (q_above - q_below) / cell_m, a ghost q of 0 outside an open face, closed boundary faces untouched.
// Every load is issued before the first store (at clamped addresses whose values a lane may not use).
constexpr int GRADIENT_ROWS = 4;
__global__ __launch_bounds__(256) void k_subtract_gradient(plm_grid_info g, Tri<float> f,
                                                           const float *__restrict__ q) {
    constexpr int R = GRADIENT_ROWS;
    const Dims n = cell_dims(g);
    const int k = blockIdx.x * BX + threadIdx.x, i = blockIdx.z;
    const int j0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.y * BY + threadIdx.y) * R));
    if (k >= n.n2 || j0 >= n.n1) return;
    const Dims nu = face_dims(g, 0), nv = face_dims(g, 1), nw = face_dims(g, 2);
    // q rows j0 - 1 .. j0 + R - 1 (clamped), and the i and k neighbours of each row
    float q_row[R + 1], q_im[R], q_km[R], u_lo[R], u_hi[R], v_lo[R], v_hi[R], w_lo[R], w_hi[R];
#pragma unroll
    for (int m = 0; m <= R; m++) q_row[m] = q[n.at(i, min(max(j0 - 1 + m, 0), n.n1 - 1), k)];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = min(j0 + r, n.n1 - 1);
        q_im[r] = q[n.at(max(i - 1, 0), j, k)];
        q_km[r] = q[n.at(i, j, max(k - 1, 0))];
        // a high face is read only by the cell it belongs to (a boundary cell); others re-read their low face
        const bool hx = i == n.n0 - 1, hy = j == n.n1 - 1, hz = k == n.n2 - 1;
        u_lo[r] = f.p[0][nu.at(i, j, k)];
        u_hi[r] = f.p[0][nu.at(hx ? i + 1 : i, j, k)];
        v_lo[r] = f.p[1][nv.at(i, j, k)];
        v_hi[r] = f.p[1][nv.at(i, hy ? j + 1 : j, k)];
        w_lo[r] = f.p[2][nw.at(i, j, k)];
        w_hi[r] = f.p[2][nw.at(i, j, hz ? k + 1 : k)];
    }
    const float cell_m = g.cell_m;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = j0 + r;
        if (j >= n.n1) break;
        const float q_cell = q_row[r + 1];
        const int cell[3] = {i, j, k};
        const float lower[3] = {q_im[r], q_row[r], q_km[r]};
        const float low_face[3] = {u_lo[r], v_lo[r], w_lo[r]}, high_face[3] = {u_hi[r], v_hi[r], w_hi[r]};
        float *low_at[3] = {f.p[0] + nu.at(i, j, k), f.p[1] + nv.at(i, j, k), f.p[2] + nw.at(i, j, k)};
        float *high_at[3] = {f.p[0] + nu.at(i + 1, j, k), f.p[1] + nv.at(i, j + 1, k), f.p[2] + nw.at(i, j, k + 1)};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (cell[a] > 0 || g.open_lo[a] == 1) {
                const float q_below = cell[a] > 0 ? lower[a] : 0.0f;
                *low_at[a] = low_face[a] - (q_cell - q_below) / cell_m;
            }
            if (cell[a] == g.dims[a] - 1 && g.open_hi[a] == 1) *high_at[a] = high_face[a] - (0.0f - q_cell) / cell_m;
        }
    }
}

inline dim3 gradient_grid(const plm_grid_info &g) {
    const Dims n = cell_dims(g);
    return dim3((n.n2 + BX - 1) / BX, (n.n1 + BY * GRADIENT_ROWS - 1) / (BY * GRADIENT_ROWS), n.n0);
}

// ------------------------------------------------------------------ multigrid transfers

template <bool UNIT>
__global__ __launch_bounds__(256) void k_residual(LevelArgs L, const float *__restrict__ q,
                                                  const float *__restrict__ rhs, float *__restrict__ r) {
    const Dims n = cell_dims(L.g);
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    float t0, t1;
    neighbor_sum<UNIT>(L, i, j, k, GlobalQ{q, n}, t0, t1);
    const int c = n.at(i, j, k);
    r[c] = residual_value(L, i, j, k, q[c], rhs[c], t0, t1);
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
// coarse_q (may be NULL) is set to 0, the reference's coarse.q.zero_() (sim/multigrid.py:143).
template <bool UNIT, typename T>
__device__ __forceinline__ void residual_restrict_cell(const LevelArgs &L, const T *__restrict__ q,
                                                       const T *__restrict__ rhs, T *__restrict__ coarse_rhs,
                                                       T *__restrict__ coarse_q, Dims nc, dim3 block,
                                                       const double *gate = nullptr) {
    const Dims nf = cell_dims(L.g);
    const int k = block.x * BX + threadIdx.x, j = block.y * BY + threadIdx.y, i = block.z;
    if (!(k < nc.n2 && j < nc.n1 && i < nc.n0)) return;
    float total = 0.0f;
    int count = 0;
    if constexpr (UNIT) {
        // Unit-coupling level: the children's 4 x 4 x 4 neighbourhood (fine coordinates 2i - 1 ... 2i + 2 per
        // axis) read at indices clamped into the level and masked (+0.0 outside, what neighbor_sum adds for a
        // ghost or a closed face); a child past the level (an overhang child) adds +0.0 to a total that starts
        // at +0.0 and is not counted. No load sits in a branch and no wavefront runs two paths; the same
        // values as neighbor_sum + residual_value over the in-range children.
        constexpr uint32_t B = sizeof(T);
        const uint32_t si = (uint32_t)(nf.n1 * nf.n2) * B, sj = (uint32_t)nf.n2 * B;
        uint32_t oi[4], oj[4], ok[4];
        bool in_i[4], in_j[4], in_k[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const int x = 2 * i - 1 + m, y = 2 * j - 1 + m, z = 2 * k - 1 + m;
            in_i[m] = x >= 0 && x < nf.n0;
            in_j[m] = y >= 0 && y < nf.n1;
            in_k[m] = z >= 0 && z < nf.n2;
            oi[m] = (uint32_t)max(0, min(x, nf.n0 - 1)) * si;
            oj[m] = (uint32_t)max(0, min(y, nf.n1 - 1)) * sj;
            ok[m] = (uint32_t)max(0, min(z, nf.n2 - 1)) * B;
        }
        auto Q = [&](int a, int b, int c) { return load_bytes(q, oi[a] + oj[b] + ok[c]); };
        const float scale = uniform(L.g.cell_m) * uniform(L.g.cell_m);  // (1 cell_m) cell_m
#pragma unroll
        for (int a = 1; a <= 2; a++)
#pragma unroll
            for (int b = 1; b <= 2; b++)
#pragma unroll
                for (int c = 1; c <= 2; c++) {
                    const float qn[6] = {Q(a - 1, b, c), Q(a + 1, b, c), Q(a, b - 1, c),
                                         Q(a, b + 1, c), Q(a, b, c - 1), Q(a, b, c + 1)};
                    const bool in[6] = {in_i[a - 1], in_i[a + 1], in_j[b - 1], in_j[b + 1], in_k[c - 1], in_k[c + 1]};
                    const int couples = unit_couples(L.g, 0, in[0], in[1]) + unit_couples(L.g, 1, in[2], in[3]) +
                                        unit_couples(L.g, 2, in[4], in[5]);
                    const float r = unit_residual(Q(a, b, c), load_bytes(rhs, oi[a] + oj[b] + ok[c]), scale, qn, in,
                                                  couples);
                    const bool exists = in_i[a] && in_j[b] && in_k[c];
                    total = total + (exists ? r : 0.0f);
                    count += exists ? 1 : 0;
                }
    } else {
        // Interior coarse cells: every child's six neighbours lie inside the level and none is the last cell
        // on an axis, so every coupling is 1, the sum of couplings 6 and the volume 1; the 32 distinct q
        // values of the children and their neighbours are loaded once. The rest take the general update.
        if (i >= 1 && 2 * i + 4 <= nf.n0 && j >= 1 && 2 * j + 4 <= nf.n1 && k >= 1 && 2 * k + 4 <= nf.n2) {
            constexpr uint32_t B = sizeof(T);
            const uint32_t si = (uint32_t)(nf.n1 * nf.n2) * B, sj = (uint32_t)nf.n2 * B;
            const uint32_t o = offset_of(nf, 2 * i, 2 * j, 2 * k) * B;
            // q at child-block coordinates (a, b, c) in [-1, 2]^3 (only the block and its face neighbours)
            auto Q = [&](int a, int b, int c) { return load_bytes(q, o + a * si + b * sj + c * B); };
            const float scale = uniform(L.g.cell_m) * uniform(L.g.cell_m);  // (1 cell_m) cell_m
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++)
#pragma unroll
                    for (int c = 0; c < 2; c++) {
                        float t0 = 0.0f + Q(a - 1, b, c);
                        t0 = t0 + Q(a + 1, b, c);
                        t0 = t0 + Q(a, b - 1, c);
                        t0 = t0 + Q(a, b + 1, c);
                        t0 = t0 + Q(a, b, c - 1);
                        t0 = t0 + Q(a, b, c + 1);
                        const float q_c = Q(a, b, c), rhs_c = load_bytes(rhs, o + a * si + b * sj + c * B);
                        total = total + (rhs_c - (t0 - 6.0f * q_c) / scale);
                        count += 1;
                    }
        } else {
            for (int di = 0; di < 2; di++)
                for (int dj = 0; dj < 2; dj++)
                    for (int dk = 0; dk < 2; dk++) {
                        const int fi = 2 * i + di, fj = 2 * j + dj, fk = 2 * k + dk;
                        if (fi < nf.n0 && fj < nf.n1 && fk < nf.n2) {
                            float t0, t1;
                            neighbor_sum<UNIT>(L, fi, fj, fk, StoredQ<T>{q, nf}, t0, t1);
                            const int c = nf.at(fi, fj, fk);
                            total = total + residual_value(L, fi, fj, fk, to_f32(q[c]), to_f32(rhs[c]), t0, t1);
                            count += 1;
                        }
                    }
        }
    }
    // the device-side gate is read only before the stores, so its latency hides behind the loads (a gated-off
    // launch, rare, then does the reads and arithmetic for nothing)
    if (gated_off(gate)) return;
    const int c = nc.at(i, j, k);
    coarse_rhs[c] = from_f32<T>(total / (float)count);
    if (coarse_q != nullptr) coarse_q[c] = from_f32<T>(0.0f);
}

// residual_kernel + restrict_kernel (sim/multigrid.py:36-60), one thread per coarse cell (grid3 of the coarse level)
template <bool UNIT, typename T = float>
__global__ __launch_bounds__(256) void k_residual_restrict(LevelArgs L, const T *__restrict__ q,
                                                           const T *__restrict__ rhs, T *__restrict__ coarse_rhs,
                                                           T *__restrict__ coarse_q, Dims nc, const double *gate) {
    residual_restrict_cell<UNIT, T>(L, q, rhs, coarse_rhs, coarse_q, nc, blockIdx, gate);
}

// k_residual_restrict for levels with couplings other than 1, with one thread per child: lanes 8 m + 4 di +
// 2 dj + dk hold the children of coarse cell m of the wavefront. A child whose six neighbours are inside the
// level and none of them (nor the child) the last cell on an axis has couplings 1, a sum of couplings of 6
// and volume 1 (the interior form); the others take the general update. Lane 8 m then adds its cell's eight
// residuals in the reference's child order (an absent, overhang child adds +0.0 to a total that starts at
// +0.0, which changes no bit) and divides by the count. Eight short threads instead of one thread walking
// eight children keep the small levels' latency hidden.
template <typename T>
__device__ __forceinline__ void residual_restrict_children(const LevelArgs &L, const T *__restrict__ q,
                                                           const T *__restrict__ rhs, T *__restrict__ coarse_rhs,
                                                           Dims nc, uint32_t block, const double *gate = nullptr) {
    const Dims nf = cell_dims(L.g);
    const int child = threadIdx.x & 7;
    const int cell = (int)(block * 32u + (threadIdx.x >> 3));
    const bool valid = cell < nc.count();
    const int row = cell / nc.n2, k = cell - row * nc.n2, i = row / nc.n1, j = row - i * nc.n1;
    const int fi = 2 * i + (child >> 2), fj = 2 * j + ((child >> 1) & 1), fk = 2 * k + (child & 1);
    const bool exists = valid && fi < nf.n0 && fj < nf.n1 && fk < nf.n2;
    float r = 0.0f;
    if (exists) {
        const float q_c = to_f32(q[nf.at(fi, fj, fk)]), rhs_c = to_f32(rhs[nf.at(fi, fj, fk)]);
        if (fi >= 1 && fi + 3 <= nf.n0 && fj >= 1 && fj + 3 <= nf.n1 && fk >= 1 && fk + 3 <= nf.n2) {
            const StoredQ<T> Q{q, nf};
            float t0 = 0.0f + Q(fi - 1, fj, fk);
            t0 = t0 + Q(fi + 1, fj, fk);
            t0 = t0 + Q(fi, fj - 1, fk);
            t0 = t0 + Q(fi, fj + 1, fk);
            t0 = t0 + Q(fi, fj, fk - 1);
            t0 = t0 + Q(fi, fj, fk + 1);
            const float scale = uniform(L.g.cell_m) * uniform(L.g.cell_m);  // (1 cell_m) cell_m
            r = rhs_c - (t0 - 6.0f * q_c) / scale;
        } else {
            float t0, t1;
            neighbor_sum<false>(L, fi, fj, fk, StoredQ<T>{q, nf}, t0, t1);
            r = residual_value(L, fi, fj, fk, q_c, rhs_c, t0, t1);
        }
    }
    const int lead = (int)(threadIdx.x & 63u) & ~7;
    float total = 0.0f;
    int count = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        total = total + __shfl(r, lead + c, 64);
        count += __shfl(exists ? 1 : 0, lead + c, 64);
    }
    if (gated_off(gate)) return;  // read only before the store, as in residual_restrict_cell
    if (child == 0 && valid) coarse_rhs[cell] = from_f32<T>(total / (float)count);
}

template <typename T>
__global__ __launch_bounds__(256) void k_residual_restrict_children(LevelArgs L, const T *__restrict__ q,
                                                                    const T *__restrict__ rhs,
                                                                    T *__restrict__ coarse_rhs, Dims nc,
                                                                    const double *gate) {
    residual_restrict_children<T>(L, q, rhs, coarse_rhs, nc, blockIdx.x, gate);
}

inline dim3 children_grid(const Dims &nc) { return dim3((unsigned)((nc.count() + 31) / 32)); }

// ------------------------------------------------------------------ smoothing of a flat mid level in LDS

// Sweeps of a flat level between the simulation grid and the LDS tail (red_black_sweep_kernel for both colours,
// sim/pressure.py:80-92), whose launches cost their latency more than their bytes, run in LDS: a block holds a
// TI x TJ x TK tile (even extents at even origins, so the tile's cells are whole 2 x 2 x 2 children blocks) with
// an H-cell halo of q and rhs, +0.0 outside the level, H = 2 SWEEPS (+ 1 with RESTRICT). Sweep s (from 0) runs
// red on the tile and its rings up to H - 1 - 2s, then black up to H - 2 - 2s, so each pass reads only values
// that are already those of the whole-level sweeps, and the tile (with RESTRICT also its first ring) ends with
// the values of SWEEPS sweeps. Before the sweeps: FROM_Q loads q_in; FROM_ZERO starts from q = 0 and loads no q;
// FROM_PROLONG adds prolong_add_kernel's correction (sim/multigrid.py:63-68) to q_in, from a box of coarse cells
// staged in LDS with prolong_at_row's taps, weights and order. RESTRICT: the tile's residuals (residual_kernel,
// sim/multigrid.py:36-43) and their restriction (restrict_kernel, sim/multigrid.py:46-60: each coarse cell's
// children summed in x-major order, the overhang skipped) to the coarse rhs. On a flat level every coupling
// between two cells and every volume is 1 and an open face couples by the level's ghost value: the update and
// the residual are the tile sweep's and k_residual_restrict_children's operations (an outside neighbour adds
// the +0.0 the reference adds for a ghost or a closed face; the division by the sum of couplings is IEEE, which
// divide_flat equals), so q and the coarse rhs are the same bits as the separate launches'.
template <int TI, int TJ, int TK, int SWEEPS, int START, bool RESTRICT>
struct SmoothBox {
    static_assert(TI % 2 == 0 && TJ % 2 == 0 && TK % 2 == 0, "tiles hold whole 2 x 2 x 2 children blocks");
    static constexpr int H = 2 * SWEEPS + (RESTRICT ? 1 : 0);
    static constexpr int QI = TI + 2 * H, QJ = TJ + 2 * H, QK = TK + 2 * H, CELLS = QI * QJ * QK;
    // the coarse taps of the box's fine cells lie in QI / 2 + 2 coarse cells from the box's lowest one, and so on
    static constexpr int CI = QI / 2 + 2, CJ = QJ / 2 + 2, CK = QK / 2 + 2;
    static constexpr int COARSE = START == FROM_PROLONG ? CI * CJ * CK : 0;
    static constexpr int RES = RESTRICT ? TI * TJ * TK : 0;
    // the prolongation's x and then y interpolations: per fine i over the coarse box's (j, k), then per fine
    // (i, j) over its k
    static constexpr int PX = START == FROM_PROLONG ? QI * CJ * CK : 0, PY = START == FROM_PROLONG ? QI * QJ * CK : 0;
    static constexpr size_t BYTES = sizeof(float) * (size_t)(2 * CELLS + COARSE + RES + PX + PY);
};

// the lowest coarse tap of a fine index f >= 1 is (f - 1) / 2, of f = 0 it is 0; the box's lowest tap
__device__ __forceinline__ int coarse_origin(int f0) { return max(0, (f0 - 1) >> 1); }

template <int TI, int TJ, int TK, int THREADS, int SWEEPS, int START, bool RESTRICT>
__global__ __launch_bounds__(THREADS) void k_level_smooth(LevelArgs L, const float *__restrict__ q_in,
                                                          float *__restrict__ q_out, const float *__restrict__ rhs,
                                                          const float *__restrict__ coarse, Dims nq,
                                                          float *__restrict__ coarse_rhs, Dims nc, float relaxation,
                                                          const double *gate, FastDiv tiles_k, FastDiv tiles_j) {
    using Box = SmoothBox<TI, TJ, TK, SWEEPS, START, RESTRICT>;
    constexpr int H = Box::H, QJ = Box::QJ, QK = Box::QK, CELLS = Box::CELLS;
    constexpr int CJ = Box::CJ, CK = Box::CK, COARSE = Box::COARSE;
    constexpr int LOADS = (CELLS + THREADS - 1) / THREADS, COARSE_LOADS = (COARSE + THREADS - 1) / THREADS;
    extern __shared__ float lds[];
    float *sq = lds, *sr = lds + CELLS, *sc = sr + CELLS, *sres = sc + COARSE, *px = sres + Box::RES, *py = px + Box::PX;
    const Dims n = cell_dims(L.g);
    const int tile = (int)blockIdx.x, tr = tiles_k(tile), tk = tile - tr * (int)tiles_k.d;
    const int ti = tiles_j(tr), tj = tr - ti * (int)tiles_j.d;
    const int i0 = ti * TI, j0 = tj * TJ, k0 = tk * TK;
    const int tid = (int)threadIdx.x;
    auto inside = [&](int i, int j, int k) {
        return (unsigned)i < (unsigned)n.n0 && (unsigned)j < (unsigned)n.n1 && (unsigned)k < (unsigned)n.n2;
    };
    // the box's q and rhs (the outermost ring's rhs is never used) and the coarse box, all in flight before the
    // gate is read (uniform: the block leaves before any barrier)
    float vq[LOADS], vr[LOADS];
#pragma unroll
    for (int u = 0; u < LOADS; u++) {
        const int idx = min(tid + u * THREADS, CELLS - 1);
        const int a = idx / (QJ * QK), b = (idx / QK) % QJ, c = idx % QK;
        const int i = i0 - H + a, j = j0 - H + b, k = k0 - H + c;
        const bool in = inside(i, j, k);
        const bool ring = a > 0 && a < Box::QI - 1 && b > 0 && b < QJ - 1 && c > 0 && c < QK - 1;
        const int at = in ? n.at(i, j, k) : 0;
        vq[u] = START != FROM_ZERO && in ? q_in[at] : 0.0f;
        vr[u] = in && ring ? rhs[at] : 0.0f;
    }
    const int ci0 = coarse_origin(i0 - H), cj0 = coarse_origin(j0 - H), ck0 = coarse_origin(k0 - H);
    float vc[COARSE_LOADS > 0 ? COARSE_LOADS : 1];
#pragma unroll
    for (int u = 0; u < COARSE_LOADS; u++) {
        const int idx = min(tid + u * THREADS, COARSE - 1);
        const int ci = ci0 + idx / (CJ * CK), cj = cj0 + (idx / CK) % CJ, ck = ck0 + idx % CK;
        const bool in = ci < nq.n0 && cj < nq.n1 && ck < nq.n2;
        vc[u] = in ? coarse[nq.at(ci, cj, ck)] : 0.0f;
    }
    if (gated_off(gate)) return;
#pragma unroll
    for (int u = 0; u < LOADS; u++) {
        const int idx = tid + u * THREADS;
        if (idx < CELLS) {
            sq[idx] = vq[u];
            sr[idx] = vr[u];
        }
    }
#pragma unroll
    for (int u = 0; u < COARSE_LOADS; u++) {
        const int idx = tid + u * THREADS;
        if (idx < COARSE) sc[idx] = vc[u];
    }
    __syncthreads();
    if constexpr (START == FROM_PROLONG) {
        // q + the coarse correction at each box cell inside the level, separably: prolong_at_row's c00, c10, c01
        // and c11 are the x interpolations at the cell's i for coarse (j, k) = (jc0, kc0), (jc1, kc0), (jc0, kc1)
        // and (jc1, kc1); its two y interpolations are those at the cell's j for kc0 and kc1, and the z one joins
        // them: the same taps, weights and order, each interpolation formed once for every box cell that uses it.
        // Box rows outside the level take the clamped taps and are never used.
        constexpr int CI = Box::CI, QI = Box::QI;
        (void)CI;
        auto taps = [](int f, int extent, int &c0, int &c1, float &w) {
            const float x = clamp_((float)f * 0.5f - 0.25f, 0.0f, (float)(extent - 1));
            c0 = (int)x;
            c1 = min(c0 + 1, extent - 1);
            w = x - (float)c0;
        };
        for (int idx = tid; idx < QI * CJ * CK; idx += THREADS) {
            const int a = idx / (CJ * CK), rest = idx - a * (CJ * CK);
            int c0, c1;
            float w;
            taps(i0 - H + a, nq.n0, c0, c1, w);
            px[idx] = lerp_(sc[(c0 - ci0) * (CJ * CK) + rest], sc[(c1 - ci0) * (CJ * CK) + rest], w);
        }
        __syncthreads();
        for (int idx = tid; idx < QI * QJ * CK; idx += THREADS) {
            const int row = idx / CK, kc = idx - row * CK, a = row / QJ, b = row - a * QJ;
            int c0, c1;
            float w;
            taps(j0 - H + b, nq.n1, c0, c1, w);
            py[idx] = lerp_(px[(a * CJ + (c0 - cj0)) * CK + kc], px[(a * CJ + (c1 - cj0)) * CK + kc], w);
        }
        __syncthreads();
        for (int idx = tid; idx < CELLS; idx += THREADS) {
            const int row = idx / QK, c = idx - row * QK, a = row / QJ, b = row - a * QJ;
            const int i = i0 - H + a, j = j0 - H + b, k = k0 - H + c;
            if (!inside(i, j, k)) continue;
            int c0, c1;
            float w;
            taps(k, nq.n2, c0, c1, w);
            sq[idx] = sq[idx] + lerp_(py[row * CK + (c0 - ck0)], py[row * CK + (c1 - ck0)], w);
        }
        __syncthreads();
    }
    const float cm2 = uniform(L.g.cell_m) * uniform(L.g.cell_m);  // (1 cell_m) cell_m: every volume is 1
    const float ghost = uniform(L.ghost);
    bool open_lo[3], open_hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        open_lo[a] = uniform(L.g.open_lo[a]) == 1;
        open_hi[a] = uniform(L.g.open_hi[a]) == 1;
    }
    constexpr int SI = QJ * QK, SJ = QK;
    // neighbor_sum's t0 (the six neighbours in the reference's order) and t1 (their couplings) at LDS index p
    auto sums = [&](int i, int j, int k, int p, float &t0, float &t1) {
        t0 = 0.0f + sq[p - SI];
        t0 = t0 + sq[p + SI];
        t0 = t0 + sq[p - SJ];
        t0 = t0 + sq[p + SJ];
        t0 = t0 + sq[p - 1];
        t0 = t0 + sq[p + 1];
        const int cell[3] = {i, j, k}, extent[3] = {n.n0, n.n1, n.n2};
        float sum = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            sum = sum + (cell[a] > 0 ? 1.0f : open_lo[a] ? ghost : 0.0f);
            sum = sum + (cell[a] < extent[a] - 1 ? 1.0f : open_hi[a] ? ghost : 0.0f);
        }
        t1 = sum;
    };
    // one colour over the tile and `ring` rings: (TI + 2 ring) (TJ + 2 ring) rows of (TK + 2 ring) / 2 cells
    auto pass = [&](int color, int ring) {
        const int ri = TI + 2 * ring, rj = TJ + 2 * ring, half = (TK + 2 * ring) / 2;
        const int total = ri * rj * half;
        for (int x = tid; x < total; x += THREADS) {
            const int row = x / half, slot = x - row * half;
            const int a = row / rj, b = row - a * rj;
            const int i = i0 - ring + a, j = j0 - ring + b, kb = k0 - ring;
            const int k = kb + 2 * slot + ((color ^ (i + j + kb)) & 1);
            if (!inside(i, j, k)) continue;
            const int p = ((i - i0 + H) * QJ + (j - j0 + H)) * QK + (k - k0 + H);
            float t0, t1;
            sums(i, j, k, p, t0, t1);
            if (t1 == 0.0f) continue;
            const float q_c = sq[p];
            const float target = (t0 - cm2 * sr[p]) / t1;
            sq[p] = q_c + relaxation * (target - q_c);
        }
    };
#pragma unroll
    for (int sweep = 0; sweep < SWEEPS; sweep++) {
        pass(0, H - 1 - 2 * sweep);
        __syncthreads();
        pass(1, H - 2 - 2 * sweep);
        __syncthreads();
    }
    // the tile's q, and with RESTRICT its residuals (residual_value with volume 1)
    for (int x = tid; x < TI * TJ * TK; x += THREADS) {
        const int a = x / (TJ * TK), b = (x / TK) % TJ, c = x % TK;
        const int i = i0 + a, j = j0 + b, k = k0 + c;
        if (!inside(i, j, k)) continue;
        const int p = ((a + H) * QJ + (b + H)) * QK + (c + H);
        const float q_c = sq[p];
        if constexpr (RESTRICT) {
            float t0, t1;
            sums(i, j, k, p, t0, t1);
            sres[x] = sr[p] - (t0 - t1 * q_c) / cm2;
        }
        q_out[n.at(i, j, k)] = q_c;
    }
    if constexpr (RESTRICT) {
        __syncthreads();
        constexpr int RI = TI / 2, RJ = TJ / 2, RK = TK / 2;
        for (int x = tid; x < RI * RJ * RK; x += THREADS) {
            const int a = x / (RJ * RK), b = (x / RK) % RJ, c = x % RK;
            const int ci = i0 / 2 + a, cj = j0 / 2 + b, ck = k0 / 2 + c;
            if (ci >= nc.n0 || cj >= nc.n1 || ck >= nc.n2) continue;
            float total = 0.0f;
            int count = 0;
#pragma unroll
            for (int ch = 0; ch < 8; ch++) {
                const int di = ch >> 2, dj = (ch >> 1) & 1, dk = ch & 1;
                if (2 * ci + di < n.n0 && 2 * cj + dj < n.n1 && 2 * ck + dk < n.n2) {
                    total = total + sres[((2 * a + di) * TJ + (2 * b + dj)) * TK + (2 * c + dk)];
                    count += 1;
                }
            }
            coarse_rhs[nc.at(ci, cj, ck)] = total / (float)count;
        }
    }
}

template <int TI, int TJ, int TK, int THREADS, int SWEEPS, int START, bool RESTRICT>
int launch_level_smooth_shape(const LevelArgs &L, const float *q_in, float *q_out, const float *rhs,
                              const float *coarse, Dims nq, float *coarse_rhs, Dims nc, float relaxation,
                              const double *gate, hipStream_t s) {
    constexpr size_t bytes = SmoothBox<TI, TJ, TK, SWEEPS, START, RESTRICT>::BYTES;
    static_assert(bytes <= 160 * 1024, "a block's LDS fits one CU");
    const auto kernel = &k_level_smooth<TI, TJ, TK, THREADS, SWEEPS, START, RESTRICT>;
    if (bytes > 65536) {
        static bool allowed = false;  // more than the default 64 KB of dynamic LDS, granted once per kernel
        if (!allowed) {
            if (hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
                return fail(PLM_EINVAL, "level smoothing: %zu bytes of LDS refused", bytes);
            allowed = true;
        }
    }
    const Dims n = cell_dims(L.g);
    const int tiles_k = (n.n2 + TK - 1) / TK, tiles_j = (n.n1 + TJ - 1) / TJ, tiles_i = (n.n0 + TI - 1) / TI;
    kernel<<<dim3((unsigned)(tiles_k * tiles_j * tiles_i)), THREADS, bytes, s>>>(
        L, q_in, q_out, rhs, coarse, nq, coarse_rhs, nc, relaxation, gate, fast_div(tiles_k), fast_div(tiles_j));
    return check_launch("level smoothing");
}

// 1024 threads a tile: each pass is one or two updates a thread, so a block's chain of passes stays short. Big
// levels take 4 x 8 x 32 tiles (fewer halo loads and ring updates per cell), small ones 2 x 4 x 32 (enough
// blocks), or 2 x 4 x 16 when the level has at most 16 cells along k (5.7 and 7.1 us at 16 x 24 x 16 against 7.0
// and 8.7 with 32). Measured per launch at 64 x 96 x 64 (big) and 32 x 48 x 32 (small): the last pre-smoothing sweep with
// the restriction 8.2 and 4.7 us (256 threads: 11.1 and 7.4; 2 x 4 x 32 tiles on the big level: 13.6); both
// pre-smoothing sweeps with the restriction 17.4 and 7.0 us (a tile sweep and the one-sweep form: 13.4 and 9.6);
// the prolongation with both post-smoothing sweeps 18.4 and 9.0 us (the tile sweeps: 17.4 and 14.4); on the big
// level one sweep from zero 4.9 us (a tile sweep 5.5), the prolongation with one sweep 9.9 us (10.7).
inline bool big_smooth_tiles(const Dims &n) { return ((n.n0 + 3) / 4) * ((n.n1 + 7) / 8) * ((n.n2 + 31) / 32) >= 128; }

template <int SWEEPS, int START, bool RESTRICT>
int launch_level_smooth(const LevelArgs &L, const float *q_in, float *q_out, const float *rhs, const float *coarse,
                        Dims nq, float *coarse_rhs, Dims nc, float relaxation, const double *gate, hipStream_t s) {
    const Dims n = cell_dims(L.g);
    if (big_smooth_tiles(n))
        return launch_level_smooth_shape<4, 8, 32, 1024, SWEEPS, START, RESTRICT>(
            L, q_in, q_out, rhs, coarse, nq, coarse_rhs, nc, relaxation, gate, s);
    if (n.n2 <= 16)  // thin levels: a 32-cell k tile would be half empty
        return launch_level_smooth_shape<2, 4, 16, 1024, SWEEPS, START, RESTRICT>(L, q_in, q_out, rhs, coarse, nq,
                                                                                 coarse_rhs, nc, relaxation, gate, s);
    return launch_level_smooth_shape<2, 4, 32, 1024, SWEEPS, START, RESTRICT>(L, q_in, q_out, rhs, coarse, nq,
                                                                             coarse_rhs, nc, relaxation, gate, s);
}

// prolong_add_kernel, sim/multigrid.py:63-68
__global__ __launch_bounds__(256) void k_prolong_add(const float *__restrict__ coarse, Dims nc, float *__restrict__ fine,
                                                     Dims nf, const double *gate) {
    if (gated_off(gate)) return;
    int i, j, k;
    if (!thread_cell(nf, i, j, k)) return;
    const int c = nf.at(i, j, k);
    fine[c] = fine[c] + prolongation(coarse, nc, i, j, k);
}

// prolong_add_kernel with one wavefront per row of 64 cells along k: the row's i and j interpolation is
// uniform and only the k part varies per lane (the operations of sample_index, so the same values). T is the
// stored type (the sum is formed in FP32 and rounded once into T).
template <typename T = float>
__device__ __forceinline__ void prolong_row_cell(const T *__restrict__ coarse, Dims nc, T *__restrict__ fine, Dims nf,
                                                 dim3 block) {
    const int k = block.x * BX + threadIdx.x;
    const int j = uniform((int)(block.y * BY + threadIdx.y)), i = block.z;
    if (j >= nf.n1 || k >= nf.n2) return;
    uint32_t pk0, pk1;
    float tz;
    prolong_k<T>(nc, k, pk0, pk1, tz);
    const float p = prolong_at_row(coarse, nc, i, j, pk0, pk1, tz);
    T *cell = reinterpret_cast<T *>(reinterpret_cast<char *>(fine) + offset_of(nf, i, j, k) * (uint32_t)sizeof(T));
    *cell = from_f32<T>(to_f32(*cell) + p);
}


// prolong_add_kernel by rows (prolong_row_cell's values) with four rows per thread (rows j, j + BY, j + 2 BY, j + 3 BY of the block's 4 BY rows, each
// uniform over a wavefront): a quarter of the wavefronts, each with the four cells' loads in flight together,
// and every load issued before the gate is read (the loads are in range whatever the gate says), so the gate's
// latency overlaps theirs. The values are prolong_row_cell's.
constexpr int PROLONG_ROWS_PER_THREAD = 4;
template <typename T = float>
__global__ __launch_bounds__(256) void k_prolong_rows4(const T *__restrict__ coarse, Dims nc, T *__restrict__ fine,
                                                       Dims nf, const double *gate) {
    constexpr int R = PROLONG_ROWS_PER_THREAD;
    const int k = blockIdx.x * BX + threadIdx.x, i = blockIdx.z;
    const int j0 = uniform((int)(blockIdx.y * BY * R + threadIdx.y));
    const int kk = min(k, nf.n2 - 1);
    uint32_t pk0, pk1;
    float tz;
    prolong_k<T>(nc, kk, pk0, pk1, tz);
    float p[R], old[R];
    T *cell[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int j = min(j0 + r * BY, nf.n1 - 1);
        p[r] = prolong_at_row(coarse, nc, i, j, pk0, pk1, tz);
        cell[r] = reinterpret_cast<T *>(reinterpret_cast<char *>(fine) + offset_of(nf, i, j, kk) * (uint32_t)sizeof(T));
        old[r] = to_f32(*cell[r]);
    }
    if (gated_off(gate)) return;
#pragma unroll
    for (int r = 0; r < R; r++)
        if (k < nf.n2 && j0 + r * BY < nf.n1) *cell[r] = from_f32<T>(old[r] + p[r]);
}

inline dim3 prolong_rows4_grid(Dims n) {
    return dim3((n.n2 + BX - 1) / BX, (n.n1 + BY * PROLONG_ROWS_PER_THREAD - 1) / (BY * PROLONG_ROWS_PER_THREAD), n.n0);
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
constexpr int WAVE = 64;
constexpr int WAVE_CELLS = 128;  // levels this small run on one wavefront, whose passes need no block barrier
// cells per colour per thread on block levels, with and without stored couplings (a level with more colour
// slots than that times TAIL_THREADS stays out of the tail)
constexpr int BLOCK_PLAN = 2, BLOCK_PLAN_ONES = 3;
// dynamic LDS the tail may take (gfx950: 160 KB per CU, less the level table staged beside it)
constexpr size_t TAIL_LDS_BYTES = 150 * 1024;

// Division by a level extent through a multiply-high: m = ceil(2^32 / d) gives floor(t / d) exactly for
// t, d < 2^16 (the error t (m d - 2^32) / (d 2^32) stays below 1 / d).

// Slots of the colour enumeration on a level: n0 n1 ceil(n2 / 2); a tail thread plans a fixed number of them
int colour_slots(const plm_grid_info &g) { return g.dims[0] * g.dims[1] * ((g.dims[2] + 1) / 2); }


// Each tail level lives in LDS zero-bordered: its q, rhs and (for the level being restricted) residual as
// (n0 + 2) (n1 + 2) (n2 + 2) arrays, cell (i, j, k) at ((i + 1)(n1 + 2) + j + 1)(n2 + 2) + k + 1. The border of
// q stays +0.0, the q a pass reads for a neighbour outside the level (a ghost or a closed face, whose term the
// reference adds as 0), so a neighbour is one address add away and no mask or select is needed.
struct TailLevels {
    int count;      // levels in the tail
    int zero_top;   // the top level's q starts at 0 (a coarse level the V-cycle just cleared)
    int wave_from;  // first level (within the tail) from which every level has <= WAVE_CELLS cells; count if none
    LevelArgs level[MAX_TAIL];
    int qoff[MAX_TAIL];  // float offset in LDS of the level's padded q; its padded rhs follows at roff
    int roff[MAX_TAIL];
    int res_offset;  // the padded residual of the level being restricted
    int q_floats;    // floats from qoff[0] through the last q and rhs: cleared at the start
    FastDiv div_n1[MAX_TAIL], div_n2[MAX_TAIL], div_half[MAX_TAIL];  // by n1, n2 and (n2 + 1) / 2
};

struct BlockSync {
    __device__ __forceinline__ void operator()() const { __syncthreads(); }
};

// One wavefront's lanes run in lockstep and its LDS operations complete in order, so between two passes of
// a single wave it is enough that the compiler keeps its LDS accesses on their side of this point.
struct WaveSync {
    __device__ __forceinline__ void operator()() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

// One tail level's geometry in scalar registers for the duration of its passes (read once from the staged
// table: a per-neighbour LDS read of these would serialise every pass on LDS latency).
struct TailLevel {
    int n[3];
    int open_lo[3], open_hi[3];
    float fill[3];
    float cell_m;
    int unit;
    FastDiv div_n1, div_n2, div_half;
    // The level's coupling table in the staged LDS copy: Coupling.in[a][area][1] and Coupling.ghost, the
    // couplings that are not the area itself, read per lane with one indexed LDS load (a select among
    // uniform values costs a move and a select per value)
    const Coupling *cp;
};

// A level's extents alone, for the other level of a restrict or prolong pass
struct TailDims {
    int n[3];
    FastDiv div_n1, div_n2;
};

__device__ __forceinline__ TailDims load_dims(const TailLevels &t, int m) {
    TailDims r;
#pragma unroll
    for (int a = 0; a < 3; a++) r.n[a] = uniform(t.level[m].g.dims[a]);
    r.div_n1 = FastDiv{uniform(t.div_n1[m].d), uniform(t.div_n1[m].m)};
    r.div_n2 = FastDiv{uniform(t.div_n2[m].d), uniform(t.div_n2[m].m)};
    return r;
}

__device__ __forceinline__ TailLevel load_level(const TailLevels &t, int m) {
    const LevelArgs &L = t.level[m];
    TailLevel r;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r.n[a] = uniform(L.g.dims[a]);
        r.open_lo[a] = uniform(L.g.open_lo[a]);
        r.open_hi[a] = uniform(L.g.open_hi[a]);
        r.fill[a] = uniform(L.g.fill_hi[a]);
    }
    r.cp = &L.cp;
    r.cell_m = uniform(L.g.cell_m);
    r.unit = uniform(L.unit);
    r.div_n1 = FastDiv{uniform(t.div_n1[m].d), uniform(t.div_n1[m].m)};
    r.div_n2 = FastDiv{uniform(t.div_n2[m].d), uniform(t.div_n2[m].m)};
    r.div_half = FastDiv{uniform(t.div_half[m].d), uniform(t.div_half[m].m)};
    return r;
}

// cell_fill, sim/pressure.py:39-44
__device__ __forceinline__ float fill_at(const TailLevel &L, int a, int index) {
    return index == L.n[a] - 1 ? L.fill[a] : 1.0f;
}

// the padded index of cell (i, j, k) of a level of extents n
__device__ __forceinline__ int pad_at(const int (&n)[3], int i, int j, int k) {
    // a tail level holds at most PLM_LDS_CELLS cells, so every product fits the full-rate 24-bit multiply
    return (int)__umul24(__umul24((uint32_t)(i + 1), (uint32_t)(n[1] + 2)) + (uint32_t)(j + 1), (uint32_t)(n[2] + 2)) +
           k + 1;
}

// What one tail pass needs of one cell, fixed for a visit of its level (neighbor_sum's couplings and their
// sum, sim/pressure.py:52-77; cell_volume * cell_m^2, :47-49 and :86-92): p = the cell's padded index, or -1 for
// no cell. ONES: every fill_hi of the level is 1, so every in-domain coupling is area / (0.5 (1 + 1)) = 1
// exactly, every volume is 1 and the scale is (1 cell_m) cell_m for all cells; only t1 (through the open-face
// ghosts) differs from the simulation grid's, and the couplings and scale are not stored. An outside
// neighbour's coupling multiplies the border's +0.0, adding +0.0 as the reference adds 0 for it.
template <bool ONES>
struct PlanCell;
template <>
struct PlanCell<true> {
    int p;
    float t1;
};
template <>
struct PlanCell<false> {
    int p;
    float t1;
    float scale;
    float cp[6];
};

// the scale (volume cm^2) of a planned cell: uniform on a ONES level
template <bool ONES>
__device__ __forceinline__ float plan_scale(const PlanCell<ONES> &p, float uniform_scale) {
    if constexpr (ONES)
        return uniform_scale;
    else
        return p.scale;
}

// Builds a cell's plan from the level's geometry and coupling table, with the reference's operations: the
// area (1 F_b) F_c, the in-domain coupling (the area itself unless the pair holds the axis' last cell, then
// the precomputed in1), the open-face ghost coupling, t1 summed in neighbour order, (F_0 F_1) F_2 cm cm.
template <bool ONES>
__device__ __forceinline__ PlanCell<ONES> plan_cell(const TailLevel &L, int i, int j, int k) {
    PlanCell<ONES> p;
    const int cell[3] = {i, j, k};
    const bool last[3] = {i == L.n[0] - 1, j == L.n[1] - 1, k == L.n[2] - 1};
    float t1 = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
        const int area_index = (last[b] ? 1 : 0) + (last[c] ? 2 : 0);
        float area = 1.0f, in1 = 1.0f;
        if (!ONES) {
            area = area * (last[b] ? L.fill[b] : 1.0f);
            area = area * (last[c] ? L.fill[c] : 1.0f);
            in1 = L.cp->in[a][area_index][1];
        }
        const bool in_lo = cell[a] > 0, in_hi = !last[a];
        const float cp_lo = ONES ? 1.0f : (last[a] ? in1 : area);
        const float cp_hi = ONES ? 1.0f : (cell[a] + 1 == L.n[a] - 1 ? in1 : area);
        const float ghost_lo = L.cp->ghost[a][area_index][last[a] ? 1 : 0];
        const float ghost_hi = L.cp->ghost[a][area_index][1];
        t1 = t1 + (in_lo ? cp_lo : L.open_lo[a] == 1 ? ghost_lo : 0.0f);
        t1 = t1 + (in_hi ? cp_hi : L.open_hi[a] == 1 ? ghost_hi : 0.0f);
        if constexpr (!ONES) {
            p.cp[2 * a] = cp_lo;
            p.cp[2 * a + 1] = cp_hi;
        }
    }
    if constexpr (!ONES) {
        float vol = fill_at(L, 0, i) * fill_at(L, 1, j);
        vol = vol * fill_at(L, 2, k);
        p.scale = vol * L.cell_m * L.cell_m;
    }
    p.t1 = t1;
    p.p = pad_at(L.n, i, j, k);
    return p;
}

// The thread's cells of one colour on one level: slots x = tid + u NT of the colour enumeration, (n2 + 1) / 2
// slots per row, k = 2 slot + ((colour ^ (i + j)) & 1).
template <bool ONES, int K, int NT>
struct ColourPlan {
    PlanCell<ONES> cell[K];
    __device__ __forceinline__ void build(const TailLevel &L, int color, int tid) {
        const int half = (L.n[2] + 1) / 2, total = L.n[0] * L.n[1] * half;
#pragma unroll
        for (int u = 0; u < K; u++) {
            const int x = tid + u * NT;
            const int row = L.div_half(x), slot = x - row * half;
            const int i = L.div_n1(row), j = row - i * L.n[1];
            const int k = 2 * slot + ((color ^ (i + j)) & 1);
            if (x < total && k < L.n[2]) {
                cell[u] = plan_cell<ONES>(L, i, j, k);
            } else {
                cell[u].p = -1;
                cell[u].t1 = 0.0f;
                if constexpr (!ONES) {
                    cell[u].scale = 1.0f;
                    for (int s = 0; s < 6; s++) cell[u].cp[s] = 1.0f;
                }
            }
        }
    }
};

// A level's six neighbour strides in its padded layout (x-, x+, y-, y+, z-, z+) and the uniform scale of a
// ONES level
struct LevelStrides {
    int s[6];
    float scale;
};

__device__ __forceinline__ LevelStrides level_strides(const TailLevel &L) {
    const int row = L.n[2] + 2, plane = (L.n[1] + 2) * row;
    return LevelStrides{{-plane, plane, -row, row, -1, 1}, (1.0f * L.cell_m) * L.cell_m};
}

// t0 of neighbor_sum for a planned cell: the six terms added in the reference's order
template <bool ONES>
__device__ __forceinline__ float plan_t0(const PlanCell<ONES> &p, const float *q, const int (&stride)[6]) {
    float t0 = 0.0f;
#pragma unroll
    for (int s = 0; s < 6; s++) {
        const float qn = q[p.p + stride[s]];
        if constexpr (ONES)
            t0 = t0 + qn;
        else
            t0 = t0 + p.cp[s] * qn;
    }
    return t0;
}

// One colour pass over the planned cells (red_black_sweep_kernel, sim/pressure.py:80-92)
template <bool ONES, int K, int NT>
__device__ __forceinline__ void plan_pass(const ColourPlan<ONES, K, NT> &plan, float *q, const float *rhs,
                                          const LevelStrides &st, float relaxation) {
#pragma unroll
    for (int u = 0; u < K; u++) {
        const PlanCell<ONES> &p = plan.cell[u];
        if (p.p < 0 || p.t1 == 0.0f) continue;
        const float t0 = plan_t0<ONES>(p, q, st.s);
        const float source = plan_scale<ONES>(p, st.scale) * rhs[p.p];
        const float target = (t0 - source) / p.t1;
        q[p.p] = q[p.p] + relaxation * (target - q[p.p]);
    }
}

// residual_kernel (sim/multigrid.py:36-43) of the planned cells into res (padded like q)
template <bool ONES, int K, int NT>
__device__ __forceinline__ void plan_residual(const ColourPlan<ONES, K, NT> &plan, const float *q, const float *rhs,
                                              float *res, const LevelStrides &st) {
#pragma unroll
    for (int u = 0; u < K; u++) {
        const PlanCell<ONES> &p = plan.cell[u];
        if (p.p < 0) continue;
        const float t0 = plan_t0<ONES>(p, q, st.s);
        res[p.p] = rhs[p.p] - (t0 - p.t1 * q[p.p]) / plan_scale<ONES>(p, st.scale);
    }
}

// restrict_kernel (sim/multigrid.py:46-60) from the stored residuals into the coarse rhs (the coarse q is
// already +0.0: every tail level's q starts cleared, the reference's coarse.q.zero_(), :143)
template <int NT, typename Sync>
__device__ __forceinline__ void tail_restrict(const TailDims &F, const TailDims &C, const float *res, float *crhs,
                                              int tid, Sync sync) {
    const int cells = C.n[0] * C.n[1] * C.n[2];
    for (int c = tid; c < cells; c += NT) {
        const int r = C.div_n2(c), k = c - r * C.n[2];
        const int i = C.div_n1(r), j = r - i * C.n[1];
        float total = 0.0f, count = 0.0f;
        for (int di = 0; di < 2; di++)
            for (int dj = 0; dj < 2; dj++)
                for (int dk = 0; dk < 2; dk++) {
                    const int fi = 2 * i + di, fj = 2 * j + dj, fk = 2 * k + dk;
                    if (fi < F.n[0] && fj < F.n[1] && fk < F.n[2]) {
                        total = total + res[pad_at(F.n, fi, fj, fk)];
                        count = count + 1.0f;
                    }
                }
        crhs[pad_at(C.n, i, j, k)] = total / count;
    }
    sync();
}

// The sweeps of one visit of a level with the thread's plans of both colours and, with a coarser level, the
// residual stored and restricted onto it.
template <bool ONES, int K, int NT, typename Sync>
__device__ __forceinline__ void visit_planned(const ColourPlan<ONES, K, NT> &red, const ColourPlan<ONES, K, NT> &black,
                                              const TailLevel &L, const TailDims &C, float *q, const float *rhs,
                                              float *res, float *crhs, int sweeps, float relaxation, int tid,
                                              Sync sync) {
    const LevelStrides st = level_strides(L);
    for (int s = 0; s < sweeps; s++) {
        plan_pass(red, q, rhs, st, relaxation);
        sync();
        plan_pass(black, q, rhs, st, relaxation);
        sync();
    }
    if (crhs != nullptr) {
        plan_residual(red, q, rhs, res, st);
        plan_residual(black, q, rhs, res, st);
        sync();
        tail_restrict<NT>(TailDims{{L.n[0], L.n[1], L.n[2]}, L.div_n1, L.div_n2}, C, res, crhs, tid, sync);
    }
}

template <bool ONES, int K, int NT, typename Sync>
__device__ __forceinline__ void visit(const TailLevel &L, const TailDims &C, float *q, const float *rhs, float *res,
                                      float *crhs, int sweeps, float relaxation, int tid, Sync sync) {
    ColourPlan<ONES, K, NT> red, black;
    red.build(L, 0, tid);
    black.build(L, 1, tid);
    visit_planned(red, black, L, C, q, rhs, res, crhs, sweeps, relaxation, tid, sync);
}

// prolongation (sample_index_space of the coarse correction at i / 2 - 1/4, sim/multigrid.py:63-68, with
// sim/sample.py:10-29's clamp, taps and nested lerps) read from a padded coarse q
__device__ __forceinline__ float prolong_padded(const float *cq, const TailDims &C, int i, int j, int k) {
    const float x = clamp_((float)i * 0.5f - 0.25f, 0.0f, (float)(C.n[0] - 1));
    const float y = clamp_((float)j * 0.5f - 0.25f, 0.0f, (float)(C.n[1] - 1));
    const float z = clamp_((float)k * 0.5f - 0.25f, 0.0f, (float)(C.n[2] - 1));
    const int a = (int)x, b = (int)y, c = (int)z;
    const int row = C.n[2] + 2, plane = (C.n[1] + 2) * row;
    const int di = a + 1 < C.n[0] ? plane : 0, dj = b + 1 < C.n[1] ? row : 0, dk = c + 1 < C.n[2] ? 1 : 0;
    const float tx = x - (float)a, ty = y - (float)b, tz = z - (float)c;
    const int o = pad_at(C.n, a, b, c), o1 = o + dk;
    const float c00 = lerp_(cq[o], cq[o + di], tx);
    const float c10 = lerp_(cq[o + dj], cq[o + di + dj], tx);
    const float c01 = lerp_(cq[o1], cq[o1 + di], tx);
    const float c11 = lerp_(cq[o1 + dj], cq[o1 + di + dj], tx);
    return lerp_(lerp_(c00, c10, ty), lerp_(c01, c11, ty), tz);
}

// prolong_add from level C into L (sim/multigrid.py:63-68, 145)
template <int NT, typename Sync>
__device__ __forceinline__ void tail_prolong(const TailDims &L, const TailDims &C, float *q, const float *cq,
                                             int tid, Sync sync) {
    const int cells = L.n[0] * L.n[1] * L.n[2];
    for (int c = tid; c < cells; c += NT) {
        const int r = L.div_n2(c), k = c - r * L.n[2];
        const int i = L.div_n1(r), j = r - i * L.n[1];
        const int p = pad_at(L.n, i, j, k);
        q[p] = q[p] + prolong_padded(cq, C, i, j, k);
    }
    sync();
}

__device__ __forceinline__ int coarsest_sweeps(const TailLevel &L) {
    const int longest = max(L.n[0], max(L.n[1], L.n[2]));
    return max(20, 4 * longest);  // COARSEST_SWEEPS, COARSEST_SWEEPS_PER_CELL (sim/multigrid.py:30-31)
}

// ONES: every tail level has fill_hi (1, 1, 1) (its in-domain couplings are all exactly 1). The first
// KEEP block levels build their plans once, on the way down, and keep them in registers for the way up.
template <bool ONES>
__device__ __forceinline__ void mg_tail(const TailLevels &args, float *__restrict__ q_top,
                                        const float *__restrict__ rhs_top, int sweeps, const double *gate = nullptr) {
    constexpr int K = ONES ? BLOCK_PLAN_ONES : BLOCK_PLAN;
    constexpr int KEEP = ONES ? 2 : 1;
    extern __shared__ float lds[];
    // The top level's rhs (and q) are read from global memory first, so their latency overlaps the staging of
    // the level table and the clearing of LDS below (a tail holds at most PLM_LDS_CELLS cells).
    constexpr int TOP_PER_THREAD = PLM_LDS_CELLS / TAIL_THREADS;
    static_assert(PLM_LDS_CELLS % TAIL_THREADS == 0, "the top level's cells split evenly over the threads");
    const int top_cells = args.level[0].g.dims[0] * args.level[0].g.dims[1] * args.level[0].g.dims[2];
    const bool top_q = args.zero_top == 0;
    float top_rhs[TOP_PER_THREAD], top_qv[TOP_PER_THREAD];
#pragma unroll
    for (int u = 0; u < TOP_PER_THREAD; u++) {
        const int c = (int)threadIdx.x + u * TAIL_THREADS;
        top_rhs[u] = c < top_cells ? rhs_top[c] : 0.0f;
        top_qv[u] = c < top_cells && top_q ? q_top[c] : 0.0f;
    }
    if (gated_off(gate)) return;  // read once the prefetch is in flight (uniform, before any barrier)
    // The level table is indexed by runtime level numbers; read from the kernel argument that way it would
    // be copied to per-thread scratch, so it is staged in LDS once.
    __shared__ TailLevels t;
    {
        static_assert(sizeof(TailLevels) % 4 == 0, "TailLevels is copied as 32-bit words");
        const int words = sizeof(TailLevels) / 4;
        const uint32_t *src = reinterpret_cast<const uint32_t *>(&args);
        uint32_t *dst = reinterpret_cast<uint32_t *>(&t);
        for (int w = threadIdx.x; w < words; w += TAIL_THREADS) dst[w] = src[w];
    }
    __syncthreads();
    for (int w = threadIdx.x; w < t.q_floats; w += TAIL_THREADS) lds[t.qoff[0] + w] = 0.0f;
    __syncthreads();
    const int tid = threadIdx.x;
    {
        const TailDims T = load_dims(t, 0);
#pragma unroll
        for (int u = 0; u < TOP_PER_THREAD; u++) {
            const int c = tid + u * TAIL_THREADS;
            if (c < top_cells) {
                const int r = T.div_n2(c), k = c - r * T.n[2];
                const int i = T.div_n1(r), j = r - i * T.n[1];
                const int p = pad_at(T.n, i, j, k);
                if (top_q) lds[t.qoff[0] + p] = top_qv[u];
                lds[t.roff[0] + p] = top_rhs[u];
            }
        }
    }
    __syncthreads();
    const int last = t.count - 1;
    const int block_levels = min(t.wave_from, last);
    float *res = lds + t.res_offset;
    auto q_of = [&](int m) { return lds + t.qoff[m]; };
    auto rhs_of = [&](int m) { return lds + t.roff[m]; };
    // down on the whole block: smooth, residual + restrict (sim/multigrid.py:137-143); the first KEEP levels'
    // plans stay for the way up
    ColourPlan<ONES, K, TAIL_THREADS> red0, black0, red1, black1;
    auto down_kept = [&](int m, ColourPlan<ONES, K, TAIL_THREADS> &red,
                         ColourPlan<ONES, K, TAIL_THREADS> &black) __attribute__((always_inline)) {
        const TailLevel L = load_level(t, m);
        red.build(L, 0, tid);
        black.build(L, 1, tid);
        visit_planned(red, black, L, load_dims(t, m + 1), q_of(m), rhs_of(m), res, rhs_of(m + 1), sweeps, 1.0f, tid,
                      BlockSync{});
    };
    auto up_kept = [&](int m, const ColourPlan<ONES, K, TAIL_THREADS> &red,
                       const ColourPlan<ONES, K, TAIL_THREADS> &black) __attribute__((always_inline)) {
        tail_prolong<TAIL_THREADS>(load_dims(t, m), load_dims(t, m + 1), q_of(m), q_of(m + 1), tid, BlockSync{});
        visit_planned(red, black, load_level(t, m), load_dims(t, m), q_of(m), rhs_of(m), res, nullptr, sweeps, 1.0f,
                      tid, BlockSync{});
    };
    if (block_levels > 0) down_kept(0, red0, black0);
    if (KEEP > 1 && block_levels > 1) down_kept(1, red1, black1);
    for (int m = KEEP; m < block_levels; m++)
        visit<ONES, K, TAIL_THREADS>(load_level(t, m), load_dims(t, m + 1), q_of(m), rhs_of(m), res, rhs_of(m + 1),
                                     sweeps, 1.0f, tid, BlockSync{});
    if (t.wave_from <= last) {
        // the smallest levels on wavefront 0: down, the coarsest solve, and back up
        if (tid < WAVE) {
            for (int m = t.wave_from; m < last; m++)
                visit<ONES, 1, WAVE>(load_level(t, m), load_dims(t, m + 1), q_of(m), rhs_of(m), res, rhs_of(m + 1),
                                     sweeps, 1.0f, tid, WaveSync{});
            const TailLevel B = load_level(t, last);
            visit<ONES, 1, WAVE>(B, load_dims(t, last), q_of(last), rhs_of(last), res, nullptr, coarsest_sweeps(B),
                                 1.5f, tid, WaveSync{});
            for (int m = last - 1; m >= t.wave_from; m--) {
                tail_prolong<WAVE>(load_dims(t, m), load_dims(t, m + 1), q_of(m), q_of(m + 1), tid, WaveSync{});
                visit<ONES, 1, WAVE>(load_level(t, m), load_dims(t, m), q_of(m), rhs_of(m), res, nullptr, sweeps,
                                     1.0f, tid, WaveSync{});
            }
        }
        __syncthreads();
    } else {
        const TailLevel B = load_level(t, last);
        visit<ONES, K, TAIL_THREADS>(B, load_dims(t, last), q_of(last), rhs_of(last), res, nullptr, coarsest_sweeps(B),
                                     1.5f, tid, BlockSync{});
    }
    // up on the whole block: prolong-add, smooth (sim/multigrid.py:145-147)
    for (int m = block_levels - 1; m >= KEEP; m--) {
        tail_prolong<TAIL_THREADS>(load_dims(t, m), load_dims(t, m + 1), q_of(m), q_of(m + 1), tid, BlockSync{});
        visit<ONES, K, TAIL_THREADS>(load_level(t, m), load_dims(t, m), q_of(m), rhs_of(m), res, nullptr, sweeps, 1.0f,
                                     tid, BlockSync{});
    }
    if (KEEP > 1 && block_levels > 1) up_kept(1, red1, black1);
    if (block_levels > 0) up_kept(0, red0, black0);
    {
        const TailDims T = load_dims(t, 0);
        const int cells = T.n[0] * T.n[1] * T.n[2];
        for (int c = tid; c < cells; c += TAIL_THREADS) {
            const int r = T.div_n2(c), k = c - r * T.n[2];
            const int i = T.div_n1(r), j = r - i * T.n[1];
            q_top[c] = lds[t.qoff[0] + pad_at(T.n, i, j, k)];
        }
    }
}

template <bool ONES>
__global__ __launch_bounds__(TAIL_THREADS) void k_mg_tail(TailLevels args, float *__restrict__ q_top,
                                                          const float *__restrict__ rhs_top, int sweeps,
                                                          const double *gate) {
    mg_tail<ONES>(args, q_top, rhs_top, sweeps, gate);
}

// ------------------------------------------------------------------ FP64 norms (tolerance mode)

constexpr int NORM_THREADS = 256;
constexpr int NORM_I = 4;  // cells along i per thread
constexpr int FINISH_THREADS = 1024;

inline dim3 norm_grid(const plm_grid_info &g) {
    const Dims n = cell_dims(g);
    return dim3((n.n2 + BX - 1) / BX, (n.n1 + BY - 1) / BY, (n.n0 + NORM_I - 1) / NORM_I);
}
inline int64_t norm_blocks(const plm_grid_info &g) {
    const dim3 grid = norm_grid(g);
    return (int64_t)grid.x * grid.y * grid.z;
}

// FP64 sum over the 64 lanes of a wavefront, in a fixed butterfly order
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int offset = 32; offset > 0; offset /= 2) v += __shfl_xor(v, offset, 64);
    return v;
}

// per block: sum of r^2 and rhs^2 over its 64 x 4 x NORM_I cells, in a fixed order (per thread along i, then
// a butterfly per wavefront, then the four wavefronts in order). Every load comes first and unconditionally
// (a neighbour outside the level reads the cell itself and is masked), so they issue as one batch; unit levels
// then need no branch per cell.
template <bool UNIT>
__device__ __forceinline__ void norm_partials(const LevelArgs &L, const float *__restrict__ q,
                                              const float *__restrict__ rhs, double *__restrict__ partials,
                                              dim3 block, dim3 grid, const double *gate = nullptr) {
    const Dims n = cell_dims(L.g);
    const int k = block.x * BX + threadIdx.x, j = block.y * BY + threadIdx.y;
    double rr = 0.0, bb = 0.0;
    if (k < n.n2 && j < n.n1) {
        const float scale = uniform(L.g.cell_m) * uniform(L.g.cell_m);  // (1 cell_m) cell_m on unit levels
        const uint32_t si = (uint32_t)(n.n1 * n.n2) * 4u, sj = (uint32_t)n.n2 * 4u;
        const int ib = block.z * NORM_I;
        const bool jm = j >= 1, jp = j + 1 < n.n1, km = k >= 1, kp = k + 1 < n.n2;
        const int couples_jk = unit_couples(L.g, 1, jm, jp) + unit_couples(L.g, 2, km, kp);
        const uint32_t dj_lo = jm ? sj : 0u, dj_hi = jp ? sj : 0u, dk_lo = km ? 4u : 0u, dk_hi = kp ? 4u : 0u;
        // rows past the level clamped to its last row, not counted. On unit levels the column of q from row
        // ib - 1 to ib + NORM_I (clamped into the level) gives each row its value and its i neighbours, so a
        // row's i neighbours are not loaded again: where the clamp applies the neighbour is masked.
        float qc[NORM_I], bc[NORM_I], qn[NORM_I][6];
        float column[NORM_I + 2];
        if (UNIT) {
#pragma unroll
            for (int m = 0; m < NORM_I + 2; m++) {
                const int i = max(0, min(ib - 1 + m, n.n0 - 1));
                column[m] = load_bytes(q, offset_of(n, i, j, k) * 4u);
            }
        }
#pragma unroll
        for (int t = 0; t < NORM_I; t++) {
            const int i = min(ib + t, n.n0 - 1);
            const uint32_t o = offset_of(n, i, j, k) * 4u;
            bc[t] = load_bytes(rhs, o);
            if (UNIT) {
                qc[t] = column[t + 1];
                qn[t][0] = column[t];
                qn[t][1] = column[t + 2];
                qn[t][2] = load_bytes(q, o - dj_lo);
                qn[t][3] = load_bytes(q, o + dj_hi);
                qn[t][4] = load_bytes(q, o - dk_lo);
                qn[t][5] = load_bytes(q, o + dk_hi);
            } else {
                qc[t] = load_bytes(q, o);
            }
        }
#pragma unroll
        for (int t = 0; t < NORM_I; t++) {
            const int i = ib + t;
            if (i >= n.n0) break;
            const float q_c = qc[t], rhs_c = bc[t];
            float r;
            if (UNIT) {
                const bool in[6] = {i >= 1, i + 1 < n.n0, jm, jp, km, kp};
                r = unit_residual(q_c, rhs_c, scale, qn[t], in, couples_jk + unit_couples(L.g, 0, in[0], in[1]));
            } else {
                float t0, t1;
                neighbor_sum<UNIT>(L, i, j, k, GlobalQ{q, n}, t0, t1);
                r = residual_value(L, i, j, k, q_c, rhs_c, t0, t1);
            }
            rr += (double)r * (double)r;
            bb += (double)rhs_c * (double)rhs_c;
        }
    }
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    // the device-side gate is read only now, its latency hidden behind the loads (uniform: the whole block
    // leaves before the barrier)
    if (gated_off(gate)) return;
    __shared__ double srr[NORM_THREADS / 64], sbb[NORM_THREADS / 64];
    const int wave = threadIdx.y;  // blocks are 64 x 4: one wavefront per row of j
    if (threadIdx.x == 0) {
        srr[wave] = rr;
        sbb[wave] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const int64_t b = ((int64_t)block.z * grid.y + block.y) * grid.x + block.x;
        partials[2 * b] = ((srr[0] + srr[1]) + srr[2]) + srr[3];
        partials[2 * b + 1] = ((sbb[0] + sbb[1]) + sbb[2]) + sbb[3];
    }
}

template <bool UNIT>
__global__ __launch_bounds__(NORM_THREADS) void k_norm_partials(LevelArgs L, const float *__restrict__ q,
                                                                const float *__restrict__ rhs,
                                                                double *__restrict__ partials, const double *gate) {
    norm_partials<UNIT>(L, q, rhs, partials, blockIdx, gridDim, gate);
}

// scratch layout: [0] relative residual, [1] gate (1 active / 0 converged), [2] |r|^2, [3] |rhs|^2,
// [4 ...] 2 per block partials. `mode`: 0 plain norms (no gate change), 1 initial check, 2 after a cycle.
__device__ __forceinline__ void norm_finish(double *scratch, int64_t blocks, double tolerance, int mode,
                                            int32_t *cycles_run) {
    const double *partials = scratch + 4;
    // fixed order: each thread's blocks in turn, a butterfly per wavefront, then the wavefronts in order (two
    // block barriers where an LDS tree over every thread took ten)
    constexpr int WAVES = FINISH_THREADS / 64;
    __shared__ double srr[WAVES], sbb[WAVES];
    double rr = 0.0, bb = 0.0;
    for (int64_t b = threadIdx.x; b < blocks; b += FINISH_THREADS) {
        rr += partials[2 * b];
        bb += partials[2 * b + 1];
    }
    // a closed gate (read once the partials are in flight; every thread reads the same word, which only thread
    // 0 writes, after the barrier below) leaves the norms as they are
    if (mode == 2 && scratch[1] == 0.0) return;
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    if ((threadIdx.x & 63) == 0) {
        srr[threadIdx.x / 64] = rr;
        sbb[threadIdx.x / 64] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < WAVES; w++) {
            srr[0] += srr[w];
            sbb[0] += sbb[w];
        }
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

__global__ __launch_bounds__(FINISH_THREADS) void k_norm_finish(double *scratch, int64_t blocks, double tolerance,
                                                                int mode, int32_t *cycles_run) {
    norm_finish(scratch, blocks, tolerance, mode, cycles_run);
}

// ------------------------------------------------------------------ host orchestration

int launch_norms(const LevelArgs &L, const float *q, const float *rhs, double *scratch, const double *gate,
                 double tolerance, int mode, int32_t *cycles_run, hipStream_t s) {
    if (L.unit)
        k_norm_partials<true><<<norm_grid(L.g), block3(), 0, s>>>(L, q, rhs, scratch + 4, gate);
    else
        k_norm_partials<false><<<norm_grid(L.g), block3(), 0, s>>>(L, q, rhs, scratch + 4, gate);
    PLM_TRY(check_launch("norm partials"));
    k_norm_finish<<<1, FINISH_THREADS, 0, s>>>(scratch, norm_blocks(L.g), tolerance, mode, cycles_run);
    return check_launch("norm finish");
}

template <bool UNIT, typename T = float>
int launch_tile(const LevelArgs &L, int start, const T *src, T *dst, const T *rhs, const T *coarse, Dims nc,
                float relaxation, const double *gate, hipStream_t s, int shape = -1) {
    const Dims n = cell_dims(L.g);
    if (shape < 0) shape = tile_shape_for(n);
    const TileShape &ts = TILE_SHAPES[shape];
    const int tiles_k = (n.n2 + TK - 1) / TK, tiles_j = (n.n1 + ts.tj - 1) / ts.tj, tiles_i = (n.n0 + ts.ti - 1) / ts.ti;
    const dim3 grid((unsigned)(tiles_k * tiles_j * tiles_i));
    const FastDiv div_k = fast_div(tiles_k), div_j = fast_div(tiles_j);
    // family "rb_sweep_tile_top": the plain FP32 two-colour sweep (q in, q out) on the simulation grid
    TimingScope launch_timing(!UNIT || sizeof(T) != 4 ? "rb_sweep_tile"
                              : start == FROM_Q      ? "rb_sweep_tile_top"
                                                     : "rb_sweep_tile_top_start",
                              s, true);
    const hipEvent_t e0 = launch_timing.start(), e1 = launch_timing.stop();
    if (start == FROM_ZERO)
        launch_tile_shape<UNIT, FROM_ZERO, T>(shape, grid, s, e0, e1, L, src, dst, rhs, coarse, nc, relaxation, gate,
                                             div_k, div_j);
    else if (start == FROM_PROLONG)
        launch_tile_shape<UNIT, FROM_PROLONG, T>(shape, grid, s, e0, e1, L, src, dst, rhs, coarse, nc, relaxation, gate,
                                             div_k, div_j);
    else
        launch_tile_shape<UNIT, FROM_Q, T>(shape, grid, s, e0, e1, L, src, dst, rhs, coarse, nc, relaxation, gate,
                                             div_k, div_j);
    return check_launch("rb sweep tile");
}

constexpr int64_t PROLONG_ROWS_CELLS = int64_t(1) << 20;

// `sweeps` sweeps on one level, the result in q. Before the first: FROM_Q uses q; FROM_ZERO treats q as 0
// (q need not be cleared); FROM_PROLONG adds the prolongation of `coarse` to q. With a level-sized scratch the
// sweeps ping-pong q -> scratch -> q in two-colour tiles (an odd count ends with a copy back); without one,
// the start is applied by its own kernel and the sweeps run colour by colour in place.
int launch_sweeps(const LevelArgs &L, float *q, const float *rhs, float *scratch, int sweeps, float relaxation,
                  int start, const float *coarse, Dims nc, const double *gate, hipStream_t s) {
    const Dims n = cell_dims(L.g);
    TimingScope t("rb_sweeps", s);
    if (scratch == nullptr || sweeps == 0) {
        if (start == FROM_ZERO) {
            k_zero<<<grid1(n.count()), 256, 0, s>>>(n.count(), q, gate);
            PLM_TRY(check_launch("zero level"));
        } else if (start == FROM_PROLONG) {
            k_prolong_add<<<grid3(n), block3(), 0, s>>>(coarse, nc, q, n, gate);
            PLM_TRY(check_launch("prolong add"));
        }
        for (int sw = 0; sw < sweeps; sw++)
            for (int color = 0; color < 2; color++) {
                if (L.unit)
                    k_rb_sweep<true><<<grid3(n), block3(), 0, s>>>(L, q, rhs, color, relaxation, gate);
                else
                    k_rb_sweep<false><<<grid3(n), block3(), 0, s>>>(L, q, rhs, color, relaxation, gate);
                PLM_TRY(check_launch("rb sweep"));
            }
        return 0;
    }
    // On large levels the prolongation is cheaper as its own row kernel than sampled inside the tile load.
    if (start == FROM_PROLONG && n.count() >= PROLONG_ROWS_CELLS) {
        k_prolong_rows4<float><<<prolong_rows4_grid(n), block3(), 0, s>>>(coarse, nc, q, n, gate);
        PLM_TRY(check_launch("prolong rows"));
        start = FROM_Q;
    }
    for (int sw = 0; sw < sweeps; sw++) {
        const bool to_scratch = sw % 2 == 0;
        const float *src = to_scratch ? q : scratch;
        float *dst = to_scratch ? scratch : q;
        const int first = sw == 0 ? start : FROM_Q;
        if (L.unit)
            PLM_TRY(launch_tile<true>(L, first, src, dst, rhs, coarse, nc, relaxation, gate, s));
        else
            PLM_TRY(launch_tile<false>(L, first, src, dst, rhs, coarse, nc, relaxation, gate, s));
    }
    if (sweeps % 2 == 1) {
        k_copy<<<grid1(n.count()), 256, 0, s>>>(n.count(), scratch, q, gate);
        PLM_TRY(check_launch("copy back"));
    }
    return 0;
}

// Times every tile shape of the plain two-colour sweep on one level and pins the fastest. Each launch carries
// its own begin and end timestamps (events bound to the dispatch through hipExtLaunchKernelGGL), so a shape is
// judged by its kernel time, not by the gaps between back-to-back launches; the shapes take turns, one launch
// each per pass (drift in clocks or caches touches all of them alike), and each one's figure is the median of
// its 2 x repeats launches after one untimed pass.
template <bool UNIT>
int autotune_sweep(const LevelArgs &L, float *q, const float *rhs, float *scratch, int repeats, float *times_us,
                   int32_t *chosen, hipStream_t s) {
    const Dims n = cell_dims(L.g), none{1, 1, 1};
    const int launches = std::min(2 * repeats, 64);
    std::vector<hipEvent_t> events((size_t)2 * N_TILE_SHAPES * launches, nullptr);
    int rc = 0;
    for (hipEvent_t &e : events)
        if (rc == 0 && hipEventCreate(&e) != hipSuccess) rc = fail(PLM_EINVAL, "plm_autotune_sweep: event creation failed");
    for (int r = 0; r <= launches && rc == 0; r++) {
        for (int v = 0; v < N_TILE_SHAPES && rc == 0; v++) {
            const TileShape &ts = TILE_SHAPES[v];
            const int tiles_k = (n.n2 + TK - 1) / TK, tiles_j = (n.n1 + ts.tj - 1) / ts.tj;
            const int tiles_i = (n.n0 + ts.ti - 1) / ts.ti;
            const bool even = (r * N_TILE_SHAPES + v) % 2 == 0;
            const size_t slot = 2 * ((size_t)v * launches + (size_t)(r - 1));
            const hipEvent_t e0 = r > 0 ? events[slot] : nullptr, e1 = r > 0 ? events[slot + 1] : nullptr;
            launch_tile_shape<UNIT, FROM_Q, float>(v, dim3((unsigned)(tiles_k * tiles_j * tiles_i)), s, e0, e1, L,
                                                   even ? q : scratch, even ? scratch : q, rhs, nullptr, none, 1.0f,
                                                   nullptr, fast_div(tiles_k), fast_div(tiles_j));
            rc = check_launch("plm_autotune_sweep");
        }
    }
    if (rc == 0 && hipStreamSynchronize(s) != hipSuccess) rc = fail(PLM_EINVAL, "plm_autotune_sweep: synchronise failed");
    int best = DEFAULT_TILE_SHAPE;
    float best_us = 0.0f;
    for (int v = 0; v < N_TILE_SHAPES && rc == 0; v++) {
        std::vector<float> us((size_t)launches);
        for (int r = 0; r < launches && rc == 0; r++) {
            const size_t slot = 2 * ((size_t)v * launches + (size_t)r);
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, events[slot], events[slot + 1]) != hipSuccess)
                rc = fail(PLM_EINVAL, "plm_autotune_sweep: event timing failed");
            us[(size_t)r] = ms * 1000.0f;
        }
        std::nth_element(us.begin(), us.begin() + launches / 2, us.end());
        const float median = us[(size_t)launches / 2];
        if (times_us) times_us[v] = median;
        if (v == 0 || median < best_us) {
            best = v;
            best_us = median;
        }
    }
    for (hipEvent_t e : events)
        if (e != nullptr) (void)hipEventDestroy(e);
    PLM_TRY(rc);
    *chosen = best;
    return plm_set_sweep_tile(L.g.dims, best);
}

struct Hierarchy {
    int n_levels;
    LevelArgs level[16];
    float *const *q;
    float *const *rhs;
    float *const *residual;
    int sweeps;
    int tail_start;
    TailLevels tail;
    size_t tail_bytes;
    bool tail_ones;  // every tail level has fill_hi (1, 1, 1): k_mg_tail<true>
    const double *gate;
};

// Lets a tail launch use more than the default 64 KB of dynamic LDS (gfx950 has 160 KB per CU).
template <bool ONES>
int allow_tail_lds(size_t bytes) {
    static size_t allowed = 0;
    if (bytes <= 65536 || bytes <= allowed) return 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mg_tail<ONES>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
        return fail(PLM_EINVAL, "multigrid tail: %zu bytes of LDS refused", bytes);
    allowed = bytes;
    return 0;
}

// A flat level below the simulation grid with two sweeps a side and a level of its own below it smooths in LDS
// (k_level_smooth)
bool smooths_in_lds(const Hierarchy &h, int index) {
    const LevelArgs &L = h.level[index];
    return !L.unit && L.flat == 1 && h.sweeps == 2 && h.residual[index] != nullptr && index + 1 < h.n_levels;
}

// _v_cycle, sim/multigrid.py:130-147, from level `index`; `cleared`: this level's q is the reference's
// freshly zeroed coarse q (the caller did not write it)
int v_cycle(Hierarchy &h, int index, bool cleared, hipStream_t s) {
    if (index == h.tail_start) {
        h.tail.zero_top = cleared ? 1 : 0;
        if (h.tail_ones)
            k_mg_tail<true><<<1, TAIL_THREADS, h.tail_bytes, s>>>(h.tail, h.q[index], h.rhs[index], h.sweeps, h.gate);
        else
            k_mg_tail<false><<<1, TAIL_THREADS, h.tail_bytes, s>>>(h.tail, h.q[index], h.rhs[index], h.sweeps, h.gate);
        return check_launch("multigrid tail");
    }
    const LevelArgs &L = h.level[index];
    const Dims none{1, 1, 1};
    const int start = cleared ? FROM_ZERO : FROM_Q;
    if (index == h.n_levels - 1) {  // coarsest level outside the tail (only when the tail is disabled)
        const int longest = max(L.g.dims[0], max(L.g.dims[1], L.g.dims[2]));
        return launch_sweeps(L, h.q[index], h.rhs[index], h.residual[index], max(20, 4 * longest), 1.5f, start,
                             nullptr, none, h.gate, s);
    }
    const Dims nc = cell_dims(h.level[index + 1].g);
    // residual[index] doubles as the ping-pong buffer of this level's sweeps; the coarse q is cleared by the
    // next level's first sweep (or the tail), so it is not written here
    if (cleared && smooths_in_lds(h, index)) {
        // A small level: both pre-smoothing sweeps from zero with the residual and restriction into residual (the
        // ping-pong buffer), then the prolongation and both post-smoothing sweeps from residual into q. A big
        // level (where a two-sweep box would repeat too much of the halo work) one sweep a launch: from zero
        // into residual, the second with the residual and restriction into q; the prolongation with the first
        // post-smoothing sweep into residual, the second into q.
        float *q = h.q[index], *scratch = h.residual[index];
        const bool small = !big_smooth_tiles(cell_dims(L.g));
        if (small) {
            const int pre = launch_level_smooth<2, FROM_ZERO, true>(L, nullptr, scratch, h.rhs[index], nullptr, none,
                                                                    h.rhs[index + 1], nc, 1.0f, h.gate, s);
            PLM_TRY(pre);
        } else {
            const int first = launch_level_smooth<1, FROM_ZERO, false>(L, nullptr, scratch, h.rhs[index], nullptr,
                                                                       none, nullptr, none, 1.0f, h.gate, s);
            PLM_TRY(first);
            const int pre = launch_level_smooth<1, FROM_Q, true>(L, scratch, q, h.rhs[index], nullptr, none,
                                                                 h.rhs[index + 1], nc, 1.0f, h.gate, s);
            PLM_TRY(pre);
        }
        PLM_TRY(v_cycle(h, index + 1, true, s));
        if (small)
            return launch_level_smooth<2, FROM_PROLONG, false>(L, scratch, q, h.rhs[index], h.q[index + 1], nc,
                                                               nullptr, none, 1.0f, h.gate, s);
        const int post = launch_level_smooth<1, FROM_PROLONG, false>(L, q, scratch, h.rhs[index], h.q[index + 1], nc,
                                                                     nullptr, none, 1.0f, h.gate, s);
        PLM_TRY(post);
        return launch_level_smooth<1, FROM_Q, false>(L, scratch, q, h.rhs[index], nullptr, none, nullptr, none, 1.0f,
                                                     h.gate, s);
    }
    PLM_TRY(launch_sweeps(L, h.q[index], h.rhs[index], h.residual[index], h.sweeps, 1.0f, start, nullptr, none,
                          h.gate, s));
    if (L.unit)
        k_residual_restrict<true, float><<<grid3(nc), block3(), 0, s>>>(L, h.q[index], h.rhs[index], h.rhs[index + 1],
                                                                nullptr, nc, h.gate);
    else
        k_residual_restrict_children<float><<<children_grid(nc), 256, 0, s>>>(L, h.q[index], h.rhs[index],
                                                                             h.rhs[index + 1], nc, h.gate);
    PLM_TRY(check_launch("residual restrict"));
    PLM_TRY(v_cycle(h, index + 1, true, s));
    return launch_sweeps(L, h.q[index], h.rhs[index], h.residual[index], h.sweeps, 1.0f, FROM_PROLONG,
                         h.q[index + 1], nc, h.gate, s);
}

// ------------------------------------------------------------------ mixed-precision V-cycle

// The coarsest level of the mixed-precision solve in FP64: red_black_sweeps (sim/pressure.py:80-92) with the
// coarsest level's max(20, 4 longest) sweeps at 1.5 (sim/multigrid.py:131-135), q and rhs staged in LDS as
// doubles and neighbor_term's couplings (sim/pressure.py:52-69) formed in double from the level's geometry.
constexpr int COARSE_F64_THREADS = 1024;
constexpr int COARSE_F64_CELLS = 4096;

__device__ __forceinline__ double fill_f64(const plm_grid_info &g, int a, int index) {
    return index == g.dims[a] - 1 ? (double)g.fill_hi[a] : 1.0;
}

template <typename T, int CPT>
__global__ __launch_bounds__(COARSE_F64_THREADS) void k_coarse_f64(plm_grid_info g, T *__restrict__ q,
                                                                   const T *__restrict__ rhs, int sweeps,
                                                                   int from_zero) {
    extern __shared__ double dl[];
    const Dims n = cell_dims(g);
    const int cells = n.count();
    double *dq = dl, *drhs = dl + cells;
    for (int c = threadIdx.x; c < cells; c += COARSE_F64_THREADS) {
        dq[c] = from_zero ? 0.0 : (double)to_f32(q[c]);
        drhs[c] = (double)to_f32(rhs[c]);
    }
    __syncthreads();
    const double cm = (double)g.cell_m, gap = (double)g.ghost_gap;
    // Each thread's CPT cells (CPT x COARSE_F64_THREADS >= cells), planned once: the six neighbours in the order
    // x-, x+, y-, y+, z-, z+ (an absent one points at the cell itself with coupling 0, which adds +0 to a sum
    // that starts at +0), the sum of couplings with open-face ghosts, and vol cm² rhs.
    int nb[CPT][6], colour[CPT];
    double cp[CPT][6], t1[CPT], src[CPT];
#pragma unroll
    for (int u = 0; u < CPT; u++) {
        const int c = min((int)threadIdx.x + u * COARSE_F64_THREADS, cells - 1);
        const int k = c % n.n2, r = c / n.n2, j = r % n.n1, i = r / n.n1;
        const int cell[3] = {i, j, k};
        colour[u] = (int)threadIdx.x + u * COARSE_F64_THREADS < cells ? (i + j + k) & 1 : 2;
        double sum = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const int b = a == 0 ? 1 : 0, cc = a == 2 ? 1 : 2;
            const double area = 1.0 * fill_f64(g, b, cell[b]) * fill_f64(g, cc, cell[cc]);
            const double own = fill_f64(g, a, cell[a]);
#pragma unroll
            for (int side = 0; side < 2; side++) {
                int other[3] = {i, j, k};
                other[a] += side == 0 ? -1 : 1;
                nb[u][2 * a + side] = c;
                cp[u][2 * a + side] = 0.0;
                if (other[a] >= 0 && other[a] < g.dims[a]) {
                    const double coupling = area / (0.5 * (own + fill_f64(g, a, other[a])));
                    nb[u][2 * a + side] = n.at(other[0], other[1], other[2]);
                    cp[u][2 * a + side] = coupling;
                    sum += coupling;
                } else if (side == 0 ? g.open_lo[a] == 1 : g.open_hi[a] == 1) {
                    sum += area / (0.5 * own + gap);
                }
            }
        }
        t1[u] = sum;
        const double vol = fill_f64(g, 0, i) * fill_f64(g, 1, j) * fill_f64(g, 2, k);
        src[u] = vol * cm * cm * drhs[c];
        if (sum == 0.0) colour[u] = 2;  // nothing couples: q stays
    }
    for (int s = 0; s < sweeps; s++)
        for (int color = 0; color < 2; color++) {
#pragma unroll
            for (int u = 0; u < CPT; u++) {
                if (colour[u] != color) continue;
                const int c = (int)threadIdx.x + u * COARSE_F64_THREADS;
                double t0 = 0.0;
#pragma unroll
                for (int m = 0; m < 6; m++) t0 += cp[u][m] * dq[nb[u][m]];
                const double target = (t0 - src[u]) / t1[u];
                dq[c] = dq[c] + 1.5 * (target - dq[c]);
            }
            __syncthreads();
        }
    for (int c = threadIdx.x; c < cells; c += COARSE_F64_THREADS) q[c] = from_f32<T>((float)dq[c]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_t(int64_t count, const T *__restrict__ src, T *__restrict__ dst) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x)
        dst[c] = src[c];
}

// dst = T(src * scale): the FP32 residual into the storage type of the correction solve
template <typename T>
__global__ __launch_bounds__(256) void k_to_storage(int64_t count, const float *__restrict__ src, float scale,
                                                    T *__restrict__ dst) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x)
        dst[c] = from_f32<T>(src[c] * scale);
}

// x += float(e) * inv_scale: the correction back into the FP32 solution
template <typename T>
__global__ __launch_bounds__(256) void k_add_storage(int64_t count, float *__restrict__ x, const T *__restrict__ e,
                                                     float inv_scale) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x)
        x[c] = x[c] + to_f32(e[c]) * inv_scale;
}

// x + float(e) * inv_scale at a cell: the FP32 pressure with the last round's correction added (x alone
// without CORRECT), the value k_add_storage would store there
template <bool CORRECT, typename T>
struct CorrectedQ {
    const float *x;
    const T *e;
    float inv_scale;
    Dims n;
    __device__ __forceinline__ float at(int c) const {
        const float v = load_at(x, (uint32_t)c);
        return CORRECT ? v + load_at(e, (uint32_t)c) * inv_scale : v;
    }
    __device__ float operator()(int i, int j, int k) const { return at(n.at(i, j, k)); }
};

// One refinement step, fused: x_out = x + float(e) * inv_scale, r = rhs - A x_out in FP32 (residual_kernel,
// sim/multigrid.py:36-43, the same values as k_residual on the corrected pressure), r_out = T(r * r_scale),
// and per block the FP64 sums of r^2 and rhs^2 in k_norm_partials' order. Blocks of 64 x 4 (k, j) threads,
// NORM_I cells along i per thread; the corrected column i - 1 ... i + NORM_I of a thread is formed once.
// x and x_out must not overlap (the neighbours read x while x_out is written); x_out may be null when e is.
template <bool UNIT, bool CORRECT, typename T>
__global__ __launch_bounds__(NORM_THREADS) void k_refine_residual(LevelArgs L, const float *__restrict__ x,
                                                                  const T *__restrict__ e, float inv_scale,
                                                                  const float *__restrict__ rhs,
                                                                  float *__restrict__ x_out, T *__restrict__ r_out,
                                                                  float r_scale, double *__restrict__ partials) {
    const Dims n = cell_dims(L.g);
    const int k = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    double rr = 0.0, bb = 0.0;
    if (k < n.n2 && j < n.n1) {
        const CorrectedQ<CORRECT, T> q{x, e, inv_scale, n};
        const float scale = uniform(L.g.cell_m) * uniform(L.g.cell_m);  // (1 cell_m) cell_m on unit levels
        const int si = n.n1 * n.n2, sj = n.n2;
        const int ib = blockIdx.z * NORM_I;
        const bool jm = j >= 1, jp = j + 1 < n.n1, km = k >= 1, kp = k + 1 < n.n2;
        const int couples_jk = unit_couples(L.g, 1, jm, jp) + unit_couples(L.g, 2, km, kp);
        const int dj_lo = jm ? sj : 0, dj_hi = jp ? sj : 0, dk_lo = km ? 1 : 0, dk_hi = kp ? 1 : 0;
        // column[t + 1] is the corrected q at i = ib + t (rows past the level clamped to its last row). Every
        // load is unconditional (a neighbour outside the level reads the cell itself and is masked), so the
        // loads issue as one batch.
        float column[NORM_I + 2], bc[NORM_I], side[NORM_I][4];
#pragma unroll
        for (int t = -1; t <= NORM_I; t++) {
            const int i = max(0, min(ib + t, n.n0 - 1));
            column[t + 1] = q.at(n.at(i, j, k));
        }
#pragma unroll
        for (int t = 0; t < NORM_I; t++) {
            const int c = n.at(min(ib + t, n.n0 - 1), j, k);
            bc[t] = load_at(rhs, (uint32_t)c);
            if (UNIT) {
                side[t][0] = q.at(c - dj_lo);
                side[t][1] = q.at(c + dj_hi);
                side[t][2] = q.at(c - dk_lo);
                side[t][3] = q.at(c + dk_hi);
            }
        }
#pragma unroll
        for (int t = 0; t < NORM_I; t++) {
            const int i = ib + t;
            if (i >= n.n0) break;
            const float q_c = column[t + 1], rhs_c = bc[t];
            float r;
            if (UNIT) {
                const float qn[6] = {column[t], column[t + 2], side[t][0], side[t][1], side[t][2], side[t][3]};
                const bool in[6] = {i >= 1, i + 1 < n.n0, jm, jp, km, kp};
                r = unit_residual(q_c, rhs_c, scale, qn, in, couples_jk + unit_couples(L.g, 0, in[0], in[1]));
            } else {
                float t0, t1;
                neighbor_sum<UNIT>(L, i, j, k, q, t0, t1);
                r = residual_value(L, i, j, k, q_c, rhs_c, t0, t1);
            }
            const uint32_t c = (uint32_t)(i * si + j * sj + k);
            if (x_out != nullptr) x_out[c] = q_c;
            if (r_out != nullptr) r_out[c] = from_f32<T>(r * r_scale);
            rr += (double)r * (double)r;
            bb += (double)rhs_c * (double)rhs_c;
        }
    }
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    __shared__ double srr[NORM_THREADS / 64], sbb[NORM_THREADS / 64];
    const int wave = threadIdx.y;
    if (threadIdx.x == 0) {
        srr[wave] = rr;
        sbb[wave] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const int64_t block = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        partials[2 * block] = ((srr[0] + srr[1]) + srr[2]) + srr[3];
        partials[2 * block + 1] = ((sbb[0] + sbb[1]) + sbb[2]) + sbb[3];
    }
}

template <typename T>
int launch_refine(const LevelArgs &L, const float *x, const void *e, float inv_scale, const float *rhs, float *x_out,
                  void *r_out, float r_scale, double *scratch, hipStream_t s) {
    const T *et = static_cast<const T *>(e);
    T *rt = static_cast<T *>(r_out);
    const dim3 grid = norm_grid(L.g), block = block3();
    if (L.unit && et)
        k_refine_residual<true, true, T><<<grid, block, 0, s>>>(L, x, et, inv_scale, rhs, x_out, rt, r_scale, scratch + 4);
    else if (L.unit)
        k_refine_residual<true, false, T><<<grid, block, 0, s>>>(L, x, et, inv_scale, rhs, x_out, rt, r_scale, scratch + 4);
    else if (et)
        k_refine_residual<false, true, T><<<grid, block, 0, s>>>(L, x, et, inv_scale, rhs, x_out, rt, r_scale, scratch + 4);
    else
        k_refine_residual<false, false, T><<<grid, block, 0, s>>>(L, x, et, inv_scale, rhs, x_out, rt, r_scale, scratch + 4);
    PLM_TRY(check_launch("refine residual"));
    k_norm_finish<<<1, FINISH_THREADS, 0, s>>>(scratch, norm_blocks(L.g), 0.0, 0, nullptr);
    return check_launch("norm finish");
}

// ------------------------------------------------------------------ packed FP16 sweep (simulation level)
//
// The two-colour tile sweep of an FP16 correction solve on a flat level (the simulation level and the coarser
// levels of even-sized grids) with its arithmetic in packed FP16: a lane updates two cells of one colour at once (v_pk_add_f16, v_pk_fma_f16), which halves the
// instructions, the LDS accesses and the LDS footprint of the FP32-arithmetic tile sweep, whose cost at large
// sizes is its issue and latency, not its HBM bytes. The update is red_black_sweep_kernel's at relaxation 1
// (sim/pressure.py:86-92), q = (sum of the six neighbours - cell_m² rhs) / t1, evaluated as
// sum * (1 / t1) + (-cell_m² rhs / t1) with both per-cell factors formed in FP32 and rounded to FP16 once. A
// neighbour outside the level is +0 in the tile, and t1 counts the level's neighbours plus the open boundary
// faces' ghost couplings (1 on the simulation level, 1 / (0.5 + ghost_gap) below it). The results are FP16 roundings of the same sweep, not the FP32-arithmetic sweep's
// bits.
//
// LDS rows, (TI + 4) x (TJ + 4) of them, hold a row's 68 cells k0 - 2 .. k0 + 65 as halves, colour-split: the
// cell kk = k - k0 + 2 at half (kk & 1) PK_HP + (kk >> 1) + 1. The 32 cells of one parity inside the tile
// (kk = 2 .. 65) are then dwords 1 .. 16 of their part, a lane's pair one dword: its i and j neighbour pairs
// are the same dword of the rows around, one k-neighbour pair the other part's dword of the same index and the
// other a half-dword off (v_alignbit of two dwords). Rows are PK_RW halves apart, so that the four rows a
// wavefront updates together (j, j + 2, j + 4, j + 6 of one i and parity) start 16 banks apart.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
constexpr int PK_HP = 36, PK_RW = 80;

__device__ __forceinline__ half2_t as_half2(uint32_t v) { return __builtin_bit_cast(half2_t, v); }
__device__ __forceinline__ uint32_t as_bits(half2_t v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t pack_half2(float lo, float hi) {
    return as_bits(half2_t{(_Float16)lo, (_Float16)hi});
}
__device__ __forceinline__ float half_lo(uint32_t v) { return (float)as_half2(v).x; }
__device__ __forceinline__ float half_hi(uint32_t v) { return (float)as_half2(v).y; }
__device__ __forceinline__ int pk_slot(int kk) { return (kk & 1) * PK_HP + (kk >> 1) + 1; }

// levels the packed sweep takes: flat levels (fill_hi (1, 1, 1): every coupling between two cells and every
// volume 1, an open boundary face coupling by the level's ghost value; the simulation level is one), rows of
// whole 4-cell groups (so a row's 8-byte loads are aligned and a group lies wholly inside or outside the
// level) and every cell coupled to something
bool packed_f16_level(const LevelArgs &L) {
    const Dims n = cell_dims(L.g);
    return (L.unit || L.flat == 1) && n.n2 % 4 == 0 && n.n0 >= 2 && n.n1 >= 2 && n.n2 >= 4;
}

template <int TI, int TJ, int THREADS, int START, bool EDGE>
__device__ __forceinline__ void sweep_tile_f16x2(const LevelArgs &L, uint32_t *__restrict__ s,
                                                 const __half *__restrict__ q_in, __half *__restrict__ q_out,
                                                 const __half *__restrict__ rhs, int i0, int j0, int k0) {
    constexpr int WAVES = THREADS / 64, LI = TI + 2 * TH, LJ = TJ + 2 * TH, RI = TI + 2, RJ = TJ + 2;
    constexpr int ROWD = PK_RW / 2, PARTD = PK_HP / 2;  // dwords per row, dwords from a row's part 0 to part 1
    constexpr int RED_GROUPS = RI * RJ / 4, BLACK_GROUPS = TI * TJ / 4;
    constexpr int RED_ITERS = (RED_GROUPS + WAVES - 1) / WAVES, BLACK_ITERS = (BLACK_GROUPS + WAVES - 1) / WAVES;
    constexpr int LOAD_GROUPS = LI * LJ / 4, LOAD_ITERS = (LOAD_GROUPS + WAVES - 1) / WAVES;
    constexpr int STORE_ITERS = (BLACK_GROUPS + WAVES - 1) / WAVES;
    constexpr int HALO = LI * LJ * 2, HALO_ITERS = (HALO + THREADS - 1) / THREADS;
    static_assert(TJ == 8 && RI % 2 == 0 && RI * RJ <= THREADS,
                  "rows group by four; one thread per ring row for the k-halo red cells");
    uint16_t *s16 = reinterpret_cast<uint16_t *>(s);
    const Dims n = cell_dims(L.g);
    const int lane = threadIdx.x & 63, wave = uniform((int)(threadIdx.x >> 6));
    const int g = lane >> 4, t = lane & 15;  // the lane's row of a group of four, its pair (dword) in the row
    const float cm2 = uniform(L.g.cell_m) * uniform(L.g.cell_m);
    const float ghost = uniform(L.ghost);
    bool open_lo[3], open_hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        open_lo[a] = uniform(L.g.open_lo[a]) == 1;
        open_hi[a] = uniform(L.g.open_hi[a]) == 1;
    }
    auto row_inside = [&](int i, int j) { return (unsigned)i < (unsigned)n.n0 && (unsigned)j < (unsigned)n.n1; };
    auto couple = [&](bool in, bool open) { return in ? 1.0f : open ? ghost : 0.0f; };
    auto t1_xy = [&](int i, int j) {
        if (!EDGE) return 4.0f;  // a ring row of an inner tile lies strictly inside the level
        return ((couple(i > 0, open_lo[0]) + couple(i < n.n0 - 1, open_hi[0])) + couple(j > 0, open_lo[1])) +
               couple(j < n.n1 - 1, open_hi[1]);
    };
    auto t1_z = [&](int k) { return couple(k > 0, open_lo[2]) + couple(k < n.n2 - 1, open_hi[2]); };
    auto row_at = [&](int i, int j) { return (uint32_t)((i * n.n1 + j) * n.n2); };
    // The lane's row in group G of a colour: its pair's dword `base` (dword 1 + t of part p), the row's parity
    // p and its cells' (i, j). Black group G: tile row i0 + (G >> 1), j = j0 + 2 g + (G & 1), so a group's four
    // rows are j, j + 2, j + 4, j + 6 of one i and parity. Red groups cover the (TI + 2) x 10 ring rows: the
    // same for G < 2 RI (ring i = i0 - 1 + (G >> 1), j = j0 - 1 + 2 g + (G & 1)), then the ring's last two
    // columns (j = j0 + 7 + c) in groups of four (i, c) pairs. The wave-uniform part of each index is a scalar
    // (G is), the lane's part a constant.
    struct Row {
        int base, p, i, j;
    };
    const int lane_pair = 2 * g * ROWD + 1 + t;
    const int left_row = (g >> 1) * LJ + (g & 1), left_p = ((g >> 1) + (g & 1)) & 1;
    auto row_of = [&](int colour, int G) {
        Row r;
        if (colour == 1 || G < 2 * RI) {
            const int ii = (G >> 1) + (colour == 1 ? TH : 1), jj = (G & 1) + (colour == 1 ? TH : 1);
            r.p = (ii + jj + colour) & 1;
            r.base = (ii * LJ + jj) * ROWD + r.p * PARTD + lane_pair;
            r.i = i0 - TH + ii;
            r.j = j0 - TH + jj + 2 * g;
        } else {
            const int ii = 2 * (G - 2 * RI) + 1, jj = TJ + 1;
            r.p = left_p;
            r.base = (ii * LJ + jj + left_row) * ROWD + left_p * PARTD + 1 + t;
            r.i = i0 - TH + ii + (g >> 1);
            r.j = j0 - TH + jj + (g & 1);
        }
        return r;
    };
    // 1 / t1 of the lane's two cells by parity on an inner tile's rows (t1 = 4 + its z couplings)
    float lane_rcp[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
    if (!EDGE) {
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int c = 0; c < 2; c++) lane_rcp[p][c] = __builtin_amdgcn_rcpf(4.0f + t1_z(k0 + 4 * t + p + 2 * c));
    }
    // Per-pair factors (1 / t1, -cell_m² rhs / t1) of each iteration's pair, formed before the first barrier;
    // a pair that is not updated (outside the level) keeps write false.
    auto pair_factors = [&](int colour, int group, uint32_t &rcp2, uint32_t &nsr2, bool &write) {
        const Row r = row_of(colour, group);
        write = !EDGE || (row_inside(r.i, r.j) && k0 + 4 * t < n.n2);
        float ra = 0.0f, rb = 0.0f, ca = 0.0f, cb = 0.0f;
        if (write) {
            const uint2 r4 = *reinterpret_cast<const uint2 *>(rhs + row_at(r.i, r.j) + (uint32_t)(k0 + 4 * t));
            ra = r.p ? half_hi(r4.x) : half_lo(r4.x);
            rb = r.p ? half_hi(r4.y) : half_lo(r4.y);
            if (EDGE) {
                const float xy = t1_xy(r.i, r.j);
                ca = __builtin_amdgcn_rcpf(xy + t1_z(k0 + 4 * t + r.p));
                cb = __builtin_amdgcn_rcpf(xy + t1_z(k0 + 4 * t + r.p + 2));
            } else {
                ca = r.p ? lane_rcp[1][0] : lane_rcp[0][0];
                cb = r.p ? lane_rcp[1][1] : lane_rcp[0][1];
            }
        }
        rcp2 = pack_half2(ca, cb);
        nsr2 = pack_half2((ra * -cm2) * ca, (rb * -cm2) * cb);
    };
    uint32_t red_rcp[RED_ITERS], red_nsr[RED_ITERS], black_rcp[BLACK_ITERS], black_nsr[BLACK_ITERS];
    bool red_write[RED_ITERS], black_write[BLACK_ITERS];
#pragma unroll
    for (int u = 0; u < RED_ITERS; u++) {
        red_write[u] = false;
        if (wave + u * WAVES < RED_GROUPS) pair_factors(0, wave + u * WAVES, red_rcp[u], red_nsr[u], red_write[u]);
    }
#pragma unroll
    for (int u = 0; u < BLACK_ITERS; u++) {
        black_write[u] = false;
        if (wave + u * WAVES < BLACK_GROUPS)
            pair_factors(1, wave + u * WAVES, black_rcp[u], black_nsr[u], black_write[u]);
    }
    // the red cell of ring row threadIdx.x in the k halo: k0 - 1 when i + j is odd, else k0 + 64
    int e_i = 0, e_j = 0, e_kk = 0;
    bool e_in = false;
    float e_rhs = 0.0f;
    if (threadIdx.x < RI * RJ) {
        e_i = i0 - 1 + (int)threadIdx.x / RJ;
        e_j = j0 - 1 + (int)threadIdx.x % RJ;
        e_kk = ((e_i + e_j) & 1) ? 1 : TK + 2;
        const int k = k0 - TH + e_kk;
        e_in = row_inside(e_i, e_j) && k >= 0 && k < n.n2;
        if (e_in) e_rhs = (float)*reinterpret_cast<const _Float16 *>(rhs + row_at(e_i, e_j) + (uint32_t)k);
    }
    if (START != FROM_ZERO) {
        // the rows' 64 cells, four a lane (one 8-byte load), split into the two parts; then the k halo, two
        // cells (one dword) per load
        uint2 v[LOAD_ITERS];
#pragma unroll
        for (int u = 0; u < LOAD_ITERS; u++) {
            v[u] = make_uint2(0u, 0u);
            const int r = 4 * (wave + u * WAVES) + g, i = i0 - TH + r / LJ, j = j0 - TH + r % LJ;
            if (wave + u * WAVES < LOAD_GROUPS && (!EDGE || (row_inside(i, j) && k0 + 4 * t < n.n2)))
                v[u] = *reinterpret_cast<const uint2 *>(q_in + row_at(i, j) + (uint32_t)(k0 + 4 * t));
        }
        uint32_t h[HALO_ITERS];
#pragma unroll
        for (int u = 0; u < HALO_ITERS; u++) {
            const int tt = min((int)threadIdx.x + u * THREADS, HALO - 1), r = tt >> 1;
            const int i = i0 - TH + r / LJ, j = j0 - TH + r % LJ, k = (tt & 1) ? k0 + TK : k0 - TH;
            h[u] = 0u;
            if ((!EDGE || row_inside(i, j)) && k >= 0 && k + 1 < n.n2)
                h[u] = *reinterpret_cast<const uint32_t *>(q_in + row_at(i, j) + (uint32_t)k);
        }
#pragma unroll
        for (int u = 0; u < LOAD_ITERS; u++) {
            if (wave + u * WAVES >= LOAD_GROUPS) continue;
            const int r = 4 * (wave + u * WAVES) + g;
            s[r * ROWD + 1 + t] = __builtin_amdgcn_perm(v[u].y, v[u].x, 0x05040100u);
            s[r * ROWD + PARTD + 1 + t] = __builtin_amdgcn_perm(v[u].y, v[u].x, 0x07060302u);
        }
#pragma unroll
        for (int u = 0; u < HALO_ITERS; u++) {
            const int tt = threadIdx.x + u * THREADS, r = tt >> 1, kk = (tt & 1) ? TK + 2 : 0;
            if (tt >= HALO) continue;
            s16[r * PK_RW + pk_slot(kk)] = (uint16_t)(h[u] & 0xffffu);
            s16[r * PK_RW + pk_slot(kk + 1)] = (uint16_t)(h[u] >> 16);
        }
        __syncthreads();
    }
    // one pass of colour c over its rows: a lane's pair of row r (parity p) is dword 1 + t of part p; the
    // other part's dwords t + p and t + p + 1 hold its k neighbours, one aligned, one across the two
    auto pass = [&](int colour, int group, uint32_t rcp2, uint32_t nsr2, bool write) {
        const Row r = row_of(colour, group);
        const int p = r.p, base = r.base, other = base + (p ? -PARTD : PARTD - 1);
        uint32_t value = 0u;
        if (colour == 0 && START == FROM_ZERO) {
            value = nsr2;  // every neighbour of a red cell is black, and black starts at 0
        } else {
            const half2_t im = as_half2(s[base - LJ * ROWD]), ip = as_half2(s[base + LJ * ROWD]);
            const half2_t jm = as_half2(s[base - ROWD]), jp = as_half2(s[base + ROWD]);
            const uint32_t x = s[other], y = s[other + 1];
            const half2_t aligned = as_half2(p ? x : y), across = as_half2(__builtin_amdgcn_alignbit(y, x, 16));
            const half2_t sum = ((im + ip) + (jm + jp)) + (aligned + across);
            value = as_bits(__builtin_elementwise_fma(sum, as_half2(rcp2), as_half2(nsr2)));
        }
        // a red row outside the level keeps +0 (from zero: written, since LDS was not filled)
        if (!EDGE || write) s[base] = value;
        else if (colour == 0 && START == FROM_ZERO) s[base] = 0u;
    };
#pragma unroll
    for (int u = 0; u < RED_ITERS; u++)
        if (wave + u * WAVES < RED_GROUPS) pass(0, wave + u * WAVES, red_rcp[u], red_nsr[u], red_write[u]);
    if (threadIdx.x < RI * RJ) {
        const int lr = (e_i - i0 + TH) * LJ + (e_j - j0 + TH), at = lr * PK_RW + pk_slot(e_kk);
        float value = 0.0f;
        if (e_in) {
            float sum = 0.0f;
            if (START != FROM_ZERO) {
                auto h16 = [&](int index) { return (float)*reinterpret_cast<const _Float16 *>(s16 + index); };
                sum = ((((h16(at - LJ * PK_RW) + h16(at + LJ * PK_RW)) + h16(at - PK_RW)) + h16(at + PK_RW)) +
                       h16(lr * PK_RW + pk_slot(e_kk - 1))) +
                      h16(lr * PK_RW + pk_slot(e_kk + 1));
            }
            const float t1 = t1_xy(e_i, e_j) + t1_z(k0 - TH + e_kk);
            value = (sum - cm2 * e_rhs) * __builtin_amdgcn_rcpf(t1);
        }
        if (e_in || START == FROM_ZERO) {
            const _Float16 hv = (_Float16)value;
            s16[at] = __builtin_bit_cast(uint16_t, hv);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < BLACK_ITERS; u++)
        if (wave + u * WAVES < BLACK_GROUPS) pass(1, wave + u * WAVES, black_rcp[u], black_nsr[u], black_write[u]);
    __syncthreads();
    // the tile's rows, four cells a lane: parts 0 and 1 interleaved back into one 8-byte store
#pragma unroll
    for (int u = 0; u < STORE_ITERS; u++) {
        if (wave + u * WAVES >= BLACK_GROUPS) continue;
        const int r = 4 * (wave + u * WAVES) + g, i = i0 + r / TJ, j = j0 + r % TJ;
        const int lr = (r / TJ + TH) * LJ + (r % TJ + TH);
        const uint32_t a = s[lr * ROWD + 1 + t], b = s[lr * ROWD + PARTD + 1 + t];
        if (EDGE && !(row_inside(i, j) && k0 + 4 * t < n.n2)) continue;
        *reinterpret_cast<uint2 *>(q_out + row_at(i, j) + (uint32_t)(k0 + 4 * t)) =
            make_uint2(__builtin_amdgcn_perm(b, a, 0x05040100u), __builtin_amdgcn_perm(b, a, 0x07060302u));
    }
}

template <int TI, int TJ, int THREADS, int START>
__global__ __launch_bounds__(THREADS) void k_rb_sweep_f16x2(LevelArgs L, const __half *__restrict__ q_in,
                                                            __half *__restrict__ q_out,
                                                            const __half *__restrict__ rhs, FastDiv tiles_k,
                                                            FastDiv tiles_j) {
    __shared__ uint32_t s[(TI + 2 * TH) * (TJ + 2 * TH) * PK_RW / 2];
    const Dims n = cell_dims(L.g);
    const int tile = xcd_contiguous(blockIdx.x, gridDim.x);
    const int tr = tiles_k(tile), tk = tile - tr * (int)tiles_k.d, ti = tiles_j(tr), tj = tr - ti * (int)tiles_j.d;
    const int k0 = tk * TK, j0 = tj * TJ, i0 = ti * TI;
    if (i0 >= TH && i0 + TI + TH <= n.n0 && j0 >= TH && j0 + TJ + TH <= n.n1 && k0 + TK <= n.n2)
        sweep_tile_f16x2<TI, TJ, THREADS, START, false>(L, s, q_in, q_out, rhs, i0, j0, k0);
    else
        sweep_tile_f16x2<TI, TJ, THREADS, START, true>(L, s, q_in, q_out, rhs, i0, j0, k0);
}

// The prolongation of a packed level in FP16 storage (the values of k_prolong_rows4<__half>, bit for bit):
// q = half(float(q) + p) with p the trilinear interpolation of the coarse correction
// (prolongation_kernel, sim/multigrid.py:63-80). A thread takes four fine cells along k, k = 4t .. 4t + 3,
// whose interpolation reads coarse k 2t - 1 .. 2t + 2: three dwords of each of the four coarse rows around
// (i, j). Each coarse k's x-then-y interpolation is formed once and shared by the fine cells that read it;
// the z step clamps at the level's two ends as the scalar form does. Needs the fine k extent a multiple of 4.
__global__ __launch_bounds__(256) void k_prolong_f16x2(const __half *__restrict__ coarse, Dims nc,
                                                       __half *__restrict__ fine, Dims nf) {
    const int t = blockIdx.x * BX + threadIdx.x;
    const int j = uniform((int)(blockIdx.y * BY + threadIdx.y)), i = blockIdx.z;
    if (j >= nf.n1 || 4 * t >= nf.n2) return;
    const float x = clamp_((float)i * 0.5f - 0.25f, 0.0f, (float)(nc.n0 - 1));
    const float y = clamp_((float)j * 0.5f - 0.25f, 0.0f, (float)(nc.n1 - 1));
    const int ic0 = (int)x, jc0 = (int)y, ic1 = min(ic0 + 1, nc.n0 - 1), jc1 = min(jc0 + 1, nc.n1 - 1);
    const float tx = x - (float)ic0, ty = y - (float)jc0;
    // coarse dwords 2t - 2 .. 2t + 3 of a row, clamped into the row (a clamped value is never selected)
    const int last_dword = nc.n2 / 2 - 1;
    const int d0 = max(t - 1, 0), d2 = min(t + 1, last_dword);
    auto row = [&](int ic, int jc, float (&v)[6]) {
        const uint32_t *r = reinterpret_cast<const uint32_t *>(coarse + (uint32_t)((ic * nc.n1 + jc) * nc.n2));
        const uint32_t a = r[d0], b = r[t], c = r[d2];
        v[0] = half_lo(a), v[1] = half_hi(a), v[2] = half_lo(b), v[3] = half_hi(b), v[4] = half_lo(c);
        v[5] = half_hi(c);
    };
    float A[6], B[6], C[6], D[6];
    row(ic0, jc0, A);
    row(ic1, jc0, B);
    row(ic0, jc1, C);
    row(ic1, jc1, D);
    // Y[m]: coarse k 2t - 2 + m interpolated in x, then y (prolong_at_row's c00 / c10 then their y step)
    float Y[5];
#pragma unroll
    for (int m = 1; m < 5; m++) Y[m] = lerp_(lerp_(A[m], B[m], tx), lerp_(C[m], D[m], tx), ty);
    // fine k = 4t + s: coarse (2t - 1, 2t) at 0.75, (2t, 2t + 1) at 0.25 and 0.75, (2t + 1, 2t + 2) at 0.25;
    // at k = 0 the clamped z = 0 takes (0, 1) at 0, at the last coarse k (2t + 1 = n2c - 1) z clamps to it
    const bool first = t == 0, last = 2 * t + 1 == nc.n2 - 1;
    const float p0 = first ? lerp_(Y[2], Y[3], 0.0f) : lerp_(Y[1], Y[2], 0.75f);
    const float p1 = lerp_(Y[2], Y[3], 0.25f), p2 = lerp_(Y[2], Y[3], 0.75f);
    const float p3 = last ? lerp_(Y[3], Y[3], 0.0f) : lerp_(Y[3], Y[4], 0.25f);
    uint2 *cell = reinterpret_cast<uint2 *>(fine + (uint32_t)((i * nf.n1 + j) * nf.n2 + 4 * t));
    const uint2 q = *cell;
    *cell = make_uint2(pack_half2(half_lo(q.x) + p0, half_hi(q.x) + p1),
                       pack_half2(half_lo(q.y) + p2, half_hi(q.y) + p3));
}

// residual_kernel + restrict_kernel (sim/multigrid.py:36-60) on a packed level in FP16 storage, one thread per
// coarse cell: the values of k_residual_restrict<true, __half> on the simulation level and of
// k_residual_restrict_children<__half> on coarser flat levels, bit for bit. The children's rows are read as
// dwords: the four centre rows at fine k 2k - 2 .. 2k + 3, the eight side rows and the four rhs rows at 2k,
// 2k + 1. As in the scalar simulation-level kernel every read is unconditional (rows clamped into the level,
// neighbours outside it masked to +0.0, an overhang child adding +0.0 uncounted), so no wavefront runs two
// paths. A child's couplings on a flat level are 1 to a neighbour in the level, the ghost value across an open
// face and 0 across a closed one, summed in the reference's order; its volume is 1.
template <bool UNIT>
__global__ __launch_bounds__(256) void k_residual_restrict_f16x2(LevelArgs L, const __half *__restrict__ q,
                                                                 const __half *__restrict__ rhs,
                                                                 __half *__restrict__ coarse_rhs, Dims nc) {
    const Dims nf = cell_dims(L.g);
    const int k = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y, i = blockIdx.z;
    if (!(k < nc.n2 && j < nc.n1 && i < nc.n0)) return;
    const float scale = uniform(L.g.cell_m) * uniform(L.g.cell_m);  // (1 cell_m) cell_m
    const float ghost = UNIT ? 1.0f : uniform(L.ghost);
    bool open_lo[3], open_hi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        open_lo[a] = uniform(L.g.open_lo[a]) == 1;
        open_hi[a] = uniform(L.g.open_hi[a]) == 1;
    }
    // fine rows x = 2i - 1 + m and y = 2j - 1 + m (m = 0 .. 3), clamped; fine k 2k - 1 + m, the middle two
    // always in the level (its k extent is even)
    int ri[4], rj[4];
    bool in_i[4], in_j[4], in_k[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int x = 2 * i - 1 + m, y = 2 * j - 1 + m;
        in_i[m] = x >= 0 && x < nf.n0;
        in_j[m] = y >= 0 && y < nf.n1;
        ri[m] = max(0, min(x, nf.n0 - 1));
        rj[m] = max(0, min(y, nf.n1 - 1));
    }
    in_k[0] = k >= 1;
    in_k[1] = in_k[2] = true;
    in_k[3] = 2 * k + 2 < nf.n2;
    const int dk_lo = max(k - 1, 0), dk_hi = min(k + 1, nf.n2 / 2 - 1);
    auto row = [&](const __half *f, int mi, int mj) {
        return reinterpret_cast<const uint32_t *>(f + (uint32_t)((ri[mi] * nf.n1 + rj[mj]) * nf.n2));
    };
    // q at child-block coordinates (a, b, c) in [-1, 2]^3 on the cross around the block: the centre rows' four
    // k values and the side rows' two
    float z[2][2][4], xs[2][2][2], ys[2][2][2], rc[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const uint32_t *r = row(q, a + 1, b + 1);
            const uint32_t lo = r[dk_lo], mid = r[k], hi = r[dk_hi];
            z[a][b][0] = half_hi(lo), z[a][b][1] = half_lo(mid), z[a][b][2] = half_hi(mid), z[a][b][3] = half_lo(hi);
            const uint32_t h = row(rhs, a + 1, b + 1)[k];
            rc[a][b][0] = half_lo(h), rc[a][b][1] = half_hi(h);
            const uint32_t x = row(q, a == 0 ? 0 : 3, b + 1)[k];  // the x side beyond child a of row b
            xs[a][b][0] = half_lo(x), xs[a][b][1] = half_hi(x);
            const uint32_t y = row(q, a + 1, b == 0 ? 0 : 3)[k];  // the y side beyond child b of row a
            ys[a][b][0] = half_lo(y), ys[a][b][1] = half_hi(y);
        }
    auto Q = [&](int a, int b, int c) {
        if (a < 0 || a > 1) return xs[a < 0 ? 0 : 1][b][c];
        if (b < 0 || b > 1) return ys[a][b < 0 ? 0 : 1][c];
        return z[a][b][c + 1];
    };
    auto couple = [&](bool in, bool open) { return in ? 1.0f : open ? ghost : 0.0f; };
    float total = 0.0f;
    int count = 0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float qn[6] = {Q(a - 1, b, c), Q(a + 1, b, c), Q(a, b - 1, c),
                                     Q(a, b + 1, c), Q(a, b, c - 1), Q(a, b, c + 1)};
                const bool in[6] = {in_i[a], in_i[a + 2], in_j[b], in_j[b + 2], in_k[c], in_k[c + 2]};
                float t0 = 0.0f, t1 = 0.0f;
#pragma unroll
                for (int m = 0; m < 6; m++) {
                    t0 = t0 + (in[m] ? qn[m] : 0.0f);
                    t1 = t1 + couple(in[m], (m & 1) ? open_hi[m >> 1] : open_lo[m >> 1]);
                }
                const float r = rc[a][b][c] - (t0 - t1 * Q(a, b, c)) / scale;
                const bool exists = in_i[a + 1] && in_j[b + 1];
                total = total + (exists ? r : 0.0f);
                count += exists ? 1 : 0;
            }
    coarse_rhs[nc.at(i, j, k)] = from_f32<__half>(total / (float)count);
}

// The packed shapes: TI, TJ (k is 64 wide), threads per block.
constexpr TileShape F16X2_SHAPES[] = {{8, 8, 256}, {8, 8, 512}, {4, 8, 256}};
constexpr int N_F16X2_SHAPES = 3;
int f16x2_shape = 1;  // the shape packed sweeps launch with (plm_set_f16x2_shape; 8 x 8 on 512 measured fastest)

template <int START>
int launch_f16x2(const LevelArgs &L, const __half *src, __half *dst, const __half *rhs, hipStream_t s) {
    const Dims n = cell_dims(L.g);
    const TileShape &ts = F16X2_SHAPES[f16x2_shape];
    const int tiles_k = (n.n2 + TK - 1) / TK, tiles_j = (n.n1 + ts.tj - 1) / ts.tj, tiles_i = (n.n0 + ts.ti - 1) / ts.ti;
    const dim3 grid((unsigned)(tiles_k * tiles_j * tiles_i));
    const FastDiv div_k = fast_div(tiles_k), div_j = fast_div(tiles_j);
    TimingScope launch_timing("rb_sweep_f16x2", s, true);
    const hipEvent_t e0 = launch_timing.start(), e1 = launch_timing.stop();
#define PLM_F16X2(TI, TJ, THREADS)                                                                                   \
    do {                                                                                                             \
        if (e0 != nullptr)                                                                                           \
            hipExtLaunchKernelGGL(k_rb_sweep_f16x2<TI, TJ, THREADS, START>, grid, dim3(THREADS), 0, s, e0, e1, 0u,  \
                                  L, src, dst, rhs, div_k, div_j);                                                   \
        else                                                                                                         \
            k_rb_sweep_f16x2<TI, TJ, THREADS, START><<<grid, THREADS, 0, s>>>(L, src, dst, rhs, div_k, div_j);     \
    } while (0)
    switch (f16x2_shape) {
    case 1: PLM_F16X2(8, 8, 512); break;
    case 2: PLM_F16X2(4, 8, 256); break;
    default: PLM_F16X2(8, 8, 256); break;
    }
#undef PLM_F16X2
    return check_launch("packed FP16 sweep");
}

template <typename T>
struct LpHierarchy {
    int n_levels;
    LevelArgs level[16];
    T *const *q;
    T *const *rhs;
    T *const *scratch;
    int sweeps;
    bool packed;  // the simulation level's sweeps in packed FP16 arithmetic (T = __half)
};

// q += prolongation(coarse) on one level in storage T (k_prolong_f16x2 on a packed level)
template <typename T>
int lp_prolong(const LevelArgs &L, T *q, const T *coarse, Dims nc, hipStream_t s, bool packed) {
    const Dims n = cell_dims(L.g);
    if constexpr (std::is_same<T, __half>::value) {
        if (packed && packed_f16_level(L)) {
            const dim3 grid((unsigned)((n.n2 / 4 + BX - 1) / BX), (unsigned)((n.n1 + BY - 1) / BY), (unsigned)n.n0);
            k_prolong_f16x2<<<grid, block3(), 0, s>>>(coarse, nc, q, n);
            return check_launch("packed prolongation");
        }
    }
    k_prolong_rows4<T><<<prolong_rows4_grid(n), block3(), 0, s>>>(coarse, nc, q, n, nullptr);
    return check_launch("prolong rows");
}

// coarse_rhs = restrict(rhs - A q) of one level in storage T (k_residual_restrict_f16x2 on a packed level)
template <typename T>
int lp_residual_restrict(const LevelArgs &L, const T *q, const T *rhs, T *coarse_rhs, Dims nc, hipStream_t s,
                         bool packed) {
    if constexpr (std::is_same<T, __half>::value) {
        // large levels (the per-child kernel keeps small levels' latency lower)
        if (packed && packed_f16_level(L) && cell_dims(L.g).count() >= PROLONG_ROWS_CELLS) {
            if (L.unit) k_residual_restrict_f16x2<true><<<grid3(nc), block3(), 0, s>>>(L, q, rhs, coarse_rhs, nc);
            else k_residual_restrict_f16x2<false><<<grid3(nc), block3(), 0, s>>>(L, q, rhs, coarse_rhs, nc);
            return check_launch("packed residual restrict");
        }
    }
    if (L.unit)
        k_residual_restrict<true, T><<<grid3(nc), block3(), 0, s>>>(L, q, rhs, coarse_rhs, nullptr, nc, nullptr);
    else
        k_residual_restrict_children<T><<<children_grid(nc), 256, 0, s>>>(L, q, rhs, coarse_rhs, nc, nullptr);
    return check_launch("residual restrict");
}

// `sweeps` (>= 1) two-colour tile sweeps of one level in storage T, ping-ponging q and scratch
template <typename T>
int lp_sweeps(const LevelArgs &L, T *q, const T *rhs, T *scratch, int sweeps, int start, const T *coarse, Dims nc,
              hipStream_t s, bool packed = false) {
    // packed FP16 arithmetic where the level allows it (the unit-coupling simulation level)
    packed = packed && std::is_same<T, __half>::value && packed_f16_level(L);
    // large levels, and every packed level: the prolongation as its own row kernel (q + p rounded into T), as
    // in launch_sweeps
    if (start == FROM_PROLONG && (packed || cell_dims(L.g).count() >= PROLONG_ROWS_CELLS)) {
        PLM_TRY(lp_prolong<T>(L, q, coarse, nc, s, packed));
        start = FROM_Q;
    }
    for (int sw = 0; sw < sweeps; sw++) {
        const bool to_scratch = sw % 2 == 0;
        const T *src = to_scratch ? q : scratch;
        T *dst = to_scratch ? scratch : q;
        const int first = sw == 0 ? start : FROM_Q;
        if constexpr (std::is_same<T, __half>::value) {
            if (packed) {
                if (first == FROM_ZERO) PLM_TRY(launch_f16x2<FROM_ZERO>(L, src, dst, rhs, s));
                else PLM_TRY(launch_f16x2<FROM_Q>(L, src, dst, rhs, s));
                continue;
            }
        }
        if (L.unit)
            PLM_TRY((launch_tile<true, T>(L, first, src, dst, rhs, coarse, nc, 1.0f, nullptr, s)));
        else
            PLM_TRY((launch_tile<false, T>(L, first, src, dst, rhs, coarse, nc, 1.0f, nullptr, s)));
    }
    if (sweeps % 2 == 1) {
        const int64_t count = cell_dims(L.g).count();
        k_copy_t<T><<<grid1(count), 256, 0, s>>>(count, scratch, q);
        PLM_TRY(check_launch("copy back"));
    }
    return 0;
}

// _v_cycle (sim/multigrid.py:130-147) with every level stored in T and the coarsest solved in FP64
template <typename T>
int v_cycle_lp(const LpHierarchy<T> &h, int index, bool cleared, hipStream_t s) {
    const LevelArgs &L = h.level[index];
    const Dims n = cell_dims(L.g);
    if (index == h.n_levels - 1) {
        const int longest = max(n.n0, max(n.n1, n.n2));
        const size_t bytes = 2 * sizeof(double) * (size_t)n.count();
        const int sweeps = max(20, 4 * longest), zero = cleared ? 1 : 0;
        if (n.count() <= COARSE_F64_THREADS)
            k_coarse_f64<T, 1><<<1, COARSE_F64_THREADS, bytes, s>>>(L.g, h.q[index], h.rhs[index], sweeps, zero);
        else if (n.count() <= 2 * COARSE_F64_THREADS)
            k_coarse_f64<T, 2><<<1, COARSE_F64_THREADS, bytes, s>>>(L.g, h.q[index], h.rhs[index], sweeps, zero);
        else
            k_coarse_f64<T, 4><<<1, COARSE_F64_THREADS, bytes, s>>>(L.g, h.q[index], h.rhs[index], sweeps, zero);
        return check_launch("coarsest FP64 solve");
    }
    const Dims none{1, 1, 1};
    PLM_TRY(lp_sweeps<T>(L, h.q[index], h.rhs[index], h.scratch[index], h.sweeps, cleared ? FROM_ZERO : FROM_Q,
                         nullptr, none, s, h.packed));
    const Dims nc = cell_dims(h.level[index + 1].g);
    PLM_TRY(lp_residual_restrict<T>(L, h.q[index], h.rhs[index], h.rhs[index + 1], nc, s, h.packed));
    PLM_TRY(v_cycle_lp(h, index + 1, true, s));
    return lp_sweeps<T>(L, h.q[index], h.rhs[index], h.scratch[index], h.sweeps, FROM_PROLONG, h.q[index + 1], nc, s,
                        h.packed);
}

template <typename T>
int vcycle_lp(int32_t n_levels, const plm_grid_info *infos, void *const *q, void *const *rhs, void *const *scratch,
              int32_t cycles, int32_t sweeps, int32_t zero_start, hipStream_t s, bool packed = false) {
    LpHierarchy<T> h{};
    h.n_levels = n_levels;
    h.packed = packed;
    for (int l = 0; l < n_levels; l++) h.level[l] = level_args(infos[l]);
    h.q = reinterpret_cast<T *const *>(q);
    h.rhs = reinterpret_cast<T *const *>(rhs);
    h.scratch = reinterpret_cast<T *const *>(scratch);
    h.sweeps = sweeps;
    const Dims nb = cell_dims(infos[n_levels - 1]);
    const size_t coarse_bytes = 2 * sizeof(double) * (size_t)nb.count();
    if (coarse_bytes > 65536) {
        const void *kernels[3] = {reinterpret_cast<const void *>(&k_coarse_f64<T, 1>),
                                  reinterpret_cast<const void *>(&k_coarse_f64<T, 2>),
                                  reinterpret_cast<const void *>(&k_coarse_f64<T, 4>)};
        for (const void *kernel : kernels)
            if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)coarse_bytes) != hipSuccess)
                return fail(PLM_EINVAL, "plm_vcycle_lp: %zu bytes of LDS refused", coarse_bytes);
    }
    TimingScope t("multigrid_lp", s);
    for (int c = 0; c < cycles; c++) PLM_TRY(v_cycle_lp<T>(h, 0, c == 0 && zero_start, s));
    return 0;
}

// The level table of one solve: LevelArgs per level, the arrays, and the LDS tail (the first level from which
// every coarser level fits in LDS together, with its staged table and size). Host only; allocates nothing.
// A flat level below the simulation grid with more cells than this smooths faster in LDS tiles (k_level_smooth,
// two launches a cycle) than inside the tail's one workgroup: at 128x192x128, 16x24x16 out of the tail takes the
// tail from 38.5 to 21.4 us a cycle for 12.8 us of its own.
constexpr int TAIL_SMOOTH_CELLS = 2048;

// `lds_smoothing`: the single-grid V-cycle, which smooths flat levels with two sweeps a side in LDS tiles
int build_hierarchy(Hierarchy &h, int32_t n_levels, const plm_grid_info *infos, float *const *q, float *const *rhs,
                    float *const *residual, int32_t sweeps, const double *gate, bool lds_smoothing) {
    h = Hierarchy{};
    h.n_levels = n_levels;
    for (int l = 0; l < n_levels; l++) h.level[l] = level_args(infos[l]);
    h.q = q;
    h.rhs = rhs;
    h.residual = residual;
    h.sweeps = sweeps;
    h.gate = gate;
    // the tail starts at the first level from which every coarser level fits in LDS together
    // (and whose colour slots fit the tail's per-thread plans; levels are added coarsest first, so the one
    // just added is the largest)
    int total = 0;
    bool ones = true;
    h.tail_start = n_levels;
    for (int l = n_levels - 1; l >= 0; l--) {
        const plm_grid_info &g = infos[l];
        const int cells = g.dims[0] * g.dims[1] * g.dims[2];
        const bool with = ones && g.fill_hi[0] == 1.0f && g.fill_hi[1] == 1.0f && g.fill_hi[2] == 1.0f;
        const int plan = with ? BLOCK_PLAN_ONES : BLOCK_PLAN;
        if (total + cells > PLM_LDS_CELLS || n_levels - l > MAX_TAIL || colour_slots(g) > plan * TAIL_THREADS) break;
        if (lds_smoothing && l + 1 < n_levels && cells > TAIL_SMOOTH_CELLS && sweeps == 2 && !h.level[l].unit &&
            h.level[l].flat == 1)
            break;
        total += cells;
        ones = with;
        h.tail_start = l;
    }
    // The tail's LDS: each level's q and rhs zero-bordered ((n0 + 2)(n1 + 2)(n2 + 2) floats each) and the padded
    // residual of the largest level that restricts; a tail that would not fit drops its largest level.
    auto padded = [&](int l) {
        return (infos[l].dims[0] + 2) * (infos[l].dims[1] + 2) * (infos[l].dims[2] + 2);
    };
    auto tail_floats = [&](int from) {
        int floats = 0, restricted = 0;
        for (int l = from; l < n_levels; l++) {
            floats += 2 * padded(l);
            if (l + 1 < n_levels) restricted = std::max(restricted, padded(l));
        }
        return floats + restricted;
    };
    while (h.tail_start < n_levels && sizeof(float) * (size_t)tail_floats(h.tail_start) > TAIL_LDS_BYTES)
        h.tail_start++;
    ones = true;
    for (int l = h.tail_start; l < n_levels; l++)
        ones = ones && infos[l].fill_hi[0] == 1.0f && infos[l].fill_hi[1] == 1.0f && infos[l].fill_hi[2] == 1.0f;
    h.tail_ones = ones;
    if (h.tail_start < n_levels) {
        h.tail.count = n_levels - h.tail_start;
        int offset = 0, restricted = 0;
        for (int m = 0; m < h.tail.count; m++) {
            h.tail.level[m] = h.level[h.tail_start + m];
            const plm_grid_info &g = infos[h.tail_start + m];
            h.tail.qoff[m] = offset;
            h.tail.roff[m] = offset + padded(h.tail_start + m);
            offset += 2 * padded(h.tail_start + m);
            h.tail.div_n1[m] = fast_div(g.dims[1]);
            h.tail.div_n2[m] = fast_div(g.dims[2]);
            h.tail.div_half[m] = fast_div((g.dims[2] + 1) / 2);
            if (m + 1 < h.tail.count) restricted = std::max(restricted, padded(h.tail_start + m));
        }
        h.tail.q_floats = offset;
        h.tail.res_offset = offset;
        h.tail_bytes = sizeof(float) * (size_t)(offset + restricted);
        h.tail.wave_from = h.tail.count;
        for (int m = h.tail.count - 1; m >= 0; m--) {
            const plm_grid_info &g = infos[h.tail_start + m];
            if (g.dims[0] * g.dims[1] * g.dims[2] > WAVE_CELLS || colour_slots(g) > WAVE) break;
            h.tail.wave_from = m;
        }
        PLM_TRY(h.tail_ones ? allow_tail_lds<true>(h.tail_bytes) : allow_tail_lds<false>(h.tail_bytes));
    }

    return 0;
}


// ------------------------------------------------------------------ batched multigrid (ensemble members)

// The V-cycles of many independent solves of uneven size, advanced together: each launch covers one stage of
// the V-cycle (a pass of sweeps at one depth, a restriction, the tails, a prolongation, the norms) for every
// member that has it, as a persistent grid of blocks walking a list of (member, tile) work units. A member's
// own operations and their order are exactly those of plm_vcycle (the same kernel bodies, the prolongation
// as the row kernel of the large levels, then plain sweeps), so each member's result is bit for bit that of
// its own solve. Every unit first reads its member's gate: a member whose residual met its tolerance, or that
// ran its cap of cycles, is skipped while the others go on.

struct MemberGate {
    const double *gate;     // norm_scratch + 1: 1 while the residual is above tolerance, 0 once it is not
    const int32_t *cycles;  // cycles_run[0]: the cycles the current solve has run
    int32_t cap;            // the member's most cycles
    int32_t pad;
};

__device__ __forceinline__ bool member_off(const MemberGate &m) { return m.gate[0] == 0.0 || m.cycles[0] >= m.cap; }

struct SweepJob {
    LevelArgs L;
    const float *q_in;
    float *q_out;
    const float *rhs;
    FastDiv tiles_k, tiles_j;
    int member, first;  // its gate, and its first work unit in the launch
};

// A stage whose work units are blocks of one of the grid-wide kernels: a restriction, a prolongation, norms
struct UnitJob {
    LevelArgs L;         // the level worked on (the fine level of a restriction or a prolongation)
    const float *a;      // q (restriction, norms) or the coarse q (prolongation)
    const float *b;      // rhs (restriction, norms)
    float *out;          // the coarse rhs (restriction) or the fine q (prolongation)
    double *scratch;     // the member's norm scratch (norms)
    int32_t *cycles_run; // the member's cycle counters (norms)
    int64_t blocks;      // the member's norm partials
    Dims nc;             // the coarse level
    int gx, gy, gz;      // the stage's blocks for this member
    int member, first;
};

struct TailJob {
    TailLevels tail;
    float *q_top;
    const float *rhs_top;
    int member, first;
};

// The job holding work unit u: `firsts` holds each job's first unit in increasing order (at most 64 jobs), one
// per lane in a single load, and the jobs that start at or before u are counted by a ballot.
__device__ __forceinline__ int job_of(const int *firsts, int n_jobs, int u) {
    const int lane = (int)__lane_id();
    const int first = lane < n_jobs ? firsts[lane] : 0x7fffffff;
    return __popcll(__builtin_amdgcn_ballot_w64(first <= u)) - 1;
}

// This block's units in a persistent launch: u = first, first + step, ... while u < end. The blocks of XCD x
// (block b runs on XCD b mod 8) own the x-th eighth of the units and take them in turn, so the blocks running
// together on one XCD work on neighbouring tiles (shared halos stay in that XCD's L2), and every block gets
// nearly the same number of units.
__device__ __forceinline__ void unit_range(int units, int &first, int &end, int &step) {
    const int blocks = (int)gridDim.x, x = (int)blockIdx.x % XCDS;
    step = blocks / XCDS + (x < blocks % XCDS ? 1 : 0);
    first = (int)((int64_t)units * x / XCDS) + (int)blockIdx.x / XCDS;
    end = (int)((int64_t)units * (x + 1) / XCDS);
}

// The unit before the first one of this block's stride at or past the next job (the loop's `u += step` then
// lands there): a member that is off is skipped with one gate read per block
__device__ __forceinline__ int skip_job(int u, int j, const int *firsts, int n_jobs, int end, int step) {
    const int next = j + 1 < n_jobs ? firsts[j + 1] : end;
    return u + ((next - u + step - 1) / step - 1) * step;
}

__device__ __forceinline__ dim3 unit_block(const UnitJob &J, int unit) {
    const int v = unit - J.first;
    return dim3((unsigned)(v % J.gx), (unsigned)((v / J.gx) % J.gy), (unsigned)(v / (J.gx * J.gy)));
}

template <int TI, int TJ, int THREADS, bool UNIT, int START>
__global__ __launch_bounds__(THREADS) void k_batch_sweep(const SweepJob *jobs, const int *firsts, int n_jobs,
                                                         int units, const MemberGate *gates) {
    __shared__ float s[(TI + 2 * TH) * (TJ + 2 * TH) * LK];
    __shared__ float cps[UNIT ? 1 : CP_WORDS];
    int first, end, step;
    unit_range(units, first, end, step);
    for (int u = first; u < end; u += step) {
        const int j = job_of(firsts, n_jobs, u);
        const SweepJob &J = jobs[j];
        if (member_off(gates[J.member])) {
            u = skip_job(u, j, firsts, n_jobs, end, step);
            continue;
        }
        sweep_tile<TI, TJ, THREADS, UNIT, START, float>(J.L, J.q_in, J.q_out, J.rhs, nullptr, Dims{1, 1, 1}, 1.0f,
                                                        u - J.first, J.tiles_k, J.tiles_j, s, cps);
        __syncthreads();  // the next tile reuses the LDS
    }
}

// residual + restriction onto the coarse rhs: unit levels one thread per coarse cell (64 x 4 blocks), the rest
// eight threads per coarse cell (256-thread blocks), as plm_vcycle launches them
__global__ __launch_bounds__(256) void k_batch_restrict_unit(const UnitJob *jobs, const int *firsts, int n_jobs, int units,
                                                             const MemberGate *gates) {
    int first, end, step;
    unit_range(units, first, end, step);
    for (int u = first; u < end; u += step) {
        const int j = job_of(firsts, n_jobs, u);
        const UnitJob &J = jobs[j];
        if (member_off(gates[J.member])) {
            u = skip_job(u, j, firsts, n_jobs, end, step);
            continue;
        }
        residual_restrict_cell<true, float>(J.L, J.a, J.b, J.out, nullptr, J.nc, unit_block(J, u));
    }
}

__global__ __launch_bounds__(256) void k_batch_restrict_children(const UnitJob *jobs, const int *firsts, int n_jobs, int units,
                                                                 const MemberGate *gates) {
    int first, end, step;
    unit_range(units, first, end, step);
    for (int u = first; u < end; u += step) {
        const int j = job_of(firsts, n_jobs, u);
        const UnitJob &J = jobs[j];
        if (member_off(gates[J.member])) {
            u = skip_job(u, j, firsts, n_jobs, end, step);
            continue;
        }
        residual_restrict_children<float>(J.L, J.a, J.b, J.out, J.nc, (uint32_t)(u - J.first));
    }
}

__global__ __launch_bounds__(256) void k_batch_prolong(const UnitJob *jobs, const int *firsts, int n_jobs, int units,
                                                       const MemberGate *gates) {
    int first, end, step;
    unit_range(units, first, end, step);
    for (int u = first; u < end; u += step) {
        const int j = job_of(firsts, n_jobs, u);
        const UnitJob &J = jobs[j];
        if (member_off(gates[J.member])) {
            u = skip_job(u, j, firsts, n_jobs, end, step);
            continue;
        }
        prolong_row_cell<float>(J.a, J.nc, J.out, cell_dims(J.L.g), unit_block(J, u));
    }
}

// One member's tail per block
template <bool ONES>
__global__ __launch_bounds__(TAIL_THREADS) void k_batch_tail(const TailJob *jobs, const MemberGate *gates, int sweeps) {
    const TailJob &J = jobs[blockIdx.x];
    if (member_off(gates[J.member])) return;
    mg_tail<ONES>(J.tail, J.q_top, J.rhs_top, sweeps);
}

// The FP64 norm partials of every member's simulation grid (`check`: skip the members that are off)
__global__ __launch_bounds__(NORM_THREADS) void k_batch_norm_partials(const UnitJob *jobs, const int *firsts, int n_jobs, int units,
                                                                      const MemberGate *gates, int check) {
    int first, end, step;
    unit_range(units, first, end, step);
    for (int u = first; u < end; u += step) {
        const int j = job_of(firsts, n_jobs, u);
        const UnitJob &J = jobs[j];
        if (check && member_off(gates[J.member])) {
            u = skip_job(u, j, firsts, n_jobs, end, step);
            continue;
        }
        norm_partials<true>(J.L, J.a, J.b, J.scratch + 4, unit_block(J, u), dim3(J.gx, J.gy, J.gz));
        __syncthreads();  // the next unit reuses the per-wavefront sums
    }
}

// One member's norm reduction per block, in `mode` (1 the solve's first check, 2 after a cycle)
__global__ __launch_bounds__(FINISH_THREADS) void k_batch_norm_finish(const UnitJob *jobs, const MemberGate *gates,
                                                                      double tolerance, int mode, int check) {
    const UnitJob &J = jobs[blockIdx.x];
    if (check && member_off(gates[J.member])) return;
    norm_finish(J.scratch, J.blocks, tolerance, mode, J.cycles_run);
}

enum BatchKind { B_SWEEP, B_RESTRICT_UNIT, B_RESTRICT_CHILDREN, B_PROLONG, B_TAIL, B_NORM, B_FINISH };

// One launch of the batched solve
struct BatchStage {
    int kind = B_SWEEP;
    int shape = 0, unit = 1, start = FROM_Q;  // sweeps: tile shape, unit couplings, start
    int ones = 1;                             // tails: k_mg_tail<ONES>
    int mode = 2, check = 1;                  // norms
    int n_jobs = 0, units = 0;
    size_t offset = 0;                        // its jobs in the device table
    size_t firsts = 0;                        // and each job's first unit (int32)
    size_t lds = 0;                           // tails: dynamic LDS
};

// Blocks of a persistent launch, at most one per unit: for the tile sweeps as many as are resident on the chip
// at once (the kernel's occupancy times the CUs: a block that starts only when another ends would run its
// whole share late), for the kernels without LDS eight per CU (32 wavefronts of 256 threads). A converged
// member's units then cost each block one gate read, where one block per unit would dispatch them all.
int device_cus() {
    int device = 0, cus = 0;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1)
        cus = 256;
    return cus;
}

template <typename Kernel>
unsigned launch_blocks(Kernel kernel, int threads, size_t lds, int units) {
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, lds) != hipSuccess || per < 1) per = 1;
    return (unsigned)std::max(1, std::min(units, device_cus() * per));
}

inline unsigned stream_blocks(int units) { return (unsigned)std::max(1, std::min(units, device_cus() * 8)); }

template <bool UNIT, int START>
int launch_batch_sweep(int shape, const void *jobs, const int *firsts, int n_jobs, int units, const MemberGate *gates,
                       hipStream_t s) {
#define PLM_BATCH_TILE(TI, TJ, THREADS)                                                                            \
    {                                                                                                              \
        auto kernel = k_batch_sweep<TI, TJ, THREADS, UNIT, START>;                                                 \
        kernel<<<launch_blocks(kernel, THREADS, 0, units), THREADS, 0, s>>>(                                   \
            static_cast<const SweepJob *>(jobs), firsts, n_jobs, units, gates);                                    \
        break;                                                                                                     \
    }
    switch (shape) {
    case 1: PLM_BATCH_TILE(2, 8, 256)
    case 2: PLM_BATCH_TILE(8, 4, 256)
    case 3: PLM_BATCH_TILE(4, 8, 512)
    case 4: PLM_BATCH_TILE(8, 8, 512)
    case 5: PLM_BATCH_TILE(2, 4, 256)
    default: PLM_BATCH_TILE(4, 4, 256)
    }
#undef PLM_BATCH_TILE
    return check_launch("batched sweep");
}

// A batched solve: the members' level tables, their gates, the tile shape per depth, and the device table of
// every stage's jobs with the launch lists (the first check, and one V-cycle with its norms)
struct BatchPlan {
    std::vector<Hierarchy> h;
    std::vector<MemberGate> gates;
    std::vector<double *> norm_scratch;
    std::vector<int32_t *> cycles_run;
    std::vector<int> shapes;
    int members = 0, depth = 0, sweeps = 0;
    double tolerance = 0.0;
    void *table = nullptr;
    size_t gate_offset = 0;
    std::vector<BatchStage> first, cycle;
};

// Adds `jobs` to the byte table (16-byte aligned) and returns their offset
template <typename Job>
size_t append_jobs(std::vector<char> &table, const std::vector<Job> &jobs) {
    const size_t offset = (table.size() + 15) & ~size_t(15);
    table.resize(offset + jobs.size() * sizeof(Job));
    if (!jobs.empty()) std::memcpy(table.data() + offset, jobs.data(), jobs.size() * sizeof(Job));
    return offset;
}

// A stage's jobs and the list of their first units
template <typename Job>
void append_stage(std::vector<char> &table, BatchStage &st, const std::vector<Job> &jobs) {
    st.offset = append_jobs(table, jobs);
    std::vector<int> firsts;
    for (const Job &J : jobs) firsts.push_back(J.first);
    st.firsts = append_jobs(table, firsts);
}

int launch_stage(const void *table, const MemberGate *gates, int sweeps, double tolerance, const BatchStage &st,
                 hipStream_t s) {
    const char *base = static_cast<const char *>(table);
    const void *jobs = base + st.offset;
    const int *firsts = reinterpret_cast<const int *>(base + st.firsts);
    if (st.n_jobs == 0) return 0;
    switch (st.kind) {
    case B_SWEEP:
        if (st.unit)
            return st.start == FROM_ZERO ? launch_batch_sweep<true, FROM_ZERO>(st.shape, jobs, firsts, st.n_jobs, st.units, gates, s)
                                         : launch_batch_sweep<true, FROM_Q>(st.shape, jobs, firsts, st.n_jobs, st.units, gates, s);
        return st.start == FROM_ZERO ? launch_batch_sweep<false, FROM_ZERO>(st.shape, jobs, firsts, st.n_jobs, st.units, gates, s)
                                     : launch_batch_sweep<false, FROM_Q>(st.shape, jobs, firsts, st.n_jobs, st.units, gates, s);
    case B_RESTRICT_UNIT:
        k_batch_restrict_unit<<<stream_blocks(st.units), block3(), 0, s>>>(
            static_cast<const UnitJob *>(jobs), firsts, st.n_jobs, st.units, gates);
        return check_launch("batched restriction");
    case B_RESTRICT_CHILDREN:
        k_batch_restrict_children<<<stream_blocks(st.units), 256, 0, s>>>(
            static_cast<const UnitJob *>(jobs), firsts, st.n_jobs, st.units, gates);
        return check_launch("batched restriction");
    case B_PROLONG:
        k_batch_prolong<<<stream_blocks(st.units), block3(), 0, s>>>(
            static_cast<const UnitJob *>(jobs), firsts, st.n_jobs, st.units, gates);
        return check_launch("batched prolongation");
    case B_TAIL:
        if (st.ones)
            k_batch_tail<true><<<st.n_jobs, TAIL_THREADS, st.lds, s>>>(static_cast<const TailJob *>(jobs), gates, sweeps);
        else
            k_batch_tail<false><<<st.n_jobs, TAIL_THREADS, st.lds, s>>>(static_cast<const TailJob *>(jobs), gates, sweeps);
        return check_launch("batched tail");
    case B_NORM:
        k_batch_norm_partials<<<stream_blocks(st.units), block3(), 0, s>>>(
            static_cast<const UnitJob *>(jobs), firsts, st.n_jobs, st.units, gates, st.check);
        return check_launch("batched norm partials");
    case B_FINISH:
        k_batch_norm_finish<<<st.n_jobs, FINISH_THREADS, 0, s>>>(static_cast<const UnitJob *>(jobs), gates, tolerance,
                                                                 st.mode, st.check);
        return check_launch("batched norm finish");
    default:
        return fail(PLM_EINVAL, "batched solve: unknown stage");
    }
}

// The sweep stage of depth d with tile shape `shape` over every member that has that level: q -> scratch
// (to_scratch) or scratch -> q, from q or from 0
BatchStage batch_sweep_stage(const BatchPlan &b, std::vector<char> &table, int d, int shape, int start,
                             bool to_scratch) {
    const TileShape &ts = TILE_SHAPES[shape];
    std::vector<SweepJob> jobs;
    int units = 0;
    bool unit = true;
    for (int m = 0; m < b.members; m++) {
        const Hierarchy &h = b.h[m];
        if (d >= h.tail_start) continue;
        const LevelArgs &L = h.level[d];
        const Dims n = cell_dims(L.g);
        const int tk = (n.n2 + TK - 1) / TK, tj = (n.n1 + ts.tj - 1) / ts.tj, ti = (n.n0 + ts.ti - 1) / ts.ti;
        SweepJob J{};
        J.L = L;
        J.q_in = to_scratch ? h.q[d] : h.residual[d];
        J.q_out = to_scratch ? h.residual[d] : h.q[d];
        J.rhs = h.rhs[d];
        J.tiles_k = fast_div(tk);
        J.tiles_j = fast_div(tj);
        J.member = m;
        J.first = units;
        units += tk * tj * ti;
        unit = unit && L.unit;
        jobs.push_back(J);
    }
    BatchStage st;
    st.kind = B_SWEEP;
    st.shape = shape;
    st.unit = unit ? 1 : 0;  // a level with other couplings takes the general template, which serves both
    st.start = start;
    st.n_jobs = (int)jobs.size();
    st.units = units;
    append_stage(table, st, jobs);
    return st;
}

UnitJob unit_job(const LevelArgs &L, int m, int &units, int gx, int gy, int gz) {
    UnitJob J{};
    J.L = L;
    J.member = m;
    J.first = units;
    J.gx = gx;
    J.gy = gy;
    J.gz = gz;
    units += gx * gy * gz;
    return J;
}

// Builds the launch lists and their jobs for b.shapes and puts the table on the device (replacing the old one)
int build_tables(BatchPlan &b) {
    std::vector<char> table;
    b.gate_offset = append_jobs(table, b.gates);
    std::vector<BatchStage> first, cycle;
    for (int d = 0; d < b.depth; d++) {
        // pre-smoothing: q -> scratch -> q ..., the first sweep from q on the simulation grid, from 0 below it
        for (int sw = 0; sw < b.sweeps; sw++)
            cycle.push_back(batch_sweep_stage(b, table, d, b.shapes[d], sw > 0 ? FROM_Q : d == 0 ? FROM_Q : FROM_ZERO,
                                              sw % 2 == 0));
        // residual + restriction onto level d + 1, by the coupling form of each member's level
        std::vector<UnitJob> unit_jobs, child_jobs;
        int unit_units = 0, child_units = 0;
        for (int m = 0; m < b.members; m++) {
            const Hierarchy &h = b.h[m];
            if (d >= h.tail_start) continue;
            const LevelArgs &L = h.level[d];
            const Dims nc = cell_dims(h.level[d + 1].g);
            UnitJob J;
            if (L.unit) {
                const dim3 g = grid3(nc);
                J = unit_job(L, m, unit_units, (int)g.x, (int)g.y, (int)g.z);
            } else {
                J = unit_job(L, m, child_units, (int)children_grid(nc).x, 1, 1);
            }
            J.a = h.q[d];
            J.b = h.rhs[d];
            J.out = h.rhs[d + 1];
            J.nc = nc;
            (L.unit ? unit_jobs : child_jobs).push_back(J);
        }
        for (int form = 0; form < 2; form++) {
            BatchStage st;
            st.kind = form == 0 ? B_RESTRICT_UNIT : B_RESTRICT_CHILDREN;
            const std::vector<UnitJob> &jobs = form == 0 ? unit_jobs : child_jobs;
            st.n_jobs = (int)jobs.size();
            st.units = form == 0 ? unit_units : child_units;
            append_stage(table, st, jobs);
            if (st.n_jobs > 0) cycle.push_back(st);
        }
    }
    // every member's LDS tail (one block each), in two launches by coupling form
    for (int ones = 1; ones >= 0; ones--) {
        std::vector<TailJob> jobs;
        size_t lds = 0;
        for (int m = 0; m < b.members; m++) {
            const Hierarchy &h = b.h[m];
            if ((h.tail_ones ? 1 : 0) != ones) continue;
            TailJob J{};
            J.tail = h.tail;
            J.q_top = h.q[h.tail_start];
            J.rhs_top = h.rhs[h.tail_start];
            J.member = m;
            J.first = (int)jobs.size();
            jobs.push_back(J);
            lds = std::max(lds, h.tail_bytes);
        }
        BatchStage st;
        st.kind = B_TAIL;
        st.ones = ones;
        st.n_jobs = (int)jobs.size();
        st.units = st.n_jobs;
        st.lds = lds;
        append_stage(table, st, jobs);
        if (st.n_jobs > 0) cycle.push_back(st);
    }
    for (int d = b.depth - 1; d >= 0; d--) {
        // prolong-add from level d + 1 (the row kernel), then post-smoothing from q
        std::vector<UnitJob> jobs;
        int units = 0;
        for (int m = 0; m < b.members; m++) {
            const Hierarchy &h = b.h[m];
            if (d >= h.tail_start) continue;
            const LevelArgs &L = h.level[d];
            const dim3 g = grid3(cell_dims(L.g));
            UnitJob J = unit_job(L, m, units, (int)g.x, (int)g.y, (int)g.z);
            J.a = h.q[d + 1];
            J.out = h.q[d];
            J.nc = cell_dims(h.level[d + 1].g);
            jobs.push_back(J);
        }
        BatchStage st;
        st.kind = B_PROLONG;
        st.n_jobs = (int)jobs.size();
        st.units = units;
        append_stage(table, st, jobs);
        cycle.push_back(st);
        for (int sw = 0; sw < b.sweeps; sw++)
            cycle.push_back(batch_sweep_stage(b, table, d, b.shapes[d], FROM_Q, sw % 2 == 0));
    }
    // the FP64 norms of every member's simulation grid: partials over the units of each member's norm_grid, then
    // one finishing block per member
    std::vector<UnitJob> partial_jobs, finish_jobs;
    int units = 0;
    for (int m = 0; m < b.members; m++) {
        const LevelArgs &L = b.h[m].level[0];
        const dim3 g = norm_grid(L.g);
        UnitJob J = unit_job(L, m, units, (int)g.x, (int)g.y, (int)g.z);
        J.a = b.h[m].q[0];
        J.b = b.h[m].rhs[0];
        J.scratch = b.norm_scratch[m];
        J.cycles_run = b.cycles_run[m];
        J.blocks = norm_blocks(L.g);
        partial_jobs.push_back(J);
        J.first = m;
        finish_jobs.push_back(J);
    }
    BatchStage partials, finish;
    partials.kind = B_NORM;
    partials.n_jobs = b.members;
    partials.units = units;
    append_stage(table, partials, partial_jobs);
    finish.kind = B_FINISH;
    finish.n_jobs = b.members;
    finish.units = b.members;
    append_stage(table, finish, finish_jobs);
    for (int mode = 1; mode <= 2; mode++) {
        partials.check = finish.check = mode == 2 ? 1 : 0;
        finish.mode = mode;
        std::vector<BatchStage> &list = mode == 1 ? first : cycle;
        list.push_back(partials);
        list.push_back(finish);
    }
    void *device = nullptr;
    if (hipMalloc(&device, table.size()) != hipSuccess)
        return fail(PLM_EINVAL, "batched solve: %zu bytes of device memory refused", table.size());
    if (hipMemcpy(device, table.data(), table.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(device);
        return fail(PLM_EINVAL, "batched solve: table copy failed");
    }
    if (b.table != nullptr) (void)hipFree(b.table);
    b.table = device;
    b.first = std::move(first);
    b.cycle = std::move(cycle);
    return 0;
}

const MemberGate *device_gates(const BatchPlan &b) {
    return reinterpret_cast<const MemberGate *>(static_cast<const char *>(b.table) + b.gate_offset);
}

// The time of one FROM_ZERO sweep stage at depth d in shape v, every member on (gates that are always open), writing
// only the members' scratch levels: microseconds per launch over 2 * repeats launches
int time_batch_sweep(const BatchPlan &b, int d, int v, int repeats, hipStream_t s, float &us) {
    std::vector<char> table;
    const size_t one_offset = append_jobs(table, std::vector<double>{1.0});
    const size_t zero_offset = append_jobs(table, std::vector<int32_t>{0});
    std::vector<MemberGate> open((size_t)b.members);
    const size_t gates_offset = append_jobs(table, open);
    const BatchStage st = batch_sweep_stage(b, table, d, v, FROM_ZERO, true);
    void *device = nullptr;
    if (hipMalloc(&device, table.size()) != hipSuccess) return fail(PLM_EINVAL, "batched autotune: allocation refused");
    for (int m = 0; m < b.members; m++)
        open[m] = MemberGate{reinterpret_cast<const double *>(static_cast<char *>(device) + one_offset),
                             reinterpret_cast<const int32_t *>(static_cast<char *>(device) + zero_offset), 1, 0};
    std::memcpy(table.data() + gates_offset, open.data(), open.size() * sizeof(MemberGate));
    int rc = hipMemcpy(device, table.data(), table.size(), hipMemcpyHostToDevice) == hipSuccess
                 ? 0
                 : fail(PLM_EINVAL, "batched autotune: table copy failed");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(PLM_EINVAL, "batched autotune: event creation failed");
    const MemberGate *gates = reinterpret_cast<const MemberGate *>(static_cast<char *>(device) + gates_offset);
    if (rc == 0) rc = launch_stage(device, gates, b.sweeps, b.tolerance, st, s);  // warm-up
    if (rc == 0 && hipEventRecord(e0, s) != hipSuccess) rc = fail(PLM_EINVAL, "batched autotune: event record failed");
    for (int r = 0; r < 2 * repeats && rc == 0; r++) rc = launch_stage(device, gates, b.sweeps, b.tolerance, st, s);
    if (rc == 0 && hipEventRecord(e1, s) != hipSuccess) rc = fail(PLM_EINVAL, "batched autotune: event record failed");
    if (rc == 0 && hipEventSynchronize(e1) != hipSuccess) rc = fail(PLM_EINVAL, "batched autotune: synchronise failed");
    float ms = 0.0f;
    if (rc == 0) (void)hipEventElapsedTime(&ms, e0, e1);
    us = ms * 1000.0f / (float)(2 * repeats);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (hipStreamSynchronize(s) != hipSuccess && rc == 0) rc = fail(PLM_EINVAL, "batched autotune: synchronise failed");
    (void)hipFree(device);
    return rc;
}

// k_batch_tail needs the dynamic LDS of its largest tail allowed once
template <bool ONES>
int allow_batch_tail_lds(size_t bytes) {
    static size_t allowed = 0;
    if (bytes <= 65536 || bytes <= allowed) return 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&k_batch_tail<ONES>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
        return fail(PLM_EINVAL, "batched tail: %zu bytes of LDS refused", bytes);
    allowed = bytes;
    return 0;
}
}  // namespace
}  // namespace plm

using namespace plm;

// The handle of a batched solve (its plan lives in the library's internal namespace)
struct plm_mg_batch {
    void *plan = nullptr;
};

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

int plm_close_divergence(const plm_grid_info *info, float *u, float *v, float *w, float *divergence, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && divergence, "plm_close_divergence: NULL array");
    k_close_divergence<<<divergence_grid(*info), block3(), 0, as_stream(stream)>>>(*info, u, v, w, divergence);
    return check_launch("plm_close_divergence");
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
    const LevelArgs L = level_args(*info);
    const Dims n = cell_dims(*info);
    if (L.unit)
        k_rb_sweep<true><<<grid3(n), block3(), 0, as_stream(stream)>>>(L, q, rhs, color, relaxation, nullptr);
    else
        k_rb_sweep<false><<<grid3(n), block3(), 0, as_stream(stream)>>>(L, q, rhs, color, relaxation, nullptr);
    return check_launch("plm_rb_sweep");
}

int plm_rb_sweeps(const plm_grid_info *info, float *q, const float *rhs, float *scratch, int32_t sweeps,
                  float relaxation, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs, "plm_rb_sweeps: NULL array");
    PLM_REQUIRE(sweeps >= 0, "plm_rb_sweeps: sweeps must be >= 0");
    return launch_sweeps(level_args(*info), q, rhs, scratch, sweeps, relaxation, FROM_Q, nullptr, Dims{1, 1, 1},
                         nullptr, as_stream(stream));
}

int plm_subtract_gradient(const plm_grid_info *info, float *u, float *v, float *w, const float *q, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && q, "plm_subtract_gradient: NULL array");
    k_subtract_gradient<<<gradient_grid(*info), block3(), 0, as_stream(stream)>>>(*info, Tri<float>{{u, v, w}}, q);
    return check_launch("plm_subtract_gradient");
}

int plm_residual(const plm_grid_info *info, const float *q, const float *rhs, float *residual, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs && residual, "plm_residual: NULL array");
    const LevelArgs L = level_args(*info);
    const Dims n = cell_dims(*info);
    if (L.unit)
        k_residual<true><<<grid3(n), block3(), 0, as_stream(stream)>>>(L, q, rhs, residual);
    else
        k_residual<false><<<grid3(n), block3(), 0, as_stream(stream)>>>(L, q, rhs, residual);
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

int32_t plm_sweep_tile_shapes(int32_t *ti, int32_t *tj, int32_t *threads, int32_t capacity) {
    for (int v = 0; v < N_TILE_SHAPES && v < capacity; v++) {
        if (ti) ti[v] = TILE_SHAPES[v].ti;
        if (tj) tj[v] = TILE_SHAPES[v].tj;
        if (threads) threads[v] = TILE_SHAPES[v].threads;
    }
    return N_TILE_SHAPES;
}

int plm_set_sweep_tile(const int32_t dims[3], int32_t shape) {
    PLM_TRY(validate_dims(dims, "plm_set_sweep_tile"));
    PLM_REQUIRE(shape >= -1 && shape < N_TILE_SHAPES, "plm_set_sweep_tile: shape %d out of range", shape);
    TileChoices &c = tile_choices();
    std::lock_guard<std::mutex> lock(c.mutex);
    if (shape < 0)
        c.shape.erase({dims[0], dims[1], dims[2]});
    else
        c.shape[{dims[0], dims[1], dims[2]}] = shape;
    return 0;
}

int plm_autotune_sweep(const plm_grid_info *info, float *q, const float *rhs, float *scratch, int32_t repeats,
                       float *times_us, int32_t *chosen, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs && scratch && chosen, "plm_autotune_sweep: NULL argument");
    PLM_REQUIRE(repeats >= 1, "plm_autotune_sweep: repeats must be >= 1");
    hipStream_t s = as_stream(stream);
    const LevelArgs L = level_args(*info);
    return L.unit ? autotune_sweep<true>(L, q, rhs, scratch, repeats, times_us, chosen, s)
                  : autotune_sweep<false>(L, q, rhs, scratch, repeats, times_us, chosen, s);
}

int64_t plm_norm_scratch_doubles(const plm_grid_info *info) {
    if (validate_info(info) != 0) return -1;
    return 4 + 2 * norm_blocks(*info);
}

int plm_residual_norms(const plm_grid_info *info, const float *q, const float *rhs, double *scratch, double *out,
                       void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs && scratch && out, "plm_residual_norms: NULL array");
    hipStream_t s = as_stream(stream);
    PLM_TRY(launch_norms(level_args(*info), q, rhs, scratch, nullptr, 0.0, 0, nullptr, s));
    PLM_REQUIRE(hipMemcpyAsync(out, scratch + 2, 2 * sizeof(double), hipMemcpyDeviceToDevice, s) == hipSuccess,
                "plm_residual_norms: copy failed");
    return 0;
}

int plm_vcycle(int32_t n_levels, const plm_grid_info *infos, float *const *q, float *const *rhs,
               float *const *residual, int32_t cycles, int32_t sweeps, double tolerance, int32_t resume,
               double *norm_scratch, int32_t *cycles_run, void *stream) {
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

    Hierarchy h;
    PLM_TRY(build_hierarchy(h, n_levels, infos, q, rhs, residual, sweeps, tolerance_mode ? norm_scratch + 1 : nullptr,
                            true));

    TimingScope t("multigrid", s);
    if (tolerance_mode && !resume)
        PLM_TRY(launch_norms(h.level[0], q[0], rhs[0], norm_scratch, nullptr, tolerance, 1, cycles_run, s));
    for (int c = 0; c < cycles; c++) {
        PLM_TRY(v_cycle(h, 0, false, s));
        if (tolerance_mode)
            PLM_TRY(launch_norms(h.level[0], q[0], rhs[0], norm_scratch, h.gate, tolerance, 2, cycles_run, s));
    }
    return 0;
}


int plm_vcycle_lp(int32_t n_levels, const plm_grid_info *infos, int32_t storage, void *const *q, void *const *rhs,
                  void *const *scratch, int32_t cycles, int32_t sweeps, int32_t zero_start, void *stream) {
    PLM_REQUIRE(n_levels >= 2 && n_levels <= 16 && infos && q && rhs && scratch, "plm_vcycle_lp: bad level arrays");
    PLM_REQUIRE(storage >= PLM_STORAGE_F32 && storage <= PLM_STORAGE_F16_PACKED, "plm_vcycle_lp: storage %d unknown",
                storage);
    PLM_REQUIRE(cycles >= 0 && sweeps >= 1, "plm_vcycle_lp: cycles must be >= 0 and sweeps >= 1");
    for (int l = 0; l < n_levels; l++) {
        PLM_TRY(validate_info(&infos[l]));
        PLM_REQUIRE(q[l] && rhs[l] && (scratch[l] || l == n_levels - 1), "plm_vcycle_lp: NULL array at level %d", l);
        if (l > 0)
            for (int a = 0; a < 3; a++)
                PLM_REQUIRE(infos[l].dims[a] == (infos[l - 1].dims[a] + 1) / 2,
                            "plm_vcycle_lp: level %d does not halve level %d", l, l - 1);
    }
    const plm_grid_info &last = infos[n_levels - 1];
    PLM_REQUIRE(last.dims[0] * last.dims[1] * last.dims[2] <= COARSE_F64_CELLS,
                "plm_vcycle_lp: the coarsest level (FP64, one workgroup) holds at most %d cells", COARSE_F64_CELLS);
    hipStream_t s = as_stream(stream);
    if (storage == PLM_STORAGE_BF16)
        return vcycle_lp<__hip_bfloat16>(n_levels, infos, q, rhs, scratch, cycles, sweeps, zero_start, s);
    if (storage == PLM_STORAGE_F16 || storage == PLM_STORAGE_F16_PACKED)
        return vcycle_lp<__half>(n_levels, infos, q, rhs, scratch, cycles, sweeps, zero_start, s,
                                 storage == PLM_STORAGE_F16_PACKED);
    return vcycle_lp<float>(n_levels, infos, q, rhs, scratch, cycles, sweeps, zero_start, s);
}

int plm_rb_sweeps_lp(const plm_grid_info *info, int32_t storage, void *q, const void *rhs, void *scratch,
                     int32_t sweeps, int32_t zero_start, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs && scratch && sweeps >= 1, "plm_rb_sweeps_lp: NULL array or sweeps < 1");
    const LevelArgs L = level_args(*info);
    hipStream_t s = as_stream(stream);
    const int start = zero_start ? FROM_ZERO : FROM_Q;
    const Dims none{1, 1, 1};
    if (storage == PLM_STORAGE_BF16)
        return lp_sweeps<__hip_bfloat16>(L, (__hip_bfloat16 *)q, (const __hip_bfloat16 *)rhs, (__hip_bfloat16 *)scratch,
                                         sweeps, start, nullptr, none, s);
    if (storage == PLM_STORAGE_F16 || storage == PLM_STORAGE_F16_PACKED)
        return lp_sweeps<__half>(L, (__half *)q, (const __half *)rhs, (__half *)scratch, sweeps, start, nullptr, none, s,
                                 storage == PLM_STORAGE_F16_PACKED);
    if (storage == PLM_STORAGE_F32)
        return lp_sweeps<float>(L, (float *)q, (const float *)rhs, (float *)scratch, sweeps, start, nullptr, none, s);
    return fail(PLM_EINVAL, "plm_rb_sweeps_lp: storage %d unknown", storage);
}

int plm_residual_restrict_lp(const plm_grid_info *info, int32_t storage, const void *q, const void *rhs,
                             void *coarse_rhs, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && rhs && coarse_rhs, "plm_residual_restrict_lp: NULL array");
    const LevelArgs L = level_args(*info);
    const Dims n = cell_dims(*info), nc{(n.n0 + 1) / 2, (n.n1 + 1) / 2, (n.n2 + 1) / 2};
    hipStream_t s = as_stream(stream);
    if (storage == PLM_STORAGE_BF16)
        return lp_residual_restrict<__hip_bfloat16>(L, (const __hip_bfloat16 *)q, (const __hip_bfloat16 *)rhs,
                                                    (__hip_bfloat16 *)coarse_rhs, nc, s, false);
    if (storage == PLM_STORAGE_F16 || storage == PLM_STORAGE_F16_PACKED)
        return lp_residual_restrict<__half>(L, (const __half *)q, (const __half *)rhs, (__half *)coarse_rhs, nc, s,
                                            storage == PLM_STORAGE_F16_PACKED);
    if (storage == PLM_STORAGE_F32)
        return lp_residual_restrict<float>(L, (const float *)q, (const float *)rhs, (float *)coarse_rhs, nc, s, false);
    return fail(PLM_EINVAL, "plm_residual_restrict_lp: storage %d unknown", storage);
}

int plm_prolong_add_lp(const plm_grid_info *info, int32_t storage, const void *coarse, void *q, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(q && coarse, "plm_prolong_add_lp: NULL array");
    const LevelArgs L = level_args(*info);
    const Dims n = cell_dims(*info), nc{(n.n0 + 1) / 2, (n.n1 + 1) / 2, (n.n2 + 1) / 2};
    hipStream_t s = as_stream(stream);
    if (storage == PLM_STORAGE_BF16)
        return lp_prolong<__hip_bfloat16>(L, (__hip_bfloat16 *)q, (const __hip_bfloat16 *)coarse, nc, s, false);
    if (storage == PLM_STORAGE_F16 || storage == PLM_STORAGE_F16_PACKED)
        return lp_prolong<__half>(L, (__half *)q, (const __half *)coarse, nc, s, storage == PLM_STORAGE_F16_PACKED);
    if (storage == PLM_STORAGE_F32)
        return lp_prolong<float>(L, (float *)q, (const float *)coarse, nc, s, false);
    return fail(PLM_EINVAL, "plm_prolong_add_lp: storage %d unknown", storage);
}

int plm_set_f16x2_shape(int32_t shape) {
    PLM_REQUIRE(shape >= 0 && shape < N_F16X2_SHAPES, "plm_set_f16x2_shape: shape %d out of range", shape);
    f16x2_shape = shape;
    return 0;
}

int plm_to_storage(int64_t count, const float *src, float scale, int32_t storage, void *dst, void *stream) {
    if (storage == PLM_STORAGE_F16_PACKED) storage = PLM_STORAGE_F16;  // the same FP16 storage
    PLM_REQUIRE(count >= 0 && src && dst, "plm_to_storage: bad arguments");
    hipStream_t s = as_stream(stream);
    if (count == 0) return 0;
    if (storage == PLM_STORAGE_BF16)
        k_to_storage<__hip_bfloat16><<<grid1(count), 256, 0, s>>>(count, src, scale, (__hip_bfloat16 *)dst);
    else if (storage == PLM_STORAGE_F16)
        k_to_storage<__half><<<grid1(count), 256, 0, s>>>(count, src, scale, (__half *)dst);
    else if (storage == PLM_STORAGE_F32)
        k_to_storage<float><<<grid1(count), 256, 0, s>>>(count, src, scale, (float *)dst);
    else
        return fail(PLM_EINVAL, "plm_to_storage: storage %d unknown", storage);
    return check_launch("to storage");
}

int plm_add_storage(int64_t count, float *x, const void *e, int32_t storage, float inv_scale, void *stream) {
    if (storage == PLM_STORAGE_F16_PACKED) storage = PLM_STORAGE_F16;  // the same FP16 storage
    PLM_REQUIRE(count >= 0 && x && e, "plm_add_storage: bad arguments");
    hipStream_t s = as_stream(stream);
    if (count == 0) return 0;
    if (storage == PLM_STORAGE_BF16)
        k_add_storage<__hip_bfloat16><<<grid1(count), 256, 0, s>>>(count, x, (const __hip_bfloat16 *)e, inv_scale);
    else if (storage == PLM_STORAGE_F16)
        k_add_storage<__half><<<grid1(count), 256, 0, s>>>(count, x, (const __half *)e, inv_scale);
    else if (storage == PLM_STORAGE_F32)
        k_add_storage<float><<<grid1(count), 256, 0, s>>>(count, x, (const float *)e, inv_scale);
    else
        return fail(PLM_EINVAL, "plm_add_storage: storage %d unknown", storage);
    return check_launch("add storage");
}

int plm_refine_residual(const plm_grid_info *info, const float *x, const void *e, float inv_scale, int32_t storage,
                        const float *rhs, float *x_out, void *r_out, float r_scale, double *scratch, double *out,
                        void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(x && rhs && scratch && out, "plm_refine_residual: NULL array");
    PLM_REQUIRE(x_out || !e, "plm_refine_residual: a correction e needs x_out");
    const int64_t count = cell_dims(*info).count();
    const uintptr_t xa = reinterpret_cast<uintptr_t>(x), xb = reinterpret_cast<uintptr_t>(x_out);
    const uintptr_t bytes = (uintptr_t)count * sizeof(float);
    PLM_REQUIRE(!x_out || xa + bytes <= xb || xb + bytes <= xa, "plm_refine_residual: x and x_out overlap");
    hipStream_t s = as_stream(stream);
    const LevelArgs L = level_args(*info);
    TimingScope t("refine_residual", s);
    if (storage == PLM_STORAGE_F16_PACKED) storage = PLM_STORAGE_F16;  // the same FP16 storage
    if (storage == PLM_STORAGE_BF16)
        PLM_TRY(launch_refine<__hip_bfloat16>(L, x, e, inv_scale, rhs, x_out, r_out, r_scale, scratch, s));
    else if (storage == PLM_STORAGE_F16)
        PLM_TRY(launch_refine<__half>(L, x, e, inv_scale, rhs, x_out, r_out, r_scale, scratch, s));
    else if (storage == PLM_STORAGE_F32)
        PLM_TRY(launch_refine<float>(L, x, e, inv_scale, rhs, x_out, r_out, r_scale, scratch, s));
    else
        return fail(PLM_EINVAL, "plm_refine_residual: storage %d unknown", storage);
    PLM_REQUIRE(hipMemcpyAsync(out, scratch + 2, 2 * sizeof(double), hipMemcpyDeviceToDevice, s) == hipSuccess,
                "plm_refine_residual: copy failed");
    return 0;
}


int plm_mg_batch_create(int32_t members, const int32_t *n_levels, const plm_grid_info *const *infos,
                        float *const *const *q, float *const *const *rhs, float *const *const *residual,
                        int32_t sweeps, double tolerance, const int32_t *max_cycles, double *const *norm_scratch,
                        int32_t *const *cycles_run, plm_mg_batch **out) {
    PLM_REQUIRE(out != nullptr, "plm_mg_batch_create: NULL out");
    *out = nullptr;
    PLM_REQUIRE(members >= 1 && members <= 64 && n_levels && infos && q && rhs && residual && max_cycles &&
                    norm_scratch && cycles_run,
                "plm_mg_batch_create: bad member arrays (1 to 64 members)");
    PLM_REQUIRE(sweeps >= 2 && sweeps % 2 == 0, "plm_mg_batch_create: sweeps must be even and positive (got %d)",
                sweeps);
    PLM_REQUIRE(tolerance > 0.0, "plm_mg_batch_create: the batched solve runs in tolerance mode (tolerance > 0)");
    std::unique_ptr<BatchPlan> plan(new BatchPlan());
    BatchPlan &bp = *plan;
    bp.members = members;
    bp.sweeps = sweeps;
    bp.tolerance = tolerance;
    bp.h.resize((size_t)members);
    for (int m = 0; m < members; m++) {
        const int n = n_levels[m];
        PLM_REQUIRE(n >= 1 && n <= 16 && infos[m] && q[m] && rhs[m] && residual[m] && norm_scratch[m] && cycles_run[m],
                    "plm_mg_batch_create: bad arrays of member %d", m);
        PLM_REQUIRE(max_cycles[m] >= 0, "plm_mg_batch_create: negative cycle cap of member %d", m);
        for (int l = 0; l < n; l++) {
            PLM_TRY(validate_info(&infos[m][l]));
            PLM_REQUIRE(q[m][l] && rhs[m][l] && residual[m][l], "plm_mg_batch_create: NULL array, member %d level %d",
                        m, l);
            if (l > 0)
                for (int a = 0; a < 3; a++)
                    PLM_REQUIRE(infos[m][l].dims[a] == (infos[m][l - 1].dims[a] + 1) / 2,
                                "plm_mg_batch_create: member %d level %d does not halve level %d", m, l, l - 1);
        }
        Hierarchy &h = bp.h[m];
        PLM_TRY(build_hierarchy(h, n, infos[m], q[m], rhs[m], residual[m], sweeps, norm_scratch[m] + 1, false));
        PLM_REQUIRE(h.tail_start < n, "plm_mg_batch_create: member %d has no LDS tail", m);
        PLM_TRY(h.tail_ones ? allow_batch_tail_lds<true>(h.tail_bytes) : allow_batch_tail_lds<false>(h.tail_bytes));
        // the reference's freshly zeroed coarse q, unless the whole hierarchy is the tail (warm start)
        h.tail.zero_top = h.tail_start > 0 ? 1 : 0;
        bp.gates.push_back(MemberGate{norm_scratch[m] + 1, cycles_run[m], max_cycles[m], 0});
        bp.norm_scratch.push_back(norm_scratch[m]);
        bp.cycles_run.push_back(cycles_run[m]);
        bp.depth = std::max(bp.depth, h.tail_start);
    }
    // each depth starts with the shape pinned for its largest member level (plm_mg_batch_autotune retunes)
    for (int d = 0; d < bp.depth; d++) {
        int largest = -1, cells = -1;
        for (int m = 0; m < members; m++)
            if (d < bp.h[m].tail_start && cell_dims(bp.h[m].level[d].g).count() > cells) {
                cells = cell_dims(bp.h[m].level[d].g).count();
                largest = m;
            }
        bp.shapes.push_back(tile_shape_for(cell_dims(bp.h[largest].level[d].g)));
    }
    PLM_TRY(build_tables(bp));
    plm_mg_batch *b = new plm_mg_batch();
    b->plan = plan.release();
    *out = b;
    return 0;
}

int plm_mg_batch_autotune(plm_mg_batch *batch, int32_t repeats, float *times_us, void *stream) {
    PLM_REQUIRE(batch != nullptr && batch->plan != nullptr && repeats >= 1, "plm_mg_batch_autotune: bad arguments");
    BatchPlan &bp = *static_cast<BatchPlan *>(batch->plan);
    hipStream_t s = as_stream(stream);
    for (int d = 0; d < bp.depth; d++) {
        int best = 0;
        float best_us = 0.0f;
        for (int v = 0; v < N_TILE_SHAPES; v++) {
            float us = 0.0f;
            PLM_TRY(time_batch_sweep(bp, d, v, repeats, s, us));
            if (times_us) times_us[d * N_TILE_SHAPES + v] = us;
            if (v == 0 || us < best_us) {
                best = v;
                best_us = us;
            }
        }
        bp.shapes[d] = best;
    }
    return build_tables(bp);
}

int32_t plm_mg_batch_depth(const plm_mg_batch *batch) {
    if (batch == nullptr || batch->plan == nullptr) return -1;
    return static_cast<const BatchPlan *>(batch->plan)->depth;
}

int plm_mg_batch_solve(plm_mg_batch *batch, int32_t cycles, void *stream) {
    PLM_REQUIRE(batch != nullptr && batch->plan != nullptr && cycles >= 0, "plm_mg_batch_solve: bad arguments");
    const BatchPlan &bp = *static_cast<const BatchPlan *>(batch->plan);
    hipStream_t s = as_stream(stream);
    const MemberGate *gates = device_gates(bp);
    TimingScope t("multigrid_batch", s);
    for (const BatchStage &st : bp.first) PLM_TRY(launch_stage(bp.table, gates, bp.sweeps, bp.tolerance, st, s));
    for (int c = 0; c < cycles; c++)
        for (const BatchStage &st : bp.cycle) PLM_TRY(launch_stage(bp.table, gates, bp.sweeps, bp.tolerance, st, s));
    return 0;
}

int plm_mg_batch_destroy(plm_mg_batch *batch) {
    if (batch == nullptr) return 0;
    BatchPlan *plan = static_cast<BatchPlan *>(batch->plan);
    const hipError_t e = plan != nullptr && plan->table != nullptr ? hipFree(plan->table) : hipSuccess;
    delete plan;
    delete batch;
    return e == hipSuccess ? 0 : fail(PLM_EINVAL, "plm_mg_batch_destroy: free failed");
}

}  // extern "C"
