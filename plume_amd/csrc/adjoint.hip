// Hand-written adjoints of the differentiable substep and of the thickness / heat channels of the ray march.
//
// The reference has no backward pass (it sets enable_backward = False, sim/device.py:28); these kernels are
// the reverse-mode derivatives of this library's own forward kernels, operation by operation:
//   - buoyancy_kernel (sim/forces.py:13-31): linear in v, T and s; the smoke-weight term gives the alpha
//     derivative;
//   - the four confinement passes (sim/vorticity.py:17-84): centre average, curl by central differences,
//     normalize(grad |curl|) x curl, face average; the normalisations differentiated where they are smooth
//     (zero push below the 1e-12 guard, as the forward);
//   - clamped MacCormack (sim/advect.py:24-68): both semi-Lagrangian passes through the trilinear weights
//     (field role) and through the midpoint trace's positions (velocity role); the clamp passes the gradient
//     to the unclamped expression when lo <= value <= hi (ties included), else to the tap it selected;
//   - projection (sim/pressure.py:144-154) as the linear map u -> Z u - G A^-1 D Z u: its transpose needs
//     G^T and D^T (here) and one more solve with the same symmetric Poisson operator (the caller's multigrid);
//   - raymarch_kernel's heat and thickness (render/raymarch.py:86-140): a recording pass keeps each sample's
//     transmittance, opacity and field values per pixel, and the reverse pass walks them back to front.
// Scatters of trilinear samples use float atomics; every transpose of a fixed stencil is a gather.
#include "common.h"

namespace plm {
namespace {

__global__ __launch_bounds__(256) void k_fill_zero(int64_t count, float *__restrict__ f) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x)
        f[c] = 0.0f;
}

int zero_fill(float *f, int64_t count, hipStream_t s) {
    if (count <= 0) return 0;
    const int64_t blocks = (count + 255) / 256;
    k_fill_zero<<<(unsigned)(blocks < 16384 ? blocks : 16384), 256, 0, s>>>(count, f);
    return check_launch("zero fill");
}

// ------------------------------------------------------------------ trilinear sampling with derivatives

// sample_index's corners and weights at (gx, gy, gz); free[a]: the position is inside [0, n - 1] on axis a,
// where the clamped sample moves with it
struct TriW {
    int i0, i1, j0, j1, k0, k1;
    float tx, ty, tz;
    bool free[3];
};

__device__ __forceinline__ TriW tri_setup(Dims n, float gx, float gy, float gz) {
    TriW t;
    const float x = clamp_(gx, 0.0f, (float)(n.n0 - 1)), y = clamp_(gy, 0.0f, (float)(n.n1 - 1)),
                z = clamp_(gz, 0.0f, (float)(n.n2 - 1));
    t.free[0] = gx >= 0.0f && gx <= (float)(n.n0 - 1);
    t.free[1] = gy >= 0.0f && gy <= (float)(n.n1 - 1);
    t.free[2] = gz >= 0.0f && gz <= (float)(n.n2 - 1);
    t.i0 = (int)x;
    t.j0 = (int)y;
    t.k0 = (int)z;
    t.i1 = min(t.i0 + 1, n.n0 - 1);
    t.j1 = min(t.j0 + 1, n.n1 - 1);
    t.k1 = min(t.k0 + 1, n.n2 - 1);
    t.tx = x - (float)t.i0;
    t.ty = y - (float)t.j0;
    t.tz = z - (float)t.k0;
    return t;
}

// d(sample)/d(position) of the nested lerp (x, then y, then z)
__device__ __forceinline__ void tri_grad(const float *__restrict__ f, Dims n, const TriW &t, float &dx, float &dy,
                                         float &dz) {
    const float f000 = f[n.at(t.i0, t.j0, t.k0)], f100 = f[n.at(t.i1, t.j0, t.k0)];
    const float f010 = f[n.at(t.i0, t.j1, t.k0)], f110 = f[n.at(t.i1, t.j1, t.k0)];
    const float f001 = f[n.at(t.i0, t.j0, t.k1)], f101 = f[n.at(t.i1, t.j0, t.k1)];
    const float f011 = f[n.at(t.i0, t.j1, t.k1)], f111 = f[n.at(t.i1, t.j1, t.k1)];
    const float c00 = lerp_(f000, f100, t.tx), c10 = lerp_(f010, f110, t.tx);
    const float c01 = lerp_(f001, f101, t.tx), c11 = lerp_(f011, f111, t.tx);
    dx = t.free[0] ? lerp_(lerp_(f100 - f000, f110 - f010, t.ty), lerp_(f101 - f001, f111 - f011, t.ty), t.tz)
                   : 0.0f;
    dy = t.free[1] ? lerp_(c10 - c00, c11 - c01, t.tz) : 0.0f;
    dz = t.free[2] ? lerp_(c01, c11, t.ty) - lerp_(c00, c10, t.ty) : 0.0f;
}

// adjoint of sample_index: g times each corner's weight added into fbar
__device__ __forceinline__ void tri_scatter(float *__restrict__ fbar, Dims n, const TriW &t, float g) {
    if (g == 0.0f) return;
    const float wx[2] = {1.0f - t.tx, t.tx}, wy[2] = {1.0f - t.ty, t.ty}, wz[2] = {1.0f - t.tz, t.tz};
    const int ii[2] = {t.i0, t.i1}, jj[2] = {t.j0, t.j1}, kk[2] = {t.k0, t.k1};
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float w = wx[a] * wy[b] * wz[c];
                if (w != 0.0f) atomicAdd(fbar + n.at(ii[a], jj[b], kk[c]), g * w);
            }
}

struct Vel {
    const float *u, *v, *w;
};
struct VelBar {
    float *u, *v, *w;
};

__device__ __forceinline__ Dims comp_dims(const plm_grid_info &g, int c) { return face_dims(g, c); }
__device__ __forceinline__ void comp_offset(int c, float &ox, float &oy, float &oz) {
    ox = c == 0 ? 0.0f : 0.5f;
    oy = c == 1 ? 0.0f : 0.5f;
    oz = c == 2 ? 0.0f : 0.5f;
}
__device__ __forceinline__ const float *comp(const Vel &U, int c) { return c == 0 ? U.u : (c == 1 ? U.v : U.w); }
__device__ __forceinline__ float *comp(const VelBar &U, int c) { return c == 0 ? U.u : (c == 1 ? U.v : U.w); }

// adjoint of trace (sim/advect.py:16-21) from p with step dto: sbar is the adjoint of the traced point; the
// velocity samples at the midpoint and at p receive their parts
__device__ void trace_adjoint(const plm_grid_info &g, const Vel &U, const VelBar &Ubar, float px, float py,
                              float pz, float dto, const float sbar[3]) {
    float vp[3];
    velocity_at(g, U.u, U.v, U.w, px, py, pz, vp[0], vp[1], vp[2]);
    const float half = 0.5f * dto;
    const float m[3] = {px - half * vp[0], py - half * vp[1], pz - half * vp[2]};
    float mbar[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float vmbar = -dto * sbar[c];
        float ox, oy, oz;
        comp_offset(c, ox, oy, oz);
        const Dims n = comp_dims(g, c);
        const TriW t = tri_setup(n, m[0] - ox, m[1] - oy, m[2] - oz);
        tri_scatter(comp(Ubar, c), n, t, vmbar);
        float dx, dy, dz;
        tri_grad(comp(U, c), n, t, dx, dy, dz);
        mbar[0] += vmbar * dx;
        mbar[1] += vmbar * dy;
        mbar[2] += vmbar * dz;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float ox, oy, oz;
        comp_offset(c, ox, oy, oz);
        const Dims n = comp_dims(g, c);
        const TriW t = tri_setup(n, px - ox, py - oy, pz - oz);
        tri_scatter(comp(Ubar, c), n, t, -half * mbar[c]);
    }
}

struct FieldDesc {
    Dims n;
    float ox, oy, oz;
    float outside;
    int uses_outside;
};

__device__ __forceinline__ FieldDesc face_field(const plm_grid_info &g, int axis) {
    FieldDesc f;
    f.n = face_dims(g, axis);
    comp_offset(axis, f.ox, f.oy, f.oz);
    f.outside = 0.0f;
    f.uses_outside = 0;
    return f;
}

// ------------------------------------------------------------------ clamped MacCormack

// Pass 1 at one sample: the clamp and the backward semi-Lagrangian pass. hat holds the forward pass's values.
__device__ void maccormack_adjoint_1(const plm_grid_info &g, const FieldDesc &f, const float *__restrict__ original,
                                     const float *__restrict__ hat, float next_bar, float *__restrict__ orig_bar,
                                     float *__restrict__ hat_bar, const Vel &U, const VelBar &Ubar, int i, int j,
                                     int k, float dto) {
    if (next_bar == 0.0f) return;
    const float px = (float)i + f.ox, py = (float)j + f.oy, pz = (float)k + f.oz;
    float sx, sy, sz;
    trace(g, U.u, U.v, U.w, px, py, pz, dto, sx, sy, sz);
    if (f.uses_outside == 1 && outside_open_face(g, sx, sy, sz)) return;  // next = outside
    float bx, by, bz;
    trace(g, U.u, U.v, U.w, px, py, pz, -dto, bx, by, bz);
    const bool back_outside = f.uses_outside == 1 && outside_open_face(g, bx, by, bz);
    const TriW tb = tri_setup(f.n, bx - f.ox, by - f.oy, bz - f.oz);
    const float back_c = back_outside ? f.outside : sample_index(hat, f.n, bx - f.ox, by - f.oy, bz - f.oz);
    const int c = f.n.at(i, j, k);
    const float corrected = hat[c] + 0.5f * (original[c] - back_c);
    // range_index at the source, keeping the tap that fmin_ / fmax_ select (the later one on ties)
    const TriW ts = tri_setup(f.n, sx - f.ox, sy - f.oy, sz - f.oz);
    const int taps[8] = {f.n.at(ts.i0, ts.j0, ts.k0), f.n.at(ts.i0, ts.j0, ts.k1), f.n.at(ts.i0, ts.j1, ts.k0),
                         f.n.at(ts.i0, ts.j1, ts.k1), f.n.at(ts.i1, ts.j0, ts.k0), f.n.at(ts.i1, ts.j0, ts.k1),
                         f.n.at(ts.i1, ts.j1, ts.k0), f.n.at(ts.i1, ts.j1, ts.k1)};
    float lo = original[taps[0]], hi = lo;
    int lo_tap = taps[0], hi_tap = taps[0];
    for (int t = 0; t < 8; t++) {
        const float value = original[taps[t]];
        if (!(lo < value)) {
            lo = value;
            lo_tap = taps[t];
        }
        if (!(hi > value)) {
            hi = value;
            hi_tap = taps[t];
        }
    }
    // The clamp passes the gradient through wherever lo <= corrected <= hi, and also in a numerically flat
    // stencil: where the eight taps agree to rounding, the forward's corrected value lands inside or a few
    // ulps outside [lo, hi] by rounding alone, while a perturbation moves the taps far more than that and the
    // clamp stays inactive. Beyond that the gradient goes to the tap the clamp selected.
    const float flat = 8.0f * __FLT_EPSILON__ * fmaxf(fabsf(lo), fabsf(hi));
    const bool rounding_only = hi - lo <= flat && corrected <= hi + flat && corrected >= lo - flat;
    if (corrected > hi && !rounding_only) {
        atomicAdd(orig_bar + hi_tap, next_bar);
        return;
    }
    if (corrected < lo && !rounding_only) {
        atomicAdd(orig_bar + lo_tap, next_bar);
        return;
    }
    atomicAdd(hat_bar + c, next_bar);
    atomicAdd(orig_bar + c, 0.5f * next_bar);
    const float back_bar = -0.5f * next_bar;
    if (back_outside) return;
    tri_scatter(hat_bar, f.n, tb, back_bar);
    float dx, dy, dz;
    tri_grad(hat, f.n, tb, dx, dy, dz);
    const float sbar[3] = {back_bar * dx, back_bar * dy, back_bar * dz};
    trace_adjoint(g, U, Ubar, px, py, pz, -dto, sbar);
}

// Pass 2 at one sample: the forward semi-Lagrangian pass, with this sample's complete hat adjoint
__device__ void maccormack_adjoint_2(const plm_grid_info &g, const FieldDesc &f, const float *__restrict__ original,
                                     float hat_bar, float *__restrict__ orig_bar, const Vel &U, const VelBar &Ubar,
                                     int i, int j, int k, float dto) {
    if (hat_bar == 0.0f) return;
    const float px = (float)i + f.ox, py = (float)j + f.oy, pz = (float)k + f.oz;
    float sx, sy, sz;
    trace(g, U.u, U.v, U.w, px, py, pz, dto, sx, sy, sz);
    if (f.uses_outside == 1 && outside_open_face(g, sx, sy, sz)) return;
    const TriW t = tri_setup(f.n, sx - f.ox, sy - f.oy, sz - f.oz);
    tri_scatter(orig_bar, f.n, t, hat_bar);
    float dx, dy, dz;
    tri_grad(original, f.n, t, dx, dy, dz);
    const float sbar[3] = {hat_bar * dx, hat_bar * dy, hat_bar * dz};
    trace_adjoint(g, U, Ubar, px, py, pz, dto, sbar);
}

__global__ __launch_bounds__(256) void k_adj_velocity_1(plm_grid_info g, Tri<const float> old, Tri<const float> hat,
                                                        Tri<const float> next_bar, Tri<float> in_bar,
                                                        Tri<float> hat_bar, float dto) {
    int axis, i, j, k;
    Dims d;
    if (!thread_face(g, axis, d, i, j, k)) return;
    const FieldDesc f = face_field(g, axis);
    const Vel U{old.p[0], old.p[1], old.p[2]};
    const VelBar Ub{in_bar.p[0], in_bar.p[1], in_bar.p[2]};
    maccormack_adjoint_1(g, f, old.p[axis], hat.p[axis], next_bar.p[axis][d.at(i, j, k)], in_bar.p[axis],
                         hat_bar.p[axis], U, Ub, i, j, k, dto);
}

__global__ __launch_bounds__(256) void k_adj_velocity_2(plm_grid_info g, Tri<const float> old,
                                                        Tri<const float> hat_bar, Tri<float> in_bar, float dto) {
    int axis, i, j, k;
    Dims d;
    if (!thread_face(g, axis, d, i, j, k)) return;
    const FieldDesc f = face_field(g, axis);
    const Vel U{old.p[0], old.p[1], old.p[2]};
    const VelBar Ub{in_bar.p[0], in_bar.p[1], in_bar.p[2]};
    maccormack_adjoint_2(g, f, old.p[axis], hat_bar.p[axis][d.at(i, j, k)], in_bar.p[axis], U, Ub, i, j, k, dto);
}

struct ScalarPair {
    const float *data[2];
    const float *hat[2];
    const float *next_bar[2];
    float *data_bar[2];
    float *hat_bar[2];
    float outside[2];
};

__global__ __launch_bounds__(256) void k_adj_scalars_1(plm_grid_info g, ScalarPair sp, Tri<const float> vel,
                                                       Tri<float> vel_bar, float dto) {
    const Dims n = cell_dims(g);
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    const Vel U{vel.p[0], vel.p[1], vel.p[2]};
    const VelBar Ub{vel_bar.p[0], vel_bar.p[1], vel_bar.p[2]};
    for (int q = 0; q < 2; q++) {
        const FieldDesc f{n, 0.5f, 0.5f, 0.5f, sp.outside[q], 1};
        maccormack_adjoint_1(g, f, sp.data[q], sp.hat[q], sp.next_bar[q][n.at(i, j, k)], sp.data_bar[q],
                             sp.hat_bar[q], U, Ub, i, j, k, dto);
    }
}

__global__ __launch_bounds__(256) void k_adj_scalars_2(plm_grid_info g, ScalarPair sp, Tri<const float> vel,
                                                       Tri<float> vel_bar, float dto) {
    const Dims n = cell_dims(g);
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    const Vel U{vel.p[0], vel.p[1], vel.p[2]};
    const VelBar Ub{vel_bar.p[0], vel_bar.p[1], vel_bar.p[2]};
    for (int q = 0; q < 2; q++) {
        const FieldDesc f{n, 0.5f, 0.5f, 0.5f, sp.outside[q], 1};
        maccormack_adjoint_2(g, f, sp.data[q], sp.hat_bar[q][n.at(i, j, k)], sp.data_bar[q], U, Ub, i, j, k, dto);
    }
}

// forward semi-Lagrangian values (the hat of each field), recomputed for the adjoint
__global__ __launch_bounds__(256) void k_sl_faces(plm_grid_info g, Tri<const float> old, Tri<float> hat, float dto) {
    int axis, i, j, k;
    Dims d;
    if (!thread_face(g, axis, d, i, j, k)) return;
    const FieldDesc f = face_field(g, axis);
    const float px = (float)i + f.ox, py = (float)j + f.oy, pz = (float)k + f.oz;
    float sx, sy, sz;
    trace(g, old.p[0], old.p[1], old.p[2], px, py, pz, dto, sx, sy, sz);
    hat.p[axis][d.at(i, j, k)] = sample_index(old.p[axis], f.n, sx - f.ox, sy - f.oy, sz - f.oz);
}

__global__ __launch_bounds__(256) void k_sl_scalars(plm_grid_info g, ScalarPair sp, Tri<const float> vel,
                                                    float *hat0, float *hat1, float dto) {
    const Dims n = cell_dims(g);
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    const float px = (float)i + 0.5f, py = (float)j + 0.5f, pz = (float)k + 0.5f;
    float sx, sy, sz;
    trace(g, vel.p[0], vel.p[1], vel.p[2], px, py, pz, dto, sx, sy, sz);
    const bool out = outside_open_face(g, sx, sy, sz);
    const int c = n.at(i, j, k);
    hat0[c] = out ? sp.outside[0] : sample_index(sp.data[0], n, sx - 0.5f, sy - 0.5f, sz - 0.5f);
    hat1[c] = out ? sp.outside[1] : sample_index(sp.data[1], n, sx - 0.5f, sy - 0.5f, sz - 0.5f);
}

// ------------------------------------------------------------------ buoyancy

// v += gravity (hb (T - amb) / amb - sw s) dt on v faces, T and s averaged over the cells j - 1 and j
// (clamped). Gathers into the cells: cell j is `above` for face j and `below` for face j + 1.
__global__ __launch_bounds__(256) void k_adj_buoyancy(plm_grid_info g, const float *__restrict__ v_bar,
                                                      const float *__restrict__ smoke, float gravity,
                                                      float heat_buoyancy, float smoke_weight, float ambient_k,
                                                      float dt, float *__restrict__ smoke_bar,
                                                      float *__restrict__ temperature_bar,
                                                      float *__restrict__ weight_terms) {
    const Dims nv = face_dims(g, 1), nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nv, i, j, k)) return;
    const int f = nv.at(i, j, k);
    const int below = max(j - 1, 0), above = min(j, g.dims[1] - 1);
    const float density = 0.5f * (smoke[nc.at(i, below, k)] + smoke[nc.at(i, above, k)]);
    // d v / d smoke_weight
    weight_terms[f] = v_bar[f] * (-gravity * density * dt);
    if (j >= g.dims[1]) return;  // cell rows: the thread of face j serves cell j
    const int c = nc.at(i, j, k);
    // faces touching cell j: j (as above; as below too when j = 0) and j + 1 (as below; as above too at the top)
    float sum = 0.0f;
    const float face_lo = v_bar[nv.at(i, j, k)], face_hi = v_bar[nv.at(i, j + 1, k)];
    sum += 0.5f * face_lo * (j == 0 ? 2.0f : 1.0f);
    sum += 0.5f * face_hi * (j + 1 == g.dims[1] ? 2.0f : 1.0f);
    const float lift_bar = gravity * dt * sum;
    temperature_bar[c] += lift_bar * heat_buoyancy / ambient_k;
    smoke_bar[c] += -lift_bar * smoke_weight;
}

// ------------------------------------------------------------------ confinement

__device__ __forceinline__ void center_at(const plm_grid_info &g, const Vel &U, int i, int j, int k, float c[3]) {
    const Dims nu = face_dims(g, 0), nv = face_dims(g, 1), nw = face_dims(g, 2);
    c[0] = 0.5f * (U.u[nu.at(i, j, k)] + U.u[nu.at(i + 1, j, k)]);
    c[1] = 0.5f * (U.v[nv.at(i, j, k)] + U.v[nv.at(i, j + 1, k)]);
    c[2] = 0.5f * (U.w[nw.at(i, j, k)] + U.w[nw.at(i, j, k + 1)]);
}

// curl (vec3 per cell) of the centre velocity, as k_curl
__global__ __launch_bounds__(256) void k_adj_curl_forward(plm_grid_info g, Tri<const float> vel,
                                                          float *__restrict__ curl) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const Vel U{vel.p[0], vel.p[1], vel.p[2]};
    const int cell[3] = {i, j, k};
    float d[3][3];
    for (int a = 0; a < 3; a++) {
        int lo[3] = {i, j, k}, hi[3] = {i, j, k};
        lo[a] = max(cell[a] - 1, 0);
        hi[a] = min(cell[a] + 1, g.dims[a] - 1);
        const float span = (float)(hi[a] - lo[a]);
        if (span == 0.0f) {
            d[a][0] = d[a][1] = d[a][2] = 0.0f;
            continue;
        }
        float h[3], l[3];
        center_at(g, U, hi[0], hi[1], hi[2], h);
        center_at(g, U, lo[0], lo[1], lo[2], l);
        const float denom = span * g.cell_m;
        for (int q = 0; q < 3; q++) d[a][q] = (h[q] - l[q]) / denom;
    }
    const int c = nc.at(i, j, k);
    curl[3 * c] = d[1][2] - d[2][1];
    curl[3 * c + 1] = d[2][0] - d[0][2];
    curl[3 * c + 2] = d[0][1] - d[1][0];
}

__device__ __forceinline__ float spin(const float *__restrict__ curl, int c) {
    return length3(curl[3 * c], curl[3 * c + 1], curl[3 * c + 2]);
}

// push_bar per cell from the faces' gradients (the transpose of add_face_push's averages, times dt)
__global__ __launch_bounds__(256) void k_adj_push_bar(plm_grid_info g, Tri<const float> f_bar, float dt,
                                                      float *__restrict__ push_bar) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const int cell[3] = {i, j, k};
    const int c = nc.at(i, j, k);
    for (int a = 0; a < 3; a++) {
        const Dims nf = face_dims(g, a);
        int lo[3] = {i, j, k}, hi[3] = {i, j, k};
        hi[a] = cell[a] + 1;
        // face cell[a] has this cell as `above` (and as `below` too when it is face 0); face cell[a] + 1 has
        // it as `below` (and as `above` too when it is the last face)
        float sum = 0.5f * f_bar.p[a][nf.at(lo[0], lo[1], lo[2])] * (cell[a] == 0 ? 2.0f : 1.0f);
        sum += 0.5f * f_bar.p[a][nf.at(hi[0], hi[1], hi[2])] * (cell[a] + 1 == g.dims[a] ? 2.0f : 1.0f);
        push_bar[3 * c + a] = sum * dt;
    }
}

// From push_bar: the curl's direct part (into curl_bar) and the spin gradient's adjoint (into toward_bar)
__global__ __launch_bounds__(256) void k_adj_push(plm_grid_info g, const float *__restrict__ curl,
                                                  const float *__restrict__ push_bar, float strength,
                                                  float *__restrict__ curl_bar, float *__restrict__ spin_bar) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const int cell[3] = {i, j, k};
    float toward[3] = {0.0f, 0.0f, 0.0f};
    int lo_c[3], hi_c[3];
    float denom[3];
    for (int a = 0; a < 3; a++) {
        int lo[3] = {i, j, k}, hi[3] = {i, j, k};
        lo[a] = max(cell[a] - 1, 0);
        hi[a] = min(cell[a] + 1, g.dims[a] - 1);
        const float span = (float)(hi[a] - lo[a]);
        lo_c[a] = nc.at(lo[0], lo[1], lo[2]);
        hi_c[a] = nc.at(hi[0], hi[1], hi[2]);
        denom[a] = span * g.cell_m;
        if (span > 0.0f) toward[a] = (spin(curl, hi_c[a]) - spin(curl, lo_c[a])) / denom[a];
    }
    const int c = nc.at(i, j, k);
    const float length = length3(toward[0], toward[1], toward[2]);
    if (length < 1.0e-12f) return;
    const float N[3] = {toward[0] / length, toward[1] / length, toward[2] / length};
    const float W[3] = {curl[3 * c], curl[3 * c + 1], curl[3 * c + 2]};
    const float s = strength * g.cell_m;
    const float P[3] = {push_bar[3 * c] * s, push_bar[3 * c + 1] * s, push_bar[3 * c + 2] * s};
    // push = s (N x W): N_bar = W x P, W_bar = P x N
    const float Nb[3] = {W[1] * P[2] - W[2] * P[1], W[2] * P[0] - W[0] * P[2], W[0] * P[1] - W[1] * P[0]};
    atomicAdd(curl_bar + 3 * c, P[1] * N[2] - P[2] * N[1]);
    atomicAdd(curl_bar + 3 * c + 1, P[2] * N[0] - P[0] * N[2]);
    atomicAdd(curl_bar + 3 * c + 2, P[0] * N[1] - P[1] * N[0]);
    // N = toward / |toward|
    const float dot = N[0] * Nb[0] + N[1] * Nb[1] + N[2] * Nb[2];
    for (int a = 0; a < 3; a++) {
        const float tb = (Nb[a] - N[a] * dot) / length;
        if (denom[a] > 0.0f) {
            atomicAdd(spin_bar + hi_c[a], tb / denom[a]);
            atomicAdd(spin_bar + lo_c[a], -tb / denom[a]);
        }
    }
}

// spin = |curl|: curl_bar += spin_bar curl / |curl|
__global__ __launch_bounds__(256) void k_adj_spin(plm_grid_info g, const float *__restrict__ curl,
                                                  const float *__restrict__ spin_bar, float *__restrict__ curl_bar) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const int c = nc.at(i, j, k);
    const float m = spin(curl, c);
    if (m <= 0.0f || spin_bar[c] == 0.0f) return;
    for (int q = 0; q < 3; q++) curl_bar[3 * c + q] += spin_bar[c] * curl[3 * c + q] / m;
}

// curl of the centre velocity: centre_bar (scattered), then the faces (gathered in k_adj_center)
__global__ __launch_bounds__(256) void k_adj_curl(plm_grid_info g, const float *__restrict__ curl_bar,
                                                  float *__restrict__ center_bar) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const int cell[3] = {i, j, k};
    const int c = nc.at(i, j, k);
    const float cb[3] = {curl_bar[3 * c], curl_bar[3 * c + 1], curl_bar[3 * c + 2]};
    // curl = (d[1][2] - d[2][1], d[2][0] - d[0][2], d[0][1] - d[1][0]): d_bar[a][q]
    float db[3][3] = {{0.0f, cb[2], -cb[1]}, {-cb[2], 0.0f, cb[0]}, {cb[1], -cb[0], 0.0f}};
    for (int a = 0; a < 3; a++) {
        int lo[3] = {i, j, k}, hi[3] = {i, j, k};
        lo[a] = max(cell[a] - 1, 0);
        hi[a] = min(cell[a] + 1, g.dims[a] - 1);
        const float span = (float)(hi[a] - lo[a]);
        if (span == 0.0f) continue;
        const float denom = span * g.cell_m;
        const int ch = nc.at(hi[0], hi[1], hi[2]), cl = nc.at(lo[0], lo[1], lo[2]);
        for (int q = 0; q < 3; q++) {
            if (db[a][q] == 0.0f) continue;
            atomicAdd(center_bar + 3 * ch + q, db[a][q] / denom);
            atomicAdd(center_bar + 3 * cl + q, -db[a][q] / denom);
        }
    }
}

// centre = face averages: each face gathers half of the centre_bar of the cells on both sides
__global__ __launch_bounds__(256) void k_adj_center(plm_grid_info g, const float *__restrict__ center_bar,
                                                    Tri<float> f_bar) {
    int axis, i, j, k;
    Dims d;
    if (!thread_face(g, axis, d, i, j, k)) return;
    const Dims nc = cell_dims(g);
    const int face[3] = {i, j, k};
    float sum = 0.0f;
    if (face[axis] > 0) {
        int b[3] = {i, j, k};
        b[axis] -= 1;
        sum += 0.5f * center_bar[3 * nc.at(b[0], b[1], b[2]) + axis];
    }
    if (face[axis] < g.dims[axis]) sum += 0.5f * center_bar[3 * nc.at(i, j, k) + axis];
    f_bar.p[axis][d.at(i, j, k)] += sum;
}

// ------------------------------------------------------------------ projection transposes

// rhs = G^T f_bar: G q at a face not closed = (q above - q below) / cell_m, ghost q = 0 past open faces;
// closed boundary faces are never written by the gradient, so they contribute nothing
__global__ __launch_bounds__(256) void k_adj_gradient_t(plm_grid_info g, Tri<const float> f_bar,
                                                        float *__restrict__ rhs) {
    const Dims nc = cell_dims(g);
    int i, j, k;
    if (!thread_cell(nc, i, j, k)) return;
    const int cell[3] = {i, j, k};
    float sum = 0.0f;
    for (int a = 0; a < 3; a++) {
        const Dims nf = face_dims(g, a);
        int hi[3] = {i, j, k};
        hi[a] = cell[a] + 1;
        const bool lo_closed = cell[a] == 0 && g.open_lo[a] == 0;
        const bool hi_closed = cell[a] + 1 == g.dims[a] && g.open_hi[a] == 0;
        if (!lo_closed) sum += f_bar.p[a][nf.at(i, j, k)];
        if (!hi_closed) sum -= f_bar.p[a][nf.at(hi[0], hi[1], hi[2])];
    }
    rhs[nc.at(i, j, k)] = sum / g.cell_m;
}

// f_bar -= D^T y: divergence = sum over axes of (f[hi] - f[lo]) / cell_m; closed faces end at 0
__global__ __launch_bounds__(256) void k_adj_divergence_t(plm_grid_info g, const float *__restrict__ y,
                                                          Tri<float> f_bar) {
    int axis, i, j, k;
    Dims d;
    if (!thread_face(g, axis, d, i, j, k)) return;
    const Dims nc = cell_dims(g);
    const int face[3] = {i, j, k};
    const bool closed = (face[axis] == 0 && g.open_lo[axis] == 0) || (face[axis] == g.dims[axis] && g.open_hi[axis] == 0);
    float *p = f_bar.p[axis] + d.at(i, j, k);
    if (closed) {
        *p = 0.0f;
        return;
    }
    float dt = 0.0f;
    if (face[axis] > 0) {  // the face is the `hi` face of the cell below
        int b[3] = {i, j, k};
        b[axis] -= 1;
        dt += y[nc.at(b[0], b[1], b[2])];
    }
    if (face[axis] < g.dims[axis]) dt -= y[nc.at(i, j, k)];
    *p = *p - dt / g.cell_m;
}

// ------------------------------------------------------------------ ray march: heat and thickness

// The orthographic pixel ray of raymarch_kernel (render/raymarch.py:72-83) and its span through the box
__device__ __forceinline__ bool pixel_ray(const plm_grid_info &g, const plm_view &view, int px, int py,
                                          float origin[3], float dir[3], float &span0, float &span1) {
    const float plane_x = view.left_m + ((float)px + 0.5f) * view.pixel_m;
    const float plane_y = view.top_m - ((float)py + 0.5f) * view.pixel_m;
    const float start_m = ortho_start_m(view);
    for (int a = 0; a < 3; a++) {
        origin[a] = view.right[a] * plane_x + view.up[a] * plane_y + view.toward[a] * start_m;
        dir[a] = -view.toward[a];
    }
    float box_max[3];
    for (int a = 0; a < 3; a++) box_max[a] = g.origin_m[a] + (float)g.dims[a] * g.cell_m;
    slab_range(origin, dir, g.origin_m, box_max, span0, span1);
    return span1 > span0;
}

// Per pixel and sample the recording pass keeps the transmittance before the sample, the opacity, the
// density (before the max with 0) and the temperature (max_steps samples per ray, fixed).
__global__ __launch_bounds__(256) void k_adj_raymarch(plm_grid_info g, const float *__restrict__ smoke,
                                                      const float *__restrict__ temperature, plm_view view,
                                                      plm_look look, const float *__restrict__ image_bar, int height,
                                                      int width, int max_steps, int record_stride,
                                                      float *__restrict__ record, float *__restrict__ smoke_bar,
                                                      float *__restrict__ temperature_bar) {
    const int px = blockIdx.x * 16 + threadIdx.x % 16, py = blockIdx.y * 16 + threadIdx.x / 16;
    if (px >= width || py >= height) return;
    const Dims n = cell_dims(g);
    float origin[3], dir[3], span0, span1;
    if (!pixel_ray(g, view, px, py, origin, dir, span0, span1)) return;
    const size_t pixel = (size_t)py * width + px;
    float *rec = record + pixel * (size_t)record_stride * 4;
    const float length = span1 - span0;
    int steps = (int)ceilf(length / (0.5f * g.cell_m));
    if (max_steps > 0) steps = max_steps;
    steps = min(steps, record_stride);  // the host sizes the stride for the longest ray
    const float step_m = length / (float)steps;
    const float hides = look.smoke_hides_heat;
    float tr = 1.0f;
    for (int s = 0; s < steps; s++) {
        const float t = span0 + ((float)s + 0.5f) * step_m;
        float gp[3];
        for (int a = 0; a < 3; a++) gp[a] = (origin[a] + dir[a] * t - g.origin_m[a]) / g.cell_m - 0.5f;
        const float rho = sample_index(smoke, n, gp[0], gp[1], gp[2]);
        const float tau = sample_index(temperature, n, gp[0], gp[1], gp[2]);
        const float alpha = 1.0f - expf(-look.opacity_per_density * fmax_(rho, 0.0f) * step_m);
        rec[4 * s] = tr;
        rec[4 * s + 1] = alpha;
        rec[4 * s + 2] = rho;
        rec[4 * s + 3] = tau;
        tr = tr * (1.0f - alpha);
    }
    const float heat_bar = image_bar[pixel * 5], thick_bar = image_bar[pixel * 5 + 1];
    float tr_bar = -thick_bar;  // thickness = 1 - final transmittance
    for (int s = steps - 1; s >= 0; s--) {
        const float trs = rec[4 * s], alpha = rec[4 * s + 1], rho = rec[4 * s + 2], tau = rec[4 * s + 3];
        // tr_{s+1} = tr_s (1 - alpha)
        const float alpha_bar = -tr_bar * trs;
        float trs_bar = tr_bar * (1.0f - alpha);
        // heat += max(tau - glow, 0) step tr_s^hides
        const float glow = fmax_(tau - look.glow_start_k, 0.0f);
        const float hw = hides == 0.0f ? 1.0f : (hides == 1.0f ? trs : powf(trs, hides));
        const float tau_bar = tau > look.glow_start_k ? heat_bar * step_m * hw : 0.0f;
        if (hides != 0.0f && trs > 0.0f)
            trs_bar += heat_bar * glow * step_m * (hides == 1.0f ? 1.0f : hides * powf(trs, hides - 1.0f));
        // alpha = 1 - exp(-opd max(rho, 0) step)
        // max(rho, 0): slope 1 above 0, 0 below, and at exactly 0 (a sample of empty space) the mean of the
        // two one-sided slopes, which is what a central difference there sees
        const float slope = rho > 0.0f ? 1.0f : rho == 0.0f ? 0.5f : 0.0f;
        const float rho_bar = slope * alpha_bar * look.opacity_per_density * step_m * (1.0f - alpha);
        const float t = span0 + ((float)s + 0.5f) * step_m;
        float gp[3];
        for (int a = 0; a < 3; a++) gp[a] = (origin[a] + dir[a] * t - g.origin_m[a]) / g.cell_m - 0.5f;
        const TriW w = tri_setup(n, gp[0], gp[1], gp[2]);
        tri_scatter(smoke_bar, n, w, rho_bar);
        tri_scatter(temperature_bar, n, w, tau_bar);
        tr_bar = trs_bar;
    }
}

}  // namespace
}  // namespace plm

using namespace plm;

extern "C" {

int plm_adj_buoyancy(const plm_grid_info *info, const float *v_bar, const float *smoke, float gravity,
                     float heat_buoyancy, float smoke_weight, float ambient_k, float dt, float *smoke_bar,
                     float *temperature_bar, float *weight_terms, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(v_bar && smoke && smoke_bar && temperature_bar && weight_terms, "plm_adj_buoyancy: NULL array");
    PLM_REQUIRE(ambient_k != 0.0f, "plm_adj_buoyancy: ambient_k is 0");
    const Dims nv = face_dims(*info, 1);
    k_adj_buoyancy<<<grid3(nv), block3(), 0, as_stream(stream)>>>(*info, v_bar, smoke, gravity, heat_buoyancy,
                                                                  smoke_weight, ambient_k, dt, smoke_bar,
                                                                  temperature_bar, weight_terms);
    return check_launch("adjoint buoyancy");
}

int plm_adj_confinement(const plm_grid_info *info, const float *u, const float *v, const float *w, float strength,
                        float dt, float *u_bar, float *v_bar, float *w_bar, float *curl, float *push_bar,
                        float *curl_bar, float *spin_bar, float *center_bar, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && u_bar && v_bar && w_bar && curl && push_bar && curl_bar && spin_bar && center_bar,
                "plm_adj_confinement: NULL array");
    if (strength == 0.0f) return 0;
    hipStream_t s = as_stream(stream);
    const Dims nc = cell_dims(*info);
    const int64_t cells = (int64_t)nc.count();
    PLM_TRY(zero_fill(curl_bar, 3 * cells, s));
    PLM_TRY(zero_fill(spin_bar, cells, s));
    PLM_TRY(zero_fill(center_bar, 3 * cells, s));
    k_adj_curl_forward<<<grid3(nc), block3(), 0, s>>>(*info, Tri<const float>{{u, v, w}}, curl);
    PLM_TRY(check_launch("adjoint curl forward"));
    k_adj_push_bar<<<grid3(nc), block3(), 0, s>>>(*info, Tri<const float>{{u_bar, v_bar, w_bar}}, dt, push_bar);
    PLM_TRY(check_launch("adjoint push bar"));
    k_adj_push<<<grid3(nc), block3(), 0, s>>>(*info, curl, push_bar, strength, curl_bar, spin_bar);
    PLM_TRY(check_launch("adjoint push"));
    k_adj_spin<<<grid3(nc), block3(), 0, s>>>(*info, curl, spin_bar, curl_bar);
    PLM_TRY(check_launch("adjoint spin"));
    k_adj_curl<<<grid3(nc), block3(), 0, s>>>(*info, curl_bar, center_bar);
    PLM_TRY(check_launch("adjoint curl"));
    k_adj_center<<<grid_faces(*info), block3(), 0, s>>>(*info, center_bar, Tri<float>{{u_bar, v_bar, w_bar}});
    return check_launch("adjoint centre");
}

int plm_adj_advect_velocity(const plm_grid_info *info, const float *u, const float *v, const float *w, float *hat_u,
                            float *hat_v, float *hat_w, const float *next_u_bar, const float *next_v_bar,
                            const float *next_w_bar, float *u_bar, float *v_bar, float *w_bar, float *hat_u_bar,
                            float *hat_v_bar, float *hat_w_bar, float dto, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && hat_u && hat_v && hat_w && next_u_bar && next_v_bar && next_w_bar && u_bar && v_bar &&
                    w_bar && hat_u_bar && hat_v_bar && hat_w_bar,
                "plm_adj_advect_velocity: NULL array");
    hipStream_t s = as_stream(stream);
    const Tri<const float> old{{u, v, w}};
    k_sl_faces<<<grid_faces(*info), block3(), 0, s>>>(*info, old, Tri<float>{{hat_u, hat_v, hat_w}}, dto);
    PLM_TRY(check_launch("adjoint velocity forward"));
    float *hb[3] = {hat_u_bar, hat_v_bar, hat_w_bar};
    for (int a = 0; a < 3; a++) PLM_TRY(zero_fill(hb[a], face_dims(*info, a).count(), s));
    k_adj_velocity_1<<<grid_faces(*info), block3(), 0, s>>>(*info, old, Tri<const float>{{hat_u, hat_v, hat_w}},
                                                            Tri<const float>{{next_u_bar, next_v_bar, next_w_bar}},
                                                            Tri<float>{{u_bar, v_bar, w_bar}},
                                                            Tri<float>{{hat_u_bar, hat_v_bar, hat_w_bar}}, dto);
    PLM_TRY(check_launch("adjoint velocity pass 1"));
    k_adj_velocity_2<<<grid_faces(*info), block3(), 0, s>>>(
        *info, old, Tri<const float>{{hat_u_bar, hat_v_bar, hat_w_bar}}, Tri<float>{{u_bar, v_bar, w_bar}}, dto);
    return check_launch("adjoint velocity pass 2");
}

int plm_adj_advect_scalars(const plm_grid_info *info, const float *u, const float *v, const float *w,
                           const float *a_data, const float *b_data, float a_outside, float b_outside, float *a_hat,
                           float *b_hat, const float *a_next_bar, const float *b_next_bar, float *a_bar,
                           float *b_bar, float *u_bar, float *v_bar, float *w_bar, float *a_hat_bar,
                           float *b_hat_bar, float dto, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && a_data && b_data && a_hat && b_hat && a_next_bar && b_next_bar && a_bar && b_bar &&
                    u_bar && v_bar && w_bar && a_hat_bar && b_hat_bar,
                "plm_adj_advect_scalars: NULL array");
    hipStream_t s = as_stream(stream);
    const Dims nc = cell_dims(*info);
    ScalarPair sp{};
    sp.data[0] = a_data;
    sp.data[1] = b_data;
    sp.hat[0] = a_hat;
    sp.hat[1] = b_hat;
    sp.next_bar[0] = a_next_bar;
    sp.next_bar[1] = b_next_bar;
    sp.data_bar[0] = a_bar;
    sp.data_bar[1] = b_bar;
    sp.hat_bar[0] = a_hat_bar;
    sp.hat_bar[1] = b_hat_bar;
    sp.outside[0] = a_outside;
    sp.outside[1] = b_outside;
    const Tri<const float> vel{{u, v, w}};
    k_sl_scalars<<<grid3(nc), block3(), 0, s>>>(*info, sp, vel, a_hat, b_hat, dto);
    PLM_TRY(check_launch("adjoint scalars forward"));
    PLM_TRY(zero_fill(a_hat_bar, nc.count(), s));
    PLM_TRY(zero_fill(b_hat_bar, nc.count(), s));
    k_adj_scalars_1<<<grid3(nc), block3(), 0, s>>>(*info, sp, vel, Tri<float>{{u_bar, v_bar, w_bar}}, dto);
    PLM_TRY(check_launch("adjoint scalars pass 1"));
    k_adj_scalars_2<<<grid3(nc), block3(), 0, s>>>(*info, sp, vel, Tri<float>{{u_bar, v_bar, w_bar}}, dto);
    return check_launch("adjoint scalars pass 2");
}

int plm_adj_gradient_t(const plm_grid_info *info, const float *u_bar, const float *v_bar, const float *w_bar,
                       float *rhs, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u_bar && v_bar && w_bar && rhs, "plm_adj_gradient_t: NULL array");
    k_adj_gradient_t<<<grid3(cell_dims(*info)), block3(), 0, as_stream(stream)>>>(
        *info, Tri<const float>{{u_bar, v_bar, w_bar}}, rhs);
    return check_launch("adjoint gradient transpose");
}

int plm_adj_divergence_t(const plm_grid_info *info, const float *y, float *u_bar, float *v_bar, float *w_bar,
                         void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(y && u_bar && v_bar && w_bar, "plm_adj_divergence_t: NULL array");
    k_adj_divergence_t<<<grid_faces(*info), block3(), 0, as_stream(stream)>>>(*info, y,
                                                                             Tri<float>{{u_bar, v_bar, w_bar}});
    return check_launch("adjoint divergence transpose");
}

int plm_adj_raymarch_record_stride(const plm_grid_info *info, int32_t max_steps) {
    if (info == nullptr) return -1;
    if (max_steps > 0) return max_steps;
    // the longest chord of the box, in half cells, plus one
    const double d = sqrt((double)info->dims[0] * info->dims[0] + (double)info->dims[1] * info->dims[1] +
                          (double)info->dims[2] * info->dims[2]);
    return (int32_t)ceil(2.0 * d) + 2;
}

int plm_adj_raymarch(const plm_grid_info *info, const float *smoke, const float *temperature, const plm_view *view,
                     const plm_look *look, const float *image_bar, int32_t height, int32_t width, int32_t max_steps,
                     float *record, float *smoke_bar, float *temperature_bar, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(smoke && temperature && view && look && image_bar && record && smoke_bar && temperature_bar,
                "plm_adj_raymarch: NULL argument");
    PLM_REQUIRE(height > 0 && width > 0 && max_steps >= 0, "plm_adj_raymarch: bad image size or step count");
    PLM_REQUIRE(view->perspective == PLM_VIEW_ORTHO || view->perspective == PLM_VIEW_ORTHO_NEAR,
                "plm_adj_raymarch: orthographic views only");
    const int stride = plm_adj_raymarch_record_stride(info, max_steps);
    const dim3 grid((width + 15) / 16, (height + 15) / 16);
    k_adj_raymarch<<<grid, 256, 0, as_stream(stream)>>>(*info, smoke, temperature, *view, *look, image_bar, height,
                                                        width, max_steps, stride, record, smoke_bar,
                                                        temperature_bar);
    return check_launch("adjoint raymarch");
}

}  // extern "C"
