// The volume ray march and its per-frame volume passes: render/raymarch.py, render/light_volume.py,
// render/normal_field.py.
//
// One thread per camera pixel (16 x 16 pixel tiles, so neighbouring rays share cache lines) and one per cell
// for the light volume. Smoke, temperature and light share one grid, so each march sample computes the
// trilinear corner indices and weights once and gathers the three fields with them (the reference samples
// each field separately; the arithmetic per field is the same).
#include "common.h"

namespace plm {
namespace {


// The corner offsets and weights of sample_index_space (sim/sample.py:10-29) at one point.
struct Trilinear {
    int c[8];  // (i0,j0,k0) (i1,j0,k0) (i0,j1,k0) (i1,j1,k0) (i0,j0,k1) (i1,j0,k1) (i0,j1,k1) (i1,j1,k1)
    float tx, ty, tz;
    __device__ __forceinline__ Trilinear(Dims n, float gx, float gy, float gz) {
        const float x = clamp_(gx, 0.0f, (float)(n.n0 - 1));
        const float y = clamp_(gy, 0.0f, (float)(n.n1 - 1));
        const float z = clamp_(gz, 0.0f, (float)(n.n2 - 1));
        const int i0 = (int)x, j0 = (int)y, k0 = (int)z;
        const int i1 = min(i0 + 1, n.n0 - 1), j1 = min(j0 + 1, n.n1 - 1), k1 = min(k0 + 1, n.n2 - 1);
        tx = x - (float)i0;
        ty = y - (float)j0;
        tz = z - (float)k0;
        c[0] = n.at(i0, j0, k0);
        c[1] = n.at(i1, j0, k0);
        c[2] = n.at(i0, j1, k0);
        c[3] = n.at(i1, j1, k0);
        c[4] = n.at(i0, j0, k1);
        c[5] = n.at(i1, j0, k1);
        c[6] = n.at(i0, j1, k1);
        c[7] = n.at(i1, j1, k1);
    }
    __device__ __forceinline__ float operator()(const float *__restrict__ f) const {
        const float c00 = lerp_(f[c[0]], f[c[1]], tx);
        const float c10 = lerp_(f[c[2]], f[c[3]], tx);
        const float c01 = lerp_(f[c[4]], f[c[5]], tx);
        const float c11 = lerp_(f[c[6]], f[c[7]], tx);
        return lerp_(lerp_(c00, c10, ty), lerp_(c01, c11, ty), tz);
    }
    __device__ __forceinline__ float vec(const float *__restrict__ f, int q) const {
        const float c00 = lerp_(f[3 * c[0] + q], f[3 * c[1] + q], tx);
        const float c10 = lerp_(f[3 * c[2] + q], f[3 * c[3] + q], tx);
        const float c01 = lerp_(f[3 * c[4] + q], f[3 * c[5] + q], tx);
        const float c11 = lerp_(f[3 * c[6] + q], f[3 * c[7] + q], tx);
        return lerp_(lerp_(c00, c10, ty), lerp_(c01, c11, ty), tz);
    }
};

// pow(transmittance, smoke_hides_heat) with IEEE pow's exact cases pow(x, 1) = x and pow(x, 0) = 1
__device__ __forceinline__ float hide_weight(float tr, float hides) {
    if (hides == 1.0f) return tr;
    if (hides == 0.0f) return 1.0f;
    return powf(tr, hides);
}

// light_volume_kernel, render/light_volume.py:18-33
__global__ __launch_bounds__(256) void k_light_volume(Dims n, const float *__restrict__ smoke, float lx, float ly,
                                                      float lz, float cell_m, float opacity_per_density,
                                                      int max_steps, float *__restrict__ light) {
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    const float start[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
    const float dir[3] = {lx, ly, lz};
    const float zero[3] = {0.0f, 0.0f, 0.0f}, box_max[3] = {(float)n.n0, (float)n.n1, (float)n.n2};
    float t_enter, length;
    slab_range(start, dir, zero, box_max, t_enter, length);
    int steps = (int)ceilf(length / 0.5f);
    if (max_steps > 0 && steps > 0) steps = max_steps;
    float column = 0.0f;
    if (steps > 0) {
        const float step = length / (float)steps;
        auto sample = [&](int s) {
            const float t = ((float)s + 0.5f) * step;
            const float px = start[0] + lx * t, py = start[1] + ly * t, pz = start[2] + lz * t;
            return sample_index(smoke, n, px - 0.5f, py - 0.5f, pz - 0.5f);
        };
        // four samples' taps in flight at once, then added to the column in step order
        constexpr int UNROLL = 4;
        int s = 0;
        for (; s + UNROLL <= steps; s += UNROLL) {
            float v[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) v[u] = sample(s + u);
#pragma unroll
            for (int u = 0; u < UNROLL; u++) column = column + fmax_(v[u], 0.0f) * step;
        }
        for (; s < steps; s++) column = column + fmax_(sample(s), 0.0f) * step;
    }
    light[n.at(i, j, k)] = expf(-opacity_per_density * column * cell_m);
}

// blur_axis_kernel, render/normal_field.py:12-22
__global__ __launch_bounds__(256) void k_blur_axis(Dims n, const float *__restrict__ src, float *__restrict__ dst,
                                                   const float *__restrict__ weights, int radius, int axis) {
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    const int cell[3] = {i, j, k};
    const int extent = axis == 0 ? n.n0 : (axis == 1 ? n.n1 : n.n2);
    float total = 0.0f;
    for (int off = -radius; off <= radius; off++) {
        int nb[3] = {i, j, k};
        nb[axis] = min(max(cell[axis] + off, 0), extent - 1);
        total = total + weights[off + radius] * src[n.at(nb[0], nb[1], nb[2])];
    }
    dst[n.at(i, j, k)] = total;
}

// gradient_kernel, render/normal_field.py:25-38
__global__ __launch_bounds__(256) void k_gradient(Dims n, const float *__restrict__ src, float *__restrict__ gradient) {
    int i, j, k;
    if (!thread_cell(n, i, j, k)) return;
    const int cell[3] = {i, j, k};
    const int extent[3] = {n.n0, n.n1, n.n2};
    float result[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int lo[3] = {i, j, k}, hi[3] = {i, j, k};
        lo[a] = max(cell[a] - 1, 0);
        hi[a] = min(cell[a] + 1, extent[a] - 1);
        const float span = (float)(hi[a] - lo[a]);
        if (span > 0.0f) result[a] = (src[n.at(hi[0], hi[1], hi[2])] - src[n.at(lo[0], lo[1], lo[2])]) / span;
    }
    const int c = n.at(i, j, k);
    gradient[3 * c] = result[0];
    gradient[3 * c + 1] = result[1];
    gradient[3 * c + 2] = result[2];
}

// raymarch_kernel, render/raymarch.py:86-140 with pixel_ray (:72-83)
__global__ __launch_bounds__(256) void k_raymarch(plm_grid_info g, const float *__restrict__ smoke,
                                                  const float *__restrict__ temperature,
                                                  const float *__restrict__ gradient, Dims ng,
                                                  const float *__restrict__ light, Dims nl, plm_view view,
                                                  plm_look look, float *__restrict__ image, int height, int width,
                                                  int max_steps, bool shared_grid) {
    const int px = blockIdx.x * 16 + threadIdx.x % 16, py = blockIdx.y * 16 + threadIdx.x / 16;
    if (px >= width || py >= height) return;
    const Dims n = cell_dims(g);
    const float plane_x = view.left_m + ((float)px + 0.5f) * view.pixel_m;
    const float plane_y = view.top_m - ((float)py + 0.5f) * view.pixel_m;
    float origin[3], dir[3];
    const float start_m = ortho_start_m(view);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        origin[a] = view.right[a] * plane_x + view.up[a] * plane_y + view.toward[a] * start_m;
        dir[a] = -view.toward[a];
    }
    if (view.perspective == PLM_VIEW_PERSPECTIVE) {
        float d[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            origin[a] = view.eye[a];
            d[a] = view.right[a] * plane_x + view.up[a] * plane_y - view.toward[a];
        }
        const float l = length3(d[0], d[1], d[2]);
#pragma unroll
        for (int a = 0; a < 3; a++) dir[a] = l > 0.0f ? d[a] / l : 0.0f;
    }
    float box_max[3];
#pragma unroll
    for (int a = 0; a < 3; a++) box_max[a] = g.origin_m[a] + (float)g.dims[a] * g.cell_m;
    float span0, span1;
    slab_range(origin, dir, g.origin_m, box_max, span0, span1);
    if (view.perspective == PLM_VIEW_PERSPECTIVE) span0 = fmax_(span0, 0.0f);

    float tr = 1.0f, heat = 0.0f, shade_sum = 0.0f;
    float ns[3] = {0.0f, 0.0f, 0.0f};
    if (span1 > span0) {
        const float length = span1 - span0;
        int steps = (int)ceilf(length / (0.5f * g.cell_m));
        if (max_steps > 0) steps = max_steps;
        const float step_m = length / (float)steps;
        for (int s = 0; s < steps; s++) {
            const float t = span0 + ((float)s + 0.5f) * step_m;
            float gp[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float position = origin[a] + dir[a] * t;
                gp[a] = (position - g.origin_m[a]) / g.cell_m - 0.5f;
            }
            const Trilinear tri(n, gp[0], gp[1], gp[2]);
            const float density = fmax_(tri(smoke), 0.0f);
            const float alpha = 1.0f - expf(-look.opacity_per_density * density * step_m);
            const float glow = fmax_(tri(temperature) - look.glow_start_k, 0.0f);
            heat = heat + glow * step_m * hide_weight(tr, look.smoke_hides_heat);
            if (look.shade_layout == 1) {
                const float sampled = shared_grid ? tri(light) : sample_index(light, nl, gp[0], gp[1], gp[2]);
                const float lit = look.shade_fill + (1.0f - look.shade_fill) * sampled;
                shade_sum = shade_sum + tr * alpha * lit;
            } else {
                float slope[3];
                if (shared_grid) {
                    slope[0] = tri.vec(gradient, 0);
                    slope[1] = tri.vec(gradient, 1);
                    slope[2] = tri.vec(gradient, 2);
                } else {
                    const Trilinear tg(ng, gp[0], gp[1], gp[2]);
                    slope[0] = tg.vec(gradient, 0);
                    slope[1] = tg.vec(gradient, 1);
                    slope[2] = tg.vec(gradient, 2);
                }
                const float ls = length3(slope[0], slope[1], slope[2]);
                if (alpha > 0.0f && ls > 1.0e-9f) {
                    const float wgt = tr * alpha;
#pragma unroll
                    for (int a = 0; a < 3; a++) ns[a] = ns[a] - wgt * (slope[a] / ls);
                }
            }
            tr = tr * (1.0f - alpha);
        }
    }
    const float thickness = 1.0f - tr;
    float *pix = image + ((size_t)py * width + px) * 5;
    pix[0] = heat;
    pix[1] = thickness;
    if (look.shade_layout == 1) {
        pix[2] = shade_sum;
        pix[3] = 0.0f;
        pix[4] = 0.0f;
        return;
    }
    float normal[3] = {0.0f, 0.0f, 1.0f};
    const float cam[3] = {ns[0] * view.right[0] + ns[1] * view.right[1] + ns[2] * view.right[2],
                          ns[0] * view.up[0] + ns[1] * view.up[1] + ns[2] * view.up[2],
                          ns[0] * view.toward[0] + ns[1] * view.toward[1] + ns[2] * view.toward[2]};
    const float lc = length3(cam[0], cam[1], cam[2]);
    if (lc > 1.0e-9f) {
        const float nn[3] = {cam[0] / lc, cam[1] / lc, fabsf(cam[2] / lc)};
        const float weight = clamp_(thickness / look.normal_fade_thickness, 0.0f, 1.0f);
        const float facing[3] = {0.0f, 0.0f, 1.0f};
        float mixed[3];
#pragma unroll
        for (int a = 0; a < 3; a++) mixed[a] = lerp_(facing[a], nn[a], weight);
        const float lm = length3(mixed[0], mixed[1], mixed[2]);
#pragma unroll
        for (int a = 0; a < 3; a++) normal[a] = lm > 0.0f ? mixed[a] / lm : 0.0f;
    }
    pix[2] = normal[0];
    pix[3] = normal[1];
    pix[4] = normal[2];
}

}  // namespace
}  // namespace plm

using namespace plm;

extern "C" {

int plm_light_volume(const int32_t shape[3], const float *smoke, const float toward_light[3], float cell_m,
                     float opacity_per_density, int32_t max_steps, float *light, void *stream) {
    PLM_TRY(validate_dims(shape, "plm_light_volume"));
    PLM_REQUIRE(smoke && toward_light && light, "plm_light_volume: NULL argument");
    PLM_REQUIRE(max_steps >= 0, "plm_light_volume: max_steps must be >= 0");
    const Dims n = dims_of(shape);
    hipStream_t s = as_stream(stream);
    TimingScope t("light_volume", s);
    k_light_volume<<<grid3(n), block3(), 0, s>>>(n, smoke, toward_light[0], toward_light[1], toward_light[2], cell_m,
                                                opacity_per_density, max_steps, light);
    return check_launch("plm_light_volume");
}

int plm_blur_axis(const int32_t shape[3], const float *src, float *dst, const float *weights, int32_t radius,
                  int32_t axis, void *stream) {
    PLM_TRY(validate_dims(shape, "plm_blur_axis"));
    PLM_REQUIRE(src && dst && weights, "plm_blur_axis: NULL array");
    PLM_REQUIRE(radius >= 0 && axis >= 0 && axis < 3, "plm_blur_axis: bad radius or axis");
    const Dims n = dims_of(shape);
    k_blur_axis<<<grid3(n), block3(), 0, as_stream(stream)>>>(n, src, dst, weights, radius, axis);
    return check_launch("plm_blur_axis");
}

int plm_gradient(const int32_t shape[3], const float *src, float *gradient, void *stream) {
    PLM_TRY(validate_dims(shape, "plm_gradient"));
    PLM_REQUIRE(src && gradient, "plm_gradient: NULL array");
    const Dims n = dims_of(shape);
    k_gradient<<<grid3(n), block3(), 0, as_stream(stream)>>>(n, src, gradient);
    return check_launch("plm_gradient");
}

int plm_raymarch(const plm_grid_info *info, const float *smoke, const float *temperature, const float *gradient,
                 const int32_t gradient_shape[3], const float *light, const int32_t light_shape[3],
                 const plm_view *view, const plm_look *look, float *image, int32_t height, int32_t width,
                 int32_t max_steps, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_TRY(validate_dims(gradient_shape, "plm_raymarch gradient"));
    PLM_TRY(validate_dims(light_shape, "plm_raymarch light"));
    PLM_REQUIRE(smoke && temperature && gradient && light && view && look && image, "plm_raymarch: NULL argument");
    PLM_REQUIRE(height > 0 && width > 0 && max_steps >= 0, "plm_raymarch: bad image size or max_steps");
    PLM_REQUIRE(look->shade_layout == 0 || look->shade_layout == 1, "plm_raymarch: shade_layout must be 0 or 1");
    PLM_REQUIRE(look->shade_layout == 1 || look->normal_fade_thickness > 0.0f,
                "plm_raymarch: normal_fade_thickness must be positive");
    const Dims n = cell_dims(*info), ng = dims_of(gradient_shape), nl = dims_of(light_shape);
    const Dims used = look->shade_layout == 1 ? nl : ng;
    const bool shared = used.n0 == n.n0 && used.n1 == n.n1 && used.n2 == n.n2;
    hipStream_t s = as_stream(stream);
    TimingScope t("raymarch", s);
    const dim3 grid((width + 15) / 16, (height + 15) / 16);
    k_raymarch<<<grid, 256, 0, s>>>(*info, smoke, temperature, gradient, ng, light, nl, *view, *look, image, height,
                                    width, max_steps, shared);
    return check_launch("plm_raymarch");
}

}  // extern "C"
