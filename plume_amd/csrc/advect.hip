// Clamped MacCormack advection, sim/advect.py.
//
// One thread per output sample; every gather goes through sample_index / range_index (common.h), which
// restate sim/sample.py. The three face arrays of the velocity advance in one launch per pass
// (grid_faces), so velocity advection is 2 launches (forward SL, then backward SL fused into the
// correction) instead of the reference's 9.
#include "common.h"

namespace plm {
namespace {

struct FieldDesc {
    Dims n;
    float ox, oy, oz;
    float outside;
    int uses_outside;
};

__device__ __forceinline__ FieldDesc face_desc(const plm_grid_info &g, int axis) {
    FieldDesc f;
    f.n = Dims{g.dims[0] + (axis == 0), g.dims[1] + (axis == 1), g.dims[2] + (axis == 2)};
    f.ox = axis == 0 ? 0.0f : 0.5f;
    f.oy = axis == 1 ? 0.0f : 0.5f;
    f.oz = axis == 2 ? 0.0f : 0.5f;
    f.outside = 0.0f;
    f.uses_outside = 0;
    return f;
}

// semi_lagrangian_kernel body, sim/advect.py:29-31
__device__ __forceinline__ float sl_value(const plm_grid_info &g, const float *__restrict__ src, const FieldDesc &f,
                                          const float *__restrict__ u, const float *__restrict__ v,
                                          const float *__restrict__ w, int i, int j, int k, float dto) {
    float px = (float)i + f.ox, py = (float)j + f.oy, pz = (float)k + f.oz, sx, sy, sz;
    trace(g, u, v, w, px, py, pz, dto, sx, sy, sz);
    return sample_field(src, f.n, g, sx, sy, sz, f.ox, f.oy, f.oz, f.outside, f.uses_outside);
}

// maccormack_correct_kernel body, sim/advect.py:40-50; `back` is the backward SL value at this sample
__device__ __forceinline__ float correct_value(const plm_grid_info &g, const float *__restrict__ original,
                                               float hat_c, float back_c, const FieldDesc &f,
                                               const float *__restrict__ u, const float *__restrict__ v,
                                               const float *__restrict__ w, int i, int j, int k, float dto) {
    const int c = f.n.at(i, j, k);
    float corrected = hat_c + 0.5f * (original[c] - back_c);
    float px = (float)i + f.ox, py = (float)j + f.oy, pz = (float)k + f.oz, sx, sy, sz;
    trace(g, u, v, w, px, py, pz, dto, sx, sy, sz);
    float lo = f.outside, hi = f.outside;
    if (f.uses_outside == 0 || !outside_open_face(g, sx, sy, sz))
        range_index(original, f.n, sx - f.ox, sy - f.oy, sz - f.oz, lo, hi);
    return clamp_(corrected, lo, hi);
}

__global__ __launch_bounds__(256) void k_advect_sl(plm_grid_info g, const float *__restrict__ src,
                                                   float *__restrict__ dst, FieldDesc f, const float *__restrict__ u,
                                                   const float *__restrict__ v, const float *__restrict__ w,
                                                   float dto) {
    int i, j, k;
    if (!thread_cell(f.n, i, j, k)) return;
    dst[f.n.at(i, j, k)] = sl_value(g, src, f, u, v, w, i, j, k, dto);
}

__global__ __launch_bounds__(256) void k_maccormack_correct(plm_grid_info g, const float *__restrict__ original,
                                                            const float *__restrict__ hat,
                                                            const float *__restrict__ back, float *__restrict__ dst,
                                                            FieldDesc f, const float *__restrict__ u,
                                                            const float *__restrict__ v,
                                                            const float *__restrict__ w, float dto) {
    int i, j, k;
    if (!thread_cell(f.n, i, j, k)) return;
    const int c = f.n.at(i, j, k);
    dst[c] = correct_value(g, original, hat[c], back[c], f, u, v, w, i, j, k, dto);
}

// Backward pass fused into the correction: back = SL(hat, -dto) at this sample, computed in registers.
__global__ __launch_bounds__(256) void k_maccormack_fused(plm_grid_info g, const float *__restrict__ original,
                                                          const float *__restrict__ hat, float *__restrict__ dst,
                                                          FieldDesc f, const float *__restrict__ u,
                                                          const float *__restrict__ v, const float *__restrict__ w,
                                                          float dto) {
    int i, j, k;
    if (!thread_cell(f.n, i, j, k)) return;
    const int c = f.n.at(i, j, k);
    const float back_c = sl_value(g, hat, f, u, v, w, i, j, k, -dto);
    dst[c] = correct_value(g, original, hat[c], back_c, f, u, v, w, i, j, k, dto);
}

__global__ __launch_bounds__(256) void k_velocity_sl(plm_grid_info g, Tri<const float> src, Tri<float> dst,
                                                     float dto) {
    int axis, i, j, k;
    Dims d;
    if (!thread_face(g, axis, d, i, j, k)) return;
    const FieldDesc f = face_desc(g, axis);
    dst.p[axis][d.at(i, j, k)] = sl_value(g, src.p[axis], f, src.p[0], src.p[1], src.p[2], i, j, k, dto);
}

__global__ __launch_bounds__(256) void k_velocity_correct(plm_grid_info g, Tri<const float> orig,
                                                          Tri<const float> hat, Tri<float> dst, float dto) {
    int axis, i, j, k;
    Dims d;
    if (!thread_face(g, axis, d, i, j, k)) return;
    const FieldDesc f = face_desc(g, axis);
    const float *u = orig.p[0], *v = orig.p[1], *w = orig.p[2];
    const int c = d.at(i, j, k);
    const float back_c = sl_value(g, hat.p[axis], f, u, v, w, i, j, k, -dto);
    dst.p[axis][c] = correct_value(g, orig.p[axis], hat.p[axis][c], back_c, f, u, v, w, i, j, k, dto);
}

int field_desc(const int32_t *shape, const float *offset, float outside, int32_t uses_outside, FieldDesc &f) {
    PLM_TRY(validate_dims(shape, "advect"));
    PLM_REQUIRE(offset != nullptr, "advect: offset is NULL");
    PLM_REQUIRE(uses_outside == 0 || uses_outside == 1, "advect: uses_outside must be 0 or 1");
    f.n = dims_of(shape);
    f.ox = offset[0];
    f.oy = offset[1];
    f.oz = offset[2];
    f.outside = outside;
    f.uses_outside = uses_outside;
    return 0;
}

}  // namespace
}  // namespace plm

using namespace plm;

extern "C" {

int plm_advect_sl(const plm_grid_info *info, const float *src, float *dst, const int32_t shape[3],
                  const float offset[3], float outside, int32_t uses_outside, const float *u, const float *v,
                  const float *w, float dt_over_cell, void *stream) {
    PLM_TRY(validate_info(info));
    FieldDesc f;
    PLM_TRY(field_desc(shape, offset, outside, uses_outside, f));
    PLM_REQUIRE(src && dst && u && v && w, "plm_advect_sl: NULL array");
    TimingScope t("advect", as_stream(stream));
    k_advect_sl<<<grid3(f.n), block3(), 0, as_stream(stream)>>>(*info, src, dst, f, u, v, w, dt_over_cell);
    return check_launch("plm_advect_sl");
}

int plm_maccormack_correct(const plm_grid_info *info, const float *original, const float *hat, const float *back,
                           float *dst, const int32_t shape[3], const float offset[3], float outside,
                           int32_t uses_outside, const float *u, const float *v, const float *w,
                           float dt_over_cell, void *stream) {
    PLM_TRY(validate_info(info));
    FieldDesc f;
    PLM_TRY(field_desc(shape, offset, outside, uses_outside, f));
    PLM_REQUIRE(original && hat && back && dst && u && v && w, "plm_maccormack_correct: NULL array");
    TimingScope t("advect", as_stream(stream));
    k_maccormack_correct<<<grid3(f.n), block3(), 0, as_stream(stream)>>>(*info, original, hat, back, dst, f, u, v,
                                                                         w, dt_over_cell);
    return check_launch("plm_maccormack_correct");
}

int plm_advect_maccormack(const plm_grid_info *info, const float *data, float *hat, float *back, float *next,
                          const int32_t shape[3], const float offset[3], float outside, int32_t uses_outside,
                          const float *u, const float *v, const float *w, float dt_over_cell, int32_t fuse_back,
                          void *stream) {
    PLM_TRY(validate_info(info));
    FieldDesc f;
    PLM_TRY(field_desc(shape, offset, outside, uses_outside, f));
    PLM_REQUIRE(data && hat && next && u && v && w, "plm_advect_maccormack: NULL array");
    PLM_REQUIRE(fuse_back || back, "plm_advect_maccormack: back is NULL");
    hipStream_t s = as_stream(stream);
    TimingScope t("advect", s);
    k_advect_sl<<<grid3(f.n), block3(), 0, s>>>(*info, data, hat, f, u, v, w, dt_over_cell);
    PLM_TRY(check_launch("plm_advect_maccormack forward"));
    if (fuse_back) {
        k_maccormack_fused<<<grid3(f.n), block3(), 0, s>>>(*info, data, hat, next, f, u, v, w, dt_over_cell);
        return check_launch("plm_advect_maccormack correct");
    }
    k_advect_sl<<<grid3(f.n), block3(), 0, s>>>(*info, hat, back, f, u, v, w, -dt_over_cell);
    PLM_TRY(check_launch("plm_advect_maccormack backward"));
    k_maccormack_correct<<<grid3(f.n), block3(), 0, s>>>(*info, data, hat, back, next, f, u, v, w, dt_over_cell);
    return check_launch("plm_advect_maccormack correct");
}

int plm_advect_velocity(const plm_grid_info *info, const float *u, const float *v, const float *w, float *u_hat,
                        float *v_hat, float *w_hat, float *u_next, float *v_next, float *w_next,
                        float dt_over_cell, void *stream) {
    PLM_TRY(validate_info(info));
    PLM_REQUIRE(u && v && w && u_hat && v_hat && w_hat && u_next && v_next && w_next,
                "plm_advect_velocity: NULL array");
    hipStream_t s = as_stream(stream);
    Tri<const float> orig{{u, v, w}};
    Tri<float> hat{{u_hat, v_hat, w_hat}};
    Tri<const float> hat_in{{u_hat, v_hat, w_hat}};
    Tri<float> next{{u_next, v_next, w_next}};
    TimingScope t("advect", s);
    k_velocity_sl<<<grid_faces(*info), block3(), 0, s>>>(*info, orig, hat, dt_over_cell);
    PLM_TRY(check_launch("plm_advect_velocity forward"));
    k_velocity_correct<<<grid_faces(*info), block3(), 0, s>>>(*info, orig, hat_in, next, dt_over_cell);
    return check_launch("plm_advect_velocity correct");
}

}  // extern "C"
