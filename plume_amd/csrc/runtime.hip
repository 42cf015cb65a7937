// Error state, argument validation and the benchmark's per-kernel-family event timing.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace plm {

static thread_local char g_last_error[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list args;
    va_start(args, fmt);
    vsnprintf(g_last_error, sizeof g_last_error, fmt, args);
    va_end(args);
    return code;
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail((int)e, "%s: launch failed: %s", what, hipGetErrorString(e));
    return 0;
}

int validate_dims(const int32_t *shape, const char *what) {
    if (shape == nullptr) return fail(PLM_EINVAL, "%s: shape is NULL", what);
    int64_t count = 1;
    for (int a = 0; a < 3; a++) {
        if (shape[a] <= 0) return fail(PLM_EINVAL, "%s: shape[%d] = %d must be positive", what, a, shape[a]);
        count *= shape[a] + 1;
    }
    if (3 * count >= (int64_t(1) << 31)) return fail(PLM_EINVAL, "%s: grid too large for 32-bit indexing", what);
    if (int64_t(shape[0] + 1) * (shape[1] + 1) >= (int64_t(1) << 24))
        return fail(PLM_EINVAL, "%s: nx * ny = %d * %d too large for 24-bit row indexing", what, shape[0], shape[1]);
    if (shape[0] + 1 > 65535 / 3) return fail(PLM_EINVAL, "%s: nx = %d exceeds the launch limit", what, shape[0]);
    return 0;
}

int validate_info(const plm_grid_info *info) {
    if (info == nullptr) return fail(PLM_EINVAL, "grid info is NULL");
    PLM_TRY(validate_dims(info->dims, "grid info"));
    if (!(info->cell_m > 0.0f)) return fail(PLM_EINVAL, "grid info: cell_m must be positive");
    for (int a = 0; a < 3; a++)
        if ((info->open_lo[a] != 0 && info->open_lo[a] != 1) || (info->open_hi[a] != 0 && info->open_hi[a] != 1))
            return fail(PLM_EINVAL, "grid info: open flags must be 0 or 1");
    return 0;
}

// ------------------------------------------------------------------ timing

namespace {
struct Timing {
    std::mutex mutex;
    std::string family;
    bool enabled = false;
    std::vector<hipEvent_t> starts, stops;
    size_t used = 0;
};
Timing &timing() {
    static Timing t;
    return t;
}
}  // namespace

TimingScope::TimingScope(const char *family, hipStream_t s, bool bind)
    : active(false), bound(bind), stream(s), slot(-1) {
    Timing &t = timing();
    std::lock_guard<std::mutex> lock(t.mutex);
    if (!t.enabled || t.family != family) return;
    if (t.used == t.starts.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return;
        }
        t.starts.push_back(a);
        t.stops.push_back(b);
    }
    slot = (int)t.used++;
    active = bound || hipEventRecord(t.starts[slot], stream) == hipSuccess;
}

hipEvent_t TimingScope::start() const {
    if (!active || !bound) return nullptr;
    Timing &t = timing();
    std::lock_guard<std::mutex> lock(t.mutex);
    return t.starts[slot];
}

hipEvent_t TimingScope::stop() const {
    if (!active || !bound) return nullptr;
    Timing &t = timing();
    std::lock_guard<std::mutex> lock(t.mutex);
    return t.stops[slot];
}

TimingScope::~TimingScope() {
    if (!active || bound) return;
    Timing &t = timing();
    std::lock_guard<std::mutex> lock(t.mutex);
    (void)hipEventRecord(t.stops[slot], stream);
}

}  // namespace plm

extern "C" {

const char *plm_last_error(void) { return plm::g_last_error; }

int plm_abi_version(void) { return 2; }

int plm_timing_enable(const char *family) {
    PLM_REQUIRE(family != nullptr, "timing family is NULL");
    plm::Timing &t = plm::timing();
    std::lock_guard<std::mutex> lock(t.mutex);
    t.family = family;
    t.enabled = true;
    t.used = 0;
    return 0;
}

int plm_timing_disable(void) {
    plm::Timing &t = plm::timing();
    std::lock_guard<std::mutex> lock(t.mutex);
    t.enabled = false;
    return 0;
}

int plm_timing_read(double *total_ms, int64_t *launches) {
    PLM_REQUIRE(total_ms != nullptr && launches != nullptr, "timing outputs are NULL");
    plm::Timing &t = plm::timing();
    std::lock_guard<std::mutex> lock(t.mutex);
    double total = 0.0;
    for (size_t s = 0; s < t.used; s++) {
        hipError_t e = hipEventSynchronize(t.stops[s]);
        if (e != hipSuccess) return plm::fail((int)e, "timing: %s", hipGetErrorString(e));
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, t.starts[s], t.stops[s]);
        if (e != hipSuccess) return plm::fail((int)e, "timing: %s", hipGetErrorString(e));
        total += ms;
    }
    *total_ms = total;
    *launches = (int64_t)t.used;
    return 0;
}

}  // extern "C"
