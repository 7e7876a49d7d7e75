// Library-wide C-ABI pieces: version, thread-local errors, host-side hooks.
#include "vs_distance.h"
#include "vs_exact.h"
#include "vs_internal.h"

namespace vs {

static thread_local char g_err[1024];

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

}  // namespace vs

extern "C" {

int vs_version(void) { return 100; }

int vs_last_error(char* buf, int64_t capacity) {
    const int n = (int)strlen(vs::g_err);
    if (buf && capacity > 0) {
        const int64_t k = n < capacity - 1 ? n : capacity - 1;
        memcpy(buf, vs::g_err, (size_t)k);
        buf[k] = '\0';
    }
    return n;
}

int vs_host_orient2d(double ax, double ay, double bx, double by, double cx, double cy) {
    return vs::orient2d(ax, ay, bx, by, cx, cy);
}

int vs_host_point_location(double px, double py, const double* x, const double* y,
                           const int32_t* geom_offsets, const int32_t* part_offsets,
                           const int32_t* ring_offsets, int64_t row) {
    if (!(px == px && py == py)) return 0;
    return vs::point_location_f64(px, py, x, y, geom_offsets, part_offsets, ring_offsets, row);
}

int vs_host_point_point_within(double px, double py, double qx, double qy, double d) {
    return vs::point_point_within(px, py, qx, qy, d) ? 1 : 0;
}

int vs_host_point_segment_within(double px, double py, double ax, double ay, double bx, double by, double d) {
    return vs::point_segment_within(px, py, ax, ay, bx, by, d) ? 1 : 0;
}

int vs_host_segments_touch(double ax, double ay, double bx, double by, double cx, double cy,
                           double dx, double dy) {
    return vs::segments_touch(ax, ay, bx, by, cx, cy, dx, dy) ? 1 : 0;
}

}  // extern "C"
