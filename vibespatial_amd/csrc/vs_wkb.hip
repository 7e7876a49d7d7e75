// WKB ingest: binary geometry records (GeoParquet / GeoArrow WKB columns) ->
// owned SoA point and polygon buffers on the device, without host objects.
//
// Replaces the reference's device WKB decode used by its GeoParquet reader
// (kernels/core/wkb_decode.py:849 decode_wkb_device_pipeline, reached from
// io/wkb.py:5049 decode_wkb_owned and io/geoparquet.py:3430 read_geoparquet).
//
// Records: OGC WKB in either byte order, ISO or EWKB type words (an EWKB SRID
// is skipped); Point (1), Polygon (3) and MultiPolygon (6) in 2D.  Other types
// and Z / M variants are reported as errors, as are truncated records.  A
// zero-length record or a cleared validity byte is a null row.
//
//   scan  one thread per record: kind (0 null, 1 point, 2 polygonal) and the
//         part / ring / coordinate counts, bounds-checked against the record's
//         bytes (a record with bytes after its geometry is an error too); the first bad record and its error go to the status word
//         (atomic min of row << 8 | code).
//   sizes exclusive scans of the per-family counts (family row index, part,
//         ring and coordinate bases), totals on the device.
//   fill  one thread per record re-reads it and writes its coordinates and
//         offsets at those bases; a last kernel writes the closing offsets.
#include "vs_internal.h"

namespace vs {

int exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, long long* d_total, void* ws, hipStream_t s);
int64_t scan_workspace_bytes(int64_t n);

enum { kWkbOk = 0, kWkbTruncated = 1, kWkbType = 2, kWkbByteOrder = 3, kWkbMember = 4, kWkbCount = 5, kWkbTrailing = 6 };
enum { kKindNull = 0, kKindPoint = 1, kKindPoly = 2 };

struct WkbCursor {
    const uint8_t* p;
    int64_t pos, end;
    bool le, ok;

    __device__ __forceinline__ bool need(int64_t k) {
        if (pos + k > end) ok = false;
        return ok;
    }
    __device__ __forceinline__ uint32_t u32() {
        if (!need(4)) return 0;
        const uint8_t* b = p + pos;
        pos += 4;
        return le ? ((uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24))
                  : ((uint32_t)b[3] | ((uint32_t)b[2] << 8) | ((uint32_t)b[1] << 16) | ((uint32_t)b[0] << 24));
    }
    __device__ __forceinline__ double f64() {
        if (!need(8)) return 0.0;
        const uint8_t* b = p + pos;
        pos += 8;
        uint64_t v = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) v |= (uint64_t)b[le ? i : 7 - i] << (8 * i);
        return __longlong_as_double((long long)v);
    }
    // byte order + type word (+ EWKB SRID); returns the 2D base type or -code
    __device__ int header() {
        if (!need(1)) return -kWkbTruncated;
        const uint8_t bo = p[pos++];
        if (bo > 1) return -kWkbByteOrder;
        le = bo == 1;
        const uint32_t t = u32();
        if (!ok) return -kWkbTruncated;
        if (t & 0xC0000000u) return -kWkbType;  // EWKB Z / M
        if (t & 0x20000000u) {                  // EWKB SRID
            u32();
            if (!ok) return -kWkbTruncated;
        }
        const uint32_t base = t & 0x1FFFFFFFu;
        if (base >= 1000u) return -kWkbType;  // ISO Z / M / ZM
        return (int)base;
    }
};

__device__ __forceinline__ bool row_bytes(const int64_t* off, const uint8_t* valid, int64_t row, int64_t& b0,
                                          int64_t& b1) {
    b0 = off[row];
    b1 = off[row + 1];
    return (!valid || valid[row]) && b1 > b0;
}

__device__ __forceinline__ void report(unsigned long long* status, int64_t row, int code) {
    atomicMin(status, ((unsigned long long)row << 8) | (unsigned long long)code);
}

// One polygon body (after its header), skipping the coordinates: its ring
// count, with its coordinate count added to coords, or -1 when it overruns
// the record.
__device__ int64_t skip_polygon_body(WkbCursor& c, int64_t& coords) {
    const int64_t nr = c.u32();
    if (!c.ok) return -1;
    int64_t nc = 0;
    for (int64_t r = 0; r < nr; ++r) {
        const int64_t np = c.u32();
        if (!c.ok || !c.need(16 * np)) return -1;
        c.pos += 16 * np;
        nc += np;
    }
    coords += nc;
    return nr;
}

__global__ __launch_bounds__(256) void wkb_scan_kernel(const uint8_t* __restrict__ data,
                                                      const int64_t* __restrict__ off,
                                                      const uint8_t* __restrict__ valid, int64_t n,
                                                      int32_t* __restrict__ is_point, int32_t* __restrict__ is_poly,
                                                      int32_t* __restrict__ parts, int32_t* __restrict__ rings,
                                                      int32_t* __restrict__ coords,
                                                      unsigned long long* __restrict__ status) {
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (row >= n) return;
    int kind = kKindNull;
    int64_t np = 0, nr = 0, nc = 0;
    int64_t b0, b1;
    int err = kWkbOk;
    if (row_bytes(off, valid, row, b0, b1)) {
        WkbCursor c = {data, b0, b1, true, true};
        const int t = c.header();
        if (t < 0) {
            err = -t;
        } else if (t == 1) {
            c.f64();
            c.f64();
            if (!c.ok) err = kWkbTruncated;
            kind = kKindPoint;
        } else if (t == 3) {
            kind = kKindPoly;
            const int64_t k = skip_polygon_body(c, nc);
            if (k < 0) err = kWkbTruncated;
            nr = k;
            np = k > 0 ? 1 : 0;  // an empty polygon has no parts
        } else if (t == 6) {
            kind = kKindPoly;
            const int64_t k = c.u32();
            if (!c.ok) err = kWkbTruncated;
            for (int64_t j = 0; j < k && err == kWkbOk; ++j) {
                const int tj = c.header();
                if (tj < 0) {
                    err = -tj;
                } else if (tj != 3) {
                    err = kWkbMember;
                } else {
                    const int64_t kr = skip_polygon_body(c, nc);
                    if (kr < 0) err = kWkbTruncated;
                    else nr += kr;
                }
            }
            np = k;
        } else {
            err = kWkbType;
        }
        if (err == kWkbOk && c.pos != c.end) err = kWkbTrailing;
        if (err == kWkbOk && (np > INT32_MAX || nr > INT32_MAX || nc > INT32_MAX)) err = kWkbCount;
    }
    if (err != kWkbOk) {
        report(status, row, err);
        kind = kKindNull;
        np = nr = nc = 0;
    }
    is_point[row] = kind == kKindPoint;
    is_poly[row] = kind == kKindPoly;
    parts[row] = (int32_t)np;
    rings[row] = (int32_t)nr;
    coords[row] = kind == kKindPoly ? (int32_t)nc : 0;
}

struct WkbWs {
    int32_t* is_point;
    int32_t* is_poly;
    int32_t* parts;
    int32_t* rings;
    int32_t* coords;
    int32_t* pidx;
    int32_t* gidx;
    int32_t* part_base;
    int32_t* ring_base;
    int32_t* coord_base;
    long long* totals;  // [0] points [1] polygon rows [2] parts [3] rings [4] coordinates
    unsigned long long* status;
    void* scan_ws;
};

static WkbWs carve_wkb(Carver& c, int64_t n) {
    WkbWs W;
    W.is_point = c.take<int32_t>(n);
    W.is_poly = c.take<int32_t>(n);
    W.parts = c.take<int32_t>(n);
    W.rings = c.take<int32_t>(n);
    W.coords = c.take<int32_t>(n);
    W.pidx = c.take<int32_t>(n);
    W.gidx = c.take<int32_t>(n);
    W.part_base = c.take<int32_t>(n);
    W.ring_base = c.take<int32_t>(n);
    W.coord_base = c.take<int32_t>(n);
    W.totals = c.take<long long>(8);
    W.status = c.take<unsigned long long>(1);
    W.scan_ws = c.take<char>(scan_workspace_bytes(n));
    return W;
}

__global__ void wkb_init_kernel(unsigned long long* status, long long* totals) {
    if (threadIdx.x == 0) *status = ~0ull;
    if (threadIdx.x < 8) totals[threadIdx.x] = 0;
}

// The polygon buffers take int32 offsets: column totals above INT32_MAX are
// refused (nothing is filled).
__device__ __forceinline__ bool totals_overflow(const long long* totals) {
    return totals[2] > INT32_MAX || totals[3] > INT32_MAX || totals[4] > INT32_MAX;
}

// result[0..4] = totals, [5] first bad row or -1, [6] its error code,
// [7] 1 when the totals overflow int32 offsets.
__global__ void wkb_publish_kernel(const long long* totals, const unsigned long long* status, int64_t* result) {
    if (threadIdx.x == 0) {
        for (int k = 0; k < 5; ++k) result[k] = totals[k];
        const unsigned long long st = *status;
        result[5] = st == ~0ull ? -1 : (int64_t)(st >> 8);
        result[6] = st == ~0ull ? 0 : (int64_t)(st & 0xFF);
        result[7] = totals_overflow(totals) ? 1 : 0;
    }
}

// Polygon body at the cursor: ring offsets and coordinates at the bases.
__device__ void fill_polygon_body(WkbCursor& c, int32_t& ring, int32_t& coord, int32_t* ring_offsets, double* x,
                                  double* y) {
    const uint32_t nr = c.u32();
    for (uint32_t r = 0; r < nr; ++r) {
        const uint32_t np = c.u32();
        ring_offsets[ring++] = coord;
        for (uint32_t k = 0; k < np; ++k) {
            x[coord] = c.f64();
            y[coord] = c.f64();
            ++coord;
        }
    }
}

__global__ __launch_bounds__(256) void wkb_fill_kernel(const uint8_t* __restrict__ data,
                                                      const int64_t* __restrict__ off,
                                                      const uint8_t* __restrict__ valid, int64_t n, WkbWs W,
                                                      double* __restrict__ px, double* __restrict__ py,
                                                      int64_t* __restrict__ point_rows, double* __restrict__ x,
                                                      double* __restrict__ y, int32_t* __restrict__ geom_offsets,
                                                      int32_t* __restrict__ part_offsets,
                                                      int32_t* __restrict__ ring_offsets,
                                                      int64_t* __restrict__ polygon_rows) {
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (row >= n || totals_overflow(W.totals)) return;
    int64_t b0, b1;
    if (W.is_point[row]) {
        row_bytes(off, valid, row, b0, b1);
        WkbCursor c = {data, b0, b1, true, true};
        c.header();
        const int32_t i = W.pidx[row];
        px[i] = c.f64();
        py[i] = c.f64();
        point_rows[i] = row;
        return;
    }
    if (!W.is_poly[row]) return;
    row_bytes(off, valid, row, b0, b1);
    WkbCursor c = {data, b0, b1, true, true};
    const int t = c.header();
    const int32_t g = W.gidx[row];
    int32_t part = W.part_base[row], ring = W.ring_base[row], coord = W.coord_base[row];
    geom_offsets[g] = part;
    polygon_rows[g] = row;
    if (t == 3) {
        if (W.parts[row] > 0) part_offsets[part++] = ring;
        fill_polygon_body(c, ring, coord, ring_offsets, x, y);
    } else {
        const uint32_t k = c.u32();
        for (uint32_t j = 0; j < k; ++j) {
            c.header();
            part_offsets[part++] = ring;
            fill_polygon_body(c, ring, coord, ring_offsets, x, y);
        }
    }
}

__global__ void wkb_close_kernel(const long long* totals, int32_t* geom_offsets, int32_t* part_offsets,
                                 int32_t* ring_offsets) {
    if (threadIdx.x == 0 && !totals_overflow(totals)) {
        geom_offsets[totals[1]] = (int32_t)totals[2];
        part_offsets[totals[2]] = (int32_t)totals[3];
        ring_offsets[totals[3]] = (int32_t)totals[4];
    }
}

int64_t wkb_workspace_bytes(int64_t n) {
    Carver c(nullptr);
    carve_wkb(c, n);
    return c.bytes;
}

}  // namespace vs

extern "C" {

int64_t vs_wkb_workspace_bytes(int64_t n_rows) { return vs::wkb_workspace_bytes(n_rows); }

int vs_wkb_scan(const uint8_t* payload, const int64_t* offsets, const uint8_t* validity, int64_t n_rows,
                int64_t* d_result, void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_rows >= 0 && n_rows < (int64_t)INT32_MAX, "vs_wkb_scan: n_rows out of range");
    VS_REQUIRE(workspace_bytes >= vs::wkb_workspace_bytes(n_rows), "vs_wkb_scan: workspace too small");
    VS_REQUIRE(d_result != nullptr && (n_rows == 0 || offsets != nullptr), "vs_wkb_scan: null pointer");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    const vs::WkbWs W = vs::carve_wkb(c, n_rows);
    hipLaunchKernelGGL(vs::wkb_init_kernel, dim3(1), dim3(64), 0, s, W.status, W.totals);
    VS_LAUNCH_CHECK();
    if (n_rows > 0) {
        hipLaunchKernelGGL(vs::wkb_scan_kernel, dim3((unsigned)vs::ceil_div(n_rows, 256)), dim3(256), 0, s, payload,
                           offsets, validity, n_rows, W.is_point, W.is_poly, W.parts, W.rings, W.coords, W.status);
        VS_LAUNCH_CHECK();
        VS_TRY(vs::exclusive_scan_i32(W.is_point, W.pidx, n_rows, W.totals + 0, W.scan_ws, s));
        VS_TRY(vs::exclusive_scan_i32(W.is_poly, W.gidx, n_rows, W.totals + 1, W.scan_ws, s));
        VS_TRY(vs::exclusive_scan_i32(W.parts, W.part_base, n_rows, W.totals + 2, W.scan_ws, s));
        VS_TRY(vs::exclusive_scan_i32(W.rings, W.ring_base, n_rows, W.totals + 3, W.scan_ws, s));
        VS_TRY(vs::exclusive_scan_i32(W.coords, W.coord_base, n_rows, W.totals + 4, W.scan_ws, s));
    }
    hipLaunchKernelGGL(vs::wkb_publish_kernel, dim3(1), dim3(64), 0, s, W.totals, W.status, d_result);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int vs_wkb_fill(const uint8_t* payload, const int64_t* offsets, const uint8_t* validity, int64_t n_rows,
                const void* workspace, int64_t workspace_bytes, double* point_x, double* point_y,
                int64_t* point_rows, double* x, double* y, int32_t* geom_offsets, int32_t* part_offsets,
                int32_t* ring_offsets, int64_t* polygon_rows, void* stream) {
    vs::clear_error();
    VS_REQUIRE(n_rows >= 0 && n_rows < (int64_t)INT32_MAX, "vs_wkb_fill: n_rows out of range");
    VS_REQUIRE(workspace_bytes >= vs::wkb_workspace_bytes(n_rows), "vs_wkb_fill: workspace too small");
    VS_REQUIRE(geom_offsets && part_offsets && ring_offsets, "vs_wkb_fill: null offsets");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(const_cast<void*>(workspace));
    const vs::WkbWs W = vs::carve_wkb(c, n_rows);
    if (n_rows > 0) {
        hipLaunchKernelGGL(vs::wkb_fill_kernel, dim3((unsigned)vs::ceil_div(n_rows, 256)), dim3(256), 0, s, payload,
                           offsets, validity, n_rows, W, point_x, point_y, point_rows, x, y, geom_offsets,
                           part_offsets, ring_offsets, polygon_rows);
        VS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(vs::wkb_close_kernel, dim3(1), dim3(64), 0, s, W.totals, geom_offsets, part_offsets,
                       ring_offsets);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

}  // extern "C"
