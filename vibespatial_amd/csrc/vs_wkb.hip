// WKB ingest: binary geometry records (GeoParquet / GeoArrow WKB columns) ->
// owned SoA point and polygon buffers on the device, without host objects.
//
// Replaces the reference's device WKB decode used by its GeoParquet reader
// (kernels/core/wkb_decode.py:849 decode_wkb_device_pipeline, reached from
// io/wkb.py:5049 decode_wkb_owned and io/geoparquet.py:3430 read_geoparquet).
//
// Records: OGC WKB in either byte order, ISO or EWKB type words (an EWKB SRID
// is skipped); Point (1), Polygon (3) and MultiPolygon (6) in 2D.  Other types
// and Z / M variants are reported as errors, as are truncated records and
// records with bytes after their geometry.  A zero-length record or a cleared
// validity byte is a null row.
//
// Rows are processed in tiles of kWkbTile = 4096 records (256 threads x 16
// rows, row = tile * 4096 + j * 256 + thread, so the loads of offsets and
// record heads are coalesced).  No per-row arrays are kept: the rows pass
// re-derives each row's counts from its record head, which costs about what
// reading a stored count would.
//   count  per row: kind and the part / ring / coordinate counts (record head
//          and ring headers only, bounds-checked against the record's
//          bytes); per tile: the five column sums.  The first bad record and
//          its error go to the status word (atomic min of row << 8 | code).
//   tiles  one block: exclusive prefixes of the tile sums, the column totals
//          and the result words.
//   rows   per tile: counts again, block prefix over the tile + tile prefix
//          = each row's point index / polygon index / part / ring /
//          coordinate base; point rows write x, y and their record number,
//          polygonal rows their record number, geometry offset and first part
//          and ring offsets.
//   polys  one wave per polygonal record: every lane walks the ring headers,
//          coordinates are copied 64 at a time (lane k takes coordinate k).
#include "vs_internal.h"

namespace vs {

enum { kWkbOk = 0, kWkbTruncated = 1, kWkbType = 2, kWkbByteOrder = 3, kWkbMember = 4, kWkbCount = 5,
       kWkbTrailing = 6 };
enum { kKindNull = 0, kKindPoint = 1, kKindPoly = 2 };

constexpr int kWkbThreads = 256;
constexpr int kWkbRowsPerThread = 16;
constexpr int kWkbTile = kWkbThreads * kWkbRowsPerThread;
constexpr int kWkbCols = 5;  // point rows, polygonal rows, parts, rings, coordinates
constexpr int kWkbPolyThreads = 256;
constexpr int64_t kWkbPolyWaves = 32768;  // grid of the polys pass (grid-stride)

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
        uint32_t v;
        __builtin_memcpy(&v, p + pos, 4);
        pos += 4;
        return le ? v : __builtin_bswap32(v);
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

// x, y of the coordinate at byte position q (bounds already checked).
__device__ __forceinline__ void read_xy(const uint8_t* p, int64_t q, bool le, double& x, double& y) {
    uint64_t a, b;
    __builtin_memcpy(&a, p + q, 8);
    __builtin_memcpy(&b, p + q + 8, 8);
    if (!le) {
        a = __builtin_bswap64(a);
        b = __builtin_bswap64(b);
    }
    x = __longlong_as_double((long long)a);
    y = __longlong_as_double((long long)b);
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

struct WkbRow {
    int kind, err;
    int64_t parts, rings, coords;
    int64_t xy_pos;  // point rows: byte position of x
    bool le;
};

// Kind and counts of one record; an invalid record counts as a null row and
// carries its error code.
__device__ WkbRow wkb_row(const uint8_t* __restrict__ data, int64_t data_bytes, const int64_t* __restrict__ off,
                          const uint8_t* __restrict__ valid, int64_t row) {
    WkbRow R = {kKindNull, kWkbOk, 0, 0, 0, 0, true};
    const int64_t b0 = off[row], b1 = off[row + 1];
    if ((valid && !valid[row]) || b1 == b0) return R;
    if (b0 < 0 || b1 < b0 || b1 > data_bytes) {  // the record's bytes are not all inside the payload
        R.err = kWkbTruncated;
        return R;
    }
    WkbCursor c = {data, b0, b1, true, true};
    const int t = c.header();
    int err = kWkbOk;
    int64_t np = 0, nr = 0, nc = 0, xy = 0;
    int kind = kKindNull;
    if (t < 0) {
        err = -t;
    } else if (t == 1) {
        kind = kKindPoint;
        xy = c.pos;
        if (!c.need(16)) err = kWkbTruncated;
        c.pos += 16;
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
    if (err != kWkbOk) {
        R.err = err;
        return R;
    }
    R.kind = kind;
    R.parts = np;
    R.rings = nr;
    R.coords = nc;
    R.xy_pos = xy;
    R.le = c.le;
    return R;
}

__device__ __forceinline__ void row_columns(const WkbRow& R, long long (&v)[kWkbCols]) {
    v[0] = R.kind == kKindPoint;
    v[1] = R.kind == kKindPoly;
    v[2] = R.parts;
    v[3] = R.rings;
    v[4] = R.coords;
}

// In place: exclusive block prefix of kWkbCols columns over kWkbThreads
// threads; tot = the block's column sums.  lds: kWkbCols x (kWkbThreads / 64).
__device__ __forceinline__ void block_scan_cols(long long (&v)[kWkbCols], long long* lds,
                                                long long (&tot)[kWkbCols]) {
    constexpr int nw = kWkbThreads / 64;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    long long inc[kWkbCols];
#pragma unroll
    for (int k = 0; k < kWkbCols; ++k) {
        inc[k] = wave_inclusive_scan_i64(v[k]);
        if (l == 63) lds[k * nw + w] = inc[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kWkbCols; ++k) {
        long long o = 0, t = 0;
#pragma unroll
        for (int q = 0; q < nw; ++q) {
            const long long s = lds[k * nw + q];
            o += q < w ? s : 0;
            t += s;
        }
        v[k] = o + inc[k] - v[k];
        tot[k] = t;
    }
    __syncthreads();
}

__device__ __forceinline__ void report(unsigned long long* status, int64_t row, int code) {
    atomicMin(status, ((unsigned long long)row << 8) | (unsigned long long)code);
}

// The polygon buffers take int32 offsets: column totals above INT32_MAX are
// refused (nothing is filled).
__device__ __forceinline__ bool totals_overflow(const long long* totals) {
    return totals[2] > INT32_MAX || totals[3] > INT32_MAX || totals[4] > INT32_MAX;
}

struct WkbWs {
    long long* tile;    // [kWkbCols][n_tiles]: column sums per tile, then their exclusive prefixes
    long long* totals;  // [0] point rows [1] polygonal rows [2] parts [3] rings [4] coordinates
    unsigned long long* status;
    int64_t n_tiles;
};

static WkbWs carve_wkb(Carver& c, int64_t n) {
    WkbWs W;
    W.n_tiles = ceil_div(n, kWkbTile);
    W.tile = c.take<long long>(kWkbCols * W.n_tiles);
    W.totals = c.take<long long>(8);
    W.status = c.take<unsigned long long>(1);
    return W;
}

__global__ void wkb_init_kernel(unsigned long long* status, long long* totals) {
    if (threadIdx.x == 0) *status = ~0ull;
    if (threadIdx.x < 8) totals[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kWkbThreads) void wkb_count_kernel(const uint8_t* __restrict__ data,
                                                               int64_t data_bytes,
                                                               const int64_t* __restrict__ off,
                                                               const uint8_t* __restrict__ valid, int64_t n,
                                                               WkbWs W) {
    __shared__ long long red[kWkbCols * (kWkbThreads / 64)];
    const int64_t base = (int64_t)blockIdx.x * kWkbTile + threadIdx.x;
    long long sum[kWkbCols] = {0, 0, 0, 0, 0};
    int64_t bad = -1;
    int bad_code = 0;
    for (int j = 0; j < kWkbRowsPerThread; ++j) {
        const int64_t row = base + (int64_t)j * kWkbThreads;
        if (row >= n) break;
        const WkbRow R = wkb_row(data, data_bytes, off, valid, row);
        if (R.err != kWkbOk && bad < 0) {
            bad = row;
            bad_code = R.err;
        }
        long long v[kWkbCols];
        row_columns(R, v);
#pragma unroll
        for (int k = 0; k < kWkbCols; ++k) sum[k] += v[k];
    }
    if (bad >= 0) report(W.status, bad, bad_code);
    long long tot[kWkbCols];
    block_scan_cols(sum, red, tot);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kWkbCols; ++k) W.tile[k * W.n_tiles + blockIdx.x] = tot[k];
    }
}

// One block of 1024 threads: exclusive prefixes of the tile sums (all columns
// in one sweep), the column totals and the result words: result[0..4] =
// totals, [5] first bad row or -1, [6] its error code, [7] 1 when the totals
// overflow int32 offsets.
__global__ __launch_bounds__(1024) void wkb_tiles_kernel(WkbWs W, int64_t* __restrict__ result) {
    __shared__ long long lds[kWkbCols][17];  // wave sums, then [q] = carry + sum of waves < q, [16] = next carry
    __shared__ long long carry[kWkbCols];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (threadIdx.x < kWkbCols) carry[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < W.n_tiles; b0 += 1024) {
        const int64_t i = b0 + threadIdx.x;
        long long v[kWkbCols], inc[kWkbCols];
#pragma unroll
        for (int k = 0; k < kWkbCols; ++k) {
            v[k] = i < W.n_tiles ? W.tile[k * W.n_tiles + i] : 0;
            inc[k] = wave_inclusive_scan_i64(v[k]);
            if (l == 63) lds[k][w] = inc[k];
        }
        __syncthreads();
        if (threadIdx.x < kWkbCols) {
            const int k = threadIdx.x;
            long long run = carry[k];
            for (int q = 0; q < 16; ++q) {
                const long long s = lds[k][q];
                lds[k][q] = run;
                run += s;
            }
            lds[k][16] = run;
            carry[k] = run;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kWkbCols; ++k)
            if (i < W.n_tiles) W.tile[k * W.n_tiles + i] = lds[k][w] + inc[k] - v[k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < kWkbCols; ++k) {
            W.totals[k] = carry[k];
            result[k] = carry[k];
        }
        const unsigned long long st = *W.status;
        result[5] = st == ~0ull ? -1 : (int64_t)(st >> 8);
        result[6] = st == ~0ull ? 0 : (int64_t)(st & 0xFF);
        result[7] = totals_overflow(W.totals) ? 1 : 0;
    }
}

__global__ __launch_bounds__(kWkbThreads) void wkb_fill_rows_kernel(
    const uint8_t* __restrict__ data, int64_t data_bytes, const int64_t* __restrict__ off, const uint8_t* __restrict__ valid, int64_t n,
    WkbWs W, double* __restrict__ px, double* __restrict__ py, int64_t* __restrict__ point_rows,
    int32_t* __restrict__ geom_offsets, int32_t* __restrict__ part_offsets, int32_t* __restrict__ ring_offsets,
    int64_t* __restrict__ polygon_rows) {
    __shared__ long long lds[kWkbCols * (kWkbThreads / 64)];
    if (totals_overflow(W.totals)) return;  // the same in every block
    long long run[kWkbCols];
#pragma unroll
    for (int k = 0; k < kWkbCols; ++k) run[k] = W.tile[k * W.n_tiles + blockIdx.x];
    const int64_t tile0 = (int64_t)blockIdx.x * kWkbTile;
    for (int j = 0; j < kWkbRowsPerThread; ++j) {
        if (tile0 + (int64_t)j * kWkbThreads >= n) break;  // the same in every thread of the block
        const int64_t row = tile0 + (int64_t)j * kWkbThreads + threadIdx.x;
        WkbRow R = {kKindNull, kWkbOk, 0, 0, 0, 0, true};
        if (row < n) R = wkb_row(data, data_bytes, off, valid, row);
        long long v[kWkbCols], tot[kWkbCols];
        row_columns(R, v);
        block_scan_cols(v, lds, tot);
        if (R.kind == kKindPoint) {
            const long long i = run[0] + v[0];
            double x, y;
            read_xy(data, R.xy_pos, R.le, x, y);
            px[i] = x;
            py[i] = y;
            point_rows[i] = row;
        } else if (R.kind == kKindPoly) {
            const long long g = run[1] + v[1];
            const long long part = run[2] + v[2], ring = run[3] + v[3], coord = run[4] + v[4];
            polygon_rows[g] = row;
            geom_offsets[g] = (int32_t)part;
            if (R.parts > 0) part_offsets[part] = (int32_t)ring;
            if (R.rings > 0) ring_offsets[ring] = (int32_t)coord;
        }
#pragma unroll
        for (int k = 0; k < kWkbCols; ++k) run[k] += tot[k];
    }
}

// One wave per polygonal record (grid-stride over the device count of them).
// Every lane walks the same headers; the record's first ring and coordinate
// bases are read from the offsets the rows pass wrote for it.  Records were
// validated by the count pass.
__global__ __launch_bounds__(kWkbPolyThreads) void wkb_fill_polys_kernel(
    const uint8_t* __restrict__ data, const int64_t* __restrict__ off, WkbWs W, double* __restrict__ x,
    double* __restrict__ y, const int32_t* __restrict__ geom_offsets, int32_t* __restrict__ part_offsets,
    int32_t* __restrict__ ring_offsets, const int64_t* __restrict__ polygon_rows) {
    if (totals_overflow(W.totals)) return;
    const int lane = threadIdx.x & 63;
    const int64_t n_poly = W.totals[1];
    const int64_t waves = (int64_t)gridDim.x * (kWkbPolyThreads / 64);
    for (int64_t g = (int64_t)blockIdx.x * (kWkbPolyThreads / 64) + (threadIdx.x >> 6); g < n_poly; g += waves) {
        const int64_t row = polygon_rows[g];
        WkbCursor c = {data, off[row], off[row + 1], true, true};
        const int t = c.header();
        const int64_t members = t == 6 ? (int64_t)c.u32() : 1;
        int64_t part = geom_offsets[g];
        int64_t ring = -1, coord = -1;
        for (int64_t j = 0; j < members; ++j) {
            if (t == 6) c.header();
            const int64_t nr = c.u32();
            if (t == 3 && nr == 0) break;  // an empty polygon has no parts
            if (ring < 0) ring = part_offsets[part];
            if (lane == 0) part_offsets[part] = (int32_t)ring;
            ++part;
            for (int64_t r = 0; r < nr; ++r) {
                const int64_t np = c.u32();
                if (coord < 0) coord = ring_offsets[ring];
                if (lane == 0) ring_offsets[ring] = (int32_t)coord;
                ++ring;
                for (int64_t k = lane; k < np; k += 64) {
                    double xv, yv;
                    read_xy(data, c.pos + 16 * k, c.le, xv, yv);
                    x[coord + k] = xv;
                    y[coord + k] = yv;
                }
                c.pos += 16 * np;
                coord += np;
            }
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

int vs_wkb_scan(const uint8_t* payload, int64_t payload_bytes, const int64_t* offsets, const uint8_t* validity,
                int64_t n_rows, int64_t* d_result, void* workspace, int64_t workspace_bytes, void* stream) {
    vs::clear_error();
    VS_REQUIRE(payload_bytes >= 0, "vs_wkb_scan: negative payload size");
    VS_REQUIRE(n_rows >= 0 && n_rows < (int64_t)INT32_MAX, "vs_wkb_scan: n_rows out of range");
    VS_REQUIRE(workspace_bytes >= vs::wkb_workspace_bytes(n_rows), "vs_wkb_scan: workspace too small");
    VS_REQUIRE(d_result != nullptr && (n_rows == 0 || offsets != nullptr), "vs_wkb_scan: null pointer");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(workspace);
    const vs::WkbWs W = vs::carve_wkb(c, n_rows);
    hipLaunchKernelGGL(vs::wkb_init_kernel, dim3(1), dim3(64), 0, s, W.status, W.totals);
    VS_LAUNCH_CHECK();
    if (n_rows > 0) {
        hipLaunchKernelGGL(vs::wkb_count_kernel, dim3((unsigned)W.n_tiles), dim3(vs::kWkbThreads), 0, s, payload,
                           payload_bytes, offsets, validity, n_rows, W);
        VS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(vs::wkb_tiles_kernel, dim3(1), dim3(1024), 0, s, W, d_result);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

int vs_wkb_fill(const uint8_t* payload, int64_t payload_bytes, const int64_t* offsets, const uint8_t* validity,
                int64_t n_rows, const void* workspace, int64_t workspace_bytes, double* point_x, double* point_y,
                int64_t* point_rows, double* x, double* y, int32_t* geom_offsets, int32_t* part_offsets,
                int32_t* ring_offsets, int64_t* polygon_rows, void* stream) {
    vs::clear_error();
    VS_REQUIRE(payload_bytes >= 0, "vs_wkb_fill: negative payload size");
    VS_REQUIRE(n_rows >= 0 && n_rows < (int64_t)INT32_MAX, "vs_wkb_fill: n_rows out of range");
    VS_REQUIRE(workspace_bytes >= vs::wkb_workspace_bytes(n_rows), "vs_wkb_fill: workspace too small");
    VS_REQUIRE(geom_offsets && part_offsets && ring_offsets, "vs_wkb_fill: null offsets");
    hipStream_t s = vs::stream_of(stream);
    vs::Carver c(const_cast<void*>(workspace));
    const vs::WkbWs W = vs::carve_wkb(c, n_rows);
    if (n_rows > 0) {
        hipLaunchKernelGGL(vs::wkb_fill_rows_kernel, dim3((unsigned)W.n_tiles), dim3(vs::kWkbThreads), 0, s,
                           payload, payload_bytes, offsets, validity, n_rows, W, point_x, point_y, point_rows, geom_offsets,
                           part_offsets, ring_offsets, polygon_rows);
        VS_LAUNCH_CHECK();
        const int64_t waves = n_rows < vs::kWkbPolyWaves ? n_rows : vs::kWkbPolyWaves;
        hipLaunchKernelGGL(vs::wkb_fill_polys_kernel, dim3((unsigned)vs::ceil_div(waves, vs::kWkbPolyThreads / 64)),
                           dim3(vs::kWkbPolyThreads), 0, s, payload, offsets, W, x, y, geom_offsets, part_offsets,
                           ring_offsets, polygon_rows);
        VS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(vs::wkb_close_kernel, dim3(1), dim3(64), 0, s, W.totals, geom_offsets, part_offsets,
                       ring_offsets);
    VS_LAUNCH_CHECK();
    return VS_OK;
}

}  // extern "C"
