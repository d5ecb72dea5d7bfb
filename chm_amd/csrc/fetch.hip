// Upwind fetch (CHM fetchr, src/modules/fetchr.cpp:28-120) on the GPU.
//
// One thread per face walks the reference's steps along the upwind azimuth;
// each step needs the face whose centroid is nearest to a point
// (triangulation.cpp:170-178: a CGAL k = 1 kd-tree query).  Here centroids
// are bucketed once into a uniform grid (about two per cell, faces sorted by
// cell and by index inside a cell), and a query scans square rings of cells
// around the point's cell, on each side only the cells that can still hold
// something nearer than the best so far, and stops when the next ring lies
// farther than the best distance.  The result is the exact nearest centroid
// under the squared distance dx*dx + dy*dy (ties to the lower face index),
// identical to a brute-force scan.  Queries outside the grid (upwind of the
// domain edge) start at the first ring that reaches the grid.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pbsm3d_fetch.h"
#include "kernels.hpp"

namespace pb {
namespace {

struct CentroidGrid {
    double xmin, ymin, h;
    int64_t nx, ny;
    const int32_t* start;  // [nx*ny + 1] first slot of each cell (cell = cy * nx + cx)
    const int32_t* face;   // [F] face ids by cell, ascending inside a cell
    const double *cx, *cy;
};

__device__ __forceinline__ void scan_cell(const CentroidGrid& g, int64_t c, double qx, double qy, double& best,
                                          int32_t& bi) {
    for (int32_t s = g.start[c]; s < g.start[c + 1]; ++s) {
        const int32_t f = g.face[s];
        const double dx = qx - g.cx[f];
        const double dy = qy - g.cy[f];
        const double d2 = dx * dx + dy * dy;
        if (d2 < best || (d2 == best && f < bi)) {
            best = d2;
            bi = f;
        }
    }
}

// cells [lo, hi] along x in row cy (or along y in column cx) that may hold a
// centroid within sqrt(best) of q, given the perpendicular gap `gap`
__device__ __forceinline__ void span(double best, double gap, double q, double origin, double h, int64_t lo,
                                     int64_t hi, int64_t& a, int64_t& b) {
    a = lo;
    b = hi;
    if (best < INFINITY) {
        const double w2 = best - gap * gap;
        if (w2 < 0) {
            a = 1;
            b = 0;
            return;
        }
        const double w = sqrt(w2) * (1 + 1e-12) + 1e-9;
        a = max(lo, (int64_t)floor((q - w - origin) / h));
        b = min(hi, (int64_t)floor((q + w - origin) / h));
    }
}

__device__ int32_t nearest_face(const CentroidGrid& g, double qx, double qy) {
    if (!isfinite(qx) || !isfinite(qy)) return -1;
    const double fx = floor((qx - g.xmin) / g.h), fy = floor((qy - g.ymin) / g.h);
    const int64_t lim = (int64_t)1 << 40;
    const int64_t ix = (int64_t)fmin(fmax(fx, -(double)lim), (double)lim);
    const int64_t iy = (int64_t)fmin(fmax(fy, -(double)lim), (double)lim);
    const int64_t r0 = max(max((int64_t)0, max(-ix, ix - (g.nx - 1))), max(-iy, iy - (g.ny - 1)));
    const int64_t rmax = max(max(ix, g.nx - 1 - ix), max(iy, g.ny - 1 - iy));
    double best = INFINITY;
    int32_t bi = -1;
    for (int64_t r = r0; r <= rmax; ++r) {
        if (r == 0) {
            scan_cell(g, iy * g.nx + ix, qx, qy, best, bi);
        } else {
            int64_t a, b;
            // rows iy - r and iy + r (corners included)
            for (int side = 0; side < 2; ++side) {
                const int64_t cyr = side == 0 ? iy - r : iy + r;
                if (cyr < 0 || cyr >= g.ny) continue;
                const double y0 = g.ymin + cyr * g.h;
                const double gap = side == 0 ? fmax(0.0, qy - (y0 + g.h)) : fmax(0.0, y0 - qy);
                span(best, gap, qx, g.xmin, g.h, max((int64_t)0, ix - r), min(g.nx - 1, ix + r), a, b);
                for (int64_t c = a; c <= b; ++c) scan_cell(g, cyr * g.nx + c, qx, qy, best, bi);
            }
            // columns ix - r and ix + r (corners excluded)
            for (int side = 0; side < 2; ++side) {
                const int64_t cxr = side == 0 ? ix - r : ix + r;
                if (cxr < 0 || cxr >= g.nx) continue;
                const double x0 = g.xmin + cxr * g.h;
                const double gap = side == 0 ? fmax(0.0, qx - (x0 + g.h)) : fmax(0.0, x0 - qx);
                span(best, gap, qy, g.ymin, g.h, max((int64_t)0, iy - r + 1), min(g.ny - 1, iy + r - 1), a, b);
                for (int64_t c = a; c <= b; ++c) scan_cell(g, c * g.nx + cxr, qx, qy, best, bi);
            }
        }
        // every cell of ring r + 1 or beyond lies at least r * h from q (q is inside
        // cell (ix, iy)); the margin covers the rounding of the cell index
        if (bi >= 0) {
            const double lb = r * g.h - 1e-6 * (1.0 + g.h);
            if (lb > 0 && best <= lb * lb) break;
        }
    }
    return bi;
}

struct FetchArgs {
    int64_t F;
    int steps;
    int incl_veg;
    double max_distance, I;
    const double *cz, *canopy;  // canopy may be null (no vegetation)
    const double* vw_dir;
    double* fetch;
};

// fetchr::run (fetchr.cpp:55-118), expression order kept
__global__ void __launch_bounds__(kBlock) k_fetch(CentroidGrid g, FetchArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.F) return;
    const double size_of_step = a.max_distance / a.steps;
    const bool has_veg = a.canopy != nullptr;
    double out = a.max_distance;
    const double wind_dir = a.vw_dir[i];
    if (a.incl_veg && has_veg && a.canopy[i] > 1) {
        a.fetch[i] = 0;
        return;
    }
    if (!isfinite(wind_dir)) {  // no upwind point to search from (the -9999 default is finite)
        a.fetch[i] = out;
        return;
    }
    // point_from_bearing_UTM (coordinates.cpp:60-71)
    const double DegreesToRadians = M_PI / 180.0;
    const double bearing = wind_dir * DegreesToRadians;
    const double sb = sin(bearing), cb = cos(bearing);
    for (int j = 1; j <= a.steps; ++j) {
        const double distance = j * size_of_step;
        const double dist_x = distance * sb;
        const double dist_y = distance * cb;
        const int32_t f = nearest_face(g, g.cx[i] + dist_x, g.cy[i] + dist_y);
        double Z_CanTop = 0;
        if (a.incl_veg && has_veg) Z_CanTop = a.canopy[f];
        const double Z_test = a.cz[f] + Z_CanTop;
        const double Z_core = a.cz[i] + distance * a.I;
        const double z0_1 = 0.12 * Z_CanTop;
        const double z0_2 = 0.001;
        const double n = 1.0 / 0.8;
        const double h = 5;
        const double x_sss = pow(((33.33333333 * h - 25. * z0_2) / (log(z0_1 / z0_2) * z0_2)), n) * z0_2;
        if (Z_test >= Z_core || (a.incl_veg && distance < x_sss)) {
            out = distance;
            break;
        }
    }
    a.fetch[i] = out;
}

__global__ void __launch_bounds__(kBlock) k_nearest(CentroidGrid g, int64_t n, const double* __restrict__ qx,
                                                    const double* __restrict__ qy, int32_t* __restrict__ idx) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) idx[k] = nearest_face(g, qx[k], qy[k]);
}

}  // namespace
}  // namespace pb

using namespace pb;

struct pbsm3d_fetch_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t F = 0;
    pbsm3d_fetch_config cfg{};
    DevBuf<double> cx, cy, cz, canopy, stage_in, stage_out;
    DevBuf<int32_t> start, face;
    double xmin = 0, ymin = 0, h = 1;
    int64_t nx = 1, ny = 1;
    bool has_canopy = false;
    std::string err;
    ~pbsm3d_fetch_ctx() {
        if (stream) (void)hipStreamDestroy(stream);
    }
    CentroidGrid grid() const { return CentroidGrid{xmin, ymin, h, nx, ny, start.p, face.p, cx.p, cy.p}; }
};

namespace {
std::string g_fetch_err;  // last creation error (no context to hold it)

template <typename Fn>
int fetch_guard(pbsm3d_fetch_ctx* c, Fn&& fn) {
    try {
        fn();
        return PBSM3D_OK;
    } catch (const Error& e) {
        if (c) c->err = e.what();
        g_fetch_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        g_fetch_err = e.what();
        return PBSM3D_EINVAL;
    }
}
}  // namespace

extern "C" {

void pbsm3d_fetch_config_default(pbsm3d_fetch_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->steps = 10;
    c->incl_veg = 1;
    c->max_distance = 1000.0;
    c->I = 0.06;
}

int pbsm3d_fetch_create(pbsm3d_fetch_ctx** out, int device, int64_t F, const double* x, const double* y,
                        const double* z, const double* canopy, const pbsm3d_fetch_config* cfg) {
    if (!out || F <= 0 || F > INT32_MAX || !x || !y || !z || !cfg) {
        g_fetch_err = "invalid argument";
        return PBSM3D_EINVAL;
    }
    if (cfg->steps < 1 || !(cfg->max_distance > 0)) {
        g_fetch_err = "steps must be >= 1 and max_distance > 0";
        return PBSM3D_EINVAL;
    }
    *out = nullptr;
    auto* c = new pbsm3d_fetch_ctx();
    const int rc = fetch_guard(c, [&] {
        c->device = device;
        c->F = F;
        c->cfg = *cfg;
        PB_HIP(hipSetDevice(device));
        PB_HIP(hipStreamCreate(&c->stream));
        hipStream_t s = c->stream;
        // uniform grid of about two centroids per cell
        double xmin = x[0], xmax = x[0], ymin = y[0], ymax = y[0];
        for (int64_t i = 0; i < F; ++i) {
            if (!std::isfinite(x[i]) || !std::isfinite(y[i])) throw Error(PBSM3D_EINVAL, "non-finite centroid");
            xmin = std::min(xmin, x[i]);
            xmax = std::max(xmax, x[i]);
            ymin = std::min(ymin, y[i]);
            ymax = std::max(ymax, y[i]);
        }
        const double ext = std::max(xmax - xmin, ymax - ymin);
        double h = std::sqrt(std::max((xmax - xmin) * (ymax - ymin), 1e-300) * 2.0 / (double)F);
        if (!(h > 0) || h < ext * 1e-6) h = std::max(ext / std::sqrt((double)F), 1e-3);
        const int64_t nx = (int64_t)std::floor((xmax - xmin) / h) + 1;
        const int64_t ny = (int64_t)std::floor((ymax - ymin) / h) + 1;
        if (nx * ny > 8 * F + 64) throw Error(PBSM3D_EINVAL, "degenerate centroid extent");
        std::vector<int32_t> cell(F), start(nx * ny + 1, 0), face(F);
        for (int64_t i = 0; i < F; ++i) {
            const int64_t ix = std::min(nx - 1, (int64_t)std::floor((x[i] - xmin) / h));
            const int64_t iy = std::min(ny - 1, (int64_t)std::floor((y[i] - ymin) / h));
            cell[i] = (int32_t)(iy * nx + ix);
            start[cell[i] + 1]++;
        }
        for (int64_t k = 0; k < nx * ny; ++k) start[k + 1] += start[k];
        std::vector<int32_t> fill(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < F; ++i) face[fill[cell[i]]++] = (int32_t)i;
        c->xmin = xmin;
        c->ymin = ymin;
        c->h = h;
        c->nx = nx;
        c->ny = ny;
        auto up = [&](DevBuf<double>& d, const double* hsrc) {
            d.alloc(F);
            PB_HIP(hipMemcpyAsync(d.p, hsrc, F * sizeof(double), hipMemcpyHostToDevice, s));
        };
        up(c->cx, x);
        up(c->cy, y);
        up(c->cz, z);
        c->has_canopy = canopy != nullptr;
        if (canopy) up(c->canopy, canopy);
        c->start.alloc(nx * ny + 1);
        c->face.alloc(F);
        PB_HIP(hipMemcpyAsync(c->start.p, start.data(), start.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
        PB_HIP(hipMemcpyAsync(c->face.p, face.data(), F * sizeof(int32_t), hipMemcpyHostToDevice, s));
        c->stage_in.alloc(F);
        c->stage_out.alloc(F);
        PB_HIP(hipStreamSynchronize(s));
    });
    if (rc != PBSM3D_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return PBSM3D_OK;
}

int pbsm3d_fetch_run(pbsm3d_fetch_ctx* c, const double* vw_dir, double* fetch, int on_device) {
    if (!c || !vw_dir || !fetch) return PBSM3D_EINVAL;
    return fetch_guard(c, [&] {
        PB_HIP(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        const double* din = vw_dir;
        double* dout = fetch;
        if (!on_device) {
            PB_HIP(hipMemcpyAsync(c->stage_in.p, vw_dir, c->F * sizeof(double), hipMemcpyHostToDevice, s));
            din = c->stage_in.p;
            dout = c->stage_out.p;
        }
        FetchArgs a{c->F, c->cfg.steps, c->cfg.incl_veg, c->cfg.max_distance, c->cfg.I, c->cz.p,
                    c->has_canopy ? c->canopy.p : nullptr, din, dout};
        hipLaunchKernelGGL(k_fetch, dim3(grid_for(c->F)), dim3(kBlock), 0, s, c->grid(), a);
        PB_HIP(hipGetLastError());
        if (!on_device) PB_HIP(hipMemcpyAsync(fetch, dout, c->F * sizeof(double), hipMemcpyDeviceToHost, s));
        PB_HIP(hipStreamSynchronize(s));
    });
}

int pbsm3d_fetch_nearest(pbsm3d_fetch_ctx* c, int64_t n, const double* qx, const double* qy, int32_t* idx) {
    if (!c || n < 0 || (n > 0 && (!qx || !qy || !idx))) return PBSM3D_EINVAL;
    if (n == 0) return PBSM3D_OK;
    return fetch_guard(c, [&] {
        PB_HIP(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        DevBuf<double> dqx, dqy;
        DevBuf<int32_t> di;
        dqx.alloc(n);
        dqy.alloc(n);
        di.alloc(n);
        PB_HIP(hipMemcpyAsync(dqx.p, qx, n * sizeof(double), hipMemcpyHostToDevice, s));
        PB_HIP(hipMemcpyAsync(dqy.p, qy, n * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_nearest, dim3(grid_for(n)), dim3(kBlock), 0, s, c->grid(), n, dqx.p, dqy.p, di.p);
        PB_HIP(hipGetLastError());
        PB_HIP(hipMemcpyAsync(idx, di.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        PB_HIP(hipStreamSynchronize(s));
    });
}

void pbsm3d_fetch_destroy(pbsm3d_fetch_ctx* c) { delete c; }

const char* pbsm3d_fetch_last_error(const pbsm3d_fetch_ctx* c) { return c ? c->err.c_str() : g_fetch_err.c_str(); }

}  // extern "C"
