// Vertical wind scaling (CHM scale_wind_vert, domain mode) on the GPU.
//
// Two passes over the faces: k_point_scale applies the log / exponential
// profiles of scale_wind_vert::point_scale (scale_wind_vert.cpp:56-155) to
// every face; k_smooth then gives each face the thin-plate-spline value of its
// (at most three) neighbours' point-scaled winds at its own centroid
// (scale_wind_vert.cpp:199-233, TPSpline.cpp:39-180): an (n+1) x (n+1) system
// per face, solved in registers by Gaussian elimination with partial pivoting.
// E1 (the reference takes GSL's gsl_sf_expint_E1) is evaluated from
// Abramowitz & Stegun: the power series 5.1.11 for x <= 1 and the continued
// fraction 5.1.22 by Lentz's method above.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/pbsm3d_wind.h"
#include "kernels.hpp"

namespace pb {
namespace {

constexpr double kEulerGamma = 0.57721566490153286061;

__device__ double expint_e1(double x) {
    if (!(x > 0)) return x == 0 ? INFINITY : NAN;
    if (x <= 1.0) {
        double sum = 0, term = 1;
        for (int k = 1; k < 60; ++k) {
            term *= -x / k;
            const double add = term / k;
            sum += add;
            if (fabs(add) < 1e-18 * fabs(sum)) break;
        }
        return -kEulerGamma - log(x) - sum;
    }
    const double tiny = 1e-300;
    double b = x + 1.0, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i < 1000; ++i) {
        const double an = -(double)i * i;
        b += 2.0;
        d = 1.0 / (an * d + b);
        c = b + an / c;
        const double del = c * d;
        h *= del;
        if (fabs(del - 1.0) < 1e-16) break;
    }
    return h * exp(-x);
}

__device__ __forceinline__ double log_scale_wind(double u, double Z_in, double Z_out, double snowdepthavg,
                                                 double z0) {
    const double Z0_SNOW = z0;
    u = u * log((Z_out - (snowdepthavg + Z0_SNOW)) / Z0_SNOW) / log((Z_in - (snowdepthavg + Z0_SNOW)) / Z0_SNOW);
    return u;
}

__device__ __forceinline__ double exp_scale_wind(double u, double Z_in, double Z_out, double alpha) {
    u = u * exp(alpha * (Z_out / Z_in - 1));
    return u;
}

struct WindArgs {
    int64_t F;
    int ignore_canopy;
    double z0;
    const double *canopy, *lai;  // null: no vegetation
    const double *U_R, *snowdepthavg;
    double* out;
};

__global__ void __launch_bounds__(kBlock) k_point_scale(WindArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.F) return;
    const double U_R = a.U_R[i];
    const double Z_R = 50.0;  // Atmosphere::Z_U_R
    double Z_CanTop = 0;
    if (!a.ignore_canopy && a.canopy) Z_CanTop = a.canopy[i];
    const double Z_CanBot = Z_CanTop / 2.0;
    double snowdepthavg = a.snowdepthavg ? a.snowdepthavg[i] : 0.0;
    if (chm_missing(snowdepthavg)) snowdepthavg = 0.0;
    const double Z_2m_above_srf = snowdepthavg + 2.0;
    if (Z_2m_above_srf >= Z_R) {
        a.out[i] = U_R;
        return;
    }
    double U_2m_above_srf;
    if (!a.ignore_canopy && (Z_CanTop > 0.0) && (Z_2m_above_srf < Z_CanTop)) {
        const double alpha = a.lai[i];
        if (snowdepthavg < Z_CanTop) {
            const double U_CanTop = log_scale_wind(U_R, Z_R, Z_CanTop, snowdepthavg, a.z0);
            const double U_CanBot = exp_scale_wind(U_CanTop, Z_CanTop, Z_CanBot, alpha);
            if (Z_2m_above_srf < Z_CanBot)
                U_2m_above_srf = log_scale_wind(U_CanBot, Z_CanBot, Z_2m_above_srf, snowdepthavg, a.z0);
            else
                U_2m_above_srf = exp_scale_wind(U_CanTop, Z_CanTop, Z_2m_above_srf, alpha);
        } else {
            U_2m_above_srf = log_scale_wind(U_R, Z_R, Z_2m_above_srf, snowdepthavg, a.z0);
        }
    } else {
        U_2m_above_srf = log_scale_wind(U_R, Z_R, Z_2m_above_srf, snowdepthavg, a.z0);
    }
    a.out[i] = fmax(0.1, U_2m_above_srf);
}

__device__ __forceinline__ double tps_basis(double sx, double sy, double ex, double ey) {
    const double c = 0.577215, weight = 0.01;  // TPSpline.cpp:196-201
    const double xdiff = sx - ex, ydiff = sy - ey;
    double dij = sqrt(xdiff * xdiff + ydiff * ydiff);
    dij = (dij * weight / 2.0) * (dij * weight / 2.0);
    return -(log(dij) + c + expint_e1(dij));
}

__global__ void k_clear_flag(double* flag) { *flag = 0.0; }

// neigh[i * sf + j * sj] is neighbour j of face i (sf = 1, sj = F: the context's j-major
// copy; sf = 3, sj = 1: the ABI's face-major array).  CHECKED: the indices come straight
// from the caller's device memory, unvalidated: one outside [-1, F) counts as no
// neighbour and sets *bad = 1.
template <bool CHECKED>
__global__ void __launch_bounds__(kBlock) k_smooth(int64_t F, const int32_t* __restrict__ neigh, int64_t sf, int64_t sj,
                                                   const double* __restrict__ cx, const double* __restrict__ cy,
                                                   const double* __restrict__ pt, double* __restrict__ out,
                                                   double* __restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    double sx[3], sy[3], sz[3];
    int n = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        int32_t nb = neigh[i * sf + j * sj];
        if (CHECKED && (nb < -1 || nb >= F)) {
            *bad = 1.0;
            nb = -1;
        }
        if (nb < 0) continue;
        sx[n] = cx[nb];
        sy[n] = cy[nb];
        sz[n] = pt[nb];
        ++n;
    }
    if (n == 0) {
        out[i] = fmax(0.1, pt[i]);
        return;
    }
    const int m = n + 1;
    double A[4][5];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int cc = 0; cc < 5; ++cc) A[r][cc] = 0;
    for (int r = 0; r < n; ++r)
        for (int j = r; j < n; ++j) {
            const double xdiff = sx[r] - sx[j], ydiff = sy[r] - sy[j];
            if (xdiff == 0. && ydiff == 0.) continue;
            const double Rd = tps_basis(sx[r], sy[r], sx[j], sy[j]);
            A[r][j + 1] = Rd;
            A[j][r + 1] = Rd;
        }
    for (int r = 0; r < m; ++r) {
        A[r][0] = 1;
        A[m - 1][r] = 1;
    }
    A[m - 1][0] = 0;
    for (int r = 0; r < n; ++r) A[r][m] = sz[r];
    A[m - 1][m] = 0.0;
    for (int k = 0; k < m; ++k) {
        int p = k;
        for (int r = k + 1; r < m; ++r)
            if (fabs(A[r][k]) > fabs(A[p][k])) p = r;
        if (p != k)
            for (int cc = 0; cc <= m; ++cc) {
                const double t = A[k][cc];
                A[k][cc] = A[p][cc];
                A[p][cc] = t;
            }
        for (int r = k + 1; r < m; ++r) {
            const double f = A[r][k] / A[k][k];
            for (int cc = k; cc <= m; ++cc) A[r][cc] -= f * A[k][cc];
        }
    }
    double x[4];
    for (int k = m - 1; k >= 0; --k) {
        double s = A[k][m];
        for (int cc = k + 1; cc < m; ++cc) s -= A[k][cc] * x[cc];
        x[k] = s / A[k][k];
    }
    double z0 = x[0];
    for (int r = 1; r < m; ++r) z0 = z0 + x[r] * tps_basis(sx[r - 1], sy[r - 1], cx[i], cy[i]);
    out[i] = fmax(0.1, z0);
}

}  // namespace
}  // namespace pb

using namespace pb;

struct pbsm3d_wind_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t F = 0;
    pbsm3d_wind_config cfg{};
    DevBuf<int32_t> neigh;  // j-major [j*F + i]
    DevBuf<double> cx, cy, canopy, lai, stage_u, stage_sd, stage_out, point;
    bool veg = false;
    std::string err;
    ~pbsm3d_wind_ctx() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// Scratch of pbsm3d_wind_run_noalloc, which allocates nothing: the point-scaled wind of
// every face (the smoothing kernel reads the neighbours' values back) and, after it, one
// flag word for out-of-range neighbour indices.  32 MiB per GPU for the rest of the
// process.  Calls that overlap on one GPU share it and race.
static __device__ __align__(256) double scratch[PBSM3D_WIND_NOALLOC_MAX_FACES + 32];

namespace {
std::string g_wind_err;

// The device address of `scratch` on `device` (current), looked up once per device: each
// GPU holds its own copy at its own address, and the host-side name is not that address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

double* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(PBSM3D_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        PB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    return g_scratch_ptr[device];
}

template <typename Fn>
int wind_guard(pbsm3d_wind_ctx* c, Fn&& fn) {
    try {
        fn();
        return PBSM3D_OK;
    } catch (const Error& e) {
        if (c) c->err = e.what();
        g_wind_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        g_wind_err = e.what();
        return PBSM3D_EINVAL;
    }
}
}  // namespace

extern "C" {

void pbsm3d_wind_config_default(pbsm3d_wind_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->ignore_canopy = 0;
    c->smooth = 1;
    c->z0_snow = 0.01;
}

int pbsm3d_wind_create(pbsm3d_wind_ctx** out, int device, int64_t F, const int32_t* neigh, const double* x,
                       const double* y, const double* canopy, const double* lai, const pbsm3d_wind_config* cfg) {
    if (!out || F <= 0 || F > INT32_MAX || !neigh || !x || !y || !cfg || !(cfg->z0_snow > 0)) {
        g_wind_err = "invalid argument";
        return PBSM3D_EINVAL;
    }
    if (canopy && !lai && !cfg->ignore_canopy) {
        g_wind_err = "LAI is required with canopy heights (scale_wind_vert.cpp:110)";
        return PBSM3D_EINVAL;
    }
    for (int64_t k = 0; k < 3 * F; ++k)
        if (neigh[k] < -1 || neigh[k] >= F) {
            g_wind_err = "neighbour index out of range";
            return PBSM3D_EINVAL;
        }
    *out = nullptr;
    auto* c = new pbsm3d_wind_ctx();
    const int rc = wind_guard(c, [&] {
        c->device = device;
        c->F = F;
        c->cfg = *cfg;
        PB_HIP(hipSetDevice(device));
        PB_HIP(hipStreamCreate(&c->stream));
        hipStream_t s = c->stream;
        std::vector<int32_t> nj(3 * F);
        for (int64_t i = 0; i < F; ++i)
            for (int j = 0; j < 3; ++j) nj[j * F + i] = neigh[3 * i + j];
        c->neigh.alloc(3 * F);
        PB_HIP(hipMemcpyAsync(c->neigh.p, nj.data(), 3 * F * sizeof(int32_t), hipMemcpyHostToDevice, s));
        auto up = [&](DevBuf<double>& d, const double* h) {
            d.alloc(F);
            PB_HIP(hipMemcpyAsync(d.p, h, F * sizeof(double), hipMemcpyHostToDevice, s));
        };
        up(c->cx, x);
        up(c->cy, y);
        c->veg = canopy != nullptr;
        if (canopy) up(c->canopy, canopy);
        if (lai) up(c->lai, lai);
        for (DevBuf<double>* b : {&c->stage_u, &c->stage_sd, &c->stage_out, &c->point}) b->alloc(F);
        PB_HIP(hipStreamSynchronize(s));
    });
    if (rc != PBSM3D_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return PBSM3D_OK;
}

int pbsm3d_wind_run(pbsm3d_wind_ctx* c, const double* U_R, const double* snowdepthavg, double* U2, int on_device) {
    if (!c || !U_R || !U2) return PBSM3D_EINVAL;
    return wind_guard(c, [&] {
        PB_HIP(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        const int64_t F = c->F;
        const double* du = U_R;
        const double* dsd = snowdepthavg;
        double* dout = U2;
        if (!on_device) {
            PB_HIP(hipMemcpyAsync(c->stage_u.p, U_R, F * sizeof(double), hipMemcpyHostToDevice, s));
            du = c->stage_u.p;
            if (snowdepthavg) {
                PB_HIP(hipMemcpyAsync(c->stage_sd.p, snowdepthavg, F * sizeof(double), hipMemcpyHostToDevice, s));
                dsd = c->stage_sd.p;
            }
            dout = c->stage_out.p;
        }
        WindArgs a{F, c->cfg.ignore_canopy, c->cfg.z0_snow, c->veg ? c->canopy.p : nullptr, c->lai.p, du, dsd,
                   c->cfg.smooth ? c->point.p : dout};
        hipLaunchKernelGGL(k_point_scale, dim3(grid_for(F)), dim3(kBlock), 0, s, a);
        if (c->cfg.smooth)
            hipLaunchKernelGGL(k_smooth<false>, dim3(grid_for(F)), dim3(kBlock), 0, s, F, c->neigh.p, (int64_t)1, F,
                               c->cx.p, c->cy.p, c->point.p, dout, (double*)nullptr);
        PB_HIP(hipGetLastError());
        if (!on_device) PB_HIP(hipMemcpyAsync(U2, dout, F * sizeof(double), hipMemcpyDeviceToHost, s));
        PB_HIP(hipStreamSynchronize(s));
    });
}

int pbsm3d_wind_run_noalloc(int device, int64_t F, const int32_t* neigh, const double* x, const double* y,
                            const double* canopy, const double* lai, const pbsm3d_wind_config* cfg, const double* U_R,
                            const double* snowdepthavg, double* U2) {
    if (F <= 0 || !neigh || !x || !y || !cfg || !U_R || !U2 || !(cfg->z0_snow > 0)) {
        g_wind_err = "invalid argument";
        return PBSM3D_EINVAL;
    }
    if (F > PBSM3D_WIND_NOALLOC_MAX_FACES) {
        g_wind_err = "pbsm3d_wind_run_noalloc: " + std::to_string(F) + " faces exceed the scratch of " +
                     std::to_string((long long)PBSM3D_WIND_NOALLOC_MAX_FACES) + " faces";
        return PBSM3D_EINVAL;
    }
    if (canopy && !lai && !cfg->ignore_canopy) {
        g_wind_err = "LAI is required with canopy heights (scale_wind_vert.cpp:110)";
        return PBSM3D_EINVAL;
    }
    return wind_guard(nullptr, [&] {
        PB_HIP(hipSetDevice(device));
        double* scr = noalloc_scratch(device);  // [max faces] point-scaled winds, then the flag word
        double* bad = scr + PBSM3D_WIND_NOALLOC_MAX_FACES;
        hipStream_t s = nullptr;  // the device's null stream: creating a stream is left to the caller's set-up
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, bad);
        WindArgs a{F, cfg->ignore_canopy, cfg->z0_snow, canopy, lai, U_R, snowdepthavg, cfg->smooth ? scr : U2};
        hipLaunchKernelGGL(k_point_scale, dim3(grid_for(F)), dim3(kBlock), 0, s, a);
        if (cfg->smooth)
            hipLaunchKernelGGL(k_smooth<true>, dim3(grid_for(F)), dim3(kBlock), 0, s, F, neigh, (int64_t)3, (int64_t)1,
                               x, y, scr, U2, bad);
        PB_HIP(hipGetLastError());
        PB_HIP(hipStreamSynchronize(s));
        double flag = 0;
        PB_HIP(hipMemcpy(&flag, bad, sizeof(double), hipMemcpyDeviceToHost));
        if (flag != 0.0) throw Error(PBSM3D_EINVAL, "neighbour index out of range");
    });
}

void pbsm3d_wind_destroy(pbsm3d_wind_ctx* c) { delete c; }

const char* pbsm3d_wind_last_error(const pbsm3d_wind_ctx* c) { return c ? c->err.c_str() : g_wind_err.c_str(); }

}  // extern "C"
