// Snowcover update from PBSM3D's drift_mass (CHM snobal, src/modules/snobal.cpp:363-387,
// and the snowcover adjustment of third_party/snobal/sno.cpp) on the GPU.
//
// One thread per face; the state is field-major ([field][face]), so every load and
// store of a wave is one contiguous 512-byte segment.  The update is a few dozen
// FP64 operations per face on 19 state values: HBM-bound, 2 x 19 x 8 + 8 = 312
// bytes per face.  Expressions keep the reference's order (the library is built
// with -ffp-contract=off).
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/pbsm3d.h"
#include "../../include/pbsm3d_snow.h"
#include "kernels.hpp"

namespace pb {
namespace {

// snomacros.h:209, 522, 546 and the CP_ICE / CAL_TO_J / G_TO_KG macros (:359, 618, 608);
// RHO_W0 / RHO_ICE (:226, 239) for the initial liquid water
constexpr double kFreeze = 2.7316e2;
constexpr double kMaxSnowDensity = 750;
constexpr double kMinSnowTemp = -75;
constexpr double kRhoW0 = 999.87;
constexpr double kRhoIce = 917.0;

struct SnowParams {
    double drift_density, max_z_s_0, ic_rho, max_h2o_vol, threshold;
};

// one face's snowcover in registers
struct Pack {
    double z_s, m_s, rho, layer_count, z_s_0, z_s_l, m_s_0, m_s_l, T_s, T_s_0, T_s_l, cc_s, cc_s_0, cc_s_l,
        h2o_total, h2o_vol, h2o, h2o_max, h2o_sat;
};

// state field k of face i at st[k * F + i] (PBSM3D_SNOW_* order)
#define PB_SNOW_FIELDS(X)                                                                                      \
    X(PBSM3D_SNOW_Z_S, z_s) X(PBSM3D_SNOW_M_S, m_s) X(PBSM3D_SNOW_RHO, rho)                                     \
    X(PBSM3D_SNOW_LAYER_COUNT, layer_count) X(PBSM3D_SNOW_Z_S_0, z_s_0) X(PBSM3D_SNOW_Z_S_L, z_s_l)            \
    X(PBSM3D_SNOW_M_S_0, m_s_0) X(PBSM3D_SNOW_M_S_L, m_s_l) X(PBSM3D_SNOW_T_S, T_s) X(PBSM3D_SNOW_T_S_0, T_s_0) \
    X(PBSM3D_SNOW_T_S_L, T_s_l) X(PBSM3D_SNOW_CC_S, cc_s) X(PBSM3D_SNOW_CC_S_0, cc_s_0)                         \
    X(PBSM3D_SNOW_CC_S_L, cc_s_l) X(PBSM3D_SNOW_H2O_TOTAL, h2o_total) X(PBSM3D_SNOW_H2O_VOL, h2o_vol)          \
    X(PBSM3D_SNOW_H2O, h2o) X(PBSM3D_SNOW_H2O_MAX, h2o_max) X(PBSM3D_SNOW_H2O_SAT, h2o_sat)

__device__ __forceinline__ void load(const double* __restrict__ st, int64_t F, int64_t i, Pack& p) {
#define PB_LOAD(K, M) p.M = st[(int64_t)(K) * F + i];
    PB_SNOW_FIELDS(PB_LOAD)
#undef PB_LOAD
}
__device__ __forceinline__ void store(double* __restrict__ st, int64_t F, int64_t i, const Pack& p) {
#define PB_STORE(K, M) st[(int64_t)(K) * F + i] = p.M;
    PB_SNOW_FIELDS(PB_STORE)
#undef PB_STORE
}

// specific heat of ice (J/(kg K)) at temperature t (K)
__device__ __forceinline__ double cp_ice(double t) { return ((0.024928 + (0.00176 * t)) * 4.186798188) / (1 * 0.001); }

// cold content of a layer (sno.cpp:2321-2328, heat_stor :523-533)
__device__ __forceinline__ double cold_content(double temp, double mass) {
    if (temp < kFreeze) return cp_ice(temp) * mass * (temp - kFreeze);
    return 0.0;
}

// number of layers and their depths (sno.cpp:2365-2405)
__device__ __forceinline__ void calc_layers(Pack& p, const SnowParams& c) {
    if (p.m_s <= c.threshold) {
        p.layer_count = 0;
        p.z_s = p.z_s_0 = p.z_s_l = 0.0;
    } else if (p.z_s < c.max_z_s_0) {
        p.layer_count = 1;
        p.z_s_0 = p.z_s;
        p.z_s_l = 0.0;
    } else {
        p.layer_count = 2;
        p.z_s_0 = c.max_z_s_0;
        p.z_s_l = p.z_s - p.z_s_0;
        if (p.z_s_l * p.rho < c.threshold) {
            p.layer_count = 1;
            p.z_s_0 = p.z_s;
            p.z_s_l = 0.0;
        }
    }
}

// layer masses from the density (sno.cpp:1564-1579)
__device__ __forceinline__ void layer_mass(Pack& p) {
    if (p.layer_count == 0) {
        p.m_s_0 = 0.0;
        p.m_s_l = 0.0;
    } else {
        p.m_s_0 = p.rho * p.z_s_0;
        if (p.layer_count == 2)
            p.m_s_l = p.rho * p.z_s_l;
        else
            p.m_s_l = 0.0;
    }
}

// layers after a change of depth (sno.cpp:2617-2695)
__device__ __forceinline__ void adj_layers(Pack& p, const SnowParams& c) {
    const double prev = p.layer_count;
    calc_layers(p, c);
    if (p.layer_count == 0) {
        p.rho = 0.0;
        if (p.m_s > 0.0) p.h2o_total += p.m_s;
        p.m_s = p.cc_s = 0.0;
        p.m_s_0 = p.cc_s_0 = 0.0;
        p.T_s = p.T_s_0 = kMinSnowTemp + kFreeze;
        if (prev == 2) {
            p.m_s_l = p.cc_s_l = 0.0;
            p.T_s_l = kMinSnowTemp + kFreeze;
        }
        p.h2o_vol = p.h2o = p.h2o_max = p.h2o_sat = 0.0;
    } else {
        layer_mass(p);
        if (prev == 1 && p.layer_count == 2) {
            p.T_s_l = p.T_s;
            p.cc_s_l = cold_content(p.T_s_l, p.m_s_l);
        } else if (prev == 2 && p.layer_count == 1) {
            p.T_s_l = kMinSnowTemp + kFreeze;
            p.cc_s_l = 0.0;
        }
    }
}

// depth and mass change of the snowcover (sno.cpp:2527-2575)
__device__ __forceinline__ void adj_snow(Pack& p, const SnowParams& c, double delta_z_s, double delta_m_s) {
    p.z_s += delta_z_s;
    p.m_s += delta_m_s;
    if (p.m_s < 0 || p.z_s < 0) {
        p.m_s = p.cc_s = 0.0;
        p.m_s_0 = p.cc_s_0 = 0.0;
    }
    if (p.z_s != 0.0)
        p.rho = p.m_s / p.z_s;
    else
        p.rho = 0;
    if (p.rho > kMaxSnowDensity) {
        p.rho = kMaxSnowDensity;
        p.z_s = p.m_s / p.rho;
        adj_layers(p, c);
    } else {
        if (delta_z_s != 0.0)
            adj_layers(p, c);
        else
            layer_mass(p);
    }
}

// drift_mass consumer of snobal::run (snobal.cpp:363-387)
__global__ void __launch_bounds__(kBlock) k_snow_drift(int64_t F, SnowParams c, const double* __restrict__ drift,
                                                       double* __restrict__ st) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    Pack p;
    load(st, F, i, p);
    double mass = drift[i];
    mass = chm_missing(mass) ? 0 : mass;
    const double transport_density = mass < 0. ? p.rho : c.drift_density;
    adj_snow(p, c, mass / transport_density, mass);
    store(st, F, i, p);
}

// initial snowcover: the state set by snobal::init (snobal.cpp:120-208) and sno::init_snow
// (sno.cpp:162-236) for an initial swe (NaN / -9999: none)
__global__ void __launch_bounds__(kBlock) k_snow_init(int64_t F, SnowParams c, const double* __restrict__ swe0,
                                                      double* __restrict__ st) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    Pack p{};  // every member the reference's init does not set starts at 0
    p.h2o_sat = .3;
    p.rho = 0.;
    p.T_s = p.T_s_0 = p.T_s_l = -75. + kFreeze;
    p.z_s = 0.;
    if (swe0) {
        const double s = swe0[i];
        if (!chm_missing(s)) {
            p.rho = c.ic_rho;
            p.T_s = -10 + kFreeze;
            p.T_s_0 = -15. + kFreeze;
            p.T_s_l = -15. + kFreeze;
            p.z_s = s / p.rho;
        }
    }
    // init_snow: T_s_0 / T_s_l are never below MIN_SNOW_TEMP + FREEZE here, so its reset
    // to the air temperature (sno.cpp:166-174) does not arise
    p.m_s = p.rho * p.z_s;
    calc_layers(p, c);
    if (p.layer_count == 0) {
        if (p.m_s > 0.0) p.h2o_total += p.m_s;
        p.rho = 0.0;
        p.m_s = p.cc_s = 0.0;
        p.m_s_0 = p.cc_s_0 = 0.0;
        p.m_s_l = p.cc_s_l = 0.0;
        p.T_s = p.T_s_0 = p.T_s_l = kMinSnowTemp + kFreeze;
        p.h2o_vol = p.h2o = p.h2o_max = p.h2o_sat = 0.0;
    } else {
        layer_mass(p);
        p.cc_s_0 = cold_content(p.T_s_0, p.m_s_0);
        if (p.layer_count == 2) {
            p.cc_s_l = cold_content(p.T_s_l, p.m_s_l);
        } else {
            p.T_s_l = kMinSnowTemp + kFreeze;
            p.cc_s_l = 0.0;
        }
        p.h2o_vol = p.h2o_sat * c.max_h2o_vol;
        const double rho_dry = ((p.rho) - (p.h2o_vol) * kRhoW0) / (1 - (p.h2o_vol) * kRhoW0 / kRhoIce);
        p.h2o_max = ((c.max_h2o_vol * p.z_s * kRhoW0 * (kRhoIce - rho_dry)) / kRhoIce);
        p.h2o = p.h2o_sat * p.h2o_max;
    }
    store(st, F, i, p);
}

__global__ void __launch_bounds__(kBlock) k_snow_outputs(int64_t F, const double* __restrict__ st,
                                                         double* __restrict__ swe, double* __restrict__ depth) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    if (swe) swe[i] = st[PBSM3D_SNOW_M_S * F + i];
    if (depth) depth[i] = st[PBSM3D_SNOW_Z_S * F + i];
}

}  // namespace
}  // namespace pb

using namespace pb;

struct pbsm3d_snow_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t F = 0;
    SnowParams par{};
    DevBuf<double> state, stage_in, stage_out;
    std::string err;
    ~pbsm3d_snow_ctx() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {
std::string g_snow_err;  // last creation error (no context to hold it)

template <typename Fn>
int snow_guard(pbsm3d_snow_ctx* c, Fn&& fn) {
    try {
        fn();
        return PBSM3D_OK;
    } catch (const Error& e) {
        if (c) c->err = e.what();
        g_snow_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        g_snow_err = e.what();
        return PBSM3D_EINVAL;
    }
}
}  // namespace

extern "C" {

void pbsm3d_snow_config_default(pbsm3d_snow_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->drift_density = 300.;
    c->max_active_layer = .1;
    c->ic_rho = 300.;
    c->max_h2o_vol = .0001;
    c->mass_threshold = 0.2;
}

int pbsm3d_snow_create(pbsm3d_snow_ctx** out, int device, int64_t F, const double* swe0,
                       const pbsm3d_snow_config* cfg) {
    if (!out || F <= 0 || !cfg) {
        g_snow_err = "invalid argument";
        return PBSM3D_EINVAL;
    }
    if (!(cfg->drift_density > 0) || !(cfg->max_active_layer > 0) || !(cfg->ic_rho > 0) ||
        !(cfg->max_h2o_vol >= 0) || !(cfg->mass_threshold >= 0)) {
        g_snow_err = "drift_density, max_active_layer and ic_rho must be > 0, max_h2o_vol and mass_threshold >= 0";
        return PBSM3D_EINVAL;
    }
    *out = nullptr;
    auto* c = new pbsm3d_snow_ctx();
    const int rc = snow_guard(c, [&] {
        c->device = device;
        c->F = F;
        c->par = SnowParams{cfg->drift_density, cfg->max_active_layer, cfg->ic_rho, cfg->max_h2o_vol,
                            cfg->mass_threshold};
        PB_HIP(hipSetDevice(device));
        PB_HIP(hipStreamCreate(&c->stream));
        hipStream_t s = c->stream;
        c->state.alloc((size_t)PBSM3D_SNOW_FIELDS * F);
        c->stage_in.alloc(F);
        c->stage_out.alloc(2 * F);
        const double* d0 = nullptr;
        if (swe0) {
            PB_HIP(hipMemcpyAsync(c->stage_in.p, swe0, F * sizeof(double), hipMemcpyHostToDevice, s));
            d0 = c->stage_in.p;
        }
        hipLaunchKernelGGL(k_snow_init, dim3(grid_for(F)), dim3(kBlock), 0, s, F, c->par, d0, c->state.p);
        PB_HIP(hipGetLastError());
        PB_HIP(hipStreamSynchronize(s));
    });
    if (rc != PBSM3D_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return PBSM3D_OK;
}

int pbsm3d_snow_drift(pbsm3d_snow_ctx* c, const double* drift_mass, int on_device) {
    if (!c || !drift_mass) return PBSM3D_EINVAL;
    return snow_guard(c, [&] {
        PB_HIP(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        const double* d = drift_mass;
        if (!on_device) {
            PB_HIP(hipMemcpyAsync(c->stage_in.p, drift_mass, c->F * sizeof(double), hipMemcpyHostToDevice, s));
            d = c->stage_in.p;
        }
        hipLaunchKernelGGL(k_snow_drift, dim3(grid_for(c->F)), dim3(kBlock), 0, s, c->F, c->par, d, c->state.p);
        PB_HIP(hipGetLastError());
        PB_HIP(hipStreamSynchronize(s));
    });
}

int pbsm3d_snow_outputs(pbsm3d_snow_ctx* c, double* swe, double* depth, int on_device) {
    if (!c) return PBSM3D_EINVAL;
    if (!swe && !depth) return PBSM3D_OK;
    return snow_guard(c, [&] {
        PB_HIP(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        const int64_t F = c->F;
        double* ds = on_device ? swe : (swe ? c->stage_out.p : nullptr);
        double* dd = on_device ? depth : (depth ? c->stage_out.p + F : nullptr);
        hipLaunchKernelGGL(k_snow_outputs, dim3(grid_for(F)), dim3(kBlock), 0, s, F, c->state.p, ds, dd);
        PB_HIP(hipGetLastError());
        if (!on_device) {
            if (swe) PB_HIP(hipMemcpyAsync(swe, ds, F * sizeof(double), hipMemcpyDeviceToHost, s));
            if (depth) PB_HIP(hipMemcpyAsync(depth, dd, F * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        PB_HIP(hipStreamSynchronize(s));
    });
}

int pbsm3d_snow_get_state(pbsm3d_snow_ctx* c, double* state) {
    if (!c || !state) return PBSM3D_EINVAL;
    return snow_guard(c, [&] {
        PB_HIP(hipSetDevice(c->device));
        PB_HIP(hipMemcpyAsync(state, c->state.p, c->state.bytes(), hipMemcpyDeviceToHost, c->stream));
        PB_HIP(hipStreamSynchronize(c->stream));
    });
}

int pbsm3d_snow_set_state(pbsm3d_snow_ctx* c, const double* state) {
    if (!c || !state) return PBSM3D_EINVAL;
    return snow_guard(c, [&] {
        PB_HIP(hipSetDevice(c->device));
        PB_HIP(hipMemcpyAsync(c->state.p, state, c->state.bytes(), hipMemcpyHostToDevice, c->stream));
        PB_HIP(hipStreamSynchronize(c->stream));
    });
}

void pbsm3d_snow_destroy(pbsm3d_snow_ctx* c) { delete c; }

const char* pbsm3d_snow_last_error(const pbsm3d_snow_ctx* c) { return c ? c->err.c_str() : g_snow_err.c_str(); }

}  // extern "C"
