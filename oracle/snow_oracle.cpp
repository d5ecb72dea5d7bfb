// snow_oracle.cpp -- CPU FP64 restatement of the drift_mass consumer of CHM's
// snobal module and of the snowcover adjustment it calls.
//
// TEST INFRASTRUCTURE ONLY: the checker of chm_amd/csrc/snow.hip.
//
//   * snobal::run, drift block          src/modules/snobal.cpp:363-387
//     (mass = drift_mass, is_nan -> 0; removal at the pack density rho,
//      deposition at drift_density; _adj_snow(mass / density, mass))
//   * module_base::is_nan               src/modules/module_base.hpp:471-479
//   * sno::_adj_snow                    third_party/snobal/sno.cpp:2527-2575
//   * sno::_adj_layers                  third_party/snobal/sno.cpp:2617-2695
//   * sno::_calc_layers                 third_party/snobal/sno.cpp:2365-2405
//     (mass threshold = tstep_info[SMALL_TSTEP].threshold, snobal.cpp:190)
//   * sno::_layer_mass                  third_party/snobal/sno.cpp:1564-1579
//   * sno::_cold_content, heat_stor     third_party/snobal/sno.cpp:2321-2328, 523-533
//   * snobal::init state + init_snow    src/modules/snobal.cpp:120-208, sno.cpp:162-236
//   * constants and macros              third_party/snobal/snomacros.h: FREEZE :209,
//     MAX_SNOW_DENSITY :522, MIN_SNOW_TEMP :546, CP_ICE :359, CAL_TO_J :618,
//     G_TO_KG :608, RHO_W0 :226, RHO_ICE :239, H2O_LEFT :111, DRY_SNO_RHO :121
//
// Built with -ffp-contract=off like the rest of the oracle.

#include <cmath>
#include <cstdint>

#include "../include/pbsm3d_snow.h"

namespace {

const double FREEZE = 2.7316e2;
const double MAX_SNOW_DENSITY = 750;
const double MIN_SNOW_TEMP = -75;
const double RHO_W0 = 999.87;
const double RHO_ICE = 917.0;

struct Snowcover {
    double f[PBSM3D_SNOW_FIELDS];
    double& operator[](int k) { return f[k]; }
};

double cp_ice(double t) { return ((0.024928 + (0.00176 * (t))) * 4.186798188) / ((1) * 0.001); }
double heat_stor(double cp, double spm, double tdif) { return cp * spm * tdif; }
double cold_content(double temp, double mass) {
    if (temp < FREEZE) return heat_stor(cp_ice(temp), mass, (temp - FREEZE));
    return 0.0;
}

struct Sno {
    Snowcover& s;
    const pbsm3d_snow_config& c;

    double& z_s() { return s[PBSM3D_SNOW_Z_S]; }
    double& m_s() { return s[PBSM3D_SNOW_M_S]; }
    double& rho() { return s[PBSM3D_SNOW_RHO]; }
    double& layers() { return s[PBSM3D_SNOW_LAYER_COUNT]; }

    void calc_layers() {
        if (m_s() <= c.mass_threshold) {
            layers() = 0;
            z_s() = s[PBSM3D_SNOW_Z_S_0] = s[PBSM3D_SNOW_Z_S_L] = 0.0;
        } else if (z_s() < c.max_active_layer) {
            layers() = 1;
            s[PBSM3D_SNOW_Z_S_0] = z_s();
            s[PBSM3D_SNOW_Z_S_L] = 0.0;
        } else {
            layers() = 2;
            s[PBSM3D_SNOW_Z_S_0] = c.max_active_layer;
            s[PBSM3D_SNOW_Z_S_L] = z_s() - s[PBSM3D_SNOW_Z_S_0];
            if (s[PBSM3D_SNOW_Z_S_L] * rho() < c.mass_threshold) {
                layers() = 1;
                s[PBSM3D_SNOW_Z_S_0] = z_s();
                s[PBSM3D_SNOW_Z_S_L] = 0.0;
            }
        }
    }

    void layer_mass() {
        if (layers() == 0) {
            s[PBSM3D_SNOW_M_S_0] = 0.0;
            s[PBSM3D_SNOW_M_S_L] = 0.0;
        } else {
            s[PBSM3D_SNOW_M_S_0] = rho() * s[PBSM3D_SNOW_Z_S_0];
            s[PBSM3D_SNOW_M_S_L] = layers() == 2 ? rho() * s[PBSM3D_SNOW_Z_S_L] : 0.0;
        }
    }

    void adj_layers() {
        const double prev = layers();
        calc_layers();
        if (layers() == 0) {
            rho() = 0.0;
            if (m_s() > 0.0) s[PBSM3D_SNOW_H2O_TOTAL] += m_s();
            m_s() = s[PBSM3D_SNOW_CC_S] = 0.0;
            s[PBSM3D_SNOW_M_S_0] = s[PBSM3D_SNOW_CC_S_0] = 0.0;
            s[PBSM3D_SNOW_T_S] = s[PBSM3D_SNOW_T_S_0] = MIN_SNOW_TEMP + FREEZE;
            if (prev == 2) {
                s[PBSM3D_SNOW_M_S_L] = s[PBSM3D_SNOW_CC_S_L] = 0.0;
                s[PBSM3D_SNOW_T_S_L] = MIN_SNOW_TEMP + FREEZE;
            }
            s[PBSM3D_SNOW_H2O_VOL] = s[PBSM3D_SNOW_H2O] = s[PBSM3D_SNOW_H2O_MAX] = s[PBSM3D_SNOW_H2O_SAT] = 0.0;
        } else {
            layer_mass();
            if (prev == 1 && layers() == 2) {
                s[PBSM3D_SNOW_T_S_L] = s[PBSM3D_SNOW_T_S];
                s[PBSM3D_SNOW_CC_S_L] = cold_content(s[PBSM3D_SNOW_T_S_L], s[PBSM3D_SNOW_M_S_L]);
            } else if (prev == 2 && layers() == 1) {
                s[PBSM3D_SNOW_T_S_L] = MIN_SNOW_TEMP + FREEZE;
                s[PBSM3D_SNOW_CC_S_L] = 0.0;
            }
        }
    }

    void adj_snow(double delta_z_s, double delta_m_s) {
        z_s() += delta_z_s;
        m_s() += delta_m_s;
        if (m_s() < 0 || z_s() < 0) {
            m_s() = s[PBSM3D_SNOW_CC_S] = 0.0;
            s[PBSM3D_SNOW_M_S_0] = s[PBSM3D_SNOW_CC_S_0] = 0.0;
        }
        rho() = z_s() != 0.0 ? m_s() / z_s() : 0;
        if (rho() > MAX_SNOW_DENSITY) {
            rho() = MAX_SNOW_DENSITY;
            z_s() = m_s() / rho();
            adj_layers();
        } else if (delta_z_s != 0.0) {
            adj_layers();
        } else {
            layer_mass();
        }
    }

    void init_snow() {
        m_s() = rho() * z_s();
        calc_layers();
        if (layers() == 0) {
            if (m_s() > 0.0) s[PBSM3D_SNOW_H2O_TOTAL] += m_s();
            rho() = 0.0;
            m_s() = s[PBSM3D_SNOW_CC_S] = 0.0;
            s[PBSM3D_SNOW_M_S_0] = s[PBSM3D_SNOW_CC_S_0] = 0.0;
            s[PBSM3D_SNOW_M_S_L] = s[PBSM3D_SNOW_CC_S_L] = 0.0;
            s[PBSM3D_SNOW_T_S] = s[PBSM3D_SNOW_T_S_0] = s[PBSM3D_SNOW_T_S_L] = MIN_SNOW_TEMP + FREEZE;
            s[PBSM3D_SNOW_H2O_VOL] = s[PBSM3D_SNOW_H2O] = s[PBSM3D_SNOW_H2O_MAX] = s[PBSM3D_SNOW_H2O_SAT] = 0.0;
            return;
        }
        layer_mass();
        s[PBSM3D_SNOW_CC_S_0] = cold_content(s[PBSM3D_SNOW_T_S_0], s[PBSM3D_SNOW_M_S_0]);
        if (layers() == 2) {
            s[PBSM3D_SNOW_CC_S_L] = cold_content(s[PBSM3D_SNOW_T_S_L], s[PBSM3D_SNOW_M_S_L]);
        } else {
            s[PBSM3D_SNOW_T_S_L] = MIN_SNOW_TEMP + FREEZE;
            s[PBSM3D_SNOW_CC_S_L] = 0.0;
        }
        s[PBSM3D_SNOW_H2O_VOL] = s[PBSM3D_SNOW_H2O_SAT] * c.max_h2o_vol;
        const double sat = s[PBSM3D_SNOW_H2O_VOL];
        const double rho_dry = ((rho()) - (sat)*RHO_W0) / (1 - (sat)*RHO_W0 / RHO_ICE);  // DRY_SNO_RHO
        s[PBSM3D_SNOW_H2O_MAX] = ((c.max_h2o_vol * z_s() * RHO_W0 * (RHO_ICE - rho_dry)) / RHO_ICE);  // H2O_LEFT
        s[PBSM3D_SNOW_H2O] = s[PBSM3D_SNOW_H2O_SAT] * s[PBSM3D_SNOW_H2O_MAX];
    }
};

bool is_nan(double v) { return std::fabs(v - -9999.0) < 1e-5 || std::isnan(v); }

}  // namespace

extern "C" {

// state [PBSM3D_SNOW_FIELDS x F] field-major
void oracle_snow_init(int64_t F, const double* swe0, const pbsm3d_snow_config* cfg, double* state) {
    for (int64_t i = 0; i < F; ++i) {
        Snowcover s{};
        s[PBSM3D_SNOW_H2O_SAT] = .3;
        s[PBSM3D_SNOW_RHO] = 0.;
        s[PBSM3D_SNOW_T_S] = s[PBSM3D_SNOW_T_S_0] = s[PBSM3D_SNOW_T_S_L] = -75. + FREEZE;
        s[PBSM3D_SNOW_Z_S] = 0.;
        if (swe0 && !is_nan(swe0[i])) {
            s[PBSM3D_SNOW_RHO] = cfg->ic_rho;
            s[PBSM3D_SNOW_T_S] = -10 + FREEZE;
            s[PBSM3D_SNOW_T_S_0] = -15. + FREEZE;
            s[PBSM3D_SNOW_T_S_L] = -15. + FREEZE;
            s[PBSM3D_SNOW_Z_S] = swe0[i] / s[PBSM3D_SNOW_RHO];
        }
        Sno{s, *cfg}.init_snow();
        for (int k = 0; k < PBSM3D_SNOW_FIELDS; ++k) state[k * F + i] = s[k];
    }
}

void oracle_snow_drift(int64_t F, const pbsm3d_snow_config* cfg, const double* drift_mass, double* state) {
    for (int64_t i = 0; i < F; ++i) {
        Snowcover s;
        for (int k = 0; k < PBSM3D_SNOW_FIELDS; ++k) s[k] = state[k * F + i];
        double mass = drift_mass[i];
        mass = is_nan(mass) ? 0 : mass;
        const double transport_density = mass < 0. ? s[PBSM3D_SNOW_RHO] : cfg->drift_density;
        Sno{s, *cfg}.adj_snow(mass / transport_density, mass);
        for (int k = 0; k < PBSM3D_SNOW_FIELDS; ++k) state[k * F + i] = s[k];
    }
}

}  // extern "C"
