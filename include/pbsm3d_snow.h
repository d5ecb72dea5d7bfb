/* Snowpack update from blowing-snow redistribution on the GPU: the consumer of
 * PBSM3D's drift_mass in CHM's `snobal` module (src/modules/snobal.cpp:363-387)
 * with the snowcover adjustment it calls (third_party/snobal/sno.cpp:
 * _adj_snow 2527-2575, _adj_layers 2617-2695, _calc_layers 2365-2405,
 * _layer_mass 1564-1579, _cold_content 2321-2328), and the module's initial
 * snowcover (snobal.cpp:120-208, sno.cpp init_snow 162-236).
 *
 * Per face and step: mass = drift_mass (-9999 / NaN count as 0); snow is
 * removed at the pack's own density (mass < 0) or added at drift_density, and
 * the two-layer snowcover (depth, mass, density, layer split at
 * max_active_layer, the mass threshold below which a layer does not exist,
 * the 750 kg/m^3 density cap, layer cold contents) is adjusted as the
 * reference does.  The module then provides swe = m_s and snowdepthavg = z_s,
 * PBSM3D's swe and snowdepthavg inputs of the next step.
 *
 * The energy and mass balance of the rest of snobal (radiation, turbulent
 * fluxes, melt, runoff, compaction, precipitation) is not part of this
 * library. */
#ifndef PBSM3D_SNOW_H
#define PBSM3D_SNOW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbsm3d_snow_config {
    double drift_density;    /* 300 kg/m^3 (snobal.cpp:83)                        */
    double max_active_layer; /* 0.1 m = max_z_s_0 (snobal.cpp:101)                */
    double ic_rho;           /* 300 kg/m^3: density of an initial swe (:199)      */
    double max_h2o_vol;      /* 1e-4 (snobal.cpp:99)                              */
    double mass_threshold;   /* 0.2 kg/m^2 = tstep_info[SMALL_TSTEP].threshold (:190) */
} pbsm3d_snow_config;

/* Snowcover state, field-major [PBSM3D_SNOW_FIELDS x F] (field k of face i at
 * k * F + i), the members of the reference's snowcover that the drift update
 * reads or writes; layer_count is stored as a double (0, 1, 2). */
#define PBSM3D_SNOW_Z_S 0
#define PBSM3D_SNOW_M_S 1
#define PBSM3D_SNOW_RHO 2
#define PBSM3D_SNOW_LAYER_COUNT 3
#define PBSM3D_SNOW_Z_S_0 4
#define PBSM3D_SNOW_Z_S_L 5
#define PBSM3D_SNOW_M_S_0 6
#define PBSM3D_SNOW_M_S_L 7
#define PBSM3D_SNOW_T_S 8
#define PBSM3D_SNOW_T_S_0 9
#define PBSM3D_SNOW_T_S_L 10
#define PBSM3D_SNOW_CC_S 11
#define PBSM3D_SNOW_CC_S_0 12
#define PBSM3D_SNOW_CC_S_L 13
#define PBSM3D_SNOW_H2O_TOTAL 14
#define PBSM3D_SNOW_H2O_VOL 15
#define PBSM3D_SNOW_H2O 16
#define PBSM3D_SNOW_H2O_MAX 17
#define PBSM3D_SNOW_H2O_SAT 18
#define PBSM3D_SNOW_FIELDS 19

typedef struct pbsm3d_snow_ctx pbsm3d_snow_ctx;

void pbsm3d_snow_config_default(pbsm3d_snow_config* cfg);

/* swe0: host array of the initial swe per face (kg/m^2; -9999 / NaN: no
 * initial condition), or NULL for none anywhere -- the reference's
 * has_initial_condition("swe") branch and init_snow().  cc_s, which init_snow
 * leaves to the first energy-balance step, starts at 0. */
int pbsm3d_snow_create(pbsm3d_snow_ctx** out, int device, int64_t n_faces, const double* swe0,
                       const pbsm3d_snow_config* cfg);

/* One step's drift: drift_mass[F] (kg/m^2, PBSM3D's output).  on_device = 1:
 * a device pointer (queued on the context's stream, complete on return);
 * 0: host. */
int pbsm3d_snow_drift(pbsm3d_snow_ctx* ctx, const double* drift_mass, int on_device);

/* swe = m_s and snowdepthavg = z_s per face (either may be NULL); on_device as above. */
int pbsm3d_snow_outputs(pbsm3d_snow_ctx* ctx, double* swe, double* snowdepthavg, int on_device);

/* The whole state, host arrays [PBSM3D_SNOW_FIELDS x F] (checkpoint / restart). */
int pbsm3d_snow_get_state(pbsm3d_snow_ctx* ctx, double* state);
int pbsm3d_snow_set_state(pbsm3d_snow_ctx* ctx, const double* state);

void pbsm3d_snow_destroy(pbsm3d_snow_ctx* ctx);
/* ctx NULL: the last pbsm3d_snow_create error */
const char* pbsm3d_snow_last_error(const pbsm3d_snow_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* PBSM3D_SNOW_H */
