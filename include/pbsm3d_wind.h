/* Vertical wind scaling on the GPU: the MI355X replacement of CHM's
 * `scale_wind_vert` module in domain mode (src/modules/scale_wind_vert.cpp:28-242),
 * the producer of PBSM3D's U_2m_above_srf input.
 *
 * Per face (point_scale, scale_wind_vert.cpp:56-155): the reference-height wind
 * U_R (at Z_R = 50 m) is scaled to 2 m above the snow surface with the neutral
 * log law (z0 = 0.01 m, Atmosphere.cpp:32-38), through the canopy with
 * Inoue's exponential profile (Atmosphere.cpp:41-46, attenuation = LAI) when a
 * canopy taller than the 2 m level exists, floored at 0.1 m/s.  Then (domain
 * mode, :182-242) every face takes the thin-plate-spline interpolation of its
 * neighbours' point-scaled winds at its own centroid (TPSpline.cpp:39-180),
 * floored at 0.1; a face without neighbours keeps its own value. */
#ifndef PBSM3D_WIND_H
#define PBSM3D_WIND_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbsm3d_wind_config {
    int32_t ignore_canopy;  /* 0 (scale_wind_vert.cpp:170)                        */
    int32_t smooth;         /* 1: domain mode, neighbour spline smoothing; 0: the
                               point-mode result (point_scale only)               */
    double z0_snow;         /* 0.01 = Snow::Z0_SNOW (Snow.h:31)                    */
} pbsm3d_wind_config;

typedef struct pbsm3d_wind_ctx pbsm3d_wind_ctx;

void pbsm3d_wind_config_default(pbsm3d_wind_config* cfg);

/* Host arrays: neigh [3F] face-major (-1 = boundary) as in pbsm3d_mesh, face
 * centroids (x, y); canopy_height and lai per face, or NULL when the mesh has
 * no vegetation (lai is required when canopy_height is given and the canopy is
 * not ignored). */
int pbsm3d_wind_create(pbsm3d_wind_ctx** out, int device, int64_t n_faces, const int32_t* neigh,
                       const double* center_x, const double* center_y, const double* canopy_height,
                       const double* lai, const pbsm3d_wind_config* cfg);

/* U_2m_above_srf[i] from U_R[i] and snowdepthavg[i] (NULL: the optional
 * variable is absent, depth 0; -9999 / NaN entries count as 0).  on_device = 1:
 * device pointers (queued on the context's stream, complete on return); 0: host. */
int pbsm3d_wind_run(pbsm3d_wind_ctx* ctx, const double* U_R, const double* snowdepthavg, double* U_2m_above_srf,
                    int on_device);

/* Allocation-free variant of pbsm3d_wind_create + pbsm3d_wind_run.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (a tight serving
 * loop, a harness that forbids hipMalloc).  Every array is a device pointer the caller
 * owns: neigh [3F] face-major as in pbsm3d_wind_create (indices are checked on the
 * device: one outside [-1, F) makes the call return PBSM3D_EINVAL), centroids,
 * optional canopy / lai / snowdepthavg (NULL as in pbsm3d_wind_create / _run), U_R in,
 * U_2m_above_srf out.  The per-face intermediate lives in a fixed device array of the
 * library (32 MiB per GPU, held until the process ends), so n_faces is at most
 * PBSM3D_WIND_NOALLOC_MAX_FACES (else PBSM3D_EINVAL).  Work runs on the device's null
 * stream and is complete on return.  Calls that overlap on one GPU share that array
 * and race. */
#define PBSM3D_WIND_NOALLOC_MAX_FACES 4194304
int pbsm3d_wind_run_noalloc(int device, int64_t n_faces, const int32_t* neigh, const double* center_x,
                            const double* center_y, const double* canopy_height, const double* lai,
                            const pbsm3d_wind_config* cfg, const double* U_R, const double* snowdepthavg,
                            double* U_2m_above_srf);

void pbsm3d_wind_destroy(pbsm3d_wind_ctx* ctx);
/* ctx NULL: the last pbsm3d_wind_create error */
const char* pbsm3d_wind_last_error(const pbsm3d_wind_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* PBSM3D_WIND_H */
