/* Upwind fetch on the GPU: the MI355X replacement of CHM's `fetchr` module
 * (src/modules/fetchr.cpp:28-120), the producer of the `fetch` input that
 * PBSM3D's tanh / exp fetch forms read (PBSM3D.cpp:706-718).
 *
 * For every face, points are taken along the upwind azimuth vw_dir at
 * j * max_distance / steps (j = 1 .. steps, projected coordinates,
 * coordinates.cpp:60-71); the face whose centroid is nearest to each point
 * (triangulation.cpp:170-178, a k = 1 kd-tree query on centroids in the
 * reference, an exact grid-bucket search here) is tested for terrain or
 * canopy higher than the face's own plus distance * I (Lapen & Martz 1993,
 * fetchr.cpp:88-112); the first hit sets fetch = distance, else fetch =
 * max_distance; faces with a canopy taller than 1 m get fetch = 0
 * (fetchr.cpp:66-77).
 *
 * The search uses centroids in projected (UTM) coordinates; geographic meshes
 * (the reference's point_from_bearing_latlong branch) are not supported.
 * Nearest-face ties (equal squared distances) go to the lower face index. */
#ifndef PBSM3D_FETCH_H
#define PBSM3D_FETCH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pbsm3d_fetch_config {
    int32_t steps;        /* 10   (fetchr.cpp:36)                               */
    int32_t incl_veg;     /* 1    (fetchr.cpp:45): canopy heights count         */
    double max_distance;  /* 1000 (fetchr.cpp:38)                               */
    double I;             /* 0.06 (fetchr.cpp:43): rise per metre of distance   */
} pbsm3d_fetch_config;

typedef struct pbsm3d_fetch_ctx pbsm3d_fetch_ctx;

void pbsm3d_fetch_config_default(pbsm3d_fetch_config* cfg);

/* Host arrays of face centroids (x, y, z); canopy_height may be NULL (no face
 * has vegetation: has_vegetation() false everywhere), else one value per face.
 * Builds the centroid grid on the device.  Returns 0 or a negative error code
 * (the PBSM3D_E* codes of pbsm3d.h). */
int pbsm3d_fetch_create(pbsm3d_fetch_ctx** out, int device, int64_t n_faces, const double* center_x,
                        const double* center_y, const double* center_z, const double* canopy_height,
                        const pbsm3d_fetch_config* cfg);

/* fetch[i] for vw_dir[i] (degrees, the direction the wind blows from).
 * on_device = 1: both are device pointers (work is queued on the context's
 * stream and completed before return); 0: host arrays. */
int pbsm3d_fetch_run(pbsm3d_fetch_ctx* ctx, const double* vw_dir, double* fetch, int on_device);

/* The nearest-face query alone (exposed for testing): idx[k] = face whose
 * centroid is nearest to (qx[k], qy[k]), ties to the lower index, -1 for a
 * non-finite query point.  Host arrays.  (fetch for a non-finite vw_dir is
 * max_distance.) */
int pbsm3d_fetch_nearest(pbsm3d_fetch_ctx* ctx, int64_t n, const double* qx, const double* qy, int32_t* idx);

void pbsm3d_fetch_destroy(pbsm3d_fetch_ctx* ctx);
/* ctx NULL: the last pbsm3d_fetch_create error */
const char* pbsm3d_fetch_last_error(const pbsm3d_fetch_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* PBSM3D_FETCH_H */
