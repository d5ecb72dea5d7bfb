// fetch_oracle.cpp -- CPU FP64 restatement of CHM's fetchr module.
//
// TEST INFRASTRUCTURE ONLY: the checker of the GPU fetch kernel
// (chm_amd/csrc/fetch.hip).  Nothing in the product links or calls it.
//
//   * fetchr::fetchr   src/modules/fetchr.cpp:28-50 (steps 10, max_distance 1000,
//                      I 0.06, incl_veg true, size_of_step = max_distance / steps)
//   * fetchr::run      src/modules/fetchr.cpp:55-118
//   * face::find_closest_face(azimuth, distance)  src/mesh/triangulation.hpp:1543-1546
//   * point_from_bearing_UTM                       src/math/coordinates.cpp:60-71
//   * triangulation::find_closest_face(Point_2)    src/mesh/triangulation.cpp:170-178
//     (CGAL k = 1 neighbour search on face centroids): restated as a brute-force
//     scan over all centroids with the squared distance dx*dx + dy*dy; exact
//     ties go to the lower face index (CGAL's tie order is parity unpinned).
//   * face::has_vegetation / veg_attribute         src/mesh/triangulation.hpp:1656-1696
//     (here: canopy_height given for every face, or for none)
//
// Built with -ffp-contract=off like the rest of the oracle.

#include <cmath>
#include <cstdint>
#include <limits>

#include "../include/pbsm3d_fetch.h"

namespace {

int64_t nearest(int64_t F, const double* cx, const double* cy, double qx, double qy) {
    double best = std::numeric_limits<double>::infinity();
    int64_t bi = -1;
    for (int64_t f = 0; f < F; ++f) {
        const double dx = qx - cx[f];
        const double dy = qy - cy[f];
        const double d2 = dx * dx + dy * dy;
        if (d2 < best) {
            best = d2;
            bi = f;
        }
    }
    return bi;
}

}  // namespace

extern "C" {

void oracle_fetch_nearest(int64_t F, const double* cx, const double* cy, int64_t n, const double* qx,
                          const double* qy, int32_t* idx) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < n; ++k) idx[k] = (int32_t)nearest(F, cx, cy, qx[k], qy[k]);
}

void oracle_fetch(int64_t F, const double* cx, const double* cy, const double* cz, const double* canopy,
                  const pbsm3d_fetch_config* cfg, const double* vw_dir, double* fetch) {
    const int steps = cfg->steps;
    const double max_distance = cfg->max_distance;
    const double size_of_step = max_distance / steps;
    const double I = cfg->I;
    const bool incl_veg = cfg->incl_veg != 0;
    const bool has_veg = canopy != nullptr;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < F; ++i) {
        fetch[i] = max_distance;
        const double wind_dir = vw_dir[i];
        if (incl_veg && has_veg) {
            const double me_Z_CanTop = canopy[i];
            if (me_Z_CanTop > 1) {
                fetch[i] = 0;
                continue;
            }
        }
        if (!std::isfinite(wind_dir)) continue;  // no upwind point (undefined in the reference): max_distance
        for (int j = 1; j <= steps; ++j) {
            const double distance = j * size_of_step;
            // point_from_bearing_UTM(center, wind_dir, distance)
            const double DegreesToRadians = M_PI / 180.0;
            const double bearing = wind_dir * DegreesToRadians;
            const double dist_x = distance * std::sin(bearing);
            const double dist_y = distance * std::cos(bearing);
            const int64_t f = nearest(F, cx, cy, cx[i] + dist_x, cy[i] + dist_y);
            double Z_CanTop = 0;
            if (incl_veg && has_veg) Z_CanTop = canopy[f];
            const double Z_test = cz[f] + Z_CanTop;
            const double Z_core = cz[i] + distance * I;
            const double z0_1 = 0.12 * Z_CanTop;
            const double z0_2 = 0.001;
            const double n = 1.0 / 0.8;
            const double h = 5;
            const double x_sss = std::pow(((33.33333333 * h - 25. * z0_2) / (std::log(z0_1 / z0_2) * z0_2)), n) * z0_2;
            if (Z_test >= Z_core || (incl_veg && distance < x_sss)) {
                fetch[i] = distance;
                break;
            }
        }
    }
}

}  // extern "C"
