// wind_oracle.cpp -- CPU FP64 restatement of CHM's scale_wind_vert module
// (domain mode), the producer of PBSM3D's U_2m_above_srf input.
//
// TEST INFRASTRUCTURE ONLY: the checker of chm_amd/csrc/wind.hip.
//
//   * scale_wind_vert::point_scale   src/modules/scale_wind_vert.cpp:56-155
//   * Atmosphere::log_scale_wind     src/physics/Atmosphere.cpp:32-38 (z0 = Snow::Z0_SNOW = 0.01, Snow.h:31)
//   * Atmosphere::exp_scale_wind     src/physics/Atmosphere.cpp:41-46 (Inoue 1963)
//   * module_base::is_nan            src/modules/module_base.hpp:471-479 (-9999 or NaN)
//   * scale_wind_vert::run(mesh&)    src/modules/scale_wind_vert.cpp:182-242: point scaling
//     of every face, then each face takes the thin-plate-spline value of its
//     neighbours' point-scaled winds at its own centroid, floored at 0.1
//   * thin_plate_spline::operator()  src/interpolation/TPSpline.cpp:39-180 with
//     c = 0.577215, weight = 0.01 (TPSpline.cpp:196-201) and the LU refactored on
//     every call (reuse_LU ends false, :186-192): rows i < n: [1, R(i, 0..n-1)],
//     row n: [0, 1 ... 1], R(i, j) = -(ln d + c + E1(d)), d = (|p_i - p_j| w / 2)^2,
//     R(i, i) = 0; value = a_0 + sum_i a_i R(p_i, query).
//   * gsl_sf_expint_E1 (GSL, absent here; parity unpinned against GSL itself):
//     restated from Abramowitz & Stegun -- the power series 5.1.11 for x <= 1,
//     the continued fraction 5.1.22 (evaluated with Lentz's method) above; pinned
//     against scipy.special.exp1 in the tests.
//   * the small dense solve: Gaussian elimination with partial pivoting (the
//     reference uses Eigen's LU; the two agree to rounding).

#include <cmath>
#include <cstdint>

#include "../include/pbsm3d_wind.h"

namespace {

constexpr double kEulerGamma = 0.57721566490153286061;

double expint_e1(double x) {
    if (!(x > 0)) return x == 0 ? INFINITY : NAN;
    if (x <= 1.0) {
        // E1(x) = -gamma - ln x - sum_{k>=1} (-x)^k / (k k!)
        double sum = 0, term = 1;  // term = (-x)^k / k!
        for (int k = 1; k < 60; ++k) {
            term *= -x / k;
            const double add = term / k;
            sum += add;
            if (std::fabs(add) < 1e-18 * std::fabs(sum)) break;
        }
        return -kEulerGamma - std::log(x) - sum;
    }
    // E1(x) = e^-x * 1/(x+1- 1^2/(x+3- 2^2/(x+5- ...)))  (A&S 5.1.22, even form), Lentz
    const double tiny = 1e-300;
    double b = x + 1.0, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i < 1000; ++i) {
        const double an = -(double)i * i;
        b += 2.0;
        d = 1.0 / (an * d + b);
        c = b + an / c;
        const double del = c * d;
        h *= del;
        if (std::fabs(del - 1.0) < 1e-16) break;
    }
    return h * std::exp(-x);
}

double log_scale_wind(double u, double Z_in, double Z_out, double snowdepthavg, double z0) {
    const double Z0_SNOW = z0;
    u = u * std::log((Z_out - (snowdepthavg + Z0_SNOW)) / Z0_SNOW) / std::log((Z_in - (snowdepthavg + Z0_SNOW)) / Z0_SNOW);
    return u;
}

double exp_scale_wind(double u, double Z_in, double Z_out, double alpha) {
    u = u * std::exp(alpha * (Z_out / Z_in - 1));
    return u;
}

bool is_nan(double v) { return std::fabs(v - -9999.0) < 1e-5 || std::isnan(v); }

double point_scale(double U_R, double snowdepthavg, bool veg, double Z_CanTop_in, double LAI,
                   const pbsm3d_wind_config& cfg) {
    const double Z_R = 50.0;  // Atmosphere::Z_U_R
    const bool ignore_canopy = cfg.ignore_canopy != 0;
    double Z_CanTop = 0;
    if (!ignore_canopy && veg) Z_CanTop = Z_CanTop_in;
    const double Z_CanBot = Z_CanTop / 2.0;
    if (is_nan(snowdepthavg)) snowdepthavg = 0.0;
    const double Z_2m_above_srf = snowdepthavg + 2.0;
    double U_2m_above_srf;
    if (Z_2m_above_srf >= Z_R) return U_R;
    if (!ignore_canopy && (Z_CanTop > 0.0) && (Z_2m_above_srf < Z_CanTop)) {
        const double alpha = LAI;
        if (snowdepthavg < Z_CanTop) {
            const double U_CanTop = log_scale_wind(U_R, Z_R, Z_CanTop, snowdepthavg, cfg.z0_snow);
            const double U_CanBot = exp_scale_wind(U_CanTop, Z_CanTop, Z_CanBot, alpha);
            if (Z_2m_above_srf < Z_CanBot)
                U_2m_above_srf = log_scale_wind(U_CanBot, Z_CanBot, Z_2m_above_srf, snowdepthavg, cfg.z0_snow);
            else
                U_2m_above_srf = exp_scale_wind(U_CanTop, Z_CanTop, Z_2m_above_srf, alpha);
        } else {
            U_2m_above_srf = log_scale_wind(U_R, Z_R, Z_2m_above_srf, snowdepthavg, cfg.z0_snow);
        }
    } else {
        U_2m_above_srf = log_scale_wind(U_R, Z_R, Z_2m_above_srf, snowdepthavg, cfg.z0_snow);
    }
    return std::fmax(0.1, U_2m_above_srf);
}

double tps_basis(double sx, double sy, double ex, double ey) {
    const double c = 0.577215, weight = 0.01;
    const double xdiff = sx - ex, ydiff = sy - ey;
    double dij = std::sqrt(xdiff * xdiff + ydiff * ydiff);
    dij = (dij * weight / 2.0) * (dij * weight / 2.0);
    return -(std::log(dij) + c + expint_e1(dij));
}

// n <= 3 samples (x, y, z); value at (qx, qy)
double tps(int n, const double* sx, const double* sy, const double* sz, double qx, double qy) {
    const int m = n + 1;
    double A[4][5] = {};  // augmented [A | b]
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            const double xdiff = sx[i] - sx[j], ydiff = sy[i] - sy[j];
            if (xdiff == 0. && ydiff == 0.) continue;
            const double Rd = tps_basis(sx[i], sy[i], sx[j], sy[j]);
            A[i][j + 1] = Rd;
            A[j][i + 1] = Rd;
        }
    for (int i = 0; i < m; ++i) {
        A[i][0] = 1;
        A[m - 1][i] = 1;
    }
    A[m - 1][0] = 0;
    for (int i = 0; i < n; ++i) A[i][m] = sz[i];
    A[m - 1][m] = 0.0;
    // Gaussian elimination, partial pivoting
    for (int k = 0; k < m; ++k) {
        int p = k;
        for (int r = k + 1; r < m; ++r)
            if (std::fabs(A[r][k]) > std::fabs(A[p][k])) p = r;
        if (p != k)
            for (int cc = 0; cc <= m; ++cc) std::swap(A[k][cc], A[p][cc]);
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
    for (int i = 1; i < m; ++i) z0 = z0 + x[i] * tps_basis(sx[i - 1], sy[i - 1], qx, qy);
    return z0;
}

}  // namespace

extern "C" {

double oracle_expint_e1(double x) { return expint_e1(x); }

void oracle_scale_wind(int64_t F, const int32_t* neigh, const double* cx, const double* cy, const double* canopy,
                       const double* lai, const pbsm3d_wind_config* cfg, const double* U_R,
                       const double* snowdepthavg, double* out) {
    double* tmp = new double[F];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F; ++i)
        tmp[i] = point_scale(U_R[i], snowdepthavg ? snowdepthavg[i] : 0.0, canopy != nullptr,
                             canopy ? canopy[i] : 0.0, lai ? lai[i] : 0.0, *cfg);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F; ++i) {
        if (!cfg->smooth) {
            out[i] = tmp[i];
            continue;
        }
        double sx[3], sy[3], sz[3];
        int n = 0;
        for (int j = 0; j < 3; ++j) {
            const int32_t nb = neigh[3 * i + j];
            if (nb < 0) continue;
            sx[n] = cx[nb];
            sy[n] = cy[nb];
            sz[n] = tmp[nb];
            ++n;
        }
        out[i] = n > 0 ? std::fmax(0.1, tps(n, sx, sy, sz, cx[i], cy[i])) : std::fmax(0.1, tmp[i]);
    }
    delete[] tmp;
}

}  // extern "C"
