// Per-face coefficient, suspension-assembly, post-processing, deposition and
// drift kernels of the PBSM3D step (src/modules/PBSM3D.cpp:400-1748).
//
// One thread owns one face and walks its nLayer cells bottom to top, as the
// reference's face loop does; all per-face state (u*, z0, hs, c_salt, the wind
// direction, the geometry of the three edges) stays in registers and every
// CSR row value of the column is produced in one pass.  Row r = z*F + face is
// layer-major, so consecutive threads write consecutive rows.
//
// The arithmetic keeps the reference's expression order and is compiled with
// -ffp-contract=off, so coefficient and matrix values match a scalar CPU FP64
// evaluation up to the libm/ocml ulp differences of pow/log/exp/cos/sin.
#include "kernels.hpp"

namespace pb {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kKappa = 0.4;      // PhysConst.h:31
constexpr double kRhoIce = 917.0;   // PhysConst.h:39
constexpr double kZUR = 50.0;       // Atmosphere.h:31
constexpr double kLMax = 40.0;      // PBSM3D.cpp:227

// stdDryAirDensity of MeteoIO: standard-atmosphere pressure / (R_dry * T)
__device__ __forceinline__ double dry_air_density(double z, double T) {
    const double lapse = 0.0065, R0 = 6356766.0, Rd = 287.058;
    const double expo = 9.80665 / (lapse * Rd);
    const double p = 101325.0 * pow(1. - ((lapse * R0 * z) / (288.15 * (R0 + z))), expo);
    return p / (Rd * T);
}

// -bearing_to_cartesian(b): the direction the wind blows towards (coordinates.cpp:112-131)
__device__ __forceinline__ void downwind(double bearing, double& x, double& y) {
    double h = 450. - bearing;
    if (h > 360.) h = h - 360.;
    const double th = h * kPi / 180.;
    x = -cos(th);
    y = -sin(th);
}

// u* of a saltation candidate with the z0 feedback of Li & Pomeroy 2000 eq. 5
// (PBSM3D.cpp:691-716): the root of g(u) = u2 k / ln(2 / (c u^2 + lambda/2)) - u.
// The bracket is searched as boost::math::tools::bracket_and_solve_root does with
// guess 1, factor 1, rising = false and 500 evaluations: from g(1) >= 0 the
// bracket walks up (u *= factor), else down (u /= factor); the factor doubles
// after 32 evaluations, then after 16, 8, 4, 2 and at every evaluation after
// that; a downward walk below the smallest normal number returns the bracket
// (0, u), and running out of evaluations is a failure (no saltation).  Inside the
// bracket the root is taken to full precision by bisection: the reference's
// TOMS 748 stops when the bracket is narrower than 1e-8 and takes its midpoint,
// which lies within 5e-9 of the same root.  ok = false: no bracket.
__device__ double ustar_coupled(double u2, double lambda, bool& ok) {
    auto g = [&](double u) { return u2 * kKappa / log(2.0 / (0.6131702345e-2 * u * u + .5 * lambda)) - u; };
    auto sgn = [](double v) { return v == 0 ? 0 : (signbit(v) ? -1 : 1); };  // boost::math::sign
    const int max_iter = 500;
    double a = 1.0, b = 1.0, fa = g(a), fb = fa, factor = 1.0;
    int count = max_iter - 1, step = 32;
    ok = true;
    if (!(fa < 0)) {  // (fa < 0) == rising with rising = false: the root lies above
        while (sgn(fb) == sgn(fa)) {
            if (count == 0) { ok = false; return 0.0; }
            if ((max_iter - count) % step == 0) {
                factor *= 2;
                if (step > 1) step /= 2;
            }
            a = b;
            fa = fb;
            b *= factor;
            fb = g(b);
            --count;
        }
    } else {
        while (sgn(fb) == sgn(fa)) {
            if (fabs(a) < 2.2250738585072014e-308) return a / 2.0;  // bracket (0, a)
            if (count == 0) { ok = false; return 0.0; }
            if ((max_iter - count) % step == 0) {
                factor *= 2;
                if (step > 1) step /= 2;
            }
            b = a;
            fb = fa;
            a /= factor;
            fa = g(a);
            --count;
        }
    }
    if (fa == 0) return a;
    if (fb == 0) return b;
    // bisection on [a, b] (a < b, g(a) and g(b) of opposite signs) to the last bit
    double lo = a, hi = b, glo = fa;
    for (int it = 0; it < 200; ++it) {
        const double mid = lo + (hi - lo) / 2.0;
        if (mid <= lo || mid >= hi) break;
        const double gm = g(mid);
        if (gm == 0) return mid;
        if (sgn(gm) == sgn(glo)) {
            lo = mid;
            glo = gm;
        } else {
            hi = mid;
        }
    }
    return lo + (hi - lo) / 2.0;
}

}  // namespace

__global__ void __launch_bounds__(kBlock) k_assemble(AssembleArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.F) return;
    const int64_t F = a.F;
    const int L = a.L;
    const double dz = a.dz;

    // ---------------- per-face forcing (PBSM3D.cpp:436-475)
    double fetch = 1000;
    if (a.use_exp_fetch || a.use_tanh_fetch) fetch = a.fetch ? a.fetch[i] : 1000.0;
    const double uref = a.U_R[i];
    double sd = a.snowdepthavg[i];
    sd = chm_missing(sd) ? 0 : sd;
    const double u2 = a.U_2m[i];
    double swe = a.swe[i];
    swe = chm_missing(swe) ? 0 : swe;
    double hdiff = 0;
    if (a.enable_veg) hdiff = fmax(0.0, a.canopy[i] - sd);

    // ---------------- u* and saltation (PBSM3D.cpp:647-925)
    const double T = a.t[i];
    const double uth = 0.35 + (1.0 / 150.0) * T + (1.0 / 8200.0) * T * T;
    bool salt = false;
    const double z0_snow = a.z0_face ? a.z0_face[i] : a.z0_snow;
    double ustar = 1.3, lambda = 0, z0 = z0_snow;
    const bool water = a.is_water ? a.is_water[i] != 0 : false;
    if (hdiff <= a.cutoff && sd >= a.min_sd_trans && !water) {
        if (a.use_R94_lambda)
            lambda = 0.5 * (a.enable_veg ? a.lai[i] : 0.0) * hdiff;
        else
            lambda = (a.enable_veg ? a.stalk_n[i] * a.stalk_d[i] : 0.0) * hdiff;
        if (a.z0_ustar_coupling) {
            // a failed search leaves u* at its placeholder 1.3, which the threshold test below
            // still sees (PBSM3D.cpp:647, 710-714)
            bool root_ok = true;
            const double r = ustar_coupled(u2, lambda, root_ok);
            if (root_ok) ustar = r;
        } else {
            ustar = u2 * kKappa / log(2.0 / 0.0002);
        }
        if (ustar >= uth) {
            salt = true;
            if (a.z0_ustar_coupling) z0 = 0.6131702345e-2 * ustar * ustar + .5 * lambda;  // Li & Pomeroy eq. 4
        }
    }
    if (!salt) {
        z0 = z0_snow;
        ustar = fmax(0.01, kKappa * uref / log(kZUR / z0));
    }
    z0 = fmax(z0_snow, z0);
    ustar = fmax(0.01, ustar);
    const double hs = salt ? 0.08436 * pow(ustar, 1.27) : 0.0;

    double vx, vy;
    downwind(a.vw_dir[i], vx, vy);
    double nx[3], ny[3], el[3];
    int32_t nb[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        nx[j] = a.nx[j * F + i];
        ny[j] = a.ny[j * F + i];
        el[j] = a.elen[j * F + i];
        nb[j] = a.neigh[j * F + i];
    }
    const double area = a.area[i];

    double Qsalt = 0, c_salt = 0;
    const double tk = T + 273.15;
    if (salt) {
        const double rho_f = dry_air_density(a.cz[i], tk);
        const double mb = 0.16 * 202.0;
        const double tau_n = (mb * lambda) / (1.0 + mb * lambda);
        c_salt = rho_f / (3.29 * ustar) * (1.0 - tau_n - (uth * uth) / (ustar * ustar));
        if (c_salt < 0 || isnan(c_salt)) {
            c_salt = 0;
            salt = false;
        }
        if (a.use_exp_fetch && fetch < 500)
            c_salt *= 1.0 - exp(-3.0 * fetch / 500.0);
        else if (a.use_tanh_fetch && fetch <= 300.)
            c_salt *= 0.5 * tanh(0.1333333333e-1 * 300.0 - 2.0) + 0.5;  // uses fetch_ref, PBSM3D.cpp:839-845
        if (a.use_PomLi_probability) {
            // probability of blowing snow, Essery, Li & Pomeroy 1999 eqs 10-14 (PBSM3D.cpp:848-865),
            // with the 10 m wind log-scaled from the reference height (PBSM3D.cpp:451-463)
            const double z10 = 10. + sd;
            const double u10 = z10 < kZUR ? uref * log((z10 - (sd + z0_snow)) / z0_snow) /
                                                log((kZUR - (sd + z0_snow)) / z0_snow)
                                          : uref;
            const double A = a.p_snow_hours[i];
            const double u_mean = 11.2 + 0.365 * T + 0.00706 * T * T + 0.9 * log(A);
            const double delta = 0.145 * T + 0.00196 * T * T + 4.3;
            const double Nst = (a.enable_veg && !a.use_R94_lambda) ? a.stalk_n[i] : 0.0;
            const double dst = (a.enable_veg && !a.use_R94_lambda) ? a.stalk_d[i] : 0.0;
            const double z0v = (Nst * dst * hdiff) / 2.0;
            const double us = u10 / sqrt((1 + 340.0 * z0v));
            const double Pu10 = 1.0 / (1.0 + exp((sqrt(kPi) * (u_mean - us)) / delta));
            a.blowing_prob[i] = Pu10;
            c_salt *= Pu10;
        }
        const double uhs = 2.8 * uth;
        Qsalt = c_salt * uhs * hs;
        double mass = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) mass += -el[j] * Qsalt * (vx * nx[j] + vy * ny[j]);
        mass = mass / area * a.dt;
        if (mass < 0 && fabs(mass) > swe) {
            c_salt = 0;
            Qsalt = c_salt * uhs * hs;
        }
    }
    a.Qsalt[i] = Qsalt;
    a.salt_flag[i] = salt ? 1 : 0;
    if (a.fscal) {
        a.fscal[i] = c_salt;
        a.fscal[F + i] = hs;
        a.fscal[2 * F + i] = ustar;
    }

    // ---------------- per-layer coefficients + column assembly (PBSM3D.cpp:927-1404)
    const double rh = a.rh[i] / 100.;
    const double TA = tk - 273.15;
    const double es = 611.21 * exp((17.502 * TA) / (240.97 + TA));  // water branch: tk >= 0 always
    const double vnorm = sqrt(vx * vx + vy * vy);
    double Alat[3], alat[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        Alat[j] = el[j] * dz;
        alat[j] = a.do_lateral_diff ? Alat[j] * 0.00001 : 0.0;
    }
    const uint32_t slot = a.slots[i];
    const int pos_self = slot & 0xF;
    const int nlat = 1 + (nb[0] >= 0) + (nb[1] >= 0) + (nb[2] >= 0);
    double Vc = area * dz;
    Vc /= 5.0;  // the sink is added five times per row (PBSM3D.cpp:1222-1226)

    // running Thomas factorisation of the column (for the column preconditioner)
    double cp_prev = 0.0;

    for (int z = 0; z < L; ++z) {
        const double czl = z * dz + hs + dz / 2.;
        const double hz = czl + sd;
        double u_z;
        if (salt && czl < hdiff)
            u_z = 2.8 * uth;
        else if (czl < hdiff)
            u_z = 0.01;
        else if (hz < kZUR)
            u_z = fmax(0.01, uref * log((hz - (sd + z0)) / z0) / log((kZUR - (sd + z0)) / z0));
        else
            u_z = fmax(0.01, uref);

        const double rm = 4.6e-5 * pow(czl, -0.258);
        const double mma = 4.08 + 12.6 * czl;
        const double mm = 4. / 3. * kPi * kRhoIce * rm * rm * rm * (1.0 + 3.0 / mma + 2. / (mma * mma));
        const double r_z = pow((3.0 * mm) / (4 * kPi * kRhoIce), 0.3333333);
        const double xrz = 0.005 * pow(u_z, 1.36);
        double w = a.do_fixed_settling ? a.settling_velocity : 1.1e7 * pow(r_z, 1.8);
        w *= a.settle_face ? a.settle_face[i] : a.settling_scale;
        const double Vr = w + 3.0 * xrz * cos(kPi / 4.0);
        const double Re = 2.0 * r_z * Vr / 1.88e-5;
        const double Nu = 1.79 + 0.606 * pow(Re, 0.5);
        const double D = 2.06e-5 * pow(tk / 273.15, 1.75);
        const double lam = 0.000063 * tk + 0.00673;
        const double Ls = 2.838e6, M = 18.01, R = 8313;
        const double sigma = (rh - 1.0) * (1.019 + 0.027 * log(czl));
        const double rho = (M * es) / (R * tk);
        const double Qr = 0.9 * kPi * rm * rm * 120.0;
        const double dmdt = Nu * rho * D * (6.283185308 * Nu * R * r_z * sigma * tk * tk * lam - Ls * M * Qr + Qr * R * tk) /
                            (D * Ls * Nu * (Ls * M - R * tk) * rho + lam * tk * tk * Nu * R);
        const double csubl = a.do_sublimation ? dmdt / mm : 0.0;

        const double l = kKappa * (czl + z0) * kLMax / (kKappa * (czl + z0) + kLMax);
        // Rouault et al. 1991 coefficient in place of snow_diffusion_const (PBSM3D.cpp:1164-1172)
        const double dcoef = a.rouault_diffusion_coef ? 1.0 / (1.0 + (1.0 * w * w) / (1.56 * ustar * ustar))
                                                      : a.snow_diffusion_const;
        const double Kv = dcoef * ustar * l;
        const double a3 = area * Kv / dz;  // alpha[3] == alpha[4]
        const double ks = u_z / vnorm;
        const double ux = vx * ks, uy = vy * ks;
        const double Vs = Vc * csubl;

        // ---- lateral faces
        double diag = 0.0;
        double off[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const double um = ux * nx[f] + uy * ny[f];
            if (um > 0) {
                if (nb[f] >= 0) {
                    diag += (Vs - Alat[f] * um - alat[f]);
                    off[f] = 0.0 + alat[f];
                } else {
                    diag += (-0.1e-1 * alat[f] - 1. * Alat[f] * um + csubl * Vc);
                }
            } else {
                if (nb[f] >= 0) {
                    diag += Vs - alat[f];
                    off[f] = 0.0 + (-Alat[f] * um + alat[f]);
                } else {
                    diag += -0.1e-1 * alat[f] - .99 * Alat[f] * um + csubl * Vc;
                }
            }
        }
        // ---- vertical faces: udotm[3] = -w (top), udotm[4] = +w (bottom)
        double below = 0.0, above = 0.0, rhs = 0.0;
        const bool top_out = -w > 0;
        const bool bot_out = w > 0;
        if (z == 0) {
            const double a4 = area * Kv / (hs / 2.0 + dz / 2.0);
            diag += Vs - area * w - a4;
            rhs += -a4 * c_salt;
            if (top_out) {
                diag += Vs - area * (-w) - a3;
                above = 0.0 + a3;
            } else {
                diag += Vs - a3;
                above = 0.0 + (-area * (-w) + a3);
            }
        } else if (z == L - 1) {
            diag += top_out ? Vs - area * (-w) - a3 : Vs - a3;
            if (bot_out) {
                diag += Vs - area * w - a3;
                below = 0.0 + a3;
            } else {
                diag += Vs - a3;
                below = 0.0 + (-area * w + a3);
            }
        } else {
            if (top_out) {
                diag += Vs - area * (-w) - a3;
                above = 0.0 + a3;
            } else {
                diag += Vs - a3;
                above = 0.0 + (-area * (-w) + a3);
            }
            if (bot_out) {
                diag += Vs - area * w - a3;
                below = 0.0 + a3;
            } else {
                diag += Vs - a3;
                below = 0.0 + (-area * w + a3);
            }
        }

        // ---- scatter the row into its CSR slots: [below] [sorted lateral block] [above]
        const int64_t r = (int64_t)z * F + i;
        const int64_t base = (int64_t)a.rp[r] + (z > 0 ? 1 : 0);
        if (z > 0) a.val[base - 1] = below;
        a.val[base + pos_self] = diag;
#pragma unroll
        for (int f = 0; f < 3; ++f)
            if (nb[f] >= 0) a.val[base + ((slot >> (4 * (f + 1))) & 0xF)] = off[f];
        if (z < L - 1) a.val[base + nlat] = above;
        a.rhs[r] = rhs;
        a.x[r] = 0.0;
        a.u_z[r] = u_z;
        a.csubl[r] = csubl;

        if (a.pc_lo) {  // Thomas factors: m = d - l*c'_{z-1}, c' = u/m
            const double lo = z > 0 ? below : 0.0;
            const double m = diag - lo * cp_prev;
            const double im = 1.0 / m;
            const double cp = (z < L - 1) ? above * im : 0.0;
            a.pc_lo[r] = lo;
            a.pc_im[r] = im;
            a.pc_cp[r] = cp;
            if (a.pc_lo32) {
                a.pc_lo32[r] = (float)lo;
                a.pc_im32[r] = (float)im;
                a.pc_cp32[r] = (float)cp;
            }
            cp_prev = cp;
        }
        if (a.pc_invdiag) a.pc_invdiag[r] = 1.0 / diag;
    }
}

// Qsusp / Qsubl / Qsubl_mass / sum_subl (PBSM3D.cpp:1465-1503)
__global__ void __launch_bounds__(kBlock) k_post(PostArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.F) return;
    double Qsusp = 0, Qsubl = 0;
    for (int z = 0; z < a.L; ++z) {
        const int64_t r = (int64_t)z * a.F + i;
        double c = a.x[r];
        c = c < 0 || chm_missing(c) ? 0 : c;
        Qsusp += c * a.u_z[r] * a.dz;
        Qsubl += a.csubl[r] * c * a.dz;
    }
    a.Qsusp[i] = Qsusp;
    a.Qsubl[i] = Qsubl;
    const double qm = Qsubl * a.dt;
    a.Qsubl_mass[i] = qm;
    a.sum_subl[i] += qm;
}

// Deposition system (PBSM3D.cpp:1516-1658): diag = area + sum eps*E/dx, off =
// -eps*E/dx, RHS = -sum_j E_j (Qsusp+Qsalt)_upwind (u_hat . m_j).
__global__ void __launch_bounds__(kBlock) k_deposition(DepositionArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.F) return;
    const int64_t F = a.F;
    double vx, vy;
    downwind(a.vw_dir[i], vx, vy);
    const uint32_t slot = a.slots[i];
    const int64_t base = a.rp[i];
    double diag = a.area[i];
    double rhs = 0.0;
    const double qt_self = a.Qsusp[i], qs_self = a.Qsalt[i];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const double um = vx * a.nx[j * F + i] + vy * a.ny[j * F + i];
        const double E = a.elen[j * F + i];
        const int32_t n = a.neigh[j * F + i];
        double qt = qt_self, qs = qs_self;
        if (!(um > 0) && n >= 0) {
            qt = a.Qsusp[n];
            qs = a.Qsalt[n];
            if (chm_missing(qs)) qs = 0;
        }
        if (n >= 0) {
            const double wj = a.dep_w[j * F + i];  // eps*E/dx, precomputed at init
            diag += wj;
            a.val[base + ((slot >> (4 * (j + 1))) & 0xF)] = 0.0 + (-wj);
        }
        rhs += -E * (qt + qs) * um;
    }
    a.val[base + (slot & 0xF)] = diag;
    a.rhs[i] = rhs;
    a.x[i] = 0.0;
    if (a.invdiag) a.invdiag[i] = 1.0 / diag;
}

// drift_mass / sum_drift / pbsm_more_than_avail (PBSM3D.cpp:1708-1740)
__global__ void __launch_bounds__(kBlock) k_drift(DriftArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.F) return;
    if (a.face_member && !a.member_on[a.face_member[i]]) return;  // that member had no deposition
    const double q = a.qdep[i];
    double mass = (chm_missing(q) ? 0 : q) * a.dt;
    double swe = a.swe[i];
    swe = chm_missing(swe) ? 0 : swe;
    if (mass < 0 && fabs(mass) > swe) {
        a.more[i] = 1;
        mass = -swe;
    }
    if (mass < 0 && !a.salt_flag[i]) mass = 0;
    a.drift_mass[i] = mass;
    a.sum_drift[i] += mass;
}

}  // namespace pb
