// pbsm3d_oracle.cpp -- CPU FP64 restatement of the reference PBSM3D step.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP path is checked
// against and the CPU baseline that bench.py times beside it.  Nothing in the
// product (chm_amd/, include/) links or calls it; only tests/, the smoke check
// and the CPU-baseline leg of bench.py load it.
//
// It restates, in plain scalar C++ with the reference's expression order, so
// that floating-point results agree to the last bit where the arithmetic is the
// same:
//   * face geometry     src/mesh/triangulation.hpp:1443-1498 (edge, edge_unit_normal,
//                       edge_length), :1577-1589 (center), :1774-1780 (get_z),
//                       :1830-1856 (get_area)
//   * physics helpers   src/physics/Atmosphere.cpp:32-37 (log_scale_wind),
//                       :62-80 (saturatedVapourPressure); src/math/coordinates.cpp:100-131
//   * PBSM3D::init      src/modules/PBSM3D.cpp:221-398
//   * PBSM3D::run       src/modules/PBSM3D.cpp:400-1748
//   * NearestNeighborProblem src/math/LinearAlgebra.cpp:31-275 (CSR pattern with
//                       sorted columns, zeroSystem, sumInto, Solve, getRhsMax)
//
// Third-party arithmetic restated from published algorithms:
//   * MeteoIO Atmosphere::stdDryAirDensity (PBSM3D.cpp:784-786): standard
//     atmosphere pressure over (R_dry * T).  MeteoIO is not in the reference
//     tree; the restatement is pinned against the U.S. Standard Atmosphere 1976
//     table (pressure and density at geometric altitudes 0 - 5 km and at the
//     11 km geopotential tropopause) in tests/test_oracle.py.
// Parity unpinned (nothing in the reference or in published tables fixes the
// iterates; only the stopping rule is matched, see the tests' docstrings):
//   * Belos GMRES (restart 30, max 1000 iterations, tol 1e-8 on ||b-Ax||/||b||,
//     right preconditioning) -- restarted GMRES of Saad & Schultz (1986) with
//     classical Gram-Schmidt plus one re-orthogonalisation pass.
//   * Ifpack2 ILUT (drop tolerance 1e-4, level-of-fill 3.0) -- ILUT(p, tau) of
//     Saad (1994) as in his SPARSKIT ilut: multipliers dropped when |l_ik| <= tau,
//     U entries when |u_ij| <= tau*mean|a_ij| of the row; keep the largest (row count
//     + fill) entries in each of L and U, fill = ceil((lof-1)*nnz/(2n)) (Ifpack2 meaning).
//
// Built with -O3 -ffp-contract=off (no FMA contraction), mirroring the
// reference's -O3 -frounding-math build (CMakeLists.txt:204,214-220).

#include "../include/pbsm3d.h"

#include <algorithm>
#include <set>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <queue>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;  // M_PI
constexpr double kKappa = 0.4;                    // PhysConst::kappa, PhysConst.h:31
constexpr double kRhoIce = 917.0;                 // PhysConst::rho_ice, PhysConst.h:39
constexpr double kZUR = 50.0;                     // Atmosphere::Z_U_R, Atmosphere.h:31
constexpr double kSuspDepth = 5.0;                // PBSM3D.cpp:225
constexpr double kLMax = 40.0;                    // PBSM3D.cpp:227
constexpr double kPresentThreshold = 1e-12;       // PBSM3D.hpp:417-418

// MeteoIO standard-atmosphere constants (recalled published values)
constexpr double kStdPress = 101325.0;
constexpr double kLapse = 0.0065;
constexpr double kSeaLevelT = 288.15;
constexpr double kGravity = 9.80665;
constexpr double kRDry = 287.058;
constexpr double kEarthR0 = 6356766.0;

thread_local std::string g_err;

// module_base::is_nan (module_base.hpp:471-479): within 1e-5 of the -9999 default, or NaN
inline bool chm_is_nan(double v) { return std::fabs(v - -9999.0) < 1e-5 || std::isnan(v); }

double std_dry_air_density(double altitude, double temperature) {
    const double expo = kGravity / (kLapse * kRDry);
    const double p = kStdPress * std::pow(1. - ((kLapse * kEarthR0 * altitude) / (kSeaLevelT * (kEarthR0 + altitude))), expo);
    return p / (kRDry * temperature);
}

double saturated_vapour_pressure(double T) {  // T in K; the reference compares K with 0 C
    double TA = T - 273.15;
    if (T >= 0.)
        return 611.21 * std::exp((17.502 * TA) / (240.97 + TA));
    return 611.15 * std::exp((22.452 * TA) / (272.55 + TA));
}

double log_scale_wind(double u, double z_in, double z_out, double sd, double z0) {
    return u * std::log((z_out - (sd + z0)) / z0) / std::log((z_in - (sd + z0)) / z0);
}

// boost::math::sign: 0, or -1 / +1 from the sign bit
int boost_sign(double v) { return v == 0 ? 0 : (std::signbit(v) ? -1 : 1); }

// u* with the z0 feedback of Li & Pomeroy 2000 eq. 5 (PBSM3D.cpp:691-716): the root
// of g(u) = u2 k / ln(2 / (0.6131702345e-2 u^2 + lambda/2)) - u, bracketed as
// boost::math::tools::bracket_and_solve_root(g, guess 1, factor 1, rising false,
// max_iter 500) brackets it (Boost.Math's documented algorithm: walk up from the guess
// when g(guess) >= 0, else down, multiplying / dividing by the factor, which doubles
// after 32 evaluations and then ever sooner; a downward walk below DBL_MIN returns
// (0, a); running out of evaluations throws, *ok = false).  The reference then runs
// TOMS 748 to a bracket narrower than 1e-8 and takes its midpoint; the restatement
// bisects the bracket to the last bit (the two differ by < 5e-9: parity with Boost's
// iterates is unpinned, the root itself is pinned in tests/test_oracle.py).
double ustar_li_pomeroy(double u2, double lambda, bool* ok) {
    auto g = [&](double u) { return u2 * kKappa / std::log(2.0 / (0.6131702345e-2 * u * u + .5 * lambda)) - u; };
    const std::uintmax_t max_iter = 500;
    double factor = 1.0;
    double a = 1.0, b = 1.0, fa = g(a), fb = fa;
    std::uintmax_t count = max_iter - 1;
    int step = 32;
    *ok = true;
    if ((fa < 0) == false) {
        while (boost_sign(fb) == boost_sign(fa)) {
            if (count == 0) { *ok = false; return 0; }
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
        while (boost_sign(fb) == boost_sign(fa)) {
            if (std::fabs(a) < 2.2250738585072014e-308) return 0 + (a - 0) / 2.0;
            if (count == 0) { *ok = false; return 0; }
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
    double lo = a, hi = b, glo = fa;
    for (int it = 0; it < 200; ++it) {
        const double mid = lo + (hi - lo) / 2.0;
        if (mid <= lo || mid >= hi) break;
        const double gm = g(mid);
        if (gm == 0) return mid;
        if (boost_sign(gm) == boost_sign(glo)) {
            lo = mid;
            glo = gm;
        } else {
            hi = mid;
        }
    }
    return lo + (hi - lo) / 2.0;
}

void neg_bearing_to_cartesian(double bearing, double* x, double* y) {
    double h = 450. - bearing;
    if (h > 360.) h = h - 360.;
    double theta = h * kPi / 180.;
    *x = -std::cos(theta);
    *y = -std::sin(theta);
}

// ---------------------------------------------------------------------------
// CSR (row_ptr / col sorted ascending per row, values)
// ---------------------------------------------------------------------------
struct Csr {
    int64_t n = 0;
    std::vector<int32_t> rp, col;
    std::vector<double> val;

    int64_t find(int64_t r, int64_t c) const {
        const int32_t* b = col.data() + rp[r];
        const int32_t* e = col.data() + rp[r + 1];
        const int32_t* p = std::lower_bound(b, e, (int32_t)c);
        if (p == e || *p != c) return -1;
        return p - col.data();
    }
    void sum_into(int64_t r, int64_t c, double v) { val[find(r, c)] += v; }
    void replace(int64_t r, int64_t c, double v) { val[find(r, c)] = v; }
};

// LinearAlgebra.cpp:67-149: per row the list [self, lateral neighbours in edge
// order, below (z>0), above (z<L-1)]; Tpetra's fillComplete sorts columns.
void build_pattern(int64_t F, int L, const int32_t* neigh, Csr& A) {
    const int64_t N = F * L;
    A.n = N;
    A.rp.assign(N + 1, 0);
    for (int z = 0; z < L; ++z)
        for (int64_t i = 0; i < F; ++i) {
            int cnt = 1;
            for (int f = 0; f < 3; ++f) cnt += neigh[3 * i + f] >= 0;
            cnt += (z > 0) + (z < L - 1);
            A.rp[z * F + i + 1] = cnt;
        }
    for (int64_t r = 0; r < N; ++r) A.rp[r + 1] += A.rp[r];
    A.col.resize(A.rp[N]);
    for (int z = 0; z < L; ++z)
        for (int64_t i = 0; i < F; ++i) {
            const int64_t r = z * F + i;
            int32_t* c = A.col.data() + A.rp[r];
            int k = 0;
            c[k++] = (int32_t)r;
            for (int f = 0; f < 3; ++f)
                if (neigh[3 * i + f] >= 0) c[k++] = (int32_t)(z * F + neigh[3 * i + f]);
            if (z > 0) c[k++] = (int32_t)(r - F);
            if (z < L - 1) c[k++] = (int32_t)(r + F);
            std::sort(c, c + k);
        }
    A.val.assign(A.rp[N], 0.0);
}

// ---------------------------------------------------------------------------
// threading helpers
// ---------------------------------------------------------------------------
double dot(int64_t n, const double* a, const double* b) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
double nrm2(int64_t n, const double* a) { return std::sqrt(dot(n, a, a)); }

void spmv(const Csr& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < A.n; ++r) {
        double s = 0;
        for (int32_t k = A.rp[r]; k < A.rp[r + 1]; ++k) s += A.val[k] * x[A.col[k]];
        y[r] = s;
    }
}

// ---------------------------------------------------------------------------
// preconditioners (right preconditioning: z = M^-1 r)
// ---------------------------------------------------------------------------
struct Precond {
    virtual ~Precond() = default;
    virtual void compute(const Csr& A) = 0;
    virtual void apply(const double* r, double* z) const = 0;
};

struct PcNone : Precond {
    int64_t n = 0;
    void compute(const Csr& A) override { n = A.n; }
    void apply(const double* r, double* z) const override { std::memcpy(z, r, n * sizeof(double)); }
};

struct PcJacobi : Precond {
    std::vector<double> inv;
    void compute(const Csr& A) override {
        inv.assign(A.n, 1.0);
        for (int64_t r = 0; r < A.n; ++r) {
            int64_t p = A.find(r, r);
            double d = p >= 0 ? A.val[p] : 0.0;
            inv[r] = d != 0.0 ? 1.0 / d : 1.0;
        }
    }
    void apply(const double* r, double* z) const override {
        const int64_t n = (int64_t)inv.size();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) z[i] = r[i] * inv[i];
    }
};

// Block Jacobi over vertical columns: for each face, the L x L tridiagonal
// formed by the diagonal and the (r, r-F) / (r, r+F) couplings, solved by the
// Thomas algorithm.
struct PcColumn : Precond {
    int64_t F;
    int L;
    std::vector<double> lo, di, up;
    PcColumn(int64_t F_, int L_) : F(F_), L(L_) {}
    void compute(const Csr& A) override {
        const int64_t N = F * L;
        lo.assign(N, 0.0);
        di.assign(N, 1.0);
        up.assign(N, 0.0);
        for (int64_t r = 0; r < N; ++r) {
            int64_t p = A.find(r, r);
            di[r] = A.val[p];
            if (r >= F) lo[r] = A.val[A.find(r, r - F)];
            if (r + F < N) up[r] = A.val[A.find(r, r + F)];
        }
    }
    void apply(const double* r, double* z) const override {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < F; ++i) {
            double cp[64], dp[64];
            // forward sweep
            double b0 = di[i];
            cp[0] = up[i] / b0;
            dp[0] = r[i] / b0;
            for (int k = 1; k < L; ++k) {
                const int64_t q = k * F + i;
                double m = di[q] - lo[q] * cp[k - 1];
                cp[k] = (k < L - 1) ? up[q] / m : 0.0;
                dp[k] = (r[q] - lo[q] * dp[k - 1]) / m;
            }
            z[(L - 1) * F + i] = dp[L - 1];
            for (int k = L - 2; k >= 0; --k) {
                dp[k] = dp[k] - cp[k] * dp[k + 1];
                z[k * F + i] = dp[k];
            }
        }
    }
};

// Incomplete LU with the sparsity of A (IKJ variant), unit lower L.
struct PcIlu0 : Precond {
    Csr LU;
    std::vector<int64_t> diag;
    void compute(const Csr& A) override {
        LU = A;
        const int64_t n = A.n;
        diag.assign(n, -1);
        for (int64_t r = 0; r < n; ++r) diag[r] = LU.find(r, r);
        std::vector<int64_t> pos(n, -1);
        for (int64_t i = 0; i < n; ++i) {
            for (int32_t k = LU.rp[i]; k < LU.rp[i + 1]; ++k) pos[LU.col[k]] = k;
            for (int32_t k = LU.rp[i]; k < LU.rp[i + 1]; ++k) {
                const int64_t c = LU.col[k];
                if (c >= i) break;
                double piv = LU.val[diag[c]];
                double lik = LU.val[k] / piv;
                LU.val[k] = lik;
                for (int32_t q = (int32_t)diag[c] + 1; q < LU.rp[c + 1]; ++q) {
                    int64_t p = pos[LU.col[q]];
                    if (p >= 0) LU.val[p] -= lik * LU.val[q];
                }
            }
            for (int32_t k = LU.rp[i]; k < LU.rp[i + 1]; ++k) pos[LU.col[k]] = -1;
        }
    }
    void apply(const double* r, double* z) const override {
        const int64_t n = LU.n;
        for (int64_t i = 0; i < n; ++i) {
            double s = r[i];
            for (int32_t k = LU.rp[i]; k < diag[i]; ++k) s -= LU.val[k] * z[LU.col[k]];
            z[i] = s;
        }
        for (int64_t i = n - 1; i >= 0; --i) {
            double s = z[i];
            for (int32_t k = (int32_t)diag[i] + 1; k < LU.rp[i + 1]; ++k) s -= LU.val[k] * z[LU.col[k]];
            z[i] = s / LU.val[diag[i]];
        }
    }
};

// ILUT(p, tau), Saad (1994).  Parameters default to the reference's Ifpack2
// setup (LinearAlgebra.cpp:178-188); oracle_set_ilut overrides them for tests.
double g_ilut_tau = 1e-4, g_ilut_lof = 3.0;
struct PcIlut : Precond {
    double tau = g_ilut_tau, lof = g_ilut_lof;
    std::vector<int64_t> lrp, urp;  // L strictly lower (unit diag implied); U with diagonal first
    std::vector<int32_t> lcol, ucol;
    std::vector<double> lval, uval;
    int64_t n = 0;

    void compute(const Csr& A) override {
        n = A.n;
        const double fill = std::ceil(((lof - 1.0) * (double)A.rp[n]) / (2.0 * (double)n));
        const int64_t fillx = fill > 0 ? (int64_t)fill : 0;
        lrp.assign(n + 1, 0);
        urp.assign(n + 1, 0);
        lcol.clear(); lval.clear(); ucol.clear(); uval.clear();
        lcol.reserve(A.rp[n] * 2); lval.reserve(A.rp[n] * 2);
        ucol.reserve(A.rp[n] * 2); uval.reserve(A.rp[n] * 2);
        std::vector<double> w(n, 0.0);
        std::vector<char> inw(n, 0);
        std::vector<int64_t> nz;
        std::priority_queue<int64_t, std::vector<int64_t>, std::greater<int64_t>> heap;
        std::vector<std::pair<double, int64_t>> keepL, keepU;
        for (int64_t i = 0; i < n; ++i) {
            nz.clear();
            double rn = 0;
            int64_t nl = 0, nu = 0;
            for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
                const int64_t c = A.col[k];
                w[c] = A.val[k];
                inw[c] = 1;
                nz.push_back(c);
                rn += std::fabs(A.val[k]);
                if (c < i) { heap.push(c); ++nl; }
                if (c > i) ++nu;
            }
            if (!inw[i]) { w[i] = 0.0; inw[i] = 1; nz.push_back(i); }
            const double thr = tau * rn / (double)(A.rp[i + 1] - A.rp[i]);
            keepL.clear();
            while (!heap.empty()) {
                const int64_t k = heap.top();
                heap.pop();
                while (!heap.empty() && heap.top() == k) heap.pop();
                const int64_t u0 = urp[k];
                double wk = w[k] / uval[u0];
                w[k] = wk;
                if (std::fabs(wk) <= tau) { w[k] = 0.0; continue; }
                keepL.emplace_back(std::fabs(wk), k);
                for (int64_t q = u0 + 1; q < urp[k + 1]; ++q) {
                    const int64_t j = ucol[q];
                    if (!inw[j]) {
                        inw[j] = 1;
                        w[j] = 0.0;
                        nz.push_back(j);
                        if (j < i) heap.push(j);
                    }
                    w[j] -= wk * uval[q];
                }
            }
            // L: largest (nl + fill)
            const size_t capL = (size_t)(nl + fillx);
            if (keepL.size() > capL) {
                std::nth_element(keepL.begin(), keepL.begin() + capL, keepL.end(),
                                 [](const auto& a, const auto& b) { return a.first > b.first; });
                keepL.resize(capL);
            }
            std::sort(keepL.begin(), keepL.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
            for (auto& e : keepL) { lcol.push_back((int32_t)e.second); lval.push_back(w[e.second]); }
            lrp[i + 1] = (int64_t)lcol.size();
            // U: diagonal + largest (nu + fill) off-diagonals above threshold
            keepU.clear();
            for (int64_t c : nz)
                if (c > i && std::fabs(w[c]) > thr) keepU.emplace_back(std::fabs(w[c]), c);
            const size_t capU = (size_t)(nu + fillx);
            if (keepU.size() > capU) {
                std::nth_element(keepU.begin(), keepU.begin() + capU, keepU.end(),
                                 [](const auto& a, const auto& b) { return a.first > b.first; });
                keepU.resize(capU);
            }
            std::sort(keepU.begin(), keepU.end(), [](const auto& a, const auto& b) { return a.second < b.second; });
            double d = w[i];
            if (d == 0.0) d = thr > 0 ? thr : 1.0;
            ucol.push_back((int32_t)i);
            uval.push_back(d);
            for (auto& e : keepU) { ucol.push_back((int32_t)e.second); uval.push_back(w[e.second]); }
            urp[i + 1] = (int64_t)ucol.size();
            for (int64_t c : nz) { w[c] = 0.0; inw[c] = 0; }
        }
    }
    void apply(const double* r, double* z) const override {
        for (int64_t i = 0; i < n; ++i) {
            double s = r[i];
            for (int64_t k = lrp[i]; k < lrp[i + 1]; ++k) s -= lval[k] * z[lcol[k]];
            z[i] = s;
        }
        for (int64_t i = n - 1; i >= 0; --i) {
            double s = z[i];
            for (int64_t k = urp[i] + 1; k < urp[i + 1]; ++k) s -= uval[k] * z[ucol[k]];
            z[i] = s / uval[urp[i]];
        }
    }
};

// Face colours of the multicolour D-ILU (rule in the comment of PcDilu), from
// the lateral couplings of the first F rows of Am.
std::vector<int> dilu_colors(const Csr& Am, int64_t F) {
    std::vector<int> color(F, -1);
    for (int64_t i = 0; i < F; ++i) {
        unsigned used = 0;
        for (int32_t k = Am.rp[i]; k < Am.rp[i + 1]; ++k) {
            const int64_t c = Am.col[k];
            if (c < F && c != i && color[c] >= 0) used |= 1u << color[c];
        }
        int c = 0;
        while (used & (1u << c)) ++c;
        color[i] = c;
    }
    auto nbrs = [&](int64_t i) {
        std::vector<int64_t> v;
        for (int32_t k = Am.rp[i]; k < Am.rp[i + 1]; ++k)
            if (Am.col[k] < F && Am.col[k] != i) v.push_back(Am.col[k]);
        return v;
    };
    auto neighbour_of_colour = [&](int64_t v, int c) {
        for (int64_t u : nbrs(v))
            if (color[u] == c) return u;
        return (int64_t)-1;
    };
    auto repair = [&](int64_t v) {
        std::set<int> present;
        for (int64_t u : nbrs(v)) present.insert(color[u]);
        for (int c = 0; c < 3; ++c)
            if (!present.count(c)) {
                color[v] = c;
                return true;
            }
        const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (const auto& pr : pairs) {
            const int ca = pr[0], cb = pr[1];
            const int64_t a = neighbour_of_colour(v, ca), b = neighbour_of_colour(v, cb);
            // component of {ca, cb} faces containing a (depth first, capped)
            std::set<int64_t> comp{a};
            std::vector<int64_t> todo{a};
            bool too_big = false;
            while (!todo.empty() && !too_big) {
                const int64_t u = todo.back();
                todo.pop_back();
                for (int64_t w : nbrs(u))
                    if ((color[w] == ca || color[w] == cb) && !comp.count(w)) {
                        comp.insert(w);
                        todo.push_back(w);
                        if ((int64_t)comp.size() > 4096) too_big = true;
                    }
            }
            if (too_big || comp.count(b)) continue;
            for (int64_t u : comp) color[u] = color[u] == ca ? cb : ca;
            color[v] = ca;
            return true;
        }
        return false;
    };
    for (int64_t v = 0; v < F; ++v) {
        if (color[v] != 3 || repair(v)) continue;
        for (int c = 0; c < 3; ++c) {  // trade colours with the neighbour of colour c
            const int64_t a = neighbour_of_colour(v, c);
            color[v] = c;
            color[a] = 3;
            if (repair(a)) break;
            color[a] = c;
            color[v] = 3;
        }
    }
    return color;
}

// D-ILU(0) in the multicolour order (face colour, layer, face): faces coloured
// greedily in index order with the smallest colour no lower-indexed neighbour
// holds, then every face left with colour 3 is repaired in index order: a free
// colour below 3 if one exists, else the first Kempe swap that frees one (for
// colour pairs (0,1), (0,2), (1,2): the component of the {ca, cb}-coloured faces
// reachable from v's ca-neighbour, if it does not contain v's cb-neighbour and
// has at most 4096 faces, has ca and cb exchanged and v takes ca), else v trades
// colours with its neighbour of colour c (c = 0, 1, 2 in turn) when that
// neighbour can then be repaired the same way;
// d~_r = a_rr - sum over earlier k of a_rk a_kr / d~_k (Pommerell's diagonal
// ILU); M = (D~ + L) D~^-1 (D~ + U) in that order.
struct PcDilu : Precond {
    int64_t F;
    int L;
    const Csr* A = nullptr;
    std::vector<int64_t> ord, key;
    std::vector<double> dt;
    PcDilu(int64_t F_, int L_) : F(F_), L(L_) {}
    void compute(const Csr& Am) override {
        A = &Am;
        const int64_t N = Am.n;
        const std::vector<int> color = dilu_colors(Am, F);
        key.resize(N);
        for (int64_t r = 0; r < N; ++r) key[r] = ((int64_t)color[r % F] * L + r / F) * F + r % F;
        ord.resize(N);
        for (int64_t r = 0; r < N; ++r) ord[r] = r;
        std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
        dt.assign(N, 0.0);
        for (int64_t r : ord) {
            double d = 0;
            for (int32_t k = Am.rp[r]; k < Am.rp[r + 1]; ++k)
                if (Am.col[k] == r) d = Am.val[k];
            for (int32_t k = Am.rp[r]; k < Am.rp[r + 1]; ++k) {
                const int64_t c = Am.col[k];
                if (c != r && key[c] < key[r]) d -= Am.val[k] * Am.val[Am.find(c, r)] / dt[c];
            }
            dt[r] = d != 0.0 ? d : 1.0;
        }
    }
    void apply(const double* r, double* z) const override {
        const Csr& Am = *A;
        for (int64_t q : ord) {
            double s = r[q];
            for (int32_t k = Am.rp[q]; k < Am.rp[q + 1]; ++k) {
                const int64_t c = Am.col[k];
                if (c != q && key[c] < key[q]) s -= Am.val[k] * z[c];
            }
            z[q] = s / dt[q];
        }
        for (auto it = ord.rbegin(); it != ord.rend(); ++it) {
            const int64_t q = *it;
            double s = 0;
            for (int32_t k = Am.rp[q]; k < Am.rp[q + 1]; ++k) {
                const int64_t c = Am.col[k];
                if (c != q && key[c] > key[q]) s += Am.val[k] * z[c];
            }
            z[q] = z[q] - s / dt[q];
        }
    }
};

std::unique_ptr<Precond> make_precond(int kind, int64_t F, int L) {
    switch (kind) {
        case PBSM3D_PC_MCILU0: return std::make_unique<PcDilu>(F, L);
        case PBSM3D_PC_NONE: return std::make_unique<PcNone>();
        case PBSM3D_PC_JACOBI: return std::make_unique<PcJacobi>();
        case PBSM3D_PC_COLUMN: return std::make_unique<PcColumn>(F, L);
        case PBSM3D_PC_ILU0: return std::make_unique<PcIlu0>();
        default: return std::make_unique<PcIlut>();
    }
}

struct SolveResult {
    int iters = 0;
    double resid = 0;
    bool converged = false;
};

// Restarted, right-preconditioned GMRES(m); convergence on ||b - A x|| / ||b||
// (implicit residual inside a cycle, confirmed by the explicit one).
SolveResult gmres(const Csr& A, const double* b, double* x, const Precond& M, int m, int maxit, double tol) {
    SolveResult res;
    const int64_t n = A.n;
    const double nb = nrm2(n, b);
    if (nb == 0.0) {
        std::fill(x, x + n, 0.0);
        res.converged = true;
        return res;
    }
    std::vector<double> V((size_t)(m + 1) * n), w(n), z(n), r(n);
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), h(m + 1), y(m);
    auto resid = [&]() {
        spmv(A, x, r.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
        return nrm2(n, r.data());
    };
    double beta = resid();
    res.resid = beta / nb;
    if (res.resid <= tol) { res.converged = true; return res; }
    while (res.iters < maxit) {
        double* v0 = V.data();
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) v0[i] = r[i] / beta;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        for (int j = 0; j < m && res.iters < maxit; ++j) {
            M.apply(V.data() + (size_t)j * n, z.data());
            spmv(A, z.data(), w.data());
            // classical Gram-Schmidt, two passes
            for (int i = 0; i <= j; ++i) h[i] = 0.0;
            for (int pass = 0; pass < 2; ++pass) {
                for (int i = 0; i <= j; ++i) {
                    double hi = dot(n, V.data() + (size_t)i * n, w.data());
                    H[(size_t)i * m + j] = hi;
                }
                for (int i = 0; i <= j; ++i) {
                    const double hi = H[(size_t)i * m + j];
                    h[i] += hi;
                    const double* vi = V.data() + (size_t)i * n;
#pragma omp parallel for schedule(static)
                    for (int64_t q = 0; q < n; ++q) w[q] -= hi * vi[q];
                }
            }
            double hn = nrm2(n, w.data());
            for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = h[i];
            H[(size_t)(j + 1) * m + j] = hn;
            if (hn > 0) {
                double* vj1 = V.data() + (size_t)(j + 1) * n;
#pragma omp parallel for schedule(static)
                for (int64_t q = 0; q < n; ++q) vj1[q] = w[q] / hn;
            }
            for (int i = 0; i < j; ++i) {
                double a = H[(size_t)i * m + j], bb = H[(size_t)(i + 1) * m + j];
                H[(size_t)i * m + j] = cs[i] * a + sn[i] * bb;
                H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * bb;
            }
            double a = H[(size_t)j * m + j], bb = H[(size_t)(j + 1) * m + j];
            double rr = std::hypot(a, bb);
            cs[j] = rr > 0 ? a / rr : 1.0;
            sn[j] = rr > 0 ? bb / rr : 0.0;
            H[(size_t)j * m + j] = rr;
            H[(size_t)(j + 1) * m + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++res.iters;
            k = j + 1;
            if (std::fabs(g[j + 1]) / nb <= tol || hn == 0.0) break;
        }
        for (int i = k - 1; i >= 0; --i) {
            double s = g[i];
            for (int q = i + 1; q < k; ++q) s -= H[(size_t)i * m + q] * y[q];
            y[i] = s / H[(size_t)i * m + i];
        }
        std::fill(w.begin(), w.end(), 0.0);
        for (int i = 0; i < k; ++i) {
            const double* vi = V.data() + (size_t)i * n;
            const double yi = y[i];
#pragma omp parallel for schedule(static)
            for (int64_t q = 0; q < n; ++q) w[q] += yi * vi[q];
        }
        M.apply(w.data(), z.data());
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < n; ++q) x[q] += z[q];
        beta = resid();
        res.resid = beta / nb;
        if (res.resid <= tol) { res.converged = true; return res; }
    }
    return res;
}

// Dense Gaussian elimination with partial pivoting (small systems only).
bool dense_solve(const Csr& A, const double* b, double* x) {
    const int64_t n = A.n;
    std::vector<double> M((size_t)n * n, 0.0), rhs(b, b + n);
    for (int64_t r = 0; r < n; ++r)
        for (int32_t k = A.rp[r]; k < A.rp[r + 1]; ++k) M[(size_t)r * n + A.col[k]] += A.val[k];
    for (int64_t c = 0; c < n; ++c) {
        int64_t p = c;
        for (int64_t r = c + 1; r < n; ++r)
            if (std::fabs(M[(size_t)r * n + c]) > std::fabs(M[(size_t)p * n + c])) p = r;
        if (M[(size_t)p * n + c] == 0.0) return false;
        if (p != c) {
            for (int64_t q = 0; q < n; ++q) std::swap(M[(size_t)p * n + q], M[(size_t)c * n + q]);
            std::swap(rhs[p], rhs[c]);
        }
        for (int64_t r = c + 1; r < n; ++r) {
            double f = M[(size_t)r * n + c] / M[(size_t)c * n + c];
            if (f == 0.0) continue;
            for (int64_t q = c; q < n; ++q) M[(size_t)r * n + q] -= f * M[(size_t)c * n + q];
            rhs[r] -= f * rhs[c];
        }
    }
    for (int64_t r = n - 1; r >= 0; --r) {
        double s = rhs[r];
        for (int64_t q = r + 1; q < n; ++q) s -= M[(size_t)r * n + q] * x[q];
        x[r] = s / M[(size_t)r * n + r];
    }
    return true;
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct oracle_ctx {
    int64_t F = 0;
    int L = 0;
    pbsm3d_config cfg{};
    double dz = 0;
    bool enable_veg = true;
    // geometry (face-major edges)
    std::vector<int32_t> neigh;
    std::vector<double> nx, ny, elen, area, cx, cy, cz;
    std::vector<double> canopy, lai, stalk_n, stalk_d;
    std::vector<uint8_t> water;
    // systems
    Csr susp, dep;
    std::vector<double> rhs_s, rhs_d, x_s, x_d;
    // per-step traces
    std::vector<double> ustar, hs, c_salt, z0, saltation, uth;
    std::vector<double> u_z, csubl;  // [L*F]
    // face variables / state
    std::vector<double> Qsalt, Qsusp, Qsubl, Qsubl_mass, sum_subl, drift_mass, sum_drift, more_than_avail;
    std::vector<double> blowing_prob;
    // solver options
    int direct = 0;
    int threads = 1;
    // replay: the next step takes these as the suspension / deposition solutions instead
    // of solving (a checker feeds another solver's x to get the rest of the step and the
    // deposition system that follows from it); empty = solve
    std::vector<double> replay_s, replay_d;
    std::string err;
};

extern "C" {

int oracle_face_geometry(int64_t F, const double* vxyz, const int32_t* elem, const double* area_param, double* nx,
                         double* ny, double* elen, double* area, double* cx, double* cy, double* cz) {
    for (int64_t i = 0; i < F; ++i) {
        const double* p[3];
        for (int k = 0; k < 3; ++k) p[k] = vxyz + 3 * (int64_t)elem[3 * i + k];
        // edge(j) = vertex(cw(j)) - vertex(ccw(j)), ccw(j)=(j+1)%3, cw(j)=(j+2)%3
        double ex[3], ey[3];
        for (int j = 0; j < 3; ++j) {
            const double* a = p[(j + 1) % 3];
            const double* b = p[(j + 2) % 3];
            ex[j] = b[0] - a[0];
            ey[j] = b[1] - a[1];
        }
        for (int j = 0; j < 3; ++j) {
            double n_x = ey[j], n_y = -ex[j];
            const int j1 = (j + 1) % 3;
            double D = ex[j1] * n_x + ey[j1] * n_y;
            if (D > 0) { n_x = -n_x; n_y = -n_y; }
            const double s = std::sqrt(n_x * n_x + n_y * n_y);
            nx[3 * i + j] = n_x / s;
            ny[3 * i + j] = n_y / s;
            elen[3 * i + j] = std::sqrt(ex[j] * ex[j] + ey[j] * ey[j]);
        }
        if (area_param) {
            area[i] = area_param[i];
        } else {
            const double v1x = p[1][0] - p[0][0], v1y = p[1][1] - p[0][1];
            const double v2x = p[2][0] - p[0][0], v2y = p[2][1] - p[0][1];
            area[i] = (v1x * v2y - v2x * v1y) / 2;
        }
        cx[i] = (p[0][0] + p[1][0] + p[2][0]) / 3;
        cy[i] = (p[0][1] + p[1][1] + p[2][1]) / 3;
        cz[i] = (p[0][2] + p[1][2] + p[2][2]) / 3;
    }
    return 0;
}

int64_t oracle_csr_nnz(int64_t F, int L, const int32_t* neigh) {
    int64_t deg = 0;
    for (int64_t i = 0; i < 3 * F; ++i) deg += neigh[i] >= 0;
    return (int64_t)L * (F + deg) + (L > 1 ? 2 * F * (L - 1) : 0);
}

int oracle_csr_pattern(int64_t F, int L, const int32_t* neigh, int32_t* rp, int32_t* col) {
    Csr A;
    build_pattern(F, L, neigh, A);
    std::memcpy(rp, A.rp.data(), (A.n + 1) * sizeof(int32_t));
    std::memcpy(col, A.col.data(), A.col.size() * sizeof(int32_t));
    return 0;
}

// face colours of the multicolour D-ILU; returns the number of colours
int oracle_face_colors(int64_t F, const int32_t* neigh, int32_t* color) {
    Csr A;
    build_pattern(F, 1, neigh, A);
    const std::vector<int> c = dilu_colors(A, F);
    int nc = 0;
    for (int64_t i = 0; i < F; ++i) {
        color[i] = c[i];
        nc = std::max(nc, c[i] + 1);
    }
    return nc;
}

void oracle_config_default(pbsm3d_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->nLayer = 10;
    c->do_fixed_settling = 0;
    c->settling_velocity = 0.5;
    c->do_sublimation = 1;
    c->do_lateral_diff = 1;
    c->smooth_coeff = 820;
    c->min_sd_trans = 0.1;
    c->cutoff = 0.3;
    c->snow_diffusion_const = 0.3;
    c->enable_veg = 1;
    c->use_tanh_fetch = 1;
    c->use_exp_fetch = 0;
    c->use_R94_lambda = 1;
    c->z0_snow = 0.01;
    c->settling_scale = 1.0;
    c->tol = 1e-8;
    c->max_iters = 1000;
    c->restart = 30;
    c->precision = PBSM3D_FP64;
    c->precond = PBSM3D_PC_ILUT;
    c->warm_start = PBSM3D_WS_ZERO;
    c->dep_precond = PBSM3D_PC_ILUT;
    c->ortho_passes = 2;
    c->block_size = 256;
}

int oracle_create(oracle_ctx** out, const pbsm3d_mesh* m, const pbsm3d_config* cfg) {
    *out = nullptr;
    if (!m || !cfg || m->n_faces <= 0 || cfg->nLayer < 2 || cfg->nLayer > 64) return PBSM3D_EINVAL;
    if (cfg->settling_velocity < 0) return PBSM3D_EINVAL;         // PBSM3D.cpp:239-242
    if (cfg->use_exp_fetch && cfg->use_tanh_fetch) return PBSM3D_EINVAL;  // PBSM3D.cpp:132-135
    auto c = std::make_unique<oracle_ctx>();
    const int64_t F = m->n_faces;
    c->F = F;
    c->L = cfg->nLayer;
    c->cfg = *cfg;
    c->dz = kSuspDepth / (double)cfg->nLayer;
    c->neigh.assign(m->neigh, m->neigh + 3 * F);
    c->nx.assign(m->normal_x, m->normal_x + 3 * F);
    c->ny.assign(m->normal_y, m->normal_y + 3 * F);
    c->elen.assign(m->edge_length, m->edge_length + 3 * F);
    c->area.assign(m->area, m->area + F);
    c->cx.assign(m->center_x, m->center_x + F);
    c->cy.assign(m->center_y, m->center_y + F);
    c->cz.assign(m->center_z, m->center_z + F);
    // PBSM3D.cpp:284-324: missing veg parameters -> zeros and veg disabled
    c->enable_veg = cfg->enable_veg != 0;
    const bool has_veg = m->canopy_height != nullptr;
    c->canopy.assign(F, 0.0);
    c->lai.assign(F, 0.0);
    c->stalk_n.assign(F, 0.0);
    c->stalk_d.assign(F, 0.0);
    if (has_veg && c->enable_veg) {
        for (int64_t i = 0; i < F; ++i) {
            c->canopy[i] = m->canopy_height[i];
            if (cfg->use_R94_lambda) {
                c->lai[i] = m->lai ? m->lai[i] : 0.0;
            } else {
                // both stalk attributes or the defaults N = 1, dv = 0.8 (PBSM3D.cpp:302-309)
                const bool both = m->stalk_number && m->stalk_diameter;
                c->stalk_n[i] = both ? m->stalk_number[i] : 1.0;
                c->stalk_d[i] = both ? m->stalk_diameter[i] : 0.8;
            }
        }
    } else {
        c->enable_veg = false;
    }
    c->water.assign(F, 0);
    if (m->is_water) c->water.assign(m->is_water, m->is_water + F);
    build_pattern(F, c->L, m->neigh, c->susp);
    build_pattern(F, 1, m->neigh, c->dep);
    const int64_t N = F * c->L;
    c->rhs_s.assign(N, 0.0);
    c->x_s.assign(N, 0.0);
    c->rhs_d.assign(F, 0.0);
    c->x_d.assign(F, 0.0);
    for (auto* v : {&c->ustar, &c->hs, &c->c_salt, &c->z0, &c->saltation, &c->uth}) v->assign(F, 0.0);
    c->u_z.assign(N, 0.0);
    c->csubl.assign(N, 0.0);
    for (auto* v : {&c->Qsalt, &c->Qsusp, &c->Qsubl, &c->Qsubl_mass, &c->drift_mass, &c->more_than_avail,
                    &c->blowing_prob})
        v->assign(F, -9999.0);  // face variables default to -9999 (variablestorage.hpp:269-272)
    c->sum_subl.assign(F, 0.0);
    c->sum_drift.assign(F, 0.0);  // PBSM3D.cpp:391
    *out = c.release();
    return 0;
}

void oracle_destroy(oracle_ctx* c) { delete c; }

int oracle_set_replay(oracle_ctx* c, const double* x_susp, const double* x_dep) {
    if (x_susp) c->replay_s.assign(x_susp, x_susp + c->F * c->L); else c->replay_s.clear();
    if (x_dep) c->replay_d.assign(x_dep, x_dep + c->F); else c->replay_d.clear();
    return 0;
}

int oracle_set_solver(oracle_ctx* c, int direct, int threads, int precond, int dep_precond) {
    c->direct = direct;
    c->threads = threads > 0 ? threads : 1;
    if (precond >= 0) c->cfg.precond = precond;
    if (dep_precond >= 0) c->cfg.dep_precond = dep_precond;
    return 0;
}

int oracle_step(oracle_ctx* c, const pbsm3d_forcing* fo, int32_t dt_i, pbsm3d_out* out, pbsm3d_stats* st) {
    if (c->cfg.use_PomLi_probability && !fo->p_snow_hours) return PBSM3D_EINVAL;
#ifdef _OPENMP
    omp_set_num_threads(c->threads);
#endif
    const int64_t F = c->F;
    const int L = c->L;
    const double dz = c->dz;
    const pbsm3d_config& cfg = c->cfg;
    const double dt = (double)dt_i;
    const double z0snow = cfg.z0_snow;
    pbsm3d_stats stats{};
    auto t_start = std::chrono::steady_clock::now();

    // zeroSystem x2 (LinearAlgebra.cpp:199-206)
    std::fill(c->susp.val.begin(), c->susp.val.end(), 0.0);
    std::fill(c->dep.val.begin(), c->dep.val.end(), 0.0);
    std::fill(c->rhs_s.begin(), c->rhs_s.end(), 0.0);
    std::fill(c->rhs_d.begin(), c->rhs_d.end(), 0.0);
    std::fill(c->x_s.begin(), c->x_s.end(), 0.0);
    std::fill(c->x_d.begin(), c->x_d.end(), 0.0);

    // ---------------- hot loop A: per face u*, saltation, per layer coefficients + assembly
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F; ++i) {
        double fetch = 1000;
        if (cfg.use_exp_fetch || cfg.use_tanh_fetch) fetch = fo->fetch ? fo->fetch[i] : 1000.0;
        const double uref = fo->U_R[i];
        double snow_depth = fo->snowdepthavg[i];
        snow_depth = chm_is_nan(snow_depth) ? 0 : snow_depth;
        const double u2 = fo->U_2m_above_srf[i];
        double swe = fo->swe[i];
        swe = chm_is_nan(swe) ? 0 : swe;
        double height_diff = std::max(0.0, c->canopy[i] - snow_depth);
        if (!c->enable_veg) height_diff = 0;
        const double min_sd_trans_avg = cfg.min_sd_trans;

        double ustar = 1.3;
        double lambda = 0;
        bool saltation = false;
        double z0 = 0;  // d.z0 is always assigned below before use
        const double T = fo->t[i];
        const double uth = 0.35 + (1.0 / 150.0) * T + (1.0 / 8200.0) * T * T;  // PBSM3D.cpp:671-672

        if (height_diff <= cfg.cutoff && snow_depth >= min_sd_trans_avg && !c->water[i]) {
            if (cfg.use_R94_lambda)
                lambda = 0.5 * c->lai[i] * height_diff;
            else
                lambda = c->stalk_n[i] * c->stalk_d[i] * height_diff;
            if (cfg.z0_ustar_coupling) {
                // PBSM3D.cpp:691-716: a failed search keeps the placeholder u* = 1.3
                bool ok = true;
                const double r = ustar_li_pomeroy(u2, lambda, &ok);
                if (ok) ustar = r;
            } else {
                ustar = u2 * kKappa / std::log(2.0 / 0.0002);
            }
            if (ustar >= uth) {
                saltation = true;
                if (cfg.z0_ustar_coupling)
                    z0 = 0.6131702345e-2 * ustar * ustar + .5 * lambda;  // PBSM3D.cpp:727-737
                else
                    z0 = z0snow;
            }
        }
        if (!saltation) {
            z0 = z0snow;
            ustar = std::max(0.01, kKappa * uref / std::log(kZUR / z0));
        }
        z0 = std::max(z0snow, z0);
        ustar = std::max(0.01, ustar);
        double hs = 0;
        if (saltation) hs = 0.08436 * std::pow(ustar, 1.27);

        double Qsalt = 0;
        double c_salt = 0;
        const double t = fo->t[i] + 273.15;
        if (saltation) {
            const double rho_f = std_dry_air_density(c->cz[i], t);
            const double Beta = 202.0;
            const double m = 0.16;
            const double tau_n_ratio = (m * Beta * lambda) / (1.0 + m * Beta * lambda);
            c_salt = rho_f / (3.29 * ustar) * (1.0 - tau_n_ratio - (uth * uth) / (ustar * ustar));
            if (c_salt < 0 || std::isnan(c_salt)) {
                c_salt = 0;
                saltation = false;
            }
            if (cfg.use_exp_fetch && fetch < 500) {
                const double fetch_ref = 500;
                const double mu = 3.0;
                c_salt *= 1.0 - std::exp(-mu * fetch / fetch_ref);
            } else if (cfg.use_tanh_fetch && fetch <= 300.) {
                const double fetch_ref = 300;
                const double Lc = 0.5 * std::tanh(0.1333333333e-1 * fetch_ref - 2.0) + 0.5;
                c_salt *= Lc;
            }
            if (cfg.use_PomLi_probability) {
                // PBSM3D.cpp:848-865, Essery, Li & Pomeroy 1999 eqs 10-14; u10 from
                // PBSM3D.cpp:451-463 (log_scale_wind with the snow roughness)
                const double z10 = 10. + snow_depth;
                const double u10 = z10 < kZUR ? log_scale_wind(uref, kZUR, z10, snow_depth, z0snow) : uref;
                const double A = fo->p_snow_hours[i];
                const double u_mean = 11.2 + 0.365 * T + 0.00706 * T * T + 0.9 * std::log(A);
                const double delta = 0.145 * T + 0.00196 * T * T + 4.3;
                const double z0v = (c->stalk_n[i] * c->stalk_d[i] * height_diff) / 2.0;
                const double us = u10 / std::sqrt((1 + 340.0 * z0v));
                const double Pu10 = 1.0 / (1.0 + std::exp((std::sqrt(kPi) * (u_mean - us)) / delta));
                c->blowing_prob[i] = Pu10;
                c_salt *= Pu10;
            }
            const double uhs = 2.8 * uth;
            Qsalt = c_salt * uhs * hs;
            double mass = 0;
            double vx, vy;
            neg_bearing_to_cartesian(fo->vw_dir[i], &vx, &vy);
            const double V = c->area[i];
            for (int j = 0; j < 3; ++j) {
                const double udotm = vx * c->nx[3 * i + j] + vy * c->ny[3 * i + j];
                const double E = c->elen[3 * i + j];
                mass += -E * Qsalt * udotm;
            }
            mass = mass / V * dt;
            if (mass < 0 && std::fabs(mass) > swe) {
                c_salt = 0;
                Qsalt = c_salt * uhs * hs;
            }
        }
        c->Qsalt[i] = Qsalt;
        c->ustar[i] = ustar;
        c->hs[i] = hs;
        c->c_salt[i] = c_salt;
        c->z0[i] = z0;
        c->saltation[i] = saltation ? 1.0 : 0.0;
        c->uth[i] = uth;

        const double rh = fo->rh[i] / 100.;
        const double es = saturated_vapour_pressure(t);

        double A[5];
        for (int j = 0; j < 3; ++j) A[j] = c->elen[3 * i + j] * dz;
        A[3] = A[4] = c->area[i];
        double vwx, vwy;
        neg_bearing_to_cartesian(fo->vw_dir[i], &vwx, &vwy);
        const double vnorm = std::sqrt(vwx * vwx + vwy * vwy);

        for (int z = 0; z < L; ++z) {
            const double czl = z * dz + hs + dz / 2.;
            double u_z = 0;
            const double hz = czl + snow_depth;
            if (saltation && czl < height_diff) {
                u_z = 2.8 * uth;
            } else if (czl < height_diff) {
                u_z = 0.01;
            } else {
                if (hz < kZUR)
                    u_z = std::max(0.01, log_scale_wind(uref, kZUR, hz, snow_depth, z0));
                else
                    u_z = std::max(0.01, uref);
            }
            c->u_z[z * F + i] = u_z;

            const double rm = 4.6e-5 * std::pow(czl, -0.258);
            const double mm_alpha = 4.08 + 12.6 * czl;
            const double mm = 4. / 3. * kPi * kRhoIce * rm * rm * rm *
                              (1.0 + 3.0 / mm_alpha + 2. / (mm_alpha * mm_alpha));
            const double r_z = std::pow((3.0 * mm) / (4 * kPi * kRhoIce), 0.3333333);
            const double xrz = 0.005 * std::pow(u_z, 1.36);
            double omega = cfg.settling_velocity;
            if (!cfg.do_fixed_settling) omega = 1.1e7 * std::pow(r_z, 1.8);
            omega *= cfg.settling_scale;  // extension key, 1.0 reproduces the reference
            const double Vr = omega + 3.0 * xrz * std::cos(kPi / 4.0);
            const double v = 1.88e-5;
            const double Re = 2.0 * r_z * Vr / v;
            const double Nu = 1.79 + 0.606 * std::pow(Re, 0.5);
            const double Sh = Nu;
            const double D = 2.06e-5 * std::pow(t / 273.15, 1.75);
            const double lambda_t = 0.000063 * t + 0.00673;
            const double Ls = 2.838e6;
            const double M = 18.01;
            const double R = 8313;
            const double sigma = (rh - 1.0) * (1.019 + 0.027 * std::log(czl));
            const double rho = (M * es) / (R * t);
            const double Qr = 0.9 * kPi * rm * rm * 120.0;
            const double dmdtz = Sh * rho * D *
                                 (6.283185308 * Nu * R * r_z * sigma * t * t * lambda_t - Ls * M * Qr + Qr * R * t) /
                                 (D * Ls * Sh * (Ls * M - R * t) * rho + lambda_t * t * t * Nu * R);
            double csubl = dmdtz / mm;

            double alpha[5] = {0, 0, 0, 0, 0};
            double K[5] = {0, 0, 0, 0, 0};
            if (cfg.do_lateral_diff) {
                for (int a = 0; a < 3; ++a) {
                    alpha[a] = A[a];
                    K[a] = 0.00001;
                    alpha[a] *= K[a];
                }
            }
            const double l = kKappa * (czl + z0) * kLMax / (kKappa * (czl + z0) + kLMax);
            const double w = omega;
            double diffusion_coeff = cfg.snow_diffusion_const;
            if (cfg.rouault_diffusion_coef) {  // PBSM3D.cpp:1166-1172
                const double c2 = 1.0;
                diffusion_coeff = 1.0 / (1.0 + (c2 * w * w) / (1.56 * ustar * ustar));
            }
            K[3] = K[4] = diffusion_coeff * ustar * l;
            alpha[3] = A[3] * K[3] / dz;
            alpha[4] = A[4] * K[4] / dz;

            const double kscale = u_z / vnorm;
            const double ux = vwx * kscale, uy = vwy * kscale;
            double udotm[5];
            for (int j = 0; j < 3; ++j) udotm[j] = ux * c->nx[3 * i + j] + uy * c->ny[3 * i + j];
            udotm[3] = -w;
            udotm[4] = w;

            const int64_t idx = z * F + i;
            double V = c->area[i] * dz;
            V /= 5.0;
            if (!cfg.do_sublimation) csubl = 0.0;
            c->csubl[idx] = csubl;

            Csr& S = c->susp;
            for (int f = 0; f < 3; ++f) {
                const int32_t nb = c->neigh[3 * i + f];
                if (udotm[f] > 0) {
                    if (nb >= 0) {
                        S.sum_into(idx, idx, (V * csubl - A[f] * udotm[f] - alpha[f]));
                        S.sum_into(idx, z * F + nb, (alpha[f]));
                    } else {
                        S.sum_into(idx, idx, (-0.1e-1 * alpha[f] - 1. * A[f] * udotm[f] + csubl * V));
                    }
                } else {
                    if (nb >= 0) {
                        S.sum_into(idx, idx, V * csubl - alpha[f]);
                        S.sum_into(idx, z * F + nb, -A[f] * udotm[f] + alpha[f]);
                    } else {
                        S.sum_into(idx, idx, -0.1e-1 * alpha[f] - .99 * A[f] * udotm[f] + csubl * V);
                    }
                }
            }
            if (z == 0) {
                const double alpha4 = A[4] * K[4] / (hs / 2.0 + dz / 2.0);
                S.sum_into(idx, idx, V * csubl - A[4] * udotm[4] - alpha4);
                c->rhs_s[idx] += -alpha4 * c_salt;
                const int64_t nidx = F * (z + 1) + i;
                if (udotm[3] > 0) {
                    S.sum_into(idx, idx, V * csubl - A[3] * udotm[3] - alpha[3]);
                    S.sum_into(idx, nidx, alpha[3]);
                } else {
                    S.sum_into(idx, idx, V * csubl - alpha[3]);
                    S.sum_into(idx, nidx, -A[3] * udotm[3] + alpha[3]);
                }
            } else if (z == L - 1) {
                const double cprecip = 0;
                if (udotm[3] > 0) {
                    S.sum_into(idx, idx, V * csubl - A[3] * udotm[3] - alpha[3]);
                    c->rhs_s[idx] += -alpha[3] * cprecip;
                } else {
                    S.sum_into(idx, idx, V * csubl - alpha[3]);
                    c->rhs_s[idx] += A[3] * cprecip * udotm[3] - alpha[3] * cprecip;
                }
                const int64_t nidx = F * (z - 1) + i;
                if (udotm[4] > 0) {
                    S.sum_into(idx, idx, V * csubl - A[4] * udotm[4] - alpha[4]);
                    S.sum_into(idx, nidx, alpha[4]);
                } else {
                    S.sum_into(idx, idx, V * csubl - alpha[4]);
                    S.sum_into(idx, nidx, -A[4] * udotm[4] + alpha[4]);
                }
            } else {
                int64_t nidx = F * (z + 1) + i;
                if (udotm[3] > 0) {
                    S.sum_into(idx, idx, V * csubl - A[3] * udotm[3] - alpha[3]);
                    S.sum_into(idx, nidx, alpha[3]);
                } else {
                    S.sum_into(idx, idx, V * csubl - alpha[3]);
                    S.sum_into(idx, nidx, -A[3] * udotm[3] + alpha[3]);
                }
                nidx = F * (z - 1) + i;
                if (udotm[4] > 0) {
                    S.sum_into(idx, idx, V * csubl - A[4] * udotm[4] - alpha[4]);
                    S.sum_into(idx, nidx, alpha[4]);
                } else {
                    S.sum_into(idx, idx, V * csubl - alpha[4]);
                    S.sum_into(idx, nidx, -A[4] * udotm[4] + alpha[4]);
                }
            }
        }
    }
    auto t_asm = std::chrono::steady_clock::now();

    // getRhsMax + suspension solve (PBSM3D.cpp:1424-1459)
    const int64_t N = F * L;
    double rmax = 0;
    for (int64_t r = 0; r < N; ++r) rmax = std::max(rmax, std::fabs(c->rhs_s[r]));
    stats.susp_rhs_max = rmax;
    const bool suspension_present = rmax > kPresentThreshold;
    stats.suspension_present = suspension_present;
    int rc = 0;
    if (suspension_present) {
        if (!c->replay_s.empty()) {
            c->x_s = c->replay_s;
        } else if (c->direct) {
            if (!dense_solve(c->susp, c->rhs_s.data(), c->x_s.data())) rc = PBSM3D_ENOCONV;
            stats.susp_iters = 0;
        } else {
            auto M = make_precond(cfg.precond, F, L);
            M->compute(c->susp);
            SolveResult sr = gmres(c->susp, c->rhs_s.data(), c->x_s.data(), *M, cfg.restart, cfg.max_iters, cfg.tol);
            stats.susp_iters = sr.iters;
            stats.susp_resid = sr.resid;
            if (!sr.converged) rc = PBSM3D_ENOCONV;
        }
    }
    auto t_susp = std::chrono::steady_clock::now();
    if (rc) {
        c->err = "suspension solver failed to converge";
        if (st) *st = stats;
        return rc;
    }

    // post-processing (PBSM3D.cpp:1465-1503)
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F; ++i) {
        double Qsusp = 0, Qsubl = 0;
        for (int z = 0; z < L; ++z) {
            double cc = c->x_s[F * z + i];
            cc = cc < 0 || chm_is_nan(cc) ? 0 : cc;
            const double u_z = c->u_z[z * F + i];
            Qsusp += cc * u_z * dz;
            Qsubl += c->csubl[z * F + i] * cc * dz;
        }
        c->Qsusp[i] = Qsusp;
        c->Qsubl[i] = Qsubl;
        c->Qsubl_mass[i] = Qsubl * dt;
        c->sum_subl[i] += c->Qsubl_mass[i];
    }

    // deposition assembly (PBSM3D.cpp:1516-1658)
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < F; ++i) {
        double vx, vy;
        neg_bearing_to_cartesian(fo->vw_dir[i], &vx, &vy);
        const double V = c->area[i];
        Csr& D = c->dep;
        D.replace(i, i, V);
        for (int j = 0; j < 3; ++j) {
            const double udotm = vx * c->nx[3 * i + j] + vy * c->ny[3 * i + j];
            const double E = c->elen[3 * i + j];
            const int32_t nb = c->neigh[3 * i + j];
            double Qtj = 0, Qsj = 0;
            if (udotm > 0) {
                Qtj = c->Qsusp[i];
                Qsj = c->Qsalt[i];
            } else {
                if (nb >= 0) {
                    Qtj = c->Qsusp[nb];
                    Qsj = c->Qsalt[nb];
                    if (chm_is_nan(Qsj)) Qsj = 0;
                } else {
                    Qtj = c->Qsusp[i];
                    Qsj = c->Qsalt[i];
                }
            }
            if (nb >= 0) {
                const double ddx = c->cx[i] - c->cx[nb], ddy = c->cy[i] - c->cy[nb];
                const double dxj = std::sqrt(ddx * ddx + ddy * ddy);
                const double eps = cfg.smooth_coeff;
                D.sum_into(i, i, eps * E / dxj);
                D.sum_into(i, nb, -eps * E / dxj);
            }
            c->rhs_d[i] += -E * (Qtj + Qsj) * udotm;
        }
    }
    double dmax = 0;
    for (int64_t r = 0; r < F; ++r) dmax = std::max(dmax, std::fabs(c->rhs_d[r]));
    stats.dep_rhs_max = dmax;
    const bool deposition_present = suspension_present && dmax > kPresentThreshold;
    stats.deposition_present = deposition_present;
    auto t_post = std::chrono::steady_clock::now();
    if (deposition_present) {
        if (!c->replay_d.empty()) {
            c->x_d = c->replay_d;
        } else if (c->direct) {
            if (!dense_solve(c->dep, c->rhs_d.data(), c->x_d.data())) rc = PBSM3D_ENOCONV;
        } else {
            auto M = make_precond(cfg.dep_precond, F, 1);
            M->compute(c->dep);
            SolveResult sr = gmres(c->dep, c->rhs_d.data(), c->x_d.data(), *M, cfg.restart, cfg.max_iters, cfg.tol);
            stats.dep_iters = sr.iters;
            stats.dep_resid = sr.resid;
            if (!sr.converged) rc = PBSM3D_ENOCONV;
        }
        if (rc) {
            c->err = "deposition solver failed to converge";
            if (st) *st = stats;
            return rc;
        }
        // drift mass (PBSM3D.cpp:1708-1740)
        for (int64_t i = 0; i < F; ++i) {
            const double qdep = chm_is_nan(c->x_d[i]) ? 0 : c->x_d[i];
            double mass = qdep * dt;
            double swe = fo->swe[i];
            swe = chm_is_nan(swe) ? 0 : swe;
            if (mass < 0 && std::fabs(mass) > swe) {
                c->more_than_avail[i] = 1;
                mass = -swe;
            }
            if (mass < 0 && c->saltation[i] == 0.0) mass = 0;
            c->drift_mass[i] = mass;
            c->sum_drift[i] += mass;
        }
    }
    auto t_end = std::chrono::steady_clock::now();
    auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
    stats.ms_assembly = ms(t_start, t_asm);
    stats.ms_susp_solve = ms(t_asm, t_susp);
    stats.ms_post = ms(t_susp, t_post);
    stats.ms_dep_solve = ms(t_post, t_end);
    stats.ms_total = ms(t_start, t_end);
    if (st) *st = stats;
    if (out) {
        auto cp = [&](double* dst, const std::vector<double>& src) {
            if (dst) std::memcpy(dst, src.data(), src.size() * sizeof(double));
        };
        cp(out->Qsalt, c->Qsalt);
        cp(out->Qsusp, c->Qsusp);
        cp(out->Qsubl, c->Qsubl);
        cp(out->Qsubl_mass, c->Qsubl_mass);
        cp(out->sum_subl, c->sum_subl);
        cp(out->drift_mass, c->drift_mass);
        cp(out->sum_drift, c->sum_drift);
        cp(out->pbsm_more_than_avail, c->more_than_avail);
        cp(out->blowingsnow_probability, c->blowing_prob);
        cp(out->conc, c->x_s);
        cp(out->qdep, c->x_d);
    }
    return 0;
}

int64_t oracle_system_nnz(oracle_ctx* c, int which) { return which == 0 ? (int64_t)c->susp.col.size() : (int64_t)c->dep.col.size(); }

int oracle_export_system(oracle_ctx* c, int which, int32_t* rp, int32_t* col, double* val, double* rhs) {
    const Csr& A = which == 0 ? c->susp : c->dep;
    const std::vector<double>& b = which == 0 ? c->rhs_s : c->rhs_d;
    if (rp) std::memcpy(rp, A.rp.data(), (A.n + 1) * sizeof(int32_t));
    if (col) std::memcpy(col, A.col.data(), A.col.size() * sizeof(int32_t));
    if (val) std::memcpy(val, A.val.data(), A.val.size() * sizeof(double));
    if (rhs) std::memcpy(rhs, b.data(), b.size() * sizeof(double));
    return 0;
}

// per-face scalars of the last step: u*, hs, c_salt, z0, saltation flag, u*_th
int oracle_export_face_trace(oracle_ctx* c, double* ustar, double* hs, double* c_salt, double* z0, double* salt,
                             double* uth) {
    const size_t b = c->F * sizeof(double);
    if (ustar) std::memcpy(ustar, c->ustar.data(), b);
    if (hs) std::memcpy(hs, c->hs.data(), b);
    if (c_salt) std::memcpy(c_salt, c->c_salt.data(), b);
    if (z0) std::memcpy(z0, c->z0.data(), b);
    if (salt) std::memcpy(salt, c->saltation.data(), b);
    if (uth) std::memcpy(uth, c->uth.data(), b);
    return 0;
}

int oracle_export_cell_trace(oracle_ctx* c, double* u_z, double* csubl) {
    const size_t b = c->u_z.size() * sizeof(double);
    if (u_z) std::memcpy(u_z, c->u_z.data(), b);
    if (csubl) std::memcpy(csubl, c->csubl.data(), b);
    return 0;
}

int oracle_get_state(oracle_ctx* c, double* sum_drift, double* sum_subl) {
    if (sum_drift) std::memcpy(sum_drift, c->sum_drift.data(), c->F * sizeof(double));
    if (sum_subl) std::memcpy(sum_subl, c->sum_subl.data(), c->F * sizeof(double));
    return 0;
}

int oracle_set_state(oracle_ctx* c, const double* sum_drift, const double* sum_subl) {
    if (sum_drift) std::memcpy(c->sum_drift.data(), sum_drift, c->F * sizeof(double));
    if (sum_subl) std::memcpy(c->sum_subl.data(), sum_subl, c->F * sizeof(double));
    return 0;
}

const char* oracle_last_error(oracle_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

// Solve a given CSR system with the oracle's Krylov/preconditioner (for tests
// of the GPU solver pieces).  F/L are used by the column preconditioner.
int oracle_solve_csr(int64_t n, const int32_t* rp, const int32_t* col, const double* val, const double* b, double* x,
                     int precond, int restart, int maxit, double tol, int64_t F, int L, int* iters, double* resid) {
    Csr A;
    A.n = n;
    A.rp.assign(rp, rp + n + 1);
    A.col.assign(col, col + rp[n]);
    A.val.assign(val, val + rp[n]);
    auto M = make_precond(precond, F, L);
    M->compute(A);
    SolveResult sr = gmres(A, b, x, *M, restart, maxit, tol);
    if (iters) *iters = sr.iters;
    if (resid) *resid = sr.resid;
    return sr.converged ? 0 : PBSM3D_ENOCONV;
}

int oracle_dense_solve(int64_t n, const int32_t* rp, const int32_t* col, const double* val, const double* b, double* x) {
    Csr A;
    A.n = n;
    A.rp.assign(rp, rp + n + 1);
    A.col.assign(col, col + rp[n]);
    A.val.assign(val, val + rp[n]);
    return dense_solve(A, b, x) ? 0 : PBSM3D_ENOCONV;
}

// z = M^-1 r for one preconditioner kind (parity of the GPU preconditioners).
int oracle_precond_apply(int64_t n, const int32_t* rp, const int32_t* col, const double* val, int precond, int64_t F,
                         int L, const double* r, double* z) {
    Csr A;
    A.n = n;
    A.rp.assign(rp, rp + n + 1);
    A.col.assign(col, col + rp[n]);
    A.val.assign(val, val + rp[n]);
    auto M = make_precond(precond, F, L);
    M->compute(A);
    M->apply(r, z);
    return 0;
}

// MeteoIO stdDryAirDensity as restated above (PBSM3D.cpp:784-786), for the table checks
double oracle_std_dry_air_density(double altitude, double temperature) {
    return std_dry_air_density(altitude, temperature);
}

// module_base::is_nan as restated above (module_base.hpp:471-479)
int oracle_is_nan(double v) { return chm_is_nan(v) ? 1 : 0; }

double oracle_ustar_coupled(double u2, double lambda, int* ok) {
    bool b = true;
    const double r = ustar_li_pomeroy(u2, lambda, &b);
    *ok = b ? 1 : 0;
    return r;
}

void oracle_set_ilut(double tau, double level_of_fill) {
    g_ilut_tau = tau;
    g_ilut_lof = level_of_fill;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
