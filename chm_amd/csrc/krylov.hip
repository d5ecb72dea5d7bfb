// GMRES(m) with classical Gram-Schmidt and a conditional second pass (the
// DGKS test ||w_after|| < ||w_before|| / sqrt(2)), as iterated classical
// Gram-Schmidt does.  The basis lives in HBM; each Arnoldi step is
//   z = M^-1 v_j ; v_{j+1} = A z ; h = V^T v_{j+1} (one batched pass) ;
//   v_{j+1} -= V h (one pass, fused norm) ; v_{j+1} /= ||v_{j+1}||
// and only the (j+2) Hessenberg entries travel to the host, where the Givens
// rotations and the least-squares solve run in FP64.
#include <algorithm>
#include <cmath>
#include <vector>

#include "krylov.hpp"

namespace pb {
namespace {

// device scalar slots
inline int off_h() { return 0; }
inline int off_ss(int m) { return m + 2; }
inline int off_y(int m) { return m + 4; }

constexpr int kVecBlocks = 512;  // two workgroups per CU

inline int reduce_blocks(int64_t n) {
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b > kReduceBlocks) b = kReduceBlocks;
    return (int)(b < 1 ? 1 : b);
}

}  // namespace

template <typename T>
void GmresWork<T>::ensure(int64_t n_, int m_) {
    if (n == n_ && m == m_ && V.p) return;
    n = n_;
    m = m_;
    V.alloc((size_t)(m + 1) * n);
    z.alloc(n);
    u.alloc(n);
    // residual partials: one per block of kBlock rows, or ceil(F / kBlock) * L <= n / kBlock + L
    // for the launches that run layer by layer (L <= 64)
    const size_t spmv_blocks = (size_t)((n + kBlock - 1) / kBlock) + 64;
    size_t pbytes = (size_t)kReduceBlocks * (m + 2);
    if (pbytes < spmv_blocks) pbytes = spmv_blocks;
    partial.alloc(pbytes);
    dsc.alloc(2 * m + 8);
    hsc.alloc(2 * m + 8);
}

template <typename T>
int gmres_cycle(hipStream_t s, GmresWork<T>& w, const CsrOp<T>& A, const PcFn<T>& M, const T* r0, double beta,
                double stop_abs, int max_steps, T* out, double* resid_est) {
    const int64_t n = A.n;
    const int m = w.m;
    const int nblk = reduce_blocks(n);
    T* V = w.V.p;
    double* d_h = w.dsc.p + off_h();
    double* d_ss = w.dsc.p + off_ss(m);
    double* d_y = w.dsc.p + off_y(m);
    double* hs = w.hsc.p;
    if (max_steps > m) max_steps = m;

    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0), h(m + 2, 0.0), y(m, 0.0);
    scale_host<T>(s, n, r0, 1.0 / beta, V);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < max_steps; ++j) {
        T* vj = V + (size_t)j * n;
        T* vj1 = V + (size_t)(j + 1) * n;
        {
            ProfScope ps(A.prof, s, PBSM3D_K_PRECOND);
            M(s, vj, w.z.p);
        }
        {
            ProfScope ps(A.prof, s, sizeof(T) == 8 ? PBSM3D_K_SPMV64 : PBSM3D_K_SPMV32);
            op_spmv<T>(s, A, w.z.p, vj1);
        }
        for (int i = 0; i <= j; ++i) h[i] = 0.0;
        double ss_before = 0, ss_after = 0;
        for (int pass = 0; pass < 2; ++pass) {
            {
                ProfScope ps(A.prof, s, PBSM3D_K_ORTHO);
                mdot<T>(s, n, V, j + 1, vj1, w.partial.p, nblk);
                reduce_cols(s, w.partial.p, nblk, j + 2, j + 2, d_h);
                gs_update<T>(s, n, V, j + 1, d_h, vj1, w.partial.p, nblk);
                reduce_cols(s, w.partial.p, nblk, 1, 1, d_ss);
            }
            PB_HIP(hipMemcpyAsync(hs, w.dsc.p, sizeof(double) * (m + 3), hipMemcpyDeviceToHost, s));
            PB_HIP(hipStreamSynchronize(s));
            for (int i = 0; i <= j; ++i) h[i] += hs[off_h() + i];
            if (pass == 0) ss_before = hs[off_h() + j + 1];
            ss_after = hs[off_ss(m)];
            if (w.passes < 2 || !(ss_after < 0.5 * ss_before)) break;  // no re-orthogonalisation needed
        }
        const double hn = std::sqrt(ss_after);
        if (hn > 0) scale_by_inv<T>(s, n, vj1, d_ss, vj1);
        for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = h[i];
        H[(size_t)(j + 1) * m + j] = hn;
        for (int i = 0; i < j; ++i) {
            const double a = H[(size_t)i * m + j], b = H[(size_t)(i + 1) * m + j];
            H[(size_t)i * m + j] = cs[i] * a + sn[i] * b;
            H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
        }
        const double a = H[(size_t)j * m + j], b = H[(size_t)(j + 1) * m + j];
        const double rr = std::hypot(a, b);
        cs[j] = rr > 0 ? a / rr : 1.0;
        sn[j] = rr > 0 ? b / rr : 0.0;
        H[(size_t)j * m + j] = rr;
        H[(size_t)(j + 1) * m + j] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        k = j + 1;
        if (std::fabs(g[j + 1]) <= stop_abs || hn == 0.0) break;
    }
    for (int i = k - 1; i >= 0; --i) {
        double sacc = g[i];
        for (int q = i + 1; q < k; ++q) sacc -= H[(size_t)i * m + q] * y[q];
        y[i] = sacc / H[(size_t)i * m + i];
    }
    for (int i = 0; i < k; ++i) hs[off_y(m) + i] = y[i];
    PB_HIP(hipMemcpyAsync(d_y, hs + off_y(m), sizeof(double) * k, hipMemcpyHostToDevice, s));
    combine<T>(s, n, V, k, d_y, w.u.p);
    {
        ProfScope ps(A.prof, s, PBSM3D_K_PRECOND);
        M(s, w.u.p, out);
    }
    // the pinned staging buffer is reused by the next cycle: wait for the copy
    PB_HIP(hipStreamSynchronize(s));
    *resid_est = std::fabs(g[k]);
    return k;
}

double dev_norm2(hipStream_t s, GmresWork<double>& w, int64_t n, const double* v) {
    const int nblk = reduce_blocks(n);
    mdot<double>(s, n, nullptr, 0, v, w.partial.p, nblk);
    reduce_cols(s, w.partial.p, nblk, 1, 1, w.dsc.p + off_ss(w.m) + 1);
    PB_HIP(hipMemcpyAsync(w.hsc.p, w.dsc.p + off_ss(w.m) + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    PB_HIP(hipStreamSynchronize(s));
    return std::sqrt(w.hsc.p[0]);
}

namespace {
// A start whose residual exceeds ||b|| (that of x0 = 0) is dropped: x is reset
// to zero and the caller recomputes the residual.  Records ||r0|| / ||b||.  The
// margin keeps x0 = 0 itself, whose ||r0|| and ||b|| differ only by the order of
// their two reductions, from counting as a fallback.
bool drop_worse_start(hipStream_t s, SolveInfo& info, double beta, double nb, double* x, int64_t n) {
    info.resid0 = beta / nb;
    if (!(info.resid0 > 1.0 + 1e-12)) return false;
    PB_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    info.cold_fallback = true;
    return true;
}

template <typename T>
double residual_norm(hipStream_t s, GmresWork<T>& w, const CsrOp<double>& A, const double* x, const double* b,
                     double* r) {
    int nblk = 0;
    {
        ProfScope ps(A.prof, s, PBSM3D_K_RESIDUAL);
        nblk = op_residual(s, A, x, b, r, w.partial.p);
    }
    reduce_cols(s, w.partial.p, nblk, 1, 1, w.dsc.p + off_ss(w.m) + 1);
    PB_HIP(hipMemcpyAsync(w.hsc.p, w.dsc.p + off_ss(w.m) + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    PB_HIP(hipStreamSynchronize(s));
    return std::sqrt(w.hsc.p[0]);
}

template <typename T>
double norm_of(hipStream_t s, GmresWork<T>& w, int64_t n, const double* v) {
    const int nblk = reduce_blocks(n);
    mdot<double>(s, n, nullptr, 0, v, w.partial.p, nblk);
    reduce_cols(s, w.partial.p, nblk, 1, 1, w.dsc.p + off_ss(w.m) + 1);
    PB_HIP(hipMemcpyAsync(w.hsc.p, w.dsc.p + off_ss(w.m) + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    PB_HIP(hipStreamSynchronize(s));
    return std::sqrt(w.hsc.p[0]);
}
}  // namespace

SolveInfo gmres_solve(hipStream_t s, GmresWork<double>& w, const CsrOp<double>& A, const PcFn<double>& M,
                      const double* b, double* x, double* r, int m, int maxit, double tol) {
    SolveInfo info;
    w.ensure(A.n, m);
    const double nb = norm_of(s, w, A.n, b);
    if (nb == 0.0) {
        PB_HIP(hipMemsetAsync(x, 0, sizeof(double) * A.n, s));
        info.converged = true;
        return info;
    }
    double beta = residual_norm(s, w, A, x, b, r);
    if (drop_worse_start(s, info, beta, nb, x, A.n)) beta = residual_norm(s, w, A, x, b, r);
    while (true) {
        info.resid = beta / nb;
        if (info.resid <= tol) {
            info.converged = true;
            break;
        }
        if (info.iters >= maxit) break;
        double est = 0;
        const int steps = gmres_cycle<double>(s, w, A, M, r, beta, tol * nb, maxit - info.iters, r, &est);
        axpy_f64(s, A.n, 1.0, r, x);
        info.iters += steps;
        info.cycles += 1;
        beta = residual_norm(s, w, A, x, b, r);
    }
    return info;
}

SolveInfo mixed_solve(hipStream_t s, GmresWork<float>& w, const CsrOp<double>& A64, const CsrOp<float>& A32,
                      const PcFn<float>& M32, const double* b, double* x, double* r, float* r32, float* d32, int m,
                      int maxit, double tol, double inner_rtol) {
    SolveInfo info;
    w.ensure(A64.n, m);
    const double nb = norm_of(s, w, A64.n, b);
    if (nb == 0.0) {
        PB_HIP(hipMemsetAsync(x, 0, sizeof(double) * A64.n, s));
        info.converged = true;
        return info;
    }
    double beta = residual_norm(s, w, A64, x, b, r);
    if (drop_worse_start(s, info, beta, nb, x, A64.n)) beta = residual_norm(s, w, A64, x, b, r);
    while (true) {
        info.resid = beta / nb;
        if (info.resid <= tol) {
            info.converged = true;
            break;
        }
        if (info.iters >= maxit) break;
        // inner FP32 cycle on the normalised FP64 residual
        scale_f64_to_f32_host(s, A64.n, r, 1.0 / beta, r32);
        double est = 0;
        double stop = 0.5 * tol * nb / beta;
        if (stop < inner_rtol) stop = inner_rtol;
        const int steps = gmres_cycle<float>(s, w, A32, M32, r32, 1.0, stop, maxit - info.iters, d32, &est);
        axpy_f32_into_f64(s, A64.n, beta, d32, x);
        info.iters += steps;
        info.cycles += 1;
        info.refinements += 1;
        beta = residual_norm(s, w, A64, x, b, r);
    }
    return info;
}

// ---------------------------------------------------------------------------
// BiCGStab
// ---------------------------------------------------------------------------
template <typename T>
void BicgWork<T>::ensure(int64_t n_) {
    if (n == n_ && r.p) return;
    n = n_;
    for (DevBuf<T>* b : {&r, &rh, &p, &v, &s, &t, &ph, &sh}) b->alloc(n);
    // two partial sums per block of the fused SpMV + dot launches, or of the vector kernels
    partial.alloc(std::max<size_t>(4 * (size_t)kReduceBlocks, 2 * (size_t)pos_spmv_dots_max_blocks(n)));
    dsc.alloc(SC_COUNT);
    ss_partial.alloc(kReduceBlocks);
    xr_partial.alloc(2 * (size_t)kReduceBlocks);
    hsc.alloc(2 * SC_COUNT);
    PB_HIP(hipHostGetDevicePointer((void**)&hsc_dev, hsc.p, 0));
    for (hipEvent_t& e : ev)
        if (!e) PB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

namespace {
// y = A x with the partial sums (w.y, y.y) of each block (two per block); returns the block count
// (stop: the split operator's launches return at entry once *stop != 0)
template <typename T>
int64_t spmv_dots(hipStream_t s, const CsrOp<T>& A, const T* x, T* y, const T* w, double* partial, int nblk,
                  const double* stop) {
    if (A.split) return split_apply_dots<T>(s, *A.co, A.F, A.L, *A.split, x, y, w, partial, A.mode, stop);
    if (A.pos) return pos_spmv_dots<T>(s, A.variant, *A.co, A.F, A.L, *A.pos, x, y, w, partial);
    op_spmv<T>(s, A, x, y);
    mdot<T>(s, A.n, w, 1, y, partial, nblk);
    return nblk;
}

// One BiCGStab run on A x = b from the x given; stops when ||r|| <= stop_abs,
// on a breakdown (the caller restarts from the true residual) or after maxit
// iterations; returns the iterations.  alpha, omega, rho and the stop test live
// on the device, and the host queues iteration k+1 before it waits for the
// scalars of iteration k, so the device never idles on the host round trip;
// after the stop every kernel of the queued iteration returns at entry (x and r
// keep the values of the stopping iteration).
// An empty M means no preconditioner (the operator already carries it, as the
// split operator does).  stop_rel: stop_abs is relative to the starting ||r||.
template <typename T>
int bicg_run(hipStream_t s, BicgWork<T>& w, const CsrOp<T>& A, const PcFn<T>& M, const T* b, T* x, bool x_zero,
             double stop_abs, int maxit, double* res_out, bool stop_rel = false, double restart_eps = 0.0,
             int* restarts = nullptr) {
    const int64_t n = A.n;
    const int nblk = reduce_blocks(n);
    // The s and x/r kernels: every workgroup first sums the operator's dot partials itself,
    // so they run as few workgroups as fill the chip (grid-stride over the vectors).
    const int nvec = std::min(nblk, kVecBlocks);
    const int spmv_id = A.prof_class >= 0 ? A.prof_class : (sizeof(T) == 8 ? PBSM3D_K_SPMV64 : PBSM3D_K_SPMV32);
    const bool pc = static_cast<bool>(M);
    T* ph = pc ? w.ph.p : w.p.p;  // the preconditioned search directions
    T* sh = pc ? w.sh.p : w.s.p;
    double* sc = w.dsc.p;
    double* h = w.hsc.p;
    auto fetch = [&]() {
        PB_HIP(hipMemcpyAsync(h, sc, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, s));
        PB_HIP(hipStreamSynchronize(s));
    };
    // r = b - A x, rh = r, (rh.r, r.r)
    if (x_zero) {
        PB_HIP(hipMemcpyAsync(w.r.p, b, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    } else {
        {
            ProfScope ps(A.prof, s, spmv_id);
            op_spmv<T>(s, A, x, w.v.p);
        }
        axpy_dots<T>(s, n, 1.0, w.v.p, b, w.r.p, nullptr, w.partial.p, nblk);
    }
    PB_HIP(hipMemcpyAsync(w.rh.p, w.r.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    mdot<T>(s, n, w.rh.p, 1, w.r.p, w.partial.p, nblk);
    bicg_stage(s, w.partial.p, nblk, 2, 0, sc);
    fetch();
    double rr = std::sqrt(h[SC_RR]);
    *res_out = rr;
    if (stop_rel) stop_abs *= rr;
    if (rr <= stop_abs) return 0;
    const double* stop = sc + SC_STOP;
    auto iterate = [&](int it) {
        if (it == 1) {
            PB_HIP(hipMemcpyAsync(w.p.p, w.r.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
        } else {
            ProfScope ps(A.prof, s, PBSM3D_K_ORTHO);
            bicg_p_dev<T>(s, n, w.r.p, w.v.p, sc, w.p.p, w.rh.p);
        }
        if (pc) {
            ProfScope ps(A.prof, s, PBSM3D_K_PRECOND);
            M(s, w.p.p, w.ph.p);
        }
        int64_t nb = 0;
        {
            ProfScope ps(A.prof, s, spmv_id);
            nb = spmv_dots<T>(s, A, ph, w.v.p, w.rh.p, w.partial.p, nblk, stop);  // v = A ph; rh.v, v.v
        }
        {
            ProfScope ps(A.prof, s, PBSM3D_K_ORTHO);
            bicg_s_dev<T>(s, n, w.v.p, w.r.p, sc, w.s.p, w.ss_partial.p, nvec, w.partial.p, nb);  // alpha, s, s.s
        }
        if (pc) {
            ProfScope ps(A.prof, s, PBSM3D_K_PRECOND);
            M(s, w.s.p, w.sh.p);
        }
        {
            ProfScope ps(A.prof, s, spmv_id);
            nb = spmv_dots<T>(s, A, sh, w.t.p, w.s.p, w.partial.p, nblk, stop);  // t = A sh; s.t, t.t
        }
        {
            ProfScope ps(A.prof, s, PBSM3D_K_ORTHO);
            // omega, x, r, (rh.r, r.r)
            bicg_xr_dev<T>(s, n, ph, sh, w.s.p, w.t.p, w.rh.p, sc, x, w.r.p, w.xr_partial.p, nvec, w.partial.p, nb,
                           w.ss_partial.p, nvec);
            // rho, rr, the stop test, and the scalars into host slot it % 2 (no copy launch)
            bicg_stage(s, w.xr_partial.p, nvec, 2, 1, sc, stop_abs, w.hsc_dev + SC_COUNT * (it & 1), restart_eps);
        }
        PB_HIP(hipEventRecord(w.ev[it & 1], s));
    };
    if (maxit <= 0) return 0;
    iterate(1);
    for (int it = 1;; ++it) {
        if (it < maxit) iterate(it + 1);
        PB_HIP(hipEventSynchronize(w.ev[it & 1]));
        const double* hk = h + SC_COUNT * (it & 1);
        rr = std::sqrt(hk[SC_RR]);
        *res_out = rr;
        if (hk[SC_STOP] != 0.0 || it >= maxit) {
            if (restarts) *restarts += (int)hk[SC_NRESTART];
            return it;
        }
    }
}
}  // namespace

SolveInfo bicgstab_solve(hipStream_t s, BicgWork<double>& w, GmresWork<double>& gw, const CsrOp<double>& A,
                         const PcFn<double>& M, const double* b, double* x, int maxit, double tol) {
    SolveInfo info;
    w.ensure(A.n);
    GmresWork<double>& dummy = gw;
    const double nb = norm_of(s, dummy, A.n, b);
    if (nb == 0.0) {
        PB_HIP(hipMemsetAsync(x, 0, sizeof(double) * A.n, s));
        info.converged = true;
        return info;
    }
    double beta = residual_norm(s, dummy, A, x, b, w.r.p);
    if (drop_worse_start(s, info, beta, nb, x, A.n)) beta = residual_norm(s, dummy, A, x, b, w.r.p);
    while (true) {
        info.resid = beta / nb;
        if (info.resid <= tol) {
            info.converged = true;
            break;
        }
        if (info.iters >= maxit) break;
        double est = 0;
        info.iters += bicg_run<double>(s, w, A, M, b, x, false, 0.9 * tol * nb, maxit - info.iters, &est);
        info.cycles += 1;
        beta = residual_norm(s, dummy, A, x, b, w.r.p);
    }
    return info;
}

namespace {
void scale_into(hipStream_t s, int64_t n, const double* r, double a, double* out) { scale_host<double>(s, n, r, a, out); }
void scale_into(hipStream_t s, int64_t n, const double* r, double a, float* out) { scale_f64_to_f32_host(s, n, r, a, out); }
void add_into(hipStream_t s, int64_t n, double a, const double* d, double* x) { axpy_f64(s, n, a, d, x); }
void add_into(hipStream_t s, int64_t n, double a, const float* d, double* x) { axpy_f32_into_f64(s, n, a, d, x); }
}  // namespace

template <typename T>
SolveInfo bicgstab_split(hipStream_t s, BicgWork<T>& w, GmresWork<T>& gw, const CsrOp<double>& A64,
                         const CsrOp<T>& Ahat, const double* b, double* x, double* r, T* work1, T* work2, T* work3,
                         int maxit, double tol, double inner_rtol, double inner_rtol_cold, const CoarseFn& coarse,
                         double restart_eps32) {
    SolveInfo info;
    const int64_t n = A64.n;
    w.ensure(n);
    const double nb = norm_of(s, gw, n, b);
    if (nb == 0.0) {
        PB_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
        info.converged = true;
        return info;
    }
    double beta = residual_norm(s, gw, A64, x, b, r);
    if (drop_worse_start(s, info, beta, nb, x, n)) {
        beta = residual_norm(s, gw, A64, x, b, r);
        if (inner_rtol_cold > 0) inner_rtol = inner_rtol_cold;  // the start is x0 = 0 after all
    }
    const int pclass = Ahat.prof_class >= 0 ? Ahat.prof_class : PBSM3D_K_PRECOND;
    while (true) {
        info.resid = beta / nb;
        if (info.resid <= tol) {
            info.converged = true;
            break;
        }
        if (info.iters >= maxit) break;
        if (coarse) {
            coarse(s, r, x);
            beta = residual_norm(s, gw, A64, x, b, r);
            if (info.cycles == 0) info.resid_coarse = beta / nb;
            info.resid = beta / nb;
            if (info.resid <= tol) {
                info.converged = true;
                break;
            }
        }
        scale_into(s, n, r, 1.0 / beta, work1);  // r / ||r||
        {
            ProfScope ps(Ahat.prof, s, pclass);
            split_op<T>(1, s, *Ahat.co, Ahat.F, Ahat.L, *Ahat.split, work1, work2, Ahat.mode);  // rhat
        }
        PB_HIP(hipMemsetAsync(work3, 0, n * sizeof(T), s));
        double stop = 0.5 * tol * nb / beta;
        if (stop < inner_rtol) stop = inner_rtol;
        double est = 0;
        info.iters += bicg_run<T>(s, w, Ahat, PcFn<T>(), work2, work3, true, stop, maxit - info.iters, &est, true,
                                  sizeof(T) == 4 ? restart_eps32 : 0.0, &info.restarts);
        {
            ProfScope ps(Ahat.prof, s, pclass);
            split_op<T>(2, s, *Ahat.co, Ahat.F, Ahat.L, *Ahat.split, work3, work1, Ahat.mode);  // d = (D~+U)^-1 dhat
        }
        add_into(s, n, beta, work1, x);
        info.cycles += 1;
        info.refinements += 1;
        beta = residual_norm(s, gw, A64, x, b, r);
    }
    return info;
}

template SolveInfo bicgstab_split<double>(hipStream_t, BicgWork<double>&, GmresWork<double>&, const CsrOp<double>&,
                                          const CsrOp<double>&, const double*, double*, double*, double*, double*,
                                          double*, int, double, double, double, const CoarseFn&, double);
template SolveInfo bicgstab_split<float>(hipStream_t, BicgWork<float>&, GmresWork<float>&, const CsrOp<double>&,
                                         const CsrOp<float>&, const double*, double*, double*, float*, float*, float*,
                                         int, double, double, double, const CoarseFn&, double);

SolveInfo bicgstab_mixed(hipStream_t s, BicgWork<float>& w, GmresWork<float>& gw, const CsrOp<double>& A64,
                         const CsrOp<float>& A32, const PcFn<float>& M32, const double* b, double* x, double* r,
                         float* r32, float* d32, int maxit, double tol, double inner_rtol, double restart_eps32) {
    SolveInfo info;
    w.ensure(A64.n);
    const double nb = norm_of(s, gw, A64.n, b);
    if (nb == 0.0) {
        PB_HIP(hipMemsetAsync(x, 0, sizeof(double) * A64.n, s));
        info.converged = true;
        return info;
    }
    double beta = residual_norm(s, gw, A64, x, b, r);
    if (drop_worse_start(s, info, beta, nb, x, A64.n)) beta = residual_norm(s, gw, A64, x, b, r);
    while (true) {
        info.resid = beta / nb;
        if (info.resid <= tol) {
            info.converged = true;
            break;
        }
        if (info.iters >= maxit) break;
        scale_f64_to_f32_host(s, A64.n, r, 1.0 / beta, r32);
        PB_HIP(hipMemsetAsync(d32, 0, A64.n * sizeof(float), s));
        double stop = 0.5 * tol * nb / beta;
        if (stop < inner_rtol) stop = inner_rtol;
        double est = 0;
        info.iters += bicg_run<float>(s, w, A32, M32, r32, d32, true, stop, maxit - info.iters, &est, false,
                                      restart_eps32, &info.restarts);
        axpy_f32_into_f64(s, A64.n, beta, d32, x);
        info.cycles += 1;
        info.refinements += 1;
        beta = residual_norm(s, gw, A64, x, b, r);
    }
    return info;
}

template struct BicgWork<double>;
template struct BicgWork<float>;
template struct GmresWork<double>;
template struct GmresWork<float>;
template int gmres_cycle<double>(hipStream_t, GmresWork<double>&, const CsrOp<double>&, const PcFn<double>&,
                                 const double*, double, double, int, double*, double*);
template int gmres_cycle<float>(hipStream_t, GmresWork<float>&, const CsrOp<float>&, const PcFn<float>&, const float*,
                                double, double, int, float*, double*);

}  // namespace pb
