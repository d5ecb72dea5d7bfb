// Restarted, right-preconditioned GMRES on the device, and FP64 iterative
// refinement around FP32 GMRES cycles (the mixed-precision solve).
//
// Replaces Belos "GMRES" (Num Blocks 30, Maximum Iterations 1000, Convergence
// Tolerance 1e-8, right preconditioner) of src/math/LinearAlgebra.cpp:164-195,
// 228-252.  Convergence is ||b - A x||_2 / ||b||_2 <= tol, the denominator
// being ||b|| whatever the initial guess, so a warm start never loosens the
// target.
#pragma once

#include <functional>

#include "kernels.hpp"

namespace pb {

template <typename T>
struct CsrOp {
    int64_t n = 0;
    const int32_t* rp = nullptr;
    const int32_t* col = nullptr;
    const T* val = nullptr;
    Profiler* prof = nullptr;  // optional launch timing
    // optional extruded-prism structure: when set, SpMV walks face columns (spmv_col)
    int64_t F = 0;
    int L = 0;
    const int32_t* neigh = nullptr;
    const uint16_t* slots = nullptr;
    // optional operator held in the D-ILU position order (diag/lat/vert of the
    // factorisation); when set, every vector of the solve is in position order
    const ColorOrder* co = nullptr;
    const DiluFactors<T>* pos = nullptr;
    int variant = 0;  // launch shape of the position-order SpMV (tune_pos_spmv)
    // optional split (Eisenstat) D-ILU operator Ahat = D~ (D~ + L)^-1 A (D~ + U)^-1 (split_op)
    DiluFactors<T>* split = nullptr;
    int mode = 0;          // split_op mode of the split operator's sweeps (XCD mapping, NT loads)
    int prof_class = -1;   // profile class of its applications (-1: the SpMV class)
};

template <typename T>
inline void op_spmv(hipStream_t s, const CsrOp<T>& A, const T* x, T* y) {
    if (A.split)
        split_op<T>(0, s, *A.co, A.F, A.L, *A.split, x, y, A.mode);
    else if (A.pos)
        pos_spmv_variant<T>(s, A.variant, *A.co, A.F, A.L, *A.pos, x, y);
    else if (A.neigh)
        spmv_col<T>(s, A.F, A.L, A.rp, A.val, A.neigh, A.slots, x, nullptr, y, nullptr);
    else
        spmv<T>(s, A.n, A.rp, A.col, A.val, x, y);
}

// r = b - A x with per-block partial sums of r.r; returns the number of partials
inline int op_residual(hipStream_t s, const CsrOp<double>& A, const double* x, const double* b, double* r,
                       double* partial) {
    if (A.pos) {
        pos_spmv<double>(s, *A.co, A.F, A.L, *A.pos, x, b, r, partial);
        return (int)pos_spmv_blocks(A.F, A.L);
    }
    if (A.neigh) {
        spmv_col<double>(s, A.F, A.L, A.rp, A.val, A.neigh, A.slots, x, b, r, partial);
        return (int)spmv_col_blocks(A.F);
    }
    const int nblk = (int)((A.n + kBlock - 1) / kBlock);
    residual<double>(s, A.n, A.rp, A.col, A.val, x, b, r, partial, nblk);
    return nblk;
}

// z = M^-1 r on the stream
template <typename T>
using PcFn = std::function<void(hipStream_t, const T*, T*)>;

struct SolveInfo {
    int iters = 0;
    int cycles = 0;
    int refinements = 0;
    double resid = 0;  // explicit ||b - A x|| / ||b||
    double resid0 = 1;  // the same for the initial guess
    bool converged = false;
    bool cold_fallback = false;  // the initial guess was worse than x0 = 0 and was dropped
    double resid_coarse = 0;     // ||b - A x|| / ||b|| after the first coarse correction (0: none)
    int restarts = 0;            // shadow-residual restarts of the FP32 BiCGStab runs
};

// Device workspace of one GMRES instance of storage type T.
template <typename T>
struct GmresWork {
    int64_t n = 0;
    int m = 0;
    int passes = 2;  // Gram-Schmidt passes allowed per Arnoldi step
    DevBuf<T> V, z, u;
    DevBuf<double> partial;  // kReduceBlocks * (m + 2)
    DevBuf<double> dsc;      // device scalars
    HostBuf<double> hsc;     // pinned mirror
    void ensure(int64_t n_, int m_);
};

// One Arnoldi cycle of at most m steps from v0 = r0 / beta; accumulates
// x_update = M^-1 V y.  Stops early when |g_{j+1}| <= stop_abs.
// Returns the number of steps taken; *resid_est = |g_k| (implicit residual).
template <typename T>
int gmres_cycle(hipStream_t s, GmresWork<T>& w, const CsrOp<T>& A, const PcFn<T>& M, const T* r0, double beta,
                double stop_abs, int max_steps, T* x_update, double* resid_est);

// Full FP64 restarted GMRES: x holds the initial guess on entry.
SolveInfo gmres_solve(hipStream_t s, GmresWork<double>& w, const CsrOp<double>& A, const PcFn<double>& M,
                      const double* b, double* x, double* r, int m, int maxit, double tol);

// FP64 iterative refinement around FP32 GMRES cycles.
SolveInfo mixed_solve(hipStream_t s, GmresWork<float>& w, const CsrOp<double>& A64, const CsrOp<float>& A32,
                      const PcFn<float>& M32, const double* b, double* x, double* r, float* r32, float* d32, int m,
                      int maxit, double tol, double inner_rtol);

// device norm helpers (synchronous: return to host)
double dev_norm2(hipStream_t s, GmresWork<double>& w, int64_t n, const double* v);

// BiCGStab (van der Vorst 1992), right-preconditioned, as an alternative
// Krylov method: two SpMV + two preconditioner applications per iteration and
// a fixed set of vectors, so no Gram-Schmidt cost growing with the iteration.
template <typename T>
struct BicgWork {
    int64_t n = 0;
    DevBuf<T> r, rh, p, v, s, t, ph, sh;
    DevBuf<double> partial, dsc;
    DevBuf<double> ss_partial;  // per-block s.s of k_bicg_s_dev (kept apart from the operator's dots)
    DevBuf<double> xr_partial;  // per-block (rh.r, r.r) of k_bicg_xr_dev (read by the closing stage)
    HostBuf<double> hsc;        // two slots of the scalars (iterations k and k+1 in flight)
    double* hsc_dev = nullptr;  // hsc in the device address space (the stage kernel writes it)
    hipEvent_t ev[2] = {};
    BicgWork() = default;
    BicgWork(const BicgWork&) = delete;
    BicgWork& operator=(const BicgWork&) = delete;
    ~BicgWork() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    void ensure(int64_t n_);
};

// Default restart threshold of the shadow residual in FP32 BiCGStab runs: the vectors carry
// 24-bit mantissas, and once |rh.r| < 1e-5 ||rh|| ||r|| rho keeps about two significant digits,
// alpha = rho / rh.v swings by orders of magnitude and the residual can wander for hundreds of
// iterations (a near breakdown of the Lanczos recurrence); restarting with rh = r removes it.
// FP64 runs (the deposition solve) converge through ratios of 1e-10 and never restart.
constexpr double kBicgRestartEps32 = 1e-5;

SolveInfo bicgstab_solve(hipStream_t s, BicgWork<double>& w, GmresWork<double>& gw, const CsrOp<double>& A,
                         const PcFn<double>& M, const double* b, double* x, int maxit, double tol);
// BiCGStab on the split-preconditioned system (Eisenstat's trick) inside FP64
// iterative refinement: each cycle takes the FP64 residual r = b - A64 x,
// transforms r / ||r|| with Ahat's left factor, runs unpreconditioned BiCGStab on
// Ahat (storage T) from zero to a relative inner_rtol, maps the correction back
// with the right factor and adds it to x in FP64.  Ahat must have .split set;
// work1, work2 are n-vectors of T.  inner_rtol_cold (> 0) replaces inner_rtol when
// the given start is dropped for x0 = 0 (a warm start worse than zero).  coarse (may be
// empty): a correction of x from the FP64 residual r, applied at the start of every
// cycle before the inner run (the learned coarse correction); the residual is then
// recomputed, and SolveInfo::resid_coarse records it after the first correction.
// restart_eps32: shadow-residual restart threshold of the runs when T = float (bicg_stage;
// FP64 runs never restart).
using CoarseFn = std::function<void(hipStream_t, const double*, double*)>;
template <typename T>
SolveInfo bicgstab_split(hipStream_t s, BicgWork<T>& w, GmresWork<T>& gw, const CsrOp<double>& A64,
                         const CsrOp<T>& Ahat, const double* b, double* x, double* r, T* work1, T* work2, T* work3,
                         int maxit, double tol, double inner_rtol, double inner_rtol_cold = 0.0,
                         const CoarseFn& coarse = CoarseFn(), double restart_eps32 = kBicgRestartEps32);
// FP64 refinement around FP32 BiCGStab runs (FP64 residuals, FP64 dot accumulation)
SolveInfo bicgstab_mixed(hipStream_t s, BicgWork<float>& w, GmresWork<float>& gw, const CsrOp<double>& A64,
                         const CsrOp<float>& A32, const PcFn<float>& M32, const double* b, double* x, double* r,
                         float* r32, float* d32, int maxit, double tol, double inner_rtol,
                         double restart_eps32 = kBicgRestartEps32);

}  // namespace pb
