// Batched parameter ensemble (include/pbsm3d_batch.h).
//
// Members of equal nLayer are concatenated into one union: faces member-major, so the
// union's matrices are block diagonal and a union context (pbsm3d_create on the
// concatenated mesh) assembles every member in one launch, with the per-member
// parameters (z0_snow, settling_scale) as per-face arrays.  The solves keep every
// member apart: the union's colour order is member-major inside each colour, the
// sweeps run from workgroup tables that never span two members (split_apply_batch),
// the element-wise BiCGStab kernels from chunk tables that never span two members, and
// one stage kernel (a workgroup per member) turns each member's partial sums into that
// member's alpha, omega, rho, residual norm and stop flag.  A member that has converged
// stops: its workgroups return at entry.  The FP64 refinement loop around the inner
// runs is per member as well (its own ||b||, residual, inner target and convergence),
// as the single-member solve does it (bicgstab_split in krylov.hip).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <exception>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pbsm3d.h"
#include "../../include/pbsm3d_batch.h"
#include "context.hpp"
#include "kernels.hpp"

namespace pb {
namespace {

// per-member scalars of the batched BiCGStab, BS_COUNT doubles per member
enum BatchSc {
    BS_RHO_PREV = 0, BS_RHO, BS_ALPHA, BS_OMEGA, BS_RHV, BS_SS, BS_ST, BS_TT, BS_RR, BS_FLAG, BS_STOP, BS_STOPABS,
    BS_ITERS, BS_RHNN, BS_RESTART, BS_NRESTART,  // shadow-residual restarts as in bicg_stage
    BS_COUNT
};
enum StageMode { ST_START = 0, ST_ALPHA, ST_OMEGA, ST_RHO };
constexpr int kChunk = 1024;       // elements per workgroup of the element-wise kernels (4 per thread)
constexpr int kStage = 256;        // threads of a stage workgroup
constexpr double kPresent = 1e-12; // getRhsMax threshold (PBSM3D.hpp:417-418)

// ---------------------------------------------------------------------------
// element-wise kernels over chunk tables ({q0, q1, member, 0}, at most kChunk elements)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_seg_dots2(const int4* __restrict__ ch, const T* __restrict__ a,
                                                      const T* __restrict__ b, double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    const int4 e = ch[blockIdx.x];
    double d1 = 0, d2 = 0;
    for (int64_t q = (int64_t)e.x + threadIdx.x; q < e.y; q += kBlock) {
        const double av = a[q], bv = b[q];
        d1 = fma(av, bv, d1);
        d2 = fma(bv, bv, d2);
    }
    d1 = block_sum(d1, red);
    d2 = block_sum(d2, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = d1;
        partial[2 * blockIdx.x + 1] = d2;
    }
}

// p = r + beta (p - omega v), beta = (rho / rho_prev) (alpha / omega) of the member; after a
// restart of the member's shadow residual, p = r and rh = r
template <typename T>
__global__ void __launch_bounds__(kBlock) k_seg_p(const int4* __restrict__ ch, const T* __restrict__ r,
                                                  const T* __restrict__ v, const double* __restrict__ sc,
                                                  T* __restrict__ p, T* __restrict__ rh) {
    const int4 e = ch[blockIdx.x];
    const double* m = sc + (int64_t)e.z * BS_COUNT;
    if (m[BS_STOP] != 0.0) return;
    if (m[BS_RESTART] != 0.0) {
        for (int64_t q = (int64_t)e.x + threadIdx.x; q < e.y; q += kBlock) p[q] = rh[q] = r[q];
        return;
    }
    const double omega = m[BS_OMEGA];
    const double beta = (m[BS_RHO] / m[BS_RHO_PREV]) * (m[BS_ALPHA] / omega);
    for (int64_t q = (int64_t)e.x + threadIdx.x; q < e.y; q += kBlock)
        p[q] = (T)fma(beta, fma(-omega, (double)v[q], (double)p[q]), (double)r[q]);
}

// s = r - alpha v; partial[chunk] = s.s
template <typename T>
__global__ void __launch_bounds__(kBlock) k_seg_s(const int4* __restrict__ ch, const T* __restrict__ v,
                                                  const T* __restrict__ r, const double* __restrict__ sc,
                                                  T* __restrict__ s_out, double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    const int4 e = ch[blockIdx.x];
    const double* m = sc + (int64_t)e.z * BS_COUNT;
    if (m[BS_STOP] != 0.0) return;  // uniform over the workgroup; the stage skips stopped members
    const double alpha = m[BS_ALPHA];
    double ss = 0;
    for (int64_t q = (int64_t)e.x + threadIdx.x; q < e.y; q += kBlock) {
        const T o = (T)fma(-alpha, (double)v[q], (double)r[q]);
        s_out[q] = o;
        ss = fma((double)o, (double)o, ss);
    }
    ss = block_sum(ss, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = ss;
}

// x += alpha ph + omega sh;  r = s - omega t;  partial[2 chunk] = (rh.r, r.r)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_seg_xr(const int4* __restrict__ ch, const T* __restrict__ ph,
                                                   const T* __restrict__ sh, const T* __restrict__ sv,
                                                   const T* __restrict__ tv, const T* __restrict__ rh,
                                                   const double* __restrict__ sc, T* __restrict__ x, T* __restrict__ r,
                                                   double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    const int4 e = ch[blockIdx.x];
    const double* m = sc + (int64_t)e.z * BS_COUNT;
    if (m[BS_STOP] != 0.0) return;
    const double alpha = m[BS_ALPHA], omega = m[BS_OMEGA];
    double d1 = 0, d2 = 0;
    for (int64_t q = (int64_t)e.x + threadIdx.x; q < e.y; q += kBlock) {
        x[q] = (T)fma(omega, (double)sh[q], fma(alpha, (double)ph[q], (double)x[q]));
        const T o = (T)fma(-omega, (double)tv[q], (double)sv[q]);
        r[q] = o;
        d1 = fma((double)rh[q], (double)o, d1);
        d2 = fma((double)o, (double)o, d2);
    }
    d1 = block_sum(d1, red);
    d2 = block_sum(d2, red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = d1;
        partial[2 * blockIdx.x + 1] = d2;
    }
}

// partial[chunk] = sum v^2 (MAXABS: max |v|) of an FP64 vector
template <bool MAXABS>
__global__ void __launch_bounds__(kBlock) k_seg_norm_part(const int4* __restrict__ ch, const double* __restrict__ v,
                                                          double* __restrict__ partial) {
    __shared__ double red[kBlock / 64];
    const int4 e = ch[blockIdx.x];
    double a = 0;
    for (int64_t q = (int64_t)e.x + threadIdx.x; q < e.y; q += kBlock) {
        const double x = v[q];
        a = MAXABS ? fmax(a, fabs(x)) : fma(x, x, a);
    }
    if (MAXABS) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off, 64));
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0) red[wid] = a;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mx = 0;
            for (int k = 0; k < kBlock / 64; ++k) mx = fmax(mx, red[k]);
            partial[blockIdx.x] = mx;
        }
    } else {
        a = block_sum(a, red);
        if (threadIdx.x == 0) partial[blockIdx.x] = a;
    }
}

// out[m] = sum (MAXABS: max) of the member's chunk partials, in a fixed order
template <bool MAXABS>
__global__ void __launch_bounds__(kStage) k_seg_reduce(const int64_t* __restrict__ chunk_begin,
                                                       const double* __restrict__ partial, double* __restrict__ out) {
    __shared__ double red[kStage / 64];
    const int m = blockIdx.x;
    double a = 0;
    for (int64_t b = chunk_begin[m] + threadIdx.x; b < chunk_begin[m + 1]; b += kStage)
        a = MAXABS ? fmax(a, partial[b]) : a + partial[b];
    if (MAXABS) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off, 64));
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0) red[wid] = a;
        __syncthreads();
        if (threadIdx.x == 0) {
            double mx = 0;
            for (int k = 0; k < kStage / 64; ++k) mx = fmax(mx, red[k]);
            out[m] = mx;
        }
    } else {
        a = block_sum(a, red);
        if (threadIdx.x == 0) out[m] = a;
    }
}

// out = f[member] * in (FP64 in, T out)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_seg_scale(const int4* __restrict__ ch, const double* __restrict__ in,
                                                      const double* __restrict__ f, T* __restrict__ out) {
    const int4 e = ch[blockIdx.x];
    const double a = f[e.z];
    for (int64_t q = (int64_t)e.x + threadIdx.x; q < e.y; q += kBlock) out[q] = (T)(a * in[q]);
}

// x += a[member] * d (FP64 x)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_seg_axpy(const int4* __restrict__ ch, const double* __restrict__ a,
                                                     const T* __restrict__ d, double* __restrict__ x) {
    const int4 e = ch[blockIdx.x];
    const double am = a[e.z];
    if (am == 0.0) return;
    for (int64_t q = (int64_t)e.x + threadIdx.x; q < e.y; q += kBlock) x[q] = fma(am, (double)d[q], x[q]);
}

// ---------------------------------------------------------------------------
// stage: one workgroup per member turns that member's partials into its scalars
// ---------------------------------------------------------------------------
__device__ __forceinline__ double stage_sum(double v, double* red, double* bc) {
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) *bc = s;
    __syncthreads();
    const double r = *bc;
    __syncthreads();
    return r;
}

// mode ST_START: (rh.r, r.r) chunk pairs of the starting residual; ST_ALPHA: the first
// operator application's (rh.v) pairs; ST_OMEGA: the second's (s.t, t.t) pairs and the
// s.s chunk partials; ST_RHO: the x/r kernel's (rh.r, r.r) chunk pairs.  op_rng holds per
// member kMaxColors [begin, end) ranges of operator pairs (empty when a colour has no
// faces of the member).  target[m] (ST_START): relative stop target, < 0 = inactive.
__global__ void __launch_bounds__(kStage) k_seg_stage(int mode, const int64_t* __restrict__ chunk_begin,
                                                      const int64_t* __restrict__ op_rng,
                                                      const double* __restrict__ pairs, const double* __restrict__ ssp,
                                                      double* __restrict__ sc, const double* __restrict__ target,
                                                      double* __restrict__ host_slot, double restart_eps) {
    __shared__ double red[kStage / 64];
    __shared__ double bc;
    const int m = blockIdx.x;
    double* S = sc + (int64_t)m * BS_COUNT;
    const bool stopped = mode != ST_START && S[BS_STOP] != 0.0;  // uniform over the workgroup
    if (!stopped) {
        const int64_t c0 = chunk_begin[m], c1 = chunk_begin[m + 1];
        if (mode == ST_START || mode == ST_RHO) {
            double a = 0, b = 0;
            for (int64_t k = c0 + threadIdx.x; k < c1; k += kStage) {
                a += pairs[2 * k];
                b += pairs[2 * k + 1];
            }
            a = stage_sum(a, red, &bc);
            b = stage_sum(b, red, &bc);
            if (threadIdx.x == 0) {
                if (mode == ST_START) {
                    S[BS_RHO] = a;
                    S[BS_RR] = b;
                    S[BS_RHO_PREV] = 1;
                    S[BS_ALPHA] = 1;
                    S[BS_OMEGA] = 1;
                    S[BS_FLAG] = 0;
                    S[BS_ITERS] = 0;
                    S[BS_RHNN] = a;  // rh = r
                    S[BS_RESTART] = 0;
                    S[BS_NRESTART] = 0;
                    const double tg = target[m], rr = sqrt(b);
                    S[BS_STOPABS] = tg < 0 ? 0.0 : tg * rr;
                    S[BS_STOP] = (tg < 0 || rr <= S[BS_STOPABS] || !isfinite(rr)) ? 1.0 : 0.0;
                } else {
                    S[BS_RHO_PREV] = S[BS_RHO];
                    S[BS_RHO] = a;
                    S[BS_RR] = b;
                    S[BS_ITERS] += 1;
                    if (!isfinite(b)) S[BS_FLAG] = 3;
                    const double rr = sqrt(b);
                    if (S[BS_FLAG] != 0.0 || rr <= S[BS_STOPABS] || !isfinite(rr)) S[BS_STOP] = 1;
                    const bool restart =
                        S[BS_STOP] == 0.0 && restart_eps > 0.0 && fabs(a) < restart_eps * sqrt(S[BS_RHNN] * b);
                    S[BS_RESTART] = restart ? 1.0 : 0.0;
                    if (restart) {
                        S[BS_RHO] = b;
                        S[BS_RHNN] = b;
                        S[BS_NRESTART] += 1.0;
                    }
                }
            }
        } else {
            const int64_t* rg = op_rng + (int64_t)m * 2 * kMaxColors;
            double a = 0, b = 0;
            for (int c = 0; c < kMaxColors; ++c)
                for (int64_t k = rg[2 * c] + threadIdx.x; k < rg[2 * c + 1]; k += kStage) {
                    a += pairs[2 * k];
                    b += pairs[2 * k + 1];
                }
            a = stage_sum(a, red, &bc);
            if (mode == ST_ALPHA) {
                if (threadIdx.x == 0) {
                    const bool ok = a != 0.0 && isfinite(a);
                    S[BS_RHV] = a;
                    S[BS_ALPHA] = ok ? S[BS_RHO] / a : 0.0;
                    if (!ok) S[BS_FLAG] = 1;  // breakdown: the member stops at the end of the iteration
                }
            } else {
                b = stage_sum(b, red, &bc);
                double ss = 0;
                for (int64_t k = chunk_begin[m] + threadIdx.x; k < chunk_begin[m + 1]; k += kStage) ss += ssp[k];
                ss = stage_sum(ss, red, &bc);
                if (threadIdx.x == 0) {
                    // minimal-residual omega, enlarged when |cos(t, s)| < 0.7 (as the single solve)
                    const double st = a, tt = b;
                    double omega = 0;
                    int flag = 0;
                    if (tt > 0.0 && isfinite(st) && isfinite(tt)) {
                        omega = st / tt;
                        const double cs = st / sqrt(tt * ss);
                        if (isfinite(cs) && fabs(cs) < 0.7 && fabs(cs) > 0.0) omega *= 0.7 / fabs(cs);
                        if (omega == 0.0) flag = -2;
                    } else {
                        flag = 2;
                    }
                    S[BS_ST] = st;
                    S[BS_TT] = tt;
                    S[BS_SS] = ss;
                    S[BS_OMEGA] = omega;
                    if (flag == 2 || (flag == -2 && S[BS_FLAG] == 0.0)) S[BS_FLAG] = 2;
                }
            }
        }
    }
    if (host_slot) {  // every member's scalars into the host-visible slot the host waits on
        __syncthreads();
        if (threadIdx.x < BS_COUNT) host_slot[(int64_t)m * BS_COUNT + threadIdx.x] = S[threadIdx.x];
    }
}

// ---------------------------------------------------------------------------
// tables of one system (the suspension system of a union, or its deposition system)
// ---------------------------------------------------------------------------
struct Segments {
    int M = 0;
    int64_t F = 0, N = 0;
    int L = 0;
    SweepTables sweep;
    DevBuf<int4> chunks;
    int64_t nchunk = 0;
    DevBuf<int64_t> chunk_begin;  // [M + 1]
    DevBuf<int64_t> op_rng;       // [M][kMaxColors][2]
    DevBuf<double> part;          // nchunk doubles (norm / max partials)
    DevBuf<double> red;           // M doubles
    HostBuf<double> h_red;        // M doubles
};

// color: colour of every union face; face_off: member m owns faces [face_off[m], face_off[m+1])
void build_segments(Segments& sg, const ColorOrder& co, const std::vector<int32_t>& color,
                    const std::vector<int64_t>& face_off, int L, hipStream_t s) {
    const int M = (int)face_off.size() - 1;
    const int64_t F = face_off[M];
    const int nc = co.ncolor;
    if (nc > kMaxColors) throw Error(PBSM3D_EINVAL, "too many colours for the batch tables");
    if (F * L >= INT32_MAX) throw Error(PBSM3D_EINVAL, "batch union too large for 32-bit positions");
    sg.M = M;
    sg.F = F;
    sg.L = L;
    sg.N = F * L;
    // faces of colour c in member m, and the start of that member's run of positions
    std::vector<int64_t> cnt((size_t)nc * M, 0), kb((size_t)nc * M, 0);
    for (int m = 0; m < M; ++m)
        for (int64_t i = face_off[m]; i < face_off[m + 1]; ++i) cnt[(size_t)color[i] * M + m]++;
    for (int c = 0; c < nc; ++c) {
        int64_t k = co.off[c];
        for (int m = 0; m < M; ++m) {
            kb[(size_t)c * M + m] = k;
            k += cnt[(size_t)c * M + m];
        }
        if (k != co.off[c + 1]) throw Error(PBSM3D_EINVAL, "batch colour order is not member-major");
    }
    // sweep workgroup tables and each member's operator partial ranges
    std::vector<int64_t> rng((size_t)M * 2 * kMaxColors, 0);
    sg.sweep.ncolor = nc;
    int64_t off = 0;
    for (int c = 0; c < nc; ++c) {
        std::vector<int4> tab;
        for (int m = 0; m < M; ++m) {
            const int64_t k0 = kb[(size_t)c * M + m], n = cnt[(size_t)c * M + m];
            const int64_t b0 = (int64_t)tab.size();
            // workgroups start on tile boundaries (a), so each wave stays inside one tile
            if (n > 0)
                for (int64_t a = tile_floor(k0); a < k0 + n; a += kBlock)
                    tab.push_back(make_int4((int)std::max(a, k0), (int)std::min(a + kBlock, k0 + n), m, (int)a));
            rng[((size_t)m * kMaxColors + c) * 2] = off + b0;
            rng[((size_t)m * kMaxColors + c) * 2 + 1] = off + (int64_t)tab.size();
        }
        sg.sweep.nblk[c] = (int64_t)tab.size();
        sg.sweep.part_off[c] = off;
        off += (int64_t)tab.size();
        if (!tab.empty()) {
            sg.sweep.tab[c].alloc(tab.size());
            PB_HIP(hipMemcpyAsync(sg.sweep.tab[c].p, tab.data(), tab.size() * sizeof(int4), hipMemcpyHostToDevice, s));
            PB_HIP(hipStreamSynchronize(s));
        }
    }
    sg.sweep.total = off;
    // element chunks, member-major: the runs of consecutive rows (position order: tiles of
    // kTile positions, each tile layer by layer) that belong to one member, walked in row
    // order and cut at kChunk rows; a tile inside one member's run of a colour is one run
    std::vector<int32_t> member_of(F);
    for (int c = 0; c < nc; ++c)
        for (int m = 0; m < M; ++m) {
            const int64_t k0 = kb[(size_t)c * M + m];
            std::fill(member_of.begin() + k0, member_of.begin() + k0 + cnt[(size_t)c * M + m], (int32_t)m);
        }
    std::vector<std::vector<int4>> per(M);
    auto add_run = [&](int m, int64_t q0, int64_t q1) {
        std::vector<int4>& v = per[m];
        while (q0 < q1) {
            if (!v.empty() && v.back().y == q0 && v.back().y - v.back().x < kChunk) {
                const int64_t e = std::min<int64_t>(q1, (int64_t)v.back().x + kChunk);
                v.back().y = (int)e;
                q0 = e;
            } else {
                const int64_t e = std::min<int64_t>(q1, q0 + kChunk);
                v.push_back(make_int4((int)q0, (int)e, m, 0));
                q0 = e;
            }
        }
    };
    for (int64_t k0 = 0; k0 < F; k0 += kTile) {
        const int64_t w = std::min<int64_t>(kTile, F - k0);
        for (int z = 0; z < L; ++z) {
            const int64_t q0 = pos_row(k0, z, F, L);
            for (int64_t j = 0; j < w;) {
                const int m = member_of[k0 + j];
                int64_t e = j + 1;
                while (e < w && member_of[k0 + e] == m) ++e;
                add_run(m, q0 + j, q0 + e);
                j = e;
            }
        }
    }
    std::vector<int4> ch;
    std::vector<int64_t> cb(M + 1, 0);
    for (int m = 0; m < M; ++m) {
        cb[m] = (int64_t)ch.size();
        ch.insert(ch.end(), per[m].begin(), per[m].end());
    }
    cb[M] = (int64_t)ch.size();
    sg.nchunk = (int64_t)ch.size();
    sg.chunks.alloc(ch.size());
    sg.chunk_begin.alloc(M + 1);
    sg.op_rng.alloc(rng.size());
    sg.part.alloc(std::max<int64_t>(1, sg.nchunk));
    sg.red.alloc(M);
    sg.h_red.alloc(M);
    PB_HIP(hipMemcpyAsync(sg.chunks.p, ch.data(), ch.size() * sizeof(int4), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(sg.chunk_begin.p, cb.data(), cb.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(sg.op_rng.p, rng.data(), rng.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipStreamSynchronize(s));
}

// per-member ||v||_2 (MAXABS: max |v|) of an FP64 vector in position order, to the host
template <bool MAXABS>
std::vector<double> seg_reduce(hipStream_t s, Segments& sg, const double* v) {
    if (sg.nchunk > 0) {
        hipLaunchKernelGGL((k_seg_norm_part<MAXABS>), dim3((unsigned)sg.nchunk), dim3(kBlock), 0, s, sg.chunks.p, v,
                           sg.part.p);
        hipLaunchKernelGGL((k_seg_reduce<MAXABS>), dim3(sg.M), dim3(kStage), 0, s, sg.chunk_begin.p, sg.part.p,
                           sg.red.p);
    }
    PB_HIP(hipGetLastError());
    PB_HIP(hipMemcpyAsync(sg.h_red.p, sg.red.p, sg.M * sizeof(double), hipMemcpyDeviceToHost, s));
    PB_HIP(hipStreamSynchronize(s));
    std::vector<double> out(sg.h_red.p, sg.h_red.p + sg.M);
    if (!MAXABS)
        for (double& x : out) x = std::sqrt(x);
    return out;
}

// per-member factors to the device (a small synchronous upload)
void upload_factors(hipStream_t s, DevBuf<double>& d, const std::vector<double>& h) {
    d.alloc(h.size());
    PB_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
    PB_HIP(hipStreamSynchronize(s));
}

// Workspace of the batched BiCGStab on storage T.
template <typename T>
struct BatchWork {
    DevBuf<T> r, rh, p, v, sv, t;
    DevBuf<double> op_part, ss_part, xr_part, sc, tgt, fac;
    HostBuf<double> hsc;  // six slots of every member's scalars, device-visible: 0 / 1 for
                          // the stream-launched iterations, 2 .. 5 for the two captured graphs
    double* hsc_dev = nullptr;
    hipEvent_t ev[2] = {};
    int M = 0;
    // two captured graphs of two iterations each (slots 2, 3 and 4, 5), replayed in turn;
    // captured for one x and one set of factors
    hipGraphExec_t pair[2] = {};
    const void* pair_key[2] = {};
    BatchWork() = default;
    BatchWork(const BatchWork&) = delete;
    BatchWork& operator=(const BatchWork&) = delete;
    ~BatchWork() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        drop_graphs();
    }
    void drop_graphs() {
        for (hipGraphExec_t& g : pair)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
        pair_key[0] = pair_key[1] = nullptr;
    }
    void ensure(const Segments& sg) {
        for (DevBuf<T>* b : {&r, &rh, &p, &v, &sv, &t}) b->alloc(sg.N);
        op_part.alloc(std::max<int64_t>(1, 2 * sg.sweep.total));
        ss_part.alloc(std::max<int64_t>(1, sg.nchunk));
        xr_part.alloc(std::max<int64_t>(1, 2 * sg.nchunk));
        sc.alloc((size_t)sg.M * BS_COUNT);
        tgt.alloc(sg.M);
        if (!hsc.p || M != sg.M) {
            drop_graphs();
            hsc.alloc((size_t)6 * sg.M * BS_COUNT);
            PB_HIP(hipHostGetDevicePointer((void**)&hsc_dev, hsc.p, 0));
        }
        for (hipEvent_t& e : ev)
            if (!e) PB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        M = sg.M;
    }
};

struct RunCount {
    int64_t launches = 0, iterations = 0;
};

// The launches of one batched BiCGStab iteration after the first (the first copies p = r):
// p, v = Ahat p (with rh.v partials), alpha, s, t = Ahat s (with s.t, t.t), omega, x / r
// (with rh.r, r.r), rho and the stop test, whose scalars go to host slot `slot`.
template <typename T>
void enqueue_iteration(hipStream_t s, Segments& sg, const ColorOrder& co, DiluFactors<T>& d, BatchWork<T>& w, T* x,
                       int slot, bool first) {
    const int M = sg.M;
    const unsigned nch = (unsigned)sg.nchunk;
    const double* stop = w.sc.p + BS_STOP;
    if (first)
        PB_HIP(hipMemcpyAsync(w.p.p, w.r.p, sg.N * sizeof(T), hipMemcpyDeviceToDevice, s));
    else
        hipLaunchKernelGGL((k_seg_p<T>), dim3(nch), dim3(kBlock), 0, s, sg.chunks.p, w.r.p, w.v.p, w.sc.p, w.p.p,
                           w.rh.p);
    split_apply_batch<T>(s, co, sg.F, sg.L, d, sg.sweep, w.p.p, w.v.p, w.rh.p, w.op_part.p, stop, BS_COUNT);
    hipLaunchKernelGGL(k_seg_stage, dim3(M), dim3(kStage), 0, s, (int)ST_ALPHA, sg.chunk_begin.p, sg.op_rng.p,
                       w.op_part.p, (const double*)nullptr, w.sc.p, (const double*)nullptr, (double*)nullptr, 0.0);
    hipLaunchKernelGGL((k_seg_s<T>), dim3(nch), dim3(kBlock), 0, s, sg.chunks.p, w.v.p, w.r.p, w.sc.p, w.sv.p,
                       w.ss_part.p);
    split_apply_batch<T>(s, co, sg.F, sg.L, d, sg.sweep, w.sv.p, w.t.p, w.sv.p, w.op_part.p, stop, BS_COUNT);
    hipLaunchKernelGGL(k_seg_stage, dim3(M), dim3(kStage), 0, s, (int)ST_OMEGA, sg.chunk_begin.p, sg.op_rng.p,
                       w.op_part.p, w.ss_part.p, w.sc.p, (const double*)nullptr, (double*)nullptr, 0.0);
    hipLaunchKernelGGL((k_seg_xr<T>), dim3(nch), dim3(kBlock), 0, s, sg.chunks.p, w.p.p, w.sv.p, w.sv.p, w.t.p,
                       w.rh.p, w.sc.p, x, w.r.p, w.xr_part.p);
    hipLaunchKernelGGL(k_seg_stage, dim3(M), dim3(kStage), 0, s, (int)ST_RHO, sg.chunk_begin.p, sg.op_rng.p,
                       w.xr_part.p, (const double*)nullptr, w.sc.p, (const double*)nullptr,
                       w.hsc_dev + (size_t)M * BS_COUNT * slot, sizeof(T) == 4 ? kBicgRestartEps32 : 0.0);
    PB_HIP(hipGetLastError());
}

// launches per iteration: p (or the copy), three stages, s, x/r and two operator applications
inline int64_t iteration_launches(const Segments& sg, const ColorOrder& co) {
    int nc = 0;
    for (int c = 0; c < co.ncolor; ++c) nc += sg.sweep.nblk[c] > 0;
    return 6 + 2 * (2 * nc - 1);
}

// Batched BiCGStab on the split operator Ahat from x = 0: each member runs until its own
// ||r|| <= target[m] ||r0|| (target < 0: the member does not run).  iters[m] receives
// the member's iterations; at most maxit lock-step iterations.
// graphs: iterations 2, 3, ... run as replays of two captured graphs of two iterations each
// (the same launches; at most maxit + 1 lock-step iterations then).
template <typename T>
void batch_bicg_run(hipStream_t s, Segments& sg, const ColorOrder& co, DiluFactors<T>& d, BatchWork<T>& w, const T* b,
                    T* x, const std::vector<double>& target, int maxit, std::vector<int>& iters, RunCount& rc,
                    bool graphs, std::vector<int>& restarts) {
    const int M = sg.M;
    const int64_t n = sg.N;
    const unsigned nch = (unsigned)sg.nchunk;
    iters.assign(M, 0);
    restarts.assign(M, 0);
    if (maxit <= 0 || nch == 0) return;
    PB_HIP(hipMemcpyAsync(w.r.p, b, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    PB_HIP(hipMemcpyAsync(w.rh.p, b, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    PB_HIP(hipMemcpyAsync(w.tgt.p, target.data(), M * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_seg_dots2<T>), dim3(nch), dim3(kBlock), 0, s, sg.chunks.p, w.rh.p, w.r.p, w.xr_part.p);
    hipLaunchKernelGGL(k_seg_stage, dim3(M), dim3(kStage), 0, s, (int)ST_START, sg.chunk_begin.p, sg.op_rng.p,
                       w.xr_part.p, (const double*)nullptr, w.sc.p, w.tgt.p, w.hsc_dev, 0.0);
    PB_HIP(hipGetLastError());
    PB_HIP(hipStreamSynchronize(s));
    auto all_stopped = [&](const double* slot) {
        for (int m = 0; m < M; ++m)
            if (slot[(size_t)m * BS_COUNT + BS_STOP] == 0.0) return false;
        return true;
    };
    if (all_stopped(w.hsc.p)) return;
    const int64_t per_it = iteration_launches(sg, co);
    const size_t slot_len = (size_t)M * BS_COUNT;
    if (!graphs || maxit < 3) {
        auto iterate = [&](int it) {
            enqueue_iteration<T>(s, sg, co, d, w, x, it & 1, it == 1);
            PB_HIP(hipEventRecord(w.ev[it & 1], s));
            rc.launches += per_it;
            rc.iterations += 1;
        };
        iterate(1);
        for (int it = 1;; ++it) {
            if (it < maxit) iterate(it + 1);
            PB_HIP(hipEventSynchronize(w.ev[it & 1]));
            if (all_stopped(w.hsc.p + slot_len * (it & 1)) || it >= maxit) break;
        }
    } else {
        const void* key = static_cast<const void*>(x);
        for (int g = 0; g < 2; ++g)
            if (!w.pair[g] || w.pair_key[g] != key) {
                if (w.pair[g]) (void)hipGraphExecDestroy(w.pair[g]);
                w.pair[g] = nullptr;
                hipGraph_t graph = nullptr;
                PB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                enqueue_iteration<T>(s, sg, co, d, w, x, 2 + 2 * g, false);
                enqueue_iteration<T>(s, sg, co, d, w, x, 3 + 2 * g, false);
                PB_HIP(hipStreamEndCapture(s, &graph));
                const hipError_t ie = hipGraphInstantiate(&w.pair[g], graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                PB_HIP(ie);
                w.pair_key[g] = key;
            }
        enqueue_iteration<T>(s, sg, co, d, w, x, 1, true);  // iteration 1 (slot 1)
        PB_HIP(hipEventRecord(w.ev[1], s));
        rc.launches += per_it;
        rc.iterations += 1;
        int launched = 0;  // pair k covers iterations 2k + 2, 2k + 3 and records ev[k & 1]
        auto launch_pair = [&]() {
            PB_HIP(hipGraphLaunch(w.pair[launched & 1], s));
            PB_HIP(hipEventRecord(w.ev[launched & 1], s));
            rc.launches += 2 * per_it;
            rc.iterations += 2;
            ++launched;
        };
        launch_pair();
        PB_HIP(hipEventSynchronize(w.ev[1]));  // iteration 1 (recorded before pair 0 re-records ev[0])
        if (!all_stopped(w.hsc.p + slot_len * 1))
            for (int k = 0;; ++k) {
                const int last = 2 * k + 3;  // the last iteration of pair k
                if (last < maxit) launch_pair();
                PB_HIP(hipEventSynchronize(w.ev[k & 1]));
                if (all_stopped(w.hsc.p + slot_len * (3 + 2 * (k & 1))) || last >= maxit) break;
            }
    }
    PB_HIP(hipStreamSynchronize(s));
    std::vector<double> h((size_t)M * BS_COUNT);
    PB_HIP(hipMemcpyAsync(h.data(), w.sc.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    PB_HIP(hipStreamSynchronize(s));
    for (int m = 0; m < M; ++m) {
        iters[m] = (int)h[(size_t)m * BS_COUNT + BS_ITERS];
        restarts[m] = (int)h[(size_t)m * BS_COUNT + BS_NRESTART];
    }
}

struct MemberSolve {
    int iters = 0, refinements = 0, restarts = 0;
    double resid = 0, resid0 = 1;
    bool converged = false, cold_fallback = false;
};

// Per-member FP64 refinement around the batched inner runs (bicgstab_split of krylov.hip
// member by member): r = b - A x in FP64 (A: the operator held by d64 in position order),
// each active member's r / ||r|| is transformed with its split factors, solved to its
// inner target, mapped back and added to x.  present[m] = 0: the member is not solved (x
// stays as given).  inner[m]: inner target (0: 0.5 tol ||b|| / ||r|| alone); inner_cold[m]
// replaces it when a warm start is dropped.
template <typename T>
void batch_split_solve(hipStream_t s, Segments& sg, const ColorOrder& co, DiluFactors<double>& d64,
                       DiluFactors<T>& dT, BatchWork<T>& w, const double* b, double* x, double* r, T* w1, T* w2,
                       T* w3, DevBuf<double>& spmv_part, const std::vector<int>& present, bool warm, int maxit,
                       double tol, std::vector<double> inner, const std::vector<double>& inner_cold,
                       std::vector<MemberSolve>& out, RunCount& rc, bool graphs) {
    const int M = sg.M;
    const int64_t n = sg.N;
    const unsigned nch = (unsigned)sg.nchunk;
    out.assign(M, MemberSolve());
    spmv_part.alloc(std::max<int64_t>(1, pos_spmv_blocks(sg.F, sg.L)));
    const std::vector<double> nb = seg_reduce<false>(s, sg, b);
    auto residual = [&]() {
        pos_spmv<double>(s, co, sg.F, sg.L, d64, x, b, r, spmv_part.p);
        return seg_reduce<false>(s, sg, r);
    };
    std::vector<double> beta = residual();
    if (warm) {  // a start worse than x0 = 0 is dropped, member by member
        std::vector<double> keep(M, 1.0);
        bool any = false;
        for (int m = 0; m < M; ++m)
            if (present[m] && nb[m] > 0 && beta[m] > nb[m]) {
                keep[m] = 0.0;
                out[m].cold_fallback = true;
                inner[m] = inner_cold[m];
                any = true;
            }
        if (any) {
            upload_factors(s, w.fac, keep);
            hipLaunchKernelGGL((k_seg_scale<double>), dim3(nch), dim3(kBlock), 0, s, sg.chunks.p, x, w.fac.p, x);
            beta = residual();
        }
    }
    for (int m = 0; m < M; ++m) out[m].resid0 = nb[m] > 0 ? beta[m] / nb[m] : 0.0;
    std::vector<double> f(M), target(M), a(M);
    std::vector<int> it_run, rs_run;
    while (true) {
        bool any = false;
        for (int m = 0; m < M; ++m) {
            MemberSolve& o = out[m];
            o.resid = nb[m] > 0 ? beta[m] / nb[m] : 0.0;
            o.converged = !present[m] || nb[m] == 0.0 || o.resid <= tol;
            const bool active = !o.converged && o.iters < maxit;
            f[m] = active ? 1.0 / beta[m] : 0.0;
            a[m] = active ? beta[m] : 0.0;
            double stop = active ? 0.5 * tol * nb[m] / beta[m] : -1.0;
            if (active && stop < inner[m]) stop = inner[m];
            target[m] = stop;
            any |= active;
        }
        if (!any) break;
        int left = maxit;
        for (int m = 0; m < M; ++m)
            if (target[m] >= 0) left = std::min(left, maxit - out[m].iters);
        upload_factors(s, w.fac, f);
        hipLaunchKernelGGL((k_seg_scale<T>), dim3(nch), dim3(kBlock), 0, s, sg.chunks.p, r, w.fac.p, w1);
        split_op<T>(1, s, co, sg.F, sg.L, dT, w1, w2, 0);  // left transform of r / ||r||
        PB_HIP(hipMemsetAsync(w3, 0, n * sizeof(T), s));
        batch_bicg_run<T>(s, sg, co, dT, w, w2, w3, target, left, it_run, rc, graphs, rs_run);
        split_op<T>(2, s, co, sg.F, sg.L, dT, w3, w1, 0);  // correction back to the unknowns
        upload_factors(s, w.fac, a);
        hipLaunchKernelGGL((k_seg_axpy<T>), dim3(nch), dim3(kBlock), 0, s, sg.chunks.p, w.fac.p, w1, x);
        PB_HIP(hipGetLastError());
        for (int m = 0; m < M; ++m)
            if (target[m] >= 0) {
                out[m].iters += it_run[m];
                out[m].refinements += 1;
                out[m].restarts += rs_run[m];
            }
        beta = residual();
    }
}

// zero x on the members whose flag is 0 (x in position order)
void zero_members(hipStream_t s, Segments& sg, DevBuf<double>& fac, const std::vector<int>& keep, double* x) {
    std::vector<double> f(sg.M);
    bool any = false;
    for (int m = 0; m < sg.M; ++m) {
        f[m] = keep[m] ? 1.0 : 0.0;
        any |= !keep[m];
    }
    if (!any || sg.nchunk == 0) return;
    upload_factors(s, fac, f);
    hipLaunchKernelGGL((k_seg_scale<double>), dim3((unsigned)sg.nchunk), dim3(kBlock), 0, s, sg.chunks.p, x, fac.p, x);
    PB_HIP(hipGetLastError());
}

}  // namespace
}  // namespace pb

using namespace pb;

// One union: the members of one nLayer.
struct BatchGroup {
    int L = 0;
    std::vector<int> members;        // batch member indices
    std::vector<int64_t> face_off;   // [M + 1]
    pbsm3d_ctx* ctx = nullptr;
    Segments susp, dep;
    BatchWork<float> w32;
    BatchWork<double> w64, wd;
    DevBuf<double> spmv_part_s, spmv_part_d, fac;
    DevBuf<int32_t> face_member, member_on;
    bool stepped = false;  // a step has run: the staged forcing and solutions are this step's
    bool pairs_ready = false;  // c->feat / c->pred hold this step's training pairs of the union
    ~BatchGroup() {
        if (ctx) pbsm3d_destroy(ctx);
    }
};

struct pbsm3d_batch {
    int device = 0;
    int M = 0;
    pbsm3d_config cfg{};  // the shared config (member 0's)
    std::vector<std::unique_ptr<BatchGroup>> groups;
    std::vector<int> group_of, index_in_group;
    std::vector<int64_t> faces;
    pbsm3d_mlp* warm = nullptr;
    bool graphs = true;  // iterations 2, 3, ... replayed from captured hipGraphs
    RunCount last;
    std::string err;
};

namespace {
std::string g_batch_err;

template <typename Fn>
int batch_guard(pbsm3d_batch* b, Fn&& fn) {
    try {
        fn();
        return PBSM3D_OK;
    } catch (const Error& e) {
        if (b) b->err = e.what();
        g_batch_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        if (b) b->err = e.what();
        g_batch_err = e.what();
        return PBSM3D_EINVAL;
    }
}

// the config fields every member must share (all but nLayer, z0_snow, settling_scale)
bool same_shared_config(const pbsm3d_config& a, const pbsm3d_config& b) {
    pbsm3d_config x = a, y = b;
    x.nLayer = y.nLayer = 0;
    x.z0_snow = y.z0_snow = 0;
    x.settling_scale = y.settling_scale = 0;
    return std::memcmp(&x, &y, sizeof(x)) == 0;
}

// concatenated host arrays of a group's member meshes (optional arrays: zeros / defaults
// where a member lacks them, which gives that member the reference's no-vegetation step)
struct UnionMesh {
    std::vector<int32_t> neigh;
    std::vector<double> nx, ny, el, area, cx, cy, cz, canopy, lai, sn, sd, z0, settle;
    std::vector<uint8_t> water;
    bool any_veg = false, any_stalk = false;
};

void build_union(UnionMesh& u, const std::vector<int>& members, const pbsm3d_mesh* meshes, const pbsm3d_config* cfgs,
                 std::vector<int64_t>& face_off) {
    face_off.assign(1, 0);
    for (int mi : members) {
        if (meshes[mi].n_faces <= 0) throw Error(PBSM3D_EINVAL, "member " + std::to_string(mi) + " has no faces");
        face_off.push_back(face_off.back() + meshes[mi].n_faces);
        u.any_veg |= meshes[mi].canopy_height != nullptr;
        u.any_stalk |= meshes[mi].stalk_number != nullptr && meshes[mi].stalk_diameter != nullptr;
    }
    const int64_t F = face_off.back();
    u.neigh.resize(3 * F);
    for (auto* v : {&u.nx, &u.ny, &u.el}) v->resize(3 * F);
    for (auto* v : {&u.area, &u.cx, &u.cy, &u.cz, &u.z0, &u.settle}) v->resize(F);
    u.water.assign(F, 0);
    if (u.any_veg) {
        u.canopy.assign(F, 0.0);
        u.lai.assign(F, 0.0);
        u.sn.assign(F, 0.0);
        u.sd.assign(F, 0.0);
    }
    for (size_t k = 0; k < members.size(); ++k) {
        const pbsm3d_mesh& m = meshes[members[k]];
        const pbsm3d_config& c = cfgs[members[k]];
        const int64_t o = face_off[k], n = m.n_faces;
        if (!m.neigh || !m.normal_x || !m.normal_y || !m.edge_length || !m.area || !m.center_x || !m.center_y ||
            !m.center_z)
            throw Error(PBSM3D_EINVAL, "member " + std::to_string(members[k]) + ": mesh geometry pointer is null");
        for (int64_t j = 0; j < 3 * n; ++j) {
            const int32_t nb = m.neigh[j];
            if (nb < -1 || nb >= n)
                throw Error(PBSM3D_EINVAL, "member " + std::to_string(members[k]) + ": neighbour index out of range");
            u.neigh[3 * o + j] = nb < 0 ? -1 : (int32_t)(nb + o);
        }
        std::copy(m.normal_x, m.normal_x + 3 * n, u.nx.begin() + 3 * o);
        std::copy(m.normal_y, m.normal_y + 3 * n, u.ny.begin() + 3 * o);
        std::copy(m.edge_length, m.edge_length + 3 * n, u.el.begin() + 3 * o);
        std::copy(m.area, m.area + n, u.area.begin() + o);
        std::copy(m.center_x, m.center_x + n, u.cx.begin() + o);
        std::copy(m.center_y, m.center_y + n, u.cy.begin() + o);
        std::copy(m.center_z, m.center_z + n, u.cz.begin() + o);
        if (m.is_water) std::copy(m.is_water, m.is_water + n, u.water.begin() + o);
        std::fill(u.z0.begin() + o, u.z0.begin() + o + n, c.z0_snow);
        std::fill(u.settle.begin() + o, u.settle.begin() + o + n, c.settling_scale);
        if (u.any_veg && m.canopy_height) {
            std::copy(m.canopy_height, m.canopy_height + n, u.canopy.begin() + o);
            if (m.lai) std::copy(m.lai, m.lai + n, u.lai.begin() + o);
            const bool both = m.stalk_number && m.stalk_diameter;  // else N = 1, dv = 0.8 (PBSM3D.cpp:302-309)
            for (int64_t i = 0; i < n; ++i) {
                u.sn[o + i] = both ? m.stalk_number[i] : 1.0;
                u.sd[o + i] = both ? m.stalk_diameter[i] : 0.8;
            }
        }
    }
}

void upload_vec(DevBuf<double>& d, const std::vector<double>& h, hipStream_t s) {
    d.alloc(h.size());
    PB_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
    PB_HIP(hipStreamSynchronize(s));
}

// Forcing of the group's members into the union staging buffer (field k at k * F).
void stage_group_forcing(BatchGroup& g, const pbsm3d_forcing* fo, const double* fdev[9]) {
    pbsm3d_ctx* c = g.ctx;
    hipStream_t s = c->stream;
    const int64_t F = c->F;
    for (int k = 0; k < 9; ++k) {
        bool any = false, all = true;
        for (int mi : g.members) {
            const double* fin[9] = {fo[mi].U_R, fo[mi].U_2m_above_srf, fo[mi].vw_dir, fo[mi].swe, fo[mi].t,
                                    fo[mi].rh, fo[mi].snowdepthavg, fo[mi].fetch, fo[mi].p_snow_hours};
            any |= fin[k] != nullptr;
            all &= fin[k] != nullptr;
        }
        if (!any) {
            fdev[k] = nullptr;
            continue;
        }
        if (!all) throw Error(PBSM3D_EINVAL, "a forcing field is given for some members only");
        double* dst = c->forcing_stage.p + (size_t)k * F;
        for (size_t j = 0; j < g.members.size(); ++j) {
            const pbsm3d_forcing& f = fo[g.members[j]];
            const double* fin[9] = {f.U_R, f.U_2m_above_srf, f.vw_dir, f.swe, f.t, f.rh, f.snowdepthavg, f.fetch,
                                    f.p_snow_hours};
            const int64_t o = g.face_off[j], n = g.face_off[j + 1] - o;
            PB_HIP(hipMemcpyAsync(dst + o, fin[k], n * sizeof(double),
                                  f.on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
        }
        fdev[k] = dst;
    }
}

template <typename T>
void group_suspension_solve(pbsm3d_batch* B, BatchGroup& g, BatchWork<T>& w, DiluFactors<T>& dT, T* const* wk,
                            const std::vector<int>& present, bool warm, std::vector<MemberSolve>& res, RunCount& rc) {
    pbsm3d_ctx* c = g.ctx;
    const pbsm3d_config& cfg = c->cfg;
    const int M = (int)g.members.size();
    w.ensure(g.susp);
    constexpr double kCold = 2e-3, kLearned = 2e-4;  // the single solve's defaults (api.hip)
    std::vector<double> inner(M), cold(M);
    for (int m = 0; m < M; ++m) {
        if (cfg.inner_rtol > 0) {
            inner[m] = cold[m] = cfg.inner_rtol;
        } else if (cfg.precision == PBSM3D_MIXED) {
            inner[m] = warm ? kLearned : kCold;
            cold[m] = kCold;
        } else {
            inner[m] = cold[m] = 0.0;
        }
    }
    batch_split_solve<T>(c->stream, g.susp, c->co, c->dl_s, dT, w, c->bp.p, c->xp.p, c->r_s.p, wk[0], wk[1], wk[2],
                         g.spmv_part_s, present, warm, cfg.max_iters, cfg.tol, inner, cold, res, rc, B->graphs);
}

void group_step(pbsm3d_batch* B, BatchGroup& g, const pbsm3d_forcing* fo, int32_t dt_s, pbsm3d_stats* st,
                RunCount& rc) {
    pbsm3d_ctx* c = g.ctx;
    hipStream_t s = c->stream;
    const pbsm3d_config& cfg = c->cfg;
    const int64_t F = c->F;
    const int L = c->L;
    const int64_t N = F * L;
    const int M = (int)g.members.size();
    g.stepped = g.pairs_ready = false;  // until this step completes
    PB_HIP(hipSetDevice(c->device));
    PB_HIP(hipEventRecord(c->ev[0], s));
    const double* fdev[9];
    stage_group_forcing(g, fo, fdev);
    if ((cfg.use_exp_fetch || cfg.use_tanh_fetch) && !fdev[7])
        throw Error(PBSM3D_EINVAL, "fetch forcing required by the fetch mode");
    if (cfg.use_PomLi_probability && !fdev[8])
        throw Error(PBSM3D_EINVAL, "p_snow_hours forcing required by use_PomLi_probability");
    for (int k = 0; k < 7; ++k)
        if (!fdev[k]) throw Error(PBSM3D_EINVAL, "forcing pointer is null");
    // ---- coefficients + suspension assembly of every member
    AssembleArgs a = assemble_args(c, fdev, dt_s);
    hipLaunchKernelGGL(k_assemble, dim3(grid_for(F)), dim3(kBlock), 0, s, a);
    PB_HIP(hipGetLastError());
    dilu_factor<double>(s, c->co, F, L, c->susp.rp.p, c->susp.val.p, c->slots.p, c->dl_s);
    if (cfg.precision == PBSM3D_MIXED) dilu_narrow(s, N, c->dl_s, c->dl32);
    to_pos<double, double>(s, c->co, F, L, c->susp.rhs.p, c->bp.p);
    PB_HIP(hipEventRecord(c->ev[1], s));
    // ---- per-member presence (getRhsMax, PBSM3D.cpp:1424-1427) and the suspension solves
    const std::vector<double> smax = seg_reduce<true>(s, g.susp, c->bp.p);
    std::vector<int> present(M);
    bool any_present = false;
    for (int m = 0; m < M; ++m) {
        present[m] = smax[m] > kPresent;
        any_present |= present[m] != 0;
    }
    const bool warm = cfg.warm_start == PBSM3D_WS_LEARNED && c->warm != nullptr;
    if (warm && any_present) {
        c->feat.alloc(F * MlpDims::IN);
        c->pred.alloc(F * MlpDims::OUT);
        warm_features(c, fdev[2], fdev[4]);
        c->warm->predict(c->feat.p, c->pred.p, F, s);
        warm_apply(c, 0);
    }
    to_pos<double, double>(s, c->co, F, L, c->susp.x.p, c->xp.p);
    zero_members(s, g.susp, g.fac, present, c->xp.p);  // no suspension: x stays 0 (zeroSystem)
    std::vector<MemberSolve> sres(M), dres(M);
    if (any_present) {
        if (cfg.precision == PBSM3D_MIXED) {
            float* wk[3] = {c->split32[0].p, c->split32[1].p, c->split32[2].p};
            group_suspension_solve<float>(B, g, g.w32, c->dl32, wk, present, warm, sres, rc);
        } else {
            double* wk[3] = {c->split64[0].p, c->split64[1].p, c->split64[2].p};
            group_suspension_solve<double>(B, g, g.w64, c->dl_s, wk, present, warm, sres, rc);
        }
    }
    for (int m = 0; m < M; ++m)
        if (present[m] && !sres[m].converged)
            throw Error(PBSM3D_ENOCONV, "batch member " + std::to_string(g.members[m]) +
                                            ": suspension solver failed to converge");
    from_pos<double, double>(s, c->co, F, L, c->xp.p, c->susp.x.p);
    PB_HIP(hipEventRecord(c->ev[2], s));
    // ---- Qsusp / Qsubl and deposition assembly of every member
    PostArgs p{F, L, c->dz, (double)dt_s, c->susp.x.p, c->u_z.p, c->csubl.p,
               c->Qsusp.p, c->Qsubl.p, c->Qsubl_mass.p, c->sum_subl.p};
    DepositionArgs d{F, fdev[2], c->nx.p, c->ny.p, c->elen.p, c->area.p, c->dep_w.p, c->Qsusp.p, c->Qsalt.p,
                     c->neigh.p, c->dep.rp.p, c->slots.p, c->dep.val.p, c->dep.rhs.p, c->dep.x.p, c->invdiag_d.p};
    hipLaunchKernelGGL(k_post, dim3(grid_for(F)), dim3(kBlock), 0, s, p);
    hipLaunchKernelGGL(k_deposition, dim3(grid_for(F)), dim3(kBlock), 0, s, d);
    PB_HIP(hipGetLastError());
    dilu_factor<double>(s, c->co, F, 1, c->dep.rp.p, c->dep.val.p, c->slots.p, c->dl_d);
    to_pos<double, double>(s, c->co, F, 1, c->dep.rhs.p, c->dsplit[0].p);
    const std::vector<double> dmax = seg_reduce<true>(s, g.dep, c->dsplit[0].p);
    std::vector<int> dep_present(M);
    bool any_dep = false;
    for (int m = 0; m < M; ++m) {
        dep_present[m] = present[m] && dmax[m] > kPresent;  // PBSM3D.cpp:1661-1664
        any_dep |= dep_present[m] != 0;
    }
    if (warm && any_dep) warm_apply(c, 1);
    to_pos<double, double>(s, c->co, F, 1, c->dep.x.p, c->dsplit[1].p);
    zero_members(s, g.dep, g.fac, dep_present, c->dsplit[1].p);
    PB_HIP(hipEventRecord(c->ev[3], s));
    if (any_dep) {
        g.wd.ensure(g.dep);
        std::vector<double> zero(M, 0.0);
        batch_split_solve<double>(s, g.dep, c->co, c->dl_d, c->dl_d, g.wd, c->dsplit[0].p, c->dsplit[1].p, c->r_d.p,
                                  c->dsplit[2].p, c->dsplit[3].p, c->dsplit[4].p, g.spmv_part_d, dep_present, warm,
                                  cfg.max_iters, cfg.tol, zero, zero, dres, rc, B->graphs);
        for (int m = 0; m < M; ++m)
            if (dep_present[m] && !dres[m].converged)
                throw Error(PBSM3D_ENOCONV, "batch member " + std::to_string(g.members[m]) +
                                                ": deposition solver failed to converge");
    }
    from_pos<double, double>(s, c->co, F, 1, c->dsplit[1].p, c->dep.x.p);
    // ---- drift of the members whose deposition was present (PBSM3D.cpp:1708-1740)
    if (any_dep) {
        PB_HIP(hipMemcpyAsync(g.member_on.p, dep_present.data(), M * sizeof(int32_t), hipMemcpyHostToDevice, s));
        DriftArgs dr{F, (double)dt_s, c->dep.x.p, fdev[3], c->salt_flag.p, c->drift_mass.p, c->sum_drift.p, c->more.p,
                     g.face_member.p, g.member_on.p};
        hipLaunchKernelGGL(k_drift, dim3(grid_for(F)), dim3(kBlock), 0, s, dr);
        PB_HIP(hipGetLastError());
    }
    PB_HIP(hipEventRecord(c->ev[4], s));
    PB_HIP(hipStreamSynchronize(s));
    g.stepped = true;
    g.pairs_ready = false;
    for (int m = 0; m < M; ++m) {
        pbsm3d_stats& o = st[g.members[m]];
        o = pbsm3d_stats{};
        o.suspension_present = present[m];
        o.deposition_present = dep_present[m];
        o.susp_iters = sres[m].iters;
        o.dep_iters = dres[m].iters;
        o.susp_resid = sres[m].resid;
        o.dep_resid = dres[m].resid;
        o.susp_rhs_max = smax[m];
        o.dep_rhs_max = dmax[m];
        o.susp_refinements = sres[m].refinements;
        o.susp_restarts = sres[m].restarts;
        o.susp_resid0 = sres[m].resid0;
        o.susp_cold_fallback = sres[m].cold_fallback ? 1 : 0;
    }
}

void group_outputs(BatchGroup& g, pbsm3d_out* outs) {
    pbsm3d_ctx* c = g.ctx;
    hipStream_t s = c->stream;
    const int64_t F = c->F;
    for (size_t j = 0; j < g.members.size(); ++j) {
        pbsm3d_out* o = &outs[g.members[j]];
        const int64_t off = g.face_off[j], n = g.face_off[j + 1] - off;
        const hipMemcpyKind kind = o->on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
        auto cp = [&](double* dst, const DevBuf<double>& src) {
            if (dst) PB_HIP(hipMemcpyAsync(dst, src.p + off, n * sizeof(double), kind, s));
        };
        cp(o->Qsalt, c->Qsalt);
        cp(o->Qsusp, c->Qsusp);
        cp(o->Qsubl, c->Qsubl);
        cp(o->Qsubl_mass, c->Qsubl_mass);
        cp(o->sum_subl, c->sum_subl);
        cp(o->drift_mass, c->drift_mass);
        cp(o->sum_drift, c->sum_drift);
        cp(o->pbsm_more_than_avail, c->more);
        cp(o->blowingsnow_probability, c->blowprob);
        cp(o->qdep, c->dep.x);
        if (o->conc)
            for (int z = 0; z < c->L; ++z)
                PB_HIP(hipMemcpyAsync(o->conc + (size_t)z * n, c->susp.x.p + (size_t)z * F + off, n * sizeof(double),
                                      kind, s));
    }
    PB_HIP(hipStreamSynchronize(s));
}
}  // namespace

extern "C" {

int pbsm3d_batch_create(pbsm3d_batch** out, int32_t n_members, const pbsm3d_mesh* meshes, const pbsm3d_config* cfgs,
                        int device) {
    if (!out || n_members <= 0 || !meshes || !cfgs) {
        g_batch_err = "invalid argument";
        return PBSM3D_EINVAL;
    }
    *out = nullptr;
    auto B = std::make_unique<pbsm3d_batch>();
    const int rc = batch_guard(B.get(), [&] {
        B->device = device;
        B->M = n_members;
        B->cfg = cfgs[0];
        for (int m = 0; m < n_members; ++m) {
            const pbsm3d_config& c = cfgs[m];
            if (!same_shared_config(c, cfgs[0]))
                throw Error(PBSM3D_EINVAL, "member " + std::to_string(m) +
                                               ": only nLayer, z0_snow and settling_scale may differ between members");
            if (c.krylov != PBSM3D_KRYLOV_BICGSTAB_SPLIT || c.precond != PBSM3D_PC_MCILU0 ||
                c.dep_precond != PBSM3D_PC_MCILU0)
                throw Error(PBSM3D_EINVAL, "the batch solves with the split D-ILU BiCGStab (krylov "
                                           "PBSM3D_KRYLOV_BICGSTAB_SPLIT, precond / dep_precond PBSM3D_PC_MCILU0)");
            if (c.nLayer < 2 || c.nLayer > 16) throw Error(PBSM3D_EINVAL, "batch members need nLayer in [2, 16]");
            if (!(c.z0_snow > 0) || !(c.settling_scale >= 0)) throw Error(PBSM3D_EINVAL, "bad z0_snow / settling_scale");
            if (c.coarse_tiles != 0) throw Error(PBSM3D_EINVAL, "the batch has no learned coarse correction");
        }
        std::vector<int> Ls;
        for (int m = 0; m < n_members; ++m)
            if (std::find(Ls.begin(), Ls.end(), cfgs[m].nLayer) == Ls.end()) Ls.push_back(cfgs[m].nLayer);
        std::sort(Ls.begin(), Ls.end());
        B->group_of.assign(n_members, -1);
        B->index_in_group.assign(n_members, -1);
        B->faces.assign(n_members, 0);
        for (int L : Ls) {
            auto g = std::make_unique<BatchGroup>();
            g->L = L;
            for (int m = 0; m < n_members; ++m)
                if (cfgs[m].nLayer == L) {
                    B->group_of[m] = (int)B->groups.size();
                    B->index_in_group[m] = (int)g->members.size();
                    B->faces[m] = meshes[m].n_faces;
                    g->members.push_back(m);
                }
            UnionMesh u;
            build_union(u, g->members, meshes, cfgs, g->face_off);
            const int64_t F = g->face_off.back();
            pbsm3d_mesh um{};
            um.n_faces = F;
            um.neigh = u.neigh.data();
            um.normal_x = u.nx.data();
            um.normal_y = u.ny.data();
            um.edge_length = u.el.data();
            um.area = u.area.data();
            um.center_x = u.cx.data();
            um.center_y = u.cy.data();
            um.center_z = u.cz.data();
            um.canopy_height = u.any_veg ? u.canopy.data() : nullptr;
            um.lai = u.any_veg ? u.lai.data() : nullptr;
            um.stalk_number = u.any_veg ? u.sn.data() : nullptr;
            um.stalk_diameter = u.any_veg ? u.sd.data() : nullptr;
            um.is_water = u.water.data();
            pbsm3d_config uc = cfgs[g->members[0]];
            const int crc = pbsm3d_create(&g->ctx, &um, &uc, device);
            if (crc != PBSM3D_OK) throw Error(crc, std::string("union of nLayer ") + std::to_string(L) + ": " +
                                                       pbsm3d_last_error(nullptr));
            pbsm3d_ctx* c = g->ctx;
            hipStream_t s = c->stream;
            upload_vec(c->z0_face, u.z0, s);
            upload_vec(c->settle_face, u.settle, s);
            build_segments(g->susp, c->co, c->co.h_color, g->face_off, L, s);
            build_segments(g->dep, c->co, c->co.h_color, g->face_off, 1, s);
            std::vector<int32_t> fm(F);
            for (size_t j = 0; j + 1 < g->face_off.size(); ++j)
                std::fill(fm.begin() + g->face_off[j], fm.begin() + g->face_off[j + 1], (int32_t)j);
            g->face_member.alloc(F);
            g->member_on.alloc(g->members.size());
            PB_HIP(hipMemcpyAsync(g->face_member.p, fm.data(), F * sizeof(int32_t), hipMemcpyHostToDevice, s));
            PB_HIP(hipStreamSynchronize(s));
            B->groups.push_back(std::move(g));
        }
    });
    if (rc != PBSM3D_OK) return rc;
    *out = B.release();
    return PBSM3D_OK;
}

int pbsm3d_batch_step(pbsm3d_batch* B, const pbsm3d_forcing* forcings, int32_t dt_s, pbsm3d_out* outs,
                      pbsm3d_stats* stats) {
    if (!B || !forcings) return PBSM3D_EINVAL;
    return batch_guard(B, [&] {
        PB_HIP(hipSetDevice(B->device));
        std::vector<pbsm3d_stats> st(B->M);
        const size_t G = B->groups.size();
        std::vector<RunCount> rc(G);
        std::vector<std::exception_ptr> err(G);
        const auto t0 = std::chrono::steady_clock::now();
        // one host thread per union: each drives its own context and stream, so the unions'
        // kernels overlap on the GPU and one union's host waits do not stall the others
        auto work = [&](size_t k) {
            try {
                PB_HIP(hipSetDevice(B->device));
                group_step(B, *B->groups[k], forcings, dt_s, st.data(), rc[k]);
                if (outs) group_outputs(*B->groups[k], outs);
            } catch (...) {
                err[k] = std::current_exception();
            }
        };
        if (G == 1) {
            work(0);
        } else {
            std::vector<std::thread> th;
            for (size_t k = 0; k < G; ++k) th.emplace_back(work, k);
            for (auto& t : th) t.join();
        }
        for (auto& e : err)
            if (e) std::rethrow_exception(e);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        B->last = RunCount();
        for (const RunCount& r : rc) {
            B->last.launches += r.launches;
            B->last.iterations += r.iterations;
        }
        if (stats)
            for (int m = 0; m < B->M; ++m) {
                stats[m] = st[m];
                stats[m].ms_total = ms;
            }
    });
}

int pbsm3d_batch_set_warm_start(pbsm3d_batch* B, pbsm3d_mlp* net) {
    if (!B) return PBSM3D_EINVAL;
    B->warm = net;
    for (auto& g : B->groups)
        if (g->L == 10) {
            const int rc = pbsm3d_set_warm_start(g->ctx, net);
            if (rc != PBSM3D_OK) return rc;
        }
    return PBSM3D_OK;
}

int pbsm3d_batch_use_graphs(pbsm3d_batch* B, int enable) {
    if (!B) return PBSM3D_EINVAL;
    B->graphs = enable != 0;
    return PBSM3D_OK;
}

int pbsm3d_batch_info(const pbsm3d_batch* B, int32_t* unions, int64_t* launches, int64_t* iterations) {
    if (!B) return PBSM3D_EINVAL;
    if (unions) *unions = (int32_t)B->groups.size();
    if (launches) *launches = B->last.launches;
    if (iterations) *iterations = B->last.iterations;
    return PBSM3D_OK;
}

int pbsm3d_batch_get_state(pbsm3d_batch* B, int32_t m, double* sum_drift, double* sum_subl) {
    if (!B || m < 0 || m >= B->M) return PBSM3D_EINVAL;
    return batch_guard(B, [&] {
        BatchGroup& g = *B->groups[B->group_of[m]];
        const int j = B->index_in_group[m];
        const int64_t off = g.face_off[j], n = g.face_off[j + 1] - off;
        pbsm3d_ctx* c = g.ctx;
        PB_HIP(hipSetDevice(c->device));
        if (sum_drift)
            PB_HIP(hipMemcpyAsync(sum_drift, c->sum_drift.p + off, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (sum_subl)
            PB_HIP(hipMemcpyAsync(sum_subl, c->sum_subl.p + off, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        PB_HIP(hipStreamSynchronize(c->stream));
    });
}

int pbsm3d_batch_training_pairs(pbsm3d_batch* B, int32_t m, float* X, float* Y) {
    if (!B || m < 0 || m >= B->M || !X || !Y) return PBSM3D_EINVAL;
    return batch_guard(B, [&] {
        BatchGroup& g = *B->groups[B->group_of[m]];
        const int j = B->index_in_group[m];
        const int64_t off = g.face_off[j], n = g.face_off[j + 1] - off;
        pbsm3d_ctx* c = g.ctx;
        if (c->L != 10) throw Error(PBSM3D_EINVAL, "training pairs need nLayer 10");
        if (!g.stepped) throw Error(PBSM3D_EINVAL, "no batch step has run yet");
        PB_HIP(hipSetDevice(c->device));
        const int64_t F = c->F;
        if (!g.pairs_ready) {
            // the union's features (neighbour indices stay inside each member) and targets,
            // once per step; vw_dir and t are the step's staged forcing
            c->feat.alloc(F * MlpDims::IN);
            c->pred.alloc(F * MlpDims::OUT);
            training_pairs_into(c, c->forcing_stage.p + 2 * F, c->forcing_stage.p + 4 * F, c->feat.p, c->pred.p);
            g.pairs_ready = true;
        }
        PB_HIP(hipMemcpyAsync(X, c->feat.p + off * MlpDims::IN, n * MlpDims::IN * sizeof(float),
                              hipMemcpyDeviceToDevice, c->stream));
        PB_HIP(hipMemcpyAsync(Y, c->pred.p + off * MlpDims::OUT, n * MlpDims::OUT * sizeof(float),
                              hipMemcpyDeviceToDevice, c->stream));
        PB_HIP(hipStreamSynchronize(c->stream));
    });
}

int pbsm3d_batch_set_state(pbsm3d_batch* B, int32_t m, const double* sum_drift, const double* sum_subl) {
    if (!B || m < 0 || m >= B->M) return PBSM3D_EINVAL;
    return batch_guard(B, [&] {
        BatchGroup& g = *B->groups[B->group_of[m]];
        const int j = B->index_in_group[m];
        const int64_t off = g.face_off[j], n = g.face_off[j + 1] - off;
        pbsm3d_ctx* c = g.ctx;
        PB_HIP(hipSetDevice(c->device));
        if (sum_drift)
            PB_HIP(hipMemcpyAsync(c->sum_drift.p + off, sum_drift, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        if (sum_subl)
            PB_HIP(hipMemcpyAsync(c->sum_subl.p + off, sum_subl, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        PB_HIP(hipStreamSynchronize(c->stream));
    });
}

void pbsm3d_batch_destroy(pbsm3d_batch* B) { delete B; }

const char* pbsm3d_batch_last_error(const pbsm3d_batch* B) { return B ? B->err.c_str() : g_batch_err.c_str(); }

}  // extern "C"
