// Sparse and dense vector kernels of the Krylov solves: CSR SpMV with the
// row block's values and column indices staged through LDS, fused residual,
// batched Gram-Schmidt dot products / updates with wavefront-shuffle
// reductions (FP64 accumulation for both FP64 and FP32 storage), and the
// Jacobi and per-column tridiagonal preconditioners.
//
// Reductions are two-stage and deterministic: every block writes its partial
// sums, then reduce_cols adds the partials in a fixed order.
#include <initializer_list>

#include "kernels.hpp"

namespace pb {

// zero operand of the CSR column kernel's gathers for absent neighbours
template <typename T>
__device__ const T kZeroVal = T(0);
namespace {

constexpr int kRowsPerBlock = kBlock;  // one row per thread
constexpr int kMaxRowLen = 6;          // self + 3 lateral + below + above

// ---------------------------------------------------------------------------
// SpMV: a block of 256 rows copies its contiguous slice of val/col into LDS
// with coalesced loads, then each thread forms its row's dot product.
// Requires every row to have at most kMaxRowLen entries (checked on the host).
// ---------------------------------------------------------------------------
template <typename T, bool RESID>
__global__ void __launch_bounds__(kBlock) k_spmv_lds(int64_t n, const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ col, const T* __restrict__ val,
                                                     const T* __restrict__ x, const T* __restrict__ b, T* __restrict__ y,
                                                     double* __restrict__ partial) {
    __shared__ T s_val[kRowsPerBlock * kMaxRowLen];
    __shared__ int32_t s_col[kRowsPerBlock * kMaxRowLen];
    __shared__ int32_t s_rp[kRowsPerBlock + 1];
    __shared__ double s_red[kBlock / 64];
    const int64_t row0 = (int64_t)blockIdx.x * kRowsPerBlock;
    const int64_t r = row0 + threadIdx.x;
    const int64_t last = min(row0 + kRowsPerBlock, n);
    if (r <= last) s_rp[threadIdx.x] = rp[r];
    if (threadIdx.x == 0 && row0 + kRowsPerBlock <= n) s_rp[kRowsPerBlock] = rp[row0 + kRowsPerBlock];
    __syncthreads();
    const int32_t k0 = s_rp[0];
    const int32_t cnt = s_rp[last - row0] - k0;
    for (int k = threadIdx.x; k < cnt; k += kBlock) {
        s_val[k] = val[k0 + k];
        s_col[k] = col[k0 + k];
    }
    __syncthreads();
    double rr = 0.0;
    if (r < n) {
        const int b0 = s_rp[threadIdx.x] - k0, b1 = s_rp[threadIdx.x + 1] - k0;
        T acc = 0;
        for (int k = b0; k < b1; ++k) acc = fma(s_val[k], x[s_col[k]], acc);
        if (RESID) {
            const T res = b[r] - acc;
            y[r] = res;
            rr = (double)res * (double)res;
        } else {
            y[r] = acc;
        }
    }
    if (RESID) {
        const double s = block_sum(rr, s_red);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

// Column-walking SpMV of the extruded operator: thread = face, loop over the
// face's rows z*F + i bottom to top.  Row layout in the CSR values:
// [below (z>0)] [lateral block, sorted] [above (z<L-1)]; slot nibbles give the
// position of the face itself and of neighbour j inside the lateral block.
template <typename T, bool RESID>
__global__ void __launch_bounds__(kBlock) k_spmv_col(int64_t F, int L, const int32_t* __restrict__ rp,
                                                     const T* __restrict__ val, const int32_t* __restrict__ neigh,
                                                     const uint16_t* __restrict__ slots, const T* __restrict__ x,
                                                     const T* __restrict__ b, T* __restrict__ y,
                                                     double* __restrict__ partial) {
    __shared__ double s_red[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double rr = 0.0;
    if (i < F) {
        const uint32_t sl = slots[i];
        int32_t nb[3];
        int pos[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            nb[f] = neigh[f * F + i];
            pos[f] = (sl >> (4 * (f + 1))) & 0xF;
        }
        const int ps = sl & 0xF;
        const int nlat = 1 + (nb[0] >= 0) + (nb[1] >= 0) + (nb[2] >= 0);
        T xb = 0;       // x of the cell below
        T xs = x[i];    // x of this cell
        for (int z = 0; z < L; ++z) {
            const int64_t r = (int64_t)z * F + i;
            const int64_t base = (int64_t)rp[r] + (z > 0 ? 1 : 0);
            T acc = val[base + ps] * xs;
            if (z > 0) acc = fma(val[base - 1], xb, acc);
#pragma unroll
            for (int f = 0; f < 3; ++f)
            {
                // unconditional loads (absent neighbours read a zero), so they issue together
                const bool on = nb[f] >= 0;
                acc = fma(*(on ? &val[base + pos[f]] : &kZeroVal<T>), *(on ? &x[(int64_t)z * F + nb[f]] : &kZeroVal<T>),
                          acc);
            }
            T xa = 0;
            if (z < L - 1) {
                xa = x[r + F];
                acc = fma(val[base + nlat], xa, acc);
            }
            if (RESID) {
                const T res = b[r] - acc;
                y[r] = res;
                rr = fma((double)res, (double)res, rr);
            } else {
                y[r] = acc;
            }
            xb = xs;
            xs = xa;
        }
    }
    if (RESID) {
        const double sum = block_sum(rr, s_red);
        if (threadIdx.x == 0) partial[blockIdx.x] = sum;
    }
}

// Generic CSR SpMV (any row length), one row per thread.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_spmv_plain(int64_t n, const int32_t* __restrict__ rp,
                                                       const int32_t* __restrict__ col, const T* __restrict__ val,
                                                       const T* __restrict__ x, T* __restrict__ y) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    T acc = 0;
    for (int32_t k = rp[r]; k < rp[r + 1]; ++k) acc = fma(val[k], x[col[k]], acc);
    y[r] = acc;
}

// Batched dot products of w against nv <= kNV basis vectors plus w.w: w is
// read once, each basis vector once; FP64 accumulation.  Each wavefront
// shuffle-reduces all nv+1 sums, then the block combines its wavefronts in LDS.
constexpr int kNV = 32;

// four consecutive elements per thread: one 16-byte (float) or two 16-byte
// (double) loads per lane, so each wavefront load moves 1-2 KB
__device__ __forceinline__ void load4(const double* p, double (&o)[4]) {
    const double4 v = *reinterpret_cast<const double4*>(p);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
}
__device__ __forceinline__ void load4(const float* p, double (&o)[4]) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
}
__device__ __forceinline__ void store4(double* p, const double (&o)[4]) {
    *reinterpret_cast<double4*>(p) = double4{o[0], o[1], o[2], o[3]};
}
__device__ __forceinline__ void store4(float* p, const double (&o)[4]) {
    *reinterpret_cast<float4*>(p) = float4{(float)o[0], (float)o[1], (float)o[2], (float)o[3]};
}

// VEC: n % 4 == 0, every vector 4-element aligned
template <typename T, bool VEC>
__global__ void __launch_bounds__(kBlock) k_mdot(int64_t n, const T* __restrict__ V, int nv, const T* __restrict__ w,
                                                 double* __restrict__ partial) {
    __shared__ double s_acc[kBlock / 64][kNV + 1];
    double acc[kNV + 1];
#pragma unroll
    for (int c = 0; c <= kNV; ++c) acc[c] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if constexpr (VEC) {
        const int64_t n4 = n / 4;
        for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n4; g += stride) {
            double wv[4];
            load4(w + 4 * g, wv);
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[kNV] = fma(wv[q], wv[q], acc[kNV]);
#pragma unroll
            for (int c = 0; c < kNV; ++c)
                if (c < nv) {
                    double vv[4];
                    load4(V + (int64_t)c * n + 4 * g, vv);
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[c] = fma(vv[q], wv[q], acc[c]);
                }
        }
    } else {
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
            const double wk = (double)w[k];
            acc[kNV] = fma(wk, wk, acc[kNV]);
#pragma unroll
            for (int c = 0; c < kNV; ++c)
                if (c < nv) acc[c] = fma((double)V[(int64_t)c * n + k], wk, acc[c]);
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c <= kNV; ++c) {
        if (c < nv || c == kNV) {
            const double v = wave_sum(acc[c]);
            if (lane == 0) s_acc[wid][c] = v;
        }
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t <= nv) {
        const int src = t == nv ? kNV : t;
        double s = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += s_acc[q][src];
        partial[(int64_t)blockIdx.x * (nv + 1) + t] = s;
    }
}

__global__ void __launch_bounds__(kBlock) k_reduce_cols(const double* __restrict__ partial, int nblk, int stride,
                                                        double* __restrict__ out) {
    __shared__ double s_red[kBlock / 64];
    const int c = blockIdx.x;
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partial[(int64_t)b * stride + c];
    s = block_sum(s, s_red);
    if (threadIdx.x == 0) out[c] = s;
}

template <typename T, bool VEC>
__global__ void __launch_bounds__(kBlock) k_gs_update(int64_t n, const T* __restrict__ V, int nv,
                                                      const double* __restrict__ h, T* __restrict__ w,
                                                      double* __restrict__ partial) {
    __shared__ double s_red[kBlock / 64];
    double hc[kNV];
#pragma unroll
    for (int c = 0; c < kNV; ++c) hc[c] = c < nv ? h[c] : 0.0;
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if constexpr (VEC) {
        const int64_t n4 = n / 4;
        for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n4; g += stride) {
            double wv[4];
            load4(w + 4 * g, wv);
#pragma unroll
            for (int c = 0; c < kNV; ++c)
                if (c < nv) {
                    double vv[4];
                    load4(V + (int64_t)c * n + 4 * g, vv);
#pragma unroll
                    for (int q = 0; q < 4; ++q) wv[q] = fma(-hc[c], vv[q], wv[q]);
                }
            double wr[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                wr[q] = (double)(T)wv[q];  // the value actually stored
                acc = fma(wr[q], wr[q], acc);
            }
            store4(w + 4 * g, wr);
        }
    } else {
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
            double wk = (double)w[k];
#pragma unroll
            for (int c = 0; c < kNV; ++c)
                if (c < nv) wk = fma(-hc[c], (double)V[(int64_t)c * n + k], wk);
            const T wt = (T)wk;
            w[k] = wt;
            acc = fma((double)wt, (double)wt, acc);
        }
    }
    const double s = block_sum(acc, s_red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_scale_by_inv(int64_t n, const T* __restrict__ w, const double* nrm,
                                                         T* __restrict__ v) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double inv = 1.0 / sqrt(*nrm);
    v[k] = (T)((double)w[k] * inv);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_combine(int64_t n, const T* __restrict__ V, int nv,
                                                    const double* __restrict__ y, T* __restrict__ u) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double s = 0.0;
    for (int c = 0; c < nv; ++c) s = fma(y[c], (double)V[(int64_t)c * n + k], s);
    u[k] = (T)s;
}

__global__ void __launch_bounds__(kBlock) k_absmax(int64_t n, const double* __restrict__ b, double* __restrict__ partial) {
    __shared__ double s_red[kBlock / 64];
    double m = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) m = fmax(m, fabs(b[k]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) s_red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r = fmax(r, s_red[k]);
        partial[blockIdx.x] = r;
    }
}

__global__ void __launch_bounds__(kBlock) k_reduce_max(const double* __restrict__ partial, int nblk, double* out) {
    __shared__ double s_red[kBlock / 64];
    double m = 0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) m = fmax(m, partial[b]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) s_red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) r = fmax(r, s_red[k]);
        *out = r;
    }
}

// Per-face vertical tridiagonal solve with precomputed Thomas factors
// (m_z = d_z - l_z c'_{z-1}, c'_z = u_z / m_z): forward d_z = (r_z - l_z d_{z-1}) / m_z,
// backward z_k = d_k - c'_k z_{k+1}.  LL > 0: column held in registers.
template <typename T, int LL>
__global__ void __launch_bounds__(kBlock) k_pc_column(int64_t F, int Ldyn, const T* __restrict__ lo,
                                                      const T* __restrict__ cp, const T* __restrict__ im,
                                                      const T* __restrict__ r, T* __restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F) return;
    if constexpr (LL > 0) {
        T d[LL];
        T prev = 0;
#pragma unroll
        for (int k = 0; k < LL; ++k) {
            const int64_t q = (int64_t)k * F + i;
            prev = (r[q] - lo[q] * prev) * im[q];
            d[k] = prev;
        }
        T next = d[LL - 1];
        z[(int64_t)(LL - 1) * F + i] = next;
#pragma unroll
        for (int k = LL - 2; k >= 0; --k) {
            next = d[k] - cp[(int64_t)k * F + i] * next;
            z[(int64_t)k * F + i] = next;
        }
    } else {
        T prev = 0;
        for (int k = 0; k < Ldyn; ++k) {
            const int64_t q = (int64_t)k * F + i;
            prev = (r[q] - lo[q] * prev) * im[q];
            z[q] = prev;
        }
        T next = z[(int64_t)(Ldyn - 1) * F + i];
        for (int k = Ldyn - 2; k >= 0; --k) {
            const int64_t q = (int64_t)k * F + i;
            next = z[q] - cp[q] * next;
            z[q] = next;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_pc_jacobi(int64_t n, const double* __restrict__ inv, const T* __restrict__ r,
                                                      T* __restrict__ z) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) z[k] = (T)((double)r[k] * inv[k]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_axpy_dots(int64_t n, double a, const T* __restrict__ x,
                                                      const T* __restrict__ y, T* __restrict__ out,
                                                      const T* __restrict__ rh, double* __restrict__ partial) {
    __shared__ double s_red[kBlock / 64];
    double d1 = 0, d2 = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const T o = (T)fma(-a, (double)x[k], (double)y[k]);
        out[k] = o;
        const double od = (double)o;
        if (rh) d1 = fma((double)rh[k], od, d1);
        d2 = fma(od, od, d2);
    }
    const double s1 = block_sum(d1, s_red);
    const double s2 = block_sum(d2, s_red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s1;
        partial[2 * blockIdx.x + 1] = s2;
    }
}

// ---- BiCGStab with device-resident scalars (slots pb::BicgSc) ----
// The vector kernels walk their vectors four elements per thread and step (16-byte loads
// for FP32) in a grid-stride loop, then the n % 4 tail one element per thread; the
// per-element arithmetic is the same either way.
template <typename T>
using Vec4 = T __attribute__((ext_vector_type(4)));

// p = r + beta (p - omega v), beta = (rho / rho_prev) (alpha / omega) of the previous step;
// after a restart of the shadow residual, p = r and rh = r
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_p_dev(int64_t n, const T* __restrict__ r, const T* __restrict__ v,
                                                       const double* __restrict__ sc, T* __restrict__ p,
                                                       T* __restrict__ rh) {
    if (sc[SC_STOP] != 0.0) return;  // uniform over the grid
    if (sc[SC_RESTART] != 0.0) {     // uniform over the grid
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const int64_t n4 = n / 4;
        for (int64_t i = t0; i < n4; i += stride) {
            const Vec4<T> rv = reinterpret_cast<const Vec4<T>*>(r)[i];
            reinterpret_cast<Vec4<T>*>(p)[i] = rv;
            reinterpret_cast<Vec4<T>*>(rh)[i] = rv;
        }
        for (int64_t k = 4 * n4 + t0; k < n; k += stride) p[k] = rh[k] = r[k];
        return;
    }
    const double omega = sc[SC_OMEGA];
    const double beta = (sc[SC_RHO] / sc[SC_RHO_PREV]) * (sc[SC_ALPHA] / omega);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n4 = n / 4;
    for (int64_t i = t0; i < n4; i += stride) {
        const Vec4<T> rv = reinterpret_cast<const Vec4<T>*>(r)[i];
        const Vec4<T> vv = reinterpret_cast<const Vec4<T>*>(v)[i];
        Vec4<T> pv = reinterpret_cast<const Vec4<T>*>(p)[i];
#pragma unroll
        for (int j = 0; j < 4; ++j) pv[j] = (T)fma(beta, fma(-omega, (double)vv[j], (double)pv[j]), (double)rv[j]);
        reinterpret_cast<Vec4<T>*>(p)[i] = pv;
    }
    for (int64_t k = 4 * n4 + t0; k < n; k += stride)
        p[k] = (T)fma(beta, fma(-omega, (double)v[k], (double)p[k]), (double)r[k]);
}

// Column c of nblk per-block partials (stride doubles apart), summed in a fixed order:
// every workgroup that calls this gets the same value, in thread 0.  Thread t adds entries
// t, t + blockDim.x, ... in turn; the loads of eight turns are issued before their adds, so
// the latency of one L2 round trip is paid once per eight entries, not per entry.
__device__ __forceinline__ double sum_partials(const double* __restrict__ partial, int64_t nblk, int stride, int c,
                                               double* s_red) {
    double a = 0;
    const int64_t step = blockDim.x;
    int64_t b = threadIdx.x;
    for (; b + 7 * step < nblk; b += 8 * step) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = partial[(b + u * step) * stride + c];
#pragma unroll
        for (int u = 0; u < 8; ++u) a += v[u];
    }
    for (; b < nblk; b += step) a += partial[b * stride + c];
    return block_sum(a, s_red);
}
// Both columns of nblk (x, y) pairs in one pass, each summed in the order sum_partials uses.
__device__ __forceinline__ void sum_pairs(const double* __restrict__ partial, int64_t nblk, double* s_red, double& x,
                                          double& y) {
    double a = 0, c = 0;
    const int64_t step = blockDim.x;
    const double2* pr = reinterpret_cast<const double2*>(partial);
    int64_t b = threadIdx.x;
    for (; b + 7 * step < nblk; b += 8 * step) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = pr[b + u * step];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a += v[u].x;
            c += v[u].y;
        }
    }
    for (; b < nblk; b += step) {
        a += pr[b].x;
        c += pr[b].y;
    }
    x = block_sum(a, s_red);
    y = block_sum(c, s_red);
}

// alpha = rho / (rh . v) from the operator's partials (rh.v in column 0 of stride-2 pairs),
// computed by every workgroup alike and recorded by workgroup 0; breakdown (rh.v = 0 or not
// finite): alpha = 0 and flag 1.  Returns alpha to every thread.
__device__ double bicg_alpha(const double* __restrict__ dots, int64_t dots_nb, double* __restrict__ sc,
                             double* s_red, double* s_bc) {
    const double rhv = sum_partials(dots, dots_nb, 2, 0, s_red);
    if (threadIdx.x == 0) {
        const bool ok = rhv != 0.0 && isfinite(rhv);
        const double alpha = ok ? sc[SC_RHO] / rhv : 0.0;
        *s_bc = alpha;
        if (blockIdx.x == 0) {
            sc[SC_RHV] = rhv;
            sc[SC_ALPHA] = alpha;
            if (!ok) sc[SC_FLAG] = 1;  // breakdown: the caller restarts
        }
    }
    __syncthreads();
    return *s_bc;
}

// s = r - alpha v, partial[b] = s.s; alpha from the operator's rh.v partials
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_s_dev(int64_t n, const T* __restrict__ v, const T* __restrict__ r,
                                                       double* __restrict__ sc, T* __restrict__ s_out,
                                                       double* __restrict__ partial, const double* __restrict__ dots,
                                                       int64_t dots_nb) {
    __shared__ double s_red[kBlock / 64];
    __shared__ double s_bc;
    if (sc[SC_STOP] != 0.0) return;  // uniform over the grid
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n4 = n / 4;
    // FP64: operands loaded one step ahead, the first step's while alpha is summed (measured
    // faster for FP64 vectors and slightly slower for FP32, which keeps the plain loop)
    constexpr bool kAhead = sizeof(T) == 8;
    Vec4<T> v_next = {}, r_next = {};
    if (kAhead && t0 < n4) {
        v_next = reinterpret_cast<const Vec4<T>*>(v)[t0];
        r_next = reinterpret_cast<const Vec4<T>*>(r)[t0];
    }
    const double alpha = bicg_alpha(dots, dots_nb, sc, s_red, &s_bc);
    double ss = 0;
    for (int64_t i = t0; i < n4; i += stride) {
        Vec4<T> vv, rv;
        if constexpr (kAhead) {
            vv = v_next;
            rv = r_next;
            if (i + stride < n4) {
                v_next = reinterpret_cast<const Vec4<T>*>(v)[i + stride];
                r_next = reinterpret_cast<const Vec4<T>*>(r)[i + stride];
            }
        } else {
            vv = reinterpret_cast<const Vec4<T>*>(v)[i];
            rv = reinterpret_cast<const Vec4<T>*>(r)[i];
        }
        Vec4<T> o;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            o[j] = (T)fma(-alpha, (double)vv[j], (double)rv[j]);
            ss = fma((double)o[j], (double)o[j], ss);
        }
        reinterpret_cast<Vec4<T>*>(s_out)[i] = o;
    }
    for (int64_t k = 4 * n4 + t0; k < n; k += stride) {
        const T o = (T)fma(-alpha, (double)v[k], (double)r[k]);
        s_out[k] = o;
        ss = fma((double)o, (double)o, ss);
    }
    const double t = block_sum(ss, s_red);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// omega from the operator's (s.t, t.t) partials and the s.s partials of k_bicg_s_dev: the
// minimal-residual value, enlarged when the angle between t and s is wide -- |cos(t, s)| <
// 0.7 -> omega *= 0.7 / |cos| (Sleijpen & van der Vorst 1995), which keeps omega away from 0
// where plain BiCGStab stagnates.  Computed by every workgroup alike, recorded by workgroup
// 0; omega = 0 or non-finite dots: flag 2.  Returns omega to every thread.
__device__ double bicg_omega(const double* __restrict__ dots, int64_t dots_nb, const double* __restrict__ ss_part,
                             int64_t ss_nb, double* __restrict__ sc, double* s_red, double* s_bc) {
    double st, tt;
    sum_pairs(dots, dots_nb, s_red, st, tt);
    const double ss = sum_partials(ss_part, ss_nb, 1, 0, s_red);
    if (threadIdx.x == 0) {
        double omega = 0;
        int flag = 0;
        if (tt > 0.0 && isfinite(st) && isfinite(tt)) {
            omega = st / tt;
            const double cs = st / sqrt(tt * ss);
            if (isfinite(cs) && fabs(cs) < 0.7 && fabs(cs) > 0.0) omega *= 0.7 / fabs(cs);
            if (omega == 0.0) flag = -2;  // only when nothing else is flagged
        } else {
            flag = 2;
        }
        *s_bc = omega;
        if (blockIdx.x == 0) {
            sc[SC_ST] = st;
            sc[SC_TT] = tt;
            sc[SC_SS] = ss;
            sc[SC_OMEGA] = omega;
            if (flag == 2 || (flag == -2 && sc[SC_FLAG] == 0.0)) sc[SC_FLAG] = 2;
        }
    }
    __syncthreads();
    return *s_bc;
}

// x += alpha ph + omega sh;  r = s - omega t;  partial[2b] = rh.r, partial[2b+1] = r.r
// (omega from the operator's s.t, t.t partials and the s.s partials, see bicg_omega)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bicg_xr_dev(int64_t n, const T* __restrict__ ph, const T* __restrict__ sh,
                                                        const T* __restrict__ sv, const T* __restrict__ tv,
                                                        const T* __restrict__ rh, double* __restrict__ sc,
                                                        T* __restrict__ x, T* __restrict__ r,
                                                        double* __restrict__ partial, const double* __restrict__ dots,
                                                        int64_t dots_nb, const double* __restrict__ ss_part,
                                                        int64_t ss_nb) {
    __shared__ double s_red[kBlock / 64];
    __shared__ double s_bc;
    if (sc[SC_STOP] != 0.0) return;  // uniform over the grid
    const double alpha = sc[SC_ALPHA];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n4 = n / 4;
    // FP64: operands one step ahead, the first step's arriving while omega is summed (as in
    // k_bicg_s_dev, FP32 keeps the plain loop)
    constexpr bool kAhead = sizeof(T) == 8;
    Vec4<T> nx[6] = {};
    auto load = [&](int64_t i) {
        nx[0] = reinterpret_cast<const Vec4<T>*>(ph)[i];
        nx[1] = reinterpret_cast<const Vec4<T>*>(sh)[i];
        nx[2] = reinterpret_cast<const Vec4<T>*>(sv)[i];
        nx[3] = reinterpret_cast<const Vec4<T>*>(tv)[i];
        nx[4] = reinterpret_cast<const Vec4<T>*>(rh)[i];
        nx[5] = reinterpret_cast<const Vec4<T>*>(x)[i];
    };
    if (kAhead && t0 < n4) load(t0);
    const double omega = bicg_omega(dots, dots_nb, ss_part, ss_nb, sc, s_red, &s_bc);
    double d1 = 0, d2 = 0;
    for (int64_t i = t0; i < n4; i += stride) {
        if (!kAhead) load(i);
        const Vec4<T> phv = nx[0], shv = nx[1], svv = nx[2], tvv = nx[3], rhv = nx[4];
        Vec4<T> xv = nx[5];
        if (kAhead && i + stride < n4) load(i + stride);
        Vec4<T> o;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            xv[j] = (T)fma(omega, (double)shv[j], fma(alpha, (double)phv[j], (double)xv[j]));
            o[j] = (T)fma(-omega, (double)tvv[j], (double)svv[j]);
            d1 = fma((double)rhv[j], (double)o[j], d1);
            d2 = fma((double)o[j], (double)o[j], d2);
        }
        reinterpret_cast<Vec4<T>*>(x)[i] = xv;
        reinterpret_cast<Vec4<T>*>(r)[i] = o;
    }
    for (int64_t k = 4 * n4 + t0; k < n; k += stride) {
        x[k] = (T)fma(omega, (double)sh[k], fma(alpha, (double)ph[k], (double)x[k]));
        const T o = (T)fma(-omega, (double)tv[k], (double)sv[k]);
        r[k] = o;
        d1 = fma((double)rh[k], (double)o, d1);
        d2 = fma((double)o, (double)o, d2);
    }
    const double a = block_sum(d1, s_red);
    const double b = block_sum(d2, s_red);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

// Sums the (rh.r, r.r) pairs of partial (stride, fixed order) in one block and
// updates the scalars of the given BiCGStab stage (see bicg_stage()).
constexpr int kStageThreads = 1024;
__global__ void __launch_bounds__(kStageThreads) k_bicg_stage(const double* __restrict__ partial, int64_t nblk,
                                                              int stride, int stage, double* __restrict__ sc,
                                                              double stop_abs, double* __restrict__ host_slot,
                                                              double restart_eps) {
    __shared__ double s_red[kStageThreads / 64];
    if (stage != 0 && sc[SC_STOP] != 0.0) return;  // uniform over the block
    double a, b;  // rh.r, r.r
    if (stride == 2) {
        sum_pairs(partial, nblk, s_red, a, b);
    } else {
        a = sum_partials(partial, nblk, stride, 0, s_red);
        b = sum_partials(partial, nblk, stride, 1, s_red);
    }
    if (threadIdx.x != 0) return;
    if (stage == 0) {  // the starting residual
        sc[SC_RHO] = a;
        sc[SC_RR] = b;
        sc[SC_RHO_PREV] = 1;
        sc[SC_ALPHA] = 1;
        sc[SC_OMEGA] = 1;
        sc[SC_FLAG] = 0;
        sc[SC_STOP] = 0;
        sc[SC_RHNN] = a;  // rh = r
        sc[SC_RESTART] = 0;
        sc[SC_NRESTART] = 0;
    } else {  // the new residual of an iteration, and the stop test
        sc[SC_RHO_PREV] = sc[SC_RHO];
        sc[SC_RHO] = a;
        sc[SC_RR] = b;
        if (!isfinite(b)) sc[SC_FLAG] = 3;
        const double rr = sqrt(b);
        if (sc[SC_FLAG] != 0.0 || rr <= stop_abs || !isfinite(rr)) sc[SC_STOP] = 1;
        // near breakdown: rh is nearly orthogonal to r, so rho carries few significant digits
        // and alpha = rho / rh.v swings; take rh = r (rho = r.r) and restart the directions
        const bool restart = sc[SC_STOP] == 0.0 && restart_eps > 0.0 && fabs(a) < restart_eps * sqrt(sc[SC_RHNN] * b);
        sc[SC_RESTART] = restart ? 1.0 : 0.0;
        if (restart) {
            sc[SC_RHO] = b;
            sc[SC_RHNN] = b;
            sc[SC_NRESTART] += 1.0;
        }
    }
    if (host_slot)  // the scalars straight into the host-visible slot the host waits on
        for (int k = 0; k < SC_COUNT; ++k) host_slot[k] = sc[k];
}

template <typename T>
__global__ void k_scale_host(int64_t n, const T* a, double alpha, T* b) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = (T)((double)a[k] * alpha);
}

__global__ void k_to_f32(int64_t n, const double* a, float* b) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = (float)a[k];
}
__global__ void k_axpy_f32_f64(int64_t n, double alpha, const float* x, double* y) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = fma(alpha, (double)x[k], y[k]);
}
__global__ void k_scale_f64_f32(int64_t n, const double* r, double alpha, float* out) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = (float)(r[k] * alpha);
}
__global__ void k_axpy_dev_f32(int64_t n, const double* scale, const float* x, double* y) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = fma(*scale, (double)x[k], y[k]);
}
__global__ void k_axpy_f64(int64_t n, double alpha, const double* x, double* y) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) y[k] = fma(alpha, x[k], y[k]);
}
__global__ void k_copy_f64(int64_t n, const double* a, double* b) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = a[k];
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
void spmv(hipStream_t s, int64_t n, const int32_t* rp, const int32_t* col, const T* val, const T* x, T* y) {
    const unsigned g = (unsigned)((n + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL((k_spmv_lds<T, false>), dim3(g), dim3(kBlock), 0, s, n, rp, col, val, x, (const T*)nullptr, y,
                       (double*)nullptr);
}

template <typename T>
void residual(hipStream_t s, int64_t n, const int32_t* rp, const int32_t* col, const T* val, const T* x, const T* b,
              T* r, double* partial, int nblk) {
    const unsigned g = (unsigned)((n + kRowsPerBlock - 1) / kRowsPerBlock);
    (void)nblk;
    hipLaunchKernelGGL((k_spmv_lds<T, true>), dim3(g), dim3(kBlock), 0, s, n, rp, col, val, x, b, r, partial);
}

template <typename T>
void spmv_col(hipStream_t s, int64_t F, int L, const int32_t* rp, const T* val, const int32_t* neigh,
              const uint16_t* slots, const T* x, const T* b, T* y, double* partial) {
    const dim3 g((unsigned)spmv_col_blocks(F)), blk(kBlock);
    if (b)
        hipLaunchKernelGGL((k_spmv_col<T, true>), g, blk, 0, s, F, L, rp, val, neigh, slots, x, b, y, partial);
    else
        hipLaunchKernelGGL((k_spmv_col<T, false>), g, blk, 0, s, F, L, rp, val, neigh, slots, x, b, y, partial);
}

// 4-wide loads need n % 4 == 0 and 32-byte aligned vector starts
template <typename T>
static bool vec4_ok(int64_t n, const void* a, const void* b) {
    return n % 4 == 0 && ((uintptr_t)a % 32) == 0 && ((uintptr_t)b % 32) == 0;
}

template <typename T>
void mdot(hipStream_t s, int64_t n, const T* V, int nv, const T* w, double* partial, int nblk) {
    if (vec4_ok<T>(n, V, w))
        hipLaunchKernelGGL((k_mdot<T, true>), dim3(nblk), dim3(kBlock), 0, s, n, V, nv, w, partial);
    else
        hipLaunchKernelGGL((k_mdot<T, false>), dim3(nblk), dim3(kBlock), 0, s, n, V, nv, w, partial);
}

void reduce_cols(hipStream_t s, const double* partial, int nblk, int stride, int ncols, double* out) {
    hipLaunchKernelGGL(k_reduce_cols, dim3(ncols), dim3(kBlock), 0, s, partial, nblk, stride, out);
}

template <typename T>
void gs_update(hipStream_t s, int64_t n, const T* V, int nv, const double* h, T* w, double* partial, int nblk) {
    if (vec4_ok<T>(n, V, w))
        hipLaunchKernelGGL((k_gs_update<T, true>), dim3(nblk), dim3(kBlock), 0, s, n, V, nv, h, w, partial);
    else
        hipLaunchKernelGGL((k_gs_update<T, false>), dim3(nblk), dim3(kBlock), 0, s, n, V, nv, h, w, partial);
}

template <typename T>
void scale_by_inv(hipStream_t s, int64_t n, const T* w, const double* nrm, T* v) {
    hipLaunchKernelGGL((k_scale_by_inv<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, w, nrm, v);
}

template <typename T>
void combine(hipStream_t s, int64_t n, const T* V, int nv, const double* y, T* u) {
    hipLaunchKernelGGL((k_combine<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, V, nv, y, u);
}

void absmax(hipStream_t s, int64_t n, const double* b, double* partial, int nblk) {
    hipLaunchKernelGGL(k_absmax, dim3(nblk), dim3(kBlock), 0, s, n, b, partial);
}
void reduce_max(hipStream_t s, const double* partial, int nblk, double* out) {
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(kBlock), 0, s, partial, nblk, out);
}

template <typename T>
void pc_column(hipStream_t s, int64_t F, int L, const T* lo, const T* cp, const T* im, const T* r, T* z) {
    const dim3 g(grid_for(F)), b(kBlock);
    switch (L) {
#define PB_COL_CASE(LL) \
    case LL: hipLaunchKernelGGL((k_pc_column<T, LL>), g, b, 0, s, F, L, lo, cp, im, r, z); break;
        PB_COL_CASE(2) PB_COL_CASE(3) PB_COL_CASE(4) PB_COL_CASE(5) PB_COL_CASE(6) PB_COL_CASE(7) PB_COL_CASE(8)
        PB_COL_CASE(9) PB_COL_CASE(10) PB_COL_CASE(11) PB_COL_CASE(12) PB_COL_CASE(16) PB_COL_CASE(20)
#undef PB_COL_CASE
        default: hipLaunchKernelGGL((k_pc_column<T, 0>), g, b, 0, s, F, L, lo, cp, im, r, z); break;
    }
}

template <typename T>
void pc_jacobi(hipStream_t s, int64_t n, const double* invdiag, const T* r, T* z) {
    hipLaunchKernelGGL((k_pc_jacobi<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, invdiag, r, z);
}

template <typename T>
void axpy_dots(hipStream_t s, int64_t n, double a, const T* x, const T* y, T* out, const T* rh, double* partial,
               int nblk) {
    hipLaunchKernelGGL((k_axpy_dots<T>), dim3(nblk), dim3(kBlock), 0, s, n, a, x, y, out, rh, partial);
}
// the BiCGStab vector kernels read and write four elements at a time
template <typename T>
void check_vec4_aligned(std::initializer_list<const T*> ptrs) {
    for (const T* p : ptrs)
        if (reinterpret_cast<uintptr_t>(p) % (4 * sizeof(T)) != 0)
            throw Error(PBSM3D_EINVAL, "BiCGStab vectors must be aligned to four elements");
}
template <typename T>
void bicg_p_dev(hipStream_t s, int64_t n, const T* r, const T* v, const double* sc, T* p, T* rh) {
    check_vec4_aligned<T>({r, v, p, rh});
    hipLaunchKernelGGL((k_bicg_p_dev<T>), dim3(grid_for((n + 3) / 4, kBlock, 8192)), dim3(kBlock), 0, s, n, r, v, sc,
                       p, rh);
}
template <typename T>
void bicg_s_dev(hipStream_t s, int64_t n, const T* v, const T* r, double* sc, T* s_out, double* partial, int nblk,
                const double* dots, int64_t dots_nb) {
    check_vec4_aligned<T>({v, r, s_out});
    if (partial == dots) throw Error(PBSM3D_EINVAL, "s.s partials must not overwrite the dots they follow");
    hipLaunchKernelGGL((k_bicg_s_dev<T>), dim3(nblk), dim3(kBlock), 0, s, n, v, r, sc, s_out, partial, dots, dots_nb);
}
template <typename T>
void bicg_xr_dev(hipStream_t s, int64_t n, const T* ph, const T* sh, const T* sv, const T* tv, const T* rh, double* sc,
                 T* x, T* r, double* partial, int nblk, const double* dots, int64_t dots_nb, const double* ss_part,
                 int64_t ss_nb) {
    check_vec4_aligned<T>({ph, sh, sv, tv, rh, x, r});
    if (partial == dots || partial == ss_part)
        throw Error(PBSM3D_EINVAL, "x/r partials must not overwrite the partials they read");
    if (reinterpret_cast<uintptr_t>(dots) % 16 != 0) throw Error(PBSM3D_EINVAL, "dot pairs must be 16-byte aligned");
    hipLaunchKernelGGL((k_bicg_xr_dev<T>), dim3(nblk), dim3(kBlock), 0, s, n, ph, sh, sv, tv, rh, sc, x, r, partial,
                       dots, dots_nb, ss_part, ss_nb);
}
void bicg_stage(hipStream_t s, const double* partial, int64_t nblk, int stride, int stage, double* sc,
                double stop_abs, double* host_slot, double restart_eps) {
    if (stride == 2 && reinterpret_cast<uintptr_t>(partial) % 16 != 0)
        throw Error(PBSM3D_EINVAL, "dot pairs must be 16-byte aligned");
    hipLaunchKernelGGL(k_bicg_stage, dim3(1), dim3(kStageThreads), 0, s, partial, nblk, stride, stage, sc, stop_abs,
                       host_slot, restart_eps);
}

template <typename T>
void scale_host(hipStream_t s, int64_t n, const T* a, double alpha, T* b) {
    hipLaunchKernelGGL((k_scale_host<T>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, alpha, b);
}

void to_f32(hipStream_t s, int64_t n, const double* a, float* b) {
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, b);
}
void axpy_f32_into_f64(hipStream_t s, int64_t n, double alpha, const float* x, double* y) {
    hipLaunchKernelGGL(k_axpy_f32_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, n, alpha, x, y);
}
void scale_f64_to_f32_host(hipStream_t s, int64_t n, const double* r, double alpha, float* out) {
    hipLaunchKernelGGL(k_scale_f64_f32, dim3(grid_for(n)), dim3(kBlock), 0, s, n, r, alpha, out);
}
void axpy_dev_scaled_f32(hipStream_t s, int64_t n, const double* scale, const float* x, double* y) {
    hipLaunchKernelGGL(k_axpy_dev_f32, dim3(grid_for(n)), dim3(kBlock), 0, s, n, scale, x, y);
}
void axpy_f64(hipStream_t s, int64_t n, double alpha, const double* x, double* y) {
    hipLaunchKernelGGL(k_axpy_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, n, alpha, x, y);
}
void copy_f64(hipStream_t s, int64_t n, const double* a, double* b) {
    hipLaunchKernelGGL(k_copy_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, b);
}

#define PB_INST(T)                                                                                                 \
    template void spmv<T>(hipStream_t, int64_t, const int32_t*, const int32_t*, const T*, const T*, T*);           \
    template void residual<T>(hipStream_t, int64_t, const int32_t*, const int32_t*, const T*, const T*, const T*,   \
                              T*, double*, int);                                                                   \
    template void mdot<T>(hipStream_t, int64_t, const T*, int, const T*, double*, int);                            \
    template void gs_update<T>(hipStream_t, int64_t, const T*, int, const double*, T*, double*, int);              \
    template void scale_by_inv<T>(hipStream_t, int64_t, const T*, const double*, T*);                              \
    template void combine<T>(hipStream_t, int64_t, const T*, int, const double*, T*);                              \
    template void pc_column<T>(hipStream_t, int64_t, int, const T*, const T*, const T*, const T*, T*);             \
    template void pc_jacobi<T>(hipStream_t, int64_t, const double*, const T*, T*);                                \
    template void scale_host<T>(hipStream_t, int64_t, const T*, double, T*);                                       \
    template void spmv_col<T>(hipStream_t, int64_t, int, const int32_t*, const T*, const int32_t*, const uint16_t*, \
                              const T*, const T*, T*, double*);                                                    \
    template void axpy_dots<T>(hipStream_t, int64_t, double, const T*, const T*, T*, const T*, double*, int);      \
    template void bicg_p_dev<T>(hipStream_t, int64_t, const T*, const T*, const double*, T*, T*);                     \
    template void bicg_s_dev<T>(hipStream_t, int64_t, const T*, const T*, double*, T*, double*, int,             \
                                const double*, int64_t);                                                          \
    template void bicg_xr_dev<T>(hipStream_t, int64_t, const T*, const T*, const T*, const T*, const T*, double*, \
                                 T*, T*, double*, int, const double*, int64_t, const double*, int64_t);
PB_INST(double)
PB_INST(float)
#undef PB_INST

}  // namespace pb
