// Learned coarse correction of the suspension solve (no counterpart in the
// reference, whose solve starts from x0 = 0 with ILUT as the only preconditioner,
// LinearAlgebra.cpp:178-206).
//
// The network's predicted profile x0 (the learned warm start), restricted to T
// tiles of faces, spans a T-dimensional coarse space: column a of W holds x0 on
// the rows (every layer) of tile a's faces and zero elsewhere.  At the start of
// every FP64 refinement cycle of the split BiCGStab the iterate is corrected by
// the Galerkin projection
//     x += W (W^T A W)^-1 W^T r,   r = b - A x,
// after which the residual is orthogonal to every column of W (deflation of the
// coarse space).  Tiles are ranges of face indices, compact patches on meshes
// numbered along a space-filling curve.  Every reduction runs in a fixed order,
// so the correction is deterministic.
#include <algorithm>
#include <cmath>
#include <vector>

#include "context.hpp"

namespace pb {
namespace {

constexpr int kInvThreads = 1024;

// G[a][b] = sum over the rows r of tile a's faces of w_r A_rj w_j, j on tile b's faces;
// one workgroup per tile a, b among a's adjacent tiles (slot 0 = a itself)
__global__ void __launch_bounds__(kBlock) k_coarse_gram(int64_t F, int L, int T, const int32_t* __restrict__ rp,
                                                        const int32_t* __restrict__ col,
                                                        const double* __restrict__ val,
                                                        const double* __restrict__ w,
                                                        const int64_t* __restrict__ tile_off,
                                                        const int32_t* __restrict__ tile_of_face,
                                                        const int32_t* __restrict__ adj, double* __restrict__ G) {
    const int a = blockIdx.x;
    __shared__ int32_t s_adj[kCoarseAdj];
    __shared__ double s_red[kBlock / 64];
    if (threadIdx.x < kCoarseAdj) s_adj[threadIdx.x] = adj[(int64_t)a * kCoarseAdj + threadIdx.x];
    __syncthreads();
    const int64_t f0 = tile_off[a], nf = tile_off[a + 1] - f0;
    double acc[kCoarseAdj];
#pragma unroll
    for (int s = 0; s < kCoarseAdj; ++s) acc[s] = 0.0;
    for (int64_t idx = threadIdx.x; idx < nf * L; idx += kBlock) {
        const int64_t z = idx / nf;
        const int64_t r = z * F + f0 + (idx - z * nf);
        const double wr = w[r];
        if (wr == 0.0) continue;
        for (int32_t e = rp[r]; e < rp[r + 1]; ++e) {
            const int64_t j = col[e];
            const double wj = w[j];
            if (wj == 0.0) continue;
            const int32_t tj = tile_of_face[j % F];
            const double v = wr * val[e] * wj;
#pragma unroll
            for (int s = 0; s < kCoarseAdj; ++s) acc[s] += s_adj[s] == tj ? v : 0.0;
        }
    }
#pragma unroll
    for (int s = 0; s < kCoarseAdj; ++s) {
        const double sum = block_sum(acc[s], s_red);
        if (threadIdx.x == 0 && s_adj[s] >= 0) {
            // a tile where the network predicts no suspension (w = 0 on all its rows) is
            // decoupled: G[a][a] = 1, and its restriction is 0
            G[(int64_t)a * T + s_adj[s]] = (s == 0 && sum == 0.0) ? 1.0 : sum;
        }
    }
}

// Ginv = G^-1 by Gauss-Jordan elimination with partial pivoting (largest |pivot|,
// smallest row on ties) on M = [G | I] (T x 2T); one workgroup.  A zero pivot
// column leaves its row of Ginv zero (that coarse unknown is not corrected).
__global__ void __launch_bounds__(kInvThreads) k_coarse_invert(int T, const double* __restrict__ G,
                                                               double* __restrict__ M, double* __restrict__ Ginv) {
    __shared__ double s_val[kInvThreads];
    __shared__ int s_idx[kInvThreads];
    __shared__ double s_f[kCoarseMaxTiles];
    __shared__ double s_piv;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int64_t W2 = 2 * (int64_t)T, TW = (int64_t)T * W2;
    const int R = nt / (int)W2;  // T <= kCoarseMaxTiles: W2 <= 512, so R >= 2
    const int ti = tid / (int)W2, tj = tid - ti * (int)W2;
    for (int64_t e = tid; e < TW; e += nt) {
        const int64_t i = e / W2, j = e - i * W2;
        M[e] = j < T ? G[i * T + j] : (j - T == i ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int k = 0; k < T; ++k) {
        double best = -1.0;
        int bi = T;
        for (int i = k + tid; i < T; i += nt) {
            const double v = fabs(M[i * W2 + k]);
            if (v > best) {
                best = v;
                bi = i;
            }
        }
        s_val[tid] = best;
        s_idx[tid] = bi;
        __syncthreads();
        for (int off = nt / 2; off > 0; off >>= 1) {
            if (tid < off) {
                const double v = s_val[tid + off];
                const int ix = s_idx[tid + off];
                if (v > s_val[tid] || (v == s_val[tid] && ix < s_idx[tid])) {
                    s_val[tid] = v;
                    s_idx[tid] = ix;
                }
            }
            __syncthreads();
        }
        const int p = s_idx[0];
        if (p != k && p < T)
            for (int64_t j = tid; j < W2; j += nt) {
                const double t = M[k * W2 + j];
                M[k * W2 + j] = M[p * W2 + j];
                M[p * W2 + j] = t;
            }
        __syncthreads();
        if (tid == 0) {
            const double pv = M[k * W2 + k];
            s_piv = pv != 0.0 ? 1.0 / pv : 0.0;
        }
        __syncthreads();
        for (int64_t j = tid; j < W2; j += nt) M[k * W2 + j] *= s_piv;
        __syncthreads();
        for (int i = tid; i < T; i += nt) s_f[i] = i == k ? 0.0 : M[i * W2 + k];
        __syncthreads();
        // rows ti, ti + R, ... and column tj of this thread (R = nt / W2 >= 2 rows per pass)
        if (tid < R * W2) {
            const double mk = M[k * W2 + tj];
            for (int i = ti; i < T; i += R) {
                const double f = s_f[i];
                if (f != 0.0) M[i * W2 + tj] -= f * mk;
            }
        }
        __syncthreads();
    }
    for (int64_t e = tid; e < (int64_t)T * T; e += nt) {
        const int64_t i = e / T, j = e - i * T;
        Ginv[e] = M[i * W2 + T + j];
    }
}

// c[a] = sum over the rows of tile a of w_r r_r (reference order)
__global__ void __launch_bounds__(kBlock) k_coarse_restrict(int64_t F, int L, const int64_t* __restrict__ tile_off,
                                                            const double* __restrict__ w,
                                                            const double* __restrict__ r, double* __restrict__ c) {
    __shared__ double s_red[kBlock / 64];
    const int a = blockIdx.x;
    const int64_t f0 = tile_off[a], nf = tile_off[a + 1] - f0;
    double acc = 0.0;
    for (int64_t idx = threadIdx.x; idx < nf * L; idx += kBlock) {
        const int64_t z = idx / nf;
        const int64_t q = z * F + f0 + (idx - z * nf);
        acc = fma(w[q], r[q], acc);
    }
    const double s = block_sum(acc, s_red);
    if (threadIdx.x == 0) c[a] = s;
}

// y = Ginv c, one thread per coarse unknown (T <= kBlock)
__global__ void __launch_bounds__(kBlock) k_coarse_solve(int T, const double* __restrict__ Ginv,
                                                         const double* __restrict__ c, double* __restrict__ y) {
    const int i = threadIdx.x;
    if (i >= T) return;
    double acc = 0.0;
    for (int j = 0; j < T; ++j) acc = fma(Ginv[(int64_t)i * T + j], c[j], acc);
    y[i] = acc;
}

// x += W y in position order: x[q] += wpos[q] y[tile of the face at position k] for the
// rows q = pos_row(k, z) of position k
__global__ void __launch_bounds__(kBlock) k_coarse_prolong(int64_t F, int L, const int32_t* __restrict__ tile_of_pos,
                                                           const double* __restrict__ wpos,
                                                           const double* __restrict__ y, double* __restrict__ x) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < F; k += (int64_t)gridDim.x * blockDim.x) {
        const double yk = y[tile_of_pos[k]];
        const int64_t qb = pos_base64(k, L), qs = pos_stride64(k, F);
        for (int z = 0; z < L; ++z) {
            const int64_t q = qb + (int64_t)z * qs;
            x[q] = fma(wpos[q], yk, x[q]);
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_coarse_fill(int64_t n, double v, double* __restrict__ a) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

}  // namespace

void coarse_setup(pbsm3d_ctx* c, int T, const int32_t* nb) {
    CoarseSpace& cs = c->coarse;
    if (cs.T == T) return;
    if (T < 1 || T > kCoarseMaxTiles) throw Error(PBSM3D_EINVAL, "coarse_tiles must be in [0, 256]");
    const int64_t F = c->F, N = F * c->L;
    if (T > F) throw Error(PBSM3D_EINVAL, "coarse_tiles exceeds the face count");
    if ((int64_t)c->co.h_faces.size() != F) throw Error(PBSM3D_EINVAL, "the coarse correction needs the D-ILU order");
    hipStream_t s = c->stream;
    std::vector<int64_t> off(T + 1);
    for (int a = 0; a <= T; ++a) off[a] = (int64_t)a * F / T;
    std::vector<int32_t> tof(F);
    for (int a = 0; a < T; ++a) std::fill(tof.begin() + off[a], tof.begin() + off[a + 1], a);
    // tile adjacency through the face neighbours (host mesh array, face-major [3 i + j])
    const std::vector<int32_t>& faces = c->co.h_faces;
    std::vector<std::vector<int32_t>> adj(T);
    for (int a = 0; a < T; ++a) adj[a].push_back(a);
    for (int64_t i = 0; i < F; ++i)
        for (int j = 0; j < 3; ++j) {
            const int32_t n = nb[3 * i + j];
            if (n < 0) continue;
            auto& v = adj[tof[i]];
            if (std::find(v.begin(), v.end(), tof[n]) == v.end()) v.push_back(tof[n]);
        }
    std::vector<int32_t> adjf((size_t)T * kCoarseAdj, -1);
    for (int a = 0; a < T; ++a) {
        if ((int)adj[a].size() > kCoarseAdj)
            throw Error(PBSM3D_EINVAL, "a coarse tile touches more than 15 others: use fewer coarse_tiles");
        std::sort(adj[a].begin() + 1, adj[a].end());
        std::copy(adj[a].begin(), adj[a].end(), adjf.begin() + (size_t)a * kCoarseAdj);
    }
    std::vector<int32_t> top(F);
    for (int64_t k = 0; k < F; ++k) top[k] = tof[faces[k]];
    cs.tile_off.alloc(T + 1);
    cs.tile_of_face.alloc(F);
    cs.tile_of_pos.alloc(F);
    cs.adj.alloc((size_t)T * kCoarseAdj);
    cs.w.alloc(N);
    cs.wpos.alloc(N);
    cs.rref.alloc(N);
    cs.G.alloc((size_t)T * T);
    cs.M.alloc((size_t)T * 2 * T);
    cs.Ginv.alloc((size_t)T * T);
    cs.c.alloc(T);
    cs.y.alloc(T);
    PB_HIP(hipMemcpyAsync(cs.tile_off.p, off.data(), (T + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(cs.tile_of_face.p, tof.data(), F * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(cs.tile_of_pos.p, top.data(), F * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(cs.adj.p, adjf.data(), adjf.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipStreamSynchronize(s));
    cs.T = T;
}

void coarse_prepare(pbsm3d_ctx* c) {
    CoarseSpace& cs = c->coarse;
    hipStream_t s = c->stream;
    const int64_t F = c->F, N = F * c->L;
    const int T = cs.T;
    // the basis: the network's predicted profile, as the learned start wrote it into x
    PB_HIP(hipMemcpyAsync(cs.w.p, c->susp.x.p, N * sizeof(double), hipMemcpyDeviceToDevice, s));
    to_pos<double, double>(s, c->co, F, c->L, cs.w.p, cs.wpos.p);
    hipLaunchKernelGGL(k_coarse_fill, dim3(grid_for((int64_t)T * T)), dim3(kBlock), 0, s, (int64_t)T * T, 0.0, cs.G.p);
    hipLaunchKernelGGL(k_coarse_gram, dim3(T), dim3(kBlock), 0, s, F, c->L, T, c->susp.rp.p, c->susp.col.p,
                       c->susp.val.p, cs.w.p, cs.tile_off.p, cs.tile_of_face.p, cs.adj.p, cs.G.p);
    hipLaunchKernelGGL(k_coarse_invert, dim3(1), dim3(kInvThreads), 0, s, T, cs.G.p, cs.M.p, cs.Ginv.p);
    PB_HIP(hipGetLastError());
    cs.ready = true;
}

void coarse_correct(pbsm3d_ctx* c, hipStream_t s, const double* r_pos, double* x_pos) {
    CoarseSpace& cs = c->coarse;
    const int64_t F = c->F;
    const int L = c->L;
    from_pos<double, double>(s, c->co, F, L, r_pos, cs.rref.p);
    hipLaunchKernelGGL(k_coarse_restrict, dim3(cs.T), dim3(kBlock), 0, s, F, L, cs.tile_off.p, cs.w.p, cs.rref.p,
                       cs.c.p);
    hipLaunchKernelGGL(k_coarse_solve, dim3(1), dim3(kBlock), 0, s, cs.T, cs.Ginv.p, cs.c.p, cs.y.p);
    hipLaunchKernelGGL(k_coarse_prolong, dim3(grid_for(F)), dim3(kBlock), 0, s, F, L, cs.tile_of_pos.p, cs.wpos.p,
                       cs.y.p, x_pos);
    PB_HIP(hipGetLastError());
}

}  // namespace pb

int pbsm3d_coarse_system(pbsm3d_ctx* c, int32_t* tile_of_face, double* w, double* G) {
    if (!c) return PBSM3D_EINVAL;
    try {
        const pb::CoarseSpace& cs = c->coarse;
        if (!cs.ready) throw pb::Error(PBSM3D_EINVAL, "no step has used the coarse correction yet");
        PB_HIP(hipSetDevice(c->device));
        const int64_t F = c->F, N = F * c->L;
        if (tile_of_face)
            PB_HIP(hipMemcpyAsync(tile_of_face, cs.tile_of_face.p, F * sizeof(int32_t), hipMemcpyDeviceToHost,
                                  c->stream));
        if (w) PB_HIP(hipMemcpyAsync(w, cs.w.p, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (G)
            PB_HIP(hipMemcpyAsync(G, cs.G.p, (size_t)cs.T * cs.T * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        PB_HIP(hipStreamSynchronize(c->stream));
        return PBSM3D_OK;
    } catch (const pb::Error& e) {
        c->err = e.what();
        return e.code;
    }
}
