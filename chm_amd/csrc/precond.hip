// Multicolour D-ILU(0) for the extruded prism-column systems.
//
// Rows are taken in the order (face colour, layer, face).  A face's lateral
// neighbours all carry other colours, so within one colour the face columns
// are independent: a colour is one kernel launch in which a thread owns a
// face and walks its column -- bottom to top for the factorisation and the
// forward solve, top to bottom for the backward solve.  Two triangles sharing
// an edge have no third common neighbour except around a degree-3 vertex, so
// ILU(0) in this order changes only the diagonal (D-ILU):
//   d~_r = a_rr - sum_{k lower} a_rk a_kr / d~_k
//   M = (D~ + L) D~^-1 (D~ + U),  L / U the strict lower / upper parts of A.
//
// Layout: faces are given colour-major positions k, and the factor data
// (lateral and vertical couplings, 1/d~, the intermediate y) are stored at
// q = z*F + k, so a colour phase touches one contiguous stretch per layer.
// The factorisation gathers each row from the CSR once per step; the apply
// gathers r and scatters z once each.  The same kernels serve the 2-D
// deposition system (one layer).
#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace pb {
namespace {

__device__ __forceinline__ int nib(uint32_t slot, int k) { return (slot >> (4 * k)) & 0xF; }

template <typename T>
__global__ void __launch_bounds__(kBlock) k_dilu_factor(int64_t F, int L, int64_t k0, int64_t k1,
                                                        const int32_t* __restrict__ faces,
                                                        const int32_t* __restrict__ nbp,
                                                        const uint8_t* __restrict__ rev,
                                                        const uint8_t* __restrict__ lowmask,
                                                        const int32_t* __restrict__ rp, const T* __restrict__ val,
                                                        const uint16_t* __restrict__ slots, T* __restrict__ lat,
                                                        T* __restrict__ vert, T* __restrict__ dinv) {
    const int64_t k = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= k1) return;
    const int64_t N = F * L;
    const int64_t i = faces[k];
    const uint32_t si = slots[i];
    const uint32_t lm = lowmask[k];
    int32_t np_[3], nf[3];
    uint32_t sn[3];
    int rv[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        np_[f] = nbp[f * F + k];
        const bool low = (lm >> f) & 1u;
        nf[f] = low ? faces[np_[f]] : 0;
        sn[f] = low ? slots[nf[f]] : 0u;
        rv[f] = low ? rev[f * F + k] : 0;
    }
    T dprev = 0;
    for (int z = 0; z < L; ++z) {
        const int64_t r = (int64_t)z * F + i;
        const int64_t q = (int64_t)z * F + k;
        const int64_t base = (int64_t)rp[r] + (z > 0 ? 1 : 0);
        T d = val[base + nib(si, 0)];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const T a_in = np_[f] >= 0 ? val[base + nib(si, f + 1)] : T(0);
            lat[f * N + q] = a_in;
            if ((lm >> f) & 1u) {
                const int64_t rn = (int64_t)z * F + nf[f];
                const int64_t bn = (int64_t)rp[rn] + (z > 0 ? 1 : 0);
                const T a_ni = val[bn + nib(sn[f], rv[f] + 1)];
                d -= a_in * a_ni * dinv[(int64_t)z * F + np_[f]];
            }
        }
        const T a_ib = z > 0 ? val[rp[r]] : T(0);            // (r, r-F): first entry of row r
        const T a_ia = z < L - 1 ? val[rp[r + 1] - 1] : T(0);  // (r, r+F): last entry of row r
        vert[q] = a_ib;
        vert[N + q] = a_ia;
        if (z > 0) {
            const T a_bi = val[rp[r - F + 1] - 1];  // (r-F, r): last entry of row r-F
            d -= a_ib * a_bi * dprev;
        }
        const T di = d != T(0) ? T(1) / d : T(1);
        dinv[q] = di;
        dprev = di;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_dilu_fwd(int64_t F, int L, int64_t k0, int64_t k1,
                                                     const int32_t* __restrict__ faces, const int32_t* __restrict__ nbp,
                                                     const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                                     const T* __restrict__ vert, const T* __restrict__ dinv,
                                                     const T* __restrict__ rin, T* __restrict__ y) {
    const int64_t k = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= k1) return;
    const int64_t N = F * L;
    const int64_t i = faces[k];
    const uint32_t lm = lowmask[k];
    int32_t np_[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) np_[f] = nbp[f * F + k];
    T prev = 0;
    for (int z = 0; z < L; ++z) {
        const int64_t q = (int64_t)z * F + k;
        T acc = rin[(int64_t)z * F + i];
#pragma unroll
        for (int f = 0; f < 3; ++f)
            if ((lm >> f) & 1u) acc -= lat[f * N + q] * y[(int64_t)z * F + np_[f]];
        if (z > 0) acc -= vert[q] * prev;
        prev = acc * dinv[q];
        y[q] = prev;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_dilu_bwd(int64_t F, int L, int64_t k0, int64_t k1,
                                                     const int32_t* __restrict__ faces, const int32_t* __restrict__ nbp,
                                                     const uint8_t* __restrict__ lowmask, const T* __restrict__ lat,
                                                     const T* __restrict__ vert, const T* __restrict__ dinv,
                                                     T* __restrict__ y, T* __restrict__ zout) {
    const int64_t k = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= k1) return;
    const int64_t N = F * L;
    const int64_t i = faces[k];
    const uint32_t lm = lowmask[k];
    int32_t np_[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) np_[f] = nbp[f * F + k];
    T next = 0;
    for (int z = L - 1; z >= 0; --z) {
        const int64_t q = (int64_t)z * F + k;
        T acc = 0;
#pragma unroll
        for (int f = 0; f < 3; ++f)
            if (np_[f] >= 0 && !((lm >> f) & 1u)) acc += lat[f * N + q] * y[(int64_t)z * F + np_[f]];
        if (z < L - 1) acc += vert[N + q] * next;
        next = y[q] - dinv[q] * acc;
        y[q] = next;
        zout[(int64_t)z * F + i] = next;
    }
}

}  // namespace

void build_color_order(ColorOrder& co, int64_t F, const int32_t* nb, hipStream_t s) {
    std::vector<int32_t> color(F, -1);
    int nc = 0;
    for (int64_t i = 0; i < F; ++i) {
        unsigned used = 0;
        for (int f = 0; f < 3; ++f) {
            const int32_t n = nb[3 * i + f];
            if (n >= 0 && color[n] >= 0) used |= 1u << color[n];
        }
        int c = 0;
        while (used & (1u << c)) ++c;
        color[i] = c;
        nc = std::max(nc, c + 1);
    }
    co.ncolor = nc;
    co.off.assign(nc + 1, 0);
    for (int64_t i = 0; i < F; ++i) co.off[color[i] + 1]++;
    for (int c = 0; c < nc; ++c) co.off[c + 1] += co.off[c];
    std::vector<int32_t> faces(F), pos(F);
    std::vector<int64_t> fill(co.off.begin(), co.off.end() - 1);
    for (int64_t i = 0; i < F; ++i) {
        const int64_t k = fill[color[i]]++;
        faces[k] = (int32_t)i;
        pos[i] = (int32_t)k;
    }
    std::vector<int32_t> nbp(3 * F, -1);
    std::vector<uint8_t> rev(3 * F, 0), low(F, 0);
    for (int64_t k = 0; k < F; ++k) {
        const int64_t i = faces[k];
        for (int f = 0; f < 3; ++f) {
            const int32_t n = nb[3 * i + f];
            if (n < 0) continue;
            nbp[f * F + k] = pos[n];
            for (int g = 0; g < 3; ++g)
                if (nb[3 * (int64_t)n + g] == (int32_t)i) rev[f * F + k] = (uint8_t)g;
            if (color[n] < color[i]) low[k] |= (uint8_t)(1u << f);
        }
    }
    co.faces.alloc(F);
    co.nbp.alloc(3 * F);
    co.rev.alloc(3 * F);
    co.lowmask.alloc(F);
    PB_HIP(hipMemcpyAsync(co.faces.p, faces.data(), F * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(co.nbp.p, nbp.data(), 3 * F * sizeof(int32_t), hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(co.rev.p, rev.data(), 3 * F, hipMemcpyHostToDevice, s));
    PB_HIP(hipMemcpyAsync(co.lowmask.p, low.data(), F, hipMemcpyHostToDevice, s));
    PB_HIP(hipStreamSynchronize(s));
    co.h_color = std::move(color);
}

template <typename T>
void dilu_factor(hipStream_t s, const ColorOrder& co, int64_t F, int L, const int32_t* rp, const T* val,
                 const uint16_t* slots, DiluFactors<T>& d) {
    for (int c = 0; c < co.ncolor; ++c) {
        const int64_t k0 = co.off[c], k1 = co.off[c + 1];
        if (k1 == k0) continue;
        hipLaunchKernelGGL((k_dilu_factor<T>), dim3(grid_for(k1 - k0)), dim3(kBlock), 0, s, F, L, k0, k1, co.faces.p,
                           co.nbp.p, co.rev.p, co.lowmask.p, rp, val, slots, d.lat.p, d.vert.p, d.dinv.p);
    }
}

template <typename T>
void dilu_apply(hipStream_t s, const ColorOrder& co, int64_t F, int L, DiluFactors<T>& d, const T* r, T* z) {
    for (int c = 0; c < co.ncolor; ++c) {
        const int64_t k0 = co.off[c], k1 = co.off[c + 1];
        if (k1 == k0) continue;
        hipLaunchKernelGGL((k_dilu_fwd<T>), dim3(grid_for(k1 - k0)), dim3(kBlock), 0, s, F, L, k0, k1, co.faces.p,
                           co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, r, d.y.p);
    }
    for (int c = co.ncolor - 1; c >= 0; --c) {
        const int64_t k0 = co.off[c], k1 = co.off[c + 1];
        if (k1 == k0) continue;
        hipLaunchKernelGGL((k_dilu_bwd<T>), dim3(grid_for(k1 - k0)), dim3(kBlock), 0, s, F, L, k0, k1, co.faces.p,
                           co.nbp.p, co.lowmask.p, d.lat.p, d.vert.p, d.dinv.p, d.y.p, z);
    }
}

template void dilu_factor<double>(hipStream_t, const ColorOrder&, int64_t, int, const int32_t*, const double*,
                                  const uint16_t*, DiluFactors<double>&);
template void dilu_factor<float>(hipStream_t, const ColorOrder&, int64_t, int, const int32_t*, const float*,
                                 const uint16_t*, DiluFactors<float>&);
template void dilu_apply<double>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<double>&, const double*,
                                 double*);
template void dilu_apply<float>(hipStream_t, const ColorOrder&, int64_t, int, DiluFactors<float>&, const float*,
                                float*);

}  // namespace pb
