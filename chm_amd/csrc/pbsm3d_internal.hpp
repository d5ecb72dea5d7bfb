// Internal types of the MI355X PBSM3D library: device buffers, the context,
// and the launch helpers shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pbsm3d.h"

namespace pb {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);

#define PB_HIP(x)                                                   \
    do {                                                            \
        hipError_t pb_e_ = (x);                                     \
        if (pb_e_ != hipSuccess) ::pb::throw_hip(pb_e_, #x, __FILE__, __LINE__); \
    } while (0)

// Owning device allocation (hipMalloc / hipFree).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(PBSM3D_ENOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Pinned host staging buffer (hipHostMalloc / hipHostFree).
template <typename T>
struct HostBuf {
    T* p = nullptr;
    size_t n = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
    void alloc(size_t count) {
        if (count <= n && p) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        PB_HIP(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
};

// CSR pattern (row_ptr, col sorted per row) plus FP64 values and vectors.
struct System {
    int64_t n = 0, nnz = 0;
    DevBuf<int32_t> rp, col;
    DevBuf<double> val, rhs, x;
    std::vector<int32_t> h_rp, h_col;  // host copy of the pattern (export, checks)
};

// Thomas factors of the per-face vertical tridiagonal blocks (FP64 and FP32).
struct ColumnFactors {
    DevBuf<double> lo, cp, im;  // sub-diagonal, c', 1/m  (each N, layer-major)
    DevBuf<float> lo32, cp32, im32;
};

// HIP-event bracketing of kernel launches on one stream (off by default).
struct Profiler {
    bool on = false;
    uint32_t mask = ~0u;  // bit k: class k is bracketed (each pair of events costs a few us)
    int every = 1;        // bracket only every every-th launch of each class
    int64_t seen[PBSM3D_K_COUNT] = {};
    std::vector<hipEvent_t> pool;
    std::vector<int> ids;
    size_t used = 0;
    int64_t count[PBSM3D_K_COUNT] = {};
    double ms[PBSM3D_K_COUNT] = {};
    ~Profiler() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
    bool begin(hipStream_t s, int id) {
        if (!on || !((mask >> id) & 1u)) return false;
        if (seen[id]++ % every != 0) return false;
        while (pool.size() < used + 2) {
            hipEvent_t e;
            PB_HIP(hipEventCreate(&e));
            pool.push_back(e);
        }
        PB_HIP(hipEventRecord(pool[used], s));
        ids.push_back(id);
        return true;
    }
    void end(hipStream_t s) {
        if (!on) return;
        PB_HIP(hipEventRecord(pool[used + 1], s));
        used += 2;
    }
    // call after the stream has been synchronised
    void collect() {
        const size_t pairs = ids.size() < used / 2 ? ids.size() : used / 2;  // an aborted step leaves one open
        for (size_t k = 0; k < pairs; ++k) {
            float t = 0;
            if (hipEventElapsedTime(&t, pool[2 * k], pool[2 * k + 1]) == hipSuccess) {
                ms[ids[k]] += t;
                count[ids[k]] += 1;
            }
        }
        ids.clear();
        used = 0;
    }
};

struct ProfScope {
    Profiler* p;
    hipStream_t s;
    bool active = false;
    ProfScope(Profiler* p_, hipStream_t s_, int id) : p(p_), s(s_) {
        if (p) active = p->begin(s, id);
    }
    ~ProfScope() noexcept(false) {
        if (active) p->end(s);
    }
};

constexpr int kBlock = 256;
constexpr int kMaxRestart = 64;

inline unsigned grid_for(int64_t n, int block = kBlock, int64_t cap = 1 << 20) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// number of blocks used by grid-stride reductions
constexpr int kReduceBlocks = 2048;

}  // namespace pb
