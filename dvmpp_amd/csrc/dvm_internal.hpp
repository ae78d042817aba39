// Internal declarations shared by the libdvm_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dvm_hip.h"

struct dvm_ctx_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int num_cus = 256;
    std::string err;
    // Growable scratch arena: at most one arena per context, reused by every
    // call on the context's stream (calls on one stream run in order).
    void* ws = nullptr;
    size_t ws_bytes = 0;
    // Pinned host staging for small device->host reads.
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // Launch-shape cache keyed by (kernel family, size bucket).
    std::map<std::pair<int, int>, int> tune;
    // Kernel timing (dvm_prof_*): pending event pairs and totals per kernel.
    bool prof = false;
    struct ProfEntry {
        std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
        double total_ms = 0.0;
        int64_t count = 0;
    };
    std::map<std::string, ProfEntry> prof_map;
    std::vector<hipEvent_t> event_pool;
};

namespace dvmi {

constexpr double kRpi2 = 1.0 / (2.0 * 3.14159265358979323846);  // src/BaseTypes.hpp:21

struct Error {
    dvm_status st;
    std::string msg;
};

[[noreturn]] inline void fail(dvm_status s, const std::string& m) { throw Error{s, m}; }

#define DVM_HIP_CHECK(expr)                                                            \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            ::dvmi::fail(e_ == hipErrorOutOfMemory ? DVM_E_OOM : DVM_E_HIP,            \
                         std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

#define DVM_LAUNCH_CHECK() DVM_HIP_CHECK(hipGetLastError())

template <class F>
dvm_status guard(dvm_ctx_t ctx, F&& f)
{
    try {
        f();
        return DVM_OK;
    } catch (const Error& e) {
        if (ctx) ctx->err = e.msg;
        return e.st;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->err = "host out of memory";
        return DVM_E_OOM;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return DVM_E_STATE;
    }
}

inline void require(bool cond, dvm_status s, const char* msg)
{
    if (!cond) fail(s, msg);
}

// Brackets the launches in its scope with timing events when profiling is on.
class ProfScope {
public:
    ProfScope(dvm_ctx_t ctx, const char* name);
    ~ProfScope();
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;

private:
    dvm_ctx_t ctx_;
    const char* name_;
    hipEvent_t start_ = nullptr;
};

// Scratch arena of at least `bytes`; contents are not preserved on growth.
void* ws_get(dvm_ctx_t ctx, size_t bytes);
// Pinned host buffer of at least `bytes`.
void* pinned_get(dvm_ctx_t ctx, size_t bytes);

// Carves aligned sub-buffers out of one scratch allocation.
struct Carver {
    char* base;
    size_t off = 0;
    explicit Carver(void* b) : base(static_cast<char*>(b)) {}
    template <class T>
    T* take(size_t n)
    {
        off = (off + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + off);
        off += n * sizeof(T);
        return p;
    }
    static size_t need(std::initializer_list<size_t> bytes)
    {
        size_t s = 0;
        for (size_t b : bytes) s = ((s + 255) & ~size_t(255)) + b;
        return s + 256;
    }
};

inline unsigned cdiv(int64_t a, int64_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// Biot-Savart entry points shared with the sheet and simulation code.
void bs_self_launch(dvm_ctx_t ctx, int prec, int64_t n, const double* x, const double* z,
                    const double* c, const double* s, double* u, double* w);
void bs_targets_launch(dvm_ctx_t ctx, int prec, int64_t ns, const double* sx, const double* sz,
                       const double* sc, const double* ss, int64_t nt, const double* tx,
                       const double* tz, int guard, double r2_skip, double* u, double* w);

}  // namespace dvmi
