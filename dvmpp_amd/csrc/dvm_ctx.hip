// Context, errors, memory helpers and scratch arena of libdvm_hip.so.
#include "dvm_internal.hpp"

#include <cstring>

namespace dvmi {

void* ws_get(dvm_ctx_t ctx, size_t bytes)
{
    if (ctx->ws_fixed) {
        if (bytes > ctx->ws_fixed_bytes)
            fail(DVM_E_SIZE, "this call needs " + std::to_string(bytes) +
                                 " bytes of scratch, more than the " +
                                 std::to_string(ctx->ws_fixed_bytes) +
                                 "-byte array of the allocation-free entry point");
        return ctx->ws_fixed;
    }
    if (bytes <= ctx->ws_bytes && ctx->ws) return ctx->ws;
    // Earlier work on this stream may still read the old arena.
    DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->ws) DVM_HIP_CHECK(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
    const size_t grow = bytes + bytes / 4 + (1u << 20);
    DVM_HIP_CHECK(hipMalloc(&ctx->ws, grow));
    ctx->ws_bytes = grow;
    return ctx->ws;
}

void* pinned_get(dvm_ctx_t ctx, size_t bytes)
{
    if (bytes <= ctx->pinned_bytes && ctx->pinned) return ctx->pinned;
    DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->pinned) DVM_HIP_CHECK(hipHostFree(ctx->pinned));
    ctx->pinned = nullptr;
    ctx->pinned_bytes = 0;
    const size_t grow = bytes < 4096 ? 4096 : bytes;
    DVM_HIP_CHECK(hipHostMalloc(&ctx->pinned, grow, hipHostMallocDefault));
    ctx->pinned_bytes = grow;
    return ctx->pinned;
}

static hipEvent_t event_take(dvm_ctx_t ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    DVM_HIP_CHECK(hipEventCreate(&e));
    return e;
}

ProfScope::ProfScope(dvm_ctx_t ctx, const char* name) : ctx_(ctx), name_(name)
{
    if (!ctx_ || !ctx_->prof) return;  // no context: work on another stream, not timed
    start_ = event_take(ctx_);
    DVM_HIP_CHECK(hipEventRecord(start_, ctx_->stream));
}

ProfScope::~ProfScope()
{
    if (!start_) return;
    // A destructor must not throw: on failure the sample is dropped.
    hipEvent_t stop = nullptr;
    if (!ctx_->event_pool.empty()) {
        stop = ctx_->event_pool.back();
        ctx_->event_pool.pop_back();
    } else if (hipEventCreate(&stop) != hipSuccess) {
        ctx_->event_pool.push_back(start_);
        return;
    }
    if (hipEventRecord(stop, ctx_->stream) != hipSuccess) {
        ctx_->event_pool.push_back(start_);
        ctx_->event_pool.push_back(stop);
        return;
    }
    ctx_->prof_map[name_].pending.emplace_back(start_, stop);
}

static void prof_drain(dvm_ctx_t ctx)
{
    for (auto& kv : ctx->prof_map) {
        for (auto& ev : kv.second.pending) {
            DVM_HIP_CHECK(hipEventSynchronize(ev.second));
            float ms = 0.0f;
            DVM_HIP_CHECK(hipEventElapsedTime(&ms, ev.first, ev.second));
            kv.second.total_ms += ms;
            kv.second.count += 1;
            ctx->event_pool.push_back(ev.first);
            ctx->event_pool.push_back(ev.second);
        }
        kv.second.pending.clear();
    }
}

}  // namespace dvmi

using namespace dvmi;

extern "C" {

const char* dvm_version(void) { return "dvm_hip 0.1 gfx950"; }

const char* dvm_status_string(dvm_status s)
{
    switch (s) {
    case DVM_OK: return "ok";
    case DVM_E_INVALID: return "invalid argument";
    case DVM_E_SIZE: return "size mismatch";
    case DVM_E_NAN: return "non-finite release counts";
    case DVM_E_HIP: return "HIP runtime error";
    case DVM_E_OOM: return "out of memory";
    case DVM_E_STATE: return "invalid state";
    }
    return "unknown status";
}

dvm_status dvm_ctx_create(int device, void* stream, dvm_ctx_t* out)
{
    if (!out) return DVM_E_INVALID;
    *out = nullptr;
    dvm_ctx_t ctx = new (std::nothrow) dvm_ctx_s();
    if (!ctx) return DVM_E_OOM;
    dvm_status st = guard(ctx, [&] {
        int count = 0;
        DVM_HIP_CHECK(hipGetDeviceCount(&count));
        require(device >= 0 && device < count, DVM_E_INVALID, "device index out of range");
        DVM_HIP_CHECK(hipSetDevice(device));
        ctx->device = device;
        hipDeviceProp_t prop;
        DVM_HIP_CHECK(hipGetDeviceProperties(&prop, device));
        ctx->num_cus = prop.multiProcessorCount;
        const char* tune = std::getenv("DVM_TUNE");
        if (tune && std::string(tune) == "measure") ctx->tuning = DVM_TUNE_MEASURE;
        if (stream) {
            ctx->stream = static_cast<hipStream_t>(stream);
            ctx->own_stream = false;
        } else {
            DVM_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            ctx->own_stream = true;
        }
    });
    if (st != DVM_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return DVM_OK;
}

dvm_status dvm_ctx_destroy(dvm_ctx_t ctx)
{
    if (!ctx) return DVM_OK;
    dvm_status st = guard(ctx, [&] {
        DVM_HIP_CHECK(hipSetDevice(ctx->device));
        DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        prof_drain(ctx);
        for (hipEvent_t e : ctx->event_pool) DVM_HIP_CHECK(hipEventDestroy(e));
        ctx->event_pool.clear();
        if (ctx->ws) DVM_HIP_CHECK(hipFree(ctx->ws));
        if (ctx->pinned) DVM_HIP_CHECK(hipHostFree(ctx->pinned));
        if (ctx->own_stream) DVM_HIP_CHECK(hipStreamDestroy(ctx->stream));
    });
    delete ctx;
    return st;
}

dvm_status dvm_ctx_sync(dvm_ctx_t ctx)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] { DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream)); });
}

void* dvm_ctx_stream(dvm_ctx_t ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

dvm_status dvm_prof_enable(dvm_ctx_t ctx, int on)
{
    if (!ctx) return DVM_E_INVALID;
    ctx->prof = on != 0;
    return DVM_OK;
}

dvm_status dvm_prof_reset(dvm_ctx_t ctx)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        prof_drain(ctx);
        ctx->prof_map.clear();
    });
}

dvm_status dvm_prof_read(dvm_ctx_t ctx, const char* kernel, double* total_ms, int64_t* launches)
{
    if (!ctx || !kernel) return DVM_E_INVALID;
    return guard(ctx, [&] {
        prof_drain(ctx);
        auto it = ctx->prof_map.find(kernel);
        if (total_ms) *total_ms = it == ctx->prof_map.end() ? 0.0 : it->second.total_ms;
        if (launches) *launches = it == ctx->prof_map.end() ? 0 : it->second.count;
    });
}

const char* dvm_last_error(dvm_ctx_t ctx) { return ctx ? ctx->err.c_str() : "null context"; }

dvm_status dvm_malloc(dvm_ctx_t ctx, size_t bytes, void** dptr)
{
    if (!ctx || !dptr) return DVM_E_INVALID;
    return guard(ctx, [&] {
        *dptr = nullptr;
        if (bytes) DVM_HIP_CHECK(hipMalloc(dptr, bytes));
    });
}

dvm_status dvm_free(dvm_ctx_t ctx, void* dptr)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        if (dptr) {
            DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
            DVM_HIP_CHECK(hipFree(dptr));
        }
    });
}

dvm_status dvm_memcpy_h2d(dvm_ctx_t ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        if (bytes) {
            DVM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
            DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
    });
}

dvm_status dvm_memcpy_d2h(dvm_ctx_t ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        if (bytes) {
            DVM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
            DVM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
    });
}

dvm_status dvm_memcpy_d2d(dvm_ctx_t ctx, void* dst, const void* src, size_t bytes)
{
    if (!ctx) return DVM_E_INVALID;
    return guard(ctx, [&] {
        if (bytes) DVM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    });
}


struct dvm_event_s {
    hipEvent_t ev = nullptr;
};

dvm_status dvm_event_create(dvm_ctx_t ctx, dvm_event_t* out)
{
    if (!ctx || !out) return DVM_E_INVALID;
    *out = nullptr;
    dvm_event_t e = new (std::nothrow) dvm_event_s();
    if (!e) return DVM_E_OOM;
    dvm_status st = guard(ctx, [&] {
        DVM_HIP_CHECK(hipSetDevice(ctx->device));
        DVM_HIP_CHECK(hipEventCreateWithFlags(&e->ev, hipEventDisableTiming));
    });
    if (st != DVM_OK) {
        delete e;
        return st;
    }
    *out = e;
    return DVM_OK;
}

dvm_status dvm_event_record(dvm_ctx_t ctx, dvm_event_t e)
{
    if (!ctx || !e) return DVM_E_INVALID;
    return guard(ctx, [&] { DVM_HIP_CHECK(hipEventRecord(e->ev, ctx->stream)); });
}

dvm_status dvm_event_wait(dvm_event_t e)
{
    if (!e) return DVM_E_INVALID;
    return hipEventSynchronize(e->ev) == hipSuccess ? DVM_OK : DVM_E_HIP;
}

dvm_status dvm_event_query(dvm_event_t e, int* done)
{
    if (!e || !done) return DVM_E_INVALID;
    const hipError_t r = hipEventQuery(e->ev);
    if (r == hipSuccess) {
        *done = 1;
        return DVM_OK;
    }
    if (r == hipErrorNotReady) {
        *done = 0;
        return DVM_OK;
    }
    return DVM_E_HIP;
}

dvm_status dvm_event_destroy(dvm_event_t e)
{
    if (!e) return DVM_OK;
    const hipError_t r = e->ev ? hipEventDestroy(e->ev) : hipSuccess;
    delete e;
    return r == hipSuccess ? DVM_OK : DVM_E_HIP;
}

dvm_status dvm_host_alloc(size_t bytes, void** ptr)
{
    if (!ptr) return DVM_E_INVALID;
    *ptr = nullptr;
    if (bytes == 0) return DVM_OK;
    const hipError_t r = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (r == hipErrorOutOfMemory) return DVM_E_OOM;
    return r == hipSuccess ? DVM_OK : DVM_E_HIP;
}

dvm_status dvm_host_free(void* ptr)
{
    if (!ptr) return DVM_OK;
    return hipHostFree(ptr) == hipSuccess ? DVM_OK : DVM_E_HIP;
}

}  // extern "C"
