// Internal declarations shared by the arecon HIP translation units.
// Not part of the public C ABI (that is include/arecon.h).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

#include "arecon.h"

namespace arecon {

// ---------------------------------------------------------------------------
// Error state.  Every entry point records its outcome in a thread-local
// slot that arecon_last_error() reads; the reference's void entry points
// keep their void signatures and report through this slot only.
// ---------------------------------------------------------------------------
void set_error(int code, const std::string& msg);
void clear_error();

struct Failure {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg);

#define ARECON_HIP(expr)                                                        \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess)                                                   \
            ::arecon::fail(ARECON_ERR_HIP, std::string(#expr " failed: ") +      \
                                               hipGetErrorString(e_));          \
    } while (0)

#define ARECON_REQUIRE(cond, msg)                                               \
    do {                                                                        \
        if (!(cond)) ::arecon::fail(ARECON_ERR_ARG, (msg));                     \
    } while (0)

// Run f(), translating Failure into the thread-local error slot.
template <class F>
int guarded(F&& f) {
    clear_error();
    try {
        f();
        return ARECON_OK;
    } catch (const Failure& e) {
        set_error(e.code, e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(ARECON_ERR_INTERNAL, e.what());
        return ARECON_ERR_INTERNAL;
    }
}

// Device used by the host-pointer drop-ins: ARECON_DEVICE (default 0).
int drop_in_device();
// Library-owned stream for the host-pointer drop-ins on the current device.
hipStream_t drop_in_stream();

// RAII device buffer (stream-ordered allocation).
struct DevBuf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    DevBuf() = default;
    DevBuf(size_t bytes, hipStream_t stream);
    ~DevBuf();
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------------------
// Stage launchers (device pointers, stream-ordered, no synchronisation).
// ---------------------------------------------------------------------------
// Output views [view_begin, view_begin + view_count) (all views when
// view_count < 0); the input holds all ViewN views.
void launch_rebin(float* out, const float* in, int YL, int ZL, int ViewN,
                  float YLC, float dYA, float DeltaTheta, float PLC, float DeltaT,
                  float DeltaFai, float SO, int out_layout, hipStream_t s,
                  int view_begin = 0, int view_count = -1);

void launch_filter(void* out, int out_dtype, int out_layout, const float* in, int YL, int ZL,
                   int ViewN, float PLC, float ZLC, float dYL, float dZL, float SO,
                   hipStream_t s);
// Same kernels; the tables are built on the stream in a fixed per-device array
// (no allocation, no host copy).
void launch_filter_noalloc(void* out, int out_dtype, int out_layout, const float* in, int YL,
                           int ZL, int ViewN, float PLC, float ZLC, float dYL, float dZL,
                           float SO, hipStream_t s);

// Per-view / per-slice / per-offset tables of the backprojector, built on the
// host in FP64 and cast to FP32 (see backproject.hip).
struct BPPlan;
BPPlan* bp_plan_create(const arecon_bp_geom& g, const double* w);
void bp_plan_destroy(BPPlan* p);
int bp_plan_nz(const BPPlan* p);
void bp_plan_launch(const BPPlan* p, float* img, const void* fp, int fp_dtype, int fp_layout,
                    hipStream_t s, int block_begin = 0, int block_end = -1);
int bp_plan_nblocks(const BPPlan* p);
int bp_plan_inside(const BPPlan* p);
void bp_plan_stats(const BPPlan* p, unsigned long long* staged, unsigned long long* direct);
void bp_plan_block_views(const BPPlan* p, int block, int* pmin, int* pmax);
struct BPBatch;
BPBatch* bp_batch_create(const BPPlan* const* plans, float* const* imgs, const void* const* fps,
                         int n, int fp_dtype, int fp_layout);
void bp_batch_destroy(BPBatch* b);
void bp_batch_launch(const BPBatch* b, hipStream_t s);
void bp_launch_f64(double* img, const double* fp, const arecon_bp_geom& g,
                   const double* w, hipStream_t s);

size_t dtype_size(int dtype);

}  // namespace arecon
