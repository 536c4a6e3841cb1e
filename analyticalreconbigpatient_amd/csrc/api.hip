// C ABI of the arecon library: the reference's extern "C" drop-ins and the native
// device-resident entry points declared in include/arecon.h.
#include "arecon_internal.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

namespace arecon {

namespace {
thread_local int t_err_code = ARECON_OK;
thread_local std::string t_err_msg;

std::mutex g_stream_mu;
std::map<int, hipStream_t> g_streams;
}  // namespace

void set_error(int code, const std::string& msg) {
    t_err_code = code;
    t_err_msg = msg;
}

void clear_error() {
    t_err_code = ARECON_OK;
    t_err_msg.clear();
}

void fail(int code, const std::string& msg) { throw Failure{code, msg}; }

size_t dtype_size(int dtype) {
    switch (dtype) {
    case ARECON_F32: return 4;
    case ARECON_F16:
    case ARECON_BF16: return 2;
    default: fail(ARECON_ERR_ARG, "unknown dtype");
    }
}

int drop_in_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        fail(ARECON_ERR_NOGPU, "no HIP device is visible");
    int dev = 0;
    if (const char* e = std::getenv("ARECON_DEVICE")) dev = std::atoi(e);
    if (dev < 0 || dev >= n) fail(ARECON_ERR_ARG, "ARECON_DEVICE is out of range");
    return dev;
}

hipStream_t drop_in_stream() {
    const int dev = drop_in_device();
    ARECON_HIP(hipSetDevice(dev));
    std::lock_guard<std::mutex> lk(g_stream_mu);
    auto it = g_streams.find(dev);
    if (it != g_streams.end()) return it->second;
    hipStream_t s = nullptr;
    ARECON_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    g_streams[dev] = s;
    return s;
}

DevBuf::DevBuf(size_t bytes, hipStream_t stream) : s(stream) {
    if (bytes == 0) bytes = 1;
    ARECON_HIP(hipMallocAsync(&p, bytes, stream));
}

DevBuf::~DevBuf() {
    if (p) (void)hipFreeAsync(p, s);
}

namespace {

template <class T>
void h2d(void* d, const T* h, size_t n, hipStream_t s) {
    ARECON_HIP(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
}

template <class T>
void d2h(T* h, const void* d, size_t n, hipStream_t s) {
    ARECON_HIP(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
}

void fdk_weights(double* w, int N_2pi, double HSCoef, int delta) {
    // src/Parallel_FDK_Helical_3DWeighting.cpp:146,156-170; the taper uses the
    // file-level pi literal of :16.
    ARECON_REQUIRE(w != nullptr && N_2pi > 0, "fdk_weights: bad arguments");
    const double pi16 = 3.14159265358979;
    const double N_pi = N_2pi / 2.0;
    const int L = (int)(2.0 * std::ceil(N_pi * HSCoef) + 1.0);
    const int Shift = (int)(N_pi - std::ceil(N_pi * HSCoef));
    ARECON_REQUIRE(Shift >= 0 && Shift + L <= N_2pi,
                   "fdk_weights: the window does not fit in w[N_2pi] (HSCoef too large)");
    for (int i = 0; i < N_2pi; ++i) w[i] = 0.0;
    for (int k = 0; k < L; ++k) {
        double v;
        if (k < delta) {
            const double c = std::cos((pi16 / 2) * (delta - k - 0.5) / delta);
            v = c * c;
        } else if (L - delta <= k) {
            const double c = std::cos((pi16 / 2) * (k - (L - delta) + 0.5) / delta);
            v = c * c;
        } else {
            v = 1.0;
        }
        w[k + Shift] = v;
    }
}

}  // namespace
}  // namespace arecon

using namespace arecon;

// An instantiated HIP graph of a captured pass (include/arecon.h).
struct arecon_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
    int device;
};

extern "C" {

int arecon_last_error(char* buf, size_t len) {
    if (buf && len > 0) {
        const size_t n = std::min(len - 1, t_err_msg.size());
        std::memcpy(buf, t_err_msg.data(), n);
        buf[n] = '\0';
    }
    return t_err_code;
}

int arecon_version(void) { return 100; }

int arecon_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int arecon_fdk_weights(double* w, int N_2pi, double HSCoef, int delta) {
    return guarded([&] { fdk_weights(w, N_2pi, HSCoef, delta); });
}

// ---------------------------------------------------------------------------
// Reference drop-ins (host pointers, synchronous).
// ---------------------------------------------------------------------------
void ParallelRebinningCBCurve_GPU(float* outputProj, const float* Proj, const int YL,
                                  const int ZL, const int ViewN, const float YLC,
                                  const float dYA, const float DeltaTheta, const float PLC,
                                  const float DeltaT, const float DeltaFai, const float SO) {
    guarded([&] {
        ARECON_REQUIRE(outputProj && Proj, "ParallelRebinningCBCurve_GPU: null pointer");
        ARECON_REQUIRE(YL > 0 && ZL > 0 && ViewN > 0,
                       "ParallelRebinningCBCurve_GPU: sizes must be positive");
        hipStream_t s = drop_in_stream();
        const size_t n = (size_t)YL * ZL * ViewN;
        DevBuf din(sizeof(float) * n, s), dout(sizeof(float) * n, s);
        h2d(din.p, Proj, n, s);
        launch_rebin(dout.as<float>(), din.as<float>(), YL, ZL, ViewN, YLC, dYA, DeltaTheta,
                     PLC, DeltaT, DeltaFai, SO, ARECON_LAYOUT_VYZ, s);
        d2h(outputProj, dout.p, n, s);
        ARECON_HIP(hipStreamSynchronize(s));
    });
}

void filtering(float* hfpwd, const float* hProj, const int YL, const int ZL, const int ViewN,
               const float PLC, const float ZLC, const float dYL, const float dZL,
               const float SO) {
    guarded([&] {
        ARECON_REQUIRE(hfpwd && hProj, "filtering: null pointer");
        ARECON_REQUIRE(YL > 0 && ZL > 0 && ViewN > 0, "filtering: sizes must be positive");
        hipStream_t s = drop_in_stream();
        const size_t n = (size_t)YL * ZL * ViewN;
        DevBuf d(sizeof(float) * n, s);
        h2d(d.p, hProj, n, s);
        launch_filter(d.p, ARECON_F32, ARECON_FP_ROWS, d.as<float>(), YL, ZL, ViewN, PLC, ZLC,
                      dYL, dZL, SO, s);
        d2h(hfpwd, d.p, n, s);
        ARECON_HIP(hipStreamSynchronize(s));
    });
}

void ParallekFDKHelical3DWeighting_GPU_float(
    float* hctimage, const float* hfp, const float* hw, const int N_2pi, const int XN,
    const int YN, const int ZN, const float XNC, const float YNC, const float ZNC,
    const float dx, const float dy, const float dz, const int YL, const int ZL, const int PN,
    const float YLC, const float ZLC, const float dYL, const float dZL, const float h,
    const float BetaS, const float DeltaFai, const float N_pi, const float HSCoef,
    const float SO, const float SD, const float k1) {
    guarded([&] {
        ARECON_REQUIRE(hctimage && hfp && hw, "ParallekFDKHelical3DWeighting_GPU_float: null pointer");
        arecon_bp_geom g{};
        g.N_2pi = N_2pi; g.XN = XN; g.YN = YN; g.ZN = ZN; g.YL = YL; g.ZL = ZL; g.PN = PN;
        g.XNC = XNC; g.YNC = YNC; g.ZNC = ZNC; g.dx = dx; g.dy = dy; g.dz = dz;
        g.YLC = YLC; g.ZLC = ZLC; g.dYL = dYL; g.dZL = dZL; g.h = h; g.BetaS = BetaS;
        g.DeltaFai = DeltaFai; g.N_pi = N_pi; g.HSCoef = HSCoef; g.SO = SO; g.SD = SD;
        g.k1 = k1;
        ARECON_REQUIRE(N_2pi > 0, "ParallekFDKHelical3DWeighting_GPU_float: N_2pi must be positive");
        std::vector<double> w(hw, hw + N_2pi);
        hipStream_t s = drop_in_stream();
        BPPlan* plan = bp_plan_create(g, w.data());
        struct PlanGuard { BPPlan* p; ~PlanGuard() { bp_plan_destroy(p); } } pg{plan};
        const size_t nfp = (size_t)YL * ZL * PN, nimg = (size_t)XN * YN * ZN;
        DevBuf dfp(sizeof(float) * nfp, s), dimg(sizeof(float) * nimg, s);
        h2d(dfp.p, hfp, nfp, s);
        bp_plan_launch(plan, dimg.as<float>(), dfp.p, ARECON_F32, ARECON_FP_ROWS, s);
        d2h(hctimage, dimg.p, nimg, s);
        ARECON_HIP(hipStreamSynchronize(s));
    });
}

void ParallekFDKHelical3DWeighting_GPU_double(
    double* hctimage, const double* hfp, const double* hw, const int N_2pi, const int XN,
    const int YN, const int ZN, const double XNC, const double YNC, const double ZNC,
    const double dx, const double dy, const double dz, const int YL, const int ZL,
    const int PN, const double YLC, const double ZLC, const double dYL, const double dZL,
    const double h, const double BetaS, const double DeltaFai, const double N_pi,
    const double HSCoef, const double SO, const double SD, const double k1) {
    guarded([&] {
        ARECON_REQUIRE(hctimage && hfp && hw, "ParallekFDKHelical3DWeighting_GPU_double: null pointer");
        arecon_bp_geom g{};
        g.N_2pi = N_2pi; g.XN = XN; g.YN = YN; g.ZN = ZN; g.YL = YL; g.ZL = ZL; g.PN = PN;
        g.XNC = XNC; g.YNC = YNC; g.ZNC = ZNC; g.dx = dx; g.dy = dy; g.dz = dz;
        g.YLC = YLC; g.ZLC = ZLC; g.dYL = dYL; g.dZL = dZL; g.h = h; g.BetaS = BetaS;
        g.DeltaFai = DeltaFai; g.N_pi = N_pi; g.HSCoef = HSCoef; g.SO = SO; g.SD = SD;
        g.k1 = k1;
        hipStream_t s = drop_in_stream();
        const size_t nfp = (size_t)YL * ZL * PN, nimg = (size_t)XN * YN * ZN;
        DevBuf dfp(sizeof(double) * nfp, s), dimg(sizeof(double) * nimg, s),
            dw(sizeof(double) * N_2pi, s);
        h2d(dfp.p, hfp, nfp, s);
        h2d(dw.p, hw, (size_t)N_2pi, s);
        bp_launch_f64(dimg.as<double>(), dfp.as<double>(), g, dw.as<double>(), s);
        d2h(hctimage, dimg.p, nimg, s);
        ARECON_HIP(hipStreamSynchronize(s));
    });
}

void ParallelFDKHelical3DWeighting(double* ctimage, const double* fp, const double SO,
                                   const double DO, const int YL, const int ZL,
                                   const double DecWidth, const double DecHeigh,
                                   const double YLC, const double ZLC, const double h,
                                   const double BetaS, const int PN, const int N_2pi,
                                   const double ObjR, const int XN, const int YN, const int ZN,
                                   const double XNC, const double YNC, const double ZNC,
                                   const int delta, const double HSCoef, const double k1,
                                   const int useGPU) {
    guarded([&] {
        ARECON_REQUIRE(useGPU == 1,
                       "ParallelFDKHelical3DWeighting: only useGPU == 1 is provided by this "
                       "library (it has no CPU implementation)");
        ARECON_REQUIRE(ctimage && fp, "ParallelFDKHelical3DWeighting: null pointer");
        ARECON_REQUIRE(XN > 0 && YN > 0 && ZN > 0 && YL > 0 && ZL > 0 && PN > 0 && N_2pi > 0,
                       "ParallelFDKHelical3DWeighting: sizes must be positive");
        // Derived quantities, src/Parallel_FDK_Helical_3DWeighting.cpp:145-154.
        const double PI = 3.141592653589793;
        const double N_pi = N_2pi / 2.0;
        std::vector<double> w(N_2pi);
        fdk_weights(w.data(), N_2pi, HSCoef, delta);
        // The useGPU == 1 branch (:173-190) hands FP32 copies of every geometry
        // value and of fp to the FP32 kernel and widens the image back to double.
        arecon_bp_geom g{};
        g.N_2pi = N_2pi; g.XN = XN; g.YN = YN; g.ZN = ZN; g.YL = YL; g.ZL = ZL; g.PN = PN;
        g.XNC = (float)XNC; g.YNC = (float)YNC; g.ZNC = (float)ZNC;
        g.dx = (float)(2.0 * ObjR / XN); g.dy = (float)(2.0 * ObjR / YN);
        g.dz = (float)(2.0 * ObjR / ZN);
        g.YLC = (float)YLC; g.ZLC = (float)ZLC;
        g.dYL = (float)(DecWidth / YL); g.dZL = (float)(DecHeigh / ZL);
        g.h = (float)h; g.BetaS = (float)BetaS; g.DeltaFai = (float)(2.0 * PI / N_2pi);
        g.N_pi = (float)N_pi; g.HSCoef = (float)HSCoef; g.SO = (float)SO;
        g.SD = (float)(SO + DO); g.k1 = (float)k1;
        std::vector<double> wf(N_2pi);
        for (int i = 0; i < N_2pi; ++i) wf[i] = (float)w[i];
        hipStream_t s = drop_in_stream();
        BPPlan* plan = bp_plan_create(g, wf.data());
        struct PlanGuard { BPPlan* p; ~PlanGuard() { bp_plan_destroy(p); } } pg{plan};
        const size_t nfp = (size_t)YL * ZL * PN, nimg = (size_t)XN * YN * ZN;
        std::vector<float> ffp(fp, fp + nfp);
        std::vector<float> fimg(nimg);
        DevBuf dfp(sizeof(float) * nfp, s), dimg(sizeof(float) * nimg, s);
        h2d(dfp.p, ffp.data(), nfp, s);
        bp_plan_launch(plan, dimg.as<float>(), dfp.p, ARECON_F32, ARECON_FP_ROWS, s);
        d2h(fimg.data(), dimg.p, nimg, s);
        ARECON_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < nimg; ++i) ctimage[i] = fimg[i];
    });
}

// ---------------------------------------------------------------------------
// Native device-resident API.
// ---------------------------------------------------------------------------
int arecon_rebin_dev(float* d_out, const float* d_in, int YL, int ZL, int ViewN, float YLC,
                     float dYA, float DeltaTheta, float PLC, float DeltaT, float DeltaFai,
                     float SO, int out_layout, void* stream) {
    return guarded([&] {
        launch_rebin(d_out, d_in, YL, ZL, ViewN, YLC, dYA, DeltaTheta, PLC, DeltaT, DeltaFai,
                     SO, out_layout, static_cast<hipStream_t>(stream));
    });
}

int arecon_rebin_views_dev(float* d_out, const float* d_in, int YL, int ZL, int ViewN,
                           float YLC, float dYA, float DeltaTheta, float PLC, float DeltaT,
                           float DeltaFai, float SO, int view_begin, int view_count,
                           void* stream) {
    return guarded([&] {
        launch_rebin(d_out, d_in, YL, ZL, ViewN, YLC, dYA, DeltaTheta, PLC, DeltaT, DeltaFai,
                     SO, ARECON_LAYOUT_VZY, static_cast<hipStream_t>(stream), view_begin,
                     view_count);
    });
}

int arecon_filter_dev(void* d_out, int out_dtype, int out_layout, const float* d_in, int YL,
                      int ZL, int ViewN, float PLC, float ZLC, float dYL, float dZL, float SO,
                      void* stream) {
    return guarded([&] {
        launch_filter(d_out, out_dtype, out_layout, d_in, YL, ZL, ViewN, PLC, ZLC, dYL, dZL, SO,
                      static_cast<hipStream_t>(stream));
    });
}

int arecon_filter_dev_noalloc(void* d_out, int out_dtype, int out_layout, const float* d_in,
                              int YL, int ZL, int ViewN, float PLC, float ZLC, float dYL,
                              float dZL, float SO, void* stream) {
    return guarded([&] {
        launch_filter_noalloc(d_out, out_dtype, out_layout, d_in, YL, ZL, ViewN, PLC, ZLC, dYL,
                              dZL, SO, static_cast<hipStream_t>(stream));
    });
}

int arecon_bp_plan_create(const arecon_bp_geom* g, const double* w, arecon_bp_plan** out) {
    return guarded([&] {
        ARECON_REQUIRE(g && w && out, "arecon_bp_plan_create: null pointer");
        *out = reinterpret_cast<arecon_bp_plan*>(bp_plan_create(*g, w));
    });
}

int arecon_bp_plan_destroy(arecon_bp_plan* plan) {
    return guarded([&] { bp_plan_destroy(reinterpret_cast<BPPlan*>(plan)); });
}

int arecon_bp_plan_nz(const arecon_bp_plan* plan) {
    return bp_plan_nz(reinterpret_cast<const BPPlan*>(plan));
}

int arecon_bp_plan_execute(const arecon_bp_plan* plan, float* d_img, const void* d_fp,
                           int fp_dtype, int fp_layout, void* stream) {
    return guarded([&] {
        bp_plan_launch(reinterpret_cast<const BPPlan*>(plan), d_img, d_fp, fp_dtype, fp_layout,
                       static_cast<hipStream_t>(stream));
    });
}

int arecon_bp_batch_create(const arecon_bp_plan* const* plans, float* const* d_imgs,
                           const void* const* d_fps, int n, int fp_dtype, int fp_layout,
                           arecon_bp_batch** out) {
    return guarded([&] {
        ARECON_REQUIRE(out != nullptr, "arecon_bp_batch_create: null output");
        *out = reinterpret_cast<arecon_bp_batch*>(
            bp_batch_create(reinterpret_cast<const BPPlan* const*>(plans), d_imgs, d_fps, n,
                            fp_dtype, fp_layout));
    });
}

int arecon_bp_batch_destroy(arecon_bp_batch* batch) {
    return guarded([&] { bp_batch_destroy(reinterpret_cast<BPBatch*>(batch)); });
}

int arecon_bp_batch_execute(const arecon_bp_batch* batch, void* stream) {
    return guarded([&] {
        bp_batch_launch(reinterpret_cast<const BPBatch*>(batch), static_cast<hipStream_t>(stream));
    });
}

int arecon_bp_plan_nblocks(const arecon_bp_plan* plan) {
    return bp_plan_nblocks(reinterpret_cast<const BPPlan*>(plan));
}

int arecon_bp_plan_block_views(const arecon_bp_plan* plan, int block, int* pmin, int* pmax) {
    return guarded([&] {
        bp_plan_block_views(reinterpret_cast<const BPPlan*>(plan), block, pmin, pmax);
    });
}

int arecon_bp_plan_inside(const arecon_bp_plan* plan) {
    return bp_plan_inside(reinterpret_cast<const BPPlan*>(plan));
}

int arecon_bp_plan_stats(const arecon_bp_plan* plan, unsigned long long* staged,
                         unsigned long long* direct) {
    return guarded([&] { bp_plan_stats(reinterpret_cast<const BPPlan*>(plan), staged, direct); });
}

int arecon_bp_plan_execute_blocks(const arecon_bp_plan* plan, float* d_img, const void* d_fp,
                                  int fp_dtype, int fp_layout, int block_begin, int block_end,
                                  void* stream) {
    return guarded([&] {
        bp_plan_launch(reinterpret_cast<const BPPlan*>(plan), d_img, d_fp, fp_dtype, fp_layout,
                       static_cast<hipStream_t>(stream), block_begin, block_end);
    });
}

int arecon_graph_capture_begin(void* stream) {
    return guarded([&] {
        ARECON_REQUIRE(stream != nullptr, "arecon_graph_capture_begin: the null stream cannot "
                                          "be captured");
        ARECON_HIP(hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                         hipStreamCaptureModeThreadLocal));
    });
}

int arecon_graph_capture_end(void* stream, arecon_graph** out) {
    return guarded([&] {
        ARECON_REQUIRE(stream != nullptr && out != nullptr,
                       "arecon_graph_capture_end: null argument");
        *out = nullptr;
        hipGraph_t graph = nullptr;
        // end the capture first, whatever follows: a stream left capturing
        // would fail every later call on it
        ARECON_HIP(hipStreamEndCapture(static_cast<hipStream_t>(stream), &graph));
        ARECON_REQUIRE(graph != nullptr, "arecon_graph_capture_end: nothing was captured");
        auto* g = new arecon_graph{graph, nullptr, 0};
        const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            delete g;
            ARECON_HIP(e);
        }
        (void)hipGetDevice(&g->device);
        *out = g;
    });
}

int arecon_graph_launch(const arecon_graph* graph, void* stream) {
    return guarded([&] {
        ARECON_REQUIRE(graph != nullptr && graph->exec != nullptr,
                       "arecon_graph_launch: null graph");
        ARECON_HIP(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)));
    });
}

int arecon_graph_destroy(arecon_graph* graph) {
    return guarded([&] {
        if (!graph) return;
        if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
        if (graph->graph) (void)hipGraphDestroy(graph->graph);
        delete graph;
    });
}

int arecon_bp_dev_f64(double* d_img, const double* d_fp, const arecon_bp_geom* g,
                      const double* w, void* stream) {
    return guarded([&] {
        ARECON_REQUIRE(g && w, "arecon_bp_dev_f64: null pointer");
        hipStream_t s = static_cast<hipStream_t>(stream);
        DevBuf dw(sizeof(double) * g->N_2pi, s);
        h2d(dw.p, w, (size_t)g->N_2pi, s);
        bp_launch_f64(d_img, d_fp, *g, dw.as<double>(), s);
        // w is the caller's host buffer: return only once it has been consumed.
        ARECON_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
