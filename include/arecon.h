/*
 * arecon.h -- C ABI of the MI355X (gfx950) helical cone-beam FDK library.
 *
 * Two layers:
 *
 *  1. Drop-ins for the reference's extern "C" host functions.  Same names
 *     (including the reference's "Parallek" spelling), same argument lists,
 *     same memory layouts, host pointers, synchronous on return.  They run on
 *     the device named by the ARECON_DEVICE environment variable (default 0)
 *     on a library-owned stream, and report failures through
 *     arecon_last_error() (the reference's functions return void and check
 *     nothing; these never call exit()).
 *
 *  2. Native device-resident entry points (arecon_*): device pointers, an
 *     explicit hipStream_t passed as void*, int status (0 = ok).  These are
 *     what a device-resident pipeline calls; nothing is copied to the host.
 *
 * Layouts (all row-major, last index fastest):
 *   fan / rebinned sinogram, reference MEX layout   [ViewN][YL][ZL]
 *   filter input / output, BP input                 [ViewN][ZL][YL]
 *   reconstructed volume                            [YN][XN][ZN]
 */
#ifndef ARECON_H_
#define ARECON_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define ARECON_OK 0
#define ARECON_ERR_ARG 1      /* invalid argument or unsupported shape */
#define ARECON_ERR_HIP 2      /* a HIP runtime call failed */
#define ARECON_ERR_NOGPU 3    /* no usable gfx950 device */
#define ARECON_ERR_INTERNAL 4

/* ---- element types for filtered projections ------------------------------- */
#define ARECON_F32 0
#define ARECON_F16 1
#define ARECON_BF16 2

/* ---- rebinning output layouts --------------------------------------------- */
#define ARECON_LAYOUT_VYZ 0   /* [ViewN][YL][ZL], the reference's rebin layout */
#define ARECON_LAYOUT_VZY 1   /* [ViewN][ZL][YL], the filter's input layout */

/* Layouts of the filtered sinogram (filter output, backprojection input). */
#define ARECON_FP_ROWS 0      /* [ViewN][ZL][YL], the reference's layout */
#define ARECON_FP_ROWPAIRS 1  /* [ViewN][ZL-1][YL][2]: element (v, r, c) holds rows r
                                 and r+1 of channel c, so one 4-sample load fetches
                                 the 2x2 bilinear neighbourhood of the backprojector */
#define ARECON_FP_QUADS 2     /* [ViewN][ZL-1][YL-1][4]: element (v, r, c) holds
                                 f(r,c), f(r+1,c)-f(r,c), f(r,c+1)-f(r,c) and
                                 f(r+1,c+1)-f(r+1,c)-f(r,c+1)+f(r,c): the bilinear
                                 sample at (r+fr, c+fc) is then three FMAs */
#define ARECON_FP_ROWDIFF 3   /* [ViewN][ZL-1][YL][2]: element (v, r, c) holds
                                 f(r,c) and f(r+1,c)-f(r,c), the row pair with its
                                 difference precomputed: the bilinear sample is
                                 four instructions instead of six (in FP32
                                 bit-identical to the row-pair blend); in FP16 /
                                 BF16 half the bytes per channel of the quads */

/* ========================================================================== */
/* 1. Reference drop-ins                                                      */
/* ========================================================================== */

/* Fan-to-parallel rebinning on the curved detector.
 * Replaces src/ParallelRebinningCBCurve.h:4-8 (definition
 * src/ParallelRebinningCBCurveGPU.cu:106-115).  YLC and PLC are 0-based, as
 * the MEX gateway passes them (src/ParallelRebinningCBCurve.cpp:121-122).
 * Layout [ViewN][YL][ZL] in and out.  Edge clamps follow the reference's CPU
 * path (src/ParallelRebinningCBCurve.cpp:55-85). */
void ParallelRebinningCBCurve_GPU(float* outputProj, const float* Proj, const int YL,
                                  const int ZL, const int ViewN, const float YLC,
                                  const float dYA, const float DeltaTheta,
                                  const float PLC, const float DeltaT,
                                  const float DeltaFai, const float SO);

/* Cone/row preweighting followed by the Shepp-Logan-type ramp filter
 * h(n) = -2/(pi^2 (4n^2-1))/dYL as an exact linear convolution.
 * Replaces src/filtering.h:12-18 (definition src/filtering.cu:270-287).
 * PLC and ZLC are 0-based (src/analyticalReconBigPatient.m:139).
 * Layout [ViewN][ZL][YL] in and out. */
void filtering(float* hfpwd, const float* hProj, const int YL, const int ZL,
               const int ViewN, const float PLC, const float ZLC, const float dYL,
               const float dZL, const float SO);

/* 3D-weighted helical FDK backprojection, FP32.
 * Replaces src/ParallelFDKHelical3DWeightingGPU.h:13-22 (definition
 * src/ParallelFDKHelical3DWeightingGPU.cu:144-159).  The view loop covers the
 * voxel's window [s1, s2] as the reference's CPU path does
 * (src/Parallel_FDK_Helical_3DWeighting.cpp:71).  hw holds N_2pi weights. */
void ParallekFDKHelical3DWeighting_GPU_float(
    float* hctimage, const float* hfp, const float* hw, const int N_2pi, const int XN,
    const int YN, const int ZN, const float XNC, const float YNC, const float ZNC,
    const float dx, const float dy, const float dz, const int YL, const int ZL,
    const int PN, const float YLC, const float ZLC, const float dYL, const float dZL,
    const float h, const float BetaS, const float DeltaFai, const float N_pi,
    const float HSCoef, const float SO, const float SD, const float k1);

/* Same in FP64 throughout.
 * Replaces src/ParallelFDKHelical3DWeightingGPU.h:4-12 (definition
 * src/ParallelFDKHelical3DWeightingGPU.cu:125-140). */
void ParallekFDKHelical3DWeighting_GPU_double(
    double* hctimage, const double* hfp, const double* hw, const int N_2pi,
    const int XN, const int YN, const int ZN, const double XNC, const double YNC,
    const double ZNC, const double dx, const double dy, const double dz, const int YL,
    const int ZL, const int PN, const double YLC, const double ZLC, const double dYL,
    const double dZL, const double h, const double BetaS, const double DeltaFai,
    const double N_pi, const double HSCoef, const double SO, const double SD,
    const double k1);

/* Mid-level entry: derives dx, dy, dz (= 2 ObjR / N), dYL, dZL, DeltaFai and
 * SD, builds the weight table w[N_2pi], and backprojects.
 * Replaces src/Parallel_FDK_Helical_3DWeighting.cpp:117-143.  useGPU == 1
 * runs the FP32 GPU path exactly as the reference's useGPU == 1 branch does
 * (FP32 copy of fp, result widened to double).  Any other useGPU value is an
 * error: this library is the GPU path and has no CPU implementation. */
void ParallelFDKHelical3DWeighting(double* ctimage, const double* fp, const double SO,
                                   const double DO, const int YL, const int ZL,
                                   const double DecWidth, const double DecHeigh,
                                   const double YLC, const double ZLC, const double h,
                                   const double BetaS, const int PN, const int N_2pi,
                                   const double ObjR, const int XN, const int YN,
                                   const int ZN, const double XNC, const double YNC,
                                   const double ZNC, const int delta,
                                   const double HSCoef, const double k1,
                                   const int useGPU);

/* ========================================================================== */
/* 2. Native API                                                              */
/* ========================================================================== */

/* Last error of the calling thread: returns its code (0 = none) and copies a
 * NUL-terminated message into buf (if buf != NULL and len > 0). */
int arecon_last_error(char* buf, size_t len);

/* Library version as 10000*major + 100*minor + patch. */
int arecon_version(void);

/* Number of HIP devices visible (0 if none), without creating contexts. */
int arecon_device_count(void);

/* Rebinning on device buffers.  out_layout selects ARECON_LAYOUT_VYZ (the
 * reference layout) or ARECON_LAYOUT_VZY (written transposed, ready for the
 * filter; removes the host permutes of src/Parallel_Rebinning_CB_Curve.m:20,23
 * and src/analyticalReconBigPatient.m:138).  in and out must not alias. */
int arecon_rebin_dev(float* d_out, const float* d_in, int YL, int ZL, int ViewN,
                     float YLC, float dYA, float DeltaTheta, float PLC, float DeltaT,
                     float DeltaFai, float SO, int out_layout, void* stream);

/* Rebinning of the output views [view_begin, view_begin + view_count) only,
 * written in ARECON_LAYOUT_VZY into the full-size d_out (d_in holds all ViewN
 * views).  Lets a caller filter and backproject early views while later ones
 * are still being rebinned. */
int arecon_rebin_views_dev(float* d_out, const float* d_in, int YL, int ZL, int ViewN,
                           float YLC, float dYA, float DeltaTheta, float PLC, float DeltaT,
                           float DeltaFai, float SO, int view_begin, int view_count,
                           void* stream);

/* Preweight + ramp filter on device buffers; input layout [ViewN][ZL][YL],
 * output layout out_layout (ARECON_FP_ROWS, ARECON_FP_ROWPAIRS, ARECON_FP_QUADS
 * or ARECON_FP_ROWDIFF).  d_out may equal d_in only for
 * ARECON_F32 rows.  For ARECON_FP_QUADS and ARECON_FP_ROWDIFF with YL <= 1024
 * the output is formed from the filtered rows in LDS and d_in is left unchanged; above that the filter runs in place in
 * d_in (which then holds the FP32 filtered rows) and the output is derived from
 * it, so callers must treat d_in as scratch for these two layouts. */
int arecon_filter_dev(void* d_out, int out_dtype, int out_layout, const float* d_in, int YL,
                      int ZL, int ViewN, float PLC, float ZLC, float dYL, float dZL, float SO,
                      void* stream);

/* Allocation-free variant of arecon_filter_dev: same arguments, layouts,
 * kernels and results.  No counterpart in the reference; it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Instead of the per-device table
 * caches arecon_filter_dev fills on first use, its twiddle, spectrum and
 * preweight tables are built by a kernel on `stream`, in FP64 rounded to FP32,
 * in a fixed 1 MiB device array that each device holds until the process ends.
 * Limit: ZL * YL <= 249,856 (and YL <= 2048 as for arecon_filter_dev); larger
 * shapes return ARECON_ERR_ARG.  Calls no hipMalloc/hipFree and does not
 * synchronise.  Every call on a device rewrites that array: calls that overlap
 * on one device (on different streams or threads) race; calls on one stream
 * are ordered. */
int arecon_filter_dev_noalloc(void* d_out, int out_dtype, int out_layout, const float* d_in,
                              int YL, int ZL, int ViewN, float PLC, float ZLC, float dYL,
                              float dZL, float SO, void* stream);

/* Backprojection geometry: the derived quantities the reference passes to its
 * kernel (src/ParallelFDKHelical3DWeightingGPU.h:13-22), kept in FP64. */
typedef struct arecon_bp_geom {
    int N_2pi, XN, YN, ZN, YL, ZL, PN;
    double XNC, YNC, ZNC;       /* 0-based volume centre */
    double dx, dy, dz;
    double YLC, ZLC;            /* 0-based detector centre */
    double dYL, dZL;
    double h;                   /* helical pitch (table travel per 2 pi) */
    double BetaS, DeltaFai;     /* first view angle, view step */
    double N_pi, HSCoef;        /* N_2pi / 2, half-scan coefficient */
    double SO, SD, k1;
    int nz;                     /* slices per thread: 0 = automatic; variants 0/1:
                                   1/2/4/8/16; variant 2: 8/16/32 */
    int order;                  /* workgroup order: 0 = slice-major, 1 = column-major,
                                   2 = slice-major with the tiles of a slice block
                                   column by column in bands of 6 tile rows */
    int variant;                /* 0 = view-table kernel (default), 1 = per-voxel weight,
                                   2 = z-column kernel (slice blocks with view masks),
                                   3 = z-column kernel reading each view's detector
                                   footprint from LDS (nz 8 or 16; FP32 row pairs or
                                   16-bit quads; ZL <= 250) */
    int group;                  /* z-column kernel: 1/2/4 = slices per branch-free
                                   group; 8 = each slice behind its own window test,
                                   all gathers of a view issued before any blend
                                   (nz <= 16); 0 = automatic */
} arecon_bp_geom;

typedef struct arecon_bp_plan arecon_bp_plan;

/* Builds the per-view, per-slice and per-offset tables (FP64 on the host,
 * stored FP32 on the device of the current HIP context) for geometry g and
 * weight table w[N_2pi].  *out receives the plan. */
int arecon_bp_plan_create(const arecon_bp_geom* g, const double* w,
                          arecon_bp_plan** out);
int arecon_bp_plan_destroy(arecon_bp_plan* plan);

/* Slices per thread the plan chose (1 .. 32); 0 for a null plan. */
int arecon_bp_plan_nz(const arecon_bp_plan* plan);

/* Backprojects d_fp (PN views in layout fp_layout, element type fp_dtype) into
 * d_img ([YN][XN][ZN], FP32; every voxel is written).  The row-pair, quad and
 * row-difference layouts need a z-column kernel (variant 2, or the LDS-staged
 * variant 3, which reads FP32 row pairs, row differences of any element type
 * or 16-bit quads).  Variant 3 builds its footprint table (device memory owned
 * by the plan, about 4 bytes per view of each slice block and 16x16-column
 * tile) at the first execute with each record layout, and synchronizes the
 * stream once then: run one execute before capturing it in a graph. */
int arecon_bp_plan_execute(const arecon_bp_plan* plan, float* d_img, const void* d_fp,
                           int fp_dtype, int fp_layout, void* stream);

/* The plan's slices come in blocks of arecon_bp_plan_nz consecutive slices.
 * arecon_bp_plan_nblocks: their number; arecon_bp_plan_block_views: the views
 * [*pmin, *pmax] block b reads (pmin > pmax: none); execute_blocks: writes only
 * the slices of blocks [block_begin, block_end) (slice-major order only), so a
 * caller can start early blocks once their views are filtered. */
int arecon_bp_plan_nblocks(const arecon_bp_plan* plan);
int arecon_bp_plan_block_views(const arecon_bp_plan* plan, int block, int* pmin, int* pmax);
int arecon_bp_plan_execute_blocks(const arecon_bp_plan* plan, float* d_img, const void* d_fp,
                                  int fp_dtype, int fp_layout, int block_begin, int block_end,
                                  void* stream);

/* 1 when every column of the plan's grid lies well inside the source circle
 * and projects strictly inside the detector, so the z-column kernels (variants
 * 2 and 3) run without per-column validity tests; 0 otherwise or for a null
 * plan.  No counterpart in the reference (a diagnostic of this library). */
int arecon_bp_plan_inside(const arecon_bp_plan* plan);

/* Diagnostics of the LDS-staged kernel (variant 3; no counterpart in the
 * reference): workgroups that read their views from LDS (*staged) and those
 * that fell back to global-memory gathers because a view's footprint did not
 * fit (*direct), summed over every execute of the plan so far.  Synchronises
 * the device.  Both are 0 for other variants. */
int arecon_bp_plan_stats(const arecon_bp_plan* plan, unsigned long long* staged,
                         unsigned long long* direct);

/* Batched backprojection (no counterpart in the reference, whose MEX
 * backprojects one volume per call): n plans -- z-slabs of one or more scans,
 * all z-column plans of one variant (2, gathering, or 3, LDS-staged, nz 8 or
 * 16) with the same nz and group -- bound to their volumes d_imgs[i] and
 * filtered sinograms d_fps[i] (all of fp_dtype and fp_layout).  For variant 3
 * create builds each plan's footprint table for the layout (and synchronizes
 * the device).  One execute is one kernel launch over every slab, with no
 * allocation or host synchronisation, so it can be captured in a HIP graph.
 * The buffers must outlive the batch. */
typedef struct arecon_bp_batch arecon_bp_batch;
int arecon_bp_batch_create(const arecon_bp_plan* const* plans, float* const* d_imgs,
                           const void* const* d_fps, int n, int fp_dtype, int fp_layout,
                           arecon_bp_batch** out);
int arecon_bp_batch_destroy(arecon_bp_batch* batch);
int arecon_bp_batch_execute(const arecon_bp_batch* batch, void* stream);

/* HIP-graph capture of a device-resident pass (no counterpart in the
 * reference, whose functions are synchronous; for callers that replay the same
 * pass many times, such as the batched reconstruction of several scans).
 * arecon_graph_capture_begin starts capturing `stream` (not the null stream);
 * the pass's calls on that stream (arecon_rebin_dev, arecon_filter_dev,
 * arecon_bp_plan_execute, arecon_bp_batch_execute, ...) are then recorded, not
 * run; arecon_graph_capture_end stops the capture and instantiates the graph
 * into *out (on failure the capture is ended and *out stays null).
 * arecon_graph_launch replays it on any stream of the same device.  Run the
 * pass once before capturing it: first calls build and cache tables (device
 * allocations, a stream synchronization), which a capture cannot hold and
 * which then fail with an error.  The graph reads and writes the buffers the
 * captured calls named; they must outlive it. */
typedef struct arecon_graph arecon_graph;
int arecon_graph_capture_begin(void* stream);
int arecon_graph_capture_end(void* stream, arecon_graph** out);
int arecon_graph_launch(const arecon_graph* graph, void* stream);
int arecon_graph_destroy(arecon_graph* graph);

/* FP64 backprojection on device buffers (same math in double throughout).
 * Synchronises `stream` before returning (w is a host buffer). */
int arecon_bp_dev_f64(double* d_img, const double* d_fp, const arecon_bp_geom* g,
                      const double* w, void* stream);

/* Synthetic data (no counterpart in the reference, which ships no data):
 * exact line integrals of nE ellipsoids (rows of 8 floats: cx, cy, cz, a, b,
 * c, phi [rad, rotation about z], mu [1/mm]) along the rays of a helical
 * curved-detector scan, written as [ViewN][YL][ZL] FP32 (the rebinning input
 * layout, channels already in the flipped order of
 * src/analyticalReconBigPatient.m:30).  View i has source angle
 * BetaS + i*dBeta and source height h*angle/(2 pi); channel c has fan angle
 * (c - YLC0)*dYA; row k has height (k - ZLC0)*dZL at distance SD.
 * Synchronises `stream` before returning. */
int arecon_project_ellipsoids_dev(float* d_out, const float* ellipsoids, int nE, int YL, int ZL,
                                  int ViewN, double BetaS, double dBeta, double SO, double SD,
                                  double h, double YLC0, double dYA, double ZLC0, double dZL,
                                  void* stream);

/* Bandwidth probes of the backprojector's roofs (diagnostics, no counterpart in
 * the reference).  arecon_probe_lds: `blocks` workgroups of 256 lanes, each
 * lane reading bytes_per_lane (8 or 16) bytes of LDS `iters` times (a multiple
 * of 4) at conflict-free addresses.  arecon_probe_l2: each lane reads 16 bytes
 * `iters` times at scattered 16-byte offsets of d_src (span_bytes long; choose
 * it small enough to stay in L2).  d_sink holds blocks * 256 floats.  Both
 * only launch (time them with events on `stream`) and report the bytes they
 * read in *bytes_read. */
int arecon_probe_lds(float* d_sink, int bytes_per_lane, int iters, int blocks,
                     double* bytes_read, void* stream);
int arecon_probe_l2(float* d_sink, const void* d_src, long long span_bytes, int iters,
                    int blocks, double* bytes_read, void* stream);
/* Streaming copy of `bytes` (a multiple of 16) from d_src to d_dst: the
 * reference kernel for calibrating HBM byte counters (reads and writes `bytes`). */
int arecon_probe_copy(void* d_dst, const void* d_src, long long bytes, void* stream);
/* Issue rate of one VALU instruction form (0 FMA, 1 FMA with a scalar operand,
 * 2 mixed-precision FMA, 3 mixed FMA with an FP16 result, 4 packed FP16 dot
 * product, 5 packed FP16 FMA, 6 floor conversion, 7 shift-add, 8 reciprocal):
 * `blocks` workgroups of 256 lanes; *instructions receives the wavefront
 * instructions issued (time the launch with events on `stream`).  d_sink holds
 * blocks * 256 floats. */
int arecon_probe_valu(float* d_sink, int form, int blocks, double* instructions, void* stream);

/* Builds the reference's weight table w[N_2pi]
 * (src/Parallel_FDK_Helical_3DWeighting.cpp:156-170) on the host. */
int arecon_fdk_weights(double* w, int N_2pi, double HSCoef, int delta);

#ifdef __cplusplus
}
#endif

#endif /* ARECON_H_ */
