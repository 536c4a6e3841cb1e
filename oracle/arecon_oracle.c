/*
 * arecon_oracle.c -- CPU restatement of the reference's helical FDK path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the timed
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library (libarecon.so) never links,
 * loads or calls it, and has no CPU fallback.
 *
 * Every function restates the reference's CPU code (or, for the filter, the
 * MATLAB CPU branch) line for line, in the reference's own arithmetic types:
 *
 *   oracle_rebin_f32 / _f64    src/ParallelRebinningCBCurve.cpp:22-97
 *   oracle_preweight_f64       src/filtering.cu:202-205,
 *                              src/analyticalReconBigPatient.m:109-115
 *   oracle_filter_nn           src/filtering.cu:65, analyticalReconBigPatient.m:121
 *   oracle_filter_kernel       src/analyticalReconBigPatient.m:119-125
 *                              (== src/filtering.cu:16-25,62-79)
 *   oracle_filter_f64          src/analyticalReconBigPatient.m:107-134
 *   oracle_fdk_weights         src/Parallel_FDK_Helical_3DWeighting.cpp:156-170
 *   oracle_fdk_core            src/Parallel_FDK_Helical_3DWeighting.cpp:42-113
 *   oracle_fdk                 src/Parallel_FDK_Helical_3DWeighting.cpp:117-200
 *                              (useGPU == 0 branch)
 *
 * How each function is pinned:
 *   reference  bit for bit against the reference's own CPU functions,
 *              compiled from its source files by build_ref.sh into
 *              _ref/libarecon_ref.so: ParallelRebinningCBCurve_CPU<float/double>
 *              and ParallelFDKHelical3DWeightingCPU, on seeded inputs at the
 *              368x16 and 888x64 geometries, with off-central-ray voxels,
 *              clipped windows and edge clamps.      tests/test_oracle_ref.py
 *              (the filter's reference is MATLAB and cuFFT, which cannot run
 *              here; the weight table lives inside the MEX-side function.)
 *   rebin      closed forms (constant in -> constant out; inputs affine in
 *              (fan view, channel) reproduced at the textbook fan->parallel
 *              coordinates beta = theta + asin(t/SO), gamma = asin(t/SO));
 *              row independence; edge clamps.            tests/test_oracle_rebin.py
 *   preweight  t = 0 reduces to the cone cosine SO/sqrt(SO^2+b^2); b = 0 -> 1.
 *   filter     numpy FFT restatement of the MATLAB lines; direct O(YL^2)
 *              convolution; the telescoping closed form of a constant row;
 *              the Shepp-Logan frequency response |sin(w/2)|/(pi*dYL).
 *                                                        tests/test_oracle_filter.py
 *   weights    cos^2 taper closed form, mirror symmetry, support.
 *   fdk        zero/linearity; windows outside [0,PN) give exactly 0; a
 *              constant projection gives exactly pi*c (conjugate-pair partition
 *              of unity; exact for any k1 when h = 1 mm because of the +-0.5 mm
 *              conjugate offset); the central-ray (Yr = 0) detector row equals
 *              the textbook magnification SD/(SO+UU).   tests/test_oracle_fdk.py
 *              The off-central-ray (Yr != 0) factor SO/sqrt(SO^2-Yr^2) in the
 *              detector row coordinate (Parallel_FDK_Helical_3DWeighting.cpp:80)
 *              has no closed form; the reference comparison pins it.
 */
#include <math.h>
#include <stdlib.h>

/* Planted-mistake builds for the tests that show each pin can fail
 * (tests/test_oracle_mutations.py compiles this file with
 * -DORACLE_MUTATION=n into a temporary library).  0 = the restatement. */
#ifndef ORACLE_MUTATION
#define ORACLE_MUTATION 0
#endif
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Rebinning: src/ParallelRebinningCBCurve.cpp:22-97.                       */
/* The reference instantiates the template with T = float and calls the    */
/* double-precision libm (only <cmath> is included, calls are unqualified), */
/* so every transcendental is evaluated in double and rounded to T.  The    */
/* float version reproduces that: each statement rounds to float exactly    */
/* where the reference assigns to a T (built with -ffp-contract=off, SSE    */
/* arithmetic, so float expressions are evaluated in float).                */
/* ------------------------------------------------------------------------ */
void oracle_rebin_f32(float* out, const float* in, int YL, int ZL, int ViewN,
                      float YLC, float dYA, float DeltaTheta, float PLC,
                      float DeltaT, float DeltaFai, float SO)
{
    for (int i = 0; i != ViewN; ++i) {
        float Theta = (float)i * DeltaTheta;           /* :34 */
        for (int j = 0; j != YL; ++j) {
            float t = ((float)j - PLC) * DeltaT;          /* :37 */
            float q = t / SO;
            float Beta = (float)asin((double)q);          /* :38 */
            float Fai = (ORACLE_MUTATION == 1) ? Theta - Beta : Theta + Beta; /* :39 */
            float r2 = SO * SO - t * t;
            float a = (float)atan((double)t / sqrt((double)r2)); /* :40 */
            float FaiIndex = Fai / DeltaFai;              /* :41 */
            float ad = a / dYA;
            float UIndex = ad + YLC;                      /* :42 */
            int FI = (int)ceil((double)FaiIndex);                  /* :43 */
            int UI = (int)ceil((double)UIndex);                    /* :44 */
            float coeXB = (float)FI - FaiIndex;           /* :45 */
            float coeXU = (float)(1.0 - (double)coeXB);   /* :46 */
            float coeYB = (float)UI - UIndex;             /* :47 */
            float coeYU = (float)(1.0 - (double)coeYB);   /* :48 */
            int IXU, IXB, IYU, IYB;
            if (FI <= 0) { IXU = 0; IXB = 0; }                     /* :55-59 */
            else if ((ORACLE_MUTATION == 2) ? FI >= ViewN - 1 : FI > ViewN - 1) {
                IXU = ViewN - 1; IXB = ViewN - 1;                  /* :60-64 */
            }
            else { IXU = FI; IXB = (int)((double)FI - 1.0); }      /* :65-69 */
            if (UI <= 0) { IYU = 0; IYB = 0; }                     /* :71-75 */
            else if (UI > YL - 1) { IYU = YL - 1; IYB = YL - 1; }  /* :76-80 */
            else { IYU = UI; IYB = UI - 1; }                       /* :81-85 */
            const float c00 = coeXB * coeYB, c10 = coeXU * coeYB;
            const float c01 = coeXB * coeYU, c11 = coeXU * coeYU;
            const float* pBB = in + ((size_t)IXB * YL + IYB) * ZL;
            const float* pUB = in + ((size_t)IXU * YL + IYB) * ZL;
            const float* pBU = in + ((size_t)IXB * YL + IYU) * ZL;
            const float* pUU = in + ((size_t)IXU * YL + IYU) * ZL;
            float* o = out + ((size_t)i * YL + j) * ZL;
            for (int k = 0; k != ZL; ++k) {                        /* :87-94 */
                o[k] = c00 * pBB[k] + c10 * pUB[k] + c01 * pBU[k] + c11 * pUU[k];
            }
        }
    }
}

void oracle_rebin_f64(double* out, const double* in, int YL, int ZL, int ViewN,
                      double YLC, double dYA, double DeltaTheta, double PLC,
                      double DeltaT, double DeltaFai, double SO)
{
    for (int i = 0; i != ViewN; ++i) {
        const double Theta = i * DeltaTheta;
        for (int j = 0; j != YL; ++j) {
            const double t = (j - PLC) * DeltaT;
            const double Beta = asin(t / SO);
            const double Fai = Theta + Beta;
            const double a = atan(t / sqrt(SO * SO - t * t));
            const double FaiIndex = Fai / DeltaFai;
            const double UIndex = a / dYA + YLC;
            const int FI = (int)ceil(FaiIndex);
            const int UI = (int)ceil(UIndex);
            const double coeXB = FI - FaiIndex, coeXU = 1.0 - coeXB;
            const double coeYB = UI - UIndex, coeYU = 1.0 - coeYB;
            int IXU, IXB, IYU, IYB;
            if (FI <= 0) { IXU = 0; IXB = 0; }
            else if (FI > ViewN - 1) { IXU = ViewN - 1; IXB = ViewN - 1; }
            else { IXU = FI; IXB = FI - 1; }
            if (UI <= 0) { IYU = 0; IYB = 0; }
            else if (UI > YL - 1) { IYU = YL - 1; IYB = YL - 1; }
            else { IYU = UI; IYB = UI - 1; }
            for (int k = 0; k != ZL; ++k) {
                out[((size_t)i * YL + j) * ZL + k] =
                    coeXB * coeYB * in[((size_t)IXB * YL + IYB) * ZL + k] +
                    coeXU * coeYB * in[((size_t)IXU * YL + IYB) * ZL + k] +
                    coeXB * coeYU * in[((size_t)IXB * YL + IYU) * ZL + k] +
                    coeXU * coeYU * in[((size_t)IXU * YL + IYU) * ZL + k];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Preweighting + ramp filter.                                              */
/* The reference has no C++ CPU filter; its CPU twin is the MATLAB branch   */
/* src/analyticalReconBigPatient.m:107-134, computed in double.  Layout is  */
/* the MEX layout of src/reWeigAdFiltr_GPU.cpp:47-50: [(v*ZL + k)*YL + j],  */
/* YL fastest; PLC and ZLC are 0-based (the caller passes PLC-1, ZLC-1,     */
/* analyticalReconBigPatient.m:139).                                        */
/* ------------------------------------------------------------------------ */
void oracle_preweight_f64(double* p, int YL, int ZL, int ViewN, double PLC,
                          double ZLC, double dYL, double dZL, double SO)
{
    for (int v = 0; v < ViewN; ++v)
        for (int k = 0; k < ZL; ++k) {
            const double b = (k - ZLC) * dZL;                      /* m:112 */
            double* row = p + ((size_t)v * ZL + k) * YL;
            for (int j = 0; j < YL; ++j) {
                const double t = (j - PLC) * dYL;                  /* m:110 */
                /* m:113  Proj*SO*SO/sqrt(SO^4+(SO*b)^2-(b*t)^2) */
                row[j] = row[j] * SO * SO /
                         sqrt(SO * SO * SO * SO + (SO * b) * (SO * b) - (b * t) * (b * t));
            }
        }
}

/* NN = 2^ceil(log2(3*YL)) -- filtering.cu:65, analyticalReconBigPatient.m:121 */
int oracle_filter_nn(int YL)
{
    return (int)pow(2.0, ceil(log(YL * 3.0) / log(2.0)));
}

/* In-place iterative radix-2 complex FFT (sign -1 forward, +1 inverse,     */
/* unnormalised).  Arrays hold interleaved (re, im).                        */
static void fft_c2c(double* a, int n, int sign)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = a[2 * i], ti = a[2 * i + 1];
            a[2 * i] = a[2 * j]; a[2 * i + 1] = a[2 * j + 1];
            a[2 * j] = tr; a[2 * j + 1] = ti;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = sign * 2.0 * M_PI / len;
        const int half = len >> 1;
        for (int k = 0; k < half; ++k) {
            const double wr = cos(ang * k), wi = sin(ang * k);
            for (int i = k; i < n; i += len) {
                const int j = i + half;
                const double xr = a[2 * j] * wr - a[2 * j + 1] * wi;
                const double xi = a[2 * j] * wi + a[2 * j + 1] * wr;
                a[2 * j] = a[2 * i] - xr; a[2 * j + 1] = a[2 * i + 1] - xi;
                a[2 * i] += xr; a[2 * i + 1] += xi;
            }
        }
    }
}

/* HS(1:YL+1) = hsl(YL+1:2YL+1); HS(end-YL+1:end) = hsl(1:YL) with          */
/* hsl = -2./((pi^2)*(4n^2-1))/dYL, n = -YL:YL   (m:119-124).               */
void oracle_filter_kernel(double* HS, int YL, double dYL, int NN)
{
    memset(HS, 0, sizeof(double) * (size_t)NN);
    for (int n = 0; n <= YL; ++n)
        HS[n] = -2.0 / ((M_PI * M_PI) * (4.0 * n * n - 1.0)) / dYL;
    for (int n = -YL; n <= -1; ++n)
        HS[NN + n - (ORACLE_MUTATION == 3)] = -2.0 / ((M_PI * M_PI) * (4.0 * n * n - 1.0)) / dYL;
}

/* Preweight then filter every (view, row) -- m:107-134.  in and out may    */
/* alias.  nthreads > 1 splits views over OpenMP threads; rows are          */
/* independent, so the result is bit-identical to the serial run.           */
void oracle_filter_f64(double* out, const double* in, int YL, int ZL, int ViewN,
                       double PLC, double ZLC, double dYL, double dZL, double SO,
                       int nthreads)
{
    const int NN = oracle_filter_nn(YL);
    double* FtS = (double*)malloc(sizeof(double) * 2 * (size_t)NN);
    double* HS = (double*)malloc(sizeof(double) * (size_t)NN);
    oracle_filter_kernel(HS, YL, dYL, NN);
    for (int i = 0; i < NN; ++i) { FtS[2 * i] = HS[i]; FtS[2 * i + 1] = 0.0; }
    fft_c2c(FtS, NN, -1);                                          /* m:125 */
    free(HS);
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
    {
        double* buf = (double*)malloc(sizeof(double) * 2 * (size_t)NN);
#pragma omp for schedule(static)
        for (int v = 0; v < ViewN; ++v)
            for (int k = 0; k < ZL; ++k) {
                const size_t off = ((size_t)v * ZL + k) * YL;
                const double b = (k - ZLC) * dZL;
                for (int j = 0; j < YL; ++j) {
                    const double t = (j - PLC) * dYL;
                    buf[2 * j] = in[off + j] * SO * SO /
                        sqrt(SO * SO * SO * SO + (SO * b) * (SO * b) - (b * t) * (b * t));
                    buf[2 * j + 1] = 0.0;
                }
                for (int j = YL; j < NN; ++j) { buf[2 * j] = 0.0; buf[2 * j + 1] = 0.0; }
                fft_c2c(buf, NN, -1);                              /* m:129 */
                for (int f = 0; f < NN; ++f) {                     /* m:130 FtS.*FtP */
                    const double xr = buf[2 * f], xi = buf[2 * f + 1];
                    buf[2 * f] = FtS[2 * f] * xr - FtS[2 * f + 1] * xi;
                    buf[2 * f + 1] = FtS[2 * f] * xi + FtS[2 * f + 1] * xr;
                }
                fft_c2c(buf, NN, +1);                              /* m:130 ifft */
                for (int j = 0; j < YL; ++j)                       /* m:131 real(tep(1:YL)) */
                    out[off + j] = buf[2 * j] / NN;
            }
        free(buf);
    }
    free(FtS);
}

/* ------------------------------------------------------------------------ */
/* 3D-weighted helical FDK: src/Parallel_FDK_Helical_3DWeighting.cpp.       */
/* ------------------------------------------------------------------------ */
/* File-level pi of the reference, :16 (truncated literal).                 */
static const double ref_pi = 3.14159265358979;

/* Weight table w[N_2pi], :156-170.  Uses the file-level truncated pi.      */
void oracle_fdk_weights(double* w, int N_2pi, double HSCoef, int delta)
{
    const double N_pi = N_2pi / 2.0;                               /* :146 */
    const int L = (int)(2.0 * ceil(N_pi * HSCoef) + 1.0);          /* :157 */
    const int Shift = (int)(N_pi - ceil(N_pi * HSCoef));           /* :158 */
    for (int i = 0; i < N_2pi; ++i) w[i] = 0.0;
    for (int k = 0; k < L; k++) {
        if (0 <= k && k < delta)
            w[k + Shift] = pow(cos((ref_pi / 2) * (delta - k - 0.5) / delta), 2);
        else if (L - delta <= k && k < L)
            w[k + Shift] = pow(cos((ref_pi / 2) * (k - (L - delta) + 0.5) / delta), 2);
        else
            w[k + Shift] = 1;
    }
}

/* Voxel-driven backprojection, :42-113.  Loops y -> x -> z as the          */
/* reference does; nthreads > 1 splits the outer y loop (voxels are         */
/* independent, so the result is bit-identical to the serial run).          */
/* zi_begin/zi_end restrict the z range that is computed (the rest of       */
/* ctimage is left untouched); pass 0, ZN for the full volume.              */
void oracle_fdk_core(double* ctimage, const double* fp, const double* w, int N_2pi,
                     int XN, int YN, int ZN, double XNC, double YNC, double ZNC,
                     double dx, double dy, double dz, int YL, int ZL, int PN,
                     double YLC, double ZLC, double dYL, double dZL, double h,
                     double BetaS, double DeltaFai, double N_pi, double HSCoef,
                     double SO, double SD, double k1, int zi_begin, int zi_end,
                     int nthreads)
{
    (void)N_2pi;
    const double invPI = 1.0 / (2.0 * 3.141592653589793);        /* :51 */
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (int yi = 0; yi < YN; yi++) {
        const double y = (yi - YNC) * dy;                          /* :54 */
        for (int xi = 0; xi < XN; xi++) {
            const double x = (xi - XNC) * dx;                      /* :57 */
            for (int zi = zi_begin; zi < zi_end; ++zi) {
                const double z = (zi - ZNC) * dz;                  /* :61 */
                const double Beta0 = 2.0 * ref_pi * z / h;         /* :62 */
                const int s0 = (int)ceil((Beta0 - BetaS) / DeltaFai - 0.5); /* :63 */
                int s1 = (int)(s0 - ceil(N_pi * HSCoef));          /* :64 */
                int s2 = (int)(s0 + ceil(N_pi * HSCoef));          /* :65 */
                double res = 0;
                if ((s1 < PN) && (s2 > 0)) {                       /* :67 */
                    s1 = (s1 < 0) ? 0 : s1;
                    s2 = (s2 > PN - 1) ? PN - 1 : s2;
                    for (int ProjInd = s1; ProjInd <= s2; ++ProjInd) {
                        const double View = BetaS + ProjInd * DeltaFai;       /* :73 */
                        const int d1 = (int)(N_pi - (s0 - ProjInd));          /* :74 */
                        const int d2 = (ORACLE_MUTATION == 6)
                            ? d1
                            : (int)((ProjInd < s0) ? (d1 + N_pi) : (d1 - N_pi)); /* :75 */
                        const double UU = -x * cos(View) - y * sin(View);     /* :76 */
                        const double Yr = -x * sin(View) + y * cos(View);     /* :77 */
                        const double temp1 = sqrt(SO * SO - Yr * Yr);         /* :78 */
                        const double temp2 = (z - h * View * invPI);          /* :79 */
                        const double Zr = (ORACLE_MUTATION == 9)
                            ? temp2 * SD / (SO + UU)
                            : temp2 * (SO * SD) / (UU * temp1 + SO * SO - Yr * Yr); /* :80 */
                        const double U1 = ((ORACLE_MUTATION == 8) ? -Yr : Yr) / dYL + YLC; /* :81 */
                        const int U = (int)ceil(U1);
                        const double V1 = Zr / dZL + ZLC;                     /* :83 */
                        const int V = (int)ceil(V1);
                        const double Dey = U - U1;
                        const double Dez = V - V1;
                        if ((U > 0) && (U < YL) && (V > 0) && (V < ZL)) {     /* :88 */
                            const double touying =
                                Dey * Dez * fp[((size_t)ProjInd * ZL + (V - 1)) * YL + (U - 1)] +
                                Dey * (1.0 - Dez) * fp[((size_t)ProjInd * ZL + (V)) * YL + (U - 1)] +
                                (1.0 - Dey) * Dez * fp[((size_t)ProjInd * ZL + (V - 1)) * YL + U] +
                                (1.0 - Dey) * (1.0 - Dez) * fp[((size_t)ProjInd * ZL + V) * YL + U];
                            const double weight1 = w[d1];                     /* :95 */
                            const double weight2 = w[d2];
                            const double sgn = (ORACLE_MUTATION == 5) ? -1.0 : 1.0;
                            const double den1 = (ORACLE_MUTATION == 7) ? temp1 - UU : temp1 + UU;
                            const double den2 = (ORACLE_MUTATION == 7) ? temp1 + UU : temp1 - UU;
                            const double Gama = fabs(temp2 / den1);           /* :98 */
                            const double Gama_C = (ProjInd < s0)
                                ? fabs((temp2 - sgn * 0.5) / den2)
                                : fabs((temp2 + sgn * 0.5) / den2);           /* :99 */
                            const double m1 = pow(Gama, k1);                  /* :100 */
                            const double m2 = pow(Gama_C, k1);
                            const double weight = (weight1 * m2) / (weight2 * m1 + weight1 * m2); /* :102 */
                            res += weight * touying * DeltaFai;               /* :104 */
                        }
                    }
                }
                ctimage[((size_t)yi * XN + xi) * ZN + zi] = res;     /* :109 */
            }
        }
    }
}

/* Mid-level entry, :117-200 (the useGPU == 0 branch).                      */
void oracle_fdk(double* ctimage, const double* fp, double SO, double DO, int YL,
                int ZL, double DecWidth, double DecHeigh, double YLC, double ZLC,
                double h, double BetaS, int PN, int N_2pi, double ObjR, int XN,
                int YN, int ZN, double XNC, double YNC, double ZNC, int delta,
                double HSCoef, double k1, int nthreads)
{
    const double PI = 3.141592653589793;                           /* :145 */
    const double N_pi = N_2pi / 2.0;
    const double dx = 2.0 * ObjR / XN;
    const double dy = 2.0 * ObjR / YN;
    const double dz = 2.0 * ObjR / ZN;
    const double dYL = DecWidth / YL;
    const double dZL = DecHeigh / ZL;
    const double DeltaFai = 2.0 * PI / N_2pi;
    const double SD = SO + DO;
    double* w = (double*)malloc(sizeof(double) * (size_t)N_2pi);
    oracle_fdk_weights(w, N_2pi, HSCoef, delta);
    oracle_fdk_core(ctimage, fp, w, N_2pi, XN, YN, ZN, XNC, YNC, ZNC, dx, dy, dz,
                    YL, ZL, PN, YLC, ZLC, dYL, dZL, h, BetaS, DeltaFai, N_pi,
                    HSCoef, SO, SD, k1, 0, ZN, nthreads);
    free(w);
}
