// Gaussian and Gaussian-copula (GCMI) mutual-information kernels for gfx950.
//
// Both estimators need, for every atom pair, a 6x6 block of a second-moment
// matrix C = V^T V / N of a [N frames][3n components] matrix V:
//   Gaussian  V = mean-centred coordinates (lib/correlation_core.cpp:140-166)
//   GCMI      V = van der Waerden copula scores (lib/gcmi.cpp:95-146, 199-211)
// C is one dense GEMM (3n x N x 3n).  It runs on the matrix cores as a
// split-precision FP16 product with FP32 accumulation (after Ootomo & Yokota
// 2022): v*s = hi + lo*2^-11 with hi = fp16(v*s) and lo = fp16((v*s - hi)*2^11),
// C*s^2 ~= hi.hi + (hi.lo + lo.hi)*2^-11, the hi.hi and cross terms kept in
// separate FP32 accumulators.  The per-pair log-determinants, the entropies
// and the copula's inverse normal stay in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include "dense_internal.h"
#include "sort_network.h"

namespace gmxcorr {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));

constexpr double kPi = 3.14159265358979323846;

// Determinant by LU with partial pivoting, the row operations of
// lib/correlation_core.cpp:49-98 (identical in lib/gcmi.cpp:152-184).
template <int n>
__device__ double lu_det(double* m)
{
    double sign = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        double best = fabs(m[col * n + col]);
        for (int r = col + 1; r < n; ++r) {
            const double v = fabs(m[r * n + col]);
            if (v > best) { best = v; piv = r; }
        }
        if (best == 0.0) return 0.0;
        if (piv != col) {
            for (int j = 0; j < n; ++j) {
                const double t = m[col * n + j];
                m[col * n + j] = m[piv * n + j];
                m[piv * n + j] = t;
            }
            sign = -sign;
        }
        const double diag = m[col * n + col];
        for (int r = col + 1; r < n; ++r) {
            const double f = m[r * n + col] / diag;
            for (int j = col + 1; j < n; ++j) m[r * n + j] -= f * m[col * n + j];
        }
    }
    double det = sign;
    for (int i = 0; i < n; ++i) det *= m[i * n + i];
    return det;
}

// Inverse error function: Acklam's published rational approximation of the
// inverse normal CDF at p = (x+1)/2 (Acklam 2003), divided by sqrt 2, then
// one Newton step on erf.  The method lib/gcmi.cpp:26-79 uses.
__device__ double erfinv_acklam(double x)
{
    if (x >= 1.0) return 1e30;
    if (x <= -1.0) return -1e30;
    if (x == 0.0) return 0.0;
    const double a0 = -3.969683028665376e+01, a1 = 2.209460984245205e+02,
                 a2 = -2.759285104469687e+02, a3 = 1.383577518672690e+02,
                 a4 = -3.066479806614716e+01, a5 = 2.506628277459239e+00;
    const double b0 = -5.447609879822406e+01, b1 = 1.615858368580409e+02,
                 b2 = -1.556989798598866e+02, b3 = 6.680131188771972e+01,
                 b4 = -1.328068155288572e+01;
    const double c0 = -7.784894002430293e-03, c1 = -3.223964580411365e-01,
                 c2 = -2.400758277161838e+00, c3 = -2.549732539343734e+00,
                 c4 = 4.374664141464968e+00, c5 = 2.938163982698783e+00;
    const double d0 = 7.784695709041462e-03, d1 = 3.224671290700398e-01,
                 d2 = 2.445134137142996e+00, d3 = 3.754408661907416e+00;
    const double p = (x + 1.0) * 0.5;
    double r;
    if (p < 0.02425) {
        const double q = sqrt(-2.0 * log(p));
        r = (((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5) /
            ((((d0 * q + d1) * q + d2) * q + d3) * q + 1.0);
    } else if (p <= 0.97575) {
        const double q = p - 0.5, s = q * q;
        r = (((((a0 * s + a1) * s + a2) * s + a3) * s + a4) * s + a5) * q /
            (((((b0 * s + b1) * s + b2) * s + b3) * s + b4) * s + 1.0);
    } else {
        const double q = sqrt(-2.0 * log(1.0 - p));
        r = -(((((c0 * q + c1) * q + c2) * q + c3) * q + c4) * q + c5) /
            ((((d0 * q + d1) * q + d2) * q + d3) * q + 1.0);
    }
    r /= sqrt(2.0);
    r -= (erf(r) - x) * sqrt(kPi) * 0.5 * exp(r * r);
    return r;
}

}  // namespace

// ---------------------------------------------------------------------------
// Split of V (FP64, component rows [3n][N] with row stride ldv) into scaled
// FP16 hi/lo rows [3n][Kp]; frames N..Kp-1 are zero.
// ---------------------------------------------------------------------------
__global__ void dense_split_kernel(const double* __restrict__ V, int N, long long ldv, int Kp,
                                   double scale, _Float16* __restrict__ hi,
                                   _Float16* __restrict__ lo)
{
    const long long c = blockIdx.y;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < Kp; f += gridDim.x * blockDim.x) {
        _Float16 h = (_Float16)0.0f, l = (_Float16)0.0f;
        if (f < N) {
            const double v = V[c * ldv + f] * scale;
            h = (_Float16)v;
            l = (_Float16)((v - (double)h) * 2048.0);
        }
        hi[c * Kp + f] = h;
        lo[c * Kp + f] = l;
    }
}

// ---------------------------------------------------------------------------
// Split-precision second-moment GEMM on the matrix cores.
// C = (hi.hi + (hi.lo + lo.hi) 2^-11) * out_scale over the upper-triangle
// 128x128 tiles (ti <= tj), K split into nsplit ranges whose partial tiles go
// to Cp[split] (mirrored into the lower triangle off the diagonal);
// dense_cov_reduce_kernel sums them in split order.
// Block = 4 waves; wave w owns the 64x64 sub-tile (w/2, w%2) as 2x2 blocks of
// v_mfma_f32_32x32x16_f16.  Each 32-frame step of the four operand panels
// (A hi/lo, B hi/lo: 128 components x 32 frames) is staged in LDS, double
// buffered, one barrier per step; the next step's panels are loaded into
// registers while the current one is multiplied.  MFMA operands: lane l
// supplies component l%32 and the 8 consecutive frames 8*(l/32).. of each
// 16-frame step (both operands read the component-major rows, since
// B[k][j] = V[k][j] = rows[j][k]); result register q of lane l is row
// 8*(q/4) + 4*(l/32) + q%4, column l%32 of the 32x32 block.
// ---------------------------------------------------------------------------
constexpr int kGTile = 128;            // block tile (components)
constexpr int kGStep = 32;             // frames per staged step
constexpr int kGPitch = kGStep;        // LDS row pitch in halves (64 B, four 16-byte units)
constexpr int kGPanel = kGTile * kGPitch;

// LDS swizzle of the 16-byte unit q (frames 8q..8q+7 of a step) in panel row
// `row`: unit q ^ gsw(row).  The staging writes (16 lanes: four rows x four
// units) and the operand reads (16 lanes: 16 consecutive rows, one unit) then
// fall on distinct banks for 32 or 64 banks of 4 bytes; rows 32 apart share
// the swizzle.
__device__ __forceinline__ int gsw(int row) { return ((row >> 1) ^ (row >> 2)) & 3; }

__global__ __launch_bounds__(256, 2) void dense_cov_tile_kernel(const _Float16* __restrict__ hi,
                                                             const _Float16* __restrict__ lo,
                                                             int M, int Kp, int ntile, int nsplit,
                                                             float out_scale,
                                                             float* __restrict__ Cp)
{
    __shared__ __attribute__((aligned(16))) _Float16 sm[2][4][kGPanel];
    int t = blockIdx.x / nsplit;
    const int split = blockIdx.x % nsplit;
    int ti = 0;
    while (t >= ntile - ti) { t -= ntile - ti; ++ti; }
    const int tj = ti + t;
    const int nsteps = Kp / kGStep;
    const int s0 = (int)((long long)nsteps * split / nsplit);
    const int s1 = (int)((long long)nsteps * (split + 1) / nsplit);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // loader: chunks c = tid, tid + 256 of each panel, chunk c = component
    // c/4, frames 8*(c%4).. of the step
    // element offsets into hi/lo (M * Kp < 2^31 is checked by the launcher)
    uint32_t aoff[2], boff[2];
    int soff[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int c = tid + 256 * h;
        const int ra = min(ti * kGTile + c / 4, M - 1), rb = min(tj * kGTile + c / 4, M - 1);
        aoff[h] = (uint32_t)ra * (uint32_t)Kp + 8u * (c % 4);
        boff[h] = (uint32_t)rb * (uint32_t)Kp + 8u * (c % 4);
        soff[h] = (c / 4) * kGPitch + 8 * ((c % 4) ^ gsw(c / 4));
    }
    const int wr = (w >> 1) * 64, wc = (w & 1) * 64;   // wave sub-tile in the block tile
    const int r = lane & 31, kq = (lane >> 5) * 8;
    float16v h00 = {}, h01 = {}, h10 = {}, h11 = {}, x00 = {}, x01 = {}, x10 = {}, x11 = {};
    // registers of the next step's panels: (A hi, A lo, B hi, B lo) x (chunk 0, 1)
    uint4 g0, g1, g2, g3, g4, g5, g6, g7;
#define GMX_FETCH(step)                                                                  \
    do {                                                                                 \
        const uint32_t k_ = (uint32_t)(step) * kGStep;                                   \
        g0 = *reinterpret_cast<const uint4*>(hi + aoff[0] + k_);                         \
        g1 = *reinterpret_cast<const uint4*>(hi + aoff[1] + k_);                         \
        g2 = *reinterpret_cast<const uint4*>(lo + aoff[0] + k_);                         \
        g3 = *reinterpret_cast<const uint4*>(lo + aoff[1] + k_);                         \
        g4 = *reinterpret_cast<const uint4*>(hi + boff[0] + k_);                         \
        g5 = *reinterpret_cast<const uint4*>(hi + boff[1] + k_);                         \
        g6 = *reinterpret_cast<const uint4*>(lo + boff[0] + k_);                         \
        g7 = *reinterpret_cast<const uint4*>(lo + boff[1] + k_);                         \
    } while (0)
#define GMX_STAGE(buf)                                                                   \
    do {                                                                                 \
        *reinterpret_cast<uint4*>(&sm[buf][0][soff[0]]) = g0;                            \
        *reinterpret_cast<uint4*>(&sm[buf][0][soff[1]]) = g1;                            \
        *reinterpret_cast<uint4*>(&sm[buf][1][soff[0]]) = g2;                            \
        *reinterpret_cast<uint4*>(&sm[buf][1][soff[1]]) = g3;                            \
        *reinterpret_cast<uint4*>(&sm[buf][2][soff[0]]) = g4;                            \
        *reinterpret_cast<uint4*>(&sm[buf][2][soff[1]]) = g5;                            \
        *reinterpret_cast<uint4*>(&sm[buf][3][soff[0]]) = g6;                            \
        *reinterpret_cast<uint4*>(&sm[buf][3][soff[1]]) = g7;                            \
    } while (0)
#define GMX_STEP(buf)                                                                    \
    do {                                                                                 \
        const _Float16* pa = sm[buf][0];                                                 \
        _Pragma("unroll") for (int k16 = 0; k16 < kGStep; k16 += 16)                     \
        {                                                                                \
            const int qu = (k16 + kq) >> 3;                                              \
            const int ao = (wr + r) * kGPitch + 8 * (qu ^ gsw(wr + r));                  \
            const int bo = (wc + r) * kGPitch + 8 * (qu ^ gsw(wc + r));                  \
            const half8 ah0 = *reinterpret_cast<const half8*>(pa + ao);                  \
            const half8 ah1 = *reinterpret_cast<const half8*>(pa + ao + 32 * kGPitch);   \
            const half8 al0 = *reinterpret_cast<const half8*>(pa + kGPanel + ao);        \
            const half8 al1 =                                                            \
                *reinterpret_cast<const half8*>(pa + kGPanel + ao + 32 * kGPitch);       \
            const half8 bh0 = *reinterpret_cast<const half8*>(pa + 2 * kGPanel + bo);    \
            const half8 bh1 =                                                            \
                *reinterpret_cast<const half8*>(pa + 2 * kGPanel + bo + 32 * kGPitch);   \
            const half8 bl0 = *reinterpret_cast<const half8*>(pa + 3 * kGPanel + bo);    \
            const half8 bl1 =                                                            \
                *reinterpret_cast<const half8*>(pa + 3 * kGPanel + bo + 32 * kGPitch);   \
            h00 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah0, bh0, h00, 0, 0, 0);        \
            h01 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah0, bh1, h01, 0, 0, 0);        \
            h10 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah1, bh0, h10, 0, 0, 0);        \
            h11 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah1, bh1, h11, 0, 0, 0);        \
            x00 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah0, bl0, x00, 0, 0, 0);        \
            x01 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah0, bl1, x01, 0, 0, 0);        \
            x10 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah1, bl0, x10, 0, 0, 0);        \
            x11 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah1, bl1, x11, 0, 0, 0);        \
            x00 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al0, bh0, x00, 0, 0, 0);        \
            x01 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al0, bh1, x01, 0, 0, 0);        \
            x10 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al1, bh0, x10, 0, 0, 0);        \
            x11 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al1, bh1, x11, 0, 0, 0);        \
        }                                                                                \
    } while (0)
    GMX_FETCH(s0);
    GMX_STAGE(0);
    __syncthreads();
    // steps s0 .. s1-2 stage the next step while multiplying; the last one
    // only multiplies
    for (int step = s0; step < s1 - 1; ++step) {
        const int cur = (step - s0) & 1;
        GMX_FETCH(step + 1);
        GMX_STEP(cur);
        GMX_STAGE(cur ^ 1);
        __syncthreads();
    }
    GMX_STEP((s1 - 1 - s0) & 1);
#undef GMX_STEP
#undef GMX_FETCH
#undef GMX_STAGE
    float* C = Cp + (size_t)split * M * M;
    const bool mirror = ti != tj;
    // one 32x32 block at a time: result register q of lane l is row
    // 8*(q/4) + 4*(l/32) + q%4, column l%32
#define GMX_STORE(H, X, U, V)                                                            \
    do {                                                                                 \
        const int j_ = tj * kGTile + wc + 32 * (V) + (lane & 31);                        \
        _Pragma("unroll") for (int q = 0; q < 16; ++q)                                   \
        {                                                                                \
            const int i_ = ti * kGTile + wr + 32 * (U) + 8 * (q >> 2) + 4 * (lane >> 5) + \
                           (q & 3);                                                      \
            if (i_ < M && j_ < M) {                                                      \
                const float val = ((H)[q] + (X)[q] * (1.0f / 2048.0f)) * out_scale;      \
                C[(size_t)i_ * M + j_] = val;                                            \
                if (mirror) C[(size_t)j_ * M + i_] = val;                                \
            }                                                                            \
        }                                                                                \
    } while (0)
    GMX_STORE(h00, x00, 0, 0);
    GMX_STORE(h01, x01, 0, 1);
    GMX_STORE(h10, x10, 1, 0);
    GMX_STORE(h11, x11, 1, 1);
#undef GMX_STORE
}

// C[e] = sum over splits of Cp[s][e], in split order.
__global__ void dense_cov_reduce_kernel(const float* __restrict__ Cp, long long n, int nsplit,
                                        float* __restrict__ C)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        float v = Cp[e];
        for (int s = 1; s < nsplit; ++s) v += Cp[(size_t)s * n + e];
        C[e] = v;
    }
}

// ---------------------------------------------------------------------------
// Per-pair MI from C (FP32 -> FP64), one thread per pair (a1 < a2), pairs
// enumerated row by row of the upper triangle.
// mode 0: Gaussian, MI = H(C11) + H(C22) - H(CJ), H = 0.5(n(1+log 2pi) + log det)
//         or 0 when det <= 0 (lib/correlation_core.cpp:100-111, 280-311);
// mode 1: GCMI, MI = 0.5(log d11 + log d22 - log dJ), 0 if any det <= 0
//         (lib/gcmi.cpp:212-231).
// ---------------------------------------------------------------------------
__global__ void dense_pair_mi_kernel(const float* __restrict__ C, int natoms, int mode,
                                     double* __restrict__ mat)
{
    const long long npairs = (long long)natoms * (natoms - 1) / 2;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    // decode p -> (a1, a2) with a1 < a2, row-major over the upper triangle
    int a1 = (int)((2.0 * natoms - 1 - sqrt((2.0 * natoms - 1) * (2.0 * natoms - 1) - 8.0 * p)) / 2);
    auto start = [&](int a) { return (long long)a * (2 * natoms - a - 1) / 2; };
    while (a1 > 0 && start(a1) > p) --a1;
    while (start(a1 + 1) <= p) ++a1;
    const int a2 = a1 + 1 + (int)(p - start(a1));
    const int M = 3 * natoms;
    double c1[9], c2[9], cj[36];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const double x11 = C[(size_t)(3 * a1 + r) * M + 3 * a1 + c];
            const double x22 = C[(size_t)(3 * a2 + r) * M + 3 * a2 + c];
            const double x12 = C[(size_t)(3 * a1 + r) * M + 3 * a2 + c];
            c1[r * 3 + c] = x11;
            c2[r * 3 + c] = x22;
            cj[r * 6 + c] = x11;
            cj[(r + 3) * 6 + c + 3] = x22;
            cj[r * 6 + c + 3] = x12;
            cj[(c + 3) * 6 + r] = x12;
        }
    const double d1 = lu_det<3>(c1), d2 = lu_det<3>(c2), dj = lu_det<6>(cj);
    double mi;
    if (mode == 0) {
        auto H = [](double det, int n) {
            return det <= 0.0 ? 0.0 : 0.5 * (n * (1.0 + log(2.0 * kPi)) + log(det));
        };
        mi = H(d1, 3) + H(d2, 3) - H(dj, 6);
    } else {
        mi = (d1 <= 0.0 || d2 <= 0.0 || dj <= 0.0) ? 0.0 : 0.5 * (log(d1) + log(d2) - log(dj));
    }
    mat[(size_t)a1 * natoms + a2] = mi;
    mat[(size_t)a2 * natoms + a1] = mi;
    if (a2 == a1 + 1) mat[(size_t)a1 * natoms + a1] = 2000.0;
    if (a1 == natoms - 2 && a2 == natoms - 1) mat[(size_t)a2 * natoms + a2] = 2000.0;
}

// ---------------------------------------------------------------------------
// Copula scores: block per component row; the row is staged in LDS when it
// fits.  Rank by counting (exact FP64 comparisons): lt = #{g: v_g < v_f},
// eq = #{g: v_g == v_f}; the score is the mean over t in [lt, lt+eq) of
// sqrt(2) erfinv(2(t + 0.5)/N - 1), summed in ascending t as the reference.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void dense_copula_kernel(const double* __restrict__ V, int N,
                                                           long long ldv,
                                                           const double* __restrict__ ztab,
                                                           double* __restrict__ Z,
                                                           long long ldz, int stage)
{
    extern __shared__ double srow[];
    const long long c = blockIdx.x;
    const double* row = V + c * ldv;
    if (stage) {
        for (int f = threadIdx.x; f < N; f += blockDim.x) srow[f] = row[f];
        __syncthreads();
        row = srow;
    }
    for (int f = threadIdx.x; f < N; f += blockDim.x) {
        const double v = row[f];
        int lt = 0, eq = 0;
        for (int g = 0; g < N; ++g) {
            const double u = row[g];
            lt += u < v;
            eq += u == v;
        }
        double z = 0.0;
        for (int t = lt; t < lt + eq; ++t) z += ztab[t];
        Z[c * ldz + f] = z / eq;
    }
}

// Normal score of rank t, the same for every component row:
// ztab[t] = sqrt(2) erfinv(2 (t + 0.5) / N - 1), the term lib/gcmi.cpp:130-135
// adds per rank (a product, then the sum: same rounding as the reference).
__global__ void dense_score_table_kernel(int N, double* __restrict__ ztab)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const double pr = (t + 0.5) / N;
    ztab[t] = sqrt(2.0) * erfinv_acklam(2.0 * pr - 1.0);
}

// Order-preserving 64-bit key of a finite double; v + 0.0 maps -0.0 to +0.0,
// so values that compare equal share one key.
__device__ __forceinline__ uint64_t order_key(double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v + 0.0);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// ---------------------------------------------------------------------------
// Copula scores by sorting (N <= kCopulaSortFrames): block of 1024 threads per
// component row.  The row's M = 1024 E >= N order keys (padding keys all-ones
// sort last) are bitonic-sorted in registers, E per thread (sort_network.h:
// thread distances of 64 and more through LDS), and stored to LDS in the
// network's conflict-free layout (position p at sortnet::slot(p)); then per
// frame lt = #{g: v_g < v_f} and eq = #{g: v_g == v_f} by binary
// search.  Same scores as dense_copula_kernel, in O(N log^2 N) compare-
// exchanges per row, most of them in registers.
// ---------------------------------------------------------------------------
constexpr int kCopulaThreads = 1024;

template <int E>
__global__ __launch_bounds__(kCopulaThreads) void dense_copula_sort_kernel(
    const double* __restrict__ V, int N, long long ldv, const double* __restrict__ ztab,
    double* __restrict__ Z, long long ldz)
{
    constexpr int M = kCopulaThreads * E;
    extern __shared__ uint64_t keys[];   // M keys
    const long long c = blockIdx.x;
    const double* row = V + c * ldv;
    const int t = threadIdx.x;
    uint64_t r[E];
    // the network sorts any arrangement: frames e * 1024 + t (coalesced loads)
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int f = e * kCopulaThreads + t;
        r[e] = f < N ? order_key(row[f]) : ~0ull;
    }
    sortnet::sort_blocked<uint64_t, E, kCopulaThreads>(r, keys, t);
    __syncthreads();   // the network's last LDS readers are done
#pragma unroll
    for (int e = 0; e < E; ++e) keys[e * kCopulaThreads + t] = r[e];   // position E t + e
    __syncthreads();
    for (int f = t; f < N; f += kCopulaThreads) {
        const uint64_t key = order_key(row[f]);
        int lo = 0, hi = N;
        while (lo < hi) {   // first position with keys >= key
            const int mid = (lo + hi) >> 1;
            if (keys[sortnet::slot<E, kCopulaThreads>(mid)] < key) lo = mid + 1; else hi = mid;
        }
        const int lt = lo;
        // position lt holds the frame's own key; most values are unique, so
        // look at the next position before searching for the first key above
        lo = lt + 1;
        if (lo < N && keys[sortnet::slot<E, kCopulaThreads>(lo)] <= key) {
            hi = N;
            while (lo < hi) {   // first position with keys > key
                const int mid = (lo + hi) >> 1;
                if (keys[sortnet::slot<E, kCopulaThreads>(mid)] <= key) lo = mid + 1; else hi = mid;
            }
        }
        const int eq = lo - lt;
        double z = 0.0;
        for (int t = lt; t < lt + eq; ++t) z += ztab[t];
        Z[c * ldz + f] = z / eq;
    }
}

// traj [n][N][3] -> component rows [3n][N]
__global__ void dense_rows_kernel(const double* __restrict__ traj, int natoms, int N,
                                  double* __restrict__ rows)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)natoms * N * 3;
    if (t >= total) return;
    const long long a = t / (3LL * N);
    const long long rem = t % (3LL * N);
    const int f = (int)(rem / 3), d = (int)(rem % 3);
    rows[(a * 3 + d) * N + f] = traj[t];
}

hipError_t launch_dense_rows(const double* traj, int natoms, int N, double* rows, hipStream_t s)
{
    const long long total = (long long)natoms * N * 3;
    hipLaunchKernelGGL(dense_rows_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       traj, natoms, N, rows);
    return hipGetLastError();
}

hipError_t launch_dense_split(const double* V, int M, int N, long long ldv, int Kp, double scale,
                              _Float16* hi, _Float16* lo, hipStream_t s)
{
    const int bx = (Kp + 255) / 256 < 64 ? (Kp + 255) / 256 : 64;
    hipLaunchKernelGGL(dense_split_kernel, dim3(bx, M), dim3(256), 0, s, V, N, ldv, Kp, scale, hi,
                       lo);
    return hipGetLastError();
}

int dense_cov_splits(int M, int Kp)
{
    const long long ntile = (M + kGTile - 1) / kGTile;
    const long long tiles = ntile * (ntile + 1) / 2;
    int ns = (int)((1024 + tiles - 1) / tiles);
    ns = std::min(ns, std::min(8, std::max(1, Kp / kGStep / 4)));
    return std::max(ns, 1);
}

hipError_t launch_dense_cov(const _Float16* hi, const _Float16* lo, int M, int Kp,
                            float out_scale, float* Cp, int nsplit, float* C, hipStream_t s,
                            int nreduce)
{
    if (Kp % kGStep != 0 || nsplit < 1 || Kp / kGStep < nsplit ||
        (long long)M * Kp >= (1LL << 31))
        return hipErrorInvalidValue;
    const int ntile = (M + kGTile - 1) / kGTile;
    const unsigned nblk = (unsigned)(ntile * (ntile + 1) / 2 * nsplit);
    hipLaunchKernelGGL(dense_cov_tile_kernel, dim3(nblk), dim3(256), 0, s, hi, lo, M, Kp, ntile,
                       nsplit, out_scale, Cp);
    if (nreduce == 0 || (nsplit == 1 && Cp == C)) return hipGetLastError();
    return launch_dense_reduce(Cp, M, nreduce < 0 ? nsplit : nreduce, C, s);
}

hipError_t launch_dense_reduce(const float* Cp, int M, int nparts, float* C, hipStream_t s)
{
    const long long n = (long long)M * M;
    hipLaunchKernelGGL(dense_cov_reduce_kernel, dim3(1024), dim3(256), 0, s, Cp, n, nparts, C);
    return hipGetLastError();
}

hipError_t launch_dense_pairs(const float* C, int natoms, int mode, double* mat, hipStream_t s)
{
    const long long npairs = (long long)natoms * (natoms - 1) / 2;
    hipLaunchKernelGGL(dense_pair_mi_kernel, dim3((unsigned)((npairs + 255) / 256)), dim3(256), 0,
                       s, C, natoms, mode, mat);
    return hipGetLastError();
}

hipError_t launch_dense_copula(const double* V, int M, int N, long long ldv, double* Z,
                               long long ldz, double* ztab, hipStream_t s)
{
    hipLaunchKernelGGL(dense_score_table_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                       s, N, ztab);
    if (N > kCopulaMaxFrames) {
        // the row does not fit in LDS: the counting kernel reads it from
        // global memory (L2), O(N^2) per row
        hipLaunchKernelGGL(dense_copula_kernel, dim3(M), dim3(256), 0, s, V, N, ldv,
                           (const double*)ztab, Z, ldz, 0);
        return hipGetLastError();
    }
    if (N <= kCopulaSortFrames) {
        static_assert(kCopulaSortFrames == 16 * kCopulaThreads, "E = 16 covers the sort range");
        auto run = [&](auto kernel, int E) -> hipError_t {
            const size_t kbytes = (size_t)kCopulaThreads * E * sizeof(uint64_t);
            hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)kbytes);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(kernel, dim3(M), dim3(kCopulaThreads), kbytes, s, V, N, ldv,
                               (const double*)ztab, Z, ldz);
            return hipGetLastError();
        };
        if (N <= kCopulaThreads) return run(dense_copula_sort_kernel<1>, 1);
        if (N <= 2 * kCopulaThreads) return run(dense_copula_sort_kernel<2>, 2);
        if (N <= 4 * kCopulaThreads) return run(dense_copula_sort_kernel<4>, 4);
        if (N <= 8 * kCopulaThreads) return run(dense_copula_sort_kernel<8>, 8);
        return run(dense_copula_sort_kernel<16>, 16);
    }
    const size_t bytes = (size_t)N * sizeof(double);
    hipError_t e = hipFuncSetAttribute((const void*)dense_copula_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(dense_copula_kernel, dim3(M), dim3(256), bytes, s, V, N, ldv,
                       (const double*)ztab, Z, ldz, 1);
    return hipGetLastError();
}

}  // namespace gmxcorr
