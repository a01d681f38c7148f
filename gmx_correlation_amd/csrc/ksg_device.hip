// KSG-2 mutual-information kernels for gfx950 (MI355X).
//
// The estimator is the reference's mir_xnyn (lib/kraskov.cpp:1228-1331) for
// two 3-D atoms, evaluated for every frame i of every atom pair (a, b):
//   dA(i,j) = max_d |x_{a,d}(i) - x_{a,d}(j)|, dB likewise, dJ = max(dA, dB)
//   S_i     = the k frames j != i with the smallest dJ
//   epsx    = max_{j in S_i} dA(i,j), epsy likewise
//   nx      = #{j != i : dA(i,j) <= epsx}, ny likewise
//   MI      = psi[N] + (psi[k] - 1/k) - (1/N) * sum_i (psi[nx] + psi[ny])
//
// Layout in HBM: every atom's N x N matrix of dA values is built once in FP32
// (row pitch P floats, padding = +inf) from the FP64 coordinates.  Rounding
// to FP32 is monotone, so an FP32 comparison can only disagree with FP64 when
// two FP32 values are *equal*; such rows are detected exactly (a tie at the
// (k+1)-th joint key, or more than one frame at the marginal radius) and
// recounted from the FP64 coordinates by ksg_fixup_kernel.  All other rows
// are bit-exact against the FP64 definition by construction.
//
// Pair kernel: one wavefront per (pair, frame row).  The two rows are
// streamed with coalesced 16-byte loads into registers (lane l holds frames
// 256*m + 4*l + c), the (k+1)-th smallest joint key (self included, key 0)
// is found by a three-digit radix select (11 + 10 + 10 bits of the
// non-negative FP32 bit pattern) with a wave-private LDS histogram, then one
// pass takes the marginal radii over the selected set and one pass counts.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ksg_internal.h"

namespace gmxcorr {

namespace {

constexpr int kWave = 64;
constexpr int kPairWaves = 4;
constexpr uint32_t kInfBits = 0x7f800000u;

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane)
{
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// Locate the bin holding the t-th smallest element (1-based) of a
// histogram of 64*BPL bins.  Lane l owns bins [l*BPL, (l+1)*BPL).  Returns the
// bin, the rank of the target inside it and the bin's population.
template <int BPL>
__device__ __forceinline__ void find_bin(const uint32_t* hist, int lane, int t, int& bin, int& rank,
                                         int& pop)
{
    int s = 0;
    // rotated order keeps the 64 lanes on distinct LDS banks
#pragma unroll
    for (int q = 0; q < BPL; ++q) s += (int)hist[lane * BPL + ((q + lane) & (BPL - 1))];
    const int incl = wave_incl_scan(s, lane);
    const int excl = incl - s;
    const uint64_t bal = __ballot(excl < t && t <= incl);
    const int owner = __ffsll((unsigned long long)bal) - 1;
    const int base = __shfl(excl, owner, kWave);
    // the owner's BPL bins spread over the first BPL lanes
    const int v = lane < BPL ? (int)hist[owner * BPL + lane] : 0;
    const int vi = wave_incl_scan(v, lane);
    const int r = t - base;
    const uint64_t b2 = __ballot(lane < BPL && vi - v < r && r <= vi);
    const int sub = __ffsll((unsigned long long)b2) - 1;
    bin = owner * BPL + sub;
    rank = r - __shfl(vi - v, sub, kWave);
    pop = __shfl(v, sub, kWave);
}

template <int NB>
__device__ __forceinline__ void clear_hist(uint32_t* hist, int lane)
{
#pragma unroll
    for (int q = 0; q < NB / (4 * kWave); ++q)
        reinterpret_cast<uint4*>(hist)[q * kWave + lane] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace

// ---------------------------------------------------------------------------
// Per-atom distance rows: D[a][i][j] = fl32(max_d |X[3a+d][i] - X[3a+d][j]|)
// computed in FP64 exactly as the reference does (fabs of the difference,
// lib/kraskov.cpp:131-132, 280-281), then rounded once.  j in [N, P) = +inf.
// One block per row (a, i).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ksg_build_rows_kernel(const double* __restrict__ X, int N,
                                                             long long P, float* __restrict__ D)
{
    const long long row = blockIdx.x;
    const int a = (int)(row / N), i = (int)(row % N);
    const double* x0 = X + (size_t)(3 * a) * N;
    const double* x1 = x0 + N;
    const double* x2 = x1 + N;
    const double q0 = x0[i], q1 = x1[i], q2 = x2[i];
    float4* out = reinterpret_cast<float4*>(D + row * P);
    const int nc = (int)(P / 4);
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int j = 4 * c + e;
            if (j < N) {
                double d = fabs(q0 - x0[j]);
                d = fmax(d, fabs(q1 - x1[j]));
                d = fmax(d, fabs(q2 - x2[j]));
                v[e] = (float)d;
            } else {
                v[e] = __int_as_float((int)kInfBits);
            }
        }
        out[c] = make_float4(v[0], v[1], v[2], v[3]);
    }
}

// ---------------------------------------------------------------------------
// Pair kernel.  Block = 4 waves = one pair; wave w takes frames w, w+4, ...
// S4 = 16-byte chunks per lane per row (N <= 256*S4).
// ---------------------------------------------------------------------------
template <int S4>
__global__ __launch_bounds__(256) void ksg_pair_kernel(KsgArgs args)
{
    __shared__ uint32_t hist_all[kPairWaves][2048];
    __shared__ double wsum[kPairWaves];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    uint32_t* hist = hist_all[w];
    const long long p = blockIdx.x;
    const int N = args.N, K = args.K, T = args.K + 1;
    const long long P = args.P;
    const float* Da = args.D + (size_t)args.a1[p] * N * P;
    const float* Db = args.D + (size_t)args.a2[p] * N * P;
    const int nchunk = (N + 3) / 4;
    double acc = 0.0;

    for (int i = w; i < N; i += kPairWaves) {
        const float4* ra = reinterpret_cast<const float4*>(Da + (size_t)i * P);
        const float4* rb = reinterpret_cast<const float4*>(Db + (size_t)i * P);
        float a[4 * S4], b[4 * S4];
#pragma unroll
        for (int m = 0; m < S4; ++m) {
            const int c = m * kWave + lane;
            float4 va = make_float4(INFINITY, INFINITY, INFINITY, INFINITY), vb = va;
            if (c < nchunk) {
                va = ra[c];
                vb = rb[c];
            }
            a[4 * m + 0] = va.x; a[4 * m + 1] = va.y; a[4 * m + 2] = va.z; a[4 * m + 3] = va.w;
            b[4 * m + 0] = vb.x; b[4 * m + 1] = vb.y; b[4 * m + 2] = vb.z; b[4 * m + 3] = vb.w;
        }

        // ---- radix select of the T-th smallest joint key -----------------
        // digit 1: bits 30..20 (2048 bins)
        clear_hist<2048>(hist, lane);
        wave_sync();
#pragma unroll
        for (int e = 0; e < 4 * S4; ++e) {
            const uint32_t key = __float_as_uint(fmaxf(a[e], b[e]));
            if (key < kInfBits) atomicAdd(&hist[key >> 20], 1u);
        }
        wave_sync();
        int bin, rank, pop;
        find_bin<32>(hist, lane, T, bin, rank, pop);
        uint32_t prefix = (uint32_t)bin;
        wave_sync();
        // digit 2: bits 19..10 (1024 bins)
        clear_hist<1024>(hist, lane);
        wave_sync();
#pragma unroll
        for (int e = 0; e < 4 * S4; ++e) {
            const uint32_t key = __float_as_uint(fmaxf(a[e], b[e]));
            if ((key >> 20) == prefix) atomicAdd(&hist[(key >> 10) & 1023u], 1u);
        }
        wave_sync();
        find_bin<16>(hist, lane, rank, bin, rank, pop);
        prefix = (prefix << 10) | (uint32_t)bin;
        wave_sync();
        // digit 3: bits 9..0 (1024 bins)
        clear_hist<1024>(hist, lane);
        wave_sync();
#pragma unroll
        for (int e = 0; e < 4 * S4; ++e) {
            const uint32_t key = __float_as_uint(fmaxf(a[e], b[e]));
            if ((key >> 10) == prefix) atomicAdd(&hist[key & 1023u], 1u);
        }
        wave_sync();
        find_bin<16>(hist, lane, rank, bin, rank, pop);
        const uint32_t eps_key = (prefix << 10) | (uint32_t)bin;
        wave_sync();
        // keys < eps: T - rank; keys <= eps: that + pop.  Unambiguous iff == T.
        const bool sel_tie = (T - rank + pop) != T;

        // ---- marginal radii over the selected set -------------------------
        float ex = 0.0f, ey = 0.0f;
#pragma unroll
        for (int e = 0; e < 4 * S4; ++e) {
            const bool in = __float_as_uint(fmaxf(a[e], b[e])) <= eps_key;
            ex = fmaxf(ex, in ? a[e] : 0.0f);
            ey = fmaxf(ey, in ? b[e] : 0.0f);
        }
        ex = wave_max(ex);
        ey = wave_max(ey);

        // ---- marginal counts (self included, removed below) ---------------
        int cx = 0, cy = 0;
#pragma unroll
        for (int e = 0; e < 4 * S4; ++e) {
            cx += (a[e] <= ex) + ((a[e] == ex) << 16);
            cy += (b[e] <= ey) + ((b[e] == ey) << 16);
        }
        cx = wave_sum(cx);
        cy = wave_sum(cy);
        const int nx = (cx & 0xffff) - 1, ny = (cy & 0xffff) - 1;
        const bool flag = sel_tie || (cx >> 16) > 1 || (cy >> 16) > 1;

        if (lane == 0) {
            if (flag) {
                const unsigned slot = atomicAdd(&args.counters[0], 1u);
                if (slot < args.fixcap) {
                    FixEntry fe;
                    fe.pair = (int)p;
                    fe.row = i;
                    fe.eps_bits = eps_key;
                    fe.pad = 0;
                    args.fixq[slot] = fe;
                }
            } else {
                acc += args.psi[nx] + args.psi[ny];
            }
            if (args.rc_nx) {
                const size_t o = (size_t)p * N + i;
                args.rc_nx[o] = flag ? -1 : nx;
                args.rc_ny[o] = flag ? -1 : ny;
                args.rc_flag[o] = 0;
            }
        }
    }
    if (lane == 0) wsum[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) args.partial[p] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

// ---------------------------------------------------------------------------
// FP64 recount of queued rows: one wave per entry.  Recomputes dA, dB in
// FP64 from the coordinates.  The FP32 key eps_bits fixes the group G of
// frames whose joint distance rounds to it; the k-th neighbour distance is
// the r-th smallest FP64 value in G with r = k - #{j != i : key < eps_bits}.
// More than r frames at that FP64 value is an exact tie: all of them enter
// the radius set (the tie is counted).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ksg_fixup_kernel(KsgArgs args, unsigned nent)
{
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned e = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (e >= nent) return;
    const FixEntry fe = args.fixq[e];
    const int p = fe.pair, i = fe.row, N = args.N, K = args.K;
    const double* xa = args.X + (size_t)(3 * args.a1[p]) * N;
    const double* xb = args.X + (size_t)(3 * args.a2[p]) * N;
    const double qa0 = xa[i], qa1 = xa[N + i], qa2 = xa[2 * N + i];
    const double qb0 = xb[i], qb1 = xb[N + i], qb2 = xb[2 * N + i];
    auto dist = [&](int j, double& A, double& B) {
        A = fmax(fmax(fabs(qa0 - xa[j]), fabs(qa1 - xa[N + j])), fabs(qa2 - xa[2 * N + j]));
        B = fmax(fmax(fabs(qb0 - xb[j]), fabs(qb1 - xb[N + j])), fabs(qb2 - xb[2 * N + j]));
    };
    const uint32_t ek = fe.eps_bits;

    int clt = 0;
    for (int j = lane; j < N; j += kWave) {
        if (j == i) continue;
        double A, B;
        dist(j, A, B);
        clt += __float_as_uint((float)fmax(A, B)) < ek;
    }
    clt = wave_sum(clt);
    const int r = K - clt;  // >= 1

    // walk the distinct FP64 values of G upwards until r are covered
    double prev = -1.0, eps64 = 0.0;
    int cum = 0;
    bool tie = false;
    for (;;) {
        double vmin = INFINITY;
        for (int j = lane; j < N; j += kWave) {
            if (j == i) continue;
            double A, B;
            dist(j, A, B);
            const double J = fmax(A, B);
            if (__float_as_uint((float)J) == ek && J > prev) vmin = fmin(vmin, J);
        }
        vmin = wave_min(vmin);
        int m = 0;
        for (int j = lane; j < N; j += kWave) {
            if (j == i) continue;
            double A, B;
            dist(j, A, B);
            m += fmax(A, B) == vmin;
        }
        m = wave_sum(m);
        if (m == 0 || cum + m >= r) {  // m == 0 cannot happen for a consistent entry
            eps64 = vmin;
            tie = cum + m > r;
            break;
        }
        cum += m;
        prev = vmin;
    }

    double ex = 0.0, ey = 0.0;
    for (int j = lane; j < N; j += kWave) {
        if (j == i) continue;
        double A, B;
        dist(j, A, B);
        if (fmax(A, B) <= eps64) {
            ex = fmax(ex, A);
            ey = fmax(ey, B);
        }
    }
    ex = wave_max(ex);
    ey = wave_max(ey);
    int nx = 0, ny = 0;
    for (int j = lane; j < N; j += kWave) {
        if (j == i) continue;
        double A, B;
        dist(j, A, B);
        nx += A <= ex;
        ny += B <= ey;
    }
    nx = wave_sum(nx);
    ny = wave_sum(ny);
    if (lane == 0) {
        atomicAdd(&args.fixsum[p], args.psi[nx] + args.psi[ny]);
        if (tie) atomicAdd(&args.counters[1], 1u);
        if (args.rc_nx) {
            const size_t o = (size_t)p * N + i;
            args.rc_nx[o] = nx;
            args.rc_ny[o] = ny;
            args.rc_flag[o] = tie ? 2 : 1;
        }
    }
}

__global__ void ksg_clear_kernel(double* fixsum, long long npairs, unsigned* counters)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < npairs) fixsum[t] = 0.0;
    if (t < 2) counters[t] = 0u;
}

// MI = psi[N] + phi[K] - sum/N as lib/kraskov.cpp:1318.
__global__ void ksg_finalize_kernel(KsgArgs args)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < args.npairs) args.mi[p] = args.psiN_phiK - (args.partial[p] + args.fixsum[p]) / args.N;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_build_rows(const double* X, int natoms, int N, long long P, float* D,
                             hipStream_t s)
{
    const long long rows = (long long)natoms * N;
    hipLaunchKernelGGL(ksg_build_rows_kernel, dim3((unsigned)rows), dim3(256), 0, s, X, N, P, D);
    return hipGetLastError();
}

hipError_t launch_ksg_clear(const KsgArgs& a, hipStream_t s)
{
    const long long n = a.npairs > 2 ? a.npairs : 2;
    hipLaunchKernelGGL(ksg_clear_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       a.fixsum, a.npairs, a.counters);
    return hipGetLastError();
}

hipError_t launch_ksg_pairs(const KsgArgs& a, hipStream_t s)
{
    const int nchunk = (a.N + 3) / 4;
    const int s4 = (nchunk + kWave - 1) / kWave;
    const dim3 grid((unsigned)a.npairs), block(kWave * kPairWaves);
    if (s4 <= 1)
        hipLaunchKernelGGL(ksg_pair_kernel<1>, grid, block, 0, s, a);
    else if (s4 <= 2)
        hipLaunchKernelGGL(ksg_pair_kernel<2>, grid, block, 0, s, a);
    else if (s4 <= 4)
        hipLaunchKernelGGL(ksg_pair_kernel<4>, grid, block, 0, s, a);
    else if (s4 <= 8)
        hipLaunchKernelGGL(ksg_pair_kernel<8>, grid, block, 0, s, a);
    else if (s4 <= 16)
        hipLaunchKernelGGL(ksg_pair_kernel<16>, grid, block, 0, s, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_ksg_fixup(const KsgArgs& a, unsigned nqueued, hipStream_t s)
{
    if (nqueued == 0) return hipSuccess;
    const unsigned per_block = 4;
    hipLaunchKernelGGL(ksg_fixup_kernel, dim3((nqueued + per_block - 1) / per_block),
                       dim3(kWave * per_block), 0, s, a, nqueued);
    return hipGetLastError();
}

hipError_t launch_ksg_finalize(const KsgArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(ksg_finalize_kernel, dim3((unsigned)((a.npairs + 255) / 256)), dim3(256), 0,
                       s, a);
    return hipGetLastError();
}

}  // namespace gmxcorr
