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
// (k+1)-th joint key, or more than one frame at a marginal radius) and
// recounted from the FP64 coordinates by ksg_fixup_kernel.  All other rows
// are bit-exact against the FP64 definition by construction.
//
// Pair kernel, one wavefront per (pair, frame row).  The two rows stream into
// registers with 16-byte loads (lane l holds frames 256*m + 4*l + c).  The
// (k+1)-th smallest joint key (self included, key 0) is located in three
// steps, all on the non-negative FP32 bit pattern:
//   sweep A  histogram of 2048 bins, 64 per octave over [2^-20, 2^12), in a
//            wave-private LDS array;
//   sweep B  frames below the target bin are in S for sure: their marginal
//            maxima are taken on the fly; the few frames inside the target
//            bin are appended to lane-private LDS slots;
//   select   the candidates are compacted one per lane and the exact key is
//            found by a 17-step ballot radix over the bits below the bin.
// Sweep C counts the marginal neighbours.  Rows whose target bin is an edge
// bin or holds more than 64 frames go to the FP64 recount unresolved.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ksg_internal.h"
#include "sort_network.h"

namespace gmxcorr {

namespace {

constexpr int kWave = 64;
constexpr int kPairWaves = 4;
constexpr uint32_t kInfBits = 0x7f800000u;
// Row padding (frames N..P-1): 3072.0f.  The host accepts only rows whose
// coordinate spread is below 2048, so every real distance is below 2^11 and
// the padding is never a neighbour; all keys stay below 2^12.
constexpr uint32_t kPadBits = 0x45400000u;
constexpr uint32_t kUnknownEps = 0xffffffffu;
// Sweep-A binning: bin = clamp((key >> 17) - kBinBase, 0, 2047); key >> 17
// is (exponent << 6 | top 6 mantissa bits), bin 0 starts at 2^-20.
constexpr int kBinShift = 17;
constexpr int kBinBase = (127 - 20) << 6;
constexpr int kBins = 2048;
constexpr int kCandCap = 4;  // candidate slots per lane
constexpr int kBinRot = (kBinBase / (kBins / kWave)) % kWave;  // raw-bin rotation in lane chunks

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Orders a wave's accesses to its own LDS region: the LDS unit executes a
// wave's LDS instructions in issue order, so only compiler reordering has to
// be prevented (no s_waitcnt is emitted).
__device__ __forceinline__ void lds_order()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// DPP lane exchange; lanes without a valid source (or outside row_mask) read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}

__device__ __forceinline__ int readlane(int v, int lane)
{
    return __builtin_amdgcn_readlane(v, lane);
}

// Copy of a wave-uniform value that the compiler knows to be uniform (held
// in scalar registers): each 32-bit word through readfirstlane.
template <typename T>
__device__ __forceinline__ T uniform_copy(const T& v)
{
    static_assert(sizeof(T) % 4 == 0, "whole words");
    int w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
}

// Sum over the wave, returned uniform: quad swaps and row mirrors reduce each
// row of 16 lanes, four readlanes combine the rows.
__device__ __forceinline__ int wave_sum(int v)
{
    v += dpp<0xb1>(v);   // quad_perm [1,0,3,2]
    v += dpp<0x4e>(v);   // quad_perm [2,3,0,1]
    v += dpp<0x141>(v);  // row_half_mirror
    v += dpp<0x140>(v);  // row_mirror
    return (readlane(v, 0) + readlane(v, 16)) + (readlane(v, 32) + readlane(v, 48));
}

__device__ __forceinline__ int wave_max_i(int v)
{
    v = max(v, dpp<0xb1>(v));
    v = max(v, dpp<0x4e>(v));
    v = max(v, dpp<0x141>(v));
    v = max(v, dpp<0x140>(v));
    return max(max(readlane(v, 0), readlane(v, 16)), max(readlane(v, 32), readlane(v, 48)));
}

__device__ __forceinline__ float wave_max(float f)
{
    int v = __float_as_int(f);
    v = __float_as_int(fmaxf(__int_as_float(v), __int_as_float(dpp<0xb1>(v))));
    v = __float_as_int(fmaxf(__int_as_float(v), __int_as_float(dpp<0x4e>(v))));
    v = __float_as_int(fmaxf(__int_as_float(v), __int_as_float(dpp<0x141>(v))));
    v = __float_as_int(fmaxf(__int_as_float(v), __int_as_float(dpp<0x140>(v))));
    const float r0 = __int_as_float(readlane(v, 0)), r1 = __int_as_float(readlane(v, 16));
    const float r2 = __int_as_float(readlane(v, 32)), r3 = __int_as_float(readlane(v, 48));
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

// Inclusive prefix sum over the 64 lanes (row shifts, then row broadcasts).
__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += dpp<0x111>(v);        // row_shr:1
    v += dpp<0x112>(v);        // row_shr:2
    v += dpp<0x114>(v);        // row_shr:4
    v += dpp<0x118>(v);        // row_shr:8
    v += dpp<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u(uint32_t u)
{
    int v = (int)u;
    v = (int)max((uint32_t)v, (uint32_t)dpp<0xb1>(v));
    v = (int)max((uint32_t)v, (uint32_t)dpp<0x4e>(v));
    v = (int)max((uint32_t)v, (uint32_t)dpp<0x141>(v));
    v = (int)max((uint32_t)v, (uint32_t)dpp<0x140>(v));
    const uint32_t r0 = (uint32_t)readlane(v, 0), r1 = (uint32_t)readlane(v, 16);
    const uint32_t r2 = (uint32_t)readlane(v, 32), r3 = (uint32_t)readlane(v, 48);
    return max(max(r0, r1), max(r2, r3));
}

// Number of set bits of m below this lane.
__device__ __forceinline__ int mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Sweep-A bin of a key: the key clamped to [2^-20, 2^12) keeps its exponent
// and top 6 mantissa bits.
__device__ __forceinline__ int bin_of(uint32_t key)
{
    constexpr uint32_t kLoKey = (uint32_t)kBinBase << kBinShift;
    constexpr uint32_t kHiKey = ((uint32_t)(kBinBase + kBins) << kBinShift) - 1u;
    return (int)(min(max(key, kLoKey), kHiKey) >> kBinShift) - kBinBase;
}

// Global load at a uniform base plus a 32-bit byte offset (selects the
// scalar-base addressing form).
template <class T>
__device__ __forceinline__ T ld_lane(const void* base, uint32_t byte_off)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

typedef __attribute__((address_space(3))) uint32_t lds_u32;

// LDS byte address of a wave's histogram (8 KB aligned by declaration).
__device__ __forceinline__ uint32_t lds_addr(uint32_t* hist)
{
    return (uint32_t)(uintptr_t)(lds_u32*)hist;
}

// Sweep-A increment for the joint key max(a, b) < 2^12: the key raised to
// 2^-20, bits 27..17, i.e. the rotated bin (bin + kBinBase) mod 2048, OR-ed
// into the 8 KB-aligned histogram base (three VALU operations).
__device__ __forceinline__ void hist_add(uint32_t base, uint32_t a, uint32_t b)
{
    constexpr uint32_t kLoKey = (uint32_t)kBinBase << kBinShift;
    const uint32_t addr =
        ((max(max(a, b), kLoKey) >> (kBinShift - 2)) & (uint32_t)(4 * kBins - 4)) | base;
    __hip_atomic_fetch_add((lds_u32*)(uintptr_t)addr, 1u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// FP64 helpers for the recount kernel (rare rows; plain shuffles suffice).
__device__ __forceinline__ double wave_max_d(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ double wave_min_d(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}

// Bin holding the t-th smallest element (1-based) of a histogram of 64*BPL
// bins; lane l owns bins [l*BPL, (l+1)*BPL), stored from word
// ((l + ROT) % 64) * BPL of the array (the histogram is rotated by ROT*BPL
// words).  Returns the bin, the rank of the target inside it and the bin's
// population.
template <int BPL, int ROT = 0>
__device__ __forceinline__ void find_bin(const uint32_t* hist, int lane, int t, int& bin,
                                         int& rank, int& pop)
{
    static_assert(BPL % 4 == 0, "16-byte reads");
    const uint4* mine = reinterpret_cast<const uint4*>(hist + ((lane + ROT) & (kWave - 1)) * BPL);
    if constexpr (BPL == 4 && ROT == 0) {
        // four bins per lane: the owner's bins are in its registers, and the
        // bin within them follows from three scalar comparisons
        const uint4 v = mine[0];
        const int s = (int)((v.x + v.y) + (v.z + v.w));
        const int incl = wave_incl_scan(s);
        const int excl = incl - s;
        // excl < t <= incl as one unsigned compare (a compound predicate
        // costs the ballot two more VALU instructions)
        const uint64_t bal = __ballot((uint32_t)(t - 1 - excl) < (uint32_t)s);
        const int owner = bal ? __ffsll((unsigned long long)bal) - 1 : 0;
        const int r = t - readlane(excl, owner);
        const int h0 = readlane((int)v.x, owner), h1 = readlane((int)v.y, owner);
        const int h2 = readlane((int)v.z, owner), h3 = readlane((int)v.w, owner);
        const int c1 = h0 + h1, c2 = c1 + h2, c3 = c2 + h3;
        int sub = 0, below = 0, p = h0;
        if (r > h0 && r <= c1) { sub = 1; below = h0; p = h1; }
        else if (r > c1 && r <= c2) { sub = 2; below = c1; p = h2; }
        else if (r > c2 && r <= c3) { sub = 3; below = c2; p = h3; }
        bin = owner * BPL + sub;
        rank = r - below;
        pop = p;
        return;
    }
    int s = 0;
    // 16-byte reads, issued together; the rotated order spreads the lanes
    // over the LDS banks
#pragma unroll
    for (int q = 0; q < BPL / 4; ++q) {
        const uint4 v = mine[(q + lane) & (BPL / 4 - 1)];
        s += (int)((v.x + v.y) + (v.z + v.w));
    }
    const int incl = wave_incl_scan(s);
    const int excl = incl - s;
    const uint64_t bal = __ballot((uint32_t)(t - 1 - excl) < (uint32_t)s);
    const int owner = bal ? __ffsll((unsigned long long)bal) - 1 : 0;
    const int base = readlane(excl, owner);
    const uint32_t* own = hist + ((owner + ROT) & (kWave - 1)) * BPL;
    const int v = lane < BPL ? (int)own[lane] : 0;
    const int vi = wave_incl_scan(v);
    const int r = t - base;
    // vi - v < r <= vi; lanes >= BPL hold v = 0 and never pass
    const uint64_t b2 = __ballot((uint32_t)(r - 1 - (vi - v)) < (uint32_t)v);
    const int sub = b2 ? __ffsll((unsigned long long)b2) - 1 : 0;
    bin = owner * BPL + sub;
    rank = r - readlane(vi - v, sub);
    pop = readlane(v, sub);
}

// Destination of a row's result: the psi table, the recount queue and the
// optional per-row count output.
struct RowOut {
    const double* psi;
    unsigned* counters;
    FixEntry* fixq;
    unsigned fixcap;
    FixEntry* longq;    // nullptr: long rows go to the FP64 recount unresolved
    unsigned longcap;
    int* rc_nx;
    int* rc_ny;
    unsigned char* rc_flag;
};

__device__ __forceinline__ void queue_row(const RowOut& o, long long item, int row, uint32_t eps)
{
    const unsigned slot = atomicAdd(&o.counters[0], 1u);
    if (slot < o.fixcap) {
        FixEntry fe;
        fe.pair = (int)item;
        fe.row = row;
        fe.eps_bits = eps;
        fe.pad = 0;
        o.fixq[slot] = fe;
    }
}

// A row whose (k+1)-th joint neighbour lies beyond the sorted prefix.
__device__ __forceinline__ void queue_long(const RowOut& o, long long item, int row)
{
    if (!o.longq) {
        queue_row(o, item, row, kUnknownEps);
        return;
    }
    const unsigned slot = atomicAdd(&o.counters[2], 1u);
    if (slot < o.longcap) {
        FixEntry fe;
        fe.pair = (int)item;
        fe.row = row;
        fe.eps_bits = kUnknownEps;
        fe.pad = 0;
        o.longq[slot] = fe;
    }
}

__device__ __forceinline__ RowOut row_out(const KsgArgs& a)
{
    RowOut o;
    o.psi = a.psi;
    o.counters = a.counters;
    o.fixq = a.fixq;
    o.fixcap = a.fixcap;
    o.longq = a.longq;
    o.longcap = a.longcap;
    o.rc_nx = a.rc_nx;
    o.rc_ny = a.rc_ny;
    o.rc_flag = a.rc_flag;
    return o;
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
// lib/kraskov.cpp:131-132, 280-281), then rounded once.  j in [N, P) = 3072.
// One block per row (a, i); used where no sorted rows are kept (N >
// kSortedMaxFrames).  With sorted rows, ksg_build_sorted_wave_kernel below
// writes the same D and also
//   SD[a][i][0..64)        0 (front pad)
//   SD[a][i][64 + t]       t-th smallest value, t < N; 3072 up to the pitch
//   SM[a][i][m]            SD value t = 64m + 63 (3072 past N), m < 64
//   SI[a][i][t]            frame of the t-th smallest value, t < kPrefix (0 past N)
// so that #{j : D[a][i][j] <= x} is found with two 64-lane loads, and the
// kPrefix nearest frames of i are known in order (ties in frame order).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ksg_build_rows_kernel(const double* __restrict__ X, int N,
                                                             long long P, long long nrows,
                                                             float* __restrict__ D)
{
    // grid-stride over rows: a launch's work-item count must stay below 2^32
    for (long long row = blockIdx.x; row < nrows; row += gridDim.x) {
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
                    v[e] = __int_as_float((int)kPadBits);
                }
            }
            out[c] = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
}

// ---------------------------------------------------------------------------
// Sorted-row build (N <= kSortedMaxFrames): one block of 256 threads per row
// (a, i).  Thread t computes frames E*t .. E*t+E-1 of the row (M = 256*E >=
// N, a power of two; frames >= N hold the padding 3072) and keeps them in
// registers, so:
//   - D[a][i][j] (j < P) is written with vector stores;
//   - the row's 32-bit values are bitonic-sorted in registers: strides below
//     E inside a thread (min/max), strides up to 64E across the wave's lanes
//     (ds_bpermute), the larger ones through LDS;
//   - the kPrefix smallest (value, frame) keys, ties in frame order, are
//     selected below the kPrefix-th value T (frames at T taken in frame
//     order by a block scan) and sorted as 64-bit keys in LDS for SI.
// Outputs as described at ksg_build_rows_kernel.  Sorting values alone is
// min/max work; only the prefix needs its frames.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Sorted rows (N <= kSortedMaxFrames): D as above and SD, SM, SI, one wave
// per row (a, i): lane l holds frames l + 64 e (e < EW, coalesced loads and
// distance-row stores), N <= 64 EW.  The 64 EW keys are sorted by the register network with the
// whole row inside the wave, so every pass is in registers or a lane
// exchange (no LDS passes, no block barriers); afterwards lane l holds sorted
// positions EW l .. EW l + EW - 1.  The kPrefix smallest (value, frame) keys
// are the frames below the value T at position kPrefix - 1 and the first
// frames at T in frame order (e-major, lane-minor, counted by ballots),
// compacted to the wave's LDS and sorted as 8 keys per lane.
// ---------------------------------------------------------------------------
template <int EW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void ksg_build_sorted_wave_kernel(const double* __restrict__ X,
                                                                    int N, long long P,
                                                                    long long nrows,
                                                                    float* __restrict__ D,
                                                                    float* __restrict__ SD,
                                                                    float* __restrict__ SM,
                                                                    uint16_t* __restrict__ SI)
{
    constexpr int M = 64 * EW;
    static_assert(EW % 4 == 0 && kPrefix == 8 * kWave, "whole 16-byte groups; 8 prefix keys per lane");
    __shared__ uint64_t sp_all[4][kPrefix];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    uint64_t* sp = sp_all[w];
    const long long ps = sorted_pitch(N);
    const int nP = N < kPrefix ? N : kPrefix;
    for (long long row = (long long)blockIdx.x * 4 + w; row < nrows;
         row += (long long)gridDim.x * 4) {   // uniform over the wave
        const int a = (int)(row / N), i = (int)(row % N);
        const double* x0 = X + (size_t)(3 * a) * N;
        const double* x1 = x0 + N;
        const double* x2 = x1 + N;
        const double q0 = x0[i], q1 = x1[i], q2 = x2[i];
        uint32_t raw[EW], r[EW];
#pragma unroll
        for (int e = 0; e < EW; ++e) {
            const int j = lane + kWave * e;
            uint32_t v = kPadBits;
            if (j < N) {
                double d = fabs(q0 - x0[j]);
                d = fmax(d, fabs(q1 - x1[j]));
                d = fmax(d, fabs(q2 - x2[j]));
                v = __float_as_uint((float)d);
            }
            raw[e] = v;
            r[e] = v;
        }
        // distance row, pitch P <= M
        uint32_t* drow = reinterpret_cast<uint32_t*>(D + row * P);
#pragma unroll
        for (int e = 0; e < EW; ++e)
            if (lane + kWave * e < P) drow[lane + kWave * e] = raw[e];
        sortnet::sort_blocked<uint32_t, EW, kWave>(r, reinterpret_cast<uint32_t*>(sp), lane);
        // sorted row: 64 zeros, the N values, padding to the pitch
        uint32_t* srow = reinterpret_cast<uint32_t*>(SD) + row * ps;
        srow[lane] = 0u;
#pragma unroll
        for (int e = 0; e < EW; e += 4) {
            const int q = EW * lane + e;   // 64 + q < ps covers q .. q + 3 (ps % 64 == 0)
            if (64 + q < ps)
                *reinterpret_cast<uint4*>(srow + 64 + q) =
                    make_uint4(q < N ? r[e] : kPadBits, q + 1 < N ? r[e + 1] : kPadBits,
                               q + 2 < N ? r[e + 2] : kPadBits, q + 3 < N ? r[e + 3] : kPadBits);
        }
        for (long long q = 64 + M + lane; q < ps; q += kWave) srow[q] = kPadBits;
        // summary: sorted value 64m + 63, the last element of lane l when
        // EW l + EW - 1 = 64m + 63
        uint32_t* sm = reinterpret_cast<uint32_t*>(SM) + row * kSumPitch;
        if ((EW * lane + EW - 1) % 64 == 63) {
            const int q = EW * lane + EW - 1;
            sm[q / 64] = q < N ? r[EW - 1] : kPadBits;
        }
        for (int m = M / 64 + lane; m < kSumPitch; m += kWave) sm[m] = kPadBits;
        // the kPrefix-th smallest value T (sorted position nP - 1)
        const int pT = nP - 1;
        uint32_t tv = 0u;
#pragma unroll
        for (int e = 0; e < EW; ++e) tv = e == pT % EW ? r[e] : tv;
        const uint32_t T = (uint32_t)readlane((int)tv, pT / EW);
        // frames below T, and the first frames at T in frame order
        int LT = 0;
#pragma unroll
        for (int e = 0; e < EW; ++e) LT += __popcll(__ballot(lane + kWave * e < N && raw[e] < T));
        const int need_eq = nP - LT;
        int base = 0, eqseen = 0;
#pragma unroll
        for (int e = 0; e < EW; ++e) {
            const int j = lane + kWave * e;
            const bool real = j < N;
            const uint64_t eqm = __ballot(real && raw[e] == T);
            const bool eq = (eqm >> lane) & 1ull;
            const bool take = (real && raw[e] < T) || (eq && eqseen + mbcnt(eqm) < need_eq);
            eqseen += __popcll(eqm);
            const uint64_t tm = __ballot(take);
            if (take) sp[base + mbcnt(tm)] = ((uint64_t)raw[e] << 32) | (uint32_t)j;
            base += __popcll(tm);
        }
        for (int q = base + lane; q < kPrefix; q += kWave) sp[q] = ~0ull;
        lds_order();
        uint64_t pk[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) pk[u] = sp[8 * lane + u];
        lds_order();
        sortnet::sort_blocked<uint64_t, 8, kWave>(pk, sp, lane);
        uint32_t f[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) f[u] = 8 * lane + u < N ? (uint32_t)(pk[u] & 0xffffu) : 0u;
        *reinterpret_cast<uint4*>(SI + row * kPrefix + 8 * lane) =
            make_uint4(f[0] | (f[1] << 16), f[2] | (f[3] << 16), f[4] | (f[5] << 16),
                       f[6] | (f[7] << 16));
    }
}

// ---------------------------------------------------------------------------
// Prefix stores of the long-trajectory path from the distance rows: one block
// of 512 threads per row (a, i), thread t holding frames 16t .. 16t+15
// (P <= 8192).  A radix select over v - lo, lo and hi the row's smallest and
// largest value, 8 bits per pass from the top set bit of hi - lo, finds the
// value v* of the 1024th smallest frame.  (Binning the raw value bits would
// put a row's values, which share a few exponents, into a few bins of the
// first pass, and their LDS atomics would serialise.)  The frames below v*,
// and the first frames at v* in frame order, are the prefix (the 1024
// smallest in (value, frame) order).  They are bitonic-sorted by (value,
// frame) in registers, then
//   LV[a][i][0..64) = 0, LV[a][i][64 + t] = t-th smallest value, t < kLongPrefix
//   LI[a][i][t]     = its frame
//   LS[a][i][m]     = LV value 64m + 63
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(512) void ksg_build_prefix_kernel(const float* __restrict__ D, int N,
                                                               long long P, long long nrows,
                                                               float* __restrict__ LV,
                                                               uint16_t* __restrict__ LI,
                                                               float* __restrict__ LS)
{
    __shared__ uint64_t sk[kLongPrefix];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t wtot[8];
    __shared__ uint32_t sel[3];   // chosen bin, rank left within it, output slot counter
    __shared__ uint32_t wlo[8], whi[8];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
    const int f0 = 16 * tid;
    // grid-stride over rows: a launch's work-item count must stay below 2^32
    for (long long row = blockIdx.x; row < nrows; row += gridDim.x) {
    const uint32_t* d = reinterpret_cast<const uint32_t*>(D) + row * P;
    uint32_t v[16];
    if (f0 < P) {   // P is a multiple of 256: all 16 frames or none
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint4 w = reinterpret_cast<const uint4*>(d + f0)[c];
            v[4 * c] = w.x; v[4 * c + 1] = w.y; v[4 * c + 2] = w.z; v[4 * c + 3] = w.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = 0xffffffffu;   // above every value and the padding
    }
    // ---- value range of the real frames (padding and frames past P excluded)
    uint32_t mn = 0xffffffffu, mx = 0u;
#pragma unroll
    for (int e = 0; e < 16; ++e)
        if (f0 + e < N) {
            mn = min(mn, v[e]);
            mx = max(mx, v[e]);
        }
    mn = ~wave_max_u(~mn);
    mx = wave_max_u(mx);
    if (lane == 0) {
        wlo[wv] = mn;
        whi[wv] = mx;
    }
    __syncthreads();
    uint32_t lo = wlo[0], hi = whi[0];
#pragma unroll
    for (int u = 1; u < 8; ++u) {
        lo = min(lo, wlo[u]);
        hi = max(hi, whi[u]);
    }
    // ---- radix select of rank kLongPrefix - 1 (N > kLongPrefix) over
    //      v - lo: before the pass with top bit T, pre = (v - lo) >> T of
    //      the frames still in the running
    const int nb = hi > lo ? 32 - __builtin_clz(hi - lo) : 0;
    uint32_t pre = 0u;
    uint32_t rank = kLongPrefix - 1;
    for (int top = nb; top > 0;) {
        const int shift = max(top - 8, 0);
        if (tid < 256) hist[tid] = 0u;
        if (tid == 0) sel[2] = 0u;
        __syncthreads();
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const uint32_t dv = v[e] - lo;
            if (v[e] <= hi && (dv >> top) == pre) atomicAdd(&hist[(dv >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid < kWave) {
            const uint4 h = reinterpret_cast<const uint4*>(hist)[lane];
            const uint32_t own = h.x + h.y + h.z + h.w;
            uint32_t incl = own;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t up = (uint32_t)__shfl_up((int)incl, o, kWave);
                if (lane >= o) incl += up;
            }
            const uint32_t excl = incl - own;
            if (excl <= rank && rank < incl) {
                uint32_t r = rank - excl, bin = 4u * lane;
                if (r >= h.x) { r -= h.x; ++bin;
                    if (r >= h.y) { r -= h.y; ++bin;
                        if (r >= h.z) { r -= h.z; ++bin; } } }
                sel[0] = bin;
                sel[1] = r;
            }
        }
        __syncthreads();
        pre = (pre << (top - shift)) | sel[0];
        rank = sel[1];
        top = shift;
        __syncthreads();   // sel and hist are rewritten by the next pass
    }
    pre += lo;   // v*
    // pre = v*; the prefix holds the frames below v* and the first rank + 1
    // frames at v* in frame order: their order among ties comes from an
    // exclusive scan of the per-thread tie counts
    int eq = 0;
#pragma unroll
    for (int e = 0; e < 16; ++e) eq += v[e] == pre ? 1 : 0;
    int incl = eq;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int up = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += up;
    }
    if (lane == kWave - 1) wtot[wv] = (uint32_t)incl;
    __syncthreads();
    int tie = incl - eq;
    for (int u = 0; u < wv; ++u) tie += (int)wtot[u];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        bool take = v[e] < pre;
        if (v[e] == pre) {
            take = tie <= (int)rank;
            ++tie;
        }
        if (take) {
            const uint32_t slot = atomicAdd(&sel[2], 1u);
            if (slot < kLongPrefix) sk[slot] = ((uint64_t)v[e] << 32) | (uint32_t)(f0 + e);
        }
    }
    __syncthreads();
    // ---- sort of the kLongPrefix selected (value, frame) keys (distinct:
    //      one per frame), two per thread, in registers (sort_network.h),
    //      written back in order
    static_assert(kLongPrefix == 2 * 512, "two keys per thread of the 512-thread block");
    uint64_t pk[2] = {sk[2 * tid], sk[2 * tid + 1]};
    sortnet::sort_blocked<uint64_t, 2, 512>(pk, sk, tid);
    __syncthreads();   // the network's last LDS readers are done
    sk[2 * tid] = pk[0];
    sk[2 * tid + 1] = pk[1];
    __syncthreads();
    uint32_t* lv = reinterpret_cast<uint32_t*>(LV) + row * kLongPitch;
    for (int t = tid; t < kLongPitch; t += blockDim.x)
        lv[t] = t < 64 ? 0u : (uint32_t)(sk[t - 64] >> 32);
    for (int t = tid; t < kLongPrefix; t += blockDim.x)
        LI[row * kLongPrefix + t] = (uint16_t)(sk[t] & 0xffffu);
    if (tid < kLongSum)
        reinterpret_cast<uint32_t*>(LS)[row * kLongSum + tid] = (uint32_t)(sk[64 * tid + 63] >> 32);
    __syncthreads();   // sk is rewritten by the next row
    }
}

// ---------------------------------------------------------------------------
// Frame rows i0, i0+step, ... < i1 of one pair, one wave (N <= 256*S4).
// Returns the sum of psi[nx] + psi[ny] over the rows not queued to the FP64
// recount (uniform over the wave).  hist: the wave's 2048-word, 8 KB
// aligned LDS array.
//
// Per row: sweep A (histogram), find_bin, sweep B (sure members' marginal
// maxima; candidates of the target bin appended to LDS), select (ballot
// radix).  The marginal counts nx = #{j : dA(i,j) <= epsx} come from the
// sorted row of atom a (SD, SM): the summary gives the 64-block, the block
// gives the position; a tie at the radius shows as an equal predecessor.
// Those loads are software-pipelined: summary loads issued after row i's
// select are consumed after row i+1's sweep A, block loads issued there are
// consumed after row i+1's select, and psi loads issued then are added
// after row i+2's sweep A.
// ---------------------------------------------------------------------------
struct SortedRows {
    const float* SDa;   // sorted rows of atom a (row i at i * PS)
    const float* SDb;
    const float* SMa;   // summaries (row i at i * kSumPitch)
    const float* SMb;
    long long PS;
};

// Count and tie flag of one marginal from its sorted block pair (lanes
// 0..31: block c1-1, lanes 32..63: block c1, two values per lane).
__device__ __forceinline__ int sorted_count(uint2 v, int c1, uint32_t r, int lane, bool& tie)
{
    // lanes 32..63 only, masked on the scalar unit
    const uint64_t hi = 0xffffffff00000000ull;
    const int c2 = __popcll(__ballot(v.x <= r) & hi) + __popcll(__ballot(v.y <= r) & hi);
    const int u = 64 * c1 + c2;   // #{values <= r}, self included
    const int q = c2 + 62;        // position of the value before the last <= r
    const uint32_t prev = (uint32_t)readlane((int)((q & 1) ? v.y : v.x), q >> 1);
    tie = u >= 2 && prev == r;
    return u;
}

template <int S4>
__device__ __forceinline__ double rows_single(const float* Da, const float* Db, const SortedRows& sr,
                                              long long P, int N, int T, int i0, int i1, int step,
                                              long long item, uint32_t* hist, int lane,
                                              const RowOut& o)
{
    const int nblk = (int)(P / 256);   // 64-lane chunks of 4 frames in a row
    const uint32_t hbase = lds_addr(hist);
    uint32_t lane16 = (uint32_t)lane * 16u;
    double acc = 0.0;
    // pipeline: row r1 waits for its summary loads, row r2 for its block
    // loads, (px, py) are psi values waiting to be added
    int r1 = -1, r2 = -1;
    uint32_t ex1 = 0u, ey1 = 0u, eps1 = 0u, ex2 = 0u, ey2 = 0u, eps2 = 0u;
    uint32_t sa = 0u, sb = 0u;
    uint2 ba = make_uint2(0u, 0u), bb = ba;
    int c1a = 0, c1b = 0;
    double px = 0.0, py = 0.0;
    for (int i = i0;; i += step) {
        const bool have = i < i1;
        // keeps the lane offset a 32-bit VGPR (global_load with a scalar base)
        asm volatile("" : "+v"(lane16));
        // Bit patterns: non-negative floats order as unsigned ints.
        uint32_t a[4 * S4], b[4 * S4];
        if (have) {
            const float* ra = Da + (size_t)i * P;
            const float* rb = Db + (size_t)i * P;
            // P is a multiple of 256 frames, so whole 64-lane chunks of 4
            // frames lie in the row or past it: chunk m is skipped, by a
            // uniform test, when m >= nblk; frames [N, P) hold padding
#pragma unroll
            for (int m = 0; m < S4; ++m) {
                if (m < nblk) {
                    const uint4 va = ld_lane<uint4>(ra, lane16 + m * 1024u);
                    const uint4 vb = ld_lane<uint4>(rb, lane16 + m * 1024u);
                    a[4 * m + 0] = va.x; a[4 * m + 1] = va.y; a[4 * m + 2] = va.z; a[4 * m + 3] = va.w;
                    b[4 * m + 0] = vb.x; b[4 * m + 1] = vb.y; b[4 * m + 2] = vb.z; b[4 * m + 3] = vb.w;
                }
            }
            // ---- sweep A: fine histogram of the joint keys ----------------
            clear_hist<kBins>(hist, lane);
            wave_sync();
#pragma unroll
            for (int e = 0; e < 4 * S4; ++e)
                if (e / 4 < nblk) hist_add(hbase, a[e], b[e]);
            wave_sync();
        }
        // ---- row r1: block position from the summaries, block loads -------
        acc += px + py;
        px = py = 0.0;
        if (r1 >= 0) {
            c1a = __popcll(__ballot(sa <= ex1));
            c1b = __popcll(__ballot(sb <= ey1));
            ba = ld_lane<uint2>(sr.SDa + (size_t)r1 * sr.PS + 64 * c1a, lane16 >> 1);
            bb = ld_lane<uint2>(sr.SDb + (size_t)r1 * sr.PS + 64 * c1b, lane16 >> 1);
            r2 = r1; ex2 = ex1; ey2 = ey1; eps2 = eps1;
            r1 = -1;
        }
        uint32_t ex = 0u, ey = 0u, eps_key = kUnknownEps;
        bool queued = false;
        if (have) {
            int b1, r0, pop;
            find_bin<kBins / kWave, kBinRot>(hist, lane, T, b1, r0, pop);
            wave_sync();
            const bool unknown = b1 == 0 || b1 == kBins - 1 || pop > kWave;
            const uint32_t lo = (uint32_t)(b1 + kBinBase) << kBinShift;
            // The keys are recomputed below rather than kept in registers.
#pragma unroll
            for (int e = 0; e < 4 * S4; ++e)
                if (e / 4 < nblk) asm volatile("" : "+v"(a[e]), "+v"(b[e]));
            queued = unknown;
            if (!unknown) {
                // ---- sweep B: sure members' marginal maxima; the frames in
                //      the target bin are appended, compacted, to LDS
                uint2* cand = reinterpret_cast<uint2*>(hist);
                int nstored = 0;
#pragma unroll
                for (int e = 0; e < 4 * S4; ++e) {
                    if (e / 4 >= nblk) continue;
                    const uint32_t key = max(a[e], b[e]);
                    const bool sure = key < lo;
                    ex = max(ex, sure ? a[e] : 0u);
                    ey = max(ey, sure ? b[e] : 0u);
                    const bool inbin = key - lo < (1u << kBinShift);
                    const uint64_t cm = __ballot(inbin);
                    if (cm) {
                        if (inbin) cand[nstored + mbcnt(cm)] = make_uint2(a[e], b[e]);
                        nstored += __popcll(cm);
                    }
                }
                wave_sync();
                // ---- select: r0-th smallest of the candidates' low bits, a
                //      ballot radix that stops once one candidate is left
                const bool act = lane < pop;
                uint32_t ca = 0u, cb = 0u, low = 0xffffffffu;
                if (act) {
                    const uint2 c = cand[lane];
                    ca = c.x;
                    cb = c.y;
                    low = max(ca, cb) - lo;   // < 2^17
                }
                uint64_t live = __ballot(act);
                uint32_t pre = 0u;
                int r = r0;
                for (int bit = kBinShift - 1; bit >= 0 && __popcll(live) > 1; --bit) {
                    const uint64_t z = __ballot((low >> bit) == (pre << 1));
                    const int c0 = __popcll(z);
                    if (r > c0) {
                        r -= c0;
                        pre = (pre << 1) | 1u;
                        live &= ~z;
                    } else {
                        pre <<= 1;
                        live = z;
                    }
                }
                const uint32_t sel =
                    (uint32_t)readlane((int)low, __ffsll((unsigned long long)live) - 1);
                eps_key = lo + sel;
                const bool in = low <= sel;   // inactive lanes: low = 0xffffffff
                queued = __popcll(__ballot(in)) != r0;
                ex = wave_max_u(max(ex, in ? ca : 0u));
                ey = wave_max_u(max(ey, in ? cb : 0u));
            }
        }
        // ---- row r2: counts from the blocks, psi loads or recount ---------
        if (r2 >= 0) {
            bool tx, ty;
            const int nx = sorted_count(ba, c1a, ex2, lane, tx) - 1;
            const int ny = sorted_count(bb, c1b, ey2, lane, ty) - 1;
            const bool flag = tx || ty;
            if (!flag) {
                px = o.psi[nx];
                py = o.psi[ny];
            }
            if (lane == 0) {
                if (flag) queue_row(o, item, r2, eps2);
                if (o.rc_nx) {
                    const size_t off = (size_t)item * N + r2;
                    o.rc_nx[off] = flag ? -1 : nx;
                    o.rc_ny[off] = flag ? -1 : ny;
                    o.rc_flag[off] = 0;
                }
            }
            r2 = -1;
        }
        if (!have) break;
        // ---- row i: queue it, or issue its summary loads ----------------------
        if (queued) {
            if (lane == 0) {
                queue_row(o, item, i, eps_key);
                if (o.rc_nx) {
                    const size_t off = (size_t)item * N + i;
                    o.rc_nx[off] = -1;
                    o.rc_ny[off] = -1;
                    o.rc_flag[off] = 0;
                }
            }
        } else {
            sa = ld_lane<uint32_t>(sr.SMa + (size_t)i * kSumPitch, lane16 >> 2);
            sb = ld_lane<uint32_t>(sr.SMb + (size_t)i * kSumPitch, lane16 >> 2);
            r1 = i; ex1 = ex; ey1 = ey; eps1 = eps_key;
        }
    }
    return acc + (px + py);
}

// ---------------------------------------------------------------------------
// Sorted-prefix rows (kind 2): frame rows i0, i0+step, ... < i1 of one pair,
// one wave, at most kPrefix rows.  For row i, let P be the atom whose sorted
// row is larger at position 255 (the sparser neighbourhood) and Q the other.
// Lane l holds the sorted P values at positions 4l+c and 256+4l+c (c < 4)
// and gathers the Q distances of the same frames from Q's row (staged in
// LDS), so the wave
// sees the joint keys J = max(dP, dQ) of the kPrefix nearest frames in P.
// Every frame outside that prefix has dP >= H = SP[kPrefix - 1], hence
// J >= H: if the (k+1)-th smallest key of the prefix is below H it is the
// row's, and the prefix holds every frame at or below it.  Otherwise the
// row goes to the long-row queue (streaming path) or, without one, to the
// FP64 recount unresolved.
//   histogram  256 bins over [L, H), L = SP[k] (the (k+1)-th smallest dP, a
//              lower bound of the key), bin 0 below L, bin 255 above H;
//   sweep B    frames below the target bin are in S; its frames go to LDS;
//   select     ballot radix over the bits below the bin.
// nP = #{dP <= epsP} is counted in registers (epsP < H), nQ from Q's sorted
// row (summary, then one block, consumed one row later).  The next row's
// prefix, Q summary and whole Q row are loaded while a row is processed.
// All LDS traffic is the wave's own region, ordered by lds_order().
// ---------------------------------------------------------------------------
constexpr int kPBins = 256;

struct PrefixRows {
    const uint32_t* Da;    // distance rows of atom a (row i at i * P)
    const uint32_t* Db;
    const uint32_t* SDa;   // sorted rows (row i at i * PS, value t at 64 + t)
    const uint32_t* SDb;
    const uint32_t* SMa;   // summaries (row i at i * kSumPitch)
    const uint32_t* SMb;
    const uint16_t* SIa;   // prefix frame indices (row i at i * kPrefix)
    const uint16_t* SIb;
    long long P, PS;
};

// Per-wave LDS of the sorted-prefix path: histogram / candidate slots, then
// the Q distance row of the current frame.
struct PrefixLds {
    uint32_t hist[kPBins];
    uint32_t qrow[kSortedMaxFrames];
};

__device__ __forceinline__ double rows_prefix(const PrefixRows& pr, int N, int T, int i0, int i1,
                                              int step, long long item, PrefixLds& lds, int lane,
                                              const RowOut& o)
{
    // R <= kPrefix holds for every caller (N <= kSortedMaxFrames rows split
    // in nsplit ranges, or stepped by 4 in the batch kernel); the side
    // bits cover kPrefix rows.
    const int R = min((i1 - i0 + step - 1) / step, kPrefix);
    const int nblk = (int)(pr.P / 256);   // 64-lane chunks of 4 frames in a row
    uint32_t* hist = lds.hist;
    const uint32_t hbase = lds_addr(hist);
    uint32_t lane4 = (uint32_t)lane * 4u;
    // side bits: bit u of lane l is set when row t = l + 64u takes atom a as P
    uint32_t sidebits = 0u;
#pragma unroll
    for (int u = 0; u < kPrefix / kWave; ++u) {
        const int t = lane + kWave * u;
        if (t < R) {
            const int row = i0 + t * step;   // < N: 32-bit offsets below
            const uint32_t sa = pr.SMa[(uint32_t)row * (uint32_t)kSumPitch + 3];
            const uint32_t sb = pr.SMb[(uint32_t)row * (uint32_t)kSumPitch + 3];
            sidebits |= (sa >= sb ? 1u : 0u) << u;
        }
    }
    auto side_of = [&](int t) -> bool {
        return ((uint32_t)readlane((int)sidebits, t & (kWave - 1)) >> (t >> 6)) & 1u;
    };
    // A prefetched row: P values and frame indices, SP[T-1], Q summary.  Two
    // sets, A and B, taken in turn: row t is read from one while row t + 1
    // loads into the other, so no row's registers are copied.  The Q row
    // (nq) is staged to LDS before the next one loads: one set.
    struct NextRow {
        uint4 v0, v1, i;
        uint32_t sm, L;
    };
    NextRow rA, rB;
    rA.v0 = make_uint4(0u, 0u, 0u, 0u);
    rA.v1 = rA.i = rA.v0;
    rA.sm = rA.L = 0u;
    rB = rA;
    uint4 nq[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) nq[m] = rA.v0;
    // Each load takes the atom's base pointer (selected by the side, held in
    // scalar registers) and a 32-bit byte offset holding the row: one vector
    // add per load instead of 64-bit scalar pointer arithmetic per row (the
    // offsets stay below 2^32: rows < 2048, pitches < 2300 words).
    auto issue = [&](int t, NextRow& n) {
        const uint32_t row = (uint32_t)(i0 + t * step);   // < N
        const bool sa = side_of(t);
        const uint32_t* sd = sa ? pr.SDa : pr.SDb;
        const uint32_t so = (row * (uint32_t)pr.PS + 64u) * 4u;
        n.v0 = ld_lane<uint4>(sd, so + lane4 * 4u);
        n.v1 = ld_lane<uint4>(sd, so + lane4 * 4u + 1024u);
        const uint16_t* si = sa ? pr.SIa : pr.SIb;
        const uint32_t io = row * (uint32_t)kPrefix * 2u;
        const uint2 i01 = ld_lane<uint2>(si, io + lane4 * 2u);
        const uint2 i23 = ld_lane<uint2>(si, io + lane4 * 2u + 512u);
        n.i = make_uint4(i01.x, i01.y, i23.x, i23.y);
        if (lane == 0) n.L = ld_lane<uint32_t>(sd, so + (uint32_t)(T - 1) * 4u);
        n.sm = ld_lane<uint32_t>(sa ? pr.SMb : pr.SMa, row * (uint32_t)kSumPitch * 4u + lane4);
        const uint32_t* rq = sa ? pr.Db : pr.Da;
        const uint32_t qo = row * (uint32_t)pr.P * 4u + lane4 * 4u;
        // all 8 chunks: past the row's last chunk the last one again (a
        // row of P < 2048 floats), staged to LDS slots no gather reads
#pragma unroll
        for (int m = 0; m < 8; ++m)
            nq[m] = ld_lane<uint4>(rq, qo + (uint32_t)min(m, nblk - 1) * 1024u);
    };
    double acc = 0.0, px = 0.0, py = 0.0;
    // the previous row, waiting for its Q block
    int prow = -1;
    bool pside = false;
    uint32_t ptv = 0u, pexp = 0u;   // value before the previous row's radius, and the radius
    int pnp = 0, pc1 = 0;
    uint32_t pexq = 0u, peps = 0u;
    uint2 pblk = make_uint2(0u, 0u);
    // one row: false to go on, true past the last row
    // the previous row: Q count from its block, psi loads
    auto finish_prev = [&]() {
        acc += px + py;
        px = py = 0.0;
        if (prow >= 0) {
            bool tq;
            const int nq_ = sorted_count(pblk, pc1, pexq, lane, tq) - 1;
            const int np = pnp - 1;
            const bool ptie = pnp >= 2 && (uint32_t)readlane((int)ptv, 0) == pexp;
            const bool flag = tq || ptie;
            const int nx = pside ? np : nq_, ny = pside ? nq_ : np;
            if (!flag && lane == 0) {   // the sum is kept on lane 0
                px = o.psi[nx];
                py = o.psi[ny];
            }
            if (lane == 0) {
                if (flag) queue_row(o, item, (int)prow, peps);
                if (o.rc_nx) {
                    const size_t off = (size_t)item * N + prow;
                    o.rc_nx[off] = flag ? -1 : nx;
                    o.rc_ny[off] = flag ? -1 : ny;
                    o.rc_flag[off] = 0;
                }
            }
            prow = -1;
        }
    };
    // one row: false to go on, true past the last row
    auto one_row = [&](int t, NextRow& cur, NextRow& nxt) -> bool {
        if (t >= R) {
            finish_prev();
            return true;
        }
        asm volatile("" : "+v"(lane4));
        const int row = i0 + t * step;
        const bool sideA = side_of(t);
        const uint32_t v[8] = {cur.v0.x, cur.v0.y, cur.v0.z, cur.v0.w,
                               cur.v1.x, cur.v1.y, cur.v1.z, cur.v1.w};
        const uint32_t idx[4] = {cur.i.x, cur.i.y, cur.i.z, cur.i.w};
        const uint32_t sm = cur.sm;
        const uint32_t L = (uint32_t)readlane((int)cur.L, 0);
        // stage the Q row (frames 256m + 4l + c at lane l, chunk m)
#pragma unroll
        for (int m = 0; m < 8; ++m) reinterpret_cast<uint4*>(lds.qrow)[m * kWave + lane] = nq[m];
        lds_order();
        uint32_t q[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) q[e] = lds.qrow[(idx[e >> 1] >> (16 * (e & 1))) & 0xffffu];
        if (t + 1 < R) issue(t + 1, nxt);
        finish_prev();
        // ---- histogram of the prefix's joint keys over [L, H) --------------
        const uint32_t H = min((uint32_t)readlane((int)v[7], kWave - 1), kPadBits);
        // smallest sh with ((H - L - 1) >> sh) <= kPBins - 3
        const uint32_t span1 = H > L ? H - L - 1u : 0u;
        int sh = max(0, 24 - (int)__builtin_clz(span1 | 1u));
        if ((span1 >> sh) > (uint32_t)(kPBins - 3)) ++sh;
        const uint32_t Lm = L - min(L, 1u << sh);
        // Bin 0 (keys below B1) and the keys at or above H are counted apart:
        // most keys fall there, and atomics on one address serialise.  Bin 0
        // takes the count of the keys below B1; bin 255 a count that reaches
        // T, so that a target beyond H is reported as bin 255.
        const uint32_t B1 = Lm + (1u << sh);
        int cb = 0;
#pragma unroll
        for (int e = 0; e < 8; ++e) cb += __popcll(__ballot(max(v[e], q[e]) < B1));
        reinterpret_cast<uint4*>(hist)[lane] =
            make_uint4(lane == 0 ? (uint32_t)cb : 0u, 0u, 0u, lane == kWave - 1 ? 4096u : 0u);
        lds_order();
        // one unsigned compare: keys below B1 wrap above the span
        const uint32_t span = H > B1 ? H - B1 : 0u;
        const uint32_t hbin1 = hbase + 4u;   // address of bin 1
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t dk = max(v[e], q[e]) - B1;
            if (dk < span)   // bin (dk >> sh) + 1, in 1 .. 254
                __hip_atomic_fetch_add((lds_u32*)(uintptr_t)(((dk >> sh) << 2) + hbin1), 1u,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
        lds_order();
        int b1, r0, pop;
        find_bin<kPBins / kWave>(hist, lane, T, b1, r0, pop);
        lds_order();
        bool queued = b1 == 0 || b1 == kPBins - 1 || pop > kWave;
        uint32_t eps_key = kUnknownEps, ex = 0u, ey = 0u;
        if (!queued) {
            // ---- sweep B: sure members' marginal maxima; the target bin's
            //      frames appended, compacted, to LDS
            const uint32_t lo = Lm + ((uint32_t)b1 << sh);
            const uint32_t width = 1u << sh;
            // the bin's keys as counted, [lo, min(lo + width, H)): one
            // unsigned compare (keys below lo wrap)
            const uint32_t lim = min(width, H > lo ? H - lo : 0u);
            uint2* cand = reinterpret_cast<uint2*>(hist);
            int nstored = 0;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t key = max(v[e], q[e]);
                const bool sure = key < lo;
                ex = max(ex, sure ? v[e] : 0u);
                ey = max(ey, sure ? q[e] : 0u);
                const bool inbin = key - lo < lim;
                const uint64_t cm = __ballot(inbin);
                if (cm) {
                    if (inbin) cand[nstored + mbcnt(cm)] = make_uint2(v[e], q[e]);
                    nstored += __popcll(cm);
                }
            }
            lds_order();
            // ---- select: r0-th smallest of the candidates' offsets in the bin
            const bool act = lane < pop;
            uint32_t ca = 0u, cb = 0u, low = 0xffffffffu;
            if (act) {
                const uint2 c = cand[lane];
                ca = c.x;
                cb = c.y;
                low = max(ca, cb) - lo;   // < width
            }
            uint64_t live = __ballot(act);
            uint32_t pre = 0u;
            int r = r0;
            for (int bit = sh - 1; bit >= 0 && __popcll(live) > 1; --bit) {
                const uint64_t z = __ballot((low >> bit) == (pre << 1));
                const int c0 = __popcll(z);
                if (r > c0) {
                    r -= c0;
                    pre = (pre << 1) | 1u;
                    live &= ~z;
                } else {
                    pre <<= 1;
                    live = z;
                }
            }
            const uint32_t sel = (uint32_t)readlane((int)low, __ffsll((unsigned long long)live) - 1);
            eps_key = lo + sel;
            const bool in = low <= sel;   // inactive lanes: low = 0xffffffff
            queued = __popcll(__ballot(in)) != r0 || eps_key >= H;
            ex = wave_max_u(max(ex, in ? ca : 0u));
            ey = wave_max_u(max(ey, in ? cb : 0u));
            lds_order();
        }
        if (queued) {
            // beyond the prefix: the target is the overflow bin or the key
            // is not below H; otherwise an FP32 tie or a crowded bin
            const bool beyond = b1 == kPBins - 1 || (eps_key != kUnknownEps && eps_key >= H);
            if (lane == 0) {
                if (beyond)
                    queue_long(o, item, (int)row);
                else
                    queue_row(o, item, (int)row, eps_key);
                if (o.rc_nx && !(beyond && o.longq)) {
                    const size_t off = (size_t)item * N + row;
                    o.rc_nx[off] = -1;
                    o.rc_ny[off] = -1;
                    o.rc_flag[off] = 0;
                }
            }
            return false;
        }
        // ---- nP from the prefix in registers (epsP < H); Q block load -------
        // one compare per value into a lane mask, counted on the scalar unit
        int c = 0;
#pragma unroll
        for (int e = 0; e < 8; ++e) c += __popcll(__ballot(v[e] <= ex));
        pnp = c;
        // ex is the sorted prefix's value at position c - 1; a tie at it
        // shows as an equal value at position c - 2 (compared one row later)
        const uint32_t ro = (uint32_t)row * (uint32_t)pr.PS * 4u;   // byte offset of the row
        ptv = ld_lane<uint32_t>(sideA ? pr.SDa : pr.SDb, ro + 256u + (uint32_t)max(c - 2, 0) * 4u);
        pexp = ex;
        pc1 = __popcll(__ballot(sm <= ey));
        pblk = ld_lane<uint2>(sideA ? pr.SDb : pr.SDa, ro + (uint32_t)pc1 * 256u + lane4 * 2u);
        prow = row;
        pside = sideA;
        pexq = ey;
        peps = eps_key;
        return false;
    };
    if (R > 0) issue(0, rA);
    for (int t = 0;; t += 2) {
        if (one_row(t, rA, rB)) break;
        if (one_row(t + 1, rB, rA)) break;
    }
    return acc + (px + py);
}

// ---------------------------------------------------------------------------
// Long-trajectory prefix rows (kind 3, kSortedMaxFrames < N <= kMaxFrames):
// the sorted-prefix method of rows_prefix over the kLongPrefix prefix stores.
// Lane l holds P's prefix positions 256c + 4l + e (c, e < 4, 16 per lane) and
// gathers Q's distances of those frames straight from Q's distance row.
// nQ = #{dQ <= epsQ} comes from Q's prefix store (its 16-value summary picks
// a 128-value block) when epsQ is below Q's 1024th value, else from a
// streaming count over Q's row; either count is consumed one row later.
// ---------------------------------------------------------------------------
struct LongRows {
    const uint32_t* Da;    // distance rows (row i at i * P)
    const uint32_t* Db;
    const uint32_t* LVa;   // prefix values (row i at i * kLongPitch, value t at 64 + t)
    const uint32_t* LVb;
    const uint16_t* LIa;   // prefix frames (row i at i * kLongPrefix)
    const uint16_t* LIb;
    const uint32_t* LSa;   // prefix summaries (row i at i * kLongSum)
    const uint32_t* LSb;
    long long P;
};

__device__ __forceinline__ LongRows long_rows(const float* D, const float* LV, const uint16_t* LI,
                                              const float* LS, long long P, int N, int a1, int a2)
{
    LongRows lr;
    lr.P = P;
    lr.Da = reinterpret_cast<const uint32_t*>(D + (size_t)a1 * N * P);
    lr.Db = reinterpret_cast<const uint32_t*>(D + (size_t)a2 * N * P);
    lr.LVa = reinterpret_cast<const uint32_t*>(LV + (size_t)a1 * N * kLongPitch);
    lr.LVb = reinterpret_cast<const uint32_t*>(LV + (size_t)a2 * N * kLongPitch);
    lr.LIa = LI + (size_t)a1 * N * kLongPrefix;
    lr.LIb = LI + (size_t)a2 * N * kLongPrefix;
    lr.LSa = reinterpret_cast<const uint32_t*>(LS + (size_t)a1 * N * kLongSum);
    lr.LSb = reinterpret_cast<const uint32_t*>(LS + (size_t)a2 * N * kLongSum);
    return lr;
}

template <int C>
__device__ __forceinline__ double rows_long(const LongRows& lr, int N, int T, int i0, int i1,
                                            int step, long long item, uint32_t* hist, int lane,
                                            const RowOut& o)
{
    // Callers pass at most kPrefix rows (the side bits cover kPrefix rows).
    const int R = min((i1 - i0 + step - 1) / step, kPrefix);
    const int nblk = (int)(lr.P / 256);
    const uint32_t hbase = lds_addr(hist);
    uint32_t lane4 = (uint32_t)lane * 4u;
    uint32_t sidebits = 0u;
#pragma unroll
    for (int u = 0; u < kPrefix / kWave; ++u) {
        const int t = lane + kWave * u;
        if (t < R) {
            const int row = i0 + t * step;   // < N: 32-bit offsets below
            const uint32_t sa = lr.LVa[(uint32_t)row * (uint32_t)kLongPitch + 64 + 255];
            const uint32_t sb = lr.LVb[(uint32_t)row * (uint32_t)kLongPitch + 64 + 255];
            sidebits |= (sa >= sb ? 1u : 0u) << u;
        }
    }
    auto side_of = [&](int t) -> bool {
        return ((uint32_t)readlane((int)sidebits, t & (kWave - 1)) >> (t >> 6)) & 1u;
    };
    // the next row: P prefix values and frames, SP[T-1], Q summary
    uint4 nv[C];
    uint2 ni[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        nv[c] = make_uint4(0u, 0u, 0u, 0u);
        ni[c] = make_uint2(0u, 0u);
    }
    uint32_t nL = 0u, nls = 0xffffffffu;
    auto issue = [&](int t) {
        const int row = i0 + t * step;   // < N: 32-bit offsets below
        const bool sa = side_of(t);
        const uint32_t* base = (sa ? lr.LVa : lr.LVb) + (uint32_t)row * (uint32_t)kLongPitch + 64;
        const uint16_t* ib = (sa ? lr.LIa : lr.LIb) + (uint32_t)row * (uint32_t)kLongPrefix;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            nv[c] = ld_lane<uint4>(base, lane4 * 4u + c * 1024u);
            ni[c] = ld_lane<uint2>(ib, lane4 * 2u + c * 512u);
        }
        if (lane == 0) nL = ld_lane<uint32_t>(base, (uint32_t)(T - 1) * 4u);
        if (lane < kLongSum)
            nls = ld_lane<uint32_t>((sa ? lr.LSb : lr.LSa) + (uint32_t)row * (uint32_t)kLongSum,
                                    lane4);
    };
    double acc = 0.0, px = 0.0, py = 0.0;
    // the previous row: its P count, and either its Q block (pending) or its
    // Q count from the streaming path (ready)
    int prow = -1;
    bool pside = false, pready = false, pqtie = false;
    uint32_t ptv = 0u, pexp = 0u;   // value before the previous row's radius, and the radius
    int pnp = 0, pc1 = 0, pnq = 0;
    uint32_t pexq = 0u, peps = 0u;
    uint2 pblk = make_uint2(0u, 0u);
    if (R > 0) issue(0);
    for (int t = 0;; ++t) {
        const bool have = t < R;
        asm volatile("" : "+v"(lane4));
        uint32_t v[4 * C], q[4 * C];
        uint32_t lsq = 0xffffffffu, L = 0u;
        bool sideA = false;
        int row = 0;
        const uint32_t* rq = nullptr;
        if (have) {
            row = i0 + t * step;
            sideA = side_of(t);
            uint32_t idx[2 * C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                v[4 * c + 0] = nv[c].x; v[4 * c + 1] = nv[c].y;
                v[4 * c + 2] = nv[c].z; v[4 * c + 3] = nv[c].w;
                idx[2 * c] = ni[c].x;
                idx[2 * c + 1] = ni[c].y;
            }
            lsq = nls;
            L = (uint32_t)readlane((int)nL, 0);
            rq = (sideA ? lr.Db : lr.Da) + (uint32_t)row * (uint32_t)lr.P;
#pragma unroll
            for (int e = 0; e < 4 * C; ++e)
                q[e] = ld_lane<uint32_t>(rq, ((idx[e >> 1] >> (16 * (e & 1))) & 0xffffu) * 4u);
            if (t + 1 < R) issue(t + 1);
        }
        // ---- previous row: Q count, psi loads ------------------------------
        acc += px + py;
        px = py = 0.0;
        if (prow >= 0) {
            bool tq = pqtie;
            const int nq_ = (pready ? pnq : sorted_count(pblk, pc1, pexq, lane, tq)) - 1;
            const int np = pnp - 1;
            const bool ptie = pnp >= 2 && (uint32_t)readlane((int)ptv, 0) == pexp;
            const bool flag = tq || ptie;
            const int nx = pside ? np : nq_, ny = pside ? nq_ : np;
            if (!flag && lane == 0) {
                px = o.psi[nx];
                py = o.psi[ny];
            }
            if (lane == 0) {
                if (flag) queue_row(o, item, (int)prow, peps);
                if (o.rc_nx) {
                    const size_t off = (size_t)item * N + prow;
                    o.rc_nx[off] = flag ? -1 : nx;
                    o.rc_ny[off] = flag ? -1 : ny;
                    o.rc_flag[off] = 0;
                }
            }
            prow = -1;
        }
        if (!have) break;
        // ---- histogram of the prefix's joint keys over [L, H) --------------
        const uint32_t H = min((uint32_t)readlane((int)v[4 * C - 1], kWave - 1), kPadBits);
        const uint32_t span1 = H > L ? H - L - 1u : 0u;
        int sh = max(0, 24 - (int)__builtin_clz(span1 | 1u));
        if ((span1 >> sh) > (uint32_t)(kPBins - 3)) ++sh;
        const uint32_t Lm = L - min(L, 1u << sh);
        const uint32_t B1 = Lm + (1u << sh);
        if constexpr (C == 2) {
            // as rows_prefix: bin 0 and the keys at or above H counted apart
            int cb = 0;
#pragma unroll
            for (int e = 0; e < 4 * C; ++e) cb += __popcll(__ballot(max(v[e], q[e]) < B1));
            reinterpret_cast<uint4*>(hist)[lane] =
                make_uint4(lane == 0 ? (uint32_t)cb : 0u, 0u, 0u, lane == kWave - 1 ? 4096u : 0u);
            lds_order();
            const uint32_t span = H > B1 ? H - B1 : 0u;
            const uint32_t hbin1 = hbase + 4u;   // address of bin 1
#pragma unroll
            for (int e = 0; e < 4 * C; ++e) {
                const uint32_t dk = max(v[e], q[e]) - B1;
                if (dk < span)   // bin (dk >> sh) + 1, in 1 .. 254
                    __hip_atomic_fetch_add((lds_u32*)(uintptr_t)(((dk >> sh) << 2) + hbin1), 1u,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
        } else {
            // 1024 frames: few keys reach H, and every key is counted (the
            // split above measured 28% slower here)
            reinterpret_cast<uint4*>(hist)[lane] = make_uint4(0u, 0u, 0u, 0u);
            lds_order();
#pragma unroll
            for (int e = 0; e < 4 * C; ++e) {
                const uint32_t key = max(v[e], q[e]);
                const uint32_t bin = min(__builtin_elementwise_sub_sat(key, Lm) >> sh,
                                         (uint32_t)(kPBins - 1));
                __hip_atomic_fetch_add((lds_u32*)(uintptr_t)((bin << 2) | hbase), 1u,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
        }
        lds_order();
        int b1, r0, pop;
        find_bin<kPBins / kWave>(hist, lane, T, b1, r0, pop);
        lds_order();
        bool queued = b1 == 0 || b1 == kPBins - 1 || pop > kWave;
        uint32_t eps_key = kUnknownEps, ex = 0u, ey = 0u;
        if (!queued) {
            const uint32_t lo = Lm + ((uint32_t)b1 << sh);
            const uint32_t width = 1u << sh;
            const uint32_t lim = C == 2 ? min(width, H > lo ? H - lo : 0u) : width;
            uint2* cand = reinterpret_cast<uint2*>(hist);
            int nstored = 0;
#pragma unroll
            for (int e = 0; e < 4 * C; ++e) {
                const uint32_t key = max(v[e], q[e]);
                const bool sure = key < lo;
                ex = max(ex, sure ? v[e] : 0u);
                ey = max(ey, sure ? q[e] : 0u);
                // as counted (keys at or above H are not, with C == 2)
                const bool inbin = key - lo < lim;
                const uint64_t cm = __ballot(inbin);
                if (cm) {
                    if (inbin) cand[nstored + mbcnt(cm)] = make_uint2(v[e], q[e]);
                    nstored += __popcll(cm);
                }
            }
            lds_order();
            const bool act = lane < pop;
            uint32_t ca = 0u, cb = 0u, low = 0xffffffffu;
            if (act) {
                const uint2 c = cand[lane];
                ca = c.x;
                cb = c.y;
                low = max(ca, cb) - lo;
            }
            uint64_t live = __ballot(act);
            uint32_t pre = 0u;
            int r = r0;
            for (int bit = sh - 1; bit >= 0 && __popcll(live) > 1; --bit) {
                const uint64_t z = __ballot((low >> bit) == (pre << 1));
                const int c0 = __popcll(z);
                if (r > c0) {
                    r -= c0;
                    pre = (pre << 1) | 1u;
                    live &= ~z;
                } else {
                    pre <<= 1;
                    live = z;
                }
            }
            const uint32_t sel = (uint32_t)readlane((int)low, __ffsll((unsigned long long)live) - 1);
            eps_key = lo + sel;
            const bool in = low <= sel;   // inactive lanes: low = 0xffffffff
            queued = __popcll(__ballot(in)) != r0 || eps_key >= H;
            ex = wave_max_u(max(ex, in ? ca : 0u));
            ey = wave_max_u(max(ey, in ? cb : 0u));
            lds_order();
        }
        if (queued) {
            const bool beyond = b1 == kPBins - 1 || (eps_key != kUnknownEps && eps_key >= H);
            if (lane == 0) {
                if (beyond)
                    queue_long(o, item, (int)row);
                else
                    queue_row(o, item, (int)row, eps_key);
                if (o.rc_nx && !(beyond && o.longq)) {
                    const size_t off = (size_t)item * N + row;
                    o.rc_nx[off] = -1;
                    o.rc_ny[off] = -1;
                    o.rc_flag[off] = 0;
                }
            }
            continue;
        }
        // ---- nP from the prefix in registers; nQ from Q's prefix store or
        //      a streaming count over Q's row ---------------------------------
        int c = 0;
#pragma unroll
        for (int e = 0; e < 4 * C; ++e) c += __popcll(__ballot(v[e] <= ex));
        pnp = c;
        // a tie at ex shows as an equal value at prefix position c - 2
        ptv = ld_lane<uint32_t>((sideA ? lr.LVa : lr.LVb) + (uint32_t)row * (uint32_t)kLongPitch + 64,
                                (uint32_t)max(c - 2, 0) * 4u);
        pexp = ex;
        const uint32_t qtop = (uint32_t)readlane((int)lsq, kLongSum - 1);
        if (ey < qtop) {
            pready = false;
            pqtie = false;
            pc1 = __popcll(__ballot(lsq <= ey) & ((1ull << kLongSum) - 1));
            const uint32_t* sq = (sideA ? lr.LVb : lr.LVa) + (uint32_t)row * (uint32_t)kLongPitch + 64 * pc1;
            pblk = ld_lane<uint2>(sq, lane4 * 2u);
        } else {
            // frames [N, P) hold padding above every distance
            int cq = 0;
            for (int m = 0; m < nblk; ++m) {
                const uint4 w = ld_lane<uint4>(rq, lane4 * 4u + (uint32_t)m * 1024u);
                cq += (w.x <= ey ? 1 : 0) + (w.x == ey ? 0x10000 : 0);
                cq += (w.y <= ey ? 1 : 0) + (w.y == ey ? 0x10000 : 0);
                cq += (w.z <= ey ? 1 : 0) + (w.z == ey ? 0x10000 : 0);
                cq += (w.w <= ey ? 1 : 0) + (w.w == ey ? 0x10000 : 0);
            }
            cq = wave_sum(cq);
            pready = true;
            pnq = cq & 0xffff;
            pqtie = (cq >> 16) > 1;
        }
        prow = row;
        pside = sideA;
        pexq = ey;
        peps = eps_key;
    }
    return acc + (px + py);
}

// ---------------------------------------------------------------------------
// One frame row processed by the 4 waves of a block together, wave w holding
// frames [w*256*S4, (w+1)*256*S4) (N <= 1024*S4).  Same three sweeps with a
// block-shared histogram; wave 0 resolves the boundary candidates.  Five
// barriers per row.  Returns the row's psi sum on thread 0.
// ---------------------------------------------------------------------------
struct CoopLds {
    uint32_t hist[kBins];                 // histogram, then candidate slots
    uint32_t dense[kPairWaves][kWave][2]; // per-wave compacted candidates
    int nc[kPairWaves], over[kPairWaves], cx[kPairWaves], cy[kPairWaves];
    float ex[kPairWaves], ey[kPairWaves];
    float rex, rey;
    int rflag;
    uint32_t reps;
};

template <int S4>
__device__ __forceinline__ double row_coop(const float* Da, const float* Db, long long P, int N,
                                           int T, int i, long long item, CoopLds& sh, int lane,
                                           int w, const RowOut& o)
{
    const int nchunk = (N + 3) / 4;
    uint32_t* hist = sh.hist;
    uint32_t* slots = hist + w * 512;
    const float4* ra = reinterpret_cast<const float4*>(Da + (size_t)i * P);
    const float4* rb = reinterpret_cast<const float4*>(Db + (size_t)i * P);
    float a[4 * S4], b[4 * S4];
#pragma unroll
    for (int m = 0; m < S4; ++m) {
        const int c = (w * S4 + m) * kWave + lane;
        float4 va = make_float4(INFINITY, INFINITY, INFINITY, INFINITY), vb = va;
        if (c < nchunk) {
            va = ra[c];
            vb = rb[c];
        }
        a[4 * m + 0] = va.x; a[4 * m + 1] = va.y; a[4 * m + 2] = va.z; a[4 * m + 3] = va.w;
        b[4 * m + 0] = vb.x; b[4 * m + 1] = vb.y; b[4 * m + 2] = vb.z; b[4 * m + 3] = vb.w;
    }
    reinterpret_cast<uint4*>(hist)[threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
    reinterpret_cast<uint4*>(hist)[threadIdx.x + 256] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 4 * S4; ++e) {
        const uint32_t key = __float_as_uint(fmaxf(a[e], b[e]));
        const int bin = min(max((int)(key >> kBinShift) - kBinBase, 0), kBins - 1);
        atomicAdd(&hist[bin], 1u);
    }
    __syncthreads();
    int b1, r0, pop;
    find_bin<kBins / kWave>(hist, lane, T, b1, r0, pop);
    const bool edge = b1 == 0 || b1 == kBins - 1 || pop > kWave;
    const uint32_t lo = (uint32_t)(b1 + kBinBase) << kBinShift;
    __syncthreads();

    float ex = 0.0f, ey = 0.0f;
    int nc = 0;
#pragma unroll
    for (int e = 0; e < 4 * S4; ++e) {
        const uint32_t key = __float_as_uint(fmaxf(a[e], b[e]));
        const bool sure = key < lo;
        ex = fmaxf(ex, sure ? a[e] : 0.0f);
        ey = fmaxf(ey, sure ? b[e] : 0.0f);
        if (key - lo < (1u << kBinShift)) {
            if (nc < kCandCap) {
                slots[(nc * kWave + lane) * 2 + 0] = __float_as_uint(a[e]);
                slots[(nc * kWave + lane) * 2 + 1] = __float_as_uint(b[e]);
            }
            ++nc;
        }
    }
    const int ncmax = wave_max_i(nc);
    const int ncw = wave_sum(nc);
    if (ncmax <= kCandCap && !edge) {
        const int off = wave_incl_scan(nc) - nc;
#pragma unroll
        for (int t = 0; t < kCandCap; ++t)
            if (t < nc && off + t < kWave) {
                sh.dense[w][off + t][0] = slots[(t * kWave + lane) * 2 + 0];
                sh.dense[w][off + t][1] = slots[(t * kWave + lane) * 2 + 1];
            }
    }
    ex = wave_max(ex);
    ey = wave_max(ey);
    if (lane == 0) {
        sh.nc[w] = ncw;
        sh.over[w] = ncmax > kCandCap;
        sh.ex[w] = ex;
        sh.ey[w] = ey;
    }
    __syncthreads();

    if (w == 0) {
        const bool unknown = edge || sh.over[0] || sh.over[1] || sh.over[2] || sh.over[3];
        float rex = fmaxf(fmaxf(sh.ex[0], sh.ex[1]), fmaxf(sh.ex[2], sh.ex[3]));
        float rey = fmaxf(fmaxf(sh.ey[0], sh.ey[1]), fmaxf(sh.ey[2], sh.ey[3]));
        uint32_t eps_key = kUnknownEps;
        bool sel_tie = false;
        if (!unknown) {
            const int c0 = sh.nc[0], c1 = c0 + sh.nc[1], c2 = c1 + sh.nc[2];
            const bool act = lane < pop;
            float ca = 0.0f, cb = 0.0f;
            uint32_t low = 0xffffffffu;
            if (act) {
                const int q = lane < c0 ? 0 : lane < c1 ? 1 : lane < c2 ? 2 : 3;
                const int l = lane - (q == 0 ? 0 : q == 1 ? c0 : q == 2 ? c1 : c2);
                ca = __uint_as_float(sh.dense[q][l][0]);
                cb = __uint_as_float(sh.dense[q][l][1]);
                low = __float_as_uint(fmaxf(ca, cb)) - lo;
            }
            uint32_t sel = 0;
            int r = r0;
            for (int bit = kBinShift - 1; bit >= 0; --bit) {
                const bool zero = act && (low >> (bit + 1)) == sel && ((low >> bit) & 1u) == 0u;
                const int cz = __popcll(__ballot(zero));
                sel <<= 1;
                if (r > cz) {
                    r -= cz;
                    sel |= 1u;
                }
            }
            eps_key = lo + sel;
            const bool in = low <= sel;   // inactive lanes: low = 0xffffffff
            sel_tie = __popcll(__ballot(in)) != r0;
            rex = fmaxf(rex, wave_max(in ? ca : 0.0f));
            rey = fmaxf(rey, wave_max(in ? cb : 0.0f));
        }
        if (lane == 0) {
            sh.rex = rex;
            sh.rey = rey;
            sh.rflag = (unknown || sel_tie) ? 1 : 0;
            sh.reps = eps_key;
        }
    }
    __syncthreads();
    ex = sh.rex;
    ey = sh.rey;
    int cx = 0, cy = 0;
#pragma unroll
    for (int e = 0; e < 4 * S4; ++e) {
        cx += (a[e] <= ex) + ((a[e] == ex) << 16);
        cy += (b[e] <= ey) + ((b[e] == ey) << 16);
    }
    cx = wave_sum(cx);
    cy = wave_sum(cy);
    if (lane == 0) {
        sh.cx[w] = cx;
        sh.cy[w] = cy;
    }
    __syncthreads();
    double contrib = 0.0;
    if (threadIdx.x == 0) {
        const int tx = sh.cx[0] + sh.cx[1] + sh.cx[2] + sh.cx[3];
        const int ty = sh.cy[0] + sh.cy[1] + sh.cy[2] + sh.cy[3];
        const int nx = (tx & 0xffff) - 1, ny = (ty & 0xffff) - 1;
        const bool flag = sh.rflag || (tx >> 16) > 1 || (ty >> 16) > 1;
        if (flag)
            queue_row(o, item, i, sh.reps);
        else
            contrib = o.psi[nx] + o.psi[ny];
        if (o.rc_nx) {
            const size_t off = (size_t)item * N + i;
            o.rc_nx[off] = flag ? -1 : nx;
            o.rc_ny[off] = flag ? -1 : ny;
            o.rc_flag[off] = 0;
        }
    }
    return contrib;
}

// Work item of virtual block vb = vb0 + blockIdx.x (vb below roundup(n, 8);
// vb0 and every grid but the last a multiple of 8): each XCD (vb mod 8)
// receives a contiguous range of items, so neighbouring items keep their
// shared atoms' rows in that XCD's L2.
// Work items of a launch: all n, or the first `limit` of them (limit > 0: a
// timing probe of the same block order over fewer frame rows).
__host__ __device__ __forceinline__ long long limit_items(long long n, long long limit)
{
    return limit > 0 && limit < n ? limit : n;
}

__device__ __forceinline__ long long xcd_item(long long vb, long long n)
{
    const long long per_xcd = (n + 7) / 8;
    return (vb % 8) * per_xcd + vb / 8;
}

// ---------------------------------------------------------------------------
// Single-trajectory pair kernels: block = one pair.
// ---------------------------------------------------------------------------
template <int S4>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void ksg_pair_kernel(
    KsgArgs args)
{
    __shared__ __attribute__((aligned(4 * kBins))) uint32_t lds[kPairWaves][kBins];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    // block = (4 pairs, one range of frame rows), wave w runs one pair: the
    // quads table's entry (a 2x2 tile of atoms, so each A and each B row is
    // read by two waves at about the same time), or pair 4q+w (consecutive
    // pairs share their first atom).
    const long long nquads = args.quads ? args.nquads : (args.npairs + kPairWaves - 1) / kPairWaves;
    const int nsplit = args.nsplit;
    const long long nitems = limit_items(nquads * nsplit, args.item_limit);
    // virtual block: the launcher splits a grid above 2^22 blocks into
    // launches at offsets vb0 (a multiple of 8)
    const long long vb = args.vb0 + blockIdx.x;
    const long long item = xcd_item(vb, nitems);
    if (item >= nitems) return;
    const long long q = item / nsplit;
    const long long p = args.quads ? (long long)args.quads[q * kPairWaves + w] : q * kPairWaves + w;
    const int part = (int)(item % nsplit);
    if (p < 0 || p >= args.npairs) return;   // no block barriers below
    const int N = args.N, T = args.K + 1;
    const int i0 = (int)((long long)N * part / nsplit);
    const int i1 = (int)((long long)N * (part + 1) / nsplit);
    const float* Da = args.D + (size_t)args.a1[p] * N * args.P;
    const float* Db = args.D + (size_t)args.a2[p] * N * args.P;
    const RowOut o = row_out(args);
    SortedRows sr;
    sr.PS = sorted_pitch(N);
    sr.SDa = args.SD + (size_t)args.a1[p] * N * sr.PS;
    sr.SDb = args.SD + (size_t)args.a2[p] * N * sr.PS;
    sr.SMa = args.SM + (size_t)args.a1[p] * N * kSumPitch;
    sr.SMb = args.SM + (size_t)args.a2[p] * N * kSumPitch;
    const double acc = rows_single<S4>(Da, Db, sr, args.P, N, T, i0, i1, 1, p, lds[w], lane, o);
    if (lane == 0) args.partial[p * nsplit + part] = acc;
}

__device__ __forceinline__ PrefixRows prefix_rows(const float* D, const float* SD, const float* SM,
                                                  const uint16_t* SI, long long P, int N, int a1,
                                                  int a2)
{
    PrefixRows pr;
    pr.P = P;
    pr.PS = sorted_pitch(N);
    pr.Da = reinterpret_cast<const uint32_t*>(D + (size_t)a1 * N * P);
    pr.Db = reinterpret_cast<const uint32_t*>(D + (size_t)a2 * N * P);
    pr.SDa = reinterpret_cast<const uint32_t*>(SD + (size_t)a1 * N * pr.PS);
    pr.SDb = reinterpret_cast<const uint32_t*>(SD + (size_t)a2 * N * pr.PS);
    pr.SMa = reinterpret_cast<const uint32_t*>(SM + (size_t)a1 * N * kSumPitch);
    pr.SMb = reinterpret_cast<const uint32_t*>(SM + (size_t)a2 * N * kSumPitch);
    pr.SIa = SI + (size_t)a1 * N * kPrefix;
    pr.SIb = SI + (size_t)a2 * N * kPrefix;
    return pr;
}

// Sorted-prefix pair kernel (kind 2): block = (4 pairs, one range of frame
// rows) as ksg_pair_kernel, each wave one pair.
template <bool Detail>
__global__ __launch_bounds__(256) void ksg_pair_prefix_kernel(KsgArgs args)
{
    __shared__ __attribute__((aligned(4 * kPBins))) PrefixLds lds[kPairWaves];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const long long nquads = args.quads ? args.nquads : (args.npairs + kPairWaves - 1) / kPairWaves;
    const int nsplit = args.nsplit;
    const long long nitems = limit_items(nquads * nsplit, args.item_limit);
    // virtual block: the launcher splits a grid above 2^22 blocks into
    // launches at offsets vb0 (a multiple of 8)
    const long long vb = args.vb0 + blockIdx.x;
    const long long item = xcd_item(vb, nitems);
    if (item >= nitems) return;
    // row-range major: the blocks in flight on one XCD are consecutive tiles
    // (sharing their first two atoms) over the same row range
    const long long q = item % nquads;
    const long long p = args.quads ? (long long)args.quads[q * kPairWaves + w] : q * kPairWaves + w;
    const int part = (int)(item / nquads);
    if (p < 0 || p >= args.npairs) return;   // no block barriers below
    const int N = args.N, T = args.K + 1;
    const int i0 = (int)((long long)N * part / nsplit);
    const int i1 = (int)((long long)N * (part + 1) / nsplit);
    const PrefixRows pr = prefix_rows(args.D, args.SD, args.SM, args.SI, args.P, N, args.a1[p],
                                      args.a2[p]);
    RowOut o = row_out(args);
    if (!Detail) {   // per-row count output off: its pointers leave the SGPR file
        o.rc_nx = nullptr;
        o.rc_ny = nullptr;
        o.rc_flag = nullptr;
    }
    const double acc = rows_prefix(pr, N, T, i0, i1, 1, p, lds[w], lane, o);
    if (lane == 0) args.partial[p * nsplit + part] = acc;
}

// Long-trajectory prefix kernel (kind 3): four pairs per block (a 2x2 atom
// tile, as ksg_pair_prefix_kernel), 1 KB of LDS per wave.  C: 256-frame
// chunks of the prefix read per row (2: 512 frames, 3: 768, 4: the whole
// 1024-frame store).  A block runs one chunk of args.chunk frame rows, and blocks are
// ordered chunk-major: the blocks in flight on one XCD run consecutive tiles,
// which share their first two atoms, over the same args.chunk rows, so those
// atoms' distance rows and prefixes are read from that XCD's L2.  Each
// wave's sum over its rows is added to fixsum[pair].
template <int C, bool Detail>
__global__ __launch_bounds__(256) void ksg_pair_long_kernel(KsgArgs args)
{
    __shared__ __attribute__((aligned(4 * kPBins))) uint32_t lds[kPairWaves][kPBins];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const long long nquads = args.quads ? args.nquads : (args.npairs + kPairWaves - 1) / kPairWaves;
    const int nchunk = (args.N + args.chunk - 1) / args.chunk;
    const long long nitems = limit_items(nquads * nchunk, args.item_limit);
    // virtual block: the launcher splits a grid above 2^22 blocks into
    // launches at offsets vb0 (a multiple of 8)
    const long long vb = args.vb0 + blockIdx.x;
    const long long item = xcd_item(vb, nitems);
    if (item >= nitems) return;
    const long long q = item % nquads;
    const long long p = args.quads ? (long long)args.quads[q * kPairWaves + w] : q * kPairWaves + w;
    const int chunk = (int)(item / nquads);
    if (p < 0 || p >= args.npairs) return;   // no block barriers below
    const int N = args.N, T = args.K + 1;
    const int i0 = chunk * args.chunk;
    const int i1 = min(N, i0 + args.chunk);
    const LongRows lr = long_rows(args.D, args.LV, args.LI, args.LS, args.P, N, args.a1[p],
                                  args.a2[p]);
    RowOut o = row_out(args);
    if (!Detail) {   // per-row count output off: its pointers leave the SGPR file
        o.rc_nx = nullptr;
        o.rc_ny = nullptr;
        o.rc_flag = nullptr;
    }
    const double acc = rows_long<C>(lr, N, T, i0, i1, 1, p, lds[w], lane, o);
    if (lane == 0) fix_add(&args.fixsum[p], acc);
}

// Rows the long-trajectory kernel queued to longq (the (k+1)-th neighbour
// beyond the 1024-frame prefix): one block of four waves per entry runs the
// cooperative streaming row (row_coop), its psi sum added to fixsum[pair].
__global__ __launch_bounds__(256) void ksg_long_rows_coop_kernel(KsgArgs args, unsigned nent)
{
    __shared__ CoopLds sh;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    for (unsigned e = blockIdx.x; e < nent; e += gridDim.x) {   // uniform over the block
    const FixEntry fe = args.longq[e];
    const long long p = fe.pair;
    const int N = args.N;
    const float* Da = args.D + (size_t)args.a1[p] * N * args.P;
    const float* Db = args.D + (size_t)args.a2[p] * N * args.P;
    const double acc = row_coop<kMaxFrames / 1024>(Da, Db, args.P, N, args.K + 1, fe.row, p, sh,
                                                   lane, w, row_out(args));
    if (threadIdx.x == 0) fix_add(&args.fixsum[p], acc);
    }
}

// Rows the sorted-prefix kernel queued to longq: one wave per entry runs the
// streaming one-wave path (rows_single) on that row, with the same recount
// rules; the row's psi[nx] + psi[ny] is added to fixsum[pair].
__global__ __launch_bounds__(256) void ksg_long_rows_kernel(KsgArgs args, unsigned nent)
{
    __shared__ __attribute__((aligned(4 * kBins))) uint32_t lds[kPairWaves][kBins];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    for (unsigned e = blockIdx.x * kPairWaves + w; e < nent; e += gridDim.x * kPairWaves) {
    const FixEntry fe = args.longq[e];
    const long long p = fe.pair;
    const int N = args.N;
    const int a1 = args.a1[p], a2 = args.a2[p];
    SortedRows sr;
    sr.PS = sorted_pitch(N);
    sr.SDa = args.SD + (size_t)a1 * N * sr.PS;
    sr.SDb = args.SD + (size_t)a2 * N * sr.PS;
    sr.SMa = args.SM + (size_t)a1 * N * kSumPitch;
    sr.SMb = args.SM + (size_t)a2 * N * kSumPitch;
    const double acc = rows_single<kSortedMaxFrames / 256>(
        args.D + (size_t)a1 * N * args.P, args.D + (size_t)a2 * N * args.P, sr, args.P, N,
        args.K + 1, fe.row, fe.row + 1, 1, p, lds[w], lane, row_out(args));
    if (lane == 0) fix_add(&args.fixsum[p], acc);
    }
}

template <int S4>
__global__ __launch_bounds__(256) void ksg_pair_coop_kernel(KsgArgs args)
{
    __shared__ CoopLds sh;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const long long vb = args.vb0 + blockIdx.x;
    const long long np = limit_items(args.npairs, args.item_limit);
    const long long p = xcd_item(vb, np);
    if (p >= np) return;   // uniform over the block
    const int N = args.N, T = args.K + 1;
    const float* Da = args.D + (size_t)args.a1[p] * N * args.P;
    const float* Db = args.D + (size_t)args.a2[p] * N * args.P;
    const RowOut o = row_out(args);
    double acc = 0.0;
    for (int i = 0; i < N; ++i) acc += row_coop<S4>(Da, Db, args.P, N, T, i, p, sh, lane, w, o);
    if (threadIdx.x == 0) args.partial[p] = acc;   // one slot per pair
}

// ---------------------------------------------------------------------------
// Persistent batched kernel: many trajectories of uneven atom and frame
// counts in one launch.  Blocks pull (trajectory, pair) items from a global
// counter until the queue is empty; each item runs the code path (kind, s4)
// that the tuner chose for its trajectory.
// ---------------------------------------------------------------------------
union __attribute__((aligned(4 * kBins))) BatchLds {
    uint32_t single[kPairWaves][kBins];
    PrefixLds prefix[kPairWaves];
    CoopLds coop;
};

template <int S4>
__device__ __forceinline__ double batch_single(const TrajDesc& d, const float* Da, const float* Db,
                                               long long item, BatchLds& u, double* wsum,
                                               int lane, int w, const RowOut& o)
{
    SortedRows sr;
    sr.PS = sorted_pitch(d.N);
    sr.SDa = d.SD + (Da - d.D) / d.P * sr.PS;
    sr.SDb = d.SD + (Db - d.D) / d.P * sr.PS;
    sr.SMa = d.SM + (Da - d.D) / d.P * kSumPitch;
    sr.SMb = d.SM + (Db - d.D) / d.P * kSumPitch;
    const double acc = rows_single<S4>(Da, Db, sr, d.P, d.N, d.K + 1, w, d.N, kPairWaves, item,
                                       u.single[w], lane, o);
    if (lane == 0) wsum[w] = acc;
    __syncthreads();
    return ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

__device__ __forceinline__ double batch_prefix(const TrajDesc& d, int a1, int a2, long long item,
                                               PrefixLds& lds, double* wsum, int lane, int w,
                                               const RowOut& o)
{
    const PrefixRows pr = prefix_rows(d.D, d.SD, d.SM, d.SI, d.P, d.N, a1, a2);
    const double acc = rows_prefix(pr, d.N, d.K + 1, w, d.N, kPairWaves, item, lds, lane, o);
    if (lane == 0) wsum[w] = acc;
    __syncthreads();
    return ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

// Kind 3 in the batched kernel: one work item is a 2x2 atom tile of the
// trajectory times one chunk of d.chunk frame rows (chunk-major, as
// ksg_pair_long_kernel); wave w runs the tile's pair w over the chunk and
// adds its row sum to fixsum[pair].
__device__ __forceinline__ void batch_long(const TrajDesc& d, const KsgBatchArgs& b,
                                           long long local, uint32_t* hist, int lane, int w,
                                           const RowOut& o)
{
    const long long chunk = local / d.nquads, qd = local - chunk * d.nquads;
    const int p = uniform_copy(d.quads[qd * kPairWaves + w]);
    if (p < 0) return;   // no block barriers in this path
    const long long gp = d.pair0 + p;
    const int a1 = uniform_copy(b.a1[gp]), a2 = uniform_copy(b.a2[gp]);
    const int i0 = (int)chunk * d.chunk;
    const int i1 = min(d.N, i0 + d.chunk);
    const LongRows lr = long_rows(d.D, d.LV, d.LI, d.LS, d.P, d.N, a1, a2);
    const double acc = d.s4 == 2   ? rows_long<2>(lr, d.N, d.K + 1, i0, i1, 1, gp, hist, lane, o)
                       : d.s4 == 3 ? rows_long<3>(lr, d.N, d.K + 1, i0, i1, 1, gp, hist, lane, o)
                                   : rows_long<4>(lr, d.N, d.K + 1, i0, i1, 1, gp, hist, lane, o);
    if (lane == 0) fix_add(&b.fixsum[gp], acc);
}

template <int S4>
__device__ __forceinline__ double batch_coop(const TrajDesc& d, const float* Da, const float* Db,
                                             long long item, BatchLds& u, int lane, int w,
                                             const RowOut& o)
{
    double acc = 0.0;
    for (int i = 0; i < d.N; ++i)
        acc += row_coop<S4>(Da, Db, d.P, d.N, d.K + 1, i, item, u.coop, lane, w, o);
    return acc;
}

// LDS of the batch kernel for the code paths it compiles (bit k = kind k):
// the streaming paths' union, the prefix path's staged rows, or the
// long-trajectory path's histograms alone.
template <int Paths>
struct BatchShared {
    BatchLds u;
    __device__ PrefixLds& pre(int w) { return u.prefix[w]; }
    __device__ uint32_t* hist(int w) { return u.single[w]; }
};
template <>
struct BatchShared<4> {
    PrefixLds prefix[kPairWaves];
    __device__ PrefixLds& pre(int w) { return prefix[w]; }
};
template <>
struct BatchShared<12> {
    PrefixLds prefix[kPairWaves];
    __device__ PrefixLds& pre(int w) { return prefix[w]; }
    __device__ uint32_t* hist(int w) { return prefix[w].hist; }
};
template <>
struct BatchShared<8> {
    __attribute__((aligned(4 * kPBins))) uint32_t h[kPairWaves][kPBins];
    __device__ uint32_t* hist(int w) { return h[w]; }
};

// Paths: the code paths compiled in (bit k = kind k).  A kernel's register
// and LDS allocation is that of its largest path, so the host launches the
// smallest instantiation that covers the kinds of a launch's trajectories.
template <int Paths>
__global__ __launch_bounds__(256) void ksg_batch_kernel(KsgBatchArgs b)
{
    __shared__ __attribute__((aligned(4 * kPBins))) BatchShared<Paths> sh;
    __shared__ double wsum[kPairWaves];
    __shared__ long long next;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    RowOut o;
    o.counters = b.counters;
    o.fixq = b.fixq;
    o.fixcap = b.fixcap;
    o.longq = b.longq;
    o.longcap = b.longcap;
    o.rc_nx = nullptr;
    o.rc_ny = nullptr;
    o.rc_flag = nullptr;
    // Work queues: XCD x (block id mod 8) takes the x-th eighth of every
    // trajectory's items, trajectories in order, so the blocks in flight on
    // one XCD run neighbouring items and share their atoms' rows in its L2;
    // a block whose queue is empty takes items from the other queues.
    const int xcd = (int)(blockIdx.x & 7);
    for (int k = 0; k < 8; ++k) {
        const int xq = (xcd + k) & 7;
        for (;;) {
            if (threadIdx.x == 0) next = (long long)atomicAdd(&b.queue[xq], 1ULL);
            __syncthreads();
            long long v = uniform_copy(next);
            __syncthreads();
            int tq = -1;
            long long local = 0;
            for (int q = 0; q < b.ntraj; ++q) {
                const long long n = uniform_copy(b.traj[q].nitems);
                const long long lo = n * xq / 8, hi = n * (xq + 1) / 8;
                if (v < hi - lo) {
                    tq = q;
                    local = lo + v;
                    break;
                }
                v -= hi - lo;
            }
            if (tq < 0) break;   // queue xq is empty (uniform over the block)
            // opaque per item: keeps lane-derived addresses from being hoisted
            // out of the item loop and held in registers across items
            int lane = (int)(threadIdx.x & (kWave - 1));
            asm volatile("" : "+v"(lane));
            const TrajDesc d = uniform_copy(b.traj[tq]);
            o.psi = d.psi;
            // the branches are uniform over the block (one item per block)
            if constexpr ((Paths & 8) != 0) {
                if (d.kind == 3) {
                    batch_long(d, b, local, sh.hist(w), lane, w, o);
                    __syncthreads();
                    continue;
                }
            }
            const long long gp = d.pair0 + local;   // one pair per item
            const int ia1 = uniform_copy(b.a1[gp]), ia2 = uniform_copy(b.a2[gp]);
            const float* Da = d.D + (size_t)ia1 * d.N * d.P;
            const float* Db = d.D + (size_t)ia2 * d.N * d.P;
            double acc = 0.0;
            if constexpr ((Paths & 4) != 0) {
                if (d.kind == 2) acc = batch_prefix(d, ia1, ia2, gp, sh.pre(w), wsum, lane, w, o);
            }
            if constexpr ((Paths & 1) != 0) {
                if (d.kind == 0) {
                    switch (d.s4) {
                        case 2: acc = batch_single<2>(d, Da, Db, gp, sh.u, wsum, lane, w, o); break;
                        case 4: acc = batch_single<4>(d, Da, Db, gp, sh.u, wsum, lane, w, o); break;
                        case 6: acc = batch_single<6>(d, Da, Db, gp, sh.u, wsum, lane, w, o); break;
                        default: acc = batch_single<8>(d, Da, Db, gp, sh.u, wsum, lane, w, o); break;
                    }
                }
            }
            if constexpr ((Paths & 2) != 0) {
                if (d.kind == 1) {
                    switch (d.s4) {
                        case 3: acc = batch_coop<3>(d, Da, Db, gp, sh.u, lane, w, o); break;
                        case 4: acc = batch_coop<4>(d, Da, Db, gp, sh.u, lane, w, o); break;
                        case 6: acc = batch_coop<6>(d, Da, Db, gp, sh.u, lane, w, o); break;
                        default: acc = batch_coop<8>(d, Da, Db, gp, sh.u, lane, w, o); break;
                    }
                }
            }
            if (threadIdx.x == 0) b.partial[gp] = acc;
            __syncthreads();
        }
    }
}

// Long rows of the batch (kinds 2 and 3): one block per entry runs the
// cooperative streaming row of its trajectory, the psi sum added to
// fixsum[item].  The register tile S4 covers the trajectory's frame count.
__global__ __launch_bounds__(256) void ksg_batch_long_rows_kernel(KsgBatchArgs b, FixSum* fixsum,
                                                                  unsigned nent)
{
    __shared__ CoopLds sh;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    RowOut o;
    o.counters = b.counters;
    o.fixq = b.fixq;
    o.fixcap = b.fixcap;
    o.longq = nullptr;
    o.longcap = 0u;
    o.rc_nx = nullptr;
    o.rc_ny = nullptr;
    o.rc_flag = nullptr;
    for (unsigned e = blockIdx.x; e < nent; e += gridDim.x) {   // uniform over the block
    const FixEntry fe = b.longq[e];
    const long long item = fe.pair;
    const TrajDesc d = b.traj[b.item_traj[item]];
    o.psi = d.psi;
    const float* Da = d.D + (size_t)b.a1[item] * d.N * d.P;
    const float* Db = d.D + (size_t)b.a2[item] * d.N * d.P;
    const int T = d.K + 1;
    double acc;
    if (d.N <= 3 * 1024)
        acc = row_coop<3>(Da, Db, d.P, d.N, T, fe.row, item, sh, lane, w, o);
    else if (d.N <= 4 * 1024)
        acc = row_coop<4>(Da, Db, d.P, d.N, T, fe.row, item, sh, lane, w, o);
    else if (d.N <= 6 * 1024)
        acc = row_coop<6>(Da, Db, d.P, d.N, T, fe.row, item, sh, lane, w, o);
    else
        acc = row_coop<8>(Da, Db, d.P, d.N, T, fe.row, item, sh, lane, w, o);
    if (threadIdx.x == 0) fix_add(&fixsum[item], acc);
    }
}

// ---------------------------------------------------------------------------
// FP64 KSG terms of one frame row i of one atom pair, computed from the
// coordinates by one wave (no distance rows; any N).  dA, dB are recomputed
// in FP64 on every pass.  ek: FP32 bits of the (k+1)-th smallest joint key
// (self included) when a pair kernel has found it, or kUnknownEps, in which
// case a three-digit radix select over the same FP32 keys finds it first
// (hist: the wave's 2048 LDS words).  The key fixes the group G of frames
// whose joint distance rounds to it; the k-th neighbour distance is the r-th
// smallest FP64 value in G with r = k - #{j != i : key < ek}.  More than r
// frames at that FP64 value is an exact tie: all of them enter the radius
// set (the reference CUDA kernel's rule, lib/kraskov_gpu.cu:160-171).  When
// the select finds exactly r frames of G besides i, no FP64 walk is needed:
// the set is {j != i : key <= ek}.
// ---------------------------------------------------------------------------
struct Fp64Row {
    int nx, ny;
    bool tie;
};

__device__ Fp64Row fp64_row(const double* __restrict__ xa, const double* __restrict__ xb, int N,
                            int K, int i, uint32_t ek, uint32_t* hist, int lane)
{
    const double qa0 = xa[i], qa1 = xa[N + i], qa2 = xa[2 * N + i];
    const double qb0 = xb[i], qb1 = xb[N + i], qb2 = xb[2 * N + i];
    auto dist = [&](int j, double& A, double& B) {
        A = fmax(fmax(fabs(qa0 - xa[j]), fabs(qa1 - xa[N + j])), fabs(qa2 - xa[2 * N + j]));
        B = fmax(fmax(fabs(qb0 - xb[j]), fabs(qb1 - xb[N + j])), fabs(qb2 - xb[2 * N + j]));
    };
    auto key_of = [&](int j) {
        double A, B;
        dist(j, A, B);
        return __float_as_uint((float)fmax(A, B));
    };

    int r;              // rank of the k-th distance inside G (G without i)
    bool whole = false; // G (without i) holds exactly r frames
    if (ek == kUnknownEps) {
        // (k+1)-th smallest key, self included: digits 30..20, 19..10, 9..0
        clear_hist<2048>(hist, lane);
        wave_sync();
        for (int j = lane; j < N; j += kWave) atomicAdd(&hist[key_of(j) >> 20], 1u);
        wave_sync();
        int bin, rank, pop;
        find_bin<32>(hist, lane, K + 1, bin, rank, pop);
        uint32_t prefix = (uint32_t)bin;
        wave_sync();
        clear_hist<1024>(hist, lane);
        wave_sync();
        for (int j = lane; j < N; j += kWave) {
            const uint32_t key = key_of(j);
            if ((key >> 20) == prefix) atomicAdd(&hist[(key >> 10) & 1023u], 1u);
        }
        wave_sync();
        find_bin<16>(hist, lane, rank, bin, rank, pop);
        prefix = (prefix << 10) | (uint32_t)bin;
        wave_sync();
        clear_hist<1024>(hist, lane);
        wave_sync();
        for (int j = lane; j < N; j += kWave) {
            const uint32_t key = key_of(j);
            if ((key >> 10) == prefix) atomicAdd(&hist[key & 1023u], 1u);
        }
        wave_sync();
        find_bin<16>(hist, lane, rank, bin, rank, pop);
        ek = (prefix << 10) | (uint32_t)bin;
        wave_sync();   // hist is cleared again by the wave's next row
        // K + 1 - rank keys lie below ek; i's own key 0 is among them unless
        // ek == 0, when i is in G itself
        r = ek ? rank : rank - 1;
        whole = pop == rank;
    } else {
        r = 0;   // set below
    }

    double eps64 = 0.0;
    bool tie = false;
    double ex = 0.0, ey = 0.0;
    bool radii = false;   // eps64, tie and the radii found by the one-pass form
    if (ek != kUnknownEps) {
        // The key is known: one pass counts the keys below it (frames certainly
        // in the set: rounding is monotone, so their FP64 distance is below
        // every FP64 value of key ek), takes their marginal maxima, and lists
        // the frames whose key equals ek (the group G, usually a few) with
        // their FP64 distances in LDS.  The r-th smallest FP64 value of G is
        // eps; its ties and G's members up to eps follow from the list.
        double* gJ = reinterpret_cast<double*>(hist);   // kWave entries each
        double* gA = gJ + kWave;
        double* gB = gA + kWave;
        int clt = 0, g = 0;
        double sx = 0.0, sy = 0.0;
        for (int j0 = 0; j0 < N; j0 += kWave) {   // uniform trip count
            const int j = j0 + lane;
            double A = 0.0, B = 0.0, J = 0.0;
            bool eq = false;
            if (j < N && j != i) {
                dist(j, A, B);
                J = fmax(A, B);
                const uint32_t key = __float_as_uint((float)J);
                if (key < ek) {
                    ++clt;
                    sx = fmax(sx, A);
                    sy = fmax(sy, B);
                }
                eq = key == ek;
            }
            const uint64_t m = __ballot(eq);
            if (m) {   // uniform
                const int slot = g + mbcnt(m);
                if (eq && slot < kWave) {
                    gJ[slot] = J;
                    gA[slot] = A;
                    gB[slot] = B;
                }
                g += __popcll(m);
            }
        }
        clt = wave_sum(clt);
        r = K - clt;   // >= 1 for a consistent entry
        if (g <= kWave && r >= 1 && r <= g) {
            wave_sync();   // the list is read by every lane
            double Jl = INFINITY;
            int cl = 0, el = 0;   // entries below and equal to this lane's
            if (lane < g) {
                Jl = gJ[lane];
                for (int u = 0; u < g; ++u) {
                    const double v = gJ[u];
                    cl += v < Jl;
                    el += v == Jl;
                }
            }
            // an entry with cl < r <= cl + el exists for 1 <= r <= g
            const uint64_t pick = __ballot(lane < g && cl < r && r <= cl + el);
            const int src = __ffsll((unsigned long long)pick) - 1;
            eps64 = __shfl(Jl, src, kWave);
            tie = __shfl(cl + el, src, kWave) > r;
            double mx = sx, my = sy;
            if (lane < g && Jl <= eps64) {
                mx = fmax(mx, gA[lane]);
                my = fmax(my, gB[lane]);
            }
            ex = wave_max_d(mx);
            ey = wave_max_d(my);
            radii = true;
            wave_sync();   // the list is rewritten by the wave's next row
        }
    }
    if (!radii && !whole) {
        // walk the distinct FP64 values of G upwards until r are covered
        // (the group did not fit the list, or the key was not given)
        if (ek != kUnknownEps) {
            int clt = 0;
            for (int j = lane; j < N; j += kWave) {
                if (j == i) continue;
                clt += key_of(j) < ek;
            }
            r = K - wave_sum(clt);
        }
        double prev = -1.0;
        int cum = 0;
        for (;;) {
            double vmin = INFINITY;
            for (int j = lane; j < N; j += kWave) {
                if (j == i) continue;
                double A, B;
                dist(j, A, B);
                const double J = fmax(A, B);
                if (__float_as_uint((float)J) == ek && J > prev) vmin = fmin(vmin, J);
            }
            vmin = wave_min_d(vmin);
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
    }

    if (!radii) {
        // radii: the largest marginal distances over the set.  With `whole`
        // the set is {key <= ek}, which equals {J <= max_G J} as rounding is
        // monotone.
        for (int j = lane; j < N; j += kWave) {
            if (j == i) continue;
            double A, B;
            dist(j, A, B);
            const bool in = whole ? __float_as_uint((float)fmax(A, B)) <= ek : fmax(A, B) <= eps64;
            if (in) {
                ex = fmax(ex, A);
                ey = fmax(ey, B);
            }
        }
        ex = wave_max_d(ex);
        ey = wave_max_d(ey);
    }
    int nx = 0, ny = 0;
    for (int j = lane; j < N; j += kWave) {
        if (j == i) continue;
        double A, B;
        dist(j, A, B);
        nx += A <= ex;
        ny += B <= ey;
    }
    Fp64Row out;
    out.nx = wave_sum(nx);
    out.ny = wave_sum(ny);
    out.tie = tie;
    return out;
}

// ---------------------------------------------------------------------------
// FP64 recount of queued rows: one wave per entry runs fp64_row with the
// key the pair kernel found (or kUnknownEps), its psi terms added to
// fixsum[pair].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ksg_fixup_kernel(FixupArgs args, unsigned nent)
{
    __shared__ uint32_t lds[4][kBins];
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t* hist = lds[threadIdx.x / kWave];
    for (unsigned e = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; e < nent;
         e += gridDim.x * (blockDim.x / kWave)) {
        const FixEntry fe = args.fixq[e];
        const long long p = fe.pair;
        const int i = fe.row;
        const TrajDesc d = args.traj[args.item_traj ? args.item_traj[p] : 0];
        const int N = d.N;
        const double* xa = d.X + (size_t)(3 * args.a1[p]) * N;
        const double* xb = d.X + (size_t)(3 * args.a2[p]) * N;
        const Fp64Row o = fp64_row(xa, xb, N, d.K, i, fe.eps_bits, hist, lane);
        if (lane == 0) {
            fix_add(&args.fixsum[p], d.psi[o.nx] + d.psi[o.ny]);
            if (o.tie) atomicAdd(&args.counters[1], 1u);
            if (args.rc_nx) {
                const size_t off = (size_t)p * N + i;
                args.rc_nx[off] = o.nx;
                args.rc_ny[off] = o.ny;
                args.rc_flag[off] = o.tie ? 2 : 1;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Coordinate path (kind 4): every frame row of every pair by fp64_row from
// the FP64 coordinates.  No distance rows, so device memory is O(natoms * N)
// as in the reference's kernel (lib/kraskov_gpu.cu:107-196), and there is no
// frame limit.  Work item = (pair p, row range r of nsplit); block = four
// waves, wave w taking rows i0 + w, i0 + w + 4, ... of the range.  The four
// wave sums are added in wave order into partial[p * nsplit + r], so results
// are reproducible bit for bit.  Items are pair-major: the blocks in flight
// work on few pairs, whose six coordinate rows stay in L2.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ksg_coord_kernel(KsgArgs args)
{
    __shared__ uint32_t lds[kPairWaves][kBins];
    __shared__ double wsum[kPairWaves];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int N = args.N, nsplit = args.nsplit;
    const long long nitems = limit_items(args.npairs * nsplit, args.item_limit);
    for (long long item = blockIdx.x; item < nitems; item += gridDim.x) {   // uniform
        const long long p = item / nsplit;
        const int part = (int)(item - p * nsplit);
        const int i0 = (int)((long long)N * part / nsplit);
        const int i1 = (int)((long long)N * (part + 1) / nsplit);
        const double* xa = args.X + (size_t)(3 * args.a1[p]) * N;
        const double* xb = args.X + (size_t)(3 * args.a2[p]) * N;
        double acc = 0.0;
        for (int i = i0 + w; i < i1; i += kPairWaves) {
            const Fp64Row o = fp64_row(xa, xb, N, args.K, i, kUnknownEps, lds[w], lane);
            acc += args.psi[o.nx] + args.psi[o.ny];
            if (lane == 0) {
                if (o.tie) atomicAdd(&args.counters[1], 1u);
                if (args.rc_nx) {
                    const size_t off = (size_t)p * N + i;
                    args.rc_nx[off] = o.nx;
                    args.rc_ny[off] = o.ny;
                    args.rc_flag[off] = o.tie ? 2 : 1;
                }
            }
        }
        if (lane == 0) wsum[w] = acc;
        __syncthreads();
        if (threadIdx.x == 0) args.partial[item] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
        __syncthreads();
    }
}
__global__ void ksg_clear_kernel(FixSum* fixsum, long long npairs, unsigned* counters,
                                 unsigned long long* queue, double* partial, int nqueue)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < npairs) fixsum[t] = 0ull;
    if (partial && t < npairs) partial[t] = 0.0;
    if (t < 3) counters[t] = 0u;
    if (queue && t < nqueue) queue[t] = 0ull;
}

// MI = psi[N] + phi[K] - sum/N as lib/kraskov.cpp:1318.
__global__ void ksg_finalize_kernel(FixupArgs args, long long nitems, const double* partial,
                                    int nsplit, double* mi)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nitems) return;
    const TrajDesc d = args.traj[args.item_traj ? args.item_traj[p] : 0];
    double sum = 0.0;
    for (int r = 0; r < nsplit; ++r) sum += partial[p * nsplit + r];
    mi[p] = d.psiN_phiK - (sum + fix_value(args.fixsum[p])) / d.N;
}

// MI of the device-pointer coordinate path: psi[N] + (psi[K] - 1/K) - sum/N
// (lib/kraskov.cpp:1318) with psi read on the device.
__global__ void ksg_finalize_coord_kernel(const double* __restrict__ partial, int nsplit,
                                          const double* __restrict__ psi, int N, int K,
                                          long long npairs, double* __restrict__ mi)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    double sum = 0.0;
    for (int r = 0; r < nsplit; ++r) sum += partial[p * nsplit + r];
    mi[p] = (psi[N] + (psi[K] - 1 / (1.0 * K))) - sum / N;
}

// Wave primitive self-check: out[0..63] inclusive scan of lane+1 (mod 7
// pattern), out[64] sum, out[65] max, out[66] float max bits.
__global__ void ksg_selftest_kernel(int* out)
{
    const int lane = threadIdx.x;
    const int v = (lane * 5 + 3) % 7 + 1;
    out[lane] = wave_incl_scan(v);
    const int s = wave_sum(v);
    const int m = wave_max_i((lane * 37) % 64);
    const float f = wave_max((float)((lane * 11) % 64) * 0.5f);
    if (lane == 0) {
        out[64] = s;
        out[65] = m;
        out[66] = __float_as_int(f);
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

// Grid for a grid-stride kernel over n blocks' worth of work: at most 2^22
// blocks (2^30 work-items at 256 threads, under the 2^32 work-items a
// dispatch can carry), a multiple of 8 (the XCD mapping of xcd_item).
unsigned capped_grid(long long nblocks)
{
    const long long cap = 1LL << 22;
    const long long g = (nblocks + 7) / 8 * 8;
    return (unsigned)(g < cap ? (g > 0 ? g : 8) : cap);
}
hipError_t launch_build_rows(const double* X, int natoms, int N, long long P, float* D,
                             float* SD, float* SM, uint16_t* SI, hipStream_t s)
{
    if ((SD == nullptr) != (SM == nullptr) || (SD == nullptr) != (SI == nullptr) ||
        (SD && N > kSortedMaxFrames))
        return hipErrorInvalidValue;
    const long long rows = (long long)natoms * N;
    if (!SD) {
        hipLaunchKernelGGL(ksg_build_rows_kernel, dim3(capped_grid(rows)), dim3(256), 0, s, X, N,
                           P, rows, D);
        return hipGetLastError();
    }
    // sorted rows: one wave per row, EW frames per lane, 64 * EW >= N, four
    // rows per block
    const unsigned g = capped_grid((rows + 3) / 4);
    if (N <= 256)
        hipLaunchKernelGGL(ksg_build_sorted_wave_kernel<4>, dim3(g), dim3(256), 0, s, X, N, P, rows,
                           D, SD, SM, SI);
    else if (N <= 512)
        hipLaunchKernelGGL(ksg_build_sorted_wave_kernel<8>, dim3(g), dim3(256), 0, s, X, N, P, rows,
                           D, SD, SM, SI);
    else if (N <= 1024)
        hipLaunchKernelGGL(ksg_build_sorted_wave_kernel<16>, dim3(g), dim3(256), 0, s, X, N, P,
                           rows, D, SD, SM, SI);
    else
        hipLaunchKernelGGL(ksg_build_sorted_wave_kernel<32>, dim3(g), dim3(256), 0, s, X, N, P,
                           rows, D, SD, SM, SI);
    return hipGetLastError();
}

hipError_t launch_build_prefix(const float* D, int natoms, int N, long long P, float* LV,
                               uint16_t* LI, float* LS, hipStream_t s)
{
    if (!D || !LV || !LI || !LS || N <= kSortedMaxFrames || N > kMaxFrames)
        return hipErrorInvalidValue;
    if (P % 256 != 0 || P < N || P > 16 * 512) return hipErrorInvalidValue;
    const long long rows = (long long)natoms * N;
    hipLaunchKernelGGL(ksg_build_prefix_kernel, dim3(capped_grid(rows)), dim3(512), 0, s, D, N, P,
                       rows, LV, LI, LS);
    return hipGetLastError();
}

hipError_t launch_ksg_clear(FixSum* fixsum, long long n, unsigned* counters,
                            unsigned long long* queue, hipStream_t s, double* partial, int nqueue)
{
    if (nqueue < 0 || nqueue > 256) return hipErrorInvalidValue;
    const long long m = n > 256 ? n : 256;
    hipLaunchKernelGGL(ksg_clear_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s,
                       fixsum, n, counters, queue, partial, nqueue);
    return hipGetLastError();
}

// Kernel variants: kind 0 = one wave per row (N <= 256*S4), kind 1 = four
// waves per row (N <= 1024*S4), kinds 2 and 3 the prefix paths, kind 4 the
// coordinate path.  default_variant picks the smallest that fits.
bool variant_fits(int kind, int s4, int N, int K)
{
    if (kind == 4) return K >= 1 && N > K;
    if (kind == 2) return N <= kSortedMaxFrames && K + 1 <= kPrefixMaxT;
    if (kind == 3)
        return s4 >= 2 && s4 <= 4 && N > kSortedMaxFrames && N <= kMaxFrames &&
               K + 1 <= kPrefixMaxT;
    const long long cap = (kind == 0 ? 256LL : 1024LL) * s4;
    const bool inst = kind == 0 ? (s4 == 1 || s4 == 2 || s4 == 3 || s4 == 4 || s4 == 5 ||
                                   s4 == 6 || s4 == 8)
                                : kind == 1 && (s4 >= 2 && s4 <= 8);
    return inst && N <= cap;
}

void default_variant(int N, int K, int& kind, int& s4)
{
    if (N > kMaxFrames) { kind = 4; s4 = 1; return; }
    if (variant_fits(2, 1, N, K)) { kind = 2; s4 = 1; return; }
    if (variant_fits(3, 2, N, K)) { kind = 3; s4 = long_prefix_chunks(N, K); return; }
    static const int single[] = {1, 2, 3, 4, 5, 6, 8};
    for (int v : single)
        if (N <= 256 * v) { kind = 0; s4 = v; return; }
    kind = 1;
    s4 = (N + 1023) / 1024;
    if (s4 < 2) s4 = 2;
}

long long pair_items(int kind, const KsgArgs& a)
{
    const long long nquads = a.quads ? a.nquads : (a.npairs + kPairWaves - 1) / kPairWaves;
    return kind == 1   ? a.npairs
           : kind == 3 ? nquads * ((a.N + a.chunk - 1) / a.chunk)
           : kind == 4 ? a.npairs * a.nsplit
                       : nquads * a.nsplit;
}

int variant_slots(int kind, int nsplit)
{
    return kind == 3 ? 0 : kind == 1 ? 1 : nsplit;
}

hipError_t launch_ksg_pairs(const KsgArgs& a, hipStream_t s, int kind, int s4, int* nsplit)
{
    if (kind < 0) {
        default_variant(a.N, a.K, kind, s4);
        if (kind == 3 && !(a.LV && a.LI && a.LS)) {   // prefix stores not built
            kind = 1;
            s4 = (a.N + 1023) / 1024;
            if (s4 < 2) s4 = 2;
        }
    }
    if (!variant_fits(kind, s4, a.N, a.K)) return hipErrorInvalidValue;
    if (variant_sorted(kind) && !(a.SD && a.SM && a.SI)) return hipErrorInvalidValue;
    if (variant_prefix_store(kind) && !(a.LV && a.LI && a.LS)) return hipErrorInvalidValue;
    if (kind != 4 && !a.D) return hipErrorInvalidValue;
    if (a.nsplit < 1 || a.nsplit > kMaxRowSplit || a.chunk < 1) return hipErrorInvalidValue;
    // partial sums per pair: nsplit slots, one (kind 1), or none (kind 3
    // adds into fixsum)
    if (nsplit) *nsplit = variant_slots(kind, a.nsplit);
    if (kind == 4) {
        const long long items = limit_items(a.npairs * a.nsplit, a.item_limit);
        hipLaunchKernelGGL(ksg_coord_kernel, dim3(capped_grid(items)), dim3(kWave * kPairWaves),
                           0, s, a);
        return hipGetLastError();
    }
    const long long nitems = limit_items(pair_items(kind, a), a.item_limit);
    // one block per virtual block; a dispatch carries fewer than 2^32
    // work-items, so grids above 2^22 blocks go out in several launches
    const long long nvb = (nitems + 7) / 8 * 8;
    const long long cap = 1LL << 22;
    const dim3 block(kWave * kPairWaves);
    for (long long vb0 = 0; vb0 < nvb; vb0 += cap) {
    KsgArgs c = a;
    c.vb0 = vb0;
    const dim3 grid((unsigned)(nvb - vb0 < cap ? nvb - vb0 : cap));
#define GMX_SINGLE(V) case V: hipLaunchKernelGGL(ksg_pair_kernel<V>, grid, block, 0, s, c); break;
#define GMX_COOP(V) case V: hipLaunchKernelGGL(ksg_pair_coop_kernel<V>, grid, block, 0, s, c); break;
    if (kind == 3) {
        if (s4 == 2 && c.rc_nx)
            hipLaunchKernelGGL((ksg_pair_long_kernel<2, true>), grid, block, 0, s, c);
        else if (s4 == 2)
            hipLaunchKernelGGL((ksg_pair_long_kernel<2, false>), grid, block, 0, s, c);
        else if (s4 == 3 && c.rc_nx)
            hipLaunchKernelGGL((ksg_pair_long_kernel<3, true>), grid, block, 0, s, c);
        else if (s4 == 3)
            hipLaunchKernelGGL((ksg_pair_long_kernel<3, false>), grid, block, 0, s, c);
        else if (c.rc_nx)
            hipLaunchKernelGGL((ksg_pair_long_kernel<4, true>), grid, block, 0, s, c);
        else
            hipLaunchKernelGGL((ksg_pair_long_kernel<4, false>), grid, block, 0, s, c);
    } else if (kind == 2) {
        if (c.rc_nx)
            hipLaunchKernelGGL(ksg_pair_prefix_kernel<true>, grid, block, 0, s, c);
        else
            hipLaunchKernelGGL(ksg_pair_prefix_kernel<false>, grid, block, 0, s, c);
    } else if (kind == 0) {
        switch (s4) {
            GMX_SINGLE(1) GMX_SINGLE(2) GMX_SINGLE(3) GMX_SINGLE(4) GMX_SINGLE(5) GMX_SINGLE(6)
            GMX_SINGLE(8)
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (s4) {
            GMX_COOP(2) GMX_COOP(3) GMX_COOP(4) GMX_COOP(5) GMX_COOP(6) GMX_COOP(7) GMX_COOP(8)
            default: return hipErrorInvalidValue;
        }
    }
#undef GMX_SINGLE
#undef GMX_COOP
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_ksg_coord_pairs(const double* X, int N, const double* psi, int K,
                                 long long npairs, const int* a1, const int* a2, int nsplit,
                                 double* partial, unsigned* counters, double* mi, hipStream_t s)
{
    if (!X || !psi || !a1 || !a2 || !partial || !counters || !mi || npairs < 1 || K < 1 ||
        N <= K || nsplit < 1 || nsplit > kMaxRowSplit)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(ksg_clear_kernel, dim3(1), dim3(256), 0, s, nullptr, 0LL, counters,
                       nullptr, nullptr, 0);
    KsgArgs a{};
    a.X = X;
    a.psi = psi;
    a.a1 = a1;
    a.a2 = a2;
    a.partial = partial;
    a.counters = counters;
    a.npairs = npairs;
    a.N = N;
    a.K = K;
    a.nsplit = nsplit;
    a.chunk = kLongChunk;
    hipLaunchKernelGGL(ksg_coord_kernel, dim3(capped_grid(npairs * nsplit)),
                       dim3(kWave * kPairWaves), 0, s, a);
    hipLaunchKernelGGL(ksg_finalize_coord_kernel, dim3((unsigned)((npairs + 255) / 256)),
                       dim3(256), 0, s, partial, nsplit, psi, N, K, npairs, mi);
    return hipGetLastError();
}

hipError_t launch_ksg_fixup(const FixupArgs& f, unsigned nqueued, hipStream_t s)
{
    if (nqueued == 0) return hipSuccess;
    const unsigned per_block = 4;
    hipLaunchKernelGGL(ksg_fixup_kernel, dim3(capped_grid((nqueued + per_block - 1) / per_block)),
                       dim3(kWave * per_block), 0, s, f, nqueued);
    return hipGetLastError();
}

hipError_t launch_ksg_long_rows(const KsgArgs& a, unsigned nlong, hipStream_t s)
{
    if (nlong == 0) return hipSuccess;
    if (!a.longq) return hipErrorInvalidValue;
    if (a.N > kSortedMaxFrames) {
        hipLaunchKernelGGL(ksg_long_rows_coop_kernel, dim3(capped_grid(nlong)),
                           dim3(kWave * kPairWaves), 0, s, a, nlong);
        return hipGetLastError();
    }
    if (!a.SD || !a.SM) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ksg_long_rows_kernel, dim3(capped_grid((nlong + kPairWaves - 1) / kPairWaves)),
                       dim3(kWave * kPairWaves), 0, s, a, nlong);
    return hipGetLastError();
}

hipError_t launch_ksg_finalize(const FixupArgs& f, long long nitems, const double* partial,
                               int nsplit, double* mi, hipStream_t s)
{
    hipLaunchKernelGGL(ksg_finalize_kernel, dim3((unsigned)((nitems + 255) / 256)), dim3(256), 0,
                       s, f, nitems, partial, nsplit, mi);
    return hipGetLastError();
}

bool batch_variant_ok(int kind, int s4, int N, int K)
{
    if (kind == 2 || kind == 3) return variant_fits(kind, s4, N, K);
    if (kind == 0) return (s4 == 2 || s4 == 4 || s4 == 6 || s4 == 8) && N <= 256 * s4;
    if (kind == 1) return (s4 == 3 || s4 == 4 || s4 == 6 || s4 == 8) && N <= 1024 * s4;
    return false;
}

hipError_t launch_ksg_batch_long_rows(const KsgBatchArgs& b, FixSum* fixsum, unsigned nlong,
                                      hipStream_t s)
{
    if (nlong == 0) return hipSuccess;
    if (!b.longq || !fixsum) return hipErrorInvalidValue;
    hipLaunchKernelGGL(ksg_batch_long_rows_kernel, dim3(capped_grid(nlong)),
                       dim3(kWave * kPairWaves), 0, s, b, fixsum, nlong);
    return hipGetLastError();
}

hipError_t launch_ksg_batch(const KsgBatchArgs& b, int nblocks, int kinds, hipStream_t s)
{
    if (kinds & ~0xf) return hipErrorInvalidValue;
    const dim3 grid(nblocks), block(kWave * kPairWaves);
    if (kinds == 4)
        hipLaunchKernelGGL(ksg_batch_kernel<4>, grid, block, 0, s, b);
    else if (kinds == 8)
        hipLaunchKernelGGL(ksg_batch_kernel<8>, grid, block, 0, s, b);
    else if ((kinds & 3) == 0)
        hipLaunchKernelGGL(ksg_batch_kernel<12>, grid, block, 0, s, b);
    else
        hipLaunchKernelGGL(ksg_batch_kernel<15>, grid, block, 0, s, b);
    return hipGetLastError();
}

hipError_t launch_selftest(int* out, hipStream_t s)
{
    hipLaunchKernelGGL(ksg_selftest_kernel, dim3(1), dim3(kWave), 0, s, out);
    return hipGetLastError();
}

}  // namespace gmxcorr
