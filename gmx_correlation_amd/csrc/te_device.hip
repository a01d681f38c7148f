// Transfer entropy (Frenzel-Pompe conditional MI) kernels for gfx950.
//
// TE(a2 -> a1) at lag L (lib/transfer_entropy.cpp:63-221): for query frame
// q < Nv = N - L, the 9-D point [X_a1(q+L), X_a1(q), X_a2(q)] of the raw
// mean-centred coordinates; eps = k-th smallest max-norm distance to the
// other Nv-1 points; counts with strict '<' in the (Xf,Xp), (Yp,Xp) and Xp
// subspaces; TE = mean over q of psi[k] + psi[n_xp+1] - psi[n_xfxp+1]
// - psi[n_ypxp+1] (counts clamped to Nv).
//
// One lane per query frame, 64 queries per wave.  The frame loop is uniform
// across the wave, so frame j's nine coordinates are scalar loads.
// Distances are FP64, computed exactly as the reference does.
//   pass 1  (te_heap_kernel, one wave per block, k <= 128; te_select_kernel,
//           a radix select, for any k) the k + 1 smallest FP32
//           keys fl32(dist), the query's own frame included, in a per-lane
//           9-ary max-heap (rounding to FP32 is monotone, so the root is
//           fl32(eps)): the root and its 9 children in registers, the rest
//           in LDS, (k - 9) * 256 bytes per wave (23 KB at k = 100, seven
//           waves per CU); a sift reads 9 independent children per level.
//           The root goes to HBM;
//   pass 2  (te_count_kernel, four waves per block, no LDS, so occupancy
//           is bounded by registers, not by the heap) eps and the counts
//           together: a frame whose joint key is below
//           the root lies inside eps with all its subspace distances, and a
//           subspace distance whose FP32 key is below (above) the root is
//           below (above) eps.  When exactly one frame has the root key it
//           is the k-th (eps = its FP64 distance); subspace keys of other
//           frames equal to the root (up to 4 frames per query, recorded in
//           the loop) are then compared with eps in FP64;
//   else    (several frames with the root key, or more such frames; rare)
//           eps as the r-th smallest FP64
//           distance among the frames whose key equals the root, r = k -
//           #{key < root}, then the three strict counts in FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sort_network.h"
#include "te_internal.h"

namespace gmxcorr {

namespace {

struct Frame9 {
    double f[9];
};

// Orders a wave's accesses to its own LDS region: the LDS unit executes a
// wave's LDS instructions in issue order, so only compiler reordering has to
// be prevented.
__device__ __forceinline__ void lds_order()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Nine coordinates of frame j: X_a1(j+lag), X_a1(j), X_a2(j).
__device__ __forceinline__ void load_frame(const double* __restrict__ X,
                                           const double* __restrict__ Y, int j, int lag,
                                           double* c)
{
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        c[d] = X[(size_t)(j + lag) * 3 + d];
        c[3 + d] = X[(size_t)j * 3 + d];
        c[6 + d] = Y[(size_t)j * 3 + d];
    }
}

// Inclusive prefix sum over the wave's 64 lanes.
__device__ __forceinline__ int wave_incl_scan(int v)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int up = __shfl_up(v, o, 64);
        if ((int)(threadIdx.x & 63) >= o) v += up;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_max_u(uint32_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t wave_min_u(uint32_t v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
    return v;
}

// Number of set bits of m below this lane.
__device__ __forceinline__ int mbcnt(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t readlane_u(uint32_t v, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

}  // namespace

// Pass 1.  Block b = one wave = query block b % nqblocks of pair b /
// nqblocks (pairs of this launch); root[64 b + lane] = fl32(eps) of its query.
__global__ __launch_bounds__(64) void te_heap_kernel(TeArgs args)
{
    constexpr int kD = 9;              // heap arity
    constexpr int kR = kTeHeapRegs;    // positions 0 .. kR - 1 (root and children) in registers
    static_assert(kR == kD + 1, "the registers hold the root and its children");
    extern __shared__ float heap_lds[];   // positions kR .. k: [k + 1 - kR][64]
    auto heap = [&](int t) -> float& { return heap_lds[(t - kR) * 64 + threadIdx.x]; };
    const int lane = threadIdx.x;
    const int nqb = args.nqblocks;
    const int N = args.N, lag = args.lag, k = args.k, Nv = N - lag;
    // list mode: the queries of pair p listed from position first on
    const bool listed = args.qblocks != nullptr;   // uniform
    long long p;
    int q;
    bool active;
    if (listed) {
        const int2 qbk = args.qblocks[blockIdx.x];
        p = qbk.x;
        const int idx = qbk.y + lane;
        active = idx < (int)args.qcount[p];
        q = active ? (int)args.qlist[(size_t)p * Nv + idx] : 0;
    } else {
        p = blockIdx.x / nqb;
        q = (int)(blockIdx.x % nqb) * 64 + lane;
        active = q < Nv;
    }
    const double* X = args.traj + (size_t)args.a1[p] * N * 3;
    const double* Y = args.traj + (size_t)args.a2[p] * N * 3;
    double qc[9];
    load_frame(X, Y, active ? q : 0, lag, qc);
    auto dist9 = [&](const double* c, double& dxf, double& dxp, double& dyp) {
        dxf = fmax(fmax(fabs(qc[0] - c[0]), fabs(qc[1] - c[1])), fabs(qc[2] - c[2]));
        dxp = fmax(fmax(fabs(qc[3] - c[3]), fabs(qc[4] - c[4])), fabs(qc[5] - c[5]));
        dyp = fmax(fmax(fabs(qc[6] - c[6]), fabs(qc[7] - c[7])), fabs(qc[8] - c[8]));
    };

    // ---- pass 1: k smallest keys (max-heap, root = largest kept) -------
    // The heap keeps kh = k + 1 keys and takes every frame, the query's own
    // included: its key 0 is the smallest, so the (k+1)-th smallest key with
    // it is the k-th smallest without it (also when other frames tie at 0),
    // and the frame loop needs no per-frame self test.  Lanes without a
    // query fill a heap whose root is not used.  The root and its kD
    // children live in registers (r[0] is the root), so a sift's first level
    // needs no LDS and the LDS part is kR slots shorter.
    const int kh = k + 1;
    float r[kR];
    auto reset_heap = [&]() {
#pragma unroll
        for (int t = 0; t < kR; ++t) r[t] = t < kh ? INFINITY : -INFINITY;
        for (int t = kR; t < kh; ++t) heap(t) = INFINITY;
    };
    // largest of kD values: a tree of pairwise maxima, index in ix
    auto argmax = [&](float* ch, int* ix) {
#pragma unroll
        for (int w = 8; w >= 1; w /= 2) {
#pragma unroll
            for (int c = 0; c < w; ++c) {
                if (c + w < kD) {
                    const bool gt = ch[c + w] > ch[c];
                    ch[c] = gt ? ch[c + w] : ch[c];
                    ix[c] = gt ? ix[c + w] : ix[c];
                }
            }
        }
    };
    // key below the root: replace the root and sift down, return the new root
    auto insert = [&](float key) -> float {
        // level 0 -> 1 in registers (slots past kh hold -inf)
        float ch[kD];
        int ix[kD];
#pragma unroll
        for (int c = 0; c < kD; ++c) {
            ch[c] = r[1 + c];
            ix[c] = 1 + c;
        }
        argmax(ch, ix);
        if (ch[0] <= key) {
            r[0] = key;
            return key;
        }
        r[0] = ch[0];
        int pos = ix[0];   // 1 .. kD
        // level 1 -> 2: the children of register slot pos are in LDS
        const int c1 = kD * pos + 1;   // >= kR
        bool done = c1 >= kh;
        if (!done) {
#pragma unroll
            for (int c = 0; c < kD; ++c) {
                const float h = heap(min(c1 + c, kh - 1));
                ch[c] = c1 + c < kh ? h : -INFINITY;
                ix[c] = c1 + c;
            }
            argmax(ch, ix);
            done = ch[0] <= key;
        }
        // register slot pos takes the key or the larger child
        const float up = done ? key : ch[0];
#pragma unroll
        for (int t = 1; t < kR; ++t) r[t] = t == pos ? up : r[t];
        if (done) return r[0];
        pos = ix[0];
        // levels 2 and below in LDS
        for (;;) {
            const int c0 = kD * pos + 1;
            if (c0 >= kh) break;
#pragma unroll
            for (int c = 0; c < kD; ++c) {
                const float h = heap(min(c0 + c, kh - 1));
                ch[c] = c0 + c < kh ? h : -INFINITY;
                ix[c] = c0 + c;
            }
            argmax(ch, ix);
            if (ch[0] <= key) break;
            heap(pos) = ch[0];
            pos = ix[0];
        }
        heap(pos) = key;
        return r[0];
    };
    // ---- a cap on the root from a strided sample of the frames ----------
    // The kTeSampleRank-th smallest key of every stride-th frame.  Keys at
    // or above it are not offered to the heap, which cuts the inserts (and
    // the wave-wide flushes) to less than half.  The heap still ends with
    // the k + 1 smallest keys whenever k + 1 keys lie below the cap (then
    // all of the k + 1 smallest do); a lane where fewer do ends with an
    // unfilled slot, an infinite root, and the wave runs the frame loop
    // again without the cap: a failed lane costs time, not exactness.
    const int stride = max(kTeSampleStride, kTeSampleStride * kh / 101);
    float smp[kTeSampleRank];
#pragma unroll
    for (int u = 0; u < kTeSampleRank; ++u) smp[u] = INFINITY;
    {
        double cs[9];
        load_frame(X, Y, 0, lag, cs);
        for (int j = 0; j < Nv; j += stride) {
            double c[9];
#pragma unroll
            for (int d = 0; d < 9; ++d) c[d] = cs[d];
            load_frame(X, Y, min(j + stride, Nv - 1), lag, cs);
            double a0, b0, e0;
            dist9(c, a0, b0, e0);
            const float key = (float)fmax(fmax(a0, b0), e0);
            // insert into the ascending list, the largest dropped:
            // smp[u] = min(smp[u], max(smp[u - 1], key))
#pragma unroll
            for (int u = kTeSampleRank - 1; u > 0; --u) smp[u] = fminf(smp[u], fmaxf(smp[u - 1], key));
            smp[0] = fminf(smp[0], key);
        }
    }
    float top = INFINITY;
    float cap = smp[kTeSampleRank - 1];
    // Keys below the root wait in kW registers per lane; when some lane is
    // about to run out of them, every lane inserts its waiting keys together
    // (each checked against the root again), so the wave runs the sift with
    // most lanes active instead of for one lane at a time.
    constexpr int kW = 8;
    float wk[kW];
    int nw = 0;
    auto flush = [&]() {
#pragma unroll
        for (int u = 0; u < kW; ++u) {
            if (wk[u] < top) top = fminf(insert(wk[u]), cap);
            wk[u] = INFINITY;
        }
        nw = 0;
    };
    // The buffer is a shift register, newest key in slot 0: a lane that
    // takes a key runs the shift under its own branch as plain moves (an
    // indexed write would be a compare and a select per slot).  At most kW
    // keys wait between flushes, so the shift drops none, and the flush
    // inserts them in any order with the same heap contents.
    auto take = [&](float key) {
#pragma unroll
        for (int u = kW - 1; u > 0; --u) wk[u] = wk[u - 1];
        wk[0] = key;
        ++nw;
    };
    // LDS stage of one chunk of frames after the heap: stream s (0: X at
    // j + lag, 1: X at j, 2: Y at j), frame f, coordinate d at
    // stage[s * 3 kTeStage + 3 f + d]; lane l holds elements l and l + 64
    // of each stream of the next chunk (zero past the last frame)
    constexpr int kE = 3 * kTeStage;        // doubles per stream and chunk
    constexpr int kPer = (kE + 63) / 64;    // per lane
    double* stage = reinterpret_cast<double*>(heap_lds + (size_t)(kh > kR ? kh - kR : 1) * 64);
    double pre[3][kPer];
    auto fetch = [&](int c0) {
        const int n = min(kTeStage, Nv - c0) * 3;
        const double* src[3] = {X + (size_t)(c0 + lag) * 3, X + (size_t)c0 * 3, Y + (size_t)c0 * 3};
#pragma unroll
        for (int st = 0; st < 3; ++st)
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int e = lane + 64 * u;
                pre[st][u] = e < n ? src[st][e] : 0.0;
            }
    };
    auto put = [&]() {
#pragma unroll
        for (int st = 0; st < 3; ++st)
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int e = lane + 64 * u;
                if (e < kE) stage[st * kE + e] = pre[st][u];
            }
    };
    // frames f and f + 1 of the chunk (f even): three 16-byte reads per stream
    auto frames = [&](int f, double* c0v, double* c1v) {
#pragma unroll
        for (int st = 0; st < 3; ++st) {
            const double2* p = reinterpret_cast<const double2*>(stage + st * kE) + (f / 2) * 3;
            const double2 v0 = p[0], v1 = p[1], v2 = p[2];
            c0v[3 * st] = v0.x;
            c0v[3 * st + 1] = v0.y;
            c0v[3 * st + 2] = v1.x;
            c1v[3 * st] = v1.y;
            c1v[3 * st + 1] = v2.x;
            c1v[3 * st + 2] = v2.y;
        }
    };
    for (int attempt = 0;; ++attempt) {
        reset_heap();
        top = cap;
#pragma unroll
        for (int u = 0; u < kW; ++u) wk[u] = INFINITY;
        nw = 0;
        // The frames pass through LDS in chunks of kTeStage: the wave loads
        // a chunk's three coordinate streams (X at j + lag, X at j, Y at j)
        // with coalesced vector loads into registers while it processes the
        // previous chunk, stores them to its LDS stage, and every lane then
        // reads each frame from the same LDS address (a broadcast).  Read
        // frame by frame through the scalar cache instead, the streams of
        // the waves sharing a scalar cache missed it on most lines and kept
        // it busy 90% of the time.  Two frames per iteration (two
        // independent FP64 chains); a lane takes at most two keys per
        // iteration, so the buffer is flushed once some lane holds kW - 1.
        fetch(0);
        for (int c0 = 0; c0 < Nv; c0 += kTeStage) {
            lds_order();
            put();
            lds_order();
            if (c0 + kTeStage < Nv) fetch(c0 + kTeStage);   // uniform
            const int nf = min(kTeStage, Nv - c0);
            for (int f = 0; f < nf; f += 2) {
                double c0v[9], c1v[9];
                frames(f, c0v, c1v);
                double a0, b0, e0, a1, b1, e1;
                dist9(c0v, a0, b0, e0);
                dist9(c1v, a1, b1, e1);
                const float key0 = (float)fmax(fmax(a0, b0), e0);
                const float key1 = (float)fmax(fmax(a1, b1), e1);
                if (key0 < top) take(key0);
                if (f + 1 < nf && key1 < top) take(key1);
                if (__ballot(nw >= kW - 1)) flush();   // uniform over the wave
            }
        }
        flush();
        // an unfilled slot leaves the root infinite: fewer than k + 1 keys
        // below a finite cap (the plain run, cap = inf, always fills it:
        // Nv >= k + 1 finite keys)
        const bool short_ = active && r[0] == INFINITY && cap < INFINITY;
        if (!__ballot(short_)) break;   // uniform over the wave
        cap = INFINITY;
    }
    // the root of query q of pair p at root[(p * nqb) * 64 + q]
    if (!listed) {
        args.root[(size_t)blockIdx.x * 64 + lane] = r[0];
        return;
    }
    if (!args.eps64) {
        if (active) args.root[(size_t)p * nqb * 64 + q] = r[0];
        return;
    }
    // The prefix path's listed queries: the exact FP64 radius, the r-th
    // smallest FP64 distance among the frames whose key is the root, r =
    // k + 1 - #{key < root} (fl32 is monotone: every frame below the root key
    // is nearer).  A second pass over the frames (the same LDS stage) counts
    // the keys below the root and keeps the distances of the first kRec
    // frames at it; a lane with more goes through their distinct values.
    constexpr int kRec = 4;
    const float rt = r[0];
    int clt = 0, nrec = 0;
    double rec[kRec];
#pragma unroll
    for (int u = 0; u < kRec; ++u) rec[u] = INFINITY;
    fetch(0);
    for (int c0 = 0; c0 < Nv; c0 += kTeStage) {
        lds_order();
        put();
        lds_order();
        if (c0 + kTeStage < Nv) fetch(c0 + kTeStage);   // uniform
        const int nf = min(kTeStage, Nv - c0);
        for (int f = 0; f < nf; f += 2) {
            double c0v[9], c1v[9];
            frames(f, c0v, c1v);
            double a0, b0, e0, a1, b1, e1;
            dist9(c0v, a0, b0, e0);
            dist9(c1v, a1, b1, e1);
            const double d0 = fmax(fmax(a0, b0), e0), d1 = fmax(fmax(a1, b1), e1);
            const float key0 = (float)d0, key1 = f + 1 < nf ? (float)d1 : INFINITY;
            clt += (key0 < rt) + (key1 < rt);
            if (__ballot(key0 == rt || key1 == rt)) {   // uniform, rare
#pragma unroll
                for (int u = 0; u < kRec; ++u) {
                    rec[u] = key0 == rt && nrec == u ? d0 : rec[u];
                    rec[u] = key1 == rt && nrec + (key0 == rt) == u ? d1 : rec[u];
                }
                nrec += (key0 == rt) + (key1 == rt);
            }
        }
    }
    if (!active) return;
    const int rk = k + 1 - clt;   // 1 <= rk <= nrec
    double eps = INFINITY;
    if (nrec <= kRec) {
        // the rk-th smallest of the recorded distances
        double v[kRec];
#pragma unroll
        for (int u = 0; u < kRec; ++u) v[u] = rec[u];
#pragma unroll
        for (int a = 0; a < kRec; ++a)
#pragma unroll
            for (int b = a + 1; b < kRec; ++b) {
                const double lo = fmin(v[a], v[b]), hi = fmax(v[a], v[b]);
                v[a] = lo;
                v[b] = hi;
            }
#pragma unroll
        for (int u = 0; u < kRec; ++u) eps = u == rk - 1 ? v[u] : eps;
    } else {
        // the distinct FP64 values of the root-key frames upwards
        double prev = -1.0;
        int cum = 0;
        for (;;) {
            double vmin = INFINITY;
            int m = 0;
            for (int j = 0; j < Nv; ++j) {
                double c[9];
                load_frame(X, Y, j, lag, c);
                double a, b, e;
                dist9(c, a, b, e);
                const double dd = fmax(fmax(a, b), e);
                if ((float)dd == rt && dd > prev) {
                    if (dd < vmin) { vmin = dd; m = 1; }
                    else if (dd == vmin) ++m;
                }
            }
            if (m == 0 || cum + m >= rk) { eps = vmin; break; }
            cum += m;
            prev = vmin;
        }
    }
    args.eps64[(size_t)p * nqb * 64 + q] = eps;
}

// Pass 1 for any k (the heap kernel's LDS stops at k = 128): per query, the
// (k+1)-th smallest FP32 key of all Nv frames, the query's own included (the
// heap's root), by a radix select over the key's bit pattern, one pass over
// the frames per digit.  Keys are non-negative floats, so their bit patterns
// order as unsigned integers, and bit 31 is clear: the digits are bits
// 30..20 (2048 bins), 19..10 and 9..0 (1024 bins each).  Block b = one wave
// = query block b % nqblocks of pair b / nqblocks; the wave takes its 64
// queries one after another, its lanes striding over the frames, and
// counts in an LDS histogram the frames whose key matches the digits chosen
// so far.
__global__ __launch_bounds__(64) void te_select_kernel(TeArgs args)
{
    __shared__ unsigned hist[2048];
    const int lane = threadIdx.x;
    const int nqb = args.nqblocks;
    const long long p = blockIdx.x / nqb;
    const int qb = blockIdx.x % nqb;
    const int N = args.N, lag = args.lag, Nv = N - lag;
    const double* X = args.traj + (size_t)args.a1[p] * N * 3;
    const double* Y = args.traj + (size_t)args.a2[p] * N * 3;
    const int nq = min(64, Nv - qb * 64);   // queries of this block (uniform)
    float mine = INFINITY;                  // root of query qb * 64 + lane
    for (int qi = 0; qi < nq; ++qi) {
        double qc[9];
        load_frame(X, Y, qb * 64 + qi, lag, qc);
        unsigned prefix = 0u, mask = 0u;
        unsigned want = (unsigned)args.k + 1u;   // 1-based rank among the matching keys
#pragma unroll 1
        for (int pass = 0; pass < 3; ++pass) {
            const int sh = pass == 0 ? 20 : pass == 1 ? 10 : 0;
            const int nb = pass == 0 ? 2048 : 1024;
            for (int b = lane; b < nb; b += 64) hist[b] = 0u;
            __syncthreads();
            for (int j = lane; j < Nv; j += 64) {
                double c[9];
                load_frame(X, Y, j, lag, c);
                double m = 0.0;
#pragma unroll
                for (int d = 0; d < 9; ++d) m = fmax(m, fabs(qc[d] - c[d]));
                const unsigned u = __float_as_uint((float)m);
                if ((u & mask) == prefix) atomicAdd(&hist[(u >> sh) & (unsigned)(nb - 1)], 1u);
            }
            __syncthreads();
            // the digit whose bin holds the want-th matching key: lane l sums
            // bins [l * per, (l + 1) * per) (read in a rotated order, fewer
            // bank conflicts), an inclusive scan over the lanes finds the
            // lane, which then walks its bins in order
            const int per = nb / 64;
            unsigned sum = 0u;
            for (int i = 0; i < per; ++i) sum += hist[lane * per + ((i + lane) & (per - 1))];
            unsigned inc = sum;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned t = __shfl_up(inc, o, 64);
                if (lane >= o) inc += t;
            }
            const int wl = __ffsll((unsigned long long)__ballot(inc >= want)) - 1;
            unsigned below = __shfl(inc - sum, wl, 64);
            int digit = wl * per;
            for (int i = 0; i < per; ++i) {   // uniform: every lane walks lane wl's bins
                const unsigned h = hist[wl * per + i];
                if (below + h >= want) { digit = wl * per + i; break; }
                below += h;
            }
            want -= below;
            prefix |= (unsigned)digit << sh;
            mask |= (unsigned)(nb - 1) << sh;
            __syncthreads();   // the histogram is cleared next
        }
        if (lane == qi) mine = __uint_as_float(prefix);
    }
    args.root[(size_t)blockIdx.x * 64 + lane] = mine;
}

// Pass 2 and the TE sum for G ordered pairs with one target atom a1
// (TE(a2_u -> a1), u < G): the target's two 3-D distances of a frame (a1's
// future and present) are computed once for the G pairs, the source's
// present distance once per pair.  Lane = query frame, the frame loop
// uniform over the wave (scalar loads of the frame's coordinates).
//
// Every frame, the query's own included (removed after the loop), adds its
// strict FP32 comparisons with the pair's root.  A frame with a key equal
// to the root, or with a subspace key equal to it, is handled under a
// wave-uniform branch (rarely taken): frames with the root key are
// gathered, frames of another key with a subspace key equal to the root
// recorded (the first kAmb per lane), and both resolved in FP64 after the
// loop.  fl32 is monotone, so the keys are maxima of the converted 3-D
// distances.  The partial sum of pair p0 + u goes to
// partial[(p0 + u) * nqblocks + qb].
template <int G>
__device__ __forceinline__ void te_count_group(const TeArgs& args, long long p0, int qb, int lane)
{
    // per pair and lane: the last root-key frame (its distances are
    // recomputed after the loop) and the first kAmb frames of another key
    // with a subspace key equal to the root; more of either sends the lane
    // to the FP64 pass
    constexpr int kAmb = 2;
    const int N = args.N, lag = args.lag, k = args.k, Nv = N - lag;
    const int nqb = args.nqblocks;
    const double* X = args.traj + (size_t)args.a1[p0] * N * 3;
    const double* Y[G];
#pragma unroll
    for (int u = 0; u < G; ++u) Y[u] = args.traj + (size_t)args.a2[p0 + u] * N * 3;
    const int q = qb * 64 + lane;
    const bool active = q < Nv;
    const int qs = active ? q : 0;
    double qx[6], qy[G][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        qx[d] = X[(size_t)(qs + lag) * 3 + d];
        qx[3 + d] = X[(size_t)qs * 3 + d];
#pragma unroll
        for (int u = 0; u < G; ++u) qy[u][d] = Y[u][(size_t)qs * 3 + d];
    }
    float root[G];
    int nxfxp[G], nypxp[G], nxp[G], gcnt[G], gj[G], na[G];
    int aj[G][kAmb];
#pragma unroll
    for (int u = 0; u < G; ++u) {
        root[u] = args.root[((size_t)(p0 + u) * nqb + qb) * 64 + lane];
        nxfxp[u] = nypxp[u] = nxp[u] = gcnt[u] = gj[u] = na[u] = 0;
#pragma unroll
        for (int v = 0; v < kAmb; ++v) aj[u][v] = 0;
    }
    // frame j: a1 at j + lag and j, each source at j, read as 32-bit byte
    // offsets from the atoms' base pointers (rows of 24 bytes; the host
    // keeps N * 24 below 2^32); the next frame is loaded while this one is
    // processed
    auto at = [](const double* base, uint32_t off) {
        return reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + off);
    };
    auto load = [&](int j, double* f, double (*y)[3]) {
        const uint32_t of = (uint32_t)(j + lag) * 24u, op = (uint32_t)j * 24u;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            f[d] = at(X, of)[d];
            f[3 + d] = at(X, op)[d];
#pragma unroll
            for (int u = 0; u < G; ++u) y[u][d] = at(Y[u], op)[d];
        }
    };
    auto maxnorm3 = [](const double* a, const double* b) {
        return fmax(fmax(fabs(a[0] - b[0]), fabs(a[1] - b[1])), fabs(a[2] - b[2]));
    };
    double nf[6], ny[G][3];
    load(0, nf, ny);
    for (int j = 0; j < Nv; ++j) {
        double f[6], y[G][3];
#pragma unroll
        for (int d = 0; d < 6; ++d) f[d] = nf[d];
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int d = 0; d < 3; ++d) y[u][d] = ny[u][d];
        load(min(j + 1, Nv - 1), nf, ny);
        const float fxf = (float)maxnorm3(qx, f), fxp = (float)maxnorm3(qx + 3, f + 3);
        const float k1 = fmaxf(fxf, fxp);
        float fyp[G];
        bool odd[G];
        bool any = false;
#pragma unroll
        for (int u = 0; u < G; ++u) {
            fyp[u] = (float)maxnorm3(qy[u], y[u]);
            const float k2 = fmaxf(fyp[u], fxp);
            const float rt = root[u];
            nxfxp[u] += k1 < rt;
            nypxp[u] += k2 < rt;
            nxp[u] += fxp < rt;
            // the key max(k1, fyp) or a subspace key equal to the root
            odd[u] = (k1 == rt) | (k2 == rt) | (fxp == rt) | (fyp[u] == rt);
            any |= odd[u];
        }
        if (__builtin_amdgcn_ballot_w64(any)) {   // uniform, rare
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (odd[u]) {
                    const float rt = root[u];
                    if (fmaxf(k1, fyp[u]) == rt) {
                        if (j != q) {
                            ++gcnt[u];
                            gj[u] = j;
                        }
                    } else if ((k1 == rt) | (fmaxf(fyp[u], fxp) == rt) | (fxp == rt)) {
                        // a subspace key equal to the root in a frame of another key
#pragma unroll
                        for (int v = 0; v < kAmb; ++v) aj[u][v] = na[u] == v ? j : aj[u][v];
                        ++na[u];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
        const float rt = root[u];
        double qc[9];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            qc[d] = qx[d];
            qc[3 + d] = qx[3 + d];
            qc[6 + d] = qy[u][d];
        }
        auto dist9 = [&](const double* c, double& dxf, double& dxp, double& dyp) {
            dxf = maxnorm3(qc, c);
            dxp = maxnorm3(qc + 3, c + 3);
            dyp = maxnorm3(qc + 6, c + 6);
        };
        // the self frame (all distances 0) added one to each count below a
        // positive root (at a zero root its key equals the root: not
        // counted, not recorded)
        int cx = nxfxp[u], cy = nypxp[u], cp = nxp[u];
        if (rt > 0.0f) {
            --cx;
            --cy;
            --cp;
        }
        if (!active) {
            // no query: nothing to count
        } else if (gcnt[u] == 1 && na[u] <= kAmb) {
            // the one root-key frame is the k-th: eps = its FP64 distance.
            // Its subspace keys below the root were counted in the loop; its
            // subspace distances are compared with eps instead, and so are
            // the recorded frames' subspace distances whose keys equal the
            // root (their keys below the root are counted already)
            double c[9];
            load_frame(X, Y[u], gj[u], lag, c);
            double dxf, dxp, dyp;
            dist9(c, dxf, dxp, dyp);
            const double eps = fmax(fmax(dxf, dxp), dyp);
            const double s1 = fmax(dxf, dxp), s2 = fmax(dyp, dxp);
            cx += (s1 < eps) - ((float)s1 < rt);
            cy += (s2 < eps) - ((float)s2 < rt);
            cp += (dxp < eps) - ((float)dxp < rt);
            for (int v = 0; v < kAmb; ++v) {
                if (v < na[u]) {
                    load_frame(X, Y[u], aj[u][v], lag, c);
                    dist9(c, dxf, dxp, dyp);
                    const double t1 = fmax(dxf, dxp), t2 = fmax(dyp, dxp);
                    cx += (float)t1 == rt && t1 < eps;
                    cy += (float)t2 == rt && t2 < eps;
                    cp += (float)dxp == rt && dxp < eps;
                }
            }
        } else {
            // FP32 ties at the root: eps in FP64, then the counts in FP64.
            // clt = #{j != q: fl32(dist) < root}; eps is the r-th smallest
            // FP64 distance among the root-key frames, r = k - clt
            // (1 <= r <= gcnt), found by walking their distinct values
            // upwards
            int clt = 0;
            for (int j = 0; j < Nv; ++j) {
                double c[9];
                load_frame(X, Y[u], j, lag, c);
                double dxf, dxp, dyp;
                dist9(c, dxf, dxp, dyp);
                clt += j != q && (float)fmax(fmax(dxf, dxp), dyp) < rt;
            }
            const int r = k - clt;
            double eps = 0.0, prev = -1.0;
            int cum = 0;
            for (;;) {
                double vmin = INFINITY;
                int m = 0;
                for (int j = 0; j < Nv; ++j) {
                    if (j == q) continue;
                    double c[9];
                    load_frame(X, Y[u], j, lag, c);
                    double dxf, dxp, dyp;
                    dist9(c, dxf, dxp, dyp);
                    const double dd = fmax(fmax(dxf, dxp), dyp);
                    if ((float)dd == rt && dd > prev) {
                        if (dd < vmin) { vmin = dd; m = 1; }
                        else if (dd == vmin) ++m;
                    }
                }
                if (m == 0 || cum + m >= r) { eps = vmin; break; }
                cum += m;
                prev = vmin;
            }
            cx = cy = cp = 0;
            for (int j = 0; j < Nv; ++j) {
                double c[9];
                load_frame(X, Y[u], j, lag, c);
                double dxf, dxp, dyp;
                dist9(c, dxf, dxp, dyp);
                if (j == q) continue;
                cx += fmax(dxf, dxp) < eps;
                cy += fmax(dyp, dxp) < eps;
                cp += dxp < eps;
            }
        }
        double te = 0.0;
        if (active) {
            const double* psi = args.psi;
            te = psi[k] + psi[min(cp, Nv) + 1] - psi[min(cx, Nv) + 1] - psi[min(cy, Nv) + 1];
        }
        // wave sum, fixed order
        for (int o = 32; o >= 1; o >>= 1) te += __shfl_xor(te, o, 64);
        if (lane == 0) args.partial[(size_t)(p0 + u) * nqb + qb] = te;
    }
}

// Wave w of block b runs item 4b + w = query block item % nqblocks of
// pair group item / nqblocks, the launch's pairs kTeCountGroup at a time:
// a group whose pairs share their target atom runs together, any other
// (the last, short group of a launch, or pairs of different targets) one
// pair at a time.
__global__ __launch_bounds__(256) void te_count_kernel(TeArgs args)
{
    constexpr int G = kTeCountGroup;
    const int lane = threadIdx.x & 63;
    const long long item = (long long)blockIdx.x * 4 +
                           __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nqb = args.nqblocks;
    const long long ngroups = (args.npairs + G - 1) / G;
    if (item >= ngroups * nqb) return;
    const long long g = item / nqb;
    const int qb = (int)(item - g * nqb);
    const long long p0 = g * G;
    bool shared = p0 + G <= args.npairs;
    if (shared) {
        const int t = args.a1[p0];
#pragma unroll
        for (int u = 1; u < G; ++u) shared = shared && args.a1[p0 + u] == t;
    }
    if (shared) {
        te_count_group<G>(args, p0, qb, lane);
    } else {
        const long long pe = p0 + G < args.npairs ? p0 + G : args.npairs;
        for (long long p = p0; p < pe; ++p) te_count_group<1>(args, p, qb, lane);
    }
}

// ---------------------------------------------------------------------------
// Prefix path, build: one block of 512 threads per row (target slot t,
// query q), thread `tid` holding frames tid + 512 e (e < E) in registers
// (coalesced coordinate loads).  Keys are fl32 of the FP64 target-only
// distance M(q, j) = max(|X(q+L) - X(j+L)|, |X(q) - X(j)|).  The query's own
// frame (key 0) is always stored; of the other frames, those below the
// value v* of rank kTePrefix - 2 (a radix select over v - lo, 8 bits per
// pass from the top set bit of hi - lo) and as many frames at v* as that
// rank leaves.  Which of several frames at v* are kept does not matter:
// every frame left out has a key >= v* = H, the bound the select checks
// against.  With Nv <= kTePrefix every frame is stored and the rest padded
// with +inf keys (frame 0).  The kTePrefix (key, frame) pairs are sorted
// two per thread by the register network of sort_network.h.
// ---------------------------------------------------------------------------
template <int E>
__global__ __launch_bounds__(512) void te_prefix_build_kernel(TeArgs args, TePrefixArgs pa)
{
    constexpr int BT = 512;
    constexpr uint32_t kNone = 0xffffffffu;   // the own frame and frames past Nv
    static_assert(kTePrefix == 2 * BT, "two stored keys per thread");
    __shared__ uint64_t sk[kTePrefix];
    __shared__ uint32_t hist[256];
    __shared__ uint32_t wlo[BT / 64], whi[BT / 64];
    __shared__ uint32_t sel[4];   // chosen bin, rank left within it, slot and tie counters
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int N = args.N, lag = args.lag, Nv = N - lag;
    const long long nrows = pa.ntargets * (long long)Nv;
    for (long long row = blockIdx.x; row < nrows; row += gridDim.x) {   // uniform
        const long long t = row / Nv;
        const int q = (int)(row - t * Nv);
        const double* X = args.traj + (size_t)pa.targets[t] * N * 3;
        double qf[3], qp[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            qf[d] = X[(size_t)(q + lag) * 3 + d];
            qp[d] = X[(size_t)q * 3 + d];
        }
        uint32_t v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int j = tid + BT * e;
            v[e] = kNone;
            if (j < Nv && j != q) {
                const double* f = X + (size_t)(j + lag) * 3;
                const double* c = X + (size_t)j * 3;
                const double a = fmax(fmax(fabs(qf[0] - f[0]), fabs(qf[1] - f[1])), fabs(qf[2] - f[2]));
                const double b = fmax(fmax(fabs(qp[0] - c[0]), fabs(qp[1] - c[1])), fabs(qp[2] - c[2]));
                v[e] = __float_as_uint((float)fmax(a, b));
            }
        }
        // v* and the number of frames at v* to keep (rank + 1)
        uint32_t vstar = kNone;
        int rank = -1;
        if (Nv - 1 > kTePrefix - 1) {
            uint32_t mn = kNone, mx = 0u;
#pragma unroll
            for (int e = 0; e < E; ++e)
                if (v[e] != kNone) {
                    mn = min(mn, v[e]);
                    mx = max(mx, v[e]);
                }
            mn = wave_min_u(mn);
            mx = wave_max_u(mx);
            if (lane == 0) {
                wlo[wv] = mn;
                whi[wv] = mx;
            }
            __syncthreads();
            uint32_t lo = wlo[0], hi = whi[0];
#pragma unroll
            for (int u = 1; u < BT / 64; ++u) {
                lo = min(lo, wlo[u]);
                hi = max(hi, whi[u]);
            }
            // keys are finite floats (< 2^31 as bits): hi - lo < 2^31
            const int nb = hi > lo ? 32 - __builtin_clz(hi - lo) : 0;
            uint32_t pre = 0u;
            uint32_t rk = kTePrefix - 2;   // 0-based among the other frames
            for (int top = nb; top > 0;) {   // uniform
                const int shift = max(top - 8, 0);
                if (tid < 256) hist[tid] = 0u;
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const uint32_t dv = v[e] - lo;
                    if (v[e] <= hi && (dv >> top) == pre) atomicAdd(&hist[(dv >> shift) & 255u], 1u);
                }
                __syncthreads();
                if (tid < 64) {
                    const uint4 h = reinterpret_cast<const uint4*>(hist)[lane];
                    const uint32_t own = h.x + h.y + h.z + h.w;
                    const uint32_t incl = (uint32_t)wave_incl_scan((int)own);
                    const uint32_t excl = incl - own;
                    if (excl <= rk && rk < incl) {
                        uint32_t r = rk - excl, bin = 4u * lane;
                        if (r >= h.x) { r -= h.x; ++bin;
                            if (r >= h.y) { r -= h.y; ++bin;
                                if (r >= h.z) { r -= h.z; ++bin; } } }
                        sel[0] = bin;
                        sel[1] = r;
                    }
                }
                __syncthreads();
                pre = (pre << (top - shift)) | sel[0];
                rk = sel[1];
                top = shift;
                __syncthreads();   // sel and hist are rewritten by the next pass
            }
            vstar = pre + lo;
            rank = (int)rk;
        }
        if (tid == 0) {
            sel[2] = 0u;
            sel[3] = 0u;
            sk[0] = (uint64_t)(uint32_t)q;   // the own frame, key 0
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            bool take = v[e] < vstar;
            if (v[e] == vstar && v[e] != kNone) take = (int)atomicAdd(&sel[3], 1u) <= rank;
            if (take) {
                const uint32_t slot = atomicAdd(&sel[2], 1u);
                if (slot < (uint32_t)(kTePrefix - 1))
                    sk[1 + slot] = ((uint64_t)v[e] << 32) | (uint32_t)(tid + BT * e);
            }
        }
        __syncthreads();
        const int filled = 1 + (int)min(sel[2], (uint32_t)(kTePrefix - 1));
        for (int i = filled + tid; i < kTePrefix; i += BT) sk[i] = (uint64_t)0x7f800000u << 32;
        __syncthreads();
        uint64_t pk[2] = {sk[2 * tid], sk[2 * tid + 1]};
        sortnet::sort_blocked<uint64_t, 2, BT>(pk, sk, tid);
        const size_t base = (size_t)row * kTePrefix;
        reinterpret_cast<uint2*>(pa.tv + base)[tid] =
            make_uint2((uint32_t)(pk[0] >> 32), (uint32_t)(pk[1] >> 32));
        reinterpret_cast<uint32_t*>(pa.ti + base)[tid] =
            (uint32_t)(pk[0] & 0xffffu) | ((uint32_t)(pk[1] & 0xffffu) << 16);
        __syncthreads();   // sk, hist and sel are rewritten by the next row
    }
}

// ---------------------------------------------------------------------------
// Prefix path, select: the T-th smallest of a wave's NK * 64 joint keys
// (key[m] at lane l is prefix position l + 64 m), T = k + 1 (the own frame
// included), known to lie in [L, H) when it is below H:
//   histogram  256 bins over [L, H) in the key's bit pattern, keys below L
//              counted in bin 0 by ballots, keys >= H left out (the last bin
//              holds a count that always reaches T);
//   select     the target bin's keys compacted to LDS (hist, the wave's 256
//              words), ballot radix select.
// Returns false when the T-th key is not below H, lies in bin 0, or its bin
// holds more than 64 keys.
// ---------------------------------------------------------------------------
template <int NK>
__device__ __forceinline__ bool prefix_select(const uint32_t (&key)[NK], uint32_t L, uint32_t H,
                                              int T, uint32_t* hist, int lane, uint32_t& eps_key)
{
    constexpr int kB = 256;
    const uint32_t span1 = H > L ? H - L - 1u : 0u;
    int sh = max(0, 24 - (int)__builtin_clz(span1 | 1u));
    if ((span1 >> sh) > (uint32_t)(kB - 3)) ++sh;
    const uint32_t Lm = L - min(L, 1u << sh);
    const uint32_t B1 = Lm + (1u << sh);
    int cb = 0;
#pragma unroll
    for (int m = 0; m < NK; ++m) cb += __popcll(__ballot(key[m] < B1));
    reinterpret_cast<uint4*>(hist)[lane] =
        make_uint4(lane == 0 ? (uint32_t)cb : 0u, 0u, 0u, lane == 63 ? 4096u : 0u);
    lds_order();
    const uint32_t span = H > B1 ? H - B1 : 0u;
#pragma unroll
    for (int m = 0; m < NK; ++m) {
        const uint32_t dk = key[m] - B1;   // keys below B1 wrap above the span
        if (dk < span) atomicAdd(&hist[(dk >> sh) + 1u], 1u);
    }
    lds_order();
    // bin holding the T-th key: lane l owns bins 4l .. 4l + 3
    const uint4 hv = reinterpret_cast<const uint4*>(hist)[lane];
    const int hs = (int)((hv.x + hv.y) + (hv.z + hv.w));
    const int incl = wave_incl_scan(hs);
    const int excl = incl - hs;
    const uint64_t bal = __ballot((uint32_t)(T - 1 - excl) < (uint32_t)hs);
    const int owner = bal ? __ffsll((unsigned long long)bal) - 1 : 0;
    const int r = T - __builtin_amdgcn_readlane(excl, owner);
    const int h0 = __builtin_amdgcn_readlane((int)hv.x, owner);
    const int h1 = __builtin_amdgcn_readlane((int)hv.y, owner);
    const int h2 = __builtin_amdgcn_readlane((int)hv.z, owner);
    const int h3 = __builtin_amdgcn_readlane((int)hv.w, owner);
    int sub = 0, below = 0, pop = h0;
    if (r > h0 && r <= h0 + h1) { sub = 1; below = h0; pop = h1; }
    else if (r > h0 + h1 && r <= h0 + h1 + h2) { sub = 2; below = h0 + h1; pop = h2; }
    else if (r > h0 + h1 + h2) { sub = 3; below = h0 + h1 + h2; pop = h3; }
    const int b1 = owner * 4 + sub;
    const int r0 = r - below;
    lds_order();
    if (b1 == 0 || b1 == kB - 1 || pop > 64) return false;
    const uint32_t lo = Lm + ((uint32_t)b1 << sh);
    const uint32_t lim = min(1u << sh, H > lo ? H - lo : 0u);
    int nstored = 0;
#pragma unroll
    for (int m = 0; m < NK; ++m) {
        const bool inbin = key[m] - lo < lim;   // keys below lo wrap
        const uint64_t cm = __ballot(inbin);
        if (cm) {
            if (inbin) hist[nstored + mbcnt(cm)] = key[m] - lo;
            nstored += __popcll(cm);
        }
    }
    lds_order();
    const bool act = lane < pop;
    const uint32_t low = act ? hist[lane] : 0xffffffffu;
    uint64_t live = __ballot(act);
    uint32_t pre = 0u;
    int rr = r0;
    for (int bit = sh - 1; bit >= 0 && __popcll(live) > 1; --bit) {
        const uint64_t z = __ballot((low >> bit) == (pre << 1));
        const int c0 = __popcll(z);
        if (rr > c0) {
            rr -= c0;
            pre = (pre << 1) | 1u;
            live &= ~z;
        } else {
            pre <<= 1;
            live = z;
        }
    }
    const uint32_t sv = readlane_u(low, __ffsll((unsigned long long)live) - 1);
    eps_key = lo + sv;
    // the r0-th smallest of the bin: fewer than r0 below it, r0 or more at
    // or below it
    const int lt = __popcll(__ballot(low < sv)), le = __popcll(__ballot(low <= sv));
    lds_order();
    return nstored == pop && lt < r0 && r0 <= le && eps_key < H;
}

// ---------------------------------------------------------------------------
// Prefix path, select kernel: block = query block qb of pair p (four waves,
// 16 queries each, one after another).  Lane l holds prefix positions
// l + 64 m: the target-only key from the store and the source's distance of
// the same frame, FP64 from the gathered coordinates (one 16-byte and one
// 8-byte load), joined as key = max(fl32(M), fl32(dY)).  L = the store's key
// at position k, a lower bound of the T-th joint key.  Two stages:
// positions 0 .. 511 with H = the store's key at 511, and, when the T-th key
// of those is not below it, all 1024 with H = the store's last key.  With
// every frame stored (Nv <= the stage's positions) H is one above the
// largest joint key.  The root goes to root[p nqb 64 + q]; a query that
// neither stage resolves is appended to its pair's list for the heap kernel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void te_prefix_select_kernel(TeArgs args, TePrefixArgs pa)
{
    constexpr int kM = kTePrefix / 64;   // keys per lane, both stages
    constexpr int kA = kM / 2;           // keys per lane of the first stage
    typedef double dpair __attribute__((ext_vector_type(2), aligned(8)));
    __shared__ __align__(16) uint32_t hist_all[4][256];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    uint32_t* hist = hist_all[w];
    const int nqb = args.nqblocks;
    const long long p = blockIdx.x / nqb;
    const int qb = (int)(blockIdx.x - p * nqb);
    const int N = args.N, lag = args.lag, k = args.k, Nv = N - lag;
    const int T = k + 1;
    const double* Y = args.traj + (size_t)args.a2[p] * N * 3;
    const long long slot = pa.tslot[p];
    for (int i = 0; i < 16; ++i) {
        const int q = qb * 64 + w * 16 + i;
        if (q >= Nv) break;   // uniform
        const size_t rbase = ((size_t)slot * Nv + q) * kTePrefix;
        const double y0 = Y[(size_t)q * 3], y1 = Y[(size_t)q * 3 + 1], y2 = Y[(size_t)q * 3 + 2];
        uint32_t key[kM];
        // positions m * 64 + lane for m in [m0, m1): store keys joined with
        // the source's distances
        auto join = [&](int m0, int m1) {
            uint32_t fr[kM];
#pragma unroll
            for (int m = 0; m < kM; ++m) {
                if (m < m0 || m >= m1) continue;
                key[m] = pa.tv[rbase + lane + 64 * m];
                fr[m] = pa.ti[rbase + lane + 64 * m];
            }
#pragma unroll
            for (int m = 0; m < kM; ++m) {
                if (m < m0 || m >= m1) continue;
                const double* yj = Y + (size_t)fr[m] * 3;
                const dpair xy = *reinterpret_cast<const dpair*>(yj);
                const double d = fmax(fmax(fabs(y0 - xy.x), fabs(y1 - xy.y)), fabs(y2 - yj[2]));
                key[m] = max(key[m], __float_as_uint((float)d));
            }
        };
        // one above the largest joint key of the real frames below position n
        auto above_all = [&](int nk, int n) {
            uint32_t mx = 0u;
#pragma unroll
            for (int m = 0; m < kM; ++m)
                if (m < nk && lane + 64 * m < n) mx = max(mx, key[m]);
            return wave_max_u(mx) + 1u;
        };
        // L and the stages' bounds are the store's target-only keys (the
        // joined keys are larger)
        const int km = k >> 6;   // k <= 128: m = 0, 1 or 2
        const uint32_t L = readlane_u(pa.tv[rbase + (size_t)(64 * km + lane)], k & 63);
        const uint32_t storeA = readlane_u(pa.tv[rbase + (size_t)(64 * (kA - 1) + lane)], 63);
        const uint32_t storeB = readlane_u(pa.tv[rbase + (size_t)(64 * (kM - 1) + lane)], 63);
        uint32_t eps_key = 0u;
        join(0, kA);
        const uint32_t HA = Nv <= 64 * kA ? above_all(kA, Nv) : storeA;
        // the first stage's keys as an array of their own (registers)
        uint32_t ka[kA];
#pragma unroll
        for (int m = 0; m < kA; ++m) ka[m] = key[m];
        bool ok = prefix_select<kA>(ka, L, HA, T, hist, lane, eps_key);
        int nk = kA;   // key arrays joined
        if (!ok && Nv > 64 * kA) {   // uniform
            join(kA, kM);
            const uint32_t HB = Nv <= kTePrefix ? above_all(kM, Nv) : storeB;
            ok = prefix_select<kM>(key, L, HB, T, hist, lane, eps_key);
            nk = kM;
        }
        // The exact radius: the r-th smallest FP64 joint distance among the
        // entries whose key is the root, r = T - #{key < root} (fl32 is
        // monotone, so every entry below the root key is nearer).  All of
        // them lie in the joined positions (their keys are below the bound).
        double eps = 0.0;
        if (ok) {
            int clt = 0, ctie = 0;
#pragma unroll
            for (int m = 0; m < kM; ++m) {
                if (m >= nk) continue;
                clt += __popcll(__ballot(key[m] < eps_key));
                ctie += __popcll(__ballot(key[m] == eps_key));
            }
            int r = T - clt;
            ok = ctie <= 64 && r >= 1 && r <= ctie;
            if (ok) {
                const double* X = args.traj + (size_t)args.a1[p] * N * 3;
                double qc[6];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    qc[d] = X[(size_t)(q + lag) * 3 + d];
                    qc[3 + d] = X[(size_t)q * 3 + d];
                }
                double* tied = reinterpret_cast<double*>(hist);   // up to 64
                int n = 0;
#pragma unroll
                for (int m = 0; m < kM; ++m) {
                    if (m >= nk) continue;
                    const bool t = key[m] == eps_key;
                    const uint64_t tm = __ballot(t);
                    if (tm) {
                        if (t) {
                            const size_t j = pa.ti[rbase + lane + 64 * m];
                            const double* f = X + (j + lag) * 3;
                            const double* c = X + j * 3;
                            const double* yj = Y + j * 3;
                            double d = 0.0;
#pragma unroll
                            for (int e = 0; e < 3; ++e) {
                                d = fmax(d, fabs(qc[e] - f[e]));
                                d = fmax(d, fabs(qc[3 + e] - c[e]));
                            }
                            d = fmax(d, fmax(fmax(fabs(y0 - yj[0]), fabs(y1 - yj[1])), fabs(y2 - yj[2])));
                            tied[n + mbcnt(tm)] = d;
                        }
                        n += __popcll(tm);
                    }
                }
                lds_order();
                double v = lane < n ? tied[lane] : INFINITY;
                lds_order();
                // the r-th smallest, ties counted
                for (;;) {
                    double mn = v;
#pragma unroll
                    for (int o = 32; o >= 1; o >>= 1) mn = fmin(mn, __shfl_xor(mn, o, 64));
                    const int c = __popcll(__ballot(v == mn));
                    if (r <= c) {
                        eps = mn;
                        break;
                    }
                    r -= c;
                    v = v == mn ? INFINITY : v;
                }
            }
        }
        if (lane == 0) {
            if (ok) {
                args.eps64[(size_t)p * nqb * 64 + q] = eps;
            } else {
                args.eps64[(size_t)p * nqb * 64 + q] = __builtin_nan("");
                const unsigned s = atomicAdd(&pa.qcount[p], 1u);
                pa.qlist[(size_t)p * Nv + s] = (uint16_t)q;
            }
        }
    }
}

// Counts of the prefix path from exact FP64 radii: G ordered pairs of one
// target atom, lane = query q of block qb; the strict counts
// #{j != q: max(dXf, dXp) < eps}, #{max(dYp, dXp) < eps} and #{dXp < eps}
// in FP64, the comparisons of lib/transfer_entropy.cpp:170-195 (the own
// frame, all distances 0, is removed after the loop).  No FP32 keys and no
// tie handling: eps is the exact radius, from the select kernel or, for the
// queries it listed, the heap kernel's list mode.  Wave sums in the order of
// te_count_group.
template <int G>
__device__ __forceinline__ void te_count64_group(const TeArgs& args, long long p0, int q,
                                                 bool active, int qb, int lane)
{
    const int N = args.N, lag = args.lag, k = args.k, Nv = N - lag;
    const int nqb = args.nqblocks;
    const double* X = args.traj + (size_t)args.a1[p0] * N * 3;
    const double* Y[G];
#pragma unroll
    for (int u = 0; u < G; ++u) Y[u] = args.traj + (size_t)args.a2[p0 + u] * N * 3;
    const int qs = active ? q : 0;
    double qx[6], qy[G][3], eps[G];
    int cx[G], cy[G], cp[G];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        qx[d] = X[(size_t)(qs + lag) * 3 + d];
        qx[3 + d] = X[(size_t)qs * 3 + d];
#pragma unroll
        for (int u = 0; u < G; ++u) qy[u][d] = Y[u][(size_t)qs * 3 + d];
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
        eps[u] = args.eps64[(size_t)(p0 + u) * nqb * 64 + qs];
        cx[u] = cy[u] = cp[u] = 0;
    }
    // The coordinates are read-only for the whole kernel: read through the
    // constant address space, the frame's loads (uniform over the wave) are
    // scalar loads
    typedef const __attribute__((address_space(4))) char* cchar;
    auto at = [](const double* base, uint32_t off) {
        return reinterpret_cast<const __attribute__((address_space(4))) double*>(
            (cchar)(reinterpret_cast<const char*>(base)) + off);
    };
    // frame j: a1 at j + lag and j, each source at j (32-bit byte offsets;
    // the next frame is loaded while this one is processed)
    auto load = [&](int j, double* f, double (*y)[3]) {
        const uint32_t of = (uint32_t)(j + lag) * 24u, op = (uint32_t)j * 24u;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            f[d] = at(X, of)[d];
            f[3 + d] = at(X, op)[d];
#pragma unroll
            for (int u = 0; u < G; ++u) y[u][d] = at(Y[u], op)[d];
        }
    };
    auto maxnorm3 = [](const double* a, const double* b) {
        return fmax(fmax(fabs(a[0] - b[0]), fabs(a[1] - b[1])), fabs(a[2] - b[2]));
    };
    double nf[6], ny[G][3];
    load(0, nf, ny);
    for (int j = 0; j < Nv; ++j) {
        double f[6], y[G][3];
#pragma unroll
        for (int d = 0; d < 6; ++d) f[d] = nf[d];
#pragma unroll
        for (int u = 0; u < G; ++u)
#pragma unroll
            for (int d = 0; d < 3; ++d) y[u][d] = ny[u][d];
        load(min(j + 1, Nv - 1), nf, ny);
        const double dxp = maxnorm3(qx + 3, f + 3);
        const double m1 = fmax(maxnorm3(qx, f), dxp);
#pragma unroll
        for (int u = 0; u < G; ++u) {
            const double m2 = fmax(maxnorm3(qy[u], y[u]), dxp);
            cx[u] += m1 < eps[u];
            cy[u] += m2 < eps[u];
            cp[u] += dxp < eps[u];
        }
    }
#pragma unroll
    for (int u = 0; u < G; ++u) {
        // the own frame counted in all three below a positive radius
        const int own = eps[u] > 0.0 ? 1 : 0;
        double te = 0.0;
        if (active) {
            const double* psi = args.psi;
            te = psi[k] + psi[min(cp[u] - own, Nv) + 1] - psi[min(cx[u] - own, Nv) + 1] -
                 psi[min(cy[u] - own, Nv) + 1];
        }
        for (int o = 32; o >= 1; o >>= 1) te += __shfl_xor(te, o, 64);
        if (lane == 0) args.partial[(size_t)(p0 + u) * nqb + qb] = te;
    }
}

// Wave w of block b runs item 4b + w as te_count_kernel does.
__global__ __launch_bounds__(256) void te_count64_kernel(TeArgs args)
{
    constexpr int G = kTeCountGroup;
    const int lane = threadIdx.x & 63;
    const long long item = (long long)blockIdx.x * 4 +
                           __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int nqb = args.nqblocks;
    const long long ngroups = (args.npairs + G - 1) / G;
    if (item >= ngroups * nqb) return;
    const long long g = item / nqb;
    const int qb = (int)(item - g * nqb);
    const long long p0 = g * G;
    bool shared = p0 + G <= args.npairs;
    if (shared) {
        const int t = args.a1[p0];
#pragma unroll
        for (int u = 1; u < G; ++u) shared = shared && args.a1[p0 + u] == t;
    }
    const int q = qb * 64 + lane;
    const bool active = q < args.N - args.lag;
    if (shared) {
        te_count64_group<G>(args, p0, q, active, qb, lane);
    } else {
        const long long pe = p0 + G < args.npairs ? p0 + G : args.npairs;
        for (long long p = p0; p < pe; ++p) te_count64_group<1>(args, p, q, active, qb, lane);
    }
}

__global__ void te_finalize_kernel(TeArgs args, double* mat, int natoms)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= args.npairs) return;
    double s = 0.0;
    for (int b = 0; b < args.nqblocks; ++b) s += args.partial[p * args.nqblocks + b];
    mat[(size_t)args.a1[p] * natoms + args.a2[p]] = s / (args.N - args.lag);
}

long long te_chunk_pairs(int nqblocks)
{
    const long long c = kTeChunkQueries / ((long long)nqblocks * 64);
    return c > 0 ? c : 1;
}

static size_t heap_lds_bytes(int k);

hipError_t launch_te(const TeArgs& a, hipStream_t s)
{
    if (a.k < 1 || !a.root) return hipErrorInvalidValue;
    const bool heap = a.path == 1 || (a.path == 0 && a.k <= kTeHeapMaxK);
    if (heap && a.k > kTeHeapMaxK) return hipErrorInvalidValue;
    const long long chunk = te_chunk_pairs(a.nqblocks);
    const size_t lds = heap_lds_bytes(a.k);
    for (long long p0 = 0; p0 < a.npairs; p0 += chunk) {
        TeArgs c = a;
        c.a1 = a.a1 + p0;
        c.a2 = a.a2 + p0;
        c.partial = a.partial + p0 * a.nqblocks;
        c.npairs = a.npairs - p0 < chunk ? a.npairs - p0 : chunk;
        const long long items = c.npairs * a.nqblocks;
        if (heap)
            hipLaunchKernelGGL(te_heap_kernel, dim3((unsigned)items), dim3(64), lds, s, c);
        else
            hipLaunchKernelGGL(te_select_kernel, dim3((unsigned)items), dim3(64), 0, s, c);
        const hipError_t e = launch_te_count(c, s);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launch_te_count64(const TeArgs& c, hipStream_t s)
{
    if (!c.eps64 || c.qblocks) return hipErrorInvalidValue;
    const long long citems = (c.npairs + kTeCountGroup - 1) / kTeCountGroup * c.nqblocks;
    hipLaunchKernelGGL(te_count64_kernel, dim3((unsigned)((citems + 3) / 4)), dim3(256), 0, s, c);
    return hipGetLastError();
}

hipError_t launch_te_count(const TeArgs& c, hipStream_t s)
{
    const long long citems = (c.npairs + kTeCountGroup - 1) / kTeCountGroup * c.nqblocks;
    hipLaunchKernelGGL(te_count_kernel, dim3((unsigned)((citems + 3) / 4)), dim3(256), 0, s, c);
    return hipGetLastError();
}

// heap positions kTeHeapRegs .. k in LDS (one slot when all fit in
// registers), then the frame stage
static size_t heap_lds_bytes(int k)
{
    const int slots = k + 1 > kTeHeapRegs ? k + 1 - kTeHeapRegs : 1;
    return (size_t)slots * 64 * sizeof(float) + 3 * 3 * kTeStage * sizeof(double);
}

hipError_t launch_te_heap_list(const TeArgs& a, long long nblocks, hipStream_t s)
{
    if (a.k < 1 || a.k > kTeHeapMaxK || !a.qblocks || !a.qcount || !a.qlist)
        return hipErrorInvalidValue;
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(te_heap_kernel, dim3((unsigned)nblocks), dim3(64), heap_lds_bytes(a.k), s, a);
    return hipGetLastError();
}

hipError_t launch_te_prefix_build(const TeArgs& a, const TePrefixArgs& pa, hipStream_t s)
{
    const int Nv = a.N - a.lag;
    if (Nv < 1 || Nv > kTePrefixMaxFrames || pa.ntargets < 1) return hipErrorInvalidValue;
    const long long rows = pa.ntargets * (long long)Nv;
    const unsigned grid = (unsigned)(rows < (1LL << 22) ? rows : (1LL << 22));
    if (Nv <= 512 * 4)
        hipLaunchKernelGGL(te_prefix_build_kernel<4>, dim3(grid), dim3(512), 0, s, a, pa);
    else if (Nv <= 512 * 16)
        hipLaunchKernelGGL(te_prefix_build_kernel<16>, dim3(grid), dim3(512), 0, s, a, pa);
    else
        hipLaunchKernelGGL(te_prefix_build_kernel<32>, dim3(grid), dim3(512), 0, s, a, pa);
    return hipGetLastError();
}

hipError_t launch_te_prefix_select(const TeArgs& a, const TePrefixArgs& pa, hipStream_t s)
{
    const int Nv = a.N - a.lag;
    if (a.k < 1 || a.k + 1 > kTePrefix || a.k > kTeHeapMaxK || Nv > kTePrefixMaxFrames)
        return hipErrorInvalidValue;
    const long long items = a.npairs * a.nqblocks;
    hipLaunchKernelGGL(te_prefix_select_kernel, dim3((unsigned)items), dim3(256), 0, s, a, pa);
    return hipGetLastError();
}

hipError_t launch_te_finalize(const TeArgs& a, double* mat, int natoms, hipStream_t s)
{
    hipLaunchKernelGGL(te_finalize_kernel, dim3((unsigned)((a.npairs + 255) / 256)), dim3(256), 0,
                       s, a, mat, natoms);
    return hipGetLastError();
}

}  // namespace gmxcorr
