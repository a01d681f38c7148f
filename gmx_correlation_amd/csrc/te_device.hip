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
    const long long p = blockIdx.x / nqb;
    const int qb = blockIdx.x % nqb;
    const int N = args.N, lag = args.lag, k = args.k, Nv = N - lag;
    const double* X = args.traj + (size_t)args.a1[p] * N * 3;
    const double* Y = args.traj + (size_t)args.a2[p] * N * 3;
    const int q = qb * 64 + lane;
    const bool active = q < Nv;
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
    args.root[(size_t)blockIdx.x * 64 + lane] = r[0];
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

hipError_t launch_te(const TeArgs& a, hipStream_t s)
{
    if (a.k < 1 || !a.root) return hipErrorInvalidValue;
    const bool heap = a.path == 1 || (a.path == 0 && a.k <= kTeHeapMaxK);
    if (heap && a.k > kTeHeapMaxK) return hipErrorInvalidValue;
    const long long chunk = te_chunk_pairs(a.nqblocks);
    // heap positions kTeHeapRegs .. k in LDS (one slot when all fit in registers)
    const int slots = a.k + 1 > kTeHeapRegs ? a.k + 1 - kTeHeapRegs : 1;
    const size_t lds = (size_t)slots * 64 * sizeof(float) + 3 * 3 * kTeStage * sizeof(double);
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
        const long long citems = (c.npairs + kTeCountGroup - 1) / kTeCountGroup * a.nqblocks;
        hipLaunchKernelGGL(te_count_kernel, dim3((unsigned)((citems + 3) / 4)), dim3(256), 0, s, c);
    }
    return hipGetLastError();
}

hipError_t launch_te_finalize(const TeArgs& a, double* mat, int natoms, hipStream_t s)
{
    hipLaunchKernelGGL(te_finalize_kernel, dim3((unsigned)((a.npairs + 255) / 256)), dim3(256), 0,
                       s, a, mat, natoms);
    return hipGetLastError();
}

}  // namespace gmxcorr
