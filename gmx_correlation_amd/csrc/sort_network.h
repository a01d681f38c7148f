// Register-resident bitonic sort of a block's keys (gfx950, 64-lane waves).
//
// A block of BT threads sorts M = BT * E keys held E per thread (thread t
// owns positions E*t .. E*t+E-1) into ascending order.  The network is the
// direction-free form of Batcher's bitonic sort: merge step K starts with the
// mirror pass (position x against x ^ (K - 1) inside its K-block), then the
// half-cleaners J = K/4 .. 1 (x against x ^ J); the lower position always
// keeps the smaller key, so every run is ascending and no pass carries a
// direction.  Each pass is resolved where its partner lives:
//   partner in the thread's registers   a min and a max into fixed registers;
//   partner thread in the same wave     a lane exchange (DPP for lane
//                                       distances 1 and 2 and the quad, half-row
//                                       and row mirrors, ds_bpermute otherwise);
//   partner thread in another wave      through LDS (lds: M keys), between
//                                       block barriers.
// The whole network is unrolled at compile time.  T must have a strict total
// order on the keys being sorted (equal keys must be equal values): the two
// threads of a cross-thread pair each keep one side of the comparison.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmxcorr {
namespace sortnet {

constexpr int kLanes = 64;

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_word(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, false);
}

// Value of lane lane ^ MT (MT < 64).
template <int MT>
__device__ __forceinline__ uint32_t xchg_word(uint32_t v, int lane)
{
    if constexpr (MT == 1) return dpp_word<0xB1>(v);         // quad_perm [1,0,3,2]
    else if constexpr (MT == 2) return dpp_word<0x4E>(v);    // quad_perm [2,3,0,1]
    else if constexpr (MT == 3) return dpp_word<0x1B>(v);    // quad_perm [3,2,1,0]
    else if constexpr (MT == 7) return dpp_word<0x141>(v);   // row_half_mirror
    else if constexpr (MT == 15) return dpp_word<0x140>(v);  // row_mirror
    else return (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ MT) << 2, (int)v);
}

template <int MT>
__device__ __forceinline__ uint32_t xchg(uint32_t v, int lane)
{
    return xchg_word<MT>(v, lane);
}

template <int MT>
__device__ __forceinline__ uint64_t xchg(uint64_t v, int lane)
{
    const uint32_t lo = xchg_word<MT>((uint32_t)v, lane);
    const uint32_t hi = xchg_word<MT>((uint32_t)(v >> 32), lane);
    return ((uint64_t)hi << 32) | lo;
}

template <typename T, int E, int BT, int K, int J, bool Flip>
__device__ __forceinline__ void pass(T (&r)[E], T* lds, int t)
{
    constexpr int H = Flip ? K / 2 : J;   // the position bit of the upper partner
    if constexpr (H < E) {
        constexpr int X = Flip ? K - 1 : J;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (e & H) continue;
            const T a = r[e], b = r[e ^ X];
            const bool lt = a < b;
            r[e] = lt ? a : b;
            r[e ^ X] = lt ? b : a;
        }
    } else {
        // the partner thread, and its element: e ^ (E - 1) for the mirror
        constexpr int MT = Flip ? K / E - 1 : J / E;
        constexpr int HB = H / E;
        const bool lower = (t & HB) == 0;
        T o[E];
        if constexpr (MT < kLanes) {
            const int lane = t & (kLanes - 1);
#pragma unroll
            for (int e = 0; e < E; ++e) o[e] = xchg<MT>(r[Flip ? E - 1 - e : e], lane);
        } else {
            // element e of thread t at lds[e * BT + t]: a wave's lanes touch
            // consecutive words (the partner t ^ MT keeps the lane bits), so
            // the exchange has no bank conflicts
            static_assert(MT < BT, "partner thread inside the block");
            __syncthreads();   // earlier readers of lds are done
#pragma unroll
            for (int e = 0; e < E; ++e) lds[e * BT + t] = r[e];
            __syncthreads();
            const int pt = t ^ MT;
#pragma unroll
            for (int e = 0; e < E; ++e) o[e] = lds[(Flip ? E - 1 - e : e) * BT + pt];
        }
#pragma unroll
        for (int e = 0; e < E; ++e) r[e] = lower == (o[e] < r[e]) ? o[e] : r[e];
    }
}

template <typename T, int E, int BT, int K, int J>
struct Halves {
    static __device__ __forceinline__ void run(T (&r)[E], T* lds, int t)
    {
        pass<T, E, BT, K, J, false>(r, lds, t);
        if constexpr (J > 1) Halves<T, E, BT, K, J / 2>::run(r, lds, t);
    }
};

template <typename T, int E, int BT, int K>
struct Merges {
    static __device__ __forceinline__ void run(T (&r)[E], T* lds, int t)
    {
        pass<T, E, BT, K, K / 2, true>(r, lds, t);
        if constexpr (K >= 4) Halves<T, E, BT, K, K / 4>::run(r, lds, t);
        if constexpr (K < BT * E) Merges<T, E, BT, 2 * K>::run(r, lds, t);
    }
};

// Ascending sort of the BT * E keys held E per thread by the BT threads of
// the block (E and BT powers of two, BT a multiple of 64).  lds: BT * E keys,
// used only for thread distances of 64 and more.  Ends without a barrier
// after its last LDS pass: a caller that reuses lds synchronises first.
// Callers that store the sorted keys to LDS use the same conflict-free
// layout: position p = E t + e at slot(p) = e * BT + t.
template <typename T, int E, int BT>
__device__ __forceinline__ void sort_blocked(T (&r)[E], T* lds, int t)
{
    static_assert((E & (E - 1)) == 0 && (BT & (BT - 1)) == 0 && BT >= kLanes, "powers of two");
    Merges<T, E, BT, 2>::run(r, lds, t);
}

// LDS slot of sorted position p in the layout above.
template <int E, int BT>
__device__ __forceinline__ int slot(int p)
{
    return (p & (E - 1)) * BT + p / E;
}

}  // namespace sortnet
}  // namespace gmxcorr
