// Internal interface between the KSG kernels (ksg_device.hip) and the host
// plan code (ksg_host.hip).  Not installed; the public ABI is gmxcorr.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmxcorr {

// A frame row whose FP32 boundary is ambiguous, queued for the FP64 recount.
struct FixEntry {
    int      pair;     // index into the pair list of the run
    int      row;      // query frame i
    uint32_t eps_bits; // FP32 bits of the (k+1)-th smallest joint distance (self included)
    int      pad;
};

// Device buffers and shape of one KSG run.
struct KsgArgs {
    const double* X;      // [3n][N] normalised coordinates (FP64)
    const float*  D;      // [n][N][P] per-atom max-norm distance rows (FP32)
    const double* psi;    // [N+1]
    const int*    a1;     // [npairs]
    const int*    a2;     // [npairs]
    double*       partial;   // [npairs] sum of psi[nx]+psi[ny] over unflagged rows
    double*       fixsum;    // [npairs] same over FP64-recounted rows
    double*       mi;        // [npairs]
    FixEntry*     fixq;      // [fixcap]
    unsigned*     counters;  // [0] queued rows, [1] exact ties
    unsigned      fixcap;
    int*          rc_nx;     // optional [npairs][N]
    int*          rc_ny;     // optional [npairs][N]
    unsigned char* rc_flag;  // optional [npairs][N]
    long long     npairs;
    long long     P;         // row pitch in floats (multiple of 64)
    int           N;
    int           K;
    double        psiN_phiK; // psi[N] + (psi[K] - 1/K)
};

// Largest N handled by the pair kernels (cooperative kernel, 4 x 2048 frames).
constexpr int kMaxFrames = 8192;

// Read-only view of a loaded plan's device data (for the surrogate features).
struct PlanView {
    const double* X;   // [3n][N] normalised rows
    const float* D;    // [n][N][P] distance rows
    int natoms, N, K;
    long long P;
    hipStream_t stream;
};

hipError_t launch_build_rows(const double* X, int natoms, int N, long long P, float* D,
                             hipStream_t s);
hipError_t launch_ksg_clear(const KsgArgs& a, hipStream_t s);
hipError_t launch_ksg_pairs(const KsgArgs& a, hipStream_t s);
hipError_t launch_ksg_fixup(const KsgArgs& a, unsigned nqueued, hipStream_t s);
hipError_t launch_ksg_finalize(const KsgArgs& a, hipStream_t s);
hipError_t launch_selftest(int* out, hipStream_t s);

}  // namespace gmxcorr
