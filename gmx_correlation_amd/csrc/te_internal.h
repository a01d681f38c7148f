// Internal interface of the transfer-entropy kernels (te_device.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace gmxcorr {

struct TeArgs {
    const double* traj;   // [n][N][3] mean-centred coordinates
    int N;
    int lag;
    int k;                // <= 128
    const int* a1;        // target atom of ordered pair p (TE(a2 -> a1))
    const int* a2;        // source atom
    long long npairs;
    int nqblocks;         // ceil((N - lag) / 64)
    const double* psi;    // [N - lag + 1]
    double* partial;      // [npairs][nqblocks]
};

hipError_t launch_te(const TeArgs& a, hipStream_t s);
hipError_t launch_te_finalize(const TeArgs& a, double* mat, int natoms, hipStream_t s);

}  // namespace gmxcorr
