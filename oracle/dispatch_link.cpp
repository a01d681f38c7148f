/*
 * dispatch_link.cpp -- links the reference's unchanged KSG dispatcher
 * kraskov_corrmatrix() (lib/correlation_core.cpp:324-459, compiled where it
 * lies with GMX_CORRELATION_USE_CUDA defined) against libgmxcorr_hip.so, the
 * way the reference's CMake would link its CUDA backend
 * (CMakeLists.txt:127-135).  oracle/Makefile target `dispatch` writes the
 * executable to oracle/_ref/dispatch_link.
 *
 * TEST INFRASTRUCTURE ONLY, and never shipped: it shows that the dispatcher's
 * two GPU calls (lib/kraskov_gpu.h:20,32) resolve to this library's exports,
 * and, where no gfx950 device is visible, that the dispatcher's own CPU
 * fallback (lib/correlation_core.cpp:337-348) runs.
 *
 * Usage: dispatch_link <coords.f32> <nframes> <natoms> <k> <use_gpu> <nthreads> <mat.f64>
 *   coords.f32: float32 [nframes][natoms][3] raw; mat.f64: double [natoms][natoms] raw.
 * Prints "probe=<0|1>" on stdout before the matrix call.
 */
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "correlation_core.h"
// This library's replacement header: it must declare the same two functions
// with the same types as the reference's lib/kraskov_gpu.h, which the
// dispatcher includes; a mismatch is a compile error here.
#include "../include/kraskov_gpu.h"

int main(int argc, char** argv)
{
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s coords.f32 nframes natoms k use_gpu nthreads mat.f64\n",
                     argv[0]);
        return 2;
    }
    const int T = std::atoi(argv[2]), n = std::atoi(argv[3]), k = std::atoi(argv[4]);
    const bool use_gpu = std::atoi(argv[5]) != 0;
    const int nthreads = std::atoi(argv[6]);
    std::vector<float> c((size_t)T * n * 3);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(c.data(), sizeof(float), c.size(), f) != c.size()) {
        std::fprintf(stderr, "cannot read %s\n", argv[1]);
        return 2;
    }
    std::fclose(f);
    // t_traj as the GROMACS front end fills it: float coordinates promoted to
    // double, per-atom mean from a sequential sum over frames, subtracted
    // (src/gmx_correlation.cpp:177-181, 200-223).
    t_traj t;
    t.nframes = T;
    t.natoms = n;
    snew(t.x, n);
    snew(t.xav, n);
    for (int a = 0; a < n; ++a) {
        snew(t.x[a], T);
        for (int fr = 0; fr < T; ++fr)
            for (int d = 0; d < DIM; ++d) t.xav[a][d] += (double)c[((size_t)fr * n + a) * 3 + d];
        for (int d = 0; d < DIM; ++d) t.xav[a][d] /= T;
        for (int fr = 0; fr < T; ++fr)
            for (int d = 0; d < DIM; ++d)
                t.x[a][fr][d] = (double)c[((size_t)fr * n + a) * 3 + d] - t.xav[a][d];
    }
    std::printf("probe=%d\n", kraskov_gpu_probe() ? 1 : 0);
    std::fflush(stdout);
    std::vector<double> mat((size_t)n * n);
    kraskov_corrmatrix(&t, mat.data(), k, use_gpu, nthreads);
    done_traj(&t);
    FILE* o = std::fopen(argv[7], "wb");
    if (!o || std::fwrite(mat.data(), sizeof(double), mat.size(), o) != mat.size()) {
        std::fprintf(stderr, "cannot write %s\n", argv[7]);
        return 2;
    }
    std::fclose(o);
    return 0;
}
