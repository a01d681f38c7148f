#pragma once
/*
 * kraskov_gpu.h -- drop-in replacement for the reference's GPU boundary
 * (lib/kraskov_gpu.h:15-34).  libgmxcorr_hip.so exports exactly these two
 * C++ symbols, so the unchanged dispatcher kraskov_corrmatrix()
 * (lib/correlation_core.cpp:337-348), compiled with GMX_CORRELATION_USE_CUDA
 * defined, links against the MI355X library instead of the CUDA one:
 *
 *   _Z17kraskov_gpu_probev
 *   _Z22kraskov_corrmatrix_gpuPK9t_kraskoviPdi
 */
#include "gmxcorr_types.h"

/* True when device 0 is a gfx950 (MI355X) GPU that this library's code
 * objects run on.  Replaces lib/kraskov_gpu.h:20 (lib/kraskov_gpu.cu:210-219). */
bool kraskov_gpu_probe();

/* Full KSG-2 MI matrix on the GPU.  kr: workspace from kraskov_prepare()
 * (read-only; rows kr->x[0 .. 3*natoms-1] of kr->N doubles, kr->psi[0..N]).
 * mat: natoms*natoms doubles, caller-owned; on return mat[i*n+i] = 2000 and
 * mat[a1*n+a2] = mat[a2*n+a1] = MI in nats.  1 <= k <= 128 (the caller's
 * contract) and N > k.  On any runtime error it prints to stderr and exits,
 * as the reference does (lib/kraskov_gpu.cu:202-208): it never returns a
 * partial matrix.  Replaces lib/kraskov_gpu.h:32 (lib/kraskov_gpu.cu:225-308). */
void kraskov_corrmatrix_gpu(const t_kraskov* kr, int natoms, double* mat, int k);
