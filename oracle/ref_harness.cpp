/*
 * ref_harness.cpp -- a thin extern "C" shim over the reference's own CPU
 * functions, compiled together with the reference sources in place
 * (lib/kraskov.cpp, lib/correlation_core.cpp, lib/gcmi.cpp,
 * lib/transfer_entropy.cpp) into oracle/_ref/libref.so by oracle/Makefile.
 *
 * TEST INFRASTRUCTURE ONLY: it validates the oracle restatement, writes the
 * golden fixtures and is the timed CPU baseline.  The only code here that is
 * not a call into the reference is the mean-centring that the GROMACS front
 * end performs (src/gmx_correlation.cpp:184-223), which needs GROMACS and
 * cannot be compiled here.
 */
#include <cstring>
#include <stdexcept>
#include <vector>

#include "correlation_core.h"
#include "gcmi.h"
#include "transfer_entropy.h"

namespace {
/* coords: float32 [nframes][natoms][3] -> t_traj with double fluctuations */
void make_traj(const float* coords, int nframes, int natoms, t_traj* t)
{
    t->natoms  = natoms;
    t->nframes = nframes;
    snew(t->x, natoms);
    snew(t->xav, natoms);
    for (int a = 0; a < natoms; ++a) {
        snew(t->x[a], nframes);
        for (int f = 0; f < nframes; ++f)
            for (int d = 0; d < DIM; ++d)
                t->xav[a][d] += static_cast<double>(coords[(static_cast<size_t>(f) * natoms + a) * 3 + d]);
        for (int d = 0; d < DIM; ++d) t->xav[a][d] /= nframes;
        for (int f = 0; f < nframes; ++f)
            for (int d = 0; d < DIM; ++d)
                t->x[a][f][d] = static_cast<double>(coords[(static_cast<size_t>(f) * natoms + a) * 3 + d]) - t->xav[a][d];
    }
}
} // namespace

extern "C" {

/* Mean-centred trajectory as [natoms][nframes][3] doubles. */
void ref_center(const float* coords, int nframes, int natoms, double* out)
{
    t_traj t;
    make_traj(coords, nframes, natoms, &t);
    for (int a = 0; a < natoms; ++a)
        std::memcpy(out + static_cast<size_t>(a) * nframes * 3, t.x[a], sizeof(double) * 3 * nframes);
    done_traj(&t);
}

/* kraskov_prepare output: x rows [3*natoms][nframes] and psi[nframes+1]. */
void ref_prepare(const float* coords, int nframes, int natoms, double* xrows, double* psi)
{
    t_traj t;
    make_traj(coords, nframes, natoms, &t);
    t_kraskov kr;
    kraskov_prepare(&t, &kr);
    for (int c = 0; c < kr.dim; ++c)
        std::memcpy(xrows + static_cast<size_t>(c) * nframes, kr.x[c], sizeof(double) * nframes);
    std::memcpy(psi, kr.psi, sizeof(double) * (nframes + 1));
    kraskov_done(&kr);
    done_traj(&t);
}

/* One KSG pair through kraskov_wrap on a prepared workspace. */
double ref_ksg_pair(const float* coords, int nframes, int natoms, int a1, int a2, int k)
{
    t_traj t;
    make_traj(coords, nframes, natoms, &t);
    t_kraskov kr;
    kraskov_prepare(&t, &kr);
    const double v = kraskov_wrap(&kr, a1, a2, k);
    kraskov_done(&kr);
    done_traj(&t);
    return v;
}

/* Full KSG matrix through kraskov_corrmatrix (CPU path). Returns 0, or -1
 * when the reference throws (natoms < 2 or nframes <= k). */
int ref_ksg_matrix(const float* coords, int nframes, int natoms, int k, int nthreads, double* mat)
{
    t_traj t;
    make_traj(coords, nframes, natoms, &t);
    int rc = 0;
    try {
        kraskov_corrmatrix(&t, mat, k, false, nthreads);
    } catch (const std::exception&) {
        rc = -1;
    }
    done_traj(&t);
    return rc;
}

/* KSG on a list of pairs with an OpenMP loop over the list: the bounded
 * sample that the benchmark times (pairs use kraskov_wrap as the matrix
 * loop does). */
void ref_ksg_pairs(const float* coords, int nframes, int natoms, int k, int npairs,
                   const int* a1, const int* a2, double* out, int nthreads);

void ref_gauss_matrix(const float* coords, int nframes, int natoms, double* mat)
{
    t_traj t;
    make_traj(coords, nframes, natoms, &t);
    gauss_corrmatrix(&t, mat);
    done_traj(&t);
}

int ref_gcmi_matrix(const float* coords, int nframes, int natoms, int nthreads, double* mat)
{
    t_traj t;
    make_traj(coords, nframes, natoms, &t);
    int rc = 0;
    try {
        gcmi_corrmatrix(&t, mat, false, nthreads);
    } catch (const std::exception&) {
        rc = -1;
    }
    done_traj(&t);
    return rc;
}

void ref_te_matrix(const float* coords, int nframes, int natoms, int k, int lag, int nthreads, double* mat)
{
    t_traj t;
    make_traj(coords, nframes, natoms, &t);
    te_matrix(&t, mat, k, lag, false, nthreads);
    done_traj(&t);
}

void ref_pearsify(double* mat, int n) { pearsify(mat, n, DIM); }

void ref_write_matrix(const double* mat, int n, const char* path) { write_matrix(mat, n, path); }

void ref_write_xpm(const double* mat, int n, const char* path, const char* title,
                   const char* legend)
{
    write_xpm_matrix(mat, n, path, title, legend);
}

} // extern "C"

#include <omp.h>
extern "C" void ref_ksg_pairs(const float* coords, int nframes, int natoms, int k, int npairs,
                              const int* a1, const int* a2, double* out, int nthreads)
{
    t_traj t;
    make_traj(coords, nframes, natoms, &t);
    t_kraskov kr;
    kraskov_prepare(&t, &kr);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < npairs; ++p)
        out[p] = kraskov_wrap(&kr, a1[p], a2[p], k);
    kraskov_done(&kr);
    done_traj(&t);
}
