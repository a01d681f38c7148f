/*
 * gmxcorr_types.h -- the data-boundary struct shared with the reference.
 *
 * t_kraskov is the normalised workspace that kraskov_prepare() fills and that
 * the reference passes to its GPU backend (lib/correlation_core.h:45-52).
 * The layout and the name must stay identical: the exported C++ entry point
 * kraskov_corrmatrix_gpu(const t_kraskov*, ...) is mangled with this name.
 * When the reference's own correlation_core.h has already been included
 * (it defines GMX_CORRELATION_CORE_H) its definition is used instead.
 */
#ifndef GMXCORR_TYPES_H
#define GMXCORR_TYPES_H

#ifndef GMX_CORRELATION_CORE_H
/* x[c] (c = 3*atom + d) are 3*natoms separately allocated rows of N doubles,
 * normalised and shifted to be >= 0; psi[0..N] the digamma table; scal[c] a
 * per-component scale that the 3-D estimator does not use; dim = 3*natoms. */
struct t_kraskov
{
    double** x    = nullptr;
    double*  psi  = nullptr;
    double*  scal = nullptr;
    int      N    = 0;
    int      dim  = 0;
};
#endif

#endif
