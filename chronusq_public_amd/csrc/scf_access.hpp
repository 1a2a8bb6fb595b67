// Internal view of a device-resident SCF for the serving loop, which builds
// the Fock matrices of several SCFs in one batched contraction.
#pragma once

#include "cqx.h"

namespace cqx {
// Device buffers of the SCF's current density D_S and of the J / K outputs
// its next iteration reads (NB x NB each, column-major).
int scf_fock_buffers(cqx_scf *s, const double **D, double **J, double **K, int *nb);
// The rest of one iteration once J and K hold J[D] and K[D]: G, F (with the
// non-finite check), energy, convergence test, CDIIS, eigensolve, next
// density.  ms_jk (the build's time) is charged to the SCF's Fock time.
int scf_after_jk(cqx_scf *s, double ms_jk, int *done);
// iterations done so far and the iteration limit
int scf_iterations(const cqx_scf *s, int *iter, int *max_iter);
}  // namespace cqx
