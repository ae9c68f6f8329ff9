// BiCGSTAB whose device buffers all live in a fixed file-scope device array (noalloc.hip).
#pragma once
#include "krylov.hpp"

namespace cuibm {

// largest system the scratch array holds: 2^18 unknowns, 11 vectors of it plus the
// reduction partials and scalars (23.1 MB per GPU)
constexpr int NOALLOC_MAX_N = 1 << 18;

// cusp::krylov::bicgstab with a NONE (0) or DIAGONAL (1) preconditioner on host vectors
// x (warm start and result) and b; no device or pinned allocation
SolveStat bicgstab_noalloc(const DCsr<double> &A, double *x, const double *b, int precond, double rtol, int maxit,
                           hipStream_t s);

}  // namespace cuibm
