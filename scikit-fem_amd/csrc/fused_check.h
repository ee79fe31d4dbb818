// Recheck of the fused path's suspects (entries whose fused value lies
// within the rounding bound of zero) in the reference's arithmetic; the
// kernels live in exact_local.hip (compiled with -ffp-contract=off).
#pragma once

#include "common.h"

namespace femhip {

// Recheck of the suspect queue of the last first pass: every queued
// (element, row i, column j) entry is evaluated in the reference's
// operation order; the verdict compares the exactly-zero ones with the
// plan's zero mask and count, writes check->verdict and clears the queue
// counters and flags of `state`.
int fused_check_run(const femhip_fused_check_t* check, femhip_fused_state_t* state, int form,
                    int dim, const double* p, int64_t nnodes, const double* kappa,
                    hipStream_t s);
// the same for the suspects of a P2 rows pass (P2 triangles/tetrahedra,
// laplace, mass or kappa_laplace with kappa per element; tables and
// coordinates from g and sp)
int rows_check_run(const femhip_fused_check_t* check, femhip_fused_state_t* state, int form,
                   const femhip_geom_t* g, const femhip_space_t* sp, const double* kappa,
                   hipStream_t s);
// verdict of an empty plan (no partitions): valid iff the plan has no zeros
int fused_check_empty(const femhip_fused_check_t* check, femhip_fused_state_t* state,
                      hipStream_t s);

}  // namespace femhip
