// Non-local pseudopotential kernels (nonlocal.hip).
#pragma once

#include "common.h"

namespace jrx {

// Rows of g per partial-sum chunk of the projections.
int nl_rows(int ng, int64_t P);
// Bytes of the scratch launch_nonlocal needs (partials, F, per-band e).
size_t nl_scratch_bytes(int64_t P, int np_, int nb, int ng);
// F = proj Q per problem, e_out[p][b] = sum_j |F|^2 / vol and, when qbar is
// given, qbar += (2 f / vol) F conj(proj).  proj [nk][np][ng], Q and qbar
// [P][nb][ng] (problem p = spin * nk + k), occ [P][nb].
int launch_nonlocal(const double2* proj, const double2* q, const double* occ,
                    int64_t P, int nk, int np_, int ng, int nb, double vol,
                    double2* qbar, void* scratch, double* e_out,
                    hipStream_t s);
// energies[2] and energies[4] += sum f e.
int launch_nonlocal_finalize(const double* e, const double* occ, int64_t S,
                             double* energies, hipStream_t s);

}  // namespace jrx
