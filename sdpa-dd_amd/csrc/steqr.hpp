// Implicit QL/QR iteration on a symmetric tridiagonal in dd (the routine
// mplapack/Rsteqr.cpp:32-503 with compz = 'I'), used by the Lanczos step
// length on the (k+1)-dimensional Lanczos tridiagonal.
//
// The plane rotations act on each row of the eigenvector matrix
// independently, so callers that need only some rows (the Lanczos estimate
// reads one entry of the last row) pass that row range; the rows computed
// are bit-identical to those of the full computation.
#pragma once
#include <cstdint>

#include "ddr.hpp"

namespace sdp {

// d[n], e[n-1] overwritten; z is n x n column-major (ldz) but only rows
// [r0, r1) are formed.  work: 2*n dd.  Returns info (0 ok, >0 unconverged).
int64_t steqr_rows(int64_t n, ddr *d, ddr *e, ddr *z, int64_t ldz, ddr *work, int64_t r0, int64_t r1);

}  // namespace sdp
