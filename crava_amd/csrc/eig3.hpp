// Right eigenvectors of a dense real 3x3 matrix in LAPACK DGEEV's order and
// signs (see eig3.cpp).
#pragma once

namespace crava {

// m: row-major 3x3.  wr: eigenvalues (real parts), vr: eigenvectors as
// columns (vr[row][col]), unit 2-norm.  Returns false (outputs undefined)
// for matrices outside the handled case: a zero off-diagonal entry,
// unbalanced row and column norms, non-finite entries, a complex pair or no
// convergence.
bool dgeev3(const double m[3][3], double wr[3], double vr[3][3]);

}  // namespace crava
