// Canonical FP64 linear solves used by the coupling on the host.
//
// The reference calls LAPACK dgetrf/dgetrs (Utils.cpp:288-311) without
// pinning a LAPACK build.  This library fixes one algorithm instead:
// right-looking, unblocked Gaussian elimination with partial pivoting (first
// row of maximum |a_ik|), the right-hand side eliminated alongside, and back
// substitution accumulating in increasing column order.  Both solvers below
// perform exactly the non-trivial floating-point operations of that dense
// algorithm, in the same order; they only skip operations whose operand is
// an exact zero (a - 0*b and a - l*0 leave a unchanged), so their results are
// bit-identical to the dense elimination.
#pragma once

#include <cstddef>
#include <vector>

namespace life {

// Solve B x = b for the trailing block B = A[bc:, bc:] of the row-major
// n x n matrix A; x[0..bc) = 0.  Skips rows with a zero multiplier and
// columns that are zero in the pivot row.  Returns 0, or k+1 if column k
// had no non-zero pivot.
int lu_solve_sparse(int n, int bc, const double* A, const double* b, double* x, std::vector<double>& work,
                    std::vector<int>& iwork);

// Banded storage: row r holds columns [r-kl, r+kl+ku] at
// data[r*(2*kl+ku+1) + (c - r + kl)] (room for the fill-in of partial
// pivoting).  Solves the trailing (n-bc) block like lu_solve_sparse.
struct BandMatrix {
    int n = 0, kl = 0, ku = 0, w = 0;
    std::vector<double> data;
    void resize(int n_, int kl_, int ku_);
    void zero();
    double& at(int r, int c) { return data[(size_t)r * w + (c - r + kl)]; }
    double at(int r, int c) const { return data[(size_t)r * w + (c - r + kl)]; }
};

int lu_solve_band(const BandMatrix& A, int bc, const double* b, double* x, BandMatrix& work, std::vector<double>& y);

}  // namespace life
