// qk_eigen.hpp -- small dense eigensolvers of the Krylov drivers (host).
//
//  * symtri_lowest: smallest eigenvalue of a symmetric tridiagonal matrix by
//    Sturm-sequence bisection (the Lanczos convergence test evaluates it
//    after every step, O(n) per bisection step).
//  * symtri_eig: all eigenvalues (ascending) and, on request, the first row
//    or all of the eigenvector matrix, by implicit QL iterations that chase
//    a bulge from the bottom of the active block to its top, with the shift
//    taken from the block's leading 2x2.
//  * herm_eig: complex Hermitian dense matrices (the Davidson projection)
//    by Householder reduction to a Hermitian tridiagonal, a diagonal phase
//    that makes it real, symtri_eig, and the back-transformation.
#pragma once
#include <complex>
#include <vector>

namespace qk {

// d[0..n), e[0..n-1): diagonal and off-diagonal (e[i] couples i and i+1)
double symtri_lowest(const double* d, const double* e, int n);

enum SymtriVectors { kNoVectors = 0, kFirstRow = 1, kAllVectors = 2 };
// w[0..n): eigenvalues in ascending order.
// want = kFirstRow: z[i] = first component of eigenvector i (z has n entries);
// want = kAllVectors: z[r * n + i] = component r of eigenvector i.
// Throws std::runtime_error if an eigenvalue does not converge.
void symtri_eig(const double* d, const double* e, int n, double* w, double* z, int want);

// A: n x n Hermitian, row-major. w: ascending eigenvalues; V[r * n + i] =
// component r of the unit eigenvector of w[i].
void herm_eig(int n, const std::vector<std::complex<double>>& A, std::vector<double>& w,
              std::vector<std::complex<double>>& V);

} // namespace qk
