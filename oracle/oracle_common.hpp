// TEST INFRASTRUCTURE ONLY -- CPU oracle of the SLOAM geometric hot path.
//
// This directory is the correctness oracle: a single-threaded (optionally
// OpenMP over independent clusters / problems) C++17 restatement of the
// reference algorithms, used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg only.  Nothing in sloam_amd/ links, loads or calls it.
//
// Compiled with -ffp-contract=off (see oracle/Makefile) so that FP32 and FP64
// expressions are evaluated operation by operation, in source order.
#pragma once

#include <cmath>
#include <cstdint>

namespace oracle {

// ---- Philox4x32-10 (Salmon et al., SC'11), published constants ------------
struct U4 {
  uint32_t v[4];
};

inline U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; ++round) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * ctr.v[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr.v[2];
    U4 n;
    n.v[0] = (uint32_t)(p1 >> 32) ^ ctr.v[1] ^ k0;
    n.v[1] = (uint32_t)p1;
    n.v[2] = (uint32_t)(p0 >> 32) ^ ctr.v[3] ^ k1;
    n.v[3] = (uint32_t)p0;
    ctr = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return ctr;
}

// floor(x * m / 2^32)
inline uint32_t bounded(uint32_t x, uint32_t m) { return (uint32_t)(((uint64_t)x * m) >> 32); }

// Three distinct indices of [0, n): first uniform over n, second over the n-1
// others, third over the n-2 others (skip-over construction).
inline void draw3(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t n, int32_t out[3]) {
  const uint32_t first = bounded(x0, n);
  uint32_t second = bounded(x1, n - 1);
  if (second >= first) ++second;
  const uint32_t small = first < second ? first : second;
  const uint32_t large = first < second ? second : first;
  uint32_t third = bounded(x2, n - 2);
  if (third >= small) ++third;
  if (third >= large) ++third;
  out[0] = (int32_t)first;
  out[1] = (int32_t)second;
  out[2] = (int32_t)third;
}

// ---- symmetric 3x3 eigen-decomposition (cyclic Jacobi, FP64) --------------
// Rotation formulas of the classical Jacobi method (Golub & Van Loan,
// "Matrix Computations", sec. 8.5; tau form as in Rutishauser's ALGOL
// procedure).  Sweep order (0,1), (0,2), (1,2); stop when the off-diagonal
// mass is <= 1e-17 of the initial matrix mass.
void eig3(const double a[6], double w[3], double V[3][3]);

// Column k of V with its largest-magnitude component made positive.
void eigvec(const double V[3][3], int k, double out[3]);
int argmax3(const double w[3]);
int argmin3(const double w[3]);
// Orthonormal (e1, e2) spanning the plane orthogonal to unit a.
void basis(const double a[3], double e1[3], double e2[3]);
// 3x3 linear solve, Gaussian elimination with partial pivoting.
bool solve3(double M[3][3], double b[3], double x[3]);
// n x n linear solve (row-major M, destroyed), partial pivoting.
bool solve_n(double* M, double* b, double* x, int n);
// Cholesky solve of a symmetric positive definite 5x5 (or 6x6) system;
// false if a pivot is not positive.
bool cholesky_solve5(const double A[5][5], const double b[5], double x[5]);
bool cholesky_solve6(const double A[6][6], const double b[6], double x[6]);

}  // namespace oracle
