// Exact signs of the small determinants the voxelizer's geometric tests
// reduce to, for double inputs.
//
// Method: Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
// Robust Geometric Predicates" (Discrete Comput. Geom. 18, 1997).  The
// determinant is first evaluated in floating point and its sign accepted
// when |det| exceeds the paper's static forward error bound (the filter);
// otherwise it is formed exactly as a nonoverlapping expansion -- Two-Sum,
// Two-Diff, Two-Product (with an FMA), Grow-, Expansion-Sum and
// Scale-Expansion -- whose sign is that of its largest nonzero component.
// Round-to-nearest-even without contraction is assumed (the library is
// built with -ffp-contract=off), and inputs far from underflow/overflow.
//
// Sign2 / Sign3 return -1, 0 or +1, or kUnsure when only the filter was
// asked for (kExact == false) and it could not decide.
#ifndef OCTLB_AMD_CSRC_EXACT_H_
#define OCTLB_AMD_CSRC_EXACT_H_

#include <hip/hip_runtime.h>

#include <cmath>

namespace octlb {
namespace exact {

constexpr int kUnsure = 2;
constexpr double kEps = 1.1102230246251565e-16;             // 2^-53
constexpr double kErr2 = (3.0 + 16.0 * kEps) * kEps;        // a*b - c*d of rounded differences
constexpr double kErr3 = (7.0 + 56.0 * kEps) * kEps;        // 3x3 determinant of rounded differences

__host__ __device__ inline void TwoSum(double a, double b, double& x, double& y) {
  x = a + b;
  const double bv = x - a;
  const double av = x - bv;
  y = (a - av) + (b - bv);
}

__host__ __device__ inline void FastTwoSum(double a, double b, double& x, double& y) {  // |a| >= |b|
  x = a + b;
  y = b - (x - a);
}

__host__ __device__ inline void TwoDiff(double a, double b, double& x, double& y) {
  x = a - b;
  const double bv = a - x;
  const double av = x + bv;
  y = (a - av) + (bv - b);
}

__host__ __device__ inline void TwoProd(double a, double b, double& x, double& y) {
  x = a * b;
  y = fma(a, b, -x);
}

// h = e + f (Expansion-Sum): h may alias e; returns n + m components.
__host__ __device__ inline int ExpSum(const double* e, int n, const double* f, int m, double* h) {
  if (h != e)
    for (int i = 0; i < n; ++i) h[i] = e[i];
  int len = n;
  for (int j = 0; j < m; ++j) {
    // grow h[j .. len) by f[j]; h[0 .. j) is final
    double q = f[j];
    for (int i = j; i < len; ++i) {
      double s, t;
      TwoSum(q, h[i], s, t);
      h[i] = t;
      q = s;
    }
    h[len++] = q;
  }
  return len;
}

// h = e * b (Scale-Expansion): returns 2 n components.
__host__ __device__ inline int ExpScale(const double* e, int n, double b, double* h) {
  double q;
  TwoProd(e[0], b, q, h[0]);
  for (int i = 1; i < n; ++i) {
    double p1, p0, s;
    TwoProd(e[i], b, p1, p0);
    TwoSum(q, p0, s, h[2 * i - 1]);
    FastTwoSum(p1, s, q, h[2 * i]);
  }
  h[2 * n - 1] = q;
  return 2 * n;
}

__host__ __device__ inline int ExpSign(const double* e, int n) {
  for (int i = n - 1; i >= 0; --i)
    if (e[i] != 0.0) return e[i] > 0.0 ? 1 : -1;
  return 0;
}

// h = x * y for two-component expansions: 8 components.
__host__ __device__ inline int ExpMul22(const double* x, const double* y, double* h) {
  double a[4], b[4];
  ExpScale(x, 2, y[0], a);
  ExpScale(x, 2, y[1], b);
  return ExpSum(a, 4, b, 4, h);
}

__host__ __device__ inline void Diff2(double a, double b, double* e) {  // e = a - b exactly (low, high)
  TwoDiff(a, b, e[1], e[0]);
}

__host__ __device__ inline int Sgn(double v) { return v > 0.0 ? 1 : (v < 0.0 ? -1 : 0); }

// sign of (p1u - q1u)(p2v - q2v) - (p1v - q1v)(p2u - q2u)
template <bool kExact>
__host__ __device__ inline int Sign2(double p1u, double p1v, double q1u, double q1v, double p2u, double p2v,
                                     double q2u, double q2v) {
  const double l = (p1u - q1u) * (p2v - q2v);
  const double r = (p1v - q1v) * (p2u - q2u);
  const double det = l - r;
  const double bound = kErr2 * (fabs(l) + fabs(r));
  if (det > bound || -det > bound) return Sgn(det);
  if (!kExact) return kUnsure;
  double a[2], b[2], c[2], d[2], ab[8], cd[8], s[16];
  Diff2(p1u, q1u, a);
  Diff2(p2v, q2v, b);
  Diff2(p1v, q1v, c);
  Diff2(p2u, q2u, d);
  ExpMul22(a, b, ab);
  ExpMul22(c, d, cd);
  for (int i = 0; i < 8; ++i) cd[i] = -cd[i];
  return ExpSign(s, ExpSum(ab, 8, cd, 8, s));
}

// orient3d: sign of det[A - D; B - D; C - D]
template <bool kExact>
__host__ __device__ inline int Sign3(const double* A, const double* B, const double* C, const double* D) {
  const double adx = A[0] - D[0], ady = A[1] - D[1], adz = A[2] - D[2];
  const double bdx = B[0] - D[0], bdy = B[1] - D[1], bdz = B[2] - D[2];
  const double cdx = C[0] - D[0], cdy = C[1] - D[1], cdz = C[2] - D[2];
  const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  const double cdxady = cdx * ady, adxcdy = adx * cdy;
  const double adxbdy = adx * bdy, bdxady = bdx * ady;
  const double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
  const double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
                      (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
  const double bound = kErr3 * perm;
  if (det > bound || -det > bound) return Sgn(det);
  if (!kExact) return kUnsure;
  // exact: the six triple products of the expanded differences, summed in place
  double e[3][3][2];
  for (int k = 0; k < 3; ++k) {
    Diff2(A[k], D[k], e[0][k]);
    Diff2(B[k], D[k], e[1][k]);
    Diff2(C[k], D[k], e[2][k]);
  }
  // det = sum over permutations (i, j, k) of (0, 1, 2) of sign * a_i b_j c_k
  constexpr int kPerm[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {1, 0, 2}, {2, 1, 0}};
  double acc[192];
  int len = 0;
  for (int p = 0; p < 6; ++p) {
    double ab[8], t0[16], t1[16], t[32];
    ExpMul22(e[0][kPerm[p][0]], e[1][kPerm[p][1]], ab);
    const double* c = e[2][kPerm[p][2]];
    ExpScale(ab, 8, c[0], t0);
    ExpScale(ab, 8, c[1], t1);
    ExpSum(t0, 16, t1, 16, t);
    if (p >= 3)
      for (int i = 0; i < 32; ++i) t[i] = -t[i];
    len = ExpSum(acc, len, t, 32, acc);
  }
  return ExpSign(acc, len);
}

}  // namespace exact
}  // namespace octlb

#endif  // OCTLB_AMD_CSRC_EXACT_H_
