// D3Q19 arithmetic shared by every kernel, written once for host and device.
//
// Each function keeps the operation order of the reference source it cites so
// that an FP64 build without FMA contraction (-ffp-contract=off) reproduces
// the reference's rounding; with T = float the same code is the FP32 path.
#ifndef OCTLB_AMD_CSRC_D3Q19_H_
#define OCTLB_AMD_CSRC_D3Q19_H_

#include <hip/hip_runtime.h>

namespace octlb {

constexpr int kQ = 19;

// Velocity set, opposite table and weight denominators in OpenLB order
// (descriptor/definition/common.h:166-198).
__host__ __device__ constexpr int Cx(int i) {
  constexpr int v[kQ] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
  return v[i];
}
__host__ __device__ constexpr int Cy(int i) {
  constexpr int v[kQ] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
  return v[i];
}
__host__ __device__ constexpr int Cz(int i) {
  constexpr int v[kQ] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
  return v[i];
}
__host__ __device__ constexpr int Ci(int i, int d) { return d == 0 ? Cx(i) : (d == 1 ? Cy(i) : Cz(i)); }
__host__ __device__ constexpr int Opp(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }
__host__ __device__ constexpr int WDen(int i) { return i == 0 ? 3 : ((i <= 3 || (i >= 10 && i <= 12)) ? 18 : 36); }

// Fraction::as<T>() = T(1)/T(den) (utilities/fraction.h:72-75).
template <typename T>
__host__ __device__ inline T W(int i) {
  return T(1) / T(WDen(i));
}

// BcKind (bc_kind.h:14-23).
enum : int {
  kBulk = 0,
  kSolid = 1,
  kBounceBack = 2,
  kMovingBounceBack = 3,
  kBouzidi = 4,
  kVelocity = 5,
  kPressure = 6,
  kOutflow = 7
};

__host__ __device__ inline bool ReflectsOnPull(int k) { return k == kVelocity || k == kPressure || k == kBouzidi; }
__host__ __device__ inline bool IsFd(int k) { return k == kVelocity || k == kPressure; }

// FP32 arithmetic on shifted populations.  Storing f - t pays only if the
// arithmetic stays shifted too: a density carried as rho = 1 + sum(f) in FP32
// resolves 2^-24 in absolute terms, and rho t (1 + X) - t then cancels the
// weight t and loses the low bits of every population.  The FP32 path
// therefore carries the density deviation d = rho - 1 (summed from the
// shifted populations, never adding the 1) wherever the FP64 path carries
// rho, and forms feq - t = t d + (1 + d) t X, which has no cancellation.
// FP64 keeps the reference's expressions bit for bit.  In the functions
// below a parameter `r` is rho for FP64 and rho - 1 for FP32.
template <typename T>
constexpr bool kShiftedArith = sizeof(T) == 4;

// rho from r.
template <typename T>
__host__ __device__ inline T RhoOf(T r) {
  if constexpr (kShiftedArith<T>) {
    return T(1) + r;
  } else {
    return r;
  }
}
// r of the rest density (rho = 1).
template <typename T>
__host__ __device__ inline T RestR() {
  return kShiftedArith<T> ? T(0) : T(1);
}
// r of a density given in FP64.
template <typename T>
__host__ __device__ inline T ROf(double rho) {
  if constexpr (kShiftedArith<T>) {
    return static_cast<T>(rho - 1.0);
  } else {
    return static_cast<T>(rho);
  }
}
// rho in FP64 from r.
template <typename T>
__host__ __device__ inline double RhoDouble(T r) {
  if constexpr (kShiftedArith<T>) {
    return 1.0 + static_cast<double>(r);
  } else {
    return static_cast<double>(r);
  }
}

// Terms c_i,d * x with c_i,d = 0 are skipped: for finite x they add a
// signed zero to an accumulator that starts at +0 and never becomes -0, so
// every result is bit-identical to the full sum, and the unrolled loops lose
// a third of their multiplies and adds.  (i is a compile-time constant in all
// callers after unrolling, so the tests fold away.)

// equilibrium::secondOrder on shifted populations (dynamics/lbm.h:51-62);
// r as above.
template <typename T>
__host__ __device__ inline T Feq(int i, T r, const T* u, T usqr) {
  T cu = T(0);
  if (Cx(i) != 0) cu += T(Cx(i)) * u[0];
  if (Cy(i) != 0) cu += T(Cy(i)) * u[1];
  if (Cz(i) != 0) cu += T(Cz(i)) * u[2];
  const T t = W<T>(i);
  if constexpr (kShiftedArith<T>) {
    return t * r + (T(1) + r) * t * (T(3) * cu + T(3) * T(3) * T(0.5) * cu * cu - T(3) * T(0.5) * usqr);
  } else {
    const T rho = r;
    return rho * t * (T(1) + T(3) * cu + T(3) * T(3) * T(0.5) * cu * cu - T(3) * T(0.5) * usqr) - t;
  }
}

template <typename T>
__host__ __device__ inline T NormSqr(const T* u) {
  T s = T(0);
  s += u[0] * u[0];
  s += u[1] * u[1];
  s += u[2] * u[2];
  return s;
}

// CellProxy::computeRhoU: rho starts at 1 (block_lattice.h:113-125).
// Returns r (rho for FP64, rho - 1 for FP32) in `rho`.
template <typename T>
__host__ __device__ inline void ProxyRhoU(const T* f, T& rho, T* u) {
  if constexpr (kShiftedArith<T>) {
    T d = T(0);
    u[0] = u[1] = u[2] = T(0);
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      d += f[i];
      if (Cx(i) != 0) u[0] += f[i] * T(Cx(i));
      if (Cy(i) != 0) u[1] += f[i] * T(Cy(i));
      if (Cz(i) != 0) u[2] += f[i] * T(Cz(i));
    }
    const T full = T(1) + d;
    u[0] /= full;
    u[1] /= full;
    u[2] /= full;
    rho = d;
    return;
  }
  T r = T(1);
  u[0] = u[1] = u[2] = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    r += f[i];
    if (Cx(i) != 0) u[0] += f[i] * T(Cx(i));
    if (Cy(i) != 0) u[1] += f[i] * T(Cy(i));
    if (Cz(i) != 0) u[2] += f[i] * T(Cz(i));
  }
  u[0] /= r;
  u[1] /= r;
  u[2] /= r;
  rho = r;
}

// olb::lbm::computeRhoU: rho = sum(f) + 1 (dynamics/lbm.h:290-331).
// Returns r (rho for FP64, rho - 1 for FP32) in `rho`.
template <typename T>
__host__ __device__ inline void OlbRhoU(const T* f, T& rho, T* u) {
  T r = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) r += f[i];
  if constexpr (kShiftedArith<T>) {
    u[0] = u[1] = u[2] = T(0);
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      if (Cx(i) != 0) u[0] += f[i] * T(Cx(i));
      if (Cy(i) != 0) u[1] += f[i] * T(Cy(i));
      if (Cz(i) != 0) u[2] += f[i] * T(Cz(i));
    }
    const T full = T(1) + r;
    u[0] /= full;
    u[1] /= full;
    u[2] /= full;
    rho = r;
    return;
  }
  r += T(1);
  u[0] = u[1] = u[2] = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    if (Cx(i) != 0) u[0] += f[i] * T(Cx(i));
    if (Cy(i) != 0) u[1] += f[i] * T(Cy(i));
    if (Cz(i) != 0) u[2] += f[i] * T(Cz(i));
  }
  u[0] /= r;
  u[1] /= r;
  u[2] /= r;
  rho = r;
}

// olb::lbm::bgkCollision (dynamics/lbm.h:395-403).
template <typename T>
__host__ __device__ inline void OlbBgk(T* f, T rho, const T* u, T omega) {
  const T usqr = NormSqr(u);
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    f[i] *= T(1) - omega;
    f[i] += omega * Feq(i, rho, u, usqr);
  }
}

// CollideBgkAt (block_lattice.cpp:328-341).
template <typename T>
__host__ __device__ inline void CollideBgkAt(T* f, T omega) {
  T rho, u[3];
  ProxyRhoU(f, rho, u);
  const T usqr = NormSqr(u);
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const T fe = Feq(i, rho, u, usqr);
    f[i] += omega * (fe - f[i]);
  }
}

// CollideConstRhoBgkWithMacroscopic (block_lattice.h:133-151).  avg is the
// average density for FP64 and its deviation avg - 1 for FP32 (AvgArg).
// FP32: ratio (feq + t) - t = (rho + 1 - avg) t (1 + X) - t = t e + (1 + e) t X
// with e = (rho - 1) - (avg - 1).
template <typename T>
__host__ __device__ inline void ConstRhoBgk(T* f, T omega, T avg, T rho, const T* u) {
  const T usqr = NormSqr(u);
  if constexpr (kShiftedArith<T>) {
    const T e = (T(1) + rho) > T(1e-12) ? rho - avg : rho;
#pragma unroll
    for (int i = 0; i < kQ; ++i) {
      const T fe = Feq(i, rho, u, usqr);
      const T scaled = Feq(i, e, u, usqr);
      f[i] = scaled + (T(1) - omega) * (f[i] - fe);
    }
    return;
  }
  const T delta = T(1) - avg;
  const T ratio = rho > T(1e-12) ? T(1) + delta / rho : T(1);
#pragma unroll
  for (int i = 0; i < kQ; ++i) {
    const T fe = Feq(i, rho, u, usqr);
    const T t = W<T>(i);
    f[i] = ratio * (fe + t) - t + (T(1) - omega) * (f[i] - fe);
  }
}

// equilibrium::fromPiToFneq (dynamics/lbm.h:248-270); pi = xx,xy,xz,yy,yz,zz.
template <typename T>
__host__ __device__ inline T PiToFneq(int i, const T* pi) {
  T fneq = T(0);
  int ipi = 0;
  for (int a = 0; a < 3; ++a) {
    for (int b = a; b < 3; ++b) {
      T to_add = T(Ci(i, a) * Ci(i, b));
      if (a == b) {
        to_add -= T(1) / T(3);
      } else {
        if (Ci(i, a) * Ci(i, b) == 0) {
          ++ipi;
          continue;
        }
        to_add *= T(2);
      }
      to_add *= pi[ipi++];
      fneq += to_add;
    }
  }
  fneq *= W<T>(i) * T(3) * T(3) / T(2);
  return fneq;
}

// BouzidiPostCollisionPull (boundary/bouzidi_pull.h:8-26); evaluated in FP64
// for every storage type, as the reference does.
__host__ __device__ inline double BouzidiPull(double f_bb_opp, double f_interior, double f_same, double q) {
  if (q >= 1.0 - 1.0e-14) return f_bb_opp;
  if (q <= 1.0e-14) return f_bb_opp;
  if (q >= 0.5 - 1.0e-14) {
    const double inv2q = 1.0 / (2.0 * q);
    return inv2q * f_bb_opp + (1.0 - inv2q) * f_interior;
  }
  return 2.0 * q * f_bb_opp + (1.0 - 2.0 * q) * f_same;
}

// VelocityBoundaryRhoFromPop (interpolated_velocity.h:170-211) for a fixed
// face direction / orientation (compile-time, like the reference's template).
// Returns r: rho for FP64; for FP32 rho - 1 = (2 normal + on_wall - O u) /
// (1 + O u), the same quotient without the cancelling 1.
template <int D, int O, typename T>
__host__ __device__ inline T VelRhoImpl(const T* f, const T* u) {
  T on_wall = T(0), normal = T(0);
#pragma unroll
  for (int i = 0; i < kQ; ++i)
    if (Ci(i, D) == 0) on_wall += f[i];
#pragma unroll
  for (int i = 0; i < kQ; ++i)
    if (Ci(i, D) == O) normal += f[i];
  if constexpr (kShiftedArith<T>) {
    return (T(2) * normal + on_wall - T(O) * u[D]) / (T(1) + T(O) * u[D]);
  } else {
    return (T(2) * normal + on_wall + T(1)) / (T(1) + T(O) * u[D]);
  }
}

// The ConstRho average as ConstRhoBgk takes it: avg (FP64), avg - 1 (FP32).
template <typename T>
__host__ __device__ inline T AvgArg(double avg) {
  if constexpr (kShiftedArith<T>) {
    return static_cast<T>(avg - 1.0);
  } else {
    return static_cast<T>(avg);
  }
}

// Runtime dispatch; a zero orientation takes the +1 instantiation like the
// reference's dispatch (interpolated_velocity.h:192-211).
template <typename T>
__host__ __device__ inline T VelRhoFromPop(int direction, int orientation, const T* f, const T* u) {
  if (direction == 0) return orientation < 0 ? VelRhoImpl<0, -1>(f, u) : VelRhoImpl<0, 1>(f, u);
  if (direction == 1) return orientation < 0 ? VelRhoImpl<1, -1>(f, u) : VelRhoImpl<1, 1>(f, u);
  return orientation < 0 ? VelRhoImpl<2, -1>(f, u) : VelRhoImpl<2, 1>(f, u);
}

}  // namespace octlb

#endif  // OCTLB_AMD_CSRC_D3Q19_H_
