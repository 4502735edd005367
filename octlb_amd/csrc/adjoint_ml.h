// Launchers of the multi-level adjoint's device operations that share the
// production kernels' device code (defined in adjoint_ml.inc, compiled into
// lbm.hip).  Layout and semantics are described there.
#ifndef OCTLB_AMD_CSRC_ADJOINT_ML_H_
#define OCTLB_AMD_CSRC_ADJOINT_ML_H_

#include "lbm_device.h"

namespace octlb {

// Reverse FD stencils (built once per adjoint handle): for each FD block,
// phase and block cell, the reads of that cell by the phase's FD cells, in
// the order the FD VJP sums them.  code = kind | gradient row << 2 | mirror
// flips of x, y, z << 4.
struct FdRevEntry {
  int ee;       // the reading FD cell (index into the block's FD list)
  int code;
  double coef;  // stencil coefficient
};
constexpr int kFdRevOwnU = 0, kFdRevGrad = 1, kFdRevRhoPost = 2, kFdRevRhoPre = 3;
constexpr int kFdRevMax = 296 * 16;           // per FD block: <= 16 reads per FD cell
constexpr int kFdRevBeg = 3 * 512 + 1;        // per FD block: CSR offsets over (phase, cell)

void LaunchMLFdRevBuild(const short* fd_list, const int4* fd_count, const short* fd_slot, int nfd, int* rbeg,
                        FdRevEntry* rent, hipStream_t s);

template <typename S>
void LaunchMLFd(const FdArgs<S>& a, const int* fd_blocks, const short* fd_slot, int nfd, S* x, const int* boff,
                const int* cnt, double t, hipStream_t s);
template <typename S>
void LaunchMLFdVjp(const FdArgs<S>& a, const int* fd_blocks, const short* fd_slot, int nfd, const S* xpre,
                   double* lam, const int* boff, const int* cnt, double t, const int* rbeg, const FdRevEntry* rent,
                   double* dui, hipStream_t s);
template <typename S>
void LaunchMLProlong(const int* plan, const int* list, int count, int half, S* x, const int* boff, const int* cnt,
                     S* prev, double scale, hipStream_t s);
template <typename S>
void LaunchMLCouplerInit(const int* plan, int npts, const S* x, const int* boff, const int* cnt, S* prev,
                         hipStream_t s);
template <typename S>
void LaunchMLRestrict(const int* plan, const int* list, int count, S* x, const int* boff, const int* cnt,
                      double scale, hipStream_t s);
template <typename S>
void LaunchMLProlongVjp(const int* plan, const int* pts, const int* pbeg, int ngroups, int count, int half,
                        const S* xc, double* lam, const int* boff, const int* cnt, const S* prev, double* lprev,
                        double scale, double* cm, hipStream_t s);
template <typename S>
void LaunchMLCouplerInitVjp(const int* plan, const int* pts, const int* pbeg, int ngroups, const S* x, double* lam,
                            const int* boff, const int* cnt, const double* lprev, hipStream_t s);
void LaunchMLRestrictTake(const int* plan, const int* list, int count, double* lam, const int* boff, const int* cnt,
                          double* buf, hipStream_t s);
template <typename S>
void LaunchMLRestrictVjp(const int* plan, const int* list, const int2* ent, const int* ebeg, const int* ecell,
                         int ncells, const S* xf, double* lam, const int* boff, const int* cnt, const double* buf,
                         double scale, hipStream_t s);

}  // namespace octlb

#endif  // OCTLB_AMD_CSRC_ADJOINT_ML_H_
