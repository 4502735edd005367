/*
 * octlb_oracle.c -- CPU FP64 restatement of the reference's per-time-step
 * AMR D3Q19 lattice-Boltzmann update.  TEST INFRASTRUCTURE (the checker):
 * only tests/, the smoke check and bench.py's cpu_baseline leg call it.
 *
 * Every function follows a reference source location (cited per function)
 * and keeps that source's floating-point evaluation order, so that a build
 * without FMA contraction (-ffp-contract=off) reproduces the reference's
 * FP64 results bit for bit.
 *
 * Parity status of each part is recorded in the header comment of the test
 * that pins it (tests/test_oracle_*.py); no function here is unpinned.
 */
#include "octlb_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define Q 19

static __thread char g_err[512];
const char* ol_last_error(void) { return g_err; }
static int fail(const char* msg) {
  snprintf(g_err, sizeof(g_err), "%s", msg);
  return -1;
}

/* ---------------- D3Q19 tables (descriptor/definition/common.h:166-198) -- */
static const int C[Q][3] = {{0, 0, 0},   {-1, 0, 0}, {0, -1, 0}, {0, 0, -1},  {-1, -1, 0},
                            {-1, 1, 0},  {-1, 0, -1}, {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
                            {1, 0, 0},   {0, 1, 0},  {0, 0, 1},  {1, 1, 0},   {1, -1, 0},
                            {1, 0, 1},   {1, 0, -1}, {0, 1, 1},  {0, 1, -1}};
static const int OPP[Q] = {0, 10, 11, 12, 13, 14, 15, 16, 17, 18, 1, 2, 3, 4, 5, 6, 7, 8, 9};
static const int TW_DEN[Q] = {3, 18, 18, 18, 36, 36, 36, 36, 36, 36,
                              18, 18, 18, 36, 36, 36, 36, 36, 36};
/* Fraction::as<double>() = double(num)/double(den) (utilities/fraction.h:72-75) */
static double TW(int i) { return (double)1 / (double)TW_DEN[i]; }
static const double INV_CS2 = 3.0; /* cs2 = 1/3 -> inverseAs = 3/1 */

enum { K_BULK = 0, K_SOLID, K_BB, K_MBB, K_BOUZIDI, K_VEL, K_PRES, K_OUTFLOW };

/* equilibrium::secondOrder (dynamics/lbm.h:51-62), shifted storage. */
static double feq2(int i, double rho, const double* u, double usqr) {
  double cu = 0.0;
  for (int d = 0; d < 3; ++d) cu += (double)C[i][d] * u[d];
  const double t = TW(i);
  return rho * t *
             (1.0 + INV_CS2 * cu + INV_CS2 * INV_CS2 * 0.5 * cu * cu -
              INV_CS2 * 0.5 * usqr) -
         t;
}

static double norm_sqr(const double* u) {
  double s = 0.0;
  for (int d = 0; d < 3; ++d) s += u[d] * u[d];
  return s;
}

/* CellProxy::computeRhoU: rho starts at 1 (block_lattice.h:113-125). */
static void proxy_rho_u(const double* f, double* rho, double* u) {
  double r = 1.0;
  u[0] = u[1] = u[2] = 0.0;
  for (int i = 0; i < Q; ++i) {
    r += f[i];
    for (int d = 0; d < 3; ++d) u[d] += f[i] * (double)C[i][d];
  }
  for (int d = 0; d < 3; ++d) u[d] /= r;
  *rho = r;
}

/* olb::lbm::computeRhoU: rho = sum(f) + 1 (dynamics/lbm.h:290-331). */
static void olb_rho_u(const double* f, double* rho, double* u) {
  double r = 0.0;
  for (int i = 0; i < Q; ++i) r += f[i];
  r += 1.0;
  u[0] = u[1] = u[2] = 0.0;
  for (int i = 0; i < Q; ++i)
    for (int d = 0; d < 3; ++d) u[d] += f[i] * (double)C[i][d];
  for (int d = 0; d < 3; ++d) u[d] /= r;
  *rho = r;
}

/* olb::lbm::bgkCollision (dynamics/lbm.h:395-403). */
static void olb_bgk(double* f, double rho, const double* u, double omega) {
  const double usqr = norm_sqr(u);
  for (int i = 0; i < Q; ++i) {
    f[i] *= 1.0 - omega;
    f[i] += omega * feq2(i, rho, u, usqr);
  }
}

/* CollideBgkAt (block_lattice.cpp:328-341). */
static void collide_bgk_at(double* f, double omega) {
  double rho, u[3];
  proxy_rho_u(f, &rho, u);
  double usqr = 0.0;
  for (int d = 0; d < 3; ++d) usqr += u[d] * u[d];
  for (int i = 0; i < Q; ++i) {
    const double fe = feq2(i, rho, u, usqr);
    f[i] += omega * (fe - f[i]);
  }
}

/* CollideConstRhoBgkWithMacroscopic (block_lattice.h:133-151). */
static void const_rho_bgk(double* f, double omega, double avg, double rho, const double* u) {
  double usqr = 0.0;
  for (int d = 0; d < 3; ++d) usqr += u[d] * u[d];
  const double delta = 1.0 - avg;
  const double ratio = rho > 1e-12 ? 1.0 + delta / rho : 1.0;
  for (int i = 0; i < Q; ++i) {
    const double fe = feq2(i, rho, u, usqr);
    const double t = TW(i);
    f[i] = ratio * (fe + t) - t + (1.0 - omega) * (f[i] - fe);
  }
}

/* equilibrium::fromPiToFneq (dynamics/lbm.h:248-270); pi in xx,xy,xz,yy,yz,zz. */
static double pi_to_fneq(int i, const double* pi) {
  double fneq = 0.0;
  int ipi = 0;
  for (int a = 0; a < 3; ++a) {
    for (int b = a; b < 3; ++b) {
      double to_add = (double)(C[i][a] * C[i][b]);
      if (a == b) {
        to_add -= 1.0 / INV_CS2;
      } else {
        to_add *= 2.0;
      }
      to_add *= pi[ipi++];
      fneq += to_add;
    }
  }
  fneq *= TW(i) * INV_CS2 * INV_CS2 / 2.0;
  return fneq;
}

/* BouzidiPostCollisionPull (boundary/bouzidi_pull.h:8-26). */
static double bouzidi_pull(double f_bb_opp, double f_interior, double f_same, double q) {
  if (q >= 1.0 - 1.0e-14) return f_bb_opp;
  if (q <= 1.0e-14) return f_bb_opp;
  if (q >= 0.5 - 1.0e-14) {
    const double inv2q = 1.0 / (2.0 * q);
    return inv2q * f_bb_opp + (1.0 - inv2q) * f_interior;
  }
  return 2.0 * q * f_bb_opp + (1.0 - 2.0 * q) * f_same;
}

static int reflects_on_pull(int k) { return k == K_VEL || k == K_PRES || k == K_BOUZIDI; }
static int is_fd_kind(int k) { return k == K_VEL || k == K_PRES; }

/* ---------------- simulation state ------------------------------------- */
typedef struct {
  int coarse_id, ci, cj, ck, fine_id, fi, fj, fk, normal, coarse_level, fine_level, fine_slot;
} ol_point;

typedef struct {
  double rho, u[3], fneq[Q];
} ol_macro;

struct ol_sim {
  int nb, n, h, NP, nslots, ncells;
  double omega;
  int handler, const_rho, stats_scope, bouzidi, snapshot;
  int num_specs;
  octlb_bc_spec* specs;
  int32_t* level;
  double* bounds;
  double* origin; /* [nb*3] */
  double* cw;     /* [nb]   */
  int nsame;
  int32_t* same;
  int ncf;
  octlb_cf_face* cf;
  int ntb;
  int32_t* tb;
  int* tb_blocks; /* unique tree-boundary blocks in first-seen order */
  int ntb_blocks;
  int max_level;
  int** by_level;
  int* n_by_level;
  int nedges;
  int32_t* edges;
  double* f;
  double* tmp;
  double* snap;
  uint8_t* kind;
  double* q;
  int npts;
  ol_point* plan;
  ol_macro* prev;
  int nranges;
  int* lvl_begin;
  int* lvl_end;
  double prolong_scale, restrict_scale;
  int coupler_ready;
  int step;
  double average_rho;
  double t;
  int32_t* collide_count;
  int32_t* stream_count;
  int32_t* calls;
  int ncalls, cap_calls;
  int nthreads;
};

static double* blk(const ol_sim* s, int b) { return s->f + (size_t)b * s->nslots * Q; }
static int slot(const ol_sim* s, int hx, int hy, int hz) {
  return (hx * s->NP + hy) * s->NP + hz;
}
/* BlockLattice::idx: interior (ix,iy,iz) -> slot (block_lattice.h:300-304) */
static int core_slot(const ol_sim* s, int ix, int iy, int iz) {
  return slot(s, ix + s->h, iy + s->h, iz + s->h);
}
static double* cell(const ol_sim* s, int b, int ix, int iy, int iz) {
  return blk(s, b) + (size_t)core_slot(s, ix, iy, iz) * Q;
}
static int kind_at(const ol_sim* s, int b, int ix, int iy, int iz) {
  return s->kind[(size_t)b * s->ncells + (size_t)((ix * s->n + iy) * s->n + iz)];
}

/* ---------------- BC specs and inlet fields ---------------------------- */
static octlb_bc_spec default_spec(int face) {
  octlb_bc_spec sp;
  memset(&sp, 0, sizeof(sp));
  sp.face = face;
  sp.type = OCTLB_BC_NO_SLIP;
  sp.rho_target = 1.0;
  sp.inlet = OCTLB_INLET_NONE;
  return sp;
}

/* detail::FindSpec (interpolated_velocity.h:38-45) */
static octlb_bc_spec find_spec(const ol_sim* s, int face) {
  for (int i = 0; i < s->num_specs; ++i)
    if (s->specs[i].face == face) return s->specs[i];
  return default_spec(face);
}

/* Smoothstep ramp x^3(10+x(-15+6x)) (inlet_velocity_field.h:77-90). */
static double ramp(double t, double ramp_end) {
  if (ramp_end <= 0.0) return 1.0;
  double x = t / ramp_end;
  if (x < 0.0) {
    x = 0.0;
  } else if (x > 1.0) {
    x = 1.0;
  }
  const double x2 = x * x;
  const double x3 = x2 * x;
  return x3 * (10.0 + x * (-15.0 + 6.0 * x));
}

static int in_plane0(int face, int ix, int iy, int iz) {
  (void)iz;
  return (face == OCTLB_XMIN || face == OCTLB_XMAX) ? iy : ix;
}
static int in_plane1(int face, int ix, int iy, int iz) {
  (void)ix;
  return (face == OCTLB_ZMIN || face == OCTLB_ZMAX) ? iy : iz;
}
static double in_plane0_phys(int face, double x, double y) {
  return (face == OCTLB_XMIN || face == OCTLB_XMAX) ? y : x;
}
static double in_plane1_phys(int face, double y, double z) {
  return (face == OCTLB_ZMIN || face == OCTLB_ZMAX) ? y : z;
}
static double clamp01(double c) {
  if (c < 0.0) return 0.0;
  if (c > 1.0) return 1.0;
  return c;
}

/* InletVelocityField::velocity / velocity_phys (inlet_velocity_field.h:19-270). */
static void inlet_velocity(const octlb_bc_spec* sp, int ix, int iy, int iz, double t,
                           const double* origin, double cwidth, double* u) {
  if (sp->inlet == OCTLB_INLET_CHANNEL) {
    /* is_physical(): evaluated from the cell centre (interpolated_velocity.h:96-101) */
    const double px = origin[0] + (ix + 0.5) * cwidth;
    const double py = origin[1] + (iy + 0.5) * cwidth;
    const double pz = origin[2] + (iz + 0.5) * cwidth;
    const double r = ramp(t, sp->ramp_end_t);
    const double c0 =
        clamp01((in_plane0_phys(sp->inlet_face, px, py) - sp->lo0) / (sp->hi0 - sp->lo0));
    const double c1 =
        clamp01((in_plane1_phys(sp->inlet_face, py, pz) - sp->lo1) / (sp->hi1 - sp->lo1));
    const double p0 = 4.0 * c0 * (1.0 - c0);
    const double p1 = 4.0 * c1 * (1.0 - c1);
    const double mag = r * sp->u_peak * p0 * p1;
    for (int d = 0; d < 3; ++d) u[d] = mag * sp->inlet_u[d];
    return;
  }
  if (sp->inlet == OCTLB_INLET_POISEUILLE) {
    const double r = ramp(t, sp->ramp_end_t);
    const int i0 = in_plane0(sp->inlet_face, ix, iy, iz);
    const int i1 = in_plane1(sp->inlet_face, ix, iy, iz);
    const double c0 = ((double)i0 + 0.5) / (double)sp->n_inplane0;
    const double c1 = ((double)i1 + 0.5) / (double)sp->n_inplane1;
    const double p0 = 4.0 * c0 * (1.0 - c0);
    const double p1 = 4.0 * c1 * (1.0 - c1);
    const double mag = r * sp->u_peak * p0 * p1;
    for (int d = 0; d < 3; ++d) u[d] = mag * sp->inlet_u[d];
    return;
  }
  /* UNIFORM */
  const double r = ramp(t, sp->ramp_end_t);
  for (int d = 0; d < 3; ++d) u[d] = r * sp->inlet_u[d];
}

/* detail::PrescribedBoundaryU (interpolated_velocity.h:69-117): block-local
 * face tests, the first touched face whose spec has an inlet field wins,
 * otherwise the cavity lid fallback on interior top-face cells. */
static void prescribed_u(const ol_sim* s, int b, int ix, int iy, int iz, double t, double* u) {
  const int n = s->n;
  u[0] = u[1] = u[2] = 0.0;
  const int on[6] = {ix == 0, ix == n - 1, iy == 0, iy == n - 1, iz == 0, iz == n - 1};
  for (int f = 0; f < 6; ++f) {
    if (!on[f]) continue;
    const octlb_bc_spec sp = find_spec(s, f);
    if (sp.inlet != OCTLB_INLET_NONE) {
      inlet_velocity(&sp, ix, iy, iz, t, s->origin + (size_t)b * 3, s->cw[b], u);
      return;
    }
  }
  if (iy == n - 1 && ix > 0 && ix < n - 1 && iz > 0 && iz < n - 1) {
    const octlb_bc_spec sp = find_spec(s, OCTLB_YMAX);
    for (int d = 0; d < 3; ++d) u[d] = sp.u_wall[d];
  }
}

/* detail::FlatBoundaryFaceInfo (interpolated_velocity.h:134-167). */
static int flat_face_info(int n, int ix, int iy, int iz, int* direction, int* orientation) {
  int count = 0;
  *direction = 0;
  *orientation = 0;
  if (ix == 0) {
    ++count;
    *direction = 0;
    *orientation = -1;
  } else if (ix == n - 1) {
    ++count;
    *direction = 0;
    *orientation = 1;
  }
  if (iy == 0) {
    ++count;
    *direction = 1;
    *orientation = -1;
  } else if (iy == n - 1) {
    ++count;
    *direction = 1;
    *orientation = 1;
  }
  if (iz == 0) {
    ++count;
    *direction = 2;
    *orientation = -1;
  } else if (iz == n - 1) {
    ++count;
    *direction = 2;
    *orientation = 1;
  }
  return count == 1;
}

/* VelocityBoundaryRhoFromPop (interpolated_velocity.h:170-211); a zero
 * orientation dispatches to the +1 instantiation, as the reference does. */
static double vel_rho_from_pop(int direction, int orientation, const double* f, const double* u) {
  const int o = orientation < 0 ? -1 : 1;
  double on_wall = 0.0, normal = 0.0;
  for (int i = 0; i < Q; ++i)
    if (C[i][direction] == 0) on_wall += f[i];
  for (int i = 0; i < Q; ++i)
    if (C[i][direction] == o) normal += f[i];
  return (2.0 * normal + on_wall + 1.0) / (1.0 + (double)o * u[direction]);
}

/* ---------------- collide dispatch ------------------------------------- */
typedef struct {
  double sum;
  int count;
} ol_stats;

/* BoundaryCellStatisticRho + CollideDirichletBoundaryCellAt
 * (interpolated_velocity.h:815-884). */
static void collide_dirichlet(ol_sim* s, int b, int ix, int iy, int iz, double t,
                              ol_stats* st) {
  const int n = s->n;
  double u[3];
  prescribed_u(s, b, ix, iy, iz, t, u);
  double* f = cell(s, b, ix, iy, iz);
  const int fc = (ix == 0) + (ix == n - 1) + (iy == 0) + (iy == n - 1) + (iz == 0) + (iz == n - 1);
  double rho;
  if (fc >= 2) {
    double ut[3];
    proxy_rho_u(f, &rho, ut);
  } else {
    int dir, ori;
    flat_face_info(n, ix, iy, iz, &dir, &ori);
    rho = vel_rho_from_pop(dir, ori, f, u);
  }
  olb_bgk(f, rho, u, s->omega);
  if (st) {
    st->sum += rho;
    st->count++;
  }
}

/* BlockLattice::collide_const_rho_at (block_lattice.cpp:440-455). */
static void collide_const_rho_at(ol_sim* s, int b, int ix, int iy, int iz, ol_stats* st) {
  if (kind_at(s, b, ix, iy, iz) != K_BULK) return;
  double* f = cell(s, b, ix, iy, iz);
  double rho, u[3];
  proxy_rho_u(f, &rho, u);
  if (st) {
    st->sum += rho + (1.0 - s->average_rho);
    st->count++;
  }
  proxy_rho_u(f, &rho, u);
  const_rho_bgk(f, s->omega, s->average_rho, rho, u);
}

/* BcDispatcher::collide (bc_dispatcher.h:32-81). */
static void dispatch_collide(ol_sim* s, int b, int ix, int iy, int iz, double t, ol_stats* st) {
  const int k = kind_at(s, b, ix, iy, iz);
  double* f = cell(s, b, ix, iy, iz);
  switch (k) {
    case K_BULK:
      if (s->const_rho) {
        collide_const_rho_at(s, b, ix, iy, iz, st);
      } else {
        double rho, u[3];
        proxy_rho_u(f, &rho, u);
        olb_bgk(f, rho, u, s->omega);
      }
      return;
    case K_BB: {
      double tmp[Q];
      for (int i = 0; i < Q; ++i) tmp[i] = f[i];
      for (int i = 0; i < Q; ++i) f[i] = tmp[OPP[i]];
      return;
    }
    case K_VEL:
      collide_dirichlet(s, b, ix, iy, iz, t, st);
      return;
    default:
      return;
  }
}

/* One block's collide for TimeLoop::collide_level (time_loop.h:173-192):
 * concrete handler -> collide_interleaved_with (domain_boundary_handler.cpp:60-76);
 * NoOp handler -> BlockLattice::collide / collide_const_rho (block_lattice.cpp:376-499).
 * The overlap-padding BGK pass is a no-op in both collide modes because
 * OverlapPaddingMaterialNonZero is true only inside the core (block_lattice.h:46-52). */
static void collide_block(ol_sim* s, int b, double t, ol_stats* st) {
  const int n = s->n;
  if (s->handler == OCTLB_HANDLER_CONCRETE) {
    for (int ix = 0; ix < n; ++ix)
      for (int iy = 0; iy < n; ++iy)
        for (int iz = 0; iz < n; ++iz) dispatch_collide(s, b, ix, iy, iz, t, st);
    return;
  }
  for (int ix = 0; ix < n; ++ix)
    for (int iy = 0; iy < n; ++iy)
      for (int iz = 0; iz < n; ++iz) {
        if (s->const_rho) {
          collide_const_rho_at(s, b, ix, iy, iz, st);
        } else if (kind_at(s, b, ix, iy, iz) == K_BULK) {
          collide_bgk_at(cell(s, b, ix, iy, iz), s->omega);
        }
      }
}

/* ---------------- stream (block_lattice.cpp:501-721) -------------------- */
static int floor_mod(int v, int m) {
  int w = v % m;
  if (w < 0) w += m;
  return w;
}

static void stream_block(ol_sim* s, int b, double* tmp) {
  const int n = s->n, h = s->h, NP = s->NP, NS = s->nslots;
  double* f = blk(s, b);
  const int rotate = !s->bouzidi;
  if (rotate) {
    /* OpenLB CyclicColumn rotate over the padded block (:551-577, 676-688) */
    for (int d = 0; d < NS; ++d)
      for (int i = 0; i < Q; ++i) {
        const int off = C[i][0] * NP * NP + C[i][1] * NP + C[i][2];
        tmp[(size_t)d * Q + i] = f[(size_t)floor_mod(d - off, NS) * Q + i];
      }
    memcpy(f, tmp, sizeof(double) * (size_t)NS * Q);
    return;
  }
  for (int hx = 0; hx < NP; ++hx)
    for (int hy = 0; hy < NP; ++hy)
      for (int hz = 0; hz < NP; ++hz) {
        const int ix = hx - h, iy = hy - h, iz = hz - h;
        const int core = ix >= 0 && ix < n && iy >= 0 && iy < n && iz >= 0 && iz < n;
        const int dst_lin = slot(s, hx, hy, hz);
        double* out = tmp + (size_t)dst_lin * Q;
        const double* self = f + (size_t)dst_lin * Q;
        if (!core) {
          /* overlap padding: in-halo direct pull, else padded torus wrap (:595-611) */
          for (int i = 0; i < Q; ++i) {
            const int sx = hx - C[i][0], sy = hy - C[i][1], sz = hz - C[i][2];
            if (sx >= 0 && sx < NP && sy >= 0 && sy < NP && sz >= 0 && sz < NP) {
              out[i] = f[(size_t)slot(s, sx, sy, sz) * Q + i];
            } else {
              const int off = C[i][0] * NP * NP + C[i][1] * NP + C[i][2];
              out[i] = f[(size_t)floor_mod(dst_lin - off, NS) * Q + i];
            }
          }
          continue;
        }
        const int k = kind_at(s, b, ix, iy, iz);
        if (k == K_SOLID) continue;
        for (int i = 0; i < Q; ++i) {
          const int sx = hx - C[i][0], sy = hy - C[i][1], sz = hz - C[i][2];
          if (!(sx >= 0 && sx < NP && sy >= 0 && sy < NP && sz >= 0 && sz < NP)) {
            out[i] = reflects_on_pull(k) ? self[OPP[i]] : self[i];
            continue;
          }
          const int sxi = sx - h, syi = sy - h, szi = sz - h;
          const int src_interior = sxi >= 0 && sxi < n && syi >= 0 && syi < n && szi >= 0 && szi < n;
          if (reflects_on_pull(k) && !src_interior) {
            out[i] = self[OPP[i]];
            continue;
          }
          const int src_kind = src_interior ? kind_at(s, b, sxi, syi, szi) : K_BULK;
          if (src_kind == K_SOLID || src_kind == K_BOUZIDI) {
            const double qf = s->q[((size_t)b * s->ncells + (size_t)((ix * n + iy) * n + iz)) * Q + i];
            if (qf > 1.0e-10 && qf < 1.0 - 1.0e-10) {
              const double f_bb = self[OPP[i]];
              double f_int = self[i];
              const int nxh = hx + C[i][0], nyh = hy + C[i][1], nzh = hz + C[i][2];
              const int nxi = nxh - h, nyi = nyh - h, nzi = nzh - h;
              const int nk = (nxi >= 0 && nxi < n && nyi >= 0 && nyi < n && nzi >= 0 && nzi < n)
                                 ? kind_at(s, b, nxi, nyi, nzi)
                                 : K_BULK;
              if (nk == K_BULK) f_int = f[(size_t)slot(s, nxh, nyh, nzh) * Q + i];
              out[i] = bouzidi_pull(f_bb, f_int, self[i], qf);
              continue;
            }
            out[i] = self[OPP[i]];
            continue;
          }
          out[i] = f[(size_t)slot(s, sx, sy, sz) * Q + i];
        }
      }
  /* copy back every streamed slot (:708-720) */
  for (int hx = 0; hx < NP; ++hx)
    for (int hy = 0; hy < NP; ++hy)
      for (int hz = 0; hz < NP; ++hz) {
        const int ix = hx - h, iy = hy - h, iz = hz - h;
        const int core = ix >= 0 && ix < n && iy >= 0 && iy < n && iz >= 0 && iz < n;
        if (core && kind_at(s, b, ix, iy, iz) == K_SOLID) continue;
        const size_t o = (size_t)slot(s, hx, hy, hz) * Q;
        memcpy(f + o, tmp + o, sizeof(double) * Q);
      }
}

/* ---------------- halo exchange (ghost_schedule.h:121-255) -------------- */
static void face_ranges(const ol_sim* s, int dir, int ghost, int* lo, int* hi) {
  /* lo/hi per axis of the halo coordinates covered by the face layer */
  const int n = s->n, h = s->h;
  for (int a = 0; a < 3; ++a) {
    lo[a] = h;
    hi[a] = n + h;
  }
  const int axis = dir / 2, is_max = dir & 1;
  int layer;
  if (ghost) {
    layer = is_max ? n + h : h - 1; /* unpack_face (block_lattice.cpp:141-194) */
  } else {
    layer = is_max ? n + h - 1 : h; /* pack_face (:31-84) */
  }
  lo[axis] = layer;
  hi[axis] = layer + 1;
}

static void copy_face(ol_sim* s, int src_b, int dir, int dst_b) {
  /* pack src interior layer at dir, unpack into dst ghost layer at opposite(dir);
   * both traverse the two tangential axes in ascending order, so the buffer
   * maps tangential coordinates 1:1. */
  int slo[3], shi[3], dlo[3], dhi[3];
  face_ranges(s, dir, 0, slo, shi);
  face_ranges(s, dir ^ 1, 1, dlo, dhi);
  const double* fs = blk(s, src_b);
  double* fd = blk(s, dst_b);
  const int axis = dir / 2;
  for (int x = slo[0]; x < shi[0]; ++x)
    for (int y = slo[1]; y < shi[1]; ++y)
      for (int z = slo[2]; z < shi[2]; ++z) {
        int p[3] = {x, y, z};
        p[axis] = dlo[axis];
        memcpy(fd + (size_t)slot(s, p[0], p[1], p[2]) * Q,
               fs + (size_t)slot(s, x, y, z) * Q, sizeof(double) * Q);
      }
}

static void copy_edge(ol_sim* s, int a, int d1, int d2, int dd) {
  /* pack_edge / unpack_edge (block_lattice.cpp:196-263) */
  const int n = s->n, h = s->h;
  const int a1 = d1 / 2, a2 = d2 / 2, a3 = 3 - a1 - a2;
  const int max1 = d1 & 1, max2 = d2 & 1;
  const int f1 = (max1 ? n - 1 : 0) + h, f2 = (max2 ? n - 1 : 0) + h;
  const int g1 = (!max1) ? n + h : h - 1; /* opposite(d1) ghost */
  const int g2 = (!max2) ? n + h : h - 1;
  const double* fs = blk(s, a);
  double* fd = blk(s, dd);
  for (int k = 0; k < n; ++k) {
    int ps[3], pd[3];
    ps[a1] = f1;
    ps[a2] = f2;
    ps[a3] = k + h;
    pd[a1] = g1;
    pd[a2] = g2;
    pd[a3] = k + h;
    memcpy(fd + (size_t)slot(s, pd[0], pd[1], pd[2]) * Q,
           fs + (size_t)slot(s, ps[0], ps[1], ps[2]) * Q, sizeof(double) * Q);
  }
}

static void exchange(ol_sim* s) {
  for (int e = 0; e < s->nsame; ++e)
    copy_face(s, s->same[3 * e], s->same[3 * e + 1], s->same[3 * e + 2]);
  for (int e = 0; e < s->nedges; ++e)
    copy_edge(s, s->edges[4 * e], s->edges[4 * e + 1], s->edges[4 * e + 2], s->edges[4 * e + 3]);
}

/* ---------------- post-stream FD (interpolated_velocity.h:214-1036) ----- */
typedef struct {
  ol_sim* s;
  int b;
  double t;
} ol_view;

static int v_in_domain(const ol_view* v, int ix, int iy, int iz) {
  const int n = v->s->n;
  return ix >= 0 && ix < n && iy >= 0 && iy < n && iz >= 0 && iz < n;
}
static int v_padding(const ol_view* v, int ix, int iy, int iz) {
  const int n = v->s->n, h = v->s->h;
  const int st = h < 2 ? h : 2;
  const int in_st = ix >= -st && ix < n + st && iy >= -st && iy < n + st && iz >= -st && iz < n + st;
  return !v_in_domain(v, ix, iy, iz) && h > 0 && in_st;
}
static int mirror_index(int i, int n) {
  if (i < 0) return -i;
  if (i >= n) return 2 * (n - 1) - i;
  return i;
}
static void v_compute_u(const ol_view* v, int ix, int iy, int iz, double* u);

static void v_extrapolated_u(const ol_view* v, int ix, int iy, int iz, double* u) {
  int dir, ori;
  if (!flat_face_info(v->s->n, ix, iy, iz, &dir, &ori)) {
    u[0] = u[1] = u[2] = 0.0;
    return;
  }
  int nix = ix, niy = iy, niz = iz;
  if (dir == 0) {
    nix = ix - ori;
  } else if (dir == 1) {
    niy = iy - ori;
  } else {
    niz = iz - ori;
  }
  v_compute_u(v, nix, niy, niz, u);
}

static void v_interior_u(const ol_view* v, int ix, int iy, int iz, double* u) {
  const int k = kind_at(v->s, v->b, ix, iy, iz);
  if (k == K_VEL) {
    prescribed_u(v->s, v->b, ix, iy, iz, v->t, u);
    return;
  }
  if (k == K_PRES) {
    v_extrapolated_u(v, ix, iy, iz, u);
    return;
  }
  double rho;
  proxy_rho_u(cell(v->s, v->b, ix, iy, iz), &rho, u);
}

static void v_mirror_u(const ol_view* v, int ix, int iy, int iz, double* u) {
  const int n = v->s->n;
  double ui[3];
  v_interior_u(v, mirror_index(ix, n), mirror_index(iy, n), mirror_index(iz, n), ui);
  u[0] = (ix < 0 || ix >= n) ? -ui[0] : ui[0];
  u[1] = (iy < 0 || iy >= n) ? -ui[1] : ui[1];
  u[2] = (iz < 0 || iz >= n) ? -ui[2] : ui[2];
}

static double v_prescribed_rho(const ol_view* v, int ix, int iy, int iz) {
  int dir, ori;
  if (flat_face_info(v->s->n, ix, iy, iz, &dir, &ori)) {
    const int face = dir * 2 + (ori < 0 ? 0 : 1);
    return find_spec(v->s, face).rho_target;
  }
  return 1.0;
}

static double v_compute_rho(const ol_view* v, int ix, int iy, int iz) {
  if (v_padding(v, ix, iy, iz)) return 1.0;
  if (!v_in_domain(v, ix, iy, iz)) {
    const int n = v->s->n;
    return v_compute_rho(v, mirror_index(ix, n), mirror_index(iy, n), mirror_index(iz, n));
  }
  const int k = kind_at(v->s, v->b, ix, iy, iz);
  if (k == K_PRES) return v_prescribed_rho(v, ix, iy, iz);
  const double* f = cell(v->s, v->b, ix, iy, iz);
  if (k == K_VEL) {
    int dir, ori;
    if (flat_face_info(v->s->n, ix, iy, iz, &dir, &ori)) {
      double u[3];
      prescribed_u(v->s, v->b, ix, iy, iz, v->t, u);
      return vel_rho_from_pop(dir, ori, f, u);
    }
  }
  double rho, u[3];
  proxy_rho_u(f, &rho, u);
  return rho;
}

static void v_compute_u(const ol_view* v, int ix, int iy, int iz, double* u) {
  if (v_padding(v, ix, iy, iz)) {
    u[0] = u[1] = u[2] = 0.0;
    return;
  }
  if (!v_in_domain(v, ix, iy, iz)) {
    v_mirror_u(v, ix, iy, iz, u);
    return;
  }
  const int k = kind_at(v->s, v->b, ix, iy, iz);
  if (k == K_VEL) {
    prescribed_u(v->s, v->b, ix, iy, iz, v->t, u);
    return;
  }
  if (k == K_PRES) {
    v_extrapolated_u(v, ix, iy, iz, u);
    return;
  }
  double rho;
  proxy_rho_u(cell(v->s, v->b, ix, iy, iz), &rho, u);
}

static void v_compute_rho_u(const ol_view* v, int ix, int iy, int iz, double* rho, double* u) {
  if (v_padding(v, ix, iy, iz)) {
    *rho = 1.0;
    u[0] = u[1] = u[2] = 0.0;
    return;
  }
  if (!v_in_domain(v, ix, iy, iz)) {
    v_mirror_u(v, ix, iy, iz, u);
    *rho = 1.0;
    return;
  }
  const int k = kind_at(v->s, v->b, ix, iy, iz);
  if (k == K_VEL) {
    prescribed_u(v->s, v->b, ix, iy, iz, v->t, u);
    *rho = v_compute_rho(v, ix, iy, iz);
    return;
  }
  if (k == K_PRES) {
    *rho = v_prescribed_rho(v, ix, iy, iz);
    v_extrapolated_u(v, ix, iy, iz, u);
    return;
  }
  proxy_rho_u(cell(v->s, v->b, ix, iy, iz), rho, u);
}

static double boundary_gradient(double u0, double u1, double u2) {
  return (-3.0 * u0 + 4.0 * u1 - 1.0 * u2) / 2.0;
}
static double central_gradient(double up, double um) { return (up - um) / 2.0; }

/* WriteFromPi (interpolated_velocity.h:446-453) */
static void write_from_pi(double* f, double rho, const double* u, const double* pi) {
  const double usqr = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  for (int i = 0; i < Q; ++i) f[i] = feq2(i, rho, u, usqr) + pi_to_fneq(i, pi);
}

static void pi_from_grad(double rho, double omega, const double* dx, const double* dy,
                         const double* dz, double* pi) {
  const double s_to_pi = -rho / INV_CS2 / omega;
  pi[0] = 2.0 * dx[0] * s_to_pi;
  pi[3] = 2.0 * dy[1] * s_to_pi;
  pi[5] = 2.0 * dz[2] * s_to_pi;
  pi[1] = (dx[1] + dy[0]) * s_to_pi;
  pi[2] = (dx[2] + dz[0]) * s_to_pi;
  pi[4] = (dy[2] + dz[1]) * s_to_pi;
}

static void normal_grad(const ol_view* v, int ix, int iy, int iz, int axis, int normal, double* g) {
  const int ox = axis == 0 ? -normal : 0, oy = axis == 1 ? -normal : 0, oz = axis == 2 ? -normal : 0;
  double u0[3], u1[3], u2[3];
  v_compute_u(v, ix, iy, iz, u0);
  v_compute_u(v, ix + ox, iy + oy, iz + oz, u1);
  v_compute_u(v, ix + 2 * ox, iy + 2 * oy, iz + 2 * oz, u2);
  for (int d = 0; d < 3; ++d) g[d] = (double)(-normal) * boundary_gradient(u0[d], u1[d], u2[d]);
}

static void tangential_grad(const ol_view* v, int ix, int iy, int iz, int tdx, int tdy, int tdz,
                            double* g) {
  double up[3], um[3];
  v_compute_u(v, ix + tdx, iy + tdy, iz + tdz, up);
  v_compute_u(v, ix - tdx, iy - tdy, iz - tdz, um);
  for (int d = 0; d < 3; ++d) g[d] = central_gradient(up[d], um[d]);
}

/* ApplyPlaneFdBoundary (:570-636) */
static void apply_plane_fd(const ol_view* v, double* f, int ix, int iy, int iz, int direction,
                           int orientation) {
  double rho, u[3];
  v_compute_rho_u(v, ix, iy, iz, &rho, u);
  double dx[3] = {0, 0, 0}, dy[3] = {0, 0, 0}, dz[3] = {0, 0, 0};
  if (direction == 0) {
    normal_grad(v, ix, iy, iz, 0, orientation, dx);
    tangential_grad(v, ix, iy, iz, 0, 1, 0, dy);
    tangential_grad(v, ix, iy, iz, 0, 0, 1, dz);
  } else if (direction == 1) {
    tangential_grad(v, ix, iy, iz, 1, 0, 0, dx);
    normal_grad(v, ix, iy, iz, 1, orientation, dy);
    tangential_grad(v, ix, iy, iz, 0, 0, 1, dz);
  } else {
    tangential_grad(v, ix, iy, iz, 1, 0, 0, dx);
    tangential_grad(v, ix, iy, iz, 0, 1, 0, dy);
    normal_grad(v, ix, iy, iz, 2, orientation, dz);
  }
  double pi[6];
  pi_from_grad(rho, v->s->omega, dx, dy, dz, pi);
  write_from_pi(f, rho, u, pi);
}

/* ApplyCornerBoundary (:638-685) */
static void apply_corner_fd(const ol_view* v, double* f, int ix, int iy, int iz, int xn, int yn,
                            int zn) {
  const double r100 = v_compute_rho(v, ix - xn, iy, iz);
  const double r010 = v_compute_rho(v, ix, iy - yn, iz);
  const double r001 = v_compute_rho(v, ix, iy, iz - zn);
  const double r200 = v_compute_rho(v, ix - 2 * xn, iy, iz);
  const double r020 = v_compute_rho(v, ix, iy - 2 * yn, iz);
  const double r002 = v_compute_rho(v, ix, iy, iz - 2 * zn);
  const double rho = 4.0 / 9.0 * (r001 + r010 + r100) - 1.0 / 9.0 * (r002 + r020 + r200);
  double dx[3], dy[3], dz[3], u[3];
  normal_grad(v, ix, iy, iz, 0, xn, dx);
  normal_grad(v, ix, iy, iz, 1, yn, dy);
  normal_grad(v, ix, iy, iz, 2, zn, dz);
  v_compute_u(v, ix, iy, iz, u);
  double pi[6];
  pi_from_grad(rho, v->s->omega, dx, dy, dz, pi);
  write_from_pi(f, rho, u, pi);
}

/* ApplyEdgeBoundary (:687-752) */
static void apply_edge_fd(const ol_view* v, double* f, int ix, int iy, int iz, int plane, int n1,
                          int n2) {
  const int d1 = (plane + 1) % 3, d2 = (plane + 2) % 3;
  double rs[4];
  const int steps[4][2] = {{1, 0}, {0, 1}, {2, 0}, {0, 2}};
  for (int k = 0; k < 4; ++k) {
    int c[3] = {0, 0, 0};
    c[d1] = -n1 * steps[k][0];
    c[d2] = -n2 * steps[k][1];
    rs[k] = v_compute_rho(v, ix + c[0], iy + c[1], iz + c[2]);
  }
  const double rho = 2.0 / 3.0 * (rs[1] + rs[0]) - 1.0 / 6.0 * (rs[3] + rs[2]);
  double g[3][3];
  memset(g, 0, sizeof(g));
  {
    const int tdx = plane == 0, tdy = plane == 1, tdz = plane == 2;
    tangential_grad(v, ix, iy, iz, tdx, tdy, tdz, g[plane]);
  }
  normal_grad(v, ix, iy, iz, d1, n1, g[d1]);
  normal_grad(v, ix, iy, iz, d2, n2, g[d2]);
  double u[3];
  v_compute_u(v, ix, iy, iz, u);
  const double s_to_pi = -rho / INV_CS2 / v->s->omega;
  double pi[6];
  pi[0] = 2.0 * g[0][0] * s_to_pi;
  pi[3] = 2.0 * g[1][1] * s_to_pi;
  pi[5] = 2.0 * g[2][2] * s_to_pi;
  pi[1] = (g[0][1] + g[1][0]) * s_to_pi;
  pi[2] = (g[0][2] + g[2][0]) * s_to_pi;
  pi[4] = (g[1][2] + g[2][1]) * s_to_pi;
  write_from_pi(f, rho, u, pi);
}

/* ApplyInterpolatedVelocityBoundaryCells (:950-1036); the kHybrid padding
 * fill before it (:926-948) copies nothing because PaddingMaterialNonZero is
 * true only inside the core. */
static void post_stream_block(ol_sim* s, int b, double t) {
  const int n = s->n;
  ol_view v = {s, b, t};
  for (int ix = 0; ix < n; ++ix)
    for (int iy = 0; iy < n; ++iy)
      for (int iz = 0; iz < n; ++iz) {
        if (!is_fd_kind(kind_at(s, b, ix, iy, iz))) continue;
        const int oxm = ix == 0, oxp = ix == n - 1, oym = iy == 0, oyp = iy == n - 1,
                  ozm = iz == 0, ozp = iz == n - 1;
        const int fc = oxm + oxp + oym + oyp + ozm + ozp;
        double* f = cell(s, b, ix, iy, iz);
        const int xn = oxm ? -1 : oxp ? 1 : 0;
        const int yn = oym ? -1 : oyp ? 1 : 0;
        const int zn = ozm ? -1 : ozp ? 1 : 0;
        if (fc == 3) {
          apply_corner_fd(&v, f, ix, iy, iz, xn, yn, zn);
          continue;
        }
        if (fc == 2) {
          int plane, n1, n2;
          if (xn != 0 && yn != 0) {
            plane = 2;
            n1 = xn;
            n2 = yn;
          } else if (xn != 0 && zn != 0) {
            plane = 1;
            n1 = xn;
            n2 = zn;
          } else {
            plane = 0;
            n1 = yn;
            n2 = zn;
          }
          apply_edge_fd(&v, f, ix, iy, iz, plane, n1, n2);
          continue;
        }
        int face = OCTLB_XMIN;
        if (oxp) {
          face = OCTLB_XMAX;
        } else if (oyp) {
          face = OCTLB_YMAX;
        } else if (oym) {
          face = OCTLB_YMIN;
        } else if (ozp) {
          face = OCTLB_ZMAX;
        } else if (ozm) {
          face = OCTLB_ZMIN;
        }
        apply_plane_fd(&v, f, ix, iy, iz, face / 2, (face & 1) ? 1 : -1);
      }
}

/* ---------------- Lagrava coupler (level_coupler.cpp) ------------------- */
static void read_macro(const ol_sim* s, int b, int ix, int iy, int iz, ol_macro* m) {
  const double* f = cell(s, b, ix, iy, iz);
  olb_rho_u(f, &m->rho, m->u);
  const double usqr = norm_sqr(m->u);
  for (int i = 0; i < Q; ++i) m->fneq[i] = f[i] - feq2(i, m->rho, m->u, usqr);
}

static void write_prolong(ol_sim* s, int b, int ix, int iy, int iz, const ol_macro* m,
                          double scaling) {
  double* f = cell(s, b, ix, iy, iz);
  const double usqr = norm_sqr(m->u);
  for (int i = 0; i < Q; ++i) f[i] = feq2(i, m->rho, m->u, usqr) + scaling * m->fneq[i];
}

static void build_plan(ol_sim* s) {
  const int n = s->n;
  s->npts = 0;
  s->plan = (ol_point*)malloc(sizeof(ol_point) * (size_t)(s->ncf > 0 ? s->ncf : 1) * n * n);
  for (int e = 0; e < s->ncf; ++e) {
    const octlb_cf_face* fc = &s->cf[e];
    const double* cb = fc->coarse_bounds;
    const double cc[3] = {0.5 * (cb[0] + cb[3]), 0.5 * (cb[1] + cb[4]), 0.5 * (cb[2] + cb[5])};
    const int axis = fc->normal / 2, is_max = fc->normal & 1;
    const int t0 = (axis + 1) % 3, t1 = (axis + 2) % 3;
    const int cni = is_max ? n - 1 : 0, fni = is_max ? 0 : n - 1;
    const int half = n / 2;
    for (int sl = 0; sl < 4; ++sl) {
      const double* fb = fc->fine_bounds[sl];
      const double fcn[3] = {0.5 * (fb[0] + fb[3]), 0.5 * (fb[1] + fb[4]), 0.5 * (fb[2] + fb[5])};
      const int qu0 = fcn[t0] < cc[t0] ? 0 : 1;
      const int qu1 = fcn[t1] < cc[t1] ? 0 : 1;
      const int b0 = qu0 * half, b1 = qu1 * half;
      for (int a = b0; a < b0 + half; ++a)
        for (int c = b1; c < b1 + half; ++c) {
          int ci[3], fi[3];
          ci[axis] = cni;
          fi[axis] = fni;
          ci[t0] = a;
          ci[t1] = c;
          fi[t0] = a - b0;
          fi[t1] = c - b1;
          ol_point* p = &s->plan[s->npts++];
          p->coarse_id = fc->coarse_id;
          p->ci = ci[0];
          p->cj = ci[1];
          p->ck = ci[2];
          p->fine_id = fc->fine_ids[sl];
          p->fi = fi[0];
          p->fj = fi[1];
          p->fk = fi[2];
          p->normal = fc->normal;
          p->coarse_level = fc->coarse_level;
          p->fine_level = fc->fine_level;
          p->fine_slot = sl;
        }
    }
  }
  int maxl = 0;
  for (int i = 0; i < s->npts; ++i)
    if (s->plan[i].coarse_level > maxl) maxl = s->plan[i].coarse_level;
  s->nranges = maxl + 1;
  s->lvl_begin = (int*)malloc(sizeof(int) * s->nranges);
  s->lvl_end = (int*)malloc(sizeof(int) * s->nranges);
  for (int l = 0; l < s->nranges; ++l) {
    s->lvl_begin[l] = s->npts;
    s->lvl_end[l] = 0;
  }
  for (int i = 0; i < s->npts; ++i) {
    const int l = s->plan[i].coarse_level;
    if (i < s->lvl_begin[l]) s->lvl_begin[l] = i;
    if (i + 1 > s->lvl_end[l]) s->lvl_end[l] = i + 1;
  }
  s->prev = (ol_macro*)calloc((size_t)(s->npts > 0 ? s->npts : 1), sizeof(ol_macro));
}

static int range_begin(const ol_sim* s, int l) {
  if (l < 0 || l >= s->nranges) return s->npts;
  return s->lvl_begin[l];
}
static int range_end(const ol_sim* s, int l) {
  if (l < 0 || l >= s->nranges) return s->npts;
  return s->lvl_end[l];
}

/* LevelCoupler ctor reads prev from the coarse cells (level_coupler.cpp:262-270). */
static void coupler_init(ol_sim* s) {
  for (int i = 0; i < s->npts; ++i) {
    const ol_point* p = &s->plan[i];
    read_macro(s, p->coarse_id, p->ci, p->cj, p->ck, &s->prev[i]);
  }
  s->coupler_ready = 1;
}

/* ApplyProlongation (level_coupler.cpp:438-478) */
static void apply_prolongation(ol_sim* s, int begin, int end, int half_time) {
  for (int i = begin; i < end; ++i) {
    const ol_point* p = &s->plan[i];
    ol_macro m;
    read_macro(s, p->coarse_id, p->ci, p->cj, p->ck, &m);
    if (half_time) {
      const ol_macro* pv = &s->prev[i];
      m.rho = 0.5 * (pv->rho + m.rho);
      for (int d = 0; d < 3; ++d) m.u[d] = 0.5 * (pv->u[d] + m.u[d]);
      for (int k = 0; k < Q; ++k) m.fneq[k] = 0.5 * (pv->fneq[k] + m.fneq[k]);
    }
    write_prolong(s, p->fine_id, p->fi, p->fj, p->fk, &m, s->prolong_scale);
    if (!half_time) read_macro(s, p->coarse_id, p->ci, p->cj, p->ck, &s->prev[i]);
  }
}

static void log_call(ol_sim* s, int phase, int level) {
  if (s->ncalls >= s->cap_calls) {
    s->cap_calls = s->cap_calls ? 2 * s->cap_calls : 64;
    s->calls = (int32_t*)realloc(s->calls, sizeof(int32_t) * 2 * (size_t)s->cap_calls);
  }
  s->calls[2 * s->ncalls] = phase;
  s->calls[2 * s->ncalls + 1] = level;
  s->ncalls++;
}

/* LevelCoupler::restrict, local-local branch (level_coupler.cpp:613-654) */
static void apply_restriction(ol_sim* s, int fine_level) {
  const int n = s->n;
  for (int i = 0; i < s->npts; ++i) {
    const ol_point* p = &s->plan[i];
    if (p->fine_level != fine_level) continue;
    ol_macro fm;
    read_macro(s, p->fine_id, p->fi, p->fj, p->fk, &fm);
    double sum[Q];
    for (int k = 0; k < Q; ++k) sum[k] = 0.0;
    int cnt = 0;
    for (int j = 1; j < Q; ++j) {
      const int x = p->fi + C[j][0], y = p->fj + C[j][1], z = p->fk + C[j][2];
      if (x < 0 || x >= n || y < 0 || y >= n || z < 0 || z >= n) continue;
      ol_macro nm;
      read_macro(s, p->fine_id, x, y, z, &nm);
      for (int k = 0; k < Q; ++k) sum[k] += nm.fneq[k];
      ++cnt;
    }
    ol_macro avg = fm;
    for (int k = 0; k < Q; ++k) avg.fneq[k] = (fm.fneq[k] + sum[k]) / (double)(cnt + 1);
    write_prolong(s, p->coarse_id, p->ci, p->cj, p->ck, &avg, s->restrict_scale);
  }
}

/* ---------------- time loop (time_loop.h:173-246) ----------------------- */
static void advance_level(ol_sim* s, int level) {
  ol_stats st = {0.0, 0};
  ol_stats* sp = (s->const_rho && level >= s->max_level) ? &st : NULL;
  const int nbl = s->n_by_level[level];
  const int* ids = s->by_level[level];
  /* collide_level */
  if (sp != NULL) {
    for (int k = 0; k < nbl; ++k) collide_block(s, ids[k], s->t, sp);
  } else {
#pragma omp parallel for schedule(dynamic, 4) num_threads(s->nthreads)
    for (int k = 0; k < nbl; ++k) collide_block(s, ids[k], s->t, NULL);
  }
  s->collide_count[level]++;
  /* NoOp handler apply() does nothing; the concrete handler collides its
   * boundary cells inside collide_interleaved_with (openlb order). */
  exchange(s);
  if (s->snapshot) {
    for (int k = 0; k < nbl; ++k)
      memcpy(s->snap + (size_t)ids[k] * s->nslots * Q, blk(s, ids[k]),
             sizeof(double) * (size_t)s->nslots * Q);
  }
#pragma omp parallel num_threads(s->nthreads)
  {
    double* tmp = (double*)malloc(sizeof(double) * (size_t)s->nslots * Q);
#pragma omp for schedule(dynamic, 4)
    for (int k = 0; k < nbl; ++k) stream_block(s, ids[k], tmp);
    free(tmp);
  }
  s->stream_count[level]++;
  if (s->handler == OCTLB_HANDLER_CONCRETE) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads)
    for (int k = 0; k < s->ntb_blocks; ++k) post_stream_block(s, s->tb_blocks[k], s->t);
  }
  if (sp != NULL) s->average_rho = st.count > 0 ? st.sum / (double)st.count : 1.0;
  if (level >= s->max_level) return;
  apply_prolongation(s, range_begin(s, level), range_end(s, level), 1);
  log_call(s, 0, level);
  advance_level(s, level + 1);
  apply_prolongation(s, range_begin(s, level), range_end(s, level), 0);
  log_call(s, 1, level);
  advance_level(s, level + 1);
  apply_restriction(s, level + 1);
  log_call(s, 2, level);
}

int ol_advance(ol_sim* s, int32_t steps) {
  if (!s) return fail("null sim");
  if (!s->coupler_ready) coupler_init(s);
  for (int k = 0; k < steps; ++k) {
    s->t = (double)s->step;
    advance_level(s, 0);
    s->step++;
  }
  return 0;
}

/* ---------------- construction ----------------------------------------- */
ol_sim* ol_create(const octlb_topology* topo, const octlb_params* prm) {
  if (!topo || !prm) {
    fail("null argument");
    return NULL;
  }
  if (topo->n < 2 || topo->num_blocks < 1 || topo->halo < 1) {
    fail("need n >= 2, halo >= 1 and at least one block");
    return NULL;
  }
  ol_sim* s = (ol_sim*)calloc(1, sizeof(ol_sim));
  s->nb = topo->num_blocks;
  s->n = topo->n;
  s->h = topo->halo;
  s->NP = s->n + 2 * s->h;
  s->nslots = s->NP * s->NP * s->NP;
  s->ncells = s->n * s->n * s->n;
  s->omega = prm->omega;
  s->handler = prm->handler;
  s->const_rho = prm->const_rho;
  s->stats_scope = prm->stats_scope;
  s->bouzidi = prm->bouzidi;
  s->snapshot = prm->snapshot;
  s->num_specs = prm->num_specs;
  s->specs = (octlb_bc_spec*)malloc(sizeof(octlb_bc_spec) * (size_t)(prm->num_specs + 1));
  if (prm->num_specs > 0) memcpy(s->specs, prm->specs, sizeof(octlb_bc_spec) * (size_t)prm->num_specs);
  s->level = (int32_t*)malloc(sizeof(int32_t) * (size_t)s->nb);
  memcpy(s->level, topo->level, sizeof(int32_t) * (size_t)s->nb);
  s->bounds = (double*)malloc(sizeof(double) * 6 * (size_t)s->nb);
  memcpy(s->bounds, topo->bounds, sizeof(double) * 6 * (size_t)s->nb);
  s->origin = (double*)malloc(sizeof(double) * 3 * (size_t)s->nb);
  s->cw = (double*)malloc(sizeof(double) * (size_t)s->nb);
  for (int b = 0; b < s->nb; ++b) {
    const double* bb = s->bounds + 6 * (size_t)b;
    s->origin[3 * b] = bb[0];
    s->origin[3 * b + 1] = bb[1];
    s->origin[3 * b + 2] = bb[2];
    s->cw[b] = (bb[3] - bb[0]) / (double)s->n; /* cylinder3d_case.h:404-409 */
  }
  s->nsame = topo->num_same_level;
  s->same = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(s->nsame + 1));
  if (s->nsame) memcpy(s->same, topo->same_level, sizeof(int32_t) * 3 * (size_t)s->nsame);
  s->ncf = topo->num_coarse_fine;
  s->cf = (octlb_cf_face*)malloc(sizeof(octlb_cf_face) * (size_t)(s->ncf + 1));
  if (s->ncf) memcpy(s->cf, topo->coarse_fine, sizeof(octlb_cf_face) * (size_t)s->ncf);
  s->ntb = topo->num_tree_boundary;
  s->tb = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)(s->ntb + 1));
  if (s->ntb) memcpy(s->tb, topo->tree_boundary, sizeof(int32_t) * 2 * (size_t)s->ntb);
  /* apply_post_stream: each tree-boundary block once, first-seen order */
  {
    char* seen = (char*)calloc((size_t)s->nb, 1);
    s->tb_blocks = (int*)malloc(sizeof(int) * (size_t)(s->ntb + 1));
    s->ntb_blocks = 0;
    for (int e = 0; e < s->ntb; ++e) {
      const int b = s->tb[2 * e];
      if (b < 0 || b >= s->nb) continue;
      if (!seen[b]) {
        seen[b] = 1;
        s->tb_blocks[s->ntb_blocks++] = b;
      }
    }
    free(seen);
  }
  /* octants_by_level_ (time_loop.h:139-160) */
  s->max_level = 0;
  for (int b = 0; b < s->nb; ++b)
    if (s->level[b] > s->max_level) s->max_level = s->level[b];
  s->by_level = (int**)calloc((size_t)s->max_level + 1, sizeof(int*));
  s->n_by_level = (int*)calloc((size_t)s->max_level + 1, sizeof(int));
  for (int l = 0; l <= s->max_level; ++l) s->by_level[l] = (int*)malloc(sizeof(int) * (size_t)s->nb);
  for (int b = 0; b < s->nb; ++b) {
    const int l = s->level[b];
    s->by_level[l][s->n_by_level[l]++] = b;
  }
  s->collide_count = (int32_t*)calloc((size_t)s->max_level + 1, sizeof(int32_t));
  s->stream_count = (int32_t*)calloc((size_t)s->max_level + 1, sizeof(int32_t));
  /* same-rank edges by face composition (ghost_schedule.h:56-91) */
  {
    int32_t* fn = (int32_t*)malloc(sizeof(int32_t) * 6 * (size_t)s->nb);
    for (int i = 0; i < 6 * s->nb; ++i) fn[i] = -1;
    for (int e = 0; e < s->nsame; ++e) fn[s->same[3 * e] * 6 + s->same[3 * e + 1]] = s->same[3 * e + 2];
    int cap = 64;
    s->edges = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)cap);
    s->nedges = 0;
    for (int a = 0; a < s->nb; ++a)
      for (int d1 = 0; d1 < 6; ++d1)
        for (int d2 = 0; d2 < 6; ++d2) {
          if (d1 / 2 == d2 / 2) continue;
          const int bb = fn[a * 6 + d1];
          if (bb < 0 || bb >= s->nb) continue;
          const int dd = fn[bb * 6 + d2];
          if (dd < 0 || dd >= s->nb) continue;
          if (s->nedges == cap) {
            cap *= 2;
            s->edges = (int32_t*)realloc(s->edges, sizeof(int32_t) * 4 * (size_t)cap);
          }
          s->edges[4 * s->nedges] = a;
          s->edges[4 * s->nedges + 1] = d1;
          s->edges[4 * s->nedges + 2] = d2;
          s->edges[4 * s->nedges + 3] = dd;
          s->nedges++;
        }
    free(fn);
  }
  s->f = (double*)calloc((size_t)s->nb * s->nslots * Q, sizeof(double));
  s->snap = s->snapshot ? (double*)calloc((size_t)s->nb * s->nslots * Q, sizeof(double)) : NULL;
  s->kind = (uint8_t*)calloc((size_t)s->nb * s->ncells, 1);
  if (s->bouzidi) {
    const size_t nq = (size_t)s->nb * s->ncells * Q;
    s->q = (double*)malloc(sizeof(double) * nq);
    for (size_t i = 0; i < nq; ++i) s->q[i] = 1.0;
  }
  {
    const double tau = 1.0 / s->omega; /* level_coupler.cpp:226-228 */
    s->prolong_scale = (tau - 0.25) / tau;
    s->restrict_scale = tau / (tau - 0.25);
  }
  build_plan(s);
  s->average_rho = 1.0;
  s->nthreads = 1;
  if (!s->f || !s->kind) {
    fail("out of memory");
    ol_destroy(s);
    return NULL;
  }
  return s;
}

void ol_destroy(ol_sim* s) {
  if (!s) return;
  free(s->specs);
  free(s->level);
  free(s->bounds);
  free(s->origin);
  free(s->cw);
  free(s->same);
  free(s->cf);
  free(s->tb);
  free(s->tb_blocks);
  if (s->by_level)
    for (int l = 0; l <= s->max_level; ++l) free(s->by_level[l]);
  free(s->by_level);
  free(s->n_by_level);
  free(s->collide_count);
  free(s->stream_count);
  free(s->edges);
  free(s->f);
  free(s->snap);
  free(s->kind);
  free(s->q);
  free(s->plan);
  free(s->prev);
  free(s->lvl_begin);
  free(s->lvl_end);
  free(s->calls);
  free(s);
}

int ol_set_threads(ol_sim* s, int32_t n) {
  if (!s || n < 1) return fail("bad thread count");
  s->nthreads = n;
  return 0;
}

int ol_set_step(ol_sim* s, int32_t step) {
  if (!s) return fail("null sim");
  s->step = step;
  return 0;
}

int ol_set_bc_kinds(ol_sim* s, const uint8_t* kinds) {
  if (!s || !kinds) return fail("null argument");
  memcpy(s->kind, kinds, (size_t)s->nb * s->ncells);
  return 0;
}

int ol_set_links(ol_sim* s, int64_t count, const int32_t* link, const double* q) {
  if (!s) return fail("null sim");
  if (!s->q) return fail("solver built without Bouzidi links");
  for (int64_t i = 0; i < count; ++i) {
    const int b = link[3 * i], c = link[3 * i + 1], p = link[3 * i + 2];
    if (b < 0 || b >= s->nb || c < 0 || c >= s->ncells || p < 0 || p >= Q) return fail("link out of range");
    s->q[((size_t)b * s->ncells + (size_t)c) * Q + (size_t)p] = q[i];
  }
  return 0;
}

/* BlockLattice::initialize / initialize_from_material: feq(rho0,u0) on the
 * selected core cells (block_lattice.cpp:272-285, lattice_material_init.cpp:11-33). */
int ol_init_equilibrium(ol_sim* s, double rho0, const double u0[3], uint32_t kind_mask) {
  if (!s) return fail("null sim");
  const double usqr = norm_sqr(u0);
  double fe[Q];
  for (int i = 0; i < Q; ++i) fe[i] = feq2(i, rho0, u0, usqr);
  const int n = s->n;
  for (int b = 0; b < s->nb; ++b)
    for (int ix = 0; ix < n; ++ix)
      for (int iy = 0; iy < n; ++iy)
        for (int iz = 0; iz < n; ++iz) {
          if (!((kind_mask >> kind_at(s, b, ix, iy, iz)) & 1u)) continue;
          memcpy(cell(s, b, ix, iy, iz), fe, sizeof(fe));
        }
  s->coupler_ready = 0;
  return 0;
}

int ol_init_field(ol_sim* s, const double* rho, const double* u) {
  if (!s || !rho || !u) return fail("null argument");
  const int n = s->n;
  for (int b = 0; b < s->nb; ++b)
    for (int ix = 0; ix < n; ++ix)
      for (int iy = 0; iy < n; ++iy)
        for (int iz = 0; iz < n; ++iz) {
          const size_t c = (size_t)b * s->ncells + (size_t)((ix * n + iy) * n + iz);
          const double* uc = u + 3 * c;
          const double usqr = norm_sqr(uc);
          double* f = cell(s, b, ix, iy, iz);
          for (int i = 0; i < Q; ++i) f[i] = feq2(i, rho[c], uc, usqr);
        }
  s->coupler_ready = 0;
  return 0;
}

int ol_set_block(ol_sim* s, int32_t b, const double* aos) {
  if (!s || !aos || b < 0 || b >= s->nb) return fail("bad block");
  memcpy(blk(s, b), aos, sizeof(double) * (size_t)s->nslots * Q);
  s->coupler_ready = 0;
  return 0;
}

int ol_get_block(const ol_sim* s, int32_t b, double* aos) {
  if (!s || !aos || b < 0 || b >= s->nb) return fail("bad block");
  memcpy(aos, blk(s, b), sizeof(double) * (size_t)s->nslots * Q);
  return 0;
}

int ol_get_snapshot_block(const ol_sim* s, int32_t b, double* aos) {
  if (!s || !aos || b < 0 || b >= s->nb || !s->snap) return fail("no snapshot");
  memcpy(aos, s->snap + (size_t)b * s->nslots * Q, sizeof(double) * (size_t)s->nslots * Q);
  return 0;
}

int ol_get_core(const ol_sim* s, double* out) {
  if (!s || !out) return fail("null argument");
  const int n = s->n;
  for (int b = 0; b < s->nb; ++b)
    for (int ix = 0; ix < n; ++ix)
      for (int iy = 0; iy < n; ++iy)
        for (int iz = 0; iz < n; ++iz) {
          const size_t c = (size_t)b * s->ncells + (size_t)((ix * n + iy) * n + iz);
          memcpy(out + c * Q, cell(s, b, ix, iy, iz), sizeof(double) * Q);
        }
  return 0;
}

/* MomentumExchangeDrag::force_on_fluid_if per block, summed over blocks
 * (drag_postprocessor.h:66-115, cylinder3d_case.h:473-503). */
int ol_drag(const ol_sim* s, const double bbox[6], double out[3]) {
  if (!s || !bbox || !out) return fail("null argument");
  const int n = s->n;
  double tot[3] = {0.0, 0.0, 0.0};
  for (int b = 0; b < s->nb; ++b) {
    const double* org = s->origin + 3 * (size_t)b;
    const double cwb = s->cw[b];
    double fo[3] = {0.0, 0.0, 0.0};
    for (int iz = 0; iz < n; ++iz)
      for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix) {
          if (kind_at(s, b, ix, iy, iz) != K_BULK) continue;
          const size_t o = (size_t)core_slot(s, ix, iy, iz) * Q;
          const double* live = blk(s, b) + o;
          const double* src = s->snap ? s->snap + (size_t)b * s->nslots * Q + o : live;
          for (int i = 1; i < Q; ++i) {
            const int x = ix + C[i][0], y = iy + C[i][1], z = iz + C[i][2];
            if (x < 0 || x >= n || y < 0 || y >= n || z < 0 || z >= n) continue;
            const int nk = kind_at(s, b, x, y, z);
            if (nk != K_SOLID && nk != K_BOUZIDI) continue;
            const double px = org[0] + (x + 0.5) * cwb;
            const double py = org[1] + (y + 0.5) * cwb;
            const double pz = org[2] + (z + 0.5) * cwb;
            if (!(px >= bbox[0] && px <= bbox[3] && py >= bbox[1] && py <= bbox[4] &&
                  pz >= bbox[2] && pz <= bbox[5]))
              continue;
            const double fout = src[i];
            const double fbar = live[OPP[i]];
            fo[0] += (fout + fbar) * (double)C[i][0];
            fo[1] += (fout + fbar) * (double)C[i][1];
            fo[2] += (fout + fbar) * (double)C[i][2];
          }
        }
    tot[0] += -fo[0];
    tot[1] += -fo[1];
    tot[2] += -fo[2];
  }
  out[0] = tot[0];
  out[1] = tot[1];
  out[2] = tot[2];
  return 0;
}

/* Cylinder3dCase::total_mass (cylinder3d_case.h:601-622). */
double ol_mass(const ol_sim* s) {
  const int n = s->n;
  double m = 0.0;
  for (int b = 0; b < s->nb; ++b)
    for (int iz = 0; iz < n; ++iz)
      for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix) {
          if (kind_at(s, b, ix, iy, iz) != K_BULK) continue;
          double rho, u[3];
          proxy_rho_u(cell(s, b, ix, iy, iz), &rho, u);
          m += rho;
        }
  return m;
}

double ol_average_rho(const ol_sim* s) { return s->average_rho; }

int ol_counters(const ol_sim* s, int32_t* collide, int32_t* stream, int32_t max_levels,
                int32_t* calls, int32_t max_calls, int32_t* num_calls) {
  if (!s) return fail("null sim");
  for (int l = 0; l < max_levels && l <= s->max_level; ++l) {
    if (collide) collide[l] = s->collide_count[l];
    if (stream) stream[l] = s->stream_count[l];
  }
  if (num_calls) *num_calls = s->ncalls;
  if (calls)
    for (int k = 0; k < s->ncalls && k < max_calls; ++k) {
      calls[2 * k] = s->calls[2 * k];
      calls[2 * k + 1] = s->calls[2 * k + 1];
    }
  return 0;
}

int32_t ol_plan_size(const ol_sim* s) { return s ? s->npts : 0; }

int ol_plan(const ol_sim* s, int32_t* rows) {
  if (!s || !rows) return fail("null argument");
  for (int i = 0; i < s->npts; ++i) {
    const ol_point* p = &s->plan[i];
    int32_t* r = rows + 12 * (size_t)i;
    r[0] = p->coarse_id;
    r[1] = p->ci;
    r[2] = p->cj;
    r[3] = p->ck;
    r[4] = p->fine_id;
    r[5] = p->fi;
    r[6] = p->fj;
    r[7] = p->fk;
    r[8] = p->normal;
    r[9] = p->coarse_level;
    r[10] = p->fine_level;
    r[11] = p->fine_slot;
  }
  return 0;
}

int32_t ol_num_edges(const ol_sim* s) { return s ? s->nedges : 0; }

int ol_edges(const ol_sim* s, int32_t* rows) {
  if (!s || !rows) return fail("null argument");
  memcpy(rows, s->edges, sizeof(int32_t) * 4 * (size_t)s->nedges);
  return 0;
}

int ol_block_collide(ol_sim* s, int32_t b, double t) {
  if (!s || b < 0 || b >= s->nb) return fail("bad block");
  ol_stats st = {0.0, 0};
  collide_block(s, b, t, s->const_rho ? &st : NULL);
  if (s->const_rho) s->average_rho = st.count > 0 ? st.sum / (double)st.count : 1.0;
  return 0;
}

int ol_block_stream(ol_sim* s, int32_t b) {
  if (!s || b < 0 || b >= s->nb) return fail("bad block");
  double* tmp = (double*)malloc(sizeof(double) * (size_t)s->nslots * Q);
  stream_block(s, b, tmp);
  free(tmp);
  return 0;
}

int ol_block_post_stream(ol_sim* s, int32_t b, double t) {
  if (!s || b < 0 || b >= s->nb) return fail("bad block");
  post_stream_block(s, b, t);
  return 0;
}

int ol_coupler_init(ol_sim* s) {
  if (!s) return fail("null sim");
  coupler_init(s);
  return 0;
}

int ol_apply_prolongation(ol_sim* s, int32_t level, int32_t half_time) {
  if (!s) return fail("null sim");
  apply_prolongation(s, range_begin(s, level), range_end(s, level), half_time != 0);
  return 0;
}

int ol_apply_restriction(ol_sim* s, int32_t fine_level) {
  if (!s) return fail("null sim");
  apply_restriction(s, fine_level);
  return 0;
}

/* ------------------------------------------------------------------
 * Bouzidi link build from a triangle soup (bouzidi_link_data.cpp:15-233).
 * Every expression keeps the reference's operand order (built without FMA
 * contraction). */

/* ClosestPointOnSegment (bouzidi_link_data.cpp:19-40); std::clamp(t,0,1). */
static void seg_closest(const double* p, const double* a, const double* b, double* out) {
  const double abx = b[0] - a[0], aby = b[1] - a[1], abz = b[2] - a[2];
  const double apx = p[0] - a[0], apy = p[1] - a[1], apz = p[2] - a[2];
  const double ab2 = abx * abx + aby * aby + abz * abz;
  double t = 0.0;
  if (ab2 > 0.0) {
    t = (apx * abx + apy * aby + apz * abz) / ab2;
    t = t < 0.0 ? 0.0 : (1.0 < t ? 1.0 : t);
  }
  out[0] = a[0] + t * abx;
  out[1] = a[1] + t * aby;
  out[2] = a[2] + t * abz;
}

/* ClosestPointOnTriangle (bouzidi_link_data.cpp:42-76): the nearest of the
 * three edge points, first one kept on ties; v0 if none is finite. */
static void tri_closest(const double* tri, const double* p, double* out) {
  double best = DBL_MAX;
  out[0] = tri[0];
  out[1] = tri[1];
  out[2] = tri[2];
  for (int e = 0; e < 3; ++e) {
    const double* a = tri + 3 * e;
    const double* b = tri + 3 * ((e + 1) % 3);
    double c[3];
    seg_closest(p, a, b, c);
    const double dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < best) {
      best = d2;
      out[0] = c[0];
      out[1] = c[1];
      out[2] = c[2];
    }
  }
}

int64_t ol_build_links(int32_t nb, int32_t n, const double* bounds, const uint8_t* material, int64_t ntri,
                       const double* tris, int64_t cap, int32_t* links, double* q) {
  if (nb < 0 || n < 1 || (nb > 0 && (!bounds || !material)) || ntri < 0 || (ntri > 0 && !tris)) {
    fail("bad link-build arguments");
    return -1;
  }
  if (ntri == 0) return 0; /* soup.empty(): no links */
  const int64_t ncell = (int64_t)n * n * n;
  double* dense = (double*)malloc(sizeof(double) * (size_t)ncell * Q);
  int64_t count = 0;
  for (int32_t b = 0; b < nb; ++b) {
    const double* bb = bounds + 6 * (size_t)b;
    const uint8_t* mat = material + (size_t)b * ncell;
    for (int64_t e = 0; e < ncell * Q; ++e) dense[e] = -1.0;
    const double dx = (bb[3] - bb[0]) / (double)n;
    const double dy = (bb[4] - bb[1]) / (double)n;
    const double dz = (bb[5] - bb[2]) / (double)n;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          if (mat[((size_t)k * n + j) * n + i] != 0) continue; /* kFluid only */
          /* CellCenter (bouzidi_link_data.cpp:85-94) */
          const double p[3] = {bb[0] + ((double)i + 0.5) * dx, bb[1] + ((double)j + 0.5) * dy,
                               bb[2] + ((double)k + 0.5) * dz};
          double w[3] = {p[0], p[1], p[2]};
          double best = DBL_MAX;
          for (int64_t t = 0; t < ntri; ++t) {
            double c[3];
            tri_closest(tris + 9 * t, p, c);
            const double d2 = (p[0] - c[0]) * (p[0] - c[0]) + (p[1] - c[1]) * (p[1] - c[1]) +
                              (p[2] - c[2]) * (p[2] - c[2]);
            if (d2 < best) {
              best = d2;
              w[0] = c[0];
              w[1] = c[1];
              w[2] = c[2];
            }
          }
          for (int ip = 1; ip < Q; ++ip) {
            const int* e = C[OPP[ip]];
            const int ni = i + e[0], nj = j + e[1], nk = k + e[2];
            if (ni < 0 || ni >= n || nj < 0 || nj >= n || nk < 0 || nk >= n) continue; /* kFluid */
            const uint8_t m = mat[((size_t)nk * n + nj) * n + ni];
            if (m != 1 && m != 2) continue; /* kSolid, kBoundary */
            const double dir_x = (double)e[0] * dx, dir_y = (double)e[1] * dx, dir_z = (double)e[2] * dx;
            const double dn2 = dir_x * dir_x + dir_y * dir_y + dir_z * dir_z;
            if (dn2 <= 1.0e-30) continue;
            const double tx = w[0] - p[0], ty = w[1] - p[1], tz = w[2] - p[2];
            const double proj = (tx * dir_x + ty * dir_y + tz * dir_z) / dn2;
            if (proj > 1.0e-8 && proj < 1.0 - 1.0e-8) dense[(((size_t)i * n + j) * n + k) * Q + ip] = proj;
          }
        }
    for (int64_t c = 0; c < ncell; ++c)
      for (int ip = 1; ip < Q; ++ip) {
        const double v = dense[c * Q + ip];
        if (v < 0.0) continue;
        if (count < cap) {
          links[3 * count] = b;
          links[3 * count + 1] = (int32_t)c;
          links[3 * count + 2] = ip;
          q[count] = v;
        }
        ++count;
      }
  }
  free(dense);
  return count;
}

/* ---------------- discrete adjoint of the single-level step -------------
 * Not a restatement: the reference has no adjoint.  This is the exact
 * transpose (FP64) of the functions above for one level with the concrete
 * handler and Bouzidi streaming, written loop for loop against
 * collide_block / exchange / stream_block so that it checks the device
 * adjoint.  It is pinned by central differences of ol_advance itself. */

/* y = olb_bgk(x, proxy_rho_u(x)) (block_lattice.h:113-125, dynamics/lbm.h:395-403):
 * lx_j = (1-w) ly_j + w (A + sum_d G_d (c_jd - u_d) / rho) with
 * A = sum_i ly_i dfeq_i/drho and G_d = sum_i ly_i dfeq_i/du_d. */
static void bgk_vjp(const double* x, const double* ly, double omega, double* lx) {
  double rho, u[3];
  proxy_rho_u(x, &rho, u);
  const double usqr = norm_sqr(u);
  double A = 0.0, G[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < Q; ++i) {
    double cu = 0.0;
    for (int d = 0; d < 3; ++d) cu += (double)C[i][d] * u[d];
    const double t = TW(i);
    A += ly[i] * t * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usqr);
    for (int d = 0; d < 3; ++d) G[d] += ly[i] * rho * t * (3.0 * C[i][d] + 9.0 * cu * C[i][d] - 3.0 * u[d]);
  }
  for (int j = 0; j < Q; ++j) {
    double g = A;
    for (int d = 0; d < 3; ++d) g += G[d] * ((double)C[j][d] - u[d]) / rho;
    lx[j] = (1.0 - omega) * ly[j] + omega * g;
  }
}

/* transpose of stream_block's per-cell branch (Bouzidi path) for block b:
 * ly += M_b^T lx, with y the post-exchange state the forward pulled from;
 * dq (dense like s->q) += lx * d out / d q on Bouzidi rows. */
static void stream_block_adj(const ol_sim* s, int b, const double* y, const double* lx, double* ly,
                             double* dq) {
  const int n = s->n, h = s->h, NP = s->NP, NS = s->nslots;
  const double* f = y + (size_t)b * NS * Q;
  const double* lxb = lx + (size_t)b * NS * Q;
  double* lyb = ly + (size_t)b * NS * Q;
  for (int hx = 0; hx < NP; ++hx)
    for (int hy = 0; hy < NP; ++hy)
      for (int hz = 0; hz < NP; ++hz) {
        const int ix = hx - h, iy = hy - h, iz = hz - h;
        const int core = ix >= 0 && ix < n && iy >= 0 && iy < n && iz >= 0 && iz < n;
        const int dst = slot(s, hx, hy, hz);
        const double* lo = lxb + (size_t)dst * Q;
        if (!core) {
          for (int i = 0; i < Q; ++i) {
            const int sx = hx - C[i][0], sy = hy - C[i][1], sz = hz - C[i][2];
            int src;
            if (sx >= 0 && sx < NP && sy >= 0 && sy < NP && sz >= 0 && sz < NP) {
              src = slot(s, sx, sy, sz);
            } else {
              const int off = C[i][0] * NP * NP + C[i][1] * NP + C[i][2];
              src = floor_mod(dst - off, NS);
            }
            lyb[(size_t)src * Q + i] += lo[i];
          }
          continue;
        }
        const int k = kind_at(s, b, ix, iy, iz);
        if (k == K_SOLID) {
          for (int i = 0; i < Q; ++i) lyb[(size_t)dst * Q + i] += lo[i];
          continue;
        }
        double* lself = lyb + (size_t)dst * Q;
        const double* self = f + (size_t)dst * Q;
        for (int i = 0; i < Q; ++i) {
          const int sx = hx - C[i][0], sy = hy - C[i][1], sz = hz - C[i][2];
          if (!(sx >= 0 && sx < NP && sy >= 0 && sy < NP && sz >= 0 && sz < NP)) {
            lself[reflects_on_pull(k) ? OPP[i] : i] += lo[i];
            continue;
          }
          const int sxi = sx - h, syi = sy - h, szi = sz - h;
          const int src_interior = sxi >= 0 && sxi < n && syi >= 0 && syi < n && szi >= 0 && szi < n;
          if (reflects_on_pull(k) && !src_interior) {
            lself[OPP[i]] += lo[i];
            continue;
          }
          const int src_kind = src_interior ? kind_at(s, b, sxi, syi, szi) : K_BULK;
          if (src_kind == K_SOLID || src_kind == K_BOUZIDI) {
            const size_t qi = ((size_t)b * s->ncells + (size_t)((ix * n + iy) * n + iz)) * Q + i;
            const double qf = s->q[qi];
            if (qf > 1.0e-10 && qf < 1.0 - 1.0e-10) {
              const int nxh = hx + C[i][0], nyh = hy + C[i][1], nzh = hz + C[i][2];
              const int nxi = nxh - h, nyi = nyh - h, nzi = nzh - h;
              const int nk = (nxi >= 0 && nxi < n && nyi >= 0 && nyi < n && nzi >= 0 && nzi < n)
                                 ? kind_at(s, b, nxi, nyi, nzi)
                                 : K_BULK;
              const size_t int_slot = nk == K_BULK ? (size_t)slot(s, nxh, nyh, nzh) : (size_t)dst;
              const double f_bb = self[OPP[i]];
              const double f_int = f[int_slot * Q + i];
              if (qf >= 0.5 - 1.0e-14) {
                const double a = 1.0 / (2.0 * qf);
                lself[OPP[i]] += a * lo[i];
                lyb[int_slot * Q + i] += (1.0 - a) * lo[i];
                dq[qi] += lo[i] * (-(f_bb - f_int) / (2.0 * qf * qf));
              } else {
                lself[OPP[i]] += 2.0 * qf * lo[i];
                lself[i] += (1.0 - 2.0 * qf) * lo[i];
                dq[qi] += lo[i] * (2.0 * f_bb - 2.0 * self[i]);
              }
              continue;
            }
            lself[OPP[i]] += lo[i];
            continue;
          }
          lyb[(size_t)slot(s, sx, sy, sz) * Q + i] += lo[i];
        }
      }
}

/* transpose of exchange(): every copy dst <- src becomes l[src] += l[dst], l[dst] = 0 */
static void move_adj(double* l, size_t src, size_t dst) {
  for (int i = 0; i < Q; ++i) {
    l[src + i] += l[dst + i];
    l[dst + i] = 0.0;
  }
}
static void exchange_adj(const ol_sim* s, double* l) {
  const int n = s->n, h = s->h;
  const size_t bs = (size_t)s->nslots * Q;
  for (int e = s->nedges - 1; e >= 0; --e) {
    const int a = s->edges[4 * e], d1 = s->edges[4 * e + 1], d2 = s->edges[4 * e + 2], dd = s->edges[4 * e + 3];
    const int a1 = d1 / 2, a2 = d2 / 2, a3 = 3 - a1 - a2;
    const int max1 = d1 & 1, max2 = d2 & 1;
    const int f1 = (max1 ? n - 1 : 0) + h, f2 = (max2 ? n - 1 : 0) + h;
    const int g1 = (!max1) ? n + h : h - 1, g2 = (!max2) ? n + h : h - 1;
    for (int k = 0; k < n; ++k) {
      int ps[3], pd[3];
      ps[a1] = f1;
      ps[a2] = f2;
      ps[a3] = k + h;
      pd[a1] = g1;
      pd[a2] = g2;
      pd[a3] = k + h;
      move_adj(l, (size_t)a * bs + (size_t)slot(s, ps[0], ps[1], ps[2]) * Q,
               (size_t)dd * bs + (size_t)slot(s, pd[0], pd[1], pd[2]) * Q);
    }
  }
  for (int e = s->nsame - 1; e >= 0; --e) {
    const int src_b = s->same[3 * e], dir = s->same[3 * e + 1], dst_b = s->same[3 * e + 2];
    int slo[3], shi[3], dlo[3], dhi[3];
    face_ranges(s, dir, 0, slo, shi);
    face_ranges(s, dir ^ 1, 1, dlo, dhi);
    const int axis = dir / 2;
    for (int x = slo[0]; x < shi[0]; ++x)
      for (int y = slo[1]; y < shi[1]; ++y)
        for (int z = slo[2]; z < shi[2]; ++z) {
          int p[3] = {x, y, z};
          p[axis] = dlo[axis];
          move_adj(l, (size_t)src_b * bs + (size_t)slot(s, x, y, z) * Q,
                   (size_t)dst_b * bs + (size_t)slot(s, p[0], p[1], p[2]) * Q);
        }
  }
}

/* seeds d(dir . F)/d(snap) and d(dir . F)/d(live) of ol_drag's sum, times w */
static void drag_seed(const ol_sim* s, const double bbox[6], const double dir[3], double w, double* l_live,
                      double* l_snap) {
  const int n = s->n;
  for (int b = 0; b < s->nb; ++b) {
    const double* org = s->origin + 3 * (size_t)b;
    const double cwb = s->cw[b];
    for (int iz = 0; iz < n; ++iz)
      for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix) {
          if (kind_at(s, b, ix, iy, iz) != K_BULK) continue;
          const size_t o = (size_t)b * s->nslots * Q + (size_t)core_slot(s, ix, iy, iz) * Q;
          for (int i = 1; i < Q; ++i) {
            const int x = ix + C[i][0], y = iy + C[i][1], z = iz + C[i][2];
            if (x < 0 || x >= n || y < 0 || y >= n || z < 0 || z >= n) continue;
            const int nk = kind_at(s, b, x, y, z);
            if (nk != K_SOLID && nk != K_BOUZIDI) continue;
            const double px = org[0] + (x + 0.5) * cwb, py = org[1] + (y + 0.5) * cwb,
                         pz = org[2] + (z + 0.5) * cwb;
            if (!(px >= bbox[0] && px <= bbox[3] && py >= bbox[1] && py <= bbox[4] && pz >= bbox[2] &&
                  pz <= bbox[5]))
              continue;
            const double g = -w * (dir[0] * C[i][0] + dir[1] * C[i][1] + dir[2] * C[i][2]);
            l_snap[o + i] += g;
            l_live[o + OPP[i]] += g;
          }
        }
  }
}

int ol_adjoint(ol_sim* s, int32_t steps, int32_t win_begin, int32_t win_end, const double dir[3],
               const double bbox[6], double rho0, const double u0[3], uint32_t kind_mask, int64_t nlinks,
               const int32_t* link, double* J, double dJ_du0[3], double* dJ_dq) {
  if (!s || !dir || !bbox || !u0 || !J || !dJ_du0 || (nlinks > 0 && (!link || !dJ_dq)))
    return fail("null argument");
  if (s->max_level != 0) return fail("adjoint: single level only");
  if (s->handler != OCTLB_HANDLER_CONCRETE || s->const_rho || !s->bouzidi || !s->snapshot)
    return fail("adjoint: needs the concrete handler, plain BGK, Bouzidi streaming and snapshots");
  if (steps < 1 || win_begin < 0 || win_end <= win_begin || win_end > steps) return fail("adjoint: bad window");
  const size_t tot = (size_t)s->nb * s->nslots * Q;
  for (size_t c = 0; c < (size_t)s->nb * s->ncells; ++c)
    if (is_fd_kind(s->kind[c])) return fail("adjoint: FD boundary kinds are not differentiated");
  double* xs = (double*)malloc(sizeof(double) * tot * (size_t)(steps + 1));
  double* ys = (double*)malloc(sizeof(double) * tot * (size_t)(steps + 1));
  double* lx = (double*)calloc(tot, sizeof(double));
  double* ly = (double*)calloc(tot, sizeof(double));
  double* ls = (double*)calloc(tot, sizeof(double));
  double* dq = (double*)calloc((size_t)s->nb * s->ncells * Q, sizeof(double));
  if (!xs || !ys || !lx || !ly || !ls || !dq) {
    free(xs); free(ys); free(lx); free(ly); free(ls); free(dq);
    return fail("adjoint: out of memory");
  }
  /* forward from feq(rho0, u0), keeping every state */
  memset(s->f, 0, sizeof(double) * tot);
  s->step = 0;
  ol_init_equilibrium(s, rho0, u0, kind_mask);
  const double wgt = 1.0 / (double)(win_end - win_begin);
  double jsum = 0.0;
  memcpy(xs, s->f, sizeof(double) * tot);
  for (int st = 1; st <= steps; ++st) {
    ol_advance(s, 1);
    memcpy(xs + tot * st, s->f, sizeof(double) * tot);
    memcpy(ys + tot * st, s->snap, sizeof(double) * tot);
    if (st > win_begin && st <= win_end) {
      double F[3];
      ol_drag(s, bbox, F);
      jsum += dir[0] * F[0] + dir[1] * F[1] + dir[2] * F[2];
    }
  }
  *J = jsum * wgt;
  /* reverse sweep */
  for (int st = steps; st >= 1; --st) {
    memset(ls, 0, sizeof(double) * tot);
    if (st > win_begin && st <= win_end) drag_seed(s, bbox, dir, wgt, lx, ls);
    memset(ly, 0, sizeof(double) * tot);
    for (int b = 0; b < s->nb; ++b) stream_block_adj(s, b, ys + tot * st, lx, ly, dq);
    for (size_t e = 0; e < tot; ++e) ly[e] += ls[e];
    exchange_adj(s, ly);
    /* collide transpose at the pre-collide state xs[st-1] */
    const double* x = xs + tot * (st - 1);
    memcpy(lx, ly, sizeof(double) * tot);
    const int n = s->n;
    for (int b = 0; b < s->nb; ++b)
      for (int ix = 0; ix < n; ++ix)
        for (int iy = 0; iy < n; ++iy)
          for (int iz = 0; iz < n; ++iz) {
            const size_t o = (size_t)b * s->nslots * Q + (size_t)core_slot(s, ix, iy, iz) * Q;
            const int k = kind_at(s, b, ix, iy, iz);
            if (k == K_BULK) {
              bgk_vjp(x + o, ly + o, s->omega, lx + o);
            } else if (k == K_BB) {
              for (int i = 0; i < Q; ++i) lx[o + i] = ly[o + OPP[i]];
            }
          }
  }
  /* x0 = feq(rho0, u0) on the masked cells */
  dJ_du0[0] = dJ_du0[1] = dJ_du0[2] = 0.0;
  {
    const int n = s->n;
    for (int b = 0; b < s->nb; ++b)
      for (int ix = 0; ix < n; ++ix)
        for (int iy = 0; iy < n; ++iy)
          for (int iz = 0; iz < n; ++iz) {
            if (!((kind_mask >> kind_at(s, b, ix, iy, iz)) & 1u)) continue;
            const size_t o = (size_t)b * s->nslots * Q + (size_t)core_slot(s, ix, iy, iz) * Q;
            for (int i = 0; i < Q; ++i) {
              double cu = 0.0;
              for (int d = 0; d < 3; ++d) cu += (double)C[i][d] * u0[d];
              for (int d = 0; d < 3; ++d)
                dJ_du0[d] += lx[o + i] * rho0 * TW(i) * (3.0 * C[i][d] + 9.0 * cu * C[i][d] - 3.0 * u0[d]);
            }
          }
  }
  for (int64_t k = 0; k < nlinks; ++k) {
    const int b = link[3 * k], c = link[3 * k + 1], i = link[3 * k + 2];
    dJ_dq[k] = (b >= 0 && b < s->nb && c >= 0 && c < s->ncells && i >= 0 && i < Q)
                   ? dq[((size_t)b * s->ncells + (size_t)c) * Q + i]
                   : 0.0;
  }
  free(xs); free(ys); free(lx); free(ly); free(ls); free(dq);
  return 0;
}

/* ---------------- state migration after a forest change ------------------
 * The reference has no migration code; its dynamic-AMR rebuild is planned
 * (doc/prd/octlb-framework.md:180,521).  The rules rescale the
 * non-equilibrium part the way the coupler does between levels:
 * prolong_scale_ = (tau - 1/4)/tau, restrict_scale_ = tau/(tau - 1/4)
 * (level_coupler.cpp:226-228), with olb's computeRhoU / computeFneq of each
 * cell (level_coupler.cpp:187).
 *   refine:  every slot of the fine block from the parent cell under it,
 *            feq(rho, u) + s_p fneq;
 *   coarsen: each coarse cell from its eight fine cells (children in octant
 *            order, each child's cells in raster order): mean rho, u = sum
 *            rho u / sum rho, s_r times the mean fneq; halo slots repeat
 *            the nearest core cell.                                        */
static void migrate_macro(const double* f, double* rho, double* u, double* fneq) {
  olb_rho_u(f, rho, u);
  const double usqr = norm_sqr(u);
  for (int i = 0; i < Q; ++i) fneq[i] = f[i] - feq2(i, *rho, u, usqr);
}

static int floor_half(int g) { return g >= 0 ? g / 2 : -((1 - g) / 2); }

int ol_migrate_block(int32_t mode, int32_t n, int32_t h, const int32_t oct[3], const double* const* views,
                     double tau, double* out) {
  if (n <= 0 || h < 0 || !views || !out) return fail("ol_migrate_block: bad arguments");
  if (mode < 0 || mode > 2) return fail("ol_migrate_block: mode out of range");
  const int E = n + 2 * h;
  const double sp = (tau - 0.25) / tau, sr = tau / (tau - 0.25);
#define VSLOT(x, y, z) ((size_t)(((x) + h) * E + (y) + h) * E + (z) + h)
  for (int x = -h; x < n + h; ++x)
    for (int y = -h; y < n + h; ++y)
      for (int z = -h; z < n + h; ++z) {
        double* o = out + VSLOT(x, y, z) * Q;
        if (mode == 0) {
          memcpy(o, views[0] + VSLOT(x, y, z) * Q, sizeof(double) * Q);
        } else if (mode == 1) {
          const int px = floor_half(oct[0] * n + x), py = floor_half(oct[1] * n + y),
                    pz = floor_half(oct[2] * n + z);
          double rho, u[3], fneq[Q];
          migrate_macro(views[0] + VSLOT(px, py, pz) * Q, &rho, u, fneq);
          const double usqr = norm_sqr(u);
          for (int i = 0; i < Q; ++i) o[i] = feq2(i, rho, u, usqr) + sp * fneq[i];
        } else {
          const int cx = x < 0 ? 0 : (x >= n ? n - 1 : x);
          const int cy = y < 0 ? 0 : (y >= n ? n - 1 : y);
          const int cz = z < 0 ? 0 : (z >= n ? n - 1 : z);
          double acc[4 + Q] = {0.0};
          for (int c = 0; c < 8; ++c) {
            const int ox = c & 1, oy = (c >> 1) & 1, oz = c >> 2;
            /* this child's fine cells under (cx,cy,cz), in its raster order */
            for (int gx = 2 * cx; gx < 2 * cx + 2; ++gx)
              for (int gy = 2 * cy; gy < 2 * cy + 2; ++gy)
                for (int gz = 2 * cz; gz < 2 * cz + 2; ++gz) {
                  if (gx / n != ox || gy / n != oy || gz / n != oz) continue;
                  double rho, u[3], fneq[Q];
                  migrate_macro(views[c] + VSLOT(gx % n, gy % n, gz % n) * Q, &rho, u, fneq);
                  acc[0] += rho;
                  for (int d = 0; d < 3; ++d) acc[1 + d] += rho * u[d];
                  for (int i = 0; i < Q; ++i) acc[4 + i] += fneq[i];
                }
          }
          const double rho = acc[0] / 8.0;
          const double u[3] = {acc[1] / acc[0], acc[2] / acc[0], acc[3] / acc[0]};
          const double usqr = norm_sqr(u);
          for (int i = 0; i < Q; ++i) o[i] = feq2(i, rho, u, usqr) + sr * (acc[4 + i] / 8.0);
        }
      }
#undef VSLOT
  return 0;
}

/* ---------------- voxelizer (part_voxelizer.cpp:71-114) --------------------
 * The reference decides its geometric tests with CGAL's exact predicates.
 * Here every test is a sign of a small determinant of the input doubles,
 * always formed exactly as a nonoverlapping expansion (Shewchuk, "Adaptive
 * Precision Floating-Point Arithmetic and Fast Robust Geometric
 * Predicates", 1997: Two-Sum, Two-Product with fma, Expansion-Sum,
 * Scale-Expansion); the sign is that of the largest nonzero component.    */
static void vx_two_sum(double a, double b, double* x, double* y) {
  *x = a + b;
  const double bv = *x - a, av = *x - bv;
  *y = (a - av) + (b - bv);
}
static void vx_fast_two_sum(double a, double b, double* x, double* y) {
  *x = a + b;
  *y = b - (*x - a);
}
/* h = e + f, h distinct from e and f; returns n + m */
static int vx_sum(const double* e, int n, const double* f, int m, double* h) {
  for (int i = 0; i < n; ++i) h[i] = e[i];
  int len = n;
  for (int j = 0; j < m; ++j) {
    double q = f[j];
    for (int i = j; i < len; ++i) {
      double s, t;
      vx_two_sum(q, h[i], &s, &t);
      h[i] = t;
      q = s;
    }
    h[len++] = q;
  }
  return len;
}
/* h = e * b; returns 2 n */
static int vx_scale(const double* e, int n, double b, double* h) {
  double q = e[0] * b;
  h[0] = fma(e[0], b, -q);
  for (int i = 1; i < n; ++i) {
    const double p1 = e[i] * b, p0 = fma(e[i], b, -p1);
    double s;
    vx_two_sum(q, p0, &s, &h[2 * i - 1]);
    vx_fast_two_sum(p1, s, &q, &h[2 * i]);
  }
  h[2 * n - 1] = q;
  return 2 * n;
}
/* h = e * f; returns 2 n m (h needs that many slots, tmp 4 n m) */
static int vx_mul(const double* e, int n, const double* f, int m, double* h, double* tmp) {
  int len = vx_scale(e, n, f[0], h);
  for (int j = 1; j < m; ++j) {
    double* s = tmp;
    double* r = tmp + 2 * n;
    const int ls = vx_scale(e, n, f[j], s);
    len = vx_sum(h, len, s, ls, r);
    memcpy(h, r, sizeof(double) * (size_t)len);
  }
  return len;
}
static int vx_sign(const double* e, int n) {
  for (int i = n - 1; i >= 0; --i)
    if (e[i] != 0.0) return e[i] > 0.0 ? 1 : -1;
  return 0;
}
static void vx_diff(double a, double b, double* e) { /* e = a - b exactly, (low, high) */
  const double x = a - b, bv = a - x, av = x + bv;
  e[1] = x;
  e[0] = (a - av) + (bv - b);
}
/* sign of (p1u - q1u)(p2v - q2v) - (p1v - q1v)(p2u - q2u) */
static int vx_det2(double p1u, double p1v, double q1u, double q1v, double p2u, double p2v, double q2u,
                   double q2v) {
  double a[2], b[2], c[2], d[2], ab[8], cd[8], s[16], tmp[16];
  vx_diff(p1u, q1u, a);
  vx_diff(p2v, q2v, b);
  vx_diff(p1v, q1v, c);
  vx_diff(p2u, q2u, d);
  vx_mul(a, 2, b, 2, ab, tmp);
  vx_mul(c, 2, d, 2, cd, tmp);
  for (int i = 0; i < 8; ++i) cd[i] = -cd[i];
  return vx_sign(s, vx_sum(ab, 8, cd, 8, s));
}
/* sign of det[A - D; B - D; C - D] by cofactors along the first row */
static int vx_det3(const double* A, const double* B, const double* C, const double* D) {
  double r[3][3][2];
  for (int k = 0; k < 3; ++k) {
    vx_diff(A[k], D[k], r[0][k]);
    vx_diff(B[k], D[k], r[1][k]);
    vx_diff(C[k], D[k], r[2][k]);
  }
  double total[192], tmp[192];
  int len = 0;
  for (int col = 0; col < 3; ++col) {
    const int c1 = (col + 1) % 3, c2 = (col + 2) % 3; /* cyclic cofactor: + for every column */
    double p[8], q[8], minor[16], term[64], t2[32];
    vx_mul(r[1][c1], 2, r[2][c2], 2, p, t2);
    vx_mul(r[1][c2], 2, r[2][c1], 2, q, t2);
    for (int i = 0; i < 8; ++i) q[i] = -q[i];
    const int lm = vx_sum(p, 8, q, 8, minor);
    double big[128];
    const int lt = vx_mul(minor, lm, r[0][col], 2, term, big);
    len = vx_sum(total, len, term, lt, tmp);
    memcpy(total, tmp, sizeof(double) * (size_t)len);
  }
  return vx_sign(total, len);
}

typedef struct {
  double v[3][3];
  double lo[3], hi[3];
  int ns[3]; /* signs of the normal (B - A) x (C - A) */
  int degenerate;
} vx_tri;

static void vx_prep(const double* t9, vx_tri* t) {
  for (int k = 0; k < 3; ++k)
    for (int d = 0; d < 3; ++d) t->v[k][d] = t9[k * 3 + d];
  for (int d = 0; d < 3; ++d) {
    t->lo[d] = fmin(fmin(t->v[0][d], t->v[1][d]), t->v[2][d]);
    t->hi[d] = fmax(fmax(t->v[0][d], t->v[1][d]), t->v[2][d]);
  }
  for (int a = 0; a < 3; ++a) {
    const int u = (a + 1) % 3, w = (a + 2) % 3;
    t->ns[a] = vx_det2(t->v[1][u], t->v[1][w], t->v[0][u], t->v[0][w], t->v[2][u], t->v[2][w], t->v[0][u],
                       t->v[0][w]);
  }
  t->degenerate = t->ns[0] == 0 && t->ns[1] == 0 && t->ns[2] == 0;
}

/* closed ray p + t e_a (t >= 0) meets the closed non-degenerate triangle
 * (CGAL::do_intersect(Ray_3, Triangle_3)) */
static int vx_ray_hit(const double* p, int a, const vx_tri* t) {
  const double *A = t->v[0], *B = t->v[1], *C = t->v[2];
  const int u = (a + 1) % 3, w = (a + 2) % 3;
  const int na = t->ns[a];
  const int s = vx_det3(B, C, p, A); /* normal . (p - A) */
  if (na != 0) {
    if (s != 0 && s == na) return 0; /* the plane lies behind the origin */
    const int e1 = vx_det2(B[u], B[w], A[u], A[w], p[u], p[w], A[u], A[w]);
    const int e2 = vx_det2(C[u], C[w], B[u], B[w], p[u], p[w], B[u], B[w]);
    const int e3 = vx_det2(A[u], A[w], C[u], C[w], p[u], p[w], C[u], C[w]);
    return e1 != -na && e2 != -na && e3 != -na;
  }
  if (s != 0) return 0;
  /* ray in the triangle's plane: the coordinate plane (a, q) it projects onto */
  const int q = t->ns[w] != 0 ? u : w;
  for (int k = 0; k < 3; ++k) {
    const double *X = t->v[k], *Y = t->v[(k + 1) % 3];
    if ((X[q] > p[q] && Y[q] > p[q]) || (X[q] < p[q] && Y[q] < p[q])) continue;
    if (X[q] == p[q] && Y[q] == p[q]) {
      if (fmax(X[a], Y[a]) >= p[a]) return 1;
      continue;
    }
    const int num = vx_det2(X[a], X[q], p[a], p[q], Y[a], Y[q], X[a], X[q]);
    const int dq = Y[q] > X[q] ? 1 : -1;
    if (num == 0 || num == dq) return 1;
  }
  return 0;
}

/* the closed triangle (degenerate ones too) meets the closed box: no
 * separating axis among the box axes, the normal and the nine e_k x edge */
static int vx_box_hit(const double* lo, const double* hi, const vx_tri* t) {
  for (int k = 0; k < 3; ++k)
    if (t->hi[k] < lo[k] || t->lo[k] > hi[k]) return 0;
  const double *A = t->v[0], *B = t->v[1], *C = t->v[2];
  if (!t->degenerate) {
    double cmin[3], cmax[3];
    for (int k = 0; k < 3; ++k) {
      cmin[k] = t->ns[k] > 0 ? lo[k] : hi[k];
      cmax[k] = t->ns[k] > 0 ? hi[k] : lo[k];
    }
    if (vx_det3(B, C, cmin, A) > 0 || vx_det3(B, C, cmax, A) < 0) return 0;
  }
  for (int e = 0; e < 3; ++e) {
    const double *X = t->v[e], *Y = t->v[(e + 1) % 3], *Z = t->v[(e + 2) % 3];
    for (int k = 0; k < 3; ++k) {
      const int i1 = (k + 1) % 3, i2 = (k + 2) % 3;
      if (Y[i1] == X[i1] && Y[i2] == X[i2]) continue;
      const int l1 = Y[i2] > X[i2] ? -1 : (Y[i2] < X[i2] ? 1 : 0);
      const int l2 = Y[i1] > X[i1] ? 1 : (Y[i1] < X[i1] ? -1 : 0);
      const double mn1 = l1 > 0 ? lo[i1] : hi[i1], mx1 = l1 > 0 ? hi[i1] : lo[i1];
      const double mn2 = l2 > 0 ? lo[i2] : hi[i2], mx2 = l2 > 0 ? hi[i2] : lo[i2];
      if (vx_det2(Y[i1], Y[i2], X[i1], X[i2], X[i1], X[i2], mn1, mn2) < 0 &&
          vx_det2(Y[i1], Y[i2], X[i1], X[i2], Z[i1], Z[i2], mn1, mn2) < 0)
        return 0;
      if (vx_det2(Y[i1], Y[i2], X[i1], X[i2], X[i1], X[i2], mx1, mx2) > 0 &&
          vx_det2(Y[i1], Y[i2], X[i1], X[i2], Z[i1], Z[i2], mx1, mx2) > 0)
        return 0;
    }
  }
  return 1;
}

/* RayParityInsideAabb: odd hit counts on at least two of the three axes */
static int vx_inside(const double* p, const vx_tri* tris, int64_t count) {
  int votes = 0;
  for (int a = 0; a < 3; ++a) {
    const int u = (a + 1) % 3, w = (a + 2) % 3;
    int odd = 0;
    for (int64_t i = 0; i < count; ++i) {
      const vx_tri* t = &tris[i];
      if (t->degenerate || t->hi[a] < p[a] || t->lo[u] > p[u] || t->hi[u] < p[u] || t->lo[w] > p[w] ||
          t->hi[w] < p[w])
        continue;
      odd ^= vx_ray_hit(p, a, t);
    }
    votes += odd;
  }
  return votes >= 2;
}

int ol_voxelize_part(int32_t num_blocks, int32_t n, const double* bounds, int64_t num_triangles,
                     const double* triangles, int64_t num_inner, const double* inner, int32_t role,
                     uint8_t* material) {
  if (num_blocks <= 0 || n <= 0 || !bounds || !material) return fail("ol_voxelize_part: bad forest arguments");
  if (role != 0 && role != 1) return fail("ol_voxelize_part: unknown part role");
  if (num_triangles <= 0 || !triangles) return fail("geometry part has no triangles");
  const int channel = role == 0;
  if (channel && (num_inner <= 0 || !inner)) return fail("kInternalChannel requires inner_cavity_soup");
  if (!channel) num_inner = 0;
  vx_tri* outer = (vx_tri*)malloc(sizeof(vx_tri) * (size_t)num_triangles);
  vx_tri* lumen = (vx_tri*)malloc(sizeof(vx_tri) * (size_t)(num_inner > 0 ? num_inner : 1));
  if (!outer || !lumen) {
    free(outer);
    free(lumen);
    return fail("ol_voxelize_part: out of memory");
  }
  for (int64_t i = 0; i < num_triangles; ++i) vx_prep(triangles + i * 9, &outer[i]);
  for (int64_t i = 0; i < num_inner; ++i) vx_prep(inner + i * 9, &lumen[i]);
  if (channel) {
    /* validate_geometry_part: outside_sample_fraction(5) >= 0.05 over the
     * outer soup's bounding box (cgal_surface_mesh.cpp) */
    double bb[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int64_t i = 0; i < num_triangles * 3; ++i)
      for (int d = 0; d < 3; ++d) {
        bb[d] = fmin(bb[d], triangles[i * 3 + d]);
        bb[3 + d] = fmax(bb[3 + d], triangles[i * 3 + d]);
      }
    const int s = 5;
    int outside = 0;
    for (int iz = 0; iz < s; ++iz)
      for (int iy = 0; iy < s; ++iy)
        for (int ix = 0; ix < s; ++ix) {
          const double p[3] = {bb[0] + (double)ix * ((bb[3] - bb[0]) / (double)(s - 1)),
                               bb[1] + (double)iy * ((bb[4] - bb[1]) / (double)(s - 1)),
                               bb[2] + (double)iz * ((bb[5] - bb[2]) / (double)(s - 1))};
          outside += !vx_inside(p, outer, num_triangles);
        }
    if ((double)outside / (double)(s * s * s) < 0.05) {
      free(outer);
      free(lumen);
      return fail("kInternalChannel mesh has insufficient exterior/lumen volume (solid-wall channel expected)");
    }
  }
  const int64_t cells = (int64_t)n * n * n;
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t b = 0; b < num_blocks; ++b) {
    const double* ob = bounds + (size_t)b * 6;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          /* CellBounds (part_voxelizer.cpp:33-47) and the centre */
          const int ijk[3] = {i, j, k};
          double lo[3], hi[3], c[3];
          for (int d = 0; d < 3; ++d) {
            const double dx = (ob[3 + d] - ob[d]) / (double)n;
            lo[d] = ob[d] + (double)ijk[d] * dx;
            hi[d] = lo[d] + dx;
            c[d] = 0.5 * (lo[d] + hi[d]);
          }
          int hit = 0;
          for (int64_t t = 0; t < num_triangles && !hit; ++t) hit = vx_box_hit(lo, hi, &outer[t]);
          for (int64_t t = 0; t < num_inner && !hit; ++t) hit = vx_box_hit(lo, hi, &lumen[t]);
          uint8_t m;
          if (hit) {
            m = 2; /* kBoundary */
          } else if (!channel) {
            m = vx_inside(c, outer, num_triangles) ? 1 : 0; /* MapExternal */
          } else {
            const int in_lumen = vx_inside(c, lumen, num_inner);
            const int in_wall = !in_lumen && vx_inside(c, outer, num_triangles);
            m = in_lumen ? 0 : (in_wall ? 1 : 0); /* MapInternalChannel */
          }
          material[(size_t)b * (size_t)cells + ((size_t)k * n + j) * n + i] = m;
        }
  }
  free(outer);
  free(lumen);
  return 0;
}
