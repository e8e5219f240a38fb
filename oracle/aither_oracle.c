/*
 * Oracle: FP64 CPU restatement of aither's per-nonlinear-iteration hot path.
 *
 * TEST INFRASTRUCTURE ONLY -- see aither_oracle.h.  Only tests/, the smoke
 * check and bench.py's cpu_baseline leg load this code, as the checker.
 *
 * Each function cites the reference file:line whose algorithm it restates.
 * Operation order follows the reference expression by expression; the file is
 * compiled with -ffp-contract=off (no fused multiply-add).
 *
 * Data layout mirrors the reference's multiArray3d: ghost-padded arrays with i
 * fastest, AoS over equations (include/multiArray3d.hpp:104-113).
 */
#include "aither_oracle.h"

#include <ctype.h>
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Parallel loops (the OpenMP build, liboracle_omp.so): every loop marked
 * ORC_PARALLEL writes disjoint cells in each iteration of its outer loop
 * and keeps, for each cell, the order in which the serial build accumulates
 * into it; sums over cells are taken serially in the serial order.  The
 * OpenMP build therefore gives the serial build's results bit for bit. */
#ifdef _OPENMP
#define ORC_PARALLEL _Pragma("omp parallel for schedule(static)")
#else
#define ORC_PARALLEL
#endif

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

#define EPS 1.0e-30 /* include/macros.hpp.in:21 */
#define NEQ 5       /* rho, u, v, w, p (single species, no turbulence) */
#define MAXBC 64

typedef struct { double x, y, z; } v3;
typedef struct { v3 u; double m; } uv3; /* unitVec3dMag: include/vector3d.hpp:131 */

static v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 vscl(double s, v3 a) { return V(a.x * s, a.y * s, a.z * s); }
static v3 vdiv(v3 a, double s) { return V(a.x / s, a.y / s, a.z / s); }
/* std::inner_product with T(0) init: include/vector3d.hpp:303-305 */
static double vdot(v3 a, v3 b) {
  double s = 0.0;
  s = s + a.x * b.x;
  s = s + a.y * b.y;
  s = s + a.z * b.z;
  return s;
}
static double vmag(v3 a) { return sqrt(vdot(a, a)); }
/* include/vector3d.hpp:332-341 */
static v3 vcross(v3 a, v3 b) {
  return V(a.y * b.z - a.z * b.y, -1.0 * (a.x * b.z - a.z * b.x),
           a.x * b.y - a.y * b.x);
}
static uv3 make_uv3(v3 a) {
  uv3 r;
  r.u = vdiv(a, vmag(a)); /* Normalize: include/vector3d.hpp:375-377 */
  r.m = vmag(a);
  return r;
}
static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* ------------------------------------------------------------------------ */
/* boundary surfaces and states                                             */

typedef struct {
  char type[32];
  int d[7]; /* imin imax jmin jmax kmin kmax tag */
} surf_t;

typedef struct {
  char kind[32];
  int tag;
  double p, rho;
  v3 vel;
  double p0, t0;
  v3 dir;
  int nonreflecting;
  int has_mf;
  int isothermal; /* viscousWall: wall temperature given */
  double twall;
  int heatflux;   /* viscousWall: a nonzero heat flux given (include/inputStates.hpp:379-381) */
  double qwall;
  int startTag, endTag; /* periodic: the two surfaces' tags and the start */
  v3 trans;             /* surface's translation onto the end surface */
} bcstate_t;

/* src/boundaryConditions.cpp:2424-2450 */
static int surf_type(const surf_t *s) {
  if (s->d[0] == s->d[1]) return (s->d[1] == 0) ? 1 : 2;
  if (s->d[2] == s->d[3]) return (s->d[3] == 0) ? 3 : 4;
  if (s->d[4] == s->d[5]) return (s->d[5] == 0) ? 5 : 6;
  return 0;
}

/* boundarySurface::operator< : src/boundaryConditions.cpp:91-105 */
static int surf_cmp(const void *pa, const void *pb) {
  const surf_t *a = (const surf_t *)pa, *b = (const surf_t *)pb;
  int ta = surf_type(a), tb = surf_type(b);
  if (ta != tb) return ta < tb ? -1 : 1;
  for (int q = 0; q < 7; ++q)
    if (a->d[q] != b->d[q]) return a->d[q] < b->d[q] ? -1 : 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* block storage                                                            */

typedef struct {
  int ni, nj, nk, g;
  int NI, NJ, NK;
  double *vol;
  v3 *ctr;
  uv3 *fa[3];
  v3 *fc[3];
  double *cw[3];
  double *W;      /* padded, NEQ per cell */
  double *R;      /* physical, NEQ per cell */
  double *sr;     /* physical */
  double *dt;     /* physical */
  double *A, *Ainv;
  double *Ab, *Abi; /* blusgs / bdplur: main diagonal and its inverse, 25 per physical cell, row-major */
  double *consN, *consNm1;
  double *x;      /* padded, NEQ per cell */
  double *T, *mu; /* padded: temperature, viscosity (viscous cases) */
  double *vg;     /* padded: cell velocity gradient, 9 per cell, row-major d v_c / d x_r */
  int *order;     /* 3 per physical cell */
  double *wd;     /* padded: wall distance of heat-flux wall cells, found once (NAN until then) */
  int nsurf;
  surf_t *surf;
  int ictag;
  int nconn;
  struct conn_s *conn;
  unsigned char *isconn[6]; /* per boundary face of each surface: 1 on a connection */
} blk_t;

/* one side of an interblock connection, as seen from the block that owns
 * the ghost cells; patch-local cell (p1, p2) along this block's d1, d2 maps
 * to the partner's patch-local cell (q1, q2):
 *   c = swap ? (p2, p1) : (p1, p2);  q1 = rev1 ? pn1-1-c1 : c1;  q2 likewise */
typedef struct conn_s {
  int s, dir, r3, r1s, n1, r2s, n2;
  int pb, ps, pdir, pr3, pr1s, pn1, pr2s, pn2;
  int swap, rev1, rev2;
  int periodic; /* periodic connection: no geometry swap */
  v3 shift;     /* added to this patch's nodes before matching */
} conn_t;

struct orc_case {
  char dir[1024];
  char grid[512];
  int iterations, nonlinear, sweeps;
  double rRef, tRef, lRef, aRef;
  double dtIn;
  double cflStart, cflMax, cflStep, cfl;
  double theta, zeta, relax, dualCFL;
  double kappa;
  int recon;   /* 0 constant, 1 MUSCL, 2 WENO5, 3 WENO-Z */
  int limiter; /* 0 none, 1 vanAlbada, 2 minmod */
  int flux;    /* 0 roe, 1 ausm */
  int timeint; /* 0 explicitEuler 1 implicitEuler 2 crankNicholson 3 bdf2 */
  int solver;  /* 0 lusgs, 1 dplur */
  int block;   /* 5x5 block main diagonal and off-diagonals (blusgs, bdplur) */
  int roeJac;  /* inviscidFluxJacobian approximateRoe: RoeOffDiagonal */
  int implicit, multilevel, storeN, initMatrix;
  orc_gas gas;
  double gamma;
  double fluidN, fluidM;
  int nic;
  struct {
    int tag, fromFile;
    char file[512];
    double rho, p;
    v3 vel;
  } ic[16];
  int nbc;
  bcstate_t bcs[MAXBC];
  int nb;
  blk_t *blk;
  double totalCells;
  int matres_on;
  double l2first[NEQ];
  /* viscous (laminar Navier-Stokes) */
  int viscous;
  int viscRecon4; /* viscousFaceReconstruction: 0 central, 1 centralFourth */
  double suth[4]; /* Sutherland C1, S for viscosity, then conductivity */
  double muRef, kNonDim, vscale; /* reference viscosity, conductivity scale, mu/(rho a L) */
  double viscCoeff;
};

static long cidx(const blk_t *b, int i, int j, int k) {
  return (long)(i + b->g) + (long)(j + b->g) * b->NI +
         (long)(k + b->g) * b->NI * b->NJ;
}
static long fidx(const blk_t *b, int d, int i, int j, int k) {
  int FI = b->NI + (d == 0), FJ = b->NJ + (d == 1);
  return (long)(i + b->g) + (long)(j + b->g) * FI + (long)(k + b->g) * FI * FJ;
}
static long pidx(const blk_t *b, int i, int j, int k) {
  return (long)i + (long)j * b->ni + (long)k * b->ni * b->nj;
}
static long ncells_padded(const blk_t *b) { return (long)b->NI * b->NJ * b->NK; }
static long nfaces_padded(const blk_t *b, int d) {
  return (long)(b->NI + (d == 0)) * (b->NJ + (d == 1)) * (b->NK + (d == 2));
}
static long ncells(const blk_t *b) { return (long)b->ni * b->nj * b->nk; }

/* index remapping (dir, d1, d2, d3) -> (i, j, k):
 * include/multiArray3d.hpp:231-243 (dir i: d1=j d2=k; j: d1=k d2=i; k: d1=i d2=j) */
static void map3(int dir, int d3, int d1, int d2, int *i, int *j, int *k) {
  if (dir == 0) { *i = d3; *j = d1; *k = d2; }
  else if (dir == 1) { *i = d2; *j = d3; *k = d1; }
  else { *i = d1; *j = d2; *k = d3; }
}
static int ndir(const blk_t *b, int d) { return d == 0 ? b->ni : (d == 1 ? b->nj : b->nk); }

/* ------------------------------------------------------------------------ */
/* thermodynamics: calorically perfect single species                       */
/* eos: src/eos.cpp:40-115, thermo: include/thermodynamic.hpp:58-121,       */
/* src/thermodynamic.cpp:63-113, helpers include/arrayView.hpp:384-444       */

static double cps(const orc_gas *g) { return g->R * (g->n + 1.0); }
static double cvs(const orc_gas *g) { return g->R * g->n; }
double orc_gamma(const orc_gas *g) {
  /* Cp(t,mf)/Cv(t,mf) with mf = {1.0}: 0.0 + 1.0*Cp_s */
  double cp = 0.0 + 1.0 * cps(g);
  double cv = 0.0 + 1.0 * cvs(g);
  return cp / cv;
}
static double temperature(const orc_gas *g, const double *W) {
  double rhoR = 0.0;
  rhoR += W[0] * g->R;
  return W[4] / rhoR;
}
double orc_sos(const orc_gas *g, const double *W) {
  return sqrt(orc_gamma(g) * W[4] / W[0]);
}
static double velmag(const double *W) { return vmag(V(W[1], W[2], W[3])); }
double orc_enthalpy(const orc_gas *g, const double *W) {
  double t = temperature(g, W);
  double vel = velmag(W);
  double h = 0.0;
  h += 1.0 * (0.0 + cps(g) * t); /* hf = 0 */
  return h + 0.5 * vel * vel;
}
static double energy(const orc_gas *g, const double *W) {
  double t = temperature(g, W);
  double e = 0.0;
  e += 1.0 * (0.0 + cvs(g) * t);
  double vel = velmag(W);
  return e + 0.5 * vel * vel;
}
/* PrimToCons: include/primitive.hpp:151-169 */
void orc_prim_to_cons(const orc_gas *g, const double *W, double *U) {
  double rho = W[0];
  U[0] = W[0];
  U[1] = rho * W[1];
  U[2] = rho * W[2];
  U[3] = rho * W[3];
  U[4] = rho * energy(g, W);
}
/* primitive(cons, phys): include/primitive.hpp:129-148; PressFromEnergy
 * src/eos.cpp:40-52; TemperatureFromSpecEnergy src/thermodynamic.cpp:107-113 */
void orc_cons_to_prim(const orc_gas *g, const double *U, double *W) {
  double rho = 0.0 + U[0];
  W[0] = U[0];
  W[1] = U[1] / rho;
  W[2] = U[2] / rho;
  W[3] = U[3] / rho;
  double en = U[4] / rho;
  double vel = velmag(W);
  double spec = en - 0.5 * vel * vel;
  double rhoSum = 0.0 + W[0];
  double mf = W[0] / rhoSum;
  double hf = 0.0 + 0.0 * mf;
  double cv = 0.0 + mf * cvs(g);
  double t = (spec - hf) / cv;
  double p = 0.0;
  p += W[0] * g->R * t;
  W[4] = p;
}
/* UpdatePrimWithCons: include/primitive.hpp:206-231 (single species: the
 * mass-fraction clip and renormalisation leave rho unchanged) */
static void update_prim(const orc_gas *g, const double *W, const double *du,
                        double *out) {
  double U[NEQ];
  orc_prim_to_cons(g, W, U);
  for (int e = 0; e < NEQ; ++e) U[e] = U[e] + du[e];
  double rho = 0.0 + U[0];
  double mf = U[0] / rho;
  mf = dmax(mf, 0.0);
  double total = 0.0 + mf;
  mf /= total;
  U[0] = rho * mf;
  orc_cons_to_prim(g, U, out);
}

/* ------------------------------------------------------------------------ */
/* fluxes                                                                   */

/* inviscidFlux::ConstructFromPrim: include/inviscidFlux.hpp:128-159 */
void orc_phys_flux(const orc_gas *g, const double *W, const double *n,
                   double *F) {
  v3 vel = V(W[1], W[2], W[3]);
  v3 nn = V(n[0], n[1], n[2]);
  double vn = vdot(vel, nn);
  F[0] = W[0] * vn;
  double rho = 0.0 + W[0];
  F[1] = rho * vn * vel.x + W[4] * nn.x;
  F[2] = rho * vn * vel.y + W[4] * nn.y;
  F[3] = rho * vn * vel.z + W[4] * nn.z;
  F[4] = rho * vn * orc_enthalpy(g, W);
}

/* RoeFlux: include/inviscidFlux.hpp:259-382, src/inviscidFlux.cpp:26-32;
 * RoeAveragedState: include/primitive.hpp:245-280 */
void orc_roe_flux(const orc_gas *g, const double *L, const double *Rt,
                  const double *nv, double *F) {
  v3 n = V(nv[0], nv[1], nv[2]);
  double roe[NEQ];
  double denRatio = sqrt(Rt[0] / L[0]);
  roe[0] = L[0] * denRatio;
  roe[1] = (L[1] + denRatio * Rt[1]) / (1.0 + denRatio);
  roe[2] = (L[2] + denRatio * Rt[2]) / (1.0 + denRatio);
  roe[3] = (L[3] + denRatio * Rt[3]) / (1.0 + denRatio);
  roe[4] = (L[4] + denRatio * Rt[4]) / (1.0 + denRatio);

  double hR = orc_enthalpy(g, roe);
  double aR = orc_sos(g, roe);
  double rhoR = 0.0 + roe[0];
  v3 vroe = V(roe[1], roe[2], roe[3]);
  double velNormR = vdot(vroe, n);
  double mfR = roe[0] / rhoR;
  double delta[NEQ];
  for (int e = 0; e < NEQ; ++e) delta[e] = Rt[e] - L[e];
  v3 dvel = V(delta[1], delta[2], delta[3]);
  double normVelDiff = vdot(dvel, n);

  double diss[NEQ] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const double entropyFix = 0.1;

  /* left moving acoustic wave */
  double waveSpeed = fabs(velNormR - aR);
  if (waveSpeed < entropyFix)
    waveSpeed = 0.5 * (waveSpeed * waveSpeed / entropyFix + entropyFix);
  double waveStrength =
      (delta[4] - rhoR * aR * normVelDiff) / (2.0 * aR * aR);
  double wss = waveSpeed * waveStrength;
  diss[0] += wss * mfR;
  diss[1] += wss * (vroe.x - aR * n.x);
  diss[2] += wss * (vroe.y - aR * n.y);
  diss[3] += wss * (vroe.z - aR * n.z);
  diss[4] += wss * (hR - aR * velNormR);

  /* entropy wave */
  waveSpeed = fabs(velNormR);
  waveStrength = -delta[4] / (aR * aR);
  wss = waveSpeed * waveStrength;
  diss[0] += wss * mfR + waveSpeed * delta[0];
  double drho = 0.0 + delta[0];
  waveStrength = drho - delta[4] / (aR * aR);
  wss = waveSpeed * waveStrength;
  diss[1] += wss * vroe.x;
  diss[2] += wss * vroe.y;
  diss[3] += wss * vroe.z;
  diss[4] += wss * 0.5 * vdot(vroe, vroe);

  /* shear wave */
  waveStrength = rhoR;
  wss = waveSpeed * waveStrength;
  diss[1] += wss * (dvel.x - normVelDiff * n.x);
  diss[2] += wss * (dvel.y - normVelDiff * n.y);
  diss[3] += wss * (dvel.z - normVelDiff * n.z);
  diss[4] += wss * (vdot(vroe, dvel) - velNormR * normVelDiff);

  /* right moving acoustic wave */
  waveSpeed = fabs(velNormR + aR);
  if (waveSpeed < entropyFix)
    waveSpeed = 0.5 * (waveSpeed * waveSpeed / entropyFix + entropyFix);
  waveStrength = (delta[4] + rhoR * aR * normVelDiff) / (2.0 * aR * aR);
  wss = waveSpeed * waveStrength;
  diss[0] += wss * mfR;
  diss[1] += wss * (vroe.x + aR * n.x);
  diss[2] += wss * (vroe.y + aR * n.y);
  diss[3] += wss * (vroe.z + aR * n.z);
  diss[4] += wss * (hR + aR * velNormR);

  double FL[NEQ], FR[NEQ];
  orc_phys_flux(g, L, nv, FL);
  orc_phys_flux(g, Rt, nv, FR);
  for (int e = 0; e < NEQ; ++e) {
    double t = FL[e] + (FR[e] - diss[e]);
    F[e] = t * 0.5;
  }
}

static double sgn(double x) { return (x > 0.0) - (x < 0.0); } /* Sign: include/utility.hpp */

/* AUSMFlux (AUSMPW+): include/inviscidFlux.hpp:161-209, 395-481 */
void orc_ausm_flux(const orc_gas *g, const double *L, const double *Rt,
                   const double *nv, double *F) {
  v3 a = V(nv[0], nv[1], nv[2]);
  double velNormL = vdot(V(L[1], L[2], L[3]), a);
  double velNormR = vdot(V(Rt[1], Rt[2], Rt[3]), a);
  double sosL = orc_sos(g, L);
  double sosR = orc_sos(g, Rt);
  double sosStar = sqrt(sosL * sosR);
  double vel = 0.5 * (velNormL + velNormR);
  double sos = sosStar;
  if (vel < 0.0) {
    sos = sosStar * sosStar / dmax(velNormR, sosStar);
  } else if (vel > 0.0) {
    sos = sosStar * sosStar / dmax(velNormL, sosStar);
  }
  double ml = velNormL / sos;
  double mr = velNormR / sos;
  double mPlusL = fabs(ml) <= 1.0 ? 0.25 * pow(ml + 1.0, 2.0)
                                  : 0.5 * (ml + fabs(ml));
  double mMinusR = fabs(mr) <= 1.0 ? -0.25 * pow(mr - 1.0, 2.0)
                                   : 0.5 * (mr - fabs(mr));
  double pPlus = fabs(ml) <= 1.0 ? 0.25 * pow(ml + 1.0, 2.0) * (2.0 - ml)
                                 : 0.5 * (1.0 + sgn(ml));
  double pMinus = fabs(mr) <= 1.0 ? 0.25 * pow(mr - 1.0, 2.0) * (2.0 + mr)
                                  : 0.5 * (1.0 - sgn(mr));
  double ps = pPlus * L[4] + pMinus * Rt[4];
  double w = 1.0 - pow(dmin(L[4] / Rt[4], Rt[4] / L[4]), 3.0);
  double fl = fabs(ml) < 1.0 ? L[4] / ps - 1.0 : 0.0;
  double fr = fabs(mr) < 1.0 ? Rt[4] / ps - 1.0 : 0.0;
  double mavg = mPlusL + mMinusR;
  double mPlusLBar = mavg >= 0.0
                         ? mPlusL + mMinusR * ((1.0 - w) * (1.0 + fr) - fl)
                         : mPlusL * w * (1.0 + fl);
  double mMinusRBar = mavg >= 0.0
                          ? mMinusR * w * (1.0 + fr)
                          : mMinusR + mPlusL * ((1.0 - w) * (1.0 + fl) - fr);
  /* inviscidFlux::AUSMFlux member: include/inviscidFlux.hpp:161-209 */
  double vl = mPlusLBar * sos;
  F[0] = L[0] * vl;
  double rhoL = 0.0 + L[0];
  F[1] = rhoL * vl * L[1] + pPlus * L[4] * a.x;
  F[2] = rhoL * vl * L[2] + pPlus * L[4] * a.y;
  F[3] = rhoL * vl * L[3] + pPlus * L[4] * a.z;
  F[4] = rhoL * vl * orc_enthalpy(g, L);
  double vr = mMinusRBar * sos;
  F[0] += Rt[0] * vr;
  double rhoR = 0.0 + Rt[0];
  F[1] += rhoR * vr * Rt[1] + pMinus * Rt[4] * a.x;
  F[2] += rhoR * vr * Rt[2] + pMinus * Rt[4] * a.y;
  F[3] += rhoR * vr * Rt[3] + pMinus * Rt[4] * a.z;
  F[4] += rhoR * vr * orc_enthalpy(g, Rt);
}

/* ------------------------------------------------------------------------ */
/* reconstruction: include/reconstruction.hpp:109-154, src/limiter.cpp:24-53 */

static double lim_vanalbada(double r) {
  double r2 = r * r;
  double l = (r + r2) / (1.0 + r2);
  return dmax(0.0, l);
}
static double lim_minmod(double r) { return dmax(0.0, dmin(1.0, r)); }

void orc_muscl(const double *u2, const double *u1, const double *d1,
               double kappa, int limiter, double uw2, double uw, double dw,
               double *out) {
  double dPlus = (uw + uw) / (uw + dw);
  double dMinus = (uw + uw) / (uw + uw2);
  for (int e = 0; e < NEQ; ++e) {
    double r = (EPS + (d1[e] - u1[e]) * dPlus) / (EPS + (u1[e] - u2[e]) * dMinus);
    double lim, inv;
    if (limiter == 1) {
      lim = lim_vanalbada(r);
      inv = lim_vanalbada(1.0 / r);
    } else if (limiter == 2) {
      lim = lim_minmod(r);
      inv = lim_minmod(1.0 / r);
    } else {
      lim = 1.0;
      inv = 1.0;
    }
    double a = 0.25 * ((u1[e] - u2[e]) * dMinus);
    double b = (1.0 - kappa) * lim + (1.0 + kappa) * r * inv;
    out[e] = u1[e] + a * b;
  }
}

/* ------------------------------------------------------------------------ */
/* WENO5 / WENO-Z on non-uniform widths                                     */

/* StencilWidth: include/utility.hpp:103-114 (signed sum of widths in
 * [start, end), accumulated from the lower index) */
static double stencil_width(const double *cw, int start, int end) {
  double w = 0.0;
  if (end > start) {
    for (int q = start; q < end; ++q) w += cw[q];
  } else if (start > end) {
    double s = 0.0;
    for (int q = end; q < start; ++q) s += cw[q];
    w = -1.0 * s;
  }
  return w;
}

/* LagrangeCoeff: src/utility.cpp:449-483 (Shu, ICASE 97-65, eq. 2.20) */
void orc_lagrange_coeff(const double *cw, int degree, int rr, int ii, double *coeffs) {
  for (int jj = 0; jj <= degree; ++jj) {
    coeffs[jj] = 0.0;
    for (int mm = jj + 1; mm <= degree + 1; ++mm) {
      double numer = 0.0, denom = 1.0;
      for (int ll = 0; ll <= degree + 1; ++ll) {
        if (ll == mm) continue;
        double prod = 1.0;
        for (int qq = 0; qq <= degree + 1; ++qq)
          if (qq != mm && qq != ll) prod *= stencil_width(cw, ii - rr + qq, ii + 1);
        numer += prod;
        denom *= stencil_width(cw, ii - rr + ll, ii - rr + mm);
      }
      coeffs[jj] += numer / denom;
    }
    coeffs[jj] *= cw[ii - rr + jj];
  }
}

/* Derivative2nd: include/utility.hpp:116-122 */
static double deriv2nd(double x0, double x1, double x2, double y0, double y1, double y2) {
  double fwd = (y2 - y1) / (0.5 * (x2 + x1));
  double bck = (y1 - y0) / (0.5 * (x1 + x0));
  return (fwd - bck) / (0.25 * (x2 + x0) + 0.5 * x1);
}

/* BetaIntegral: include/reconstruction.hpp:157-181 */
static double beta_int_at(double d1, double d2, double dx, double x) {
  return (d1 * d1 * x + d1 * d2 * x * x + d2 * d2 * pow(x, 3.0) / 3.0) * dx +
         d2 * d2 * x * pow(dx, 3.0);
}
static double beta_int(double d1, double d2, double dx, double xl, double xh) {
  return beta_int_at(d1, d2, dx, xh) - beta_int_at(d1, d2, dx, xl);
}

/* Beta0/Beta1/Beta2: include/reconstruction.hpp:183-240; which = 0, 1, 2 */
double orc_weno_beta(int which, double x0, double x1, double x2, double y0, double y1,
                     double y2) {
  double d2 = deriv2nd(x0, x1, x2, y0, y1, y2);
  if (which == 0) {
    double d1 = (y2 - y1) / (0.5 * (x2 + x1)) + 0.5 * x2 * d2;
    return beta_int(d1, d2, x2, -0.5 * x2, 0.5 * x2);
  } else if (which == 1) {
    double d1 = (y2 - y1) / (0.5 * (x2 + x1)) - 0.5 * x1 * d2;
    return beta_int(d1, d2, x1, -0.5 * x1, 0.5 * x1);
  }
  double d1 = (y1 - y0) / (0.5 * (x1 + x0)) - 0.5 * x0 * d2;
  return beta_int(d1, d2, x0, -0.5 * x0, 0.5 * x0);
}

/* FaceReconWENO: include/reconstruction.hpp:243-310.  u3, u2, u1 are the
 * upwind cells (u1 adjacent to the face), d1, d2 the downwind cells; w[5]
 * their widths in the same order. */
void orc_weno(const double *u3, const double *u2, const double *u1, const double *d1,
              const double *d2, const double *w, int wenoZ, double *out) {
  double c0[3], c1[3], c2[3], full[5];
  orc_lagrange_coeff(w, 2, 2, 2, c0);
  orc_lagrange_coeff(w, 2, 1, 2, c1);
  orc_lagrange_coeff(w, 2, 0, 2, c2);
  orc_lagrange_coeff(w, 4, 2, 2, full);
  double lw0 = full[0] / c0[0];
  double lw1 = full[4] / c2[2];
  double lw2 = 1.0 - lw0 - lw1;
  for (int e = 0; e < NEQ; ++e) {
    double s0 = c0[0] * u3[e] + c0[1] * u2[e] + c0[2] * u1[e];
    double s1 = c1[0] * u2[e] + c1[1] * u1[e] + c1[2] * d1[e];
    double s2 = c2[0] * u1[e] + c2[1] * d1[e] + c2[2] * d2[e];
    double b0 = orc_weno_beta(0, w[0], w[1], w[2], u3[e], u2[e], u1[e]);
    double b1 = orc_weno_beta(1, w[1], w[2], w[3], u2[e], u1[e], d1[e]);
    double b2 = orc_weno_beta(2, w[2], w[3], w[4], u1[e], d1[e], d2[e]);
    double n0, n1, n2;
    if (wenoZ) {
      double tau5 = fabs(b0 - b2);
      double t0 = tau5 / (1.0e-40 + b0), t1 = tau5 / (1.0e-40 + b1), t2 = tau5 / (1.0e-40 + b2);
      n0 = lw0 * (1.0 + t0 * t0);
      n1 = lw1 * (1.0 + t1 * t1);
      n2 = lw2 * (1.0 + t2 * t2);
    } else {
      double e0 = 1.0e-6 + b0, e1 = 1.0e-6 + b1, e2 = 1.0e-6 + b2;
      n0 = lw0 / (e0 * e0);
      n1 = lw1 / (e1 * e1);
      n2 = lw2 / (e2 * e2);
    }
    double sum = n0 + n1 + n2;
    n0 /= sum;
    n1 /= sum;
    n2 /= sum;
    out[e] = n0 * s0 + n1 * s1 + n2 * s2;
  }
}

/* ------------------------------------------------------------------------ */
/* spectral radii: include/spectralRadius.hpp:43-80                        */

static double inv_cell_specrad(const orc_gas *g, const double *W, uv3 aL,
                               uv3 aR) {
  v3 s = vadd(aL.u, aR.u);
  v3 avg = vscl(0.5, s);
  v3 normAvg = vdiv(avg, vmag(avg));
  double fMag = 0.5 * (aL.m + aR.m);
  return (fabs(vdot(V(W[1], W[2], W[3]), normAvg)) + orc_sos(g, W)) * fMag;
}
double orc_inv_cell_specrad(const orc_gas *g, const double *W, const double *aL,
                            const double *aR) {
  uv3 l = {V(aL[0], aL[1], aL[2]), aL[3]};
  uv3 r = {V(aR[0], aR[1], aR[2]), aR[3]};
  return inv_cell_specrad(g, W, l, r);
}
static double inv_face_specrad(const orc_gas *g, const double *W, uv3 a) {
  return 0.5 * a.m * (fabs(vdot(V(W[1], W[2], W[3]), a.u)) + orc_sos(g, W));
}

/* RusanovScalarOffDiagonal (inviscid, no turbulence):
 * src/fluxJacobian.cpp:122-162; ConvectiveFluxUpdate
 * include/inviscidFlux.hpp:543-562 */
static void scalar_offdiag_v(const orc_gas *g, const double *Wnb, const double *du, uv3 a,
                             int positive, int viscous, double lamVisc, double *out) {
  double Wu[NEQ];
  update_prim(g, Wnb, du, Wu);
  double n[3] = {a.u.x, a.u.y, a.u.z};
  double Fold[NEQ], Fnew[NEQ];
  orc_phys_flux(g, Wnb, n, Fold);
  orc_phys_flux(g, Wu, n, Fnew);
  double half = 0.5 * a.m;
  double lam = inv_face_specrad(g, Wnb, a);
  if (viscous) lam += lamVisc;
  for (int e = 0; e < NEQ; ++e) {
    double fc = half * (Fnew[e] - Fold[e]);
    double sd = du[e] * lam;
    out[e] = positive ? fc + sd : fc - sd;
  }
}
static void scalar_offdiag(const orc_gas *g, const double *Wnb,
                           const double *du, uv3 a, int positive, double *out) {
  scalar_offdiag_v(g, Wnb, du, a, positive, 0, 0.0, out);
}
void orc_scalar_offdiag(const orc_gas *g, const double *Wnb, const double *du,
                        const double *area, int positive, double *out) {
  uv3 a = {V(area[0], area[1], area[2]), area[3]};
  scalar_offdiag(g, Wnb, du, a, positive, out);
}

/* RoeOffDiagonal, inviscid (src/fluxJacobian.cpp:240-290): the change of the
 * Roe flux through the face when the neighbour's state moves by du.  The old
 * flux always takes the neighbour as its left state and the cell (`Wd`) as
 * its right one; the new flux takes the updated neighbour as the left state
 * for a lower (positive) and as the right state for an upper neighbour.
 * The spectral radius term is zero without the viscous terms, applied as
 * +-(du * 0). */
static void roe_offdiag(const orc_gas *g, const double *Wnb, const double *Wd, const double *du, uv3 a,
                        int positive, double *out) {
  double n[3] = {a.u.x, a.u.y, a.u.z};
  double Fold[NEQ], Fnew[NEQ], Wu[NEQ];
  orc_roe_flux(g, Wnb, Wd, n, Fold);
  update_prim(g, Wnb, du, Wu);
  if (positive) orc_roe_flux(g, Wu, Wd, n, Fnew);
  else orc_roe_flux(g, Wd, Wu, n, Fnew);
  for (int e = 0; e < NEQ; ++e) {
    double fc = a.m * (Fnew[e] - Fold[e]);
    double sd = du[e] * 0.0;
    out[e] = positive ? fc + sd : fc - sd;
  }
}
void orc_roe_offdiag(const orc_gas *g, const double *Wnb, const double *Wd, const double *du,
                     const double *area, int positive, double *out) {
  uv3 a = {V(area[0], area[1], area[2]), area[3]};
  roe_offdiag(g, Wnb, Wd, du, a, positive, out);
}

/* ------------------------------------------------------------------------ */
/* block matrices (blusgs, bdplur): 5x5, row-major                          */

/* fluxJacobian::InvFluxJacobian and RusanovFluxJacobian
 * (include/fluxJacobian.hpp:447-560) for one species: 0.5 |A| dF/dU at the
 * state, then the face spectral radius added to (positive) or subtracted from
 * every entry of the identity (off-diagonal entries +/- 0) */
static void rusanov_jacobian(const orc_gas *g, const double *W, uv3 a, int positive, double *J) {
  double sr = inv_face_specrad(g, W, a);
  v3 n = a.u;
  v3 vel = V(W[1], W[2], W[3]);
  double u = W[1], v = W[2], w = W[3];
  double vn = vdot(vel, n);
  double mf = W[0] / (0.0 + W[0]);
  double gamma = orc_gamma(g);
  double gm1 = gamma - 1.0;
  double phi = 0.5 * gm1 * vdot(vel, vel);
  double a1 = gamma * energy(g, W) - phi;
  double a3 = gamma - 2.0;
  for (int q = 0; q < 25; ++q) J[q] = 0.0;
  J[0] = vn * (1.0 - mf);
  J[1] = mf * n.x;
  J[2] = mf * n.y;
  J[3] = mf * n.z;
  J[5] = phi * n.x - u * vn;
  J[10] = phi * n.y - v * vn;
  J[15] = phi * n.z - w * vn;
  J[20] = vn * (phi - a1);
  J[6] = vn - a3 * n.x * u;
  J[11] = v * n.x - gm1 * u * n.y;
  J[16] = w * n.x - gm1 * u * n.z;
  J[21] = a1 * n.x - gm1 * u * vn;
  J[7] = u * n.y - gm1 * v * n.x;
  J[12] = vn - a3 * n.y * v;
  J[17] = w * n.y - gm1 * v * n.z;
  J[22] = a1 * n.y - gm1 * v * vn;
  J[8] = u * n.z - gm1 * w * n.x;
  J[13] = v * n.z - gm1 * w * n.y;
  J[18] = vn - a3 * n.z * w;
  J[23] = a1 * n.z - gm1 * w * vn;
  J[9] = gm1 * n.x;
  J[14] = gm1 * n.y;
  J[19] = gm1 * n.z;
  J[24] = gamma * vn;
  double f = 0.5 * a.m;
  for (int q = 0; q < 25; ++q) J[q] *= f;
  for (int r = 0; r < 5; ++r)
    for (int cc = 0; cc < 5; ++cc) {
      double dis = (r == cc) ? 1.0 * sr : 0.0;
      J[r * 5 + cc] = positive ? J[r * 5 + cc] + dis : J[r * 5 + cc] - dis;
    }
}
void orc_rusanov_jacobian(const orc_gas *g, const double *W, const double *area, int positive,
                          double *J) {
  uv3 a = {V(area[0], area[1], area[2]), area[3]};
  rusanov_jacobian(g, W, a, positive, J);
}

/* ArrayMultiplication, block branch (include/fluxJacobian.hpp:50-86):
 * out[r] = sum over c in order of M[r][c] v[c], from +0 */
static void block_mult(const double *M, const double *v, double *out) {
  for (int r = 0; r < 5; ++r) {
    double s = 0.0;
    for (int cc = 0; cc < 5; ++cc) s += M[r * 5 + cc] * v[cc];
    out[r] = s;
  }
}

/* LinearCombRow (src/matrix.cpp:119-132): row r2 -= factor * row r1 */
static void lin_comb_row(double *m, int n, int r1, double factor, int r2) {
  for (int cc = 0; cc < n; ++cc) m[r2 * n + cc] = m[r2 * n + cc] - factor * m[r1 * n + cc];
}
static void swap_rows(double *m, int n, int r1, int r2) {
  if (r1 == r2) return;
  for (int cc = 0; cc < n; ++cc) {
    double t = m[r1 * n + cc];
    m[r1 * n + cc] = m[r2 * n + cc];
    m[r2 * n + cc] = t;
  }
}

/* MatrixInverse (src/matrix.cpp:57-105) with FindMaxInColumn (:145-165):
 * the pivot of row r is the largest |m[i][r]|, i >= r, taken before row r's
 * leading entries are eliminated (row 0 when every candidate vanishes); row r
 * is then eliminated against rows 0..r-1, normalised from the pivot column
 * on, and the upper triangle is cleared from the bottom up.  Returns -1 on a
 * zero pivot (the reference exits). */
static int gj_inverse(double *m, int n) {
  double I[25];
  for (int r = 0; r < n; ++r)
    for (int cc = 0; cc < n; ++cc) I[r * n + cc] = (r == cc) ? 1.0 : 0.0;
  for (int r = 0; r < n; ++r) {
    int cp = r;
    double mv = 0.0;
    int pr = 0;
    for (int ii = cp; ii <= n - 1; ++ii)
      if (fabs(m[ii * n + r]) > mv) {
        mv = fabs(m[ii * n + r]);
        pr = ii;
      }
    swap_rows(m, n, r, pr);
    swap_rows(I, n, r, pr);
    if (r != 0)
      for (int ii = 0; ii < cp; ++ii) {
        double factor = m[r * n + ii] / m[ii * n + ii];
        lin_comb_row(m, n, ii, factor, r);
        lin_comb_row(I, n, ii, factor, r);
      }
    if (m[r * n + cp] == 0.0) return -1;
    double nf = 1.0 / m[r * n + cp];
    for (int cc = cp; cc < n; ++cc) m[r * n + cc] *= nf;
    for (int cc = 0; cc < n; ++cc) I[r * n + cc] *= nf;
  }
  for (int r = n - 2; r >= 0; --r) {
    int cp = r;
    for (int ii = n - 1; ii > cp; --ii) {
      double factor = m[r * n + ii];
      lin_comb_row(m, n, ii, factor, r);
      lin_comb_row(I, n, ii, factor, r);
    }
  }
  memcpy(m, I, sizeof(double) * (size_t)(n * n));
  return 0;
}
int orc_block_inverse(double *m) { return gj_inverse(m, 5); }

/* ------------------------------------------------------------------------ */
/* geometry: src/plot3d.cpp:36-359, 490-498                                 */

static v3 P(const v3 *c, int ni, int nj, int i, int j, int k) {
  return c[(long)i + (long)j * (ni + 1) + (long)k * (ni + 1) * (nj + 1)];
}

double orc_pyramid_volume(const double *p, const double *a, const double *b,
                          const double *c, const double *d) {
  v3 vp = V(p[0], p[1], p[2]), va = V(a[0], a[1], a[2]),
     vb = V(b[0], b[1], b[2]), vc = V(c[0], c[1], c[2]),
     vd = V(d[0], d[1], d[2]);
  v3 s = vadd(vadd(vadd(vsub(va, vp), vsub(vb, vp)), vsub(vc, vp)), vsub(vd, vp));
  v3 xp = vscl(0.25, s);
  v3 xac = vsub(vc, va);
  v3 xbd = vsub(vd, vb);
  return 1.0 / 6.0 * vdot(xp, vcross(xac, xbd));
}
static double pyr(v3 p, v3 a, v3 b, v3 c, v3 d) {
  double pp[3] = {p.x, p.y, p.z}, aa[3] = {a.x, a.y, a.z},
         bb[3] = {b.x, b.y, b.z}, cc[3] = {c.x, c.y, c.z},
         dd[3] = {d.x, d.y, d.z};
  return orc_pyramid_volume(pp, aa, bb, cc, dd);
}

static v3 centroid(const v3 *c, int ni, int nj, int i, int j, int k) {
  v3 s = P(c, ni, nj, i, j, k);
  s = vadd(s, P(c, ni, nj, i + 1, j, k));
  s = vadd(s, P(c, ni, nj, i, j + 1, k));
  s = vadd(s, P(c, ni, nj, i + 1, j + 1, k));
  s = vadd(s, P(c, ni, nj, i, j, k + 1));
  s = vadd(s, P(c, ni, nj, i + 1, j, k + 1));
  s = vadd(s, P(c, ni, nj, i, j + 1, k + 1));
  s = vadd(s, P(c, ni, nj, i + 1, j + 1, k + 1));
  return vscl(0.125, s);
}

static int build_geometry(blk_t *b, const v3 *c) {
  int ni = b->ni, nj = b->nj, nk = b->nk;
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        v3 ct = centroid(c, ni, nj, i, j, k);
        double v = 0.0;
        v += pyr(ct, P(c, ni, nj, i, j, k), P(c, ni, nj, i, j, k + 1),
                 P(c, ni, nj, i, j + 1, k + 1), P(c, ni, nj, i, j + 1, k));
        v += pyr(ct, P(c, ni, nj, i + 1, j, k), P(c, ni, nj, i + 1, j + 1, k),
                 P(c, ni, nj, i + 1, j + 1, k + 1), P(c, ni, nj, i + 1, j, k + 1));
        v += pyr(ct, P(c, ni, nj, i, j, k), P(c, ni, nj, i + 1, j, k),
                 P(c, ni, nj, i + 1, j, k + 1), P(c, ni, nj, i, j, k + 1));
        v += pyr(ct, P(c, ni, nj, i, j + 1, k), P(c, ni, nj, i, j + 1, k + 1),
                 P(c, ni, nj, i + 1, j + 1, k + 1), P(c, ni, nj, i + 1, j + 1, k));
        v += pyr(ct, P(c, ni, nj, i, j, k), P(c, ni, nj, i, j + 1, k),
                 P(c, ni, nj, i + 1, j + 1, k), P(c, ni, nj, i + 1, j, k));
        v += pyr(ct, P(c, ni, nj, i, j, k + 1), P(c, ni, nj, i + 1, j, k + 1),
                 P(c, ni, nj, i + 1, j + 1, k + 1), P(c, ni, nj, i, j + 1, k + 1));
        if (!(v > 0.0)) return -1;
        b->vol[cidx(b, i, j, k)] = v;
        b->ctr[cidx(b, i, j, k)] = ct;
      }
  /* i faces */
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i <= ni; ++i) {
        v3 xac = vsub(P(c, ni, nj, i, j + 1, k + 1), P(c, ni, nj, i, j, k));
        v3 xbd = vsub(P(c, ni, nj, i, j + 1, k), P(c, ni, nj, i, j, k + 1));
        b->fa[0][fidx(b, 0, i, j, k)] = make_uv3(vscl(0.5, vcross(xbd, xac)));
        v3 s = P(c, ni, nj, i, j, k);
        s = vadd(s, P(c, ni, nj, i, j + 1, k));
        s = vadd(s, P(c, ni, nj, i, j, k + 1));
        s = vadd(s, P(c, ni, nj, i, j + 1, k + 1));
        b->fc[0][fidx(b, 0, i, j, k)] = vscl(0.25, s);
      }
  /* j faces */
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j <= nj; ++j)
      for (int i = 0; i < ni; ++i) {
        v3 xac = vsub(P(c, ni, nj, i, j, k + 1), P(c, ni, nj, i + 1, j, k));
        v3 xbd = vsub(P(c, ni, nj, i, j, k), P(c, ni, nj, i + 1, j, k + 1));
        b->fa[1][fidx(b, 1, i, j, k)] = make_uv3(vscl(0.5, vcross(xbd, xac)));
        v3 s = P(c, ni, nj, i, j, k);
        s = vadd(s, P(c, ni, nj, i + 1, j, k));
        s = vadd(s, P(c, ni, nj, i, j, k + 1));
        s = vadd(s, P(c, ni, nj, i + 1, j, k + 1));
        b->fc[1][fidx(b, 1, i, j, k)] = vscl(0.25, s);
      }
  /* k faces */
  for (int k = 0; k <= nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        v3 xac = vsub(P(c, ni, nj, i, j + 1, k), P(c, ni, nj, i + 1, j, k));
        v3 xbd = vsub(P(c, ni, nj, i + 1, j + 1, k), P(c, ni, nj, i, j, k));
        b->fa[2][fidx(b, 2, i, j, k)] = make_uv3(vscl(0.5, vcross(xbd, xac)));
        v3 s = P(c, ni, nj, i, j, k);
        s = vadd(s, P(c, ni, nj, i + 1, j, k));
        s = vadd(s, P(c, ni, nj, i, j + 1, k));
        s = vadd(s, P(c, ni, nj, i + 1, j + 1, k));
        b->fc[2][fidx(b, 2, i, j, k)] = vscl(0.25, s);
      }
  return 0;
}

/* face index along direction d for "cell" q with face array of direction d,
 * lower/upper surface (SliceArray/InsertArray with id == dir:
 * include/multiArray3d.hpp:536-580, 722-760) */
static int face_of(int q, int upper) { return upper ? q + 1 : q; }

/* procBlock::AssignGhostCellsGeom: src/procBlock.cpp:2160-2268 */
static void ghost_geometry(blk_t *b) {
  for (int layer = 1; layer <= b->g; ++layer) {
    for (int s = 0; s < b->nsurf; ++s) {
      const surf_t *sf = &b->surf[s];
      if (strcmp(sf->type, "interblock") == 0) continue;
      int st = surf_type(sf);
      int dir = (st - 1) / 2;
      int upper = (st % 2 == 0);
      int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
      int r1s = sf->d[2 * d1], r1e = sf->d[2 * d1 + 1];
      int r2s = sf->d[2 * d2], r2e = sf->d[2 * d2 + 1];
      int r3 = sf->d[2 * dir];
      int n3 = ndir(b, dir);
      int gC, iC, pC, piC, iF, piF;
      if (upper) {
        gC = r3 + layer - 1;
        iC = r3 - layer;
        pC = gC - 1;
        if (iC < 0) iC = 0;
        piC = iC + 1;
        iF = r3 - layer;
        if (iF < 0) iF = 0;
        piF = iF + 1;
      } else {
        gC = r3 - layer;
        iC = r3 + layer - 1;
        pC = gC + 1;
        if (iC >= n3) iC = n3 - 1;
        piC = iC - 1;
        iF = r3 + layer;
        if (iF > n3) iF = n3;
        piF = iF - 1;
      }
      /* distF2F and distC2C over the surface patch */
      int n1 = r1e - r1s, n2 = r2e - r2s;
      v3 *f2f = (v3 *)malloc(sizeof(v3) * (size_t)(n1 * n2));
      v3 *c2c = (v3 *)malloc(sizeof(v3) * (size_t)(n1 * n2));
      for (int q2 = 0; q2 < n2; ++q2)
        for (int q1 = 0; q1 < n1; ++q1) {
          int i, j, k, i2, j2, k2;
          map3(dir, piF, r1s + q1, r2s + q2, &i, &j, &k);
          map3(dir, iF, r1s + q1, r2s + q2, &i2, &j2, &k2);
          f2f[q1 + q2 * n1] =
              vsub(b->fc[dir][fidx(b, dir, i, j, k)], b->fc[dir][fidx(b, dir, i2, j2, k2)]);
          if (layer > 1) {
            map3(dir, piC, r1s + q1, r2s + q2, &i, &j, &k);
            map3(dir, iC, r1s + q1, r2s + q2, &i2, &j2, &k2);
            c2c[q1 + q2 * n1] = vsub(b->ctr[cidx(b, i, j, k)], b->ctr[cidx(b, i2, j2, k2)]);
          } else {
            c2c[q1 + q2 * n1] = f2f[q1 + q2 * n1];
          }
        }
      /* volumes, normal-direction face data, centroids */
      for (int q2 = 0; q2 < n2; ++q2)
        for (int q1 = 0; q1 < n1; ++q1) {
          int gi, gj, gk, ii, ij, ik, pi, pj, pk;
          map3(dir, gC, r1s + q1, r2s + q2, &gi, &gj, &gk);
          map3(dir, iC, r1s + q1, r2s + q2, &ii, &ij, &ik);
          map3(dir, pC, r1s + q1, r2s + q2, &pi, &pj, &pk);
          b->vol[cidx(b, gi, gj, gk)] = b->vol[cidx(b, ii, ij, ik)];
          /* normal-direction face area and centre */
          {
            int fi, fj, fk, si, sj, sk, ti, tj, tk;
            map3(dir, face_of(gC, upper), r1s + q1, r2s + q2, &fi, &fj, &fk);
            map3(dir, face_of(iC, upper), r1s + q1, r2s + q2, &si, &sj, &sk);
            map3(dir, face_of(pC, upper), r1s + q1, r2s + q2, &ti, &tj, &tk);
            b->fa[dir][fidx(b, dir, fi, fj, fk)] = b->fa[dir][fidx(b, dir, si, sj, sk)];
            b->fc[dir][fidx(b, dir, fi, fj, fk)] =
                vadd(f2f[q1 + q2 * n1], b->fc[dir][fidx(b, dir, ti, tj, tk)]);
          }
          b->ctr[cidx(b, gi, gj, gk)] =
              vadd(b->ctr[cidx(b, pi, pj, pk)], c2c[q1 + q2 * n1]);
        }
      /* transverse face data: range grown by one along the face direction;
       * distC2C is duplicated at the grown end (GrowInI/J/K,
       * include/multiArray3d.hpp:929-990) */
      for (int t = 1; t <= 2; ++t) {
        int fd = (dir + t) % 3;
        int m1 = n1 + (t == 1), m2 = n2 + (t == 2);
        for (int q2 = 0; q2 < m2; ++q2)
          for (int q1 = 0; q1 < m1; ++q1) {
            int gi, gj, gk, ii, ij, ik, pi, pj, pk;
            map3(dir, gC, r1s + q1, r2s + q2, &gi, &gj, &gk);
            map3(dir, iC, r1s + q1, r2s + q2, &ii, &ij, &ik);
            map3(dir, pC, r1s + q1, r2s + q2, &pi, &pj, &pk);
            int c1 = q1 < n1 ? q1 : n1 - 1, c2 = q2 < n2 ? q2 : n2 - 1;
            b->fa[fd][fidx(b, fd, gi, gj, gk)] = b->fa[fd][fidx(b, fd, ii, ij, ik)];
            b->fc[fd][fidx(b, fd, gi, gj, gk)] =
                vadd(c2c[c1 + c2 * n1], b->fc[fd][fidx(b, fd, pi, pj, pk)]);
          }
      }
      free(f2f);
      free(c2c);
    }
  }
}

/* procBlock::CalcCellWidths: src/procBlock.cpp:6397-6411 (all cells incl.
 * ghosts; edge/corner ghost cells hold unused values) */
static void cell_widths(blk_t *b) {
  int g = b->g;
  for (int k = -g; k < b->nk + g; ++k)
    for (int j = -g; j < b->nj + g; ++j)
      for (int i = -g; i < b->ni + g; ++i) {
        long c = cidx(b, i, j, k);
        b->cw[0][c] = vmag(vsub(b->fc[0][fidx(b, 0, i, j, k)], b->fc[0][fidx(b, 0, i + 1, j, k)]));
        b->cw[1][c] = vmag(vsub(b->fc[1][fidx(b, 1, i, j, k)], b->fc[1][fidx(b, 1, i, j + 1, k)]));
        b->cw[2][c] = vmag(vsub(b->fc[2][fidx(b, 2, i, j, k)], b->fc[2][fidx(b, 2, i, j, k + 1)]));
      }
}

/* ------------------------------------------------------------------------ */
/* interblock connections                                                   */

static void seterr(char *err, int errlen, const char *fmt, ...);
static double proj_c2c(const blk_t *b, int d, int i, int j, int k);
static void edge_ghosts(orc_case *c, blk_t *b, int viscous);
static const bcstate_t *find_periodic(const orc_case *c, int tag);

/* partner block and surface from the tag: src/boundaryConditions.cpp:2455-2495 */
static int tag_partner_surface(int tag) {
  if (tag < 2000) return 1;
  if (tag < 3000) return 2;
  if (tag < 4000) return 3;
  if (tag < 5000) return 4;
  if (tag < 6000) return 5;
  if (tag < 7000) return 6;
  return -1;
}

static v3 node_at(const v3 *nodes, const blk_t *b, int dir, int r3, int a1, int a2) {
  int i, j, k;
  map3(dir, r3, a1, a2, &i, &j, &k);
  return nodes[(long)i + (long)(b->ni + 1) * (j + (long)(b->nj + 1) * k)];
}

static int same_point(v3 a, v3 b) {
  /* vector3d::CompareWithTol, include/vector3d.hpp:319-327 */
  return fabs(a.x - b.x) < 1.0e-10 && fabs(a.y - b.y) < 1.0e-10 && fabs(a.z - b.z) < 1.0e-10;
}

static void conn_map(const conn_t *cn, int p1, int p2, int cells, int *q1, int *q2) {
  int c1 = cn->swap ? p2 : p1, c2 = cn->swap ? p1 : p2;
  int m1 = cells ? cn->pn1 - 1 : cn->pn1, m2 = cells ? cn->pn2 - 1 : cn->pn2;
  *q1 = cn->rev1 ? m1 - c1 : c1;
  *q2 = cn->rev2 ? m2 - c2 : c2;
}

/* patch match by three corners (TestPatchMatch, src/boundaryConditions.cpp:
 * 729-831): find how the partner's patch axes map onto this patch's */
static int orient_patch(conn_t *cn, const orc_case *c, const v3 *const *nodes, int a) {
  const blk_t *A = &c->blk[a], *B = &c->blk[cn->pb];
  int corners[3][2] = {{0, 0}, {cn->n1, 0}, {0, cn->n2}};
  for (int swap = 0; swap < 2; ++swap)
    for (int rev1 = 0; rev1 < 2; ++rev1)
      for (int rev2 = 0; rev2 < 2; ++rev2) {
        cn->swap = swap;
        cn->rev1 = rev1;
        cn->rev2 = rev2;
        if (swap ? (cn->n1 != cn->pn2 || cn->n2 != cn->pn1) : (cn->n1 != cn->pn1 || cn->n2 != cn->pn2))
          continue;
        int ok = 1;
        for (int q = 0; q < 3 && ok; ++q) {
          int b1, b2;
          conn_map(cn, corners[q][0], corners[q][1], 0, &b1, &b2);
          v3 pa = vadd(node_at(nodes[a], A, cn->dir, cn->r3, cn->r1s + corners[q][0], cn->r2s + corners[q][1]),
                       cn->shift);
          v3 pbn = node_at(nodes[cn->pb], B, cn->pdir, cn->pr3, cn->pr1s + b1, cn->pr2s + b2);
          ok = same_point(pa, pbn);
        }
        if (ok) return 0;
      }
  return -1;
}

/* this side's ghost cell at layer l (1..g) and the partner's interior cell at
 * layer l, patch-local (p1, p2): the ghost cells continue the partner block
 * as if the grid were not cut (src/boundaryConditions.cpp:3006-3185) */
static long conn_ghost(const conn_t *cn, const blk_t *A, int l, int p1, int p2) {
  int n3 = (cn->s % 2 == 1) ? -l : ndir(A, cn->dir) + l - 1;
  int i, j, k;
  map3(cn->dir, n3, cn->r1s + p1, cn->r2s + p2, &i, &j, &k);
  return cidx(A, i, j, k);
}
static long conn_source(const conn_t *cn, const blk_t *B, int l, int p1, int p2) {
  int q1, q2;
  conn_map(cn, p1, p2, 1, &q1, &q2);
  int n3 = (cn->ps % 2 == 1) ? l - 1 : ndir(B, cn->pdir) - l;
  int i, j, k;
  map3(cn->pdir, n3, cn->pr1s + q1, cn->pr2s + q2, &i, &j, &k);
  return cidx(B, i, j, k);
}

/* SwapGeomSlice / procBlock::PutGeomSlice (src/utility.cpp:213-256,
 * src/procBlock.cpp:3167-4011), restricted to the patch: each ghost cell
 * takes the partner cell's volume and centroid; its faces normal to the
 * patch (the boundary face included) take the partner's faces at the same
 * distance from the cut, with the area direction reversed when both patches
 * are lower or both upper surfaces; its faces along the patch take the
 * partner's, swapped and reversed where the partner's axis runs backwards.
 * All swaps read the geometry as it was before any of them. */
typedef struct { double *vol; v3 *ctr; uv3 *fa[3]; v3 *fc[3]; } geom_copy;

static void swap_geometry(orc_case *c) {
  geom_copy *gc = (geom_copy *)calloc((size_t)c->nb, sizeof(geom_copy));
  for (int bq = 0; bq < c->nb; ++bq) {
    blk_t *b = &c->blk[bq];
    long nc = ncells_padded(b);
    gc[bq].vol = (double *)malloc(sizeof(double) * (size_t)nc);
    gc[bq].ctr = (v3 *)malloc(sizeof(v3) * (size_t)nc);
    memcpy(gc[bq].vol, b->vol, sizeof(double) * (size_t)nc);
    memcpy(gc[bq].ctr, b->ctr, sizeof(v3) * (size_t)nc);
    for (int d = 0; d < 3; ++d) {
      long nf = nfaces_padded(b, d);
      gc[bq].fa[d] = (uv3 *)malloc(sizeof(uv3) * (size_t)nf);
      gc[bq].fc[d] = (v3 *)malloc(sizeof(v3) * (size_t)nf);
      memcpy(gc[bq].fa[d], b->fa[d], sizeof(uv3) * (size_t)nf);
      memcpy(gc[bq].fc[d], b->fc[d], sizeof(v3) * (size_t)nf);
    }
  }
  for (int a = 0; a < c->nb; ++a) {
    blk_t *A = &c->blk[a];
    for (int q = 0; q < A->nconn; ++q) {
      const conn_t *cn = &A->conn[q];
      if (cn->periodic) continue; /* periodic ghosts keep the mirrored geometry */
      const blk_t *B = &c->blk[cn->pb];
      const geom_copy *G = &gc[cn->pb];
      double f3 = ((cn->s + cn->ps) % 2 == 0) ? -1.0 : 1.0;
      int lowA = cn->s % 2 == 1, lowB = cn->ps % 2 == 1;
      int nA = ndir(A, cn->dir), nB = ndir(B, cn->pdir);
      int dA1 = (cn->dir + 1) % 3, dA2 = (cn->dir + 2) % 3;
      int dB1 = (cn->pdir + 1) % 3, dB2 = (cn->pdir + 2) % 3;
      for (int p2 = 0; p2 < cn->n2; ++p2)
        for (int p1 = 0; p1 < cn->n1; ++p1) {
          int q1, q2;
          conn_map(cn, p1, p2, 1, &q1, &q2);
          /* faces normal to the patch at distance t = 0..g from the cut */
          for (int t = 0; t <= A->g; ++t) {
            int fa3 = lowA ? -t : nA + t, fb3 = lowB ? t : nB - t;
            int i, j, k, bi, bj, bk;
            map3(cn->dir, fa3, cn->r1s + p1, cn->r2s + p2, &i, &j, &k);
            map3(cn->pdir, fb3, cn->pr1s + q1, cn->pr2s + q2, &bi, &bj, &bk);
            uv3 v = G->fa[cn->pdir][fidx(B, cn->pdir, bi, bj, bk)];
            if (f3 < 0.0) v.u = vscl(-1.0, v.u);
            A->fa[cn->dir][fidx(A, cn->dir, i, j, k)] = v;
            A->fc[cn->dir][fidx(A, cn->dir, i, j, k)] = G->fc[cn->pdir][fidx(B, cn->pdir, bi, bj, bk)];
          }
          for (int l = 1; l <= A->g; ++l) {
            long ga = conn_ghost(cn, A, l, p1, p2), sb = conn_source(cn, B, l, p1, p2);
            A->vol[ga] = G->vol[sb];
            A->ctr[ga] = G->ctr[sb];
            int na3 = lowA ? -l : nA + l - 1, nb3 = lowB ? l - 1 : nB - l;
            int gi, gj, gk, si, sj, sk;
            map3(cn->dir, na3, cn->r1s + p1, cn->r2s + p2, &gi, &gj, &gk);
            map3(cn->pdir, nb3, cn->pr1s + q1, cn->pr2s + q2, &si, &sj, &sk);
            /* faces along the patch: axis dA1 <-> (swap ? dB2 : dB1), dA2 <-> the other */
            for (int w = 0; w < 2; ++w) {
              int da = w == 0 ? dA1 : dA2;
              int useB1 = (w == 0) != (cn->swap != 0);
              int db = useB1 ? dB1 : dB2;
              int rev = useB1 ? cn->rev1 : cn->rev2;
              for (int up = 0; up < 2; ++up) {
                int ai = gi + (da == 0 && up), aj = gj + (da == 1 && up), ak = gk + (da == 2 && up);
                int bup = rev ? !up : up;
                int xi = si + (db == 0 && bup), xj = sj + (db == 1 && bup), xk = sk + (db == 2 && bup);
                uv3 v = G->fa[db][fidx(B, db, xi, xj, xk)];
                if (rev) v.u = vscl(-1.0, v.u);
                A->fa[da][fidx(A, da, ai, aj, ak)] = v;
                A->fc[da][fidx(A, da, ai, aj, ak)] = G->fc[db][fidx(B, db, xi, xj, xk)];
              }
            }
          }
        }
    }
  }
  for (int bq = 0; bq < c->nb; ++bq) {
    free(gc[bq].vol);
    free(gc[bq].ctr);
    for (int d = 0; d < 3; ++d) { free(gc[bq].fa[d]); free(gc[bq].fc[d]); }
  }
  free(gc);
}

/* pair every interblock surface with its partner patch
 * (GetBlockInterConnBCs, src/boundaryConditions.cpp:603-642) */
static int connect_blocks(orc_case *c, const v3 *const *nodes, char *err, int errlen) {
  for (int a = 0; a < c->nb; ++a) {
    blk_t *A = &c->blk[a];
    for (int q = 0; q < 6; ++q) {
      int d = (q / 2), d1 = (d + 1) % 3, d2 = (d + 2) % 3;
      int n = (d1 == 0 ? A->ni : d1 == 1 ? A->nj : A->nk) * (d2 == 0 ? A->ni : d2 == 1 ? A->nj : A->nk);
      A->isconn[q] = (unsigned char *)calloc((size_t)n, 1);
    }
    for (int s = 0; s < A->nsurf; ++s) {
      const surf_t *sf = &A->surf[s];
      int isPer = !strcmp(sf->type, "periodic");
      if (strcmp(sf->type, "interblock") && !isPer) continue;
      conn_t cn;
      memset(&cn, 0, sizeof cn);
      cn.s = surf_type(sf);
      cn.dir = (cn.s - 1) / 2;
      int d1 = (cn.dir + 1) % 3, d2 = (cn.dir + 2) % 3;
      cn.r3 = sf->d[2 * cn.dir];
      cn.r1s = sf->d[2 * d1];
      cn.n1 = sf->d[2 * d1 + 1] - cn.r1s;
      cn.r2s = sf->d[2 * d2];
      cn.n2 = sf->d[2 * d2 + 1] - cn.r2s;
      if (isPer) {
        /* periodic pairs match by geometry once the start surface is moved
         * by the translation (GetConnectionBCs, src/boundaryConditions.cpp:
         * 543-593, patch::Transform :2223-2248) */
        const bcstate_t *ps = find_periodic(c, sf->d[6]);
        if (!ps) { seterr(err, errlen, "block %d: no periodic state for tag %d", a, sf->d[6]); return -1; }
        int other = ps->startTag == sf->d[6] ? ps->endTag : ps->startTag;
        cn.periodic = 1;
        cn.shift = ps->startTag == sf->d[6] ? ps->trans : vscl(-1.0, ps->trans);
        int found = 0;
        for (int pbq = 0; pbq < c->nb && !found; ++pbq) {
          const blk_t *B = &c->blk[pbq];
          for (int t = 0; t < B->nsurf && !found; ++t) {
            const surf_t *pf = &B->surf[t];
            if (strcmp(pf->type, "periodic") || pf->d[6] != other || pf == sf) continue;
            cn.pb = pbq;
            cn.ps = surf_type(pf);
            cn.pdir = (cn.ps - 1) / 2;
            int e1 = (cn.pdir + 1) % 3, e2 = (cn.pdir + 2) % 3;
            cn.pr3 = pf->d[2 * cn.pdir];
            cn.pr1s = pf->d[2 * e1];
            cn.pn1 = pf->d[2 * e1 + 1] - cn.pr1s;
            cn.pr2s = pf->d[2 * e2];
            cn.pn2 = pf->d[2 * e2 + 1] - cn.pr2s;
            if (orient_patch(&cn, c, nodes, a) == 0) found = 1;
          }
        }
        if (!found) { seterr(err, errlen, "block %d: no matching periodic patch for tag %d", a, sf->d[6]); return -1; }
        if (ndir(&c->blk[cn.pb], cn.pdir) < A->g) {
          seterr(err, errlen, "block %d: periodic partner is thinner than the ghost layers", a);
          return -1;
        }
        A->conn = (conn_t *)realloc(A->conn, sizeof(conn_t) * (size_t)(A->nconn + 1));
        A->conn[A->nconn++] = cn;
        int n1full = d1 == 0 ? A->ni : d1 == 1 ? A->nj : A->nk;
        for (int p2 = 0; p2 < cn.n2; ++p2)
          for (int p1 = 0; p1 < cn.n1; ++p1)
            A->isconn[cn.s - 1][(cn.r1s + p1) + (long)n1full * (cn.r2s + p2)] = 1;
        continue;
      }
      cn.ps = tag_partner_surface(sf->d[6]);
      cn.pb = sf->d[6] - 1000 * cn.ps;
      if (cn.ps < 0 || cn.pb < 0 || cn.pb >= c->nb) {
        seterr(err, errlen, "block %d: bad interblock tag %d", a, sf->d[6]);
        return -1;
      }
      cn.pdir = (cn.ps - 1) / 2;
      const blk_t *B = &c->blk[cn.pb];
      int found = 0;
      for (int t = 0; t < B->nsurf && !found; ++t) {
        const surf_t *pf = &B->surf[t];
        if (strcmp(pf->type, "interblock")) continue;
        int pps = tag_partner_surface(pf->d[6]);
        if (surf_type(pf) != cn.ps || pps != cn.s || pf->d[6] - 1000 * pps != a) continue;
        if (cn.pb == a && pf == sf) continue;
        int e1 = (cn.pdir + 1) % 3, e2 = (cn.pdir + 2) % 3;
        cn.pr3 = pf->d[2 * cn.pdir];
        cn.pr1s = pf->d[2 * e1];
        cn.pn1 = pf->d[2 * e1 + 1] - cn.pr1s;
        cn.pr2s = pf->d[2 * e2];
        cn.pn2 = pf->d[2 * e2 + 1] - cn.pr2s;
        if (orient_patch(&cn, c, nodes, a) == 0) found = 1;
      }
      if (!found) {
        seterr(err, errlen, "block %d: no matching partner patch for interblock tag %d", a, sf->d[6]);
        return -1;
      }
      if (ndir(B, cn.pdir) < A->g) {
        seterr(err, errlen, "block %d: partner block %d is thinner than the ghost layers", a, cn.pb);
        return -1;
      }
      A->conn = (conn_t *)realloc(A->conn, sizeof(conn_t) * (size_t)(A->nconn + 1));
      A->conn[A->nconn++] = cn;
      int n1full = d1 == 0 ? A->ni : d1 == 1 ? A->nj : A->nk;
      for (int p2 = 0; p2 < cn.n2; ++p2)
        for (int p1 = 0; p1 < cn.n1; ++p1)
          A->isconn[cn.s - 1][(cn.r1s + p1) + (long)n1full * (cn.r2s + p2)] = 1;
    }
  }
  swap_geometry(c);
  return 0;
}

/* SwapStateSlice / SwapImplicitUpdate (src/procBlock.cpp:3050-3055,
 * src/utility.cpp:400-423): ghost cells of every connection take the
 * partner's interior values; the sources are interior cells, so the order
 * of the swaps does not matter */
/* which: 0 states, 1 implicit update, 2 cell velocity gradients
 * (SwapEddyViscAndGradients, src/gridLevel.cpp:385-387; Cartesian
 * components, so no rotation) */
static void swap_halo(orc_case *c, int which) {
  for (int a = 0; a < c->nb; ++a) {
    blk_t *A = &c->blk[a];
    for (int q = 0; q < A->nconn; ++q) {
      const conn_t *cn = &A->conn[q];
      const blk_t *B = &c->blk[cn->pb];
      double *dst = which == 2 ? A->vg : which ? A->x : A->W;
      const double *src = which == 2 ? B->vg : which ? B->x : B->W;
      int w = which == 2 ? 9 : NEQ;
      for (int l = 1; l <= A->g; ++l)
        for (int p2 = 0; p2 < cn->n2; ++p2)
          for (int p1 = 0; p1 < cn->n1; ++p1)
            memcpy(&dst[conn_ghost(cn, A, l, p1, p2) * w], &src[conn_source(cn, B, l, p1, p2) * w],
                   sizeof(double) * (size_t)w);
    }
  }
}

/* HyperplaneReorder: src/utility.cpp:377-398 */
void orc_hyperplane_order(int ni, int nj, int nk, int *out) {
  int np = ni + nj + nk - 2;
  long n = 0;
  for (int p = 0; p < np; ++p)
    for (int k = 0; k < nk; ++k)
      for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i)
          if (i + j + k == p) {
            out[3 * n] = i;
            out[3 * n + 1] = j;
            out[3 * n + 2] = k;
            ++n;
          }
}

/* ------------------------------------------------------------------------ */
/* input parsing (subset of src/input.cpp:167-643, src/inputStates.cpp)     */

static void trim(char *s) {
  /* Trim: strip whitespace, then cut at '#' (src/inputStates.cpp:231-252) */
  char *p = s;
  while (*p && isspace((unsigned char)*p)) ++p;
  memmove(s, p, strlen(p) + 1);
  size_t n = strlen(s);
  while (n > 0 && isspace((unsigned char)s[n - 1])) s[--n] = 0;
  char *h = strchr(s, '#');
  if (h) *h = 0;
  n = strlen(s);
  while (n > 0 && isspace((unsigned char)s[n - 1])) s[--n] = 0;
}

static void seterr(char *err, int errlen, const char *fmt, ...) {
  if (!err || errlen <= 0) return;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(err, (size_t)errlen, fmt, ap);
  va_end(ap);
}

/* split a <...> list body into top-level entries "name(...)" */
static int split_entries(const char *body, char out[][1024], int maxn) {
  int n = 0, depth = 0;
  const char *start = NULL;
  for (const char *p = body; *p; ++p) {
    if (*p == '(') {
      if (depth == 0) {
        const char *q = p;
        while (q > body && (isalnum((unsigned char)q[-1]) || q[-1] == '_')) --q;
        start = q;
      }
      ++depth;
    } else if (*p == ')') {
      --depth;
      if (depth == 0 && start && n < maxn) {
        size_t len = (size_t)(p - start + 1);
        if (len > 1023) len = 1023;
        memcpy(out[n], start, len);
        out[n][len] = 0;
        ++n;
        start = NULL;
      }
    }
  }
  return n;
}

static int get_param(const char *entry, const char *key, char *val, int vlen) {
  const char *lp = strchr(entry, '(');
  if (!lp) return 0;
  char buf[1024];
  snprintf(buf, sizeof buf, "%s", lp + 1);
  char *rp = strrchr(buf, ')');
  if (rp) *rp = 0;
  char *save = NULL;
  for (char *tok = strtok_r(buf, ";", &save); tok; tok = strtok_r(NULL, ";", &save)) {
    char t[512];
    snprintf(t, sizeof t, "%s", tok);
    trim(t);
    char *eq = strchr(t, '=');
    if (!eq) continue;
    *eq = 0;
    char k[128];
    snprintf(k, sizeof k, "%s", t);
    trim(k);
    if (strcmp(k, key) == 0) {
      snprintf(val, (size_t)vlen, "%s", eq + 1);
      trim(val);
      size_t L = strlen(val);
      if (L > 0 && val[L - 1] == ',') val[L - 1] = 0;
      return 1;
    }
  }
  return 0;
}

static v3 parse_vec(const char *s) {
  double a = 0, b = 0, c = 0;
  const char *p = strchr(s, '[');
  if (p) sscanf(p + 1, "%lf , %lf , %lf", &a, &b, &c);
  return V(a, b, c);
}

static void entry_name(const char *e, char *name, int n) {
  const char *lp = strchr(e, '(');
  int L = lp ? (int)(lp - e) : (int)strlen(e);
  if (L >= n) L = n - 1;
  memcpy(name, e, (size_t)L);
  name[L] = 0;
  trim(name);
}

/* fluid properties: read <name>.dat from the case directory (as
 * src/fluid.cpp:105-118 does), else the built-in values of
 * fluidDatabase/air.dat:14-15 and fluidDatabase/N2.dat */
/* fluid database entry (fluidDatabase/<name>.dat, src/fluid.cpp): n, molar
 * mass and the Sutherland coefficients of viscosity and conductivity */
static int fluid_props(const char *dir, const char *name, double *n, double *M, double *suth) {
  char path[1200];
  snprintf(path, sizeof path, "%s/%s.dat", dir, name);
  FILE *f = fopen(path, "r");
  if (f) {
    char line[512];
    int got = 0;
    static const char *sk[4] = {"sutherlandViscosityC1", "sutherlandViscosityS",
                                "sutherlandConductivityC1", "sutherlandConductivityS"};
    while (fgets(line, sizeof line, f)) {
      trim(line);
      char *c = strchr(line, ':');
      if (!c) continue;
      *c = 0;
      char key[128];
      snprintf(key, sizeof key, "%s", line);
      trim(key);
      if (strcmp(key, "n") == 0) { *n = strtod(c + 1, NULL); got |= 1; }
      if (strcmp(key, "molarMass") == 0) { *M = strtod(c + 1, NULL) / 1000.; got |= 2; }
      for (int q = 0; q < 4; ++q)
        if (strcmp(key, sk[q]) == 0) { suth[q] = strtod(c + 1, NULL); got |= 4 << q; }
    }
    fclose(f);
    if (got == 63) return 0;
  }
  if (strcmp(name, "air") == 0) {
    *n = 2.5; *M = 28.97 / 1000.;
    suth[0] = 1.458e-6; suth[1] = 110.4; suth[2] = 2.495e-3; suth[3] = 194.0;
    return 0;
  }
  if (strcmp(name, "N2") == 0) {
    *n = 2.5; *M = 28.0134 / 1000.;
    suth[0] = 1.4742e-06; suth[1] = 1.2846e+02; suth[2] = 2.6834e-03; suth[3] = 2.5615e+02;
    return 0;
  }
  return -1;
}

static int read_p3d(const char *path, double lref, int *nb, int **dims, v3 ***coords) {
  /* ReadP3dGrid: src/plot3d.cpp:363-445 */
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  int n = 0;
  if (fread(&n, sizeof(int), 1, f) != 1 || n <= 0 || n > 100000) { fclose(f); return -2; }
  int *d = (int *)malloc(sizeof(int) * 3 * (size_t)n);
  if (fread(d, sizeof(int), 3 * (size_t)n, f) != 3 * (size_t)n) { fclose(f); free(d); return -2; }
  v3 **c = (v3 **)calloc((size_t)n, sizeof(v3 *));
  for (int b = 0; b < n; ++b) {
    long np = (long)d[3 * b] * d[3 * b + 1] * d[3 * b + 2];
    c[b] = (v3 *)malloc(sizeof(v3) * (size_t)np);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)np);
    for (int comp = 0; comp < 3; ++comp) {
      if (fread(tmp, sizeof(double), (size_t)np, f) != (size_t)np) {
        fclose(f); free(tmp); return -2;
      }
      for (long q = 0; q < np; ++q) {
        double v = tmp[q] / lref;
        if (comp == 0) c[b][q].x = v;
        else if (comp == 1) c[b][q].y = v;
        else c[b][q].z = v;
      }
    }
    free(tmp);
  }
  fclose(f);
  *nb = n;
  *dims = d;
  *coords = c;
  return 0;
}

/* nearest neighbour over a point cloud (the reference uses a k-d tree,
 * src/kdtree.cpp:29-225; any exact nearest-neighbour search gives the same
 * match).  Uniform bucket grid. */
typedef struct {
  int np;
  v3 *pts;
  double lo[3], h[3];
  int nb[3];
  int *head, *next;
} nn_t;

static void nn_build(nn_t *t, v3 *pts, int np) {
  t->np = np;
  t->pts = pts;
  double hi[3];
  t->lo[0] = t->lo[1] = t->lo[2] = 1e300;
  hi[0] = hi[1] = hi[2] = -1e300;
  for (int q = 0; q < np; ++q) {
    double c[3] = {pts[q].x, pts[q].y, pts[q].z};
    for (int d = 0; d < 3; ++d) {
      if (c[d] < t->lo[d]) t->lo[d] = c[d];
      if (c[d] > hi[d]) hi[d] = c[d];
    }
  }
  double m = cbrt((double)np);
  for (int d = 0; d < 3; ++d) {
    int nbd = (int)m + 1;
    if (nbd > 256) nbd = 256;
    t->nb[d] = nbd;
    double ext = hi[d] - t->lo[d];
    t->h[d] = ext > 0 ? ext / nbd * 1.000001 : 1.0;
  }
  long nbt = (long)t->nb[0] * t->nb[1] * t->nb[2];
  t->head = (int *)malloc(sizeof(int) * (size_t)nbt);
  t->next = (int *)malloc(sizeof(int) * (size_t)np);
  for (long q = 0; q < nbt; ++q) t->head[q] = -1;
  for (int q = 0; q < np; ++q) {
    double c[3] = {pts[q].x, pts[q].y, pts[q].z};
    int bi[3];
    for (int d = 0; d < 3; ++d) {
      bi[d] = (int)((c[d] - t->lo[d]) / t->h[d]);
      if (bi[d] < 0) bi[d] = 0;
      if (bi[d] >= t->nb[d]) bi[d] = t->nb[d] - 1;
    }
    long key = bi[0] + (long)bi[1] * t->nb[0] + (long)bi[2] * t->nb[0] * t->nb[1];
    t->next[q] = t->head[key];
    t->head[key] = q;
  }
}

static int nn_query(const nn_t *t, v3 x) {
  double c[3] = {x.x, x.y, x.z};
  int bi[3];
  for (int d = 0; d < 3; ++d) {
    bi[d] = (int)floor((c[d] - t->lo[d]) / t->h[d]);
    if (bi[d] < 0) bi[d] = 0;
    if (bi[d] >= t->nb[d]) bi[d] = t->nb[d] - 1;
  }
  int best = -1;
  double bd = 1e300;
  int maxr = t->nb[0] + t->nb[1] + t->nb[2];
  for (int r = 0; r <= maxr; ++r) {
    for (int a = bi[0] - r; a <= bi[0] + r; ++a)
      for (int bq = bi[1] - r; bq <= bi[1] + r; ++bq)
        for (int cq = bi[2] - r; cq <= bi[2] + r; ++cq) {
          if (a < 0 || bq < 0 || cq < 0 || a >= t->nb[0] || bq >= t->nb[1] || cq >= t->nb[2])
            continue;
          if (abs(a - bi[0]) != r && abs(bq - bi[1]) != r && abs(cq - bi[2]) != r) continue;
          long key = a + (long)bq * t->nb[0] + (long)cq * t->nb[0] * t->nb[1];
          for (int q = t->head[key]; q >= 0; q = t->next[q]) {
            v3 dv = vsub(t->pts[q], x);
            double dd = vdot(dv, dv);
            if (dd < bd) { bd = dd; best = q; }
          }
        }
    if (best >= 0) {
      double hmin = t->h[0];
      if (t->h[1] < hmin) hmin = t->h[1];
      if (t->h[2] < hmin) hmin = t->h[2];
      double reach = r * hmin;
      if (reach * reach >= bd) break;
    }
  }
  return best;
}

static void free_case(orc_case *c);

static int alloc_block(blk_t *b) {
  long nc = ncells_padded(b), np = ncells(b);
  b->vol = (double *)calloc((size_t)nc, sizeof(double));
  b->ctr = (v3 *)calloc((size_t)nc, sizeof(v3));
  for (int d = 0; d < 3; ++d) {
    b->fa[d] = (uv3 *)calloc((size_t)nfaces_padded(b, d), sizeof(uv3));
    b->fc[d] = (v3 *)calloc((size_t)nfaces_padded(b, d), sizeof(v3));
    b->cw[d] = (double *)calloc((size_t)nc, sizeof(double));
  }
  b->W = (double *)calloc((size_t)nc * NEQ, sizeof(double));
  b->x = (double *)calloc((size_t)nc * NEQ, sizeof(double));
  b->T = (double *)calloc((size_t)nc, sizeof(double));
  b->vg = (double *)calloc((size_t)nc * 9, sizeof(double));
  b->mu = (double *)calloc((size_t)nc, sizeof(double));
  b->R = (double *)calloc((size_t)np * NEQ, sizeof(double));
  b->sr = (double *)calloc((size_t)np, sizeof(double));
  b->dt = (double *)calloc((size_t)np, sizeof(double));
  b->A = (double *)calloc((size_t)np, sizeof(double));
  b->Ainv = (double *)calloc((size_t)np, sizeof(double));
  b->Ab = (double *)calloc((size_t)np * 25, sizeof(double));
  b->Abi = (double *)calloc((size_t)np * 25, sizeof(double));
  b->consN = (double *)calloc((size_t)np * NEQ, sizeof(double));
  b->consNm1 = (double *)calloc((size_t)np * NEQ, sizeof(double));
  b->order = (int *)malloc(sizeof(int) * 3 * (size_t)np);
  return (b->vol && b->W && b->order) ? 0 : -1;
}

static const bcstate_t *find_periodic(const orc_case *c, int tag) {
  for (int q = 0; q < c->nbc; ++q)
    if (!strcmp(c->bcs[q].kind, "periodic") && (c->bcs[q].startTag == tag || c->bcs[q].endTag == tag))
      return &c->bcs[q];
  return NULL;
}

static const bcstate_t *find_bc(const orc_case *c, int tag) {
  for (int q = 0; q < c->nbc; ++q)
    if (c->bcs[q].tag == tag) return &c->bcs[q];
  return NULL;
}

orc_case *orc_load(const char *inp_path, char *err, int errlen) {
  FILE *f = fopen(inp_path, "r");
  if (!f) {
    seterr(err, errlen, "cannot open %s", inp_path);
    return NULL;
  }
  orc_case *c = (orc_case *)calloc(1, sizeof(orc_case));
  /* defaults: src/input.cpp:50-103 */
  c->iterations = 1;
  c->rRef = -1.0;
  c->tRef = -1.0;
  c->lRef = 1.0;
  c->dtIn = -1.0;
  c->cflMax = 1.0;
  c->cflStep = 0.0;
  c->cflStart = 1.0;
  c->theta = 1.0;
  c->zeta = 0.0;
  c->relax = 1.0;
  c->sweeps = 1;
  c->nonlinear = 1;
  c->dualCFL = -1.0;
  c->kappa = -2.0;
  c->recon = 0;
  c->limiter = 0;
  c->flux = 0;
  c->timeint = 0;
  c->solver = 0;
  char fluidName[64] = "air";
  {
    const char *sl = strrchr(inp_path, '/');
    if (sl) {
      size_t L = (size_t)(sl - inp_path);
      if (L >= sizeof c->dir) L = sizeof c->dir - 1;
      memcpy(c->dir, inp_path, L);
      c->dir[L] = 0;
    } else {
      strcpy(c->dir, ".");
    }
  }
  char line[4096];
  int readingBC = 0, nbcBlocks = 0, curBlk = 0, surfLeft = -1, surfIdx = 0;
  int *nsurfPer = NULL;
  surf_t **surfsPer = NULL;
  while (fgets(line, sizeof line, f)) {
    trim(line);
    if (!line[0]) continue;
    if (readingBC) {
      if (surfLeft < 0) {
        int a, b2, c3;
        if (sscanf(line, "%d %d %d", &a, &b2, &c3) != 3) {
          seterr(err, errlen, "bad BC surface count line: %s", line);
          goto fail;
        }
        nsurfPer[curBlk] = a + b2 + c3;
        surfsPer[curBlk] = (surf_t *)calloc((size_t)(a + b2 + c3), sizeof(surf_t));
        surfLeft = a + b2 + c3;
        surfIdx = 0;
      } else {
        surf_t *s = &surfsPer[curBlk][surfIdx];
        if (sscanf(line, "%31s %d %d %d %d %d %d %d", s->type, &s->d[0], &s->d[1],
                   &s->d[2], &s->d[3], &s->d[4], &s->d[5], &s->d[6]) != 8) {
          seterr(err, errlen, "bad BC surface line: %s", line);
          goto fail;
        }
        ++surfIdx;
        --surfLeft;
      }
      if (surfLeft == 0) {
        ++curBlk;
        surfLeft = -1;
        if (curBlk == nbcBlocks) readingBC = 0;
      }
      continue;
    }
    char *colon = strchr(line, ':');
    if (!colon) continue;
    *colon = 0;
    char key[128];
    snprintf(key, sizeof key, "%s", line);
    trim(key);
    char val[8192];
    snprintf(val, sizeof val, "%s", colon + 1);
    trim(val);
    if (val[0] == '<' && !strchr(val, '>')) {
      char more[4096];
      while (fgets(more, sizeof more, f)) {
        trim(more);
        strncat(val, " ", sizeof val - strlen(val) - 1);
        strncat(val, more, sizeof val - strlen(val) - 1);
        if (strchr(more, '>')) break;
      }
    }
    if (!strcmp(key, "gridName")) snprintf(c->grid, sizeof c->grid, "%s", val);
    else if (!strcmp(key, "timeStep")) c->dtIn = strtod(val, NULL);
    else if (!strcmp(key, "iterations")) c->iterations = atoi(val);
    else if (!strcmp(key, "referenceDensity")) c->rRef = strtod(val, NULL);
    else if (!strcmp(key, "referenceTemperature")) c->tRef = strtod(val, NULL);
    else if (!strcmp(key, "referenceLength")) c->lRef = strtod(val, NULL);
    else if (!strcmp(key, "timeIntegration")) {
      if (!strcmp(val, "explicitEuler")) c->timeint = 0;
      else if (!strcmp(val, "implicitEuler")) { c->timeint = 1; c->theta = 1.0; c->zeta = 0.0; }
      else if (!strcmp(val, "crankNicholson")) { c->timeint = 2; c->theta = 0.5; c->zeta = 0.0; }
      else if (!strcmp(val, "bdf2")) { c->timeint = 3; c->theta = 1.0; c->zeta = 0.5; }
      else { seterr(err, errlen, "unsupported timeIntegration %s", val); goto fail; }
    } else if (!strcmp(key, "faceReconstruction")) {
      c->recon = 1;
      if (!strcmp(val, "upwind")) c->kappa = -1.0;
      else if (!strcmp(val, "fromm")) c->kappa = 0.0;
      else if (!strcmp(val, "quick")) c->kappa = 0.5;
      else if (!strcmp(val, "central")) c->kappa = 1.0;
      else if (!strcmp(val, "thirdOrder")) c->kappa = 1.0 / 3.0;
      else if (!strcmp(val, "constant")) c->recon = 0;
      else if (!strcmp(val, "weno")) c->recon = 2;
      else if (!strcmp(val, "wenoZ")) c->recon = 3;
      else { seterr(err, errlen, "unsupported faceReconstruction %s", val); goto fail; }
    } else if (!strcmp(key, "limiter")) {
      if (!strcmp(val, "none")) c->limiter = 0;
      else if (!strcmp(val, "vanAlbada")) c->limiter = 1;
      else if (!strcmp(val, "minmod")) c->limiter = 2;
      else { seterr(err, errlen, "unsupported limiter %s", val); goto fail; }
    } else if (!strcmp(key, "equationSet")) {
      if (!strcmp(val, "euler")) c->viscous = 0;
      else if (!strcmp(val, "navierStokes")) c->viscous = 1;
      else { seterr(err, errlen, "unsupported equationSet %s", val); goto fail; }
    } else if (!strcmp(key, "viscousFaceReconstruction")) {
      if (!strcmp(val, "central")) c->viscRecon4 = 0;
      else if (!strcmp(val, "centralFourth")) c->viscRecon4 = 1;
      else { seterr(err, errlen, "unsupported viscousFaceReconstruction %s", val); goto fail; }
    } else if (!strcmp(key, "turbulenceModel")) {
      if (strcmp(val, "none")) { seterr(err, errlen, "unsupported turbulenceModel %s", val); goto fail; }
    } else if (!strcmp(key, "matrixSolver")) {
      /* src/input.cpp:713-720, 845-856 */
      if (!strcmp(val, "lusgs")) c->solver = 0, c->block = 0;
      else if (!strcmp(val, "dplur")) c->solver = 1, c->block = 0;
      else if (!strcmp(val, "blusgs")) c->solver = 0, c->block = 1;
      else if (!strcmp(val, "bdplur")) c->solver = 1, c->block = 1;
      else { seterr(err, errlen, "unsupported matrixSolver %s", val); goto fail; }
    } else if (!strcmp(key, "matrixSweeps")) c->sweeps = atoi(val);
    else if (!strcmp(key, "matrixRelaxation")) c->relax = strtod(val, NULL);
    else if (!strcmp(key, "nonlinearIterations")) c->nonlinear = atoi(val);
    else if (!strcmp(key, "cflMax")) c->cflMax = strtod(val, NULL);
    else if (!strcmp(key, "cflStep")) c->cflStep = strtod(val, NULL);
    else if (!strcmp(key, "cflStart")) c->cflStart = strtod(val, NULL);
    else if (!strcmp(key, "dualTimeCFL")) c->dualCFL = strtod(val, NULL);
    else if (!strcmp(key, "inviscidFlux")) {
      if (!strcmp(val, "roe")) c->flux = 0;
      else if (!strcmp(val, "ausm")) c->flux = 1;
      else { seterr(err, errlen, "unsupported inviscidFlux %s", val); goto fail; }
    } else if (!strcmp(key, "inviscidFluxJacobian")) {
      /* src/input.cpp:82, 372-373; OffDiagonal src/fluxJacobian.cpp:196-238 */
      if (!strcmp(val, "rusanov")) c->roeJac = 0;
      else if (!strcmp(val, "approximateRoe")) c->roeJac = 1;
      else { seterr(err, errlen, "unsupported inviscidFluxJacobian %s", val); goto fail; }
    } else if (!strcmp(key, "multigridLevels")) {
      if (atoi(val) != 1) { seterr(err, errlen, "multigrid is not supported"); goto fail; }
    } else if (!strcmp(key, "fluids")) {
      char ents[8][1024];
      int ne = split_entries(val, ents, 8);
      if (ne != 1) { seterr(err, errlen, "only single-species fluids are supported"); goto fail; }
      char nm[64];
      if (get_param(ents[0], "name", nm, sizeof nm)) snprintf(fluidName, sizeof fluidName, "%s", nm);
    } else if (!strcmp(key, "initialConditions")) {
      char ents[16][1024];
      int ne = split_entries(val, ents, 16);
      for (int q = 0; q < ne && c->nic < 16; ++q) {
        char v[512];
        c->ic[c->nic].tag = get_param(ents[q], "tag", v, sizeof v) ? atoi(v) : -1;
        if (get_param(ents[q], "file", v, sizeof v)) {
          c->ic[c->nic].fromFile = 1;
          snprintf(c->ic[c->nic].file, sizeof c->ic[c->nic].file, "%s", v);
        } else {
          if (get_param(ents[q], "pressure", v, sizeof v)) c->ic[c->nic].p = strtod(v, NULL);
          if (get_param(ents[q], "density", v, sizeof v)) c->ic[c->nic].rho = strtod(v, NULL);
          if (get_param(ents[q], "velocity", v, sizeof v)) c->ic[c->nic].vel = parse_vec(v);
        }
        c->nic++;
      }
    } else if (!strcmp(key, "boundaryStates")) {
      char ents[MAXBC][1024];
      int ne = split_entries(val, ents, MAXBC);
      for (int q = 0; q < ne && c->nbc < MAXBC; ++q) {
        bcstate_t *s = &c->bcs[c->nbc++];
        char v[512];
        entry_name(ents[q], s->kind, sizeof s->kind);
        s->tag = get_param(ents[q], "tag", v, sizeof v) ? atoi(v) : -1;
        if (get_param(ents[q], "pressure", v, sizeof v)) s->p = strtod(v, NULL);
        if (get_param(ents[q], "density", v, sizeof v)) s->rho = strtod(v, NULL);
        if (get_param(ents[q], "velocity", v, sizeof v)) s->vel = parse_vec(v);
        if (get_param(ents[q], "p0", v, sizeof v)) s->p0 = strtod(v, NULL);
        if (get_param(ents[q], "t0", v, sizeof v)) s->t0 = strtod(v, NULL);
        if (get_param(ents[q], "direction", v, sizeof v)) s->dir = parse_vec(v);
        if (get_param(ents[q], "nonreflecting", v, sizeof v)) s->nonreflecting = !strcmp(v, "true");
        if (get_param(ents[q], "temperature", v, sizeof v)) { s->isothermal = 1; s->twall = strtod(v, NULL); }
        if (get_param(ents[q], "heatFlux", v, sizeof v)) {
          if (s->isothermal) { seterr(err, errlen, "viscousWall: give a temperature or a heat flux, not both"); goto fail; }
          s->qwall = strtod(v, NULL);
          s->heatflux = s->qwall != 0.0;
        }
        if (get_param(ents[q], "startTag", v, sizeof v)) s->startTag = atoi(v);
        if (get_param(ents[q], "endTag", v, sizeof v)) s->endTag = atoi(v);
        if (get_param(ents[q], "translation", v, sizeof v)) s->trans = parse_vec(v);
        if (!strcmp(s->kind, "periodic") && (get_param(ents[q], "axis", v, sizeof v) ||
                                             get_param(ents[q], "rotation", v, sizeof v))) {
          seterr(err, errlen, "rotational periodicity is not supported");
          goto fail;
        }
        if (get_param(ents[q], "wallTreatment", v, sizeof v) && strcmp(v, "lowRe")) {
          seterr(err, errlen, "wall laws are not supported");
          goto fail;
        }
        if (s->nonreflecting) { seterr(err, errlen, "nonreflecting BCs are not supported"); goto fail; }
      }
    } else if (!strcmp(key, "boundaryConditions")) {
      nbcBlocks = atoi(val);
      nsurfPer = (int *)calloc((size_t)nbcBlocks, sizeof(int));
      surfsPer = (surf_t **)calloc((size_t)nbcBlocks, sizeof(surf_t *));
      readingBC = nbcBlocks > 0;
      curBlk = 0;
      surfLeft = -1;
    }
    /* other keys (output, decomposition, models at their defaults) are
     * accepted and ignored */
  }
  fclose(f);
  f = NULL;

  if (c->rRef <= 0 || c->tRef <= 0) { seterr(err, errlen, "reference density/temperature missing"); goto fail2; }
  if (fluid_props(c->dir, fluidName, &c->fluidN, &c->fluidM, c->suth)) {
    seterr(err, errlen, "unknown fluid %s", fluidName);
    goto fail2;
  }
  c->implicit = (c->timeint >= 1);
  c->multilevel = (c->timeint == 3);
  c->storeN = c->implicit;
  c->initMatrix = (c->solver == 1 || c->sweeps > 1);
  c->block = c->block && c->implicit; /* IsBlockMatrix: src/input.cpp:713-720 */
  /* RoeOffDiagonal forms a flux change, not a matrix, so the block solvers
   * cannot use it; with the viscous terms the reference's call passes the
   * blending coefficient (0 in laminar flow) as the centre distance
   * (src/fluxJacobian.cpp:229-230 against :240-245) */
  if (c->roeJac && c->block) { seterr(err, errlen, "approximateRoe with a block matrix solver is not supported"); goto fail2; }
  if (c->roeJac && c->viscous) { seterr(err, errlen, "approximateRoe with the viscous equations is not supported"); goto fail2; }
  if (c->timeint == 0) c->nonlinear = 1; /* src/input.cpp:928-934 */

  /* reference speed of sound and nondimensional gas constant:
   * src/input.cpp:616-621, src/fluid.cpp:89-103 */
  {
    double Ru = 8.3144598;
    double Rdim = Ru / c->fluidM;
    double gamma = (c->fluidN + 1) / c->fluidN;
    double a = 0.0;
    a += 1.0 * gamma * Rdim * c->tRef;
    c->aRef = sqrt(a);
    double RuNd = Ru / (c->aRef * c->aRef * c->rRef / (c->tRef * pow(c->lRef, 3.0)));
    double Mnd = c->fluidM / (c->rRef / pow(c->lRef, 3.0));
    c->gas.R = RuNd / Mnd;
    c->gas.n = c->fluidN;
    c->gamma = orc_gamma(&c->gas);
  }
  /* state data nondimensionalisation: src/inputStates.cpp:464-473, 590-599,
   * 674-681 (stagnationInlet direction is left as given: Normalize() returns
   * a copy, src/inputStates.cpp:594) */
  for (int q = 0; q < c->nic; ++q) {
    if (c->ic[q].fromFile) continue;
    c->ic[q].vel = vdiv(c->ic[q].vel, c->aRef);
    c->ic[q].rho /= c->rRef;
    c->ic[q].p /= c->rRef * c->aRef * c->aRef;
  }
  /* Sutherland transport and its nondimensional scaling:
   * src/transport.cpp:29-66, include/transport.hpp:40-44 */
  {
    c->muRef = c->suth[0] * pow(c->tRef, 1.5) / (c->tRef + c->suth[1]);
    c->kNonDim = (c->aRef * c->aRef * c->muRef) / c->tRef;
    c->vscale = c->muRef / (c->rRef * c->aRef * c->lRef);
    /* input::ViscousCFLCoefficient, src/input.cpp:1110-1118 */
    c->viscCoeff = c->kappa == 1.0 ? 4.0 : (c->kappa == -2.0 ? 2.0 : 1.0);
  }
  for (int q = 0; q < c->nbc; ++q) {
    bcstate_t *s = &c->bcs[q];
    if (!strcmp(s->kind, "periodic")) { /* src/inputStates.cpp:883-894 */
      s->trans = vdiv(s->trans, c->lRef);
      continue;
    }
    if (!strcmp(s->kind, "viscousWall")) { /* src/inputStates.cpp:775-783 */
      s->vel = vdiv(s->vel, c->aRef);
      s->twall /= c->tRef;
      s->qwall /= pow(c->aRef / c->lRef, 3.0);
      continue;
    }
    if (!strcmp(s->kind, "stagnationInlet")) {
      s->p0 /= c->rRef * c->aRef * c->aRef;
      s->t0 /= c->tRef;
    } else if (!strcmp(s->kind, "pressureOutlet")) {
      s->p /= c->rRef * c->aRef * c->aRef;
    } else {
      s->vel = vdiv(s->vel, c->aRef);
      s->rho /= c->rRef;
      s->p /= c->rRef * c->aRef * c->aRef;
    }
  }

  /* grid */
  {
    char path[2048];
    snprintf(path, sizeof path, "%s/%s.xyz", c->dir, c->grid);
    int nbg = 0, *dims = NULL;
    v3 **coords = NULL;
    if (read_p3d(path, c->lRef, &nbg, &dims, &coords)) {
      seterr(err, errlen, "cannot read grid %s", path);
      goto fail2;
    }
    if (nbg != nbcBlocks) {
      seterr(err, errlen, "grid has %d blocks but input has %d BC blocks", nbg, nbcBlocks);
      goto fail2;
    }
    c->nb = nbg;
    c->blk = (blk_t *)calloc((size_t)nbg, sizeof(blk_t));
    int g = (c->recon == 0) ? 1 : (c->recon == 1 ? 2 : 3); /* src/input.cpp:1127-1143 */
    if (c->viscRecon4 && g < 2) g = 2;
    c->totalCells = 0;
    for (int bq = 0; bq < nbg; ++bq) {
      blk_t *b = &c->blk[bq];
      b->ni = dims[3 * bq] - 1;
      b->nj = dims[3 * bq + 1] - 1;
      b->nk = dims[3 * bq + 2] - 1;
      b->g = g;
      b->NI = b->ni + 2 * g;
      b->NJ = b->nj + 2 * g;
      b->NK = b->nk + 2 * g;
      c->totalCells += (double)ncells(b);
      if (alloc_block(b)) { seterr(err, errlen, "out of memory"); goto fail2; }
      b->nsurf = nsurfPer[bq];
      b->surf = surfsPer[bq];
      surfsPer[bq] = NULL;

      qsort(b->surf, (size_t)b->nsurf, sizeof(surf_t), surf_cmp);
      if (build_geometry(b, coords[bq])) { seterr(err, errlen, "negative volume in block %d", bq); goto fail2; }
      ghost_geometry(b);
      orc_hyperplane_order(b->ni, b->nj, b->nk, b->order);
    }
    /* src/gridLevel.cpp:67-74 (geometry swap), then src/main.cpp:169
     * (cell widths, ghost cells included) */
    if (connect_blocks(c, (const v3 *const *)coords, err, errlen)) goto fail2;
    for (int bq = 0; bq < nbg; ++bq) {
      cell_widths(&c->blk[bq]);
      free(coords[bq]);
    }
    free(coords);
    free(dims);
  }

  /* initial conditions: procBlock::InitializeStates, src/procBlock.cpp:280-355 */
  for (int bq = 0; bq < c->nb; ++bq) {
    blk_t *b = &c->blk[bq];
    int use = -1;
    for (int q = 0; q < c->nic; ++q) {
      if (c->ic[q].tag == -1 && use < 0) use = q;
      if (c->ic[q].tag == bq) { use = q; break; }
    }
    if (use < 0) { seterr(err, errlen, "no initial condition for block %d", bq); goto fail2; }
    if (!c->ic[use].fromFile) {
      for (int k = 0; k < b->nk; ++k)
        for (int j = 0; j < b->nj; ++j)
          for (int i = 0; i < b->ni; ++i) {
            double *W = &b->W[cidx(b, i, j, k) * NEQ];
            W[0] = 1.0 * c->ic[use].rho;
            W[1] = c->ic[use].vel.x;
            W[2] = c->ic[use].vel.y;
            W[3] = c->ic[use].vel.z;
            W[4] = c->ic[use].p;
          }
    } else {
      /* CalcTreeFromCloud: src/utility.cpp:513-605 */
      char path[2048];
      snprintf(path, sizeof path, "%s/%s", c->dir, c->ic[use].file);
      FILE *cf = fopen(path, "r");
      if (!cf) { seterr(err, errlen, "cannot open IC file %s", path); goto fail2; }
      int np = 0;
      char sp[256];
      if (fscanf(cf, "%d", &np) != 1 || np <= 0) { fclose(cf); seterr(err, errlen, "bad IC file"); goto fail2; }
      if (fscanf(cf, "%255s", sp) != 1) { fclose(cf); seterr(err, errlen, "bad IC file"); goto fail2; }
      v3 *pts = (v3 *)malloc(sizeof(v3) * (size_t)np);
      double *st = (double *)malloc(sizeof(double) * NEQ * (size_t)np);
      for (int q = 0; q < np; ++q) {
        double x, y, z, r, u, v, w, p, tk, om, mf;
        if (fscanf(cf, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &x, &y, &z, &r, &u,
                   &v, &w, &p, &tk, &om, &mf) != 11) {
          fclose(cf); free(pts); free(st);
          seterr(err, errlen, "bad IC line %d", q);
          goto fail2;
        }
        pts[q] = V(x / c->lRef, y / c->lRef, z / c->lRef);
        double rho = r / c->rRef;
        st[NEQ * q + 0] = rho * mf;
        st[NEQ * q + 1] = u / c->aRef;
        st[NEQ * q + 2] = v / c->aRef;
        st[NEQ * q + 3] = w / c->aRef;
        st[NEQ * q + 4] = p / (c->rRef * c->aRef * c->aRef);
      }
      fclose(cf);
      nn_t t;
      nn_build(&t, pts, np);
      for (int k = 0; k < b->nk; ++k)
        for (int j = 0; j < b->nj; ++j)
          for (int i = 0; i < b->ni; ++i) {
            int id = nn_query(&t, b->ctr[cidx(b, i, j, k)]);
            memcpy(&b->W[cidx(b, i, j, k) * NEQ], &st[NEQ * id], sizeof(double) * NEQ);
          }
      free(t.head);
      free(t.next);
      free(pts);
      free(st);
    }
  }
  free(nsurfPer);
  free(surfsPer);
  c->matres_on = 1;
  return c;

fail:
  if (f) fclose(f);
fail2:
  if (surfsPer) {
    for (int q = 0; q < nbcBlocks; ++q) free(surfsPer[q]);
    free(surfsPer);
  }
  free(nsurfPer);
  free_case(c);
  return NULL;
}

static void free_case(orc_case *c) {
  if (!c) return;
  for (int bq = 0; bq < c->nb && c->blk; ++bq) {
    blk_t *b = &c->blk[bq];
    free(b->vol); free(b->ctr);
    for (int d = 0; d < 3; ++d) { free(b->fa[d]); free(b->fc[d]); free(b->cw[d]); }
    free(b->W); free(b->x); free(b->T); free(b->vg); free(b->mu); free(b->R); free(b->sr); free(b->dt);
    free(b->A); free(b->Ainv); free(b->Ab); free(b->Abi); free(b->consN); free(b->consNm1);
    free(b->order); free(b->surf); free(b->conn); free(b->wd);
    for (int q = 0; q < 6; ++q) free(b->isconn[q]);
  }
  free(c->blk);
  free(c);
}
void orc_free(orc_case *c) { free_case(c); }

/* ------------------------------------------------------------------------ */
/* boundary conditions                                                      */

/* ExtrapolateHoldMixture: src/ghostStates.cpp:691-709 */
static void extrap_hold(const double *bnd, double factor, const double *in,
                        double *out) {
  double bndRho = 0.0 + bnd[0];
  double bndMf = bnd[0] / bndRho;
  double intRho = 0.0 + in[0];
  double ghostRho = factor * bndRho - intRho;
  if (ghostRho <= 0.0) {
    memcpy(out, bnd, sizeof(double) * NEQ);
    return;
  }
  for (int e = 0; e < NEQ; ++e) out[e] = factor * bnd[e] - in[e];
  out[0] = dmax(ghostRho * bndMf, 0.0);
}

/* GetGhostState: src/ghostStates.cpp:62-689 (inviscid branches) */
static int ghost_state(const orc_case *c, const double *in, const char *bc,
                       v3 area, int surf, int tag, int layer, double *ghost) {
  const orc_gas *g = &c->gas;
  memcpy(ghost, in, sizeof(double) * NEQ);
  int isLower = (surf % 2 == 1);
  v3 n = isLower ? vscl(-1.0, area) : area;
  if (!strcmp(bc, "slipWall")) {
    v3 vel = V(in[1], in[2], in[3]);
    double vn = vdot(vel, n);
    v3 gv = vsub(vel, vscl(vn, vscl(2.0, n)));
    ghost[1] = gv.x;
    ghost[2] = gv.y;
    ghost[3] = gv.z;
  } else if (!strcmp(bc, "characteristic")) {
    const bcstate_t *s = find_bc(c, tag);
    if (!s) return -1;
    double fs[NEQ];
    fs[0] = s->rho * 1.0;
    fs[1] = s->vel.x;
    fs[2] = s->vel.y;
    fs[3] = s->vel.z;
    fs[4] = s->p;
    double velIntNorm = vdot(V(in[1], in[2], in[3]), n);
    double sosInt = orc_sos(g, in);
    double machInt = fabs(velIntNorm) / sosInt;
    if (machInt >= 1.0 && velIntNorm < 0.0) {
      memcpy(ghost, fs, sizeof fs);
    } else if (machInt >= 1.0 && velIntNorm >= 0.0) {
      /* interior values */
    } else if (machInt < 1.0 && velIntNorm < 0.0) {
      double rhoSoSInt = (0.0 + in[0]) * sosInt;
      v3 velDiff = vsub(V(fs[1], fs[2], fs[3]), V(in[1], in[2], in[3]));
      ghost[4] = 0.5 * (fs[4] + in[4] - rhoSoSInt * vdot(n, velDiff));
      double dp = fs[4] - ghost[4];
      double rho = (0.0 + fs[0]) - dp / (sosInt * sosInt);
      ghost[0] = rho * (fs[0] / (0.0 + fs[0]));
      ghost[1] = fs[1] - n.x * dp / rhoSoSInt;
      ghost[2] = fs[2] - n.y * dp / rhoSoSInt;
      ghost[3] = fs[3] - n.z * dp / rhoSoSInt;
    } else if (machInt < 1.0 && velIntNorm >= 0.0) {
      double rhoSoSInt = (0.0 + in[0]) * sosInt;
      double dp = in[4] - fs[4];
      double rho = (0.0 + in[0]) - dp / (sosInt * sosInt);
      ghost[0] = rho * (in[0] / (0.0 + in[0]));
      ghost[1] = in[1] + n.x * dp / rhoSoSInt;
      ghost[2] = in[2] + n.y * dp / rhoSoSInt;
      ghost[3] = in[3] + n.z * dp / rhoSoSInt;
      ghost[4] = fs[4];
    } else {
      return -2;
    }
    double t1[NEQ];
    extrap_hold(ghost, 2.0, in, t1);
    memcpy(ghost, t1, sizeof t1);
    if (layer > 1) {
      extrap_hold(ghost, (double)layer, in, t1);
      memcpy(ghost, t1, sizeof t1);
    }
  } else if (!strcmp(bc, "supersonicInflow")) {
    const bcstate_t *s = find_bc(c, tag);
    if (!s) return -1;
    ghost[0] = 0.0;
    ghost[0] = s->rho * 1.0;
    ghost[1] = s->vel.x;
    ghost[2] = s->vel.y;
    ghost[3] = s->vel.z;
    ghost[4] = s->p;
  } else if (!strcmp(bc, "supersonicOutflow")) {
    if (layer > 1)
      for (int e = 0; e < NEQ; ++e) ghost[e] = (double)layer * ghost[e] - in[e];
  } else if (!strcmp(bc, "stagnationInlet")) {
    const bcstate_t *s = find_bc(c, tag);
    if (!s) return -1;
    double gm1 = c->gamma - 1.0;
    v3 vel = V(in[1], in[2], in[3]);
    double sosI = orc_sos(g, in);
    double rNeg = vdot(vel, n) - 2.0 * sosI / gm1;
    double cosTheta = -1.0 * vdot(vel, n) / vmag(vel);
    double stagSoSsq = pow(sosI, 2.0) + 0.5 * gm1 * vdot(vel, vel);
    double sosB = -1.0 * rNeg * gm1 / (gm1 * cosTheta * cosTheta + 2.0) *
                  (1.0 + cosTheta * sqrt((gm1 * cosTheta * cosTheta + 2.0) *
                                             stagSoSsq / (gm1 * rNeg * rNeg) -
                                         0.5 * gm1));
    double tb = s->t0 * (sosB * sosB / stagSoSsq);
    double pb = s->p0 * pow(sosB * sosB / stagSoSsq, c->gamma / gm1);
    double vbMag = sqrt(2.0 / gm1 * (s->t0 - tb));
    /* DensityTP: src/eos.cpp:111-115 with mf = {1.0} */
    double Rmix = 0.0 + 1.0 * g->R;
    double rhoGhost = pb / (Rmix * tb);
    ghost[0] = 0.0;
    ghost[0] = rhoGhost * 1.0;
    ghost[1] = vbMag * s->dir.x;
    ghost[2] = vbMag * s->dir.y;
    ghost[3] = vbMag * s->dir.z;
    ghost[4] = pb;
    double t1[NEQ];
    extrap_hold(ghost, 2.0, in, t1);
    memcpy(ghost, t1, sizeof t1);
    if (layer > 1) {
      extrap_hold(ghost, (double)layer, in, t1);
      memcpy(ghost, t1, sizeof t1);
    }
  } else if (!strcmp(bc, "pressureOutlet")) {
    const bcstate_t *s = find_bc(c, tag);
    if (!s) return -1;
    double pb = s->p;
    double sosI = orc_sos(g, in);
    double rhoSoSInt = (0.0 + in[0]) * sosI;
    ghost[4] = pb;
    double dp = in[4] - ghost[4];
    double rho = (0.0 + in[0]) - dp / (sosI * sosI);
    ghost[0] = rho * (in[0] / (0.0 + in[0]));
    ghost[1] = in[1] + n.x * dp / rhoSoSInt;
    ghost[2] = in[2] + n.y * dp / rhoSoSInt;
    ghost[3] = in[3] + n.z * dp / rhoSoSInt;
    if (vdot(V(ghost[1], ghost[2], ghost[3]), n) / orc_sos(g, ghost) >= 1.0)
      memcpy(ghost, in, sizeof(double) * NEQ);
    for (int e = 0; e < NEQ; ++e) ghost[e] = 2.0 * ghost[e] - in[e];
    if (layer > 1)
      for (int e = 0; e < NEQ; ++e) ghost[e] = (double)layer * ghost[e] - in[e];
  } else {
    return -3;
  }
  return 0;
}

/* procBlock::AssignInviscidGhostCells: src/procBlock.cpp:2449-2529 */
static void assign_ghosts(orc_case *c, blk_t *b) {
  for (int layer = 1; layer <= b->g; ++layer) {
    for (int s = 0; s < b->nsurf; ++s) {
      const surf_t *sf = &b->surf[s];
      int st = surf_type(sf);
      int dir = (st - 1) / 2;
      int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
      int r1s = sf->d[2 * d1], r1e = sf->d[2 * d1 + 1];
      int r2s = sf->d[2 * d2], r2e = sf->d[2 * d2 + 1];
      int r3 = sf->d[2 * dir];
      int n3 = ndir(b, dir);
      int gC, iC, aC;
      if (st % 2 == 0) {
        gC = r3 + layer - 1;
        iC = r3 - layer;
        aC = r3 - 1;
        if (iC < 0) iC = 0;
      } else {
        gC = r3 - layer;
        iC = r3 + layer - 1;
        aC = r3;
        if (iC >= n3) iC = n3 - 1;
      }
      if (!strcmp(sf->type, "interblock") || !strcmp(sf->type, "periodic")) continue; /* swap_halo */
      char bcn[32];
      snprintf(bcn, sizeof bcn, "%s", sf->type);
      if (!strcmp(bcn, "viscousWall")) strcpy(bcn, "slipWall");
      int isSlip = !strcmp(bcn, "slipWall");
      for (int q2 = r2s; q2 < r2e; ++q2)
        for (int q1 = r1s; q1 < r1e; ++q1) {
          int fi, fj, fk, ci, cj, ck, gi, gj, gk;
          map3(dir, r3, q1, q2, &fi, &fj, &fk);
          map3(dir, isSlip ? iC : aC, q1, q2, &ci, &cj, &ck);
          map3(dir, gC, q1, q2, &gi, &gj, &gk);
          uv3 a = b->fa[dir][fidx(b, dir, fi, fj, fk)];
          double gs[NEQ];
          ghost_state(c, &b->W[cidx(b, ci, cj, ck) * NEQ], bcn, a.u, st, sf->d[6],
                      layer, gs);
          memcpy(&b->W[cidx(b, gi, gj, gk) * NEQ], gs, sizeof gs);
        }
    }
  }
}

/* gridLevel::GetBoundaryConditions: src/gridLevel.cpp:287-319 */
void orc_bc(orc_case *c) {
  for (int bq = 0; bq < c->nb; ++bq) assign_ghosts(c, &c->blk[bq]);
  swap_halo(c, 0);
  for (int bq = 0; bq < c->nb; ++bq) edge_ghosts(c, &c->blk[bq], 0);
}

/* ------------------------------------------------------------------------ */
/* viscous terms (laminar Navier-Stokes)                                     */

/* sutherland::SpeciesViscosity / SpeciesConductivity (src/transport.cpp:
 * 97-117), nondimensional */
static double visc_of_t(const orc_case *c, double t) {
  double temp = t * c->tRef;
  double mu = (c->suth[0] * pow(temp, 1.5)) / (temp + c->suth[1]);
  return mu / c->muRef;
}
static double cond_of_t(const orc_case *c, double t) {
  double temp = t * c->tRef;
  double k = (c->suth[2] * pow(temp, 1.5)) / (temp + c->suth[3]);
  return k / c->kNonDim;
}
/* thermodynamic::Prandtl (include/thermodynamic.hpp:61-64) */
static double prandtl(const orc_gas *g) {
  double gamma = orc_gamma(g);
  return (4.0 * gamma) / (9.0 * gamma - 5.0);
}
static double std_max(double a, double b) { return (a < b) ? b : a; }

/* the part of Visc{Cell,Face}SpectralRadius (include/spectralRadius.hpp:
 * 93-151) that depends on the state: max(4/(3 rho), gamma/rho) and
 * NondimScaling (mu / Pr + mut / Prt), mut = 0 */
static double visc_state_factor(const orc_case *c, const double *W, double mu, double *viscTerm) {
  double rho = 0.0 + W[0];
  double maxTerm = std_max(4.0 / (3.0 * rho), orc_gamma(&c->gas) / rho);
  *viscTerm = c->vscale * (mu / prandtl(&c->gas) + 0.0 / 0.9);
  return maxTerm;
}
static double visc_cell_specrad(const orc_case *c, const double *W, uv3 aL, uv3 aR, double vol,
                                double mu) {
  double fMag = 0.5 * (aL.m + aR.m);
  double vt;
  double mt = visc_state_factor(c, W, mu, &vt);
  return mt * vt * fMag * fMag / vol;
}
static double visc_face_specrad(const orc_case *c, const double *W, uv3 a, double dist, double mu) {
  double vt;
  double mt = visc_state_factor(c, W, mu, &vt);
  return a.m / dist * mt * vt;
}

static int at_corner(const blk_t *b, int i, int j, int k) {
  return (i < 0 || i >= b->ni) && (j < 0 || j >= b->nj) && (k < 0 || k >= b->nk);
}

/* procBlock::UpdateAuxillaryVariables (src/procBlock.cpp:6171-6189): every
 * cell but the block corners, ghost cells included */
static void update_aux(const orc_case *c, blk_t *b) {
  ORC_PARALLEL
  for (int k = -b->g; k < b->nk + b->g; ++k)
    for (int j = -b->g; j < b->nj + b->g; ++j)
      for (int i = -b->g; i < b->ni + b->g; ++i) {
        if (at_corner(b, i, j, k)) continue;
        long q = cidx(b, i, j, k);
        b->T[q] = temperature(&c->gas, &b->W[q * NEQ]);
        if (c->viscous) b->mu[q] = visc_of_t(c, b->T[q]);
      }
}

/* boundaryConditions::GetBCSurface (src/boundaryConditions.cpp:109-190): the
 * surface of the same normal direction whose patch holds face (i, j, k) */
static const surf_t *bc_surface_at(const blk_t *b, int i, int j, int k, int st) {
  int dir = (st - 1) / 2;
  int x[3] = {i, j, k};
  for (int s = 0; s < b->nsurf; ++s) {
    const surf_t *sf = &b->surf[s];
    if ((surf_type(sf) - 1) / 2 != dir) continue;
    int ok = 1;
    for (int q = 0; q < 3 && ok; ++q) {
      if (q == dir) ok = x[q] >= sf->d[2 * q] && x[q] <= sf->d[2 * q + 1];
      else ok = x[q] >= sf->d[2 * q] && x[q] < sf->d[2 * q + 1];
    }
    if (ok) return sf;
  }
  return NULL;
}

/* procBlock::AssignInviscidGhostCellsEdge (src/procBlock.cpp:2565-2714)
 * (viscous == 0) and AssignViscousGhostCellsEdge (:2874-3030): the ghost
 * cells along the 12 block edges, line by line, layer by layer */
static void edge_ghosts(orc_case *c, blk_t *b, int viscous) {
  for (int dd = 0; dd < 3; ++dd) {
    int d2 = (dd + 1) % 3, d3 = (dd + 2) % 3;
    int max1 = ndir(b, dd), max2 = ndir(b, d2), max3 = ndir(b, d3);
    int surfStart2 = 2 * d2 + 1, surfStart3 = 2 * d3 + 1;
    for (int layer3 = 1; layer3 <= b->g; ++layer3)
      for (int layer2 = 1; layer2 <= b->g; ++layer2)
        for (int cc = 0; cc < 4; ++cc) {
          int upper2 = cc > 1, upper3 = cc % 2 == 1;
          int pC2 = upper2 ? max2 + layer2 - 2 : 1 - layer2;
          int gC2 = upper2 ? pC2 + 1 : pC2 - 1;
          int pC3 = upper3 ? max3 + layer3 - 2 : 1 - layer3;
          int gC3 = upper3 ? pC3 + 1 : pC3 - 1;
          int surf2 = upper2 ? surfStart2 + 1 : surfStart2;
          int surf3 = upper3 ? surfStart3 + 1 : surfStart3;
          int cF2_2 = upper2 ? max2 : 0, cF2_3 = upper3 ? max3 - 1 : 0;
          int cF3_2 = upper2 ? max2 - 1 : 0, cF3_3 = upper3 ? max3 : 0;
          for (int d1 = 0; d1 < max1; ++d1) {
            int i, j, k;
            map3(dd, d1, cF2_2, cF2_3, &i, &j, &k);
            const surf_t *s2 = bc_surface_at(b, i, j, k, surf2);
            map3(dd, d1, cF3_2, cF3_3, &i, &j, &k);
            const surf_t *s3 = bc_surface_at(b, i, j, k, surf3);
            map3(dd, d1, cF2_2, gC3, &i, &j, &k);
            v3 fa2 = b->fa[d2][fidx(b, d2, i, j, k)].u;
            map3(dd, d1, gC2, cF3_3, &i, &j, &k);
            v3 fa3 = b->fa[d3][fidx(b, d3, i, j, k)].u;
            char bc2[32], bc3[32];
            snprintf(bc2, sizeof bc2, "%s", s2 ? s2->type : "");
            snprintf(bc3, sizeof bc3, "%s", s3 ? s3->type : "");
            if (!viscous) {
              if (!strcmp(bc2, "viscousWall")) strcpy(bc2, "slipWall");
              if (!strcmp(bc3, "viscousWall")) strcpy(bc3, "slipWall");
            }
            int slip2 = !strcmp(bc2, "slipWall"), slip3 = !strcmp(bc3, "slipWall");
            long gq, q2, q3;
            map3(dd, d1, gC2, gC3, &i, &j, &k);
            gq = cidx(b, i, j, k);
            map3(dd, d1, pC2, gC3, &i, &j, &k);
            q2 = cidx(b, i, j, k);
            map3(dd, d1, gC2, pC3, &i, &j, &k);
            q3 = cidx(b, i, j, k);
            double *W = b->W;
            if (slip2 && !slip3) {
              ghost_state(c, &W[q2 * NEQ], bc2, fa2, surf2, s2->d[6], layer2, &W[gq * NEQ]);
            } else if (!slip2 && slip3) {
              ghost_state(c, &W[q3 * NEQ], bc3, fa3, surf3, s3->d[6], layer3, &W[gq * NEQ]);
            } else if (!viscous || (!strcmp(bc2, "viscousWall") && !strcmp(bc3, "viscousWall"))) {
              if (layer2 == layer3) {
                for (int e = 0; e < NEQ; ++e) W[gq * NEQ + e] = 0.5 * (W[q2 * NEQ + e] + W[q3 * NEQ + e]);
              } else if (layer2 > layer3) {
                memcpy(&W[gq * NEQ], &W[q3 * NEQ], sizeof(double) * NEQ);
              } else {
                memcpy(&W[gq * NEQ], &W[q2 * NEQ], sizeof(double) * NEQ);
              }
            }
          }
        }
  }
}

/* GetGhostState, viscousWall with the low-Re treatment
 * (src/ghostStates.cpp:126-145, 185-190, adiabatic default): the velocity
 * mirrors about the wall velocity; an isothermal wall sets the density from
 * the mirrored temperature at the interior pressure */
static double cond_of_t(const orc_case *c, double t);
static void viscous_wall_ghost(const orc_case *c, const double *in, const bcstate_t *s, double wd, double *ghost) {
  memcpy(ghost, in, sizeof(double) * NEQ);
  v3 vw = s ? s->vel : V(0, 0, 0);
  ghost[1] = 2.0 * vw.x - in[1];
  ghost[2] = 2.0 * vw.y - in[2];
  ghost[3] = 2.0 * vw.z - in[3];
  if (s && s->isothermal) {
    double tGhost = 2.0 * s->twall - temperature(&c->gas, in);
    double R = 0.0 + 1.0 * c->gas.R;
    double rho = ghost[4] / (R * tGhost);
    ghost[0] = rho * (in[0] / (0.0 + in[0]));
  } else if (s && s->heatflux) {
    /* constant heat flux, low-Re (src/ghostStates.cpp:228-238): the
     * interior temperature less q / k times twice the wall distance of the
     * wall-adjacent cell; k = EffectiveConductivity = Sutherland k times
     * the nondimensional scaling (src/transport.cpp:192-196) */
    double t = temperature(&c->gas, in);
    double kappa = cond_of_t(c, t) * c->vscale;
    double tGhost = temperature(&c->gas, in) - s->qwall / kappa * 2.0 * wd;
    double R = 0.0 + 1.0 * c->gas.R;
    double rho = ghost[4] / (R * tGhost);
    ghost[0] = rho * (in[0] / (0.0 + in[0]));
  }
}

/* procBlock::CalcWallDistance via kdtree::NearestNeighbor
 * (src/procBlock.cpp:6030-6041, src/kdtree.cpp:123-224): the distance from
 * a point to the nearest viscousWall face centre of any block
 * (GetViscousFaceCenters, src/utility.cpp:310-366); brute force gives the
 * same minimum of the squared distances */
static double wall_distance(const orc_case *c, v3 pt) {
  double best = DBL_MAX;
  for (int bq = 0; bq < c->nb; ++bq) {
    const blk_t *b = &c->blk[bq];
    for (int s = 0; s < b->nsurf; ++s) {
      const surf_t *sf = &b->surf[s];
      if (strcmp(sf->type, "viscousWall")) continue;
      int st = surf_type(sf), dir = (st - 1) / 2;
      int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
      for (int q2 = sf->d[2 * d2]; q2 < sf->d[2 * d2 + 1]; ++q2)
        for (int q1 = sf->d[2 * d1]; q1 < sf->d[2 * d1 + 1]; ++q1) {
          int i, j, k;
          map3(dir, sf->d[2 * dir], q1, q2, &i, &j, &k);
          v3 f = b->fc[dir][fidx(b, dir, i, j, k)];
          double dx = pt.x - f.x, dy = pt.y - f.y, dz = pt.z - f.z;
          double d2s = ((0.0 + dx * dx) + dy * dy) + dz * dz;
          if (d2s < best) best = d2s;
        }
    }
  }
  return sqrt(best);
}

/* procBlock::AssignViscousGhostCells (src/procBlock.cpp:2760-2836) */
static void viscous_ghosts(orc_case *c, blk_t *b) {
  for (int layer = 1; layer <= b->g; ++layer)
    for (int s = 0; s < b->nsurf; ++s) {
      const surf_t *sf = &b->surf[s];
      if (strcmp(sf->type, "viscousWall")) continue;
      int st = surf_type(sf), dir = (st - 1) / 2;
      int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
      int r3 = sf->d[2 * dir], n3 = ndir(b, dir);
      int gC, iC;
      if (st % 2 == 0) {
        gC = r3 + layer - 1;
        iC = r3 - layer;
        if (iC < 0) iC = 0;
      } else {
        gC = r3 - layer;
        iC = r3 + layer - 1;
        if (iC >= n3) iC = n3 - 1;
      }
      const bcstate_t *bs = find_bc(c, sf->d[6]);
      /* the wall distance is the wall-adjacent cell's, whatever the layer */
      int aC = st % 2 == 0 ? r3 - 1 : r3;
      for (int q2 = sf->d[2 * d2]; q2 < sf->d[2 * d2 + 1]; ++q2)
        for (int q1 = sf->d[2 * d1]; q1 < sf->d[2 * d1 + 1]; ++q1) {
          int ii, ij, ik, gi, gj, gk, ai, aj, ak;
          map3(dir, iC, q1, q2, &ii, &ij, &ik);
          map3(dir, gC, q1, q2, &gi, &gj, &gk);
          map3(dir, aC, q1, q2, &ai, &aj, &ak);
          double wd = 0.0;
          if (bs && bs->heatflux) {
            const long ac = cidx(b, ai, aj, ak);
            if (!b->wd) {
              const long np = (long)b->NI * b->NJ * b->NK;
              b->wd = (double *)malloc(sizeof(double) * np);
              for (long q = 0; q < np; ++q) b->wd[q] = NAN;
            }
            if (isnan(b->wd[ac])) b->wd[ac] = wall_distance(c, b->ctr[ac]);
            wd = b->wd[ac];
          }
          viscous_wall_ghost(c, &b->W[cidx(b, ii, ij, ik) * NEQ], bs, wd, &b->W[cidx(b, gi, gj, gk) * NEQ]);
        }
    }
  edge_ghosts(c, b, 1);
}

/* CalcGradsI/J/K at a face (src/procBlock.cpp:5173-5788), VectorGradGG /
 * ScalarGradGG (src/utility.cpp:59-188): Green-Gauss over the control
 * volume spanning the face's lower cell L and upper cell U.  The face values
 * in direction d are L's and U's; in the other directions the average of
 * L, U and their two neighbours on that side. */
static void face_grads(const orc_case *c, const blk_t *b, int d, int i, int j, int k,
                       double vg[3][3], double tg[3]) {
  (void)c;
  int ed[3] = {d == 0, d == 1, d == 2};
  int L[3] = {i - ed[0], j - ed[1], k - ed[2]}, U[3] = {i, j, k};
  double au[3][3], al[3][3];
  for (int q = 0; q < 3; ++q) {
    int eq[3] = {q == 0, q == 1, q == 2};
    uv3 f1, f2, f3, f4;
    if (q == d) {
      f1 = b->fa[q][fidx(b, q, U[0], U[1], U[2])];
      f2 = b->fa[q][fidx(b, q, U[0] + eq[0], U[1] + eq[1], U[2] + eq[2])];
      f3 = f1;
      f4 = b->fa[q][fidx(b, q, L[0], L[1], L[2])];
    } else {
      f1 = b->fa[q][fidx(b, q, U[0] + eq[0], U[1] + eq[1], U[2] + eq[2])];
      f2 = b->fa[q][fidx(b, q, L[0] + eq[0], L[1] + eq[1], L[2] + eq[2])];
      f3 = b->fa[q][fidx(b, q, U[0], U[1], U[2])];
      f4 = b->fa[q][fidx(b, q, L[0], L[1], L[2])];
    }
    au[q][0] = 0.5 * (f1.u.x * f1.m + f2.u.x * f2.m);
    au[q][1] = 0.5 * (f1.u.y * f1.m + f2.u.y * f2.m);
    au[q][2] = 0.5 * (f1.u.z * f1.m + f2.u.z * f2.m);
    al[q][0] = 0.5 * (f3.u.x * f3.m + f4.u.x * f4.m);
    al[q][1] = 0.5 * (f3.u.y * f3.m + f4.u.y * f4.m);
    al[q][2] = 0.5 * (f3.u.z * f3.m + f4.u.z * f4.m);
  }
  long cL = cidx(b, L[0], L[1], L[2]), cU = cidx(b, U[0], U[1], U[2]);
  double vol = 0.5 * (b->vol[cL] + b->vol[cU]);
  /* values on the six faces: velocity (3) and temperature */
  double vu[3][4], vl[3][4];
  for (int q = 0; q < 3; ++q) {
    int eq[3] = {q == 0, q == 1, q == 2};
    for (int v = 0; v < 4; ++v) {
      double xl = v < 3 ? b->W[cL * NEQ + 1 + v] : b->T[cL];
      double xu = v < 3 ? b->W[cU * NEQ + 1 + v] : b->T[cU];
      if (q == d) {
        vu[q][v] = xu;
        vl[q][v] = xl;
      } else {
        long uP = cidx(b, U[0] + eq[0], U[1] + eq[1], U[2] + eq[2]);
        long lP = cidx(b, L[0] + eq[0], L[1] + eq[1], L[2] + eq[2]);
        long uM = cidx(b, U[0] - eq[0], U[1] - eq[1], U[2] - eq[2]);
        long lM = cidx(b, L[0] - eq[0], L[1] - eq[1], L[2] - eq[2]);
        double up = v < 3 ? b->W[uP * NEQ + 1 + v] : b->T[uP];
        double lp = v < 3 ? b->W[lP * NEQ + 1 + v] : b->T[lP];
        double um = v < 3 ? b->W[uM * NEQ + 1 + v] : b->T[uM];
        double lm = v < 3 ? b->W[lM * NEQ + 1 + v] : b->T[lM];
        vu[q][v] = 0.25 * (xl + xu + up + lp);
        vl[q][v] = 0.25 * (xl + xu + um + lm);
      }
    }
  }
  double invVol = 1.0 / vol;
  for (int r = 0; r < 3; ++r) {
    for (int cc = 0; cc < 3; ++cc)
      vg[r][cc] = (vu[0][cc] * au[0][r] - vl[0][cc] * al[0][r] + vu[1][cc] * au[1][r] -
                   vl[1][cc] * al[1][r] + vu[2][cc] * au[2][r] - vl[2][cc] * al[2][r]) * invVol;
    tg[r] = (vu[0][3] * au[0][r] - vl[0][3] * al[0][r] + vu[1][3] * au[1][r] -
             vl[1][3] * al[1][r] + vu[2][3] * au[2][r] - vl[2][3] * al[2][r]) * invVol;
  }
}

/* viscousFlux::CalcFlux / CalcWallFlux (src/viscousFlux.cpp:58-188) with
 * TauNormal (src/utility.cpp:426-436), laminar single species.  The wall
 * form omits the species-enthalpy term (+0). */
static void viscous_flux(const orc_case *c, const double vg[3][3], const double tg[3], v3 n,
                         const double *Wf, double muFace, int wall, double *F) {
  double mu = c->vscale * muFace, mut = c->vscale * 0.0;
  double lambda = 0.0 - (2.0 / 3.0) * (mu + mut);
  double trace = vg[0][0] + vg[1][1] + vg[2][2];
  double nn[3] = {n.x, n.y, n.z};
  double tau[3];
  for (int r = 0; r < 3; ++r) {
    double s0 = vg[r][0] + vg[0][r], s1 = vg[r][1] + vg[1][r], s2 = vg[r][2] + vg[2][r];
    double sa = s0 * nn[0] + s1 * nn[1] + s2 * nn[2];
    tau[r] = lambda * trace * nn[r] + (mu + mut) * sa;
  }
  double t = temperature(&c->gas, Wf);
  double kk = cond_of_t(c, t) * c->vscale;
  double kt = mut * cps(&c->gas) / 0.9;
  double tv = tau[0] * Wf[1] + tau[1] * Wf[2] + tau[2] * Wf[3];
  double q = (kk + kt) * (tg[0] * nn[0] + tg[1] * nn[1] + tg[2] * nn[2]);
  F[0] = 0.0;
  F[1] = tau[0];
  F[2] = tau[1];
  F[3] = tau[2];
  F[4] = wall ? tv + q : tv + q + 0.0;
}

/* fluxJacobian::ApproxTSLJacobian (include/fluxJacobian.hpp:666-758) for a
 * laminar single species (diffusion coefficient 0, mass fraction 1), times
 * DelprimitiveDelConservative (:612-662) on the right (MatrixMultiply,
 * src/matrix.cpp:186-199).  `left` selects the sign of the shear term;
 * vgr is the velocity gradient, row-major d v_c / d x_r. */
static void tsl_jacobian(const orc_case *c, const double *W, double muLam, uv3 a, double dist, int left,
                         const double *vgr, double *out) {
  const orc_gas *g = &c->gas;
  double t = temperature(g, W);
  double mu = c->vscale * muLam, mut = c->vscale * 0.0;
  v3 n = a.u;
  v3 vel = V(W[1], W[2], W[3]);
  double vn = vdot(vel, n);
  double mf = W[0] / (0.0 + W[0]);
  double rho = 0.0 + W[0];
  double k = cond_of_t(c, t) * c->vscale;
  double kt = mut * cps(g) / 0.9;
  /* TauNormal (src/utility.cpp:426-436) */
  double lambda = 0.0 - (2.0 / 3.0) * (mu + mut);
  double trace = vgr[0] + vgr[4] + vgr[8];
  double nn[3] = {n.x, n.y, n.z}, tau[3];
  for (int r = 0; r < 3; ++r) {
    double s0 = vgr[r * 3 + 0] + vgr[0 * 3 + r], s1 = vgr[r * 3 + 1] + vgr[1 * 3 + r],
           s2 = vgr[r * 3 + 2] + vgr[2 * 3 + r];
    double sa = s0 * nn[0] + s1 * nn[1] + s2 * nn[2];
    tau[r] = lambda * trace * nn[r] + (mu + mut) * sa;
  }
  double fac = left ? -1.0 : 1.0;
  const double third = 1.0 / 3.0;
  double J[25];
  for (int q = 0; q < 25; ++q) J[q] = 0.0;
  double diff = 0.0 + 0.0; /* diffNone: LaminarDiffCoeff + TurbDiffCoeff */
  J[0] = diff * (1.0 - mf) / ((mu + mut) * rho);
  double hs = cps(g) * t; /* SpeciesEnthalpy, hf = 0 */
  double speciesEnthalpy = J[0] * hs;
  J[20] = -(k + kt) * t / ((mu + mut) * rho) + speciesEnthalpy;
  J[6] = third * n.x * n.x + 1.0;
  J[11] = third * n.x * n.y;
  J[16] = third * n.x * n.z;
  J[21] = fac * 0.5 * dist / (mu + mut) * tau[0] + third * n.x * vn + W[1];
  J[7] = third * n.y * n.x;
  J[12] = third * n.y * n.y + 1.0;
  J[17] = third * n.y * n.z;
  J[22] = fac * 0.5 * dist / (mu + mut) * tau[1] + third * n.y * vn + W[2];
  J[8] = third * n.z * n.x;
  J[13] = third * n.z * n.y;
  J[18] = third * n.z * n.z + 1.0;
  J[23] = fac * 0.5 * dist / (mu + mut) * tau[2] + third * n.z * vn + W[3];
  J[24] = (k + kt) / ((mu + mut) * rho);
  double f = a.m * (mu + mut) / dist;
  for (int q = 0; q < 25; ++q) J[q] *= f;
  /* DelprimitiveDelConservative */
  double gm1 = orc_gamma(g) - 1.0;
  double invRho = 1.0 / (0.0 + W[0]);
  double P[25];
  for (int q = 0; q < 25; ++q) P[q] = 0.0;
  P[0] = 1.0;
  P[5] = -invRho * W[1];
  P[10] = -invRho * W[2];
  P[15] = -invRho * W[3];
  P[20] = 0.5 * gm1 * vdot(vel, vel);
  P[6] = invRho;
  P[21] = -gm1 * W[1];
  P[12] = invRho;
  P[22] = -gm1 * W[2];
  P[18] = invRho;
  P[23] = -gm1 * W[3];
  P[24] = gm1;
  for (int r = 0; r < 5; ++r)
    for (int ii = 0; ii < 5; ++ii) {
      double sacc = 0.0;
      for (int cc = 0; cc < 5; ++cc) sacc += J[r * 5 + cc] * P[cc * 5 + ii];
      out[r * 5 + ii] = sacc;
    }
}
void orc_viscous_flux(const orc_case *c, const double *vgr, const double *tg, const double *n, const double *Wf,
                      double muFace, int wall, double *F) {
  double vg[3][3] = {{vgr[0], vgr[1], vgr[2]}, {vgr[3], vgr[4], vgr[5]}, {vgr[6], vgr[7], vgr[8]}};
  viscous_flux(c, vg, tg, V(n[0], n[1], n[2]), Wf, muFace, wall, F);
}
void orc_get_cell_grad(const orc_case *c, int bq, double *out) {
  const blk_t *b = &c->blk[bq];
  memcpy(out, b->vg, sizeof(double) * 9 * (size_t)ncells_padded(b));
}
void orc_tsl_jacobian(const orc_case *c, const double *W, double muLam, const double *area, double dist,
                      int left, const double *vgr, double *out) {
  uv3 a = {V(area[0], area[1], area[2]), area[3]};
  tsl_jacobian(c, W, muLam, a, dist, left, vgr, out);
}

/* procBlock::CalcViscFluxI/J/K (src/procBlock.cpp:1233-2135), laminar */
static void visc_flux_dir(orc_case *c, blk_t *b, int d) {
  int ei = (d == 0), ej = (d == 1), ek = (d == 2);
  int n3 = ndir(b, d);
  /* faces of direction d touch the cells on their two sides: the k-faces
   * of one (i, j) column, or the i- and j-faces of one k plane, touch no
   * other column's (plane's) cells, and each cell sees its faces in the
   * serial order */
  const int outerN = d == 2 ? b->nj : b->nk + ek, midN = d == 2 ? b->nk + ek : b->nj + ej;
  ORC_PARALLEL
  for (int o = 0; o < outerN; ++o)
    for (int m = 0; m < midN; ++m)
      for (int i = 0; i < b->ni + ei; ++i) {
        const int k = d == 2 ? m : o, j = d == 2 ? o : m;
        int f3 = d == 0 ? i : (d == 1 ? j : k);
        double vg[3][3], tg[3];
        face_grads(c, b, d, i, j, k, vg, tg);
        int wall = 0;
        if (f3 == 0 || f3 == n3) {
          const surf_t *sf = bc_surface_at(b, i, j, k, 2 * d + (f3 == 0 ? 1 : 2));
          wall = sf && !strcmp(sf->type, "viscousWall");
        }
        long cL = cidx(b, i - ei, j - ej, k - ek), cU = cidx(b, i, j, k);
        double Wf[NEQ], muF;
        if (!c->viscRecon4) {
          /* FaceReconCentral (include/reconstruction.hpp:314-327) */
          double w[2] = {b->cw[d][cL], b->cw[d][cU]}, co[2];
          orc_lagrange_coeff(w, 1, 0, 0, co);
          for (int e = 0; e < NEQ; ++e) Wf[e] = co[0] * b->W[cU * NEQ + e] + co[1] * b->W[cL * NEQ + e];
          muF = co[0] * b->mu[cU] + co[1] * b->mu[cL];
        } else {
          /* FaceReconCentral4th (:331-349) */
          long cLL = cidx(b, i - 2 * ei, j - 2 * ej, k - 2 * ek);
          long cUU = cidx(b, i + ei, j + ej, k + ek);
          double w[4] = {b->cw[d][cLL], b->cw[d][cL], b->cw[d][cU], b->cw[d][cUU]}, co[4];
          orc_lagrange_coeff(w, 3, 1, 1, co);
          for (int e = 0; e < NEQ; ++e)
            Wf[e] = co[0] * b->W[cLL * NEQ + e] + co[1] * b->W[cL * NEQ + e] +
                    co[2] * b->W[cU * NEQ + e] + co[3] * b->W[cUU * NEQ + e];
          muF = co[0] * b->mu[cLL] + co[1] * b->mu[cL] + co[2] * b->mu[cU] + co[3] * b->mu[cUU];
        }
        uv3 a = b->fa[d][fidx(b, d, i, j, k)];
        double F[NEQ];
        viscous_flux(c, vg, tg, a.u, Wf, muF, wall, F);
        const double sixth = 1.0 / 6.0;
        double vgr[9] = {vg[0][0], vg[0][1], vg[0][2], vg[1][0], vg[1][1], vg[1][2], vg[2][0], vg[2][1], vg[2][2]};
        double J[25];
        if (f3 > 0) {
          double *R = &b->R[pidx(b, i - ei, j - ej, k - ek) * NEQ];
          for (int e = 0; e < NEQ; ++e) R[e] -= F[e] * a.m;
          /* cell velocity gradient: a sixth of each face's (:1396-1414) */
          double *G = &b->vg[cL * 9];
          for (int q = 0; q < 9; ++q) G[q] += sixth * vgr[q];
          /* block diagonal: the lower cell subtracts the TSL Jacobian of
           * the face state with the left sign (:1416-1424) */
          if (c->block) {
            tsl_jacobian(c, Wf, muF, a, proj_c2c(b, d, i, j, k), 1, vgr, J);
            double *Ab = &b->Ab[pidx(b, i - ei, j - ej, k - ek) * 25];
            for (int q = 0; q < 25; ++q) Ab[q] = Ab[q] - J[q];
          }
        }
        if (f3 < n3) {
          long p = pidx(b, i, j, k);
          double *R = &b->R[p * NEQ];
          for (int e = 0; e < NEQ; ++e) R[e] += F[e] * a.m;
          double *G = &b->vg[cU * 9];
          for (int q = 0; q < 9; ++q) G[q] += sixth * vgr[q];
          uv3 aU = b->fa[d][fidx(b, d, i + ei, j + ej, k + ek)];
          double vsr = visc_cell_specrad(c, &b->W[cU * NEQ], a, aU, b->vol[cU], b->mu[cU]);
          b->sr[p] += vsr * c->viscCoeff;
          if (c->block) {
            /* the upper cell adds it with the right sign (:1467-1474) */
            tsl_jacobian(c, Wf, muF, a, proj_c2c(b, d, i, j, k), 0, vgr, J);
            double *Ab = &b->Ab[p * 25];
            for (int q = 0; q < 25; ++q) Ab[q] = Ab[q] + J[q];
          } else if (c->implicit) {
            b->A[p] += 2.0 * vsr;
          }
        }
      }
}

/* test hooks: temperature/viscosity of every non-corner cell, and the face
 * gradients (velocity tensor row-major, then temperature) of one face */
void orc_update_aux(orc_case *c) {
  for (int bq = 0; bq < c->nb; ++bq) update_aux(c, &c->blk[bq]);
}
void orc_face_grads(const orc_case *c, int bq, int d, int i, int j, int k, double *out) {
  double vg[3][3], tg[3];
  face_grads(c, &c->blk[bq], d, i, j, k, vg, tg);
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) out[3 * r + q] = vg[r][q];
  for (int r = 0; r < 3; ++r) out[9 + r] = tg[r];
}
double orc_viscosity(const orc_case *c, double t) { return visc_of_t(c, t); }
double orc_conductivity(const orc_case *c, double t) { return cond_of_t(c, t); }

/* ProjC2CDist (src/procBlock.cpp:6316-6341): centre-to-centre distance
 * across face (i, j, k) of direction d, projected on its unit normal */
static double proj_c2c(const blk_t *b, int d, int i, int j, int k) {
  v3 cu = b->ctr[cidx(b, i, j, k)];
  v3 cl = b->ctr[cidx(b, i - (d == 0), j - (d == 1), k - (d == 2))];
  return vdot(vsub(cu, cl), b->fa[d][fidx(b, d, i, j, k)].u);
}

/* ------------------------------------------------------------------------ */
/* residual: procBlock::CalcResidualNoSource / CalcInvFluxI/J/K            */
/* src/procBlock.cpp:6111-6147, 384-769                                     */

static void inv_flux_dir(orc_case *c, blk_t *b, int d) {
  int ei = (d == 0), ej = (d == 1), ek = (d == 2);
  int n3 = ndir(b, d);
  /* as in visc_flux_dir: columns (k-faces) or planes (i-, j-faces) */
  const int outerN = d == 2 ? b->nj : b->nk + ek, midN = d == 2 ? b->nk + ek : b->nj + ej;
  ORC_PARALLEL
  for (int o = 0; o < outerN; ++o)
    for (int m = 0; m < midN; ++m)
      for (int i = 0; i < b->ni + ei; ++i) {
        const int k = d == 2 ? m : o, j = d == 2 ? o : m;
        int f3 = d == 0 ? i : (d == 1 ? j : k);
        const double *Wm1 = &b->W[cidx(b, i - ei, j - ej, k - ek) * NEQ];
        const double *W0 = &b->W[cidx(b, i, j, k) * NEQ];
        double WL[NEQ], WR[NEQ];
        if (c->recon == 0) {
          memcpy(WL, Wm1, sizeof WL);
          memcpy(WR, W0, sizeof WR);
        } else if (c->recon >= 2) {
          /* src/procBlock.cpp:416-430: lower side from cells -3..+1 around
           * the face, upper side from +2..-2 */
          const double *Wc[6];
          double wc[6];
          for (int q = 0; q < 6; ++q) {
            long cc = cidx(b, i + (q - 3) * ei, j + (q - 3) * ej, k + (q - 3) * ek);
            Wc[q] = &b->W[cc * NEQ];
            wc[q] = b->cw[d][cc];
          }
          double wl[5] = {wc[0], wc[1], wc[2], wc[3], wc[4]};
          double wr[5] = {wc[5], wc[4], wc[3], wc[2], wc[1]};
          orc_weno(Wc[0], Wc[1], Wc[2], Wc[3], Wc[4], wl, c->recon == 3, WL);
          orc_weno(Wc[5], Wc[4], Wc[3], Wc[2], Wc[1], wr, c->recon == 3, WR);
        } else {
          const double *Wm2 = &b->W[cidx(b, i - 2 * ei, j - 2 * ej, k - 2 * ek) * NEQ];
          const double *Wp1 = &b->W[cidx(b, i + ei, j + ej, k + ek) * NEQ];
          double cm2 = b->cw[d][cidx(b, i - 2 * ei, j - 2 * ej, k - 2 * ek)];
          double cm1 = b->cw[d][cidx(b, i - ei, j - ej, k - ek)];
          double c0 = b->cw[d][cidx(b, i, j, k)];
          double cp1 = b->cw[d][cidx(b, i + ei, j + ej, k + ek)];
          orc_muscl(Wm2, Wm1, W0, c->kappa, c->limiter, cm2, cm1, c0, WL);
          orc_muscl(Wp1, W0, Wm1, c->kappa, c->limiter, cp1, c0, cm1, WR);
        }
        uv3 a = b->fa[d][fidx(b, d, i, j, k)];
        double n[3] = {a.u.x, a.u.y, a.u.z};
        double F[NEQ];
        if (c->flux == 0) orc_roe_flux(&c->gas, WL, WR, n, F);
        else orc_ausm_flux(&c->gas, WL, WR, n, F);
        if (f3 > 0) {
          double *R = &b->R[pidx(b, i - ei, j - ej, k - ek) * NEQ];
          for (int e = 0; e < NEQ; ++e) R[e] += F[e] * a.m;
          /* block main diagonal: the lower cell adds the Rusanov Jacobian of
           * the face's lower state (src/procBlock.cpp:449-457) */
          if (c->block) {
            double J[25];
            rusanov_jacobian(&c->gas, WL, a, 1, J);
            double *Ab = &b->Ab[pidx(b, i - ei, j - ej, k - ek) * 25];
            for (int q = 0; q < 25; ++q) Ab[q] = Ab[q] + J[q];
          }
        }
        if (f3 < n3) {
          double *R = &b->R[pidx(b, i, j, k) * NEQ];
          for (int e = 0; e < NEQ; ++e) R[e] -= F[e] * a.m;
          uv3 aU = b->fa[d][fidx(b, d, i + ei, j + ej, k + ek)];
          double lam = inv_cell_specrad(&c->gas, W0, a, aU);
          b->sr[pidx(b, i, j, k)] += lam;
          if (c->block) {
            /* the upper cell subtracts that of the upper state (:480-486) */
            double J[25];
            rusanov_jacobian(&c->gas, WR, a, 0, J);
            double *Ab = &b->Ab[pidx(b, i, j, k) * 25];
            for (int q = 0; q < 25; ++q) Ab[q] = Ab[q] - J[q];
          } else if (c->implicit) {
            b->A[pidx(b, i, j, k)] += lam;
          }
        }
      }
}

void orc_residual(orc_case *c) {
  for (int bq = 0; bq < c->nb; ++bq) {
    blk_t *b = &c->blk[bq];
    memset(b->R, 0, sizeof(double) * NEQ * (size_t)ncells(b));
    memset(b->sr, 0, sizeof(double) * (size_t)ncells(b));
    inv_flux_dir(c, b, 0);
    inv_flux_dir(c, b, 1);
    inv_flux_dir(c, b, 2);
    if (c->viscous) {
      viscous_ghosts(c, b);
      update_aux(c, b);
      memset(b->vg, 0, sizeof(double) * 9 * (size_t)ncells_padded(b)); /* ResetGradients */
      visc_flux_dir(c, b, 0);
      visc_flux_dir(c, b, 1);
      visc_flux_dir(c, b, 2);
    }
  }
  if (c->viscous) swap_halo(c, 2);
}

/* CalcBlockTimeStep / CalcCellDt: src/procBlock.cpp:782-820 */
void orc_time_step(orc_case *c) {
  for (int bq = 0; bq < c->nb; ++bq) {
    blk_t *b = &c->blk[bq];
    ORC_PARALLEL
    for (int k = 0; k < b->nk; ++k)
      for (int j = 0; j < b->nj; ++j)
        for (int i = 0; i < b->ni; ++i) {
          long p = pidx(b, i, j, k);
          if (c->dtIn > 0.0) b->dt[p] = c->dtIn * c->aRef / c->lRef;
          else b->dt[p] = c->cfl * (b->vol[cidx(b, i, j, k)] / dmax(b->sr[p], 0.0));
        }
  }
}

/* AddDiagonalTerms + Invert: src/linearSolver.cpp:146-188, src/matrix.cpp:57-105 */
void orc_invert_diagonal(orc_case *c) {
  for (int bq = 0; bq < c->nb; ++bq) {
    blk_t *b = &c->blk[bq];
    ORC_PARALLEL
    for (int k = 0; k < b->nk; ++k)
      for (int j = 0; j < b->nj; ++j)
        for (int i = 0; i < b->ni; ++i) {
          long p = pidx(b, i, j, k);
          double dvt = (b->vol[cidx(b, i, j, k)] * (1.0 + c->zeta)) / (b->dt[p] * c->theta);
          if (c->dualCFL > 0.0) dvt += dmax(b->sr[p], 0.0) / c->dualCFL;
          if (c->block) {
            /* MultiplyFacOnDiagonal / AddFacOnDiagonal (src/matrix.cpp:
             * 201-225), then MatrixInverse of a copy */
            double *Ab = &b->Ab[p * 25], *Abi = &b->Abi[p * 25];
            for (int q = 0; q < 5; ++q) Ab[q * 6] *= c->relax;
            for (int q = 0; q < 5; ++q) Ab[q * 6] += dvt;
            memcpy(Abi, Ab, sizeof(double) * 25);
            if (gj_inverse(Abi, 5)) {
              fprintf(stderr, "ERROR: Singular matrix in Gauss-Jordan elimination!\n");
              abort();
            }
            continue;
          }
          b->A[p] *= c->relax;
          b->A[p] += dvt;
          b->Ainv[p] = 1.0 * (1.0 / b->A[p]);
        }
  }
}

void orc_reset_diagonal(orc_case *c) {
  for (int bq = 0; bq < c->nb; ++bq) {
    memset(c->blk[bq].A, 0, sizeof(double) * (size_t)ncells(&c->blk[bq]));
    memset(c->blk[bq].Ab, 0, sizeof(double) * 25 * (size_t)ncells(&c->blk[bq]));
  }
}

/* SolDeltaMmN / SolDeltaNm1: src/procBlock.cpp:1010-1034; the rhs
 * b = -R/theta + forcing + Nm1 - MmN (forcing is zero on the single level) */
static void rhs_b(const orc_case *c, const blk_t *b, int i, int j, int k,
                  int withForcing, double *out) {
  long p = pidx(b, i, j, k);
  double vol = b->vol[cidx(b, i, j, k)];
  double thetaInv = 1.0 / c->theta;
  double coeff = (vol * (1.0 + c->zeta)) / (b->dt[p] * c->theta);
  double U[NEQ];
  orc_prim_to_cons(&c->gas, &b->W[cidx(b, i, j, k) * NEQ], U);
  double nm1[NEQ] = {0, 0, 0, 0, 0};
  if (c->multilevel) {
    double c2 = (vol * c->zeta) / (b->dt[p] * c->theta);
    for (int e = 0; e < NEQ; ++e)
      nm1[e] = c2 * (b->consN[p * NEQ + e] - b->consNm1[p * NEQ + e]);
  }
  for (int e = 0; e < NEQ; ++e) {
    double mmn = coeff * (U[e] - b->consN[p * NEQ + e]);
    double v = -thetaInv * b->R[p * NEQ + e];
    if (withForcing) v = v + 0.0;
    out[e] = (v + nm1[e]) - mmn;
  }
}

/* ImplicitLower / ImplicitUpper: src/procBlock.cpp:1056-1161 */
static void implicit_lu(const orc_case *c, const blk_t *b, int i, int j, int k,
                        const double *x, int upper, double *out) {
  for (int e = 0; e < NEQ; ++e) out[e] = 0.0;
  for (int d = 0; d < 3; ++d) {
    int ei = (d == 0), ej = (d == 1), ek = (d == 2);
    int s = upper ? 1 : -1;
    int ni = i + s * ei, nj = j + s * ej, nk = k + s * ek;
    if (ni < 0 || nj < 0 || nk < 0 || ni >= b->ni || nj >= b->nj || nk >= b->nk) {
      /* a ghost neighbour counts only across a connection (BCIsConnection,
       * src/procBlock.cpp:1064,1078,1092) */
      int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
      int c1 = d1 == 0 ? i : d1 == 1 ? j : k, c2 = d2 == 0 ? i : d2 == 1 ? j : k;
      int n1 = d1 == 0 ? b->ni : d1 == 1 ? b->nj : b->nk;
      if (!b->isconn[2 * d + (upper ? 1 : 0)][c1 + (long)n1 * c2]) continue;
    }
    /* face between the cell and its neighbour */
    int fi = upper ? ni : i, fj = upper ? nj : j, fk = upper ? nk : k;
    uv3 a = b->fa[d][fidx(b, d, fi, fj, fk)];
    double o[NEQ];
    long cn = cidx(b, ni, nj, nk);
    /* FaceSpectralRadius with the viscous part (include/spectralRadius.hpp:
     * 192-203): the neighbour's state and viscosity, ProjC2CDist */
    if (c->block) {
      /* RusanovBlockOffDiagonal (src/fluxJacobian.cpp:164-194): with the
       * viscous terms the neighbour's TSL Jacobian, from its state,
       * viscosity and cell velocity gradient, is subtracted (lower) or
       * added (upper) */
      double J[25];
      rusanov_jacobian(&c->gas, &b->W[cn * NEQ], a, !upper, J);
      if (c->viscous) {
        double Vj[25];
        tsl_jacobian(c, &b->W[cn * NEQ], b->mu[cn], a, proj_c2c(b, d, fi, fj, fk), !upper, &b->vg[cn * 9], Vj);
        for (int q = 0; q < 25; ++q) J[q] = !upper ? J[q] - Vj[q] : J[q] + Vj[q];
      }
      block_mult(J, &x[cn * NEQ], o);
      for (int e = 0; e < NEQ; ++e) out[e] += o[e];
      continue;
    }
    if (c->roeJac) {
      roe_offdiag(&c->gas, &b->W[cn * NEQ], &b->W[cidx(b, i, j, k) * NEQ], &x[cn * NEQ], a, !upper, o);
      for (int e = 0; e < NEQ; ++e) out[e] += o[e];
      continue;
    }
    double lv = c->viscous ? visc_face_specrad(c, &b->W[cn * NEQ], a, proj_c2c(b, d, fi, fj, fk), b->mu[cn])
                           : 0.0;
    scalar_offdiag_v(&c->gas, &b->W[cn * NEQ], &x[cn * NEQ], a, !upper, c->viscous, lv, o);
    for (int e = 0; e < NEQ; ++e) out[e] += o[e];
  }
}

/* InitializeMatrixUpdate: src/linearSolver.cpp:111-144 */
void orc_init_update(orc_case *c) {
  for (int bq = 0; bq < c->nb; ++bq) {
    blk_t *b = &c->blk[bq];
    memset(b->x, 0, sizeof(double) * NEQ * (size_t)ncells_padded(b));
    if (!c->initMatrix) continue;
    ORC_PARALLEL
    for (int k = 0; k < b->nk; ++k)
      for (int j = 0; j < b->nj; ++j)
        for (int i = 0; i < b->ni; ++i) {
          double rb[NEQ];
          rhs_b(c, b, i, j, k, 0, rb);
          long p = pidx(b, i, j, k);
          if (c->block) block_mult(&b->Abi[p * 25], rb, &b->x[cidx(b, i, j, k) * NEQ]);
          else
            for (int e = 0; e < NEQ; ++e) b->x[cidx(b, i, j, k) * NEQ + e] = rb[e] * b->Ainv[p];
        }
  }
}

/* lusgs::LUSGS_Forward / LUSGS_Backward: src/linearSolver.cpp:341-428 */
/* the cells of one i+j+k hyperplane depend only on earlier planes: each
 * plane's cells run in parallel, the planes in order */
static long plane_end(const blk_t *b, long q0, long n) {
  const int s = b->order[3 * q0] + b->order[3 * q0 + 1] + b->order[3 * q0 + 2];
  long q = q0;
  while (q < n && b->order[3 * q] + b->order[3 * q + 1] + b->order[3 * q + 2] == s) ++q;
  return q;
}

static void lusgs_forward(orc_case *c, blk_t *b, int sweep) {
  long n = ncells(b);
  for (long q0 = 0; q0 < n;) {
  const long q1 = plane_end(b, q0, n);
  ORC_PARALLEL
  for (long q = q0; q < q1; ++q) {
    int i = b->order[3 * q], j = b->order[3 * q + 1], k = b->order[3 * q + 2];
    double L[NEQ], U[NEQ], rb[NEQ];
    implicit_lu(c, b, i, j, k, b->x, 0, L);
    if (sweep > 0 || c->initMatrix) {
      implicit_lu(c, b, i, j, k, b->x, 1, U);
      for (int e = 0; e < NEQ; ++e) L[e] -= U[e];
    }
    rhs_b(c, b, i, j, k, 1, rb);
    long p = pidx(b, i, j, k);
    if (c->block) {
      double v[NEQ];
      for (int e = 0; e < NEQ; ++e) v[e] = rb[e] + L[e];
      block_mult(&b->Abi[p * 25], v, &b->x[cidx(b, i, j, k) * NEQ]);
      continue;
    }
    for (int e = 0; e < NEQ; ++e) b->x[cidx(b, i, j, k) * NEQ + e] = (rb[e] + L[e]) * b->Ainv[p];
  }
  q0 = q1;
  }
}

static void lusgs_backward(orc_case *c, blk_t *b, int sweep) {
  long n = ncells(b);
  /* planes from the last to the first */
  long *ends = (long *)malloc(sizeof(long) * (size_t)(b->ni + b->nj + b->nk));
  int np = 0;
  for (long q0 = 0; q0 < n; q0 = ends[np++]) ends[np] = plane_end(b, q0, n);
  for (int pl = np - 1; pl >= 0; --pl) {
  const long q0 = pl > 0 ? ends[pl - 1] : 0, q1 = ends[pl];
  ORC_PARALLEL
  for (long q = q1 - 1; q >= q0; --q) {
    int i = b->order[3 * q], j = b->order[3 * q + 1], k = b->order[3 * q + 2];
    double U[NEQ];
    implicit_lu(c, b, i, j, k, b->x, 1, U);
    long p = pidx(b, i, j, k);
    double *xc = &b->x[cidx(b, i, j, k) * NEQ];
    if (sweep > 0 || c->initMatrix) {
      double L[NEQ], rb[NEQ];
      implicit_lu(c, b, i, j, k, b->x, 0, L);
      rhs_b(c, b, i, j, k, 1, rb);
      if (c->block) {
        double v[NEQ];
        for (int e = 0; e < NEQ; ++e) v[e] = (rb[e] + L[e]) - U[e];
        block_mult(&b->Abi[p * 25], v, xc);
      } else {
        for (int e = 0; e < NEQ; ++e) xc[e] = ((rb[e] + L[e]) - U[e]) * b->Ainv[p];
      }
    } else if (c->block) {
      double t[NEQ];
      block_mult(&b->Abi[p * 25], U, t);
      for (int e = 0; e < NEQ; ++e) xc[e] = xc[e] - t[e];
    } else {
      for (int e = 0; e < NEQ; ++e) xc[e] = xc[e] - U[e] * b->Ainv[p];
    }
  }
  }
  free(ends);
}

/* dplur::DPLUR: src/linearSolver.cpp:473-507 */
static void dplur_sweep(orc_case *c, blk_t *b) {
  long npad = ncells_padded(b);
  double *xold = (double *)malloc(sizeof(double) * NEQ * (size_t)npad);
  memcpy(xold, b->x, sizeof(double) * NEQ * (size_t)npad);
  ORC_PARALLEL
  for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
      for (int i = 0; i < b->ni; ++i) {
        double L[NEQ], U[NEQ], rb[NEQ];
        implicit_lu(c, b, i, j, k, xold, 0, L);
        implicit_lu(c, b, i, j, k, xold, 1, U);
        for (int e = 0; e < NEQ; ++e) L[e] -= U[e];
        rhs_b(c, b, i, j, k, 0, rb);
        long p = pidx(b, i, j, k);
        if (c->block) {
          double v[NEQ];
          for (int e = 0; e < NEQ; ++e) v[e] = (rb[e] + 0.0) + L[e];
          block_mult(&b->Abi[p * 25], v, &b->x[cidx(b, i, j, k) * NEQ]);
          continue;
        }
        for (int e = 0; e < NEQ; ++e)
          b->x[cidx(b, i, j, k) * NEQ + e] = ((rb[e] + 0.0) + L[e]) * b->Ainv[p];
      }
  free(xold);
}

/* linearSolver::Residual / AXmB (src/linearSolver.cpp:58-109) and the
 * CycleAtLevel reduction (src/mgSolution.cpp:197-206) */
static double matrix_residual(orc_case *c) {
  double l2 = 0.0;
  long total = 0;
  for (int bq = 0; bq < c->nb; ++bq) {
    blk_t *b = &c->blk[bq];
    double part = 0.0;
    double *sq = (double *)malloc(sizeof(double) * NEQ * (size_t)ncells(b));
    ORC_PARALLEL
    for (int k = 0; k < b->nk; ++k)
      for (int j = 0; j < b->nj; ++j)
        for (int i = 0; i < b->ni; ++i) {
          double L[NEQ], U[NEQ], rb[NEQ];
          implicit_lu(c, b, i, j, k, b->x, 0, L);
          implicit_lu(c, b, i, j, k, b->x, 1, U);
          for (int e = 0; e < NEQ; ++e) L[e] -= U[e];
          rhs_b(c, b, i, j, k, 0, rb);
          long p = pidx(b, i, j, k);
          double axb[NEQ];
          if (c->block) block_mult(&b->Ab[p * 25], &b->x[cidx(b, i, j, k) * NEQ], axb);
          for (int e = 0; e < NEQ; ++e) {
            double ax = c->block ? axb[e] : b->x[cidx(b, i, j, k) * NEQ + e] * b->A[p];
            double axmb = (ax - L[e]) - rb[e];
            double r = 0.0 - axmb;
            sq[p * NEQ + e] = r * r;
          }
        }
    for (long q = 0; q < NEQ * ncells(b); ++q) part += sq[q];
    free(sq);
    l2 += part;
    total += ncells_padded(b) * NEQ;
  }
  return l2 / (double)total;
}

/* lusgs::Relax / dplur::Relax (src/linearSolver.cpp:432-468, 509-535):
 * the connections' ghost updates are swapped before every pass, so the
 * blocks see each other's update lagged by one pass */
double orc_relax(orc_case *c) {
  for (int sw = 0; sw < c->sweeps; ++sw) {
    swap_halo(c, 1);
    for (int bq = 0; bq < c->nb; ++bq) {
      if (c->solver == 0) lusgs_forward(c, &c->blk[bq], sw);
      else dplur_sweep(c, &c->blk[bq]);
    }
    if (c->solver == 0) {
      swap_halo(c, 1);
      for (int bq = 0; bq < c->nb; ++bq) lusgs_backward(c, &c->blk[bq], sw);
    }
  }
  swap_halo(c, 1);
  return c->matres_on ? matrix_residual(c) : 0.0;
}

/* UpdateBlock: src/procBlock.cpp:826-871 (+ ExplicitEulerTimeAdvance
 * :880-896, ImplicitTimeAdvance :902-916) */
void orc_update(orc_case *c, int mm, double *l2, double *linf, int *linfloc) {
  for (int e = 0; e < NEQ; ++e) l2[e] = 0.0;
  *linf = 0.0;
  for (int q = 0; q < 5; ++q) linfloc[q] = 0;
  for (int bq = 0; bq < c->nb; ++bq) {
    blk_t *b = &c->blk[bq];
    ORC_PARALLEL
    for (int k = 0; k < b->nk; ++k)
      for (int j = 0; j < b->nj; ++j)
        for (int i = 0; i < b->ni; ++i) {
          long p = pidx(b, i, j, k);
          double *W = &b->W[cidx(b, i, j, k) * NEQ];
          double *R = &b->R[p * NEQ];
          double Wn[NEQ];
          if (c->timeint == 0) {
            double U[NEQ];
            orc_prim_to_cons(&c->gas, W, U);
            double f = b->dt[p] / b->vol[cidx(b, i, j, k)];
            for (int e = 0; e < NEQ; ++e) U[e] -= f * R[e];
            orc_cons_to_prim(&c->gas, U, Wn);
          } else {
            update_prim(&c->gas, W, &b->x[cidx(b, i, j, k) * NEQ], Wn);
          }
          memcpy(W, Wn, sizeof Wn);
        }
    /* norms in the serial order */
    for (int k = 0; k < b->nk; ++k)
      for (int j = 0; j < b->nj; ++j)
        for (int i = 0; i < b->ni; ++i) {
          const double *R = &b->R[pidx(b, i, j, k) * NEQ];
          for (int e = 0; e < NEQ; ++e) l2[e] += R[e] * R[e];
          for (int e = 0; e < NEQ; ++e)
            if (R[e] > *linf) {
              *linf = R[e];
              linfloc[0] = bq;
              linfloc[1] = i;
              linfloc[2] = j;
              linfloc[3] = k;
              linfloc[4] = e + 1;
            }
        }
    if (c->multilevel && mm == c->nonlinear - 1)
      memcpy(b->consNm1, b->consN, sizeof(double) * NEQ * (size_t)ncells(b));
  }
}

/* input::CalcCFL (src/input.cpp:647-649) + StoreOldSolution
 * (src/mgSolution.cpp:103-114, src/procBlock.cpp:1037-1053) */
void orc_begin_step(orc_case *c, int nn) {
  c->cfl = dmin(c->cflStart + nn * c->cflStep, c->cflMax);
  if (!c->storeN) return;
  for (int bq = 0; bq < c->nb; ++bq) {
    blk_t *b = &c->blk[bq];
    ORC_PARALLEL
    for (int k = 0; k < b->nk; ++k)
      for (int j = 0; j < b->nj; ++j)
        for (int i = 0; i < b->ni; ++i)
          orc_prim_to_cons(&c->gas, &b->W[cidx(b, i, j, k) * NEQ],
                           &b->consN[pidx(b, i, j, k) * NEQ]);
    if (c->multilevel && nn == 0)
      memcpy(b->consNm1, b->consN, sizeof(double) * NEQ * (size_t)ncells(b));
  }
}

/* mgSolution::Iterate: src/mgSolution.cpp:246-269, ImplicitUpdate :209-244 */
int orc_iterate(orc_case *c, int mm, double *l2, double *linf, int *linfloc,
                double *matres) {
  orc_bc(c);
  orc_residual(c);
  orc_time_step(c);
  double mr = 0.0;
  if (c->implicit) {
    orc_invert_diagonal(c);
    orc_init_update(c);
    mr = orc_relax(c);
    orc_update(c, mm, l2, linf, linfloc);
    orc_reset_diagonal(c);
  } else {
    orc_update(c, mm, l2, linf, linfloc);
  }
  /* src/main.cpp:269-270 */
  *matres = sqrt(mr / (c->totalCells * NEQ));
  return 0;
}

/* PrintResiduals normalisation: src/output.cpp:1028-1063 */
void orc_normalize(orc_case *c, int nn, int mm, const double *l2, double *out) {
  if (nn == 0 && mm == 0) {
    for (int e = 0; e < NEQ; ++e) c->l2first[e] = l2[e];
  } else if (nn < 5 && mm == 0) {
    if (0.0 + l2[0] > 0.0 + c->l2first[0]) c->l2first[0] = l2[0];
    for (int e = 1; e < NEQ; ++e) c->l2first[e] = dmax(c->l2first[e], l2[e]);
  }
  out[0] = ((0.0 + l2[0]) + EPS) / ((0.0 + c->l2first[0]) + EPS);
  for (int e = 1; e < NEQ; ++e) out[e] = (l2[e] + EPS) / (c->l2first[e] + EPS);
}

double orc_cfl(const orc_case *c) { return c->cfl; }

/* ------------------------------------------------------------------------ */
/* accessors                                                                */

int orc_num_blocks(const orc_case *c) { return c->nb; }
void orc_block_dims(const orc_case *c, int b, int *ni, int *nj, int *nk) {
  *ni = c->blk[b].ni;
  *nj = c->blk[b].nj;
  *nk = c->blk[b].nk;
}
int orc_num_eq(const orc_case *c) { (void)c; return NEQ; }
int orc_num_ghosts(const orc_case *c) { return c->nb ? c->blk[0].g : 0; }
int orc_iterations(const orc_case *c) { return c->iterations; }
int orc_nonlinear_iterations(const orc_case *c) { return c->nonlinear; }
void orc_refs(const orc_case *c, double *o) {
  o[0] = c->rRef; o[1] = c->tRef; o[2] = c->lRef; o[3] = c->aRef;
}
void orc_gas_of(const orc_case *c, orc_gas *g) { *g = c->gas; }
void orc_set_matrix_residual(orc_case *c, int on) { c->matres_on = on; }

void orc_get_state(const orc_case *c, int bq, double *out) {
  const blk_t *b = &c->blk[bq];
  for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
      for (int i = 0; i < b->ni; ++i)
        memcpy(&out[pidx(b, i, j, k) * NEQ], &b->W[cidx(b, i, j, k) * NEQ], sizeof(double) * NEQ);
}
void orc_set_state(orc_case *c, int bq, const double *in) {
  blk_t *b = &c->blk[bq];
  for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
      for (int i = 0; i < b->ni; ++i)
        memcpy(&b->W[cidx(b, i, j, k) * NEQ], &in[pidx(b, i, j, k) * NEQ], sizeof(double) * NEQ);
}
void orc_set_state_padded(orc_case *c, int bq, const double *in) {
  blk_t *b = &c->blk[bq];
  memcpy(b->W, in, sizeof(double) * NEQ * (size_t)ncells_padded(b));
}
void orc_get_state_padded(const orc_case *c, int bq, double *out) {
  const blk_t *b = &c->blk[bq];
  memcpy(out, b->W, sizeof(double) * NEQ * (size_t)ncells_padded(b));
}
void orc_get_residual(const orc_case *c, int bq, double *out) {
  const blk_t *b = &c->blk[bq];
  memcpy(out, b->R, sizeof(double) * NEQ * (size_t)ncells(b));
}
void orc_get_dt(const orc_case *c, int bq, double *out) {
  memcpy(out, c->blk[bq].dt, sizeof(double) * (size_t)ncells(&c->blk[bq]));
}
void orc_get_specrad(const orc_case *c, int bq, double *out) {
  memcpy(out, c->blk[bq].sr, sizeof(double) * (size_t)ncells(&c->blk[bq]));
}
void orc_get_diag(const orc_case *c, int bq, double *out) {
  memcpy(out, c->blk[bq].A, sizeof(double) * (size_t)ncells(&c->blk[bq]));
}
void orc_get_block_diag(const orc_case *c, int bq, double *out) {
  const blk_t *b = &c->blk[bq];
  memcpy(out, b->Ab, sizeof(double) * 25 * (size_t)ncells(b));
}
void orc_get_block_inverse(const orc_case *c, int bq, double *out) {
  const blk_t *b = &c->blk[bq];
  memcpy(out, b->Abi, sizeof(double) * 25 * (size_t)ncells(b));
}
int orc_is_block(const orc_case *c) { return c->block; }
void orc_get_update(const orc_case *c, int bq, double *out) {
  const blk_t *b = &c->blk[bq];
  for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
      for (int i = 0; i < b->ni; ++i)
        memcpy(&out[pidx(b, i, j, k) * NEQ], &b->x[cidx(b, i, j, k) * NEQ], sizeof(double) * NEQ);
}
void orc_set_update(orc_case *c, int bq, const double *in) {
  blk_t *b = &c->blk[bq];
  for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
      for (int i = 0; i < b->ni; ++i)
        memcpy(&b->x[cidx(b, i, j, k) * NEQ], &in[pidx(b, i, j, k) * NEQ], sizeof(double) * NEQ);
}
void orc_get_volume(const orc_case *c, int bq, double *out) {
  const blk_t *b = &c->blk[bq];
  for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
      for (int i = 0; i < b->ni; ++i) out[pidx(b, i, j, k)] = b->vol[cidx(b, i, j, k)];
}
void orc_get_center(const orc_case *c, int bq, double *out) {
  const blk_t *b = &c->blk[bq];
  for (int k = 0; k < b->nk; ++k)
    for (int j = 0; j < b->nj; ++j)
      for (int i = 0; i < b->ni; ++i) {
        v3 v = b->ctr[cidx(b, i, j, k)];
        long p = pidx(b, i, j, k);
        out[3 * p] = v.x; out[3 * p + 1] = v.y; out[3 * p + 2] = v.z;
      }
}
void orc_get_face_area(const orc_case *c, int bq, int d, double *out) {
  const blk_t *b = &c->blk[bq];
  long n = nfaces_padded(b, d);
  for (long q = 0; q < n; ++q) {
    out[4 * q] = b->fa[d][q].u.x;
    out[4 * q + 1] = b->fa[d][q].u.y;
    out[4 * q + 2] = b->fa[d][q].u.z;
    out[4 * q + 3] = b->fa[d][q].m;
  }
}
void orc_get_cell_width(const orc_case *c, int bq, int d, double *out) {
  const blk_t *b = &c->blk[bq];
  memcpy(out, b->cw[d], sizeof(double) * (size_t)ncells_padded(b));
}

/* ---- SST k-omega 2003 cell model (turbKWSst, src/turbulence.cpp:573-760,
 * include/turbulence.hpp:489-540; turbModel base src/turbulence.cpp:36-105,
 * sutherland::Lambda src/transport.cpp:172-175 with zero bulk viscosity).
 * Pointwise: the blending functions, the limited eddy viscosity, the k and
 * omega source terms with the 2003 production limiter, and the diagonal of
 * the source Jacobian, in the reference's operation order.  Tensors are in
 * the reference's data order (row-major 3 x 3). */
static const double SST_BETA_STAR = 0.09, SST_SIGMA_W2 = 0.856, SST_BETA1 = 0.075, SST_BETA2 = 0.0828,
                    SST_GAMMA1 = 5.0 / 9.0, SST_GAMMA2 = 0.44, SST_A1 = 0.31, SST_PROD2DEST = 10.0;

static double sst_ddot(const double *A, const double *B) {
  /* tensor::DoubleDotTrans: elementwise product, std::accumulate from 0 */
  double s = 0.0;
  for (int q = 0; q < 9; ++q) s = s + A[q] * B[q];
  return s;
}

static double sst_dot3(const double *a, const double *b) {
  /* vector3d::DotProd: std::inner_product from 0 */
  return ((0.0 + a[0] * b[0]) + a[1] * b[1]) + a[2] * b[2];
}

static double sst_cdkw(const double *W, const double *kg, const double *wg) {
  double v = (((2.0 * W[0]) * SST_SIGMA_W2) / W[6]) * sst_dot3(kg, wg);
  return v > 1.0e-10 ? v : 1.0e-10;
}

/* in: W[7] (rho, u, v, w, p, k, omega), G[9] velocity gradient, kg[3], wg[3],
 *     mu (laminar), wallDist, vol, scaling, invScaling (27 values)
 * out: f1, f2, mut, source k, source omega, jacobian (0,0), jacobian (1,1) */
void orc_sst_cell(const double *in, double *out) {
  const double *W = in, *G = in + 7, *kg = in + 16, *wg = in + 19;
  const double mu = in[22], d = in[23], vol = in[24], sc = in[25], isc = in[26];
  const double rho = W[0], k = W[5], om = W[6];
  /* EddyViscAndBlending (:697-727) */
  const double alpha1 = (sc * sqrt(k)) / ((SST_BETA_STAR * om) * (d + EPS));
  const double alpha2 = (((sc * sc) * 500.0) * mu) / ((((d + EPS) * (d + EPS)) * rho) * om);
  const double cdkw = sst_cdkw(W, kg, wg);
  const double alpha3 = (((4.0 * rho) * SST_SIGMA_W2) * k) / ((cdkw * (d + EPS)) * (d + EPS));
  double arg1 = alpha1 > alpha2 ? alpha1 : alpha2;
  arg1 = arg1 < alpha3 ? arg1 : alpha3;
  const double f1 = tanh(pow(arg1, 4.0));
  const double arg2 = (2.0 * alpha1) > alpha2 ? (2.0 * alpha1) : alpha2;
  const double f2 = tanh(arg2 * arg2);
  /* EddyVisc (:573-589): mean strain rate sqrt(2 S:S), S = (G + G^T) / 2 */
  double S[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) S[3 * r + c] = 0.5 * (G[3 * r + c] + G[3 * c + r]);
  const double msr = sqrt(2.0 * sst_ddot(S, S));
  const double lim = (sc * msr) * f2, aw = SST_A1 * om;
  const double mut = ((rho * SST_A1) * k) / (aw > lim ? aw : lim);
  /* CalcTurbSrc (:641-695) */
  const double gamma = f1 * SST_GAMMA1 + (1.0 - f1) * SST_GAMMA2;
  const double beta = f1 * SST_BETA1 + (1.0 - f1) * SST_BETA2;
  const double tkeDest = (isc * SST_BETA_STAR) * (((rho * k) * om) * 1.0);
  const double omgDest = (isc * beta) * ((rho * om) * om);
  /* Boussinesq Reynolds stress (:53-68): lambda tr(G) I + mut (G + G^T) -
   * 2/3 rho k I, lambda = 0 - 2/3 mut; production tau : G */
  const double lambda = 0.0 - (2.0 / 3.0) * mut;
  const double lt = lambda * (G[0] + G[4] + G[8]);
  const double c2 = ((2.0 / 3.0) * rho) * k;
  double tau[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const double I = r == c ? 1.0 : 0.0;
      tau[3 * r + c] = (lt * I + mut * (G[3 * r + c] + G[3 * c + r])) - c2 * I;
    }
  double tkeProd = sc * sst_ddot(tau, G);
  const double pl = SST_PROD2DEST * tkeDest;
  tkeProd = tkeProd < pl ? tkeProd : pl;
  tkeProd = tkeProd > 0.0 ? tkeProd : 0.0;
  double omgProd = ((gamma * rho) / mut) * tkeProd;
  omgProd = omgProd > 0.0 ? omgProd : 0.0;
  const double omgCd = (sc * (1.0 - f1)) * cdkw;
  out[0] = f1;
  out[1] = f2;
  out[2] = mut;
  out[3] = tkeProd - tkeDest;
  out[4] = (omgProd - omgDest) + omgCd;
  /* TurbSrcJac (:745-760), phi = 1 */
  out[5] = ((((-2.0 * SST_BETA_STAR) * om) * 1.0) * vol) * isc;
  out[6] = (((-2.0 * beta) * om) * vol) * isc;
}
