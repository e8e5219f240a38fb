// Device kernels of the hot path.  Launch geometry and data layout:
//
//  * state W: SoA per variable, ghost padded (ni+2g)(nj+2g)(nk+2g), i fastest;
//  * residual, spectral radius, dt, diagonal, du, time-level states: SoA over
//    physical cells, i fastest;
//  * face vectors: one double4 (unit normal, magnitude) per physical face.
#pragma once

#include "device_math.hpp"

namespace ath {

struct BcStateDev;

// The device pointers of a block point to global memory.  In the device pass
// they carry the global address space, so that the compiler emits global
// loads and stores rather than flat ones: a flat access also counts against
// the LDS wait counter, so every wait for LDS would wait for outstanding
// global loads as well.  The host pass sees plain pointers of the same size.
#ifdef __HIP_DEVICE_COMPILE__
#define ATH_GLOBAL __attribute__((address_space(1)))
#else
#define ATH_GLOBAL
#endif
// a device allocation as a block's pointer
template <typename T>
__host__ __device__ inline ATH_GLOBAL T *Dev(T *p) {
  return (ATH_GLOBAL T *)p;
}

struct DevBlock {
  int ni, nj, nk, g;
  int NI, NJ;
  long npad, ncell;
  double ATH_GLOBAL *W;             // 5 * npad
  const double ATH_GLOBAL *cw;      // 3 * npad
  const double4 ATH_GLOBAL *fa[3];  // physical faces per direction
  const double ATH_GLOBAL *vol;     // ncell
  double ATH_GLOBAL *R;             // 5 * ncell
  double ATH_GLOBAL *sr, *dt, *A;   // ncell
  double ATH_GLOBAL *Ainv;          // ncell (FP64 operator)
  float ATH_GLOBAL *Ainv32;         // ncell (FP32 operator)
  double ATH_GLOBAL *consN, *consNm1;  // 5 * ncell
  double ATH_GLOBAL *x;             // 5 * ncell (FP64 operator)
  float ATH_GLOBAL *x32;            // 5 * ncell (FP32 operator)
  double ATH_GLOBAL *xold;          // 5 * ncell, DP-LUR Jacobi copy
  float ATH_GLOBAL *xold32;
  double ATH_GLOBAL *rhs;           // 5 * ncell, LU-SGS right-hand side (pencil sweep)
  // block matrix solvers: main diagonal after AddDiagonalTerms, 25 * ncell,
  // entry (r, c) of cell p at (5 r + c) ncell + p; its inverse in the
  // tile-blocked layout of TileBase (the sweep's 8 x 8 x 4 tiles: one tile's
  // values of one entry are 256 consecutive doubles, so a wavefront reads
  // whole cache lines no other tile shares).  The off-diagonal blocks are
  // formed where they are used (NeighbourBlock).
  double ATH_GLOBAL *Ab, *Abi;
  float ATH_GLOBAL *Abi32;       // FP32 operator: Abi rounded to float, same layout
  int tni, tnj, tnk;  // tile grid of the layout
  // block solvers with the viscous terms: the velocity gradient the viscous
  // pass formed at every physical face of direction d, 9 * faces(d), entry q
  // of face f at q * faces(d) + f; the thin-shear-layer part of the block
  // diagonal reads it back (and forms the face state, viscosity and
  // conductivity again from the cells, with the same functions)
  double ATH_GLOBAL *fvis[3];
  // off-diagonal product of the ghost neighbour across each boundary face
  // that lies on an interblock connection (zero on other faces), per sweep
  // pass; null when the block has no connection.  5 * nbf, SoA.
  double ATH_GLOBAL *bt;            // FP64 operator
  float ATH_GLOBAL *bt32;           // FP32 operator
  long nbf;              // boundary faces: 2 (nj nk + nk ni + ni nj)
  // viscous cases only
  double ATH_GLOBAL *T, *mu;               // padded: temperature, viscosity
  double ATH_GLOBAL *kc;                   // padded: conductivity before scaling (block solvers only)
  const double4 ATH_GLOBAL *faP[3];        // padded faces (ghost faces included)
  const double ATH_GLOBAL *volP;           // padded volumes
  const double ATH_GLOBAL *c2c[3];         // physical faces: centre-to-centre distance projected on the normal
  const unsigned char ATH_GLOBAL *wallf;   // per boundary face (BtSlot order): 1 on a viscousWall patch
  double ATH_GLOBAL *vgrad;                // block solvers: padded cell velocity gradient, 9 * npad, d v_c / d x_r at (3 r + c) npad
  // viscous block solvers: per cell, 1 where the inviscid face states reach
  // a viscousWall ghost cell (null when the block has none)
  const unsigned char ATH_GLOBAL *invPark;
  long cellBase;         // cell offset of this block within its case (L-inf index)
  int blockId;           // block index within its case
  // per case: the CFL number of the current iteration, the device error bits
  // (1 nonphysical state, 4 singular block diagonal) and the boundary states
  // by tag
  const double ATH_GLOBAL *cfl;
  int ATH_GLOBAL *flag;
  const BcStateDev ATH_GLOBAL *states;
};

// laminar viscous terms: Sutherland transport (src/transport.cpp:29-176),
// its nondimensional scaling mu_ref / (rho_ref a_ref L_ref), Prandtl number
// 4 gamma / (9 gamma - 5), and the viscous CFL coefficient
struct ViscParams {
  int on, recon4;
  double vscale, prandtl, coeff, tRef, muRef, kNonDim;
  double c1, s1, c2, s2;  // Sutherland C1, S of viscosity, then of conductivity
};

struct Params {
  Gas gas;
  ViscParams visc;
  double kappa;
  double theta, zeta, relax, dualCFL;
  double dtFixed;  // nondimensional global time step, <= 0 for local dt
  int implicit, multilevel, initMatrix, explicitEuler;
  int block;  // matrixSolver blusgs / bdplur: 5x5 main diagonal and off-diagonals
  int fp32;   // the scalar implicit operator runs in FP32: the residual stage writes Ainv32, else Ainv
  int roeJac; // inviscidFluxJacobian approximateRoe: the off-diagonals are Roe flux changes (FP64)
  int rhsOnce; // the relax forms the right-hand side once, into rhs (the row-parallel block sweep reads it)
  int rhsInDiag; // the viscous block diagonal forms that right-hand side and du = 0 (the relax does not)
};

// ---------------------------------------------------------------------------
// laminar viscous helpers

// idealGas::Temperature (src/eos.cpp:54-60)
__device__ __forceinline__ double TempOf(const GasT<double> &g, double rho, double p) {
  const double rhoR = 0.0 + rho * g.R;
  return p / rhoR;
}
// sutherland::SpeciesViscosity / SpeciesConductivity (src/transport.cpp:
// 97-117), nondimensional; t^1.5 correctly rounded
__device__ __forceinline__ double SutherlandMu(const ViscParams &v, double t) {
  const double temp = t * v.tRef;
  const double mu = (v.c1 * Pow15CR(temp)) / (temp + v.s1);
  return mu / v.muRef;
}
__device__ __forceinline__ double SutherlandK(const ViscParams &v, double t) {
  const double temp = t * v.tRef;
  const double k = (v.c2 * Pow15CR(temp)) / (temp + v.s2);
  return k / v.kNonDim;
}
// max(4/(3 rho), gamma/rho) and NondimScaling (mu / Pr + mut / Prt), mut = 0
// (include/spectralRadius.hpp:93-151)
template <typename T>
__device__ __forceinline__ T ViscFaceSR(const GasT<T> &g, const T W[5], T mag, T dist, T mu,
                                        const ViscParams &v) {
  const T rho = T(0) + W[0];
  const T a = T(4) / (T(3) * rho), bq = g.gamma / rho;
  const T maxTerm = a < bq ? bq : a;
  const T viscTerm = T(v.vscale) * (mu / T(v.prandtl) + T(0) / T(0.9));
  return mag / dist * maxTerm * viscTerm;
}
__device__ __forceinline__ double ViscCellSR(const GasT<double> &g, const double W[5], double magL,
                                             double magR, double vol, double mu, const ViscParams &v) {
  const double fMag = 0.5 * (magL + magR);
  const double rho = 0.0 + W[0];
  const double a = 4.0 / (3.0 * rho), bq = g.gamma / rho;
  const double maxTerm = a < bq ? bq : a;
  const double viscTerm = v.vscale * (mu / v.prandtl + 0.0 / 0.9);
  return maxTerm * viscTerm * fMag * fMag / vol;
}

struct BcStateDev {
  double p, rho, u, v, w, p0, t0, dx, dy, dz;
  double tw;  // viscousWall: wall temperature when iso
  int iso;
  int flux;   // viscousWall: a given heat flux qw
  double qw;
};

struct BcTask {
  int dir, r1s, n1, r2s, n2, r3;
  int gC, srcC, stype, bctype, layer, state;
  long base;
};

__device__ __forceinline__ long PadIdx(const DevBlock &b, int i, int j, int k) {
  return (long)(i + b.g) + (long)(j + b.g) * b.NI + (long)(k + b.g) * b.NI * b.NJ;
}
__device__ __forceinline__ long FaceIdx(const DevBlock &b, int d, int i, int j, int k) {
  const int FI = b.ni + (d == 0), FJ = b.nj + (d == 1);
  return (long)i + (long)j * FI + (long)k * FI * FJ;
}
// slot of a boundary face: surfaces in the order i-lower, i-upper, j-lower,
// j-upper, k-lower, k-upper; within a surface of normal axis d the face
// (c1, c2) along axes d+1, d+2 (mod 3) is at c1 + N(d+1) c2
__device__ __forceinline__ long BtSlot(const DevBlock &b, int d, int upper, int i, int j, int k) {
  const long sI = (long)b.nj * b.nk, sJ = (long)b.nk * b.ni, sK = (long)b.ni * b.nj;
  if (d == 0) return upper * sI + j + (long)b.nj * k;
  if (d == 1) return 2 * sI + upper * sJ + k + (long)b.nk * i;
  return 2 * sI + 2 * sJ + upper * sK + i + (long)b.ni * j;
}
// tile-blocked layout of the block solvers' per-cell 5x5 blocks: entry q of
// cell (i, j, k) at TileBase(b, i, j, k) + 256 q
__device__ __forceinline__ long TileBase(const DevBlock &b, int i, int j, int k) {
  const long tile = (long)(i >> 3) + (long)b.tni * ((j >> 3) + (long)b.tnj * (k >> 2));
  return tile * (25 * 256) + ((i & 7) + 8 * ((j & 7) + 8 * (k & 3)));
}
__device__ __forceinline__ void Map3(int dir, int d3, int d1, int d2, int &i, int &j, int &k) {
  if (dir == 0) { i = d3; j = d1; k = d2; }
  else if (dir == 1) { i = d2; j = d3; k = d1; }
  else { i = d1; j = d2; k = d3; }
}

// ---------------------------------------------------------------------------
// boundary ghost cells: GetGhostState (src/ghostStates.cpp:62-689, inviscid
// branches) and ExtrapolateHoldMixture (:691-709), driven by
// procBlock::AssignInviscidGhostCells (src/procBlock.cpp:2449-2529).  One
// thread per (surface patch cell, layer); every layer depends only on interior
// cells, so all layers of all surfaces run in one launch.

__device__ inline void ExtrapHold(const double bnd[5], double factor, const double in[5],
                                  double out[5]) {
  const double ghostRho = factor * bnd[0] - in[0];
  if (ghostRho <= 0.0) {
    for (int e = 0; e < 5; ++e) out[e] = bnd[e];
    return;
  }
  const double mf = bnd[0] / bnd[0];
  for (int e = 0; e < 5; ++e) out[e] = factor * bnd[e] - in[e];
  out[0] = Dmax(ghostRho * mf, 0.0);
}

__device__ inline void GhostState(const GasT<double> &g, const double in[5], int bctype,
                                  double ax, double ay, double az, int stype, int layer,
                                  const BcStateDev *s, double gh[5]) {
  for (int e = 0; e < 5; ++e) gh[e] = in[e];
  const bool lower = (stype % 2 == 1);
  const double nx = lower ? -1.0 * ax : ax, ny = lower ? -1.0 * ay : ay,
               nz = lower ? -1.0 * az : az;
  if (bctype == 0) {  // slipWall
    const double vn = (in[1] * nx + in[2] * ny) + in[3] * nz;
    gh[1] = in[1] - (nx * 2.0) * vn;
    gh[2] = in[2] - (ny * 2.0) * vn;
    gh[3] = in[3] - (nz * 2.0) * vn;
  } else if (bctype == 1) {  // characteristic
    const double fs[5] = {s->rho * 1.0, s->u, s->v, s->w, s->p};
    const double vn = (in[1] * nx + in[2] * ny) + in[3] * nz;
    const double c = SoS(g, in[0], in[4]);
    const double mach = fabs(vn) / c;
    if (mach >= 1.0 && vn < 0.0) {
      for (int e = 0; e < 5; ++e) gh[e] = fs[e];
    } else if (mach >= 1.0) {
      // supersonic outflow: interior values
    } else if (vn < 0.0) {
      const double rc = in[0] * c;
      const double dv = ((fs[1] - in[1]) * nx + (fs[2] - in[2]) * ny) + (fs[3] - in[3]) * nz;
      gh[4] = 0.5 * (fs[4] + in[4] - rc * dv);
      const double dp = fs[4] - gh[4];
      const double rho = fs[0] - dp / (c * c);
      gh[0] = rho * (fs[0] / fs[0]);
      gh[1] = fs[1] - nx * dp / rc;
      gh[2] = fs[2] - ny * dp / rc;
      gh[3] = fs[3] - nz * dp / rc;
    } else {
      const double rc = in[0] * c;
      const double dp = in[4] - fs[4];
      const double rho = in[0] - dp / (c * c);
      gh[0] = rho * (in[0] / in[0]);
      gh[1] = in[1] + nx * dp / rc;
      gh[2] = in[2] + ny * dp / rc;
      gh[3] = in[3] + nz * dp / rc;
      gh[4] = fs[4];
    }
    double t[5];
    ExtrapHold(gh, 2.0, in, t);
    if (layer > 1) ExtrapHold(t, (double)layer, in, gh);
    else for (int e = 0; e < 5; ++e) gh[e] = t[e];
  } else if (bctype == 2) {  // supersonicInflow
    gh[0] = s->rho * 1.0;
    gh[1] = s->u;
    gh[2] = s->v;
    gh[3] = s->w;
    gh[4] = s->p;
  } else if (bctype == 3) {  // supersonicOutflow
    if (layer > 1)
      for (int e = 0; e < 5; ++e) gh[e] = (double)layer * gh[e] - in[e];
  } else if (bctype == 4) {  // stagnationInlet
    const double gm1 = g.gamma - 1.0;
    const double vn = (in[1] * nx + in[2] * ny) + in[3] * nz;
    const double c = SoS(g, in[0], in[4]);
    const double rNeg = vn - 2.0 * c / gm1;
    const double vmag2 = (in[1] * in[1] + in[2] * in[2]) + in[3] * in[3];
    const double cosT = -1.0 * vn / sqrt(vmag2);
    const double stag = c * c + 0.5 * gm1 * vmag2;
    const double sosB = -1.0 * rNeg * gm1 / (gm1 * cosT * cosT + 2.0) *
                        (1.0 + cosT * sqrt((gm1 * cosT * cosT + 2.0) * stag / (gm1 * rNeg * rNeg) -
                                           0.5 * gm1));
    const double ratio = sosB * sosB / stag;
    const double tb = s->t0 * ratio;
    const double pb = s->p0 * PowCR(ratio, g.gamma / gm1);
    const double vb = sqrt(2.0 / gm1 * (s->t0 - tb));
    gh[0] = (pb / (g.R * tb)) * 1.0;
    gh[1] = vb * s->dx;
    gh[2] = vb * s->dy;
    gh[3] = vb * s->dz;
    gh[4] = pb;
    double t[5];
    ExtrapHold(gh, 2.0, in, t);
    if (layer > 1) ExtrapHold(t, (double)layer, in, gh);
    else for (int e = 0; e < 5; ++e) gh[e] = t[e];
  } else if (bctype == 5) {  // pressureOutlet (reflecting)
    const double c = SoS(g, in[0], in[4]);
    const double rc = in[0] * c;
    gh[4] = s->p;
    const double dp = in[4] - gh[4];
    gh[0] = (in[0] - dp / (c * c)) * (in[0] / in[0]);
    gh[1] = in[1] + nx * dp / rc;
    gh[2] = in[2] + ny * dp / rc;
    gh[3] = in[3] + nz * dp / rc;
    const double gvn = (gh[1] * nx + gh[2] * ny) + gh[3] * nz;
    if (gvn / SoS(g, gh[0], gh[4]) >= 1.0)
      for (int e = 0; e < 5; ++e) gh[e] = in[e];
    for (int e = 0; e < 5; ++e) gh[e] = 2.0 * gh[e] - in[e];
    if (layer > 1)
      for (int e = 0; e < 5; ++e) gh[e] = (double)layer * gh[e] - in[e];
  }
}

// Ghost-cell work of every block in one launch: per block, its task list,
// the global index of its first item (a ghost cell, an edge line or a wall
// cell) and its item count; thread t works on the item of the last block
// whose first item is at or before t.
struct BlockTasks {
  const void *tasks;
  long first, count;
  int ntasks;  // tasks in the list (boundary patches)
  int per;     // tasks per item (edge lines)
};
__device__ __forceinline__ int FindTaskBlock(const BlockTasks *bt, int nb, long t) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bt[mid].first <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) KBoundary(const DevBlock *__restrict__ blocks,
                                                 const BlockTasks *__restrict__ bt, int nb, long total,
                                                 Gas gas) {
  const long tg = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tg >= total) return;
  const int wb = FindTaskBlock(bt, nb, tg);
  const DevBlock &b = blocks[wb];
  const BcTask *__restrict__ tasks = static_cast<const BcTask *>(bt[wb].tasks);
  const int ntasks = bt[wb].ntasks;
  const long t = tg - bt[wb].first;
  int q = 0;
  while (q + 1 < ntasks && tasks[q + 1].base <= t) ++q;
  const BcTask tk = tasks[q];
  const long loc = t - tk.base;
  const int q1 = (int)(loc % tk.n1), q2 = (int)(loc / tk.n1);
  int fi, fj, fk, ci, cj, ck, gi, gj, gk;
  Map3(tk.dir, tk.r3, tk.r1s + q1, tk.r2s + q2, fi, fj, fk);
  Map3(tk.dir, tk.srcC, tk.r1s + q1, tk.r2s + q2, ci, cj, ck);
  Map3(tk.dir, tk.gC, tk.r1s + q1, tk.r2s + q2, gi, gj, gk);
  const double4 a = b.fa[tk.dir][FaceIdx(b, tk.dir, fi, fj, fk)];
  const long src = PadIdx(b, ci, cj, ck), dst = PadIdx(b, gi, gj, gk);
  double in[5], gh[5];
  for (int e = 0; e < 5; ++e) in[e] = b.W[e * b.npad + src];
  const GasT<double> g = Cast<double>(gas);
  GhostState(g, in, tk.bctype, a.x, a.y, a.z, tk.stype, tk.layer,
             tk.state >= 0 ? &b.states[tk.state] : nullptr, gh);
  for (int e = 0; e < 5; ++e) b.W[e * b.npad + dst] = gh[e];
}

// ---------------------------------------------------------------------------
// fused residual: procBlock::CalcInvFluxI/J/K (src/procBlock.cpp:384-769),
// CalcBlockTimeStep (:782-820), linearSolver::AddDiagonalTerms + Invert
// (src/linearSolver.cpp:146-188).
//
// One workgroup per 8x8x4 tile of cells.  The workgroup evaluates every face
// flux that touches the tile exactly once (896 faces: 288 i, 288 j, 320 k),
// keeps F*|A| in LDS, then each thread assembles its cell's residual in the
// reference's accumulation order (-F_i,lo +F_i,hi -F_j,lo +F_j,hi -F_k,lo
// +F_k,hi), the cell spectral radius (i, j, k order), dt, the scalar diagonal
// and its inverse.
constexpr int kTI = 8, kTJ = 8, kTK = 4;
constexpr int kNFI = (kTI + 1) * kTJ * kTK;
constexpr int kNFJ = kTI * (kTJ + 1) * kTK;
constexpr int kNFK = kTI * kTJ * (kTK + 1);
constexpr int kNF = kNFI + kNFJ + kNFK;

// reconstructed lower (WL) and upper (WR) states of the face (i, j, k) of
// direction d (src/procBlock.cpp:384-437)
template <int RECON, int LIM>
__device__ __forceinline__ void FaceStates(const DevBlock &b, double kappa, int d, int i, int j, int k,
                                           double WL[5], double WR[5]) {
  const long s = d == 0 ? 1 : (d == 1 ? (long)b.NI : (long)b.NI * b.NJ);
  const long c0 = PadIdx(b, i, j, k);  // upper cell of the face
  if (RECON == 0) {
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      WL[e] = b.W[e * b.npad + c0 - s];
      WR[e] = b.W[e * b.npad + c0];
    }
  } else if (RECON >= 2) {
    // src/procBlock.cpp:416-430: the lower state from cells -3..+1 around
    // the face, the upper state from +2..-2
    const double *cw = b.cw + d * b.npad;
    double wc[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) wc[q] = cw[c0 + (q - 3) * s];
    const double wl[5] = {wc[0], wc[1], wc[2], wc[3], wc[4]};
    const double wr[5] = {wc[5], wc[4], wc[3], wc[2], wc[1]};
    const WenoGeom GL = MakeWenoGeom(wl), GR = MakeWenoGeom(wr);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const double *We = b.W + e * b.npad + c0;
      double u[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) u[q] = We[(q - 3) * s];
      WL[e] = Weno<RECON == 3>(GL, wl, u[0], u[1], u[2], u[3], u[4]);
      WR[e] = Weno<RECON == 3>(GR, wr, u[5], u[4], u[3], u[2], u[1]);
    }
  } else {
    const double *cw = b.cw + d * b.npad;
    const double wm2 = cw[c0 - 2 * s], wm1 = cw[c0 - s], w0 = cw[c0], wp1 = cw[c0 + s];
    const double dPL = (wm1 + wm1) / (wm1 + w0), dML = (wm1 + wm1) / (wm1 + wm2);
    const double dPR = (w0 + w0) / (w0 + wm1), dMR = (w0 + w0) / (w0 + wp1);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const double *We = b.W + e * b.npad;
      const double um2 = We[c0 - 2 * s], um1 = We[c0 - s], u0 = We[c0], up1 = We[c0 + s];
      WL[e] = Muscl<LIM>(um2, um1, u0, kappa, dPL, dML);
      WR[e] = Muscl<LIM>(up1, u0, um1, kappa, dPR, dMR);
    }
  }
}

// one side of FaceStates: the lower state WL (upper == false) or the upper
// state WR of face (i, j, k) of direction d, with the same arithmetic
template <int RECON, int LIM>
__device__ __forceinline__ void FaceSide(const DevBlock &b, double kappa, int d, int i, int j, int k, bool upper,
                                         double Wout[5]) {
  const long s = d == 0 ? 1 : (d == 1 ? (long)b.NI : (long)b.NI * b.NJ);
  const long c0 = PadIdx(b, i, j, k);
  if (RECON == 0) {
#pragma unroll
    for (int e = 0; e < 5; ++e) Wout[e] = b.W[e * b.npad + c0 - (upper ? 0 : s)];
  } else if (RECON >= 2) {
    const double *cw = b.cw + d * b.npad;
    double wc[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) wc[q] = cw[c0 + (q - 3) * s];
    const double w5[5] = {upper ? wc[5] : wc[0], upper ? wc[4] : wc[1], upper ? wc[3] : wc[2],
                          upper ? wc[2] : wc[3], upper ? wc[1] : wc[4]};
    const WenoGeom G = MakeWenoGeom(w5);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const double *We = b.W + e * b.npad + c0;
      double u[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) u[q] = We[(q - 3) * s];
      Wout[e] = upper ? Weno<RECON == 3>(G, w5, u[5], u[4], u[3], u[2], u[1])
                      : Weno<RECON == 3>(G, w5, u[0], u[1], u[2], u[3], u[4]);
    }
  } else {
    const double *cw = b.cw + d * b.npad;
    const double wm2 = cw[c0 - 2 * s], wm1 = cw[c0 - s], w0 = cw[c0], wp1 = cw[c0 + s];
    if (!upper) {
      const double dPL = (wm1 + wm1) / (wm1 + w0), dML = (wm1 + wm1) / (wm1 + wm2);
#pragma unroll
      for (int e = 0; e < 5; ++e) {
        const double *We = b.W + e * b.npad;
        Wout[e] = Muscl<LIM>(We[c0 - 2 * s], We[c0 - s], We[c0], kappa, dPL, dML);
      }
    } else {
      const double dPR = (w0 + w0) / (w0 + wm1), dMR = (w0 + w0) / (w0 + wp1);
#pragma unroll
      for (int e = 0; e < 5; ++e) {
        const double *We = b.W + e * b.npad;
        Wout[e] = Muscl<LIM>(We[c0 + s], We[c0], We[c0 - s], kappa, dPR, dMR);
      }
    }
  }
}

template <int RECON, int LIM, int FLUX>
__device__ __forceinline__ void FaceFlux(const DevBlock &b, const GasT<double> &g, double kappa,
                                         int d, int i, int j, int k, double F[5]) {
  double WL[5], WR[5];
  FaceStates<RECON, LIM>(b, kappa, d, i, j, k, WL, WR);
  const double4 a = b.fa[d][FaceIdx(b, d, i, j, k)];
  if (FLUX == 0) RoeFlux(g, WL, WR, a.x, a.y, a.z, F);
  else AusmFlux(g, WL, WR, a.x, a.y, a.z, F);
#pragma unroll
  for (int e = 0; e < 5; ++e) F[e] = F[e] * a.w;
}

// CalcCellDt (src/procBlock.cpp:782-820) and AddDiagonalTerms + Invert
// (src/linearSolver.cpp:146-188) for one cell; `sr` is the cell spectral
// radius, `Ad` the accumulated main diagonal
__device__ __forceinline__ void CellTail(const DevBlock &b, const Params &P, long p, double vol, double sr,
                                         double Ad) {
  b.sr[p] = sr;
  const double dt = P.dtFixed > 0.0 ? P.dtFixed : (*b.cfl) * (vol / Dmax(sr, 0.0));
  b.dt[p] = dt;
  if (P.implicit && !P.block) {
    double dvt = (vol * (1.0 + P.zeta)) / (dt * P.theta);
    if (P.dualCFL > 0.0) dvt += Dmax(sr, 0.0) / P.dualCFL;
    double A = Ad * P.relax;
    A += dvt;
    b.A[p] = A;
    const double ainv = 1.0 / A;
    if (!P.fp32) {
      if (b.Ainv) b.Ainv[p] = ainv;
    } else if (b.Ainv32) {
      b.Ainv32[p] = (float)ainv;
    }
  }
}

// Workgroups are dealt to the eight XCDs in turn (workgroup n to XCD n mod
// 8), and each XCD has its own L2.  Renumbering them so that every XCD
// receives one contiguous range of the work keeps neighbouring tiles, which
// read the same stencil data, in one L2.  A bijection of [0, N) for any N.
__device__ __forceinline__ unsigned XcdRemap(unsigned n, unsigned N) {
  const unsigned q = N / 8u, r = N % 8u, x = n % 8u, l = n / 8u;
  return x * q + min(x, r) + l;
}

// One launch over every block.  Per block, its grid of work units (tiles or
// ranges of cells) and the global index of its first unit; workgroup n, after
// the XCD renumbering, takes unit n of the concatenation, so each XCD works
// on one contiguous range of units across the blocks.
struct BlockMap {
  const DevBlock *blocks;
  const int4 *units;  // per block: (ux, uy, uz, first unit), ux uy uz units
  int nb;
};
struct Unit {
  const DevBlock *b;
  int x, y, z;      // the unit within its block's grid
  unsigned local;   // index within the block (x fastest)
  unsigned global;  // index in the concatenation
};
__device__ __forceinline__ Unit MapUnit(const BlockMap &m) {
  const unsigned n = XcdRemap(blockIdx.x, gridDim.x);
  int lo = 0, hi = m.nb - 1;  // the last block whose first unit is <= n
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((unsigned)m.units[mid].w <= n) lo = mid;
    else hi = mid - 1;
  }
  const int4 u = m.units[lo];
  const unsigned l = n - (unsigned)u.w;
  return {m.blocks + lo, (int)(l % (unsigned)u.x), (int)((l / (unsigned)u.x) % (unsigned)u.y),
          (int)(l / ((unsigned)u.x * (unsigned)u.y)), l, n};
}

template <int RECON, int LIM, int FLUX>
__global__ void __launch_bounds__(256) KResidual(BlockMap M, Params P) {
  __shared__ double sF[5][kNF];
  const Unit U = MapUnit(M);
  const DevBlock &b = *U.b;
  const int i0 = U.x * kTI, j0 = U.y * kTJ, k0 = U.z * kTK;
  const GasT<double> g = Cast<double>(P.gas);
  for (int f = threadIdx.x; f < kNF; f += blockDim.x) {
    int d, li, lj, lk;
    if (f < kNFI) {
      d = 0; li = f % (kTI + 1); lj = (f / (kTI + 1)) % kTJ; lk = f / ((kTI + 1) * kTJ);
    } else if (f < kNFI + kNFJ) {
      const int q = f - kNFI;
      d = 1; li = q % kTI; lj = (q / kTI) % (kTJ + 1); lk = q / (kTI * (kTJ + 1));
    } else {
      const int q = f - kNFI - kNFJ;
      d = 2; li = q % kTI; lj = (q / kTI) % kTJ; lk = q / (kTI * kTJ);
    }
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    const bool ok = (i < b.ni + (d == 0)) && (j < b.nj + (d == 1)) && (k < b.nk + (d == 2));
    if (ok) {
      double F[5];
      FaceFlux<RECON, LIM, FLUX>(b, g, P.kappa, d, i, j, k, F);
#pragma unroll
      for (int e = 0; e < 5; ++e) sF[e][f] = F[e];
    }
  }
  __syncthreads();
  const int li = threadIdx.x % kTI, lj = (threadIdx.x / kTI) % kTJ, lk = threadIdx.x / (kTI * kTJ);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  if (i >= b.ni || j >= b.nj || k >= b.nk) return;
  const int fi = lk * (kTI + 1) * kTJ + lj * (kTI + 1) + li;
  const int fj = kNFI + lk * kTI * (kTJ + 1) + lj * kTI + li;
  const int fk = kNFI + kNFJ + lk * kTI * kTJ + lj * kTI + li;
  const long p = (long)i + (long)j * b.ni + (long)k * b.ni * b.nj;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    double r = -sF[e][fi];
    r += sF[e][fi + 1];
    r -= sF[e][fj];
    r += sF[e][fj + kTI];
    r -= sF[e][fk];
    r += sF[e][fk + kTI * kTJ];
    b.R[e * b.ncell + p] = r;
  }
  const long c = PadIdx(b, i, j, k);
  const double rho = b.W[c], u = b.W[b.npad + c], v = b.W[2 * b.npad + c],
               w = b.W[3 * b.npad + c], pr = b.W[4 * b.npad + c];
  const double li_ = InvCellSpecRad(g, rho, u, v, w, pr, b.fa[0][FaceIdx(b, 0, i, j, k)],
                                    b.fa[0][FaceIdx(b, 0, i + 1, j, k)]);
  const double lj_ = InvCellSpecRad(g, rho, u, v, w, pr, b.fa[1][FaceIdx(b, 1, i, j, k)],
                                    b.fa[1][FaceIdx(b, 1, i, j + 1, k)]);
  const double lk_ = InvCellSpecRad(g, rho, u, v, w, pr, b.fa[2][FaceIdx(b, 2, i, j, k)],
                                    b.fa[2][FaceIdx(b, 2, i, j, k + 1)]);
  const double sr = (li_ + lj_) + lk_;
  if (P.visc.on) {
    // the viscous pass adds its spectral radii and forms dt and the diagonal
    b.sr[p] = sr;
    return;
  }
  CellTail(b, P, p, b.vol[p], sr, sr);
}

// ---------------------------------------------------------------------------
// laminar viscous residual: procBlock::CalcResidualNoSource's viscous branch
// (src/procBlock.cpp:6126-6138)

__device__ __forceinline__ long FacePadIdx(const DevBlock &b, int d, int i, int j, int k) {
  const int FI = b.NI + (d == 0), FJ = b.NJ + (d == 1);
  return (long)(i + b.g) + (long)(j + b.g) * FI + (long)(k + b.g) * FI * FJ;
}

// one ghost cell on a block edge (AssignInviscidGhostCellsEdge,
// src/procBlock.cpp:2565-2714, and AssignViscousGhostCellsEdge, :2874-3030):
// action 0 / 1 reflects q2 / q3 across the slip wall of unit normal n,
// 2 averages q2 and q3, 3 copies q3, 4 copies q2, 5 leaves the cell
struct EdgeTask {
  long gq, q2, q3;
  double nx, ny, nz;
  int action, stype, layer, pad;
};

// one thread per edge line; its tasks run in the reference's layer order
__global__ void __launch_bounds__(64) KEdgeGhosts(const DevBlock *__restrict__ blocks,
                                                  const BlockTasks *__restrict__ bt, int nb, long total, Gas gas) {
  const long tg = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tg >= total) return;
  const int wb = FindTaskBlock(bt, nb, tg);
  const DevBlock &b = blocks[wb];
  const EdgeTask *__restrict__ tasks = static_cast<const EdgeTask *>(bt[wb].tasks);
  const int perLine = bt[wb].per;
  const long line = tg - bt[wb].first;
  const GasT<double> g = Cast<double>(gas);
  for (int q = 0; q < perLine; ++q) {
    const EdgeTask t = tasks[(long)line * perLine + q];
    if (t.action == 5) continue;
    double a[5], c[5], gh[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      a[e] = b.W[e * b.npad + t.q2];
      c[e] = b.W[e * b.npad + t.q3];
    }
    if (t.action == 0) GhostState(g, a, 0, t.nx, t.ny, t.nz, t.stype, t.layer, nullptr, gh);
    else if (t.action == 1) GhostState(g, c, 0, t.nx, t.ny, t.nz, t.stype, t.layer, nullptr, gh);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const double v = t.action <= 1 ? gh[e] : (t.action == 2 ? 0.5 * (a[e] + c[e]) : (t.action == 3 ? c[e] : a[e]));
      b.W[e * b.npad + t.gq] = v;
    }
  }
}

// GetGhostState, viscousWall, low-Re treatment (src/ghostStates.cpp:126-145,
// 185-190): velocity mirrored about the wall velocity; an isothermal wall sets
// the density from the mirrored temperature at the interior pressure
struct WallTask {
  long src, dst;
  int state, pad;
  double wd;  // wall distance of the wall-adjacent cell (heat-flux walls)
};

__global__ void __launch_bounds__(256) KViscWall(const DevBlock *__restrict__ blocks,
                                                 const BlockTasks *__restrict__ bt, int nb, long total,
                                                 Gas gas, ViscParams vp) {
  const long tg = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tg >= total) return;
  const int wb = FindTaskBlock(bt, nb, tg);
  const DevBlock &b = blocks[wb];
  const WallTask *__restrict__ tasks = static_cast<const WallTask *>(bt[wb].tasks);
  const WallTask w = tasks[tg - bt[wb].first];
  const BcStateDev *s = w.state >= 0 ? &b.states[w.state] : nullptr;
  double in[5], gh[5];
#pragma unroll
  for (int e = 0; e < 5; ++e) gh[e] = in[e] = b.W[e * b.npad + w.src];
  const double vx = s ? s->u : 0.0, vy = s ? s->v : 0.0, vz = s ? s->w : 0.0;
  gh[1] = 2.0 * vx - in[1];
  gh[2] = 2.0 * vy - in[2];
  gh[3] = 2.0 * vz - in[3];
  if (s && s->iso) {
    const GasT<double> g = Cast<double>(gas);
    const double tGhost = 2.0 * s->tw - TempOf(g, in[0], in[4]);
    const double R = 0.0 + 1.0 * g.R;
    gh[0] = (gh[4] / (R * tGhost)) * (in[0] / (0.0 + in[0]));
  } else if (s && s->flux) {
    // constant heat flux, low-Re treatment (src/ghostStates.cpp:228-238):
    // the interior temperature less q / k times twice the wall distance of
    // the wall-adjacent cell, k = Sutherland conductivity times the scaling
    const GasT<double> g = Cast<double>(gas);
    const double t = TempOf(g, in[0], in[4]);
    const double kappa = SutherlandK(vp, t) * vp.vscale;
    const double tGhost = TempOf(g, in[0], in[4]) - s->qw / kappa * 2.0 * w.wd;
    const double R = 0.0 + 1.0 * g.R;
    gh[0] = (gh[4] / (R * tGhost)) * (in[0] / (0.0 + in[0]));
  }
#pragma unroll
  for (int e = 0; e < 5; ++e) b.W[e * b.npad + w.dst] = gh[e];
}

// UpdateAuxillaryVariables (src/procBlock.cpp:6171-6189): temperature and
// viscosity of every cell but the block corners
__global__ void __launch_bounds__(256) KAux(BlockMap M, Params P) {
  const Unit U = MapUnit(M);
  const DevBlock &b = *U.b;
  const long q = (long)U.local * blockDim.x + threadIdx.x;
  if (q >= b.npad) return;
  const int i = (int)(q % b.NI) - b.g, j = (int)((q / b.NI) % b.NJ) - b.g, k = (int)(q / ((long)b.NI * b.NJ)) - b.g;
  const bool corner = (i < 0 || i >= b.ni) && (j < 0 || j >= b.nj) && (k < 0 || k >= b.nk);
  if (corner) return;
  const double t = TempOf(Cast<double>(P.gas), b.W[q], b.W[4 * b.npad + q]);
  b.T[q] = t;
  b.mu[q] = SutherlandMu(P.visc, t);
  if (b.kc) b.kc[q] = SutherlandK(P.visc, t);  // read by the block Jacobians
}

// The cell values the viscous face terms read (velocity components and
// temperature, padded volume, state, viscosity), from the block's padded
// arrays (CellsG) or staged in LDS for a box of cells (CellsL: eight arrays
// of n cells, rho, u, v, w, p, T, volume, viscosity, the box's cell (i0, j0,
// k0) first).  The values are the same doubles either way.
struct CellsG {
  const DevBlock &b;
  __device__ __forceinline__ long At(int i, int j, int k) const { return PadIdx(b, i, j, k); }
  __device__ __forceinline__ double Var(int v, long c) const { return (v < 3 ? b.W + (v + 1) * b.npad : b.T)[c]; }
  __device__ __forceinline__ double Vol(long c) const { return b.volP[c]; }
  __device__ __forceinline__ double State(int e, long c) const { return b.W[e * b.npad + c]; }
  __device__ __forceinline__ double Mu(long c) const { return b.mu[c]; }
};
struct CellsL {
  const double *s;
  int i0, j0, k0, sx, sy, n;
  __device__ __forceinline__ long At(int i, int j, int k) const {
    return (i - i0) + sx * ((j - j0) + (long)sy * (k - k0));
  }
  __device__ __forceinline__ double Var(int v, long c) const { return s[(v < 3 ? v + 1 : 5) * n + c]; }
  __device__ __forceinline__ double Vol(long c) const { return s[6 * n + c]; }
  __device__ __forceinline__ double State(int e, long c) const { return s[e * n + c]; }
  __device__ __forceinline__ double Mu(long c) const { return s[7 * n + c]; }
};

// CalcGradsI/J/K at a face (src/procBlock.cpp:5173-5788) with VectorGradGG /
// ScalarGradGG (src/utility.cpp:59-188): Green-Gauss over the control volume
// spanning the face's lower cell L and upper cell U; velocity gradient
// vg[r][c] = d v_c / d x_r, temperature gradient tg
template <class C>
__device__ __forceinline__ void FaceGrads(const DevBlock &b, const C &cv, int d, int i, int j, int k,
                                          double vg[3][3], double tg[3]) {
  const int L[3] = {i - (d == 0), j - (d == 1), k - (d == 2)}, U[3] = {i, j, k};
  const long cL = cv.At(L[0], L[1], L[2]), cU = cv.At(U[0], U[1], U[2]);
  const double vol = 0.5 * (cv.Vol(cL) + cv.Vol(cU));
  // Green-Gauss over the control volume's three face pairs: each entry is
  // the sum, left to right, of +vu au - vl al for the pairs q = 0, 1, 2, so
  // it is accumulated one pair at a time
  double acc[3][4];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int eq[3] = {q == 0, q == 1, q == 2};
    const double4 *F = b.faP[q];
    double4 f1, f2, f3, f4;
    if (q == d) {
      f1 = F[FacePadIdx(b, q, U[0], U[1], U[2])];
      f2 = F[FacePadIdx(b, q, U[0] + eq[0], U[1] + eq[1], U[2] + eq[2])];
      f3 = f1;
      f4 = F[FacePadIdx(b, q, L[0], L[1], L[2])];
    } else {
      f1 = F[FacePadIdx(b, q, U[0] + eq[0], U[1] + eq[1], U[2] + eq[2])];
      f2 = F[FacePadIdx(b, q, L[0] + eq[0], L[1] + eq[1], L[2] + eq[2])];
      f3 = F[FacePadIdx(b, q, U[0], U[1], U[2])];
      f4 = F[FacePadIdx(b, q, L[0], L[1], L[2])];
    }
    double au[3], al[3];
    au[0] = 0.5 * (f1.x * f1.w + f2.x * f2.w);
    au[1] = 0.5 * (f1.y * f1.w + f2.y * f2.w);
    au[2] = 0.5 * (f1.z * f1.w + f2.z * f2.w);
    al[0] = 0.5 * (f3.x * f3.w + f4.x * f4.w);
    al[1] = 0.5 * (f3.y * f3.w + f4.y * f4.w);
    al[2] = 0.5 * (f3.z * f3.w + f4.z * f4.w);
    // face values of the pair: three velocity components, then temperature
    const long uP = cv.At(U[0] + eq[0], U[1] + eq[1], U[2] + eq[2]);
    const long lP = cv.At(L[0] + eq[0], L[1] + eq[1], L[2] + eq[2]);
    const long uM = cv.At(U[0] - eq[0], U[1] - eq[1], U[2] - eq[2]);
    const long lM = cv.At(L[0] - eq[0], L[1] - eq[1], L[2] - eq[2]);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double xl = cv.Var(v, cL), xu = cv.Var(v, cU);
      double vu, vl;
      if (q == d) {
        vu = xu;
        vl = xl;
      } else {
        vu = 0.25 * (xl + xu + cv.Var(v, uP) + cv.Var(v, lP));
        vl = 0.25 * (xl + xu + cv.Var(v, uM) + cv.Var(v, lM));
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        if (q == 0) acc[r][v] = vu * au[r] - vl * al[r];
        else acc[r][v] = acc[r][v] + vu * au[r] - vl * al[r];
      }
    }
  }
  const double invVol = 1.0 / vol;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int c = 0; c < 3; ++c) vg[r][c] = acc[r][c] * invVol;
    tg[r] = acc[r][3] * invVol;
  }
}

// viscous flux through face (d, i, j, k) times its area: central or
// fourth-order face state and viscosity (include/reconstruction.hpp:314-349),
// viscousFlux::CalcFlux / CalcWallFlux (src/viscousFlux.cpp:58-188),
// TauNormal (src/utility.cpp:426-436); the wall form has no +0 species term
// (the fourth-order state reads the cells two away from the face, which a
// staged box does not hold: it always reads the block's arrays)
template <bool R4, class C>
__device__ __forceinline__ void ViscFaceState(const DevBlock &b, const C &cv, int d, int i, int j, int k,
                                              double Wf[5], double &muF) {
#pragma clang fp contract(off)
  const long s = d == 0 ? 1 : (d == 1 ? (long)b.NI : (long)b.NI * b.NJ);
  const long cU = PadIdx(b, i, j, k), cL = cU - s;
  const double *cw = b.cw + d * b.npad;
  if (!R4) {
    const double w[2] = {cw[cL], cw[cU]};
    const double c0 = LagrangeCoeffN<2, 1, 0, 0>(w, 0), c1 = LagrangeCoeffN<2, 1, 0, 0>(w, 1);
    const long vU = cv.At(i, j, k), vL = cv.At(i - (d == 0), j - (d == 1), k - (d == 2));
#pragma unroll
    for (int e = 0; e < 5; ++e) Wf[e] = c0 * cv.State(e, vU) + c1 * cv.State(e, vL);
    muF = c0 * cv.Mu(vU) + c1 * cv.Mu(vL);
  } else {
    const long cLL = cL - s, cUU = cU + s;
    const double w[4] = {cw[cLL], cw[cL], cw[cU], cw[cUU]};
    double c[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) c[q] = LagrangeCoeffN<4, 3, 1, 1>(w, q);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const double *We = b.W + e * b.npad;
      Wf[e] = c[0] * We[cLL] + c[1] * We[cL] + c[2] * We[cU] + c[3] * We[cUU];
    }
    muF = c[0] * b.mu[cLL] + c[1] * b.mu[cL] + c[2] * b.mu[cU] + c[3] * b.mu[cUU];
  }
}

template <bool R4>
__device__ __forceinline__ void ViscFaceState(const DevBlock &b, int d, int i, int j, int k, double Wf[5],
                                              double &muF) {
  ViscFaceState<R4>(b, CellsG{b}, d, i, j, k, Wf, muF);
}

template <bool R4, class C>
__device__ __forceinline__ void ViscFaceFlux(const DevBlock &b, const C &cv, const Params &P, int d, int i, int j,
                                             int k, double F[5]) {
#pragma clang fp contract(off)
  const GasT<double> g = Cast<double>(P.gas);
  double vg[3][3], tg[3];
  FaceGrads(b, cv, d, i, j, k, vg, tg);
  const int n3 = d == 0 ? b.ni : (d == 1 ? b.nj : b.nk);
  const int f3 = d == 0 ? i : (d == 1 ? j : k);
  const bool wall = (f3 == 0 || f3 == n3) && b.wallf[BtSlot(b, d, f3 == 0 ? 0 : 1, i, j, k)];
  double Wf[5], muF;
  ViscFaceState<R4>(b, cv, d, i, j, k, Wf, muF);
  const long fq = FaceIdx(b, d, i, j, k);
  const double4 a = b.fa[d][fq];
  const double n[3] = {a.x, a.y, a.z};
  const double mu = P.visc.vscale * muF, mut = P.visc.vscale * 0.0;
  const double lambda = 0.0 - (2.0 / 3.0) * (mu + mut);
  const double trace = vg[0][0] + vg[1][1] + vg[2][2];
  double tau[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double s0 = vg[r][0] + vg[0][r], s1 = vg[r][1] + vg[1][r], s2 = vg[r][2] + vg[2][r];
    const double sa = s0 * n[0] + s1 * n[1] + s2 * n[2];
    tau[r] = lambda * trace * n[r] + (mu + mut) * sa;
  }
  const double t = TempOf(g, Wf[0], Wf[4]);
  const double kLam = SutherlandK(P.visc, t);
  const double kk = kLam * P.visc.vscale;
  const double kt = mut * g.cps / 0.9;
  if (b.fvis[0]) {
    const long nf = (long)(b.ni + (d == 0)) * (b.nj + (d == 1)) * (b.nk + (d == 2));
    double *o = b.fvis[d];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) o[(r * 3 + c) * nf + fq] = vg[r][c];
  }
  const double tv = tau[0] * Wf[1] + tau[1] * Wf[2] + tau[2] * Wf[3];
  const double q = (kk + kt) * (tg[0] * n[0] + tg[1] * n[1] + tg[2] * n[2]);
  F[0] = 0.0 * a.w;
  F[1] = tau[0] * a.w;
  F[2] = tau[1] * a.w;
  F[3] = tau[2] * a.w;
  F[4] = (wall ? tv + q : tv + q + 0.0) * a.w;
}

// fluxJacobian::ApproxTSLJacobian (include/fluxJacobian.hpp:666-758) for a
// laminar single species (diffusion coefficient 0, mass fraction 1), times
// DelprimitiveDelConservative (:612-662) on the right (MatrixMultiply,
// src/matrix.cpp:186-199).  `left` gives the shear term the lower cell's
// sign; vg is the velocity gradient d v_c / d x_r at [3 r + c].
// kLam is Sutherland's conductivity at the state's temperature before the
// nondimensional scaling (formed once per cell or face by the kernels that
// hold it).
__device__ __forceinline__ void TSLJacobian(const GasT<double> &g, const ViscParams &vp, const double W[5],
                                            double muLam, double kLam, double4 a, double dist, bool left,
                                            const double vg[9], double out[25]) {
#pragma clang fp contract(off)
  const double t = TempOf(g, W[0], W[4]);
  const double mu = vp.vscale * muLam, mut = vp.vscale * 0.0;
  const double nx = a.x, ny = a.y, nz = a.z;
  const double u = W[1], v = W[2], w = W[3];
  const double vn = (u * nx + v * ny) + w * nz;
  const double rho = 0.0 + W[0];
  const double k = kLam * vp.vscale;
  const double kt = mut * g.cps / 0.9;
  // TauNormal (src/utility.cpp:426-436)
  const double lambda = 0.0 - (2.0 / 3.0) * (mu + mut);
  const double trace = vg[0] + vg[4] + vg[8];
  const double n[3] = {nx, ny, nz};
  double tau[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double s0 = vg[r * 3 + 0] + vg[r], s1 = vg[r * 3 + 1] + vg[3 + r], s2 = vg[r * 3 + 2] + vg[6 + r];
    const double sa = s0 * n[0] + s1 * n[1] + s2 * n[2];
    tau[r] = lambda * trace * n[r] + (mu + mut) * sa;
  }
  const double fac = left ? -1.0 : 1.0;
  const double third = 1.0 / 3.0;
  const double mr = (mu + mut) * rho;
  double J[25];
#pragma unroll
  for (int q = 0; q < 25; ++q) J[q] = 0.0;
  J[0] = (0.0 + 0.0) * (1.0 - W[0] / rho) / mr;
  J[20] = -(k + kt) * t / mr + J[0] * (g.cps * t);
  const double sh = fac * 0.5 * dist / (mu + mut);
  J[6] = third * nx * nx + 1.0;
  J[11] = third * nx * ny;
  J[16] = third * nx * nz;
  J[21] = sh * tau[0] + third * nx * vn + u;
  J[7] = third * ny * nx;
  J[12] = third * ny * ny + 1.0;
  J[17] = third * ny * nz;
  J[22] = sh * tau[1] + third * ny * vn + v;
  J[8] = third * nz * nx;
  J[13] = third * nz * ny;
  J[18] = third * nz * nz + 1.0;
  J[23] = sh * tau[2] + third * nz * vn + w;
  J[24] = (k + kt) / mr;
  const double f = a.w * (mu + mut) / dist;
#pragma unroll
  for (int q = 0; q < 25; ++q) J[q] *= f;
  // DelprimitiveDelConservative: its nonzero entries (row-major index)
  const double gm1 = g.gamma - 1.0;
  const double invRho = 1.0 / rho;
  const double p5 = -invRho * u, p10 = -invRho * v, p15 = -invRho * w;
  const double p20 = 0.5 * gm1 * ((u * u + v * v) + w * w);
  const double p6 = invRho, p12 = invRho, p18 = invRho;
  const double p21 = -gm1 * u, p22 = -gm1 * v, p23 = -gm1 * w, p24 = gm1;
  // J times it, the row-by-column sums from +0 of the 5x5x5 product with the
  // products of structural zeros left out: row 0 of J is +-0, so is its
  // first and last column in rows 1-3, and the product matrix is zero off
  // its first column, diagonal and last row.  A sum from +0 of finite terms
  // is never -0 and is unchanged by adding +-0, so each entry equals the
  // full product's bit for bit.
#pragma unroll
  for (int q = 0; q < 25; ++q) out[q] = 0.0;
#pragma unroll
  for (int r = 1; r < 4; ++r) {
    out[r * 5 + 0] = (J[r * 5 + 1] * p5 + J[r * 5 + 2] * p10) + J[r * 5 + 3] * p15;
    out[r * 5 + 1] = J[r * 5 + 1] * p6;
    out[r * 5 + 2] = J[r * 5 + 2] * p12;
    out[r * 5 + 3] = J[r * 5 + 3] * p18;
  }
  out[20] = (((J[20] + J[21] * p5) + J[22] * p10) + J[23] * p15) + J[24] * p20;
  out[21] = J[21] * p6 + J[24] * p21;
  out[22] = J[22] * p12 + J[24] * p22;
  out[23] = J[23] * p18 + J[24] * p23;
  out[24] = J[24] * p24;
}

// procBlock::CalcViscFluxI/J/K (src/procBlock.cpp:1233-2135), laminar.  The
// same 8x8x4 tiling as KResidual: every viscous face flux of the tile once
// into LDS, then each cell adds them to its residual in the reference's order
// (+Fv_i,lo -Fv_i,hi +Fv_j,lo -Fv_j,hi +Fv_k,lo -Fv_k,hi), adds the viscous
// spectral radii (i, j, k) to its spectral radius (times the viscous CFL
// coefficient) and twice them to its diagonal, then forms dt and A.
constexpr int kVX = kTI + 2, kVY = kTJ + 2, kVZ = kTK + 2, kVN = kVX * kVY * kVZ;
template <bool R4>
__global__ void __launch_bounds__(256) KViscFlux(BlockMap M, Params P) {
  __shared__ double sF[5][kNF];
  // the tile's cells and one layer around them: every cell a face gradient
  // or a central face state of the tile reads
  __shared__ double sC[8][kVN];
  const Unit U = MapUnit(M);
  const DevBlock &b = *U.b;
  const int i0 = U.x * kTI, j0 = U.y * kTJ, k0 = U.z * kTK;
  for (int q = threadIdx.x; q < kVN; q += blockDim.x) {
    const int i = i0 - 1 + q % kVX, j = j0 - 1 + (q / kVX) % kVY, k = k0 - 1 + q / (kVX * kVY);
    // positions past a partial tile's padded extent are never read
    if (i < b.ni + b.g && j < b.nj + b.g && k < b.nk + b.g) {
      const long c = PadIdx(b, i, j, k);
#pragma unroll
      for (int e = 0; e < 5; ++e) sC[e][q] = b.W[e * b.npad + c];
      sC[5][q] = b.T[c];
      sC[6][q] = b.volP[c];
      sC[7][q] = b.mu[c];
    }
  }
  __syncthreads();
  const CellsL cv{&sC[0][0], i0 - 1, j0 - 1, k0 - 1, kVX, kVY, kVN};
  for (int f = threadIdx.x; f < kNF; f += blockDim.x) {
    int d, li, lj, lk;
    if (f < kNFI) {
      d = 0; li = f % (kTI + 1); lj = (f / (kTI + 1)) % kTJ; lk = f / ((kTI + 1) * kTJ);
    } else if (f < kNFI + kNFJ) {
      const int q = f - kNFI;
      d = 1; li = q % kTI; lj = (q / kTI) % (kTJ + 1); lk = q / (kTI * (kTJ + 1));
    } else {
      const int q = f - kNFI - kNFJ;
      d = 2; li = q % kTI; lj = (q / kTI) % kTJ; lk = q / (kTI * kTJ);
    }
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    const bool ok = (i < b.ni + (d == 0)) && (j < b.nj + (d == 1)) && (k < b.nk + (d == 2));
    if (ok) {
      double F[5];
      ViscFaceFlux<R4>(b, cv, P, d, i, j, k, F);
#pragma unroll
      for (int e = 0; e < 5; ++e) sF[e][f] = F[e];
    }
  }
  __syncthreads();
  const int li = threadIdx.x % kTI, lj = (threadIdx.x / kTI) % kTJ, lk = threadIdx.x / (kTI * kTJ);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  if (i >= b.ni || j >= b.nj || k >= b.nk) return;
  const int fi = lk * (kTI + 1) * kTJ + lj * (kTI + 1) + li;
  const int fj = kNFI + lk * kTI * (kTJ + 1) + lj * kTI + li;
  const int fk = kNFI + kNFJ + lk * kTI * kTJ + lj * kTI + li;
  const long p = (long)i + (long)j * b.ni + (long)k * b.ni * b.nj;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    double r = b.R[e * b.ncell + p];
    r += sF[e][fi];
    r -= sF[e][fi + 1];
    r += sF[e][fj];
    r -= sF[e][fj + kTI];
    r += sF[e][fk];
    r -= sF[e][fk + kTI * kTJ];
    b.R[e * b.ncell + p] = r;
  }
  const GasT<double> g = Cast<double>(P.gas);
  const long c = PadIdx(b, i, j, k);
  double W[5];
#pragma unroll
  for (int e = 0; e < 5; ++e) W[e] = b.W[e * b.npad + c];
  const double vol = b.vol[p], mu = b.mu[c];
  double sr = b.sr[p], Ad = sr;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int ei = d == 0, ej = d == 1, ek = d == 2;
    const double v = ViscCellSR(g, W, b.fa[d][FaceIdx(b, d, i, j, k)].w,
                                b.fa[d][FaceIdx(b, d, i + ei, j + ej, k + ek)].w, vol, mu, P.visc);
    sr += v * P.visc.coeff;
    if (P.implicit) Ad += 2.0 * v;
  }
  CellTail(b, P, p, vol, sr, Ad);
}
// ---------------------------------------------------------------------------
// implicit operator

// b = -R/theta + forcing + SolDeltaNm1 - SolDeltaMmN
// (src/linearSolver.cpp:360-378, src/procBlock.cpp:1010-1034); forcing is
// zero on the single grid level
// the same from values already loaded (cNm1 is read only when multilevel)
__device__ __forceinline__ void RhsFrom(const Params &P, const double W[5], double vol, double dt,
                                        const double R[5], const double cN[5], const double cNm1[5],
                                        double out[5]) {
#pragma clang fp contract(off)
  const GasT<double> g = Cast<double>(P.gas);
  double U[5];
  PrimToCons(g, W, U);
  const double coeff = (vol * (1.0 + P.zeta)) / (dt * P.theta);
  const double thetaInv = 1.0 / P.theta;
  const double c2 = P.multilevel ? (vol * P.zeta) / (dt * P.theta) : 0.0;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const double nm1 = P.multilevel ? c2 * (cN[e] - cNm1[e]) : 0.0;
    const double mmn = coeff * (U[e] - cN[e]);
    double v = -thetaInv * R[e];
    out[e] = (v + nm1) - mmn;
  }
}

__device__ __forceinline__ void Rhs(const DevBlock &b, const Params &P, long p, long c,
                                    double out[5]) {
  double W[5], R[5], cN[5], cNm1[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    W[e] = b.W[e * b.npad + c];
    R[e] = b.R[e * b.ncell + p];
    cN[e] = b.consN[e * b.ncell + p];
    if (P.multilevel) cNm1[e] = b.consNm1[e * b.ncell + p];
  }
  RhsFrom(P, W, b.vol[p], b.dt[p], R, cN, cNm1, out);
}

// Block main diagonal (src/procBlock.cpp:449-486, 588-625, 726-763):
// every face adds the Rusanov Jacobian of its lower state to its lower cell
// and subtracts that of its upper state from its upper cell, faces in i, j, k
// order, so a cell sees -J(face i-) +J(face i+) -J(j-) +J(j+) -J(k-) +J(k+)
// from a zero matrix.  Then AddDiagonalTerms (src/linearSolver.cpp:146-171)
// and the Gauss-Jordan inverse (src/linearSolver.cpp:173-184).  One thread per
// cell, after the residual kernels have written sr and dt.  A zero pivot sets
// bit 4 of `flag`.
//
// MODE 0: the inviscid sum, then AddDiagonalTerms and the inverse.
// With the viscous terms the inviscid sum must see the inviscid ghost states
// (a viscousWall is a slip wall there).  The cells whose face states reach a
// viscousWall ghost (`invPark`) form it right after the inviscid residual
// kernel and park it in Ab (MODE 1); after the viscous pass MODE 2 (central)
// or 3 (fourth-order face states) takes the parked sum or forms the sum
// itself (every other cell: its face states read no ghost cell that the
// viscous pass changes, so the sum is the same bits), adds the viscous
// faces' thin-shear-layer Jacobians, the cell as upper cell adding
// (src/procBlock.cpp:1467-1474) and as lower cell subtracting (:1416-1424),
// stores the cell velocity gradient, a sixth of each face's in the same order
// (:1396-1414, 1427-1436), and finishes.
template <int RECON, int LIM, int MODE>
__global__ void __launch_bounds__(128) KBlockDiag(BlockMap M, Params P) {
#pragma clang fp contract(off)
  const Unit U = MapUnit(M);
  const DevBlock &b = *U.b;
  const long p = (long)U.local * blockDim.x + threadIdx.x;
  if (p >= b.ncell) return;
  const bool parked = MODE >= 1 && b.invPark && b.invPark[p];
  if (MODE == 1 && !parked) return;
  const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
  const GasT<double> g = Cast<double>(P.gas);
  double D[25];
#pragma unroll
  for (int q = 0; q < 25; ++q) D[q] = MODE >= 2 && parked ? b.Ab[q * b.ncell + p] : 0.0;
  if (!(MODE >= 2 && parked)) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      // the cell is the upper cell of its lower face (upper state WR) and the
      // lower cell of its upper face (lower state WL): one side of each
      double Wf[5], J[25];
      FaceSide<RECON, LIM>(b, P.kappa, d, i, j, k, true, Wf);
      double4 a = b.fa[d][FaceIdx(b, d, i, j, k)];
      RusanovJacobian<double>(g, Wf, a.x, a.y, a.z, a.w, false, J);
#pragma unroll
      for (int q = 0; q < 25; ++q) D[q] = D[q] - J[q];
      const int ui = i + (d == 0), uj = j + (d == 1), uk = k + (d == 2);
      FaceSide<RECON, LIM>(b, P.kappa, d, ui, uj, uk, false, Wf);
      a = b.fa[d][FaceIdx(b, d, ui, uj, uk)];
      RusanovJacobian<double>(g, Wf, a.x, a.y, a.z, a.w, true, J);
#pragma unroll
      for (int q = 0; q < 25; ++q) D[q] = D[q] + J[q];
    }
  }
  if (MODE == 1) {
#pragma unroll
    for (int q = 0; q < 25; ++q) b.Ab[q * b.ncell + p] = D[q];
    return;
  }
  if (MODE >= 2) {
    constexpr double sixth = 1.0 / 6.0;
    double G[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) G[q] = 0.0;
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const int fi = i + (side && d == 0), fj = j + (side && d == 1), fk = k + (side && d == 2);
        const long fq = FaceIdx(b, d, fi, fj, fk);
        const long nf = (long)(b.ni + (d == 0)) * (b.nj + (d == 1)) * (b.nk + (d == 2));
        // the face's gradient as the viscous pass formed it; its state,
        // viscosity and conductivity formed again, as that pass forms them
        const double *fv = b.fvis[d];
        double vgr[9], Wf[5], J[25], muF;
#pragma unroll
        for (int q = 0; q < 9; ++q) vgr[q] = fv[q * nf + fq];
        ViscFaceState<MODE == 3>(b, d, fi, fj, fk, Wf, muF);
        const double kF = SutherlandK(P.visc, TempOf(g, Wf[0], Wf[4]));
        TSLJacobian(g, P.visc, Wf, muF, kF, b.fa[d][fq], b.c2c[d][fq], side == 1, vgr, J);
#pragma unroll
        for (int q = 0; q < 25; ++q) D[q] = side ? D[q] - J[q] : D[q] + J[q];
#pragma unroll
        for (int q = 0; q < 9; ++q) G[q] += sixth * vgr[q];
      }
    }
    const long c = PadIdx(b, i, j, k);
#pragma unroll
    for (int q = 0; q < 9; ++q) b.vgrad[q * b.npad + c] = G[q];
  }
  const double sr = b.sr[p];
  double dvt = (b.vol[p] * (1.0 + P.zeta)) / (b.dt[p] * P.theta);
  if (P.dualCFL > 0.0) dvt += Dmax(sr, 0.0) / P.dualCFL;
#pragma unroll
  for (int q = 0; q < 5; ++q) D[q * 6] *= P.relax;
#pragma unroll
  for (int q = 0; q < 5; ++q) D[q * 6] += dvt;
#pragma unroll
  for (int q = 0; q < 25; ++q) b.Ab[q * b.ncell + p] = D[q];
  double I[25];
  if (!GJInverse5<double>(D, I)) atomicOr(b.flag, 4);
#pragma unroll
  for (int q = 0; q < 25; ++q) b.Abi[TileBase(b, i, j, k) + 256 * q] = I[q];
  if (P.fp32 && b.Abi32) {
#pragma unroll
    for (int q = 0; q < 25; ++q) b.Abi32[TileBase(b, i, j, k) + 256 * q] = float(I[q]);
  }
  if (MODE >= 2 && P.rhsInDiag) {
    // the initialisation of a one-sweep relax (KInitUpdate with rhsOnce,
    // FP64 operator) from the values at hand: the right-hand side and du = 0
    double rb[5];
    Rhs(b, P, p, PadIdx(b, i, j, k), rb);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      b.rhs[e * b.ncell + p] = rb[e];
      b.x[e * b.ncell + p] = 0.0;
    }
  }
}

// RusanovBlockOffDiagonal (src/fluxJacobian.cpp:164-194): the neighbour's
// Rusanov Jacobian; with the viscous terms (vp != null) its thin-shear-layer
// Jacobian from the neighbour's state, viscosity `mu` and cell velocity
// gradient `vg`, at the face's projected centre distance, subtracted for a
// lower (positive) and added for an upper neighbour
template <typename T>
__device__ __forceinline__ void BlockOffDiag(const GasT<T> &g, const T W[5], const T du[5], double4 a,
                                             bool positive, T out[5], const ViscParams *vp = nullptr,
                                             double mu = 0.0, double kLam = 0.0, double dist = 1.0,
                                             const double *vg = nullptr) {
  T J[25];
  RusanovJacobian<T>(g, W, T(a.x), T(a.y), T(a.z), T(a.w), positive, J);
  if (vp && vp->on) {
    const GasT<double> gd = {double(g.R), double(g.gamma), double(g.cps), double(g.cvs)};
    double Wd[5], V[25];
#pragma unroll
    for (int e = 0; e < 5; ++e) Wd[e] = double(W[e]);
    TSLJacobian(gd, *vp, Wd, mu, kLam, a, dist, positive, vg, V);
#pragma unroll
    for (int q = 0; q < 25; ++q) J[q] = positive ? J[q] - T(V[q]) : J[q] + T(V[q]);
  }
  BlockMult<T>(J, du, out);
}

// The off-diagonal block of cell (i, j, k) for its neighbour across
// direction d (upper: the neighbour at +1, else at -1), the matrix that
// RusanovBlockOffDiagonal (src/fluxJacobian.cpp:164-194) forms: the
// neighbour's Rusanov Jacobian at the shared face, with the neighbour's
// thin-shear-layer Jacobian (its cell velocity gradient, viscosity and
// conductivity, the face's projected centre distance) subtracted for a lower
// and added for an upper neighbour.  Formed where it is used, from the same
// inputs in the same order, so every user sees the same bits.
__device__ __forceinline__ void NeighbourBlock(const DevBlock &b, const GasT<double> &g, const ViscParams &vp,
                                               int d, bool upper, int i, int j, int k, double J[25]) {
#pragma clang fp contract(off)
  const int s = upper ? 1 : -1;
  const int ni = i + (d == 0 ? s : 0), nj = j + (d == 1 ? s : 0), nk = k + (d == 2 ? s : 0);
  const long f = upper ? FaceIdx(b, d, ni, nj, nk) : FaceIdx(b, d, i, j, k);
  const double4 a = b.fa[d][f];
  const long c = PadIdx(b, ni, nj, nk);
  double W[5];
#pragma unroll
  for (int e = 0; e < 5; ++e) W[e] = b.W[e * b.npad + c];
  const bool positive = !upper;
  RusanovJacobian<double>(g, W, a.x, a.y, a.z, a.w, positive, J);
  if (vp.on) {
    double vg[9], V[25];
#pragma unroll
    for (int q = 0; q < 9; ++q) vg[q] = b.vgrad[q * b.npad + c];
    TSLJacobian(g, vp, W, b.mu[c], b.kc[c], a, b.c2c[d][f], positive, vg, V);
#pragma unroll
    for (int q = 0; q < 25; ++q) J[q] = positive ? J[q] - V[q] : J[q] + V[q];
  }
}

// BlockMult with the matrix read entry by entry from global memory (entry q
// at M[q * stride + idx]), in the same order of operations
template <typename T, typename MT>
__device__ __forceinline__ void BlockMultG(const MT *__restrict__ M, long stride, long idx, const T v[5],
                                           T out[5]) {
  // all 25 loads issued before the arithmetic
  T m[25];
#pragma unroll
  for (int q = 0; q < 25; ++q) m[q] = T(M[q * stride + idx]);
  BlockMult<T>(m, v, out);
}

// block product with the cell's main diagonal Ab (cell-major layout)
template <typename T>
__device__ __forceinline__ void CellBlockMult(const DevBlock &b, const double *M, long p, const T v[5], T out[5]) {
  BlockMultG<T, double>(M, b.ncell, p, v, out);
}
// block product with a tile-blocked per-cell block (Abi)
template <typename T, typename MT>
__device__ __forceinline__ void TileBlockMult(const DevBlock &b, const MT *M, int i, int j, int k, const T v[5],
                                              T out[5]) {
  BlockMultG<T>(M, 256, TileBase(b, i, j, k), v, out);
}

template <typename T>
struct OpBuf;
template <>
struct OpBuf<double> {
  static __device__ __forceinline__ double *Bt(const DevBlock &b) { return b.bt; }
  static __device__ __forceinline__ double *X(const DevBlock &b) { return b.x; }
  static __device__ __forceinline__ double *Xold(const DevBlock &b) { return b.xold; }
  static __device__ __forceinline__ const double *Ainv(const DevBlock &b) { return b.Ainv; }
  static __device__ __forceinline__ const double *Abi(const DevBlock &b) { return b.Abi; }
};
template <>
struct OpBuf<float> {
  static __device__ __forceinline__ float *Bt(const DevBlock &b) { return b.bt32; }
  static __device__ __forceinline__ float *X(const DevBlock &b) { return b.x32; }
  static __device__ __forceinline__ float *Xold(const DevBlock &b) { return b.xold32; }
  static __device__ __forceinline__ const float *Ainv(const DevBlock &b) { return b.Ainv32; }
  static __device__ __forceinline__ const float *Abi(const DevBlock &b) { return b.Abi32; }
};

// sum over the lower (upper == false) or upper neighbours of the Rusanov
// off-diagonal products: procBlock::ImplicitLower / ImplicitUpper
// (src/procBlock.cpp:1056-1161), neighbours in i, j, k order.  A ghost
// neighbour counts only across a connection (src/procBlock.cpp:1064, 1078,
// 1092); its product is the precomputed boundary term, zero elsewhere, and
// adding +0 leaves the sum unchanged (it starts at +0 and never becomes -0).
template <typename T>
__device__ __forceinline__ void OffDiagSum(const DevBlock &b, const GasT<T> &g, const ViscParams &vp,
                                           const T *x, int i, int j, int k, bool upper, T out[5],
                                           int block = 0, int roe = 0, const Gas *gas64 = nullptr) {
  // the block and Roe products are formed in FP64, with the FP64 gas
  const GasT<double> gd = gas64 ? Cast<double>(*gas64)
                                : GasT<double>{double(g.R), double(g.gamma), double(g.cps), double(g.cvs)};
#pragma unroll
  for (int e = 0; e < 5; ++e) out[e] = T(0);
  const T *bt = OpBuf<T>::Bt(b);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int s = upper ? 1 : -1;
    const int ni = i + (d == 0 ? s : 0), nj = j + (d == 1 ? s : 0), nk = k + (d == 2 ? s : 0);
    if (ni < 0 || nj < 0 || nk < 0 || ni >= b.ni || nj >= b.nj || nk >= b.nk) {
      if (bt) {
        const long q = BtSlot(b, d, upper ? 1 : 0, i, j, k);
#pragma unroll
        for (int e = 0; e < 5; ++e) out[e] += bt[e * b.nbf + q];
      }
      continue;
    }
    const long fq = upper ? FaceIdx(b, d, ni, nj, nk) : FaceIdx(b, d, i, j, k);
    const double4 a = b.fa[d][fq];
    const long c = PadIdx(b, ni, nj, nk);
    const long q = (long)ni + (long)nj * b.ni + (long)nk * b.ni * b.nj;
    T W[5], du[5], o[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      W[e] = T(b.W[e * b.npad + c]);
      du[e] = x[e * b.ncell + q];
    }
    if (block) {
      double J[25];
      NeighbourBlock(b, gd, vp, d, upper, i, j, k, J);
      T m[25];
#pragma unroll
      for (int q = 0; q < 25; ++q) m[q] = T(J[q]);
      BlockMult<T>(m, du, o);
    } else if (roe) {
      // RoeOffDiagonal: the cell's own state is the other side of the face
      const long cc = PadIdx(b, i, j, k);
      double Wd[5], Wn[5], dd[5], od[5];
#pragma unroll
      for (int e = 0; e < 5; ++e) {
        Wd[e] = b.W[e * b.npad + cc];
        Wn[e] = double(W[e]);
        dd[e] = double(du[e]);
      }
      RoeOffDiag(gd, Wn, Wd, dd, a, !upper, od);
#pragma unroll
      for (int e = 0; e < 5; ++e) o[e] = T(od[e]);
    } else {
      const T lv = vp.on ? ViscFaceSR<T>(g, W, T(a.w), T(b.c2c[d][fq]), T(b.mu[c]), vp) : T(0);
      ScalarOffDiag<T>(g, W, du, T(a.x), T(a.y), T(a.z), T(a.w), !upper, o, vp.on != 0, lv);
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) out[e] += o[e];
  }
}

// InitializeMatrixUpdate (src/linearSolver.cpp:111-144)
template <typename T>
__global__ void __launch_bounds__(256) KInitUpdate(BlockMap M, Params P) {
  const Unit U = MapUnit(M);
  const DevBlock &b = *U.b;
  const long p = (long)U.local * blockDim.x + threadIdx.x;
  if (p >= b.ncell) return;
  T *x = OpBuf<T>::X(b);
  if (!P.initMatrix && !P.rhsOnce) {
#pragma unroll
    for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = T(0);
    return;
  }
  const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
  double rb[5];
  Rhs(b, P, p, PadIdx(b, i, j, k), rb);
  if (P.rhsOnce) {
    // the right-hand side is fixed for the relax: formed once here, read by
    // every sweep (b.rhs is FP64; the operator rounds it to its precision)
#pragma unroll
    for (int e = 0; e < 5; ++e) b.rhs[e * b.ncell + p] = rb[e];
  }
  if (!P.initMatrix) {
#pragma unroll
    for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = T(0);
    return;
  }
  if (P.block) {
    T v[5], o[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) v[e] = T(rb[e]);
    TileBlockMult<T>(b, OpBuf<T>::Abi(b), i, j, k, v, o);
#pragma unroll
    for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = o[e];
    return;
  }
  const T ainv = OpBuf<T>::Ainv(b)[p];
#pragma unroll
  for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = T(rb[e]) * ainv;
}

// lusgs::LUSGS_Forward over one hyperplane (src/linearSolver.cpp:341-383)
template <typename T>
__global__ void __launch_bounds__(64) KLusgsForward(DevBlock b, Params P,
                                                    const int *__restrict__ cells, int n,
                                                    int useUpper) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long p = cells[t];
  const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
  const GasT<T> g = Cast<T>(P.gas);
  T *x = OpBuf<T>::X(b);
  T L[5];
  OffDiagSum<T>(b, g, P.visc, x, i, j, k, false, L, P.block, P.roeJac, &P.gas);
  if (useUpper) {
    T U[5];
    OffDiagSum<T>(b, g, P.visc, x, i, j, k, true, U, P.block, P.roeJac, &P.gas);
#pragma unroll
    for (int e = 0; e < 5; ++e) L[e] -= U[e];
  }
  double rb[5];
  Rhs(b, P, p, PadIdx(b, i, j, k), rb);
  if (P.block) {
    T v[5], o[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) v[e] = T(rb[e]) + L[e];
    TileBlockMult<T>(b, OpBuf<T>::Abi(b), i, j, k, v, o);
#pragma unroll
    for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = o[e];
    return;
  }
  const T ainv = OpBuf<T>::Ainv(b)[p];
#pragma unroll
  for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = (T(rb[e]) + L[e]) * ainv;
}

// lusgs::LUSGS_Backward over one hyperplane (src/linearSolver.cpp:385-428)
template <typename T>
__global__ void __launch_bounds__(64) KLusgsBackward(DevBlock b, Params P,
                                                     const int *__restrict__ cells, int n,
                                                     int full) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long p = cells[t];
  const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
  const GasT<T> g = Cast<T>(P.gas);
  T *x = OpBuf<T>::X(b);
  T U[5];
  OffDiagSum<T>(b, g, P.visc, x, i, j, k, true, U, P.block, P.roeJac, &P.gas);
  if (P.block) {
    T v[5], o[5];
    if (full) {
      T L[5];
      OffDiagSum<T>(b, g, P.visc, x, i, j, k, false, L, P.block, P.roeJac, &P.gas);
      double rb[5];
      Rhs(b, P, p, PadIdx(b, i, j, k), rb);
#pragma unroll
      for (int e = 0; e < 5; ++e) v[e] = (T(rb[e]) + L[e]) - U[e];
      TileBlockMult<T>(b, OpBuf<T>::Abi(b), i, j, k, v, o);
#pragma unroll
      for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = o[e];
    } else {
      TileBlockMult<T>(b, OpBuf<T>::Abi(b), i, j, k, U, o);
#pragma unroll
      for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = x[e * b.ncell + p] - o[e];
    }
    return;
  }
  const T ainv = OpBuf<T>::Ainv(b)[p];
  if (full) {
    T L[5];
    OffDiagSum<T>(b, g, P.visc, x, i, j, k, false, L, 0, P.roeJac);
    double rb[5];
    Rhs(b, P, p, PadIdx(b, i, j, k), rb);
#pragma unroll
    for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = ((T(rb[e]) + L[e]) - U[e]) * ainv;
  } else {
#pragma unroll
    for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = x[e * b.ncell + p] - U[e] * ainv;
  }
}

// dplur::DPLUR Jacobi sweep (src/linearSolver.cpp:473-507): reads the copy
// xold, writes x
template <typename T>
__global__ void __launch_bounds__(256) KDplur(DevBlock b, Params P) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= b.ncell) return;
  const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
  const GasT<T> g = Cast<T>(P.gas);
  const T *xo = OpBuf<T>::Xold(b);
  T L[5], U[5];
  OffDiagSum<T>(b, g, P.visc, xo, i, j, k, false, L, P.block, P.roeJac, &P.gas);
  OffDiagSum<T>(b, g, P.visc, xo, i, j, k, true, U, P.block, P.roeJac, &P.gas);
  double rb[5];
  Rhs(b, P, p, PadIdx(b, i, j, k), rb);
  T *x = OpBuf<T>::X(b);
  if (P.block) {
    T v[5], o[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) v[e] = (T(rb[e]) + T(0)) + (L[e] - U[e]);
    TileBlockMult<T>(b, OpBuf<T>::Abi(b), i, j, k, v, o);
#pragma unroll
    for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = o[e];
    return;
  }
  const T ainv = OpBuf<T>::Ainv(b)[p];
#pragma unroll
  for (int e = 0; e < 5; ++e) x[e * b.ncell + p] = ((T(rb[e]) + T(0)) + (L[e] - U[e])) * ainv;
}

template <typename T>
__global__ void __launch_bounds__(256) KCopyUpdate(DevBlock b) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= 5 * b.ncell) return;
  OpBuf<T>::Xold(b)[p] = OpBuf<T>::X(b)[p];
}

// ---------------------------------------------------------------------------
// deterministic reductions: each workgroup reduces its 256 values in a fixed
// tree and writes one partial; a single-workgroup kernel sums the partials in
// index order.

constexpr int kRedThreads = 256;
constexpr int kNormSlots = 7;  // 5 x sum R^2, linf value, matrix-residual sum

struct Partial {
  double l2[5];
  double linf;
  long long linfIdx;  // ((cellBase + p) * 5 + e), -1 if none
  double mres;
};

__device__ __forceinline__ bool LinfBetter(double v, long long idx, double bv, long long bidx) {
  if (idx < 0) return false;
  if (bidx < 0) return true;
  return v > bv || (v == bv && idx < bidx);
}

// linearSolver::Residual / AXmB (src/linearSolver.cpp:58-109), squared and
// summed as in mgSolution::CycleAtLevel (src/mgSolution.cpp:197-206)
template <typename T>
__global__ void __launch_bounds__(kRedThreads) KMatrixResidual(BlockMap M, Params P,
                                                               Partial *part) {
  __shared__ double s[kRedThreads];
  // the partial of unit n covers the same cells whichever XCD runs it
  const Unit U = MapUnit(M);
  const DevBlock &b = *U.b;
  const unsigned n = U.global;
  const long p = (long)U.local * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (p < b.ncell) {
    const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
    const GasT<T> g = Cast<T>(P.gas);
    const T *x = OpBuf<T>::X(b);
    T L[5], U[5];
    OffDiagSum<T>(b, g, P.visc, x, i, j, k, false, L, P.block, P.roeJac, &P.gas);
    OffDiagSum<T>(b, g, P.visc, x, i, j, k, true, U, P.block, P.roeJac, &P.gas);
    double rb[5];
    Rhs(b, P, p, PadIdx(b, i, j, k), rb);
    const double A = P.block ? 0.0 : b.A[p];
    double axb[5] = {0, 0, 0, 0, 0};
    if (P.block) {
      double xv[5];
#pragma unroll
      for (int e = 0; e < 5; ++e) xv[e] = (double)x[e * b.ncell + p];
      CellBlockMult<double>(b, b.Ab, p, xv, axb);
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const double ax = P.block ? axb[e] : (double)x[e * b.ncell + p] * A;
      const double axmb = (ax - (double)(L[e] - U[e])) - rb[e];
      const double r = 0.0 - axmb;
      acc += r * r;
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[n].mres = s[0];
}

// procBlock::UpdateBlock (src/procBlock.cpp:826-871) with
// ImplicitTimeAdvance (:902-916) or ExplicitEulerTimeAdvance (:880-896)
template <typename T>
__global__ void __launch_bounds__(kRedThreads) KUpdate(BlockMap M, Params P, Partial *part) {
  __shared__ double sl2[5][kRedThreads];
  __shared__ double sv[kRedThreads];
  __shared__ long long si[kRedThreads];
  const Unit U = MapUnit(M);
  const DevBlock &b = *U.b;
  const long p = (long)U.local * blockDim.x + threadIdx.x;
  double l2[5] = {0, 0, 0, 0, 0};
  double bv = 0.0;
  long long bi = -1;
  if (p < b.ncell) {
    const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
    const long c = PadIdx(b, i, j, k);
    const GasT<double> g = Cast<double>(P.gas);
    double W[5], Wn[5], R[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      W[e] = b.W[e * b.npad + c];
      R[e] = b.R[e * b.ncell + p];
    }
    if (P.explicitEuler) {
      double U[5];
      PrimToCons(g, W, U);
      const double f = b.dt[p] / b.vol[p];
#pragma unroll
      for (int e = 0; e < 5; ++e) U[e] -= f * R[e];
      ConsToPrim(g, U, Wn);
    } else {
      const T *x = OpBuf<T>::X(b);
      double du[5];
#pragma unroll
      for (int e = 0; e < 5; ++e) du[e] = (double)x[e * b.ncell + p];
      UpdatePrim(g, W, du, Wn);
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) b.W[e * b.npad + c] = Wn[e];
    if (!(Wn[0] > 0.0) || !(Wn[4] > 0.0)) atomicOr(b.flag, 1);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      l2[e] = R[e] * R[e];
      if (R[e] > bv) {  // strict '>' from 0: first maximum wins
        bv = R[e];
        bi = (long long)(b.cellBase + p) * 5 + e;
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 5; ++e) sl2[e][threadIdx.x] = l2[e];
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int e = 0; e < 5; ++e) sl2[e][threadIdx.x] += sl2[e][threadIdx.x + w];
      if (LinfBetter(sv[threadIdx.x + w], si[threadIdx.x + w], sv[threadIdx.x], si[threadIdx.x])) {
        sv[threadIdx.x] = sv[threadIdx.x + w];
        si[threadIdx.x] = si[threadIdx.x + w];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    Partial &o = part[U.global];
#pragma unroll
    for (int e = 0; e < 5; ++e) o.l2[e] = sl2[e][0];
    o.linf = sv[0];
    o.linfIdx = si[0];
  }
}

// the norms of one case: its partials are [first, first + n) (its blocks'
// 256-cell units, in block order), written to its history ring
struct NormCase {
  int first, n;
  double *hist;
  int *counter;
};

// one workgroup per case: sums the case's partials in index order (a tree
// that depends only on the case's own partial count, so a case's norms do
// not depend on which other cases share the launch); writes the next slot
// of the case's ring = {l2[5], linf, linfIdx (as bits), mres}
__global__ void __launch_bounds__(kRedThreads) KReduceNorms(const Partial *__restrict__ allPart,
                                                           const NormCase *__restrict__ cases,
                                                           int withMres) {
  const NormCase nc = cases[blockIdx.x];
  const Partial *part = allPart + nc.first;
  const int n = nc.n;
  double *hist = nc.hist;
  int *counter = nc.counter;
  __shared__ double sl2[5][kRedThreads];
  __shared__ double sm[kRedThreads];
  __shared__ double sv[kRedThreads];
  __shared__ long long si[kRedThreads];
  double l2[5] = {0, 0, 0, 0, 0}, m = 0.0, bv = 0.0;
  long long bi = -1;
  for (int q = threadIdx.x; q < n; q += kRedThreads) {
#pragma unroll
    for (int e = 0; e < 5; ++e) l2[e] += part[q].l2[e];
    if (withMres) m += part[q].mres;
    if (LinfBetter(part[q].linf, part[q].linfIdx, bv, bi)) {
      bv = part[q].linf;
      bi = part[q].linfIdx;
    }
  }
#pragma unroll
  for (int e = 0; e < 5; ++e) sl2[e][threadIdx.x] = l2[e];
  sm[threadIdx.x] = m;
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = kRedThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
#pragma unroll
      for (int e = 0; e < 5; ++e) sl2[e][threadIdx.x] += sl2[e][threadIdx.x + w];
      sm[threadIdx.x] += sm[threadIdx.x + w];
      if (LinfBetter(sv[threadIdx.x + w], si[threadIdx.x + w], sv[threadIdx.x], si[threadIdx.x])) {
        sv[threadIdx.x] = sv[threadIdx.x + w];
        si[threadIdx.x] = si[threadIdx.x + w];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int slot = *counter;
    double *o = hist + (long)slot * 8;
#pragma unroll
    for (int e = 0; e < 5; ++e) o[e] = sl2[e][0];
    o[5] = sv[0];
    o[6] = __longlong_as_double(si[0]);
    o[7] = sm[0];
    *counter = slot + 1;
  }
}

__global__ void KSetScalar(double *dst, double v) { *dst = v; }
__global__ void KSetInt(int *dst, int v) { *dst = v; }

// input::CalcCFL (src/input.cpp:647-649) of every case of a group at its own
// outer step, which then advances by one
struct CflCase {
  double start, step, max;
  double *cfl;
  int *nn;
};
// restart every case's norm ring at slot 0
__global__ void KZeroCounters(const NormCase *__restrict__ cs, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) *cs[q].counter = 0;
}

__global__ void KCaseCfl(const CflCase *__restrict__ cs, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const CflCase c = cs[q];
  const int nn = *c.nn;
  const double v = c.start + nn * c.step;
  *c.cfl = (c.max < v) ? c.max : v;
  *c.nn = nn + 1;
}

// StoreOldSolution (src/procBlock.cpp:1037-1053)
__global__ void __launch_bounds__(256) KBeginStep(BlockMap M, Params P, int copyNm1) {
#pragma clang fp contract(off)
  const Unit Un = MapUnit(M);
  const DevBlock &b = *Un.b;
  const long p = (long)Un.local * blockDim.x + threadIdx.x;
  if (p >= b.ncell) return;
  const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
  const long c = PadIdx(b, i, j, k);
  const GasT<double> g = Cast<double>(P.gas);
  const double W[5] = {b.W[c], b.W[b.npad + c], b.W[2 * b.npad + c], b.W[3 * b.npad + c],
                       b.W[4 * b.npad + c]};
  double U[5];
  PrimToCons(g, W, U);
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    b.consN[e * b.ncell + p] = U[e];
    if (copyNm1) b.consNm1[e * b.ncell + p] = U[e];
  }
}

// AssignSolToTimeNm1 after the last nonlinear iteration (src/gridLevel.cpp:424-428)
__global__ void __launch_bounds__(256) KCopyNm1(DevBlock b) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= 5 * b.ncell) return;
  b.consNm1[p] = b.consN[p];
}

// batch face-flux evaluation used by the parity tests
__global__ void __launch_bounds__(256) KEvalFlux(int flux, Gas gas, const double *WL,
                                                 const double *WR, const double *n, int count,
                                                 double *F) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const GasT<double> g = Cast<double>(gas);
  double L[5], R[5], out[5];
  for (int e = 0; e < 5; ++e) {
    L[e] = WL[5 * t + e];
    R[e] = WR[5 * t + e];
  }
  if (flux == 0) RoeFlux(g, L, R, n[3 * t], n[3 * t + 1], n[3 * t + 2], out);
  else AusmFlux(g, L, R, n[3 * t], n[3 * t + 1], n[3 * t + 2], out);
  for (int e = 0; e < 5; ++e) F[5 * t + e] = out[e];
}

}  // namespace ath

namespace ath {

// ---------------------------------------------------------------------------
// Publishing a sweep's du to other workgroups.
//
// A workgroup publishes a tile's (or a pencil's) du by storing it and then a
// flag or progress count that other workgroups poll; they read the du after
// an agent-scope acquire.  An agent-scope release on this target writes back
// the XCD's L2 and then waits for the wave's stores: the write-back is there
// for stores that are not coherent at agent scope, and it cost about 100 us
// of each 760 us block-sweep pass.  The sweeps therefore store du with
// agent-scope atomic stores (written through to the agent's point of
// coherence), wait for each wave's own stores to complete, and only then
// store the flag; no write-back is needed for stores that are already
// coherent.
template <typename T>
__device__ __forceinline__ void StoreCoherent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wait until every store the wave issued has completed
__device__ __forceinline__ void WaitStores() { __builtin_amdgcn_s_waitcnt(0); }

// ---------------------------------------------------------------------------
// tiled, dataflow-synchronised LU-SGS sweep (lusgs::LUSGS_Forward/Backward,
// src/linearSolver.cpp:341-428).
//
// The hyperplane order of HyperplaneReorder (src/utility.cpp:377-398) is one
// topological order of the Gauss-Seidel dependencies (a cell needs its lower
// neighbours in the forward sweep, its upper neighbours in the backward
// sweep); any other topological order gives the same result bit for bit.
// This kernel uses a blocked one: the block is cut into kLT = 8x8x4 tiles,
// tiles are taken in tile-hyperplane order from an atomic ticket by
// persistent workgroups, and a tile starts once the (at most three) tiles it
// depends on have published their du.  Inside a tile the workgroup walks the
// 18 intra-tile hyperplanes with LDS-resident state, du, right-hand side and
// face vectors.  Cross-tile visibility: du is stored agent-coherent and the
// per-tile flag, carrying the sweep's epoch, is stored once those stores have
// completed (StoreCoherent / WaitStores); readers acquire at agent scope.
// Every wait is bounded and reports through `err` instead of hanging.
constexpr int kLTI = 8, kLTJ = 8, kLTK = 4;
constexpr int kLT = kLTI * kLTJ * kLTK;  // 256 cells, one per thread
constexpr int kLTPlanes = kLTI + kLTJ + kLTK - 2;
constexpr int kHaloI = kLTJ * kLTK, kHaloJ = kLTI * kLTK, kHaloK = kLTI * kLTJ;
constexpr int kHalo = kHaloI + kHaloJ + kHaloK;
constexpr long kSpinLimit = 1L << 20;

struct SweepCtl {
  int *ticket;      // reset to 0 before each sweep
  int *epoch;       // incremented before each sweep
  int *flags;       // one per tile, holds the epoch of its last publication
  int *err;         // bit 1: a dependency wait exceeded the spin limit
  const int4 *tiles;  // (a, b, c, -) in tile-hyperplane order
  int ntiles, ta, tb, tc;  // tile count and tile grid dims
};

__global__ void KSweepPrologue(SweepCtl s) {
  *s.ticket = 0;
  *s.epoch = *s.epoch + 1;
}

__device__ __forceinline__ int TileId(const SweepCtl &s, int a, int b, int c) {
  return a + s.ta * (b + s.tb * c);
}

// off-diagonal product of one neighbour: state/du from LDS or global
template <typename T>
__device__ __forceinline__ void NbOffDiag(const GasT<T> &g, const double *Wnb, const T *du,
                                          double4 a, bool positive, T out[5],
                                          const ViscParams *vp = nullptr, double dist = 1.0,
                                          double mu = 0.0) {
  T W[5], d[5];
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    W[e] = T(Wnb[e]);
    d[e] = du[e];
  }
  const bool visc = vp && vp->on;
  const T lv = visc ? ViscFaceSR<T>(g, W, T(a.w), T(dist), T(mu), *vp) : T(0);
  ScalarOffDiag<T>(g, W, d, T(a.x), T(a.y), T(a.z), T(a.w), positive, out, visc, lv);
}

// ---------------------------------------------------------------------------
// interblock connections

// ghost cell <- partner interior cell, all layers (SwapStateSlice,
// src/procBlock.cpp:3050-3055; sources are interior cells, so the swaps are
// independent of each other)
struct HaloTask {
  int dst, src;
  long dpad, spad;
};

__global__ void __launch_bounds__(256) KHaloState(const DevBlock *__restrict__ blocks,
                                                  const HaloTask *__restrict__ tasks, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const HaloTask h = tasks[t];
  const DevBlock &A = blocks[h.dst], &B = blocks[h.src];
#pragma unroll
  for (int e = 0; e < 5; ++e) A.W[e * A.npad + h.dpad] = B.W[e * B.npad + h.spad];
}

// ghost cell velocity gradients <- partner interior cells
// (SwapEddyViscAndGradients, src/gridLevel.cpp:385-387), block solvers with
// the viscous terms
__global__ void __launch_bounds__(256) KHaloGrad(const DevBlock *__restrict__ blocks,
                                                 const HaloTask *__restrict__ tasks, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const HaloTask h = tasks[t];
  const DevBlock &A = blocks[h.dst], &B = blocks[h.src];
#pragma unroll
  for (int q = 0; q < 9; ++q) A.vgrad[q * A.npad + h.dpad] = B.vgrad[q * B.npad + h.spad];
}

// the off-diagonal product of the ghost neighbour across one connection
// face, from the partner's current du (SwapImplicitUpdate before each pass,
// src/linearSolver.cpp:444-463, then OffDiagonal with the ghost state)
struct ConnTask {
  int blk, src, d, upper;
  long gpad, scell, face, slot;
};

template <typename T>
__global__ void __launch_bounds__(256) KConnTerms(const DevBlock *__restrict__ blocks,
                                                  const ConnTask *__restrict__ tasks, int n, Params P) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const ConnTask c = tasks[t];
  const DevBlock &A = blocks[c.blk], &B = blocks[c.src];
  const T *xb = OpBuf<T>::X(B);
  double Wn[5];
  T dn[5], o[5];
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    Wn[e] = A.W[e * A.npad + c.gpad];
    dn[e] = xb[e * B.ncell + c.scell];
  }
  if (P.block) {
    // the block is formed in FP64 from the FP64 state (with either operator
    // precision)
    double dd[5], od[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) dd[e] = double(dn[e]);
    if (P.visc.on) {
      double vgn[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) vgn[q] = A.vgrad[q * A.npad + c.gpad];
      BlockOffDiag<double>(Cast<double>(P.gas), Wn, dd, A.fa[c.d][c.face], c.upper == 0, od, &P.visc,
                           A.mu[c.gpad], A.kc[c.gpad], A.c2c[c.d][c.face], vgn);
    } else {
      BlockOffDiag<double>(Cast<double>(P.gas), Wn, dd, A.fa[c.d][c.face], c.upper == 0, od);
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) o[e] = T(od[e]);
  } else if (P.roeJac) {
    // RoeOffDiagonal with the interior cell next to the ghost as the other
    // side (a lower ghost lies below it, an upper ghost above)
    const long s = c.d == 0 ? 1 : (c.d == 1 ? (long)A.NI : (long)A.NI * A.NJ);
    const long ci = c.upper ? c.gpad - s : c.gpad + s;
    double Wd[5], dd[5], od[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      Wd[e] = A.W[e * A.npad + ci];
      dd[e] = double(dn[e]);
    }
    RoeOffDiag(Cast<double>(P.gas), Wn, Wd, dd, A.fa[c.d][c.face], c.upper == 0, od);
#pragma unroll
    for (int e = 0; e < 5; ++e) o[e] = T(od[e]);
  } else if (P.visc.on)
    NbOffDiag<T>(Cast<T>(P.gas), Wn, dn, A.fa[c.d][c.face], c.upper == 0, o, &P.visc, A.c2c[c.d][c.face],
                 A.mu[c.gpad]);
  else
    NbOffDiag<T>(Cast<T>(P.gas), Wn, dn, A.fa[c.d][c.face], c.upper == 0, o);
  T *bt = OpBuf<T>::Bt(A);
#pragma unroll
  for (int e = 0; e < 5; ++e) bt[e * A.nbf + c.slot] = o[e];
}

template <typename T, bool FWD>
__global__ void __launch_bounds__(kLT) KLusgsTile(DevBlock b, Params P, SweepCtl s, int full) {
  __shared__ double sW[5][kLT];
  __shared__ T sX[5][kLT];
  __shared__ double sFa[3][4][kLT];  // face shared with the dependency neighbour
  __shared__ double sHW[5][kHalo];   // halo (outside-tile dependency) state
  __shared__ T sHX[5][kHalo];        // halo du
  __shared__ int sTicket;
  __shared__ int sEpoch;

  const int tid = threadIdx.x;
  const int li = tid % kLTI, lj = (tid / kLTI) % kLTJ, lk = tid / (kLTI * kLTJ);
  const GasT<T> g = Cast<T>(P.gas);
  T *x = OpBuf<T>::X(b);
  const T *ainvG = OpBuf<T>::Ainv(b);
  const int sgn = FWD ? -1 : 1;  // direction of the dependency neighbours

  // Thread 0 owns the ticket and the flag.  Exactly one `tid == 0` block
  // sits between two barriers at the end of each pass (publish the tile,
  // fetch the next ticket); with separate blocks on both sides of the loop's
  // back edge the compiler may merge them into a lane-0-only path around the
  // loop header, leaving the other lanes to pass the header barrier with a
  // stale ticket.
  if (tid == 0) {
    sTicket = __hip_atomic_fetch_add(s.ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sEpoch = __hip_atomic_load(s.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const int epoch = sEpoch;
  for (;;) {
    const int t = sTicket;
    if (t >= s.ntiles) break;
    const int4 tl = s.tiles[FWD ? t : s.ntiles - 1 - t];
    const int i = tl.x * kLTI + li, j = tl.y * kLTJ + lj, k = tl.z * kLTK + lk;
    const bool valid = i < b.ni && j < b.nj && k < b.nk;
    const long p = (long)i + (long)j * b.ni + (long)k * b.ni * b.nj;

    // --- static data (independent of du) before the dependency wait; the
    // right-hand side and D^-1 of the thread's own cell stay in registers
    double rb[5] = {0, 0, 0, 0, 0};
    T ainv = T(0);
    if (valid) {
      const long c = PadIdx(b, i, j, k);
#pragma unroll
      for (int e = 0; e < 5; ++e) sW[e][tid] = b.W[e * b.npad + c];
      if (FWD || full) Rhs(b, P, p, c, rb);
      ainv = ainvG[p];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        // forward: the cell's lower face; backward: its upper face
        const int fi = i + (FWD ? 0 : (d == 0)), fj = j + (FWD ? 0 : (d == 1)),
                  fk = k + (FWD ? 0 : (d == 2));
        const int n3 = d == 0 ? b.ni : (d == 1 ? b.nj : b.nk);
        const int f3 = d == 0 ? fi : (d == 1 ? fj : fk);
        double4 a = make_double4(0, 0, 0, 0);
        if (f3 > 0 && f3 < n3) a = b.fa[d][FaceIdx(b, d, fi, fj, fk)];
        sFa[d][0][tid] = a.x;
        sFa[d][1][tid] = a.y;
        sFa[d][2][tid] = a.z;
        sFa[d][3][tid] = a.w;
      }
    }
    // halo state: dependency neighbours just outside the tile
    {
      const int di = FWD ? -1 : kLTI, dj = FWD ? -1 : kLTJ, dk = FWD ? -1 : kLTK;
      if (tid < kHalo) {
        int hi, hj, hk;
        if (tid < kHaloI) { hi = di; hj = tid % kLTJ; hk = tid / kLTJ; }
        else if (tid < kHaloI + kHaloJ) { const int q = tid - kHaloI; hi = q % kLTI; hj = dj; hk = q / kLTI; }
        else { const int q = tid - kHaloI - kHaloJ; hi = q % kLTI; hj = q / kLTI; hk = dk; }
        const int gi = tl.x * kLTI + hi, gj = tl.y * kLTJ + hj, gk = tl.z * kLTK + hk;
        if (gi >= 0 && gj >= 0 && gk >= 0 && gi < b.ni && gj < b.nj && gk < b.nk) {
          const long c = PadIdx(b, gi, gj, gk);
#pragma unroll
          for (int e = 0; e < 5; ++e) sHW[e][tid] = b.W[e * b.npad + c];
        }
      }
    }

    // --- wait for the tiles this one depends on
    if (tid == 0) {
      const int deps[3][3] = {{tl.x + sgn, tl.y, tl.z}, {tl.x, tl.y + sgn, tl.z}, {tl.x, tl.y, tl.z + sgn}};
      for (int q = 0; q < 3; ++q) {
        const int a = deps[q][0], bb = deps[q][1], cc = deps[q][2];
        if (a < 0 || bb < 0 || cc < 0 || a >= s.ta || bb >= s.tb || cc >= s.tc) continue;
        const int id = TileId(s, a, bb, cc);
        long spins = 0;
        while (__hip_atomic_load(&s.flags[id], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
          // once any wait of the launch has timed out, stop waiting at all
          if ((++spins & 255) == 0 &&
              (__hip_atomic_load(s.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2))
            break;
          if (spins > kSpinLimit) {
            __hip_atomic_fetch_or(s.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      // one acquire for the workgroup, ordered before every wave's halo
      // loads by the barrier
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();

    // halo du, published by the dependency tiles
    {
      const int di = FWD ? -1 : kLTI, dj = FWD ? -1 : kLTJ, dk = FWD ? -1 : kLTK;
      if (tid < kHalo) {
        int hi, hj, hk;
        if (tid < kHaloI) { hi = di; hj = tid % kLTJ; hk = tid / kLTJ; }
        else if (tid < kHaloI + kHaloJ) { const int q = tid - kHaloI; hi = q % kLTI; hj = dj; hk = q / kLTI; }
        else { const int q = tid - kHaloI - kHaloJ; hi = q % kLTI; hj = q / kLTI; hk = dk; }
        const int gi = tl.x * kLTI + hi, gj = tl.y * kLTJ + hj, gk = tl.z * kLTK + hk;
        if (gi >= 0 && gj >= 0 && gk >= 0 && gi < b.ni && gj < b.nj && gk < b.nk) {
          const long q = (long)gi + (long)gj * b.ni + (long)gk * b.ni * b.nj;
#pragma unroll
          for (int e = 0; e < 5; ++e) sHX[e][tid] = x[e * b.ncell + q];
        }
      }
    }
    if (!FWD && valid) {
#pragma unroll
      for (int e = 0; e < 5; ++e) sX[e][tid] = x[e * b.ncell + p];  // forward result
    }
    __syncthreads();

    // --- intra-tile hyperplanes
    const int plane = li + lj + lk;
    for (int st = 0; st < kLTPlanes; ++st) {
      const int pl = FWD ? st : kLTPlanes - 1 - st;
      if (valid && plane == pl) {
        T D[5];  // dependency-side sum: L (forward) or U (backward)
#pragma unroll
        for (int e = 0; e < 5; ++e) D[e] = T(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int gi = i + (d == 0 ? sgn : 0), gj = j + (d == 1 ? sgn : 0), gk = k + (d == 2 ? sgn : 0);
          if (gi < 0 || gj < 0 || gk < 0 || gi >= b.ni || gj >= b.nj || gk >= b.nk) {
            // ghost neighbour: the connection term (zero off connections)
            const T *bt = OpBuf<T>::Bt(b);
            if (bt) {
              const long q = BtSlot(b, d, FWD ? 0 : 1, i, j, k);
#pragma unroll
              for (int e = 0; e < 5; ++e) D[e] += bt[e * b.nbf + q];
            }
            continue;
          }
          const int l3 = d == 0 ? li : (d == 1 ? lj : lk);
          const int L3 = d == 0 ? kLTI : (d == 1 ? kLTJ : kLTK);
          const bool inTile = FWD ? (l3 > 0) : (l3 < L3 - 1);
          double Wn[5];
          T dn[5];
          if (inTile) {
            const int nt = tid + sgn * (d == 0 ? 1 : (d == 1 ? kLTI : kLTI * kLTJ));
#pragma unroll
            for (int e = 0; e < 5; ++e) {
              Wn[e] = sW[e][nt];
              dn[e] = sX[e][nt];
            }
          } else {
            const int h = d == 0 ? lj + kLTJ * lk : (d == 1 ? kHaloI + li + kLTI * lk : kHaloI + kHaloJ + li + kLTI * lj);
#pragma unroll
            for (int e = 0; e < 5; ++e) {
              Wn[e] = sHW[e][h];
              dn[e] = sHX[e][h];
            }
          }
          const double4 a = make_double4(sFa[d][0][tid], sFa[d][1][tid], sFa[d][2][tid], sFa[d][3][tid]);
          T o[5];
          if (P.visc.on) {
            const long fq = FWD ? FaceIdx(b, d, i, j, k) : FaceIdx(b, d, gi, gj, gk);
            NbOffDiag<T>(g, Wn, dn, a, FWD, o, &P.visc, b.c2c[d][fq], b.mu[PadIdx(b, gi, gj, gk)]);
          } else {
            NbOffDiag<T>(g, Wn, dn, a, FWD, o);
          }
#pragma unroll
          for (int e = 0; e < 5; ++e) D[e] += o[e];
        }
        T xn[5];
        if (full) {
          // the other side, from the previous sweep's values in global memory
          T O[5];
          OffDiagSum<T>(b, g, P.visc, x, i, j, k, FWD, O);
          if (FWD) {
#pragma unroll
            for (int e = 0; e < 5; ++e) xn[e] = (T(rb[e]) + (D[e] - O[e])) * ainv;
          } else {
#pragma unroll
            for (int e = 0; e < 5; ++e) xn[e] = ((T(rb[e]) + O[e]) - D[e]) * ainv;
          }
        } else if (FWD) {
#pragma unroll
          for (int e = 0; e < 5; ++e) xn[e] = (T(rb[e]) + D[e]) * ainv;
        } else {
#pragma unroll
          for (int e = 0; e < 5; ++e) xn[e] = sX[e][tid] - D[e] * ainv;
        }
#pragma unroll
        for (int e = 0; e < 5; ++e) {
          sX[e][tid] = xn[e];
          StoreCoherent(&x[e * b.ncell + p], xn[e]);
        }
      }
      __syncthreads();
    }

    // --- publish: every wave's du stores have completed before the barrier,
    // and thread 0 stores the flag after it; the same thread takes the next
    // ticket
    WaitStores();
    __syncthreads();
    if (tid == 0) {
      __hip_atomic_store(&s.flags[TileId(s, tl.x, tl.y, tl.z)], epoch, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      sTicket = __hip_atomic_fetch_add(s.ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Block LU-SGS sweep over all blocks in one launch (blusgs).
//
// The blocks of a pass are independent (a connection's partner update enters
// lagged, through the precomputed connection terms), so one persistent launch
// sweeps them together: the tiles of every block are queued by their tile
// hyperplane (then block, then k, j, i), each workgroup takes the next ticket,
// waits for the (at most three) tiles of the same block it depends on, and
// walks the tile's cell hyperplanes.  Per cell, the dependency side's three
// off-diagonal blocks (formed once per relax) and D^-1 are read before the
// wait and stay in registers; the arithmetic of each cell is that of
// KLusgsForward / KLusgsBackward, so every order gives the same bits.

struct MultiSweepCtl {
  int *ticket;          // reset to 0 before each sweep
  int *epoch;           // incremented before each sweep
  int *flags;           // one per tile of every block, the epoch of its publication
  int *err;             // bit 1: a dependency wait exceeded the spin limit
  const int4 *tiles;    // (a, b, c, block) in pass order
  const int4 *grid;     // per block: tile grid dims and the block's first flag
  int ntiles;
};

__global__ void KMultiSweepPrologue(MultiSweepCtl s) {
  *s.ticket = 0;
  *s.epoch = *s.epoch + 1;
}

// tiles of 8 x 8 x 4 cells (four wavefronts).  Measured on the multiblock
// channel: 4 x 4 x 4 tiles (more tickets and flags) take 1.3 times as long
// per pass, 16 x 4 x 4 tiles (whole cache lines per tile row, but a longer
// walk of 22 cell hyperplanes) 1.12 times as long.
constexpr int kBTI = 8, kBTJ = 8, kBTK = 4;
constexpr int kBT = kBTI * kBTJ * kBTK;
constexpr int kBTPlanes = kBTI + kBTJ + kBTK - 2;
constexpr int kBHaloI = kBTJ * kBTK, kBHaloJ = kBTI * kBTK, kBHaloK = kBTI * kBTJ;
constexpr int kBHalo = kBHaloI + kBHaloJ + kBHaloK;

template <typename T, bool FWD>
__global__ void __launch_bounds__(kBT, sizeof(T) == 4 ? 2 : 1) KBlockSweepTiles(const DevBlock *__restrict__ blocks, Params P,
                                                        MultiSweepCtl s, int full) {
  __shared__ T sA[5][kBT + kBHalo];  // du of the tile's cells, then of the halo columns
  __shared__ int sTicket, sNext;
  __shared__ int sEpoch;

  const int tid = threadIdx.x;
  const int li = tid % kBTI, lj = (tid / kBTI) % kBTJ, lk = tid / (kBTI * kBTJ);
  const GasT<T> g = Cast<T>(P.gas);
  const int sgn = FWD ? -1 : 1;  // direction of the dependency neighbours

  // thread 0 owns the ticket and the flag (see KLusgsTile)
  if (tid == 0) {
    sTicket = __hip_atomic_fetch_add(s.ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sEpoch = __hip_atomic_load(s.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const int epoch = sEpoch;
  for (;;) {
    const int t = sTicket;
    if (t >= s.ntiles) break;
    const int4 tl = s.tiles[FWD ? t : s.ntiles - 1 - t];
    const DevBlock &b = blocks[tl.w];
    const int4 gr = s.grid[tl.w];
    T *x = OpBuf<T>::X(b);
    const int i = tl.x * kBTI + li, j = tl.y * kBTJ + lj, k = tl.z * kBTK + lk;
    const bool valid = i < b.ni && j < b.nj && k < b.nk;
    const long p = (long)i + (long)j * b.ni + (long)k * b.ni * b.nj;

    // --- static data before the dependency wait
    const int ni = b.ni, nj = b.nj, nk = b.nk;
    const long ncell = b.ncell, nbf = b.nbf;
    const T *bt = OpBuf<T>::Bt(b);
    // connection term (static during the pass) of the face of cell (ci, cj,
    // ck) in direction d on the dependency side; zero without connections,
    // and zero in the first forward pass, where every du starts at zero (J 0
    // sums to +0), so the connection terms are not formed for it
    const bool connZero = FWD && !full;
    auto conn = [&](int d, int ci, int cj, int ck, int e) -> T {
      return (bt && !connZero) ? bt[e * nbf + BtSlot(b, d, FWD ? 0 : 1, ci, cj, ck)] : T(0);
    };
    // LDS column of the dependency neighbour in each direction: a cell of
    // the tile or a halo column
    int nb[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int l3 = d == 0 ? li : (d == 1 ? lj : lk);
      const int L3 = d == 0 ? kBTI : (d == 1 ? kBTJ : kBTK);
      const bool inTile = FWD ? (l3 > 0) : (l3 < L3 - 1);
      nb[d] = inTile ? tid + sgn * (d == 0 ? 1 : (d == 1 ? kBTI : kBTI * kBTJ))
                     : kBT + (d == 0 ? lj + kBTJ * lk
                                     : (d == 1 ? kBHaloI + li + kBTI * lk : kBHaloI + kBHaloJ + li + kBTI * lj));
    }
    // the right-hand side is formed in FP64 and rounded once to the
    // operator's precision; the dependency blocks are formed in FP64 and
    // held in the operator's precision, as is D^-1
    T rb[5] = {T(0), T(0), T(0), T(0), T(0)};
    T mdep[3][25], mai[25];
    T Ost[5];  // full sweeps: the other side's sum, from the previous sweep's du
    if (valid) {
      if (FWD || full) {
        double r64[5];
        Rhs(b, P, p, PadIdx(b, i, j, k), r64);
#pragma unroll
        for (int e = 0; e < 5; ++e) rb[e] = T(r64[e]);
      }
      const GasT<double> gd = Cast<double>(P.gas);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int gi = i + (d == 0 ? sgn : 0), gj = j + (d == 1 ? sgn : 0), gk = k + (d == 2 ? sgn : 0);
        const bool inside = gi >= 0 && gj >= 0 && gk >= 0 && gi < ni && gj < nj && gk < nk;
        // a neighbour outside the block enters through its connection term,
        // which its LDS column holds; the identity block passes it on
        // unchanged (x + 0 y = x, and the sum never starts from -0)
        if (inside) {
          double J[25];
          NeighbourBlock(b, gd, P.visc, d, !FWD, i, j, k, J);
#pragma unroll
          for (int q = 0; q < 25; ++q) mdep[d][q] = T(J[q]);
        } else {
#pragma unroll
          for (int q = 0; q < 25; ++q) mdep[d][q] = q % 6 == 0 ? T(1) : T(0);
        }
      }
      // the other side's du is not written before this tile publishes: its
      // cells in other tiles wait for this one, those in it are stored after
      // the walk
      if (full) OffDiagSum<T>(b, g, P.visc, x, i, j, k, FWD, Ost, 1, 0, &P.gas);
      const auto *Mi = OpBuf<T>::Abi(b) + TileBase(b, i, j, k);
#pragma unroll
      for (int q = 0; q < 25; ++q) mai[q] = T(Mi[256 * q]);
      // backward: the forward result, which only this tile overwrites
      if (!FWD) {
#pragma unroll
        for (int e = 0; e < 5; ++e) sA[e][tid] = x[e * ncell + p];
      }
    } else if (!FWD) {
      // a partial tile's position just past the block's upper face along one
      // axis is the dependency neighbour of the face cell below it
      const int oi = i >= ni, oj = j >= nj, ok = k >= nk;
      const int d = oi ? 0 : (oj ? 1 : 2);
      const bool edge = oi + oj + ok == 1 && (oi ? i == ni : (oj ? j == nj : k == nk));
#pragma unroll
      for (int e = 0; e < 5; ++e)
        sA[e][tid] = edge ? conn(d, i - (d == 0), j - (d == 1), k - (d == 2), e) : T(0);
    }
    // halo columns: positions outside the block hold connection terms now,
    // those inside the du of the dependency tiles after the wait
    int hi = 0, hj = 0, hk = 0, hd = 0;
    bool hin = false;
    if (tid < kBHalo) {
      const int di = FWD ? -1 : kBTI, dj = FWD ? -1 : kBTJ, dk = FWD ? -1 : kBTK;
      if (tid < kBHaloI) { hd = 0; hi = di; hj = tid % kBTJ; hk = tid / kBTJ; }
      else if (tid < kBHaloI + kBHaloJ) { const int q = tid - kBHaloI; hd = 1; hi = q % kBTI; hj = dj; hk = q / kBTI; }
      else { const int q = tid - kBHaloI - kBHaloJ; hd = 2; hi = q % kBTI; hj = q / kBTI; hk = dk; }
      hi += tl.x * kBTI; hj += tl.y * kBTJ; hk += tl.z * kBTK;
      hin = hi >= 0 && hj >= 0 && hk >= 0 && hi < ni && hj < nj && hk < nk;
      if (!hin) {
        // the tile cell next to it along hd
        const int ci = hi - (hd == 0 ? sgn : 0), cj = hj - (hd == 1 ? sgn : 0), ck = hk - (hd == 2 ? sgn : 0);
        const bool face = ci >= 0 && cj >= 0 && ck >= 0 && ci < ni && cj < nj && ck < nk;
#pragma unroll
        for (int e = 0; e < 5; ++e) sA[e][kBT + tid] = face ? conn(hd, ci, cj, ck, e) : T(0);
      }
    }

    // --- wait for the tiles of the same block this one depends on
    if (tid == 0) {
      const int deps[3][3] = {{tl.x + sgn, tl.y, tl.z}, {tl.x, tl.y + sgn, tl.z}, {tl.x, tl.y, tl.z + sgn}};
      for (int q = 0; q < 3; ++q) {
        const int a = deps[q][0], bb = deps[q][1], cc = deps[q][2];
        if (a < 0 || bb < 0 || cc < 0 || a >= gr.x || bb >= gr.y || cc >= gr.z) continue;
        const int id = gr.w + a + gr.x * (bb + gr.y * cc);
        long spins = 0;
        while (__hip_atomic_load(&s.flags[id], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
          if ((++spins & 255) == 0 &&
              (__hip_atomic_load(s.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2))
            break;
          if (spins > kSpinLimit) {
            __hip_atomic_fetch_or(s.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      // one acquire for the workgroup: the barrier below orders it before
      // every wave's halo loads (each wave's own fence would invalidate the
      // XCD's L2 once more)
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      // the next ticket now, so that its atomic overlaps this tile's walk (a
      // held ticket only ever waits on smaller ones, so no cycle can form)
      sNext = __hip_atomic_fetch_add(s.ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();

    // halo du, published by the dependency tiles
    if (hin) {
      const long q = (long)hi + (long)hj * ni + (long)hk * ni * nj;
#pragma unroll
      for (int e = 0; e < 5; ++e) sA[e][kBT + tid] = x[e * ncell + q];
    }
    __syncthreads();

    // --- intra-tile hyperplanes
    const int plane = li + lj + lk;
    for (int st = 0; st < kBTPlanes; ++st) {
      const int pl = FWD ? st : kBTPlanes - 1 - st;
      if (valid && plane == pl) {
        T D[5];  // dependency-side sum: L (forward) or U (backward), i, j, k order
#pragma unroll
        for (int e = 0; e < 5; ++e) D[e] = T(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          T dn[5], o[5];
#pragma unroll
          for (int e = 0; e < 5; ++e) dn[e] = sA[e][nb[d]];
          BlockMult<T>(mdep[d], dn, o);
#pragma unroll
          for (int e = 0; e < 5; ++e) D[e] += o[e];
        }
        T xn[5], v[5];
        const T *ai = mai;
        if (full) {
          if (FWD) {
#pragma unroll
            for (int e = 0; e < 5; ++e) v[e] = T(rb[e]) + (D[e] - Ost[e]);
          } else {
#pragma unroll
            for (int e = 0; e < 5; ++e) v[e] = (T(rb[e]) + Ost[e]) - D[e];
          }
          BlockMult<T>(ai, v, xn);
        } else if (FWD) {
#pragma unroll
          for (int e = 0; e < 5; ++e) v[e] = T(rb[e]) + D[e];
          BlockMult<T>(ai, v, xn);
        } else {
          T o[5];
          BlockMult<T>(ai, D, o);
#pragma unroll
          for (int e = 0; e < 5; ++e) xn[e] = sA[e][tid] - o[e];
        }
        // in LDS only during the walk: per-step global stores of scattered
        // cells would write partial cache lines
#pragma unroll
        for (int e = 0; e < 5; ++e) sA[e][tid] = xn[e];
      }
      __syncthreads();
    }
    if (valid) {
#pragma unroll
      for (int e = 0; e < 5; ++e) StoreCoherent(&x[e * ncell + p], sA[e][tid]);
    }

    // --- publish: every wave's du stores have completed before the barrier,
    // and thread 0 stores the flag after it, then takes the next ticket
    WaitStores();
    __syncthreads();
    if (tid == 0) {
      __hip_atomic_store(&s.flags[gr.w + tl.x + gr.x * (tl.y + gr.y * tl.z)], epoch, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      sTicket = sNext;
    }
    __syncthreads();
  }
}

// A full block LU-SGS sweep's other-side sums (OffDiagSum over the upper
// neighbours in a forward pass, the lower ones in a backward pass, from the
// previous sweep's du and this pass's connection terms), formed for every
// cell before the pass into xold: none of the du they read is written by the
// pass before the cell's own tile publishes, so the values are those the
// sweep would form in place.
template <typename T, bool FWD>
__global__ void __launch_bounds__(256) KOtherSide(BlockMap M, Params P) {
  const Unit U = MapUnit(M);
  const DevBlock &b = *U.b;
  const long p = (long)U.local * blockDim.x + threadIdx.x;
  if (p >= b.ncell) return;
  const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
  const GasT<T> g = Cast<T>(P.gas);
  T o[5];
  OffDiagSum<T>(b, g, P.visc, OpBuf<T>::X(b), i, j, k, FWD, o, 1, 0, &P.gas);
  T *xo = OpBuf<T>::Xold(b);
#pragma unroll
  for (int e = 0; e < 5; ++e) xo[e * b.ncell + p] = o[e];
}

// ---------------------------------------------------------------------------
// Row-parallel block LU-SGS sweep over all blocks (blusgs).
//
// The tiles, queue and flags of KBlockSweepTiles (8 x 8 x 4 cells), with two
// lanes per cell: lane h of a cell's pair holds rows 3h .. 3h+2 (rows 3 and
// 4 for h = 1) of its three dependency-side blocks, of D^-1, of the
// right-hand side and of the forward result, in registers.
//
// * Forming: lane h forms the block of direction 2h (i or k) and both lanes
//   form the j block; a lane keeps its rows and hands the pair's other rows
//   to its partner with a lane shuffle.  The right-hand side is formed once
//   per relax (by the viscous block diagonal for a one-sweep FP64 relax,
//   else by KInitUpdate), a full sweep's other-side sums by KOtherSide
//   before the pass.
// * The walk: per cell hyperplane a lane forms its rows of the dependency
//   sum and of v, the pair exchanges v, and each lane forms its rows of
//   D^-1 v.  Pairs are given to cells in hyperplane order, so a hyperplane
//   occupies one or two of the eight wavefronts.
//
// Each row is summed in the order BlockMult sums it and the dependency sum
// adds the directions in i, j, k order from +0, so the result is that of
// KLusgsForward / KLusgsBackward bit for bit.
constexpr int kPThreads = 2 * kBT;
static_assert(kPThreads == 512 && kBHalo <= kBT, "row sweep layout");

template <typename T, bool FWD>
__global__ void __launch_bounds__(kPThreads) KBlockSweepRows(const DevBlock *__restrict__ blocks, Params P,
                                                             MultiSweepCtl s, int full) {
  __shared__ T sA[5][kBT + kBHalo];  // du of the tile's cells, then of the halo columns
  __shared__ unsigned char sCell[kBT];  // pair -> cell, in hyperplane order
  __shared__ int sTicket, sNext, sEpoch;

  const int tid = threadIdx.x;
  const int sgn = FWD ? -1 : 1;  // direction of the dependency neighbours
  const int g = tid >> 1, h = tid & 1;
  const int r0 = 3 * h;           // this lane's first row
  const int nrow = h ? 2 : 3;     // its rows r0 .. r0 + nrow - 1
  const bool connZero = FWD && !full;

  if (tid < kBT) {
    // rank of cell tid among the tile's cells by (hyperplane, index)
    const int li = tid % kBTI, lj = (tid / kBTI) % kBTJ, lk = tid / (kBTI * kBTJ);
    const int pl = li + lj + lk;
    int rank = 0;
    for (int q = 0; q < kBT; ++q) {
      const int qp = q % kBTI + (q / kBTI) % kBTJ + q / (kBTI * kBTJ);
      rank += (qp < pl || (qp == pl && q < tid)) ? 1 : 0;
    }
    sCell[rank] = (unsigned char)tid;
  }
  if (tid == 0) {
    sTicket = __hip_atomic_fetch_add(s.ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sEpoch = __hip_atomic_load(s.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  const int epoch = sEpoch;
  const GasT<double> gd = Cast<double>(P.gas);
  const int wc = sCell[g];
  const int li = wc % kBTI, lj = (wc / kBTI) % kBTJ, lk = wc / (kBTI * kBTJ);
  const int plane = li + lj + lk;
  // LDS column of the dependency neighbour in each direction: a cell of the
  // tile or a halo column
  int nb[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int l3 = d == 0 ? li : (d == 1 ? lj : lk);
    const int L3 = d == 0 ? kBTI : (d == 1 ? kBTJ : kBTK);
    const bool inTile = FWD ? (l3 > 0) : (l3 < L3 - 1);
    nb[d] = inTile ? wc + sgn * (d == 0 ? 1 : (d == 1 ? kBTI : kBTI * kBTJ))
                   : kBT + (d == 0 ? lj + kBTJ * lk
                                   : (d == 1 ? kBHaloI + li + kBTI * lk : kBHaloI + kBHaloJ + li + kBTI * lj));
  }

  for (;;) {
    const int t = sTicket;
    if (t >= s.ntiles) break;
    const int4 tl = s.tiles[FWD ? t : s.ntiles - 1 - t];
    const DevBlock &b = blocks[tl.w];
    const int4 gr = s.grid[tl.w];
    T *x = OpBuf<T>::X(b);
    const int ni = b.ni, nj = b.nj, nk = b.nk;
    const long ncell = b.ncell, nbf = b.nbf;
    const T *bt = OpBuf<T>::Bt(b);
    auto conn = [&](int d, int ci, int cj, int ck, int e) -> T {
      return (bt && !connZero) ? bt[e * nbf + BtSlot(b, d, FWD ? 0 : 1, ci, cj, ck)] : T(0);
    };

    // --- static data before the dependency wait
    const int i = tl.x * kBTI + li, j = tl.y * kBTJ + lj, k = tl.z * kBTK + lk;
    const bool valid = i < ni && j < nj && k < nk;
    const long p = (long)i + (long)j * ni + (long)k * ni * nj;
    T m[3][3][5];  // [direction][slot][column]: rows r0 + slot of the dependency blocks
    T ai[3][5], rb[3], ost[3], xf[3];
#pragma unroll
    for (int sl = 0; sl < 3; ++sl) {
      rb[sl] = ost[sl] = xf[sl] = T(0);
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        ai[sl][q] = T(0);
#pragma unroll
        for (int d = 0; d < 3; ++d) m[d][sl][q] = T(0);
      }
    }
    if (valid) {
      // one dependency-side block of direction d, formed in FP64 and held in
      // the operator's precision; a neighbour outside the block enters
      // through its connection term in the LDS column, the identity block
      // passing it on unchanged (x + 0 y = x, and the sum never starts from -0)
      auto form = [&](int d, T Jt[25]) {
        const int gi = i + (d == 0 ? sgn : 0), gj = j + (d == 1 ? sgn : 0), gk = k + (d == 2 ? sgn : 0);
        const bool inside = gi >= 0 && gj >= 0 && gk >= 0 && gi < ni && gj < nj && gk < nk;
        if (inside) {
          double J[25];
          NeighbourBlock(b, gd, P.visc, d, !FWD, i, j, k, J);
#pragma unroll
          for (int q = 0; q < 25; ++q) Jt[q] = T(J[q]);
        } else {
#pragma unroll
          for (int q = 0; q < 25; ++q) Jt[q] = q % 6 == 0 ? T(1) : T(0);
        }
      };
      // (every register index below is a compile-time constant; the lane's
      // half selects between values)
      {
        // lane h forms direction 2h; each keeps its rows and sends the
        // partner's rows (rows 3-4 from lane 0, rows 0-2 from lane 1)
        T J[25];
        form(2 * h, J);
        T send[15], own[3][5];
#pragma unroll
        for (int q = 0; q < 15; ++q) send[q] = h ? J[q] : (q < 10 ? J[15 + q] : T(0));
#pragma unroll
        for (int sl = 0; sl < 3; ++sl)
#pragma unroll
          for (int q = 0; q < 5; ++q) own[sl][q] = h ? (sl < 2 ? J[15 + 5 * sl + q] : T(0)) : J[5 * sl + q];
        T recv[15];
#pragma unroll
        for (int q = 0; q < 15; ++q) recv[q] = __shfl_xor(send[q], 1, 64);
        // lane 0 receives direction k, rows 0-2; lane 1 direction i, rows 3-4
#pragma unroll
        for (int sl = 0; sl < 3; ++sl)
#pragma unroll
          for (int q = 0; q < 5; ++q) {
            const T got = (sl == 2) ? (h ? T(0) : recv[5 * sl + q]) : recv[5 * sl + q];
            m[0][sl][q] = h ? got : own[sl][q];
            m[2][sl][q] = h ? own[sl][q] : got;
          }
      }
      {
        // both lanes form direction j and keep their rows
        T J[25];
        form(1, J);
#pragma unroll
        for (int sl = 0; sl < 3; ++sl)
#pragma unroll
          for (int q = 0; q < 5; ++q) m[1][sl][q] = h ? (sl < 2 ? J[15 + 5 * sl + q] : T(0)) : J[5 * sl + q];
      }
      const auto *Mi = OpBuf<T>::Abi(b) + TileBase(b, i, j, k);
#pragma unroll
      for (int sl = 0; sl < 3; ++sl) {
        const int row = r0 + sl;
        if (row < 5) {
#pragma unroll
          for (int q = 0; q < 5; ++q) ai[sl][q] = T(Mi[256 * (5 * row + q)]);
          if (FWD || full) rb[sl] = T(b.rhs[row * ncell + p]);
          if (full) ost[sl] = OpBuf<T>::Xold(b)[row * ncell + p];
          if (!FWD && !full) xf[sl] = x[row * ncell + p];
        }
      }
    }
    // backward: a partial tile's position just past the block's upper face
    // along one axis is the dependency neighbour of the face cell below it
    if (!FWD && tid < kBT) {
      const int pi = tl.x * kBTI + tid % kBTI, pj = tl.y * kBTJ + (tid / kBTI) % kBTJ,
                pk = tl.z * kBTK + tid / (kBTI * kBTJ);
      if (!(pi < ni && pj < nj && pk < nk)) {
        const int oi = pi >= ni, oj = pj >= nj, ok = pk >= nk;
        const int d = oi ? 0 : (oj ? 1 : 2);
        const bool edge = oi + oj + ok == 1 && (oi ? pi == ni : (oj ? pj == nj : pk == nk));
#pragma unroll
        for (int e = 0; e < 5; ++e)
          sA[e][tid] = edge ? conn(d, pi - (d == 0), pj - (d == 1), pk - (d == 2), e) : T(0);
      }
    }
    // halo columns: positions outside the block hold connection terms now,
    // those inside the du of the dependency tiles after the wait
    int hi = 0, hj = 0, hk = 0, hd = 0;
    bool hin = false;
    if (tid < kBHalo) {
      const int di = FWD ? -1 : kBTI, dj = FWD ? -1 : kBTJ, dk = FWD ? -1 : kBTK;
      if (tid < kBHaloI) { hd = 0; hi = di; hj = tid % kBTJ; hk = tid / kBTJ; }
      else if (tid < kBHaloI + kBHaloJ) { const int q = tid - kBHaloI; hd = 1; hi = q % kBTI; hj = dj; hk = q / kBTI; }
      else { const int q = tid - kBHaloI - kBHaloJ; hd = 2; hi = q % kBTI; hj = q / kBTI; hk = dk; }
      hi += tl.x * kBTI; hj += tl.y * kBTJ; hk += tl.z * kBTK;
      hin = hi >= 0 && hj >= 0 && hk >= 0 && hi < ni && hj < nj && hk < nk;
      if (!hin) {
        const int ci = hi - (hd == 0 ? sgn : 0), cj = hj - (hd == 1 ? sgn : 0), ck = hk - (hd == 2 ? sgn : 0);
        const bool face = ci >= 0 && cj >= 0 && ck >= 0 && ci < ni && cj < nj && ck < nk;
#pragma unroll
        for (int e = 0; e < 5; ++e) sA[e][kBT + tid] = face ? conn(hd, ci, cj, ck, e) : T(0);
      }
    }

    // --- wait for the tiles of the same block this one depends on
    if (tid == 0) {
      const int deps[3][3] = {{tl.x + sgn, tl.y, tl.z}, {tl.x, tl.y + sgn, tl.z}, {tl.x, tl.y, tl.z + sgn}};
      for (int q = 0; q < 3; ++q) {
        const int a = deps[q][0], bb = deps[q][1], cc = deps[q][2];
        if (a < 0 || bb < 0 || cc < 0 || a >= gr.x || bb >= gr.y || cc >= gr.z) continue;
        const int id = gr.w + a + gr.x * (bb + gr.y * cc);
        long spins = 0;
        while (__hip_atomic_load(&s.flags[id], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
          if ((++spins & 255) == 0 &&
              (__hip_atomic_load(s.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 2))
            break;
          if (spins > kSpinLimit) {
            __hip_atomic_fetch_or(s.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      // one acquire for the workgroup; the barrier below orders it before
      // every wave's halo loads
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      // the next ticket now, so that its atomic overlaps this tile's walk
      sNext = __hip_atomic_fetch_add(s.ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (hin) {
      const long q = (long)hi + (long)hj * ni + (long)hk * ni * nj;
#pragma unroll
      for (int e = 0; e < 5; ++e) sA[e][kBT + tid] = x[e * ncell + q];
    }
    __syncthreads();

    // --- intra-tile hyperplanes
    for (int st = 0; st < kBTPlanes; ++st) {
      const int pl = FWD ? st : kBTPlanes - 1 - st;
      if (valid && plane == pl) {
        T dn[3][5];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
          for (int q = 0; q < 5; ++q) dn[d][q] = sA[q][nb[d]];
        // this lane's rows of the dependency-side sum, L (forward) or U
        // (backward), and of v (of D, backward)
        T v[3];
#pragma unroll
        for (int sl = 0; sl < 3; ++sl) {
          T D = T(0);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            T o = T(0);
#pragma unroll
            for (int q = 0; q < 5; ++q) o += m[d][sl][q] * dn[d][q];
            D += o;
          }
          if (full) v[sl] = FWD ? rb[sl] + (D - ost[sl]) : (rb[sl] + ost[sl]) - D;
          else v[sl] = FWD ? rb[sl] + D : D;
        }
        // the pair's five rows of v
        T w[3];
#pragma unroll
        for (int sl = 0; sl < 3; ++sl) w[sl] = __shfl_xor(v[sl], 1, 64);
        const T vv[5] = {h ? w[0] : v[0], h ? w[1] : v[1], h ? w[2] : v[2], h ? v[0] : w[0], h ? v[1] : w[1]};
#pragma unroll
        for (int sl = 0; sl < 3; ++sl) {
          const int row = r0 + sl;
          if (sl < nrow) {
            T o = T(0);
#pragma unroll
            for (int q = 0; q < 5; ++q) o += ai[sl][q] * vv[q];
            sA[row < 5 ? row : 4][wc] = (!full && !FWD) ? xf[sl] - o : o;
          }
        }
      }
      __syncthreads();
    }
    for (int q = tid; q < 5 * kBT; q += kPThreads) {
      const int e = q / kBT, c = q % kBT;
      const int ci = tl.x * kBTI + c % kBTI, cj = tl.y * kBTJ + (c / kBTI) % kBTJ, ck = tl.z * kBTK + c / (kBTI * kBTJ);
      if (ci < ni && cj < nj && ck < nk)
        StoreCoherent(&x[e * ncell + (long)ci + (long)cj * ni + (long)ck * ni * nj], sA[e][c]);
    }

    // --- publish: every wave's du stores have completed before the barrier,
    // and thread 0 stores the flag after it, then takes the next ticket
    WaitStores();
    __syncthreads();
    if (tid == 0) {
      __hip_atomic_store(&s.flags[gr.w + tl.x + gr.x * (tl.y + gr.y * tl.z)], epoch, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      sTicket = sNext;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Pencil-wavefront LU-SGS sweep.
//
// The block is cut into kPW x kPW columns ("pencils") that run the full k
// extent.  One workgroup of five waves owns a pencil; lane l of every wave
// stands for column (li, lj) = (l % kPW, l / kPW).  At pencil step s the lanes
// update the cells of local hyperplane li + lj + k = s (forward; mirrored for
// the backward sweep), so a pencil walks nk + 2 kPW - 2 steps:
//
//  * waves 0, 1, 2 evaluate the dependency-side off-diagonal product of the
//    i, j and k neighbour (and, when the sweep also needs it, the other side's
//    product from the previous sweep's du);
//  * wave 3 sums the products in the order of ImplicitLower/ImplicitUpper,
//    writes du and publishes the pencil's progress;
//  * wave 4 loads the right-hand side and D^-1 and stages them in LDS.
//
// Everything that does not depend on this sweep's du (neighbour states, face
// vectors, right-hand side, D^-1, du across a pencil face) is loaded two
// steps ahead into registers, so the step-to-step chain is LDS, barriers and
// arithmetic.  In-pencil neighbours come from LDS (the previous step's du).
// The neighbour across a pencil face comes from global memory once that
// pencil's progress counter covers it; the i- and j-neighbour pencils run at
// least kPW steps ahead, so a 64^3 block's dependency chain stays at the
// 3*64-2 hyperplanes of HyperplaneReorder plus 14 pencil handoffs.
//
// Progress: wave 3 stores du with agent-coherent stores, waits for them to
// complete, then stores the step count with an agent-scope store.  A consumer reads the
// count with agent-scope loads and issues its loads of that du only after it
// has seen a count that covers them; those loads are agent-scope (coherent)
// loads, so no acquire fence (which would drain the prefetch loads in flight)
// is needed on the consumer side.
//
// Pencils are taken from an atomic ticket in pencil-hyperplane order, so a
// pencil only ever waits on pencils whose workgroups are already running: no
// co-residency assumption.  Every wait is bounded.  Two divergent regions with
// the same condition are always separated by a barrier, so the compiler cannot
// merge them across the loop's back edge.
constexpr int kPW = 8;
constexpr int kPenLanes = kPW * kPW;


// one block's view of a pencil sweep
struct PencilCtl {
  int *prog;          // per pencil of the block: steps completed in the current sweep
  int *err;           // bit 1: a dependency wait exceeded the spin limit
  int pa, pb;         // pencil grid of the block
  void *sinkv;        // per pencil 5 x 64 operator words: stores of idle lanes
};

// the pencils of every block in one launch: entry t = (a, b, block, first
// pencil of the block), ordered by pencil hyperplane a + b, then block, then
// b, a.  A pencil depends only on pencils of its own block one hyperplane
// lower (forward; higher in the backward pass, which takes the entries in
// reverse), all of which hold earlier tickets.
struct PencilGroup {
  int *ticket;          // reset to 0 before each sweep
  int *prog;            // per pencil of every block
  int *err;
  const int4 *entries;
  int nent;             // pencils of all blocks
  void *sinkv;          // 5 x 64 operator words per pencil
  const DevBlock *blocks;
};

__global__ void __launch_bounds__(256) KPencilPrologue(PencilGroup s) {
  if (threadIdx.x == 0) *s.ticket = 0;
  for (int q = threadIdx.x; q < s.nent; q += blockDim.x) s.prog[q] = 0;
}

// pencil width (cells) in i or j
__device__ __forceinline__ int PenWidth(int n, int a) { return min(kPW, n - a * kPW); }

template <typename T>
__device__ __forceinline__ T CoherentLoad(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// spin until pencil `id` has completed `need` steps; uniform across the wave
__device__ __forceinline__ void PencilSpin(const PencilCtl &s, int id, int need, int &known) {
  long spins = 0;
  while (known < need) {
    known = __builtin_amdgcn_readfirstlane(CoherentLoad(&s.prog[id]));
    if (known >= need) break;
    if ((++spins & 255) == 0 && (__builtin_amdgcn_readfirstlane(CoherentLoad(s.err)) & 2)) break;
    if (spins > kSpinLimit) {
      __hip_atomic_fetch_or(s.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// per-lane geometry of a pencil
struct PenLane {
  int i, j, li, lj, wi, wj, off, nsteps;
  bool col;
  // k of the lane's cell at step s, or -1 when the lane is idle
  __device__ __forceinline__ int K(int s, int nk, bool fwd) const {
    const int kk = s - off;
    if (!col || kk < 0 || kk >= nk) return -1;
    return fwd ? kk : nk - 1 - kk;
  }
};

// The right-hand side b of every cell, computed before the pencil sweeps
// (it depends only on data the sweeps do not modify)
__global__ void __launch_bounds__(256) KLusgsRhs(BlockMap M, Params P) {
  const Unit U = MapUnit(M);
  const DevBlock &b = *U.b;
  const long p = (long)U.local * blockDim.x + threadIdx.x;
  if (p >= b.ncell) return;
  const int i = (int)(p % b.ni), j = (int)((p / b.ni) % b.nj), k = (int)(p / ((long)b.ni * b.nj));
  double rb[5];
  Rhs(b, P, p, PadIdx(b, i, j, k), rb);
#pragma unroll
  for (int e = 0; e < 5; ++e) b.rhs[e * b.ncell + p] = rb[e];
}

// ---- decoupled waves of the pencil sweep.
//
// A pencil's workgroup has eight waves that run their own loops and
// synchronise through step counters in LDS (no workgroup barrier in the
// loop), so a wave that waits on global memory holds up only the waves that
// need its result:
//
//  * waves 4-7 (stagers): stager j prepares every step s = j mod 4 into LDS
//    slot j.  It loads, for each lane, the three dependency-side neighbours
//    and computes everything of their off-diagonal products that does not
//    depend on this sweep's du (conservative state U, flux F(W), the face
//    normal, half the face area, the spectral radius term), the du across the
//    i and j pencil faces (after the neighbour pencil's progress covers it),
//    the right-hand side (or the cell's forward du) and D^-1, and in the
//    two-sided sweep the whole other-side sum.  It waits only for the slot to
//    be released (the owner finished step s-4) and for the neighbour pencils.
//  * waves 0-2 (neighbours): wave d finishes the dependency-side product of
//    direction d: U + du, primitive state, flux, and the Rusanov combination
//    (the operations of RusanovScalarOffDiagonal, in the same order).
//  * wave 3 (owner): sums the products in the order of ImplicitLower /
//    ImplicitUpper, forms du, passes it to the neighbour waves through LDS,
//    stores it (agent-coherent), and every kPubEvery steps waits for its
//    stores and publishes the pencil's progress.
//
// Counters are monotone; each has one writer and is written with a
// workgroup-scope release after the data it covers, read with an acquire.
// Every wait is bounded (an exceeded bound sets err bit 1 and stops waiting).
constexpr int kStagers = 4;
constexpr int kPenWaves = 4 + kStagers;
constexpr int kPubEvery = 2;
constexpr long kLdsSpinLimit = 1L << 24;

template <typename T, bool FULL>
struct PenSlot {
  T U[3][5][kPenLanes];   // neighbour conservative state
  T Fo[3][5][kPenLanes];  // neighbour flux through the face
  T n[3][3][kPenLanes];   // face unit normal
  T half[3][kPenLanes];   // 0.5 |A|
  T lam[3][kPenLanes];    // 0.5 |A| (|v.n| + c) of the neighbour
  T halo[2][5][kPenLanes];  // du across the i and j pencil faces
  double v5[5][kPenLanes];  // right-hand side, or the cell's forward du
  T ainv[kPenLanes];
  T O[FULL ? 5 : 1][kPenLanes];  // other-side sum (two-sided sweep)
};

template <typename T, bool FULL>
struct PenShared {
  PenSlot<T, FULL> slot[kStagers];
  T sX[5][kPenLanes];     // du of the owner's last step
  T sO[3][5][kPenLanes];  // dependency-side products of the current step
  int stageReady[kStagers];  // slot j holds step stageReady[j] - 1
  int oReady[3];             // wave d has written sO for oReady[d] steps
  int xReady;                // the owner has finished xReady steps
  int ticket;
};

__device__ __forceinline__ int LdsLoad(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// publish a step counter for the whole wave: the wave barrier is a
// convergent point the compiler cannot duplicate into a lane-divergent
// branch, the release fence waits for every lane's earlier LDS and global
// writes, and then one lane stores the count
__device__ __forceinline__ void LdsStore(int *p, int v) {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  if (__lane_id() == 0) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// wait until an LDS counter reaches `need`; uniform across the wave
__device__ __forceinline__ void LdsWait(const int *p, int need, int *err) {
  long spins = 0;
  while (__builtin_amdgcn_readfirstlane(LdsLoad(p)) < need) {
    if (++spins > kLdsSpinLimit) {
      __hip_atomic_fetch_or(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

__device__ __forceinline__ int Clamp(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

__device__ __forceinline__ bool InBlock(const DevBlock &b, int i, int j, int k) {
  return i >= 0 && j >= 0 && k >= 0 && i < b.ni && j < b.nj && k < b.nk;
}

// Lane divergence and the step counters.  A counter is stored after data
// written by every lane of the wave.  If the counter store sat in a
// lane-divergent region, the compiler could duplicate it into both sides of
// a branch and run the idle lanes' copy first, publishing the counter before
// the active lanes have written.  Two things prevent that: LdsStore joins
// the wave at a wave barrier and fences before a single lane stores the
// counter, and everything from a wave's loads to its counter store is
// straight-line code (idle lanes and absent neighbours compute on clamped,
// valid addresses, their results are selected away, and their stores go to
// a per-pencil sink).

// stager: prepare step st into `sl`
template <typename T, bool FWD, bool FULL>
__device__ __forceinline__ void PenStage(const DevBlock &b, const Params &P, const PencilCtl &s,
                                         PenSlot<T, FULL> &sl, const PenLane &L, int lane, int st,
                                         const int nbId[2], const int nbW[2], const int nbTotal[2], int known[2]) {
  const int sgn = FWD ? -1 : 1;
  const GasT<T> g = Cast<T>(P.gas);
  const T *x = OpBuf<T>::X(b);
  // the du across the i and j pencil faces must be published first (uniform)
  bool fresh = false;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int need = min(st + nbW[h], nbTotal[h]);
    if (nbId[h] >= 0 && known[h] < need) {
      PencilSpin(s, nbId[h], need, known[h]);
      fresh = true;
    }
  }
  if (fresh) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  const int k = L.K(st, b.nk, FWD);
  const int i = Clamp(L.i, b.ni), j = Clamp(L.j, b.nj), kc = k < 0 ? 0 : k;
  const long p = (long)i + (long)j * b.ni + (long)kc * b.ni * b.nj;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int ci = Clamp(i + (d == 0 ? sgn : 0), b.ni), cj = Clamp(j + (d == 1 ? sgn : 0), b.nj),
              ck = Clamp(kc + (d == 2 ? sgn : 0), b.nk);
    const long fq = FWD ? FaceIdx(b, d, i, j, kc) : FaceIdx(b, d, ci, cj, ck);
    const double4 a = b.fa[d][fq];
    const long c = PadIdx(b, ci, cj, ck);
    T W[5], U[5], Fo[5];
#pragma unroll
    for (int e = 0; e < 5; ++e) W[e] = T(b.W[e * b.npad + c]);
    const T nx = T(a.x), ny = T(a.y), nz = T(a.z), mag = T(a.w);
    if constexpr (sizeof(T) == 4) {
      // the FP32 operator's flux change is formed from the primitive state
      // (DeltaFlux): the slot of U carries W
#pragma unroll
      for (int e = 0; e < 5; ++e) {
        U[e] = W[e];
        Fo[e] = T(0);
      }
    } else {
      PrimToCons(g, W, U);
      PhysFlux(g, W, nx, ny, nz, Fo);
    }
    const T half = T(0.5) * mag;
    T lam = T(0.5) * mag * (fabs((W[1] * nx + W[2] * ny) + W[3] * nz) + SoS(g, W[0], W[4]));
    if (P.visc.on) lam += ViscFaceSR<T>(g, W, mag, T(b.c2c[d][fq]), T(b.mu[c]), P.visc);
    // a neighbour outside the block: the slot of Fo carries the connection
    // term instead (zero off connections), selected by the neighbour wave
    const bool outside = !InBlock(b, i + (d == 0 ? sgn : 0), j + (d == 1 ? sgn : 0), kc + (d == 2 ? sgn : 0));
    T bv[5] = {T(0), T(0), T(0), T(0), T(0)};
    const T *bt = OpBuf<T>::Bt(b);
    if (bt) {
      const long q = BtSlot(b, d, FWD ? 0 : 1, i, j, kc);
#pragma unroll
      for (int e = 0; e < 5; ++e) bv[e] = bt[e * b.nbf + q];
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      sl.U[d][e][lane] = U[e];
      sl.Fo[d][e][lane] = outside ? bv[e] : Fo[e];
    }
    sl.n[d][0][lane] = nx;
    sl.n[d][1][lane] = ny;
    sl.n[d][2][lane] = nz;
    sl.half[d][lane] = half;
    sl.lam[d][lane] = lam;
    if (d < 2) {
      // du of the neighbour; used only across the pencil face
      const long q = (long)ci + (long)cj * b.ni + (long)ck * b.ni * b.nj;
#pragma unroll
      for (int e = 0; e < 5; ++e) sl.halo[d][e][lane] = x[e * b.ncell + q];
    }
  }
#pragma unroll
  for (int e = 0; e < 5; ++e) sl.v5[e][lane] = (FWD || FULL) ? b.rhs[e * b.ncell + p] : double(x[e * b.ncell + p]);
  sl.ainv[lane] = OpBuf<T>::Ainv(b)[p];
  if (FULL) {
    // other side from the previous sweep's du, as in OffDiagSum (a missing
    // neighbour adds +0, which leaves the sum unchanged: it starts at +0 and
    // a sum of finite terms reached from +0 is never -0)
    T O[5] = {T(0), T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int hi = i - (d == 0 ? sgn : 0), hj = j - (d == 1 ? sgn : 0), hk = kc - (d == 2 ? sgn : 0);
      const bool ok = k >= 0 && InBlock(b, hi, hj, hk);
      const int ui = Clamp(hi, b.ni), uj = Clamp(hj, b.nj), uk = Clamp(hk, b.nk);
      const long fq = FWD ? FaceIdx(b, d, ui, uj, uk) : FaceIdx(b, d, i, j, kc);
      const double4 a = b.fa[d][fq];
      const long c = PadIdx(b, ui, uj, uk);
      const long q = (long)ui + (long)uj * b.ni + (long)uk * b.ni * b.nj;
      double Wn[5];
      T dn[5], u[5];
#pragma unroll
      for (int e = 0; e < 5; ++e) {
        Wn[e] = b.W[e * b.npad + c];
        dn[e] = x[e * b.ncell + q];
      }
      if (P.visc.on) NbOffDiag<T>(g, Wn, dn, a, !FWD, u, &P.visc, b.c2c[d][fq], b.mu[c]);
      else NbOffDiag<T>(g, Wn, dn, a, !FWD, u);
      T bv[5] = {T(0), T(0), T(0), T(0), T(0)};
      const T *bt = OpBuf<T>::Bt(b);
      if (bt) {
        const long qb = BtSlot(b, d, FWD ? 1 : 0, i, j, kc);
#pragma unroll
        for (int e = 0; e < 5; ++e) bv[e] = bt[e * b.nbf + qb];
      }
#pragma unroll
      for (int e = 0; e < 5; ++e) O[e] += ok ? u[e] : bv[e];
    }
#pragma unroll
    for (int e = 0; e < 5; ++e) sl.O[FULL ? e : 0][lane] = O[e];
  }
}

template <typename T, bool FWD, bool FULL>
__global__ void __launch_bounds__(kPenWaves * kPenLanes) KLusgsPencil(PencilGroup G, Params P) {
  __shared__ PenShared<T, FULL> sh;
  const int tid = threadIdx.x, wave = tid / kPenLanes, lane = tid % kPenLanes;
  if (tid == 0) sh.ticket = __hip_atomic_fetch_add(G.ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (tid < kStagers) sh.stageReady[tid] = 0;
  if (tid < 3) sh.oReady[tid] = 0;
  if (tid == 0) sh.xReady = 0;
  __syncthreads();
  const int t = __builtin_amdgcn_readfirstlane(sh.ticket);
  if (t >= G.nent) return;
  const int4 ent = G.entries[FWD ? t : G.nent - 1 - t];
  const DevBlock b = G.blocks[ent.z];
  PencilCtl s;
  s.prog = G.prog + ent.w;
  s.err = G.err;
  s.pa = (b.ni + kPW - 1) / kPW;
  s.pb = (b.nj + kPW - 1) / kPW;
  s.sinkv = static_cast<T *>(G.sinkv) + (size_t)ent.w * 5 * kPenLanes;
  const int2 pc = make_int2(ent.x, ent.y);
  PenLane L;
  L.li = lane % kPW;
  L.lj = lane / kPW;
  L.wi = PenWidth(b.ni, pc.x);
  L.wj = PenWidth(b.nj, pc.y);
  L.nsteps = L.wi + L.wj + b.nk - 2;
  L.i = pc.x * kPW + L.li;
  L.j = pc.y * kPW + L.lj;
  L.col = L.li < L.wi && L.lj < L.wj;
  L.off = FWD ? L.li + L.lj : (L.wi - 1 - L.li) + (L.wj - 1 - L.lj);
  const int pid = pc.x + s.pa * pc.y;
  const int sgn = FWD ? -1 : 1;
  const int ns = L.nsteps;

  if (wave >= 4) {
    // ---- stager j: steps j, j+4, ...
    const int jst = wave - 4;
    int nbId[2], nbW[2], nbTotal[2], known[2] = {0, 0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int na = pc.x + (h == 0 ? sgn : 0), nb = pc.y + (h == 1 ? sgn : 0);
      nbId[h] = -1;
      nbW[h] = nbTotal[h] = 0;
      if (na >= 0 && nb >= 0 && na < s.pa && nb < s.pb) {
        nbId[h] = na + s.pa * nb;
        const int nwi = PenWidth(b.ni, na), nwj = PenWidth(b.nj, nb);
        nbW[h] = h == 0 ? nwi : nwj;
        nbTotal[h] = nwi + nwj + b.nk - 2;
      }
    }
    for (int st = jst; st < ns; st += kStagers) {
      // the slot is free once the owner has finished step st - kStagers
      LdsWait(&sh.xReady, st - kStagers + 1, s.err);
      PenStage<T, FWD, FULL>(b, P, s, sh.slot[jst], L, lane, st, nbId, nbW, nbTotal, known);
      LdsStore(&sh.stageReady[jst], st + 1);
    }
  } else if (wave < 3) {
    // ---- neighbour wave d
    const int d = wave;
    const GasT<T> g = Cast<T>(P.gas);
    // in-pencil neighbour lane (clamped: lanes across the pencil face read a
    // valid slot and use the staged du instead)
    const int nl = min(max(lane + sgn * (d == 0 ? 1 : (d == 1 ? kPW : 0)), 0), kPenLanes - 1);
    const bool cross = (d == 0 && (FWD ? L.li == 0 : L.li == L.wi - 1)) ||
                       (d == 1 && (FWD ? L.lj == 0 : L.lj == L.wj - 1));
    for (int st = 0; st < ns; ++st) {
      const PenSlot<T, FULL> &sl = sh.slot[st % kStagers];
      LdsWait(&sh.stageReady[st % kStagers], st + 1, s.err);
      LdsWait(&sh.xReady, st, s.err);
      const int k = L.K(st, b.nk, FWD);
      const bool dep = k >= 0 && InBlock(b, L.i + (d == 0 ? sgn : 0), L.j + (d == 1 ? sgn : 0), k + (d == 2 ? sgn : 0));
      T du[5], U[5], Wu[5], Fn[5];
#pragma unroll
      for (int e = 0; e < 5; ++e) {
        const T hx = sl.halo[d < 2 ? d : 0][e][lane], px = sh.sX[e][nl];
        du[e] = cross ? hx : px;
        U[e] = sl.U[d][e][lane] + du[e];
      }
      if constexpr (sizeof(T) == 4) {
        // flux change from the staged primitive state (DeltaFlux)
        T Wn[5];
#pragma unroll
        for (int e = 0; e < 5; ++e) Wn[e] = sl.U[d][e][lane];
        DeltaFlux(g, Wn, du, sl.n[d][0][lane], sl.n[d][1][lane], sl.n[d][2][lane], Fn);
      } else {
        // UpdatePrim + PhysFlux of the updated state + the Rusanov terms
        ConsToPrim(g, U, Wu);
        PhysFlux(g, Wu, sl.n[d][0][lane], sl.n[d][1][lane], sl.n[d][2][lane], Fn);
      }
      const T half = sl.half[d][lane], lam = sl.lam[d][lane];
      // a neighbour outside the block contributes the staged connection
      // term (zero off connections); idle steps contribute zero
      const bool outside = k >= 0 && !dep;
#pragma unroll
      for (int e = 0; e < 5; ++e) {
        const T fc = half * (sizeof(T) == 4 ? Fn[e] : Fn[e] - sl.Fo[d][e][lane]);
        const T o = FWD ? fc + du[e] * lam : fc - du[e] * lam;
        sh.sO[d][e][lane] = dep ? o : (outside ? sl.Fo[d][e][lane] : T(0));
      }
      LdsStore(&sh.oReady[d], st + 1);
    }
  } else {
    // ---- owner
    T *x = OpBuf<T>::X(b);
    T *sink = static_cast<T *>(s.sinkv) + (size_t)pid * 5 * kPenLanes;
    for (int st = 0; st < ns; ++st) {
      const PenSlot<T, FULL> &sl = sh.slot[st % kStagers];
      LdsWait(&sh.oReady[0], st + 1, s.err);
      LdsWait(&sh.oReady[1], st + 1, s.err);
      LdsWait(&sh.oReady[2], st + 1, s.err);
      const int k = L.K(st, b.nk, FWD);
      const T ainv = sl.ainv[lane];
      // products of absent neighbours are +0 and leave D unchanged (D starts
      // at +0 and is never -0)
      T D[5], xn[5];
#pragma unroll
      for (int e = 0; e < 5; ++e) D[e] = ((T(0) + sh.sO[0][e][lane]) + sh.sO[1][e][lane]) + sh.sO[2][e][lane];
      if (FULL) {
        if (FWD) {
#pragma unroll
          for (int e = 0; e < 5; ++e) xn[e] = (T(sl.v5[e][lane]) + (D[e] - sl.O[FULL ? e : 0][lane])) * ainv;
        } else {
#pragma unroll
          for (int e = 0; e < 5; ++e) xn[e] = ((T(sl.v5[e][lane]) + sl.O[FULL ? e : 0][lane]) - D[e]) * ainv;
        }
      } else if (FWD) {
#pragma unroll
        for (int e = 0; e < 5; ++e) xn[e] = (T(sl.v5[e][lane]) + D[e]) * ainv;
      } else {
#pragma unroll
        for (int e = 0; e < 5; ++e) xn[e] = T(sl.v5[e][lane]) - D[e] * ainv;
      }
#pragma unroll
      for (int e = 0; e < 5; ++e) sh.sX[e][lane] = xn[e];
      LdsStore(&sh.xReady, st + 1);
      const long p = (long)L.i + (long)L.j * b.ni + (long)(k < 0 ? 0 : k) * b.ni * b.nj;
#pragma unroll
      for (int e = 0; e < 5; ++e) StoreCoherent(k >= 0 ? &x[e * b.ncell + p] : &sink[e * kPenLanes + lane], xn[e]);
      if ((st + 1) % kPubEvery == 0 || st + 1 == ns) {
        // the owner wave's du stores have completed before the count
        WaitStores();
        if (lane == 0)
          __hip_atomic_store(&s.prog[pid], st + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

}  // namespace ath
