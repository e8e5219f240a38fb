// Host side of the drop-in boundary: the reference's input file, Plot3D grid,
// boundary-condition tables and FP64 cell geometry, assembled on the CPU and
// handed to the device solver.
//
// Geometry is computed once, in FP64, with the operation order of
// src/plot3d.cpp:36-359 and src/procBlock.cpp:2160-2268, 6397-6411, so that the
// device sees the same volumes, face vectors and cell widths as the reference.
#pragma once

#include <array>
#include <set>
#include <string>
#include <vector>

namespace ath {

constexpr int kNeq = 5;  // rho, u, v, w, p (single species, no turbulence)

enum class Recon { Constant, MUSCL, WENO, WENOZ };
enum class Limiter { None = 0, VanAlbada = 1, Minmod = 2 };
enum class Flux { Roe = 0, AUSM = 1 };
enum class TimeInt { ExplicitEuler = 0, ImplicitEuler = 1, CrankNicholson = 2, BDF2 = 3 };
enum class Solver { LUSGS = 0, DPLUR = 1 };

// boundary condition types handled by the inviscid ghost pass
enum BcType : int {
  BC_SLIPWALL = 0,  // viscousWall is treated as slipWall in the inviscid pass
  BC_CHARACTERISTIC = 1,
  BC_SUPERSONIC_INFLOW = 2,
  BC_SUPERSONIC_OUTFLOW = 3,
  BC_STAGNATION_INLET = 4,
  BC_PRESSURE_OUTLET = 5,
};

struct BcState {  // boundaryStates entry, nondimensional
  std::string kind;
  int tag = -1;
  double p = 0, rho = 0;
  std::array<double, 3> vel{0, 0, 0};
  double p0 = 0, t0 = 0;
  std::array<double, 3> dir{0, 0, 0};
  bool isothermal = false;  // viscousWall with a given wall temperature
  double twall = 0;
  bool heatFlux = false;    // viscousWall with a given nonzero heat flux
  double qwall = 0;
  int startTag = -1, endTag = -1;             // periodic: the paired surfaces' tags
  std::array<double, 3> translation{0, 0, 0};  // periodic: start surface onto end surface
};

struct Surface {  // boundarySurface: bc type, i/j/k ranges, tag
  std::string type;
  std::array<int, 7> d{};  // imin imax jmin jmax kmin kmax tag
  int SurfType() const;     // 1..6 = imin, imax, jmin, jmax, kmin, kmax
};

struct Vec3 {
  double x = 0, y = 0, z = 0;
};

struct HostBlock {
  int ni = 0, nj = 0, nk = 0, g = 0;
  int NI = 0, NJ = 0, NK = 0;  // padded cell dims
  // padded arrays (ghost layers included), i fastest
  std::vector<double> vol;
  std::vector<Vec3> ctr;
  std::array<std::vector<std::array<double, 4>>, 3> fa;  // unit normal + mag
  std::array<std::vector<Vec3>, 3> fc;
  std::array<std::vector<double>, 3> cw;
  std::vector<double> W0;  // initial primitive state, physical cells, AoS
  std::vector<Surface> surf;

  long Cell(int i, int j, int k) const {
    return (long)(i + g) + (long)(j + g) * NI + (long)(k + g) * NI * NJ;
  }
  long Face(int d, int i, int j, int k) const {
    const int FI = NI + (d == 0), FJ = NJ + (d == 1);
    return (long)(i + g) + (long)(j + g) * FI + (long)(k + g) * FI * FJ;
  }
  long NumFaces(int d) const {
    return (long)(NI + (d == 0)) * (NJ + (d == 1)) * (NK + (d == 2));
  }
  long NumCells() const { return (long)ni * nj * nk; }
  long NumPadded() const { return (long)NI * NJ * NK; }
  int N(int d) const { return d == 0 ? ni : (d == 1 ? nj : nk); }
};

// One side of an interblock connection, seen from the block that owns the
// ghost cells (boundarySurface pairs of src/boundaryConditions.cpp:603-642).
// Patch-local cell (p1, p2) along this block's in-surface axes d1 = dir+1,
// d2 = dir+2 (mod 3) continues in the partner's patch-local cell (q1, q2):
//   c = swap ? (p2, p1) : (p1, p2);  q1 = rev1 ? pn1-1-c1 : c1;  q2 likewise.
// Ghost layer l (1..g) continues in the partner's interior layer l.
struct Connection {
  int block = 0;                   // owner of the ghost cells
  int s = 0, dir = 0, r3 = 0;      // surface type 1..6, normal axis, face index
  int r1s = 0, n1 = 0, r2s = 0, n2 = 0;
  int pb = 0, ps = 0, pdir = 0, pr3 = 0;  // partner block, surface, axis, face index
  int pr1s = 0, pn1 = 0, pr2s = 0, pn2 = 0;
  bool swap = false, rev1 = false, rev2 = false;
  bool periodic = false;  // periodic: ghost geometry stays the mirror

  void Map(int p1, int p2, bool cells, int *q1, int *q2) const;
  // (i, j, k) of this block's ghost cell at layer l and of the partner's
  // interior source cell
  void Ghost(const struct HostBlock &a, int l, int p1, int p2, int *i, int *j, int *k) const;
  void Source(const struct HostBlock &b, int l, int p1, int p2, int *i, int *j, int *k) const;
};

struct HostCase {
  std::string dir, grid;
  std::string simName;      // input file name without directory and extension (input::SimNameRoot)
  std::string fluidName = "air";
  // outputVariables (a sorted set, as the reference's std::set), with its
  // default (src/input.cpp:106)
  std::set<std::string> outputVars{"density", "vel_x", "vel_y", "vel_z", "pressure"};
  int iterations = 1, nonlinear = 1, sweeps = 1;
  double rRef = -1, tRef = -1, lRef = 1, aRef = 0;
  double dtIn = -1;
  double cflStart = 1, cflMax = 1, cflStep = 0;
  double theta = 1, zeta = 0, relax = 1, dualCFL = -1;
  double kappa = -2;
  Recon recon = Recon::Constant;
  Limiter limiter = Limiter::None;
  Flux flux = Flux::Roe;
  TimeInt timeint = TimeInt::ExplicitEuler;
  Solver solver = Solver::LUSGS;
  bool block = false;  // blusgs / bdplur: 5x5 block main diagonal and off-diagonals
  bool roeJac = false;  // inviscidFluxJacobian approximateRoe: RoeOffDiagonal off-diagonals
  double gasR = 0, gasN = 0, gamma = 0;
  // laminar viscous terms (equationSet navierStokes)
  bool viscous = false;
  bool viscRecon4 = false;                // viscousFaceReconstruction centralFourth
  std::array<double, 4> suth{};           // Sutherland C1, S (viscosity, conductivity)
  double muRef = 0, kNonDim = 0, vscale = 0, prandtl = 0, viscCoeff = 1;
  double totalCells = 0;
  std::vector<BcState> bcs;
  std::vector<HostBlock> blocks;
  std::vector<Connection> conns;

  bool Implicit() const { return timeint != TimeInt::ExplicitEuler; }
  bool MultiLevel() const { return timeint == TimeInt::BDF2; }
  bool InitMatrix() const { return solver == Solver::DPLUR || sweeps > 1; }
  // src/input.cpp:1127-1143
  int Ghosts() const {
    const int g = recon == Recon::Constant ? 1 : (recon == Recon::MUSCL ? 2 : 3);
    return viscRecon4 && g < 2 ? 2 : g;
  }
  double Cfl(int nn) const;
};

// Parse `inp_path` and its grid; returns an empty string on success or an
// error message.  Supported subset: see the README of the package.
std::string LoadCase(const std::string &inp_path, HostCase *out);

// HyperplaneReorder (src/utility.cpp:377-398): cells sorted by i+j+k, then
// k, j, i ascending.  Returns packed linear physical indices and the plane
// offsets (size planes+1).
void HyperplaneOrder(int ni, int nj, int nk, std::vector<int> *cells,
                     std::vector<int> *offsets);

}  // namespace ath
