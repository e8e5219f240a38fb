// C-ABI implementation: context, device arena, launch orchestration, graphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <fstream>
#include <limits>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <string>
#include <vector>

#include "../../include/aither_amd.h"
#include "host_case.hpp"
#include "io.hpp"
#include "kernels.hpp"
#include "sst.hpp"
#include "wall_tree.hpp"

namespace {

thread_local std::string g_err;

int Fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_OK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return Fail(ATH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));    \
  } while (0)

struct BlockRes {
  ath::DevBlock d{};
  std::vector<int> planeOff;  // host copy of hyperplane offsets
  int *cells = nullptr;       // device hyperplane cell list
  int nparts = 0;             // reduction partials of this block
  int partBase = 0;
  std::vector<ath::BcTask> tasks;
  ath::BcTask *dtasks = nullptr;
  long bcTotal = 0;
  ath::SweepCtl sweep{};  // tiled LU-SGS dataflow state
  int sweepGrid = 0;
  // viscous cases: edge ghost tasks (inviscid and viscous pass) and wall tasks
  ath::EdgeTask *edgeInv = nullptr, *edgeVisc = nullptr;
  int edgeLines = 0, edgePerLine = 0;
  ath::WallTask *wall = nullptr;
  int nwall = 0;
};

// One case resident on the device: its blocks, its device memory and the
// per-case scalars every kernel reaches through its blocks (CFL, error bits,
// boundary states), its norm history and the PrintResiduals normalisation.
struct CaseRes {
  ath::HostCase own;                  // a served case (geometry dropped after upload)
  const ath::HostCase *hc = nullptr;  // the case description (own, or the context's)
  std::vector<BlockRes> blocks;
  std::vector<void *> allocs;         // device memory owned by this case
  size_t bytes = 0;
  int partFirst = 0, nparts = 0;      // reduction partials in the current group
  ath::BcStateDev *dstates = nullptr;
  double *dcfl = nullptr;
  int *dflag = nullptr;               // bit 1 nonphysical, bit 4 singular block diagonal
  int *dnn = nullptr;                 // served cases: outer step of the next iteration
  double *dhist = nullptr;            // per-iteration norms ring
  int *dcounter = nullptr;
  int histCap = 0;
  double l2first[ATH_MAX_EQ]{};
  bool restart = false;  // started from a restart file: the normalisation is the file's
  double cfl = 0;
  double reserved = 0;  // served cases: the estimate admission reserved
  std::vector<ath::HaloTask> halo;  // interblock connections, block indices within the case
  std::vector<ath::ConnTask> conn;
  // served cases
  int id = -1, state = 0, error = 0, iters = 0, maxIters = 0;
  double drop = 0;
  ath_norms last{};
  ~CaseRes() {
    for (void *p : allocs) hipFree(p);
  }
};

}  // namespace

struct ath_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int anyPark = 0;  // some block parks its wall cells' inviscid block sums (DevBlock::invPark)
  int rhsFormed = 0;  // the last residual stage formed the relax's right-hand side (RhsInDiag)
  ath::HostCase hc;  // the case (a single-case context) or the numerics every case shares (a server)
  ath::Params P{};
  // the resident cases; blk concatenates the blocks of the cases that
  // iterate (all of them in a single-case context)
  std::vector<std::unique_ptr<CaseRes>> cases;
  std::vector<CaseRes *> active;
  std::vector<BlockRes> blk;
  std::vector<void *> allocs;       // device memory of the context itself
  std::vector<void *> groupAllocs;  // tables of the current set of active cases
  std::vector<void *> *allocTo = nullptr;  // where Alloc records (default allocs)
  size_t *bytesTo = nullptr;
  ath::Partial *dpart = nullptr;
  int nparts = 0;
  ath::NormCase *dnorm = nullptr;   // per active case: its partials and norm ring
  ath::CflCase *dcflc = nullptr;    // per active served case: its CFL ramp
  int *dflag = nullptr;             // bit 2: a sweep dependency wait exceeded its spin limit
  // all-block pencil sweep
  ath::PencilGroup pen{};
  int fp32 = 0;
  int matres = 1;
  int sweepOrder = ATH_SWEEP_PENCIL;  // ATH_SWEEP_*
  double cfl = 0;
  double l2first[ATH_MAX_EQ]{};
  // graph of one whole nonlinear iteration (ghost cells to norms)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  int graphFp32 = -1, graphMatres = -1, graphOrder = -1;
  // interblock connections: a device copy of every block's DevBlock, the
  // ghost-state gather list and the per-face connection terms
  ath::DevBlock *dblocks = nullptr;
  ath::HaloTask *dhalo = nullptr;
  int nhalo = 0;
  ath::ConnTask *dconn = nullptr;
  int nconn = 0;
  // block solvers: one tiled sweep over all blocks, 8 x 8 x 4 tiles with a
  // lane per cell (msweep, ATH_SWEEP_TILED) or two lanes per cell, each
  // holding some of the cell's rows (rsweep, ATH_SWEEP_ROWS)
  ath::MultiSweepCtl msweep{};
  int msweepGrid = 0;
  ath::MultiSweepCtl rsweep{};
  int rsweepGrid = 0;
  // unit grids of the all-block launches (ath::BlockMap), kNumMaps of them,
  // nb entries each, and each one's total unit count
  int4 *dmaps = nullptr;
  int mapTotal[4] = {0, 0, 0, 0};
  // the ghost-cell task lists of every block (ath::BlockTasks), kNumTaskMaps
  // tables of nb entries, and each one's item total
  ath::BlockTasks *dbt = nullptr;
  long taskTotal[4] = {0, 0, 0, 0};
  const ath::BlockTasks *Tasks(int q) const { return dbt + (size_t)q * blk.size(); }
  ath::BlockMap Map(int q) const { return {dblocks, dmaps + (size_t)q * blk.size(), (int)blk.size()}; }
  // timing
  int timing = 0;
  std::vector<hipEvent_t> evs;
  ath_timing tm{};
  // residual and timing logs (<simName>.resid, .tme)
  std::unique_ptr<std::ofstream> logResid, logTime;
  std::chrono::steady_clock::time_point logStart;

  ~ath_ctx() {
    DropGraph();
    for (auto e : evs) hipEventDestroy(e);
    cases.clear();
    for (void *p : groupAllocs) hipFree(p);
    for (void *p : allocs) hipFree(p);
    if (stream) hipStreamDestroy(stream);
  }
  void DropGraph() {
    if (graphExec) hipGraphExecDestroy(graphExec);
    if (graph) hipGraphDestroy(graph);
    graphExec = nullptr;
    graph = nullptr;
  }
  template <typename T>
  int Alloc(T **p, size_t n) {
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (hipMalloc(&q, bytes) != hipSuccess)
      return Fail(ATH_ERR_NOMEM, "device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed");
    (allocTo ? allocTo : &allocs)->push_back(q);
    if (bytesTo) *bytesTo += bytes;
    *p = (T *)q;
    return ATH_OK;
  }
};

namespace {

constexpr int kBlk = 256;

int Grid(long n, int b = kBlk) { return (int)((n + b - 1) / b); }

// the unit grids of the all-block launches: 8x8x4 tiles (residual and
// viscous flux), ranges of 128 cells (block diagonal), of 256 cells (matrix
// residual, update, initialisation; their partials are the reductions'
// slots) and of 256 padded cells (auxiliary variables)
enum { kMapTile = 0, kMapCell128 = 1, kMapCell256 = 2, kMapPad256 = 3, kNumMaps = 4 };
// the ghost-cell task tables: boundary patches, edge lines of the inviscid
// and of the viscous pass, no-slip wall cells
enum { kTaskBc = 0, kTaskEdgeInv = 1, kTaskEdgeVisc = 2, kTaskWall = 3, kNumTaskMaps = 4 };

// boundary-face slot, as BtSlot in the kernels
long HostBtSlot(const ath::HostBlock &h, int d, int upper, int i, int j, int k) {
  const long sI = (long)h.nj * h.nk, sJ = (long)h.nk * h.ni, sK = (long)h.ni * h.nj;
  if (d == 0) return upper * sI + j + (long)h.nj * k;
  if (d == 1) return 2 * sI + upper * sJ + k + (long)h.nk * i;
  return 2 * sI + 2 * sJ + upper * sK + i + (long)h.ni * j;
}

// boundaryConditions::GetBCSurface (src/boundaryConditions.cpp:109-190): the
// surface of the same normal direction whose patch holds face (i, j, k)
const ath::Surface *SurfaceAt(const ath::HostBlock &h, int i, int j, int k, int st) {
  const int dir = (st - 1) / 2, x[3] = {i, j, k};
  for (const auto &sf : h.surf) {
    if ((sf.SurfType() - 1) / 2 != dir) continue;
    bool ok = true;
    for (int q = 0; q < 3 && ok; ++q)
      ok = q == dir ? (x[q] >= sf.d[2 * q] && x[q] <= sf.d[2 * q + 1]) : (x[q] >= sf.d[2 * q] && x[q] < sf.d[2 * q + 1]);
    if (ok) return &sf;
  }
  return nullptr;
}

void HostMap3(int dir, int d3, int d1, int d2, int *i, int *j, int *k) {
  if (dir == 0) { *i = d3; *j = d1; *k = d2; }
  else if (dir == 1) { *i = d2; *j = d3; *k = d1; }
  else { *i = d1; *j = d2; *k = d3; }
}

// procBlock::CalcWallDistance through kdtree::NearestNeighbor
// (src/procBlock.cpp:6030-6041, src/kdtree.cpp:123-224): the distance from a
// point to the nearest centre of a viscousWall face of any block
// (GetViscousFaceCenters, src/utility.cpp:310-366), from a k-d tree whose
// search gives the full scan's minimum bit for bit (wall_tree.hpp)
std::vector<ath::Vec3> WallFaceCentres(const ath::HostCase &hc) {
  std::vector<ath::Vec3> pts;
  for (const auto &h : hc.blocks)
    for (const auto &sf : h.surf) {
      if (sf.type != "viscousWall") continue;
      const int st = sf.SurfType(), dir = (st - 1) / 2;
      const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
      for (int q2 = sf.d[2 * d2]; q2 < sf.d[2 * d2 + 1]; ++q2)
        for (int q1 = sf.d[2 * d1]; q1 < sf.d[2 * d1 + 1]; ++q1) {
          int i, j, k;
          HostMap3(dir, sf.d[2 * dir], q1, q2, &i, &j, &k);
          pts.push_back(h.fc[dir][h.Face(dir, i, j, k)]);
        }
    }
  return pts;
}

// padded geometry, projected centre distances, wall flags and the ghost-cell
// task lists of a viscous case
int UploadViscous(ath_ctx *c, const ath::HostCase &hc, const ath::HostBlock &h, BlockRes &br) {
  using namespace ath;
  DevBlock &d = br.d;
  const long np = d.npad;
  if (int r = c->Alloc(&d.T, np)) return r;
  if (int r = c->Alloc(&d.mu, np)) return r;
  HIP_OK(hipMemset(d.T, 0, np * sizeof(double)));
  HIP_OK(hipMemset(d.mu, 0, np * sizeof(double)));
  {
    double *v = nullptr;
    if (int r = c->Alloc(&v, np)) return r;
    HIP_OK(hipMemcpy(v, h.vol.data(), np * sizeof(double), hipMemcpyHostToDevice));
    d.volP = ath::Dev(v);
  }
  for (int dd = 0; dd < 3; ++dd) {
    std::vector<double4> fa(h.NumFaces(dd));
    for (size_t q = 0; q < fa.size(); ++q) {
      const auto &a = h.fa[dd][q];
      fa[q] = make_double4(a[0], a[1], a[2], a[3]);
    }
    double4 *dfa = nullptr;
    if (int r = c->Alloc(&dfa, fa.size())) return r;
    HIP_OK(hipMemcpy(dfa, fa.data(), fa.size() * sizeof(double4), hipMemcpyHostToDevice));
    d.faP[dd] = ath::Dev(dfa);
    // ProjC2CDist (src/procBlock.cpp:6316-6341) at every physical face
    const int FI = h.ni + (dd == 0), FJ = h.nj + (dd == 1), FK = h.nk + (dd == 2);
    std::vector<double> cc((size_t)FI * FJ * FK);
    for (int k = 0; k < FK; ++k)
      for (int j = 0; j < FJ; ++j)
        for (int i = 0; i < FI; ++i) {
          const Vec3 &u = h.ctr[h.Cell(i, j, k)];
          const Vec3 &l = h.ctr[h.Cell(i - (dd == 0), j - (dd == 1), k - (dd == 2))];
          const auto &a = h.fa[dd][h.Face(dd, i, j, k)];
          const double dx = u.x - l.x, dy = u.y - l.y, dz = u.z - l.z;
          cc[(size_t)i + (size_t)j * FI + (size_t)k * FI * FJ] = (dx * a[0] + dy * a[1]) + dz * a[2];
        }
    double *dc = nullptr;
    if (int r = c->Alloc(&dc, cc.size())) return r;
    HIP_OK(hipMemcpy(dc, cc.data(), cc.size() * sizeof(double), hipMemcpyHostToDevice));
    d.c2c[dd] = ath::Dev(dc);
  }
  // viscousWall flag of every boundary face (BtSlot order)
  {
    std::vector<unsigned char> wf((size_t)d.nbf, 0);
    for (int dd = 0; dd < 3; ++dd)
      for (int up = 0; up < 2; ++up) {
        const int d1 = (dd + 1) % 3, d2 = (dd + 2) % 3;
        for (int c2 = 0; c2 < h.N(d2); ++c2)
          for (int c1 = 0; c1 < h.N(d1); ++c1) {
            int i, j, k;
            HostMap3(dd, up ? h.N(dd) : 0, c1, c2, &i, &j, &k);
            const Surface *sf = SurfaceAt(h, i, j, k, 2 * dd + 1 + up);
            int ci = i, cj = j, ck = k;
            if (up) { ci -= dd == 0; cj -= dd == 1; ck -= dd == 2; }
            wf[HostBtSlot(h, dd, up, ci, cj, ck)] = sf && sf->type == "viscousWall";
          }
      }
    unsigned char *dw = nullptr;
    if (int r = c->Alloc(&dw, wf.size())) return r;
    HIP_OK(hipMemcpy(dw, wf.data(), wf.size(), hipMemcpyHostToDevice));
    d.wallf = ath::Dev(dw);
  }
  // block solvers: the cells within the ghost depth of a viscousWall face
  // along its normal, whose inviscid face states read the wall's ghost cells
  // (which the viscous pass replaces); their inviscid block sum is formed
  // before the viscous pass
  if (hc.block) {
    std::vector<unsigned char> pk((size_t)h.ni * h.nj * h.nk, 0);
    bool any = false;
    for (int dd = 0; dd < 3; ++dd)
      for (int up = 0; up < 2; ++up) {
        const int d1 = (dd + 1) % 3, d2 = (dd + 2) % 3;
        for (int c2 = 0; c2 < h.N(d2); ++c2)
          for (int c1 = 0; c1 < h.N(d1); ++c1) {
            int i, j, k;
            HostMap3(dd, up ? h.N(dd) : 0, c1, c2, &i, &j, &k);
            const Surface *sf = SurfaceAt(h, i, j, k, 2 * dd + 1 + up);
            if (!sf || sf->type != "viscousWall") continue;
            for (int layer = 0; layer < h.g && layer < h.N(dd); ++layer) {
              int ci, cj, ck;
              HostMap3(dd, up ? h.N(dd) - 1 - layer : layer, c1, c2, &ci, &cj, &ck);
              pk[(size_t)ci + (size_t)cj * h.ni + (size_t)ck * h.ni * h.nj] = 1;
              any = true;
            }
          }
      }
    if (any) {
      unsigned char *dp = nullptr;
      if (int r = c->Alloc(&dp, pk.size())) return r;
      HIP_OK(hipMemcpy(dp, pk.data(), pk.size(), hipMemcpyHostToDevice));
      d.invPark = ath::Dev(dp);
      c->anyPark = 1;
    }
  }
  // edge ghost cells, per line in the reference's layer order
  // (src/procBlock.cpp:2565-2714, 2874-3030)
  {
    const int g = h.g;
    int lines = 0;
    for (int dd = 0; dd < 3; ++dd) lines += 4 * h.N(dd);
    br.edgeLines = lines;
    br.edgePerLine = g * g;
    std::vector<EdgeTask> inv((size_t)lines * g * g), vis((size_t)lines * g * g);
    int line = 0;
    for (int dd = 0; dd < 3; ++dd) {
      const int d2 = (dd + 1) % 3, d3 = (dd + 2) % 3;
      const int max1 = h.N(dd), max2 = h.N(d2), max3 = h.N(d3);
      for (int cc = 0; cc < 4; ++cc)
        for (int d1 = 0; d1 < max1; ++d1, ++line) {
          int q = 0;
          for (int layer3 = 1; layer3 <= g; ++layer3)
            for (int layer2 = 1; layer2 <= g; ++layer2, ++q) {
              const bool upper2 = cc > 1, upper3 = cc % 2 == 1;
              const int pC2 = upper2 ? max2 + layer2 - 2 : 1 - layer2, gC2 = upper2 ? pC2 + 1 : pC2 - 1;
              const int pC3 = upper3 ? max3 + layer3 - 2 : 1 - layer3, gC3 = upper3 ? pC3 + 1 : pC3 - 1;
              const int surf2 = 2 * d2 + 1 + (upper2 ? 1 : 0), surf3 = 2 * d3 + 1 + (upper3 ? 1 : 0);
              const int cF2_2 = upper2 ? max2 : 0, cF2_3 = upper3 ? max3 - 1 : 0;
              const int cF3_2 = upper2 ? max2 - 1 : 0, cF3_3 = upper3 ? max3 : 0;
              int i, j, k;
              HostMap3(dd, d1, cF2_2, cF2_3, &i, &j, &k);
              const Surface *s2 = SurfaceAt(h, i, j, k, surf2);
              HostMap3(dd, d1, cF3_2, cF3_3, &i, &j, &k);
              const Surface *s3 = SurfaceAt(h, i, j, k, surf3);
              HostMap3(dd, d1, cF2_2, gC3, &i, &j, &k);
              const auto fa2 = h.fa[d2][h.Face(d2, i, j, k)];
              HostMap3(dd, d1, gC2, cF3_3, &i, &j, &k);
              const auto fa3 = h.fa[d3][h.Face(d3, i, j, k)];
              EdgeTask t{};
              HostMap3(dd, d1, gC2, gC3, &i, &j, &k);
              t.gq = h.Cell(i, j, k);
              HostMap3(dd, d1, pC2, gC3, &i, &j, &k);
              t.q2 = h.Cell(i, j, k);
              HostMap3(dd, d1, gC2, pC3, &i, &j, &k);
              t.q3 = h.Cell(i, j, k);
              const std::string b2 = s2 ? s2->type : "", b3 = s3 ? s3->type : "";
              const int avg = layer2 == layer3 ? 2 : (layer2 > layer3 ? 3 : 4);
              auto fill = [&](EdgeTask &e, bool slip2, bool slip3, bool keep) {
                e = t;
                if (slip2 && !slip3) {
                  e.action = 0;
                  e.nx = fa2[0]; e.ny = fa2[1]; e.nz = fa2[2];
                  e.stype = surf2;
                  e.layer = layer2;
                } else if (!slip2 && slip3) {
                  e.action = 1;
                  e.nx = fa3[0]; e.ny = fa3[1]; e.nz = fa3[2];
                  e.stype = surf3;
                  e.layer = layer3;
                } else {
                  e.action = keep ? avg : 5;
                }
              };
              const bool w2 = b2 == "slipWall" || b2 == "viscousWall", w3 = b3 == "slipWall" || b3 == "viscousWall";
              fill(inv[(size_t)line * g * g + q], w2, w3, true);
              fill(vis[(size_t)line * g * g + q], b2 == "slipWall", b3 == "slipWall",
                   b2 == "viscousWall" && b3 == "viscousWall");
            }
        }
    }
    if (int r = c->Alloc(&br.edgeInv, inv.size())) return r;
    if (int r = c->Alloc(&br.edgeVisc, vis.size())) return r;
    HIP_OK(hipMemcpy(br.edgeInv, inv.data(), inv.size() * sizeof(EdgeTask), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(br.edgeVisc, vis.data(), vis.size() * sizeof(EdgeTask), hipMemcpyHostToDevice));
  }
  // viscousWall ghost cells, all layers (src/procBlock.cpp:2760-2836)
  {
    std::vector<WallTask> wt;
    // heat-flux walls: the wall distance of the cell next to the wall, the
    // same for every ghost layer, found once per cell
    std::unordered_map<long, double> wallDist;
    std::unique_ptr<ath::NearestPoints> tree;
    for (int layer = 1; layer <= h.g; ++layer)
      for (const auto &sf : h.surf) {
        if (sf.type != "viscousWall") continue;
        const int st = sf.SurfType(), dir = (st - 1) / 2;
        const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
        const int r3 = sf.d[2 * dir], n3 = h.N(dir);
        const int gC = st % 2 == 0 ? r3 + layer - 1 : r3 - layer;
        const int iC = st % 2 == 0 ? std::max(r3 - layer, 0) : std::min(r3 + layer - 1, n3 - 1);
        int state = -1;
        for (size_t q = 0; q < hc.bcs.size(); ++q)
          if (hc.bcs[q].tag == sf.d[6]) { state = (int)q; break; }
        const bool flux = state >= 0 && hc.bcs[state].heatFlux;
        // the cell next to the wall, whatever the layer
        const int aC = st % 2 == 0 ? r3 - 1 : r3;
        for (int q2 = sf.d[2 * d2]; q2 < sf.d[2 * d2 + 1]; ++q2)
          for (int q1 = sf.d[2 * d1]; q1 < sf.d[2 * d1 + 1]; ++q1) {
            int ii, ij, ik, gi, gj, gk, ai, aj, ak;
            HostMap3(dir, iC, q1, q2, &ii, &ij, &ik);
            HostMap3(dir, gC, q1, q2, &gi, &gj, &gk);
            HostMap3(dir, aC, q1, q2, &ai, &aj, &ak);
            double wd = 0.0;
            if (flux) {
              const long ac = h.Cell(ai, aj, ak);
              auto it = wallDist.find(ac);
              if (it == wallDist.end()) {
                if (!tree) tree = std::make_unique<ath::NearestPoints>(WallFaceCentres(hc));
                it = wallDist.emplace(ac, tree->Distance(h.ctr[ac])).first;
              }
              wd = it->second;
            }
            wt.push_back({h.Cell(ii, ij, ik), h.Cell(gi, gj, gk), state, 0, wd});
          }
      }
    br.nwall = (int)wt.size();
    if (!wt.empty()) {
      if (int r = c->Alloc(&br.wall, wt.size())) return r;
      HIP_OK(hipMemcpy(br.wall, wt.data(), wt.size() * sizeof(WallTask), hipMemcpyHostToDevice));
    }
  }
  return ATH_OK;
}

// gather lists of a case's interblock connections (block indices within the
// case; the group tables add the case's first block)
void CaseConnections(const ath::HostCase &hc, CaseRes &cr) {
  using namespace ath;
  auto &halo = cr.halo;
  auto &conn = cr.conn;
  halo.clear();
  conn.clear();
  for (const auto &cn : hc.conns) {
    const HostBlock &A = hc.blocks[cn.block], &B = hc.blocks[cn.pb];
    const bool upper = cn.s % 2 == 0;
    for (int p2 = 0; p2 < cn.n2; ++p2)
      for (int p1 = 0; p1 < cn.n1; ++p1) {
        for (int l = 1; l <= A.g; ++l) {
          int gi, gj, gk, si, sj, sk;
          cn.Ghost(A, l, p1, p2, &gi, &gj, &gk);
          cn.Source(B, l, p1, p2, &si, &sj, &sk);
          halo.push_back({cn.block, cn.pb, A.Cell(gi, gj, gk), B.Cell(si, sj, sk)});
          if (l > 1) continue;
          ConnTask t{};
          t.blk = cn.block;
          t.src = cn.pb;
          t.d = cn.dir;
          t.upper = upper ? 1 : 0;
          t.gpad = A.Cell(gi, gj, gk);
          t.scell = si + (long)sj * B.ni + (long)sk * B.ni * B.nj;
          // the block's own boundary cell and face
          int bi, bj, bk;
          {
            const int n3 = upper ? A.N(cn.dir) - 1 : 0, c1 = cn.r1s + p1, c2 = cn.r2s + p2;
            bi = cn.dir == 0 ? n3 : (cn.dir == 1 ? c2 : c1);
            bj = cn.dir == 0 ? c1 : (cn.dir == 1 ? n3 : c2);
            bk = cn.dir == 0 ? c2 : (cn.dir == 1 ? c1 : n3);
          }
          const int fi = bi + (cn.dir == 0 && upper), fj = bj + (cn.dir == 1 && upper),
                    fk = bk + (cn.dir == 2 && upper);
          const int FI = A.ni + (cn.dir == 0), FJ = A.nj + (cn.dir == 1);
          t.face = fi + (long)fj * FI + (long)fk * FI * FJ;
          t.slot = HostBtSlot(A, cn.dir, t.upper, bi, bj, bk);
          conn.push_back(t);
        }
      }
  }
}

int RebuildGroup(ath_ctx *c);

int AllocFp32(ath_ctx *c, ath::DevBlock &d) {
  if (d.x32) return ATH_OK;
  if (int r = c->Alloc(&d.Ainv32, d.ncell)) return r;
  if (int r = c->Alloc(&d.x32, 5 * d.ncell)) return r;
  if (int r = c->Alloc(&d.xold32, 5 * d.ncell)) return r;
  HIP_OK(hipMemset(d.x32, 0, 5 * d.ncell * sizeof(float)));
  if (d.bt) {
    if (int r = c->Alloc(&d.bt32, 5 * d.nbf)) return r;
    HIP_OK(hipMemset(d.bt32, 0, 5 * d.nbf * sizeof(float)));
  }
  if (d.Abi) {
    // block solvers: D^-1 rounded to float in the tile-blocked layout
    const long tb = 25L * 256 * d.tni * d.tnj * d.tnk;
    if (int r = c->Alloc(&d.Abi32, tb)) return r;
    HIP_OK(hipMemset(d.Abi32, 0, tb * sizeof(float)));
  }
  return ATH_OK;
}

int EnsureFp32(ath_ctx *c) {
  for (auto &cp : c->cases) {
    c->allocTo = &cp->allocs;
    c->bytesTo = &cp->bytes;
    for (auto &br : cp->blocks) {
      const int r = AllocFp32(c, br.d);
      if (r) {
        c->allocTo = nullptr;
        c->bytesTo = nullptr;
        return r;
      }
    }
  }
  c->allocTo = nullptr;
  c->bytesTo = nullptr;
  if (int r = RebuildGroup(c)) return r;
  HIP_OK(hipDeviceSynchronize());
  return ATH_OK;
}

// Everything a case owns on the device, recorded in cr.allocs: per block the
// state, geometry, residual and operator arrays, task lists and sweep state;
// per case the boundary states by tag, the CFL number, the error bits and
// the norm history.  The tables that span the active cases are built by
// RebuildGroup.
int UploadCase(ath_ctx *c, CaseRes &cr, bool fp32) {
  using namespace ath;
  const HostCase &hc = *cr.hc;
  c->allocTo = &cr.allocs;
  c->bytesTo = &cr.bytes;
  struct Restore {
    ath_ctx *c;
    ~Restore() { c->allocTo = nullptr; c->bytesTo = nullptr; }
  } restore{c};
  // boundary states by tag
  std::vector<BcStateDev> st;
  for (const auto &s : hc.bcs)
    st.push_back({s.p, s.rho, s.vel[0], s.vel[1], s.vel[2], s.p0, s.t0, s.dir[0], s.dir[1], s.dir[2], s.twall,
                  s.isothermal ? 1 : 0, s.heatFlux ? 1 : 0, s.qwall});
  if (int r = c->Alloc(&cr.dstates, st.size())) return r;
  if (!st.empty())
    HIP_OK(hipMemcpy(cr.dstates, st.data(), st.size() * sizeof(BcStateDev), hipMemcpyHostToDevice));
  if (int r = c->Alloc(&cr.dcfl, 1)) return r;
  if (int r = c->Alloc(&cr.dflag, 1)) return r;
  if (int r = c->Alloc(&cr.dnn, 1)) return r;
  if (int r = c->Alloc(&cr.dcounter, 1)) return r;
  HIP_OK(hipMemset(cr.dflag, 0, sizeof(int)));
  HIP_OK(hipMemset(cr.dnn, 0, sizeof(int)));
  HIP_OK(hipMemset(cr.dcounter, 0, sizeof(int)));
  cr.histCap = 4096;
  if (int r = c->Alloc(&cr.dhist, 8 * (size_t)cr.histCap)) return r;
  cr.cfl = hc.Cfl(0);
  HIP_OK(hipMemcpy(cr.dcfl, &cr.cfl, sizeof(double), hipMemcpyHostToDevice));

  long cellBase = 0;
  cr.blocks.resize(hc.blocks.size());
  for (size_t bq = 0; bq < hc.blocks.size(); ++bq) {
    const HostBlock &h = hc.blocks[bq];
    BlockRes &br = cr.blocks[bq];
    DevBlock &d = br.d;
    d.cfl = ath::Dev(cr.dcfl);
    d.flag = ath::Dev(cr.dflag);
    d.states = ath::Dev(cr.dstates);
    d.ni = h.ni; d.nj = h.nj; d.nk = h.nk; d.g = h.g;
    d.NI = h.NI; d.NJ = h.NJ;
    d.npad = h.NumPadded();
    d.ncell = h.NumCells();
    d.cellBase = cellBase;
    d.blockId = (int)bq;
    cellBase += d.ncell;
    const long np = d.npad, nc = d.ncell;
    // state: SoA padded
    {
      std::vector<double> W(5 * (size_t)np, 0.0);
      for (int k = 0; k < h.nk; ++k)
        for (int j = 0; j < h.nj; ++j)
          for (int i = 0; i < h.ni; ++i) {
            const long p = i + (long)j * h.ni + (long)k * h.ni * h.nj;
            for (int e = 0; e < 5; ++e) W[e * np + h.Cell(i, j, k)] = h.W0[5 * p + e];
          }
      if (int r = c->Alloc(&d.W, 5 * np)) return r;
      HIP_OK(hipMemcpy(d.W, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    {
      std::vector<double> cw(3 * (size_t)np);
      for (int dd = 0; dd < 3; ++dd) std::copy(h.cw[dd].begin(), h.cw[dd].end(), cw.begin() + dd * np);
      double *dcw = nullptr;
      if (int r = c->Alloc(&dcw, 3 * np)) return r;
      HIP_OK(hipMemcpy(dcw, cw.data(), cw.size() * sizeof(double), hipMemcpyHostToDevice));
      d.cw = ath::Dev(dcw);
    }
    for (int dd = 0; dd < 3; ++dd) {
      const int FI = h.ni + (dd == 0), FJ = h.nj + (dd == 1), FK = h.nk + (dd == 2);
      std::vector<double4> fa((size_t)FI * FJ * FK);
      for (int k = 0; k < FK; ++k)
        for (int j = 0; j < FJ; ++j)
          for (int i = 0; i < FI; ++i) {
            const auto &a = h.fa[dd][h.Face(dd, i, j, k)];
            fa[(size_t)i + (size_t)j * FI + (size_t)k * FI * FJ] = make_double4(a[0], a[1], a[2], a[3]);
          }
      double4 *dfa = nullptr;
      if (int r = c->Alloc(&dfa, fa.size())) return r;
      HIP_OK(hipMemcpy(dfa, fa.data(), fa.size() * sizeof(double4), hipMemcpyHostToDevice));
      d.fa[dd] = ath::Dev(dfa);
    }
    {
      std::vector<double> vol((size_t)nc);
      for (int k = 0; k < h.nk; ++k)
        for (int j = 0; j < h.nj; ++j)
          for (int i = 0; i < h.ni; ++i) vol[i + (size_t)j * h.ni + (size_t)k * h.ni * h.nj] = h.vol[h.Cell(i, j, k)];
      double *dv = nullptr;
      if (int r = c->Alloc(&dv, nc)) return r;
      HIP_OK(hipMemcpy(dv, vol.data(), vol.size() * sizeof(double), hipMemcpyHostToDevice));
      d.vol = ath::Dev(dv);
    }
    if (int r = c->Alloc(&d.R, 5 * nc)) return r;
    if (int r = c->Alloc(&d.sr, nc)) return r;
    if (int r = c->Alloc(&d.dt, nc)) return r;
    if (int r = c->Alloc(&d.A, nc)) return r;
    if (int r = c->Alloc(&d.Ainv, nc)) return r;
    if (int r = c->Alloc(&d.consN, 5 * nc)) return r;
    if (int r = c->Alloc(&d.consNm1, 5 * nc)) return r;
    if (int r = c->Alloc(&d.x, 5 * nc)) return r;
    if (int r = c->Alloc(&d.xold, 5 * nc)) return r;
    if (int r = c->Alloc(&d.rhs, 5 * nc)) return r;
    HIP_OK(hipMemset(d.x, 0, 5 * nc * sizeof(double)));
    HIP_OK(hipMemset(d.A, 0, nc * sizeof(double)));
    d.Ab = d.Abi = nullptr;
    d.vgrad = nullptr;
    d.invPark = nullptr;
    d.kc = nullptr;
    for (int dd = 0; dd < 3; ++dd) d.fvis[dd] = nullptr;
    d.tni = (h.ni + 7) / 8;
    d.tnj = (h.nj + 7) / 8;
    d.tnk = (h.nk + 3) / 4;
    if (hc.block) {
      // tile-blocked per-cell blocks: 25 entries of 256 cells per 8x8x4 tile
      const long tb = 25L * 256 * d.tni * d.tnj * d.tnk;
      if (int r = c->Alloc(&d.Ab, 25 * nc)) return r;
      if (int r = c->Alloc(&d.Abi, tb)) return r;
      for (int dd = 0; dd < 3; ++dd) {
        const long nf = (long)(h.ni + (dd == 0)) * (h.nj + (dd == 1)) * (h.nk + (dd == 2));
        if (hc.viscous)
          if (int r = c->Alloc(&d.fvis[dd], 9 * nf)) return r;
      }
      if (hc.viscous) {
        if (int r = c->Alloc(&d.vgrad, 9 * np)) return r;
        HIP_OK(hipMemset(d.vgrad, 0, 9 * np * sizeof(double)));
        if (int r = c->Alloc(&d.kc, np)) return r;
        HIP_OK(hipMemset(d.kc, 0, np * sizeof(double)));
      }
    }
    d.Ainv32 = nullptr;
    d.Abi32 = nullptr;
    d.x32 = nullptr;
    d.xold32 = nullptr;
    d.nbf = 2 * ((long)h.nj * h.nk + (long)h.nk * h.ni + (long)h.ni * h.nj);
    d.bt = nullptr;
    d.bt32 = nullptr;
    bool hasConn = false;
    for (const auto &cn : hc.conns) hasConn = hasConn || cn.block == (int)bq;
    if (hasConn) {
      if (int r = c->Alloc(&d.bt, 5 * d.nbf)) return r;
      HIP_OK(hipMemset(d.bt, 0, 5 * d.nbf * sizeof(double)));
    }
    d.T = d.mu = nullptr;
    d.volP = nullptr;
    d.wallf = nullptr;
    for (int dd = 0; dd < 3; ++dd) {
      d.faP[dd] = nullptr;
      d.c2c[dd] = nullptr;
    }
    if (hc.viscous)
      if (int r = UploadViscous(c, hc, h, br)) return r;
    // hyperplane order
    std::vector<int> cells;
    HyperplaneOrder(h.ni, h.nj, h.nk, &cells, &br.planeOff);
    if (int r = c->Alloc(&br.cells, cells.size())) return r;
    HIP_OK(hipMemcpy(br.cells, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
    // tiled sweep: tiles in tile-hyperplane order, flags, ticket, epoch
    {
      SweepCtl &sc = br.sweep;
      sc.ta = (h.ni + kLTI - 1) / kLTI;
      sc.tb = (h.nj + kLTJ - 1) / kLTJ;
      sc.tc = (h.nk + kLTK - 1) / kLTK;
      sc.ntiles = sc.ta * sc.tb * sc.tc;
      std::vector<int4> tl;
      tl.reserve(sc.ntiles);
      for (int pl = 0; pl < sc.ta + sc.tb + sc.tc - 2; ++pl)
        for (int cc = 0; cc < sc.tc; ++cc)
          for (int bb = 0; bb < sc.tb; ++bb)
            for (int aa = 0; aa < sc.ta; ++aa)
              if (aa + bb + cc == pl) tl.push_back(make_int4(aa, bb, cc, 0));
      int4 *dtl = nullptr;
      if (int r = c->Alloc(&dtl, tl.size())) return r;
      HIP_OK(hipMemcpy(dtl, tl.data(), tl.size() * sizeof(int4), hipMemcpyHostToDevice));
      sc.tiles = dtl;
      if (int r = c->Alloc(&sc.flags, sc.ntiles)) return r;
      if (int r = c->Alloc(&sc.ticket, 2)) return r;
      HIP_OK(hipMemset(sc.flags, 0, sc.ntiles * sizeof(int)));
      HIP_OK(hipMemset(sc.ticket, 0, 2 * sizeof(int)));
      sc.epoch = sc.ticket + 1;
      sc.err = c->dflag;
      hipDeviceProp_t prop;
      HIP_OK(hipGetDeviceProperties(&prop, c->device));
      br.sweepGrid = std::max(1, std::min(sc.ntiles, 2 * prop.multiProcessorCount));
    }
    // reduction partials: the block's 256-cell units
    br.nparts = Grid(nc, kRedThreads);
    // boundary tasks (procBlock::AssignInviscidGhostCells, src/procBlock.cpp:2449-2529)
    long total = 0;
    for (int layer = 1; layer <= h.g; ++layer)
      for (const auto &sf : h.surf) {
        if (sf.type == "interblock" || sf.type == "periodic") continue;  // filled by the halo gather
        const int stp = sf.SurfType();
        const int dir = (stp - 1) / 2;
        const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
        const int r3 = sf.d[2 * dir];
        const int n3 = h.N(dir);
        BcTask t{};
        t.dir = dir;
        t.r1s = sf.d[2 * d1];
        t.n1 = sf.d[2 * d1 + 1] - t.r1s;
        t.r2s = sf.d[2 * d2];
        t.n2 = sf.d[2 * d2 + 1] - t.r2s;
        t.r3 = r3;
        int iC, aC;
        if (stp % 2 == 0) {
          t.gC = r3 + layer - 1;
          iC = std::max(r3 - layer, 0);
          aC = r3 - 1;
        } else {
          t.gC = r3 - layer;
          iC = std::min(r3 + layer - 1, n3 - 1);
          aC = r3;
        }
        std::string ty = sf.type == "viscousWall" ? "slipWall" : sf.type;
        t.bctype = ty == "slipWall" ? BC_SLIPWALL
                   : ty == "characteristic" ? BC_CHARACTERISTIC
                   : ty == "supersonicInflow" ? BC_SUPERSONIC_INFLOW
                   : ty == "supersonicOutflow" ? BC_SUPERSONIC_OUTFLOW
                   : ty == "stagnationInlet" ? BC_STAGNATION_INLET
                                             : BC_PRESSURE_OUTLET;
        t.srcC = t.bctype == BC_SLIPWALL ? iC : aC;
        t.stype = stp;
        t.layer = layer;
        t.state = -1;
        for (size_t q = 0; q < hc.bcs.size(); ++q)
          if (hc.bcs[q].tag == sf.d[6]) { t.state = (int)q; break; }
        t.base = total;
        total += (long)t.n1 * t.n2;
        if (t.n1 > 0 && t.n2 > 0) br.tasks.push_back(t);
      }
    br.bcTotal = total;
    if (!br.tasks.empty()) {
      if (int r = c->Alloc(&br.dtasks, br.tasks.size())) return r;
      HIP_OK(hipMemcpy(br.dtasks, br.tasks.data(), br.tasks.size() * sizeof(BcTask), hipMemcpyHostToDevice));
    }
  }
  CaseConnections(hc, cr);
  if (fp32)
    for (auto &br : cr.blocks)
      if (int r = AllocFp32(c, br.d)) return r;
  return ATH_OK;
}

// The tables of the active cases, rebuilt whenever that set or a block's
// buffers change: the concatenated block list and its device copy, the
// connection gather lists, the unit grids of the all-block launches and the
// reduction partials (each case's 256-cell units are one contiguous range),
// the ghost-cell task tables, the pencil and block sweep queues and the
// per-case norm and CFL tables.  Captured graphs are dropped.
int RebuildGroup(ath_ctx *c) {
  using namespace ath;
  c->DropGraph();
  for (void *p : c->groupAllocs) hipFree(p);
  c->groupAllocs.clear();
  c->allocTo = &c->groupAllocs;
  struct Restore {
    ath_ctx *c;
    ~Restore() { c->allocTo = nullptr; }
  } restore{c};
  c->blk.clear();
  std::vector<int> first;  // first block of each active case
  for (CaseRes *cr : c->active) {
    first.push_back((int)c->blk.size());
    for (const auto &br : cr->blocks) c->blk.push_back(br);
  }
  const int nb = (int)c->blk.size();
  c->dblocks = nullptr;
  c->dmaps = nullptr;
  c->dbt = nullptr;
  c->dhalo = nullptr;
  c->dconn = nullptr;
  c->dpart = nullptr;
  c->dnorm = nullptr;
  c->dcflc = nullptr;
  c->nhalo = c->nconn = 0;
  c->nparts = 0;
  c->pen = PencilGroup{};
  c->msweep = MultiSweepCtl{};
  c->rsweep = MultiSweepCtl{};
  for (int q = 0; q < kNumMaps; ++q) c->mapTotal[q] = 0;
  for (int q = 0; q < kNumTaskMaps; ++q) c->taskTotal[q] = 0;
  if (nb == 0) return ATH_OK;
  // the device block table
  {
    std::vector<DevBlock> v;
    for (const auto &br : c->blk) v.push_back(br.d);
    if (int r = c->Alloc(&c->dblocks, v.size())) return r;
    HIP_OK(hipMemcpy(c->dblocks, v.data(), v.size() * sizeof(DevBlock), hipMemcpyHostToDevice));
  }
  // connections
  {
    std::vector<HaloTask> halo;
    std::vector<ConnTask> conn;
    for (size_t q = 0; q < c->active.size(); ++q) {
      const int f = first[q];
      for (HaloTask t : c->active[q]->halo) {
        t.dst += f;
        t.src += f;
        halo.push_back(t);
      }
      for (ConnTask t : c->active[q]->conn) {
        t.blk += f;
        t.src += f;
        conn.push_back(t);
      }
    }
    c->nhalo = (int)halo.size();
    c->nconn = (int)conn.size();
    if (!halo.empty()) {
      if (int r = c->Alloc(&c->dhalo, halo.size())) return r;
      HIP_OK(hipMemcpy(c->dhalo, halo.data(), halo.size() * sizeof(HaloTask), hipMemcpyHostToDevice));
    }
    if (!conn.empty()) {
      if (int r = c->Alloc(&c->dconn, conn.size())) return r;
      HIP_OK(hipMemcpy(c->dconn, conn.data(), conn.size() * sizeof(ConnTask), hipMemcpyHostToDevice));
    }
  }
  // unit grids of the all-block launches
  std::vector<int4> v((size_t)kNumMaps * nb);
  for (int q = 0; q < kNumMaps; ++q) {
    int f = 0;
    for (int b = 0; b < nb; ++b) {
      const auto &d = c->blk[b].d;
      int4 u;
      if (q == kMapTile)
        u = make_int4((d.ni + ath::kTI - 1) / ath::kTI, (d.nj + ath::kTJ - 1) / ath::kTJ,
                      (d.nk + ath::kTK - 1) / ath::kTK, f);
      else if (q == kMapCell128) u = make_int4(Grid(d.ncell, 128), 1, 1, f);
      else if (q == kMapCell256) u = make_int4(Grid(d.ncell, ath::kRedThreads), 1, 1, f);
      else u = make_int4(Grid(d.npad, 256), 1, 1, f);
      v[(size_t)q * nb + b] = u;
      f += u.x * u.y * u.z;
    }
    c->mapTotal[q] = f;
  }
  if (int r = c->Alloc(&c->dmaps, v.size())) return r;
  HIP_OK(hipMemcpy(c->dmaps, v.data(), v.size() * sizeof(int4), hipMemcpyHostToDevice));
  // the reductions' partial slots are the 256-cell units; each case's are
  // contiguous
  for (int b = 0; b < nb; ++b) c->blk[b].partBase = v[(size_t)kMapCell256 * nb + b].w;
  c->nparts = c->mapTotal[kMapCell256];
  if (int r = c->Alloc(&c->dpart, c->nparts)) return r;
  {
    std::vector<NormCase> nc;
    for (size_t q = 0; q < c->active.size(); ++q) {
      CaseRes &cr = *c->active[q];
      cr.partFirst = c->blk[first[q]].partBase;
      cr.nparts = 0;
      for (int b = first[q]; b < first[q] + (int)cr.blocks.size(); ++b) cr.nparts += c->blk[b].nparts;
      nc.push_back({cr.partFirst, cr.nparts, cr.dhist, cr.dcounter});
    }
    if (int r = c->Alloc(&c->dnorm, nc.size())) return r;
    HIP_OK(hipMemcpy(c->dnorm, nc.data(), nc.size() * sizeof(NormCase), hipMemcpyHostToDevice));
    std::vector<CflCase> cc;
    for (CaseRes *cr : c->active)
      cc.push_back({cr->hc->cflStart, cr->hc->cflStep, cr->hc->cflMax, cr->dcfl, cr->dnn});
    if (int r = c->Alloc(&c->dcflc, cc.size())) return r;
    HIP_OK(hipMemcpy(c->dcflc, cc.data(), cc.size() * sizeof(CflCase), hipMemcpyHostToDevice));
  }
  // ghost-cell task tables
  std::vector<ath::BlockTasks> bt((size_t)kNumTaskMaps * nb);
  for (int q = 0; q < kNumTaskMaps; ++q) {
    long f = 0;
    for (int b = 0; b < nb; ++b) {
      const auto &br = c->blk[b];
      ath::BlockTasks e{};
      e.first = f;
      if (q == kTaskBc) { e.tasks = br.dtasks; e.count = br.bcTotal; e.ntasks = (int)br.tasks.size(); }
      else if (q == kTaskEdgeInv) { e.tasks = br.edgeInv; e.count = br.edgeLines; e.per = br.edgePerLine; }
      else if (q == kTaskEdgeVisc) { e.tasks = br.edgeVisc; e.count = br.edgeLines; e.per = br.edgePerLine; }
      else { e.tasks = br.wall; e.count = br.nwall; }
      if (!e.tasks) e.count = 0;
      bt[(size_t)q * nb + b] = e;
      f += e.count;
    }
    c->taskTotal[q] = f;
  }
  if (int r = c->Alloc(&c->dbt, bt.size())) return r;
  HIP_OK(hipMemcpy(c->dbt, bt.data(), bt.size() * sizeof(ath::BlockTasks), hipMemcpyHostToDevice));
  // pencil sweep queue: every block's kPW x kPW columns by column hyperplane
  // a + b, then block, then b, a
  {
    std::vector<int4> ent;
    std::vector<int> base(nb);
    int total = 0, maxPlane = 0;
    for (int b = 0; b < nb; ++b) {
      const auto &d = c->blk[b].d;
      const int pa = (d.ni + kPW - 1) / kPW, pb = (d.nj + kPW - 1) / kPW;
      base[b] = total;
      total += pa * pb;
      maxPlane = std::max(maxPlane, pa + pb - 1);
    }
    for (int pl = 0; pl < maxPlane; ++pl)
      for (int b = 0; b < nb; ++b) {
        const auto &d = c->blk[b].d;
        const int pa = (d.ni + kPW - 1) / kPW, pb = (d.nj + kPW - 1) / kPW;
        for (int bb = 0; bb < pb; ++bb)
          for (int aa = 0; aa < pa; ++aa)
            if (aa + bb == pl) ent.push_back(make_int4(aa, bb, b, base[b]));
      }
    int4 *de = nullptr;
    if (int r = c->Alloc(&de, ent.size())) return r;
    HIP_OK(hipMemcpy(de, ent.data(), ent.size() * sizeof(int4), hipMemcpyHostToDevice));
    PencilGroup &g = c->pen;
    g.entries = de;
    g.nent = (int)ent.size();
    if (int r = c->Alloc(&g.prog, total)) return r;
    if (int r = c->Alloc(&g.ticket, 1)) return r;
    double *sink = nullptr;
    if (int r = c->Alloc(&sink, (size_t)total * 5 * kPenLanes)) return r;
    g.sinkv = sink;
    HIP_OK(hipMemset(g.prog, 0, total * sizeof(int)));
    HIP_OK(hipMemset(g.ticket, 0, sizeof(int)));
    g.err = c->dflag;
    g.blocks = c->dblocks;
  }
  // block solvers: one tiled sweep over all blocks, tiles by tile hyperplane,
  // then block, then k, j, i; per block its tile grid and first flag
  if (c->hc.block) {
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, c->device));
    // the tiles of every block by tile hyperplane, then block, then k, j, i;
    // per block its tile grid and first flag
    auto queue = [&](int TI, int TJ, int TK, MultiSweepCtl &m) -> int {
      std::vector<int4> grid, tiles;
      int base = 0, maxPlane = 0;
      for (auto &br : c->blk) {
        const int ta = (br.d.ni + TI - 1) / TI, tb = (br.d.nj + TJ - 1) / TJ, tc = (br.d.nk + TK - 1) / TK;
        grid.push_back(make_int4(ta, tb, tc, base));
        base += ta * tb * tc;
        maxPlane = std::max(maxPlane, ta + tb + tc - 2);
      }
      for (int pl = 0; pl < maxPlane; ++pl)
        for (size_t q = 0; q < grid.size(); ++q)
          for (int cc = 0; cc < grid[q].z; ++cc)
            for (int bb = 0; bb < grid[q].y; ++bb)
              for (int aa = 0; aa < grid[q].x; ++aa)
                if (aa + bb + cc == pl) tiles.push_back(make_int4(aa, bb, cc, (int)q));
      m.ntiles = (int)tiles.size();
      int4 *dt = nullptr, *dg = nullptr;
      if (int r = c->Alloc(&dt, tiles.size())) return r;
      if (int r = c->Alloc(&dg, grid.size())) return r;
      HIP_OK(hipMemcpy(dt, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(dg, grid.data(), grid.size() * sizeof(int4), hipMemcpyHostToDevice));
      m.tiles = dt;
      m.grid = dg;
      if (int r = c->Alloc(&m.flags, base)) return r;
      if (int r = c->Alloc(&m.ticket, 2)) return r;
      HIP_OK(hipMemset(m.flags, 0, base * sizeof(int)));
      HIP_OK(hipMemset(m.ticket, 0, 2 * sizeof(int)));
      m.epoch = m.ticket + 1;
      m.err = c->dflag;
      return ATH_OK;
    };
    if (int r = queue(kBTI, kBTJ, kBTK, c->msweep)) return r;
    if (int r = queue(kBTI, kBTJ, kBTK, c->rsweep)) return r;
    c->msweepGrid = std::max(1, std::min(c->msweep.ntiles, 2 * prop.multiProcessorCount));
    // one 512-thread workgroup fits a CU (registers)
    c->rsweepGrid = std::max(1, std::min(c->rsweep.ntiles, prop.multiProcessorCount));
  }
  return ATH_OK;
}

// kernel parameters of a case (every case of a server must produce the same)
void FillParams(const ath::HostCase &hc, ath::Params *out) {
  ath::Params P;
  std::memset(&P, 0, sizeof P);
  P.gas.R = hc.gasR;
  P.gas.n = hc.gasN;
  P.gas.gamma = hc.gamma;
  P.gas.cps = hc.gasR * (hc.gasN + 1.0);
  P.gas.cvs = hc.gasR * hc.gasN;
  P.kappa = hc.kappa;
  P.theta = hc.theta;
  P.zeta = hc.zeta;
  P.relax = hc.relax;
  P.dualCFL = hc.dualCFL;
  P.dtFixed = hc.dtIn > 0.0 ? hc.dtIn * hc.aRef / hc.lRef : -1.0;
  P.implicit = hc.Implicit() ? 1 : 0;
  P.multilevel = hc.MultiLevel() ? 1 : 0;
  P.initMatrix = hc.InitMatrix() ? 1 : 0;
  P.explicitEuler = hc.timeint == ath::TimeInt::ExplicitEuler ? 1 : 0;
  P.roeJac = hc.roeJac ? 1 : 0;
  P.block = hc.block ? 1 : 0;
  P.visc.on = hc.viscous ? 1 : 0;
  P.visc.recon4 = hc.viscRecon4 ? 1 : 0;
  P.visc.vscale = hc.vscale;
  P.visc.prandtl = hc.prandtl;
  P.visc.coeff = hc.viscCoeff;
  P.visc.tRef = hc.tRef;
  P.visc.muRef = hc.muRef;
  P.visc.kNonDim = hc.kNonDim;
  P.visc.c1 = hc.suth[0];
  P.visc.s1 = hc.suth[1];
  P.visc.c2 = hc.suth[2];
  P.visc.s2 = hc.suth[3];
  *out = P;
}

// a single-case context: the case is the context's own
int Upload(ath_ctx *c) {
  FillParams(c->hc, &c->P);
  if (int r = c->Alloc(&c->dflag, 1)) return r;
  HIP_OK(hipMemset(c->dflag, 0, sizeof(int)));
  auto cr = std::make_unique<CaseRes>();
  cr->hc = &c->hc;
  c->cases.push_back(std::move(cr));
  CaseRes &k = *c->cases.back();
  if (int r = UploadCase(c, k, false)) return r;
  c->active.assign(1, &k);
  return RebuildGroup(c);
}

// --- launch helpers (all on c->stream) -------------------------------------

// gridLevel::GetBoundaryConditions (src/gridLevel.cpp:287-319): ghost
// states of the boundary patches, then the connections' ghost cells (edge
// ghost cells, which only viscous terms read, are not formed)
void LaunchBoundary(ath_ctx *c) {
  const int nb = (int)c->blk.size();
  if (c->taskTotal[kTaskBc] > 0)
    hipLaunchKernelGGL(ath::KBoundary, dim3(Grid(c->taskTotal[kTaskBc])), dim3(kBlk), 0, c->stream, c->dblocks,
                       c->Tasks(kTaskBc), nb, c->taskTotal[kTaskBc], c->P.gas);
  if (c->nhalo > 0)
    hipLaunchKernelGGL(ath::KHaloState, dim3(Grid(c->nhalo)), dim3(kBlk), 0, c->stream, c->dblocks, c->dhalo,
                       c->nhalo);
  // edge ghost cells: read only by the viscous gradients
  if (c->P.visc.on && c->taskTotal[kTaskEdgeInv] > 0)
    hipLaunchKernelGGL(ath::KEdgeGhosts, dim3(Grid(c->taskTotal[kTaskEdgeInv], 64)), dim3(64), 0, c->stream,
                       c->dblocks, c->Tasks(kTaskEdgeInv), nb, c->taskTotal[kTaskEdgeInv], c->P.gas);
}

// connection terms from the partners' current du, before each pass
template <typename T>
void LaunchConnTerms(ath_ctx *c) {
  if (c->nconn > 0)
    hipLaunchKernelGGL(ath::KConnTerms<T>, dim3(Grid(c->nconn)), dim3(kBlk), 0, c->stream, c->dblocks, c->dconn,
                       c->nconn, c->P);
}

int EffectiveOrder(const ath_ctx *c);

// the relax's initialisation (right-hand side once per relax, du = 0) is
// formed by the viscous block diagonal: one-sweep block LU-SGS with the FP64
// operator in the row-parallel order, viscous terms on.  A one-sweep relax
// never reads its initial du, so a relax repeated without a new residual
// gives the same result.
bool RhsInDiag(const ath_ctx *c) {
  return c->P.block && c->P.visc.on && !c->fp32 && c->hc.solver == ath::Solver::LUSGS &&
         !c->hc.InitMatrix() && EffectiveOrder(c) == ATH_SWEEP_ROWS;
}

template <int RECON, int LIM, int FLUX>
void LaunchResidualT(ath_ctx *c) {
  const ath::Params &P = c->P;
  const ath::BlockMap mt = c->Map(kMapTile), mc = c->Map(kMapCell128), mp = c->Map(kMapPad256);
  const dim3 tiles(c->mapTotal[kMapTile]), cells(c->mapTotal[kMapCell128]);
  hipLaunchKernelGGL((ath::KResidual<RECON, LIM, FLUX>), tiles, dim3(256), 0, c->stream, mt, P);
  if (P.block) {
    // inviscid block sum from the inviscid ghost states: finished here, or,
    // for the cells next to a viscous wall, parked in Ab until the viscous
    // pass has formed dt (the other cells form it then)
    if (!P.visc.on)
      hipLaunchKernelGGL((ath::KBlockDiag<RECON, LIM, 0>), cells, dim3(128), 0, c->stream, mc, P);
    else if (c->anyPark)
      hipLaunchKernelGGL((ath::KBlockDiag<RECON, LIM, 1>), cells, dim3(128), 0, c->stream, mc, P);
  }
  if (!P.visc.on) return;
  // viscous ghost cells, temperature and viscosity, viscous fluxes
  // (src/procBlock.cpp:6126-6138); every block's ghost cells are its own, so
  // the blocks' passes need no order among them
  const int nb = (int)c->blk.size();
  if (c->taskTotal[kTaskWall] > 0)
    hipLaunchKernelGGL(ath::KViscWall, dim3(Grid(c->taskTotal[kTaskWall])), dim3(kBlk), 0, c->stream, c->dblocks,
                       c->Tasks(kTaskWall), nb, c->taskTotal[kTaskWall], P.gas, P.visc);
  if (c->taskTotal[kTaskEdgeVisc] > 0)
    hipLaunchKernelGGL(ath::KEdgeGhosts, dim3(Grid(c->taskTotal[kTaskEdgeVisc], 64)), dim3(64), 0, c->stream,
                       c->dblocks, c->Tasks(kTaskEdgeVisc), nb, c->taskTotal[kTaskEdgeVisc], P.gas);
  hipLaunchKernelGGL(ath::KAux, dim3(c->mapTotal[kMapPad256]), dim3(kBlk), 0, c->stream, mp, P);
  if (P.visc.recon4)
    hipLaunchKernelGGL((ath::KViscFlux<true>), tiles, dim3(256), 0, c->stream, mt, P);
  else
    hipLaunchKernelGGL((ath::KViscFlux<false>), tiles, dim3(256), 0, c->stream, mt, P);
  // the block diagonal needs dt, which the viscous pass completes
  if (P.block) {
    ath::Params PD = P;
    PD.rhsInDiag = RhsInDiag(c) ? 1 : 0;
    c->rhsFormed = PD.rhsInDiag;
    if (P.visc.recon4)
      hipLaunchKernelGGL((ath::KBlockDiag<RECON, LIM, 3>), cells, dim3(128), 0, c->stream, mc, PD);
    else
      hipLaunchKernelGGL((ath::KBlockDiag<RECON, LIM, 2>), cells, dim3(128), 0, c->stream, mc, PD);
  }
}

// after every block's residual: the connections' ghost velocity gradients,
// which the viscous block off-diagonals read
void LaunchAfterResiduals(ath_ctx *c) {
  if (c->P.block && c->P.visc.on && c->nhalo > 0)
    hipLaunchKernelGGL(ath::KHaloGrad, dim3(Grid(c->nhalo)), dim3(kBlk), 0, c->stream, c->dblocks, c->dhalo,
                       c->nhalo);
}

// the residual stage of every block: one launch per kernel over all blocks
void LaunchResidual(ath_ctx *c) {
  c->rhsFormed = 0;
  const auto &hc = c->hc;
  const int recon = (int)hc.recon;
  const int lim = (int)hc.limiter;
  const int flux = (int)hc.flux;
#define ATH_RES(R, L, F) if (recon == R && lim == L && flux == F) return LaunchResidualT<R, L, F>(c)
  ATH_RES(0, 0, 0); ATH_RES(0, 0, 1);
  ATH_RES(1, 0, 0); ATH_RES(1, 1, 0); ATH_RES(1, 2, 0);
  ATH_RES(1, 0, 1); ATH_RES(1, 1, 1); ATH_RES(1, 2, 1);
  // WENO ignores the limiter
  if (recon == 2) return flux == 0 ? LaunchResidualT<2, 0, 0>(c) : LaunchResidualT<2, 0, 1>(c);
  if (recon == 3) return flux == 0 ? LaunchResidualT<3, 0, 0>(c) : LaunchResidualT<3, 0, 1>(c);
  // constant reconstruction ignores the limiter
  if (recon == 0) return flux == 0 ? LaunchResidualT<0, 0, 0>(c) : LaunchResidualT<0, 0, 1>(c);
#undef ATH_RES
}

// the order the sweeps run in: the block solvers have the row-parallel,
// tiled and hyperplane orders (the pencil order maps to the row-parallel
// one with the FP64 operator and to the tiled one with the FP32 operator,
// the faster of the two for each on the 1 M-cell channel), scalar LU-SGS
// the pencil, tiled and hyperplane orders (the row-parallel order maps to
// the tiled one); the Roe flux-change off-diagonals (approximateRoe) run in
// the hyperplane order only
int EffectiveOrder(const ath_ctx *c) {
  if (c->P.roeJac) return ATH_SWEEP_HYPERPLANE;
  if (c->P.block && c->sweepOrder == ATH_SWEEP_PENCIL) return c->fp32 ? ATH_SWEEP_TILED : ATH_SWEEP_ROWS;
  if (!c->P.block && c->sweepOrder == ATH_SWEEP_ROWS) return ATH_SWEEP_TILED;
  return c->sweepOrder;
}

// One relax.  When `timed`, an event pair brackets every sweep pass (a
// forward or backward LU-SGS pass, or one DP-LUR sweep, over all blocks)
// from c->evs[evBase]; returns the number of sweep kernel launches.
template <typename T>
int LaunchRelaxT(ath_ctx *c, bool timed, size_t evBase) {
  const auto &hc = c->hc;
  int launches = 0, pass = 0;
  auto mark = [&](bool begin) {
    if (timed) hipEventRecord(c->evs[evBase + 2 * pass + (begin ? 0 : 1)], c->stream);
    if (!begin) ++pass;
  };
  // the row-parallel block sweep reads the right-hand side that the
  // initialisation forms once per relax
  ath::Params P0 = c->P;
  P0.rhsOnce = (hc.solver == ath::Solver::LUSGS && c->P.block && EffectiveOrder(c) == ATH_SWEEP_ROWS) ? 1 : 0;
  // (formed by the residual stage unless the sweep order changed since)
  if (!(RhsInDiag(c) && c->rhsFormed))
    hipLaunchKernelGGL(ath::KInitUpdate<T>, dim3(c->mapTotal[kMapCell256]), dim3(kBlk), 0, c->stream,
                       c->Map(kMapCell256), P0);
  for (int sw = 0; sw < hc.sweeps; ++sw) {
    const int full = (sw > 0 || hc.InitMatrix()) ? 1 : 0;
    const int order = EffectiveOrder(c);
    if (hc.solver == ath::Solver::LUSGS && order == ATH_SWEEP_PENCIL) {
      if (sw == 0)
        hipLaunchKernelGGL(ath::KLusgsRhs, dim3(c->mapTotal[kMapCell256]), dim3(kBlk), 0, c->stream,
                           c->Map(kMapCell256), c->P);
      for (int dir = 0; dir < 2; ++dir) {
        LaunchConnTerms<T>(c);
        mark(true);
        // every block's pencils in one launch
        hipLaunchKernelGGL(ath::KPencilPrologue, dim3(1), dim3(256), 0, c->stream, c->pen);
        const dim3 grid(c->pen.nent), blk(ath::kPenWaves * ath::kPenLanes);
        if (dir == 0 && full)
          hipLaunchKernelGGL((ath::KLusgsPencil<T, true, true>), grid, blk, 0, c->stream, c->pen, c->P);
        else if (dir == 0)
          hipLaunchKernelGGL((ath::KLusgsPencil<T, true, false>), grid, blk, 0, c->stream, c->pen, c->P);
        else if (full)
          hipLaunchKernelGGL((ath::KLusgsPencil<T, false, true>), grid, blk, 0, c->stream, c->pen, c->P);
        else
          hipLaunchKernelGGL((ath::KLusgsPencil<T, false, false>), grid, blk, 0, c->stream, c->pen, c->P);
        ++launches;
        mark(false);
      }
    } else if (hc.solver == ath::Solver::LUSGS && c->P.block &&
               (order == ATH_SWEEP_TILED || order == ATH_SWEEP_ROWS)) {
      for (int dir = 0; dir < 2; ++dir) {
        // the first forward pass starts from du = 0: its connection terms
        // are zero and the sweep does not read them
        if (dir > 0 || full) LaunchConnTerms<T>(c);
        mark(true);
        if (order == ATH_SWEEP_ROWS) {
          if (full) {
            if (dir == 0)
              hipLaunchKernelGGL((ath::KOtherSide<T, true>), dim3(c->mapTotal[kMapCell256]), dim3(kBlk), 0,
                                 c->stream, c->Map(kMapCell256), c->P);
            else
              hipLaunchKernelGGL((ath::KOtherSide<T, false>), dim3(c->mapTotal[kMapCell256]), dim3(kBlk), 0,
                                 c->stream, c->Map(kMapCell256), c->P);
          }
          hipLaunchKernelGGL(ath::KMultiSweepPrologue, dim3(1), dim3(1), 0, c->stream, c->rsweep);
          if (dir == 0)
            hipLaunchKernelGGL((ath::KBlockSweepRows<T, true>), dim3(c->rsweepGrid), dim3(ath::kPThreads), 0,
                               c->stream, c->dblocks, c->P, c->rsweep, full);
          else
            hipLaunchKernelGGL((ath::KBlockSweepRows<T, false>), dim3(c->rsweepGrid), dim3(ath::kPThreads), 0,
                               c->stream, c->dblocks, c->P, c->rsweep, full);
        } else {
          hipLaunchKernelGGL(ath::KMultiSweepPrologue, dim3(1), dim3(1), 0, c->stream, c->msweep);
          if (dir == 0)
            hipLaunchKernelGGL((ath::KBlockSweepTiles<T, true>), dim3(c->msweepGrid), dim3(ath::kBT), 0,
                               c->stream, c->dblocks, c->P, c->msweep, full);
          else
            hipLaunchKernelGGL((ath::KBlockSweepTiles<T, false>), dim3(c->msweepGrid), dim3(ath::kBT), 0,
                               c->stream, c->dblocks, c->P, c->msweep, full);
        }
        ++launches;
        mark(false);
      }
    } else if (hc.solver == ath::Solver::LUSGS && order == ATH_SWEEP_TILED) {
      for (int dir = 0; dir < 2; ++dir) {
        LaunchConnTerms<T>(c);
        mark(true);
        for (auto &br : c->blk) {
          hipLaunchKernelGGL(ath::KSweepPrologue, dim3(1), dim3(1), 0, c->stream, br.sweep);
          if (dir == 0)
            hipLaunchKernelGGL((ath::KLusgsTile<T, true>), dim3(br.sweepGrid), dim3(ath::kLT), 0, c->stream, br.d,
                               c->P, br.sweep, full);
          else
            hipLaunchKernelGGL((ath::KLusgsTile<T, false>), dim3(br.sweepGrid), dim3(ath::kLT), 0, c->stream, br.d,
                               c->P, br.sweep, full);
          ++launches;
        }
        mark(false);
      }
    } else if (hc.solver == ath::Solver::LUSGS) {
      LaunchConnTerms<T>(c);
      mark(true);
      for (auto &br : c->blk) {
        const int np = (int)br.planeOff.size() - 1;
        for (int p = 0; p < np; ++p, ++launches) {
          const int n = br.planeOff[p + 1] - br.planeOff[p];
          hipLaunchKernelGGL(ath::KLusgsForward<T>, dim3(Grid(n, 64)), dim3(64), 0, c->stream, br.d, c->P,
                             br.cells + br.planeOff[p], n, full);
        }
      }
      mark(false);
      LaunchConnTerms<T>(c);
      mark(true);
      for (auto &br : c->blk) {
        const int np = (int)br.planeOff.size() - 1;
        for (int p = np - 1; p >= 0; --p, ++launches) {
          const int n = br.planeOff[p + 1] - br.planeOff[p];
          hipLaunchKernelGGL(ath::KLusgsBackward<T>, dim3(Grid(n, 64)), dim3(64), 0, c->stream, br.d, c->P,
                             br.cells + br.planeOff[p], n, full);
        }
      }
      mark(false);
    } else {
      LaunchConnTerms<T>(c);
      mark(true);
      for (auto &br : c->blk) {
        hipLaunchKernelGGL(ath::KCopyUpdate<T>, dim3(Grid(5 * br.d.ncell)), dim3(kBlk), 0, c->stream, br.d);
        hipLaunchKernelGGL(ath::KDplur<T>, dim3(Grid(br.d.ncell)), dim3(kBlk), 0, c->stream, br.d, c->P);
        ++launches;
      }
      mark(false);
    }
  }
  if (c->matres) {
    LaunchConnTerms<T>(c);
    hipLaunchKernelGGL(ath::KMatrixResidual<T>, dim3(c->mapTotal[kMapCell256]), dim3(ath::kRedThreads), 0,
                       c->stream, c->Map(kMapCell256), c->P, c->dpart);
  }
  return launches;
}

template <typename T>
void LaunchUpdateT(ath_ctx *c) {
  hipLaunchKernelGGL(ath::KUpdate<T>, dim3(c->mapTotal[kMapCell256]), dim3(ath::kRedThreads), 0, c->stream,
                     c->Map(kMapCell256), c->P, c->dpart);
  hipLaunchKernelGGL(ath::KReduceNorms, dim3(c->active.size()), dim3(ath::kRedThreads), 0, c->stream, c->dpart,
                     c->dnorm, (c->P.implicit && c->matres) ? 1 : 0);
}

// everything after the residual: relax (implicit) + update + reduction
// returns the number of sweep kernel launches
int LaunchTail(ath_ctx *c, bool timed, size_t evBase) {
  int launches = 0;
  if (c->P.implicit) {
    if (c->fp32) launches = LaunchRelaxT<float>(c, timed, evBase);
    else launches = LaunchRelaxT<double>(c, timed, evBase);
  }
  if (c->fp32) LaunchUpdateT<float>(c);
  else LaunchUpdateT<double>(c);
  return launches;
}

// sweep passes of one relax (event pairs needed to time it)
int SweepPasses(const ath_ctx *c) {
  if (!c->P.implicit) return 0;
  return c->hc.solver == ath::Solver::LUSGS ? 2 * c->hc.sweeps : c->hc.sweeps;
}

void LaunchCopyNm1(ath_ctx *c) {
  for (auto &br : c->blk)
    hipLaunchKernelGGL(ath::KCopyNm1, dim3(Grid(5 * br.d.ncell)), dim3(kBlk), 0, c->stream, br.d);
}

int CheckLaunch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return Fail(ATH_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return ATH_OK;
}

// the norms of one iteration of case `cr` from its history slot values h
// (sums of R^2, L-inf value and index, matrix-residual sum), with the
// PrintResiduals normalisation (src/output.cpp:1028-1063) advanced for outer
// step nn, nonlinear iteration mm
void NormsFrom(CaseRes &cr, const ath::Params &P, const double h[8], int nn, int mm, double cfl,
               ath_norms *out) {
  ath_norms n{};
  n.num_equations = 5;
  for (int e = 0; e < 5; ++e) {
    n.l2_sq[e] = h[e];
    n.l2[e] = std::sqrt(h[e]);
  }
  n.linf = h[5];
  long long idx;
  std::memcpy(&idx, &h[6], sizeof idx);
  if (idx >= 0) {
    const long long cell = idx / 5;
    n.linf_eqn = (int)(idx % 5) + 1;
    long long base = 0;
    for (const auto &br : cr.blocks) {
      if (cell < base + br.d.ncell) {
        const long long p = cell - base;
        n.linf_block = br.d.blockId;
        n.linf_i = (int)(p % br.d.ni);
        n.linf_j = (int)((p / br.d.ni) % br.d.nj);
        n.linf_k = (int)(p / ((long long)br.d.ni * br.d.nj));
        break;
      }
      base += br.d.ncell;
    }
  }
  // mgSolution::CycleAtLevel divides by the padded size of x
  // (src/mgSolution.cpp:197-206); main divides by cells * neq
  // (src/main.cpp:269-270)
  double padded = 0;
  for (const auto &br : cr.blocks) padded += (double)br.d.npad * 5;
  const double cyc = P.implicit ? h[7] / padded : 0.0;
  n.matrix_resid = std::sqrt(cyc / (cr.hc->totalCells * 5));
  n.cfl = cfl;
  // PrintResiduals (src/output.cpp:1028-1063)
  constexpr double eps = 1.0e-30;
  if (cr.restart) {
    // a restart keeps the normalisation read from the file
  } else if (nn == 0 && mm == 0) {
    for (int e = 0; e < 5; ++e) cr.l2first[e] = n.l2[e];
  } else if (nn < 5 && mm == 0) {
    if (n.l2[0] > cr.l2first[0]) cr.l2first[0] = n.l2[0];
    for (int e = 1; e < 5; ++e) cr.l2first[e] = std::max(cr.l2first[e], n.l2[e]);
  }
  for (int e = 0; e < 5; ++e) n.normalized[e] = (n.l2[e] + eps) / (cr.l2first[e] + eps);
  if (out) *out = n;
}

// the device error bits of case `cr` and of the context (spin limits),
// read and cleared: a device error is reported once, so that the context
// can go on (for example after ath_set_state restores a state)
int TakeFlags(ath_ctx *c, CaseRes &cr, int *flag) {
  int fc = 0, fg = 0;
  HIP_OK(hipMemcpy(&fc, cr.dflag, sizeof fc, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(&fg, c->dflag, sizeof fg, hipMemcpyDeviceToHost));
  const int zero = 0;
  if (fc) HIP_OK(hipMemcpy(cr.dflag, &zero, sizeof zero, hipMemcpyHostToDevice));
  if (fg) HIP_OK(hipMemcpy(c->dflag, &zero, sizeof zero, hipMemcpyHostToDevice));
  *flag = fc | fg;
  return ATH_OK;
}

int FlagStatus(int flag) {
  if (flag & 2) return Fail(ATH_ERR_DEVICE, "a sweep dependency wait exceeded its spin limit");
  if (flag & 4) return Fail(ATH_ERR_DEVICE, "singular block main diagonal in Gauss-Jordan elimination");
  if (flag & 1) return Fail(ATH_ERR_NONPHYSICAL, "nonphysical state (rho <= 0, p <= 0 or NaN) after update");
  return ATH_OK;
}

// single-case context: read norms slot `slot` and fill `out`
int ReadNorms(ath_ctx *c, int slot, int nn, int mm, ath_norms *out) {
  CaseRes &cr = *c->cases[0];
  double h[8];
  HIP_OK(hipMemcpy(h, cr.dhist + 8 * (size_t)slot, sizeof h, hipMemcpyDeviceToHost));
  NormsFrom(cr, c->P, h, nn, mm, cr.cfl, out);
  int flag = 0;
  if (int r = TakeFlags(c, cr, &flag)) return r;
  return FlagStatus(flag);
}

// one line of the residual log for outer step nn, iteration mm
void LogLine(ath_ctx *c, const ath_norms &n, int nn, int mm, double iterSeconds) {
  if (!c->logResid) return;
  const double v = c->hc.dtIn > 0.0 ? c->hc.dtIn : n.cfl;
  ath::PrintResidualLine(c->hc, nn, mm, v, n.normalized, n.linf_eqn, n.linf_block, n.linf_i, n.linf_j, n.linf_k,
                         n.linf, n.matrix_resid, *c->logResid);
  if (mm == c->hc.nonlinear - 1 && c->logTime) {
    const double sim = std::chrono::duration<double>(std::chrono::steady_clock::now() - c->logStart).count();
    ath::PrintTimeLine(nn, iterSeconds, sim, *c->logTime);
  }
}

int BuildGraph(ath_ctx *c) {
  if (c->graphExec && c->graphFp32 == c->fp32 && c->graphMatres == c->matres && c->graphOrder == c->sweepOrder)
    return ATH_OK;
  if (c->graphExec) { hipGraphExecDestroy(c->graphExec); c->graphExec = nullptr; }
  if (c->graph) { hipGraphDestroy(c->graph); c->graph = nullptr; }
  // the whole nonlinear iteration: ghost cells, residual stage, relax,
  // update and norms
  HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  LaunchBoundary(c);
  LaunchResidual(c);
  LaunchAfterResiduals(c);
  LaunchTail(c, false, 0);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess) return Fail(ATH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  c->graph = g;
  HIP_OK(hipGraphInstantiate(&c->graphExec, c->graph, nullptr, nullptr, 0));
  c->graphFp32 = c->fp32;
  c->graphMatres = c->matres;
  c->graphOrder = c->sweepOrder;
  return ATH_OK;
}


}  // namespace

extern "C" {

const char *ath_last_error(void) { return g_err.c_str(); }

int ath_open(const char *inp_path, int device, ath_ctx **out) {
  if (!inp_path || !out) return Fail(ATH_ERR_ARG, "null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return Fail(ATH_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return Fail(ATH_ERR_ARG, "device index out of range");
  auto c = std::make_unique<ath_ctx>();
  std::string err = ath::LoadCase(inp_path, &c->hc);
  if (!err.empty()) return Fail(err.find("cannot open") == 0 ? ATH_ERR_IO : ATH_ERR_CONFIG, err);
  c->device = device;
  HIP_OK(hipSetDevice(device));
  HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  if (int r = Upload(c.get())) return r;
  HIP_OK(hipDeviceSynchronize());
  *out = c.release();
  return ATH_OK;
}

void ath_close(ath_ctx *ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  delete ctx;
}

static void FillInfo(const ath::HostCase &hc, ath_info *o);

int ath_get_info(const ath_ctx *c, ath_info *o) {
  if (!c || !o) return Fail(ATH_ERR_ARG, "null argument");
  FillInfo(c->hc, o);
  return ATH_OK;
}

int ath_block_dims(const ath_ctx *c, int b, int *ni, int *nj, int *nk) {
  if (!c || b < 0 || b >= (int)c->blk.size()) return Fail(ATH_ERR_ARG, "bad block index");
  *ni = c->blk[b].d.ni;
  *nj = c->blk[b].d.nj;
  *nk = c->blk[b].d.nk;
  return ATH_OK;
}

int ath_set_implicit_precision(ath_ctx *c, int precision) {
  if (!c || (precision != ATH_IMPLICIT_FP64 && precision != ATH_IMPLICIT_FP32))
    return Fail(ATH_ERR_ARG, "bad precision");
  HIP_OK(hipSetDevice(c->device));
  if (precision == ATH_IMPLICIT_FP32 && c->P.roeJac)
    return Fail(ATH_ERR_ARG, "the FP32 operator is not available with the approximateRoe flux Jacobian");
  if (precision == ATH_IMPLICIT_FP32)
    if (int r = EnsureFp32(c)) return r;
  c->fp32 = precision;
  c->P.fp32 = precision == ATH_IMPLICIT_FP32 ? 1 : 0;
  return ATH_OK;
}

int ath_set_sweep_ordering(ath_ctx *c, int ordering) {
  if (!c || (ordering != ATH_SWEEP_TILED && ordering != ATH_SWEEP_HYPERPLANE &&
             ordering != ATH_SWEEP_PENCIL && ordering != ATH_SWEEP_ROWS))
    return Fail(ATH_ERR_ARG, "bad sweep ordering");
  c->sweepOrder = ordering;
  return ATH_OK;
}

int ath_get_sweep_ordering(const ath_ctx *c, int *ordering) {
  if (!c || !ordering) return Fail(ATH_ERR_ARG, "null argument");
  *ordering = EffectiveOrder(c);
  return ATH_OK;
}

int ath_set_matrix_residual(ath_ctx *c, int on) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  c->matres = on ? 1 : 0;
  return ATH_OK;
}

static int CopyState(ath_ctx *c, int b, double *host, bool toHost) {
  if (!c || !host || b < 0 || b >= (int)c->blk.size()) return Fail(ATH_ERR_ARG, "bad argument");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  const auto &d = c->blk[b].d;
  std::vector<double> W(5 * (size_t)d.npad);
  HIP_OK(hipMemcpy(W.data(), d.W, W.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < d.nk; ++k)
    for (int j = 0; j < d.nj; ++j)
      for (int i = 0; i < d.ni; ++i) {
        const long p = i + (long)j * d.ni + (long)k * d.ni * d.nj;
        const long q = (i + d.g) + (long)(j + d.g) * d.NI + (long)(k + d.g) * d.NI * d.NJ;
        for (int e = 0; e < 5; ++e) {
          if (toHost) host[5 * p + e] = W[e * d.npad + q];
          else W[e * d.npad + q] = host[5 * p + e];
        }
      }
  if (!toHost) HIP_OK(hipMemcpy(d.W, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice));
  return ATH_OK;
}

int ath_get_state(ath_ctx *c, int b, double *prim) { return CopyState(c, b, prim, true); }
int ath_set_state(ath_ctx *c, int b, const double *prim) {
  return CopyState(c, b, const_cast<double *>(prim), false);
}

int ath_begin_step(ath_ctx *c, int nn) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  CaseRes &cr = *c->cases[0];
  cr.cfl = c->hc.Cfl(nn);
  hipLaunchKernelGGL(ath::KSetScalar, dim3(1), dim3(1), 0, c->stream, cr.dcfl, cr.cfl);
  // StoreOldSolution only when the scheme needs U^n
  // (input::NeedToStoreTimeN, include/input.hpp:174-176)
  if (c->P.implicit) {
    const int copyNm1 = (c->P.multilevel && nn == 0) ? 1 : 0;
    hipLaunchKernelGGL(ath::KBeginStep, dim3(c->mapTotal[kMapCell256]), dim3(kBlk), 0, c->stream,
                       c->Map(kMapCell256), c->P, copyNm1);
  }
  return CheckLaunch();
}

int ath_assign_ghost_cells(ath_ctx *c) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  LaunchBoundary(c);
  if (int r = CheckLaunch()) return r;
  HIP_OK(hipStreamSynchronize(c->stream));
  return ATH_OK;
}

int ath_calc_residual(ath_ctx *c) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  LaunchResidual(c);
  LaunchAfterResiduals(c);
  if (int r = CheckLaunch()) return r;
  HIP_OK(hipStreamSynchronize(c->stream));
  return ATH_OK;
}

int ath_relax(ath_ctx *c, double *cycle_value) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  if (!c->P.implicit) return Fail(ATH_ERR_ARG, "relax requires an implicit time integration");
  HIP_OK(hipSetDevice(c->device));
  if (c->fp32) LaunchRelaxT<float>(c, false, 0);
  else LaunchRelaxT<double>(c, false, 0);
  if (int r = CheckLaunch()) return r;
  HIP_OK(hipStreamSynchronize(c->stream));
  if (cycle_value) {
    std::vector<ath::Partial> part(c->nparts);
    HIP_OK(hipMemcpy(part.data(), c->dpart, part.size() * sizeof(ath::Partial), hipMemcpyDeviceToHost));
    double s = 0, padded = 0;
    for (auto &p : part) s += p.mres;
    for (const auto &br : c->blk) padded += (double)br.d.npad * 5;
    *cycle_value = c->matres ? s / padded : 0.0;
  }
  return ATH_OK;
}

static int ResetCounterIfFull(ath_ctx *c, int need) {
  int slot = 0;
  CaseRes &cr = *c->cases[0];
  HIP_OK(hipMemcpyAsync(&slot, cr.dcounter, sizeof slot, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (slot + need > cr.histCap) {
    hipLaunchKernelGGL(ath::KSetInt, dim3(1), dim3(1), 0, c->stream, cr.dcounter, 0);
    slot = 0;
  }
  return slot;
}

int ath_update(ath_ctx *c, int mm, ath_norms *norms) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  const int slot = ResetCounterIfFull(c, 1);
  if (slot < 0) return slot;
  if (c->fp32) LaunchUpdateT<float>(c);
  else LaunchUpdateT<double>(c);
  if (c->P.multilevel && mm == c->hc.nonlinear - 1) LaunchCopyNm1(c);
  if (int r = CheckLaunch()) return r;
  HIP_OK(hipStreamSynchronize(c->stream));
  // raw norms; the first-iteration normalisation state is only advanced by
  // ath_iterate / ath_run, which know the outer step
  ath_norms n{};
  double *l2first = c->cases[0]->l2first;
  const double keep[5] = {l2first[0], l2first[1], l2first[2], l2first[3], l2first[4]};
  const int r = ReadNorms(c, slot, 1 << 30, mm, &n);
  std::memcpy(l2first, keep, sizeof keep);
  if (norms) *norms = n;
  return r;
}

int ath_iterate(ath_ctx *c, int nn, int mm, ath_norms *norms) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  HIP_OK(hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  const int slot = ResetCounterIfFull(c, 1);
  if (slot < 0) return slot;
  LaunchBoundary(c);
  LaunchResidual(c);
  LaunchAfterResiduals(c);
  LaunchTail(c, false, 0);
  if (c->P.multilevel && mm == c->hc.nonlinear - 1) LaunchCopyNm1(c);
  if (int r = CheckLaunch()) return r;
  HIP_OK(hipStreamSynchronize(c->stream));
  ath_norms n{};
  const int r = ReadNorms(c, slot, nn, mm, &n);
  if (norms) *norms = n;
  LogLine(c, n, nn, mm, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  return r;
}

int ath_run(ath_ctx *c, int nn0, int nsteps, ath_norms *last) {
  if (!c || nsteps < 0) return Fail(ATH_ERR_ARG, "bad argument");
  HIP_OK(hipSetDevice(c->device));
  const int nl = c->hc.nonlinear;
  CaseRes &cr = *c->cases[0];
  const long iters = (long)nsteps * nl;
  // nsteps == 0 only captures and instantiates the iteration's graph
  if (int r = BuildGraph(c)) return r;
  if (iters == 0) return ATH_OK;
  const int nflux = 1;  // one residual stage over all blocks
  // timed runs launch the tail eagerly with an event pair around every sweep
  // pass; untimed runs replay the captured graph
  const int npass = SweepPasses(c);
  const size_t perIt = 2 * (size_t)nflux + 2 * (size_t)npass;
  int sweepLaunches = 0;
  if (c->timing) {
    const size_t need = (size_t)iters * perIt + 2;
    while (c->evs.size() < need) {
      hipEvent_t e;
      HIP_OK(hipEventCreate(&e));
      c->evs.push_back(e);
    }
  }
  hipLaunchKernelGGL(ath::KSetInt, dim3(1), dim3(1), 0, c->stream, cr.dcounter, 0);
  if (c->timing) HIP_OK(hipEventRecord(c->evs[0], c->stream));
  const auto t0 = std::chrono::steady_clock::now();
  // iterations [q0, q1) of this run, from ring slots q0 % histCap on, in
  // order: the PrintResiduals normalisation advances through every one of
  // them, the log gets a line for each, `last` the last one
  auto consume = [&](long q0, long q1) -> int {
    std::vector<double> h(8 * (size_t)(q1 - q0));
    HIP_OK(hipMemcpy(h.data(), cr.dhist + 8 * (size_t)(q0 % cr.histCap), h.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
    const double per = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / (double)q1;
    for (long q = q0; q < q1; ++q) {
      const int qnn = nn0 + (int)(q / nl), qmm = (int)(q % nl);
      ath_norms n{};
      NormsFrom(cr, c->P, &h[8 * (size_t)(q - q0)], qnn, qmm, c->hc.Cfl(qnn), &n);
      LogLine(c, n, qnn, qmm, per);
      if (q == iters - 1 && last) *last = n;
    }
    return ATH_OK;
  };
  long it = 0;
  for (int s = 0; s < nsteps; ++s) {
    const int nn = nn0 + s;
    if (int r = ath_begin_step(c, nn)) return r;
    for (int mm = 0; mm < nl; ++mm, ++it) {
      if (c->timing) {
        LaunchBoundary(c);
        const size_t eb = 2 + (size_t)it * perIt;
        hipEventRecord(c->evs[eb], c->stream);
        LaunchResidual(c);
        hipEventRecord(c->evs[eb + 1], c->stream);
        LaunchAfterResiduals(c);
        sweepLaunches += LaunchTail(c, true, eb + 2 * nflux);
      } else {
        HIP_OK(hipGraphLaunch(c->graphExec, c->stream));
      }
      if (c->P.multilevel && mm == nl - 1) LaunchCopyNm1(c);
      if (int r = CheckLaunch()) return r;
      // the norm history holds histCap iterations; drain it when full
      if ((it + 1) % cr.histCap == 0 && it + 1 < iters) {
        HIP_OK(hipStreamSynchronize(c->stream));
        if (int r = consume(it + 1 - cr.histCap, it + 1)) return r;
        hipLaunchKernelGGL(ath::KSetInt, dim3(1), dim3(1), 0, c->stream, cr.dcounter, 0);
      }
    }
  }
  if (c->timing) HIP_OK(hipEventRecord(c->evs[1], c->stream));
  HIP_OK(hipStreamSynchronize(c->stream));
  if (int r = consume((iters - 1) / cr.histCap * cr.histCap, iters)) return r;
  cr.cfl = c->hc.Cfl(nn0 + nsteps - 1);
  int flag = 0;
  if (int r2 = TakeFlags(c, cr, &flag)) return r2;
  const int r = FlagStatus(flag);
  if (c->timing) {
    ath_timing t{};
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, c->evs[0], c->evs[1]));
    t.run_ms = ms;
    for (long q = 0; q < iters; ++q) {
      const size_t eb = 2 + (size_t)q * perIt;
      for (int b = 0; b < nflux; ++b) {
        HIP_OK(hipEventElapsedTime(&ms, c->evs[eb + 2 * b], c->evs[eb + 2 * b + 1]));
        t.flux_ms += ms;
        t.flux_launches++;
      }
      for (int ps = 0; ps < npass; ++ps) {
        HIP_OK(hipEventElapsedTime(&ms, c->evs[eb + 2 * nflux + 2 * ps], c->evs[eb + 2 * nflux + 2 * ps + 1]));
        t.sweep_ms += ms;
      }
    }
    t.sweep_launches = sweepLaunches;
    t.sweep_passes = (int)(iters * npass);
    c->tm = t;
  }
  return r;
}

// Autotuning of the sweep execution order: every order the case admits
// (pencil, tiled and hyperplane for scalar LU-SGS; the row-parallel sweep,
// the tiled sweep with one or two workgroups per CU and the hyperplane order
// for the block solvers) runs `iterations`
// nonlinear iterations from the current state, timed by events around each
// sweep pass; the fastest is kept.  All orders give the same results bit for
// bit (topological orders of the same dependencies), so the choice changes
// speed only.  The state, the old time levels and the norm normalisation are
// restored afterwards.
int ath_autotune(ath_ctx *c, int iterations, int *chosen, double *pass_ms) {
  if (!c || iterations <= 0) return Fail(ATH_ERR_ARG, "bad argument");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  CaseRes &cr = *c->cases[0];
  // snapshot
  struct Snap { std::vector<double> W, N, Nm1; };
  std::vector<Snap> snap(c->blk.size());
  for (size_t b = 0; b < c->blk.size(); ++b) {
    const auto &d = c->blk[b].d;
    snap[b].W.resize(5 * (size_t)d.npad);
    snap[b].N.resize(5 * (size_t)d.ncell);
    snap[b].Nm1.resize(5 * (size_t)d.ncell);
    HIP_OK(hipMemcpy(snap[b].W.data(), d.W, snap[b].W.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(snap[b].N.data(), d.consN, snap[b].N.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(snap[b].Nm1.data(), d.consNm1, snap[b].Nm1.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  double l2keep[ATH_MAX_EQ];
  std::memcpy(l2keep, cr.l2first, sizeof l2keep);
  auto restore = [&]() -> int {
    for (size_t b = 0; b < c->blk.size(); ++b) {
      const auto &d = c->blk[b].d;
      HIP_OK(hipMemcpy(d.W, snap[b].W.data(), snap[b].W.size() * sizeof(double), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(d.consN, snap[b].N.data(), snap[b].N.size() * sizeof(double), hipMemcpyHostToDevice));
      HIP_OK(hipMemcpy(d.consNm1, snap[b].Nm1.data(), snap[b].Nm1.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    std::memcpy(cr.l2first, l2keep, sizeof l2keep);
    return ATH_OK;
  };
  struct Cand { int order, gridMul; };
  std::vector<Cand> cands;
  if (!c->P.implicit) return Fail(ATH_ERR_ARG, "autotuning needs an implicit time integration");
  if (c->P.roeJac || c->hc.solver != ath::Solver::LUSGS) {
    cands.push_back({EffectiveOrder(c), 0});
  } else if (c->P.block) {
    cands.push_back({ATH_SWEEP_ROWS, 0});
    cands.push_back({ATH_SWEEP_TILED, 1});
    cands.push_back({ATH_SWEEP_TILED, 2});
    cands.push_back({ATH_SWEEP_HYPERPLANE, 0});
  } else {
    cands.push_back({ATH_SWEEP_PENCIL, 0});
    cands.push_back({ATH_SWEEP_TILED, 0});
    cands.push_back({ATH_SWEEP_HYPERPLANE, 0});
  }
  hipDeviceProp_t prop;
  HIP_OK(hipGetDeviceProperties(&prop, c->device));
  const int keepTiming = c->timing;
  const int nl = c->hc.nonlinear;
  const int steps = std::max(1, (iterations + nl - 1) / nl);
  double best = 1e300;
  Cand pick = cands[0];
  for (const Cand &k : cands) {
    c->sweepOrder = k.order;
    if (k.gridMul > 0)
      c->msweepGrid = std::max(1, std::min(c->msweep.ntiles, k.gridMul * prop.multiProcessorCount));
    c->timing = 1;
    const int r = ath_run(c, 0, steps, nullptr);
    c->timing = keepTiming;
    if (int q = restore()) return q;
    if (r != ATH_OK && r != ATH_ERR_NONPHYSICAL) return r;
    const double per = c->tm.sweep_ms / std::max(1, c->tm.sweep_passes);
    if (per < best) {
      best = per;
      pick = k;
    }
  }
  c->sweepOrder = pick.order;
  if (pick.gridMul > 0)
    c->msweepGrid = std::max(1, std::min(c->msweep.ntiles, pick.gridMul * prop.multiProcessorCount));
  c->DropGraph();
  if (chosen) *chosen = pick.order;
  if (pass_ms) *pass_ms = best;
  return ATH_OK;
}

int ath_enable_timing(ath_ctx *c, int on) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  c->timing = on ? 1 : 0;
  return ATH_OK;
}

int ath_get_timing(const ath_ctx *c, ath_timing *out) {
  if (!c || !out) return Fail(ATH_ERR_ARG, "null argument");
  *out = c->tm;
  return ATH_OK;
}

static int GetPhys(ath_ctx *c, int b, double *out, int which) {
  if (!c || !out || b < 0 || b >= (int)c->blk.size()) return Fail(ATH_ERR_ARG, "bad argument");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  const auto &d = c->blk[b].d;
  const long n = d.ncell;
  if (which == 0 || which == 4) {  // residual / update: SoA -> AoS
    std::vector<double> t(5 * (size_t)n);
    if (which == 0) {
      HIP_OK(hipMemcpy(t.data(), d.R, t.size() * sizeof(double), hipMemcpyDeviceToHost));
    } else if (c->fp32) {
      std::vector<float> f(5 * (size_t)n);
      HIP_OK(hipMemcpy(f.data(), d.x32, f.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t q = 0; q < f.size(); ++q) t[q] = f[q];
    } else {
      HIP_OK(hipMemcpy(t.data(), d.x, t.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (long p = 0; p < n; ++p)
      for (int e = 0; e < 5; ++e) out[5 * p + e] = t[e * n + p];
    return ATH_OK;
  }
  const double *src = which == 1 ? d.dt : which == 2 ? d.sr : d.A;
  HIP_OK(hipMemcpy(out, src, n * sizeof(double), hipMemcpyDeviceToHost));
  return ATH_OK;
}

int ath_get_residual(ath_ctx *c, int b, double *r) { return GetPhys(c, b, r, 0); }
int ath_get_dt(ath_ctx *c, int b, double *dt) { return GetPhys(c, b, dt, 1); }
int ath_get_spectral_radius(ath_ctx *c, int b, double *sr) { return GetPhys(c, b, sr, 2); }
int ath_get_diagonal(ath_ctx *c, int b, double *a) {
  if (c && c->P.block) return Fail(ATH_ERR_ARG, "the case uses a block matrix: use ath_get_block_diagonal");
  return GetPhys(c, b, a, 3);
}

int ath_get_wall_distance(ath_ctx *c, int b, double *a) {
  if (!c || !a || b < 0 || b >= (int)c->hc.blocks.size()) return Fail(ATH_ERR_ARG, "bad argument");
  const ath::NearestPoints tree(WallFaceCentres(c->hc));
  const auto &h = c->hc.blocks[b];
  long n = 0;
  for (int k = 0; k < h.nk; ++k)
    for (int j = 0; j < h.nj; ++j)
      for (int i = 0; i < h.ni; ++i) a[n++] = tree.Distance(h.ctr[h.Cell(i, j, k)]);
  return ATH_OK;
}

int ath_get_block_diagonal(ath_ctx *c, int b, int inverse, double *a) {
  if (!c || !a || b < 0 || b >= (int)c->blk.size()) return Fail(ATH_ERR_ARG, "bad argument");
  if (!c->P.block) return Fail(ATH_ERR_ARG, "the case does not use a block matrix");
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  const auto &d = c->blk[b].d;
  const long n = d.ncell;
  if (!inverse) {
    std::vector<double> t(25 * (size_t)n);
    HIP_OK(hipMemcpy(t.data(), d.Ab, t.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (long p = 0; p < n; ++p)
      for (int q = 0; q < 25; ++q) a[25 * p + q] = t[q * n + p];
    return ATH_OK;
  }
  // the inverse is kept in the sweep's tile-blocked layout
  std::vector<double> t(25L * 256 * d.tni * d.tnj * d.tnk);
  HIP_OK(hipMemcpy(t.data(), d.Abi, t.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < d.nk; ++k)
    for (int j = 0; j < d.nj; ++j)
      for (int i = 0; i < d.ni; ++i) {
        const long tile = (long)(i >> 3) + (long)d.tni * ((j >> 3) + (long)d.tnj * (k >> 2));
        const long base = tile * (25 * 256) + ((i & 7) + 8 * ((j & 7) + 8 * (k & 3)));
        const long p = i + (long)j * d.ni + (long)k * d.ni * d.nj;
        for (int q = 0; q < 25; ++q) a[25 * p + q] = t[base + 256 * q];
      }
  return ATH_OK;
}
int ath_get_update(ath_ctx *c, int b, double *du) { return GetPhys(c, b, du, 4); }

// the cell-centre fields of every block of a single-case context, physical
// cells, AoS (what the output files are written from)
static int GatherIo(ath_ctx *c, std::vector<ath::IoBlock> *out) {
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  const auto &hc = c->hc;
  out->assign(c->blk.size(), ath::IoBlock());
  for (size_t b = 0; b < c->blk.size(); ++b) {
    const auto &d = c->blk[b].d;
    ath::IoBlock &o = (*out)[b];
    o.ni = d.ni;
    o.nj = d.nj;
    o.nk = d.nk;
    const long nc = d.ncell;
    o.W.resize(5 * (size_t)nc);
    if (int r = CopyState(c, (int)b, o.W.data(), true)) return r;
    std::vector<double> soa(5 * (size_t)nc);
    auto aos = [&](const double *dev, std::vector<double> *dst) -> int {
      HIP_OK(hipMemcpy(soa.data(), dev, soa.size() * sizeof(double), hipMemcpyDeviceToHost));
      dst->resize(soa.size());
      for (long p = 0; p < nc; ++p)
        for (int e = 0; e < 5; ++e) (*dst)[5 * p + e] = soa[e * nc + p];
      return ATH_OK;
    };
    if (int r = aos(d.R, &o.R)) return r;
    if (c->P.multilevel)
      if (int r = aos(d.consNm1, &o.consNm1)) return r;
    o.dt.resize(nc);
    HIP_OK(hipMemcpy(o.dt.data(), d.dt, nc * sizeof(double), hipMemcpyDeviceToHost));
    if (d.vgrad) {
      std::vector<double> vg(9 * (size_t)d.npad);
      HIP_OK(hipMemcpy(vg.data(), d.vgrad, vg.size() * sizeof(double), hipMemcpyDeviceToHost));
      o.vgrad.resize(9 * (size_t)nc);
      for (int k = 0; k < d.nk; ++k)
        for (int j = 0; j < d.nj; ++j)
          for (int i = 0; i < d.ni; ++i) {
            const long p = i + (long)j * d.ni + (long)k * d.ni * d.nj;
            const long q = (i + d.g) + (long)(j + d.g) * d.NI + (long)(k + d.g) * d.NI * d.NJ;
            for (int e = 0; e < 9; ++e) o.vgrad[9 * p + e] = vg[e * d.npad + q];
          }
    }
    const ath::HostBlock &h = hc.blocks[b];
    o.ctr.resize(3 * (size_t)nc);
    for (int k = 0; k < d.nk; ++k)
      for (int j = 0; j < d.nj; ++j)
        for (int i = 0; i < d.ni; ++i) {
          const long p = i + (long)j * d.ni + (long)k * d.ni * d.nj;
          const ath::Vec3 &v = h.ctr[h.Cell(i, j, k)];
          o.ctr[3 * p] = v.x;
          o.ctr[3 * p + 1] = v.y;
          o.ctr[3 * p + 2] = v.z;
        }
  }
  return ATH_OK;
}

static std::string Join(const char *dir, const std::string &name) {
  std::string d = dir && *dir ? dir : ".";
  return d.back() == '/' ? d + name : d + "/" + name;
}

int ath_write_restart(ath_ctx *c, const char *path, int iteration) {
  if (!c || !path) return Fail(ATH_ERR_ARG, "null argument");
  std::vector<ath::IoBlock> io;
  if (int r = GatherIo(c, &io)) return r;
  const std::string e = ath::WriteRestart(path, c->hc, io, iteration, c->cases[0]->l2first);
  return e.empty() ? ATH_OK : Fail(ATH_ERR_IO, e);
}

int ath_read_restart(ath_ctx *c, const char *path, int *iteration) {
  if (!c || !path) return Fail(ATH_ERR_ARG, "null argument");
  HIP_OK(hipSetDevice(c->device));
  std::vector<ath::IoBlock> io(c->blk.size());
  for (size_t b = 0; b < io.size(); ++b) {
    io[b].ni = c->blk[b].d.ni;
    io[b].nj = c->blk[b].d.nj;
    io[b].nk = c->blk[b].d.nk;
  }
  int iter = 0;
  bool haveNm1 = false;
  double l2[5];
  const std::string e = ath::ReadRestart(path, c->hc, &io, &iter, l2, &haveNm1);
  if (!e.empty()) return Fail(e.find("cannot open") == 0 ? ATH_ERR_IO : ATH_ERR_CONFIG, e);
  for (size_t b = 0; b < io.size(); ++b) {
    if (int r = CopyState(c, (int)b, io[b].W.data(), false)) return r;
    if (c->P.multilevel) {
      // U^{n-1} from the file; with one time level the reference assigns
      // the solution to both n and n-1 (src/output.cpp:887-893)
      const auto &d = c->blk[b].d;
      const long nc = d.ncell;
      std::vector<double> soa(5 * (size_t)nc);
      if (haveNm1) {
        for (long p = 0; p < nc; ++p)
          for (int q = 0; q < 5; ++q) soa[q * nc + p] = io[b].consNm1[5 * p + q];
        HIP_OK(hipMemcpy(d.consNm1, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
      } else {
        hipLaunchKernelGGL(ath::KBeginStep, dim3(c->mapTotal[kMapCell256]), dim3(kBlk), 0, c->stream,
                           c->Map(kMapCell256), c->P, 1);
        HIP_OK(hipStreamSynchronize(c->stream));
      }
    }
  }
  CaseRes &cr = *c->cases[0];
  for (int q = 0; q < 5; ++q) cr.l2first[q] = l2[q];
  cr.restart = true;
  if (iteration) *iteration = iter;
  return ATH_OK;
}

int ath_write_output(ath_ctx *c, const char *dir, int iteration) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  std::vector<ath::IoBlock> io;
  if (int r = GatherIo(c, &io)) return r;
  const std::string grid = c->hc.grid + "_center.xyz";
  const std::string fun = c->hc.simName + "_" + std::to_string(iteration) + "_center.fun";
  std::string e = ath::WriteCenters(Join(dir, grid), c->hc, io);
  if (e.empty()) e = ath::WriteFunction(Join(dir, fun), c->hc, io);
  if (e.empty()) e = ath::WriteMeta(Join(dir, c->hc.simName + "_center.p3d"), c->hc, grid, fun, iteration);
  if (e.empty()) return ATH_OK;
  return Fail(e.find("not available") != std::string::npos || e.find("needs") != std::string::npos
                  ? ATH_ERR_CONFIG : ATH_ERR_IO, e);
}

int ath_open_logs(ath_ctx *c, const char *dir, int append) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  const auto mode = append ? std::ios::app : std::ios::out;
  auto r = std::make_unique<std::ofstream>(Join(dir, c->hc.simName + ".resid"), mode);
  auto t = std::make_unique<std::ofstream>(Join(dir, c->hc.simName + ".tme"), std::ios::out);
  if (!*r || !*t) return Fail(ATH_ERR_IO, "cannot open the residual or timing log in " + std::string(dir ? dir : "."));
  ath::PrintHeaders(c->hc, *r);
  ath::PrintTimeHeader(*t);
  c->logResid = std::move(r);
  c->logTime = std::move(t);
  c->logStart = std::chrono::steady_clock::now();
  return ATH_OK;
}

int ath_close_logs(ath_ctx *c) {
  if (!c) return Fail(ATH_ERR_ARG, "null context");
  c->logResid.reset();
  c->logTime.reset();
  return ATH_OK;
}

static void FillInfo(const ath::HostCase &hc, ath_info *o) {
  o->num_blocks = (int)hc.blocks.size();
  o->num_equations = 5;
  o->num_ghosts = hc.Ghosts();
  o->iterations = hc.iterations;
  o->nonlinear_iterations = hc.nonlinear;
  o->implicit = hc.Implicit() ? 1 : 0;
  o->total_cells = hc.totalCells;
  o->r_ref = hc.rRef;
  o->t_ref = hc.tRef;
  o->l_ref = hc.lRef;
  o->a_ref = hc.aRef;
  o->gas_constant = hc.gasR;
  o->gamma = hc.gamma;
  o->block_matrix = hc.block ? 1 : 0;
}

int ath_inspect_info(const char *inp_path, ath_info *out) {
  if (!inp_path || !out) return Fail(ATH_ERR_ARG, "null argument");
  ath::HostCase hc;
  const std::string err = ath::LoadCase(inp_path, &hc);
  if (!err.empty()) return Fail(err.find("cannot open") == 0 ? ATH_ERR_IO : ATH_ERR_CONFIG, err);
  FillInfo(hc, out);
  return ATH_OK;
}

int ath_inspect_field(const char *inp_path, int block, int which, double *out, long count, long *needed) {
  if (!inp_path || which < 0 || which > 7) return Fail(ATH_ERR_ARG, "bad argument");
  ath::HostCase hc;
  const std::string err = ath::LoadCase(inp_path, &hc);
  if (!err.empty()) return Fail(err.find("cannot open") == 0 ? ATH_ERR_IO : ATH_ERR_CONFIG, err);
  if (block < 0 || block >= (int)hc.blocks.size()) return Fail(ATH_ERR_ARG, "bad block index");
  const auto &b = hc.blocks[block];
  std::vector<double> v;
  if (which == 0) {
    for (int k = 0; k < b.nk; ++k)
      for (int j = 0; j < b.nj; ++j)
        for (int i = 0; i < b.ni; ++i) v.push_back(b.vol[b.Cell(i, j, k)]);
  } else if (which <= 3) {
    v = b.cw[which - 1];
  } else if (which <= 6) {
    for (const auto &a : b.fa[which - 4]) v.insert(v.end(), a.begin(), a.end());
  } else {
    v = b.W0;
  }
  if (needed) *needed = (long)v.size();
  if (out) {
    if (count < (long)v.size()) return Fail(ATH_ERR_ARG, "output buffer too small");
    std::copy(v.begin(), v.end(), out);
  }
  return ATH_OK;
}

int ath_eval_face_flux(int device, int flux, double gas_r, double gas_n, const double *WL,
                       const double *WR, const double *n, int count, double *F) {
  if (count <= 0 || !WL || !WR || !n || !F || (flux != 0 && flux != 1)) return Fail(ATH_ERR_ARG, "bad argument");
  HIP_OK(hipSetDevice(device));
  ath::Gas g{};
  g.R = gas_r;
  g.n = gas_n;
  const double cp = 0.0 + 1.0 * (gas_r * (gas_n + 1.0)), cv = 0.0 + 1.0 * (gas_r * gas_n);
  g.gamma = cp / cv;
  g.cps = gas_r * (gas_n + 1.0);
  g.cvs = gas_r * gas_n;
  double *buf = nullptr;
  const size_t nd = (size_t)count * 18;
  HIP_OK(hipMalloc(&buf, nd * sizeof(double)));
  double *dL = buf, *dR = buf + 5 * (size_t)count, *dn = buf + 10 * (size_t)count, *dF = buf + 13 * (size_t)count;
  hipError_t e = hipMemcpy(dL, WL, 5 * (size_t)count * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dR, WR, 5 * (size_t)count * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dn, n, 3 * (size_t)count * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(ath::KEvalFlux, dim3(Grid(count)), dim3(kBlk), 0, 0, flux, g, dL, dR, dn, count, dF);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(F, dF, 5 * (size_t)count * sizeof(double), hipMemcpyDeviceToHost);
  hipFree(buf);
  if (e != hipSuccess) return Fail(ATH_ERR_HIP, std::string("eval_face_flux: ") + hipGetErrorString(e));
  return ATH_OK;
}

int ath_eval_sst_cells(int device, const double *in, int count, double *out) {
  if (count <= 0 || !in || !out) return Fail(ATH_ERR_ARG, "bad argument");
  HIP_OK(hipSetDevice(device));
  double *buf = nullptr;
  const size_t ni = (size_t)count * ath::sst::kIn, no = (size_t)count * ath::sst::kOut;
  HIP_OK(hipMalloc(&buf, (ni + no) * sizeof(double)));
  hipError_t e = hipMemcpy(buf, in, ni * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(ath::sst::KCells, dim3(Grid(count)), dim3(kBlk), 0, 0, buf, count, buf + ni);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, buf + ni, no * sizeof(double), hipMemcpyDeviceToHost);
  hipFree(buf);
  if (e != hipSuccess) return Fail(ATH_ERR_HIP, std::string("eval_sst_cells: ") + hipGetErrorString(e));
  return ATH_OK;
}

}  // extern "C"

// ---- face-flux evaluation without allocation -------------------------------

// Device staging of ath_eval_face_flux_static: 18 doubles per face (left and
// right states, normal, flux) for at most kStaticFaces faces, 9 MiB on every
// device the library runs on, held until the process ends.  Host code never
// uses the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device (each device holds its own copy).
static constexpr int kStaticFaces = 1 << 16;
static __device__ __align__(256) double scratch[(size_t)kStaticFaces * 18];

namespace {

std::mutex g_scratchMu;
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};

// the device address of `scratch` on `device` (current), looked up on the
// first call on that device
int ScratchOn(int device, double **out) {
  std::lock_guard<std::mutex> lock(g_scratchMu);
  if (device < 0 || device >= kMaxDevices) return Fail(ATH_ERR_ARG, "device index out of range");
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return Fail(ATH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<double *>(p);
  }
  *out = g_scratch[device];
  return ATH_OK;
}

}  // namespace

extern "C" int ath_eval_face_flux_static(int device, int flux, double gas_r, double gas_n, const double *WL,
                                         const double *WR, const double *n, int count, double *F) {
  if (count <= 0 || !WL || !WR || !n || !F || (flux != 0 && flux != 1)) return Fail(ATH_ERR_ARG, "bad argument");
  if (count > kStaticFaces)
    return Fail(ATH_ERR_ARG, "ath_eval_face_flux_static takes at most " + std::to_string(kStaticFaces) +
                                 " faces per call");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return Fail(ATH_ERR_ARG, "device index out of range");
  HIP_OK(hipSetDevice(device));
  double *buf = nullptr;
  if (int r = ScratchOn(device, &buf)) return r;
  ath::Gas g{};
  g.R = gas_r;
  g.n = gas_n;
  const double cp = 0.0 + 1.0 * (gas_r * (gas_n + 1.0)), cv = 0.0 + 1.0 * (gas_r * gas_n);
  g.gamma = cp / cv;
  g.cps = gas_r * (gas_n + 1.0);
  g.cvs = gas_r * gas_n;
  double *dL = buf, *dR = buf + 5 * (size_t)count, *dn = buf + 10 * (size_t)count, *dF = buf + 13 * (size_t)count;
  HIP_OK(hipMemcpy(dL, WL, 5 * (size_t)count * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dR, WR, 5 * (size_t)count * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dn, n, 3 * (size_t)count * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(ath::KEvalFlux, dim3(Grid(count)), dim3(kBlk), 0, 0, flux, g, dL, dR, dn, count, dF);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(F, dF, 5 * (size_t)count * sizeof(double), hipMemcpyDeviceToHost));
  return ATH_OK;
}

// ---- serving --------------------------------------------------------------

struct ath_server {
  std::unique_ptr<ath_ctx> c;
  int precision = ATH_IMPLICIT_FP64;
  int maxSlots = 0;
  double budget = 0;  // bytes the resident cases may hold (0: no limit but free HBM)
  bool haveNumerics = false;
  bool dirty = true;  // the set of running cases changed since the group tables were built
  int nextId = 0;
  double used = 0;    // device bytes reserved by the resident cases (their estimates)
};

namespace {

// device bytes a case will hold (an upper bound of what UploadCase
// allocates: state, geometry, residual and operator arrays, the block
// operator and viscous arrays when selected, sweep tables, norm history)
double EstimateBytes(const ath::HostCase &hc, bool fp32) {
  double t = 4096.0 * 64 + 4096;  // norm history, scalars, states
  for (const auto &h : hc.blocks) {
    const double np = (double)h.NumPadded(), nc = (double)h.NumCells();
    double nf = 0, nfp = 0;
    for (int d = 0; d < 3; ++d) {
      nf += (double)(h.ni + (d == 0)) * (h.nj + (d == 1)) * (h.nk + (d == 2));
      nfp += (double)h.NumFaces(d);
    }
    const double tiles = (double)((h.ni + 7) / 8) * ((h.nj + 7) / 8) * ((h.nk + 3) / 4);
    const double nbf = 2.0 * ((double)h.nj * h.nk + (double)h.nk * h.ni + (double)h.ni * h.nj);
    t += np * 8 * 8 + nf * 32 + nc * 8 * 35 + nc * 4 + nbf * 80 + tiles * 32 + 8192;
    if (fp32) t += nc * 4 * 11 + nbf * 20;
    if (hc.block) t += nc * 25 * 8 + tiles * 256 * 25 * 8;
    if (hc.viscous) t += np * 8 * 3 + nfp * 32 + nf * 8 + nbf + (double)h.g * h.g * 4 * (h.ni + h.nj + h.nk) * 2 * 64 +
                         nbf * 40.0 * h.g;
    if (hc.block && hc.viscous) t += nf * 9 * 8 + np * 10 * 8;
  }
  return t * 1.05 + (1 << 20);
}

// the numerics every case of a server shares: the kernel parameters plus
// the choices the launches dispatch on
bool SameNumerics(const ath::HostCase &a, const ath::HostCase &b) {
  ath::Params pa, pb;
  FillParams(a, &pa);
  FillParams(b, &pb);
  return std::memcmp(&pa, &pb, sizeof pa) == 0 && a.recon == b.recon && a.limiter == b.limiter &&
         a.flux == b.flux && a.timeint == b.timeint && a.solver == b.solver && a.sweeps == b.sweeps &&
         a.nonlinear == b.nonlinear && a.block == b.block && a.viscous == b.viscous && a.roeJac == b.roeJac;
}

// keep what a resident case still needs on the host (dimensions, patches,
// boundary states); drop the geometry and the initial state
void DropGeometry(ath::HostCase &hc) {
  for (auto &h : hc.blocks) {
    std::vector<double>().swap(h.vol);
    std::vector<ath::Vec3>().swap(h.ctr);
    for (int d = 0; d < 3; ++d) {
      std::vector<std::array<double, 4>>().swap(h.fa[d]);
      std::vector<ath::Vec3>().swap(h.fc[d]);
      std::vector<double>().swap(h.cw[d]);
    }
    std::vector<double>().swap(h.W0);
  }
}

CaseRes *FindCase(ath_server *s, int id) {
  for (auto &cp : s->c->cases)
    if (cp->id == id) return cp.get();
  return nullptr;
}

// the residual drop of a case: the largest normalised L2 over the equations
// whose first residual carries a signal (an equation whose first L2 is below
// 1e-10 of the largest, such as the spanwise momentum of a flow without
// spanwise gradients, holds round-off only and is left out)
double Drop(const CaseRes &cr, const ath_norms &n) {
  double big = 0;
  for (int e = 0; e < 5; ++e) big = std::max(big, cr.l2first[e]);
  double d = 0;
  for (int e = 0; e < 5; ++e)
    if (cr.l2first[e] >= 1e-10 * big) d = std::max(d, n.normalized[e]);
  return d;
}

// one nonlinear iteration of every running case, each at its own outer
// step (KCaseCfl advances it): StoreOldSolution, ghost cells, residual,
// relax, update and the per-case norm reductions
void LaunchServeIteration(ath_ctx *c) {
  const int na = (int)c->active.size();
  hipLaunchKernelGGL(ath::KCaseCfl, dim3(Grid(na, 64)), dim3(64), 0, c->stream, c->dcflc, na);
  if (c->P.implicit)
    hipLaunchKernelGGL(ath::KBeginStep, dim3(c->mapTotal[kMapCell256]), dim3(kBlk), 0, c->stream,
                       c->Map(kMapCell256), c->P, 0);
  LaunchBoundary(c);
  LaunchResidual(c);
  LaunchAfterResiduals(c);
  LaunchTail(c, false, 0);
}

int BuildServeGraph(ath_ctx *c) {
  if (c->graphExec) return ATH_OK;
  HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  LaunchServeIteration(c);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess) return Fail(ATH_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  c->graph = g;
  HIP_OK(hipGraphInstantiate(&c->graphExec, c->graph, nullptr, nullptr, 0));
  return ATH_OK;
}

}  // namespace

extern "C" {

int ath_server_open(int device, double hbm_budget, int max_slots, int precision, ath_server **out) {
  if (!out || max_slots <= 0 || hbm_budget < 0 ||
      (precision != ATH_IMPLICIT_FP64 && precision != ATH_IMPLICIT_FP32))
    return Fail(ATH_ERR_ARG, "bad argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return Fail(ATH_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return Fail(ATH_ERR_ARG, "device index out of range");
  auto s = std::make_unique<ath_server>();
  s->c = std::make_unique<ath_ctx>();
  ath_ctx *c = s->c.get();
  c->device = device;
  HIP_OK(hipSetDevice(device));
  HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  if (int r = c->Alloc(&c->dflag, 1)) return r;
  HIP_OK(hipMemset(c->dflag, 0, sizeof(int)));
  c->fp32 = precision == ATH_IMPLICIT_FP32 ? 1 : 0;
  // served cases report no matrix residual (the reference computes it for
  // its log column only); it would cost a fourth off-diagonal pass
  c->matres = 0;
  s->precision = precision;
  s->maxSlots = max_slots;
  s->budget = hbm_budget;
  *out = s.release();
  return ATH_OK;
}

void ath_server_close(ath_server *s) {
  if (!s) return;
  hipSetDevice(s->c->device);
  if (s->c->stream) hipStreamSynchronize(s->c->stream);
  delete s;
}

int ath_estimate_case_bytes(const char *inp_path, int precision, double *bytes) {
  if (!inp_path || !bytes) return Fail(ATH_ERR_ARG, "null argument");
  ath::HostCase hc;
  std::string err = ath::LoadCase(inp_path, &hc);
  if (!err.empty()) return Fail(err.find("cannot open") == 0 ? ATH_ERR_IO : ATH_ERR_CONFIG, err);
  *bytes = EstimateBytes(hc, precision == ATH_IMPLICIT_FP32);
  return ATH_OK;
}

int ath_server_memory(const ath_server *s, double *used, double *free_hbm) {
  if (!s) return Fail(ATH_ERR_ARG, "null server");
  HIP_OK(hipSetDevice(s->c->device));
  size_t fr = 0, tot = 0;
  HIP_OK(hipMemGetInfo(&fr, &tot));
  if (used) *used = s->used;
  if (free_hbm) *free_hbm = (double)fr;
  return ATH_OK;
}

int ath_admit(ath_server *s, const char *inp_path, double residual_drop, int max_iterations, int *slot) {
  if (!s || !inp_path || !slot || residual_drop <= 0 || max_iterations <= 0)
    return Fail(ATH_ERR_ARG, "bad argument");
  *slot = -1;
  ath_ctx *c = s->c.get();
  HIP_OK(hipSetDevice(c->device));
  auto cr = std::make_unique<CaseRes>();
  std::string err = ath::LoadCase(inp_path, &cr->own);
  if (!err.empty()) return Fail(err.find("cannot open") == 0 ? ATH_ERR_IO : ATH_ERR_CONFIG, err);
  const ath::HostCase &hc = cr->own;
  if (!hc.Implicit() || hc.nonlinear != 1 || hc.MultiLevel() || hc.dtIn > 0.0)
    return Fail(ATH_ERR_CONFIG, "a served case must be a steady solve: implicit local time stepping, one "
                                "nonlinear iteration per step");
  if (s->haveNumerics && !SameNumerics(c->hc, hc))
    return Fail(ATH_ERR_CONFIG, "the case's numerics differ from those of the resident cases");
  int running = 0;
  for (auto &cp : c->cases) running += cp->state == ATH_SLOT_RUNNING;
  if (running >= s->maxSlots) return Fail(ATH_ERR_NOMEM, "no free slot");
  const bool fp32 = s->precision == ATH_IMPLICIT_FP32;
  const double need = EstimateBytes(hc, fp32);
  size_t fr = 0, tot = 0;
  HIP_OK(hipMemGetInfo(&fr, &tot));
  const double reserve = 2.0 * (1 << 30);
  if (need + reserve > (double)fr) return Fail(ATH_ERR_NOMEM, "not enough free device memory for the case");
  if (s->budget > 0 && s->used + need > s->budget)
    return Fail(ATH_ERR_NOMEM, "the case does not fit in the remaining memory budget");
  if (!s->haveNumerics) {
    // the shared numerics (no geometry)
    c->hc = hc;
    DropGeometry(c->hc);
    c->hc.blocks.clear();
    c->hc.conns.clear();
    FillParams(c->hc, &c->P);
    c->P.fp32 = fp32 ? 1 : 0;
    s->haveNumerics = true;
  }
  cr->hc = &cr->own;
  if (int r = UploadCase(c, *cr, fp32)) return r;
  HIP_OK(hipDeviceSynchronize());
  DropGeometry(cr->own);
  cr->id = s->nextId++;
  cr->state = ATH_SLOT_RUNNING;
  cr->drop = residual_drop;
  cr->maxIters = max_iterations;
  cr->reserved = need;
  s->used += need;
  *slot = cr->id;
  c->cases.push_back(std::move(cr));
  s->dirty = true;
  return ATH_OK;
}

int ath_step_batch(ath_server *s, int iterations, int *running) {
  if (!s || iterations <= 0) return Fail(ATH_ERR_ARG, "bad argument");
  ath_ctx *c = s->c.get();
  HIP_OK(hipSetDevice(c->device));
  if (s->dirty) {
    c->active.clear();
    for (auto &cp : c->cases)
      if (cp->state == ATH_SLOT_RUNNING) c->active.push_back(cp.get());
    if (int r = RebuildGroup(c)) return r;
    s->dirty = false;
  }
  if (running) *running = (int)c->active.size();
  if (c->active.empty()) return ATH_OK;
  int cap = 1 << 30;
  for (CaseRes *cr : c->active) cap = std::min(cap, cr->histCap);
  if (iterations > cap) return Fail(ATH_ERR_ARG, "more iterations per batch step than the norm history holds");
  if (int r = BuildServeGraph(c)) return r;
  const int na = (int)c->active.size();
  hipLaunchKernelGGL(ath::KZeroCounters, dim3(Grid(na, 64)), dim3(64), 0, c->stream, c->dnorm, na);
  for (int it = 0; it < iterations; ++it) HIP_OK(hipGraphLaunch(c->graphExec, c->stream));
  if (int r = CheckLaunch()) return r;
  HIP_OK(hipStreamSynchronize(c->stream));
  int left = 0;
  int status = ATH_OK;
  std::vector<double> h(8 * (size_t)iterations);
  for (CaseRes *cr : c->active) {
    HIP_OK(hipMemcpy(h.data(), cr->dhist, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    int flag = 0;
    if (int r = TakeFlags(c, *cr, &flag)) return r;
    for (int q = 0; q < iterations; ++q) {
      const int nn = cr->iters + q;
      NormsFrom(*cr, c->P, &h[8 * (size_t)q], nn, 0, cr->hc->Cfl(nn), &cr->last);
    }
    cr->iters += iterations;
    const double d = Drop(*cr, cr->last);
    if (flag & 2) status = FlagStatus(flag);  // a spin limit: the launch itself failed
    if (flag & 5) {
      cr->state = ATH_SLOT_FAILED;
      cr->error = (flag & 1) ? ATH_ERR_NONPHYSICAL : ATH_ERR_DEVICE;
    } else if (!(d == d)) {
      cr->state = ATH_SLOT_FAILED;
      cr->error = ATH_ERR_NONPHYSICAL;
    } else if (d <= cr->drop) {
      cr->state = ATH_SLOT_CONVERGED;
    } else if (cr->iters >= cr->maxIters) {
      cr->state = ATH_SLOT_LIMIT;
    }
    if (cr->state != ATH_SLOT_RUNNING) s->dirty = true;
    else ++left;
  }
  if (running) *running = left;
  return status;
}

int ath_get_slot_info(ath_server *s, int slot, ath_slot_info *out) {
  if (!s || !out) return Fail(ATH_ERR_ARG, "null argument");
  CaseRes *cr = FindCase(s, slot);
  if (!cr) return Fail(ATH_ERR_ARG, "no such slot");
  ath_slot_info o{};
  o.state = cr->state;
  o.error = cr->error;
  o.iterations = cr->iters;
  o.num_blocks = (int)cr->blocks.size();
  o.cells = cr->hc->totalCells;
  o.device_bytes = (double)cr->bytes;
  o.drop = cr->iters > 0 ? Drop(*cr, cr->last) : 1.0;
  o.last = cr->last;
  *out = o;
  return ATH_OK;
}

static int SlotState(ath_server *s, int slot, int b, double *prim, bool toHost) {
  if (!s || !prim) return Fail(ATH_ERR_ARG, "null argument");
  CaseRes *cr = FindCase(s, slot);
  if (!cr) return Fail(ATH_ERR_ARG, "no such slot");
  if (b < 0 || b >= (int)cr->blocks.size()) return Fail(ATH_ERR_ARG, "bad block index");
  ath_ctx *c = s->c.get();
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  const auto &d = cr->blocks[b].d;
  std::vector<double> W(5 * (size_t)d.npad);
  HIP_OK(hipMemcpy(W.data(), d.W, W.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < d.nk; ++k)
    for (int j = 0; j < d.nj; ++j)
      for (int i = 0; i < d.ni; ++i) {
        const long p = i + (long)j * d.ni + (long)k * d.ni * d.nj;
        const long q = (i + d.g) + (long)(j + d.g) * d.NI + (long)(k + d.g) * d.NI * d.NJ;
        for (int e = 0; e < 5; ++e) {
          if (toHost) prim[5 * p + e] = W[e * d.npad + q];
          else W[e * d.npad + q] = prim[5 * p + e];
        }
      }
  if (!toHost) HIP_OK(hipMemcpy(d.W, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice));
  return ATH_OK;
}

int ath_slot_block_dims(ath_server *s, int slot, int block, int *ni, int *nj, int *nk) {
  if (!s || !ni || !nj || !nk) return Fail(ATH_ERR_ARG, "null argument");
  CaseRes *cr = FindCase(s, slot);
  if (!cr) return Fail(ATH_ERR_ARG, "no such slot");
  if (block < 0 || block >= (int)cr->blocks.size()) return Fail(ATH_ERR_ARG, "bad block index");
  *ni = cr->blocks[block].d.ni;
  *nj = cr->blocks[block].d.nj;
  *nk = cr->blocks[block].d.nk;
  return ATH_OK;
}

int ath_get_slot_state(ath_server *s, int slot, int block, double *prim) {
  return SlotState(s, slot, block, prim, true);
}

int ath_set_slot_state(ath_server *s, int slot, int block, const double *prim) {
  return SlotState(s, slot, block, const_cast<double *>(prim), false);
}

int ath_retire(ath_server *s, int slot) {
  if (!s) return Fail(ATH_ERR_ARG, "null server");
  ath_ctx *c = s->c.get();
  HIP_OK(hipSetDevice(c->device));
  HIP_OK(hipStreamSynchronize(c->stream));
  for (size_t q = 0; q < c->cases.size(); ++q)
    if (c->cases[q]->id == slot) {
      const bool wasActive = std::find(c->active.begin(), c->active.end(), c->cases[q].get()) != c->active.end();
      s->used -= c->cases[q]->reserved;
      if (wasActive) {
        // the group tables point at the case's memory: rebuild them before
        // anything launches again
        c->active.erase(std::find(c->active.begin(), c->active.end(), c->cases[q].get()));
        s->dirty = true;
        c->DropGraph();
      }
      c->cases.erase(c->cases.begin() + q);
      if (c->cases.empty()) s->haveNumerics = false;
      return ATH_OK;
    }
  return Fail(ATH_ERR_ARG, "no such slot");
}

}  // extern "C"
