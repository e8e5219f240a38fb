// Output layer files: restart, Plot3D function and centre-grid files, meta
// file, residual and timing log lines.  The formats are the reference's
// (cited per function); the code is this library's.
#include "io.hpp"

#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace ath {

namespace {

template <typename T>
void Put(std::ofstream &f, T v) {
  f.write(reinterpret_cast<const char *>(&v), sizeof v);
}
template <typename T>
bool Get(std::ifstream &f, T *v) {
  f.read(reinterpret_cast<char *>(v), sizeof *v);
  return (bool)f;
}

// WriteBlockDims (include/output.hpp:109-130): block count, then per block
// imax, jmax, kmax and, when asked, the number of variables
void PutDims(std::ofstream &f, const std::vector<IoBlock> &blocks, int numVars) {
  Put<int32_t>(f, (int32_t)blocks.size());
  for (const auto &b : blocks) {
    Put<int32_t>(f, b.ni);
    Put<int32_t>(f, b.nj);
    Put<int32_t>(f, b.nk);
    if (numVars > 0) Put<int32_t>(f, numVars);
  }
}

// calorically perfect single-species thermodynamics of the nondimensional
// state (src/eos.cpp:54-84, src/thermodynamic.cpp:63-113)
struct Thermo {
  double R, n;
  double Cp() const { return R * (n + 1.0); }
  double Cv() const { return R * n; }
  double Gamma() const { return (0.0 + 1.0 * Cp()) / (0.0 + 1.0 * Cv()); }
  double T(const double *W) const { return W[4] / (0.0 + W[0] * R); }
  double SoS(const double *W) const { return std::sqrt(Gamma() * W[4] / W[0]); }
  static double Vel(const double *W) { return std::sqrt((W[1] * W[1] + W[2] * W[2]) + W[3] * W[3]); }
  double Energy(const double *W) const {
    const double v = Vel(W);
    return (0.0 + 1.0 * (0.0 + Cv() * T(W))) + 0.5 * v * v;
  }
  double Enthalpy(const double *W) const {
    const double v = Vel(W);
    return (0.0 + 1.0 * (0.0 + Cp() * T(W))) + 0.5 * v * v;
  }
};

// the restart variables of a single-species laminar or inviscid case
// (src/output.cpp:633-645): density, velocity, pressure, mass fraction
std::vector<std::string> RestartVars(const HostCase &hc) {
  return {"density", "vel_x", "vel_y", "vel_z", "pressure", "mf_" + hc.fluidName};
}

}  // namespace

// WriteRestart (src/output.cpp:591-755)
std::string WriteRestart(const std::string &path, const HostCase &hc, const std::vector<IoBlock> &blocks,
                         int iteration, const double l2first[5]) {
  std::ofstream f(path, std::ios::out | std::ios::binary);
  if (!f) return "cannot open restart file " + path + " for writing";
  const bool two = hc.MultiLevel();
  Put<int32_t>(f, two ? 2 : 1);
  Put<int32_t>(f, iteration);
  Put<int32_t>(f, kNeq);
  Put<int32_t>(f, 1);
  Put<uint64_t>(f, hc.fluidName.size());
  f.write(hc.fluidName.data(), (std::streamsize)hc.fluidName.size());
  for (int e = 0; e < kNeq; ++e) Put<double>(f, l2first[e]);
  const auto vars = RestartVars(hc);
  PutDims(f, blocks, (int)vars.size());
  const double rR = hc.rRef, aR = hc.aRef;
  for (const auto &b : blocks)
    for (long p = 0; p < b.Cells(); ++p) {
      const double *W = &b.W[5 * p];
      Put<double>(f, W[0] * rR);
      Put<double>(f, W[1] * aR);
      Put<double>(f, W[2] * aR);
      Put<double>(f, W[3] * aR);
      Put<double>(f, W[4] * (rR * aR * aR));
      Put<double>(f, 1.0);  // mass fraction of the single species
    }
  if (two)
    for (const auto &b : blocks)
      for (long p = 0; p < b.Cells(); ++p) {
        const double *U = b.consNm1.empty() ? nullptr : &b.consNm1[5 * p];
        const double u[5] = {U ? U[0] : 0.0, U ? U[1] : 0.0, U ? U[2] : 0.0, U ? U[3] : 0.0, U ? U[4] : 0.0};
        Put<double>(f, u[0] * rR);
        Put<double>(f, u[1] * (aR * rR));
        Put<double>(f, u[2] * (aR * rR));
        Put<double>(f, u[3] * (aR * rR));
        Put<double>(f, u[4] * (aR * aR * rR));
        Put<double>(f, 1.0);
      }
  if (!f) return "error writing restart file " + path;
  return "";
}

// ReadRestart (src/output.cpp:756-899), ReadSolFromRestart and
// ReadSolNm1FromRestart (:1177-1290)
std::string ReadRestart(const std::string &path, const HostCase &hc, std::vector<IoBlock> *blocks, int *iteration,
                        double l2first[5], bool *haveNm1) {
  std::ifstream f(path, std::ios::in | std::ios::binary);
  if (!f) return "cannot open restart file " + path;
  int32_t nsol = 0, iter = 0, neq = 0, nsp = 0;
  if (!Get(f, &nsol) || !Get(f, &iter) || !Get(f, &neq) || !Get(f, &nsp)) return "truncated restart header";
  if (neq != kNeq || nsp != 1)
    return "restart file holds " + std::to_string(neq) + " equations and " + std::to_string(nsp) +
           " species; this case has " + std::to_string(kNeq) + " and 1";
  uint64_t len = 0;
  if (!Get(f, &len) || len > 4096) return "bad species name in restart file";
  std::string name(len, ' ');
  f.read(&name[0], (std::streamsize)len);
  if (name != hc.fluidName) return "restart file species " + name + " is not the case's fluid " + hc.fluidName;
  for (int e = 0; e < kNeq; ++e)
    if (!Get(f, &l2first[e])) return "truncated restart residuals";
  int32_t nb = 0;
  if (!Get(f, &nb) || nb != (int32_t)blocks->size())
    return "restart file has " + std::to_string(nb) + " blocks, the grid " + std::to_string(blocks->size());
  for (auto &b : *blocks) {
    int32_t ni, nj, nk, nv;
    if (!Get(f, &ni) || !Get(f, &nj) || !Get(f, &nk) || !Get(f, &nv)) return "truncated restart block sizes";
    if (ni != b.ni || nj != b.nj || nk != b.nk || nv - 1 != neq)
      return "restart block size or variable count does not match the grid";
  }
  const double rR = hc.rRef, aR = hc.aRef;
  for (auto &b : *blocks) {
    b.W.assign(5 * (size_t)b.Cells(), 0.0);
    for (long p = 0; p < b.Cells(); ++p) {
      double v[6];
      for (double &x : v)
        if (!Get(f, &x)) return "truncated restart solution";
      double *W = &b.W[5 * p];
      const double rho = v[0] / rR;
      W[1] = v[1] / aR;
      W[2] = v[2] / aR;
      W[3] = v[3] / aR;
      W[4] = v[4] / (rR * aR * aR);
      W[0] = rho * v[5];  // species density = density x mass fraction
    }
  }
  *haveNm1 = false;
  if (nsol == 2) {
    for (auto &b : *blocks) {
      b.consNm1.assign(5 * (size_t)b.Cells(), 0.0);
      for (long p = 0; p < b.Cells(); ++p) {
        double v[6];
        for (double &x : v)
          if (!Get(f, &x)) return "truncated restart solution at time n-1";
        double *U = &b.consNm1[5 * p];
        const double rho = v[0] / rR;
        U[1] = v[1] / (aR * rR);
        U[2] = v[2] / (aR * rR);
        U[3] = v[3] / (aR * rR);
        U[4] = v[4] / (rR * aR * aR);
        U[0] = rho * v[5];
      }
    }
    *haveNm1 = true;
  }
  *iteration = iter;
  return "";
}

// WriteFunFile (src/output.cpp:206-439) for the cell-centre variables this
// library holds
std::string WriteFunction(const std::string &path, const HostCase &hc, const std::vector<IoBlock> &blocks) {
  for (const auto &v : hc.outputVars) {
    static const char *known[] = {"density", "vel_x", "vel_y", "vel_z", "pressure", "mach", "sos", "dt",
                                  "temperature", "energy", "enthalpy", "cp", "cv", "viscosity", "rank",
                                  "globalPosition", "resid_mass", "resid_mom_x", "resid_mom_y",
                                  "resid_mom_z", "resid_energy", "velGrad_ux", "velGrad_vx", "velGrad_wx",
                                  "velGrad_uy", "velGrad_vy", "velGrad_wy", "velGrad_uz", "velGrad_vz",
                                  "velGrad_wz"};
    bool ok = v == "mf_" + hc.fluidName;
    for (const char *k : known) ok = ok || v == k;
    if (!ok) return "output variable " + v + " is not available";
    if (v == "viscosity" && !hc.viscous) return "output variable viscosity needs the viscous equations";
    if (v.rfind("velGrad_", 0) == 0)
      for (const auto &b : blocks)
        if (b.vgrad.empty()) return "output variable " + v + " needs the cell velocity gradients (block solver, viscous)";
  }
  std::ofstream f(path, std::ios::out | std::ios::binary);
  if (!f) return "cannot open function file " + path + " for writing";
  PutDims(f, blocks, (int)hc.outputVars.size());
  const Thermo th{hc.gasR, hc.gasN};
  const double rR = hc.rRef, aR = hc.aRef, lR = hc.lRef, tR = hc.tRef;
  for (size_t bq = 0; bq < blocks.size(); ++bq) {
    const IoBlock &b = blocks[bq];
    for (const auto &v : hc.outputVars)
      for (long p = 0; p < b.Cells(); ++p) {
        const double *W = &b.W[5 * p];
        const double *R = b.R.empty() ? nullptr : &b.R[5 * p];
        double x = 0.0;
        if (v == "density") x = W[0] * rR;
        else if (v == "vel_x") x = W[1] * aR;
        else if (v == "vel_y") x = W[2] * aR;
        else if (v == "vel_z") x = W[3] * aR;
        else if (v == "pressure") x = W[4] * (rR * aR * aR);
        else if (v == "mach") x = Thermo::Vel(W) / th.SoS(W);
        else if (v == "sos") x = th.SoS(W) * aR;
        else if (v == "dt") x = (b.dt.empty() ? 0.0 : b.dt[p]) / (aR * lR);
        else if (v == "temperature") x = th.T(W) * tR;
        else if (v == "energy") x = th.Energy(W) * (aR * aR);
        else if (v == "enthalpy") x = th.Enthalpy(W) * (aR * aR);
        else if (v == "cp") x = (0.0 + 1.0 * th.Cp()) * (aR * aR / tR);
        else if (v == "cv") x = (0.0 + 1.0 * th.Cv()) * (aR * aR / tR);
        else if (v == "viscosity") {
          const double t = th.T(W) * tR;
          x = (hc.suth[0] * std::pow(t, 1.5)) / (t + hc.suth[1]);
        } else if (v == "rank") x = 0.0;
        else if (v == "globalPosition") x = (double)bq;
        else if (v == "resid_mass") x = (R ? R[0] : 0.0) * (rR * aR * lR * lR);
        else if (v == "resid_mom_x") x = (R ? R[1] : 0.0) * (rR * aR * aR * lR * lR);
        else if (v == "resid_mom_y") x = (R ? R[2] : 0.0) * (rR * aR * aR * lR * lR);
        else if (v == "resid_mom_z") x = (R ? R[3] : 0.0) * (rR * aR * aR * lR * lR);
        else if (v == "resid_energy") x = (R ? R[4] : 0.0) * (rR * std::pow(aR, 3.0) * lR * lR);
        else if (v.rfind("velGrad_", 0) == 0) {
          // velGrad_<c><r>: d(vel_c)/d(x_r)
          const int c = v[8] == 'u' ? 0 : (v[8] == 'v' ? 1 : 2);
          const int r = v[9] == 'x' ? 0 : (v[9] == 'y' ? 1 : 2);
          x = b.vgrad[9 * p + 3 * r + c] * (aR / lR);
        } else x = 1.0;  // mass fraction of the single species
        Put<double>(f, x);
      }
  }
  if (!f) return "error writing function file " + path;
  return "";
}

// WriteCellCenter (src/output.cpp:58-98): a Plot3D grid of the cell centres
std::string WriteCenters(const std::string &path, const HostCase &hc, const std::vector<IoBlock> &blocks) {
  std::ofstream f(path, std::ios::out | std::ios::binary);
  if (!f) return "cannot open grid file " + path + " for writing";
  PutDims(f, blocks, 0);
  for (const auto &b : blocks)
    for (int d = 0; d < 3; ++d)
      for (long p = 0; p < b.Cells(); ++p) Put<double>(f, b.ctr[3 * p + d] * hc.lRef);
  if (!f) return "error writing grid file " + path;
  return "";
}

// WriteMeta (src/output.cpp:903-956), steady form: one grid and function
// file at `iteration`
std::string WriteMeta(const std::string &path, const HostCase &hc, const std::string &gridFile,
                      const std::string &funFile, int iteration) {
  std::ofstream m(path, std::ios::out);
  if (!m) return "cannot open meta file " + path + " for writing";
  m << "{\n\"auto-detect-format\" : true,\n\"format\" : \"binary\",\n\"language\" : \"C\",\n";
  m << "\"filenames\" : [{ \"time\" : " << iteration << ", \"xyz\" : \"" << gridFile << "\", \"function\" : \""
    << funFile << "\" }],\n";
  m << "\"function-names\" : [ ";
  size_t q = 0;
  for (const auto &v : hc.outputVars) {
    m << "\"" << v << "\"";
    if (++q < hc.outputVars.size()) m << ", ";
  }
  m << " ]\n}\n";
  if (!m) return "error writing meta file " + path;
  return "";
}

// PrintHeaders (src/output.cpp:1007-1026)
void PrintHeaders(const HostCase &hc, std::ostream &os) {
  os << std::left << std::setw(7) << "Step" << std::setw(8) << "NL-Iter";
  if (hc.dtIn > 0.0) os << std::left << std::setw(12) << "Time-Step";
  else os << std::left << std::setw(12) << "CFL";
  os << std::left << std::setw(12) << "Res-Mass" << std::setw(12) << "Res-Mom-X" << std::setw(12) << "Res-Mom-Y"
     << std::setw(12) << "Res-Mom-Z" << std::setw(12) << "Res-Energy";
  os << std::left << std::setw(8) << "Max-Eqn" << std::setw(8) << "Max-Blk" << std::setw(8) << "Max-I"
     << std::setw(8) << "Max-J" << std::setw(8) << "Max-K" << std::setw(12) << "Max-Res" << std::setw(12)
     << "Res-Matrix" << "\n";
}

// PrintResiduals (src/output.cpp:1028-1085), the line it prints
void PrintResidualLine(const HostCase &hc, int nn, int mm, double cflOrDt, const double normalized[5], int linfEqn,
                       int linfBlock, int linfI, int linfJ, int linfK, double linf, double matrixResid,
                       std::ostream &os) {
  (void)hc;
  os << std::left << std::setw(7) << nn << std::setw(8) << mm;
  os << std::left << std::setw(12) << std::setprecision(4) << std::scientific << cflOrDt;
  os << std::left << std::setw(12) << normalized[0] << std::setw(12) << normalized[1] << std::setw(12)
     << normalized[2] << std::setw(12) << normalized[3] << std::setw(12) << normalized[4];
  os.unsetf(std::ios::fixed | std::ios::scientific);
  os << std::left << std::setw(8) << linfEqn << std::setw(8) << linfBlock << std::setw(8) << linfI << std::setw(8)
     << linfJ << std::setw(8) << linfK << std::setw(12) << std::setprecision(4) << std::scientific << linf
     << std::setw(12) << matrixResid << "\n";
  os.unsetf(std::ios::fixed | std::ios::scientific);
}

// the timing log (src/logFileManager.cpp:51-57, 102-113)
void PrintTimeHeader(std::ostream &os) {
  os << std::left << std::setw(7) << "Step" << std::setw(16) << "Iter-Time" << std::setw(16) << "Sim-Time" << "\n";
}
void PrintTimeLine(int nn, double iterSeconds, double simSeconds, std::ostream &os) {
  os << std::left << std::setw(7) << nn << std::setw(16) << std::setprecision(6) << std::scientific << iterSeconds
     << std::setw(16) << simSeconds << "\n";
  os.unsetf(std::ios::fixed | std::ios::scientific);
}

}  // namespace ath
