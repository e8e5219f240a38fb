// Files of the reference's output layer, written and read on the host from
// cell-centre fields the device hands over: the binary restart file
// (src/output.cpp:591-899), the Plot3D function file of cell-centre variables
// and its grid of cell centres (src/output.cpp:94-135, 206-439), the
// ParaView meta file (:903-956), and the lines of the residual and timing
// logs (src/output.cpp:1007-1085, src/logFileManager.cpp:34-113).
#pragma once

#include <ostream>
#include <string>
#include <vector>

#include "host_case.hpp"

namespace ath {

// one block's cell-centre fields, physical cells, i fastest
struct IoBlock {
  int ni = 0, nj = 0, nk = 0;
  std::vector<double> W;        // 5 per cell: nondimensional rho, u, v, w, p
  std::vector<double> consNm1;  // 5 per cell: nondimensional conserved U^{n-1} (BDF2), or empty
  std::vector<double> R;        // 5 per cell: residual (flux x area), or empty
  std::vector<double> dt;       // per cell, or empty
  std::vector<double> vgrad;    // 9 per cell: d v_c / d x_r at 3 r + c, or empty
  std::vector<double> ctr;      // 3 per cell: nondimensional cell centre
  long Cells() const { return (long)ni * nj * nk; }
};

// Every function returns an empty string on success, else the error.
std::string WriteRestart(const std::string &path, const HostCase &hc, const std::vector<IoBlock> &blocks,
                         int iteration, const double l2first[5]);
// `blocks` holds the case's dimensions on entry; W (and consNm1 when the
// file holds a second time level) are filled
std::string ReadRestart(const std::string &path, const HostCase &hc, std::vector<IoBlock> *blocks, int *iteration,
                        double l2first[5], bool *haveNm1);
std::string WriteFunction(const std::string &path, const HostCase &hc, const std::vector<IoBlock> &blocks);
std::string WriteCenters(const std::string &path, const HostCase &hc, const std::vector<IoBlock> &blocks);
std::string WriteMeta(const std::string &path, const HostCase &hc, const std::string &gridFile,
                      const std::string &funFile, int iteration);

// residual log (PrintHeaders / PrintResiduals): `cflOrDt` is the CFL number
// of the iteration, or the time step when one is given; normalised L2 of the
// five equations, L-inf equation (1-based), block, i, j, k and value, matrix
// residual
void PrintHeaders(const HostCase &hc, std::ostream &os);
void PrintResidualLine(const HostCase &hc, int nn, int mm, double cflOrDt, const double normalized[5], int linfEqn,
                       int linfBlock, int linfI, int linfJ, int linfK, double linf, double matrixResid,
                       std::ostream &os);
void PrintTimeHeader(std::ostream &os);
void PrintTimeLine(int nn, double iterSeconds, double simSeconds, std::ostream &os);

}  // namespace ath
