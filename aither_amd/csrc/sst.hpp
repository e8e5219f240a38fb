// SST k-omega 2003 cell model (turbKWSst, src/turbulence.cpp:573-760,
// include/turbulence.hpp:489-540; the turbModel base, src/turbulence.cpp:
// 36-105; sutherland::Lambda with zero bulk viscosity, src/transport.cpp:
// 172-175): per cell, the blending functions F1 and F2, the limited eddy
// viscosity, the k and omega source terms with the 2003 production limiter,
// and the diagonal of the source Jacobian, in the reference's operation
// order.  Tensors are in the reference's data order (row-major 3 x 3).
#pragma once
#include <hip/hip_runtime.h>

namespace ath {
namespace sst {

constexpr double kBetaStar = 0.09, kSigmaW2 = 0.856, kBeta1 = 0.075, kBeta2 = 0.0828, kGamma1 = 5.0 / 9.0,
                 kGamma2 = 0.44, kA1 = 0.31, kProd2Dest = 10.0, kEps = 1.0e-30;
constexpr int kIn = 27, kOut = 7;

// tensor::DoubleDotTrans: elementwise product summed from 0 in data order
__device__ __forceinline__ double DDot(const double *A, const double *B) {
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 9; ++q) s = s + A[q] * B[q];
  return s;
}

// in: W[7] (rho, u, v, w, p, k, omega), velocity gradient[9], k gradient[3],
//     omega gradient[3], mu (laminar), wall distance, volume, scaling,
//     1/scaling
// out: f1, f2, mut, source k, source omega, source Jacobian (0,0), (1,1)
__device__ __forceinline__ void Cell(const double *in, double *out) {
#pragma clang fp contract(off)
  const double *W = in, *G = in + 7, *kg = in + 16, *wg = in + 19;
  const double mu = in[22], d = in[23], vol = in[24], sc = in[25], isc = in[26];
  const double rho = W[0], k = W[5], om = W[6];
  // EddyViscAndBlending
  const double alpha1 = (sc * sqrt(k)) / ((kBetaStar * om) * (d + kEps));
  const double alpha2 = (((sc * sc) * 500.0) * mu) / ((((d + kEps) * (d + kEps)) * rho) * om);
  const double dotkw = ((0.0 + kg[0] * wg[0]) + kg[1] * wg[1]) + kg[2] * wg[2];
  double cdkw = (((2.0 * rho) * kSigmaW2) / om) * dotkw;
  cdkw = cdkw > 1.0e-10 ? cdkw : 1.0e-10;
  const double alpha3 = (((4.0 * rho) * kSigmaW2) * k) / ((cdkw * (d + kEps)) * (d + kEps));
  double arg1 = alpha1 > alpha2 ? alpha1 : alpha2;
  arg1 = arg1 < alpha3 ? arg1 : alpha3;
  const double f1 = tanh(pow(arg1, 4.0));
  const double arg2 = (2.0 * alpha1) > alpha2 ? (2.0 * alpha1) : alpha2;
  const double f2 = tanh(arg2 * arg2);
  // EddyVisc: mean strain rate sqrt(2 S:S), S = (G + G^T) / 2
  double S[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) S[3 * r + c] = 0.5 * (G[3 * r + c] + G[3 * c + r]);
  const double msr = sqrt(2.0 * DDot(S, S));
  const double lim = (sc * msr) * f2, aw = kA1 * om;
  const double mut = ((rho * kA1) * k) / (aw > lim ? aw : lim);
  // CalcTurbSrc
  const double gamma = f1 * kGamma1 + (1.0 - f1) * kGamma2;
  const double beta = f1 * kBeta1 + (1.0 - f1) * kBeta2;
  const double tkeDest = (isc * kBetaStar) * (((rho * k) * om) * 1.0);
  const double omgDest = (isc * beta) * ((rho * om) * om);
  // Boussinesq Reynolds stress lambda tr(G) I + mut (G + G^T) - 2/3 rho k I
  // (lambda = 0 - 2/3 mut) and the production tau : G
  const double lambda = 0.0 - (2.0 / 3.0) * mut;
  const double lt = lambda * ((G[0] + G[4]) + G[8]);
  const double c2 = ((2.0 / 3.0) * rho) * k;
  double tau[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double I = r == c ? 1.0 : 0.0;
      tau[3 * r + c] = (lt * I + mut * (G[3 * r + c] + G[3 * c + r])) - c2 * I;
    }
  double tkeProd = sc * DDot(tau, G);
  const double pl = kProd2Dest * tkeDest;
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
  // TurbSrcJac (phi = 1)
  out[5] = ((((-2.0 * kBetaStar) * om) * 1.0) * vol) * isc;
  out[6] = (((-2.0 * beta) * om) * vol) * isc;
}

__global__ void __launch_bounds__(256) KCells(const double *__restrict__ in, int n, double *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double x[kIn], y[kOut];
#pragma unroll
  for (int q = 0; q < kIn; ++q) x[q] = in[(size_t)t * kIn + q];
  Cell(x, y);
#pragma unroll
  for (int q = 0; q < kOut; ++q) out[(size_t)t * kOut + q] = y[q];
}

}  // namespace sst
}  // namespace ath
