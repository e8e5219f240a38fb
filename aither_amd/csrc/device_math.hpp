// Per-face and per-cell arithmetic of the hot path, as device functions.
//
// Each function restates the reference routine it names, keeping the
// reference's operation order.  The residual path (reconstruction, inviscid
// flux, spectral radius, boundary states) is FP64; the implicit-operator
// helpers are templated on the working precision T (double or float).
#pragma once

#include <hip/hip_runtime.h>

namespace ath {

constexpr double kEps = 1.0e-30;  // include/macros.hpp.in:21

struct Gas {
  double R;      // nondimensional gas constant
  double n;      // Cv / R
  double gamma;  // Cp / Cv
  double cps;    // R (n + 1)
  double cvs;    // R n
};

template <typename T>
struct GasT {
  T R, gamma, cps, cvs;
};

template <typename T>
__device__ __forceinline__ GasT<T> Cast(const Gas &g) {
  return {T(g.R), T(g.gamma), T(g.cps), T(g.cvs)};
}

__device__ __forceinline__ double Dmax(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ double Dmin(double a, double b) { return (b < a) ? b : a; }  // std::min

// idealGas::Temperature (src/eos.cpp:103-111), SoS (include/arrayView.hpp:384)
template <typename T>
__device__ __forceinline__ T Temperature(const GasT<T> &g, T rho, T p) {
#pragma clang fp contract(off)
  return p / (rho * g.R);
}
template <typename T>
__device__ __forceinline__ T SoS(const GasT<T> &g, T rho, T p) {
  return sqrt(g.gamma * p / rho);
}
template <typename T>
__device__ __forceinline__ T VelMag(T u, T v, T w) {
#pragma clang fp contract(off)
  return sqrt((u * u + v * v) + w * w);
}
// EnthalpyFunc / InternalEnergy (include/arrayView.hpp:420-444; hf = 0)
template <typename T>
__device__ __forceinline__ T Enthalpy(const GasT<T> &g, T rho, T u, T v, T w, T p) {
#pragma clang fp contract(off)
  const T vel = VelMag(u, v, w);
  return g.cps * Temperature(g, rho, p) + T(0.5) * vel * vel;
}
template <typename T>
__device__ __forceinline__ T TotalEnergy(const GasT<T> &g, T rho, T u, T v, T w, T p) {
#pragma clang fp contract(off)
  const T vel = VelMag(u, v, w);
  return g.cvs * Temperature(g, rho, p) + T(0.5) * vel * vel;
}

// PrimToCons (include/primitive.hpp:151-169)
template <typename T>
__device__ __forceinline__ void PrimToCons(const GasT<T> &g, const T W[5], T U[5]) {
#pragma clang fp contract(off)
  U[0] = W[0];
  U[1] = W[0] * W[1];
  U[2] = W[0] * W[2];
  U[3] = W[0] * W[3];
  U[4] = W[0] * TotalEnergy(g, W[0], W[1], W[2], W[3], W[4]);
}
// primitive(cons, phys) + PressFromEnergy (include/primitive.hpp:129-148,
// src/eos.cpp:40-52, src/thermodynamic.cpp:107-113)
template <typename T>
__device__ __forceinline__ void ConsToPrim(const GasT<T> &g, const T U[5], T W[5]) {
#pragma clang fp contract(off)
  const T rho = U[0];
  W[0] = U[0];
  W[1] = U[1] / rho;
  W[2] = U[2] / rho;
  W[3] = U[3] / rho;
  const T en = U[4] / rho;
  const T vel = VelMag(W[1], W[2], W[3]);
  const T t = (en - T(0.5) * vel * vel) / g.cvs;
  W[4] = (W[0] * g.R) * t;
}
// UpdatePrimWithCons (include/primitive.hpp:206-231); with a single species
// the mass-fraction clip leaves the density unchanged
template <typename T>
__device__ __forceinline__ void UpdatePrim(const GasT<T> &g, const T W[5], const T du[5], T out[5]) {
#pragma clang fp contract(off)
  T U[5];
  PrimToCons(g, W, U);
#pragma unroll
  for (int e = 0; e < 5; ++e) U[e] = U[e] + du[e];
  ConsToPrim(g, U, out);
}

// inviscidFlux::ConstructFromPrim (include/inviscidFlux.hpp:128-159)
template <typename T>
__device__ __forceinline__ void PhysFlux(const GasT<T> &g, const T W[5], T nx, T ny, T nz, T F[5]) {
  const T vn = (W[1] * nx + W[2] * ny) + W[3] * nz;
  const T rvn = W[0] * vn;
  F[0] = W[0] * vn;
  F[1] = rvn * W[1] + W[4] * nx;
  F[2] = rvn * W[2] + W[4] * ny;
  F[3] = rvn * W[3] + W[4] * nz;
  F[4] = rvn * Enthalpy(g, W[0], W[1], W[2], W[3], W[4]);
}

// RoeFlux (include/inviscidFlux.hpp:259-382, src/inviscidFlux.cpp:26-32),
// RoeAveragedState (include/primitive.hpp:245-280)
__device__ inline void RoeFlux(const GasT<double> &g, const double L[5], const double R[5],
                               double nx, double ny, double nz, double F[5]) {
  const double denRatio = sqrt(R[0] / L[0]);
  const double rr = L[0] * denRatio;
  const double inv = 1.0 + denRatio;
  const double ur = (L[1] + denRatio * R[1]) / inv;
  const double vr = (L[2] + denRatio * R[2]) / inv;
  const double wr = (L[3] + denRatio * R[3]) / inv;
  const double pr = (L[4] + denRatio * R[4]) / inv;
  const double hR = Enthalpy(g, rr, ur, vr, wr, pr);
  const double aR = SoS(g, rr, pr);
  const double velNormR = (ur * nx + vr * ny) + wr * nz;
  const double mfR = rr / rr;
  const double d0 = R[0] - L[0], d1 = R[1] - L[1], d2 = R[2] - L[2], d3 = R[3] - L[3],
               d4 = R[4] - L[4];
  const double normVelDiff = (d1 * nx + d2 * ny) + d3 * nz;
  const double aR2 = aR * aR;
  constexpr double entropyFix = 0.1;
  double D0, D1, D2, D3, D4;

  // left moving acoustic wave
  double ws = fabs(velNormR - aR);
  if (ws < entropyFix) ws = 0.5 * (ws * ws / entropyFix + entropyFix);
  double wst = (d4 - rr * aR * normVelDiff) / (2.0 * aR * aR);
  double wss = ws * wst;
  D0 = wss * mfR;
  D1 = wss * (ur - aR * nx);
  D2 = wss * (vr - aR * ny);
  D3 = wss * (wr - aR * nz);
  D4 = wss * (hR - aR * velNormR);

  // entropy wave
  ws = fabs(velNormR);
  wst = -d4 / aR2;
  wss = ws * wst;
  D0 += wss * mfR + ws * d0;
  wst = d0 - d4 / aR2;
  wss = ws * wst;
  D1 += wss * ur;
  D2 += wss * vr;
  D3 += wss * wr;
  D4 += wss * 0.5 * ((ur * ur + vr * vr) + wr * wr);

  // shear wave
  wss = ws * rr;
  D1 += wss * (d1 - normVelDiff * nx);
  D2 += wss * (d2 - normVelDiff * ny);
  D3 += wss * (d3 - normVelDiff * nz);
  D4 += wss * (((ur * d1 + vr * d2) + wr * d3) - velNormR * normVelDiff);

  // right moving acoustic wave
  ws = fabs(velNormR + aR);
  if (ws < entropyFix) ws = 0.5 * (ws * ws / entropyFix + entropyFix);
  wst = (d4 + rr * aR * normVelDiff) / (2.0 * aR * aR);
  wss = ws * wst;
  D0 += wss * mfR;
  D1 += wss * (ur + aR * nx);
  D2 += wss * (vr + aR * ny);
  D3 += wss * (wr + aR * nz);
  D4 += wss * (hR + aR * velNormR);

  double FL[5], FR[5];
  PhysFlux(g, L, nx, ny, nz, FL);
  PhysFlux(g, R, nx, ny, nz, FR);
  F[0] = (FL[0] + (FR[0] - D0)) * 0.5;
  F[1] = (FL[1] + (FR[1] - D1)) * 0.5;
  F[2] = (FL[2] + (FR[2] - D2)) * 0.5;
  F[3] = (FL[3] + (FR[3] - D3)) * 0.5;
  F[4] = (FL[4] + (FR[4] - D4)) * 0.5;
}

// ---------------------------------------------------------------------------
// correctly rounded pow for the few places the reference calls std::pow with
// a non-integer exponent (stagnation-inlet state) or an exponent of 3
// (AUSMPW+ pressure weight).  The host libm returns the correctly rounded
// result in all but vanishingly rare cases; these routines evaluate x^y in
// double-double arithmetic (about 100 significant bits) and round once, so
// device and host agree to the last bit.
struct DD {
  double hi, lo;
};
__host__ __device__ __forceinline__ DD TwoSum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ __forceinline__ DD QuickTwoSum(double a, double b) {
  const double s = a + b;
  return {s, b - (s - a)};
}
__host__ __device__ __forceinline__ DD TwoProd(double a, double b) {
  const double p = a * b;
  return {p, fma(a, b, -p)};
}
__host__ __device__ __forceinline__ DD DDAdd(DD a, DD b) {
  DD s = TwoSum(a.hi, b.hi);
  DD t = TwoSum(a.lo, b.lo);
  s.lo += t.hi;
  s = QuickTwoSum(s.hi, s.lo);
  s.lo += t.lo;
  return QuickTwoSum(s.hi, s.lo);
}
__host__ __device__ __forceinline__ DD DDMul(DD a, DD b) {
  DD p = TwoProd(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return QuickTwoSum(p.hi, p.lo);
}
__host__ __device__ __forceinline__ DD DDMulD(DD a, double b) {
  DD p = TwoProd(a.hi, b);
  p.lo += a.lo * b;
  return QuickTwoSum(p.hi, p.lo);
}
__host__ __device__ __forceinline__ DD DDDiv(DD a, DD b) {
  const double q1 = a.hi / b.hi;
  DD r = DDAdd(a, DDMulD(b, -q1));
  const double q2 = r.hi / b.hi;
  r = DDAdd(r, DDMulD(b, -q2));
  const double q3 = r.hi / b.hi;
  return DDAdd(QuickTwoSum(q1, q2), DD{q3, 0.0});
}

// ln 2 to double-double precision
constexpr double kLn2Hi = 0x1.62e42fefa39efp-1;
constexpr double kLn2Lo = 0x1.abc9e3b39803fp-56;

// log(x), x > 0 finite: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log(m) = 2 atanh(s), s = (m - 1)/(m + 1), |s| < 0.172
__host__ __device__ inline DD DDLog(double x) {
  int e;
  double m = frexp(x, &e);
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  const DD s = DDDiv(DD{m - 1.0, 0.0}, TwoSum(m, 1.0));
  const DD s2 = DDMul(s, s);
  // 2 (s + s^3/3 + s^5/5 + ...), Horner in s^2, 24 terms (|s|^49 < 1e-37)
  DD acc{1.0 / 49.0, 0.0};
  for (int k = 23; k >= 0; --k)
    acc = DDAdd(DDMul(acc, s2), DDDiv(DD{1.0, 0.0}, DD{(double)(2 * k + 1), 0.0}));
  DD lm = DDMulD(DDMul(acc, s), 2.0);
  const DD eln2 = DDAdd(TwoProd((double)e, kLn2Hi), DD{(double)e * kLn2Lo, 0.0});
  return DDAdd(eln2, lm);
}

// exp(t) for |t| < 700: t = k ln2 + r, exp(r) = (exp(r / 256))^256
__host__ __device__ inline DD DDExp(DD t) {
  const double k = rint(t.hi / kLn2Hi);
  DD r = DDAdd(t, DDAdd(TwoProd(-k, kLn2Hi), DD{-k * kLn2Lo, 0.0}));
  r = DD{ldexp(r.hi, -8), ldexp(r.lo, -8)};
  // Taylor series of exp(r) - 1, |r| < 0.0014: 12 terms
  DD term = r, sum = r;
  for (int n = 2; n <= 12; ++n) {
    term = DDDiv(DDMul(term, r), DD{(double)n, 0.0});
    sum = DDAdd(sum, term);
  }
  // (1 + sum)^256 via exp(2a) - 1 = (exp(a) - 1)(exp(a) + 1)
  for (int q = 0; q < 8; ++q) sum = DDMul(sum, DDAdd(sum, DD{2.0, 0.0}));
  DD res = DDAdd(sum, DD{1.0, 0.0});
  return DD{ldexp(res.hi, (int)k), ldexp(res.lo, (int)k)};
}

// x^y for x > 0, y finite
__host__ __device__ inline double PowCR(double x, double y) {
  const DD t = DDMulD(DDLog(x), y);
  const DD p = DDExp(t);
  return p.hi + p.lo;
}

// x^1.5 for x > 0 (Sutherland's law): sqrt(x) to double-double (the exact
// residual x - s^2 by one fused multiply-add, then one Newton correction),
// times x in double-double, rounded once.  The same ~100-bit accuracy as
// PowCR(x, 1.5), so both round to the same double, at a few dozen operations
// instead of the series of PowCR.
__host__ __device__ __forceinline__ double Pow15CR(double x) {
  const double s = sqrt(x);
  const double e = fma(-s, s, x);
  const DD S = QuickTwoSum(s, e / (2.0 * s));
  const DD p = DDMulD(S, x);
  return p.hi + p.lo;
}

// x^3
__host__ __device__ __forceinline__ double CubeCR(double x) {
  const DD x2 = TwoProd(x, x);
  const DD x3 = DDMulD(x2, x);
  return x3.hi + x3.lo;
}

__device__ __forceinline__ double Sign(double x) { return (double)((0.0 < x) - (x < 0.0)); }

// AUSMFlux / AUSMPW+ (include/inviscidFlux.hpp:395-481, member :161-209)
__device__ inline void AusmFlux(const GasT<double> &g, const double L[5], const double R[5],
                                double nx, double ny, double nz, double F[5]) {
  const double velNormL = (L[1] * nx + L[2] * ny) + L[3] * nz;
  const double velNormR = (R[1] * nx + R[2] * ny) + R[3] * nz;
  const double sosStar = sqrt(SoS(g, L[0], L[4]) * SoS(g, R[0], R[4]));
  const double vel = 0.5 * (velNormL + velNormR);
  double sos = sosStar;
  if (vel < 0.0) sos = sosStar * sosStar / Dmax(velNormR, sosStar);
  else if (vel > 0.0) sos = sosStar * sosStar / Dmax(velNormL, sosStar);
  const double ml = velNormL / sos, mr = velNormR / sos;
  const double mlp = (ml + 1.0) * (ml + 1.0), mrm = (mr - 1.0) * (mr - 1.0);
  const double mPlusL = fabs(ml) <= 1.0 ? 0.25 * mlp : 0.5 * (ml + fabs(ml));
  const double mMinusR = fabs(mr) <= 1.0 ? -0.25 * mrm : 0.5 * (mr - fabs(mr));
  const double pPlus = fabs(ml) <= 1.0 ? 0.25 * mlp * (2.0 - ml) : 0.5 * (1.0 + Sign(ml));
  const double pMinus = fabs(mr) <= 1.0 ? 0.25 * mrm * (2.0 + mr) : 0.5 * (1.0 - Sign(mr));
  const double ps = pPlus * L[4] + pMinus * R[4];
  const double w = 1.0 - CubeCR(Dmin(L[4] / R[4], R[4] / L[4]));
  const double fl = fabs(ml) < 1.0 ? L[4] / ps - 1.0 : 0.0;
  const double fr = fabs(mr) < 1.0 ? R[4] / ps - 1.0 : 0.0;
  const double mavg = mPlusL + mMinusR;
  const double mPlusLBar = mavg >= 0.0 ? mPlusL + mMinusR * ((1.0 - w) * (1.0 + fr) - fl)
                                       : mPlusL * w * (1.0 + fl);
  const double mMinusRBar = mavg >= 0.0 ? mMinusR * w * (1.0 + fr)
                                        : mMinusR + mPlusL * ((1.0 - w) * (1.0 + fl) - fr);
  const double vl = mPlusLBar * sos, vr = mMinusRBar * sos;
  const double rvl = L[0] * vl, rvr = R[0] * vr;
  F[0] = L[0] * vl + R[0] * vr;
  F[1] = (rvl * L[1] + pPlus * L[4] * nx) + (rvr * R[1] + pMinus * R[4] * nx);
  F[2] = (rvl * L[2] + pPlus * L[4] * ny) + (rvr * R[2] + pMinus * R[4] * ny);
  F[3] = (rvl * L[3] + pPlus * L[4] * nz) + (rvr * R[3] + pMinus * R[4] * nz);
  F[4] = rvl * Enthalpy(g, L[0], L[1], L[2], L[3], L[4]) +
         rvr * Enthalpy(g, R[0], R[1], R[2], R[3], R[4]);
}

// LimiterVanAlbada / LimiterMinmod (src/limiter.cpp:24-53)
__device__ __forceinline__ double VanAlbada(double r) {
  const double r2 = r * r;
  return Dmax(0.0, (r + r2) / (1.0 + r2));
}
__device__ __forceinline__ double Minmod(double r) { return Dmax(0.0, Dmin(1.0, r)); }

// FaceReconMUSCL for one variable (include/reconstruction.hpp:109-154)
template <int LIM>
__device__ __forceinline__ double Muscl(double u2, double u1, double d1, double kappa,
                                        double dPlus, double dMinus) {
  const double r = (kEps + (d1 - u1) * dPlus) / (kEps + (u1 - u2) * dMinus);
  double lim = 1.0, inv = 1.0;
  if (LIM == 1) {
    lim = VanAlbada(r);
    inv = VanAlbada(1.0 / r);
  } else if (LIM == 2) {
    lim = Minmod(r);
    inv = Minmod(1.0 / r);
  }
  return u1 + (0.25 * ((u1 - u2) * dMinus)) * ((1.0 - kappa) * lim + (1.0 + kappa) * r * inv);
}

// ---------------------------------------------------------------------------
// WENO5 / WENO-Z on non-uniform widths (include/reconstruction.hpp:155-310).
// A stencil is five cells, three upwind (w[0..2], w[2] next to the face) and
// two downwind; all loop bounds below are compile-time constants, so each
// coefficient unrolls to straight-line arithmetic in the reference's order.

// StencilWidth (include/utility.hpp:103-114) over N widths
template <int N>
__device__ __forceinline__ double StencilWN(const double *w, int start, int end) {
  double s = 0.0;
  if (end > start) {
#pragma unroll
    for (int q = 0; q < N; ++q)
      if (q >= start && q < end) s += w[q];
  } else if (start > end) {
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < N; ++q)
      if (q >= end && q < start) t += w[q];
    s = -1.0 * t;
  }
  return s;
}

// coefficient jj of LagrangeCoeff(w, DEG, RR, II) (src/utility.cpp:449-483)
template <int N, int DEG, int RR, int II>
__device__ __forceinline__ double LagrangeCoeffN(const double *w, int jj) {
  double c = 0.0;
#pragma unroll
  for (int mm = 1; mm <= DEG + 1; ++mm) {
    if (mm < jj + 1) continue;
    double numer = 0.0, denom = 1.0;
#pragma unroll
    for (int ll = 0; ll <= DEG + 1; ++ll) {
      if (ll == mm) continue;
      double prod = 1.0;
#pragma unroll
      for (int qq = 0; qq <= DEG + 1; ++qq)
        if (qq != mm && qq != ll) prod *= StencilWN<N>(w, II - RR + qq, II + 1);
      numer += prod;
      denom *= StencilWN<N>(w, II - RR + ll, II - RR + mm);
    }
    c += numer / denom;
  }
  return c * w[II - RR + jj];
}

// the WENO stencils: five widths, first upwind cell at index 2
template <int DEG, int RR>
__device__ __forceinline__ double LagrangeCoeffAt(const double w[5], int jj) {
  return LagrangeCoeffN<5, DEG, RR, 2>(w, jj);
}

// the width-only part of one WENO reconstruction
struct WenoGeom {
  double c0[3], c1[3], c2[3];
  double lw0, lw1, lw2;
};

__device__ __forceinline__ WenoGeom MakeWenoGeom(const double w[5]) {
  WenoGeom G;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    G.c0[q] = LagrangeCoeffAt<2, 2>(w, q);
    G.c1[q] = LagrangeCoeffAt<2, 1>(w, q);
    G.c2[q] = LagrangeCoeffAt<2, 0>(w, q);
  }
  // linear weights as the reference forms them (:273-276): lw1 is the
  // weight of the downwind stencil and is applied to the central one
  G.lw0 = LagrangeCoeffAt<4, 2>(w, 0) / G.c0[0];
  G.lw1 = LagrangeCoeffAt<4, 2>(w, 4) / G.c2[2];
  G.lw2 = 1.0 - G.lw0 - G.lw1;
  return G;
}

// Derivative2nd (include/utility.hpp:116-122)
__device__ __forceinline__ double Deriv2nd(double x0, double x1, double x2, double y0, double y1,
                                           double y2) {
  const double fwd = (y2 - y1) / (0.5 * (x2 + x1));
  const double bck = (y1 - y0) / (0.5 * (x1 + x0));
  return (fwd - bck) / (0.25 * (x2 + x0) + 0.5 * x1);
}

// BetaIntegral (include/reconstruction.hpp:157-181); x^3 correctly rounded
__device__ __forceinline__ double BetaIntAt(double d1, double d2, double dx, double x) {
  return (d1 * d1 * x + d1 * d2 * x * x + d2 * d2 * CubeCR(x) / 3.0) * dx +
         d2 * d2 * x * CubeCR(dx);
}

// Beta0 / Beta1 / Beta2 (include/reconstruction.hpp:183-240)
template <int WHICH>
__device__ __forceinline__ double WenoBeta(double x0, double x1, double x2, double y0, double y1,
                                           double y2) {
  const double d2 = Deriv2nd(x0, x1, x2, y0, y1, y2);
  double d1, h;
  if (WHICH == 0) {
    d1 = (y2 - y1) / (0.5 * (x2 + x1)) + 0.5 * x2 * d2;
    h = x2;
  } else if (WHICH == 1) {
    d1 = (y2 - y1) / (0.5 * (x2 + x1)) - 0.5 * x1 * d2;
    h = x1;
  } else {
    d1 = (y1 - y0) / (0.5 * (x1 + x0)) - 0.5 * x0 * d2;
    h = x0;
  }
  return BetaIntAt(d1, d2, h, 0.5 * h) - BetaIntAt(d1, d2, h, -0.5 * h);
}

// FaceReconWENO for one variable (include/reconstruction.hpp:243-310)
template <bool WENOZ>
__device__ __forceinline__ double Weno(const WenoGeom &G, const double w[5], double u3, double u2,
                                       double u1, double d1, double d2) {
  const double s0 = G.c0[0] * u3 + G.c0[1] * u2 + G.c0[2] * u1;
  const double s1 = G.c1[0] * u2 + G.c1[1] * u1 + G.c1[2] * d1;
  const double s2 = G.c2[0] * u1 + G.c2[1] * d1 + G.c2[2] * d2;
  const double b0 = WenoBeta<0>(w[0], w[1], w[2], u3, u2, u1);
  const double b1 = WenoBeta<1>(w[1], w[2], w[3], u2, u1, d1);
  const double b2 = WenoBeta<2>(w[2], w[3], w[4], u1, d1, d2);
  double n0, n1, n2;
  if (WENOZ) {
    const double tau5 = fabs(b0 - b2);
    const double t0 = tau5 / (1.0e-40 + b0), t1 = tau5 / (1.0e-40 + b1),
                 t2 = tau5 / (1.0e-40 + b2);
    n0 = G.lw0 * (1.0 + t0 * t0);
    n1 = G.lw1 * (1.0 + t1 * t1);
    n2 = G.lw2 * (1.0 + t2 * t2);
  } else {
    const double e0 = 1.0e-6 + b0, e1 = 1.0e-6 + b1, e2 = 1.0e-6 + b2;
    n0 = G.lw0 / (e0 * e0);
    n1 = G.lw1 / (e1 * e1);
    n2 = G.lw2 / (e2 * e2);
  }
  const double sum = n0 + n1 + n2;
  n0 /= sum;
  n1 /= sum;
  n2 /= sum;
  return n0 * s0 + n1 * s1 + n2 * s2;
}

// InvCellSpectralRadius (include/spectralRadius.hpp:43-64)
__device__ __forceinline__ double InvCellSpecRad(const GasT<double> &g, double rho, double u,
                                                 double v, double w, double p, double4 aL,
                                                 double4 aR) {
  const double sx = 0.5 * (aL.x + aR.x), sy = 0.5 * (aL.y + aR.y), sz = 0.5 * (aL.z + aR.z);
  const double m = sqrt((sx * sx + sy * sy) + sz * sz);
  const double nx = sx / m, ny = sy / m, nz = sz / m;
  const double fMag = 0.5 * (aL.w + aR.w);
  return (fabs((u * nx + v * ny) + w * nz) + SoS(g, rho, p)) * fMag;
}

// RoeOffDiagonal for the inviscid equations (src/fluxJacobian.cpp:240-290):
// the change of the Roe flux through the face when the neighbour's state
// moves by du.  The old flux always takes the neighbour as its left state and
// the cell (Wd) as its right one; the new flux takes the updated neighbour as
// the left state for a lower (positive) and as the right state for an upper
// neighbour.  The spectral radius term is zero without the viscous terms,
// applied as +-(du 0).
__device__ __forceinline__ void RoeOffDiag(const GasT<double> &g, const double Wnb[5], const double Wd[5],
                                           const double du[5], double4 a, bool positive, double out[5]) {
  double Fo[5], Fn[5], Wu[5];
  RoeFlux(g, Wnb, Wd, a.x, a.y, a.z, Fo);
  UpdatePrim(g, Wnb, du, Wu);
  if (positive) RoeFlux(g, Wu, Wd, a.x, a.y, a.z, Fn);
  else RoeFlux(g, Wd, Wu, a.x, a.y, a.z, Fn);
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const double fc = a.w * (Fn[e] - Fo[e]);
    const double sd = du[e] * 0.0;
    out[e] = positive ? fc + sd : fc - sd;
  }
}

// F(prim(cons(W) + du)) - F(W), the flux change the off-diagonal products
// take, with every term a product of the increments and the base state, so
// that nothing cancels: rho' = rho + drho, v' = v + dv with
// dv = (d(rho v) - v drho) / rho', dp = (gamma - 1)(d(rho E) - (drho |v'|^2 +
// rho dv.(2 v + dv)) / 2), and
//   dF_rho   = rho dv_n + drho v_n'
//   dF_rhov  = dF_rho v' + rho v_n dv + dp n
//   dF_rhoE  = (rho E + p) dv_n + (d(rho E) + dp) v_n'
// The FP32 operator takes this form: its du falls far below the precision
// of the state once the iteration converges, and F(U + du) - F(U) formed in
// FP32 would round it away.  (The FP64 operator forms both fluxes, as the
// reference's RusanovScalarOffDiagonal does.)
template <typename T>
__device__ __forceinline__ void DeltaFlux(const GasT<T> &g, const T W[5], const T du[5], T nx, T ny, T nz,
                                          T dF[5]) {
  const T rho = W[0], u = W[1], v = W[2], w = W[3], p = W[4];
  const T gm1 = g.gamma - T(1);
  const T dr = du[0], r1 = rho + dr;
  const T dux = (du[1] - u * dr) / r1, duy = (du[2] - v * dr) / r1, duz = (du[3] - w * dr) / r1;
  const T u1 = u + dux, v1 = v + duy, w1 = w + duz;
  const T q1 = (u1 * u1 + v1 * v1) + w1 * w1;
  const T dq = (dux * (T(2) * u + dux) + duy * (T(2) * v + duy)) + duz * (T(2) * w + duz);
  const T dp = gm1 * (du[4] - T(0.5) * (dr * q1 + rho * dq));
  const T vn = (u * nx + v * ny) + w * nz, dvn = (dux * nx + duy * ny) + duz * nz, vn1 = vn + dvn;
  const T dm = rho * dvn + dr * vn1;
  const T rE = p / gm1 + T(0.5) * rho * ((u * u + v * v) + w * w);
  const T rvn = rho * vn;
  dF[0] = dm;
  dF[1] = dm * u1 + rvn * dux + dp * nx;
  dF[2] = dm * v1 + rvn * duy + dp * ny;
  dF[3] = dm * w1 + rvn * duz + dp * nz;
  dF[4] = (rE + p) * dvn + (du[4] + dp) * vn1;
}

// RusanovScalarOffDiagonal for the inviscid equations
// (src/fluxJacobian.cpp:122-162): 0.5|A| (F(U_nb + du) - F(U_nb)) +/- lambda du
// with lambda = InvFaceSpectralRadius (include/spectralRadius.hpp:66-80)
// With the viscous terms FaceSpectralRadius adds the viscous face spectral
// radius `lamv` (include/spectralRadius.hpp:192-203).
template <typename T>
__device__ __forceinline__ void ScalarOffDiag(const GasT<T> &g, const T Wnb[5], const T du[5],
                                              T nx, T ny, T nz, T mag, bool positive, T out[5],
                                              bool viscous = false, T lamv = T(0)) {
  T dF[5];
  if constexpr (sizeof(T) == 4) {
    DeltaFlux(g, Wnb, du, nx, ny, nz, dF);
  } else {
    T Wu[5], Fo[5], Fn[5];
    UpdatePrim(g, Wnb, du, Wu);
    PhysFlux(g, Wnb, nx, ny, nz, Fo);
    PhysFlux(g, Wu, nx, ny, nz, Fn);
#pragma unroll
    for (int e = 0; e < 5; ++e) dF[e] = Fn[e] - Fo[e];
  }
  const T half = T(0.5) * mag;
  T lam = T(0.5) * mag * (fabs((Wnb[1] * nx + Wnb[2] * ny) + Wnb[3] * nz) + SoS(g, Wnb[0], Wnb[4]));
  if (viscous) lam += lamv;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const T fc = half * dF[e];
    out[e] = positive ? fc + du[e] * lam : fc - du[e] * lam;
  }
}

// ---------------------------------------------------------------------------
// 5x5 blocks (matrixSolver blusgs / bdplur), row-major

// fluxJacobian::InvFluxJacobian and RusanovFluxJacobian
// (include/fluxJacobian.hpp:447-560) for one species (mass fraction 1):
// 0.5 |A| dF/dU at the state, then the face spectral radius added to
// (positive) or subtracted from each entry of the identity, the off-diagonal
// entries +/- 0 included
template <typename T>
__device__ __forceinline__ void RusanovJacobian(const GasT<T> &g, const T W[5], T nx, T ny, T nz, T mag,
                                                bool positive, T J[25]) {
#pragma clang fp contract(off)
  const T u = W[1], v = W[2], w = W[3];
  const T vn = (u * nx + v * ny) + w * nz;
  const T sr = T(0.5) * mag * (fabs(vn) + SoS(g, W[0], W[4]));
  const T gamma = g.gamma, gm1 = gamma - T(1);
  const T phi = T(0.5) * gm1 * ((u * u + v * v) + w * w);
  const T a1 = gamma * TotalEnergy(g, W[0], u, v, w, W[4]) - phi;
  const T a3 = gamma - T(2);
  J[0] = vn * T(0);
  J[1] = nx;
  J[2] = ny;
  J[3] = nz;
  J[4] = T(0);
  J[5] = phi * nx - u * vn;
  J[10] = phi * ny - v * vn;
  J[15] = phi * nz - w * vn;
  J[20] = vn * (phi - a1);
  J[6] = vn - a3 * nx * u;
  J[11] = v * nx - gm1 * u * ny;
  J[16] = w * nx - gm1 * u * nz;
  J[21] = a1 * nx - gm1 * u * vn;
  J[7] = u * ny - gm1 * v * nx;
  J[12] = vn - a3 * ny * v;
  J[17] = w * ny - gm1 * v * nz;
  J[22] = a1 * ny - gm1 * v * vn;
  J[8] = u * nz - gm1 * w * nx;
  J[13] = v * nz - gm1 * w * ny;
  J[18] = vn - a3 * nz * w;
  J[23] = a1 * nz - gm1 * w * vn;
  J[9] = gm1 * nx;
  J[14] = gm1 * ny;
  J[19] = gm1 * nz;
  J[24] = gamma * vn;
  const T f = T(0.5) * mag;
#pragma unroll
  for (int q = 0; q < 25; ++q) {
    const T dis = (q % 6 == 0) ? sr : T(0);
    const T jq = J[q] * f;
    J[q] = positive ? jq + dis : jq - dis;
  }
}

// ArrayMultiplication, block branch (include/fluxJacobian.hpp:50-86):
// out[r] = sum over c in order of M[r][c] v[c], from +0
template <typename T>
__device__ __forceinline__ void BlockMult(const T M[25], const T v[5], T out[5]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < 5; ++c) s += M[r * 5 + c] * v[c];
    out[r] = s;
  }
}

// MatrixInverse (src/matrix.cpp:57-105) with FindMaxInColumn (:145-165),
// in registers: row r's pivot is the largest |m[i][r]|, i >= r, taken before
// row r's leading entries are eliminated (row 0 when all vanish); row r is
// eliminated against rows 0..r-1 (LinearCombRow over the whole row),
// normalised from the pivot column on, and the upper triangle is cleared from
// the bottom row up.  The identity takes every row operation.  Returns false
// on a zero pivot.
template <typename T>
__device__ __forceinline__ void SwapRows5(T m[25], int r1, int r2) {
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const T t = m[r1 * 5 + c];
    m[r1 * 5 + c] = m[r2 * 5 + c];
    m[r2 * 5 + c] = t;
  }
}
template <typename T>
__device__ __forceinline__ bool GJInverse5(T m[25], T I[25]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int q = 0; q < 25; ++q) I[q] = (q % 6 == 0) ? T(1) : T(0);
  bool ok = true;
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    T mv = T(0);
    int pr = 0;
#pragma unroll
    for (int i = r; i < 5; ++i) {
      const T a = fabs(m[i * 5 + r]);
      if (a > mv) {
        mv = a;
        pr = i;
      }
    }
    // the swap with a runtime partner, as selects over the candidate rows
#pragma unroll
    for (int i = 0; i < 5; ++i)
      if (i != r && pr == i) {
        SwapRows5(m, r, i);
        SwapRows5(I, r, i);
      }
#pragma unroll
    for (int i = 0; i < r; ++i) {
      const T factor = m[r * 5 + i] / m[i * 5 + i];
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        m[r * 5 + c] = m[r * 5 + c] - factor * m[i * 5 + c];
        I[r * 5 + c] = I[r * 5 + c] - factor * I[i * 5 + c];
      }
    }
    if (m[r * 5 + r] == T(0)) ok = false;
    const T nf = T(1) / m[r * 5 + r];
#pragma unroll
    for (int c = r; c < 5; ++c) m[r * 5 + c] *= nf;
#pragma unroll
    for (int c = 0; c < 5; ++c) I[r * 5 + c] *= nf;
  }
#pragma unroll
  for (int r = 3; r >= 0; --r)
#pragma unroll
    for (int i = 4; i > r; --i) {
      const T factor = m[r * 5 + i];
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        m[r * 5 + c] = m[r * 5 + c] - factor * m[i * 5 + c];
        I[r * 5 + c] = I[r * 5 + c] - factor * I[i * 5 + c];
      }
    }
  return ok;
}

}  // namespace ath
