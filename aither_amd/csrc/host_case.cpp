// Input file, Plot3D grid, FP64 geometry and initial state for the device
// solver.  See host_case.hpp.
#include "host_case.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>

namespace ath {
namespace {

// --- small vector algebra, in the reference's operation order -------------
// vector3d::DotProd is std::inner_product from T(0)
// (include/vector3d.hpp:303-305); CrossProd include/vector3d.hpp:332-341.
inline Vec3 Add(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Vec3 Sub(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline Vec3 Scale(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double Dot(Vec3 a, Vec3 b) {
  double s = 0.0;
  s += a.x * b.x;
  s += a.y * b.y;
  s += a.z * b.z;
  return s;
}
inline double Mag(Vec3 a) { return std::sqrt(Dot(a, a)); }
inline Vec3 Cross(Vec3 a, Vec3 b) {
  return {a.y * b.z - a.z * b.y, -1.0 * (a.x * b.z - a.z * b.x),
          a.x * b.y - a.y * b.x};
}
inline std::array<double, 4> UnitMag(Vec3 a) {  // unitVec3dMag ctor
  const double m = Mag(a);
  return {a.x / m, a.y / m, a.z / m, Mag(a)};
}

std::string Trim(std::string s) {
  // Trim (src/inputStates.cpp:231-252): strip whitespace, drop comments
  auto notsp = [](unsigned char c) { return !std::isspace(c); };
  s.erase(s.begin(), std::find_if(s.begin(), s.end(), notsp));
  s.erase(std::find_if(s.rbegin(), s.rend(), notsp).base(), s.end());
  const auto h = s.find('#');
  if (h != std::string::npos) s = s.substr(0, h);
  s.erase(std::find_if(s.rbegin(), s.rend(), notsp).base(), s.end());
  return s;
}

// entries "name(k=v; k=v)" of a <...> list
std::vector<std::string> ListEntries(const std::string &body) {
  std::vector<std::string> out;
  int depth = 0;
  size_t start = std::string::npos;
  for (size_t p = 0; p < body.size(); ++p) {
    if (body[p] == '(') {
      if (depth == 0) {
        size_t q = p;
        while (q > 0 && (std::isalnum((unsigned char)body[q - 1]) || body[q - 1] == '_')) --q;
        start = q;
      }
      ++depth;
    } else if (body[p] == ')') {
      --depth;
      if (depth == 0 && start != std::string::npos) {
        out.push_back(body.substr(start, p - start + 1));
        start = std::string::npos;
      }
    }
  }
  return out;
}

std::string EntryName(const std::string &e) { return Trim(e.substr(0, e.find('('))); }

std::map<std::string, std::string> EntryParams(const std::string &e) {
  std::map<std::string, std::string> m;
  const auto lp = e.find('('), rp = e.rfind(')');
  if (lp == std::string::npos || rp == std::string::npos) return m;
  std::stringstream ss(e.substr(lp + 1, rp - lp - 1));
  std::string tok;
  while (std::getline(ss, tok, ';')) {
    const auto eq = tok.find('=');
    if (eq == std::string::npos) continue;
    auto v = Trim(tok.substr(eq + 1));
    if (!v.empty() && v.back() == ',') v.pop_back();
    m[Trim(tok.substr(0, eq))] = v;
  }
  return m;
}

std::array<double, 3> ParseVec(const std::string &s) {
  std::array<double, 3> v{0, 0, 0};
  const auto lb = s.find('[');
  if (lb == std::string::npos) return v;
  std::sscanf(s.c_str() + lb + 1, "%lf , %lf , %lf", &v[0], &v[1], &v[2]);
  return v;
}

// fluid properties: <name>.dat in the case directory first
// (src/fluid.cpp:105-118), then the values of fluidDatabase/air.dat and N2.dat
// fluid database entry (fluidDatabase/<name>.dat): n, molar mass and the
// Sutherland coefficients of viscosity and conductivity
bool FluidProps(const std::string &dir, const std::string &name, double *n,
                double *molar, std::array<double, 4> *suth) {
  static const char *sk[4] = {"sutherlandViscosityC1", "sutherlandViscosityS",
                              "sutherlandConductivityC1", "sutherlandConductivityS"};
  std::ifstream f(dir + "/" + name + ".dat");
  if (f) {
    std::string line;
    int got = 0;
    while (std::getline(f, line)) {
      line = Trim(line);
      const auto c = line.find(':');
      if (c == std::string::npos) continue;
      const auto key = Trim(line.substr(0, c));
      if (key == "n") { *n = std::stod(line.substr(c + 1)); got |= 1; }
      if (key == "molarMass") { *molar = std::stod(line.substr(c + 1)) / 1000.; got |= 2; }
      for (int q = 0; q < 4; ++q)
        if (key == sk[q]) { (*suth)[q] = std::stod(line.substr(c + 1)); got |= 4 << q; }
    }
    if (got == 63) return true;
  }
  if (name == "air") {
    *n = 2.5;
    *molar = 28.97 / 1000.;
    *suth = {1.458e-6, 110.4, 2.495e-3, 194.0};
    return true;
  }
  if (name == "N2") {
    *n = 2.5;
    *molar = 28.0134 / 1000.;
    *suth = {1.4742e-06, 1.2846e+02, 2.6834e-03, 2.5615e+02};
    return true;
  }
  return false;
}

// (dir, d3, d1, d2) -> (i, j, k); direction 1/2 follow the cyclic order of
// include/multiArray3d.hpp:231-243
inline void Map3(int dir, int d3, int d1, int d2, int *i, int *j, int *k) {
  if (dir == 0) { *i = d3; *j = d1; *k = d2; }
  else if (dir == 1) { *i = d2; *j = d3; *k = d1; }
  else { *i = d1; *j = d2; *k = d3; }
}

// plot3dBlock geometry (src/plot3d.cpp:36-359, 490-498)
double Pyramid(Vec3 p, Vec3 a, Vec3 b, Vec3 c, Vec3 d) {
  Vec3 s = Add(Add(Add(Sub(a, p), Sub(b, p)), Sub(c, p)), Sub(d, p));
  const Vec3 xp = Scale(s, 0.25);
  const Vec3 xac = Sub(c, a), xbd = Sub(d, b);
  return 1.0 / 6.0 * Dot(xp, Cross(xac, xbd));
}

std::string BuildGeometry(HostBlock &b, const std::vector<Vec3> &X) {
  const int ni = b.ni, nj = b.nj, nk = b.nk;
  auto P = [&](int i, int j, int k) {
    return X[(size_t)i + (size_t)j * (ni + 1) + (size_t)k * (ni + 1) * (nj + 1)];
  };
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        Vec3 c = P(i, j, k);
        c = Add(c, P(i + 1, j, k));
        c = Add(c, P(i, j + 1, k));
        c = Add(c, P(i + 1, j + 1, k));
        c = Add(c, P(i, j, k + 1));
        c = Add(c, P(i + 1, j, k + 1));
        c = Add(c, P(i, j + 1, k + 1));
        c = Add(c, P(i + 1, j + 1, k + 1));
        c = Scale(c, 0.125);
        double v = 0.0;
        v += Pyramid(c, P(i, j, k), P(i, j, k + 1), P(i, j + 1, k + 1), P(i, j + 1, k));
        v += Pyramid(c, P(i + 1, j, k), P(i + 1, j + 1, k), P(i + 1, j + 1, k + 1), P(i + 1, j, k + 1));
        v += Pyramid(c, P(i, j, k), P(i + 1, j, k), P(i + 1, j, k + 1), P(i, j, k + 1));
        v += Pyramid(c, P(i, j + 1, k), P(i, j + 1, k + 1), P(i + 1, j + 1, k + 1), P(i + 1, j + 1, k));
        v += Pyramid(c, P(i, j, k), P(i, j + 1, k), P(i + 1, j + 1, k), P(i + 1, j, k));
        v += Pyramid(c, P(i, j, k + 1), P(i + 1, j, k + 1), P(i + 1, j + 1, k + 1), P(i, j + 1, k + 1));
        if (!(v > 0.0)) return "negative volume in Plot3D block";
        b.vol[b.Cell(i, j, k)] = v;
        b.ctr[b.Cell(i, j, k)] = c;
      }
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i <= ni; ++i) {
        const Vec3 xac = Sub(P(i, j + 1, k + 1), P(i, j, k));
        const Vec3 xbd = Sub(P(i, j + 1, k), P(i, j, k + 1));
        b.fa[0][b.Face(0, i, j, k)] = UnitMag(Scale(Cross(xbd, xac), 0.5));
        Vec3 s = P(i, j, k);
        s = Add(s, P(i, j + 1, k));
        s = Add(s, P(i, j, k + 1));
        s = Add(s, P(i, j + 1, k + 1));
        b.fc[0][b.Face(0, i, j, k)] = Scale(s, 0.25);
      }
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j <= nj; ++j)
      for (int i = 0; i < ni; ++i) {
        const Vec3 xac = Sub(P(i, j, k + 1), P(i + 1, j, k));
        const Vec3 xbd = Sub(P(i, j, k), P(i + 1, j, k + 1));
        b.fa[1][b.Face(1, i, j, k)] = UnitMag(Scale(Cross(xbd, xac), 0.5));
        Vec3 s = P(i, j, k);
        s = Add(s, P(i + 1, j, k));
        s = Add(s, P(i, j, k + 1));
        s = Add(s, P(i + 1, j, k + 1));
        b.fc[1][b.Face(1, i, j, k)] = Scale(s, 0.25);
      }
  for (int k = 0; k <= nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) {
        const Vec3 xac = Sub(P(i, j + 1, k), P(i + 1, j, k));
        const Vec3 xbd = Sub(P(i + 1, j + 1, k), P(i, j, k));
        b.fa[2][b.Face(2, i, j, k)] = UnitMag(Scale(Cross(xbd, xac), 0.5));
        Vec3 s = P(i, j, k);
        s = Add(s, P(i + 1, j, k));
        s = Add(s, P(i, j + 1, k));
        s = Add(s, P(i + 1, j + 1, k));
        b.fc[2][b.Face(2, i, j, k)] = Scale(s, 0.25);
      }
  return "";
}

// procBlock::AssignGhostCellsGeom (src/procBlock.cpp:2160-2268): ghost
// volumes and face vectors are copies of the mirrored interior cell; ghost
// centroids and face centres are shifted by the interior face-to-face
// (first layer) or centre-to-centre (deeper layers) distance.  Face data in
// the two in-surface directions covers one extra face, whose shift repeats the
// last one (GrowInI/J/K, include/multiArray3d.hpp:929-990).
void GhostGeometry(HostBlock &b) {
  for (int layer = 1; layer <= b.g; ++layer) {
    for (const auto &sf : b.surf) {
      if (sf.type == "interblock") continue;
      const int st = sf.SurfType();
      const int dir = (st - 1) / 2;
      const bool upper = (st % 2 == 0);
      const int d1 = (dir + 1) % 3, d2 = (dir + 2) % 3;
      const int r1s = sf.d[2 * d1], n1 = sf.d[2 * d1 + 1] - r1s;
      const int r2s = sf.d[2 * d2], n2 = sf.d[2 * d2 + 1] - r2s;
      const int r3 = sf.d[2 * dir];
      const int n3 = b.N(dir);
      int gC, iC, pC, piC, iF, piF;
      if (upper) {
        gC = r3 + layer - 1;
        iC = std::max(r3 - layer, 0);
        pC = gC - 1;
        piC = iC + 1;
        iF = std::max(r3 - layer, 0);
        piF = iF + 1;
      } else {
        gC = r3 - layer;
        iC = std::min(r3 + layer - 1, n3 - 1);
        pC = gC + 1;
        piC = iC - 1;
        iF = std::min(r3 + layer, n3);
        piF = iF - 1;
      }
      const int fu = upper ? 1 : 0;  // face index offset of the normal face
      std::vector<Vec3> f2f((size_t)n1 * n2), c2c((size_t)n1 * n2);
      for (int q2 = 0; q2 < n2; ++q2)
        for (int q1 = 0; q1 < n1; ++q1) {
          int i, j, k, i2, j2, k2;
          Map3(dir, piF, r1s + q1, r2s + q2, &i, &j, &k);
          Map3(dir, iF, r1s + q1, r2s + q2, &i2, &j2, &k2);
          const size_t q = (size_t)q1 + (size_t)q2 * n1;
          f2f[q] = Sub(b.fc[dir][b.Face(dir, i, j, k)], b.fc[dir][b.Face(dir, i2, j2, k2)]);
          if (layer > 1) {
            Map3(dir, piC, r1s + q1, r2s + q2, &i, &j, &k);
            Map3(dir, iC, r1s + q1, r2s + q2, &i2, &j2, &k2);
            c2c[q] = Sub(b.ctr[b.Cell(i, j, k)], b.ctr[b.Cell(i2, j2, k2)]);
          } else {
            c2c[q] = f2f[q];
          }
        }
      for (int q2 = 0; q2 < n2; ++q2)
        for (int q1 = 0; q1 < n1; ++q1) {
          const size_t q = (size_t)q1 + (size_t)q2 * n1;
          int gi, gj, gk, ii, ij, ik, pi, pj, pk;
          Map3(dir, gC, r1s + q1, r2s + q2, &gi, &gj, &gk);
          Map3(dir, iC, r1s + q1, r2s + q2, &ii, &ij, &ik);
          Map3(dir, pC, r1s + q1, r2s + q2, &pi, &pj, &pk);
          b.vol[b.Cell(gi, gj, gk)] = b.vol[b.Cell(ii, ij, ik)];
          int fi, fj, fk, si, sj, sk, ti, tj, tk;
          Map3(dir, gC + fu, r1s + q1, r2s + q2, &fi, &fj, &fk);
          Map3(dir, iC + fu, r1s + q1, r2s + q2, &si, &sj, &sk);
          Map3(dir, pC + fu, r1s + q1, r2s + q2, &ti, &tj, &tk);
          b.fa[dir][b.Face(dir, fi, fj, fk)] = b.fa[dir][b.Face(dir, si, sj, sk)];
          b.fc[dir][b.Face(dir, fi, fj, fk)] = Add(f2f[q], b.fc[dir][b.Face(dir, ti, tj, tk)]);
          b.ctr[b.Cell(gi, gj, gk)] = Add(b.ctr[b.Cell(pi, pj, pk)], c2c[q]);
        }
      for (int t = 1; t <= 2; ++t) {
        const int fd = (dir + t) % 3;
        const int m1 = n1 + (t == 1), m2 = n2 + (t == 2);
        for (int q2 = 0; q2 < m2; ++q2)
          for (int q1 = 0; q1 < m1; ++q1) {
            int gi, gj, gk, ii, ij, ik, pi, pj, pk;
            Map3(dir, gC, r1s + q1, r2s + q2, &gi, &gj, &gk);
            Map3(dir, iC, r1s + q1, r2s + q2, &ii, &ij, &ik);
            Map3(dir, pC, r1s + q1, r2s + q2, &pi, &pj, &pk);
            const size_t q = (size_t)std::min(q1, n1 - 1) + (size_t)std::min(q2, n2 - 1) * n1;
            b.fa[fd][b.Face(fd, gi, gj, gk)] = b.fa[fd][b.Face(fd, ii, ij, ik)];
            b.fc[fd][b.Face(fd, gi, gj, gk)] = Add(c2c[q], b.fc[fd][b.Face(fd, pi, pj, pk)]);
          }
      }
    }
  }
}

// procBlock::CalcCellWidths (src/procBlock.cpp:6397-6411)
void CellWidths(HostBlock &b) {
  const int g = b.g;
  for (int k = -g; k < b.nk + g; ++k)
    for (int j = -g; j < b.nj + g; ++j)
      for (int i = -g; i < b.ni + g; ++i) {
        const long c = b.Cell(i, j, k);
        b.cw[0][c] = Mag(Sub(b.fc[0][b.Face(0, i, j, k)], b.fc[0][b.Face(0, i + 1, j, k)]));
        b.cw[1][c] = Mag(Sub(b.fc[1][b.Face(1, i, j, k)], b.fc[1][b.Face(1, i, j + 1, k)]));
        b.cw[2][c] = Mag(Sub(b.fc[2][b.Face(2, i, j, k)], b.fc[2][b.Face(2, i, j, k + 1)]));
      }
}

// exact nearest neighbour over the IC point cloud (the reference builds a
// k-d tree, src/kdtree.cpp:29-225; any exact search returns the same point)
class CloudSearch {
 public:
  explicit CloudSearch(const std::vector<Vec3> &pts) : pts_(pts) {
    double hi[3] = {-1e300, -1e300, -1e300};
    lo_[0] = lo_[1] = lo_[2] = 1e300;
    for (const auto &p : pts_) {
      const double c[3] = {p.x, p.y, p.z};
      for (int d = 0; d < 3; ++d) {
        lo_[d] = std::min(lo_[d], c[d]);
        hi[d] = std::max(hi[d], c[d]);
      }
    }
    const int nb = std::min(256, (int)std::cbrt((double)pts_.size()) + 1);
    for (int d = 0; d < 3; ++d) {
      nb_[d] = nb;
      const double ext = hi[d] - lo_[d];
      h_[d] = ext > 0 ? ext / nb * 1.000001 : 1.0;
    }
    head_.assign((size_t)nb_[0] * nb_[1] * nb_[2], -1);
    next_.assign(pts_.size(), -1);
    for (size_t q = 0; q < pts_.size(); ++q) {
      const long key = Key(pts_[q]);
      next_[q] = head_[key];
      head_[key] = (int)q;
    }
  }
  int Nearest(Vec3 x) const {
    int b[3];
    Bucket(x, b);
    int best = -1;
    double bd = 1e300;
    const double hmin = std::min(h_[0], std::min(h_[1], h_[2]));
    for (int r = 0; r <= nb_[0] + nb_[1] + nb_[2]; ++r) {
      for (int a = b[0] - r; a <= b[0] + r; ++a)
        for (int c = b[1] - r; c <= b[1] + r; ++c)
          for (int e = b[2] - r; e <= b[2] + r; ++e) {
            if (a < 0 || c < 0 || e < 0 || a >= nb_[0] || c >= nb_[1] || e >= nb_[2]) continue;
            if (std::abs(a - b[0]) != r && std::abs(c - b[1]) != r && std::abs(e - b[2]) != r) continue;
            for (int q = head_[a + (long)c * nb_[0] + (long)e * nb_[0] * nb_[1]]; q >= 0; q = next_[q]) {
              const Vec3 d = Sub(pts_[q], x);
              const double dd = Dot(d, d);
              if (dd < bd) { bd = dd; best = q; }
            }
          }
      if (best >= 0 && (r * hmin) * (r * hmin) >= bd) break;
    }
    return best;
  }

 private:
  void Bucket(Vec3 x, int *b) const {
    const double c[3] = {x.x, x.y, x.z};
    for (int d = 0; d < 3; ++d) b[d] = std::max(0, std::min(nb_[d] - 1, (int)std::floor((c[d] - lo_[d]) / h_[d])));
  }
  long Key(Vec3 x) const {
    int b[3];
    Bucket(x, b);
    return b[0] + (long)b[1] * nb_[0] + (long)b[2] * nb_[0] * nb_[1];
  }
  const std::vector<Vec3> &pts_;
  double lo_[3], h_[3];
  int nb_[3];
  std::vector<int> head_, next_;
};

// ---------------------------------------------------------------------------
// interblock connections

// partner surface from the tag (src/boundaryConditions.cpp:2468-2495); the
// partner block is tag - 1000 * surface
int TagPartnerSurface(int tag) {
  for (int s = 1; s <= 6; ++s)
    if (tag < 1000 * (s + 1)) return s;
  return -1;
}

// vector3d::CompareWithTol (include/vector3d.hpp:319-327)
bool SamePoint(const Vec3 &a, const Vec3 &b) {
  return std::fabs(a.x - b.x) < 1.0e-10 && std::fabs(a.y - b.y) < 1.0e-10 && std::fabs(a.z - b.z) < 1.0e-10;
}

Vec3 NodeAt(const std::vector<Vec3> &X, const HostBlock &b, int dir, int r3, int a1, int a2) {
  int i, j, k;
  Map3(dir, r3, a1, a2, &i, &j, &k);
  return X[(size_t)i + (size_t)(b.ni + 1) * ((size_t)j + (size_t)(b.nj + 1) * k)];
}

// the patch axes' correspondence, from three corners (TestPatchMatch,
// src/boundaryConditions.cpp:729-831)
bool OrientPatch(Connection &cn, const HostBlock &A, const std::vector<Vec3> &XA, const HostBlock &B,
                 const std::vector<Vec3> &XB, const std::array<double, 3> &shift = {0, 0, 0}) {
  const int corners[3][2] = {{0, 0}, {cn.n1, 0}, {0, cn.n2}};
  for (int sw = 0; sw < 2; ++sw)
    for (int r1 = 0; r1 < 2; ++r1)
      for (int r2 = 0; r2 < 2; ++r2) {
        cn.swap = sw;
        cn.rev1 = r1;
        cn.rev2 = r2;
        if (sw ? (cn.n1 != cn.pn2 || cn.n2 != cn.pn1) : (cn.n1 != cn.pn1 || cn.n2 != cn.pn2)) continue;
        bool ok = true;
        for (int q = 0; q < 3 && ok; ++q) {
          int b1, b2;
          cn.Map(corners[q][0], corners[q][1], false, &b1, &b2);
          Vec3 pa = NodeAt(XA, A, cn.dir, cn.r3, cn.r1s + corners[q][0], cn.r2s + corners[q][1]);
          pa.x += shift[0];
          pa.y += shift[1];
          pa.z += shift[2];
          ok = SamePoint(pa, NodeAt(XB, B, cn.pdir, cn.pr3, cn.pr1s + b1, cn.pr2s + b2));
        }
        if (ok) return true;
      }
  return false;
}

// SwapGeomSlice / PutGeomSlice (src/utility.cpp:213-256,
// src/procBlock.cpp:3167-4011) over the patch: a ghost cell takes its
// partner cell's volume and centroid; its faces normal to the patch (the
// boundary face included) take the partner's faces at the same distance from
// the cut, the area direction reversed when both patches are lower or both
// upper surfaces; its faces along the patch take the partner's, lower and
// upper exchanged and the area reversed where the partner's axis runs
// backwards.  Every swap reads the geometry as it was before any of them.
void SwapGeometry(HostCase &c) {
  const std::vector<HostBlock> before = c.blocks;
  for (const auto &cn : c.conns) {
    if (cn.periodic) continue;  // periodic ghost cells keep the mirrored geometry
    HostBlock &A = c.blocks[cn.block];
    const HostBlock &B = before[cn.pb];
    const bool flip3 = (cn.s + cn.ps) % 2 == 0;
    const bool lowA = cn.s % 2 == 1, lowB = cn.ps % 2 == 1;
    const int nA = A.N(cn.dir), nB = B.N(cn.pdir);
    const int dA[2] = {(cn.dir + 1) % 3, (cn.dir + 2) % 3};
    const int dB[2] = {(cn.pdir + 1) % 3, (cn.pdir + 2) % 3};
    auto neg = [](std::array<double, 4> v) {
      v[0] = -1.0 * v[0];
      v[1] = -1.0 * v[1];
      v[2] = -1.0 * v[2];
      return v;
    };
    for (int p2 = 0; p2 < cn.n2; ++p2)
      for (int p1 = 0; p1 < cn.n1; ++p1) {
        int q1, q2;
        cn.Map(p1, p2, true, &q1, &q2);
        for (int t = 0; t <= A.g; ++t) {
          int i, j, k, bi, bj, bk;
          Map3(cn.dir, lowA ? -t : nA + t, cn.r1s + p1, cn.r2s + p2, &i, &j, &k);
          Map3(cn.pdir, lowB ? t : nB - t, cn.pr1s + q1, cn.pr2s + q2, &bi, &bj, &bk);
          const auto &v = B.fa[cn.pdir][B.Face(cn.pdir, bi, bj, bk)];
          A.fa[cn.dir][A.Face(cn.dir, i, j, k)] = flip3 ? neg(v) : v;
          A.fc[cn.dir][A.Face(cn.dir, i, j, k)] = B.fc[cn.pdir][B.Face(cn.pdir, bi, bj, bk)];
        }
        for (int l = 1; l <= A.g; ++l) {
          int gi, gj, gk, si, sj, sk;
          cn.Ghost(A, l, p1, p2, &gi, &gj, &gk);
          cn.Source(B, l, p1, p2, &si, &sj, &sk);
          A.vol[A.Cell(gi, gj, gk)] = B.vol[B.Cell(si, sj, sk)];
          A.ctr[A.Cell(gi, gj, gk)] = B.ctr[B.Cell(si, sj, sk)];
          for (int w = 0; w < 2; ++w) {
            // this block's axis dA[w] runs along the partner's axis dB[w]
            // (or dB[1-w] when the patch axes are swapped)
            const bool useB1 = (w == 0) != cn.swap;
            const int da = dA[w], db = dB[useB1 ? 0 : 1];
            const bool rev = useB1 ? cn.rev1 : cn.rev2;
            for (int up = 0; up < 2; ++up) {
              const int bup = rev ? 1 - up : up;
              const long fa = A.Face(da, gi + (da == 0 ? up : 0), gj + (da == 1 ? up : 0), gk + (da == 2 ? up : 0));
              const long fb = B.Face(db, si + (db == 0 ? bup : 0), sj + (db == 1 ? bup : 0), sk + (db == 2 ? bup : 0));
              A.fa[da][fa] = rev ? neg(B.fa[db][fb]) : B.fa[db][fb];
              A.fc[da][fa] = B.fc[db][fb];
            }
          }
        }
      }
  }
}

// pair every interblock surface with its partner patch
// (GetBlockInterConnBCs, src/boundaryConditions.cpp:603-642)
std::string ConnectBlocks(HostCase &c, const std::vector<std::vector<Vec3>> &nodes) {
  for (int a = 0; a < (int)c.blocks.size(); ++a) {
    const HostBlock &A = c.blocks[a];
    for (const auto &sf : A.surf) {
      if (sf.type != "interblock" && sf.type != "periodic") continue;
      Connection cn;
      cn.block = a;
      cn.s = sf.SurfType();
      cn.dir = (cn.s - 1) / 2;
      const int d1 = (cn.dir + 1) % 3, d2 = (cn.dir + 2) % 3;
      cn.r3 = sf.d[2 * cn.dir];
      cn.r1s = sf.d[2 * d1];
      cn.n1 = sf.d[2 * d1 + 1] - cn.r1s;
      cn.r2s = sf.d[2 * d2];
      cn.n2 = sf.d[2 * d2 + 1] - cn.r2s;
      if (sf.type == "periodic") {
        // periodic pairs match by geometry once the start surface is moved by
        // the translation (GetConnectionBCs, src/boundaryConditions.cpp:543-593;
        // patch::Transform, :2223-2248)
        const BcState *ps = nullptr;
        for (const auto &st : c.bcs)
          if (st.kind == "periodic" && (st.startTag == sf.d[6] || st.endTag == sf.d[6])) {
            ps = &st;
            break;
          }
        const std::string where = "block " + std::to_string(a) + ": periodic tag " + std::to_string(sf.d[6]);
        if (!ps) return where + " has no periodic boundary state";
        const int other = ps->startTag == sf.d[6] ? ps->endTag : ps->startTag;
        const double sg = ps->startTag == sf.d[6] ? 1.0 : -1.0;
        const std::array<double, 3> shift = {sg * ps->translation[0], sg * ps->translation[1],
                                             sg * ps->translation[2]};
        cn.periodic = true;
        bool found = false;
        for (int pbq = 0; pbq < (int)c.blocks.size() && !found; ++pbq) {
          const HostBlock &B = c.blocks[pbq];
          for (const auto &pf : B.surf) {
            if (pf.type != "periodic" || pf.d[6] != other || &pf == &sf) continue;
            cn.pb = pbq;
            cn.ps = pf.SurfType();
            cn.pdir = (cn.ps - 1) / 2;
            const int e1 = (cn.pdir + 1) % 3, e2 = (cn.pdir + 2) % 3;
            cn.pr3 = pf.d[2 * cn.pdir];
            cn.pr1s = pf.d[2 * e1];
            cn.pn1 = pf.d[2 * e1 + 1] - cn.pr1s;
            cn.pr2s = pf.d[2 * e2];
            cn.pn2 = pf.d[2 * e2 + 1] - cn.pr2s;
            if (OrientPatch(cn, A, nodes[a], B, nodes[pbq], shift)) {
              found = true;
              break;
            }
          }
        }
        if (!found) return where + " has no matching periodic patch";
        if (c.blocks[cn.pb].N(cn.pdir) < A.g) return where + ": partner block is thinner than the ghost layers";
        c.conns.push_back(cn);
        continue;
      }
      cn.ps = TagPartnerSurface(sf.d[6]);
      cn.pb = sf.d[6] - 1000 * cn.ps;
      const std::string where = "block " + std::to_string(a) + ": interblock tag " + std::to_string(sf.d[6]);
      if (cn.ps < 0 || cn.pb < 0 || cn.pb >= (int)c.blocks.size()) return where + " names no block";
      cn.pdir = (cn.ps - 1) / 2;
      const HostBlock &B = c.blocks[cn.pb];
      bool found = false;
      for (const auto &pf : B.surf) {
        if (pf.type != "interblock" || pf.SurfType() != cn.ps) continue;
        const int pps = TagPartnerSurface(pf.d[6]);
        if (pps != cn.s || pf.d[6] - 1000 * pps != a) continue;
        if (cn.pb == a && &pf == &sf) continue;
        const int e1 = (cn.pdir + 1) % 3, e2 = (cn.pdir + 2) % 3;
        cn.pr3 = pf.d[2 * cn.pdir];
        cn.pr1s = pf.d[2 * e1];
        cn.pn1 = pf.d[2 * e1 + 1] - cn.pr1s;
        cn.pr2s = pf.d[2 * e2];
        cn.pn2 = pf.d[2 * e2 + 1] - cn.pr2s;
        if (OrientPatch(cn, A, nodes[a], B, nodes[cn.pb])) {
          found = true;
          break;
        }
      }
      if (!found) return where + " has no matching interblock patch in the partner block";
      if (B.N(cn.pdir) < A.g) return where + ": partner block is thinner than the ghost layers";
      c.conns.push_back(cn);
    }
  }
  SwapGeometry(c);
  return "";
}

}  // namespace

void Connection::Map(int p1, int p2, bool cells, int *q1, int *q2) const {
  const int c1 = swap ? p2 : p1, c2 = swap ? p1 : p2;
  const int m1 = cells ? pn1 - 1 : pn1, m2 = cells ? pn2 - 1 : pn2;
  *q1 = rev1 ? m1 - c1 : c1;
  *q2 = rev2 ? m2 - c2 : c2;
}

void Connection::Ghost(const HostBlock &a, int l, int p1, int p2, int *i, int *j, int *k) const {
  Map3(dir, s % 2 == 1 ? -l : a.N(dir) + l - 1, r1s + p1, r2s + p2, i, j, k);
}

void Connection::Source(const HostBlock &b, int l, int p1, int p2, int *i, int *j, int *k) const {
  int q1, q2;
  Map(p1, p2, true, &q1, &q2);
  Map3(pdir, ps % 2 == 1 ? l - 1 : b.N(pdir) - l, pr1s + q1, pr2s + q2, i, j, k);
}

int Surface::SurfType() const {  // src/boundaryConditions.cpp:2424-2450
  if (d[0] == d[1]) return d[1] == 0 ? 1 : 2;
  if (d[2] == d[3]) return d[3] == 0 ? 3 : 4;
  if (d[4] == d[5]) return d[5] == 0 ? 5 : 6;
  return 0;
}

double HostCase::Cfl(int nn) const {  // input::CalcCFL, src/input.cpp:647-649
  const double c = cflStart + nn * cflStep;
  return (cflMax < c) ? cflMax : c;
}

void HyperplaneOrder(int ni, int nj, int nk, std::vector<int> *cells,
                     std::vector<int> *offsets) {
  const int np = ni + nj + nk - 2;
  cells->clear();
  cells->reserve((size_t)ni * nj * nk);
  offsets->assign(1, 0);
  // count per plane, then fill in (k, j, i) order within each plane
  std::vector<int> count(np, 0);
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) ++count[i + j + k];
  for (int p = 0; p < np; ++p) offsets->push_back(offsets->back() + count[p]);
  cells->resize((size_t)ni * nj * nk);
  std::vector<int> fill(offsets->begin(), offsets->end() - 1);
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < nj; ++j)
      for (int i = 0; i < ni; ++i) (*cells)[fill[i + j + k]++] = i + j * ni + k * ni * nj;
}

std::string LoadCase(const std::string &inp_path, HostCase *out) {
  HostCase &c = *out;
  c = HostCase();
  std::ifstream f(inp_path);
  if (!f) return "cannot open input file " + inp_path;
  const auto sl = inp_path.rfind('/');
  c.dir = (sl == std::string::npos) ? "." : inp_path.substr(0, sl);
  {
    // input::SimNameRoot (src/input.cpp:872-874): the name up to its first '.'
    const std::string base = sl == std::string::npos ? inp_path : inp_path.substr(sl + 1);
    c.simName = base.substr(0, base.find('.'));
  }
  std::string fluid = "air";
  struct Ic { int tag = -1; bool file = false; std::string fname; double rho = 0, p = 0; std::array<double, 3> vel{0, 0, 0}; };
  std::vector<Ic> ics;
  std::vector<std::vector<Surface>> surfs;
  int bcBlocks = 0, curBlk = 0, left = -1;
  bool readingBC = false;
  std::string line;
  // input::ReadInput key handling and defaults: src/input.cpp:50-103, 167-643
  while (std::getline(f, line)) {
    line = Trim(line);
    if (line.empty()) continue;
    if (readingBC) {
      std::istringstream ss(line);
      if (left < 0) {
        int a, b, d;
        if (!(ss >> a >> b >> d)) return "bad boundary-condition count line: " + line;
        left = a + b + d;
        surfs.emplace_back();
      } else {
        Surface s;
        if (!(ss >> s.type >> s.d[0] >> s.d[1] >> s.d[2] >> s.d[3] >> s.d[4] >> s.d[5] >> s.d[6]))
          return "bad boundary-condition surface line: " + line;
        surfs.back().push_back(s);
        --left;
      }
      if (left == 0) {
        left = -1;
        if (++curBlk == bcBlocks) readingBC = false;
      }
      continue;
    }
    const auto colon = line.find(':');
    if (colon == std::string::npos) continue;
    const std::string key = Trim(line.substr(0, colon));
    std::string val = Trim(line.substr(colon + 1));
    if (!val.empty() && val[0] == '<' && val.find('>') == std::string::npos) {
      std::string more;
      while (std::getline(f, more)) {
        more = Trim(more);
        val += " " + more;
        if (more.find('>') != std::string::npos) break;
      }
    }
    try {
      if (key == "gridName") c.grid = val;
      else if (key == "outputVariables") {
        // ReadStringList: <a, b, c>
        c.outputVars.clear();
        std::string body = val;
        if (!body.empty() && body.front() == '<') body = body.substr(1);
        const auto close = body.find('>');
        if (close != std::string::npos) body = body.substr(0, close);
        std::istringstream ls(body);
        std::string item;
        while (std::getline(ls, item, ',')) {
          item = Trim(item);
          if (!item.empty()) c.outputVars.insert(item);
        }
      }
      else if (key == "timeStep") c.dtIn = std::stod(val);
      else if (key == "iterations") c.iterations = std::stoi(val);
      else if (key == "referenceDensity") c.rRef = std::stod(val);
      else if (key == "referenceTemperature") c.tRef = std::stod(val);
      else if (key == "referenceLength") c.lRef = std::stod(val);
      else if (key == "timeIntegration") {
        if (val == "explicitEuler") c.timeint = TimeInt::ExplicitEuler;
        else if (val == "implicitEuler") { c.timeint = TimeInt::ImplicitEuler; c.theta = 1.0; c.zeta = 0.0; }
        else if (val == "crankNicholson") { c.timeint = TimeInt::CrankNicholson; c.theta = 0.5; c.zeta = 0.0; }
        else if (val == "bdf2") { c.timeint = TimeInt::BDF2; c.theta = 1.0; c.zeta = 0.5; }
        else return "unsupported timeIntegration " + val;
      } else if (key == "faceReconstruction") {
        c.recon = Recon::MUSCL;
        if (val == "upwind") c.kappa = -1.0;
        else if (val == "fromm") c.kappa = 0.0;
        else if (val == "quick") c.kappa = 0.5;
        else if (val == "central") c.kappa = 1.0;
        else if (val == "thirdOrder") c.kappa = 1.0 / 3.0;
        else if (val == "constant") c.recon = Recon::Constant;
        else if (val == "weno") c.recon = Recon::WENO;
        else if (val == "wenoZ") c.recon = Recon::WENOZ;
        else return "unsupported faceReconstruction " + val;
      } else if (key == "limiter") {
        if (val == "none") c.limiter = Limiter::None;
        else if (val == "vanAlbada") c.limiter = Limiter::VanAlbada;
        else if (val == "minmod") c.limiter = Limiter::Minmod;
        else return "unsupported limiter " + val;
      } else if (key == "equationSet") {
        if (val == "euler") c.viscous = false;
        else if (val == "navierStokes") c.viscous = true;
        else return "unsupported equationSet " + val;
      } else if (key == "viscousFaceReconstruction") {
        if (val == "central") c.viscRecon4 = false;
        else if (val == "centralFourth") c.viscRecon4 = true;
        else return "unsupported viscousFaceReconstruction " + val;
      } else if (key == "turbulenceModel") {
        if (val != "none") return "unsupported turbulenceModel " + val;
      } else if (key == "matrixSolver") {
        // src/input.cpp:713-720, 845-856
        if (val == "lusgs") c.solver = Solver::LUSGS, c.block = false;
        else if (val == "dplur") c.solver = Solver::DPLUR, c.block = false;
        else if (val == "blusgs") c.solver = Solver::LUSGS, c.block = true;
        else if (val == "bdplur") c.solver = Solver::DPLUR, c.block = true;
        else return "unsupported matrixSolver " + val;
      } else if (key == "matrixSweeps") c.sweeps = std::stoi(val);
      else if (key == "matrixRelaxation") c.relax = std::stod(val);
      else if (key == "nonlinearIterations") c.nonlinear = std::stoi(val);
      else if (key == "cflMax") c.cflMax = std::stod(val);
      else if (key == "cflStep") c.cflStep = std::stod(val);
      else if (key == "cflStart") c.cflStart = std::stod(val);
      else if (key == "dualTimeCFL") c.dualCFL = std::stod(val);
      else if (key == "inviscidFlux") {
        if (val == "roe") c.flux = Flux::Roe;
        else if (val == "ausm") c.flux = Flux::AUSM;
        else return "unsupported inviscidFlux " + val;
      } else if (key == "inviscidFluxJacobian") {
        // src/input.cpp:82, 372-373; OffDiagonal (src/fluxJacobian.cpp:196-238)
        if (val == "rusanov") c.roeJac = false;
        else if (val == "approximateRoe") c.roeJac = true;
        else return "unsupported inviscidFluxJacobian " + val;
      } else if (key == "multigridLevels") {
        if (std::stoi(val) != 1) return "multigrid is not supported";
      } else if (key == "fluids") {
        const auto e = ListEntries(val);
        if (e.size() != 1) return "only single-species fluids are supported";
        const auto p = EntryParams(e[0]);
        if (p.count("name")) fluid = p.at("name");
      } else if (key == "initialConditions") {
        for (const auto &e : ListEntries(val)) {
          const auto p = EntryParams(e);
          Ic ic;
          if (p.count("tag")) ic.tag = std::stoi(p.at("tag"));
          if (p.count("file")) { ic.file = true; ic.fname = p.at("file"); }
          if (p.count("pressure")) ic.p = std::stod(p.at("pressure"));
          if (p.count("density")) ic.rho = std::stod(p.at("density"));
          if (p.count("velocity")) ic.vel = ParseVec(p.at("velocity"));
          ics.push_back(ic);
        }
      } else if (key == "boundaryStates") {
        for (const auto &e : ListEntries(val)) {
          const auto p = EntryParams(e);
          BcState s;
          s.kind = EntryName(e);
          if (p.count("tag")) s.tag = std::stoi(p.at("tag"));
          if (p.count("pressure")) s.p = std::stod(p.at("pressure"));
          if (p.count("density")) s.rho = std::stod(p.at("density"));
          if (p.count("velocity")) s.vel = ParseVec(p.at("velocity"));
          if (p.count("p0")) s.p0 = std::stod(p.at("p0"));
          if (p.count("t0")) s.t0 = std::stod(p.at("t0"));
          if (p.count("direction")) s.dir = ParseVec(p.at("direction"));
          if (p.count("nonreflecting") && p.at("nonreflecting") == "true")
            return "nonreflecting boundary states are not supported";
          if (p.count("temperature")) {
            s.isothermal = true;
            s.twall = std::stod(p.at("temperature"));
          }
          // viscousWall::IsConstantHeatFlux (include/inputStates.hpp:379-381)
          if (p.count("heatFlux")) {
            if (p.count("temperature")) return "viscousWall: give a temperature or a heat flux, not both";
            s.qwall = std::stod(p.at("heatFlux"));
            s.heatFlux = s.qwall != 0.0;
          }
          if (p.count("wallTreatment") && p.at("wallTreatment") != "lowRe")
            return "wall laws are not supported";
          if (p.count("startTag")) s.startTag = std::stoi(p.at("startTag"));
          if (p.count("endTag")) s.endTag = std::stoi(p.at("endTag"));
          if (p.count("translation")) s.translation = ParseVec(p.at("translation"));
          if (s.kind == "periodic" && (p.count("axis") || p.count("rotation")))
            return "rotational periodicity is not supported";
          c.bcs.push_back(s);
        }
      } else if (key == "boundaryConditions") {
        bcBlocks = std::stoi(val);
        readingBC = bcBlocks > 0;
        curBlk = 0;
        left = -1;
      }
    } catch (const std::exception &) {
      return "cannot parse value of " + key + ": " + val;
    }
  }
  if (c.rRef <= 0 || c.tRef <= 0) return "referenceDensity and referenceTemperature are required";
  if (c.timeint == TimeInt::ExplicitEuler) c.nonlinear = 1;  // src/input.cpp:928-934
  c.block = c.block && c.Implicit();  // input::IsBlockMatrix (src/input.cpp:713-720)
  // RoeOffDiagonal forms a flux change, not a matrix, so the block solvers
  // cannot use it; with the viscous terms the reference's call passes the
  // blending coefficient (0 in laminar flow) as the centre distance
  // (src/fluxJacobian.cpp:229-230 against :240-245)
  if (c.roeJac && c.block) return "approximateRoe with a block matrix solver is not supported";
  if (c.roeJac && c.viscous) return "approximateRoe with the viscous equations is not supported";
  double fn = 0, fm = 0;
  if (!FluidProps(c.dir, fluid, &fn, &fm, &c.suth)) return "unknown fluid " + fluid;
  c.fluidName = fluid;

  // reference speed of sound (src/input.cpp:616-621) and nondimensional gas
  // constant (src/fluid.cpp:89-103)
  {
    const double Ru = 8.3144598;
    const double Rdim = Ru / fm;
    const double gm = (fn + 1) / fn;
    double a = 0.0;
    a += 1.0 * gm * Rdim * c.tRef;
    c.aRef = std::sqrt(a);
    const double RuNd = Ru / (c.aRef * c.aRef * c.rRef / (c.tRef * std::pow(c.lRef, 3.0)));
    const double Mnd = fm / (c.rRef / std::pow(c.lRef, 3.0));
    c.gasR = RuNd / Mnd;
    c.gasN = fn;
    const double cp = 0.0 + 1.0 * (c.gasR * (fn + 1.0));
    const double cv = 0.0 + 1.0 * (c.gasR * fn);
    c.gamma = cp / cv;
  }
  // Sutherland transport and its nondimensional scaling
  // (src/transport.cpp:29-66, include/transport.hpp:40-44); Prandtl number
  // (include/thermodynamic.hpp:61-64); viscous CFL coefficient
  // (src/input.cpp:1110-1118)
  c.muRef = c.suth[0] * std::pow(c.tRef, 1.5) / (c.tRef + c.suth[1]);
  c.kNonDim = (c.aRef * c.aRef * c.muRef) / c.tRef;
  c.vscale = c.muRef / (c.rRef * c.aRef * c.lRef);
  c.prandtl = (4.0 * c.gamma) / (9.0 * c.gamma - 5.0);
  c.viscCoeff = c.kappa == 1.0 ? 4.0 : (c.kappa == -2.0 ? 2.0 : 1.0);
  // state nondimensionalisation (src/inputStates.cpp:464-473, 590-599,
  // 674-681); the stagnation direction stays as given (src/inputStates.cpp:594
  // discards the normalised copy)
  const double pScale = c.rRef * c.aRef * c.aRef;
  for (auto &ic : ics) {
    if (ic.file) continue;
    for (auto &v : ic.vel) v /= c.aRef;
    ic.rho /= c.rRef;
    ic.p /= pScale;
  }
  for (auto &s : c.bcs) {
    if (s.kind == "periodic") {  // src/inputStates.cpp:883-894
      for (auto &v : s.translation) v /= c.lRef;
      continue;
    }
    if (s.kind == "viscousWall") {  // src/inputStates.cpp:775-783
      for (auto &v : s.vel) v /= c.aRef;
      s.twall /= c.tRef;
      s.qwall /= std::pow(c.aRef / c.lRef, 3.0);  // src/inputStates.cpp:775-783
      continue;
    }
    if (s.kind == "stagnationInlet") {
      s.p0 /= pScale;
      s.t0 /= c.tRef;
    } else if (s.kind == "pressureOutlet") {
      s.p /= pScale;
    } else {
      for (auto &v : s.vel) v /= c.aRef;
      s.rho /= c.rRef;
      s.p /= pScale;
    }
  }

  // grid: ReadP3dGrid (src/plot3d.cpp:363-445)
  const std::string gpath = c.dir + "/" + c.grid + ".xyz";
  std::ifstream g(gpath, std::ios::binary);
  if (!g) return "cannot open grid " + gpath;
  int nb = 0;
  g.read(reinterpret_cast<char *>(&nb), sizeof nb);
  if (!g || nb <= 0 || nb > 100000) return "bad Plot3D header in " + gpath;
  if (nb != bcBlocks || (int)surfs.size() != nb) return "grid block count does not match the boundaryConditions table";
  std::vector<std::array<int, 3>> dims(nb);
  for (auto &d : dims) g.read(reinterpret_cast<char *>(d.data()), 3 * sizeof(int));
  if (!g) return "truncated Plot3D header";
  const int ng = c.Ghosts();
  c.blocks.resize(nb);
  std::vector<std::vector<Vec3>> nodes(nb);
  c.totalCells = 0;
  for (int bq = 0; bq < nb; ++bq) {
    HostBlock &b = c.blocks[bq];
    b.ni = dims[bq][0] - 1;
    b.nj = dims[bq][1] - 1;
    b.nk = dims[bq][2] - 1;
    if (b.ni < 1 || b.nj < 1 || b.nk < 1) return "degenerate Plot3D block";
    b.g = ng;
    b.NI = b.ni + 2 * ng;
    b.NJ = b.nj + 2 * ng;
    b.NK = b.nk + 2 * ng;
    c.totalCells += (double)b.NumCells();
    const size_t np = (size_t)dims[bq][0] * dims[bq][1] * dims[bq][2];
    std::vector<Vec3> X(np);
    std::vector<double> tmp(np);
    for (int comp = 0; comp < 3; ++comp) {
      g.read(reinterpret_cast<char *>(tmp.data()), (std::streamsize)(np * sizeof(double)));
      if (!g) return "truncated Plot3D coordinates";
      for (size_t q = 0; q < np; ++q) {
        const double v = tmp[q] / c.lRef;
        (comp == 0 ? X[q].x : comp == 1 ? X[q].y : X[q].z) = v;
      }
    }
    b.vol.assign(b.NumPadded(), 0.0);
    b.ctr.assign(b.NumPadded(), Vec3{});
    for (int d = 0; d < 3; ++d) {
      b.fa[d].assign(b.NumFaces(d), {0, 0, 0, 0});
      b.fc[d].assign(b.NumFaces(d), Vec3{});
      b.cw[d].assign(b.NumPadded(), 0.0);
    }
    b.surf = surfs[bq];
    for (const auto &s : b.surf) {

      const int st = s.SurfType();
      if (st == 0) return "boundary surface is neither an i, j nor k surface";
      // the surface must lie on the block boundary and its patch inside it
      const int dir = (st - 1) / 2;
      const int r3 = s.d[2 * dir];
      if (r3 != (st % 2 == 1 ? 0 : b.N(dir))) return "boundary surface does not lie on the block boundary";
      for (int t = 1; t <= 2; ++t) {
        const int dd = (dir + t) % 3;
        if (s.d[2 * dd] < 0 || s.d[2 * dd + 1] > b.N(dd) || s.d[2 * dd] >= s.d[2 * dd + 1])
          return "boundary surface range outside the block";
      }
    }
    // boundarySurface ordering (src/boundaryConditions.cpp:91-105)
    std::sort(b.surf.begin(), b.surf.end(), [](const Surface &x, const Surface &y) {
      if (x.SurfType() != y.SurfType()) return x.SurfType() < y.SurfType();
      return x.d < y.d;
    });
    auto err = BuildGeometry(b, X);
    if (!err.empty()) return err;
    GhostGeometry(b);
    nodes[bq] = std::move(X);
  }
  // geometry swap across connections (src/gridLevel.cpp:67-74), then the
  // cell widths, ghost cells included (src/main.cpp:169)
  {
    auto err = ConnectBlocks(c, nodes);
    if (!err.empty()) return err;
  }
  for (auto &b : c.blocks) CellWidths(b);

  // initial state (procBlock::InitializeStates, src/procBlock.cpp:280-355)
  for (int bq = 0; bq < nb; ++bq) {
    HostBlock &b = c.blocks[bq];
    const Ic *use = nullptr;
    for (const auto &ic : ics) {
      if (ic.tag == -1 && !use) use = &ic;
      if (ic.tag == bq) { use = &ic; break; }
    }
    if (!use) return "no initial condition for block " + std::to_string(bq);
    b.W0.assign((size_t)b.NumCells() * kNeq, 0.0);
    if (!use->file) {
      for (long p = 0; p < b.NumCells(); ++p) {
        double *W = &b.W0[(size_t)p * kNeq];
        W[0] = 1.0 * use->rho;
        W[1] = use->vel[0];
        W[2] = use->vel[1];
        W[3] = use->vel[2];
        W[4] = use->p;
      }
    } else {
      // point cloud (CalcTreeFromCloud, src/utility.cpp:513-605)
      std::ifstream cf(c.dir + "/" + use->fname);
      if (!cf) return "cannot open initial-condition file " + use->fname;
      int npts = 0;
      std::string species;
      cf >> npts >> species;
      if (!cf || npts <= 0) return "bad initial-condition file header";
      std::vector<Vec3> pts(npts);
      std::vector<double> st((size_t)npts * kNeq);
      for (int q = 0; q < npts; ++q) {
        double x, y, z, r, u, v, w, p, tk, om, mf;
        if (!(cf >> x >> y >> z >> r >> u >> v >> w >> p >> tk >> om >> mf))
          return "bad initial-condition line " + std::to_string(q);
        pts[q] = {x / c.lRef, y / c.lRef, z / c.lRef};
        const double rho = r / c.rRef;
        st[(size_t)q * kNeq + 0] = rho * mf;
        st[(size_t)q * kNeq + 1] = u / c.aRef;
        st[(size_t)q * kNeq + 2] = v / c.aRef;
        st[(size_t)q * kNeq + 3] = w / c.aRef;
        st[(size_t)q * kNeq + 4] = p / (c.rRef * c.aRef * c.aRef);
      }
      CloudSearch search(pts);
      for (int k = 0; k < b.nk; ++k)
        for (int j = 0; j < b.nj; ++j)
          for (int i = 0; i < b.ni; ++i) {
            const int id = search.Nearest(b.ctr[b.Cell(i, j, k)]);
            const long p = i + (long)j * b.ni + (long)k * b.ni * b.nj;
            std::copy(&st[(size_t)id * kNeq], &st[(size_t)id * kNeq] + kNeq, &b.W0[(size_t)p * kNeq]);
          }
    }
  }
  // every boundary surface must resolve to a supported type and state
  for (const auto &b : c.blocks)
    for (const auto &s : b.surf) {
      const auto &t = s.type;
      if (t != "slipWall" && t != "viscousWall" && t != "characteristic" && t != "supersonicInflow" &&
          t != "supersonicOutflow" && t != "stagnationInlet" && t != "pressureOutlet" && t != "interblock" &&
          t != "periodic")
        return "unsupported boundary condition " + t;
      if (t == "characteristic" || t == "supersonicInflow" || t == "stagnationInlet" || t == "pressureOutlet") {
        bool found = false;
        for (const auto &bs : c.bcs) found = found || bs.tag == s.d[6];
        if (!found) return "no boundaryStates entry for tag " + std::to_string(s.d[6]);
      }
    }
  return "";
}

}  // namespace ath
