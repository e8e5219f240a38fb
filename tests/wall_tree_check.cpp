// Check of ath::NearestPoints (the wall-distance k-d tree): for many point
// sets and queries the tree's smallest squared distance equals the full
// scan's, bit for bit.  Built and run by tests/test_wall_tree.py.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "wall_tree.hpp"

using ath::NearestPoints;
using ath::Vec3;

static double Brute(const std::vector<Vec3> &pts, const Vec3 &q) {
  double best = std::numeric_limits<double>::max();
  for (const auto &p : pts) best = std::min(best, NearestPoints::Dist2(q, p));
  return best;
}

int main() {
  std::mt19937_64 rng(20261203);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  long checked = 0, bad = 0;
  auto check = [&](const std::vector<Vec3> &pts, const char *name, int nq) {
    const NearestPoints tree(pts);
    for (int q = 0; q < nq; ++q) {
      Vec3 v;
      if (q % 4 == 0 && !pts.empty()) {
        v = pts[rng() % pts.size()];  // a point of the set: distance 0
      } else if (q % 4 == 1) {
        v = {10 * u(rng), 10 * u(rng), 10 * u(rng)};  // mostly far away
      } else {
        v = {u(rng), u(rng), u(rng)};
      }
      const double a = tree.Distance2(v), b = Brute(pts, v);
      ++checked;
      if (std::memcmp(&a, &b, sizeof(double)) != 0) {
        if (++bad < 10) std::printf("mismatch %s n=%zu q=%d: %.17g vs %.17g\n", name, pts.size(), q, a, b);
      }
    }
  };
  // small sets: leaves only, and the first splits
  for (int n = 0; n <= 40; ++n) {
    std::vector<Vec3> pts;
    for (int i = 0; i < n; ++i) pts.push_back({u(rng), u(rng), u(rng)});
    check(pts, "small", 50);
  }
  // uniform in a box
  {
    std::vector<Vec3> pts;
    for (int i = 0; i < 5000; ++i) pts.push_back({u(rng), u(rng), u(rng)});
    check(pts, "box", 4000);
  }
  // a curved wall: face centres of a stretched surface grid
  {
    std::vector<Vec3> pts;
    for (int i = 0; i < 120; ++i)
      for (int k = 0; k < 30; ++k) {
        const double x = -1.0 + 2.0 * i / 119.0, z = -0.5 + k / 29.0;
        pts.push_back({x, 0.05 * std::sin(6.0 * x) * (1.0 + 0.1 * z), z});
      }
    check(pts, "wall", 4000);
  }
  // ties: an integer lattice with repeated points, queried at lattice and
  // half-lattice points where many distances are equal
  {
    std::vector<Vec3> pts;
    for (int r = 0; r < 2; ++r)
      for (int i = -4; i <= 4; ++i)
        for (int j = -4; j <= 4; ++j)
          for (int k = -2; k <= 2; ++k) pts.push_back({0.25 * i, 0.25 * j, 0.25 * k});
    const NearestPoints tree(pts);
    for (int i = -9; i <= 9; ++i)
      for (int j = -9; j <= 9; ++j)
        for (int k = -5; k <= 5; ++k) {
          const Vec3 v{0.125 * i, 0.125 * j, 0.125 * k};
          const double a = tree.Distance2(v), b = Brute(pts, v);
          ++checked;
          if (std::memcmp(&a, &b, sizeof(double)) != 0 && ++bad < 10)
            std::printf("mismatch ties: %.17g vs %.17g\n", a, b);
        }
  }
  std::printf("%s %ld queries, %ld mismatches\n", bad ? "FAIL" : "ok", checked, bad);
  return bad ? 1 : 0;
}
