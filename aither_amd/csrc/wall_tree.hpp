// Nearest-point distance over a fixed set of points (the viscous-wall face
// centres of a case), exact to the bit of a full scan.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <limits>
#include <vector>

#include "host_case.hpp"

namespace ath {

// A k-d tree: median splits along the axis of the depth (x, y, z in turn),
// leaves of up to kLeaf points searched in full.  The search skips a subtree
// beyond a splitting plane only when the squared distance to the plane
// exceeds the best so far.  Every point beyond the plane has a coordinate
// difference at least as large in magnitude, rounding is monotonic, and a
// rounded sum of non-negative terms is at least each term, so its rounded
// squared distance is at least the plane's: the minimum found is the full
// scan's minimum, bit for bit.  Squared distances are summed as
// ((0 + dx dx) + dy dy) + dz dz.
class NearestPoints {
 public:
  explicit NearestPoints(std::vector<Vec3> pts) : pts_(std::move(pts)) { Build(0, pts_.size(), 0); }
  size_t size() const { return pts_.size(); }
  // the smallest squared distance from q to a point (the largest double
  // when there are none)
  double Distance2(const Vec3 &q) const {
    double best = std::numeric_limits<double>::max();
    Search(q, 0, pts_.size(), 0, best);
    return best;
  }
  double Distance(const Vec3 &q) const { return std::sqrt(Distance2(q)); }
  static double Dist2(const Vec3 &q, const Vec3 &f) {
    const double dx = q.x - f.x, dy = q.y - f.y, dz = q.z - f.z;
    return ((0.0 + dx * dx) + dy * dy) + dz * dz;
  }

 private:
  static constexpr size_t kLeaf = 8;
  std::vector<Vec3> pts_;
  static double Coord(const Vec3 &v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
  void Build(size_t lo, size_t hi, int depth) {
    if (hi - lo <= kLeaf) return;
    const size_t mid = lo + (hi - lo) / 2;
    const int a = depth % 3;
    std::nth_element(pts_.begin() + lo, pts_.begin() + mid, pts_.begin() + hi,
                     [a](const Vec3 &u, const Vec3 &v) { return Coord(u, a) < Coord(v, a); });
    Build(lo, mid, depth + 1);
    Build(mid + 1, hi, depth + 1);
  }
  void Search(const Vec3 &q, size_t lo, size_t hi, int depth, double &best) const {
    if (hi - lo <= kLeaf) {
      for (size_t n = lo; n < hi; ++n) best = std::min(best, Dist2(q, pts_[n]));
      return;
    }
    const size_t mid = lo + (hi - lo) / 2;
    const int a = depth % 3;
    best = std::min(best, Dist2(q, pts_[mid]));
    // points before mid have this coordinate <= the split, points after >= it
    const double diff = Coord(q, a) - Coord(pts_[mid], a);
    const bool below = diff < 0.0;
    if (below) Search(q, lo, mid, depth + 1, best);
    else Search(q, mid + 1, hi, depth + 1, best);
    if (!(diff * diff > best)) {
      if (below) Search(q, mid + 1, hi, depth + 1, best);
      else Search(q, lo, mid, depth + 1, best);
    }
  }
};

}  // namespace ath
