// exact_predicates.h -- signs of the 2D and 3D orientation determinants of
// double-precision points, exact for every input.
//
// Method: a floating-point evaluation whose rounding error is bounded by a
// multiple of the determinant's permanent decides the sign when the value
// clears that bound; otherwise the determinant is evaluated exactly as a
// floating-point expansion (a sum of non-overlapping doubles), following
// J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
// Robust Geometric Predicates", Discrete Comput. Geom. 18 (1997): two-sum
// (Knuth), two-product (here by fused multiply-add), expansion growth,
// expansion sum and scaling.  The sign of an expansion is the sign of its
// largest component.
//
//   orient2d(a, b, c) = det [a - c; b - c]           (x, y only)
//   orient3d(a, b, c, d) = det [a - d; b - d; c - d]
#ifndef PD_HOST_EXACT_PREDICATES_H
#define PD_HOST_EXACT_PREDICATES_H

#include <math.h>

#include <vector>

namespace exact {

typedef std::vector<double> Expansion;  // increasing magnitude, no zeros

inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bv = s - a;
  const double av = s - bv;
  e = (a - av) + (b - bv);
}

inline void two_prod(double a, double b, double& p, double& e) {
  p = a * b;
  e = fma(a, b, -p);
}

// e + b
inline Expansion grow(const Expansion& e, double b) {
  Expansion h;
  h.reserve(e.size() + 1);
  double q = b;
  for (double c : e) {
    double s, t;
    two_sum(q, c, s, t);
    if (t != 0.0) h.push_back(t);
    q = s;
  }
  if (q != 0.0) h.push_back(q);
  return h;
}

inline Expansion add(const Expansion& e, const Expansion& f) {
  Expansion h = e;
  for (double c : f) h = grow(h, c);
  return h;
}

// e * b
inline Expansion scale(const Expansion& e, double b) {
  Expansion h;
  if (e.empty() || b == 0.0) return h;
  h.reserve(2 * e.size());
  double q, t;
  two_prod(e[0], b, q, t);
  if (t != 0.0) h.push_back(t);
  for (size_t i = 1; i < e.size(); ++i) {
    double p1, p0, s, u;
    two_prod(e[i], b, p1, p0);
    two_sum(q, p0, s, u);
    if (u != 0.0) h.push_back(u);
    two_sum(p1, s, q, u);
    if (u != 0.0) h.push_back(u);
  }
  if (q != 0.0) h.push_back(q);
  return h;
}

inline Expansion mul(const Expansion& e, const Expansion& f) {
  Expansion h;
  for (double c : f) h = add(h, scale(e, c));
  return h;
}

inline Expansion neg(Expansion e) {
  for (double& c : e) c = -c;
  return e;
}

// a - b exactly
inline Expansion diff(double a, double b) {
  double s, t;
  two_sum(a, -b, s, t);
  Expansion h;
  if (t != 0.0) h.push_back(t);
  if (s != 0.0) h.push_back(s);
  return h;
}

inline int sign(const Expansion& e) {
  if (e.empty()) return 0;
  return e.back() > 0.0 ? 1 : -1;
}

constexpr double kEps = 1.1102230246251565e-16;  // 2^-53

inline int orient2d(const double* a, const double* b, const double* c) {
  const double acx = a[0] - c[0], bcx = b[0] - c[0];
  const double acy = a[1] - c[1], bcy = b[1] - c[1];
  const double l = acx * bcy, r = acy * bcx;
  const double det = l - r;
  const double bound = 8.0 * kEps * (fabs(l) + fabs(r));
  if (det > bound) return 1;
  if (-det > bound) return -1;
  const Expansion L = mul(diff(a[0], c[0]), diff(b[1], c[1]));
  const Expansion R = mul(diff(a[1], c[1]), diff(b[0], c[0]));
  return sign(add(L, neg(R)));
}

inline int orient3d(const double* a, const double* b, const double* c, const double* d) {
  const double adx = a[0] - d[0], ady = a[1] - d[1], adz = a[2] - d[2];
  const double bdx = b[0] - d[0], bdy = b[1] - d[1], bdz = b[2] - d[2];
  const double cdx = c[0] - d[0], cdy = c[1] - d[1], cdz = c[2] - d[2];
  const double det = adx * (bdy * cdz - bdz * cdy) + bdx * (cdy * adz - cdz * ady) +
                     cdx * (ady * bdz - adz * bdy);
  const double perm = fabs(adx) * (fabs(bdy * cdz) + fabs(bdz * cdy)) +
                      fabs(bdx) * (fabs(cdy * adz) + fabs(cdz * ady)) +
                      fabs(cdx) * (fabs(ady * bdz) + fabs(adz * bdy));
  const double bound = 16.0 * kEps * perm;
  if (det > bound) return 1;
  if (-det > bound) return -1;
  const Expansion ax = diff(a[0], d[0]), ay = diff(a[1], d[1]), az = diff(a[2], d[2]);
  const Expansion bx = diff(b[0], d[0]), by = diff(b[1], d[1]), bz = diff(b[2], d[2]);
  const Expansion cx = diff(c[0], d[0]), cy = diff(c[1], d[1]), cz = diff(c[2], d[2]);
  const Expansion m1 = add(mul(by, cz), neg(mul(bz, cy)));
  const Expansion m2 = add(mul(cy, az), neg(mul(cz, ay)));
  const Expansion m3 = add(mul(ay, bz), neg(mul(az, by)));
  return sign(add(add(mul(ax, m1), mul(bx, m2)), mul(cx, m3)));
}

}  // namespace exact

#endif
