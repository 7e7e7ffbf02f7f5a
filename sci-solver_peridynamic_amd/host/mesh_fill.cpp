// mesh_fill.cpp -- OBJ reader and exact point-in-mesh test.
#include "mesh_fill.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "exact_predicates.h"

bool load_obj(const char* path, ObjMesh& mesh, std::string& err) {
  FILE* fp = fopen(path, "r");
  if (!fp) {
    err = std::string("cannot open OBJ file ") + path;
    return false;
  }
  mesh.v.clear();
  mesh.f.clear();
  std::vector<long> poly;
  std::vector<std::pair<size_t, long>> pending;  // (index into f, raw index) for checks
  char line[4096];
  int lineno = 0;
  bool ok = true;
  while (ok && fgets(line, sizeof(line), fp)) {
    ++lineno;
    if (line[0] == 'v' && (line[1] == ' ' || line[1] == '\t')) {
      float a = 0.f, b = 0.f, c = 0.f;
      if (sscanf(line + 2, "%f %f %f", &a, &b, &c) != 3) {
        err = "bad vertex line " + std::to_string(lineno);
        ok = false;
        break;
      }
      mesh.v.push_back(a);
      mesh.v.push_back(b);
      mesh.v.push_back(c);
    } else if (line[0] == 'f' && (line[1] == ' ' || line[1] == '\t')) {
      poly.clear();
      char* s = line + 2;
      while (*s) {
        while (*s == ' ' || *s == '\t') ++s;
        if (!*s || *s == '\n' || *s == '\r') break;
        char* end = nullptr;
        const long idx = strtol(s, &end, 10);
        if (end == s || idx == 0) {
          err = "bad face token on line " + std::to_string(lineno);
          ok = false;
          break;
        }
        poly.push_back(idx);
        s = end;
        while (*s && *s != ' ' && *s != '\t' && *s != '\n' && *s != '\r') ++s;  // skip /t/n
      }
      if (!ok) break;
      const long nv = (long)(mesh.v.size() / 3);
      for (long& idx : poly) {
        idx = idx > 0 ? idx - 1 : nv + idx;
        if (idx < 0 || idx >= nv) {
          // forward references are legal in OBJ: checked again at the end
          pending.push_back({mesh.f.size(), idx});
        }
      }
      // fan from the first vertex: (v0, v_{k-1}, v_k)
      for (size_t k = 2; k < poly.size(); ++k) {
        mesh.f.push_back((int)poly[0]);
        mesh.f.push_back((int)poly[k - 1]);
        mesh.f.push_back((int)poly[k]);
      }
    }
  }
  fclose(fp);
  if (!ok) return false;
  const long nv = (long)(mesh.v.size() / 3);
  for (int idx : mesh.f)
    if (idx < 0 || idx >= nv) {
      err = "face index out of range in " + std::string(path);
      return false;
    }
  if (mesh.f.empty()) {
    err = std::string("no faces in ") + path;
    return false;
  }
  return true;
}

MeshInside::MeshInside(const std::vector<double>& xyz, const std::vector<int>& tri)
    : x_(xyz), tri_(tri) {
  const size_t nt = tri.size() / 3;
  for (int a = 0; a < 3; ++a) {
    lo_[a] = INFINITY;
    hi_[a] = -INFINITY;
  }
  zmax_.assign(nt, -INFINITY);
  for (size_t t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) {
      const double* v = &xyz[3 * (size_t)tri[3 * t + k]];
      for (int a = 0; a < 3; ++a) {
        lo_[a] = std::min(lo_[a], v[a]);
        hi_[a] = std::max(hi_[a], v[a]);
      }
      zmax_[t] = std::max(zmax_[t], v[2]);
    }
  empty_ = nt == 0;
  if (empty_) return;
  // about two triangles per cell on average
  const double ex = hi_[0] - lo_[0], ey = hi_[1] - lo_[1];
  const double area = std::max(ex * ey, 1e-300);
  const double target = sqrt(area / std::max<double>(1.0, 0.5 * (double)nt));
  nx_ = std::max(1, std::min(4096, (int)ceil(ex / target)));
  ny_ = std::max(1, std::min(4096, (int)ceil(ey / target)));
  cell_[0] = ex > 0 ? ex / nx_ : 1.0;
  cell_[1] = ey > 0 ? ey / ny_ : 1.0;
  std::vector<int> count((size_t)nx_ * ny_ + 1, 0);
  auto range = [&](size_t t, int* c0, int* c1) {
    for (int a = 0; a < 2; ++a) {
      double mn = INFINITY, mx = -INFINITY;
      for (int k = 0; k < 3; ++k) {
        const double v = xyz[3 * (size_t)tri[3 * t + k] + a];
        mn = std::min(mn, v);
        mx = std::max(mx, v);
      }
      c0[a] = cell_of(mn, a);
      c1[a] = cell_of(mx, a);
    }
  };
  for (size_t t = 0; t < nt; ++t) {
    int c0[2], c1[2];
    range(t, c0, c1);
    for (int j = c0[1]; j <= c1[1]; ++j)
      for (int i = c0[0]; i <= c1[0]; ++i) ++count[(size_t)j * nx_ + i];
  }
  start_.assign((size_t)nx_ * ny_ + 1, 0);
  for (size_t c = 0; c < (size_t)nx_ * ny_; ++c) start_[c + 1] = start_[c] + count[c];
  items_.assign(start_.back(), 0);
  std::vector<int> fill(start_.begin(), start_.end() - 1);
  for (size_t t = 0; t < nt; ++t) {
    int c0[2], c1[2];
    range(t, c0, c1);
    for (int j = c0[1]; j <= c1[1]; ++j)
      for (int i = c0[0]; i <= c1[0]; ++i) items_[fill[(size_t)j * nx_ + i]++] = (int)t;
  }
}

// Monotone in v, so a point on the boundary of a triangle's box falls in a
// cell the triangle was binned into.
int MeshInside::cell_of(double v, int axis) const {
  const int n = axis == 0 ? nx_ : ny_;
  const double c = floor((v - lo_[axis]) / cell_[axis]);
  if (!(c >= 0.0)) return 0;
  if (c >= (double)(n - 1)) return n - 1;
  return (int)c;
}

// Does the ray from p along +z cross triangle t (counted once per crossing
// over the whole mesh)?  q, r, s are put in counter-clockwise xy order; p
// must lie strictly below the triangle's plane and inside its xy projection,
// with points on an edge or a vertex assigned to one side by the half-open
// rules below.
bool MeshInside::crosses(const double p[3], int t) const {
  const double* q = &x_[3 * (size_t)tri_[3 * (size_t)t]];
  const double* r = &x_[3 * (size_t)tri_[3 * (size_t)t + 1]];
  const double* s = &x_[3 * (size_t)tri_[3 * (size_t)t + 2]];
  const int qrs = exact::orient2d(q, r, s);
  if (qrs == 0) return false;  // degenerate in the xy projection
  if (qrs < 0) std::swap(r, s);
  if (exact::orient3d(p, q, r, s) >= 0) return false;  // on or above the triangle
  const int pqr = exact::orient2d(p, q, r);
  if (pqr < 0) return false;
  const int prs = exact::orient2d(p, r, s);
  if (prs < 0) return false;
  const int psq = exact::orient2d(p, s, q);
  if (psq < 0) return false;
  if (pqr > 0 && prs > 0 && psq > 0) return true;
  // on the open edge from a to b: counted when a is above b in y, or level
  // with it and left of it
  auto edge = [](const double* a, const double* b) {
    if (a[1] != b[1]) return a[1] > b[1];
    return a[0] < b[0];
  };
  if (pqr == 0 && prs > 0 && psq > 0) return edge(q, r);
  if (prs == 0 && pqr > 0 && psq > 0) return edge(r, s);
  if (psq == 0 && pqr > 0 && prs > 0) return edge(s, q);
  // on a vertex a with neighbours b (next) and c (previous): counted when
  // the vertex is the lower end of the edge to c and not below b
  auto vertex = [](const double* a, const double* b, const double* c) {
    return a[1] >= b[1] && a[1] < c[1];
  };
  if (p[0] == q[0] && p[1] == q[1]) return vertex(q, r, s);
  if (p[0] == r[0] && p[1] == r[1]) return vertex(r, s, q);
  if (p[0] == s[0] && p[1] == s[1]) return vertex(s, q, r);
  return false;
}

bool MeshInside::inside(const double p[3]) const {
  if (empty_) return false;
  for (int a = 0; a < 3; ++a)
    if (!(lo_[a] <= p[a] && p[a] <= hi_[a])) return false;
  const int i = cell_of(p[0], 0), j = cell_of(p[1], 1);
  const size_t c = (size_t)j * nx_ + i;
  int hits = 0;
  for (int k = start_[c]; k < start_[c + 1]; ++k) {
    const int t = items_[k];
    if (p[2] > zmax_[t]) continue;  // the ray starts above the triangle
    if (crosses(p, t)) ++hits;
  }
  return (hits & 1) != 0;
}
