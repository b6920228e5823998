// Mesh side of the boundary: synthetic closed surfaces and RWG construction.
//
// The reference reads an ASCII STL (modReadStl.f90:1378-1442) and builds RWG
// edges (modRWG.f90:8-174) and triangle geometry (modRWG.f90:270-328) in
// host code before calling inicializar_MLFMA.  These routines produce the same
// arrays, in the same order and with the same floating-point operations, so a
// caller that has a triangle list gets bit-identical geometry:
//   - edges are visited as half-edges (v1,v2),(v1,v3),(v2,v3) per triangle;
//     the first later half-edge with the same vertex pair closes the edge and
//     its triangle becomes the "minus" half (llenar_todo, modRWG.f90:59-131);
//   - centroids and sub-triangle centroids use the single-precision literals
//     (1./3.) and (2./3.) promoted to double, as the reference does.
// A hash map replaces the reference's O(E^2) search; the result is the same.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "lywell.h"

namespace {

struct Tri {
  int a, b, c;
};

void icosa(std::vector<double>& p, std::vector<Tri>& t) {
  const double g = (1.0 + std::sqrt(5.0)) / 2.0;
  const double v[12][3] = {{-1, g, 0}, {1, g, 0},  {-1, -g, 0}, {1, -g, 0},
                           {0, -1, g}, {0, 1, g},  {0, -1, -g}, {0, 1, -g},
                           {g, 0, -1}, {g, 0, 1},  {-g, 0, -1}, {-g, 0, 1}};
  for (auto& q : v) {
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    p.push_back(q[0] / n);
    p.push_back(q[1] / n);
    p.push_back(q[2] / n);
  }
  const int f[20][3] = {{0, 11, 5}, {0, 5, 1},   {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                        {1, 5, 9},  {5, 11, 4},  {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                        {3, 9, 4},  {3, 4, 2},   {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                        {4, 9, 5},  {2, 4, 11},  {6, 2, 10},  {8, 6, 7}, {9, 8, 1}};
  for (auto& q : f) t.push_back({q[0], q[1], q[2]});
}

int midpoint(std::vector<double>& p, std::unordered_map<uint64_t, int>& cache, int a, int b) {
  uint64_t key = a < b ? ((uint64_t)a << 32) | (uint32_t)b : ((uint64_t)b << 32) | (uint32_t)a;
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  double m[3];
  for (int k = 0; k < 3; ++k) m[k] = 0.5 * (p[3 * a + k] + p[3 * b + k]);
  double n = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  int id = (int)(p.size() / 3);
  for (int k = 0; k < 3; ++k) p.push_back(m[k] / n);
  cache.emplace(key, id);
  return id;
}

void build_unit_icosphere(int subdiv, std::vector<double>& p, std::vector<Tri>& t) {
  icosa(p, t);
  for (int s = 0; s < subdiv; ++s) {
    std::unordered_map<uint64_t, int> cache;
    std::vector<Tri> nt;
    nt.reserve(t.size() * 4);
    for (const Tri& q : t) {
      int ab = midpoint(p, cache, q.a, q.b);
      int bc = midpoint(p, cache, q.b, q.c);
      int ca = midpoint(p, cache, q.c, q.a);
      nt.push_back({q.a, ab, ca});
      nt.push_back({q.b, bc, ab});
      nt.push_back({q.c, ca, bc});
      nt.push_back({ab, bc, ca});
    }
    t.swap(nt);
  }
}

}  // namespace

extern "C" {

int lw_mesh_icosphere_size(int subdiv, int* np, int* nt) {
  if (subdiv < 0 || subdiv > 9) return LW_ERR_ARG;
  long long f = 20;
  for (int s = 0; s < subdiv; ++s) f *= 4;
  *nt = (int)f;
  *np = (int)(f / 2 + 2);
  return LW_OK;
}

int lw_mesh_icosphere(int subdiv, double ax, double ay, double az, double* p_coord, int* t_p,
                      double* t_normal) {
  if (subdiv < 0 || subdiv > 9 || !(ax > 0) || !(ay > 0) || !(az > 0)) return LW_ERR_ARG;
  std::vector<double> p;
  std::vector<Tri> t;
  build_unit_icosphere(subdiv, p, t);
  int np = (int)(p.size() / 3);
  for (int i = 0; i < np; ++i) {
    p_coord[3 * i + 0] = ax * p[3 * i + 0];
    p_coord[3 * i + 1] = ay * p[3 * i + 1];
    p_coord[3 * i + 2] = az * p[3 * i + 2];
  }
  for (size_t i = 0; i < t.size(); ++i) {
    int v[3] = {t[i].a, t[i].b, t[i].c};
    const double *A = p_coord + 3 * v[0], *B = p_coord + 3 * v[1], *C = p_coord + 3 * v[2];
    double e1[3], e2[3], n[3];
    for (int k = 0; k < 3; ++k) {
      e1[k] = B[k] - A[k];
      e2[k] = C[k] - A[k];
    }
    n[0] = e1[1] * e2[2] - e1[2] * e2[1];
    n[1] = e1[2] * e2[0] - e1[0] * e2[2];
    n[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double s = n[0] * (A[0] + B[0] + C[0]) + n[1] * (A[1] + B[1] + C[1]) + n[2] * (A[2] + B[2] + C[2]);
    if (s < 0) {  // keep outward orientation
      std::swap(v[1], v[2]);
      for (int k = 0; k < 3; ++k) n[k] = -n[k];
    }
    double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) {
      t_p[3 * i + k] = v[k] + 1;
      if (t_normal) t_normal[3 * i + k] = n[k] / nn;
    }
  }
  return LW_OK;
}

int lw_mesh_perturb_sh(int np, double* p_coord, int ncoef, const int* deg, const int* ord,
                       const double* coef) {
  // Radial perturbation r -> r (1 + sum_c coef_c Y_{deg_c, ord_c}(theta, phi)) with
  // orthonormal real spherical harmonics (ord < 0: sin(|m| phi) branch).
  if (np <= 0 || ncoef < 0) return LW_ERR_ARG;
  for (int c = 0; c < ncoef; ++c) {
    const int l = deg[c], m = ord[c] < 0 ? -ord[c] : ord[c];
    if (l < 0 || l > 64 || m > l) return LW_ERR_ARG;
  }
  for (int i = 0; i < np; ++i) {
    double* q = p_coord + 3 * i;
    double r = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    if (!(r > 0)) return LW_ERR_ARG;
    double th = std::acos(std::max(-1.0, std::min(1.0, q[2] / r))), ph = std::atan2(q[1], q[0]);
    double f = 1.0;
    for (int c = 0; c < ncoef; ++c) {
      int l = deg[c], m = ord[c] < 0 ? -ord[c] : ord[c];
      double y = std::sph_legendre((unsigned)l, (unsigned)m, th);
      if (m > 0) y *= std::sqrt(2.0) * (ord[c] < 0 ? std::sin(m * ph) : std::cos(m * ph));
      f += coef[c] * y;
    }
    for (int k = 0; k < 3; ++k) q[k] *= f;
  }
  return LW_OK;
}

int lw_rwg_build(int np, const double* p_coord, int nt, const int* t_p, int* e_p, int* e_t,
                 int* e_po, double* e_long, double* e_centro) {
  if (np <= 0 || nt <= 0) return LW_ERR_ARG;
  const int nall = 3 * nt;
  std::vector<int> a(nall), b(nall);
  for (int it = 0; it < nt; ++it) {
    const int* q = t_p + 3 * it;
    a[3 * it + 0] = q[0];
    b[3 * it + 0] = q[1];
    a[3 * it + 1] = q[0];
    b[3 * it + 1] = q[2];
    a[3 * it + 2] = q[1];
    b[3 * it + 2] = q[2];
  }
  // Live half-edges waiting for a partner, keyed by unordered vertex pair.
  // Processing i in increasing order, the first later live half-edge with the
  // same key is found by scanning forward once: every key's occurrences are
  // kept in increasing order.
  std::unordered_map<uint64_t, std::vector<int>> occ;
  occ.reserve((size_t)nall);
  auto key = [&](int i) {
    uint32_t lo = (uint32_t)std::min(a[i], b[i]), hi = (uint32_t)std::max(a[i], b[i]);
    return ((uint64_t)lo << 32) | hi;
  };
  for (int i = 0; i < nall; ++i) occ[key(i)].push_back(i);
  std::vector<char> dead(nall, 0);
  int ne = 0;
  for (int i = 0; i < nall; ++i) {
    if (dead[i]) continue;
    const std::vector<int>& v = occ[key(i)];
    int partner = -1;
    for (int j : v)
      if (j > i && !dead[j]) {
        partner = j;
        break;
      }
    if (partner < 0) {
      dead[i] = 1;
      continue;
    }
    dead[partner] = 1;
    int x = i / 3 + 1, y = partner / 3 + 1;
    int pa = a[i], pb = b[i];
    e_p[2 * ne] = pa;
    e_p[2 * ne + 1] = pb;
    e_t[2 * ne] = x;
    e_t[2 * ne + 1] = y;
    auto opp = [&](int tri) {
      const int* q = t_p + 3 * (tri - 1);
      for (int k = 0; k < 2; ++k)
        if (q[k] != pa && q[k] != pb) return q[k];
      return q[2];
    };
    e_po[2 * ne] = opp(x);
    e_po[2 * ne + 1] = opp(y);
    ++ne;
  }
  for (int i = 0; i < ne; ++i) {
    const double *A = p_coord + 3 * (e_p[2 * i] - 1), *B = p_coord + 3 * (e_p[2 * i + 1] - 1);
    double v[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    e_long[i] = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    for (int k = 0; k < 3; ++k) e_centro[3 * i + k] = 0.5 * (A[k] + B[k]);
  }
  return ne;
}

int lw_tri_geometry(int np, const double* p_coord, int nt, const int* t_p, double* t_area,
                    double* t_baric, double* t_baric_sub) {
  if (np <= 0 || nt <= 0) return LW_ERR_ARG;
  const double third = (double)(1.0f / 3.0f), two3 = (double)(2.0f / 3.0f);
  // sub-triangle vertex sets of hallar_geometria_rwg (modRWG.f90:309-317),
  // vertices numbered 1..10 as there
  static const int sub[9][3] = {{1, 4, 9}, {4, 5, 10}, {5, 2, 6}, {6, 10, 7}, {7, 8, 3},
                                {7, 8, 10}, {8, 9, 10}, {9, 4, 10}, {5, 6, 10}};
  for (int i = 0; i < nt; ++i) {
    const double* P[3] = {p_coord + 3 * (t_p[3 * i] - 1), p_coord + 3 * (t_p[3 * i + 1] - 1),
                          p_coord + 3 * (t_p[3 * i + 2] - 1)};
    double v[11][3];
    for (int k = 0; k < 3; ++k) {
      double c = third * ((P[0][k] + P[1][k]) + P[2][k]);
      t_baric[3 * i + k] = c;
      v[1][k] = P[0][k];
      v[2][k] = P[1][k];
      v[3][k] = P[2][k];
      v[4][k] = P[0][k] + third * (P[1][k] - P[0][k]);
      v[5][k] = P[0][k] + two3 * (P[1][k] - P[0][k]);
      v[6][k] = P[1][k] + third * (P[2][k] - P[1][k]);
      v[7][k] = P[1][k] + two3 * (P[2][k] - P[1][k]);
      v[8][k] = P[2][k] + third * (P[0][k] - P[2][k]);
      v[9][k] = P[2][k] + two3 * (P[0][k] - P[2][k]);
      v[10][k] = c;
    }
    for (int s = 0; s < 9; ++s)
      for (int k = 0; k < 3; ++k)
        t_baric_sub[27 * i + 3 * s + k] =
            third * ((v[sub[s][0]][k] + v[sub[s][1]][k]) + v[sub[s][2]][k]);
    double ba[3], ca[3];
    for (int k = 0; k < 3; ++k) {
      ba[k] = P[0][k] - P[1][k];
      ca[k] = P[0][k] - P[2][k];
    }
    double cr[3] = {ba[1] * ca[2] - ba[2] * ca[1], ba[2] * ca[0] - ba[0] * ca[2],
                    ba[0] * ca[1] - ba[1] * ca[0]};
    t_area[i] = 0.5 * std::sqrt((cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2]);
  }
  return LW_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// ASCII STL, the reference's input format (modReadStl.f90).
//
// Writer: coordinates in fixed-point decimal with 12 fractional digits and no
// exponent, so that the reference's digit-accumulating parser (s_to_r8,
// modReadStl.f90:133-266: an exact integer mantissa divided by an exact power
// of ten) and a correctly rounded strtod read the same doubles.
//
// Reader: the facet vertices become 3*nf nodes in file order; duplicates are
// then merged exactly as limpiarnodos does (modReadStl.f90:1444-1533): for
// each node i in turn, the first later node with identical coordinates is
// overwritten by the current last node, triangles pointing at it are moved to
// i, triangles pointing at the last node are moved to its slot, and the node
// count shrinks by one.  Hash maps replace the reference's O(P^2) scans.
#include <cstdio>
#include <map>
#include <set>
#include <string>

namespace {

struct Key3 {
  double x, y, z;
  bool operator<(const Key3& o) const {
    if (x != o.x) return x < o.x;
    if (y != o.y) return y < o.y;
    return z < o.z;
  }
};

struct StlData {
  std::vector<double> p, n;  // (3, np) nodes after merging, (3, nt) facet normals
  std::vector<int> t;        // (3, nt), 1-based
};

int parse_stl(const char* path, StlData& out) {
  FILE* f = std::fopen(path, "r");
  if (!f) return LW_ERR_ARG;
  std::vector<double> node, nrm;
  char line[512];
  int state = 0, vert = 0;
  while (std::fgets(line, sizeof line, f)) {
    char w1[64] = {0}, w2[64] = {0};
    double a = 0, b = 0, c = 0;
    int nw = std::sscanf(line, "%63s %63s", w1, w2);
    if (nw < 1) continue;
    std::string s1(w1);
    for (auto& ch : s1) ch = (char)std::tolower((unsigned char)ch);
    if (s1 == "facet") {
      if (std::sscanf(line, " %*s %*s %lf %lf %lf", &a, &b, &c) != 3) { a = b = c = 0.0; }
      nrm.push_back(a);
      nrm.push_back(b);
      nrm.push_back(c);
      state = 2;
      vert = 0;
    } else if (s1 == "vertex") {
      if (state != 2 || std::sscanf(line, " %*s %lf %lf %lf", &a, &b, &c) != 3) {
        std::fclose(f);
        return LW_ERR_ARG;
      }
      node.push_back(a);
      node.push_back(b);
      node.push_back(c);
      ++vert;
    } else if (s1 == "endfacet") {
      if (vert != 3) {
        std::fclose(f);
        return LW_ERR_ARG;
      }
      state = 1;
    }
  }
  std::fclose(f);
  const int nt = (int)(nrm.size() / 3);
  if (nt < 1 || (int)node.size() != 9 * nt) return LW_ERR_ARG;
  int np = 3 * nt;
  std::vector<int> tri(3 * (size_t)nt);
  for (int i = 0; i < 3 * nt; ++i) tri[i] = i + 1;
  // node -> triangle corners referencing it; coordinates -> live positions
  std::vector<std::vector<int>> refs((size_t)np + 1);
  for (int c = 0; c < 3 * nt; ++c) refs[tri[c]].push_back(c);
  std::map<Key3, std::set<int>> where;
  auto key = [&](int id) { return Key3{node[3 * (id - 1)], node[3 * (id - 1) + 1], node[3 * (id - 1) + 2]}; };
  for (int id = 1; id <= np; ++id) where[key(id)].insert(id);
  auto retarget = [&](int from, int to) {
    for (int c : refs[from]) tri[c] = to;
    auto& dst = refs[to];
    dst.insert(dst.end(), refs[from].begin(), refs[from].end());
    refs[from].clear();
  };
  for (int i = 1; i <= np; ++i) {
    for (;;) {
      auto& s = where[key(i)];
      auto it = s.upper_bound(i);
      if (it == s.end()) break;
      const int res = *it;
      s.erase(it);
      // node res := node np (then np is dropped)
      if (res != np) {
        where[key(np)].erase(np);
        for (int d = 0; d < 3; ++d) node[3 * (res - 1) + d] = node[3 * (np - 1) + d];
        where[key(res)].insert(res);
      }
      retarget(res, i);
      if (res != np) retarget(np, res);
      --np;
    }
  }
  out.p.assign(node.begin(), node.begin() + 3 * (size_t)np);
  out.t = tri;
  out.n = nrm;
  return LW_OK;
}

}  // namespace

extern "C" {

int lw_stl_size(const char* path, int* np, int* nt) {
  if (!path || !np || !nt) return LW_ERR_ARG;
  StlData d;
  int rc = parse_stl(path, d);
  if (rc) return rc;
  *np = (int)(d.p.size() / 3);
  *nt = (int)(d.t.size() / 3);
  return LW_OK;
}

int lw_stl_read(const char* path, double* p_coord, int* t_p, double* t_normal) {
  if (!path || !p_coord || !t_p) return LW_ERR_ARG;
  StlData d;
  int rc = parse_stl(path, d);
  if (rc) return rc;
  std::memcpy(p_coord, d.p.data(), sizeof(double) * d.p.size());
  std::memcpy(t_p, d.t.data(), sizeof(int) * d.t.size());
  if (t_normal) std::memcpy(t_normal, d.n.data(), sizeof(double) * d.n.size());
  return LW_OK;
}

int lw_stl_write(const char* path, int np, const double* p_coord, int nt, const int* t_p,
                 const double* t_normal) {
  if (!path || np < 3 || nt < 1 || !p_coord || !t_p) return LW_ERR_ARG;
  for (int i = 0; i < 3 * np; ++i)
    if (!(std::fabs(p_coord[i]) < 1.0e3)) return LW_ERR_ARG;  // keeps 16 significant digits exact
  FILE* f = std::fopen(path, "w");
  if (!f) return LW_ERR_ARG;
  std::fprintf(f, "solid lywell\n");
  for (int i = 0; i < nt; ++i) {
    const int* q = t_p + 3 * i;
    double n[3] = {0, 0, 0};
    if (t_normal) {
      for (int k = 0; k < 3; ++k) n[k] = t_normal[3 * i + k];
    } else {
      const double *A = p_coord + 3 * (q[0] - 1), *B = p_coord + 3 * (q[1] - 1), *C = p_coord + 3 * (q[2] - 1);
      double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
      n[0] = e1[1] * e2[2] - e1[2] * e2[1];
      n[1] = e1[2] * e2[0] - e1[0] * e2[2];
      n[2] = e1[0] * e2[1] - e1[1] * e2[0];
      const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int k = 0; k < 3; ++k) n[k] /= nn;
    }
    std::fprintf(f, "  facet normal %.12f %.12f %.12f\n    outer loop\n", n[0], n[1], n[2]);
    for (int k = 0; k < 3; ++k) {
      const double* P = p_coord + 3 * (q[k] - 1);
      std::fprintf(f, "      vertex %.12f %.12f %.12f\n", P[0], P[1], P[2]);
    }
    std::fprintf(f, "    endloop\n  endfacet\n");
  }
  std::fprintf(f, "endsolid lywell\n");
  return std::fclose(f) == 0 ? LW_OK : LW_ERR_ARG;
}

}  // extern "C"
