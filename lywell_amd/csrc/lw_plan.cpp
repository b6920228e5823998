// Host-side operator plan for the MLFMA matvec (see lw_plan.h).
//
// Structure and parity-critical scalars follow modMLFMA.f90 exactly; the
// algorithms used to get there are not the reference's O(N^2) searches:
//   - the octree is built level by level (boxes ordered by (parent, child
//     slot)), which is the order the reference's depth-first recursion
//     creates them in (geomMLFMA, :497-581);
//   - interaction and near lists are found among the children of the
//     parent's neighbours instead of scanning all box pairs (:835-960), then
//     ordered by box index as the reference's j-loop orders them;
//   - transfer vectors are deduplicated by their integer box offset; the
//     first pair (in (i, j) order) that produces an offset supplies the rab
//     used to evaluate it, as in the reference's linear search;
//   - near-field rows come from a basis->leaf map instead of a linear search
//     per row (llenarZesparcida, :1594-1634).
#include "lw_plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "lywell.h"

namespace lw {

void Physics::set_lambda(double l) {
  vp = 1.0 / std::sqrt(mu * epsi);
  lambda = l;
  w = 2.0 * pi * vp / l;
  kappa = 2.0 * pi / l;
}

double norm2_3(const double* v) {
  double m = 0.0, s = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double a = std::fabs(v[i]);
    if (m == 0.0) {
      m = a;
    } else if (a > m) {
      const double t = m / a, tsq = t * t;
      s *= tsq;
      s += tsq;
      m = a;
    } else {
      const double t = a / m;
      s += t * t;
    }
  }
  return m * std::sqrt(1.0 + s);
}

namespace {

// ---- spherical Hankel h_l^(2)(x) = j_l(x) - i y_l(x), Zhang & Jin's
// SPHJ/SPHY as the reference calls them (modGlobalMetod.f90:493-979): j_l
// comes from a normalised backward recurrence started at MSTA1/MSTA2, y_l
// from the forward recurrence.
double envj(int n, double x) { return 0.5 * std::log10(6.28 * n) - n * std::log10(1.36 * x / n); }

int start_order(double x, int n_or_zero, int mp, bool second) {
  const double a0 = std::fabs(x);
  double obj = mp;
  int n0, n1;
  if (!second) {
    n0 = (int)(1.1 * a0) + 1;
  } else {
    const double hmp = 0.5 * mp, ejn = envj(n_or_zero, a0);
    if (ejn <= hmp) {
      obj = mp;
      n0 = (int)(1.1 * a0);
    } else {
      obj = hmp + ejn;
      n0 = n_or_zero;
    }
  }
  double f0 = envj(n0, a0) - obj;
  n1 = n0 + 5;
  double f1 = envj(n1, a0) - obj;
  int nn = n1;
  for (int it = 0; it < 20; ++it) {
    nn = (int)(n1 - (n1 - n0) / (1.0 - f0 / f1));
    const double f = envj(nn, a0) - obj;
    if (std::abs(nn - n1) < 1) break;
    n0 = n1;
    f0 = f1;
    n1 = nn;
    f1 = f;
  }
  return second ? nn + 10 : nn;
}

double sph_j(int n, double x) {
  if (std::fabs(x) <= 1.0e-100) return n == 0 ? 1.0 : 0.0;
  const double j0 = std::sin(x) / x, j1 = (j0 - std::cos(x)) / x;
  if (n == 0) return j0;
  if (n == 1) return j1;
  int nm = n, m = start_order(x, 0, 200, false);
  if (m < n)
    nm = m;
  else
    m = start_order(x, n, 15, true);
  if (nm < n) return 0.0;  // not reached for kD >= 1 at the degrees used here
  double f0 = 0.0, f1 = 1.0 - 100.0, f = 0.0, fn = 0.0;
  for (int k = m; k >= 0; --k) {
    f = (2.0 * k + 3.0) * f1 / x - f0;
    if (k == n) fn = f;
    f0 = f1;
    f1 = f;
  }
  const double cs = (std::fabs(j0) <= std::fabs(j1)) ? j1 / f0 : j0 / f;
  return cs * fn;
}

double sph_y(int n, double x) {
  if (x < 1.0e-60) return -1.0e300;
  const double y0 = -std::cos(x) / x, y1 = (y0 - std::sin(x)) / x;
  if (n == 0) return y0;
  if (n == 1) return y1;
  double f0 = y0, f1 = y1, f = 0.0;
  for (int k = 2; k <= n; ++k) {
    f = (2.0 * k - 1.0) * f1 / x - f0;
    if (1.0e300 <= std::fabs(f)) return f;
    f0 = f1;
    f1 = f;
  }
  return f;
}

// ---- sample grids (llenar_kappas, modMLFMA.f90:453-495)
void fill_samples(LevelPlan& lv) {
  const double pi = Physics::pi;
  const int nk = lv.nk;
  lv.kap.resize(3 * (size_t)nk);
  lv.ath.resize(3 * (size_t)nk);
  lv.aph.resize(3 * (size_t)nk);
  lv.ang.resize(2 * (size_t)nk);
  lv.pes.resize(nk);
  const double dphi = 2.0 * pi / lv.nph, dth = pi / lv.nth;
  int s = 0;
  for (int j = 1; j <= lv.nph; ++j) {
    const double phi = (dphi / 2.0) * (double)((2 * j) - 1);
    for (int k = 1; k <= lv.nth; ++k, ++s) {
      const double th = (dth / 2.0) * (double)((2 * k) - 1);
      const double st = std::sin(th), ct = std::cos(th), sp = std::sin(phi), cp = std::cos(phi);
      double* kp = &lv.kap[3 * s];
      kp[0] = st * cp;
      kp[1] = st * sp;
      kp[2] = ct;
      double* at = &lv.ath[3 * s];
      at[0] = ct * cp;
      at[1] = ct * sp;
      at[2] = -st;
      double* ap = &lv.aph[3 * s];
      ap[0] = -sp;
      ap[1] = cp;
      ap[2] = 0.0;
      lv.ang[2 * s] = th;
      lv.ang[2 * s + 1] = phi;
      lv.pes[s] = (std::cos((k - 1) * dth) - std::cos(k * dth)) * dphi;
    }
  }
}

// ---- nearest fine sample along one angle, hallarKcercanos's bisection
// (modMLFMA.f90:1117-1262): the search keeps [techo, piso] and stops when the
// midpoint hits a bracket end, then picks the closer end.
long long bisect_nearest(double v, long long n, const double* grid_of /* 1-based via lambda */,
                         long long stride, long long offset) {
  auto at = [&](long long p) { return grid_of[2 * ((p - 1) * stride) + offset]; };
  const double del = at(2) - at(1);
  long long lo = 1, hi = n, mid;
  for (;;) {
    mid = (lo + hi) / 2;
    const double val = at(mid);
    if (mid == lo || mid == hi) {
      mid = (std::fabs(v - at(lo)) < std::fabs(v - at(hi))) ? lo : hi;
      break;
    }
    if (val == v) break;
    if (del > 0) {
      if (val > v) hi = mid;
      else if (val < v) lo = mid;
    } else {
      if (val > v) lo = mid;
      else if (val < v) hi = mid;
    }
  }
  return mid;
}

void neighbours(long long pos, long long n, bool increasing, long long& lower, long long& upper) {
  // hallarKcercanos wraps both angles at the grid ends (no pole reflection)
  if (increasing) {
    lower = pos == 1 ? n : pos - 1;
    upper = pos == n ? 1 : pos + 1;
  } else {
    lower = pos == n ? 1 : pos + 1;
    upper = pos == 1 ? n : pos - 1;
  }
}

void lagrange3(double x, const double* xm, double* w) {
  for (int i = 0; i < 3; ++i) {
    double ww = 1.0;
    for (int j = 0; j < 3; ++j)
      if (j != i) ww = ww * ((x - xm[j]) / (xm[i] - xm[j]));
    w[i] = ww;
  }
}

// llenarMinterp (modMLFMA.f90:998-1092): 9-point Lagrange stencil per coarse
// sample, weights w = Wtheta(m2) * Wphi(m1) in (phi, theta) nesting.
void fill_interp(LevelPlan& lc, const LevelPlan& lf) {
  const long long nth = lf.nth, nph = lf.nph;
  lc.wint.resize(9 * (size_t)lc.nk);
  lc.iint.resize(9 * (size_t)lc.nk);
  lc.sth_idx.assign(3 * (size_t)lc.nth, 0);
  lc.sth_w.assign(3 * (size_t)lc.nth, 0.0);
  lc.sph_idx.assign(3 * (size_t)lc.nph, 0);
  lc.sph_w.assign(3 * (size_t)lc.nph, 0.0);
  const bool phi_inc = lf.ang[2 * nth + 1] - lf.ang[1] > 0;
  const bool th_inc = lf.ang[2] - lf.ang[0] > 0;
  for (int i = 0; i < lc.nk; ++i) {
    const double th = lc.ang[2 * i], ph = lc.ang[2 * i + 1];
    const long long pp = bisect_nearest(ph, nph, lf.ang.data(), nth, 1);
    const long long pt = bisect_nearest(th, nth, lf.ang.data(), 1, 0);
    long long pl, pu, tl, tu;
    neighbours(pp, nph, phi_inc, pl, pu);
    neighbours(pt, nth, th_inc, tl, tu);
    const long long phis[3] = {pp, pl, pu}, thes[3] = {pt, tl, tu};
    int V[9];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) V[3 * a + b] = (int)(nth * (phis[a] - 1) + thes[b]);
    const double mt[3] = {lf.ang[2 * (V[0] - 1)], lf.ang[2 * (V[1] - 1)], lf.ang[2 * (V[2] - 1)]};
    const double mp[3] = {lf.ang[2 * (V[0] - 1) + 1], lf.ang[2 * (V[3] - 1) + 1],
                          lf.ang[2 * (V[6] - 1) + 1]};
    double wt[3], wp[3];
    lagrange3(th, mt, wt);
    lagrange3(ph, mp, wp);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        lc.wint[9 * (size_t)i + 3 * a + b] = wt[b] * wp[a];
        lc.iint[9 * (size_t)i + 3 * a + b] = V[3 * a + b] - 1;
      }
    // the stencil is a tensor product: theta part per coarse theta row, phi
    // part per coarse phi column (sample i = iphi*nth + itheta)
    const int it = i % lc.nth, ip = i / lc.nth;
    for (int b = 0; b < 3; ++b) {
      lc.sth_idx[3 * it + b] = (int)thes[b] - 1;
      lc.sth_w[3 * it + b] = wt[b];
      lc.sph_idx[3 * ip + b] = (int)phis[b] - 1;
      lc.sph_w[3 * ip + b] = wp[b];
    }
  }
  // transpose for the gather form of Anterpolar: fine sample -> (coarse row,
  // weight) in ascending coarse row order, the order the reference's scatter
  // loop accumulates in (modMLFMA.f90:1382-1388)
  lc.wt_ptr.assign((size_t)lf.nk + 1, 0);
  for (size_t e = 0; e < lc.iint.size(); ++e) lc.wt_ptr[lc.iint[e] + 1]++;
  for (int s = 0; s < lf.nk; ++s) lc.wt_ptr[s + 1] += lc.wt_ptr[s];
  lc.wt_row.resize(lc.iint.size());
  lc.wt_w.resize(lc.iint.size());
  std::vector<int> fill(lc.wt_ptr.begin(), lc.wt_ptr.end() - 1);
  for (int i = 0; i < lc.nk; ++i)
    for (int j = 0; j < 9; ++j) {
      const int s = lc.iint[9 * (size_t)i + j];
      const int at = fill[s]++;
      lc.wt_row[at] = i;
      lc.wt_w[at] = lc.wint[9 * (size_t)i + j];
    }
}

}  // namespace

int build_plan(const Geometry& g, double lambda, double ladomax, Plan& plan, std::string& msg) {
  Physics& ph = plan.phys;
  ph.set_lambda(lambda);
  const double pi = Physics::pi;
  const int ne = g.ne;
  plan.ne = ne;
  if (ne < 2 || !(ladomax > 0) || !(lambda > 0)) {
    msg = "lw_mlfma_init: need num_e >= 2, ladomax > 0, lambda > 0";
    return LW_ERR_ARG;
  }
  // bounding box of edge centres (limites_dispersor, modGlobalMetod.f90:386-434)
  int lim[6] = {0, 0, 0, 0, 0, 0};
  const double* ec = g.e_centro;
  for (int i = 1; i < ne; ++i) {
    const double* c = ec + 3 * i;
    if (c[0] > ec[3 * lim[0]]) lim[0] = i;
    if (c[0] < ec[3 * lim[1]]) lim[1] = i;
    if (c[1] > ec[3 * lim[2] + 1]) lim[2] = i;
    if (c[1] < ec[3 * lim[3] + 1]) lim[3] = i;
    if (c[2] > ec[3 * lim[4] + 2]) lim[4] = i;
    if (c[2] < ec[3 * lim[5] + 2]) lim[5] = i;
  }
  const double mx = ec[3 * lim[0]], mnx = ec[3 * lim[1]], my = ec[3 * lim[2] + 1],
               mny = ec[3 * lim[3] + 1], mz = ec[3 * lim[4] + 2], mnz = ec[3 * lim[5] + 2];
  const double dx = mx - mnx, dy = my - mny, dz = mz - mnz;
  const double roo[3] = {0.5 * (mx + mnx), 0.5 * (my + mny), 0.5 * (mz + mnz)};
  double side = (dx >= dy && dx >= dz) ? dx : (dy >= dx && dy >= dz) ? dy : dz;
  side = side + lambda / 10000.0;
  // inicializar_MLFMA (modMLFMA.f90:189): log10(2.) is a single-precision value
  const int L = (int)std::ceil(1.0 + (std::log10(side / (ladomax * lambda)) / (double)std::log10(2.0f)));
  if (L <= 3) {
    msg = "lw_mlfma_init: scatterer electrically too small for MLFMA (nivel_max <= 3)";
    return LW_ERR_OCTREE;
  }
  if (L > 12) {
    msg = "lw_mlfma_init: more than 12 octree levels requested";
    return LW_ERR_ARG;
  }
  plan.L = L;
  plan.lv.assign(L + 1, LevelPlan());
  const double sq3 = (double)std::sqrt(3.0f), third = (double)(1.0f / 3.0f);
  for (int i = 1; i <= L; ++i) {
    LevelPlan& lv = plan.lv[i];
    lv.lado = side / std::ldexp(1.0, i - 1);
    const double kd = (2.0 * pi / lambda) * (sq3 * lv.lado);
    lv.nmult = (int)std::floor(kd + 6 * std::pow(kd, third));
    lv.nth = lv.nmult;
    lv.nph = (int)(2.0 * lv.nmult);
    lv.nk = lv.nth * lv.nph;
    if (lv.nmult + 2 > Plan::NMAX) {
      msg = "lw_mlfma_init: multipole count beyond supported table size";
      return LW_ERR_ARG;
    }
    if (i >= 3) fill_samples(lv);
  }

  // ---- octree, level by level
  {
    LevelPlan& root = plan.lv[1];
    root.nb = 1;
    root.centre.assign(roo, roo + 3);
    root.parent.assign(1, -1);
    root.bases.resize(1);
    root.bases[0].resize(ne);
    for (int i = 0; i < ne; ++i) root.bases[0][i] = i + 1;
  }
  for (int n = 1; n < L; ++n) {
    LevelPlan& lp = plan.lv[n];
    LevelPlan& lc = plan.lv[n + 1];
    lp.child_ptr.assign(lp.nb + 1, 0);
    const double lcubo = lp.lado, lq = lcubo / 4.0, half = (lcubo / 2.0) / 2.0;
    for (int b = 0; b < lp.nb; ++b) {
      lp.child_ptr[b] = lc.nb;
      const double* c = &lp.centre[3 * b];
      // dividir_en_8_hijos child order (modMLFMA.f90:706-755): x slowest,
      // then z, then y fastest with sign (-,+)
      static const int sx[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
      static const int sy[8] = {-1, 1, -1, 1, -1, 1, -1, 1};
      static const int sz[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
      double ch[8][3], lim[8][6];
      for (int k = 0; k < 8; ++k) {
        ch[k][0] = sx[k] < 0 ? c[0] - lq : c[0] + lq;
        ch[k][1] = sy[k] < 0 ? c[1] - lq : c[1] + lq;
        ch[k][2] = sz[k] < 0 ? c[2] - lq : c[2] + lq;
        for (int d = 0; d < 3; ++d) {
          lim[k][2 * d] = ch[k][d] + half;
          lim[k][2 * d + 1] = ch[k][d] - half;
        }
      }
      std::vector<int> owned[8];
      for (int id : lp.bases[b]) {
        const double* e = ec + 3 * (id - 1);
        for (int k = 0; k < 8; ++k) {
          if (e[0] <= lim[k][0] && e[0] >= lim[k][1] && e[1] <= lim[k][2] && e[1] >= lim[k][3] &&
              e[2] <= lim[k][4] && e[2] >= lim[k][5]) {
            owned[k].push_back(id);
            break;
          }
        }
      }
      for (int k = 0; k < 8; ++k) {
        if (owned[k].empty()) continue;
        lc.centre.insert(lc.centre.end(), ch[k], ch[k] + 3);
        lc.parent.push_back(b);
        lc.bases.push_back(std::move(owned[k]));
        lc.nb++;
      }
    }
    lp.child_ptr[lp.nb] = lc.nb;
  }
  // every basis must land in a leaf (the root box contains all centres)
  {
    size_t tot = 0;
    for (auto& v : plan.lv[L].bases) tot += v.size();
    if ((int)tot != ne) {
      msg = "lw_mlfma_init: octree lost basis functions (degenerate geometry)";
      return LW_ERR_ARG;
    }
  }

  // ---- interaction lists, transfer vectors and (leaf) near lists
  const double tol = lambda / 100.0;
  std::vector<std::vector<int>> nearp;  // near boxes (incl. self) of level n-1
  {
    // level 2: all 8 boxes of level 2 are mutual neighbours by distance test
    LevelPlan& l2 = plan.lv[2];
    nearp.assign(l2.nb, {});
    for (int i = 0; i < l2.nb; ++i)
      for (int j = 0; j < l2.nb; ++j) {
        double r[3] = {l2.centre[3 * i] - l2.centre[3 * j], l2.centre[3 * i + 1] - l2.centre[3 * j + 1],
                       l2.centre[3 * i + 2] - l2.centre[3 * j + 2]};
        if (norm2_3(r) < (2 * l2.lado - tol)) nearp[i].push_back(j);
      }
  }
  for (int n = 3; n <= L; ++n) {
    LevelPlan& lv = plan.lv[n];
    LevelPlan& pl = plan.lv[n - 1];
    std::vector<std::vector<int>> nearc(lv.nb);
    lv.far_ptr.assign(lv.nb + 1, 0);
    std::unordered_map<int, int> offset_id;  // integer offset key -> transfer index
    std::vector<double> trans_rab;
    for (int i = 0; i < lv.nb; ++i) {
      const int pi_ = lv.parent[i];
      // candidate boxes: children of the parent's near boxes, in index order
      std::vector<int> cand;
      for (int pj : nearp[pi_])
        for (int j = pl.child_ptr[pj]; j < pl.child_ptr[pj + 1]; ++j) cand.push_back(j);
      std::sort(cand.begin(), cand.end());
      for (int j : cand) {
        const int pj = lv.parent[j];
        const double rp[3] = {pl.centre[3 * pi_] - pl.centre[3 * pj], pl.centre[3 * pi_ + 1] - pl.centre[3 * pj + 1],
                              pl.centre[3 * pi_ + 2] - pl.centre[3 * pj + 2]};
        if (!(norm2_3(rp) < (2 * pl.lado - tol))) continue;
        const double rh[3] = {lv.centre[3 * i] - lv.centre[3 * j], lv.centre[3 * i + 1] - lv.centre[3 * j + 1],
                              lv.centre[3 * i + 2] - lv.centre[3 * j + 2]};
        if (norm2_3(rh) >= (2 * lv.lado - tol)) {
          int o[3];
          for (int d = 0; d < 3; ++d) o[d] = (int)std::lround(rh[d] / lv.lado);
          const int key = ((o[0] + 8) * 17 + (o[1] + 8)) * 17 + (o[2] + 8);
          auto it = offset_id.find(key);
          int id;
          if (it == offset_id.end()) {
            id = (int)offset_id.size();
            offset_id.emplace(key, id);
            trans_rab.insert(trans_rab.end(), rh, rh + 3);
          } else {
            id = it->second;
          }
          lv.far_box.push_back(j);
          lv.far_tr.push_back(id);
        } else {
          nearc[i].push_back(j);
        }
      }
      lv.far_ptr[i + 1] = (int)lv.far_box.size();
    }
    lv.nT = (int)offset_id.size();
    if (lv.nT > 316) {
      msg = "lw_mlfma_init: more than 316 transfer vectors on a level";
      return LW_ERR_ARG;
    }
    // transfer-vector scalars (calc_functransf, modMLFMA.f90:964-994): unit
    // vector, truncation and the Hankel values; the angular sums run on the GPU
    lv.trans_u.resize(3 * (size_t)lv.nT);
    lv.trans_L.resize(lv.nT);
    lv.hstride = lv.nmult + 1;
    lv.trans_h.assign(2 * (size_t)lv.hstride * lv.nT, 0.0);
    for (int t = 0; t < lv.nT; ++t) {
      const double* rab = &trans_rab[3 * t];
      const double modul = norm2_3(rab);
      const double kD = (2.0 * pi / lambda) * modul;
      int Lt = lv.nmult;
      const int kDint = (int)kD;
      if (kDint < Lt) Lt = kDint;
      lv.trans_L[t] = Lt;
      for (int d = 0; d < 3; ++d) lv.trans_u[3 * t + d] = rab[d] / modul;
      for (int l = 0; l <= Lt; ++l) {
        lv.trans_h[2 * ((size_t)t * lv.hstride + l)] = sph_j(l, kD);
        lv.trans_h[2 * ((size_t)t * lv.hstride + l) + 1] = -sph_y(l, kD);
      }
    }
    nearp.swap(nearc);
  }
  // nearp now holds the leaf near lists (box indices, ascending)

  // ---- interpolation stencils
  for (int n = 3; n < L; ++n) fill_interp(plan.lv[n], plan.lv[n + 1]);

  // ---- leaf ordering and near CSR
  LevelPlan& lf = plan.lv[L];
  plan.slot_basis.resize(ne);
  plan.basis_slot.resize(ne);
  plan.basis_box.resize(ne);
  plan.box_slot.assign(lf.nb + 1, 0);
  {
    int s = 0;
    for (int b = 0; b < lf.nb; ++b) {
      plan.box_slot[b] = s;
      for (int id : lf.bases[b]) {
        plan.slot_basis[s] = id - 1;
        plan.basis_slot[id - 1] = s;
        plan.basis_box[id - 1] = b;
        ++s;
      }
    }
    plan.box_slot[lf.nb] = s;
  }
  std::vector<std::vector<int>> near_cols(lf.nb);
  long long nnz = 0;
  for (int b = 0; b < lf.nb; ++b) {
    auto& cols = near_cols[b];
    for (int j : nearp[b]) cols.insert(cols.end(), lf.bases[j].begin(), lf.bases[j].end());
    std::sort(cols.begin(), cols.end());
    nnz += (long long)cols.size() * (long long)lf.bases[b].size();
  }
  if (nnz >= (1LL << 31)) {
    msg = "lw_mlfma_init: near field exceeds 2^31 entries";
    return LW_ERR_ARG;
  }
  plan.nnz = nnz;
  plan.ia.resize((size_t)ne + 1);
  plan.ja.resize((size_t)nnz);
  {
    long long k = 0;
    for (int m = 0; m < ne; ++m) {
      plan.ia[m] = (int)k;
      const auto& cols = near_cols[plan.basis_box[m]];
      for (int c : cols) plan.ja[k++] = c - 1;
    }
    plan.ia[ne] = (int)k;
  }

  // ---- binomial squares for legendre (modGlobalMetod.f90:988-1016)
  const int NM = Plan::NMAX;
  plan.binom2.assign((size_t)(NM + 1) * (NM + 1), 0.0);
  std::vector<double> fact(NM + 1);
  fact[0] = 1.0;
  for (int x = 1; x <= NM; ++x) {
    double r = 1.0;
    for (int i = 1; i <= x; ++i) r = r * (double)i;
    fact[x] = r;
  }
  for (int n = 0; n <= NM; ++n)
    for (int i = 0; i <= n; ++i) {
      const double c = fact[n] / (fact[i] * fact[n - i]);
      plan.binom2[(size_t)n * (NM + 1) + i] = c * c;
    }
  return LW_OK;
}

// ---------------------------------------------------------------- ILU0

void tri_levels(int n, const std::vector<int>& jlu, const std::vector<int>& ju, bool upper,
                std::vector<int>& lev_ptr, std::vector<int>& lev_rows) {
  // jlu/ju are the 1-based MSR arrays stored with a leading pad (jlu[k] is
  // jlu(k)); row i's L part is jlu(i)..ju(i)-1, U part ju(i)..jlu(i+1)-1.
  std::vector<int> lev(n, 0);
  int maxl = 0;
  if (!upper) {
    for (int i = 1; i <= n; ++i) {
      int l = 0;
      for (int k = jlu[i]; k <= ju[i - 1] - 1; ++k) l = std::max(l, lev[jlu[k] - 1] + 1);
      lev[i - 1] = l;
      maxl = std::max(maxl, l);
    }
  } else {
    for (int i = n; i >= 1; --i) {
      int l = 0;
      for (int k = ju[i - 1]; k <= jlu[i + 1] - 1; ++k) l = std::max(l, lev[jlu[k] - 1] + 1);
      lev[i - 1] = l;
      maxl = std::max(maxl, l);
    }
  }
  lev_ptr.assign(maxl + 2, 0);
  for (int i = 0; i < n; ++i) lev_ptr[lev[i] + 1]++;
  for (int l = 0; l <= maxl; ++l) lev_ptr[l + 1] += lev_ptr[l];
  lev_rows.resize(n);
  std::vector<int> fill(lev_ptr.begin(), lev_ptr.end() - 1);
  for (int i = 0; i < n; ++i) lev_rows[fill[lev[i]]++] = i;
  // within a level, rows with more entries on the swept side first (the
  // rows of a level are independent; the factorisation deals them to its
  // waves in this order)
  for (int l = 0; l <= maxl; ++l) {
    auto side = [&](int i) {
      return upper ? jlu[i + 2] - ju[i] : ju[i] - jlu[i + 1];
    };
    std::stable_sort(lev_rows.begin() + lev_ptr[l], lev_rows.begin() + lev_ptr[l + 1],
                     [&](int a, int b) { return side(a) > side(b); });
  }
}

}  // namespace lw
