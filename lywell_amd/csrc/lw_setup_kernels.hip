// Operator-setup kernels: radiation patterns, transfer functions and the
// near-field MoM entries, computed on the GPU from the host plan.
//
// This file is compiled with -ffp-contract=off: the transfer functions go
// through the reference's binomial Legendre sum (modGlobalMetod.f90:988-1002),
// whose cancellation turns a one-ulp change of its argument into ~1e-6 of its
// value at degree 38, so its argument and its operation sequence must be the
// reference's, rounding for rounding (no fused multiply-adds).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lw_internal.h"

namespace lw {
namespace {

constexpr double kMu = Physics::mu, kEps = Physics::epsi, kPi = Physics::pi;

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// real(8)**integer as flang lowers it: binary exponentiation
__device__ inline double ipow(double x, int n) {
  double r = 1.0, b = x;
  unsigned m = (unsigned)n;
  while (m) {
    if (m & 1u) r *= b;
    b *= b;
    m >>= 1;
  }
  return r;
}

// Per-basis geometry of calcRad (modMLFMA.f90:669-699) in slot order.
__global__ void k_rad_geom(int ne, const int* __restrict__ slot_basis, const int* __restrict__ basis_box,
                           const double* __restrict__ box_c, const double* __restrict__ tb,
                           const double* __restrict__ pc, const int* __restrict__ et, const int* __restrict__ epo,
                           const double* __restrict__ el, RadGeom* __restrict__ g) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ne) return;
  const int f = slot_basis[s];
  const double* c = box_c + 3 * basis_box[f];
  const double* bp = tb + 3 * (et[2 * f] - 1);
  const double* vp = pc + 3 * (epo[2 * f] - 1);
  const double* bm = tb + 3 * (et[2 * f + 1] - 1);
  const double* vm = pc + 3 * (epo[2 * f + 1] - 1);
  RadGeom r;
  for (int d = 0; d < 3; ++d) {
    r.rp[d] = bp[d] - vp[d];
    r.rm[d] = vm[d] - bm[d];
    r.dp[d] = bp[d] - c[d];
    r.dm[d] = bm[d] - c[d];
  }
  r.len = el[f];
  g[s] = r;
}

// The phase rows of the matvec's pattern form: row p = exp(j ka khat_k .
// (c_T - C_b)) for the triangle T and leaf box b of pair p, at the leaf
// samples k < len (all of them, or the first half, whose antipodes give the
// rest), the factors calcRad multiplies, with its operation order.
__global__ void k_tri_phase(int npairs, int len, const int* __restrict__ pair_tri, const int* __restrict__ pair_box,
                            const double* __restrict__ box_c, const double* __restrict__ tb,
                            const double* __restrict__ kap, double ka, double2* __restrict__ tph) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)npairs * len) return;
  const int p = (int)(idx / len), k = (int)(idx % len);
  const double* t = tb + 3 * pair_tri[p];
  const double* c = box_c + 3 * pair_box[p];
  double d[3];
  for (int q = 0; q < 3; ++q) d[q] = t[q] - c[q];
  tph[idx] = rad_phase(kap + 3 * k, d, ka);
}

// calcRad, non-maxPrecision branch: radiation pattern of basis slot s about
// its leaf-box centre, theta and phi components.
__global__ void k_rad(int ne, int nk, const double* __restrict__ kap, const double* __restrict__ ath,
                      const double* __restrict__ aph, const RadGeom* __restrict__ geo, double ka,
                      double2* __restrict__ rad) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)ne * nk) return;
  const int s = (int)(idx / nk), k = (int)(idx % nk);
  double2 r0, r1;
  rad_eval(geo[s], kap + 3 * k, ath + 3 * k, aph + 3 * k, ka, r0, r1);
  rad[((size_t)s * 2 + 0) * nk + k] = r0;
  rad[((size_t)s * 2 + 1) * nk + k] = r1;
}

// calc_functransf (modMLFMA.f90:964-994):
// T(k) = (k0/4pi) sum_{l=0}^{Ltrue} (-i)^{l+1} (2l+1) h_l(kD) P_l(khat . u)
__global__ void k_transfer(int nk, int nT, int nmax1, const double* __restrict__ kap,
                           const double* __restrict__ tu, const int* __restrict__ tL,
                           const double* __restrict__ th, int hstride,
                           const double* __restrict__ binom2, double k0, double2* __restrict__ T) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nk * nT) return;
  const int t = (int)(idx / nk), k = (int)(idx % nk);
  const double* u = tu + 3 * t;
  const double* vk = kap + 3 * k;
  const double x = (vk[0] * u[0] + vk[1] * u[1]) + vk[2] * u[2];
  const int Lt = tL[t];
  const double* h = th + 2 * (size_t)t * hstride;
  double vr = 0.0, vi = 0.0;
  for (int l = 0; l <= Lt; ++l) {
    // legendre(l, x), modGlobalMetod.f90:988-1002
    const double* b2 = binom2 + (size_t)l * nmax1;
    double res = 0.0;
    for (int i = 0; i <= l; ++i) res = res + (b2[i] * ipow(x + 1.0, l - i)) * ipow(x - 1.0, i);
    const double P = ldexp(1.0, -l) * res;
    const double r = (double)(2.0f * (float)l + 1.0f);
    const double hr = h[2 * l], hi = h[2 * l + 1];
    double br, bi;
    switch ((l + 1) & 3) {
      case 0: br = r * hr; bi = r * hi; break;
      case 1: br = r * hi; bi = -(r * hr); break;
      case 2: br = -(r * hr); bi = -(r * hi); break;
      default: br = -(r * hi); bi = r * hr; break;
    }
    vr = vr + br * P;
    vi = vi + bi * P;
  }
  T[(size_t)t * nk + k] = make_double2((vr * k0) / (4.0 * kPi), (vi * k0) / (4.0 * kPi));
}

__device__ inline double2 green(const double* pm, const double* pn, double kfs) {
  const double d0 = pm[0] - pn[0], d1 = pm[1] - pn[1], d2 = pm[2] - pn[2];
  const double r = sqrt((d0 * d0 + d1 * d1) + d2 * d2);
  double s, c;
  sincos(-(kfs * r), &s, &c);
  return make_double2(c / r, s / r);
}

// Zmn_MoM (modMoM.f90:89-135): 9-point source quadrature per triangle against
// the test triangles' centroids; one wavefront per CSR row.
template <typename V>
__global__ void k_near(int ne, const int* __restrict__ ia, const int* __restrict__ ja,
                       const double* __restrict__ tb, const double* __restrict__ tbs,
                       const double* __restrict__ pc, const int* __restrict__ et,
                       const int* __restrict__ epo, const double* __restrict__ el, double w,
                       V* __restrict__ z) {
  const int m = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (m >= ne) return;
  const double kfs = w * sqrt(kMu * kEps);
  const double c72 = (double)(1.0f / 72.0f);
  const double* tm1 = tb + 3 * (et[2 * m] - 1);
  const double* tm2 = tb + 3 * (et[2 * m + 1] - 1);
  const double* pm1 = pc + 3 * (epo[2 * m] - 1);
  const double* pm2 = pc + 3 * (epo[2 * m + 1] - 1);
  const double lm = el[m];
  double qm1[3], qm2[3];
  for (int d = 0; d < 3; ++d) {
    qm1[d] = tm1[d] - pm1[d];
    qm2[d] = -tm2[d] + pm2[d];
  }
  for (int e = ia[m] + lane; e < ia[m + 1]; e += 64) {
    const int n = ja[e];
    const double* sn1 = tbs + 27 * (et[2 * n] - 1);
    const double* sn2 = tbs + 27 * (et[2 * n + 1] - 1);
    const double* pn1 = pc + 3 * (epo[2 * n] - 1);
    const double* pn2 = pc + 3 * (epo[2 * n + 1] - 1);
    double2 AP[3] = {}, AM[3] = {}, PhiP = {}, PhiM = {};
    for (int i = 0; i < 9; ++i) {
      const double2 g11 = green(tm1, sn1 + 3 * i, kfs), g12 = green(tm1, sn2 + 3 * i, kfs);
      const double2 g21 = green(tm2, sn1 + 3 * i, kfs), g22 = green(tm2, sn2 + 3 * i, kfs);
      for (int d = 0; d < 3; ++d) {
        const double r1 = sn1[3 * i + d] - pn1[d], r2 = pn2[d] - sn2[3 * i + d];
        AP[d].x = (AP[d].x + g11.x * r1) + g12.x * r2;
        AP[d].y = (AP[d].y + g11.y * r1) + g12.y * r2;
        AM[d].x = (AM[d].x + g21.x * r1) + g22.x * r2;
        AM[d].y = (AM[d].y + g21.y * r1) + g22.y * r2;
      }
      PhiP.x = PhiP.x + (g11.x - g12.x);
      PhiP.y = PhiP.y + (g11.y - g12.y);
      PhiM.x = PhiM.x + (g21.x - g22.x);
      PhiM.y = PhiM.y + (g21.y - g22.y);
    }
    const double ln = el[n];
    const double sa = ((kMu / kPi) * c72) * ln;
    // (-1) * Phi * l / (i 36 pi w eps) = i * Phi * l / (36 pi w eps)
    const double den = ((36.0 * kPi) * w) * kEps;
    const double2 dPhi = make_double2(PhiM.x - PhiP.x, PhiM.y - PhiP.y);
    // PintPhi = lm * (PhiM' - PhiP') with Phi' = i Phi l / den
    const double2 pint_phi = make_double2(-(dPhi.y * ln / den) * lm, (dPhi.x * ln / den) * lm);
    double2 s = {0.0, 0.0};
    for (int d = 0; d < 3; ++d) {
      s.x = s.x + (AP[d].x * sa) * qm1[d] + (AM[d].x * sa) * qm2[d];
      s.y = s.y + (AP[d].y * sa) * qm1[d] + (AM[d].y * sa) * qm2[d];
    }
    // PintA = (i 0.5 w lm) * s
    const double f = (0.5 * w) * lm;
    const double2 pint_a = make_double2(-f * s.y, f * s.x);
    const double zr = pint_phi.x + pint_a.x, zi = pint_phi.y + pint_a.y;
    V out;
    out.x = zr;
    out.y = zi;
    z[e] = out;
  }
}

}  // namespace

int launch_setup_rad(Handle* h) {
  const LevelPlan& lf = h->plan.lv[h->L];
  DevLevel& d = h->lv[h->L];
  int rc = upload(h, &h->d_ath, lf.ath);
  if (rc) return rc;
  rc = upload(h, &h->d_aph, lf.aph);
  if (rc) return rc;
  rc = dmalloc(h, &h->d_rgeo, (size_t)h->ne);
  if (rc) return rc;
  hipLaunchKernelGGL(k_rad_geom, dim3((unsigned)((h->ne + 255) / 256)), dim3(256), 0, h->stream, h->ne,
                     h->d_slot_basis, h->d_basis_box, d.centre, h->d_tb, h->d_p, h->d_et, h->d_epo, h->d_el,
                     h->d_rgeo);
  const long long tot = (long long)h->ne * lf.nk;
  const int blk = 256;
  hipLaunchKernelGGL(k_rad, dim3((unsigned)((tot + blk - 1) / blk)), dim3(blk), 0, h->stream, h->ne, lf.nk,
                     d.kap, h->d_ath, h->d_aph, h->d_rgeo, (2.0 * kPi) / h->plan.phys.lambda, h->d_rad);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

// The per-(leaf box, triangle) phase rows and each slot's two rows: the
// triangles of a box's bases in first-seen slot order.
int launch_setup_phases(Handle* h, const int* e_t) {
  const Plan& P = h->plan;
  const LevelPlan& lf = P.lv[h->L];
  std::vector<int> pair_tri, pair_box, box_row(lf.nb + 1, 0);
  std::vector<int2> slot_pair(P.ne);
  for (int b = 0; b < lf.nb; ++b) {
    const int first = (int)pair_tri.size();
    box_row[b] = first;
    for (int s = P.box_slot[b]; s < P.box_slot[b + 1]; ++s) {
      const int f = P.slot_basis[s];
      int rows[2];
      for (int side = 0; side < 2; ++side) {
        const int t = e_t[2 * f + side] - 1;
        int r = -1;
        for (int q = first; q < (int)pair_tri.size(); ++q)
          if (pair_tri[q] == t) {
            r = q;
            break;
          }
        if (r < 0) {
          r = (int)pair_tri.size();
          pair_tri.push_back(t);
          pair_box.push_back(b);
        }
        rows[side] = r;
      }
      slot_pair[s] = make_int2(rows[0], rows[1]);
    }
  }
  h->npairs = (int)pair_tri.size();
  box_row[lf.nb] = h->npairs;
  // LDS of the receive from phase rows: the box's 3-vector field and one
  // 3-vector per phase row of the box
  int maxrows = 0;
  for (int b = 0; b < lf.nb; ++b) maxrows = std::max(maxrows, box_row[b + 1] - box_row[b]);
  const long long lds = 3LL * (lf.nk + maxrows) * (long long)sizeof(double2);
  h->recv_lds = lds <= kRecvLdsMax ? (int)lds : 0;
  h->agg_lds = h->recv_lds ? 3 * maxrows * (int)sizeof(double2) : 0;
  // the factored aggregation and receive read half rows (the antipodal
  // samples from the conjugates); the pattern-by-pattern kernels full rows
  const bool half = h->recv_lds && !(h->flags & LW_PATTERN_ORDER) && lf.nph == 2 * lf.nth && lf.nk == lf.nth * lf.nph;
  if (h->recv_lds && !(h->flags & LW_PATTERN_ORDER) && !half) {
    set_error("lw_mlfma_init: leaf sample grid without antipodal pairs");
    return LW_ERR_ARG;
  }
  h->ph_len = half ? lf.nk / 2 : lf.nk;
  // the bases each phase row serves: per row, entries 2 s + side (s the
  // basis slot, side 0 = its plus triangle, 1 = its minus triangle) in slot
  // order, plus then minus
  std::vector<int> row_ptr(h->npairs + 1, 0), row_ent(2 * (size_t)P.ne);
  for (int s = 0; s < P.ne; ++s) {
    ++row_ptr[slot_pair[s].x + 1];
    ++row_ptr[slot_pair[s].y + 1];
  }
  for (int r = 0; r < h->npairs; ++r) row_ptr[r + 1] += row_ptr[r];
  {
    std::vector<int> fill(row_ptr.begin(), row_ptr.end() - 1);
    for (int s = 0; s < P.ne; ++s) {
      row_ent[fill[slot_pair[s].x]++] = 2 * s;
      row_ent[fill[slot_pair[s].y]++] = 2 * s + 1;
    }
  }
  int *d_pt = nullptr, *d_pb = nullptr;
  int rc;
  if ((rc = upload(h, &d_pt, pair_tri)) || (rc = upload(h, &d_pb, pair_box)) || (rc = upload(h, &h->d_box_row, box_row)) ||
      (rc = upload(h, &h->d_row_ptr, row_ptr)) || (rc = upload(h, &h->d_row_ent, row_ent)) ||
      (rc = upload(h, &h->d_slot_pair, slot_pair)) || (rc = dmalloc(h, &h->d_tph, (size_t)h->npairs * h->ph_len)))
    return rc;
  const long long tot = (long long)h->npairs * h->ph_len;
  hipLaunchKernelGGL(k_tri_phase, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, h->npairs, h->ph_len,
                     d_pt, d_pb, h->lv[h->L].centre, h->d_tb, h->lv[h->L].kap, (2.0 * kPi) / h->plan.phys.lambda,
                     h->d_tph);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

int launch_setup_transfer(Handle* h, int level, const double* d_binom2, const double* d_tu,
                          const int* d_tL, const double* d_th, int hstride) {
  DevLevel& d = h->lv[level];
  if (d.nT == 0) return LW_OK;
  const long long tot = (long long)d.nk * d.nT;
  const int blk = 256;
  hipLaunchKernelGGL(k_transfer, dim3((unsigned)((tot + blk - 1) / blk)), dim3(blk), 0, h->stream, d.nk, d.nT,
                     Plan::NMAX + 1, d.kap, d_tu, d_tL, d_th, hstride, d_binom2,
                     2.0 * kPi / h->plan.phys.lambda, d.T);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

int launch_setup_near(Handle* h) {
  const int rows_per_block = 4;
  dim3 grid((unsigned)((h->ne + rows_per_block - 1) / rows_per_block)), block(64 * rows_per_block);
  if (h->fp64_near)
    hipLaunchKernelGGL(k_near<double2>, grid, block, 0, h->stream, h->ne, h->d_ia, h->d_ja, h->d_tb, h->d_tbs,
                       h->d_p, h->d_et, h->d_epo, h->d_el, h->plan.phys.w, (double2*)h->d_z);
  else
    hipLaunchKernelGGL(k_near<float2>, grid, block, 0, h->stream, h->ne, h->d_ia, h->d_ja, h->d_tb, h->d_tbs,
                       h->d_p, h->d_et, h->d_epo, h->d_el, h->plan.phys.w, (float2*)h->d_z);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

}  // namespace lw
