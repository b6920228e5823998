// Multi-RHS form of the hot path: R right-hand sides share one pass over the
// operator.
//
//   far field (aggregation, interpolation, translation, anterpolation,
//   receive) in FP32 complex with FP32 accumulation;
//   near field (complex(4)-stored values, FP64 accumulation), ILU0 solves and
//   every CGS vector operation and reduction in FP64.
//
// Layouts (r = right-hand side, fastest):
//   batch vectors      V[i*R + r]                  (complex(8))
//   campo arrays       F[((k*nb + b)*2 + p)*R + r] (complex(4), sample-major so
//                      one sample's slice of a level is contiguous for the
//                      translation kernel)
//   patterns           rad32[(s*2 + p)*nk + k]     (complex(4), leaf slot order)
// The per-RHS CGS follows CGSold (modIterativo.f90:331-493) column by column:
// a column that meets its tolerance (or breaks down) is frozen, so every
// column's iterate is the one a single-RHS run would produce with this
// operator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "lw_internal.h"

namespace lw {
namespace {

constexpr double kMu = Physics::mu, kEps = Physics::epsi, kPi = Physics::pi;
constexpr int kWave = 64;

__device__ __forceinline__ float2 fmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 fmulc(float2 a, float2 b) {  // conj(a) * b
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 fadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 dmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 dadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 dsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 ddiv(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

__global__ void k_to_f32(long long n, const double2* __restrict__ in, float2* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}
__global__ void k_to_f32r(long long n, const double* __restrict__ in, float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

// ---------------------------------------------------------------- far field (FP32)

// Leaf aggregation: C_L[k][b][p][r] = sum_s rad(s; p,k) x(basis(s), r)
__global__ void __launch_bounds__(256) k_b_aggregate(int R, int nk, int nb, const int* __restrict__ box_slot,
                                                     const int* __restrict__ slot_basis,
                                                     const float2* __restrict__ rad, const double2* __restrict__ X,
                                                     float2* __restrict__ C) {
  extern __shared__ float2 xs[];  // (n_f, R)
  const int b = blockIdx.x;
  const int s0 = box_slot[b], nf = box_slot[b + 1] - s0;
  for (int e = threadIdx.x; e < nf * R; e += blockDim.x) {
    const double2 v = X[(size_t)slot_basis[s0 + e / R] * R + (e % R)];
    xs[e] = make_float2((float)v.x, (float)v.y);
  }
  __syncthreads();
  const int r = threadIdx.x % R, per = blockDim.x / R;
  const int nk2 = 2 * nk;
  for (int item = blockIdx.y * per + threadIdx.x / R; item < nk2; item += gridDim.y * per) {
    float2 acc = make_float2(0.f, 0.f);
    for (int s = 0; s < nf; ++s) acc = fadd(acc, fmul(rad[(size_t)(s0 + s) * nk2 + item], xs[s * R + r]));
    const int p = item / nk, k = item % nk;
    C[(((size_t)k * nb + b) * 2 + p) * R + r] = acc;
  }
}

// Upward pass: C_{n-1}[i][P] = sum_children ph(oct_c, i) sum_9 w_ij C_n[idx_j][c]
__global__ void __launch_bounds__(256) k_b_interp_up(int R, int nkc, int nbc, int nkf, int nbf,
                                                     const int* __restrict__ child_ptr, const int* __restrict__ oct,
                                                     const float* __restrict__ wint, const int* __restrict__ iint,
                                                     const float2* __restrict__ ph, const float2* __restrict__ Cf,
                                                     float2* __restrict__ Cc) {
  const int P = blockIdx.x;
  const long long t = (long long)blockIdx.y * blockDim.x + threadIdx.x;
  if (t >= (long long)nkc * 2 * R) return;
  const int r = (int)(t % R), p = (int)((t / R) % 2), i = (int)(t / (2 * R));
  float w[9];
  int id[9];
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    w[j] = wint[9 * (size_t)i + j];
    id[j] = iint[9 * (size_t)i + j];
  }
  float2 acc = make_float2(0.f, 0.f);
  for (int c = child_ptr[P]; c < child_ptr[P + 1]; ++c) {
    float2 a = make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const float2 v = Cf[(((size_t)id[j] * nbf + c) * 2 + p) * R + r];
      a.x += w[j] * v.x;
      a.y += w[j] * v.y;
    }
    acc = fadd(acc, fmul(a, ph[(size_t)oct[c] * nkc + i]));
  }
  Cc[(((size_t)i * nbc + P) * 2 + p) * R + r] = acc;
}

// Translation: B[k][b] = pes[k] sum_far T_t[k] C[k][far]; workgroups holding
// the same sample k run on the same XCD (workgroup id mod 8) so the k-slice
// of campoC stays in that XCD's L2.
__global__ void __launch_bounds__(256) k_b_translate(int R, int nk, int nb, int nbb, const int* __restrict__ far_ptr,
                                                     const int* __restrict__ far_box, const int* __restrict__ far_tr,
                                                     const float2* __restrict__ T, const float* __restrict__ pes,
                                                     const float2* __restrict__ C, float2* __restrict__ B) {
  const int w = blockIdx.x;
  const int g = w / 8, x = w % 8;
  const int k = (g / nbb) * 8 + x, bb = g % nbb;
  if (k >= nk) return;
  const int per = blockDim.x / (2 * R);  // boxes per workgroup
  const int b = bb * per + threadIdx.x / (2 * R);
  if (b >= nb) return;
  const int pr = threadIdx.x % (2 * R);
  const float2* Ck = C + (size_t)k * nb * 2 * R;
  float2 acc = make_float2(0.f, 0.f);
  for (int m = far_ptr[b]; m < far_ptr[b + 1]; ++m)
    acc = fadd(acc, fmul(Ck[(size_t)far_box[m] * 2 * R + pr], T[(size_t)far_tr[m] * nk + k]));
  const float pk = pes[k];
  B[((size_t)k * nb + b) * 2 * R + pr] = make_float2(acc.x * pk, acc.y * pk);
}

// Downward pass: A_n[s][c] = B_n[s][c] + sum_{(i,w) in W^T(s)} w conj(ph(oct_c, i)) A_{n-1}[i][P]
__global__ void __launch_bounds__(256) k_b_interp_down(int R, int nkc, int nbc, int nkf, int nbf,
                                                       const int* __restrict__ parent, const int* __restrict__ oct,
                                                       const int* __restrict__ wt_ptr, const int* __restrict__ wt_row,
                                                       const float* __restrict__ wt_w, const float2* __restrict__ ph,
                                                       const float2* __restrict__ Ac, const float2* __restrict__ Bf,
                                                       float2* __restrict__ Af) {
  const int c = blockIdx.x;
  const long long t = (long long)blockIdx.y * blockDim.x + threadIdx.x;
  if (t >= (long long)nkf * 2 * R) return;
  const int r = (int)(t % R), p = (int)((t / R) % 2), s = (int)(t / (2 * R));
  const int P = parent[c], o = oct[c];
  float2 acc = make_float2(0.f, 0.f);
  for (int e = wt_ptr[s]; e < wt_ptr[s + 1]; ++e) {
    const int i = wt_row[e];
    const float2 v = fmulc(ph[(size_t)o * nkc + i], Ac[(((size_t)i * nbc + P) * 2 + p) * R + r]);
    acc.x += wt_w[e] * v.x;
    acc.y += wt_w[e] * v.y;
  }
  const size_t at = (((size_t)s * nbf + c) * 2 + p) * R + r;
  Af[at] = fadd(Bf[at], acc);
}

// Receive, one workgroup per leaf box: Yf[basis(s)][r] = sum_(p,k) conj(rad(s;p,k)) A_L[k][b][p][r]
__global__ void __launch_bounds__(256) k_b_receive(int R, int nk, int nb, const int* __restrict__ box_slot,
                                                   const int* __restrict__ slot_basis,
                                                   const float2* __restrict__ rad, const float2* __restrict__ AL,
                                                   float2* __restrict__ Yf) {
  constexpr int CH = 32;
  extern __shared__ float2 la[];  // (CH, R)
  const int b = blockIdx.x;
  const int s0 = box_slot[b], nf = box_slot[b + 1] - s0;
  const int nk2 = 2 * nk;
  const int r = threadIdx.x % R, per = blockDim.x / R;
  const int sbase = blockIdx.y * 8 * per;  // this workgroup's slice of the box's bases
  if (sbase >= nf) return;
  float2 acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = make_float2(0.f, 0.f);
  for (int base = 0; base < nk2; base += CH) {
    __syncthreads();
    for (int e = threadIdx.x; e < CH * R; e += blockDim.x) {
      const int item = base + e / R;
      if (item < nk2) {
        const int p = item / nk, k = item % nk;
        la[e] = AL[(((size_t)k * nb + b) * 2 + p) * R + (e % R)];
      } else {
        la[e] = make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int s = sbase + threadIdx.x / R + q * per;
      if (s < nf) {
        const float2* rs = rad + (size_t)(s0 + s) * nk2 + base;
        const int lim = min(CH, nk2 - base);
        for (int j = 0; j < lim; ++j) acc[q] = fadd(acc[q], fmulc(rs[j], la[j * R + r]));
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int s = sbase + threadIdx.x / R + q * per;
    if (s < nf) Yf[(size_t)slot_basis[s0 + s] * R + r] = acc[q];
  }
}

// Near field (FP64 accumulation of complex(4) values) plus the scaled far field.
template <typename V>
__global__ void __launch_bounds__(256) k_b_near(int R, int ne, const int* __restrict__ ia, const int* __restrict__ ja,
                                                const V* __restrict__ z, const double2* __restrict__ X,
                                                const float2* __restrict__ Yf, double2 cc, double2* __restrict__ Y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int m = (int)(t / R), r = (int)(t % R);
  if (m >= ne) return;
  double2 acc = make_double2(0.0, 0.0);
  for (int k = ia[m]; k < ia[m + 1]; ++k) acc = dadd(acc, dmul(to_d2(z[k]), X[(size_t)ja[k] * R + r]));
  const double2 f = dmul(to_d2(Yf[(size_t)m * R + r]), cc);
  Y[(size_t)m * R + r] = dadd(f, acc);
}

// ---------------------------------------------------------------- ILU0 solves (FP64)

// lusol0C for R columns: one workgroup; a row is handled by R lanes.
template <typename V>
__global__ void __launch_bounds__(1024) k_b_lusol(int R, const double2* __restrict__ y, double2* __restrict__ x,
                                                  const V* __restrict__ alu, const int* __restrict__ jlu,
                                                  const int* __restrict__ ju, const int* __restrict__ lptr,
                                                  const int* __restrict__ lrows, int nlev_l,
                                                  const int* __restrict__ uptr, const int* __restrict__ urows,
                                                  int nlev_u) {
  const int slot = threadIdx.x / R, r = threadIdx.x % R, nslot = blockDim.x / R;
  for (int l = 0; l < nlev_l; ++l) {
    for (int q = lptr[l] + slot; q < lptr[l + 1]; q += nslot) {
      const int i = lrows[q] + 1;
      double2 s = y[(size_t)(i - 1) * R + r];
      for (int k = jlu[i]; k <= ju[i] - 1; ++k) s = dsub(s, dmul(to_d2(alu[k]), x[(size_t)(jlu[k] - 1) * R + r]));
      x[(size_t)(i - 1) * R + r] = s;
    }
    __syncthreads();
  }
  for (int l = 0; l < nlev_u; ++l) {
    for (int q = uptr[l] + slot; q < uptr[l + 1]; q += nslot) {
      const int i = urows[q] + 1;
      double2 s = x[(size_t)(i - 1) * R + r];
      for (int k = ju[i]; k <= jlu[i + 1] - 1; ++k) s = dsub(s, dmul(to_d2(alu[k]), x[(size_t)(jlu[k] - 1) * R + r]));
      x[(size_t)(i - 1) * R + r] = dmul(to_d2(alu[i]), s);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- CGS pieces (FP64, per column)

constexpr int kRedBlocks = 256;

// part[(blk*R + r)*4 + {0,1,2,3}]: per-column sums of a*b and a2*b2
__global__ void __launch_bounds__(256) k_b_dot2(int R, int n, const double2* __restrict__ a,
                                                const double2* __restrict__ b, int conj_a,
                                                const double2* __restrict__ a2, const double2* __restrict__ b2,
                                                int conj_a2, double* __restrict__ part) {
  extern __shared__ double2 sh[];  // (256) x 2
  const int r = threadIdx.x % R, rows_per = blockDim.x / R;
  double2 s1 = make_double2(0, 0), s2 = make_double2(0, 0);
  for (long long i = (long long)blockIdx.x * rows_per + threadIdx.x / R; i < n; i += (long long)gridDim.x * rows_per) {
    const double2 av = a[i * R + r], bv = b[i * R + r];
    s1 = dadd(s1, conj_a ? make_double2(av.x * bv.x + av.y * bv.y, av.x * bv.y - av.y * bv.x) : dmul(av, bv));
    if (a2) {
      const double2 cv = a2[i * R + r], dv = b2[i * R + r];
      s2 = dadd(s2, conj_a2 ? make_double2(cv.x * dv.x + cv.y * dv.y, cv.x * dv.y - cv.y * dv.x) : dmul(cv, dv));
    }
  }
  sh[threadIdx.x] = s1;
  sh[blockDim.x + threadIdx.x] = s2;
  __syncthreads();
  if (threadIdx.x < R) {
    double2 t1 = make_double2(0, 0), t2 = make_double2(0, 0);
    for (int q = 0; q < rows_per; ++q) {
      t1 = dadd(t1, sh[q * R + threadIdx.x]);
      t2 = dadd(t2, sh[blockDim.x + q * R + threadIdx.x]);
    }
    double* o = part + ((size_t)blockIdx.x * R + threadIdx.x) * 4;
    o[0] = t1.x;
    o[1] = t1.y;
    o[2] = t2.x;
    o[3] = t2.y;
  }
}

__global__ void k_b_dot_finish(int R, int nblk, const double* __restrict__ part, double2* __restrict__ out1,
                               double2* __restrict__ out2) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double2 s1 = make_double2(0, 0), s2 = make_double2(0, 0);
  for (int b = 0; b < nblk; ++b) {
    const double* o = part + ((size_t)b * R + r) * 4;
    s1 = dadd(s1, make_double2(o[0], o[1]));
    s2 = dadd(s2, make_double2(o[2], o[3]));
  }
  out1[r] = s1;
  if (out2) out2[r] = s2;
}

__global__ void k_b_cgs_up(int R, int n, const int* __restrict__ active, const int* __restrict__ first,
                           const double2* __restrict__ rho1, const double2* __restrict__ rho2,
                           const double2* __restrict__ rr, const double2* __restrict__ q, double2* __restrict__ u,
                           double2* __restrict__ p) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * R) return;
  const int r = (int)(t % R);
  if (!active[r]) return;
  if (first[r]) {
    u[t] = rr[t];
    p[t] = rr[t];
    return;
  }
  const double2 beta = ddiv(rho1[r], rho2[r]);
  const double2 ui = dadd(rr[t], dmul(beta, q[t]));
  u[t] = ui;
  p[t] = dadd(ui, dmul(beta, dadd(q[t], dmul(beta, p[t]))));
}

__global__ void k_b_cgs_q(int R, int n, const int* __restrict__ active, const double2* __restrict__ rho1,
                          const double2* __restrict__ den, double2* __restrict__ alfa_out,
                          const double2* __restrict__ u, const double2* __restrict__ V, double2* __restrict__ q,
                          double2* __restrict__ tv) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * R) return;
  const int r = (int)(t % R);
  if (!active[r]) return;
  const double2 alfa = ddiv(rho1[r], den[r]);
  if (t < R) alfa_out[r] = alfa;
  const double2 qi = dsub(u[t], dmul(alfa, V[t]));
  q[t] = qi;
  tv[t] = dadd(u[t], qi);
}

__global__ void k_b_cgs_xr(int R, int n, const int* __restrict__ active, const double2* __restrict__ rho1,
                           const double2* __restrict__ den, const double2* __restrict__ U,
                           const double2* __restrict__ Q, double2* __restrict__ x, double2* __restrict__ rr) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * R) return;
  const int r = (int)(t % R);
  if (!active[r]) return;
  const double2 a = ddiv(rho1[r], den[r]);
  x[t] = dadd(x[t], dmul(a, U[t]));
  rr[t] = dsub(rr[t], dmul(a, Q[t]));
}

__global__ void k_b_sub(long long n, const double2* __restrict__ a, const double2* __restrict__ b,
                        double2* __restrict__ o) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = dsub(a[i], b[i]);
}

__global__ void k_b_copy_masked(int R, int n, const int* __restrict__ active, const double2* __restrict__ a,
                                double2* __restrict__ o) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * R) return;
  if (active[t % R]) o[t] = a[t];
}

// ---------------------------------------------------------------- excitation and RCS per column

__global__ void k_b_plane_wave(int R, int ne, const double* __restrict__ tb, const double* __restrict__ pc,
                               const int* __restrict__ et, const int* __restrict__ epo,
                               const double* __restrict__ el, const double* __restrict__ dirs,
                               const double2* __restrict__ pols, double2 ctte, double k0, double2* __restrict__ B) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ne * R) return;
  const int i = (int)(t / R), r = (int)(t % R);
  const double* d = dirs + 3 * r;
  const double2 p0 = dmul(pols[3 * r], ctte), p1 = dmul(pols[3 * r + 1], ctte), p2 = dmul(pols[3 * r + 2], ctte);
  double2 acc = make_double2(0, 0);
  for (int side = 0; side < 2; ++side) {
    const double* c = tb + 3 * (et[2 * i + side] - 1);
    const double* v = pc + 3 * (epo[2 * i + side] - 1);
    double sn, cs;
    sincos(-(k0 * ((d[0] * c[0] + d[1] * c[1]) + d[2] * c[2])), &sn, &cs);
    const double2 ph = make_double2(cs, sn);
    const double sg = side == 0 ? 1.0 : -1.0;
    const double d0 = sg * (c[0] - v[0]), d1 = sg * (c[1] - v[1]), d2 = sg * (c[2] - v[2]);
    const double2 e0 = dmul(p0, ph), e1 = dmul(p1, ph), e2 = dmul(p2, ph);
    acc.x += (e0.x * d0 + e1.x * d1) + e2.x * d2;
    acc.y += (e0.y * d0 + e1.y * d1) + e2.y * d2;
  }
  const double f = 0.5 * el[i];
  B[t] = make_double2(acc.x * f, acc.y * f);
}

__global__ void __launch_bounds__(256) k_b_rcs(int R, int ne, const double* __restrict__ tb,
                                               const int* __restrict__ et, const double* __restrict__ el,
                                               const double2* __restrict__ cur, const double* __restrict__ robs,
                                               double kfs, double2* __restrict__ Eout) {
  __shared__ double2 sh[3][256 / kWave];
  const int r = blockIdx.x;
  const double* ro = robs + 3 * r;
  double2 E0 = make_double2(0, 0), E1 = make_double2(0, 0), E2 = make_double2(0, 0);
  for (int i = threadIdx.x; i < ne; i += blockDim.x) {
    const double* t1 = tb + 3 * (et[2 * i] - 1);
    const double* t2 = tb + 3 * (et[2 * i + 1] - 1);
    const double2 cv = cur[(size_t)i * R + r];
    const double2 Il = make_double2(cv.x * el[i], cv.y * el[i]);
    double rr[3], dm[3];
    for (int d = 0; d < 3; ++d) {
      const double rm = 0.5 * (t1[d] + t2[d]);
      dm[d] = t2[d] - t1[d];
      rr[d] = ro[d] - rm;
    }
    const double rr2 = (rr[0] * rr[0] + rr[1] * rr[1]) + rr[2] * rr[2];
    const double dr = (dm[0] * rr[0] + dm[1] * rr[1]) + dm[2] * rr[2];
    const double rn = sqrt(rr2);
    double sn, cs;
    sincos(-(kfs * rn), &sn, &cs);
    const double2 gI = dmul(make_double2(cs / rn, sn / rn), Il);
    const double c0 = dr * rr[0] / rr2 - dm[0], c1 = dr * rr[1] / rr2 - dm[1], c2 = dr * rr[2] / rr2 - dm[2];
    E0.x += gI.x * c0; E0.y += gI.y * c0;
    E1.x += gI.x * c1; E1.y += gI.y * c1;
    E2.x += gI.x * c2; E2.y += gI.y * c2;
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    E0.x += __shfl_xor(E0.x, o, kWave); E0.y += __shfl_xor(E0.y, o, kWave);
    E1.x += __shfl_xor(E1.x, o, kWave); E1.y += __shfl_xor(E1.y, o, kWave);
    E2.x += __shfl_xor(E2.x, o, kWave); E2.y += __shfl_xor(E2.y, o, kWave);
  }
  const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
  if (ln == 0) {
    sh[0][wv] = E0;
    sh[1][wv] = E1;
    sh[2][wv] = E2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double2 s = make_double2(0, 0);
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s = dadd(s, sh[threadIdx.x][w]);
    Eout[3 * r + threadIdx.x] = s;
  }
}

int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------- state

struct BatchLevel {
  int nk = 0, nb = 0;
  float2* T = nullptr;
  float* pes = nullptr;
  float* wint = nullptr;  // interpolation to this level (n < L)
  float* wt_w = nullptr;
  float2* ph = nullptr;   // (8, nk) child-offset phases at this level's samples
  int* oct = nullptr;     // octant of each box within its parent (n > 3)
  float2 *C = nullptr, *B = nullptr, *A = nullptr;
};

struct Batch {
  Handle* h = nullptr;
  int R = 0;
  float2* rad32 = nullptr;
  std::vector<BatchLevel> lv;
  double2 *X = nullptr, *Y = nullptr, *w[12] = {};
  float2* Yf = nullptr;
  double* part = nullptr;
  double2* sc = nullptr;  // 8 per-column scalar rows of R
  int *active = nullptr, *first = nullptr;
  std::vector<void*> allocs;
};

template <typename T>
static int bmalloc(Batch* b, T** p, size_t n) {
  void* q = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&q, n * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return LW_ERR_NOMEM;
  }
  b->allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return LW_OK;
}

static void free_batch(Batch* b) {
  if (!b) return;
  if (b->h && b->h->stream) (void)hipStreamSynchronize(b->h->stream);
  for (void* p : b->allocs) (void)hipFree(p);
  delete b;
}

static int batch_matvec(Batch* bt, const double2* X, double2* Y) {
  Handle* h = bt->h;
  const int R = bt->R, L = h->L;
  hipStream_t st = h->stream;
  BatchLevel& lf = bt->lv[L];
  int maxnf = 0;
  for (int b = 0; b < lf.nb; ++b) maxnf = std::max(maxnf, h->plan.box_slot[b + 1] - h->plan.box_slot[b]);
  {
    const int per = 256 / R;
    dim3 grid(lf.nb, std::max(1, std::min(16, nblk(2 * lf.nk, per))));
    hipLaunchKernelGGL(k_b_aggregate, grid, dim3(256), maxnf * R * sizeof(float2), st, R, lf.nk, lf.nb,
                       h->d_box_slot, h->d_slot_basis, bt->rad32, X, lf.C);
  }
  for (int n = L; n >= 4; --n) {
    BatchLevel& c = bt->lv[n];
    BatchLevel& p = bt->lv[n - 1];
    dim3 grid(p.nb, nblk((long long)p.nk * 2 * R, 256));
    hipLaunchKernelGGL(k_b_interp_up, grid, dim3(256), 0, st, R, p.nk, p.nb, c.nk, c.nb, h->lv[n - 1].child_ptr, c.oct,
                       p.wint, h->lv[n - 1].iint, p.ph, c.C, p.C);
  }
  for (int n = 3; n <= L; ++n) {
    BatchLevel& d = bt->lv[n];
    const int per = 256 / (2 * R);
    const int nbb = nblk(d.nb, per);
    const int groups = nblk(d.nk, 8);
    hipLaunchKernelGGL(k_b_translate, dim3(groups * 8 * nbb), dim3(256), 0, st, R, d.nk, d.nb, nbb, h->lv[n].far_ptr,
                       h->lv[n].far_box, h->lv[n].far_tr, d.T, d.pes, d.C, d.B);
  }
  for (int n = 4; n <= L; ++n) {
    BatchLevel& c = bt->lv[n];
    BatchLevel& p = bt->lv[n - 1];
    const float2* src = (n == 4) ? p.B : p.A;
    dim3 grid(c.nb, nblk((long long)c.nk * 2 * R, 256));
    hipLaunchKernelGGL(k_b_interp_down, grid, dim3(256), 0, st, R, p.nk, p.nb, c.nk, c.nb, h->lv[n].parent, c.oct,
                       h->lv[n - 1].wt_ptr, h->lv[n - 1].wt_row, p.wt_w, p.ph, src, c.B, c.A);
  }
  hipLaunchKernelGGL(k_b_receive, dim3(lf.nb, nblk(maxnf, 8 * (256 / R))), dim3(256), 32 * R * sizeof(float2), st,
                     R, lf.nk, lf.nb,
                     h->d_box_slot, h->d_slot_basis, bt->rad32, lf.A, bt->Yf);
  const double2 cc = make_double2(0.0, h->plan.phys.w * kMu / (4.0 * kPi));
  const long long tot = (long long)h->ne * R;
  if (h->fp64_near)
    hipLaunchKernelGGL(k_b_near<double2>, dim3(nblk(tot, 256)), dim3(256), 0, st, R, h->ne, h->d_ia, h->d_ja,
                       (const double2*)h->d_z, X, bt->Yf, cc, Y);
  else
    hipLaunchKernelGGL(k_b_near<float2>, dim3(nblk(tot, 256)), dim3(256), 0, st, R, h->ne, h->d_ia, h->d_ja,
                       (const float2*)h->d_z, X, bt->Yf, cc, Y);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

static int batch_lusol(Batch* bt, const double2* y, double2* x) {
  Handle* h = bt->h;
  if (!h->have_ilu) {
    set_error("batch lusol0: no ILU0 factor (call lw_ilu0 first)");
    return LW_ERR_STATE;
  }
  const int R = bt->R;
  const int threads = std::min(1024, (1024 / R) * R);
  if (h->fp64_near)
    hipLaunchKernelGGL(k_b_lusol<double2>, dim3(1), dim3(threads), 0, h->stream, R, y, x, (const double2*)h->d_alu,
                       h->d_jlu, h->d_ju, h->d_lptr, h->d_lrows, h->nlev_l, h->d_uptr, h->d_urows, h->nlev_u);
  else
    hipLaunchKernelGGL(k_b_lusol<float2>, dim3(1), dim3(threads), 0, h->stream, R, y, x, (const float2*)h->d_alu,
                       h->d_jlu, h->d_ju, h->d_lptr, h->d_lrows, h->nlev_l, h->d_uptr, h->d_urows, h->nlev_u);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

static int batch_dot2(Batch* bt, const double2* a, const double2* b, int ca, const double2* a2, const double2* b2,
                      int ca2, double2* out1, double2* out2) {
  Handle* h = bt->h;
  const int R = bt->R;
  hipLaunchKernelGGL(k_b_dot2, dim3(kRedBlocks), dim3(256), 2 * 256 * sizeof(double2), h->stream, R, h->ne, a, b, ca,
                     a2, b2, ca2, bt->part);
  hipLaunchKernelGGL(k_b_dot_finish, dim3(1), dim3(64), 0, h->stream, R, kRedBlocks, bt->part, out1, out2);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

}  // namespace lw

struct lw_batch {
  lw::Batch b;
};

extern "C" {

int lw_batch_create(lw_handle* hh, int nrhs, lw_batch** out) {
  using namespace lw;
  if (!hh || !out || !(nrhs == 8 || nrhs == 16 || nrhs == 32 || nrhs == 64)) {
    set_error("lw_batch_create: nrhs must be 8, 16, 32 or 64");
    return LW_ERR_ARG;
  }
  *out = nullptr;
  Handle* h = &hh->h;
  lw_batch* ob = new lw_batch();
  Batch* bt = &ob->b;
  bt->h = h;
  bt->R = nrhs;
  const int L = h->L, R = nrhs;
  const Plan& P = h->plan;
  int rc;
#define TRY(x)                 \
  do {                         \
    rc = (x);                  \
    if (rc) {                  \
      free_batch(bt);          \
      return rc;               \
    }                          \
  } while (0)
  hipStream_t st = h->stream;
  const long long nrad = (long long)P.lv[L].nk * 2 * h->ne;
  TRY(bmalloc(bt, &bt->rad32, (size_t)nrad));
  hipLaunchKernelGGL(k_to_f32, dim3(nblk(nrad, 256)), dim3(256), 0, st, nrad, h->d_rad, bt->rad32);
  bt->lv.assign(L + 1, BatchLevel());
  for (int n = 3; n <= L; ++n) {
    const LevelPlan& lp = P.lv[n];
    DevLevel& d = h->lv[n];
    BatchLevel& b = bt->lv[n];
    b.nk = lp.nk;
    b.nb = lp.nb;
    const long long nT = (long long)lp.nk * std::max(lp.nT, 1);
    TRY(bmalloc(bt, &b.T, (size_t)nT));
    hipLaunchKernelGGL(k_to_f32, dim3(nblk(nT, 256)), dim3(256), 0, st, nT, d.T, b.T);
    TRY(bmalloc(bt, &b.pes, (size_t)lp.nk));
    hipLaunchKernelGGL(k_to_f32r, dim3(nblk(lp.nk, 256)), dim3(256), 0, st, (long long)lp.nk, d.pes, b.pes);
    if (n < L) {
      TRY(bmalloc(bt, &b.wint, 9 * (size_t)lp.nk));
      hipLaunchKernelGGL(k_to_f32r, dim3(nblk(9LL * lp.nk, 256)), dim3(256), 0, st, 9LL * lp.nk, d.wint, b.wint);
      TRY(bmalloc(bt, &b.wt_w, lp.wt_w.size()));
      hipLaunchKernelGGL(k_to_f32r, dim3(nblk((long long)lp.wt_w.size(), 256)), dim3(256), 0, st,
                         (long long)lp.wt_w.size(), d.wt_w, b.wt_w);
      // child offset phases at this level's samples (the FP64 table, rounded)
      TRY(bmalloc(bt, &b.ph, 8 * (size_t)lp.nk));
      hipLaunchKernelGGL(k_to_f32, dim3(nblk(8LL * lp.nk, 256)), dim3(256), 0, st, 8LL * lp.nk, d.ph, b.ph);
    }
    b.oct = d.oct;
    const size_t fsz = (size_t)lp.nk * lp.nb * 2 * R;
    TRY(bmalloc(bt, &b.C, fsz));
    TRY(bmalloc(bt, &b.B, fsz));
    TRY(bmalloc(bt, &b.A, fsz));
  }
  const size_t nv = (size_t)h->ne * R;
  TRY(bmalloc(bt, &bt->X, nv));
  TRY(bmalloc(bt, &bt->Y, nv));
  for (auto& w : bt->w) TRY(bmalloc(bt, &w, nv));
  TRY(bmalloc(bt, &bt->Yf, nv));
  TRY(bmalloc(bt, &bt->part, (size_t)kRedBlocks * R * 4));
  TRY(bmalloc(bt, &bt->sc, (size_t)8 * R));
  TRY(bmalloc(bt, &bt->active, (size_t)R));
  TRY(bmalloc(bt, &bt->first, (size_t)R));
#undef TRY
  if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess) {
    set_error("lw_batch_create: conversion kernels failed");
    free_batch(bt);
    return LW_ERR_HIP;
  }
  *out = ob;
  return LW_OK;
}

int lw_batch_destroy(lw_batch* b) {
  if (b) {
    lw::Batch* bt = &b->b;
    for (void* p : bt->allocs) (void)hipFree(p);
    bt->allocs.clear();
    delete b;
  }
  return LW_OK;
}

int lw_batch_nrhs(lw_batch* b) { return b ? b->b.R : LW_ERR_ARG; }

int lw_batch_matvec(lw_batch* ob, const double* X, double* Y) {
  if (!ob || !X || !Y) return LW_ERR_ARG;
  lw::Batch* bt = &ob->b;
  lw::Handle* h = bt->h;
  const size_t bytes = sizeof(double2) * (size_t)h->ne * bt->R;
  LW_HIP(hipMemcpy(bt->X, X, bytes, hipMemcpyHostToDevice));
  int rc = lw::batch_matvec(bt, bt->X, bt->Y);
  if (rc) return rc;
  LW_HIP(hipMemcpyAsync(Y, bt->Y, bytes, hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  return LW_OK;
}

int lw_batch_lusol0(lw_batch* ob, const double* Yv, double* Xv) {
  if (!ob || !Yv || !Xv) return LW_ERR_ARG;
  lw::Batch* bt = &ob->b;
  lw::Handle* h = bt->h;
  const size_t bytes = sizeof(double2) * (size_t)h->ne * bt->R;
  LW_HIP(hipMemcpy(bt->X, Yv, bytes, hipMemcpyHostToDevice));
  int rc = lw::batch_lusol(bt, bt->X, bt->Y);
  if (rc) return rc;
  LW_HIP(hipMemcpyAsync(Xv, bt->Y, bytes, hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  return LW_OK;
}

int lw_batch_cgs(lw_batch* ob, const double* Bv, const double* guess, double errmax, int use_prec, int maxit,
                 double* Xv, int* iters, double* relres, int* converged) {
  using namespace lw;
  if (!ob || !Bv || !Xv || maxit < 1) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  Handle* h = bt->h;
  if (use_prec && !h->have_ilu) {
    set_error("lw_batch_cgs: preconditioner requested but no ILU0 factor");
    return LW_ERR_STATE;
  }
  const int R = bt->R, n = h->ne;
  const size_t nv = (size_t)n * R, bytes = sizeof(double2) * nv;
  hipStream_t st = h->stream;
  double2 *b = bt->w[0], *r = bt->w[1], *R_ = bt->w[2], *u = bt->w[3], *p = bt->w[4], *P_ = bt->w[5],
          *V = bt->w[6], *q = bt->w[7], *U = bt->w[8], *Q = bt->w[9], *t = bt->w[10], *x = bt->X;
  double2 *rho1 = bt->sc, *rho2 = bt->sc + R, *den = bt->sc + 2 * R, *alfa = bt->sc + 3 * R, *nrm = bt->sc + 4 * R,
          *rhon = bt->sc + 5 * R;
  LW_HIP(hipMemcpy(b, Bv, bytes, hipMemcpyHostToDevice));
  if (guess)
    LW_HIP(hipMemcpy(x, guess, bytes, hipMemcpyHostToDevice));
  else
    LW_HIP(hipMemset(x, 0, bytes));
  int rc = batch_matvec(bt, x, t);
  if (rc) return rc;
  const int vb = nblk((long long)nv, 256);
  hipLaunchKernelGGL(k_b_sub, dim3(vb), dim3(256), 0, st, (long long)nv, b, t, r);
  LW_HIP(hipMemcpyAsync(R_, r, bytes, hipMemcpyDeviceToDevice, st));
  if ((rc = batch_dot2(bt, b, b, 1, r, R_, 0, nrm, rhon))) return rc;
  std::vector<double2> hn(R), hr(R);
  LW_HIP(hipMemcpyAsync(hn.data(), nrm, sizeof(double2) * R, hipMemcpyDeviceToHost, st));
  LW_HIP(hipMemcpyAsync(hr.data(), rhon, sizeof(double2) * R, hipMemcpyDeviceToHost, st));
  LW_HIP(hipStreamSynchronize(st));
  std::vector<double> nb(R);
  for (int c = 0; c < R; ++c) nb[c] = std::sqrt(hn[c].x);
  LW_HIP(hipMemcpyAsync(rho1, rhon, sizeof(double2) * R, hipMemcpyDeviceToDevice, st));
  std::vector<int> act(R, 1), fst(R, 1), it(R, 0), conv(R, 0);
  std::vector<double> rr(R, 0.0);
  for (int c = 0; c < R; ++c)
    if (hr[c].x == 0.0 && hr[c].y == 0.0) act[c] = 0;
  for (int k = 0; k < maxit; ++k) {
    int nact = 0;
    for (int c = 0; c < R; ++c) nact += act[c];
    if (!nact) break;
    for (int c = 0; c < R; ++c)
      if (act[c]) it[c] += 1;
    LW_HIP(hipMemcpyAsync(bt->active, act.data(), sizeof(int) * R, hipMemcpyHostToDevice, st));
    LW_HIP(hipMemcpyAsync(bt->first, fst.data(), sizeof(int) * R, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_b_cgs_up, dim3(vb), dim3(256), 0, st, R, n, bt->active, bt->first, rho1, rho2, r, q, u, p);
    if (use_prec) {
      if ((rc = batch_lusol(bt, p, P_))) return rc;
    } else {
      hipLaunchKernelGGL(k_b_copy_masked, dim3(vb), dim3(256), 0, st, R, n, bt->active, p, P_);
    }
    if ((rc = batch_matvec(bt, P_, V))) return rc;
    if ((rc = batch_dot2(bt, R_, V, 0, nullptr, nullptr, 0, den, nullptr))) return rc;
    hipLaunchKernelGGL(k_b_cgs_q, dim3(vb), dim3(256), 0, st, R, n, bt->active, rho1, den, alfa, u, V, q, t);
    if (use_prec) {
      if ((rc = batch_lusol(bt, t, U))) return rc;
    } else {
      hipLaunchKernelGGL(k_b_copy_masked, dim3(vb), dim3(256), 0, st, R, n, bt->active, t, U);
    }
    if ((rc = batch_matvec(bt, U, Q))) return rc;
    hipLaunchKernelGGL(k_b_cgs_xr, dim3(vb), dim3(256), 0, st, R, n, bt->active, rho1, den, U, Q, x, r);
    if ((rc = batch_dot2(bt, r, r, 1, r, R_, 0, nrm, rhon))) return rc;
    LW_HIP(hipMemcpyAsync(hn.data(), nrm, sizeof(double2) * R, hipMemcpyDeviceToHost, st));
    LW_HIP(hipMemcpyAsync(hr.data(), rhon, sizeof(double2) * R, hipMemcpyDeviceToHost, st));
    LW_HIP(hipStreamSynchronize(st));
    for (int c = 0; c < R; ++c) {
      if (!act[c]) continue;
      const double nr = std::sqrt(hn[c].x);
      rr[c] = nr / nb[c];
      fst[c] = 0;
      if (nr < errmax * nb[c]) {
        conv[c] = 1;
        act[c] = 0;
      } else if (hr[c].x == 0.0 && hr[c].y == 0.0) {
        act[c] = 0;  // breakdown exit of the next iteration
      }
    }
    if (progress_every() > 0 && (k + 1) % progress_every() == 0) {
      int na = 0;
      double worst = 0.0;
      for (int c = 0; c < R; ++c) {
        na += act[c];
        if (act[c]) worst = std::max(worst, rr[c]);
      }
      std::fprintf(stderr, "[lw_batch_cgs] iteration %d: %d of %d columns active, worst relres %.3e\n", k + 1, na, R,
                   worst);
    }
    // rho2 = rho1; rho1 = next (columns that stopped no longer read them)
    LW_HIP(hipMemcpyAsync(rho2, rho1, sizeof(double2) * R, hipMemcpyDeviceToDevice, st));
    LW_HIP(hipMemcpyAsync(rho1, rhon, sizeof(double2) * R, hipMemcpyDeviceToDevice, st));
  }
  LW_HIP(hipMemcpyAsync(Xv, x, bytes, hipMemcpyDeviceToHost, st));
  LW_HIP(hipStreamSynchronize(st));
  for (int c = 0; c < R; ++c) {
    if (iters) iters[c] = it[c];
    if (relres) relres[c] = rr[c];
    if (converged) converged[c] = conv[c];
  }
  return LW_OK;
}

int lw_batch_plane_wave(lw_batch* ob, const double* dirs, const double* pols, const double* ctte, double* Bv) {
  using namespace lw;
  if (!ob || !dirs || !pols || !ctte || !Bv) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  Handle* h = bt->h;
  const int R = bt->R;
  std::vector<double> d(3 * R);
  for (int c = 0; c < R; ++c) {
    const double* v = dirs + 3 * c;
    const double nd = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    if (!(nd > 0)) return LW_ERR_ARG;
    for (int k = 0; k < 3; ++k) d[3 * c + k] = v[k] / nd;
  }
  double* d_d = nullptr;
  double2* d_p = nullptr;
  LW_HIP(hipMalloc(&d_d, sizeof(double) * d.size()));
  LW_HIP(hipMalloc(&d_p, sizeof(double2) * 3 * R));
  LW_HIP(hipMemcpy(d_d, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice));
  LW_HIP(hipMemcpy(d_p, pols, sizeof(double2) * 3 * R, hipMemcpyHostToDevice));
  const long long tot = (long long)h->ne * R;
  hipLaunchKernelGGL(k_b_plane_wave, dim3(nblk(tot, 256)), dim3(256), 0, h->stream, R, h->ne, h->d_tb, h->d_p,
                     h->d_et, h->d_epo, h->d_el, d_d, d_p, make_double2(ctte[0], ctte[1]),
                     h->plan.phys.w * std::sqrt(kMu * kEps), bt->Y);
  LW_HIP(hipGetLastError());
  LW_HIP(hipMemcpyAsync(Bv, bt->Y, sizeof(double2) * tot, hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  (void)hipFree(d_d);
  (void)hipFree(d_p);
  return LW_OK;
}

int lw_batch_rcs(lw_batch* ob, const double* cur, const double* pols, const double* ctte, const double* theta,
                 const double* phi, double* sigma) {
  using namespace lw;
  if (!ob || !cur || !pols || !ctte || !theta || !phi || !sigma) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  Handle* h = bt->h;
  const int R = bt->R;
  const Plan& P = h->plan;
  // far-zone distance and centre from the edge-centre bounding box (modRCS.f90:120-147)
  std::vector<double> ec(3 * (size_t)h->ne);
  LW_HIP(hipMemcpy(ec.data(), h->d_ec, sizeof(double) * ec.size(), hipMemcpyDeviceToHost));
  int lim[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 1; i < h->ne; ++i) {
    const double* c = &ec[3 * (size_t)i];
    if (c[0] > ec[3 * lim[0]]) lim[0] = i;
    if (c[0] < ec[3 * lim[1]]) lim[1] = i;
    if (c[1] > ec[3 * lim[2] + 1]) lim[2] = i;
    if (c[1] < ec[3 * lim[3] + 1]) lim[3] = i;
    if (c[2] > ec[3 * lim[4] + 2]) lim[4] = i;
    if (c[2] < ec[3 * lim[5] + 2]) lim[5] = i;
  }
  const double mx = ec[3 * lim[0]], mnx = ec[3 * lim[1]], my = ec[3 * lim[2] + 1], mny = ec[3 * lim[3] + 1],
               mz = ec[3 * lim[4] + 2], mnz = ec[3 * lim[5] + 2];
  const double w = P.phys.w, kfs = w * std::sqrt(kMu * kEps);
  const double lam = 2 * kPi / kfs;
  const double D = std::sqrt(((mx - mnx) * (mx - mnx) + (my - mny) * (my - mny)) + (mz - mnz) * (mz - mnz));
  double rzl = 2 * (D * D) / lam;
  if (50 * D > rzl) rzl = 50 * D;
  if (20 * lam > rzl) rzl = 20 * lam;
  rzl = rzl * 10;
  const double roo[3] = {0.5 * (mx + mnx), 0.5 * (my + mny), 0.5 * (mz + mnz)};
  std::vector<double> ro(3 * (size_t)R);
  for (int c = 0; c < R; ++c) {
    ro[3 * c + 0] = roo[0] + (rzl * std::sin(theta[c]) * std::cos(phi[c]));
    ro[3 * c + 1] = roo[1] + (rzl * std::sin(theta[c]) * std::sin(phi[c]));
    ro[3 * c + 2] = roo[2] + (rzl * std::cos(theta[c]));
  }
  const size_t bytes = sizeof(double2) * (size_t)h->ne * R;
  double* d_ro = nullptr;
  double2* d_E = nullptr;
  LW_HIP(hipMalloc(&d_ro, sizeof(double) * ro.size()));
  LW_HIP(hipMalloc(&d_E, sizeof(double2) * 3 * R));
  LW_HIP(hipMemcpy(d_ro, ro.data(), sizeof(double) * ro.size(), hipMemcpyHostToDevice));
  LW_HIP(hipMemcpy(bt->Y, cur, bytes, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_b_rcs, dim3(R), dim3(256), 0, h->stream, R, h->ne, h->d_tb, h->d_et, h->d_el, bt->Y, d_ro, kfs,
                     d_E);
  std::vector<double2> E(3 * (size_t)R);
  LW_HIP(hipMemcpyAsync(E.data(), d_E, sizeof(double2) * E.size(), hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  (void)hipFree(d_ro);
  (void)hipFree(d_E);
  const double cm = w * kMu / (4 * kPi);
  for (int c = 0; c < R; ++c) {
    double pn = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double pr = pols[6 * c + 2 * k], pi = pols[6 * c + 2 * k + 1];
      const double re = pr * ctte[0] - pi * ctte[1], im = pr * ctte[1] + pi * ctte[0];
      pn += re * re + im * im;
    }
    double e2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double2 v = E[3 * (size_t)c + k];
      e2 += (v.x * v.x + v.y * v.y) * (cm * cm);
    }
    sigma[c] = e2 * 4 * kPi * (rzl * rzl) / pn;
  }
  return LW_OK;
}

int lw_bench_batch(lw_batch* ob, int reps, double* matvec_ms, double* lusol_ms) {
  using namespace lw;
  if (!ob || reps < 1) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  Handle* h = bt->h;
  const size_t nv = (size_t)h->ne * bt->R;
  std::vector<double2> xin(nv);
  for (size_t i = 0; i < nv; ++i) xin[i] = make_double2(1.0 / (1.0 + (double)(i % 977)), 0.5 - (double)(i % 13) / 13);
  LW_HIP(hipMemcpy(bt->X, xin.data(), sizeof(double2) * nv, hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  LW_HIP(hipEventCreate(&e0));
  LW_HIP(hipEventCreate(&e1));
  int rc = batch_matvec(bt, bt->X, bt->Y);
  if (rc) return rc;
  LW_HIP(hipEventRecord(e0, h->stream));
  for (int k = 0; k < reps; ++k)
    if ((rc = batch_matvec(bt, bt->X, bt->Y))) return rc;
  LW_HIP(hipEventRecord(e1, h->stream));
  LW_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  LW_HIP(hipEventElapsedTime(&t, e0, e1));
  if (matvec_ms) *matvec_ms = t / reps;
  if (lusol_ms) {
    *lusol_ms = 0.0;
    if (h->have_ilu) {
      LW_HIP(hipEventRecord(e0, h->stream));
      if ((rc = batch_lusol(bt, bt->Y, bt->w[0]))) return rc;
      LW_HIP(hipEventRecord(e1, h->stream));
      LW_HIP(hipEventSynchronize(e1));
      LW_HIP(hipEventElapsedTime(&t, e0, e1));
      *lusol_ms = t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return LW_OK;
}

}  // extern "C"
