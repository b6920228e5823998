// MLFMA matvec, ILU0 solves, CGS, plane wave and RCS on gfx950, and the C ABI
// declared in include/lywell.h.
//
// Data layout in HBM (complex(8) = double2, complex(4) = float2):
//   campo arrays per level:  F[(box*2 + pol)*nk + sample]      (reference (K,2,B))
//   radiation patterns:      rad[(slot*2 + pol)*nk_L + sample], slots grouped
//                            by leaf box so a box's patterns are contiguous
//   transfer functions:      T[t*nk + sample]
//   near field:              CSR in the reference's row order, float2 values
//   ILU0 factor:             MSR (alu, jlu, ju) exactly as ilu0C leaves it
// Vectors stay in the reference's (natural RWG) order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "lw_internal.h"

namespace lw {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

// LW_PROGRESS=n makes the iterative solvers report every n iterations on
// stderr (long solves otherwise print nothing).
int progress_every() {
  static const int every = [] {
    const char* v = std::getenv("LW_PROGRESS");
    return v ? std::atoi(v) : 0;
  }();
  return every;
}

namespace {

constexpr double kMu = Physics::mu, kEps = Physics::epsi, kPi = Physics::pi;
constexpr int kWave = 64;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }

__device__ __forceinline__ double2 wave_sum(double2 v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, kWave);
    v.y += __shfl_xor(v.y, o, kWave);
  }
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Sum over the 16 lanes of a DPP row (lanes 16q..16q+15), every lane of the
// row receiving it: rotations by 8 and 4 within the row, then the quad
// permutations [2,3,0,1] and [1,0,3,2].  All 16 lanes must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double row16_sum(double v) {
  v += dpp_mov<0x128>(v);  // row_ror:8
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  return v;
}

// ---------------------------------------------------------------- matvec

// Workgroups are dispatched to the 8 XCDs round-robin by linear id.  This maps
// the linear id to a work index so that each XCD receives one contiguous range
// (a bijection on [0, n)): neighbouring boxes, whose fields and interaction
// lists overlap, then share one XCD's L2.

// Leaf aggregation (modMLFMA.f90:1805-1812): campoC_L(:,:,box) =
// sum_f funcRAD(:,:,f) x(f), each box's bases in box order.  The patterns are
// assembled from phase rows: one exp(j ka khat . (c_T - C)) row per (leaf
// box, triangle of one of its bases), 0.44 of the bytes of the pattern array
// at 122,880 unknowns, shared by the box's bases through L2, combined with
// each basis's RWG vectors by rad_from_phase (the function that tabulates the
// patterns for the receive, so the values are the same bits).  One workgroup
// per leaf box, one thread per sample (both polarisations).
__global__ void __launch_bounds__(256) k_aggregate_ph(int nk, const int* __restrict__ box_slot,
                                                      const int* __restrict__ slot_basis,
                                                      const int2* __restrict__ slot_pair,
                                                      const RadGeom* __restrict__ geo,
                                                      const double* __restrict__ ath,
                                                      const double* __restrict__ aph,
                                                      const double2* __restrict__ tph,
                                                      const double2* __restrict__ x, double2* __restrict__ C) {
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int s0 = box_slot[b], s1 = box_slot[b + 1];
  constexpr int kU = 4;  // bases whose phase loads are issued together
  for (int k = threadIdx.x; k < nk; k += blockDim.x) {
    double at[3], ap[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      at[d] = ath[3 * k + d];
      ap[d] = aph[3 * k + d];
    }
    double2 acc0 = make_double2(0.0, 0.0), acc1 = make_double2(0.0, 0.0);
    int s = s0;
    for (; s + kU <= s1; s += kU) {
      double2 pp[kU], pm[kU], xv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int2 pr = slot_pair[s + u];
        pp[u] = tph[(size_t)pr.x * nk + k];
        pm[u] = tph[(size_t)pr.y * nk + k];
        xv[u] = x[slot_basis[s + u]];
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        double2 r0, r1;
        rad_from_phase(geo[s + u], pp[u], pm[u], at, ap, r0, r1);
        acc0 = cadd(acc0, cmul(r0, xv[u]));
        acc1 = cadd(acc1, cmul(r1, xv[u]));
      }
    }
    for (; s < s1; ++s) {
      const int2 pr = slot_pair[s];
      double2 r0, r1;
      rad_from_phase(geo[s], tph[(size_t)pr.x * nk + k], tph[(size_t)pr.y * nk + k], at, ap, r0, r1);
      const double2 xv = x[slot_basis[s]];
      acc0 = cadd(acc0, cmul(r0, xv));
      acc1 = cadd(acc1, cmul(r1, xv));
    }
    C[((size_t)b * 2 + 0) * nk + k] = acc0;
    C[((size_t)b * 2 + 1) * nk + k] = acc1;
  }
}

// Leaf aggregation from the phase rows, factored: with funcRAD(k,p,f) =
// l/2 (e+(k) rho+ + e-(k) rho-) . u_p(k) (calcRad, modMLFMA.f90:669-699),
//   campoC_L(k,p) = sum_f funcRAD(k,p,f) x(f) = u_p(k) . sum_T e_T(k) H_T,
//   H_T = sum of l/2 x(f) rho+ over the box's bases whose plus triangle is T
//         and of l/2 x(f) rho- over those whose minus triangle is T,
// so each phase row of the box is read once and the pattern is never formed
// (the terms are summed in another order than the reference's; same value
// to rounding).  The rows hold the samples m < nk/2; the antipodal sample
// m' = nk/2 + sample_flip(m) has the conjugate phase, so one load serves
// both: a(m) = sum_T e_T(m) H_T, a(m') = sum_T conj(e_T(m)) H_T.  One
// workgroup per leaf box: H of its rows in LDS (a thread per row over the
// bases the row serves, in slot order), then one thread per sample pair.
__global__ void __launch_bounds__(256) k_aggregate_fac(int nk, int nth, const int* __restrict__ box_row,
                                                       const int* __restrict__ row_ptr,
                                                       const int* __restrict__ row_ent,
                                                       const int* __restrict__ slot_basis,
                                                       const RadGeom* __restrict__ geo,
                                                       const double* __restrict__ ath,
                                                       const double* __restrict__ aph,
                                                       const double2* __restrict__ tph,
                                                       const double2* __restrict__ x, double2* __restrict__ C) {
  extern __shared__ double2 hs[];  // [row of the box][3]
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = box_row[b], nr = box_row[b + 1] - r0;
  for (int t = threadIdx.x; t < nr; t += blockDim.x) {
    double2 h[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) h[d] = make_double2(0.0, 0.0);
    for (int e = row_ptr[r0 + t]; e < row_ptr[r0 + t + 1]; ++e) {
      const int v = row_ent[e], s = v >> 1;
      const RadGeom& g = geo[s];
      const double2 xv = x[slot_basis[s]];
      const double hl = 0.5 * g.len;
      const double2 xs = make_double2(xv.x * hl, xv.y * hl);
      const double* rho = (v & 1) ? g.rm : g.rp;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        h[d].x += xs.x * rho[d];
        h[d].y += xs.y * rho[d];
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) hs[3 * t + d] = h[d];
  }
  __syncthreads();
  const int nh = nk / 2;
  for (int m = threadIdx.x; m < nh; m += blockDim.x) {
    double2 a[3], am[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) a[d] = am[d] = make_double2(0.0, 0.0);
    constexpr int kU = 8;  // phase loads issued together
    for (int t0 = 0; t0 < nr; t0 += kU) {
      double2 e[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (t0 + u < nr) e[u] = tph[(size_t)(r0 + t0 + u) * nh + m];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (t0 + u < nr) {
          const double2* h = hs + 3 * (t0 + u);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            a[d] = cadd(a[d], cmul(e[u], h[d]));
            am[d] = cadd(am[d], cmulc(e[u], h[d]));
          }
        }
    }
    // campoC at sample k from the 3-vector v: theta-hat . v, phi-hat . v
    auto store = [&](int k, const double2 (&v)[3]) {
      double2 c0 = make_double2(0.0, 0.0), c1 = c0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double t = ath[3 * k + d], p = aph[3 * k + d];
        c0.x += t * v[d].x;
        c0.y += t * v[d].y;
        c1.x += p * v[d].x;
        c1.y += p * v[d].y;
      }
      C[((size_t)b * 2 + 0) * nk + k] = c0;
      C[((size_t)b * 2 + 1) * nk + k] = c1;
    };
    store(m, a);
    store(nh + sample_flip(m, nth), am);
  }
}

// Upward pass (modMLFMA.f90:1813-1835): for each parent box, interpolate each
// child's field to the parent's samples (Interpolar, :1343-1365), shift it
// to the parent centre and sum over children in box order.  Grid: (parent,
// chunk of 256 parent samples); the child field is staged in LDS.
__global__ void __launch_bounds__(256) k_interp_up(int nkc, int nkf, const int* __restrict__ child_ptr,
                                                   const int* __restrict__ oct, const double2* __restrict__ ph,
                                                   const double* __restrict__ wint, const int* __restrict__ iint,
                                                   const double2* __restrict__ Cf, double2* __restrict__ Cc) {
  extern __shared__ double2 lds[];  // 2 * nkf
  const int P = blockIdx.x;
  const int i = blockIdx.y * blockDim.x + threadIdx.x;
  double2 acc0 = make_double2(0.0, 0.0), acc1 = make_double2(0.0, 0.0);
  double w[9];
  int id[9];
  if (i < nkc) {
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      w[j] = wint[9 * (size_t)i + j];
      id[j] = iint[9 * (size_t)i + j];
    }
  }
  for (int c = child_ptr[P]; c < child_ptr[P + 1]; ++c) {
    __syncthreads();
    const double2* src = Cf + (size_t)c * 2 * nkf;
    // four loads per thread in flight before its stores (indices clamped:
    // the excess threads rewrite the last element with its own value)
    for (int e0 = threadIdx.x; e0 < 2 * nkf; e0 += 4 * blockDim.x) {
      double2 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = src[min(e0 + u * (int)blockDim.x, 2 * nkf - 1)];
#pragma unroll
      for (int u = 0; u < 4; ++u) lds[min(e0 + u * (int)blockDim.x, 2 * nkf - 1)] = v[u];
    }
    __syncthreads();
    if (i < nkc) {
      double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double2 v0 = lds[id[j]], v1 = lds[nkf + id[j]];
        a0.x += w[j] * v0.x;
        a0.y += w[j] * v0.y;
        a1.x += w[j] * v1.x;
        a1.y += w[j] * v1.y;
      }
      // exp(+j k khat_i . (c_child - c_parent)) from the child's octant
      const double2 e = ph[(size_t)oct[c] * nkc + i];
      acc0 = cadd(acc0, cmul(a0, e));
      acc1 = cadd(acc1, cmul(a1, e));
    }
  }
  if (i < nkc) {
    Cc[((size_t)P * 2 + 0) * nkc + i] = acc0;
    Cc[((size_t)P * 2 + 1) * nkc + i] = acc1;
  }
}

// Downward pass (modMLFMA.f90:1857-1897): shift the parent's field to the
// child centre, anterpolate (the transpose of Interpolar, :1370-1390, in
// gather form) and add the child's own translated field.  One workgroup per
// child box; the shifted parent field is staged in LDS.
__global__ void __launch_bounds__(256) k_interp_down(int nkc, int nkf, const int* __restrict__ parent,
                                                     const int* __restrict__ oct, const double2* __restrict__ ph,
                                                     const int* __restrict__ wt_ptr, const WtEntry* __restrict__ wt,
                                                     const double2* __restrict__ Ap, const double2* __restrict__ Bf,
                                                     double2* __restrict__ Af) {
  // grid: (child box, chunk of 256 fine samples), XCD-contiguous
  const int nch = (nkf + blockDim.x - 1) / blockDim.x;
  const int wb = xcd_block(blockIdx.x, gridDim.x);
  const int c = wb / nch;
  const int s = (wb % nch) * blockDim.x + threadIdx.x;
  if (s >= nkf) return;
  const int P = parent[c];
  const double2* src = Ap + (size_t)P * 2 * nkc;
  const double2* e = ph + (size_t)oct[c] * nkc;  // exp(+j k khat . r); conjugated below
  double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
  // this sample's entries, a wave's block stored entry-major
  const int q0 = wt_ptr[s / kWave], cnt = (wt_ptr[s / kWave + 1] - q0) / kWave;
  const WtEntry* tq = wt + q0 + s % kWave;
  int q = 0;
  // four stencil entries per step: their loads are issued together
  for (; q + 4 <= cnt; q += 4) {
    WtEntry t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) t[u] = tq[(q + u) * kWave];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = t[u].r;
      const double2 pr = e[r];
      const double2 v0 = cmulc(pr, src[r]), v1 = cmulc(pr, src[nkc + r]);
      a0.x += t[u].w * v0.x;
      a0.y += t[u].w * v0.y;
      a1.x += t[u].w * v1.x;
      a1.y += t[u].w * v1.y;
    }
  }
  for (; q < cnt; ++q) {
    const WtEntry t = tq[q * kWave];
    const double2 pr = e[t.r];
    const double2 v0 = cmulc(pr, src[t.r]), v1 = cmulc(pr, src[nkc + t.r]);
    a0.x += t.w * v0.x;
    a0.y += t.w * v0.y;
    a1.x += t.w * v1.x;
    a1.y += t.w * v1.y;
  }
  const size_t o0 = ((size_t)c * 2 + 0) * nkf + s, o1 = ((size_t)c * 2 + 1) * nkf + s;
  Af[o0] = cadd(Bf[o0], a0);
  Af[o1] = cadd(Bf[o1], a1);
}

// Downward pass, LDS-staged variant for levels whose parent field fits in
// LDS: one workgroup per child box stages the phase-shifted parent field once
// and every fine sample of the child gathers from it.
__global__ void __launch_bounds__(512) k_interp_down_lds(int nkc, int nkf, const int* __restrict__ parent,
                                                         const int* __restrict__ oct, const double2* __restrict__ ph,
                                                         const int* __restrict__ wt_ptr,
                                                         const WtEntry* __restrict__ wt,
                                                         const double2* __restrict__ Ap,
                                                         const double2* __restrict__ Bf, double2* __restrict__ Af) {
  extern __shared__ double2 lds[];  // 2 * nkc
  const int c = xcd_block(blockIdx.x, gridDim.x);
  const int P = parent[c];
  const double2* src = Ap + (size_t)P * 2 * nkc;
  const double2* e = ph + (size_t)oct[c] * nkc;
  // two samples per thread in flight before its stores (indices clamped:
  // the excess threads rewrite the last sample with its own value)
  for (int i0 = threadIdx.x; i0 < nkc; i0 += 2 * blockDim.x) {
    double2 pr[2], v0[2], v1[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, nkc - 1);
      pr[u] = e[i];
      v0[u] = src[i];
      v1[u] = src[nkc + i];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, nkc - 1);
      lds[i] = cmulc(pr[u], v0[u]);
      lds[nkc + i] = cmulc(pr[u], v1[u]);
    }
  }
  __syncthreads();
  for (int s = threadIdx.x; s < nkf; s += blockDim.x) {
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
    const int q0 = wt_ptr[s / kWave], cnt = (wt_ptr[s / kWave + 1] - q0) / kWave;
    const WtEntry* tq = wt + q0 + s % kWave;
    int q = 0;
    for (; q + 4 <= cnt; q += 4) {
      WtEntry t[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) t[u] = tq[(q + u) * kWave];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double2 v0 = lds[t[u].r], v1 = lds[nkc + t[u].r];
        a0.x += t[u].w * v0.x;
        a0.y += t[u].w * v0.y;
        a1.x += t[u].w * v1.x;
        a1.y += t[u].w * v1.y;
      }
    }
    for (; q < cnt; ++q) {
      const WtEntry t = tq[q * kWave];
      a0.x += t.w * lds[t.r].x;
      a0.y += t.w * lds[t.r].y;
      a1.x += t.w * lds[nkc + t.r].x;
      a1.y += t.w * lds[nkc + t.r].y;
    }
    const size_t o0 = ((size_t)c * 2 + 0) * nkf + s, o1 = ((size_t)c * 2 + 1) * nkf + s;
    Af[o0] = cadd(Bf[o0], a0);
    Af[o1] = cadd(Bf[o1], a1);
  }
}

// Downward pass for levels whose parent field does not fit LDS whole: a
// workgroup takes one child box and 256 of its fine samples, whose stencils
// read a cyclic range of the parent's phi columns (first column win[2 ch],
// count win[2 ch + 1]).  That window of the phase-shifted parent field is
// staged in LDS, and the stencil (rows relative to the window) gathers from
// it with k_interp_down_lds's operations.
__global__ void __launch_bounds__(256) k_interp_down_win(int nkc, int nkf, int nthc, int nphc,
                                                         const int* __restrict__ parent, const int* __restrict__ oct,
                                                         const double2* __restrict__ ph, const int* __restrict__ win,
                                                         const int* __restrict__ wt_ptr,
                                                         const WtEntry* __restrict__ wtw,
                                                         const double2* __restrict__ Ap,
                                                         const double2* __restrict__ Bf, double2* __restrict__ Af) {
  extern __shared__ double2 lds[];  // 2 * (window columns) * nthc
  const int nch = (nkf + 255) / 256;
  const int wb = xcd_block(blockIdx.x, gridDim.x);
  const int c = wb / nch, ch = wb % nch;
  const int P = parent[c];
  const double2* src = Ap + (size_t)P * 2 * nkc;
  const double2* e = ph + (size_t)oct[c] * nkc;
  const int ip0 = win[2 * ch], nw = win[2 * ch + 1] * nthc;
  // two window samples per thread in flight before its stores (indices
  // clamped: the excess threads rewrite the last sample with its own value)
  for (int i0 = threadIdx.x; i0 < nw; i0 += 2 * blockDim.x) {
    double2 pr[2], v0[2], v1[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, nw - 1);
      const int q = i / nthc;
      int ip = ip0 + q;
      if (ip >= nphc) ip -= nphc;
      const int r = ip * nthc + (i - q * nthc);
      pr[u] = e[r];
      v0[u] = src[r];
      v1[u] = src[nkc + r];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int i = min(i0 + u * (int)blockDim.x, nw - 1);
      lds[i] = cmulc(pr[u], v0[u]);
      lds[nw + i] = cmulc(pr[u], v1[u]);
    }
  }
  __syncthreads();
  const int s = ch * 256 + threadIdx.x;
  if (s >= nkf) return;
  double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
  const int q0 = wt_ptr[s / kWave], cnt = (wt_ptr[s / kWave + 1] - q0) / kWave;
  const WtEntry* tq = wtw + q0 + s % kWave;
  int q = 0;
  for (; q + 4 <= cnt; q += 4) {
    WtEntry t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) t[u] = tq[(q + u) * kWave];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double2 v0 = lds[t[u].r], v1 = lds[nw + t[u].r];
      a0.x += t[u].w * v0.x;
      a0.y += t[u].w * v0.y;
      a1.x += t[u].w * v1.x;
      a1.y += t[u].w * v1.y;
    }
  }
  for (; q < cnt; ++q) {
    const WtEntry t = tq[q * kWave];
    a0.x += t.w * lds[t.r].x;
    a0.y += t.w * lds[t.r].y;
    a1.x += t.w * lds[nw + t.r].x;
    a1.y += t.w * lds[nw + t.r].y;
  }
  const size_t o0 = ((size_t)c * 2 + 0) * nkf + s, o1 = ((size_t)c * 2 + 1) * nkf + s;
  Af[o0] = cadd(Bf[o0], a0);
  Af[o1] = cadd(Bf[o1], a1);
}

// Receive (modMLFMA.f90:1915-1933, funcREC = conjg(funcRAD) under EFIE) and
// near-field CSR product (:1935-1942), one wavefront per row.
template <typename V>
__global__ void __launch_bounds__(256) k_receive_near(int ne, int nk2, const int* __restrict__ slot_basis,
                                                      const int* __restrict__ basis_box,
                                                      const double2* __restrict__ rad,
                                                      const double2* __restrict__ AL, const int* __restrict__ ia,
                                                      const int* __restrict__ ja, const V* __restrict__ z,
                                                      const double2* __restrict__ x, double2 cc,
                                                      double2* __restrict__ y, double2* __restrict__ y_near) {
  // rows taken in leaf-slot order: a box's campoA is then read by neighbouring
  // wavefronts of one XCD
  const int slot = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (slot >= ne) return;
  const int m = slot_basis[slot];
  const double2* r = rad + (size_t)slot * nk2;
  const double2* a = AL + (size_t)basis_box[m] * nk2;
  double2 f = make_double2(0.0, 0.0);
  for (int e = lane; e < nk2; e += kWave) f = cadd(f, cmulc(r[e], a[e]));
  f = wave_sum(f);
  double2 nsum = make_double2(0.0, 0.0);
  for (int k = ia[m] + lane; k < ia[m + 1]; k += kWave) nsum = cadd(nsum, cmul(x[ja[k]], to_d2(z[k])));
  nsum = wave_sum(nsum);
  if (lane == 0) {
    const double2 far = cmul(f, cc);
    if (y_near) {
      y[m] = far;
      y_near[m] = nsum;
    } else {
      y[m] = cadd(far, nsum);
    }
  }
}

// Near-field CSR product alone (modMLFMA.f90:1935-1942): y_near(m) = the sum
// of row m's terms, one wavefront per row, rows in leaf-slot order (the rows
// of a box gather the same x entries).  The lane-strided partial sums and
// their wavefront reduction are those of k_receive_near, so its near sums
// are the same bits.  It reads only x, so it runs beside the far field.
template <typename V>
__global__ void __launch_bounds__(256) k_near_part(int ne, const int* __restrict__ slot_basis,
                                                   const int* __restrict__ ia, const int* __restrict__ ja,
                                                   const V* __restrict__ z, const double2* __restrict__ x,
                                                   double2* __restrict__ y_near) {
  const int slot = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  if (slot >= ne) return;
  const int m = slot_basis[slot];
  const int k0 = ia[m], k1 = ia[m + 1];
  double2 nsum = make_double2(0.0, 0.0);
  // eight entries per lane per pass: the column and value loads of the pass,
  // then its x gathers, are issued together; each lane adds its entries in
  // the order of the one-entry loop (k = k0 + lane, + 64, ...)
  constexpr int kU = 8;
  for (int kb = k0 + lane; kb < k1; kb += kU * kWave) {
    int j[kU];
    V zz[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int k = kb + u * kWave;
      j[u] = k < k1 ? ja[k] : -1;
      if (k < k1) zz[u] = z[k];
    }
    double2 xv[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (j[u] >= 0) xv[u] = x[j[u]];
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (j[u] >= 0) nsum = cadd(nsum, cmul(xv[u], to_d2(zz[u])));
  }
  nsum = wave_sum(nsum);
  if (lane == 0) y_near[m] = nsum;
}

// Near-field CSR product by leaf box: every row of a leaf box has the same
// near columns (the box's near boxes' bases, ascending), so the box's x
// entries are gathered once into LDS and each row (a wavefront) streams only
// its values.  Each lane adds the entries of the one-entry loop of
// k_near_part in its order, and the reduction is the same: the sums are the
// same bits.
template <typename V>
__global__ void __launch_bounds__(256) k_near_box(const int* __restrict__ box_slot,
                                                  const int* __restrict__ slot_basis, const int* __restrict__ ia,
                                                  const int* __restrict__ ja, const V* __restrict__ z,
                                                  const double2* __restrict__ x, double2* __restrict__ y_near) {
  extern __shared__ double2 xs[];
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int s0 = box_slot[b], s1 = box_slot[b + 1];
  const int c0 = ia[slot_basis[s0]], nc = ia[slot_basis[s0] + 1] - c0;
  // four columns per thread in flight (indices, then values) before its
  // stores (clamped: the excess threads rewrite the last column's value)
  for (int e0 = threadIdx.x; e0 < nc; e0 += 4 * blockDim.x) {
    int col[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) col[u] = ja[c0 + min(e0 + u * (int)blockDim.x, nc - 1)];
    double2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = x[col[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) xs[min(e0 + u * (int)blockDim.x, nc - 1)] = v[u];
  }
  __syncthreads();
  const int lane = threadIdx.x % kWave;
  const int nw = blockDim.x / kWave;
  constexpr int kU = 8;
  // two rows per wavefront at a time (s and s + nw): the value loads of both
  // are issued together
  for (int s = s0 + (int)(threadIdx.x / kWave); s < s1; s += 2 * nw) {
    const bool two = s + nw < s1;
    const int m0 = slot_basis[s], m1 = two ? slot_basis[s + nw] : m0;
    const V* zr0 = z + ia[m0];
    const V* zr1 = z + ia[m1];
    double2 n0 = make_double2(0.0, 0.0), n1 = n0;
    for (int eb = lane; eb < nc; eb += kU * kWave) {
      V z0[kU], z1[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (eb + u * kWave < nc) {
          z0[u] = zr0[eb + u * kWave];
          z1[u] = zr1[eb + u * kWave];
        }
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (eb + u * kWave < nc) {
          const double2 xv = xs[eb + u * kWave];
          n0 = cadd(n0, cmul(xv, to_d2(z0[u])));
          n1 = cadd(n1, cmul(xv, to_d2(z1[u])));
        }
    }
    n0 = wave_sum(n0);
    n1 = wave_sum(n1);
    if (lane == 0) {
      y_near[m0] = n0;
      if (two) y_near[m1] = n1;
    }
  }
}

// Receive (modMLFMA.f90:1915-1933) from the phase rows, one workgroup per
// leaf box.  With funcREC = conj(funcRAD) and funcRAD(k,p,f) =
// l/2 (e+(k) rho+ + e-(k) rho-) . u_p(k) (calcRad, :669-699; e+- the phase
// rows of the basis's triangles, u_1 = theta-hat, u_2 = phi-hat), the sum
// over samples and polarisations factors as
//   sum_k sum_p conj(funcRAD(k,p,f)) campoA(k,p) = l/2 (rho+ . G+ + rho- . G-),
//   G_T = sum_k conj(e_T(k)) F(k),  F(k) = sum_p u_p(k) campoA(k,p),
// so each phase row is read once and contracted with the box's 3-vector
// field F in LDS.  The rows hold the samples m < nk/2, and the antipodal
// sample m' = nk/2 + sample_flip(m) has the conjugate phase, so
//   G_T = sum_{m < nk/2} conj(e_T(m)) F(m) + e_T(m) F(m'),
// 0.22 of the pattern bytes at 122,880 unknowns.  The terms are summed in
// another order than the reference's (same value to rounding).
// y(f) = (j w mu / 4 pi) (...) + y_near(f) when y_near is given (the near
// sum of k_near_part, added as k_receive_near adds it), else the far part
// alone.
__global__ void __launch_bounds__(256) k_receive_ph(int nk, int nth, const int* __restrict__ box_slot,
                                                    const int* __restrict__ box_row,
                                                    const int* __restrict__ slot_basis,
                                                    const int2* __restrict__ slot_pair,
                                                    const RadGeom* __restrict__ geo,
                                                    const double* __restrict__ ath,
                                                    const double* __restrict__ aph,
                                                    const double2* __restrict__ tph,
                                                    const double2* __restrict__ AL, double2 cc,
                                                    const double2* __restrict__ y_near, double2* __restrict__ y) {
  extern __shared__ double2 sm[];
  const int nh = nk / 2;
  double2* F = sm;            // [3][nh]: F(m), m < nh
  double2* Fm = sm + 3 * nh;  // [3][nh]: F of m's antipode
  double2* G = sm + 3 * nk;   // [row of the box][3]
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const double2* a = AL + (size_t)b * 2 * nk;
  for (int k = threadIdx.x; k < nk; k += blockDim.x) {
    const double2 a0 = a[k], a1 = a[nk + k];
    double2* f = k < nh ? F + k : Fm + sample_flip(k - nh, nth);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double t = ath[3 * k + d], p = aph[3 * k + d];
      f[d * nh] = make_double2(t * a0.x + p * a1.x, t * a0.y + p * a1.y);
    }
  }
  __syncthreads();
  const int r0 = box_row[b], r1 = box_row[b + 1];
  // a group of 16 lanes (one DPP row) per phase row: lane j takes the
  // samples j, j + 16, ...; the group's partial sums are combined by DPP
  const int j = threadIdx.x % 16, ng = blockDim.x / 16;
  for (int r = r0 + (int)(threadIdx.x / 16); r < r1; r += ng) {
    const double2* ph = tph + (size_t)r * nh;
    double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    constexpr int kU = 8;  // row loads issued together
    for (int kb = j; kb < nh; kb += kU * 16) {
      double2 p[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u)
        if (kb + 16 * u < nh) p[u] = ph[kb + 16 * u];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int m = kb + 16 * u;
        if (m < nh) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double2 v = cmulc(p[u], F[d * nh + m]), w = cmul(p[u], Fm[d * nh + m]);
            g[2 * d] += v.x;
            g[2 * d + 1] += v.y;
            g[2 * d] += w.x;
            g[2 * d + 1] += w.y;
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) g[q] = row16_sum(g[q]);
    if (j == 0) {
      double2* o = G + 3 * (size_t)(r - r0);
#pragma unroll
      for (int d = 0; d < 3; ++d) o[d] = make_double2(g[2 * d], g[2 * d + 1]);
    }
  }
  __syncthreads();
  for (int s = box_slot[b] + threadIdx.x; s < box_slot[b + 1]; s += blockDim.x) {
    const int2 pr = slot_pair[s];
    const RadGeom& g = geo[s];
    const double2* gp = G + 3 * (size_t)(pr.x - r0);
    const double2* gm = G + 3 * (size_t)(pr.y - r0);
    double2 v = make_double2(0.0, 0.0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v.x += g.rp[d] * gp[d].x + g.rm[d] * gm[d].x;
      v.y += g.rp[d] * gp[d].y + g.rm[d] * gm[d].y;
    }
    const double hl = 0.5 * g.len;
    const double2 far = cmul(make_double2(v.x * hl, v.y * hl), cc);
    const int m = slot_basis[s];
    y[m] = y_near ? cadd(far, y_near[m]) : far;
  }
}

// Device copy of n complex(8) values (grid-stride): the bandwidth reference
// for the roofline figures and the HBM-traffic counter calibration.
__global__ void __launch_bounds__(256) k_copy(long long n, const double2* __restrict__ a, double2* __restrict__ b) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// ---------------------------------------------------------------- ILU0 solves

// ---------------------------------------------------------------- CGS pieces

constexpr int kRedBlocks = 512;

// Two grid-wide sums in fixed order: each block writes its partials (waves
// in order), and the block that arrives last at *counter sums all partials
// (lane b of wave 0 takes blocks b, b + 64, ..., then a butterfly), so a sum
// has the same bits from run to run. True in wave 0 of that block, with the
// sums in t1, t2; *counter is back at zero for the next launch. 256 threads.
__device__ bool grid_sum2(double2 s1, double2 s2, double* __restrict__ part, int* __restrict__ counter,
                          double2& t1, double2& t2) {
  __shared__ double2 sh[2][256 / kWave];
  __shared__ int last;
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
  if (ln == 0) {
    sh[0][wv] = s1;
    sh[1][wv] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 b1 = make_double2(0.0, 0.0), b2 = make_double2(0.0, 0.0);
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) {
      b1 = cadd(b1, sh[0][w]);
      b2 = cadd(b2, sh[1][w]);
    }
    part[4 * blockIdx.x + 0] = b1.x;
    part[4 * blockIdx.x + 1] = b1.y;
    part[4 * blockIdx.x + 2] = b2.x;
    part[4 * blockIdx.x + 3] = b2.y;
    // release: this block's partials before its count; acquire: the others'
    last = __hip_atomic_fetch_add(counter, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last || threadIdx.x >= kWave) return false;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  t1 = make_double2(0.0, 0.0);
  t2 = make_double2(0.0, 0.0);
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kWave) {
    t1 = cadd(t1, make_double2(part[4 * b], part[4 * b + 1]));
    t2 = cadd(t2, make_double2(part[4 * b + 2], part[4 * b + 3]));
  }
  t1 = wave_sum(t1);
  t2 = wave_sum(t2);
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// out[0] = sum(a*b) (conj_a: conj(a)*b), out[1] = sum(a2*b2) (0 without a2),
// in one pass over the vectors.
__global__ void __launch_bounds__(256) k_dot2(int n, const double2* __restrict__ a, const double2* __restrict__ b,
                                              int conj_a, const double2* __restrict__ a2,
                                              const double2* __restrict__ b2, int conj_a2,
                                              double* __restrict__ part, int* __restrict__ counter,
                                              double2* __restrict__ out) {
  double2 s1 = make_double2(0.0, 0.0), s2 = make_double2(0.0, 0.0);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    s1 = cadd(s1, conj_a ? cmulc(a[i], b[i]) : cmul(a[i], b[i]));
    if (a2) s2 = cadd(s2, conj_a2 ? cmulc(a2[i], b2[i]) : cmul(a2[i], b2[i]));
  }
  double2 t1, t2;
  if (grid_sum2(s1, s2, part, counter, t1, t2) && threadIdx.x == 0) {
    out[0] = t1;
    out[1] = t2;
  }
}

__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

// it == 1: u = r, p = u; else beta = rho1/rho2, u = r + beta q,
// p = u + beta (q + beta p)   (modIterativo.f90:396-403)
__global__ void k_cgs_up(int n, int first, const double2* __restrict__ rho1p, const double2* __restrict__ rho2p,
                         const double2* __restrict__ r, const double2* __restrict__ q, double2* __restrict__ u,
                         double2* __restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (first) {
    u[i] = r[i];
    p[i] = r[i];
    return;
  }
  const double2 beta = cdiv(*rho1p, *rho2p);
  const double2 ui = cadd(r[i], cmul(beta, q[i]));
  u[i] = ui;
  p[i] = cadd(ui, cmul(beta, cadd(q[i], cmul(beta, p[i]))));
}

// alfa = rho1 / (R_ . V_); q = u - alfa V_; t = u + q   (:419-429)
__global__ void k_cgs_q(int n, const double2* __restrict__ rho1p, const double2* __restrict__ denp,
                        double2* __restrict__ alfa_out, const double2* __restrict__ u,
                        const double2* __restrict__ V, double2* __restrict__ q, double2* __restrict__ t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double2 alfa = cdiv(*rho1p, *denp);
  if (i == 0) *alfa_out = alfa;
  if (i >= n) return;
  const double2 qi = make_double2(u[i].x - (alfa.x * V[i].x - alfa.y * V[i].y),
                                  u[i].y - (alfa.x * V[i].y + alfa.y * V[i].x));
  q[i] = qi;
  t[i] = cadd(u[i], qi);
}

// The end of an iteration. x += alfa U_; r -= alfa Q_ (:434-437), nothing
// when *stop is set (the iteration was queued before the host saw the
// previous one end the solve); |r|^2 and the next rho = sum(r*R_) summed as
// k_dot2 sums them; then, in the last block, the host's two exits decided
// again from the same sums: the stopping test |r|^2 < thr (thr chosen so that
// this is sqrt(|r|^2) < errmax ||b||, :472) and a zero next rho (the breakdown
// exit at the top of the next iteration, :392). rho2 = rho1, rho1 = next rho
// (:487); s[4] = |r|^2 and s[5] = next rho, also written to the host slot.
__global__ void __launch_bounds__(256) k_cgs_xr_dot(int n, const double2* __restrict__ U,
                                                    const double2* __restrict__ Q, double2* __restrict__ x,
                                                    double2* __restrict__ r, const double2* __restrict__ R_,
                                                    double* __restrict__ part, int* __restrict__ counter,
                                                    double thr, double2* __restrict__ s, int* __restrict__ stop,
                                                    double2* __restrict__ host) {
  // every block reads the flag before its count, the last block writes it after
  const bool skip = *stop != 0;
  const double2 a = s[3];
  double2 s1 = make_double2(0.0, 0.0), s2 = make_double2(0.0, 0.0);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double2 ri = r[i];
    if (!skip) {
      x[i] = cadd(x[i], cmul(a, U[i]));
      const double2 aq = cmul(a, Q[i]);
      ri = make_double2(ri.x - aq.x, ri.y - aq.y);
      r[i] = ri;
    }
    s1 = cadd(s1, cmulc(ri, ri));
    s2 = cadd(s2, cmul(ri, R_[i]));
  }
  double2 t1, t2;
  if (!grid_sum2(s1, s2, part, counter, t1, t2) || threadIdx.x != 0) return;
  s[4] = t1;
  s[5] = t2;
  *stop = (t1.x < thr || (t2.x == 0.0 && t2.y == 0.0)) ? 1 : 0;
  s[1] = s[0];
  s[0] = t2;
  host[0] = t1;
  host[1] = t2;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
}

__global__ void k_sub(int n, const double2* __restrict__ a, const double2* __restrict__ b, double2* __restrict__ o) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = make_double2(a[i].x - b[i].x, a[i].y - b[i].y);
}

// ---------------------------------------------------------------- excitation and RCS

// generar_onda (main.f90:1488-1525), EFIE term: b(i) = l_i/2 [E(c+).(c+ - v+) +
// E(c-).(v- - c-)], E(r) = pol ctte exp(-j k d.r)
__global__ void k_plane_wave(int ne, const double* __restrict__ tb, const double* __restrict__ pc,
                             const int* __restrict__ et, const int* __restrict__ epo,
                             const double* __restrict__ el, double ux, double uy, double uz, double k0,
                             double2 p0, double2 p1, double2 p2, double2 ctte, double2* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  const double2 pc0 = cmul(p0, ctte), pc1 = cmul(p1, ctte), pc2 = cmul(p2, ctte);
  double2 acc = make_double2(0.0, 0.0);
  for (int side = 0; side < 2; ++side) {
    const double* c = tb + 3 * (et[2 * i + side] - 1);
    const double* v = pc + 3 * (epo[2 * i + side] - 1);
    double sn, cs;
    sincos(-(k0 * ((ux * c[0] + uy * c[1]) + uz * c[2])), &sn, &cs);
    const double2 ph = make_double2(cs, sn);
    const double sg = side == 0 ? 1.0 : -1.0;
    const double d0 = sg * (c[0] - v[0]), d1 = sg * (c[1] - v[1]), d2 = sg * (c[2] - v[2]);
    const double2 e0 = cmul(pc0, ph), e1 = cmul(pc1, ph), e2 = cmul(pc2, ph);
    acc.x += (e0.x * d0 + e1.x * d1) + e2.x * d2;
    acc.y += (e0.y * d0 + e1.y * d1) + e2.y * d2;
  }
  const double f = 0.5 * el[i];
  b[i] = make_double2(acc.x * f, acc.y * f);
}

// calc_E (modRCS.f90:60-89): E(r_obs) = (j w mu / 4 pi) sum_i g(r_obs, rm_i)
// ((m_i.rrm) rrm/|rrm|^2 - m_i); one workgroup per observation point.
__global__ void __launch_bounds__(256) k_rcs(int ne, const double* __restrict__ tb, const int* __restrict__ et,
                                             const double* __restrict__ el, const double2* __restrict__ cur,
                                             const double* __restrict__ robs, double kfs,
                                             double2* __restrict__ Eout) {
  __shared__ double2 sh[3][256 / kWave];
  const double* ro = robs + 3 * blockIdx.x;
  double2 E0 = make_double2(0, 0), E1 = make_double2(0, 0), E2 = make_double2(0, 0);
  for (int i = threadIdx.x; i < ne; i += blockDim.x) {
    const double* t1 = tb + 3 * (et[2 * i] - 1);
    const double* t2 = tb + 3 * (et[2 * i + 1] - 1);
    const double2 Il = make_double2(cur[i].x * el[i], cur[i].y * el[i]);
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
    const double2 g = make_double2(cs / rn, sn / rn);
    const double2 gI = cmul(g, Il);
    // (m.rrm) rrm / |rrm|^2 - m with m = dm * I * l
    const double c0 = dr * rr[0] / rr2 - dm[0], c1 = dr * rr[1] / rr2 - dm[1], c2 = dr * rr[2] / rr2 - dm[2];
    E0.x += gI.x * c0; E0.y += gI.y * c0;
    E1.x += gI.x * c1; E1.y += gI.y * c1;
    E2.x += gI.x * c2; E2.y += gI.y * c2;
  }
  E0 = wave_sum(E0);
  E1 = wave_sum(E1);
  E2 = wave_sum(E2);
  const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
  if (ln == 0) {
    sh[0][wv] = E0;
    sh[1][wv] = E1;
    sh[2][wv] = E2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double2 s = make_double2(0, 0);
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s = cadd(s, sh[threadIdx.x][w]);
    Eout[3 * blockIdx.x + threadIdx.x] = s;
  }
}

// child-offset phase table of one level (see DevLevel::ph)
__global__ void k_phase_table64(int nk, const double* __restrict__ kap, double q, double k0,
                                double2* __restrict__ ph) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  for (int o = 0; o < 8; ++o) {
    const double rx = (o & 4) ? q : -q, ry = (o & 2) ? q : -q, rz = (o & 1) ? q : -q;
    double s, c;
    sincos(k0 * ((kap[3 * i] * rx + kap[3 * i + 1] * ry) + kap[3 * i + 2] * rz), &s, &c);
    ph[(size_t)o * nk + i] = make_double2(c, s);
  }
}

// ---------------------------------------------------------------- helpers

int blocks_for(long long n, int b) { return (int)((n + b - 1) / b); }

// parent fields up to this size are staged in LDS by the downward pass
constexpr int kDownLdsBytes = 64 * 1024;

// one full matvec on device vectors x -> y (y_near: optional split output)
}  // namespace

int run_matvec(Handle* h, const double2* x, double2* y, double2* y_near, hipEvent_t* ev) {
  const int L = h->L;
  hipStream_t st = h->stream;
  DevLevel& lf = h->lv[L];
  const bool serial = ev != nullptr;
  // the near part goes to the caller's buffer (parts) or to the handle's,
  // from which the receive adds it
  double2* yn = y_near ? y_near : h->d_ynear;
  auto aggregate = [&](hipStream_t s) {
    if (h->agg_lds)
      hipLaunchKernelGGL(k_aggregate_fac, dim3(lf.nb), dim3(h->ph_len <= 128 ? 128 : 256), h->agg_lds, s, lf.nk,
                         h->plan.lv[L].nth, h->d_box_row, h->d_row_ptr,
                         h->d_row_ent, h->d_slot_basis, h->d_rgeo, h->d_ath, h->d_aph, h->d_tph, x, lf.C);
    else
      hipLaunchKernelGGL(k_aggregate_ph, dim3(lf.nb), dim3(256), 0, s, lf.nk, h->d_box_slot, h->d_slot_basis,
                         h->d_slot_pair, h->d_rgeo, h->d_ath, h->d_aph, h->d_tph, x, lf.C);
  };
  // campoC of level n-1 from level n
  auto up = [&](int n, hipStream_t s) {
    DevLevel& c = h->lv[n];
    DevLevel& p = h->lv[n - 1];
    dim3 grid(p.nb, blocks_for(p.nk, 256));
    hipLaunchKernelGGL(k_interp_up, grid, dim3(256), 2 * c.nk * sizeof(double2), s, p.nk, c.nk, p.child_ptr, c.oct,
                       p.ph, p.wint, p.iint, c.C, p.C);
  };
  // campoA of level n from level n-1 and campoB of level n
  auto down = [&](int n, hipStream_t s) {
    DevLevel& c = h->lv[n];
    DevLevel& p = h->lv[n - 1];
    const double2* src = (n == 4) ? p.B : p.A;
    // LDS form: a thread per fine sample, 256 threads when the child has no
    // more samples, so that more boxes share a CU
    if (2 * p.nk * (int)sizeof(double2) <= kDownLdsBytes)
      hipLaunchKernelGGL(k_interp_down_lds, dim3(c.nb), dim3(c.nk <= 256 ? 256 : 512), 2 * p.nk * sizeof(double2), s,
                         p.nk, c.nk, c.parent, c.oct, p.ph, p.wt_ptr, p.wt, src, c.B, c.A);
    else if (p.win_lds)
      hipLaunchKernelGGL(k_interp_down_win, dim3(c.nb * blocks_for(c.nk, 256)), dim3(256), p.win_lds, s, p.nk, c.nk,
                         h->plan.lv[n - 1].nth, h->plan.lv[n - 1].nph, c.parent, c.oct, p.ph, p.win, p.wt_ptr, p.wtw,
                         src, c.B, c.A);
    else
      hipLaunchKernelGGL(k_interp_down, dim3(c.nb * blocks_for(c.nk, 256)), dim3(256), 0, s, p.nk, c.nk,
                         c.parent, c.oct, p.ph, p.wt_ptr, p.wt, src, c.B, c.A);
  };
  const double w = h->plan.phys.w;
  const double2 cc = make_double2(0.0, w * kMu / (4.0 * kPi));
  const int rows_per_block = 4;
  dim3 rgrid(blocks_for(h->ne, rows_per_block)), rblock(kWave * rows_per_block);
  auto near = [&](hipStream_t s) {
    if (!h->recv_lds) return;  // summed by k_receive_near with the receive
    if (h->near_lds) {
      if (h->fp64_near)
        hipLaunchKernelGGL(k_near_box<double2>, dim3(lf.nb), dim3(256), h->near_lds, s, h->d_box_slot,
                           h->d_slot_basis, h->d_ia, h->d_ja, (const double2*)h->d_z, x, yn);
      else
        hipLaunchKernelGGL(k_near_box<float2>, dim3(lf.nb), dim3(256), h->near_lds, s, h->d_box_slot,
                           h->d_slot_basis, h->d_ia, h->d_ja, (const float2*)h->d_z, x, yn);
    } else if (h->fp64_near)
      hipLaunchKernelGGL(k_near_part<double2>, rgrid, rblock, 0, s, h->ne, h->d_slot_basis, h->d_ia, h->d_ja,
                         (const double2*)h->d_z, x, yn);
    else
      hipLaunchKernelGGL(k_near_part<float2>, rgrid, rblock, 0, s, h->ne, h->d_slot_basis, h->d_ia, h->d_ja,
                         (const float2*)h->d_z, x, yn);
  };
  auto receive = [&](hipStream_t s) {
    if (h->recv_lds) {
      hipLaunchKernelGGL(k_receive_ph, dim3(lf.nb), dim3(256), h->recv_lds, s, lf.nk, h->plan.lv[L].nth,
                         h->d_box_slot, h->d_box_row,
                         h->d_slot_basis, h->d_slot_pair, h->d_rgeo, h->d_ath, h->d_aph, h->d_tph, lf.A, cc,
                         y_near ? nullptr : yn, y);
    } else if (h->fp64_near) {
      hipLaunchKernelGGL(k_receive_near<double2>, rgrid, rblock, 0, s, h->ne, 2 * lf.nk, h->d_slot_basis,
                         h->d_basis_box, h->d_rad, lf.A, h->d_ia, h->d_ja, (const double2*)h->d_z, x, cc, y, y_near);
    } else {
      hipLaunchKernelGGL(k_receive_near<float2>, rgrid, rblock, 0, s, h->ne, 2 * lf.nk, h->d_slot_basis,
                         h->d_basis_box, h->d_rad, lf.A, h->d_ia, h->d_ja, (const float2*)h->d_z, x, cc, y, y_near);
    }
  };
  if (serial) {
    LW_HIP(hipEventRecord(ev[0], st));
    aggregate(st);
    LW_HIP(hipEventRecord(ev[1], st));
    for (int n = L; n >= 4; --n) up(n, st);
    LW_HIP(hipEventRecord(ev[2], st));
    for (int n = 3; n <= L; ++n) launch_translate_parent(h, n, st);
    LW_HIP(hipEventRecord(ev[3], st));
    for (int n = 4; n <= L; ++n) down(n, st);
    LW_HIP(hipEventRecord(ev[4], st));
    near(st);
    LW_HIP(hipEventRecord(ev[5], st));
    receive(st);
    LW_HIP(hipEventRecord(ev[6], st));
  } else {
    // critical path on the main stream: aggregation, upward passes, level
    // 3's translation, downward passes, receive.  Beside it, once the
    // aggregation is done (started earlier, the near field would share HBM
    // with it): the near field and then level 4's translation (side 0), the
    // leaf translation (side 1) and the translations of levels L-1..5 (side
    // 2), each translation once its campoC is ready.  Each downward pass
    // waits only for the translation of its own level.
    hipStream_t sn = h->side[0], sa = h->side[1], sb = h->side[2];
    aggregate(st);
    LW_HIP(hipEventRecord(h->ev_fork, st));
    LW_HIP(hipStreamWaitEvent(sn, h->ev_fork, 0));
    near(sn);
    const bool leaf_side = L > 3, mid_side = L > 4;
    if (leaf_side) {
      LW_HIP(hipStreamWaitEvent(sa, h->ev_fork, 0));
      launch_translate_parent(h, L, sa);
      LW_HIP(hipEventRecord(h->ev_ta, sa));
    }
    for (int n = L; n >= 4; --n) {
      up(n, st);
      if (n - 1 >= 4) {
        // level 4 behind the near field rather than behind level 5's
        // translation, whose campoC is ready much earlier
        hipStream_t s = n - 1 == 4 ? sn : sb;
        LW_HIP(hipEventRecord(h->ev_c[n - 1], st));
        LW_HIP(hipStreamWaitEvent(s, h->ev_c[n - 1], 0));
        launch_translate_parent(h, n - 1, s);
      }
    }
    // near field and level 4's translation; levels L-1..5
    LW_HIP(hipEventRecord(h->ev_near, sn));
    if (mid_side) LW_HIP(hipEventRecord(h->ev_tb, sb));
    launch_translate_parent(h, 3, st);
    for (int n = 4; n <= L; ++n) {
      if (n == 4 && mid_side) LW_HIP(hipStreamWaitEvent(st, h->ev_near, 0));
      if (n == 5 && L > 5) LW_HIP(hipStreamWaitEvent(st, h->ev_tb, 0));
      if (n == L && leaf_side) LW_HIP(hipStreamWaitEvent(st, h->ev_ta, 0));
      down(n, st);
    }
    // the receive adds into the near field's output: with level 4 on the
    // side stream, level 4's downward pass has already waited for it (each
    // wait costs the main stream several microseconds of idle time)
    if (!mid_side) LW_HIP(hipStreamWaitEvent(st, h->ev_near, 0));
    receive(st);
  }
  LW_HIP(hipGetLastError());
  return LW_OK;
}

namespace {

int run_lusol(Handle* h, const double2* y, double2* x) {
  if (!h->have_ilu) {
    set_error("lusol0: no ILU0 factor (call lw_ilu0 first)");
    return LW_ERR_STATE;
  }
  return launch_lusol(h, y, x);
}

// the grid counter of k_dot2 and k_cgs_xr_dot, in the CGS scalar block
int* red_counter(Handle* h) { return reinterpret_cast<int*>(h->d_scal + 10); }

// dot products: out[0] = sum(a*b) (conj a if ca), out[1] = sum(a2*b2)
int run_dot2(Handle* h, const double2* a, const double2* b, int ca, const double2* a2, const double2* b2, int ca2,
             double2* out) {
  hipLaunchKernelGGL(k_dot2, dim3(kRedBlocks), dim3(256), 0, h->stream, h->ne, a, b, ca, a2, b2, ca2, h->d_part,
                     red_counter(h), out);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

}  // namespace

// LW_ERR_STATE if a dataflow triangular solve on this handle gave up a wait
// (it then finished with wrong values; it cannot happen with a factor whose
// level sets were built here); the flag is cleared.
// The least s with sqrt(s) >= t. A correctly rounded square root is
// monotone, so for s >= 0 (and NaN) s < sq_threshold(t) exactly when
// sqrt(s) < t: the CGS stopping test without a square root, the same on the
// host and on the device.
double sq_threshold(double t) {
  if (!(t > 0.0)) return 0.0;  // sqrt(s) < t is false for every s >= 0
  if (std::isinf(t)) return t;
  double c = t * t;
  if (std::sqrt(c) >= t) {
    while (c > 0.0 && std::sqrt(std::nextafter(c, 0.0)) >= t) c = std::nextafter(c, 0.0);
  } else {
    while (std::sqrt(c) < t) c = std::nextafter(c, HUGE_VAL);
  }
  return c;
}

int lusol_status(Handle* h) {
  int e = 0;
  LW_HIP(hipMemcpyAsync(&e, h->d_lusol_err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  if (!e) return LW_OK;
  LW_HIP(hipMemsetAsync(h->d_lusol_err, 0, sizeof(int), h->stream));
  set_error("lusol0: a triangular-solve wait reached its limit");
  return LW_ERR_STATE;
}

namespace {

int check_handle(lw_handle* hh) {
  if (!hh) {
    set_error("null handle");
    return LW_ERR_ARG;
  }
  return LW_OK;
}

}  // namespace

// ---------------------------------------------------------------- init

int upload_factor(Handle* h);

// calc_RCS (modRCS.f90:109-153): the far-zone distance r_zl = 10 max(2 D^2 /
// lambda, 50 D, 20 lambda) and the centre of the bounding box of the edge
// centres (limites_dispersor), D its diagonal.
static void rcs_far_zone(const double* ec, int ne, double w, double& rzl, double* roo) {
  int lim[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 1; i < ne; ++i) {
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
  const double lam = 2 * kPi / (w * std::sqrt(kMu * kEps));
  const double D = std::sqrt(((mx - mnx) * (mx - mnx) + (my - mny) * (my - mny)) + (mz - mnz) * (mz - mnz));
  rzl = 2 * (D * D) / lam;
  if (50 * D > rzl) rzl = 50 * D;
  if (20 * lam > rzl) rzl = 20 * lam;
  rzl = rzl * 10;
  roo[0] = 0.5 * (mx + mnx);
  roo[1] = 0.5 * (my + mny);
  roo[2] = 0.5 * (mz + mnz);
}

static int init_device_state(Handle* h, const Geometry& g) {
  Plan& P = h->plan;
  const int L = P.L;
  h->L = L;
  h->ne = P.ne;
  h->nnz = P.nnz;
  LW_HIP(hipGetDevice(&h->device));
  LW_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  // the matvec's side streams (default priority) and their events
  for (auto& s : h->side) LW_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  for (hipEvent_t* e : {&h->ev_fork, &h->ev_near, &h->ev_ta, &h->ev_tb})
    LW_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
  h->ev_c.assign(L + 1, nullptr);
  for (int n = 3; n <= L; ++n) LW_HIP(hipEventCreateWithFlags(&h->ev_c[n], hipEventDisableTiming));
  int rc;
#define TRY(x)          \
  do {                  \
    rc = (x);           \
    if (rc) return rc;  \
  } while (0)
  // geometry
  auto vec_d = [](const double* p, size_t n) { return std::vector<double>(p, p + n); };
  auto vec_i = [](const int* p, size_t n) { return std::vector<int>(p, p + n); };
  TRY(upload(h, &h->d_p, vec_d(g.p_coord, 3 * (size_t)g.np)));
  TRY(upload(h, &h->d_tb, vec_d(g.t_baric, 3 * (size_t)g.nt)));
  TRY(upload(h, &h->d_tbs, vec_d(g.t_baric_sub, 27 * (size_t)g.nt)));
  TRY(upload(h, &h->d_el, vec_d(g.e_long, (size_t)g.ne)));
  TRY(upload(h, &h->d_ec, vec_d(g.e_centro, 3 * (size_t)g.ne)));
  TRY(upload(h, &h->d_et, vec_i(g.e_t, 2 * (size_t)g.ne)));
  TRY(upload(h, &h->d_epo, vec_i(g.e_po, 2 * (size_t)g.ne)));
  // levels
  h->lv.assign(L + 1, DevLevel());
  double* d_binom2 = nullptr;
  TRY(upload(h, &d_binom2, P.binom2));
  for (int n = 3; n <= L; ++n) {
    LevelPlan& lp = P.lv[n];
    DevLevel& d = h->lv[n];
    d.nk = lp.nk;
    d.nb = lp.nb;
    d.nT = lp.nT;
    d.npairs = (long long)lp.far_box.size();
    TRY(upload(h, &d.kap, lp.kap));
    TRY(upload(h, &d.pes, lp.pes));
    TRY(upload(h, &d.centre, lp.centre));
    std::vector<int> par(lp.parent.begin(), lp.parent.end());
    TRY(upload(h, &d.parent, par));
    if (n < L) TRY(upload(h, &d.child_ptr, lp.child_ptr));
    TRY(dmalloc(h, &d.T, (size_t)lp.nk * std::max(lp.nT, 1)));
    const size_t fsz = (size_t)lp.nk * 2 * lp.nb;
    TRY(dmalloc(h, &d.C, fsz));
    TRY(dmalloc(h, &d.B, fsz));
    TRY(dmalloc(h, &d.A, fsz));
    if (n < L) {
      TRY(upload(h, &d.wint, lp.wint));
      TRY(upload(h, &d.iint, lp.iint));
      {
        // the transposed stencil in blocks of kWave fine samples, entry-major
        // within a block (entry u of sample s at wt_ptr[s / kWave] +
        // u * kWave + s % kWave), each block padded with zero weights (row
        // 0) to its longest run, so that a wave's entry loads are contiguous
        const int nkf = (int)lp.wt_ptr.size() - 1, nbk = (nkf + kWave - 1) / kWave;
        std::vector<int> off(nbk + 1, 0);
        for (int bk = 0; bk < nbk; ++bk) {
          int m = 0;
          for (int q = bk * kWave; q < std::min(nkf, (bk + 1) * kWave); ++q)
            m = std::max(m, lp.wt_ptr[q + 1] - lp.wt_ptr[q]);
          off[bk + 1] = off[bk] + m * kWave;
        }
        std::vector<WtEntry> wt(off[nbk], WtEntry{0.0, 0, 0});
        for (int q = 0; q < nkf; ++q)
          for (int e = lp.wt_ptr[q]; e < lp.wt_ptr[q + 1]; ++e)
            wt[off[q / kWave] + (e - lp.wt_ptr[q]) * kWave + q % kWave] = WtEntry{lp.wt_w[e], lp.wt_row[e], 0};
        TRY(upload(h, &d.wt_ptr, off));
        TRY(upload(h, &d.wt, wt));
        if (2 * lp.nk * (int)sizeof(double2) > kDownLdsBytes) {
          // LDS windows for the downward pass: per chunk of 256 fine
          // samples, the shortest cyclic range of phi columns holding every
          // row its stencils read (the complement of the longest cyclic run
          // of unread columns), and the rows renumbered within it (padding
          // entries, weight 0, read window row 0)
          const int nch = (nkf + 255) / 256, nth = lp.nth, nph = lp.nph;
          std::vector<int> win(2 * (size_t)nch);
          std::vector<WtEntry> wtw(wt);
          int maxc = 0;
          for (int ch = 0; ch < nch; ++ch) {
            const int q1 = std::min(nkf, (ch + 1) * 256);
            std::vector<char> used(nph, 0);
            for (int q = ch * 256; q < q1; ++q)
              for (int e = lp.wt_ptr[q]; e < lp.wt_ptr[q + 1]; ++e) used[lp.wt_row[e] / nth] = 1;
            int best = 0, bstart = 0, run = 0;
            for (int j = 0; j < 2 * nph; ++j) {
              run = used[j % nph] ? 0 : run + 1;
              if (run > best && run <= nph) {
                best = run;
                bstart = j - run + 1;
              }
            }
            const int start = best >= nph ? 0 : (bstart + best) % nph, cnt = nph - std::min(best, nph);
            win[2 * ch] = start;
            win[2 * ch + 1] = cnt;
            maxc = std::max(maxc, cnt);
            for (int q = ch * 256; q < q1; ++q) {
              const int base = off[q / kWave] + q % kWave, ne = lp.wt_ptr[q + 1] - lp.wt_ptr[q];
              const int m = (off[q / kWave + 1] - off[q / kWave]) / kWave;
              for (int u = 0; u < m; ++u) {
                WtEntry& t = wtw[base + (size_t)u * kWave];
                t.r = u < ne ? ((t.r / nth - start + nph) % nph) * nth + t.r % nth : 0;
              }
            }
          }
          const long long lds = 2LL * maxc * nth * (long long)sizeof(double2);
          if (lds <= kDownLdsBytes) {
            d.win_lds = (int)lds;
            TRY(upload(h, &d.win, win));
            TRY(upload(h, &d.wtw, wtw));
          }
        }
      }
      TRY(dmalloc(h, &d.ph, 8 * (size_t)lp.nk));
      hipLaunchKernelGGL(k_phase_table64, dim3(blocks_for(lp.nk, 256)), dim3(256), 0, h->stream, lp.nk, d.kap,
                         lp.lado / 4.0, 2.0 * kPi / P.phys.lambda, d.ph);
    }
    if (n > 3) {
      const LevelPlan& pp = P.lv[n - 1];
      std::vector<int> oc(lp.nb);
      for (int c = 0; c < lp.nb; ++c) {
        const double* cc = &lp.centre[3 * c];
        const double* pc = &pp.centre[3 * lp.parent[c]];
        oc[c] = ((cc[0] > pc[0]) ? 4 : 0) | ((cc[1] > pc[1]) ? 2 : 0) | ((cc[2] > pc[2]) ? 1 : 0);
      }
      TRY(upload(h, &d.oct, oc));
    }
    double* d_tu = nullptr;
    double* d_th = nullptr;
    int* d_tL = nullptr;
    TRY(upload(h, &d_tu, lp.trans_u));
    TRY(upload(h, &d_th, lp.trans_h));
    TRY(upload(h, &d_tL, lp.trans_L));
    TRY(launch_setup_transfer(h, n, d_binom2, d_tu, d_tL, d_th, lp.hstride));
  }
  // the translation's parent blocks
  TRY(setup_far_parent(h));
  // level 3's upward output and level 2 have no parents: nothing else to upload
  // leaf ordering and patterns
  TRY(upload(h, &h->d_slot_basis, P.slot_basis));
  TRY(upload(h, &h->d_basis_slot, P.basis_slot));
  TRY(upload(h, &h->d_basis_box, P.basis_box));
  TRY(upload(h, &h->d_box_slot, P.box_slot));
  TRY(dmalloc(h, &h->d_rad, (size_t)P.lv[L].nk * 2 * P.ne));
  TRY(launch_setup_rad(h));
  TRY(launch_setup_phases(h, g.e_t));
  // near field
  TRY(upload(h, &h->d_ia, P.ia));
  TRY(upload(h, &h->d_ja, P.ja));
  if (h->fp64_near) {
    double2* z = nullptr;
    TRY(dmalloc(h, &z, (size_t)P.nnz));
    h->d_z = z;
  } else {
    float2* z = nullptr;
    TRY(dmalloc(h, &z, (size_t)P.nnz));
    h->d_z = z;
  }
  TRY(launch_setup_near(h));

  // work vectors and reduction buffers
  TRY(dmalloc(h, &h->d_x, (size_t)P.ne));
  TRY(dmalloc(h, &h->d_y, (size_t)P.ne));
  TRY(dmalloc(h, &h->d_ynear, (size_t)P.ne));
  for (auto& w : h->d_w) TRY(dmalloc(h, &w, (size_t)P.ne));
  TRY(dmalloc(h, &h->d_part, 4 * (size_t)kRedBlocks));
  // the RCS current (lw_rcs_init copies into it), the far-zone distance and
  // the static scratch of the allocation-free RCS entry point
  TRY(dmalloc(h, &h->d_cur, (size_t)P.ne));
  rcs_far_zone(g.e_centro, P.ne, P.phys.w, h->rcs_rzl, h->rcs_centre);
  TRY(rcs_static_setup(h));
  TRY(dmalloc(h, &h->d_lusol_err, 1));
  LW_HIP(hipMemsetAsync(h->d_lusol_err, 0, sizeof(int), h->stream));
  h->scal_cap = 64;
  TRY(dmalloc(h, &h->d_scal, (size_t)h->scal_cap));
  // CGS scalars, stop flag and grid counter start at zero
  LW_HIP(hipMemsetAsync(h->d_scal, 0, sizeof(double2) * h->scal_cap, h->stream));
  LW_HIP(hipHostMalloc((void**)&h->h_cgs, 4 * sizeof(double2), hipHostMallocDefault));
  LW_HIP(hipHostGetDevicePointer((void**)&h->d_cgs_host, h->h_cgs, 0));
  for (hipEvent_t& e : h->ev_cgs) LW_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  LW_HIP(hipStreamSynchronize(h->stream));
  // large LDS requests for the downward kernel at the top level
  int maxk = 0;
  for (int n = 3; n <= L; ++n) maxk = std::max(maxk, P.lv[n].nk);
  if (2 * maxk * (int)sizeof(double2) > 160 * 1024) {
    set_error("lw_mlfma_init: sample count exceeds LDS capacity of the pass kernels");
    return LW_ERR_ARG;
  }
  LW_HIP(hipFuncSetAttribute((const void*)k_interp_up, hipFuncAttributeMaxDynamicSharedMemorySize,
                             2 * maxk * (int)sizeof(double2)));
  if (h->flags & LW_PATTERN_ORDER) h->recv_lds = h->agg_lds = 0;  // the pattern-by-pattern kernels
  if (h->recv_lds) {
    LW_HIP(hipFuncSetAttribute((const void*)k_receive_ph, hipFuncAttributeMaxDynamicSharedMemorySize, h->recv_lds));
    LW_HIP(hipFuncSetAttribute((const void*)k_aggregate_fac, hipFuncAttributeMaxDynamicSharedMemorySize,
                               std::max(h->agg_lds, 16)));
  }
  {
    // the near product by leaf box needs the same near columns in every row
    // of a box, and the box's x entries in LDS
    int maxc = 0;
    bool same = true;
    h->near_box_cols = 0;
    for (int b = 0; b < P.lv[L].nb && same; ++b) {
      const int m0 = P.slot_basis[P.box_slot[b]], n0 = P.ia[m0 + 1] - P.ia[m0];
      maxc = std::max(maxc, n0);
      h->near_box_cols += n0;
      for (int s = P.box_slot[b] + 1; s < P.box_slot[b + 1] && same; ++s) {
        const int m = P.slot_basis[s];
        same = P.ia[m + 1] - P.ia[m] == n0 &&
               std::equal(P.ja.begin() + P.ia[m], P.ja.begin() + P.ia[m + 1], P.ja.begin() + P.ia[m0]);
      }
    }
    const long long nl = (long long)maxc * sizeof(double2);
    h->near_lds = (same && h->recv_lds && nl <= kRecvLdsMax) ? (int)std::max<long long>(nl, 16) : 0;
    if (h->near_lds) {
      LW_HIP(hipFuncSetAttribute((const void*)k_near_box<float2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 h->near_lds));
      LW_HIP(hipFuncSetAttribute((const void*)k_near_box<double2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 h->near_lds));
    }
  }
#undef TRY
  return LW_OK;
}

}  // namespace lw

using lw::Handle;

extern "C" {

const char* lw_last_error(void) { return lw::g_err.c_str(); }

int lw_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

static void free_handle(lw_handle* hh) {
  if (!hh) return;
  Handle* h = &hh->h;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (hipStream_t s : h->side)
    if (s) (void)hipStreamSynchronize(s);
  for (void* p : h->allocs) (void)hipFree(p);
  for (hipEvent_t e : {h->ev_fork, h->ev_near, h->ev_ta, h->ev_tb})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->ev_c)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->ev_cgs)
    if (e) (void)hipEventDestroy(e);
  if (h->h_cgs) (void)hipHostFree(h->h_cgs);
  for (hipStream_t s : h->side)
    if (s) (void)hipStreamDestroy(s);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete hh;
}

int lw_mlfma_init(double lambda, double ladomax, int flags, int num_e, const double* e_centro, const int* e_t,
                  const int* e_po, const double* e_long, int num_t, const double* t_area, const double* t_baric,
                  const double* t_baric_sub, int num_p, const double* p_coord, const int* t_p, const int* e_p,
                  const double* t_normal, lw_handle** out) {
  if (!out || !e_centro || !e_t || !e_po || !e_long || !t_baric || !t_baric_sub || !p_coord || num_e < 2 ||
      num_t < 2 || num_p < 3) {
    lw::set_error("lw_mlfma_init: missing geometry array or empty mesh");
    return LW_ERR_ARG;
  }
  if (flags & ~(LW_NEAR_FP64 | LW_PATTERN_ORDER)) {
    lw::set_error("lw_mlfma_init: unknown flags");
    return LW_ERR_ARG;
  }
  *out = nullptr;
  lw::Geometry g;
  g.ne = num_e;
  g.nt = num_t;
  g.np = num_p;
  g.e_centro = e_centro;
  g.e_long = e_long;
  g.t_area = t_area;
  g.t_baric = t_baric;
  g.t_baric_sub = t_baric_sub;
  g.p_coord = p_coord;
  g.t_normal = t_normal;
  g.e_t = e_t;
  g.e_po = e_po;
  g.t_p = t_p;
  g.e_p = e_p;
  for (int i = 0; i < 2 * num_e; ++i)
    if (e_t[i] < 1 || e_t[i] > num_t || e_po[i] < 1 || e_po[i] > num_p) {
      lw::set_error("lw_mlfma_init: connectivity index out of range");
      return LW_ERR_ARG;
    }
  lw_handle* hh = new lw_handle();
  Handle* h = &hh->h;
  h->flags = flags;
  h->fp64_near = (flags & LW_NEAR_FP64) != 0;
  std::string msg;
  // the octree and the host tables come first (host code), so that argument
  // and octree errors are reported whether or not a device is present
  int rc = lw::build_plan(g, lambda, ladomax, h->plan, msg);
  if (rc) {
    lw::set_error(msg);
    delete hh;
    return rc;
  }
  if (lw_device_count() < 1) {
    lw::set_error("lw_mlfma_init: no HIP device visible");
    delete hh;
    return LW_ERR_NODEV;
  }
  rc = lw::init_device_state(h, g);
  if (rc) {
    free_handle(hh);
    return rc;
  }
  *out = hh;
  return LW_OK;
}

int lw_mlfma_destroy(lw_handle* hh) {
  free_handle(hh);
  return LW_OK;
}

int lw_mlfma_nlevels(lw_handle* hh) { return hh ? hh->h.L : LW_ERR_ARG; }

int lw_mlfma_level_info(lw_handle* hh, int level, int* nboxes, int* nmult, int* nth, int* nph, int* ntransf,
                        long long* npairs) {
  if (lw::check_handle(hh)) return LW_ERR_ARG;
  Handle* h = &hh->h;
  if (level < 1 || level > h->L) return LW_ERR_ARG;
  const lw::LevelPlan& lp = h->plan.lv[level];
  if (nboxes) *nboxes = lp.nb;
  if (nmult) *nmult = lp.nmult;
  if (nth) *nth = lp.nth;
  if (nph) *nph = lp.nph;
  if (ntransf) *ntransf = lp.nT;
  if (npairs) *npairs = (long long)lp.far_box.size();
  return LW_OK;
}

int lw_matvec_dev(lw_handle* hh, const void* x_dev, void* y_dev) {
  if (lw::check_handle(hh) || !x_dev || !y_dev) return LW_ERR_ARG;
  return lw::run_matvec(&hh->h, (const double2*)x_dev, (double2*)y_dev, nullptr, nullptr);
}

int lw_matvec(lw_handle* hh, const double* x, double* y) {
  if (lw::check_handle(hh) || !x || !y) return LW_ERR_ARG;
  Handle* h = &hh->h;
  const size_t bytes = sizeof(double2) * h->ne;
  LW_HIP(hipMemcpyAsync(h->d_x, x, bytes, hipMemcpyHostToDevice, h->stream));
  int rc = lw::run_matvec(h, h->d_x, h->d_y, nullptr, nullptr);
  if (rc) return rc;
  LW_HIP(hipMemcpyAsync(y, h->d_y, bytes, hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  return LW_OK;
}

int lw_matvec_parts(lw_handle* hh, const double* x, double* y_far, double* y_near) {
  if (lw::check_handle(hh) || !x || !y_far || !y_near) return LW_ERR_ARG;
  Handle* h = &hh->h;
  const size_t bytes = sizeof(double2) * h->ne;
  LW_HIP(hipMemcpyAsync(h->d_x, x, bytes, hipMemcpyHostToDevice, h->stream));
  int rc = lw::run_matvec(h, h->d_x, h->d_y, h->d_w[0], nullptr);
  if (rc) return rc;
  LW_HIP(hipMemcpyAsync(y_far, h->d_y, bytes, hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipMemcpyAsync(y_near, h->d_w[0], bytes, hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  return LW_OK;
}

long long lw_near_nnz(lw_handle* hh) { return hh ? hh->h.nnz : LW_ERR_ARG; }

int lw_near_csr(lw_handle* hh, int* ia, int* ja, float* z) {
  if (lw::check_handle(hh)) return LW_ERR_ARG;
  Handle* h = &hh->h;
  if (h->fp64_near && z) {
    lw::set_error("lw_near_csr: handle stores complex(8) values; use lw_near_csr64");
    return LW_ERR_ARG;
  }
  if (ia)
    for (int i = 0; i <= h->ne; ++i) ia[i] = h->plan.ia[i] + 1;
  if (ja)
    for (long long k = 0; k < h->nnz; ++k) ja[k] = h->plan.ja[k] + 1;
  if (z) {
    LW_HIP(hipMemcpyAsync(z, h->d_z, sizeof(float2) * h->nnz, hipMemcpyDeviceToHost, h->stream));
    LW_HIP(hipStreamSynchronize(h->stream));
  }
  return LW_OK;
}

int lw_near_csr64(lw_handle* hh, int* ia, int* ja, double* z) {
  if (lw::check_handle(hh)) return LW_ERR_ARG;
  Handle* h = &hh->h;
  if (ia)
    for (int i = 0; i <= h->ne; ++i) ia[i] = h->plan.ia[i] + 1;
  if (ja)
    for (long long k = 0; k < h->nnz; ++k) ja[k] = h->plan.ja[k] + 1;
  if (z) {
    if (h->fp64_near) {
      LW_HIP(hipMemcpyAsync(z, h->d_z, sizeof(double2) * h->nnz, hipMemcpyDeviceToHost, h->stream));
      LW_HIP(hipStreamSynchronize(h->stream));
    } else {
      std::vector<float> tmp(2 * (size_t)h->nnz);
      LW_HIP(hipMemcpyAsync(tmp.data(), h->d_z, sizeof(float2) * h->nnz, hipMemcpyDeviceToHost, h->stream));
      LW_HIP(hipStreamSynchronize(h->stream));
      for (size_t k = 0; k < tmp.size(); ++k) z[k] = tmp[k];
    }
  }
  return LW_OK;
}

int lw_ilu0(lw_handle* hh, int* ierr) {
  if (lw::check_handle(hh)) return LW_ERR_ARG;
  Handle* h = &hh->h;
  h->have_ilu = false;
  int piv = 0;
  if (int rc = lw::ilu0_device(h, &piv)) return rc;
  if (ierr) *ierr = piv;
  if (piv) {
    lw::set_error("ilu0: zero pivot at row " + std::to_string(piv));
    return LW_ERR_PIVOT;
  }
  h->have_ilu = true;
  return LW_OK;
}

}  // extern "C"

namespace lw {
// Device arrays of the MSR factor's structure in h->h_jlu/h_ju (1-based with
// a pad element so alu(k) is d_alu[k]; the value array is allocated here) and
// the level sets of both sweeps.
int upload_structure(Handle* h) {
  const int n = h->ne;
  const size_t na = (size_t)h->nnz + 2;
  if (!h->d_alu) {
    if (h->fp64_near) {
      double2* p = nullptr;
      if (int rc = lw::dmalloc(h, &p, na)) return rc;
      h->d_alu = p;
    } else {
      float2* p = nullptr;
      if (int rc = lw::dmalloc(h, &p, na)) return rc;
      h->d_alu = p;
    }
    if (int rc = lw::dmalloc(h, &h->d_jlu, na)) return rc;
    std::vector<int> ju1((size_t)n + 1, 0);
    if (int rc = lw::dmalloc(h, &h->d_ju, ju1.size())) return rc;
  }
  LW_HIP(hipMemcpyAsync(h->d_jlu, h->h_jlu.data(), sizeof(int) * na, hipMemcpyHostToDevice, h->stream));
  std::vector<int> ju1((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) ju1[i + 1] = h->h_ju[i];
  LW_HIP(hipMemcpyAsync(h->d_ju, ju1.data(), sizeof(int) * ju1.size(), hipMemcpyHostToDevice, h->stream));
  std::vector<int> lptr, lrows, uptr, urows;
  lw::tri_levels(n, h->h_jlu, h->h_ju, false, lptr, lrows);
  lw::tri_levels(n, h->h_jlu, h->h_ju, true, uptr, urows);
  h->nlev_l = (int)lptr.size() - 1;
  h->nlev_u = (int)uptr.size() - 1;
  h->h_lptr = lptr;
  if (int rc = lw::lusol_flow_setup(h)) return rc;
  if (int rc = lw::upload(h, &h->d_lptr, lptr)) return rc;
  if (int rc = lw::upload(h, &h->d_lrows, lrows)) return rc;
  if (int rc = lw::upload(h, &h->d_uptr, uptr)) return rc;
  if (int rc = lw::upload(h, &h->d_urows, urows)) return rc;
  LW_HIP(hipStreamSynchronize(h->stream));
  return LW_OK;
}

// Device copies of a complete host MSR factor (structure and values).
int upload_factor(Handle* h) {
  h->have_ilu = false;
  if (int rc = upload_structure(h)) return rc;
  const size_t na = (size_t)h->nnz + 2;
  if (h->fp64_near)
    LW_HIP(hipMemcpyAsync(h->d_alu, h->h_alu64.data(), sizeof(double2) * na, hipMemcpyHostToDevice, h->stream));
  else
    LW_HIP(hipMemcpyAsync(h->d_alu, h->h_alu32.data(), sizeof(float2) * na, hipMemcpyHostToDevice, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  h->have_ilu = true;
  return LW_OK;
}
}  // namespace lw

extern "C" {

int lw_ilu0_set(lw_handle* hh, const float* alu, const double* alu64, const int* jlu, const int* ju) {
  if (lw::check_handle(hh) || !jlu || !ju) return LW_ERR_ARG;
  Handle* h = &hh->h;
  const long long nnz = h->nnz;
  const int n = h->ne;
  if (jlu[0] != n + 2 || jlu[n] - 1 > nnz + 1 || jlu[n] < jlu[0]) {
    lw::set_error("lw_ilu0_set: jlu does not describe an MSR factor of this operator");
    return LW_ERR_ARG;
  }
  // every check runs on the caller's arrays before the handle is touched, so
  // a rejected factor leaves the previous one (and its host copy) in place
  for (int i = 0; i < n; ++i)
    if (jlu[i + 1] < jlu[i] || ju[i] < jlu[i] || ju[i] > jlu[i + 1]) {
      lw::set_error("lw_ilu0_set: ju out of range");
      return LW_ERR_ARG;
    }
  for (size_t k = (size_t)n + 2; k <= (size_t)jlu[n] - 1; ++k)
    if (jlu[k - 1] < 1 || jlu[k - 1] > n) {
      lw::set_error("lw_ilu0_set: column index out of range");
      return LW_ERR_ARG;
    }
  if (h->fp64_near ? !alu64 : !alu) {
    lw::set_error(h->fp64_near ? "lw_ilu0_set: the handle stores complex(8) factors: pass alu64"
                               : "lw_ilu0_set: the handle stores complex(4) factors: pass alu");
    return LW_ERR_ARG;
  }
  const size_t na = (size_t)nnz + 1;
  h->h_jlu.assign(na + 1, 0);
  std::memcpy(h->h_jlu.data() + 1, jlu, sizeof(int) * na);
  h->h_ju.assign(ju, ju + n);
  if (h->fp64_near) {
    h->h_alu64.assign(2 * (na + 1), 0.0);
    std::memcpy(h->h_alu64.data() + 2, alu64, sizeof(double) * 2 * na);
  } else {
    h->h_alu32.assign(2 * (na + 1), 0.0f);
    std::memcpy(h->h_alu32.data() + 2, alu, sizeof(float) * 2 * na);
  }
  return lw::upload_factor(h);
}

int lw_ilu0_get(lw_handle* hh, float* alu, double* alu64, int* jlu, int* ju) {
  if (lw::check_handle(hh)) return LW_ERR_ARG;
  Handle* h = &hh->h;
  if (!h->have_ilu) {
    lw::set_error("lw_ilu0_get: no factor");
    return LW_ERR_STATE;
  }
  const size_t na = (size_t)h->nnz + 1;
  if (alu) {
    if (h->fp64_near) return LW_ERR_ARG;
    std::memcpy(alu, h->h_alu32.data() + 2, sizeof(float) * 2 * na);
  }
  if (alu64) {
    if (h->fp64_near)
      std::memcpy(alu64, h->h_alu64.data() + 2, sizeof(double) * 2 * na);
    else
      for (size_t k = 0; k < 2 * na; ++k) alu64[k] = h->h_alu32[2 + k];
  }
  if (jlu) std::memcpy(jlu, h->h_jlu.data() + 1, sizeof(int) * na);
  if (ju) std::memcpy(ju, h->h_ju.data(), sizeof(int) * h->ne);
  return LW_OK;
}

int lw_lusol0(lw_handle* hh, const double* y, double* x) {
  if (lw::check_handle(hh) || !y || !x) return LW_ERR_ARG;
  Handle* h = &hh->h;
  const size_t bytes = sizeof(double2) * h->ne;
  LW_HIP(hipMemcpyAsync(h->d_x, y, bytes, hipMemcpyHostToDevice, h->stream));
  int rc = lw::run_lusol(h, h->d_x, h->d_y);
  if (rc) return rc;
  LW_HIP(hipMemcpyAsync(x, h->d_y, bytes, hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  return lw::lusol_status(h);
}

int lw_cgs(lw_handle* hh, const double* b, const double* guess, double errmax, int use_prec, int maxit, double* x,
           int* iters, double* relres, int* converged) {
  if (lw::check_handle(hh) || !b || !x || maxit < 1) return LW_ERR_ARG;
  Handle* h = &hh->h;
  if (use_prec && !h->have_ilu) {
    lw::set_error("lw_cgs: preconditioner requested but no ILU0 factor");
    return LW_ERR_STATE;
  }
  const int n = h->ne;
  const size_t bytes = sizeof(double2) * n;
  hipStream_t st = h->stream;
  // without the preconditioner P_ = p and U_ = u + q = t (modIterativo.f90:408, 424): the
  // matvecs read p and t themselves
  double2 *d_b = h->d_w[0], *r = h->d_w[1], *R_ = h->d_w[2], *u = h->d_w[3], *p = h->d_w[4], *V = h->d_w[6],
          *q = h->d_w[7], *Q = h->d_w[9], *t = h->d_w[10], *xv = h->d_x;
  double2 *P_ = use_prec ? h->d_w[5] : p, *U = use_prec ? h->d_w[8] : t;
  // device scalars: [0] rho1 [1] rho2 [2] den [3] alfa [4] |r|^2 or |b|^2 [5] next rho,
  // and the stop flag of k_cgs_xr_dot
  double2* s = h->d_scal;
  int* stop = reinterpret_cast<int*>(s + 8);
  LW_HIP(hipMemcpyAsync(d_b, b, bytes, hipMemcpyHostToDevice, st));
  if (guess)
    LW_HIP(hipMemcpyAsync(xv, guess, bytes, hipMemcpyHostToDevice, st));
  else
    LW_HIP(hipMemsetAsync(xv, 0, bytes, st));
  int rc = lw::run_matvec(h, xv, t, nullptr, nullptr);
  if (rc) return rc;
  const int vb = lw::blocks_for(n, 256);
  hipLaunchKernelGGL(lw::k_sub, dim3(vb), dim3(256), 0, st, n, d_b, t, r);
  LW_HIP(hipMemcpyAsync(R_, r, bytes, hipMemcpyDeviceToDevice, st));
  // |b|^2 (conjugated) and rho1 = sum(r*R_) (unconjugated), then |r|^2 of
  // the initial residual, so that every exit reports the residual of the
  // iterate it returns
  if ((rc = lw::run_dot2(h, d_b, d_b, 1, r, R_, 0, s + 4))) return rc;
  if ((rc = lw::run_dot2(h, r, r, 1, nullptr, nullptr, 0, s + 6))) return rc;
  double2 hs[3];
  LW_HIP(hipMemcpyAsync(hs, s + 4, sizeof(hs), hipMemcpyDeviceToHost, st));
  LW_HIP(hipStreamSynchronize(st));
  const double nb = std::sqrt(hs[0].x);
  // ||r|| / ||b||; a zero right-hand side counts a zero residual as 0
  auto relative = [nb](double nr) { return nb > 0.0 ? nr / nb : (nr == 0.0 ? 0.0 : HUGE_VAL); };
  double r2 = hs[2].x, nr = std::sqrt(r2);
  double rr = relative(nr);
  const double thr = lw::sq_threshold(errmax * nb);
  LW_HIP(hipMemcpyAsync(s + 0, s + 5, sizeof(double2), hipMemcpyDeviceToDevice, st));
  LW_HIP(hipMemsetAsync(stop, 0, sizeof(int), st));
  // one iteration (modIterativo.f90:396-437), its |r|^2 and next rho written
  // to pinned slot i % 2, then the event of that slot
  auto iteration = [&](int i) -> int {
    hipLaunchKernelGGL(lw::k_cgs_up, dim3(vb), dim3(256), 0, st, n, i == 1 ? 1 : 0, s + 0, s + 1, r, q, u, p);
    int e;
    if (use_prec && (e = lw::run_lusol(h, p, P_))) return e;
    if ((e = lw::run_matvec(h, P_, V, nullptr, nullptr))) return e;
    if ((e = lw::run_dot2(h, R_, V, 0, nullptr, nullptr, 0, s + 2))) return e;
    hipLaunchKernelGGL(lw::k_cgs_q, dim3(vb), dim3(256), 0, st, n, s + 0, s + 2, s + 3, u, V, q, t);
    if (use_prec && (e = lw::run_lusol(h, t, U))) return e;
    if ((e = lw::run_matvec(h, U, Q, nullptr, nullptr))) return e;
    hipLaunchKernelGGL(lw::k_cgs_xr_dot, dim3(lw::kRedBlocks), dim3(256), 0, st, n, U, Q, xv, r, R_, h->d_part,
                       lw::red_counter(h), thr, s, stop, h->d_cgs_host + 2 * (i & 1));
    LW_HIP(hipGetLastError());
    LW_HIP(hipEventRecord(h->ev_cgs[i & 1], st));
    return LW_OK;
  };
  int it = 1;
  bool conv = false;
  // the breakdown exit (modIterativo.f90:392-395) keeps the last iterate, which
  // has converged if its residual passes the stopping test (an exact guess, or
  // a zero right-hand side from a zero guess)
  if (hs[1].x == 0.0 && hs[1].y == 0.0) {
    conv = r2 < thr || nr == 0.0;
  } else {
    if ((rc = iteration(1))) return rc;
    for (;;) {
      // iteration it + 1 is queued before iteration it is read back, so the
      // GPU does not wait for the host; when the host stops after it,
      // k_cgs_xr_dot has set the flag that makes it leave x alone
      if (it < maxit && (rc = iteration(it + 1))) return rc;
      LW_HIP(hipEventSynchronize(h->ev_cgs[it & 1]));
      r2 = h->h_cgs[2 * (it & 1)].x;
      const double2 next = h->h_cgs[2 * (it & 1) + 1];
      nr = std::sqrt(r2);
      rr = relative(nr);
      if (r2 < thr) {
        conv = true;
        break;
      }
      if (lw::progress_every() > 0 && it % lw::progress_every() == 0)
        std::fprintf(stderr, "[lw_cgs] iteration %d: relres %.3e\n", it, rr);
      if (it == maxit) break;
      ++it;
      if (next.x == 0.0 && next.y == 0.0) {
        conv = nr == 0.0;
        break;
      }
    }
  }
  LW_HIP(hipMemcpyAsync(x, xv, bytes, hipMemcpyDeviceToHost, st));
  LW_HIP(hipStreamSynchronize(st));
  if (use_prec && (rc = lw::lusol_status(h))) return rc;
  if (iters) *iters = it;
  if (relres) *relres = rr;
  if (converged) *converged = conv ? 1 : 0;
  return LW_OK;
}

int lw_plane_wave(lw_handle* hh, const double* dir, const double* ctte, const double* pol, double* b) {
  if (lw::check_handle(hh) || !dir || !ctte || !pol || !b) return LW_ERR_ARG;
  Handle* h = &hh->h;
  const double nd = std::sqrt((dir[0] * dir[0] + dir[1] * dir[1]) + dir[2] * dir[2]);
  if (!(nd > 0)) return LW_ERR_ARG;
  const double k0 = h->plan.phys.w * std::sqrt(lw::kMu * lw::kEps);
  hipLaunchKernelGGL(lw::k_plane_wave, dim3(lw::blocks_for(h->ne, 256)), dim3(256), 0, h->stream, h->ne, h->d_tb,
                     h->d_p, h->d_et, h->d_epo, h->d_el, dir[0] / nd, dir[1] / nd, dir[2] / nd, k0,
                     make_double2(pol[0], pol[1]), make_double2(pol[2], pol[3]), make_double2(pol[4], pol[5]),
                     make_double2(ctte[0], ctte[1]), h->d_y);
  LW_HIP(hipGetLastError());
  LW_HIP(hipMemcpyAsync(b, h->d_y, sizeof(double2) * h->ne, hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  return LW_OK;
}

int lw_rcs_init(lw_handle* hh, const double* current, const double* pol, const double* ctte) {
  if (lw::check_handle(hh) || !current || !pol || !ctte) return LW_ERR_ARG;
  Handle* h = &hh->h;
  LW_HIP(hipMemcpyAsync(h->d_cur, current, sizeof(double2) * h->ne, hipMemcpyHostToDevice, h->stream));
  for (int c = 0; c < 3; ++c) h->pol[c] = make_double2(pol[2 * c], pol[2 * c + 1]);
  h->ctte = make_double2(ctte[0], ctte[1]);
  h->rcs_ready = true;
  return LW_OK;
}

int lw_rcs_sigma(lw_handle* hh, int n, const double* theta, const double* phi, double* sigma) {
  if (lw::check_handle(hh) || n < 0 || (n > 0 && (!theta || !phi || !sigma))) return LW_ERR_ARG;
  Handle* h = &hh->h;
  if (!h->rcs_ready) {
    lw::set_error("lw_rcs_sigma: call lw_rcs_init first");
    return LW_ERR_STATE;
  }
  if (n == 0) return LW_OK;
  // calc_RCS (modRCS.f90:109-153): far-zone distance from the bounding box of
  // the edge centres
  const lw::Geometry* gp = nullptr;
  (void)gp;
  const lw::Plan& P = h->plan;
  std::vector<double> ec(3 * (size_t)h->ne);
  LW_HIP(hipMemcpyAsync(ec.data(), h->d_ec, sizeof(double) * ec.size(), hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
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
  const double w = P.phys.w;
  const double lam = 2 * lw::kPi / (w * std::sqrt(lw::kMu * lw::kEps));
  const double D = std::sqrt(((mx - mnx) * (mx - mnx) + (my - mny) * (my - mny)) + (mz - mnz) * (mz - mnz));
  double rzl = 2 * (D * D) / lam;
  if (50 * D > rzl) rzl = 50 * D;
  if (20 * lam > rzl) rzl = 20 * lam;
  rzl = rzl * 10;
  const double roo[3] = {0.5 * (mx + mnx), 0.5 * (my + mny), 0.5 * (mz + mnz)};
  std::vector<double> ro(3 * (size_t)n);
  for (int a = 0; a < n; ++a) {
    ro[3 * a + 0] = roo[0] + (rzl * std::sin(theta[a]) * std::cos(phi[a]));
    ro[3 * a + 1] = roo[1] + (rzl * std::sin(theta[a]) * std::sin(phi[a]));
    ro[3 * a + 2] = roo[2] + (rzl * std::cos(theta[a]));
  }
  double* d_ro = nullptr;
  double2* d_E = nullptr;
  LW_HIP(hipMalloc(&d_ro, sizeof(double) * ro.size()));
  LW_HIP(hipMalloc(&d_E, sizeof(double2) * 3 * n));
  LW_HIP(hipMemcpyAsync(d_ro, ro.data(), sizeof(double) * ro.size(), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(lw::k_rcs, dim3(n), dim3(256), 0, h->stream, h->ne, h->d_tb, h->d_et, h->d_el, h->d_cur, d_ro,
                     w * std::sqrt(lw::kMu * lw::kEps), d_E);
  std::vector<double2> E(3 * (size_t)n);
  LW_HIP(hipMemcpyAsync(E.data(), d_E, sizeof(double2) * E.size(), hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  (void)hipFree(d_ro);
  (void)hipFree(d_E);
  // E *= j w mu / 4pi; sigma = 4 pi r^2 |E|^2 / |pol ctte|^2 (modRCS.f90:91-105)
  const double cm = w * lw::kMu / (4 * lw::kPi);
  double pn = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double re = h->pol[c].x * h->ctte.x - h->pol[c].y * h->ctte.y;
    const double im = h->pol[c].x * h->ctte.y + h->pol[c].y * h->ctte.x;
    pn += re * re + im * im;
  }
  for (int a = 0; a < n; ++a) {
    double e2 = 0.0;
    for (int c = 0; c < 3; ++c) {
      const double2 v = E[3 * (size_t)a + c];
      e2 += (v.x * v.x + v.y * v.y) * (cm * cm);
    }
    sigma[a] = e2 * 4 * lw::kPi * (rzl * rzl) / pn;
  }
  return LW_OK;
}

int lw_bench_matvec(lw_handle* hh, int reps, double* ms, double* stage_ms) {
  if (lw::check_handle(hh) || reps < 1 || !ms) return LW_ERR_ARG;
  Handle* h = &hh->h;
  hipEvent_t ev[lw::kMatvecStages + 1];
  for (auto& e : ev) LW_HIP(hipEventCreate(&e));
  // deterministic input
  std::vector<double2> xin(h->ne);
  for (int i = 0; i < h->ne; ++i) xin[i] = make_double2((i + 1) / 1.3823, 57.382 / (i + 1));
  LW_HIP(hipMemcpyAsync(h->d_x, xin.data(), sizeof(double2) * h->ne, hipMemcpyHostToDevice, h->stream));
  int rc = lw::run_matvec(h, h->d_x, h->d_y, nullptr, nullptr);
  if (rc) return rc;
  LW_HIP(hipStreamSynchronize(h->stream));
  // the stages one after the other, each bracketed by events
  double acc[lw::kMatvecStages] = {};
  for (int r = 0; r < reps; ++r) {
    if ((rc = lw::run_matvec(h, h->d_x, h->d_y, nullptr, ev))) return rc;
    LW_HIP(hipEventSynchronize(ev[lw::kMatvecStages]));
    for (int s = 0; s < lw::kMatvecStages; ++s) {
      float t = 0.f;
      LW_HIP(hipEventElapsedTime(&t, ev[s], ev[s + 1]));
      acc[s] += t;
    }
  }
  // the matvec as the solver runs it (side streams), back to back
  LW_HIP(hipEventRecord(ev[0], h->stream));
  for (int r = 0; r < reps; ++r)
    if ((rc = lw::run_matvec(h, h->d_x, h->d_y, nullptr, nullptr))) return rc;
  LW_HIP(hipEventRecord(ev[1], h->stream));
  LW_HIP(hipEventSynchronize(ev[1]));
  float t = 0.f;
  LW_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  *ms = t / reps;
  if (stage_ms)
    for (int s = 0; s < lw::kMatvecStages; ++s) stage_ms[s] = acc[s] / reps;
  return LW_OK;
}

int lw_bench_lusol(lw_handle* hh, int reps, double* ms) {
  if (lw::check_handle(hh) || reps < 1 || !ms) return LW_ERR_ARG;
  Handle* h = &hh->h;
  hipEvent_t e0, e1;
  LW_HIP(hipEventCreate(&e0));
  LW_HIP(hipEventCreate(&e1));
  int rc = lw::run_lusol(h, h->d_y, h->d_x);
  if (rc) return rc;
  LW_HIP(hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps; ++r)
    if ((rc = lw::run_lusol(h, h->d_y, h->d_x))) return rc;
  LW_HIP(hipEventRecord(e1, h->stream));
  LW_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  LW_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t / reps;
  return LW_OK;
}

int lw_bench_copy(lw_handle* hh, long long bytes, int reps, double* ms) {
  if (lw::check_handle(hh) || reps < 1 || !ms || bytes < 16 || bytes % 16) return LW_ERR_ARG;
  Handle* h = &hh->h;
  const long long n = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  LW_HIP(hipMalloc(&a, bytes));
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return LW_ERR_NOMEM;
  }
  hipEvent_t e0, e1;
  LW_HIP(hipEventCreate(&e0));
  LW_HIP(hipEventCreate(&e1));
  const int blocks = (int)std::min<long long>((n + 255) / 256, 1 << 20);
  hipLaunchKernelGGL(lw::k_copy, dim3(blocks), dim3(256), 0, h->stream, n, (const double2*)a, b);
  LW_HIP(hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(lw::k_copy, dim3(blocks), dim3(256), 0, h->stream, n, (const double2*)a, b);
  LW_HIP(hipEventRecord(e1, h->stream));
  LW_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  LW_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  *ms = t / reps;
  return LW_OK;
}

int lw_matvec_bytes(lw_handle* hh, double* total, double* aggregation, double* receive_near) {
  if (lw::check_handle(hh)) return LW_ERR_ARG;
  double sb[lw::kMatvecStages];
  int rc = lw_matvec_stage_bytes(hh, sb);
  if (rc) return rc;
  if (total) {
    *total = 0.0;
    for (double b : sb) *total += b;
  }
  if (aggregation) *aggregation = sb[0];
  if (receive_near) *receive_near = sb[4] + sb[5];
  return LW_OK;
}

int lw_matvec_stage_bytes(lw_handle* hh, double* sb) {
  if (lw::check_handle(hh) || !sb) return LW_ERR_ARG;
  Handle* h = &hh->h;
  const lw::Plan& P = h->plan;
  const int L = P.L;
  auto field = [&](int n) { return (double)P.lv[n].nk * 2 * P.lv[n].nb * 16.0; };
  // the phase rows (one per leaf box and triangle of its bases) and the
  // per-basis RWG vectors, row indices and slot order
  const double phases = (double)h->npairs * h->ph_len * 16.0 +
                        (double)P.ne * (sizeof(lw::RadGeom) + sizeof(int2) + sizeof(int));
  const double zb = h->fp64_near ? 16.0 : 8.0;
  const double csr = (zb + 4.0) * (double)P.nnz + 4.0 * (P.ne + 1);
  // aggregation: phase rows and RWG vectors, x, leaf campoC written
  sb[0] = phases + 16.0 * P.ne + field(L);
  sb[1] = 0.0;
  for (int n = L; n >= 4; --n) sb[1] += field(n) + field(n - 1) + 12.0 * 9 * P.lv[n - 1].nk + 16.0 * 8 * P.lv[n - 1].nk;
  sb[2] = 0.0;
  // parent blocks: C read and B written once, the (code, k) transfer table,
  // the flat source lists and the per-box and per-parent index arrays
  for (int n = 3; n <= L; ++n)
    sb[2] += 2 * field(n) + 16.0 * P.lv[n].nk * 343 + 8.0 * (double)h->lv[n].nps + 4.0 * P.lv[n].nb +
             8.0 * (P.lv[n - 1].nb + 1);
  sb[3] = 0.0;
  for (int n = 4; n <= L; ++n) sb[3] += field(n - 1) + 2 * field(n) + 12.0 * 9 * P.lv[n - 1].nk;
  if (h->recv_lds && h->near_lds) {
    // near field by leaf box: the CSR values and row pointers, one column
    // list per box and its x entries, the near sums written
    sb[4] = zb * (double)P.nnz + 4.0 * (P.ne + 1) + 20.0 * (double)h->near_box_cols + 16.0 * P.ne + 4.0 * P.ne;
    sb[5] = phases + field(L) + 48.0 * P.lv[L].nk + 4.0 * (P.lv[L].nb + 1) + 2.0 * 16.0 * P.ne;
  } else if (h->recv_lds) {
    // near field: the CSR, x gathered once, the near sums written
    sb[4] = csr + 16.0 * P.ne + 16.0 * P.ne + 4.0 * P.ne;
    // receive: phase rows, leaf campoA, sample directions, near sums read, y written
    sb[5] = phases + field(L) + 48.0 * P.lv[L].nk + 4.0 * (P.lv[L].nb + 1) + 2.0 * 16.0 * P.ne;
  } else {
    // one kernel: patterns, leaf campoA, the CSR, x gathered once, y written
    sb[4] = 0.0;
    sb[5] = (double)P.lv[L].nk * 2 * P.ne * 16.0 + field(L) + csr + 2.0 * 16.0 * P.ne + 8.0 * P.ne;
  }
  return LW_OK;
}

}  // extern "C"
