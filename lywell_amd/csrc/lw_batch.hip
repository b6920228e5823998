// Multi-RHS form of the hot path: R right-hand sides share one pass over the
// operator.
//
//   far field (aggregation, interpolation, translation, anterpolation,
//   receive) in FP32 complex with FP32 accumulation, or in FP64 (the same
//   kernels instantiated for double);
//   near field (complex(4)-stored values, FP64 accumulation), ILU0 solves and
//   every CGS vector operation and reduction in FP64.
//
// Layouts (r = right-hand side, fastest):
//   batch vectors      V[i*R + r]                  (complex(8))
//   campo arrays       F[((k*nb + b)*2 + p)*R + r] (far-field precision,
//                      sample-major so one sample's slice of a level is
//                      contiguous for the translation kernel)
//   patterns           rad[(s*2 + p)*nk + k]       (leaf slot order)
//   half patterns      rad_h[(s*2 + p)*nh + m]     (the samples m < nh =
//                      nk/2; sample nh + sample_flip(m) has the pattern
//                      conj(rad) (p = theta) or -conj(rad) (p = phi):
//                      opposite khat, same theta-hat, opposite phi-hat)
// The per-RHS CGS follows CGSold (modIterativo.f90:331-493) column by column:
// a column that meets its tolerance (or breaks down) is frozen, so every
// column's iterate is the one a single-RHS run would produce with this
// operator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lw_internal.h"

namespace lw {
namespace {

constexpr double kMu = Physics::mu, kEps = Physics::epsi, kPi = Physics::pi;
constexpr int kWave = 64;

__device__ __forceinline__ double2 dmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 dadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 dsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
typedef float f32x4v __attribute__((ext_vector_type(4)));
typedef double f64x4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 ddiv(double2 a, double2 b) {
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, d = b.x + b.y * r;
    return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  }
  const double r = b.x / b.y, d = b.x * r + b.y;
  return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

// ---------------------------------------------------------------- far field
//
// Every far-field kernel is a template on the scalar S of the campo arrays,
// patterns and tables: float (complex(4) far field, FP32 accumulation) or
// double (the reference's complex(8)).  Vectors, near field and CGS are FP64
// in both.

template <typename S> struct Cx;
template <> struct Cx<float> {
  typedef float2 T;
  typedef f32x4v Acc;
};
template <> struct Cx<double> {
  typedef double2 T;
  typedef f64x4v Acc;
};
template <typename S>
__device__ __forceinline__ typename Cx<S>::T mk(S x, S y) {
  typename Cx<S>::T v;
  v.x = x;
  v.y = y;
  return v;
}

// 16x16x4 MFMA: lane l supplies A[l%16][l/16] and B[l/16][l%16]; result
// register v of lane l holds D[row][l%16] with row = 4*(l/16) + v for the f32
// form and row = l/16 + 4*v for the f64 form (both measured on gfx950).
__device__ __forceinline__ f32x4v mma(float a, float b, f32x4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4v mma(double a, double b, f64x4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <typename S>
__device__ __forceinline__ int mma_row(int lk, int v) {
  return sizeof(S) == sizeof(float) ? 4 * lk + v : lk + 4 * v;
}

// out[(s*2 + p)*nh + k] = rad[(s*2 + p)*nk + k] for k < nh = nk/2, in the
// far-field precision
template <typename S>
__global__ void k_rad_half(long long n, int nk, const double2* __restrict__ rad, typename Cx<S>::T* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int nh = nk / 2;
  const double2 v = rad[(i / nh) * nk + i % nh];
  out[i] = mk<S>((S)v.x, (S)v.y);
}

template <typename S>
__global__ void k_to_prec(long long n, const double2* __restrict__ in, typename Cx<S>::T* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = mk<S>((S)in[i].x, (S)in[i].y);
}

// Leaf aggregation on MFMA: per leaf box the complex GEMM
//   C_L[m][b][r] = sum_s rad(s; m) x(basis(s), r),  m = p*nk + k,
// over the box's nf bases.  One workgroup per box; its 4 waves take the
// 16-item tiles in rounds, each wave owning 16 items and all R columns.  The
// box's x rows go through LDS (row stride R + 8 against bank conflicts),
// once per box when they fit (nf <= KC), else in chunks of KC bases per tile
// round.
template <typename S, int R, int KC>
__global__ void __launch_bounds__(256) k_b_aggregate(int nk, int nb, const int* __restrict__ box_slot,
                                                     const int* __restrict__ slot_basis,
                                                     const typename Cx<S>::T* __restrict__ rad,
                                                     const double2* __restrict__ X, typename Cx<S>::T* __restrict__ C) {
  typedef typename Cx<S>::T T2;
  typedef typename Cx<S>::Acc Acc;
  constexpr int RS = R + 8, NT = R / 16;
  __shared__ T2 xs[KC * RS];
  const int b = blockIdx.x;
  const int s0 = box_slot[b], nf = box_slot[b + 1] - s0;
  const int nk2 = 2 * nk, ntile = (nk2 + 15) / 16;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), l = threadIdx.x % kWave;
  const int li = l % 16, lk = l / 16;
  const bool one_chunk = nf <= KC;
  auto stage = [&](int c0) {
    const int nc = min(KC, nf - c0), nc4 = (nc + 3) & ~3;
    for (int e = threadIdx.x; e < nc4 * R; e += blockDim.x) {
      const int sl = e / R, r = e % R;
      T2 v = mk<S>(0, 0);
      if (sl < nc) {
        const double2 d = X[(size_t)slot_basis[s0 + c0 + sl] * R + r];
        v = mk<S>((S)d.x, (S)d.y);
      }
      xs[sl * RS + r] = v;
    }
  };
  if (one_chunk) {
    stage(0);
    __syncthreads();
  }
  for (int round = 0; round < (ntile + 3) / 4; ++round) {
    const int tile = round * 4 + wave;
    const int m = tile * 16 + li;
    Acc cr[NT], ci[NT];
#pragma unroll
    for (int jt = 0; jt < NT; ++jt) {
      cr[jt] = Acc{0, 0, 0, 0};
      ci[jt] = Acc{0, 0, 0, 0};
    }
    for (int c0 = 0; c0 < nf; c0 += KC) {
      if (!one_chunk) {
        __syncthreads();
        stage(c0);
        __syncthreads();
      }
      const int nc = min(KC, nf - c0), nc4 = (nc + 3) & ~3;
      for (int ks = 0; ks < nc4; ks += 4) {
        const int sl = ks + lk;
        // clamped, unconditional pattern load; out-of-range entries weigh 0
        const T2 av = rad[(size_t)(s0 + c0 + min(sl, nc - 1)) * nk2 + min(m, nk2 - 1)];
        const T2 a = (sl < nc && m < nk2) ? av : mk<S>(0, 0);
        const T2* xr = xs + sl * RS + li;
        T2 x[NT];
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) x[jt] = xr[jt * 16];
        // the 2*NT accumulators in turn, so that no MFMA waits on the one
        // before it (each accumulator sees its terms in the same order)
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) cr[jt] = mma(a.x, x[jt].x, cr[jt]);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) ci[jt] = mma(a.x, x[jt].y, ci[jt]);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) cr[jt] = mma(-a.y, x[jt].y, cr[jt]);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) ci[jt] = mma(a.y, x[jt].x, ci[jt]);
      }
    }
    if (tile < ntile) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int mo = tile * 16 + mma_row<S>(lk, v);
        if (mo < nk2) {
          const int p = mo / nk, k = mo % nk;
          T2* o = C + (((size_t)k * nb + b) * 2 + p) * R + li;
#pragma unroll
          for (int jt = 0; jt < NT; ++jt) o[jt * 16] = mk<S>(cr[jt][v], ci[jt][v]);
        }
      }
    }
  }
}

// Upward pass with the factored stencil (W = Wphi (x) Wtheta), one wave per
// (parent P, coarse phi column ip, block of TB coarse theta rows, 64-column
// chunk):
//   C_{n-1}[(ip, it)][P] = sum_c ph(oct_c, (ip, it))
//                          sum_a wphi[ip][a] sum_b wth[it][b] C_n[(fa, tb)][c]
// The wave walks its theta rows in order; the 3x3 window of fine samples
// slides with the theta stencil, so most steps load 3 new values (one per
// fine phi column) instead of 9, and the values of the next step are fetched
// before the current one is computed.  Accumulators stay in registers.
// Theta blocks the autotuner chooses from (kThBlock is the default).
constexpr int kThBlock = 8;

template <typename S, int TB>
__global__ void __launch_bounds__(256) k_b_interp_up(int ncol, int nthc, int nkc, int nbc, int nthf, int nbf,
                                                     int nphc, const int* __restrict__ child_ptr,
                                                     const int* __restrict__ oct, const int* __restrict__ sth_idx,
                                                     const S* __restrict__ sth_w, const int* __restrict__ sph_idx,
                                                     const S* __restrict__ sph_w,
                                                     const typename Cx<S>::T* __restrict__ ph,
                                                     const typename Cx<S>::T* __restrict__ Cf,
                                                     typename Cx<S>::T* __restrict__ Cc) {
  typedef typename Cx<S>::T T2;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
  const int nchunk = ncol / kWave, ntb = (nthc + TB - 1) / TB;
  int t = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x / kWave) + wave;
  const int chunk = t % nchunk;
  t /= nchunk;
  const int tbk = t % ntb;
  t /= ntb;
  const int ip = t % nphc, P = t / nphc;
  if (P >= nbc) return;
  const int col = chunk * kWave + lane;
  const int it0 = tbk * TB, nit = min(TB, nthc - it0);
  T2 acc[TB];
#pragma unroll
  for (int u = 0; u < TB; ++u) acc[u] = mk<S>(0, 0);
  const int f0 = sph_idx[3 * ip], f1 = sph_idx[3 * ip + 1], f2 = sph_idx[3 * ip + 2];
  const S w0 = sph_w[3 * ip], w1 = sph_w[3 * ip + 1], w2 = sph_w[3 * ip + 2];
  const size_t ks = (size_t)nbf * ncol;  // stride of one fine sample
  for (int c = child_ptr[P]; c < child_ptr[P + 1]; ++c) {
    const T2* ph_c = ph + (size_t)oct[c] * nkc + (size_t)ip * nthc + it0;
    const T2* s0 = Cf + (size_t)f0 * nthf * ks + (size_t)c * ncol + col;
    const T2* s1 = Cf + (size_t)f1 * nthf * ks + (size_t)c * ncol + col;
    const T2* s2 = Cf + (size_t)f2 * nthf * ks + (size_t)c * ncol + col;
    int cur[3];
    T2 v[3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const int tb = sth_idx[3 * it0 + b];
      cur[b] = tb;
      v[0][b] = s0[(size_t)tb * ks];
      v[1][b] = s1[(size_t)tb * ks];
      v[2][b] = s2[(size_t)tb * ks];
    }
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      if (u < nit) {
        const int it = it0 + u;
        int tn[3] = {cur[0], cur[1], cur[2]};
        T2 nv[3][3];
        if (u + 1 < nit) {
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            const int tb = sth_idx[3 * (it + 1) + b];
            tn[b] = tb;
            if (tb == cur[0]) {
              nv[0][b] = v[0][0]; nv[1][b] = v[1][0]; nv[2][b] = v[2][0];
            } else if (tb == cur[1]) {
              nv[0][b] = v[0][1]; nv[1][b] = v[1][1]; nv[2][b] = v[2][1];
            } else if (tb == cur[2]) {
              nv[0][b] = v[0][2]; nv[1][b] = v[1][2]; nv[2][b] = v[2][2];
            } else {
              nv[0][b] = s0[(size_t)tb * ks];
              nv[1][b] = s1[(size_t)tb * ks];
              nv[2][b] = s2[(size_t)tb * ks];
            }
          }
        }
        const S u0 = sth_w[3 * it], u1 = sth_w[3 * it + 1], u2 = sth_w[3 * it + 2];
        S vx = 0, vy = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const S wa = a == 0 ? w0 : (a == 1 ? w1 : w2);
          vx += wa * (u0 * v[a][0].x + u1 * v[a][1].x + u2 * v[a][2].x);
          vy += wa * (u0 * v[a][0].y + u1 * v[a][1].y + u2 * v[a][2].y);
        }
        const T2 pz = ph_c[u];
        acc[u].x += pz.x * vx - pz.y * vy;
        acc[u].y += pz.x * vy + pz.y * vx;
        if (u + 1 < nit) {
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) v[a][b] = nv[a][b];
          cur[0] = tn[0];
          cur[1] = tn[1];
          cur[2] = tn[2];
        }
      }
    }
  }
  T2* out = Cc + ((size_t)(ip * nthc + it0) * nbc + P) * ncol + col;
#pragma unroll
  for (int u = 0; u < TB; ++u)
    if (u < nit) out[(size_t)u * nbc * ncol] = acc[u];
}

// Translation fused with the downward pass, one wave per (sample k, parent
// box P, 64*CPL-column chunk: CPL columns per lane) at level n:
//   A_n[k][c] = sum_{far d} pes[k] T(off(c,d))[k] C_n[k][d]
//             + sum_{(i,w) in W^T(k)} w conj(ph_{n-1}(oct_c, i)) A_{n-1}[i][P]
// for the children c of P.  The far boxes of every child of P are the
// children of P's neighbour parents Q that do not touch it, so a wave loads
// each source row C_n[k][d] once and applies it to all children of P; the
// transfer function of a pair depends only on the integer offset
// off = 2(Q - P) + oct_d - oct_c in box units, so the k-major table Tk holds
// one entry per offset code (pes folded in).  Parent-level data come in
// through the combined weights wp[oct][e] = w_e conj(ph(oct, row_e)).
//
// Per parent the sources are one flat list of (box d, code part
// 49 (2ox+3+bx_d) + 7 (2oy+3+by_d) + (2oz+3+bz_d) - 57), padded to a multiple
// of kSrcU = 4 with entries whose codes are near offsets (T = 0).  Everything that
// selects a transfer value (k, P, the source entry, the child's octant) is
// uniform over the wave, so the values come in through scalar loads from the
// Tk row and enter the FMAs as scalar operands; the children of P are
// compacted (nch = 1..8 accumulators, one code path per count), so no FMA is
// spent on absent octants and the inner loop has no branches.  Source rows
// are loaded one group of SU (2 or 4) ahead.  Every variant adds each child's
// terms in the same order, so all give the same bits.  The parents are cut into 1, 2, 4 or 8
// contiguous ranges per level, each range on the XCDs (workgroup id mod 8)
// that take its samples (see k_b_far).
constexpr int kOffCodes = 343;  // (dx, dy, dz) in [-3, 3]^3
constexpr int kSrcU = 4;        // source-list padding: the lists are a multiple of this
constexpr int kFarCpl = 2;      // columns per lane of the default far-kernel variant
// 49*2 + 7*2 + 2: plus any child's octant part 57 - oct_code(o), an offset in
// [-1, 1]^3, which has no transfer vector (T = 0)
constexpr int kPadCode = 114;

// Offset-code part of a child octant o (bits x,y,z = 4,2,1): 49 bx + 7 by + bz.
__host__ __device__ constexpr int oct_code(int o) { return (o >> 2) * 49 + ((o >> 1) & 1) * 7 + (o & 1); }

__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
// acc += t v as four FMAs
template <typename T2>
__device__ __forceinline__ void cmac(T2& acc, const T2 t, const T2 v) {
  acc.x = fmadd(t.x, v.x, acc.x);
  acc.x = fmadd(-t.y, v.y, acc.x);
  acc.y = fmadd(t.x, v.y, acc.y);
  acc.y = fmadd(t.y, v.x, acc.y);
}

#define LW_FAR_PARAMS                                                                                 \
  int nb, int ncol, int nbc, int nwt, const int* __restrict__ pc_ptr, const int* __restrict__ bcd,     \
      const int* __restrict__ boct, const int* __restrict__ ps_ptr, const int2* __restrict__ ps_ent,   \
      const typename Cx<S>::T* __restrict__ Tk, const typename Cx<S>::T* __restrict__ C,                \
      const int* __restrict__ wt_ptr, const int* __restrict__ wt_row,                                  \
      const typename Cx<S>::T* __restrict__ wp, const typename Cx<S>::T* __restrict__ Ac,              \
      typename Cx<S>::T* __restrict__ A
#define LW_FAR_ARGS nb, ncol, nbc, nwt, pc_ptr, bcd, boct, ps_ptr, ps_ent, Tk, C, wt_ptr, wt_row, wp, Ac, A

template <typename S, int NCH, int SU, int CPL>
__device__ __forceinline__ void far_parent(int k, int P, int col, const typename Cx<S>::T* Tkk, LW_FAR_PARAMS) {
  typedef typename Cx<S>::T T2;
  constexpr int kMaxCq = oct_code(7);  // 57
  const int c0 = pc_ptr[P];
  int coff[NCH], wo[NCH];
#pragma unroll
  for (int i = 0; i < NCH; ++i) {
    coff[i] = kMaxCq - bcd[c0 + i];
    wo[i] = boct[c0 + i] * nwt;
  }
  T2 acc[NCH][CPL];
#pragma unroll
  for (int i = 0; i < NCH; ++i)
#pragma unroll
    for (int c = 0; c < CPL; ++c) acc[i][c] = mk<S>(0, 0);
  // offsets within a level's sample slice fit 32 bits (checked at set-up)
  const T2* Ck = C + (size_t)k * nb * ncol + col;
  const int s0 = ps_ptr[P], s1 = ps_ptr[P + 1];
  if (s0 < s1) {
    int2 ea[SU], eb[SU];
    T2 va[SU][CPL], vb[SU][CPL];
    auto load = [&](int s, int2(&e)[SU], T2(&v)[SU][CPL]) {
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        e[u] = ps_ent[s + u];
#pragma unroll
        for (int c = 0; c < CPL; ++c) v[u][c] = Ck[(unsigned)(e[u].x * ncol + c * kWave)];
      }
    };
    auto apply = [&](const int2(&e)[SU], const T2(&v)[SU][CPL]) {
#pragma unroll
      for (int u = 0; u < SU; ++u) {
        const T2* tb = Tkk + e[u].y;
#pragma unroll
        for (int i = 0; i < NCH; ++i) {
          const T2 t = tb[coff[i]];
#pragma unroll
          for (int c = 0; c < CPL; ++c) cmac(acc[i][c], t, v[u][c]);
        }
      }
    };
    // two buffers: the next group (or the last group again at the end, so
    // that the loads are always in bounds and need no branch) is loaded
    // before the current group's FMAs
    // (the scheduling barriers keep the compiler from sinking the loads
    // back down to their uses)
    load(s0, ea, va);
    for (int s = s0;;) {
      load(min(s + SU, s1 - SU), eb, vb);
      __builtin_amdgcn_sched_barrier(0);
      apply(ea, va);
      s += SU;
      if (s >= s1) break;
      load(min(s + SU, s1 - SU), ea, va);
      __builtin_amdgcn_sched_barrier(0);
      apply(eb, vb);
      s += SU;
      if (s >= s1) break;
    }
  }
  const int e0 = Ac ? wt_ptr[k] : 0, e1 = Ac ? wt_ptr[k + 1] : 0;
  if (e0 < e1) {
    // entries come in groups of 4 (padded with zero weights), loaded one
    // group ahead like the sources
    T2 aa[4][CPL], ab[4][CPL];
    auto load = [&](int e, T2(&a)[4][CPL]) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const T2* ar = Ac + (unsigned)((wt_row[e + u] * nbc + P) * ncol + col);
#pragma unroll
        for (int c = 0; c < CPL; ++c) a[u][c] = ar[c * kWave];
      }
    };
    auto apply = [&](int e, const T2(&a)[4][CPL]) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int i = 0; i < NCH; ++i) {
          const T2 w = wp[wo[i] + e + u];
#pragma unroll
          for (int c = 0; c < CPL; ++c) cmac(acc[i][c], w, a[u][c]);
        }
    };
    load(e0, aa);
    for (int e = e0;;) {
      load(min(e + 4, e1 - 4), ab);
      __builtin_amdgcn_sched_barrier(0);
      apply(e, aa);
      e += 4;
      if (e >= e1) break;
      load(min(e + 4, e1 - 4), aa);
      __builtin_amdgcn_sched_barrier(0);
      apply(e, ab);
      e += 4;
      if (e >= e1) break;
    }
  }
  T2* Ak = A + (size_t)k * nb * ncol + col;
#pragma unroll
  for (int i = 0; i < NCH; ++i)
#pragma unroll
    for (int c = 0; c < CPL; ++c) Ak[(unsigned)((c0 + i) * ncol + c * kWave)] = acc[i][c];
}

template <typename S, int SU, int CPL, bool TL>
__global__ void __launch_bounds__(256) k_b_far(int nk, int nP, int nchunk, int wpx, int nsplit, int order,
                                               LW_FAR_PARAMS) {
  typedef typename Cx<S>::T T2;
  // Workgroup map (order 0 and 1): within a workgroup one sample k and one
  // column chunk, 4 waves = 4 consecutive parents.  The parents are cut into
  // nsplit contiguous ranges (nsplit divides 8) and XCD x owns range
  // x % nsplit for the samples k = x / nsplit mod 8 / nsplit, so that its L2
  // holds the sources of a box range (1/nsplit of a k-slice plus a halo) for
  // the few samples it has in flight.  nsplit = 1 gives each XCD whole
  // k-slices, nsplit = 8 a range of every k-slice.  The column chunks of a
  // sample follow each other (order 0), or each chunk runs over all samples
  // before the next (order 1), which halves the rows in flight.
  // Order 2: 4 waves = 4 consecutive samples of one parent, whose downward
  // stencils share most of their parent rows A_{n-1}[i][P] (read once per
  // workgroup through L1 instead of once per wave); the parents of a range
  // follow each other for a block of samples.
  // TL (orders 0 and 1): the workgroup's transfer row Tk[k][.] is staged in
  // LDS and read from there (broadcast reads into vector registers) instead
  // of through scalar loads from the table.
  if constexpr (TL) {
    // not a TL shape (each wave of order 2 has its own sample); never
    // launched, and returning here keeps the transfer-row pointer below an
    // LDS pointer on every path
    if (order == 2) return;
  }
  const int g = blockIdx.x, x = g % 8, j = g / 8;
  const int kstep = 8 / nsplit, pr = x % nsplit;
  const int nPx = (nP + nsplit - 1) / nsplit, nPBx = wpx / nchunk;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  int k, chunk, P, P1 = min((pr + 1) * nPx, nP);
  if (order == 2) {
    const int per = nPx * nchunk;
    const int kb = (j / per) * kstep + x / nsplit;
    chunk = (j % per) / nPx;
    P = pr * nPx + (j % per) % nPx;
    k = kb * 4 + wave;
    if (P >= P1 || k >= nk) return;
  } else {
    int pb;
    if (order == 1) {
      const int per = ((nk + kstep - 1) / kstep) * nPBx;
      chunk = j / per;
      k = ((j % per) / nPBx) * kstep + x / nsplit;
      pb = (j % per) % nPBx;
    } else {
      k = (j / wpx) * kstep + x / nsplit;
      chunk = (j % wpx) / nPBx;
      pb = (j % wpx) % nPBx;
    }
    if (k >= nk) return;  // uniform over the workgroup
    P = pr * nPx + pb * 4 + wave;
  }
  const T2* Tkk = Tk + (size_t)k * kOffCodes;
  if constexpr (TL) {
    __shared__ T2 tsh[kOffCodes];
    for (int i = threadIdx.x; i < kOffCodes; i += blockDim.x) tsh[i] = Tkk[i];
    __syncthreads();
    Tkk = tsh;
  }
  if (P >= P1) return;
  const int col = chunk * kWave * CPL + (threadIdx.x % kWave);
  switch (pc_ptr[P + 1] - pc_ptr[P]) {
    case 1: far_parent<S, 1, SU, CPL>(k, P, col, Tkk, LW_FAR_ARGS); break;
    case 2: far_parent<S, 2, SU, CPL>(k, P, col, Tkk, LW_FAR_ARGS); break;
    case 3: far_parent<S, 3, SU, CPL>(k, P, col, Tkk, LW_FAR_ARGS); break;
    case 4: far_parent<S, 4, SU, CPL>(k, P, col, Tkk, LW_FAR_ARGS); break;
    case 5: far_parent<S, 5, SU, CPL>(k, P, col, Tkk, LW_FAR_ARGS); break;
    case 6: far_parent<S, 6, SU, CPL>(k, P, col, Tkk, LW_FAR_ARGS); break;
    case 7: far_parent<S, 7, SU, CPL>(k, P, col, Tkk, LW_FAR_ARGS); break;
    case 8: far_parent<S, 8, SU, CPL>(k, P, col, Tkk, LW_FAR_ARGS); break;
    default: break;
  }
}

// The downward pass of the batched form as its own kernel (the second half
// of far_parent, run after the translation has stored A_n = sum over far
// boxes): A_n[k][c] += sum_{(i,w) in W^T(k)} w conj(ph_{n-1}(oct_c, i))
// A_{n-1}[i][P] for the children c of P, each term added in far_parent's
// order, so the result has the same bits as the fused kernel.  A workgroup
// takes one parent, one 64-column chunk and kAntK consecutive samples (four
// per wave), and consecutive workgroups on one XCD take the same parent, so
// a parent row A_{n-1}[i][P], which the stencils of several neighbouring
// samples read, comes from the XCD's L2 instead of memory.
constexpr int kAntK = 16;

template <typename S, int NCH>
__device__ __forceinline__ void anterp_parent(int k, int P, int col, int nb, int ncol, int nbc, int nwt,
                                              const int* __restrict__ pc_ptr, const int* __restrict__ boct,
                                              const int* __restrict__ wt_ptr, const int* __restrict__ wt_row,
                                              const typename Cx<S>::T* __restrict__ wp,
                                              const typename Cx<S>::T* __restrict__ Ac, typename Cx<S>::T* A) {
  typedef typename Cx<S>::T T2;
  const int c0 = pc_ptr[P];
  int wo[NCH];
#pragma unroll
  for (int i = 0; i < NCH; ++i) wo[i] = boct[c0 + i] * nwt;
  T2* Ak = A + (size_t)k * nb * ncol + col;
  T2 acc[NCH];
#pragma unroll
  for (int i = 0; i < NCH; ++i) acc[i] = Ak[(unsigned)((c0 + i) * ncol)];
  const int e0 = wt_ptr[k], e1 = wt_ptr[k + 1];
  if (e0 < e1) {
    T2 aa[4], ab[4];
    auto load = [&](int e, T2(&a)[4]) {
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = Ac[(unsigned)((wt_row[e + u] * nbc + P) * ncol + col)];
    };
    auto apply = [&](int e, const T2(&a)[4]) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int i = 0; i < NCH; ++i) cmac(acc[i], wp[wo[i] + e + u], a[u]);
    };
    load(e0, aa);
    for (int e = e0;;) {
      load(min(e + 4, e1 - 4), ab);
      __builtin_amdgcn_sched_barrier(0);
      apply(e, aa);
      e += 4;
      if (e >= e1) break;
      load(min(e + 4, e1 - 4), aa);
      __builtin_amdgcn_sched_barrier(0);
      apply(e, ab);
      e += 4;
      if (e >= e1) break;
    }
  }
#pragma unroll
  for (int i = 0; i < NCH; ++i) Ak[(unsigned)((c0 + i) * ncol)] = acc[i];
}

template <typename S>
__global__ void __launch_bounds__(256) k_b_anterp(int nk, int nP, int nchunk, int nb, int ncol, int nbc, int nwt,
                                                  const int* __restrict__ pc_ptr, const int* __restrict__ boct,
                                                  const int* __restrict__ wt_ptr, const int* __restrict__ wt_row,
                                                  const typename Cx<S>::T* __restrict__ wp,
                                                  const typename Cx<S>::T* __restrict__ Ac,
                                                  typename Cx<S>::T* A) {
  const int nkb = (nk + kAntK - 1) / kAntK;
  int t = xcd_block(blockIdx.x, gridDim.x);
  const int kb = t % nkb;
  t /= nkb;
  const int chunk = t % nchunk, P = t / nchunk;
  if (P >= nP) return;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int col = chunk * kWave + (threadIdx.x % kWave);
  const int nch = pc_ptr[P + 1] - pc_ptr[P];
  for (int m = 0; m < kAntK / 4; ++m) {
    const int k = kb * kAntK + wave * (kAntK / 4) + m;
    if (k >= nk) break;
    switch (nch) {
#define LW_ANT(c)                                                                                                \
  case c:                                                                                                        \
    anterp_parent<S, c>(k, P, col, nb, ncol, nbc, nwt, pc_ptr, boct, wt_ptr, wt_row, wp, Ac, A);                 \
    break;
      LW_ANT(1) LW_ANT(2) LW_ANT(3) LW_ANT(4) LW_ANT(5) LW_ANT(6) LW_ANT(7) LW_ANT(8)
#undef LW_ANT
      default: break;
    }
  }
}

// Single right-hand side translation by parent blocks (the sum of
// modMLFMA.f90:1837-1853 regrouped as in k_b_far): one lane per sample k; a
// workgroup takes 64 consecutive samples of one parent P, its wave w the
// children w and w+4 of P.  P's flat source list is walked in groups of
// kSrcU = 4 rows: wave u fetches row u of the next group (both
// polarisations) into registers while the current group is applied from LDS,
// then stores it to the other LDS buffer before the group barrier, so every
// source row is read from memory once per workgroup.  Each row is applied to
// the wave's children with the pair's transfer value Tc[code][k] (zero for
// near offsets and pads), the group's values loaded together.  Each child's far boxes
// come in ascending box order, the order of the plan's per-box far lists,
// and each term is added as the reference adds it (campoB += campoC * T, then
// campoB * pesos), so the sums are those of the per-box form: the near and
// pad entries add exact zeros.  Lanes past nk load sample nk-1 and store
// nothing.
constexpr int kTrWaves = 4;
static_assert(kTrWaves == kSrcU, "one source row per wave and group");

template <int NW>
__device__ __forceinline__ void translate_group(const double2 (*sv)[2][kWave], int lane, int nk, int kl,
                                                const int2 (&e)[kSrcU], const int (&coff)[2],
                                                const double2* __restrict__ Tc, double2 (&a0)[2],
                                                double2 (&a1)[2]) {
  // the group's transfer values (every source, every child) are loaded
  // together, then applied source by source
  double2 t[kSrcU][NW];
#pragma unroll
  for (int u = 0; u < kSrcU; ++u)
#pragma unroll
    for (int i = 0; i < NW; ++i) t[u][i] = Tc[(size_t)(e[u].y + coff[i]) * nk + kl];
#pragma unroll
  for (int u = 0; u < kSrcU; ++u) {
    const double2 v0 = sv[u][0][lane], v1 = sv[u][1][lane];
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      a0[i] = dadd(a0[i], dmul(v0, t[u][i]));
      a1[i] = dadd(a1[i], dmul(v1, t[u][i]));
    }
  }
}

__global__ void __launch_bounds__(256) k_translate_parent(int nk, int nchk, const int* __restrict__ pc_ptr,
                                                          const int* __restrict__ bcd,
                                                          const int* __restrict__ ps_ptr,
                                                          const int2* __restrict__ ps_ent,
                                                          const double2* __restrict__ Tc,
                                                          const double* __restrict__ pes,
                                                          const double2* __restrict__ C, double2* __restrict__ B) {
  // source rows of two groups: [buffer][row][polarisation][lane]
  __shared__ double2 sv[2][kSrcU][2][kWave];
  constexpr int kMaxCq = oct_code(7);
  // grid: (parent, chunk of 64 samples), XCD-contiguous
  const int wb = xcd_block(blockIdx.x, gridDim.x);
  const int P = wb / nchk;
  const int lane = threadIdx.x % kWave;
  const int k = (wb % nchk) * kWave + lane, kl = min(k, nk - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  // wave w takes the children w and w + 4 of P, so that the four waves share
  // the children of a parent with four or fewer (most of them at the leaf
  // level)
  const int nc = pc_ptr[P + 1] - pc_ptr[P];
  const int c0 = pc_ptr[P] + w, nw = (w < nc ? 1 : 0) + (w + kTrWaves < nc ? 1 : 0);
  const int s0 = ps_ptr[P], s1 = ps_ptr[P + 1];
  int coff[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) coff[i] = kMaxCq - bcd[min(c0 + kTrWaves * i, pc_ptr[P + 1] - 1)];
  double2 a0[2], a1[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    a0[i] = make_double2(0.0, 0.0);
    a1[i] = make_double2(0.0, 0.0);
  }
  if (s0 < s1) {
    // row w of the first group
    {
      const double2* cr = C + (size_t)ps_ent[s0 + w].x * 2 * nk + kl;
      sv[0][w][0][lane] = cr[0];
      sv[0][w][1][lane] = cr[nk];
    }
    __syncthreads();
    int cur = 0;
    // the group's entries (source box, code part), loaded a group ahead
    int2 e[kSrcU];
#pragma unroll
    for (int u = 0; u < kSrcU; ++u) e[u] = ps_ent[s0 + u];
    for (int g = s0; g < s1; g += kSrcU) {
      // the next group (the last group again at the end, so that the loads
      // stay in bounds): its entries and its row w
      const int gn = min(g + kSrcU, s1 - kSrcU);
      int2 en[kSrcU];
#pragma unroll
      for (int u = 0; u < kSrcU; ++u) en[u] = ps_ent[gn + u];
      const double2* cr = C + (size_t)ps_ent[gn + w].x * 2 * nk + kl;
      const double2 n0 = cr[0], n1 = cr[nk];
      if (nw == 2)
        translate_group<2>(sv[cur], lane, nk, kl, e, coff, Tc, a0, a1);
      else if (nw == 1)
        translate_group<1>(sv[cur], lane, nk, kl, e, coff, Tc, a0, a1);
      sv[cur ^ 1][w][0][lane] = n0;
      sv[cur ^ 1][w][1][lane] = n1;
      __syncthreads();
      cur ^= 1;
#pragma unroll
      for (int u = 0; u < kSrcU; ++u) e[u] = en[u];
    }
  }
  if (k < nk) {
    const double p = pes[k];
#pragma unroll
    for (int i = 0; i < 2; ++i)
      if (i < nw) {
        double2* o = B + (size_t)(c0 + kTrWaves * i) * 2 * nk + k;
        o[0] = make_double2(a0[i].x * p, a0[i].y * p);
        o[nk] = make_double2(a1[i].x * p, a1[i].y * p);
      }
  }
}

// Tc[code][k] = T[tr(code)][k] (zero for codes without a transfer vector):
// the code-major table of the single-RHS translation
__global__ void k_tc(int nk, const int* __restrict__ code_tr, const double2* __restrict__ T,
                     double2* __restrict__ Tc) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nk * kOffCodes) return;
  const int code = (int)(t / nk), k = (int)(t % nk);
  const int tr = code_tr[code];
  if (tr < 0) {
    Tc[t] = make_double2(0.0, 0.0);
    return;
  }
  Tc[t] = T[(size_t)tr * nk + k];
}

// Tk[k][code] = pes[k] T[tr(code)][k] (zero for codes without a transfer vector)
template <typename S>
__global__ void k_b_tk(int nk, const int* __restrict__ code_tr, const double2* __restrict__ T,
                       const double* __restrict__ pes, typename Cx<S>::T* __restrict__ Tk) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nk * kOffCodes) return;
  const int k = (int)(t / kOffCodes), code = (int)(t % kOffCodes);
  const int tr = code_tr[code];
  if (tr < 0) {
    Tk[t] = mk<S>(0, 0);
    return;
  }
  const double2 v = T[(size_t)tr * nk + k];
  Tk[t] = mk<S>((S)(v.x * pes[k]), (S)(v.y * pes[k]));
}

// wp[o][e] = w_e conj(ph(o, row_e)): anterpolation weight of transposed entry e
// for a child in octant o
template <typename S>
__global__ void k_b_wp(int nkc, int nwt, const int* __restrict__ wt_row, const double* __restrict__ wt_w,
                       const double2* __restrict__ ph, typename Cx<S>::T* __restrict__ wp) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8LL * nwt) return;
  const int o = (int)(t / nwt), e = (int)(t % nwt);
  const double2 p = ph[(size_t)o * nkc + wt_row[e]];
  const double w = wt_w[e];
  wp[t] = mk<S>((S)(w * p.x), (S)(-w * p.y));
}

// Receive on MFMA, one workgroup per leaf box:
//   Yf[basis(s)][r] = sum_m conj(rad(s; m)) A_L[m][b][r],  m = p*nk + k,
// a complex GEMM with the items as the contraction.  The half pattern rows
// hold the items (p, k < nh); with a = rad(s; p, k), the antipode k' =
// nh + sample_flip(k) and sg = +1 (p = theta) or -1 (p = phi), the item and
// its antipode together add
//   conj(a) A(k) + sg a A(k') = a.x U1 + a.y U2 + j (a.x U3 - a.y U4),
//   U1 = A.x + sg A'.x, U2 = A.y - sg A'.y, U3 = A.y + sg A'.y, U4 = A.x - sg A'.x,
// so the contraction runs over nk items instead of 2 nk.  A wave owns 16 of
// the box's bases and 64 columns.  Each lane loads QB consecutive items of
// one pattern row (coalesced pieces), so within a block of 4*QB items, MFMA
// step q pairs lane group g with item QB*g + q; the A_L operands are fetched
// with the same assignment.
// With Y given, the epilogue writes Y = (j w mu / 4 pi) Yf + Yn instead of Yf
// (Yn: the near field of k_b_near_box), the sum k_b_near_box forms from a
// stored Yf, operand for operand.
template <typename S, int R>
__global__ void __launch_bounds__(128) k_b_receive(int nk, int nth, int nb, const int* __restrict__ box_slot,
                                                   const int* __restrict__ slot_basis,
                                                   const typename Cx<S>::T* __restrict__ rad,
                                                   const typename Cx<S>::T* __restrict__ AL,
                                                   typename Cx<S>::T* __restrict__ Yf, double2 cc,
                                                   const double2* __restrict__ Yn, double2* __restrict__ Y) {
  typedef typename Cx<S>::T T2;
  typedef typename Cx<S>::Acc Acc;
  constexpr int NT = 4, NCH = R / 64;
  const int b = blockIdx.x;
  const int s0 = box_slot[b], nf = box_slot[b + 1] - s0;
  const int nh = nk / 2;  // items (p, k < nh): nk of them
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), l = threadIdx.x % kWave;
  const int li = l % 16, lk = l / 16;
  const int nst = (nf + 15) / 16;
  for (int task = wave; task < nst * NCH; task += blockDim.x / kWave) {
    const int st = task / NCH, c0 = (task % NCH) * 64;
    const int sl = st * 16 + li;  // basis row this lane loads
    const T2* rrow = rad + (size_t)(s0 + min(sl, nf - 1)) * nk;
    const bool row_ok = sl < nf;
    Acc dr[NT], di[NT];
#pragma unroll
    for (int jt = 0; jt < NT; ++jt) {
      dr[jt] = Acc{0, 0, 0, 0};
      di[jt] = Acc{0, 0, 0, 0};
    }
    // one block of 4*QB items (lane group g takes items QB*g .. QB*g+QB-1):
    // the pattern pieces and the A_L operands, with clamped unconditional
    // loads (out-of-range items weigh 0), fetched a block ahead of the MFMAs
    // that use them
    constexpr int QB = 1;
    T2 aa[QB], ab[QB], xa[QB][NT], xb[QB][NT], ya[QB][NT], yb[QB][NT];
    S ga[QB], gb[QB];
    // item m: its pattern piece, A at the item and A at its antipode, and sg
    auto load = [&](int mb, T2(&a4)[QB], T2(&x)[QB][NT], T2(&y)[QB][NT], S(&sg)[QB]) {
      const int m0 = mb + QB * lk;
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        const int m = min(m0 + q, nk - 1);
        const T2 av = rrow[m];
        a4[q] = (row_ok && m0 + q < nk) ? av : mk<S>(0, 0);
        const int p = m / nh, k = m % nh, ka = nh + sample_flip(k, nth);
        sg[q] = p == 0 ? S(1) : S(-1);
        const T2* ap = AL + (((size_t)k * nb + b) * 2 + p) * R + c0 + li;
        const T2* aq = AL + (((size_t)ka * nb + b) * 2 + p) * R + c0 + li;
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) {
          x[q][jt] = ap[jt * 16];
          y[q][jt] = aq[jt * 16];
        }
      }
    };
    auto apply = [&](const T2(&a4)[QB], const T2(&x)[QB][NT], const T2(&y)[QB][NT], const S(&sg)[QB]) {
      // the 2*NT accumulators in turn, so that no MFMA waits on the one
      // before it (each accumulator sees its terms in the same order)
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        S u1[NT], u2[NT], u3[NT], u4[NT];
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) {
          u1[jt] = x[q][jt].x + sg[q] * y[q][jt].x;
          u2[jt] = x[q][jt].y - sg[q] * y[q][jt].y;
          u3[jt] = x[q][jt].y + sg[q] * y[q][jt].y;
          u4[jt] = x[q][jt].x - sg[q] * y[q][jt].x;
        }
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) dr[jt] = mma(a4[q].x, u1[jt], dr[jt]);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) di[jt] = mma(a4[q].x, u3[jt], di[jt]);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) dr[jt] = mma(a4[q].y, u2[jt], dr[jt]);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) di[jt] = mma(-a4[q].y, u4[jt], di[jt]);
      }
    };
    constexpr int MB = 4 * QB;
    load(0, aa, xa, ya, ga);
    for (int mb = 0;;) {
      load(min(mb + MB, nk - 1), ab, xb, yb, gb);
      __builtin_amdgcn_sched_barrier(0);
      apply(aa, xa, ya, ga);
      mb += MB;
      if (mb >= nk) break;
      load(min(mb + MB, nk - 1), aa, xa, ya, ga);
      __builtin_amdgcn_sched_barrier(0);
      apply(ab, xb, yb, gb);
      mb += MB;
      if (mb >= nk) break;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int so = st * 16 + mma_row<S>(lk, v);
      if (so < nf) {
        const size_t at0 = (size_t)slot_basis[s0 + so] * R + c0 + li;
        if (Y) {
#pragma unroll
          for (int jt = 0; jt < NT; ++jt) {
            const T2 f = mk<S>(dr[jt][v], di[jt][v]);
            const double2 fs = dmul(make_double2((double)f.x, (double)f.y), cc);
            const double2 yn = Yn[at0 + jt * 16];
            Y[at0 + jt * 16] = make_double2(fs.x + yn.x, fs.y + yn.y);
          }
        } else {
#pragma unroll
          for (int jt = 0; jt < NT; ++jt) Yf[at0 + jt * 16] = mk<S>(dr[jt][v], di[jt][v]);
        }
      }
    }
  }
}

// Near field plus the scaled far field (Yf null: the near field alone), one
// workgroup per leaf box.  Every
// row of a box has the same near columns (the bases of the box's near boxes,
// ascending), so the box's block of the near matrix is a dense (nf x ncol)
// matrix applied to ncol rows of X: a complex GEMM on FP64 MFMA.  A wave owns
// 16 of the box's rows and 64 right-hand sides; each lane loads QB
// consecutive matrix entries of its row, so within a block of 4*QB columns
// MFMA step q pairs lane group g with column QB*g + q, and the X rows are
// fetched with the same assignment.  The stored values (complex(4) or
// complex(8)) are widened to FP64; accumulation is FP64.
template <typename V, typename F>
__global__ void __launch_bounds__(128) k_b_near_box(int R, const int* __restrict__ box_slot,
                                                    const int* __restrict__ slot_basis,
                                                    const int* __restrict__ col_ptr, const int* __restrict__ cols,
                                                    const int* __restrict__ ia, const V* __restrict__ z,
                                                    const double2* __restrict__ X, const F* __restrict__ Yf,
                                                    double2 cc, double2* __restrict__ Y) {
  constexpr int NT = 4;
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int s0 = box_slot[b], nf = box_slot[b + 1] - s0;
  const int j0 = col_ptr[b], ncol = col_ptr[b + 1] - j0;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), l = threadIdx.x % kWave;
  const int li = l % 16, lk = l / 16;
  const int nch = R / 64, nst = (nf + 15) / 16;
  const int* cb = cols + j0;
  for (int task = wave; task < nst * nch; task += blockDim.x / kWave) {
    const int st = task / nch, c0 = (task % nch) * 64;
    const int sl = st * 16 + li;  // matrix row this lane loads
    const bool row_ok = sl < nf;
    const V* zrow = z + ia[slot_basis[s0 + min(sl, nf - 1)]];
    f64x4v dr[NT], di[NT];
#pragma unroll
    for (int jt = 0; jt < NT; ++jt) {
      dr[jt] = f64x4v{0, 0, 0, 0};
      di[jt] = f64x4v{0, 0, 0, 0};
    }
    // blocks of 4*QB columns (lane group g takes columns QB*g .. QB*g+QB-1):
    // the matrix entries and the X rows, with clamped unconditional loads
    // (out-of-range entries weigh 0), fetched a block ahead of the MFMAs
    constexpr int QB = 2, CB = 4 * QB;
    double2 aa[QB], ab[QB], xa[QB][NT], xb[QB][NT];
    auto load = [&](int jb, double2(&a4)[QB], double2(&x)[QB][NT]) {
      const int jl = jb + QB * lk;
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        const int j = min(jl + q, ncol - 1);
        const double2 av = to_d2(zrow[j]);
        a4[q] = (row_ok && jl + q < ncol) ? av : make_double2(0.0, 0.0);
        const double2* xp = X + (size_t)cb[j] * R + c0 + li;
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) x[q][jt] = xp[jt * 16];
      }
    };
    auto apply = [&](const double2(&a4)[QB], const double2(&x)[QB][NT]) {
      // the 2*NT accumulators in turn (each sees its terms in the same order)
#pragma unroll
      for (int q = 0; q < QB; ++q) {
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) dr[jt] = mma(a4[q].x, x[q][jt].x, dr[jt]);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) di[jt] = mma(a4[q].x, x[q][jt].y, di[jt]);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) dr[jt] = mma(-a4[q].y, x[q][jt].y, dr[jt]);
#pragma unroll
        for (int jt = 0; jt < NT; ++jt) di[jt] = mma(a4[q].y, x[q][jt].x, di[jt]);
      }
    };
    load(0, aa, xa);
    for (int jb = 0;;) {
      load(min(jb + CB, ncol - 1), ab, xb);
      __builtin_amdgcn_sched_barrier(0);
      apply(aa, xa);
      jb += CB;
      if (jb >= ncol) break;
      load(min(jb + CB, ncol - 1), aa, xa);
      __builtin_amdgcn_sched_barrier(0);
      apply(ab, xb);
      jb += CB;
      if (jb >= ncol) break;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int so = st * 16 + mma_row<double>(lk, v);
      if (so < nf) {
        const size_t at0 = (size_t)slot_basis[s0 + so] * R + c0 + li;
        if (Yf) {
#pragma unroll
          for (int jt = 0; jt < NT; ++jt) {
            const F f = Yf[at0 + jt * 16];
            const double2 fs = dmul(make_double2((double)f.x, (double)f.y), cc);
            Y[at0 + jt * 16] = make_double2(fs.x + dr[jt][v], fs.y + di[jt][v]);
          }
        } else {
          // the near field alone (k_b_receive adds the far field)
#pragma unroll
          for (int jt = 0; jt < NT; ++jt) Y[at0 + jt * 16] = make_double2(dr[jt][v], di[jt][v]);
        }
      }
    }
  }
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
  // the partials of eight blocks are loaded together, then added in block
  // order
  constexpr int kU = 8;
  for (int b0 = 0; b0 < nblk; b0 += kU) {
    double4 v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (b0 + u < nblk) v[u] = *reinterpret_cast<const double4*>(part + ((size_t)(b0 + u) * R + r) * 4);
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (b0 + u < nblk) {
        s1 = dadd(s1, make_double2(v[u].x, v[u].y));
        s2 = dadd(s2, make_double2(v[u].z, v[u].w));
      }
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

// Per-column CGS control on the device, so that iterations run back to back
// (and replay as one captured graph) with the host looking in only every few
// iterations.  st[c] = {active, first, converged, iterations}.
// ||r|| / ||b||; a zero right-hand side counts a zero residual as 0
__device__ __forceinline__ double rel_res(double nr, double nb) {
  return nb > 0.0 ? nr / nb : (nr == 0.0 ? 0.0 : HUGE_VAL);
}

// Column state after the initial residual: nrm = |b|^2, rhon = r.R~, nr0 =
// |r|^2.  A column whose rho is zero stops before its first update, as
// CGSold does; it has converged if its initial residual passes the test.
__global__ void k_b_cgs_init(int R, double errmax, const double2* __restrict__ nrm,
                             const double2* __restrict__ rhon, const double2* __restrict__ nr0,
                             double* __restrict__ nbv, double* __restrict__ rr, int4* __restrict__ stv,
                             double2* __restrict__ rho1) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= R) return;
  const double nb = sqrt(nrm[c].x), nr = sqrt(nr0[c].x);
  nbv[c] = nb;
  rr[c] = rel_res(nr, nb);
  const bool brk = rhon[c].x == 0.0 && rhon[c].y == 0.0;
  const bool conv = brk && (nr < errmax * nb || nr == 0.0);
  stv[c] = make_int4(brk ? 0 : 1, 1, conv ? 1 : 0, 0);
  rho1[c] = rhon[c];
}

// start of an iteration: a column that has used maxit iterations stops
__global__ void k_b_cgs_start(int R, int maxit, int4* __restrict__ stv, int* __restrict__ active,
                              int* __restrict__ first) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= R) return;
  int4 v = stv[c];
  if (v.x && v.w >= maxit) v.x = 0;
  if (v.x) v.w += 1;
  stv[c] = v;
  active[c] = v.x;
  first[c] = v.y;
}

// end of an iteration: the stopping test of CGSold on the recurrence residual
// (modIterativo.f90:453-457) and the breakdown exit (:397), then the rho shift
__global__ void k_b_cgs_check(int R, double errmax, const double2* __restrict__ nrm,
                              const double2* __restrict__ rhon, const double* __restrict__ nbv,
                              double* __restrict__ rr, int4* __restrict__ stv, double2* __restrict__ rho1,
                              double2* __restrict__ rho2) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= R) return;
  int4 v = stv[c];
  if (v.x) {
    const double nr = sqrt(nrm[c].x);
    rr[c] = rel_res(nr, nbv[c]);
    v.y = 0;
    if (nr < errmax * nbv[c]) {
      v.z = 1;
      v.x = 0;
    } else if (rhon[c].x == 0.0 && rhon[c].y == 0.0) {
      v.x = 0;
    }
    stv[c] = v;
  }
  rho2[c] = rho1[c];
  rho1[c] = rhon[c];
}

// Restart of column c (persistent sweeps): b and x of the new problem come in
// contiguous (n); tmp holds Z x.  Sets b, x, r = R~ = b - Z x in the batch
// layout, then (one workgroup) |b|, rho1 = r.R~ and a fresh state.
__global__ void k_b_col_set(int R, int n, int c, const double2* __restrict__ bc, const double2* __restrict__ xc,
                            const double2* __restrict__ zx, double2* __restrict__ b, double2* __restrict__ x,
                            double2* __restrict__ r, double2* __restrict__ rt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t at = (size_t)i * R + c;
  const double2 ri = dsub(bc[i], zx[i]);
  b[at] = bc[i];
  x[at] = xc[i];
  r[at] = ri;
  rt[at] = ri;
}

__global__ void __launch_bounds__(256) k_b_col_init(int R, int n, int c, double errmax,
                                                    const double2* __restrict__ bc, const double2* __restrict__ zx,
                                                    double* __restrict__ nbv, double* __restrict__ rr,
                                                    int4* __restrict__ stv, double2* __restrict__ rho1) {
  __shared__ double sh[4][256];
  double nb = 0.0, qr = 0.0, qi = 0.0, nr = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double2 bi = bc[i], ri = dsub(bc[i], zx[i]);
    nb += bi.x * bi.x + bi.y * bi.y;
    qr += ri.x * ri.x - ri.y * ri.y;  // unconjugated r . r
    qi += 2.0 * ri.x * ri.y;
    nr += ri.x * ri.x + ri.y * ri.y;
  }
  sh[0][threadIdx.x] = nb;
  sh[1][threadIdx.x] = qr;
  sh[2][threadIdx.x] = qi;
  sh[3][threadIdx.x] = nr;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nb = sqrt(sh[0][0]), nr = sqrt(sh[3][0]);
    nbv[c] = nb;
    rr[c] = rel_res(nr, nb);
    const bool brk = sh[1][0] == 0.0 && sh[2][0] == 0.0;
    const bool conv = brk && (nr < errmax * nb || nr == 0.0);
    stv[c] = make_int4(brk ? 0 : 1, 1, conv ? 1 : 0, 0);
    rho1[c] = make_double2(sh[1][0], sh[2][0]);
  }
}

__global__ void k_b_col_get(int R, int n, int c, const double2* __restrict__ x, double2* __restrict__ xc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xc[i] = x[(size_t)i * R + c];
}

// ---- true residuals with the FP64 operator (batches whose far field is FP32)
//
// A pass computes t = Z64 x for every column (the FP64-far-field operator)
// and, per column, mode 1 (an active column: its recurrence residual r is
// replaced by b - t and rho(1) recomputed from it, a reliable update of
// CGSold's recurrence) or mode 2 (a column that met its tolerance on the
// recurrence and has not been verified: it stays converged only if its true
// residual also meets the tolerance; otherwise CGSold starts afresh from its
// iterate, with R~ = r = b - t).
__global__ void k_b_tr_mode(int R, int replace, const int4* __restrict__ stv, const int* __restrict__ verified,
                            int* __restrict__ mode) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= R) return;
  const int4 v = stv[c];
  mode[c] = (v.x && replace) ? 1 : (v.z && !verified[c]) ? 2 : 0;
}

__global__ void k_b_tr_resid(int R, int n, const int* __restrict__ mode, const double2* __restrict__ b,
                             const double2* __restrict__ t, double2* __restrict__ r) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * R) return;
  if (mode[i % R]) r[i] = dsub(b[i], t[i]);
}

// nrm = |r|^2, rhon = r.R~ (unconjugated), rrun = r.r (unconjugated)
__global__ void k_b_tr_state(int R, double errmax, int maxit, const int* __restrict__ mode,
                             const double2* __restrict__ nrm, const double2* __restrict__ rhon,
                             const double2* __restrict__ rrun, const double* __restrict__ nbv,
                             double* __restrict__ rr, int4* __restrict__ stv, int* __restrict__ verified,
                             double2* __restrict__ rho1, int* __restrict__ restart) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= R) return;
  restart[c] = 0;
  const int m = mode[c];
  if (!m) return;
  const double nr = sqrt(nrm[c].x), nb = nbv[c];
  rr[c] = rel_res(nr, nb);
  const bool ok = nr < errmax * nb || nr == 0.0;
  int4 v = stv[c];
  if (m == 1) {
    rho1[c] = rhon[c];
    if (ok) {
      v.x = 0;
      v.z = 1;
      verified[c] = 1;
    } else if (rhon[c].x == 0.0 && rhon[c].y == 0.0) {
      v.x = 0;  // CGSold's breakdown exit at the next iteration
    }
  } else if (ok) {
    verified[c] = 1;
  } else {
    v.z = 0;
    if (v.w < maxit && !(rrun[c].x == 0.0 && rrun[c].y == 0.0)) {
      v.x = 1;
      v.y = 1;
      rho1[c] = rrun[c];
      restart[c] = 1;
    }
  }
  stv[c] = v;
}

__global__ void k_b_tr_restart(int R, int n, const int* __restrict__ restart, const double2* __restrict__ r,
                               double2* __restrict__ rt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n * R) return;
  if (restart[i % R]) rt[i] = r[i];
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

// Far-field arrays of one level; the scalar type (float or double) is the
// batch's far-field precision.
struct BatchLevel {
  int nk = 0, nb = 0;
  // factored interpolation stencil to this level from level n+1 (n < L)
  int *sth_idx = nullptr, *sph_idx = nullptr, *pcu_ptr = nullptr;  // pcu_ptr: children at n+1
  void *sth_w = nullptr, *sph_w = nullptr;
  void* ph = nullptr;  // (8, nk) child-offset phases at this level's samples (n < L)
  int* oct = nullptr;  // octant of each box within its parent (n > 3)
  // fused translation + downward pass at this level (parents at level n-1)
  int nP = 0, nwt = 0;
  int* pc_ptr = nullptr;  // (nP+1) children of each parent
  int* bcd = nullptr;     // (nb) offset-code part of each box's octant
  int* boct = nullptr;    // (nb) octant of each box within its parent
  int* ps_ptr = nullptr;  // (nP+1) flat source list of each parent, a multiple of kSrcU entries
  int2* ps_ent = nullptr; // (box d, offset-code part of the pair minus 57)
  void* Tk = nullptr;     // (nk, 343) pes-scaled transfer function per offset code
  int *wtp_ptr = nullptr, *wtp_row = nullptr;  // transposed stencil from level n-1, padded (n > 3)
  void* wp = nullptr;     // (8, nwt) anterpolation weights of the padded entries (n > 3)
  void *C = nullptr, *A = nullptr;
};

struct Batch {
  Handle* h = nullptr;
  int R = 0;
  bool fp64 = false;  // far-field precision
  int maxnf = 0;      // largest number of bases in a leaf box
  int *ncol_ptr = nullptr, *ncols = nullptr;  // near columns of each leaf box (shared by its rows)
  void* rad = nullptr;                        // leaf patterns in the far-field precision
  void* rad_h = nullptr;                      // their half rows (the receive's)
  std::vector<BatchLevel> lv;
  // tile shapes per level, set at creation and by lw_batch_tune: the parent
  // ranges of the far kernel (1, 2, 4 or 8; see k_b_far), its variant
  // (kFarVariants) and the theta block of the upward pass (kUpBlocks)
  int far_split[16] = {};
  int far_var[16] = {};
  int far_sep[16] = {};  // downward pass in its own kernel (k_b_anterp)
  int up_tb[16] = {};
  double2 *X = nullptr, *Y = nullptr, *w[12] = {};
  void* Yf = nullptr;
  double2* Yn = nullptr;  // near field of the matvec in progress (FP64)
  hipEvent_t ev_fork = nullptr, ev_near = nullptr;  // near field on the handle's side stream
  double* part = nullptr;
  double2* sc = nullptr;  // 8 per-column scalar rows of R
  int *active = nullptr, *first = nullptr;
  int4* stv = nullptr;                     // per-column CGS state
  double *nbv = nullptr, *rrv = nullptr;   // per-column |b| and relative residual
  hipGraphExec_t iter_graph[2] = {nullptr, nullptr};  // one CGS iteration, without / with ILU0
  double graph_errmax[2] = {0.0, 0.0};
  int graph_maxit[2] = {0, 0};
  hipGraphExec_t cur_graph = nullptr;  // the graph of the solve in progress
  double errmax = 0.0;                 // the stopping tolerance of the solve in progress
  int maxit = 0;                       // and its iteration cap
  // FP32 far field with FP64 true residuals: the FP64-far-field operator on
  // the same handle, the replacement period (iterations, 0 = verification
  // only) and the per-column verification state
  Batch* ref = nullptr;
  int ru_period = 0;
  long long last_ru = 0;
  int *verified = nullptr, *tr_mode = nullptr, *tr_restart = nullptr;
  long long tr_passes = 0;
  long long done = 0, next_report = 0;  // iterations replayed since the start
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

template <typename T>
static int upload_b(Batch* b, T** p, const std::vector<T>& v) {
  int rc = bmalloc(b, p, v.size());
  if (rc) return rc;
  // ordered on the handle's (non-blocking) stream and completed before
  // returning: callers pass temporaries
  hipStream_t st = b->h->stream;
  if (!v.empty() && (hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess ||
                     hipStreamSynchronize(st) != hipSuccess)) {
    set_error("lw_batch_create: upload failed");
    return LW_ERR_HIP;
  }
  return LW_OK;
}

// Interaction lists of level n regrouped by parent pairs (see k_b_far).
struct FarBlocks {
  std::vector<int> bo, q_ptr, code_tr;
  std::vector<int2> q_ent;
  std::vector<int> bcd;                   // per box: offset-code part of its octant
  std::vector<int> ps_ptr;                // per parent: flat padded source list (see k_b_far)
  std::vector<int2> ps_ent;
};

// Derives the parent-pair blocks from the plan's per-box far lists and checks
// that they reproduce them exactly: the same far boxes for every box and one
// transfer vector per offset code.  Returns false with msg on a mismatch.
static bool far_blocks(const Plan& P, int n, FarBlocks& fb, std::string& msg) {
  const LevelPlan& lp = P.lv[n];
  const LevelPlan& pp = P.lv[n - 1];
  const int nb = lp.nb, nP = pp.nb;
  fb.bo.assign(nb, 0);
  for (int c = 0; c < nb; ++c) {
    const double* cc = &lp.centre[3 * c];
    const double* pc = &pp.centre[3 * lp.parent[c]];
    fb.bo[c] = ((cc[0] > pc[0]) ? 4 : 0) | ((cc[1] > pc[1]) ? 2 : 0) | ((cc[2] > pc[2]) ? 1 : 0);
  }
  for (int Pb = 0; Pb < nP; ++Pb)
    for (int c = pp.child_ptr[Pb]; c < pp.child_ptr[Pb + 1]; ++c)
      if (lp.parent[c] != Pb) {
        msg = "children of a parent are not contiguous";
        return false;
      }
  std::vector<std::vector<int>> qs(nP);
  for (int c = 0; c < nb; ++c)
    for (int m = lp.far_ptr[c]; m < lp.far_ptr[c + 1]; ++m) qs[lp.parent[c]].push_back(lp.parent[lp.far_box[m]]);
  fb.q_ptr.assign(nP + 1, 0);
  fb.q_ent.clear();
  fb.code_tr.assign(kOffCodes, -1);
  long long generated = 0;
  std::vector<int> gen, ref;
  for (int Pb = 0; Pb < nP; ++Pb) {
    auto& q = qs[Pb];
    std::sort(q.begin(), q.end());
    q.erase(std::unique(q.begin(), q.end()), q.end());
    for (int Q : q) {
      int o[3];
      for (int k = 0; k < 3; ++k) {
        const double dv = (pp.centre[3 * Q + k] - pp.centre[3 * Pb + k]) / pp.lado;
        o[k] = (int)std::lround(dv);
        if (std::fabs(dv - o[k]) > 1e-6 || o[k] < -1 || o[k] > 1) {
          msg = "far box outside the parent's neighbourhood";
          return false;
        }
      }
      fb.q_ent.push_back(make_int2(Q, (2 * o[0] + 3) * 49 + (2 * o[1] + 3) * 7 + (2 * o[2] + 3)));
    }
    fb.q_ptr[Pb + 1] = (int)fb.q_ent.size();
    // regenerate every child's far list from the blocks and compare
    for (int c = pp.child_ptr[Pb]; c < pp.child_ptr[Pb + 1]; ++c) {
      gen.clear();
      for (int e = fb.q_ptr[Pb]; e < fb.q_ptr[Pb + 1]; ++e) {
        const int Q = fb.q_ent[e].x, oc = fb.q_ent[e].y;
        const int o[3] = {(oc / 49 - 3) / 2, ((oc / 7) % 7 - 3) / 2, (oc % 7 - 3) / 2};
        for (int d = pp.child_ptr[Q]; d < pp.child_ptr[Q + 1]; ++d) {
          int off[3], far = 0;
          for (int k = 0; k < 3; ++k) {
            const int bit = 2 - k;
            off[k] = 2 * o[k] + ((fb.bo[d] >> bit) & 1) - ((fb.bo[c] >> bit) & 1);
            far |= off[k] > 1 || off[k] < -1;
          }
          if (far) gen.push_back(d * kOffCodes + (off[0] + 3) * 49 + (off[1] + 3) * 7 + (off[2] + 3));
        }
      }
      ref.clear();
      for (int m = lp.far_ptr[c]; m < lp.far_ptr[c + 1]; ++m) {
        const int f = lp.far_box[m];
        int code = 0;
        for (int k = 0; k < 3; ++k) {
          const int off = (int)std::lround((lp.centre[3 * f + k] - lp.centre[3 * c + k]) / lp.lado);
          if (off < -3 || off > 3) {
            msg = "far box offset outside [-3, 3]";
            return false;
          }
          code = code * 7 + (off + 3);
        }
        int& slot = fb.code_tr[code];
        if (slot < 0) slot = lp.far_tr[m];
        if (slot != lp.far_tr[m]) {
          msg = "two transfer vectors for one box offset";
          return false;
        }
        ref.push_back(f * kOffCodes + code);
      }
      std::sort(gen.begin(), gen.end());
      std::sort(ref.begin(), ref.end());
      if (gen != ref) {
        msg = "interaction lists are not parent-neighbour blocks";
        return false;
      }
      generated += (long long)gen.size();
    }
  }
  if (generated != (long long)lp.far_box.size()) {
    msg = "interaction list count mismatch";
    return false;
  }
  fb.bcd.resize(nb);
  for (int c = 0; c < nb; ++c) fb.bcd[c] = oct_code(fb.bo[c]);
  // flat source lists, padded with entries whose codes are near offsets
  for (int o = 0; o < 8; ++o)
    if (fb.code_tr[kPadCode + oct_code(7) - oct_code(o)] >= 0) {
      msg = "padding code has a transfer vector";
      return false;
    }
  fb.ps_ptr.assign(nP + 1, 0);
  fb.ps_ent.clear();
  for (int Pb = 0; Pb < nP; ++Pb) {
    for (int e = fb.q_ptr[Pb]; e < fb.q_ptr[Pb + 1]; ++e) {
      const int Q = fb.q_ent[e].x;
      for (int d = pp.child_ptr[Q]; d < pp.child_ptr[Q + 1]; ++d)
        fb.ps_ent.push_back(make_int2(d, fb.q_ent[e].y + fb.bcd[d] - oct_code(7)));
    }
    while ((fb.ps_ent.size() - fb.ps_ptr[Pb]) % kSrcU) fb.ps_ent.push_back(make_int2(pp.child_ptr[Pb], kPadCode));
    fb.ps_ptr[Pb + 1] = (int)fb.ps_ent.size();
  }
  for (int Pb = 0; Pb < nP; ++Pb) {
    int mask = 0;
    for (int c = pp.child_ptr[Pb]; c < pp.child_ptr[Pb + 1]; ++c) {
      if (mask & (1 << fb.bo[c])) {
        msg = "two children of one parent in the same octant";
        return false;
      }
      mask |= 1 << fb.bo[c];
    }
  }
  return true;
}

static void release_batch(Batch* b) {
  if (!b) return;
  if (b->h && b->h->stream) (void)hipStreamSynchronize(b->h->stream);
  for (auto& g : b->iter_graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (void* p : b->allocs) (void)hipFree(p);
  b->allocs.clear();
  for (hipEvent_t* e : {&b->ev_fork, &b->ev_near})
    if (*e) {
      (void)hipEventDestroy(*e);
      *e = nullptr;
    }
  if (b->ref) {
    release_batch(b->ref);
    delete b->ref;
    b->ref = nullptr;
  }
}



// bases per LDS chunk of the aggregation kernel (x rows of R columns)
template <typename S, int R>
constexpr int agg_chunk() {
  return std::min(64, (int)(49152 / ((R + 8) * 2 * sizeof(S))) / 4 * 4);
}

// The leaf aggregation (up) or the receive; with Y, the receive combines
// its far field with the near field Yn into Y.
template <typename S, int R>
static void launch_leaf(Batch* bt, bool up, const double2* X, double2* Y = nullptr) {
  typedef typename Cx<S>::T T2;
  Handle* h = bt->h;
  BatchLevel& lf = bt->lv[h->L];
  if (up) {
    constexpr int KC = agg_chunk<S, R>();
    hipLaunchKernelGGL((k_b_aggregate<S, R, KC>), dim3(lf.nb), dim3(256), 0, h->stream, lf.nk,
                       lf.nb, h->d_box_slot, h->d_slot_basis, (const T2*)bt->rad, X, (T2*)lf.C);
  } else {
    const double2 cc = make_double2(0.0, h->plan.phys.w * kMu / (4.0 * kPi));
    hipLaunchKernelGGL((k_b_receive<S, R>), dim3(lf.nb), dim3(128), 0, h->stream, lf.nk, h->plan.lv[h->L].nth,
                       lf.nb, h->d_box_slot,
                       h->d_slot_basis, (const T2*)bt->rad_h, (const T2*)lf.A, (T2*)bt->Yf, cc,
                       (const double2*)bt->Yn, Y);
  }
}

// Tile shapes the autotuner chooses from, per level: the theta block of the
// upward pass and the (sources per group, columns per lane) of the far kernel.
constexpr int kUpBlocks[3] = {4, 8, 16};
// (sources per group, columns per lane, workgroup order (k_b_far), transfer
// row in LDS)
constexpr int kNFar = 16;
constexpr int kFarVariants[kNFar][4] = {{4, 2, 0, 0}, {2, 2, 0, 0}, {4, 1, 0, 0}, {2, 1, 0, 0},
                                        {4, 2, 1, 0}, {2, 2, 1, 0}, {4, 1, 1, 0}, {2, 1, 1, 0},
                                        {4, 2, 0, 1}, {2, 2, 0, 1}, {4, 1, 0, 1}, {4, 1, 1, 1},
                                        {4, 2, 2, 0}, {2, 2, 2, 0}, {4, 1, 2, 0}, {2, 1, 2, 0}};

template <typename S, int TB>
static void launch_up_tb(Batch* bt, int n) {
  typedef typename Cx<S>::T T2;
  Handle* h = bt->h;
  BatchLevel& c = bt->lv[n];
  BatchLevel& p = bt->lv[n - 1];
  const LevelPlan& pp = h->plan.lv[n - 1];
  const LevelPlan& cp = h->plan.lv[n];
  const int ncol = 2 * bt->R, nchunk = ncol / kWave;
  const long long waves = (long long)p.nb * pp.nph * ((pp.nth + TB - 1) / TB) * nchunk;
  hipLaunchKernelGGL((k_b_interp_up<S, TB>), dim3(nblk(waves, 4)), dim3(256), 0, h->stream, ncol, pp.nth, p.nk,
                     p.nb, cp.nth, c.nb, pp.nph, p.pcu_ptr, c.oct, p.sth_idx, (const S*)p.sth_w, p.sph_idx,
                     (const S*)p.sph_w, (const T2*)p.ph, (const T2*)c.C, (T2*)p.C);
}

// C_{n-1} from C_n (the upward pass into level n-1) with theta block tb
template <typename S>
static void launch_up(Batch* bt, int n, int tb) {
  if (tb == 4)
    launch_up_tb<S, 4>(bt, n);
  else if (tb == 16)
    launch_up_tb<S, 16>(bt, n);
  else
    launch_up_tb<S, 8>(bt, n);
}

template <typename S, int SU, int CPL, bool TL>
static void launch_far_v(Batch* bt, int n, int ns, int order, bool split) {
  typedef typename Cx<S>::T T2;
  Handle* h = bt->h;
  BatchLevel& d = bt->lv[n];
  const BatchLevel* pl = n > 3 ? &bt->lv[n - 1] : nullptr;
  // split: the translation alone here, the downward pass by k_b_anterp
  const T2* Ac = (pl && !split) ? (const T2*)pl->A : nullptr;
  const int ncol = 2 * bt->R, fnchunk = ncol / (kWave * CPL);
  const int wpx = nblk(nblk(d.nP, ns), 4) * fnchunk;
  const long long grid = order == 2 ? 8LL * nblk(nblk(d.nk, 4), 8 / ns) * nblk(d.nP, ns) * fnchunk
                                    : 8LL * nblk(d.nk, 8 / ns) * wpx;
  hipLaunchKernelGGL((k_b_far<S, SU, CPL, TL>), dim3((unsigned)grid), dim3(256), 0, h->stream, d.nk, d.nP, fnchunk,
                     wpx, ns, order, d.nb, ncol, pl ? pl->nb : 0, d.nwt, d.pc_ptr, d.bcd, d.boct, d.ps_ptr, d.ps_ent,
                     (const T2*)d.Tk, (const T2*)d.C, d.wtp_ptr, d.wtp_row, (const T2*)d.wp, Ac, (T2*)d.A);
  if (pl && split) {
    const int achunk = ncol / kWave;
    const long long ag = (long long)d.nP * achunk * nblk(d.nk, kAntK);
    hipLaunchKernelGGL(k_b_anterp<S>, dim3((unsigned)ag), dim3(256), 0, h->stream, d.nk, d.nP, achunk, d.nb, ncol,
                       pl->nb, d.nwt, d.pc_ptr, d.boct, d.wtp_ptr, d.wtp_row, (const T2*)d.wp, (const T2*)pl->A,
                       (T2*)d.A);
  }
}

// A_n (translation + downward pass at level n) with far variant v, ns
// parent ranges, and the downward pass fused or in its own kernel (split)
template <typename S>
static void launch_far(Batch* bt, int n, int v, int ns, bool split) {
  const int* f = kFarVariants[v];
  const int cout = f[2], shape = f[0] * 10 + f[1];  // cout: the workgroup order
  if (f[3]) {
    switch (shape) {
      case 22: launch_far_v<S, 2, 2, true>(bt, n, ns, cout, split); break;
      case 41: launch_far_v<S, 4, 1, true>(bt, n, ns, cout, split); break;
      default: launch_far_v<S, 4, 2, true>(bt, n, ns, cout, split); break;
    }
    return;
  }
  switch (shape) {
    case 22: launch_far_v<S, 2, 2, false>(bt, n, ns, cout, split); break;
    case 41: launch_far_v<S, 4, 1, false>(bt, n, ns, cout, split); break;
    case 21: launch_far_v<S, 2, 1, false>(bt, n, ns, cout, split); break;
    default: launch_far_v<S, 4, 2, false>(bt, n, ns, cout, split); break;
  }
}

// One batched matvec.  The near field (FP64 MFMA; it reads only X) runs on
// the handle's first side stream into Yn, beside the far field's VALU- and
// latency-bound passes on the main stream; the receive then adds it.
template <typename S>
static int batch_matvec_t(Batch* bt, const double2* X, double2* Y) {
  typedef typename Cx<S>::T T2;
  Handle* h = bt->h;
  const int R = bt->R, L = h->L;
  hipStream_t st = h->stream, sn = h->side[0];
  BatchLevel& lf = bt->lv[L];
  LW_HIP(hipEventRecord(bt->ev_fork, st));
  LW_HIP(hipStreamWaitEvent(sn, bt->ev_fork, 0));
  const double2 cc = make_double2(0.0, h->plan.phys.w * kMu / (4.0 * kPi));
  if (h->fp64_near)
    hipLaunchKernelGGL((k_b_near_box<double2, T2>), dim3(lf.nb), dim3(128), 0, sn, R, h->d_box_slot,
                       h->d_slot_basis, bt->ncol_ptr, bt->ncols, h->d_ia, (const double2*)h->d_z, X,
                       (const T2*)nullptr, cc, bt->Yn);
  else
    hipLaunchKernelGGL((k_b_near_box<float2, T2>), dim3(lf.nb), dim3(128), 0, sn, R, h->d_box_slot,
                       h->d_slot_basis, bt->ncol_ptr, bt->ncols, h->d_ia, (const float2*)h->d_z, X,
                       (const T2*)nullptr, cc, bt->Yn);
  LW_HIP(hipEventRecord(bt->ev_near, sn));
  if (R == 64)
    launch_leaf<S, 64>(bt, true, X);
  else
    launch_leaf<S, 128>(bt, true, X);
  for (int n = L; n >= 4; --n) launch_up<S>(bt, n, bt->up_tb[n]);
  for (int n = 3; n <= L; ++n) launch_far<S>(bt, n, bt->far_var[n], bt->far_split[n], bt->far_sep[n] != 0);
  LW_HIP(hipStreamWaitEvent(st, bt->ev_near, 0));
  if (R == 64)
    launch_leaf<S, 64>(bt, false, nullptr, Y);
  else
    launch_leaf<S, 128>(bt, false, nullptr, Y);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

static int batch_matvec(Batch* bt, const double2* X, double2* Y) {
  return bt->fp64 ? batch_matvec_t<double>(bt, X, Y) : batch_matvec_t<float>(bt, X, Y);
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

constexpr int kGraphGroup = 8;

// One CGSold iteration for every active column (modIterativo.f90:376-486),
// enqueued on the handle's stream; it is captured once into a graph.
static int batch_cgs_iteration(Batch* bt, bool use_prec, double errmax, int maxit) {
  Handle* h = bt->h;
  const int R = bt->R, n = h->ne;
  hipStream_t st = h->stream;
  const int vb = nblk((long long)n * R, 256);
  double2 *r = bt->w[1], *R_ = bt->w[2], *u = bt->w[3], *p = bt->w[4], *P_ = bt->w[5], *V = bt->w[6],
          *q = bt->w[7], *U = bt->w[8], *Q = bt->w[9], *t = bt->w[10], *x = bt->X;
  double2 *rho1 = bt->sc, *rho2 = bt->sc + R, *den = bt->sc + 2 * R, *alfa = bt->sc + 3 * R, *nrm = bt->sc + 4 * R,
          *rhon = bt->sc + 5 * R;
  int rc;
  hipLaunchKernelGGL(k_b_cgs_start, dim3(nblk(R, 64)), dim3(64), 0, st, R, maxit, bt->stv, bt->active, bt->first);
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
  hipLaunchKernelGGL(k_b_cgs_check, dim3(nblk(R, 64)), dim3(64), 0, st, R, errmax, nrm, rhon, bt->nbv, bt->rrv,
                     bt->stv, rho1, rho2);
  LW_HIP(hipGetLastError());
  return LW_OK;
}

// CGS on every column from b = w[0] and the initial guess in X: the initial
// residual, then captured iterations replayed in groups of kGraphGroup with
// one host check of the column states per group.
// Initial residual of every column from b = w[0] and the guess in X, and
// the iteration graph for (use_prec, errmax, maxit).
static int batch_cgs_start(Batch* bt, double errmax, bool use_prec, int maxit) {
  Handle* h = bt->h;
  const int R = bt->R;
  const long long nv = (long long)h->ne * R;
  hipStream_t st = h->stream;
  double2 *b = bt->w[0], *r = bt->w[1], *R_ = bt->w[2], *t = bt->w[10];
  double2 *rho1 = bt->sc, *nrm = bt->sc + 4 * R, *rhon = bt->sc + 5 * R, *nr0 = bt->sc + 6 * R;
  int rc = batch_matvec(bt, bt->X, t);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_sub, dim3(nblk(nv, 256)), dim3(256), 0, st, nv, b, t, r);
  LW_HIP(hipMemcpyAsync(R_, r, sizeof(double2) * nv, hipMemcpyDeviceToDevice, st));
  if ((rc = batch_dot2(bt, b, b, 1, r, R_, 0, nrm, rhon))) return rc;
  if ((rc = batch_dot2(bt, r, r, 1, nullptr, nullptr, 0, nr0, nullptr))) return rc;
  hipLaunchKernelGGL(k_b_cgs_init, dim3(nblk(R, 64)), dim3(64), 0, st, R, errmax, nrm, rhon, nr0, bt->nbv,
                     bt->rrv, bt->stv, rho1);
  LW_HIP(hipGetLastError());
  // the iteration graph bakes in errmax and maxit: recapture when they change
  hipGraphExec_t& gx = bt->iter_graph[use_prec ? 1 : 0];
  if (gx && (bt->graph_errmax[use_prec] != errmax || bt->graph_maxit[use_prec] != maxit)) {
    LW_HIP(hipGraphExecDestroy(gx));
    gx = nullptr;
  }
  if (!gx) {
    hipGraph_t g = nullptr;
    LW_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = batch_cgs_iteration(bt, use_prec, errmax, maxit);
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    LW_HIP(e);
    e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    LW_HIP(e);
    bt->graph_errmax[use_prec] = errmax;
    bt->graph_maxit[use_prec] = maxit;
  }
  bt->cur_graph = gx;
  bt->errmax = errmax;
  bt->maxit = maxit;
  bt->done = 0;
  bt->last_ru = 0;
  if (bt->ref) LW_HIP(hipMemsetAsync(bt->verified, 0, sizeof(int) * R, st));
  return LW_OK;
}

// One true-residual pass (see k_b_tr_mode); replace: also replace the
// recurrence residuals of the active columns.
static int batch_true_residual(Batch* bt, bool replace) {
  Handle* h = bt->h;
  const int R = bt->R, n = h->ne;
  hipStream_t st = h->stream;
  const int vb = nblk((long long)n * R, 256), cb = nblk(R, 64);
  double2 *b = bt->w[0], *r = bt->w[1], *R_ = bt->w[2], *t = bt->w[11];
  double2 *rho1 = bt->sc, *nrm = bt->sc + 4 * R, *rhon = bt->sc + 5 * R, *rrun = bt->sc + 7 * R;
  int rc = batch_matvec(bt->ref, bt->X, t);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_tr_mode, dim3(cb), dim3(64), 0, st, R, replace ? 1 : 0, bt->stv, bt->verified, bt->tr_mode);
  hipLaunchKernelGGL(k_b_tr_resid, dim3(vb), dim3(256), 0, st, R, n, bt->tr_mode, b, t, r);
  if ((rc = batch_dot2(bt, r, r, 1, r, R_, 0, nrm, rhon))) return rc;
  if ((rc = batch_dot2(bt, r, r, 0, nullptr, nullptr, 0, rrun, nullptr))) return rc;
  hipLaunchKernelGGL(k_b_tr_state, dim3(cb), dim3(64), 0, st, R, bt->errmax, bt->maxit, bt->tr_mode, nrm, rhon, rrun,
                     bt->nbv, bt->rrv, bt->stv, bt->verified, rho1, bt->tr_restart);
  hipLaunchKernelGGL(k_b_tr_restart, dim3(vb), dim3(256), 0, st, R, n, bt->tr_restart, r, R_);
  LW_HIP(hipGetLastError());
  ++bt->tr_passes;
  return LW_OK;
}

// Up to `iterations` replays of the iteration graph, in groups of
// kGraphGroup with one host check of the column states per group; stops
// early when no column is active.  sv/rr receive the states.
static int batch_cgs_advance(Batch* bt, int iterations, std::vector<int4>& sv, std::vector<double>& rr) {
  Handle* h = bt->h;
  const int R = bt->R;
  hipStream_t st = h->stream;
  if (!bt->cur_graph) {
    set_error("lw_batch: no solve started");
    return LW_ERR_STATE;
  }
  sv.resize(R);
  rr.resize(R);
  int left = iterations;
  while (left > 0) {
    const int g = std::min(kGraphGroup, left);
    for (int j = 0; j < g; ++j) LW_HIP(hipGraphLaunch(bt->cur_graph, st));
    left -= g;
    bt->done += g;
    LW_HIP(hipMemcpyAsync(sv.data(), bt->stv, sizeof(int4) * R, hipMemcpyDeviceToHost, st));
    LW_HIP(hipMemcpyAsync(rr.data(), bt->rrv, sizeof(double) * R, hipMemcpyDeviceToHost, st));
    LW_HIP(hipStreamSynchronize(st));
    if (bt->ref) {
      // FP32 far field: replace the active columns' residuals every
      // ru_period iterations, and verify newly converged columns with the
      // FP64 operator before anyone sees them converged
      std::vector<int> ver(R);
      LW_HIP(hipMemcpyAsync(ver.data(), bt->verified, sizeof(int) * R, hipMemcpyDeviceToHost, st));
      LW_HIP(hipStreamSynchronize(st));
      bool verify = false, any_active = false;
      for (int c = 0; c < R; ++c) {
        verify = verify || (sv[c].z && !ver[c]);
        any_active = any_active || sv[c].x;
      }
      const bool replace = any_active && bt->ru_period > 0 && bt->done - bt->last_ru >= bt->ru_period;
      if (verify || replace) {
        if (int rc = batch_true_residual(bt, replace)) return rc;
        if (replace) bt->last_ru = bt->done;
        LW_HIP(hipMemcpyAsync(sv.data(), bt->stv, sizeof(int4) * R, hipMemcpyDeviceToHost, st));
        LW_HIP(hipMemcpyAsync(rr.data(), bt->rrv, sizeof(double) * R, hipMemcpyDeviceToHost, st));
        LW_HIP(hipStreamSynchronize(st));
      }
    }
    int na = 0;
    double worst = 0.0;
    for (int c = 0; c < R; ++c)
      if (sv[c].x) {
        ++na;
        worst = std::max(worst, rr[c]);
      }
    if (progress_every() > 0 && bt->done >= bt->next_report) {
      std::fprintf(stderr, "[lw_batch_cgs] iteration %lld: %d of %d columns active, worst relres %.3e\n",
                   bt->done, na, R, worst);
      bt->next_report += progress_every();
    }
    if (!na) break;
  }
  return LW_OK;
}

// CGS on every column from b = w[0] and the initial guess in X.
static int batch_cgs_run(Batch* bt, double errmax, bool use_prec, int maxit) {
  int rc = batch_cgs_start(bt, errmax, use_prec, maxit);
  if (rc) return rc;
  bt->next_report = progress_every();
  std::vector<int4> sv;
  std::vector<double> rr;
  return batch_cgs_advance(bt, maxit, sv, rr);
}

// Far-field tables of every level in the precision S (patterns, phases,
// stencil weights, per-offset transfer functions, anterpolation weights)
// and the campo arrays.
template <typename S>
static int setup_far(Batch* bt) {
  typedef typename Cx<S>::T T2;
  Handle* h = bt->h;
  const Plan& P = h->plan;
  const int L = h->L, R = bt->R;
  hipStream_t st = h->stream;
  int rc;
  const long long nrad = (long long)P.lv[L].nk * 2 * h->ne;
  if (sizeof(S) == sizeof(double)) {
    bt->rad = h->d_rad;  // the operator's own FP64 patterns
  } else {
    T2* r = nullptr;
    if ((rc = bmalloc(bt, &r, (size_t)nrad))) return rc;
    hipLaunchKernelGGL(k_to_prec<S>, dim3(nblk(nrad, 256)), dim3(256), 0, st, nrad, h->d_rad, r);
    bt->rad = r;
  }
  // the receive's half pattern rows (samples k < nk/2) in the far-field
  // precision; the antipodal samples are their conjugates
  const LevelPlan& ll = P.lv[L];
  if (ll.nph != 2 * ll.nth || ll.nk != ll.nth * ll.nph) {
    set_error("lw_batch_create: leaf sample grid without antipodal pairs");
    return LW_ERR_ARG;
  }
  {
    const long long nh = (long long)(ll.nk / 2) * 2 * h->ne;
    T2* r = nullptr;
    if ((rc = bmalloc(bt, &r, (size_t)nh))) return rc;
    hipLaunchKernelGGL(k_rad_half<S>, dim3(nblk(nh, 256)), dim3(256), 0, st, nh, ll.nk, h->d_rad, r);
    bt->rad_h = r;
  }
  for (int n = 3; n <= L; ++n) {
    const LevelPlan& lp = P.lv[n];
    DevLevel& d = h->lv[n];
    BatchLevel& b = bt->lv[n];
    b.nk = lp.nk;
    b.nb = lp.nb;
    if (n < L) {
      std::vector<S> tw(lp.sth_w.begin(), lp.sth_w.end()), pw(lp.sph_w.begin(), lp.sph_w.end());
      S *dtw = nullptr, *dpw = nullptr;
      if ((rc = upload_b(bt, &b.sth_idx, lp.sth_idx)) || (rc = upload_b(bt, &b.sph_idx, lp.sph_idx)) ||
          (rc = upload_b(bt, &dtw, tw)) || (rc = upload_b(bt, &dpw, pw)) ||
          (rc = upload_b(bt, &b.pcu_ptr, lp.child_ptr)))
        return rc;
      b.sth_w = dtw;
      b.sph_w = dpw;
      if (sizeof(S) == sizeof(double)) {
        b.ph = d.ph;
      } else {
        T2* ph = nullptr;
        if ((rc = bmalloc(bt, &ph, 8 * (size_t)lp.nk))) return rc;
        hipLaunchKernelGGL(k_to_prec<S>, dim3(nblk(8LL * lp.nk, 256)), dim3(256), 0, st, 8LL * lp.nk, d.ph, ph);
        b.ph = ph;
      }
    }
    b.oct = d.oct;
    FarBlocks fb;
    std::string msg;
    if (!far_blocks(P, n, fb, msg)) {
      set_error("lw_batch_create: " + msg);
      return LW_ERR_STATE;
    }
    b.nP = P.lv[n - 1].nb;
    bt->far_split[n] = b.nP >= 64 ? 8 : 1;
    bt->far_var[n] = 0;
    bt->far_sep[n] = 0;
    bt->up_tb[n] = kThBlock;
    int* d_code_tr = nullptr;
    if ((rc = upload_b(bt, &b.pc_ptr, P.lv[n - 1].child_ptr)) || (rc = upload_b(bt, &b.bcd, fb.bcd)) ||
        (rc = upload_b(bt, &b.boct, fb.bo)) || (rc = upload_b(bt, &b.ps_ptr, fb.ps_ptr)) ||
        (rc = upload_b(bt, &b.ps_ent, fb.ps_ent)) || (rc = upload_b(bt, &d_code_tr, fb.code_tr)))
      return rc;
    T2* tk = nullptr;
    if ((rc = bmalloc(bt, &tk, (size_t)lp.nk * kOffCodes))) return rc;
    hipLaunchKernelGGL(k_b_tk<S>, dim3(nblk((long long)lp.nk * kOffCodes, 256)), dim3(256), 0, st, lp.nk,
                       d_code_tr, d.T, d.pes, tk);
    b.Tk = tk;
    if (n > 3) {
      // the transposed stencil with every fine sample's entries padded to a
      // multiple of 4 (weight 0, row 0)
      const LevelPlan& pp = P.lv[n - 1];
      std::vector<int> wptr(lp.nk + 1, 0), wrow;
      std::vector<double> ww;
      for (int k = 0; k < lp.nk; ++k) {
        for (int e = pp.wt_ptr[k]; e < pp.wt_ptr[k + 1]; ++e) {
          wrow.push_back(pp.wt_row[e]);
          ww.push_back(pp.wt_w[e]);
        }
        while ((wrow.size() - wptr[k]) % 4) {
          wrow.push_back(0);
          ww.push_back(0.0);
        }
        wptr[k + 1] = (int)wrow.size();
      }
      b.nwt = (int)ww.size();
      double* d_ww = nullptr;
      T2* wp = nullptr;
      if ((rc = upload_b(bt, &b.wtp_ptr, wptr)) || (rc = upload_b(bt, &b.wtp_row, wrow)) ||
          (rc = upload_b(bt, &d_ww, ww)) || (rc = bmalloc(bt, &wp, 8 * (size_t)b.nwt)))
        return rc;
      hipLaunchKernelGGL(k_b_wp<S>, dim3(nblk(8LL * b.nwt, 256)), dim3(256), 0, st, pp.nk, b.nwt, b.wtp_row, d_ww,
                         h->lv[n - 1].ph, wp);
      b.wp = wp;
    }
    const size_t fsz = (size_t)lp.nk * lp.nb * 2 * R;
    if (fsz >= (1ull << 31)) {
      // the far kernel indexes a level's campo arrays with 32-bit offsets
      set_error("lw_batch_create: a level's field array exceeds 2^31 entries");
      return LW_ERR_ARG;
    }
    T2 *c = nullptr, *a = nullptr;
    if ((rc = bmalloc(bt, &c, fsz)) || (rc = bmalloc(bt, &a, fsz))) return rc;
    b.C = c;
    b.A = a;
  }
  return LW_OK;
}

int setup_far_parent(Handle* h) {
  const Plan& P = h->plan;
  const int L = h->L;
  std::vector<FarBlocks> fbs(L + 1);
  for (int n = 3; n <= L; ++n) {
    std::string msg;
    if (!far_blocks(P, n, fbs[n], msg)) {
      set_error("lw_mlfma_init: " + msg);
      return LW_ERR_STATE;
    }
  }
  int rc;
  for (int n = 3; n <= L; ++n) {
    DevLevel& d = h->lv[n];
    const FarBlocks& fb = fbs[n];
    d.nps = (long long)fb.ps_ent.size();
    int* d_code_tr = nullptr;
    if ((rc = upload(h, &d.pc_ptr, P.lv[n - 1].child_ptr)) || (rc = upload(h, &d.bcd, fb.bcd)) ||
        (rc = upload(h, &d.ps_ptr, fb.ps_ptr)) ||
        (rc = upload(h, &d.ps_ent, fb.ps_ent)) || (rc = upload(h, &d_code_tr, fb.code_tr)) ||
        (rc = dmalloc(h, &d.Tc, (size_t)d.nk * kOffCodes)))
      return rc;
    hipLaunchKernelGGL(k_tc, dim3(nblk((long long)d.nk * kOffCodes, 256)), dim3(256), 0, h->stream, d.nk, d_code_tr,
                       d.T, d.Tc);
    LW_HIP(hipGetLastError());
  }
  LW_HIP(hipStreamSynchronize(h->stream));
  return LW_OK;
}

void launch_translate_parent(Handle* h, int n, hipStream_t st) {
  DevLevel& d = h->lv[n];
  const int nP = h->plan.lv[n - 1].nb, nchk = nblk(d.nk, kWave);
  hipLaunchKernelGGL(k_translate_parent, dim3(nP * nchk), dim3(256), 0, st, d.nk, nchk,
                     d.pc_ptr, d.bcd, d.ps_ptr, d.ps_ent, d.Tc, d.pes, d.C, d.B);
}

}  // namespace lw

struct lw_batch {
  lw::Batch b;
};

lw::Handle* lw::batch_handle(lw_batch* b) { return b ? b->b.h : nullptr; }

namespace lw {
// The operator part of a batch (far-field tables in its precision, the near
// columns of each leaf box, the far-field receive buffer).
static int batch_operator(Batch* bt) {
  Handle* h = bt->h;
  const int L = h->L, R = bt->R;
  const Plan& P = h->plan;
  int rc;
#define TRY(x)        \
  do {                \
    rc = (x);         \
    if (rc) return rc; \
  } while (0)
  bt->lv.assign(L + 1, BatchLevel());
  TRY(bt->fp64 ? setup_far<double>(bt) : setup_far<float>(bt));
  for (int b = 0; b < P.lv[L].nb; ++b) bt->maxnf = std::max(bt->maxnf, P.box_slot[b + 1] - P.box_slot[b]);
  {
    // the near columns of each leaf box, checked to be those of every row of the box
    std::vector<int> cp(P.lv[L].nb + 1, 0), cl;
    for (int b = 0; b < P.lv[L].nb; ++b) {
      const int m0 = P.slot_basis[P.box_slot[b]];
      const int n0 = P.ia[m0 + 1] - P.ia[m0];
      for (int s = P.box_slot[b]; s < P.box_slot[b + 1]; ++s) {
        const int m = P.slot_basis[s];
        bool same = P.ia[m + 1] - P.ia[m] == n0;
        for (int j = 0; same && j < n0; ++j) same = P.ja[P.ia[m] + j] == P.ja[P.ia[m0] + j];
        if (!same) {
          set_error("lw_batch_create: rows of a leaf box have different near columns");
          return LW_ERR_STATE;
        }
      }
      cl.insert(cl.end(), P.ja.begin() + P.ia[m0], P.ja.begin() + P.ia[m0 + 1]);
      cp[b + 1] = (int)cl.size();
    }
    TRY(upload_b(bt, &bt->ncol_ptr, cp));
    TRY(upload_b(bt, &bt->ncols, cl));
  }
  TRY(bmalloc(bt, (char**)&bt->Yf, (size_t)h->ne * R * (bt->fp64 ? sizeof(double2) : sizeof(float2))));
  TRY(bmalloc(bt, &bt->Yn, (size_t)h->ne * R));
  if (hipEventCreateWithFlags(&bt->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&bt->ev_near, hipEventDisableTiming) != hipSuccess) {
    set_error("lw_batch_create: hipEventCreate failed");
    return LW_ERR_HIP;
  }
#undef TRY
  return LW_OK;
}
}  // namespace lw

extern "C" {

int lw_batch_create(lw_handle* hh, int nrhs, int flags, lw_batch** out) {
  using namespace lw;
  const int known = LW_BATCH_FAR_FP64 | LW_BATCH_FP64_RESIDUAL;
  if (!hh || !out || !(nrhs == 64 || nrhs == 128) || (flags & ~known) ||
      ((flags & LW_BATCH_FAR_FP64) && (flags & LW_BATCH_FP64_RESIDUAL))) {
    set_error("lw_batch_create: nrhs must be 64 or 128; flags: LW_BATCH_FAR_FP64 or LW_BATCH_FP64_RESIDUAL");
    return LW_ERR_ARG;
  }
  *out = nullptr;
  Handle* h = &hh->h;
  lw_batch* ob = new lw_batch();
  Batch* bt = &ob->b;
  bt->h = h;
  bt->R = nrhs;
  bt->fp64 = (flags & LW_BATCH_FAR_FP64) != 0;
  const int R = nrhs;
  int rc;
#define TRY(x)                 \
  do {                         \
    rc = (x);                  \
    if (rc) {                  \
      release_batch(bt);       \
      delete ob;               \
      return rc;               \
    }                          \
  } while (0)
  hipStream_t st = h->stream;
  TRY(batch_operator(bt));
  if (flags & LW_BATCH_FP64_RESIDUAL) {
    // the FP64-far-field operator that judges the FP32 iteration
    bt->ref = new Batch();
    bt->ref->h = h;
    bt->ref->R = R;
    bt->ref->fp64 = true;
    TRY(batch_operator(bt->ref));
    TRY(bmalloc(bt, &bt->verified, (size_t)R));
    TRY(bmalloc(bt, &bt->tr_mode, (size_t)R));
    TRY(bmalloc(bt, &bt->tr_restart, (size_t)R));
  }
  const size_t nv = (size_t)h->ne * R;
  TRY(bmalloc(bt, &bt->X, nv));
  TRY(bmalloc(bt, &bt->Y, nv));
  for (auto& w : bt->w) TRY(bmalloc(bt, &w, nv));
  TRY(bmalloc(bt, &bt->part, (size_t)kRedBlocks * R * 4));
  TRY(bmalloc(bt, &bt->sc, (size_t)8 * R));
  TRY(bmalloc(bt, &bt->active, (size_t)R));
  TRY(bmalloc(bt, &bt->first, (size_t)R));
  TRY(bmalloc(bt, &bt->stv, (size_t)R));
  TRY(bmalloc(bt, &bt->nbv, (size_t)R));
  TRY(bmalloc(bt, &bt->rrv, (size_t)R));
#undef TRY
  if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess) {
    set_error("lw_batch_create: conversion kernels failed");
    release_batch(bt);
    delete ob;
    return LW_ERR_HIP;
  }
  *out = ob;
  return LW_OK;
}

int lw_batch_destroy(lw_batch* b) {
  if (b) {
    lw::release_batch(&b->b);
    delete b;
  }
  return LW_OK;
}

int lw_batch_set_replacement(lw_batch* ob, int period) {
  using namespace lw;
  if (!ob || period < 0) return LW_ERR_ARG;
  if (!ob->b.ref) {
    set_error("lw_batch_set_replacement: the batch was not created with LW_BATCH_FP64_RESIDUAL");
    return LW_ERR_STATE;
  }
  ob->b.ru_period = period;
  return LW_OK;
}

long long lw_batch_true_residual_passes(lw_batch* ob) { return ob ? ob->b.tr_passes : LW_ERR_ARG; }

int lw_batch_nrhs(lw_batch* b) { return b ? b->b.R : LW_ERR_ARG; }

int lw_batch_matvec(lw_batch* ob, const double* X, double* Y) {
  if (!ob || !X || !Y) return LW_ERR_ARG;
  lw::Batch* bt = &ob->b;
  lw::Handle* h = bt->h;
  const size_t bytes = sizeof(double2) * (size_t)h->ne * bt->R;
  LW_HIP(hipMemcpyAsync(bt->X, X, bytes, hipMemcpyHostToDevice, h->stream));
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
  LW_HIP(hipMemcpyAsync(bt->X, Yv, bytes, hipMemcpyHostToDevice, h->stream));
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
  const int R = bt->R;
  const size_t bytes = sizeof(double2) * (size_t)h->ne * R;
  hipStream_t st = h->stream;
  // on the handle's stream: it does not synchronise with the null stream
  LW_HIP(hipMemcpyAsync(bt->w[0], Bv, bytes, hipMemcpyHostToDevice, h->stream));
  if (guess)
    LW_HIP(hipMemcpyAsync(bt->X, guess, bytes, hipMemcpyHostToDevice, h->stream));
  else
    LW_HIP(hipMemsetAsync(bt->X, 0, bytes, h->stream));
  int rc = batch_cgs_run(bt, errmax, use_prec != 0, maxit);
  if (rc) return rc;
  std::vector<int4> sv(R);
  std::vector<double> rr(R);
  LW_HIP(hipMemcpyAsync(Xv, bt->X, bytes, hipMemcpyDeviceToHost, st));
  LW_HIP(hipMemcpyAsync(sv.data(), bt->stv, sizeof(int4) * R, hipMemcpyDeviceToHost, st));
  LW_HIP(hipMemcpyAsync(rr.data(), bt->rrv, sizeof(double) * R, hipMemcpyDeviceToHost, st));
  LW_HIP(hipStreamSynchronize(st));
  for (int c = 0; c < R; ++c) {
    if (iters) iters[c] = sv[c].w;
    if (relres) relres[c] = rr[c];
    if (converged) converged[c] = sv[c].z;
  }
  return LW_OK;
}

int lw_batch_cgs_start(lw_batch* ob, const double* Bv, const double* guess, double errmax, int use_prec,
                       int maxit) {
  using namespace lw;
  if (!ob || !Bv || maxit < 1) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  Handle* h = bt->h;
  if (use_prec && !h->have_ilu) {
    set_error("lw_batch_cgs_start: preconditioner requested but no ILU0 factor");
    return LW_ERR_STATE;
  }
  const size_t bytes = sizeof(double2) * (size_t)h->ne * bt->R;
  // on the handle's stream: it does not synchronise with the null stream
  LW_HIP(hipMemcpyAsync(bt->w[0], Bv, bytes, hipMemcpyHostToDevice, h->stream));
  if (guess)
    LW_HIP(hipMemcpyAsync(bt->X, guess, bytes, hipMemcpyHostToDevice, h->stream));
  else
    LW_HIP(hipMemsetAsync(bt->X, 0, bytes, h->stream));
  int rc = batch_cgs_start(bt, errmax, use_prec != 0, maxit);
  bt->next_report = progress_every();
  return rc;
}

int lw_batch_cgs_advance(lw_batch* ob, int iterations, int* state, double* relres) {
  using namespace lw;
  if (!ob || iterations < 0) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  std::vector<int4> sv;
  std::vector<double> rr;
  int rc = batch_cgs_advance(bt, iterations, sv, rr);
  if (rc) return rc;
  if (iterations == 0) {
    sv.resize(bt->R);
    rr.resize(bt->R);
    hipStream_t st = bt->h->stream;
    LW_HIP(hipMemcpyAsync(sv.data(), bt->stv, sizeof(int4) * bt->R, hipMemcpyDeviceToHost, st));
    LW_HIP(hipMemcpyAsync(rr.data(), bt->rrv, sizeof(double) * bt->R, hipMemcpyDeviceToHost, st));
    LW_HIP(hipStreamSynchronize(st));
  }
  for (int c = 0; c < bt->R; ++c) {
    if (state) {
      state[4 * c] = sv[c].x;
      state[4 * c + 1] = sv[c].y;
      state[4 * c + 2] = sv[c].z;
      state[4 * c + 3] = sv[c].w;
    }
    if (relres) relres[c] = rr[c];
  }
  return LW_OK;
}

int lw_batch_column(lw_batch* ob, int col, double* x) {
  using namespace lw;
  if (!ob || !x || col < 0 || col >= ob->b.R) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  Handle* h = bt->h;
  hipLaunchKernelGGL(k_b_col_get, dim3(nblk(h->ne, 256)), dim3(256), 0, h->stream, bt->R, h->ne, col, bt->X,
                     h->d_x);
  LW_HIP(hipGetLastError());
  LW_HIP(hipMemcpyAsync(x, h->d_x, sizeof(double2) * h->ne, hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));
  return LW_OK;
}

int lw_batch_restart_column(lw_batch* ob, int col, const double* b, const double* guess) {
  using namespace lw;
  if (!ob || !b || col < 0 || col >= ob->b.R) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  Handle* h = bt->h;
  if (!bt->cur_graph) {
    set_error("lw_batch_restart_column: no solve started");
    return LW_ERR_STATE;
  }
  const size_t bytes = sizeof(double2) * h->ne;
  hipStream_t st = h->stream;
  // the handle's single-RHS work vectors: x (guess), y (Z x), w[0] (b)
  LW_HIP(hipMemcpyAsync(h->d_w[0], b, bytes, hipMemcpyHostToDevice, st));
  if (guess)
    LW_HIP(hipMemcpyAsync(h->d_x, guess, bytes, hipMemcpyHostToDevice, st));
  else
    LW_HIP(hipMemsetAsync(h->d_x, 0, bytes, st));
  int rc = run_matvec(h, h->d_x, h->d_y, nullptr, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_col_set, dim3(nblk(h->ne, 256)), dim3(256), 0, st, bt->R, h->ne, col, h->d_w[0], h->d_x,
                     h->d_y, bt->w[0], bt->X, bt->w[1], bt->w[2]);
  hipLaunchKernelGGL(k_b_col_init, dim3(1), dim3(256), 0, st, bt->R, h->ne, col, bt->errmax, h->d_w[0], h->d_y,
                     bt->nbv, bt->rrv, bt->stv, bt->sc);
  if (bt->ref) LW_HIP(hipMemsetAsync(bt->verified + col, 0, sizeof(int), st));
  LW_HIP(hipGetLastError());
  LW_HIP(hipStreamSynchronize(st));
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
  LW_HIP(hipMemcpyAsync(d_d, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice, h->stream));
  LW_HIP(hipMemcpyAsync(d_p, pols, sizeof(double2) * 3 * R, hipMemcpyHostToDevice, h->stream));
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
  LW_HIP(hipMemcpyAsync(d_ro, ro.data(), sizeof(double) * ro.size(), hipMemcpyHostToDevice, h->stream));
  LW_HIP(hipMemcpyAsync(bt->Y, cur, bytes, hipMemcpyHostToDevice, h->stream));
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
  LW_HIP(hipMemcpyAsync(bt->X, xin.data(), sizeof(double2) * nv, hipMemcpyHostToDevice, h->stream));
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

int lw_batch_set_tiles(lw_batch* ob, const int* config) {
  using namespace lw;
  if (!ob || !config) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  const int L = bt->h->L;
  int up[16] = {}, fv[16] = {}, fs[16] = {}, fp[16] = {};
  for (int n = 3; n <= L; ++n) {
    const int tb = config[3 * n], sc = config[3 * n + 1] % 10000, ns = config[3 * n + 2];
    const int sep = config[3 * n + 1] / 10000;
    int v = -1;
    for (int i = 0; i < kNFar; ++i)
      if (sc == kFarVariants[i][3] * 1000 + kFarVariants[i][2] * 100 + kFarVariants[i][0] * 10 + kFarVariants[i][1])
        v = i;
    const bool tb_ok = n < 4 ? tb == 0 : (tb == 4 || tb == 8 || tb == 16);
    if (!tb_ok || v < 0 || !(ns == 1 || ns == 2 || ns == 4 || ns == 8) || sep < 0 || sep > (n > 3 ? 1 : 0)) {
      set_error("lw_batch_set_tiles: level " + std::to_string(n) +
                ": theta block 4, 8 or 16 (0 at level 3); a far shape of the list; ranges 1, 2, 4 or 8");
      return LW_ERR_ARG;
    }
    up[n] = n >= 4 ? tb : kThBlock, fv[n] = v, fs[n] = ns, fp[n] = sep;
  }
  LW_HIP(hipStreamSynchronize(bt->h->stream));
  for (auto& g : bt->iter_graph)
    if (g) {
      LW_HIP(hipGraphExecDestroy(g));
      g = nullptr;
    }
  bt->cur_graph = nullptr;
  for (int n = 3; n <= L; ++n)
    bt->up_tb[n] = up[n], bt->far_var[n] = fv[n], bt->far_split[n] = fs[n], bt->far_sep[n] = fp[n];
  return LW_OK;
}

int lw_batch_tune(lw_batch* ob, int reps, double* ms, int* config) {
  using namespace lw;
  if (!ob || reps < 1) return LW_ERR_ARG;
  Batch* bt = &ob->b;
  Handle* h = bt->h;
  hipStream_t st = h->stream;
  const int L = h->L;
  // a solve in progress would continue with other kernels in its captured
  // graph: the graphs are recaptured by the next lw_batch_cgs / _start
  LW_HIP(hipStreamSynchronize(st));
  for (auto& g : bt->iter_graph)
    if (g) {
      LW_HIP(hipGraphExecDestroy(g));
      g = nullptr;
    }
  bt->cur_graph = nullptr;
  double before = 0.0, after = 0.0;
  int rc = lw_bench_batch(ob, reps, &before, nullptr);  // also fills X and every level's fields
  if (rc) return rc;
  hipEvent_t e0, e1;
  LW_HIP(hipEventCreate(&e0));
  LW_HIP(hipEventCreate(&e1));
  // mean time of `reps` launches of f, after one untimed launch
  auto timed = [&](auto f, double* t) -> int {
    f();
    LW_HIP(hipEventRecord(e0, st));
    for (int k = 0; k < reps; ++k) f();
    LW_HIP(hipEventRecord(e1, st));
    LW_HIP(hipEventSynchronize(e1));
    LW_HIP(hipGetLastError());
    float v = 0.f;
    LW_HIP(hipEventElapsedTime(&v, e0, e1));
    *t = (double)v / reps;
    return LW_OK;
  };
  // every variant computes the same values: the choice changes only the time
  for (int n = L; n >= 4 && !rc; --n) {
    double best = 1e30;
    for (int tb : kUpBlocks) {
      double t = 0.0;
      rc = timed([&] { bt->fp64 ? launch_up<double>(bt, n, tb) : launch_up<float>(bt, n, tb); }, &t);
      if (rc) break;
      if (t < best) best = t, bt->up_tb[n] = tb;
    }
  }
  for (int n = 3; n <= L && !rc; ++n) {
    double best = 1e30;
    for (int sep = 0; sep < (n > 3 ? 2 : 1) && !rc; ++sep)
      for (int v = 0; v < kNFar && !rc; ++v)
        for (int ns : {1, 2, 8}) {
          if (ns > 1 && bt->lv[n].nP < 8 * ns) continue;  // ranges of a few parents only
          double t = 0.0;
          rc = timed([&] { bt->fp64 ? launch_far<double>(bt, n, v, ns, sep) : launch_far<float>(bt, n, v, ns, sep); },
                     &t);
          if (rc) break;
          if (t < best) best = t, bt->far_var[n] = v, bt->far_split[n] = ns, bt->far_sep[n] = sep;
        }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if ((rc = lw_bench_batch(ob, reps, &after, nullptr))) return rc;
  if (ms) ms[0] = before, ms[1] = after;
  if (config)
    for (int n = 0; n <= L; ++n) {
      const bool up = n >= 4, far = n >= 3;
      config[3 * n] = up ? bt->up_tb[n] : 0;
      const int* fv = kFarVariants[bt->far_var[n]];
      config[3 * n + 1] = far ? bt->far_sep[n] * 10000 + fv[3] * 1000 + fv[2] * 100 + fv[0] * 10 + fv[1] : 0;
      config[3 * n + 2] = far ? bt->far_split[n] : 0;
    }
  return LW_OK;
}

}  // extern "C"
