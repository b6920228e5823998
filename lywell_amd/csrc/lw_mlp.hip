// Learned initial guess for monostatic sweeps.
//
// A pointwise MLP maps per-unknown features of a new incidence angle (the
// phase-corrected currents of the two nearest solved angles, incidence
// direction and polarisation, leaf-box and body coordinates of the RWG edge)
// to 4 real outputs o that mix those currents:
//   guess_i = (1 + o0 + j o1) xt1_i + (o2 + j o3) (xt1_i - xt2_i)
// where xt_a = x_a * exp(-j k (d - d_a) . e_centro) is the reference's own
// warm start (main.f90:1017-1022, guess = previous current times the phase
// correction).  With o = 0 the guess IS the reference's warm start, and the
// last layer starts at zero, so training starts from it.  No counterpart in
// the reference beyond that warm start.
//
// Network: in -> H -> H -> H -> out, tanh on the hidden layers.  Weights are
// kept as FP32 masters (Adam state in FP32) and every GEMM rounds its two
// operands to bfloat16 (round to nearest even) and accumulates in FP32 on
// MFMA (v_mfma_f32_16x16x16_bf16).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "lw_internal.h"

namespace lw {
namespace {

constexpr int kWaveM = 64;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4m __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite values)
  return (unsigned short)(u >> 16);
}

// Generic GEMM C(i,j) = sum_kk A(i,kk) B(kk,j) with strided operands
// A(i,kk) = A[i*lda_i + kk*lda_k], B(kk,j) = B[kk*ldb_k + j*ldb_j].
// 64x64 tile per workgroup (4 waves of 32x32 = 2x2 MFMA tiles), K chunks
// of 32 staged in LDS as bf16.  blockIdx.z splits the kk range (partial
// results at C + z*I*ldc, summed by k_sum_parts).
// epi: 0 store, 1 tanh(acc + bias[j]), 2 acc + bias[j].
struct Gemm {
  const float* A;
  long long lda_i, lda_k;
  const float* B;
  long long ldb_k, ldb_j;
  float* C;
  long long ldc;
  int I, J, KK, kk_split;
  const float* bias;
  int epi;
};

__global__ void __launch_bounds__(256) k_gemm_bf16(Gemm g) {
  constexpr int TK = 32, PAD = 8;
  __shared__ unsigned short As[64][TK + PAD];
  __shared__ unsigned short Bs[64][TK + PAD];  // [j][kk]
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  const int kz0 = blockIdx.z * g.kk_split, kz1 = min(g.KK, kz0 + g.kk_split);
  const int wave = threadIdx.x / kWaveM, l = threadIdx.x % kWaveM;
  const int wi = wave / 2, wj = wave % 2, li = l % 16, lg = l / 16;
  f32x4m acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4m{0.f, 0.f, 0.f, 0.f};
  for (int kc = kz0; kc < kz1; kc += TK) {
    for (int e = threadIdx.x; e < 64 * TK; e += blockDim.x) {
      int ii, kk;
      if (g.lda_k == 1) {
        ii = e / TK;
        kk = e % TK;
      } else {
        ii = e % 64;
        kk = e / 64;
      }
      const int gi = i0 + ii, gk = kc + kk;
      const float v = (gi < g.I && gk < kz1) ? g.A[gi * g.lda_i + gk * g.lda_k] : 0.f;
      As[ii][kk] = to_bf16(v);
    }
    for (int e = threadIdx.x; e < 64 * TK; e += blockDim.x) {
      int jj, kk;
      if (g.ldb_k == 1) {
        jj = e / TK;
        kk = e % TK;
      } else {
        jj = e % 64;
        kk = e / 64;
      }
      const int gj = j0 + jj, gk = kc + kk;
      const float v = (gj < g.J && gk < kz1) ? g.B[gk * g.ldb_k + gj * g.ldb_j] : 0.f;
      Bs[jj][kk] = to_bf16(v);
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < TK; ks += 16) {
      bf16x4 af[2], bfr[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const unsigned short* p = &As[wi * 32 + a * 16 + li][ks + 4 * lg];
#pragma unroll
        for (int q = 0; q < 4; ++q) af[a][q] = (short)p[q];
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const unsigned short* p = &Bs[wj * 32 + b * 16 + li][ks + 4 * lg];
#pragma unroll
        for (int q = 0; q < 4; ++q) bfr[b][q] = (short)p[q];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  float* C = g.C + (size_t)blockIdx.z * g.I * g.ldc;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi * 32 + a * 16 + 4 * lg + v, j = j0 + wj * 32 + b * 16 + li;
        if (i < g.I && j < g.J) {
          float x = acc[a][b][v];
          if (g.epi == 1) x = tanhf(x + g.bias[j]);
          else if (g.epi == 2) x = x + g.bias[j];
          C[(size_t)i * g.ldc + j] = x;
        }
      }
}

__global__ void k_sum_parts(long long n, int nparts, const float* __restrict__ part, float* __restrict__ out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.f;
  for (int z = 0; z < nparts; ++z) s += part[(size_t)z * n + e];
  out[e] = s;
}

// column sums of a (rows, n) matrix: one thread per column, fixed row order
__global__ void k_colsum(int rows, int n, const float* __restrict__ a, float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s = 0.f;
  for (int i = 0; i < rows; ++i) s += a[(size_t)i * n + j];
  out[j] = s;
}

// dZ = dY * (1 - Y^2) (tanh backward), in place on dY
__global__ void k_tanh_back(long long n, const float* __restrict__ y, float* __restrict__ dy) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) dy[e] *= 1.f - y[e] * y[e];
}

__global__ void k_adam(long long n, float lr, float b1, float b2, float eps, float c1, float c2,
                       const float* __restrict__ g, float* __restrict__ p, float* __restrict__ m,
                       float* __restrict__ v) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.f - b1) * gi;
  const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------------------------ guess head

constexpr int kFeat = 19, kOut = 4;

// Per row (unknown i, column c; row = i*ncol + c): phase-corrected neighbour
// currents and the feature vector: the two normalised neighbour currents, the
// incidence direction and polarisation, the edge centre relative to its leaf
// box (box units) and to the body (half-sizes), and the unit step d - d1 from
// the nearest solved angle (which side the neighbours lie on).
__global__ void k_guess_features(int ne, int ncol, double k0, const double* __restrict__ ec,
                                 const double* __restrict__ boxc, const int* __restrict__ basis_box,
                                 double box_side, const double* __restrict__ body, const double* __restrict__ dir,
                                 const double* __restrict__ pol, const double* __restrict__ dir1,
                                 const double2* __restrict__ x1, const double* __restrict__ dir2,
                                 const double2* __restrict__ x2, float* __restrict__ feat,
                                 float2* __restrict__ xt1, float2* __restrict__ xt2) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ne * ncol) return;
  const int i = (int)(t / ncol), c = (int)(t % ncol);
  const double* e = ec + 3 * (size_t)i;
  const double* d = dir + 3 * c;
  const double* d1 = dir1 + 3 * c;
  const double* d2 = dir2 + 3 * c;
  // exp(-j k (d - d_a) . e)
  double s1, c1, s2, c2;
  sincos(-k0 * ((d[0] - d1[0]) * e[0] + (d[1] - d1[1]) * e[1] + (d[2] - d1[2]) * e[2]), &s1, &c1);
  sincos(-k0 * ((d[0] - d2[0]) * e[0] + (d[1] - d2[1]) * e[1] + (d[2] - d2[2]) * e[2]), &s2, &c2);
  const double2 a1 = x1[t], a2 = x2[t];
  const double2 t1 = make_double2(a1.x * c1 - a1.y * s1, a1.x * s1 + a1.y * c1);
  const double2 t2 = make_double2(a2.x * c2 - a2.y * s2, a2.x * s2 + a2.y * c2);
  xt1[t] = make_float2((float)t1.x, (float)t1.y);
  xt2[t] = make_float2((float)t2.x, (float)t2.y);
  const double sc = sqrt(t1.x * t1.x + t1.y * t1.y + t2.x * t2.x + t2.y * t2.y) + 1e-30;
  float* f = feat + (size_t)t * kFeat;
  f[0] = (float)(t1.x / sc);
  f[1] = (float)(t1.y / sc);
  f[2] = (float)(t2.x / sc);
  f[3] = (float)(t2.y / sc);
  f[4] = (float)d[0];
  f[5] = (float)d[1];
  f[6] = (float)d[2];
  f[7] = (float)pol[3 * c];
  f[8] = (float)pol[3 * c + 1];
  f[9] = (float)pol[3 * c + 2];
  const double* bc = boxc + 3 * (size_t)basis_box[i];
  f[10] = (float)((e[0] - bc[0]) / box_side);
  f[11] = (float)((e[1] - bc[1]) / box_side);
  f[12] = (float)((e[2] - bc[2]) / box_side);
  f[13] = (float)((e[0] - body[0]) / body[3]);
  f[14] = (float)((e[1] - body[1]) / body[3]);
  f[15] = (float)((e[2] - body[2]) / body[3]);
  const double st[3] = {d[0] - d1[0], d[1] - d1[1], d[2] - d1[2]};
  const double sn = sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2]);
  const double si = sn > 0.0 ? 1.0 / sn : 0.0;
  f[16] = (float)(st[0] * si);
  f[17] = (float)(st[1] * si);
  f[18] = (float)(st[2] * si);
}

__global__ void k_guess_apply(long long rows, const float* __restrict__ o, const float2* __restrict__ xt1,
                              const float2* __restrict__ xt2, double2* __restrict__ g) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows) return;
  const float* oo = o + (size_t)t * kOut;
  const double2 a = make_double2(xt1[t].x, xt1[t].y), b = make_double2(xt2[t].x, xt2[t].y);
  const double2 df = make_double2(a.x - b.x, a.y - b.y);
  const double ar = 1.0 + oo[0], ai = oo[1], br = oo[2], bi = oo[3];
  g[t] = make_double2(ar * a.x - ai * a.y + br * df.x - bi * df.y, ar * a.y + ai * a.x + br * df.y + bi * df.x);
}

// loss rows: w |guess - x|^2 ; dO = d(loss)/dO / rows (mean over the batch)
__global__ void k_guess_loss(int rows, const float* __restrict__ o, const float2* __restrict__ xt1,
                             const float2* __restrict__ xt2, const float2* __restrict__ xtrue,
                             const float* __restrict__ w, float* __restrict__ dO, float* __restrict__ lrow) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows) return;
  const float* oo = o + (size_t)t * kOut;
  const float2 a = xt1[t], b = xt2[t], x = xtrue[t];
  const float2 df = make_float2(a.x - b.x, a.y - b.y);
  const float ar = 1.f + oo[0], ai = oo[1], br = oo[2], bi = oo[3];
  const float gr = ar * a.x - ai * a.y + br * df.x - bi * df.y;
  const float gi = ar * a.y + ai * a.x + br * df.y + bi * df.x;
  const float er = gr - x.x, ei = gi - x.y;
  const float s = 2.f * w[t] / (float)rows;
  // d guess / d o: a, j a, df, j df;  dL/do = s Re(conj(e) * dguess)
  float* d = dO + (size_t)t * kOut;
  d[0] = s * (er * a.x + ei * a.y);
  d[1] = s * (-er * a.y + ei * a.x);
  d[2] = s * (er * df.x + ei * df.y);
  d[3] = s * (-er * df.y + ei * df.x);
  lrow[t] = w[t] * (er * er + ei * ei);
}

__global__ void k_gather_rows(int rows, long long nset, unsigned long long seed, unsigned step,
                              const float* __restrict__ feat, const float2* __restrict__ xt1,
                              const float2* __restrict__ xt2, const float2* __restrict__ xtr,
                              const float* __restrict__ w, float* __restrict__ bf, float2* __restrict__ b1,
                              float2* __restrict__ b2, float2* __restrict__ bt, float* __restrict__ bw) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows) return;
  // splitmix64 of (seed, step, t): uniform row of the training set
  unsigned long long z = seed + 0x9E3779B97F4A7C15ULL * ((unsigned long long)step * 1000003ULL + (unsigned)t + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  const long long r = (long long)(z % (unsigned long long)nset);
  for (int f = 0; f < kFeat; ++f) bf[(size_t)t * kFeat + f] = feat[(size_t)r * kFeat + f];
  b1[t] = xt1[r];
  b2[t] = xt2[r];
  bt[t] = xtr[r];
  bw[t] = w[r];
}

__global__ void k_store_true(long long rows, const double2* __restrict__ x, float2* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < rows) out[t] = make_float2((float)x[t].x, (float)x[t].y);
}

int nb256(long long n) { return (int)((n + 255) / 256); }

}  // namespace

struct Mlp {
  int in = 0, hid = 0, depth = 0, out = 0;  // depth = hidden layers
  std::vector<int> kin, nout;               // per layer
  std::vector<long long> woff, boff;        // offsets in the flat parameter vector
  long long np = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  long long adam_t = 0;
  hipStream_t st = nullptr;
  // activations of the last forward (rows x width per layer)
  int cap_rows = 0;
  std::vector<float*> act;  // act[0] input, act[l+1] output of layer l
  float *dA = nullptr, *dB = nullptr, *part = nullptr;
  long long part_cap = 0;
  // training set (device)
  long long nset = 0, set_cap = 0;
  float* s_feat = nullptr;
  float2 *s_x1 = nullptr, *s_x2 = nullptr, *s_xt = nullptr;
  float* s_w = nullptr;
  // minibatch buffers
  int bcap = 0;
  float2 *b_x1 = nullptr, *b_x2 = nullptr, *b_xt = nullptr;
  float *b_w = nullptr, *b_loss = nullptr, *b_lsum = nullptr;
  std::vector<void*> allocs;
};

template <typename T>
static int mmalloc(Mlp* m, T** p, size_t n) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    set_error("lw_mlp: device allocation failed");
    return LW_ERR_NOMEM;
  }
  m->allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return LW_OK;
}

static void mfree(Mlp* m, void* q) {
  if (!q) return;
  auto it = std::find(m->allocs.begin(), m->allocs.end(), q);
  if (it != m->allocs.end()) m->allocs.erase(it);
  (void)hipFree(q);
}

static int launch_gemm(Mlp* m, const Gemm& g0, int splits) {
  Gemm g = g0;
  splits = std::max(1, splits);
  g.kk_split = ((g.KK + splits - 1) / splits + 31) / 32 * 32;
  const int nz = (g.KK + g.kk_split - 1) / g.kk_split;
  dim3 grid((g.I + 63) / 64, (g.J + 63) / 64, std::max(1, nz));
  hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, m->st, g);
  return nz;
}

static int ensure_rows(Mlp* m, int rows) {
  if (rows <= m->cap_rows) return LW_OK;
  for (float* a : m->act) mfree(m, a);
  mfree(m, m->dA);
  mfree(m, m->dB);
  m->act.assign(m->depth + 2, nullptr);
  int rc;
  const int L = m->depth + 1;
  for (int l = 0; l <= L; ++l) {
    const int w = l == 0 ? m->in : m->nout[l - 1];
    if ((rc = mmalloc(m, &m->act[l], (size_t)rows * w))) return rc;
  }
  const int wmax = std::max(std::max(m->in, m->hid), m->out);
  if ((rc = mmalloc(m, &m->dA, (size_t)rows * wmax))) return rc;
  if ((rc = mmalloc(m, &m->dB, (size_t)rows * wmax))) return rc;
  m->cap_rows = rows;
  return LW_OK;
}

// forward over rows already in act[0]
static void mlp_forward(Mlp* m, int rows) {
  const int L = m->depth + 1;
  for (int l = 0; l < L; ++l) {
    Gemm g{};
    g.A = m->act[l];
    g.lda_i = m->kin[l];
    g.lda_k = 1;
    g.B = m->p + m->woff[l];  // W (nout, kin): B(kk, j) = W[j*kin + kk]
    g.ldb_k = 1;
    g.ldb_j = m->kin[l];
    g.C = m->act[l + 1];
    g.ldc = m->nout[l];
    g.I = rows;
    g.J = m->nout[l];
    g.KK = m->kin[l];
    g.bias = m->p + m->boff[l];
    g.epi = l < L - 1 ? 1 : 2;
    launch_gemm(m, g, 1);
  }
}

// backward from dOut (rows, out) in dA: gradients into m->g; returns the
// buffer holding d(input)
static int mlp_backward(Mlp* m, int rows, float** dx_out) {
  const int L = m->depth + 1;
  float* dcur = m->dA;  // d(output of layer l)
  float* dnext = m->dB;
  int rc;
  for (int l = L - 1; l >= 0; --l) {
    const int kin = m->kin[l], no = m->nout[l];
    if (l < L - 1)
      hipLaunchKernelGGL(k_tanh_back, dim3(nb256((long long)rows * no)), dim3(256), 0, m->st, (long long)rows * no,
                         m->act[l + 1], dcur);
    // dW (no, kin) = dZ^T X: C(i=n, j=k) = sum_m dZ[m*no + n] X[m*kin + k]
    const long long wsz = (long long)no * kin;
    const int splits = std::max(1, std::min(64, rows / 1024));
    if ((long long)splits * wsz > m->part_cap) {
      mfree(m, m->part);
      if ((rc = mmalloc(m, &m->part, (size_t)splits * wsz))) return rc;
      m->part_cap = (long long)splits * wsz;
    }
    Gemm g{};
    g.A = dcur;
    g.lda_i = 1;
    g.lda_k = no;
    g.B = m->act[l];
    g.ldb_k = kin;
    g.ldb_j = 1;
    g.C = m->part;
    g.ldc = kin;
    g.I = no;
    g.J = kin;
    g.KK = rows;
    g.epi = 0;
    const int nz = launch_gemm(m, g, splits);
    hipLaunchKernelGGL(k_sum_parts, dim3(nb256(wsz)), dim3(256), 0, m->st, wsz, nz, m->part, m->g + m->woff[l]);
    hipLaunchKernelGGL(k_colsum, dim3((no + 255) / 256), dim3(256), 0, m->st, rows, no, dcur, m->g + m->boff[l]);
    // dX (rows, kin) = dZ W: C(i=m, j=k) = sum_n dZ[m*no + n] W[n*kin + k]
    Gemm d{};
    d.A = dcur;
    d.lda_i = no;
    d.lda_k = 1;
    d.B = m->p + m->woff[l];
    d.ldb_k = kin;
    d.ldb_j = 1;
    d.C = dnext;
    d.ldc = kin;
    d.I = rows;
    d.J = kin;
    d.KK = no;
    d.epi = 0;
    launch_gemm(m, d, 1);
    std::swap(dcur, dnext);
  }
  *dx_out = dcur;
  return LW_OK;
}

}  // namespace lw

struct lw_mlp {
  lw::Mlp m;
};

using lw::Mlp;

extern "C" {

int lw_mlp_create(int in_dim, int hidden, int depth, int out_dim, unsigned long long seed, lw_mlp** out) {
  using namespace lw;
  if (!out || in_dim < 1 || hidden < 1 || depth < 1 || out_dim < 1 || in_dim > 4096 || hidden > 4096 ||
      out_dim > 4096 || depth > 16)
    return LW_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("lw_mlp_create: no HIP device");
    return LW_ERR_NODEV;
  }
  lw_mlp* o = new lw_mlp();
  Mlp* m = &o->m;
  m->in = in_dim;
  m->hid = hidden;
  m->depth = depth;
  m->out = out_dim;
  const int L = depth + 1;
  std::vector<float> host;
  std::mt19937_64 rng(seed);
  for (int l = 0; l < L; ++l) {
    const int kin = l == 0 ? in_dim : hidden, no = l == L - 1 ? out_dim : hidden;
    m->kin.push_back(kin);
    m->nout.push_back(no);
    m->woff.push_back((long long)host.size());
    // Glorot-uniform hidden layers; the output layer starts at zero
    const float a = l == L - 1 ? 0.f : std::sqrt(6.f / (float)(kin + no));
    std::uniform_real_distribution<float> U(-a, a);
    for (long long e = 0; e < (long long)kin * no; ++e) host.push_back(l == L - 1 ? 0.f : U(rng));
    m->boff.push_back((long long)host.size());
    for (int e = 0; e < no; ++e) host.push_back(0.f);
  }
  m->np = (long long)host.size();
  int rc = 0;
  if (hipStreamCreate(&m->st) != hipSuccess) rc = LW_ERR_HIP;
  if (!rc) rc = mmalloc(m, &m->p, m->np);
  if (!rc) rc = mmalloc(m, &m->g, m->np);
  if (!rc) rc = mmalloc(m, &m->m, m->np);
  if (!rc) rc = mmalloc(m, &m->v, m->np);
  if (!rc && hipMemcpy(m->p, host.data(), sizeof(float) * m->np, hipMemcpyHostToDevice) != hipSuccess) rc = LW_ERR_HIP;
  if (!rc && (hipMemset(m->m, 0, sizeof(float) * m->np) != hipSuccess ||
              hipMemset(m->v, 0, sizeof(float) * m->np) != hipSuccess))
    rc = LW_ERR_HIP;
  if (rc) {
    for (void* q : m->allocs) (void)hipFree(q);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete o;
    return rc;
  }
  *out = o;
  return LW_OK;
}

int lw_mlp_destroy(lw_mlp* o) {
  if (!o) return LW_OK;
  Mlp* m = &o->m;
  if (m->st) (void)hipStreamSynchronize(m->st);
  for (void* q : m->allocs) (void)hipFree(q);
  if (m->st) (void)hipStreamDestroy(m->st);
  delete o;
  return LW_OK;
}

long long lw_mlp_nparams(lw_mlp* o) { return o ? o->m.np : (long long)LW_ERR_ARG; }

int lw_mlp_get_params(lw_mlp* o, float* p) {
  if (!o || !p) return LW_ERR_ARG;
  LW_HIP(hipStreamSynchronize(o->m.st));
  LW_HIP(hipMemcpy(p, o->m.p, sizeof(float) * o->m.np, hipMemcpyDeviceToHost));
  return LW_OK;
}

int lw_mlp_set_params(lw_mlp* o, const float* p) {
  if (!o || !p) return LW_ERR_ARG;
  LW_HIP(hipStreamSynchronize(o->m.st));
  LW_HIP(hipMemcpy(o->m.p, p, sizeof(float) * o->m.np, hipMemcpyHostToDevice));
  return LW_OK;
}

int lw_mlp_forward(lw_mlp* o, const float* x, int rows, float* y) {
  using namespace lw;
  if (!o || !x || !y || rows < 1) return LW_ERR_ARG;
  Mlp* m = &o->m;
  int rc = ensure_rows(m, rows);
  if (rc) return rc;
  LW_HIP(hipMemcpyAsync(m->act[0], x, sizeof(float) * (size_t)rows * m->in, hipMemcpyHostToDevice, m->st));
  mlp_forward(m, rows);
  LW_HIP(hipGetLastError());
  LW_HIP(hipMemcpyAsync(y, m->act[m->depth + 1], sizeof(float) * (size_t)rows * m->out, hipMemcpyDeviceToHost,
                        m->st));
  LW_HIP(hipStreamSynchronize(m->st));
  return LW_OK;
}

int lw_mlp_backward(lw_mlp* o, const float* x, int rows, const float* dy, float* grad, float* dx) {
  using namespace lw;
  if (!o || !x || !dy || rows < 1) return LW_ERR_ARG;
  Mlp* m = &o->m;
  int rc = ensure_rows(m, rows);
  if (rc) return rc;
  LW_HIP(hipMemcpyAsync(m->act[0], x, sizeof(float) * (size_t)rows * m->in, hipMemcpyHostToDevice, m->st));
  mlp_forward(m, rows);
  LW_HIP(hipMemcpyAsync(m->dA, dy, sizeof(float) * (size_t)rows * m->out, hipMemcpyHostToDevice, m->st));
  float* dxd = nullptr;
  if ((rc = mlp_backward(m, rows, &dxd))) return rc;
  LW_HIP(hipGetLastError());
  if (grad) LW_HIP(hipMemcpyAsync(grad, m->g, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->st));
  if (dx) LW_HIP(hipMemcpyAsync(dx, dxd, sizeof(float) * (size_t)rows * m->in, hipMemcpyDeviceToHost, m->st));
  LW_HIP(hipStreamSynchronize(m->st));
  return LW_OK;
}

}  // extern "C"

namespace lw {

// Features and phase-corrected neighbour currents of ncol columns on the
// handle's mesh, into device buffers (rows = ne*ncol).
static int guess_inputs(Mlp* m, Handle* h, int ncol, const double* dir, const double* pol, const double* dir1,
                        const double* x1, const double* dir2, const double* x2, float* feat, float2* xt1,
                        float2* xt2) {
  const Plan& P = h->plan;
  const LevelPlan& lf = P.lv[h->L];
  // body centre and half-size of the edge-centre bounding box
  std::vector<double> ec(3 * (size_t)h->ne);
  LW_HIP(hipMemcpy(ec.data(), h->d_ec, sizeof(double) * ec.size(), hipMemcpyDeviceToHost));
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < h->ne; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], ec[3 * (size_t)i + k]);
      hi[k] = std::max(hi[k], ec[3 * (size_t)i + k]);
    }
  const double body[4] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2]),
                          0.5 * std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2])) + 1e-30};
  const size_t rows = (size_t)h->ne * ncol;
  std::vector<double> dirs(9 * (size_t)ncol + 3 * (size_t)ncol);
  std::memcpy(dirs.data(), dir, sizeof(double) * 3 * ncol);
  std::memcpy(dirs.data() + 3 * ncol, dir1, sizeof(double) * 3 * ncol);
  std::memcpy(dirs.data() + 6 * ncol, dir2, sizeof(double) * 3 * ncol);
  std::memcpy(dirs.data() + 9 * ncol, pol, sizeof(double) * 3 * ncol);
  double *d_dirs = nullptr, *d_boxc = nullptr, *d_body = nullptr;
  double2 *d_x1 = nullptr, *d_x2 = nullptr;
  int* d_bbox = nullptr;
  int rc = LW_OK;
  auto cleanup = [&]() {
    (void)hipFree(d_dirs);
    (void)hipFree(d_boxc);
    (void)hipFree(d_body);
    (void)hipFree(d_x1);
    (void)hipFree(d_x2);
    (void)hipFree(d_bbox);
  };
  if (hipMalloc(&d_dirs, sizeof(double) * dirs.size()) != hipSuccess ||
      hipMalloc(&d_boxc, sizeof(double) * lf.centre.size()) != hipSuccess ||
      hipMalloc(&d_body, sizeof(double) * 4) != hipSuccess ||
      hipMalloc(&d_x1, sizeof(double2) * rows) != hipSuccess ||
      hipMalloc(&d_x2, sizeof(double2) * rows) != hipSuccess ||
      hipMalloc(&d_bbox, sizeof(int) * (size_t)h->ne) != hipSuccess) {
    cleanup();
    set_error("lw_guess: device allocation failed");
    return LW_ERR_NOMEM;
  }
  if (hipMemcpy(d_dirs, dirs.data(), sizeof(double) * dirs.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_boxc, lf.centre.data(), sizeof(double) * lf.centre.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_body, body, sizeof(double) * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_x1, x1, sizeof(double2) * rows, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_x2, x2, sizeof(double2) * rows, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_bbox, P.basis_box.data(), sizeof(int) * (size_t)h->ne, hipMemcpyHostToDevice) != hipSuccess)
    rc = LW_ERR_HIP;
  if (!rc) {
    hipLaunchKernelGGL(k_guess_features, dim3(nb256((long long)rows)), dim3(256), 0, m->st, h->ne, ncol,
                       2.0 * Physics::pi / P.phys.lambda, h->d_ec, d_boxc, d_bbox, lf.lado, d_body, d_dirs,
                       d_dirs + 9 * ncol, d_dirs + 3 * ncol, d_x1, d_dirs + 6 * ncol, d_x2, feat, xt1, xt2);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(m->st) != hipSuccess) rc = LW_ERR_HIP;
  }
  cleanup();
  if (rc) set_error("lw_guess: feature kernel failed");
  return rc;
}

}  // namespace lw

extern "C" {

int lw_guess_predict(lw_mlp* o, lw_handle* hh, int ncol, const double* dir, const double* pol, const double* dir1,
                     const double* x1, const double* dir2, const double* x2, double* guess) {
  using namespace lw;
  if (!o || !hh || ncol < 1 || !dir || !pol || !dir1 || !x1 || !dir2 || !x2 || !guess) return LW_ERR_ARG;
  Mlp* m = &o->m;
  Handle* h = &hh->h;
  if (m->in != kFeat || m->out != kOut) {
    set_error("lw_guess_predict: the network must map 16 features to 4 outputs");
    return LW_ERR_ARG;
  }
  const long long rows = (long long)h->ne * ncol;
  if (rows > (1LL << 30)) return LW_ERR_ARG;
  int rc = ensure_rows(m, (int)rows);
  if (rc) return rc;
  float2 *xt1 = nullptr, *xt2 = nullptr;
  double2* g = nullptr;
  if (hipMalloc(&xt1, sizeof(float2) * rows) != hipSuccess || hipMalloc(&xt2, sizeof(float2) * rows) != hipSuccess ||
      hipMalloc(&g, sizeof(double2) * rows) != hipSuccess) {
    (void)hipFree(xt1);
    (void)hipFree(xt2);
    (void)hipFree(g);
    return LW_ERR_NOMEM;
  }
  rc = guess_inputs(m, h, ncol, dir, pol, dir1, x1, dir2, x2, m->act[0], xt1, xt2);
  if (!rc) {
    mlp_forward(m, (int)rows);
    hipLaunchKernelGGL(k_guess_apply, dim3(nb256(rows)), dim3(256), 0, m->st, rows, m->act[m->depth + 1], xt1, xt2,
                       g);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(guess, g, sizeof(double2) * rows, hipMemcpyDeviceToHost, m->st) != hipSuccess ||
        hipStreamSynchronize(m->st) != hipSuccess)
      rc = LW_ERR_HIP;
  }
  (void)hipFree(xt1);
  (void)hipFree(xt2);
  (void)hipFree(g);
  return rc;
}

int lw_guess_add_samples(lw_mlp* o, lw_handle* hh, int ncol, const double* dir, const double* pol,
                         const double* dir1, const double* x1, const double* dir2, const double* x2,
                         const double* xtrue) {
  using namespace lw;
  if (!o || !hh || ncol < 1 || !dir || !pol || !dir1 || !x1 || !dir2 || !x2 || !xtrue) return LW_ERR_ARG;
  Mlp* m = &o->m;
  Handle* h = &hh->h;
  const long long rows = (long long)h->ne * ncol;
  int rc;
  if (m->nset + rows > m->set_cap) {
    const long long cap = std::max(2 * m->set_cap, m->nset + rows);
    float* f = nullptr;
    float2 *a = nullptr, *b = nullptr, *t = nullptr;
    float* w = nullptr;
    if ((rc = mmalloc(m, &f, (size_t)cap * kFeat)) || (rc = mmalloc(m, &a, (size_t)cap)) ||
        (rc = mmalloc(m, &b, (size_t)cap)) || (rc = mmalloc(m, &t, (size_t)cap)) || (rc = mmalloc(m, &w, (size_t)cap)))
      return rc;
    if (m->nset) {
      LW_HIP(hipMemcpy(f, m->s_feat, sizeof(float) * kFeat * m->nset, hipMemcpyDeviceToDevice));
      LW_HIP(hipMemcpy(a, m->s_x1, sizeof(float2) * m->nset, hipMemcpyDeviceToDevice));
      LW_HIP(hipMemcpy(b, m->s_x2, sizeof(float2) * m->nset, hipMemcpyDeviceToDevice));
      LW_HIP(hipMemcpy(t, m->s_xt, sizeof(float2) * m->nset, hipMemcpyDeviceToDevice));
      LW_HIP(hipMemcpy(w, m->s_w, sizeof(float) * m->nset, hipMemcpyDeviceToDevice));
    }
    mfree(m, m->s_feat);
    mfree(m, m->s_x1);
    mfree(m, m->s_x2);
    mfree(m, m->s_xt);
    mfree(m, m->s_w);
    m->s_feat = f;
    m->s_x1 = a;
    m->s_x2 = b;
    m->s_xt = t;
    m->s_w = w;
    m->set_cap = cap;
  }
  const long long at = m->nset;
  rc = guess_inputs(m, h, ncol, dir, pol, dir1, x1, dir2, x2, m->s_feat + at * kFeat, m->s_x1 + at, m->s_x2 + at);
  if (rc) return rc;
  // per-row weight: ne / |x_true(column)|^2, so a column's rows average to
  // its relative squared error
  std::vector<float> w((size_t)rows);
  for (int c = 0; c < ncol; ++c) {
    double s = 0.0;
    for (int i = 0; i < h->ne; ++i) {
      const double* x = xtrue + 2 * ((size_t)i * ncol + c);
      s += x[0] * x[0] + x[1] * x[1];
    }
    const float wc = (float)((double)h->ne / std::max(s, 1e-300));
    for (int i = 0; i < h->ne; ++i) w[(size_t)i * ncol + c] = wc;
  }
  double2* d_xt = nullptr;
  LW_HIP(hipMalloc(&d_xt, sizeof(double2) * rows));
  hipError_t e = hipMemcpy(d_xt, xtrue, sizeof(double2) * rows, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_store_true, dim3(nb256(rows)), dim3(256), 0, m->st, rows, d_xt, m->s_xt + at);
    e = hipStreamSynchronize(m->st);
  }
  (void)hipFree(d_xt);
  LW_HIP(e);
  LW_HIP(hipMemcpy(m->s_w + at, w.data(), sizeof(float) * rows, hipMemcpyHostToDevice));
  m->nset += rows;
  return LW_OK;
}

long long lw_guess_nsamples(lw_mlp* o) { return o ? o->m.nset : (long long)LW_ERR_ARG; }

int lw_guess_train(lw_mlp* o, int steps, int batch_rows, float lr, unsigned long long seed, float* loss) {
  using namespace lw;
  if (!o || steps < 1 || batch_rows < 64 || batch_rows > (1 << 22) || !(lr > 0)) return LW_ERR_ARG;
  Mlp* m = &o->m;
  if (m->in != kFeat || m->out != kOut) return LW_ERR_ARG;
  if (m->nset < 1) {
    set_error("lw_guess_train: no training samples");
    return LW_ERR_STATE;
  }
  int rc = ensure_rows(m, batch_rows);
  if (rc) return rc;
  if (batch_rows > m->bcap) {
    mfree(m, m->b_x1);
    mfree(m, m->b_x2);
    mfree(m, m->b_xt);
    mfree(m, m->b_w);
    mfree(m, m->b_loss);
    mfree(m, m->b_lsum);
    if ((rc = mmalloc(m, &m->b_x1, batch_rows)) || (rc = mmalloc(m, &m->b_x2, batch_rows)) ||
        (rc = mmalloc(m, &m->b_xt, batch_rows)) || (rc = mmalloc(m, &m->b_w, batch_rows)) ||
        (rc = mmalloc(m, &m->b_loss, batch_rows)) || (rc = mmalloc(m, &m->b_lsum, 1)))
      return rc;
    m->bcap = batch_rows;
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  std::vector<float> lsum(1);
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(k_gather_rows, dim3(nb256(batch_rows)), dim3(256), 0, m->st, batch_rows, m->nset, seed,
                       (unsigned)(m->adam_t + 1), m->s_feat, m->s_x1, m->s_x2, m->s_xt, m->s_w, m->act[0], m->b_x1,
                       m->b_x2, m->b_xt, m->b_w);
    mlp_forward(m, batch_rows);
    hipLaunchKernelGGL(k_guess_loss, dim3(nb256(batch_rows)), dim3(256), 0, m->st, batch_rows,
                       m->act[m->depth + 1], m->b_x1, m->b_x2, m->b_xt, m->b_w, m->dA, m->b_loss);
    float* dx = nullptr;
    if ((rc = mlp_backward(m, batch_rows, &dx))) return rc;
    m->adam_t += 1;
    const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(nb256(m->np)), dim3(256), 0, m->st, m->np, lr, b1, b2, eps, c1, c2, m->g, m->p,
                       m->m, m->v);
    if (loss) {
      hipLaunchKernelGGL(k_colsum, dim3(1), dim3(64), 0, m->st, batch_rows, 1, m->b_loss, m->b_lsum);
      LW_HIP(hipMemcpyAsync(lsum.data(), m->b_lsum, sizeof(float), hipMemcpyDeviceToHost, m->st));
      LW_HIP(hipStreamSynchronize(m->st));
      loss[s] = lsum[0] / (float)batch_rows;
    }
  }
  LW_HIP(hipGetLastError());
  LW_HIP(hipStreamSynchronize(m->st));
  return LW_OK;
}

}  // extern "C"
