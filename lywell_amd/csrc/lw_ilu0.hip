// ILU(0) factorisation of the near-field matrix on the GPU: ilu0C
// (modIterativo.f90:1805-1976) as crear_precond calls it (main.f90:884-903),
// with the rows level-scheduled.
//
// The factor is the reference's MSR layout (1-based, a pad element first):
// alu(ii) holds the inverted diagonal of row ii, alu(jlu(ii) .. ju(ii)-1) its
// L part (ascending columns < ii) and alu(ju(ii) .. jlu(ii+1)-1) its U part.
// Row ii is eliminated against the rows jrow of its L part in ascending
// order, exactly as the reference does, so it can start once those rows are
// final: the rows of one level of the lower-triangular dependency graph are
// independent.  One launch per level spreads its rows over workgroups of four
// waves; each wave factors one row at a time with the row staged in LDS: per elimination step the wave's lanes
// take consecutive runs of jrow's U part, find their columns in row ii by
// interleaved binary searches over the row's sorted columns, and update them.  Every entry of row
// ii is updated in the same order and with the same arithmetic as the
// sequential reference (products in the storage precision, updates in
// double, contraction off), so the factor is bit-identical to it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <string>
#include <vector>

#include "lw_internal.h"

namespace lw {
namespace {

constexpr int kWaveSize = 64;
constexpr int kIluWaves = 4;  // waves (rows in flight) per workgroup
constexpr int kIluChunk = 4;  // consecutive U-part entries per lane and block

// Orders the wave's LDS accesses: every lane's writes are complete before any
// lane reads on.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The same for LDS only, leaving vector-memory loads in flight: the wave
// waits for its LDS operations (lgkmcnt 0; vmcnt and expcnt at their maxima)
// and the compiler moves no memory access across.
__device__ __forceinline__ void wave_lds_order() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// CSR (natural order, 0-based) values into the MSR positions: one wave per
// row; the diagonal goes to alu(i), the other entries keep their CSR order.
template <typename T2>
__global__ void k_ilu0_scatter(int n, const int* __restrict__ ia, const int* __restrict__ ja,
                               const T2* __restrict__ z, const int* __restrict__ jlu, T2* __restrict__ alu) {
  const int i = blockIdx.x * (blockDim.x / kWaveSize) + threadIdx.x / kWaveSize;
  const int lane = threadIdx.x % kWaveSize;
  if (i >= n) return;
  const int r0 = ia[i], r1 = ia[i + 1], p0 = jlu[i + 1];
  // CSR position of the diagonal (every row has one; checked on the host)
  int jd = INT_MAX;
  for (int j = r0 + lane; j < r1; j += kWaveSize)
    if (ja[j] == i) jd = j;
  for (int off = kWaveSize / 2; off > 0; off /= 2) jd = min(jd, __shfl_xor(jd, off));
  for (int j = r0 + lane; j < r1; j += kWaveSize) {
    if (j == jd)
      alu[i + 1] = z[j];
    else
      alu[p0 + (j - r0) - (j > jd ? 1 : 0)] = z[j];
  }
}

template <typename S>
struct IluCx;
template <>
struct IluCx<float> {
  typedef float2 T;
};
template <>
struct IluCx<double> {
  typedef double2 T;
};

// Levels l0 .. l1-1; the rows of a level are dealt to the waves of all
// workgroups (a launch per level, so the rows of one level run on as many
// CUs as it has rows; the kernel boundary orders the levels).
template <typename S>
__global__ void __launch_bounds__(kIluWaves * kWaveSize) k_ilu0_factor(int l0, int l1, const int* __restrict__ lptr,
                                                      const int* __restrict__ lrows, const int* __restrict__ jlu,
                                                      const int* __restrict__ ju, typename IluCx<S>::T* alu,
                                                      int maxrow, int* __restrict__ piv) {
#pragma clang fp contract(off)
  typedef typename IluCx<S>::T T2;
  extern __shared__ char smem[];
  const int nw = blockDim.x / kWaveSize;
  const int wave = threadIdx.x / kWaveSize, lane = threadIdx.x % kWaveSize;
  T2* dslot = reinterpret_cast<T2*>(smem) + wave;                        // the row's diagonal
  T2* vals = reinterpret_cast<T2*>(smem) + nw + (size_t)wave * maxrow;  // its other entries
  int* cols = reinterpret_cast<int*>(reinterpret_cast<T2*>(smem) + nw + (size_t)nw * maxrow) +
              (size_t)wave * maxrow;
  for (int l = l0; l < l1; ++l) {
    for (int r = lptr[l] + (int)blockIdx.x * nw + wave; r < lptr[l + 1]; r += (int)gridDim.x * nw) {
      const int ii = lrows[r] + 1;
      const int p0 = jlu[ii], len = jlu[ii + 1] - p0, pu = ju[ii];
      for (int e = lane; e < len; e += kWaveSize) {
        vals[e] = alu[p0 + e];
        cols[e] = jlu[p0 + e];
      }
      if (lane == 0) *dslot = alu[ii];
      wave_lds_sync();
      // Per elimination step the lanes need jrow's U-part bounds, its
      // inverted diagonal and their run of its U part: all final, so they are
      // fetched ahead (bounds three steps, runs two steps), with clamped
      // addresses so that the loads are unconditional and stay in flight
      // while the current step is applied.  Four stages rotate through a
      // step unrolled four times: copying a stage whose loads are in flight
      // would wait for them.
      struct Stage {
        int q0, q1;
        T2 b;
        int cc[kIluChunk];
        T2 uu[kIluChunk];
      };
      auto meta = [&](int j, Stage& st) {
        const int jr = cols[min(j, pu - 1) - p0];
        st.q0 = ju[jr];
        st.q1 = jlu[jr + 1];
        st.b = alu[jr];
      };
      auto load = [&](int b0, int q1, int(&cc)[kIluChunk], T2(&uu)[kIluChunk]) {
#pragma unroll
        for (int u = 0; u < kIluChunk; ++u) {
          const int jj = b0 + lane * kIluChunk + u;
          const int jc = min(jj, max(q1 - 1, 0));
          const int c = jlu[jc];
          uu[u] = alu[jc];
          cc[u] = jj < q1 ? c : INT_MAX;
        }
      };
      // row ii's entries at the lanes' columns of one run of jrow's U part,
      // less tl times those U entries
      auto eliminate = [&](int lo0, double tr, double ti, const int(&cc)[kIluChunk], const T2(&uu)[kIluChunk]) {
        if (cc[0] == INT_MAX) return;
        // position of each of the lane's columns in row ii: interleaved
        // branch-free lower-bound searches over the row's columns above
        // jrow (a uniform number of halvings, so no lane waits on
        // another's longer walk)
        const int n0 = len - lo0;
        int base[kIluChunk];
#pragma unroll
        for (int u = 0; u < kIluChunk; ++u) base[u] = lo0;
        for (int nn = n0; nn > 1;) {
          const int half = nn >> 1;
#pragma unroll
          for (int u = 0; u < kIluChunk; ++u) base[u] = cols[base[u] + half] < cc[u] ? base[u] + half : base[u];
          nn -= half;
        }
#pragma unroll
        for (int u = 0; u < kIluChunk; ++u) {
          const int c = cc[u];
          T2* w = nullptr;
          if (c == ii) {
            w = dslot;
          } else if (c != INT_MAX && n0 > 0) {
            const int pos = base[u] + (cols[base[u]] < c ? 1 : 0);
            if (pos < len && cols[pos] == c) w = vals + pos;
          }
          if (w) {
            const T2 wv = *w;
            T2 o;
            o.x = (S)((double)wv.x - (tr * (double)uu[u].x - ti * (double)uu[u].y));
            o.y = (S)((double)wv.y - (tr * (double)uu[u].y + ti * (double)uu[u].x));
            *w = o;
          }
        }
      };
      // step j: A holds its stage; C's runs and D's bounds are fetched
      auto step = [&](int j, const Stage& A, Stage& C, Stage& D) __attribute__((always_inline)) {
        load(C.q0, C.q1, C.cc, C.uu);
        meta(j + 3, D);
        // tl = alu(j) * alu(jrow), in the storage precision
        const T2 a = vals[j - p0];
        T2 t;
        t.x = a.x * A.b.x - a.y * A.b.y;
        t.y = a.x * A.b.y + a.y * A.b.x;
        wave_lds_order();
        if (lane == 0) vals[j - p0] = t;
        const double tr = (double)t.x, ti = (double)t.y;
        // row ii's columns above jrow start after position j - p0
        const int lo0 = j - p0 + 1;
        if (A.q0 < A.q1) eliminate(lo0, tr, ti, A.cc, A.uu);
        // rows with more than 64*kIluChunk U entries
        for (int b0 = A.q0 + kWaveSize * kIluChunk; b0 < A.q1; b0 += kWaveSize * kIluChunk) {
          int cl[kIluChunk];
          T2 ul[kIluChunk];
          load(b0, A.q1, cl, ul);
          eliminate(lo0, tr, ti, cl, ul);
        }
        wave_lds_order();
      };
      Stage s0, s1, s2, s3;
      if (p0 < pu) {
        meta(p0, s0);
        load(s0.q0, s0.q1, s0.cc, s0.uu);
        meta(p0 + 1, s1);
        load(s1.q0, s1.q1, s1.cc, s1.uu);
        meta(p0 + 2, s2);
      }
      for (int j = p0; j < pu; j += 4) {
        step(j, s0, s2, s3);
        if (j + 1 >= pu) break;
        step(j + 1, s1, s3, s0);
        if (j + 2 >= pu) break;
        step(j + 2, s2, s0, s1);
        if (j + 3 >= pu) break;
        step(j + 3, s3, s1, s2);
      }
      if (lane == 0) {
        // 1/d with Smith's scaling, in double
        const T2 d = *dslot;
        const double dr = (double)d.x, di = (double)d.y;
        if (dr == 0.0 && di == 0.0) atomicMin(piv, ii);
        double qr, qi;
        if (fabs(dr) >= fabs(di)) {
          const double q = di / dr, den = dr + di * q;
          qr = 1.0 / den;
          qi = -q / den;
        } else {
          const double q = dr / di, den = dr * q + di;
          qr = q / den;
          qi = -1.0 / den;
        }
        T2 inv;
        inv.x = (S)qr;
        inv.y = (S)qi;
        alu[ii] = inv;
      }
      for (int e = lane; e < len; e += kWaveSize) alu[p0 + e] = vals[e];
      wave_lds_sync();
    }
    __syncthreads();  // the level's rows are final before the next level reads them
  }
}

}  // namespace

// MSR structure of the factor (the index arrays ilu0C builds, without values):
// jlu(1) = n+2, row ii's off-diagonal entries at jlu(ii) .. jlu(ii+1)-1 in CSR
// order, ju(ii) = the first entry with a column above ii.  Requires every row
// to hold its diagonal and ascending columns.
int ilu0_structure(int n, const std::vector<int>& ia, const std::vector<int>& ja, std::vector<int>& jlu,
                   std::vector<int>& ju, int& maxrow) {
  const long long nnz = ia[n];
  jlu.assign((size_t)nnz + 2, 0);
  ju.assign(n, 0);
  maxrow = 0;
  int p = n + 2;
  jlu[1] = p;
  for (int i = 0; i < n; ++i) {
    bool diag = false;
    for (int j = ia[i]; j < ia[i + 1]; ++j) {
      if (j > ia[i] && ja[j] <= ja[j - 1]) {
        set_error("ilu0: the columns of row " + std::to_string(i + 1) + " are not ascending");
        return LW_ERR_STATE;
      }
      if (ja[j] == i) {
        diag = true;
        ju[i] = p;
      } else {
        jlu[p++] = ja[j] + 1;
      }
    }
    if (!diag) {
      set_error("ilu0: row " + std::to_string(i + 1) + " has no diagonal entry");
      return LW_ERR_STATE;
    }
    jlu[i + 2] = p;
    maxrow = std::max(maxrow, ia[i + 1] - ia[i] - 1);
  }
  return LW_OK;
}

// The factor of the near field on the device into h->d_alu (structure, levels
// and device index arrays from upload_factor's layout).  *piv = 0 or the
// first row with a zero pivot.
int ilu0_device(Handle* h, int* piv) {
  const int n = h->ne;
  int maxrow = 0;
  if (int rc = ilu0_structure(n, h->plan.ia, h->plan.ja, h->h_jlu, h->h_ju, maxrow)) return rc;
  if (int rc = upload_structure(h)) return rc;
  const size_t na = (size_t)h->nnz + 2;
  const size_t esz = h->fp64_near ? sizeof(double2) : sizeof(float2);
  // waves that fit the device's LDS per workgroup, each with one row
  int dev = 0, lds = 0, lds_optin = 0;
  LW_HIP(hipGetDevice(&dev));
  LW_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess)
    lds = std::max(lds, lds_optin);
  const size_t per_wave = (size_t)std::max(maxrow, 1) * (esz + sizeof(int)) + esz;
  const int nw = (int)std::min<size_t>(kIluWaves, lds / per_wave);
  if (nw < 1) {
    set_error("ilu0: a near-field row of " + std::to_string(maxrow) + " entries does not fit in LDS");
    return LW_ERR_ARG;
  }
  const size_t shm = (size_t)nw * per_wave;
  if (h->fp64_near)
    LW_HIP(hipFuncSetAttribute((const void*)k_ilu0_factor<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shm));
  else
    LW_HIP(hipFuncSetAttribute((const void*)k_ilu0_factor<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)shm));
  if (progress_every() > 0)
    fprintf(stderr, "[lw_ilu0] %d levels, %d waves, %zu bytes of LDS\n", h->nlev_l, nw, shm);
  int* d_piv = nullptr;
  LW_HIP(hipMalloc(&d_piv, sizeof(int)));
  const int none = INT_MAX;
  hipStream_t st = h->stream;
  int rc = LW_OK;
  auto fail = [&](hipError_t e) {
    set_error(std::string("ilu0: ") + hipGetErrorString(e));
    rc = LW_ERR_HIP;
  };
  hipError_t e = hipMemcpyAsync(d_piv, &none, sizeof(int), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(h->d_alu, 0, esz * na, st);
  if (e == hipSuccess) {
    const int rows_per_block = 4;
    if (h->fp64_near) {
      hipLaunchKernelGGL(k_ilu0_scatter<double2>, dim3((n + rows_per_block - 1) / rows_per_block),
                         dim3(rows_per_block * kWaveSize), 0, st, n, h->d_ia, h->d_ja, (const double2*)h->d_z,
                         h->d_jlu, (double2*)h->d_alu);
      for (int l = 0; l < h->nlev_l; ++l) {
        const int rows = h->h_lptr[l + 1] - h->h_lptr[l];
        hipLaunchKernelGGL(k_ilu0_factor<double>, dim3((rows + nw - 1) / nw), dim3(nw * kWaveSize), shm, st, l,
                           l + 1, h->d_lptr, h->d_lrows, h->d_jlu, h->d_ju, (double2*)h->d_alu, maxrow, d_piv);
      }
    } else {
      hipLaunchKernelGGL(k_ilu0_scatter<float2>, dim3((n + rows_per_block - 1) / rows_per_block),
                         dim3(rows_per_block * kWaveSize), 0, st, n, h->d_ia, h->d_ja, (const float2*)h->d_z,
                         h->d_jlu, (float2*)h->d_alu);
      for (int l = 0; l < h->nlev_l; ++l) {
        const int rows = h->h_lptr[l + 1] - h->h_lptr[l];
        hipLaunchKernelGGL(k_ilu0_factor<float>, dim3((rows + nw - 1) / nw), dim3(nw * kWaveSize), shm, st, l,
                           l + 1, h->d_lptr, h->d_lrows, h->d_jlu, h->d_ju, (float2*)h->d_alu, maxrow, d_piv);
      }
    }
    e = hipGetLastError();
  }
  int p = none;
  if (e == hipSuccess) e = hipMemcpyAsync(&p, d_piv, sizeof(int), hipMemcpyDeviceToHost, st);
  // the host copy of the factor (lw_ilu0_get and the Fortran module arrays)
  if (e == hipSuccess) {
    if (h->fp64_near) {
      h->h_alu64.assign(2 * na, 0.0);
      e = hipMemcpyAsync(h->h_alu64.data(), h->d_alu, esz * na, hipMemcpyDeviceToHost, st);
    } else {
      h->h_alu32.assign(2 * na, 0.0f);
      e = hipMemcpyAsync(h->h_alu32.data(), h->d_alu, esz * na, hipMemcpyDeviceToHost, st);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) fail(e);
  (void)hipFree(d_piv);
  if (rc) return rc;
  *piv = p == none ? 0 : p;
  return LW_OK;
}

}  // namespace lw

extern "C" {

// ilu0C on any CSR matrix (modIterativo.f90:1805-1976): the same device
// factorisation as lw_ilu0, on a transient operator-less context.
int lw_ilu0_csr(int n, const int* ia, const int* ja, const void* a, int flags, void* alu, int* jlu, int* ju,
                int* ierr) {
  using namespace lw;
  if (n < 1 || !ia || !ja || !a || !alu || !jlu || !ju || (flags & ~LW_NEAR_FP64)) {
    set_error("lw_ilu0_csr: missing array, n < 1 or unknown flags");
    return LW_ERR_ARG;
  }
  if (ierr) *ierr = 0;
  if (ia[0] != 1 || ia[n] < 1) {
    set_error("lw_ilu0_csr: ia must be 1-based with ia(1) = 1");
    return LW_ERR_ARG;
  }
  const long long nnz = (long long)ia[n] - 1;
  for (int i = 0; i < n; ++i)
    if (ia[i + 1] < ia[i]) {
      set_error("lw_ilu0_csr: ia is not monotone");
      return LW_ERR_ARG;
    }
  for (long long k = 0; k < nnz; ++k)
    if (ja[k] < 1 || ja[k] > n) {
      set_error("lw_ilu0_csr: column index out of range");
      return LW_ERR_ARG;
    }
  if (lw_device_count() < 1) {
    set_error("lw_ilu0_csr: no HIP device visible");
    return LW_ERR_NODEV;
  }
  Handle h;
  h.ne = n;
  h.nnz = nnz;
  h.fp64_near = (flags & LW_NEAR_FP64) != 0;
  h.plan.ia.resize((size_t)n + 1);
  h.plan.ja.resize((size_t)nnz);
  for (int i = 0; i <= n; ++i) h.plan.ia[i] = ia[i] - 1;
  for (long long k = 0; k < nnz; ++k) h.plan.ja[k] = ja[k] - 1;
  const size_t esz = h.fp64_near ? sizeof(double2) : sizeof(float2);
  int rc = LW_OK, piv = 0;
  auto cleanup = [&h]() {
    if (h.stream) (void)hipStreamSynchronize(h.stream);
    for (void* p : h.allocs) (void)hipFree(p);
    if (h.stream) (void)hipStreamDestroy(h.stream);
  };
  if (hipGetDevice(&h.device) != hipSuccess || hipStreamCreateWithFlags(&h.stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("lw_ilu0_csr: stream creation failed");
    cleanup();
    return LW_ERR_HIP;
  }
  void* dz = nullptr;
  if (!(rc = upload(&h, &h.d_ia, h.plan.ia)) && !(rc = upload(&h, &h.d_ja, h.plan.ja)) &&
      !(rc = dmalloc(&h, (char**)&dz, (size_t)nnz * esz))) {
    h.d_z = dz;
    if (hipMemcpyAsync(dz, a, (size_t)nnz * esz, hipMemcpyHostToDevice, h.stream) != hipSuccess) {
      set_error("lw_ilu0_csr: upload failed");
      rc = LW_ERR_HIP;
    }
  }
  if (!rc) rc = ilu0_device(&h, &piv);
  cleanup();
  if (rc) return rc;
  // the MSR arrays as ilu0C leaves them (the host copies carry a pad element)
  const size_t na = (size_t)nnz + 1;
  if (h.fp64_near)
    std::copy(h.h_alu64.begin() + 2, h.h_alu64.begin() + 2 + 2 * na, (double*)alu);
  else
    std::copy(h.h_alu32.begin() + 2, h.h_alu32.begin() + 2 + 2 * na, (float*)alu);
  std::copy(h.h_jlu.begin() + 1, h.h_jlu.begin() + 1 + na, jlu);
  std::copy(h.h_ju.begin(), h.h_ju.end(), ju);
  if (piv) {
    if (ierr) *ierr = piv;
    set_error("lw_ilu0_csr: zero pivot at row " + std::to_string(piv));
    return LW_ERR_PIVOT;
  }
  return LW_OK;
}

}  // extern "C"
