// lusol0C (modIterativo.f90:504-560) on the device without level barriers.
//
// The forward sweep solves with the unit-diagonal L (x(i) = y(i) - sum
// alu(k) x(jlu(k)) over k = jlu(i) .. ju(i)-1), the backward sweep with U
// and the inverted diagonal (x(i) = alu(i) (x(i) - sum alu(k) x(jlu(k)))
// over k = ju(i) .. jlu(i+1)-1).  One workgroup of 16 waves runs both sweeps;
// its waves take the rows of a topological order of each sweep round-robin
// (wave w: positions w, w+16, ...) and every lane waits only for the rows its
// own terms read, on a bitmap of finished rows in LDS, so rows of later
// levels start as soon as their inputs are final instead of at a barrier.
// The order is the concatenation of the level sets, so a wave never waits on
// a row that comes after its own in the order: the lowest unfinished
// position has all its inputs, and the sweep cannot deadlock.
//
// Each row is summed exactly as the level-synchronous kernel below sums it
// (lane l takes the terms l, l+64, ... in order, then the wave's DPP sum;
// this unit is compiled without contraction into FMAs), so both kernels give
// the same bits.  The level-synchronous kernel serves systems whose row
// bitmaps do not fit in LDS.
//
// A row's values and column indices do not depend on other rows: they are
// loaded one row ahead, as are the row's own y(i) or x(i) and diagonal, so
// that only the gathers of x wait on other rows.
// A lane's x loads of one row are issued together after its waits.
#include <hip/hip_runtime.h>

#include <string>

#include "lw_internal.h"

namespace lw {
namespace {

constexpr int kWave = 64;
constexpr int kAhead = 4;  // prefetched terms per lane (rows of up to 256 terms)
// sleeps of a lane that waits on one row before it gives up (a safety net:
// the order makes the wait finite; reaching the limit is reported)
constexpr int kSpinLimit = 1 << 22;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
// s + a x where the term exists, else s, without a branch: a branch would
// let the compiler move the load of x into it, behind the next row's loads
__device__ __forceinline__ double2 add_term(double2 s, double2 a, double2 x, bool in) {
  const double2 p = cmul(a, x);
  const double2 t = cadd(s, p);
  return make_double2(in ? t.x : s.x, in ? t.y : s.y);
}

// Sum over the wavefront, every lane receiving it (all 64 lanes active):
// each 16-lane DPP row sums its lanes (rotations by 8 and 4, then the quad
// permutations [2,3,0,1] and [1,0,3,2]), then the four row sums are read
// into scalars and added as (r0 + r1) + (r2 + r3).  DPP moves are VALU
// operations, so the reduction stays off the LDS pipe that the row-ready
// bitmap uses; it is the step of every row on the solve's critical path.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double read_lane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_mov<0x128>(v);  // row_ror:8
  v += dpp_mov<0x124>(v);  // row_ror:4
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}
__device__ __forceinline__ double2 wave_sum(double2 v) { return make_double2(wave_sum(v.x), wave_sum(v.y)); }

// Waits until row j's bit is set (acquire: its x is then visible to this
// lane's loads; all waves of the workgroup share the CU's L1).  A wait that
// reaches the spin limit raises the workgroup's error flag, after which no
// lane waits any more, so the kernel ends within one limit in any case.
__device__ __forceinline__ void wait_row(const unsigned* bits, int j, int* lerr) {
  const unsigned m = 1u << (j & 31);
  int spins = 0;
  while (!(__hip_atomic_load(&bits[j >> 5], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) & m)) {
    if (__hip_atomic_load(lerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
    __builtin_amdgcn_s_sleep(1);
    if (++spins > kSpinLimit) {
      __hip_atomic_store(lerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      break;
    }
  }
}

// Sets row i's bit after its x store (release: the store is complete first).
__device__ __forceinline__ void publish_row(unsigned* bits, int i) {
  __hip_atomic_fetch_or(&bits[i >> 5], 1u << (i & 31), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// A row's prefetched operands.
template <typename V>
struct RowOps {
  V a[kAhead];
  int c[kAhead];
  double2 own;
  V dia;
};

// One sweep over the rows of `order`.  upper = false: the L sweep; true: the
// U sweep with the inverted diagonal.  Term range of row i (1-based): L
// [jlu(i), ju(i)), U [ju(i), jlu(i+1)).
//
// Per row, a lane first waits for all of its terms' rows, then issues all
// its x loads together, then the next row's values and columns, then sums
// its terms in order: the waits, not the loads, are serial.  The row numbers
// and term ranges of a wave's next 64 rows are loaded at once (lane l holds
// those of the wave's l-th next row) and broadcast when their turn comes.
template <typename V, bool upper>
__device__ __forceinline__ void sweep(int n, const double2* __restrict__ y, double2* __restrict__ x,
                                      const V* __restrict__ alu, const int* __restrict__ jlu,
                                      const int* __restrict__ ju, const int* __restrict__ order, unsigned* bits,
                                      int* lerr) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave, nw = blockDim.x / kWave;
  // rows (1-based) and term ranges of positions wave + nw * (base + l)
  int mi = 0, m0 = 0, m1 = 0;
  auto meta = [&](int base) {
    const int q = wave + nw * (base + lane);
    mi = q < n ? order[q] + 1 : 0;
    if (mi) {
      m0 = upper ? ju[mi] : jlu[mi];
      m1 = upper ? jlu[mi + 1] : ju[mi];
    } else {
      m0 = m1 = 0;
    }
    // use the values here: a register whose load may still be in flight
    // after the branch that reloads it would make every row wait for all of
    // its loads (the join cannot tell whether the reload ran)
    asm volatile("" ::"v"(mi), "v"(m0), "v"(m1));
  };
  // the row's operands: its values and columns (lane l: terms l, l+64, ...),
  // its own y(i) (forward) or x(i) from the forward sweep and inverted
  // diagonal alu(i) (backward), which no other row writes.  The loads are
  // unconditional (jlu(0) is the pad; r = 0, past the wave's last row,
  // loads row 1's): a wait for a load in a branch, or for one that a
  // branch may have skipped, would drain every load in flight
  auto fetch = [&](RowOps<V>& ro, int r, int k0, int k1) {
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
      const int k = k0 + lane + u * kWave;
      const bool in = k < k1;
      ro.a[u] = alu[in ? k : 0];
      const int j = jlu[in ? k : 0];
      ro.c[u] = in ? j - 1 : -1;
    }
    const int o = max(r, 1);
    ro.own = upper ? x[o - 1] : y[o - 1];
    if (upper) ro.dia = alu[o];
  };
  meta(0);
  int t = 0;  // this wave's row count
  int i = __shfl(mi, 0, kWave), k0 = __shfl(m0, 0, kWave), k1 = __shfl(m1, 0, kWave);
  // one row, whose operands are in cur, while the next row's load into nxt;
  // false after the wave's last row.  The loop below alternates two operand
  // sets, so that no register copy waits for a load in flight
  auto step = [&](RowOps<V>& cur, RowOps<V>& nxt) __attribute__((always_inline)) -> bool {
    const int tn = t + 1;
    if (tn % kWave == 0) meta(tn);
    const int in_ = __shfl(mi, tn % kWave, kWave), n0 = __shfl(m0, tn % kWave, kWave),
              n1 = __shfl(m1, tn % kWave, kWave);
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
      if (cur.c[u] >= 0) wait_row(bits, cur.c[u], lerr);
    double2 xv[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) xv[u] = x[cur.c[u] >= 0 ? cur.c[u] : 0];
    // the next row's loads go after the gathers: loads complete in order, so
    // the sum below then waits for the gathers only, not for an HBM fetch
    __builtin_amdgcn_sched_barrier(0);
    fetch(nxt, in_, n0, n1);
    __builtin_amdgcn_sched_barrier(0);  // and all of them before the first wait
    double2 s = make_double2(0.0, 0.0);
#pragma unroll
    for (int u = 0; u < kAhead; ++u) s = add_term(s, to_d2(cur.a[u]), xv[u], cur.c[u] >= 0);
    // rows longer than the prefetched terms, kAhead terms per lane at a time
    for (int kb = k0 + kAhead * kWave; kb < k1; kb += kAhead * kWave) {
      V ab[kAhead];
      int cb[kAhead];
#pragma unroll
      for (int u = 0; u < kAhead; ++u) {
        const int k = kb + lane + u * kWave;
        const bool in = k < k1;
        ab[u] = alu[in ? k : 0];
        const int j = jlu[in ? k : 0];
        cb[u] = in ? j - 1 : -1;
      }
#pragma unroll
      for (int u = 0; u < kAhead; ++u)
        if (cb[u] >= 0) wait_row(bits, cb[u], lerr);
#pragma unroll
      for (int u = 0; u < kAhead; ++u) xv[u] = x[cb[u] >= 0 ? cb[u] : 0];
#pragma unroll
      for (int u = 0; u < kAhead; ++u) s = add_term(s, to_d2(ab[u]), xv[u], cb[u] >= 0);
    }
    s = wave_sum(s);
    if (lane == 0) {
      const double2 v = make_double2(cur.own.x - s.x, cur.own.y - s.y);
      x[i - 1] = upper ? cmul(to_d2(cur.dia), v) : v;
      publish_row(bits, i - 1);
    }
    t = tn;
    i = in_;
    k0 = n0;
    k1 = n1;
    return i != 0;
  };
  RowOps<V> ra, rb;
  if (i) {
    fetch(ra, i, k0, k1);
    while (step(ra, rb) && step(rb, ra)) {
    }
  }
}

// lusol0C (modIterativo.f90:504-560): forward solve with the unit-diagonal L
// then backward solve with U and the inverted diagonal, rows taken level by
// level (the level sets of the reference's natural order).  A single
// workgroup runs every level; the barrier between levels orders the global
// writes of x for the next level's reads (all waves share one CU's L1).
template <typename V>
__global__ void __launch_bounds__(1024) k_lusol(const double2* __restrict__ y, double2* __restrict__ x,
                                                const V* __restrict__ alu, const int* __restrict__ jlu,
                                                const int* __restrict__ ju, const int* __restrict__ lptr,
                                                const int* __restrict__ lrows, int nlev_l,
                                                const int* __restrict__ uptr, const int* __restrict__ urows,
                                                int nlev_u) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave, nw = blockDim.x / kWave;
  // alu(k), jlu(k), ju(i) are 1-based as in the reference: arrays carry a pad
  for (int l = 0; l < nlev_l; ++l) {
    for (int q = lptr[l] + wave; q < lptr[l + 1]; q += nw) {
      const int i = lrows[q] + 1;
      double2 s = make_double2(0.0, 0.0);
      for (int k = jlu[i] + lane; k <= ju[i] - 1; k += kWave) s = cadd(s, cmul(to_d2(alu[k]), x[jlu[k] - 1]));
      s = wave_sum(s);
      if (lane == 0) {
        const double2 yi = y[i - 1];
        x[i - 1] = make_double2(yi.x - s.x, yi.y - s.y);
      }
    }
    __syncthreads();
  }
  for (int l = 0; l < nlev_u; ++l) {
    for (int q = uptr[l] + wave; q < uptr[l + 1]; q += nw) {
      const int i = urows[q] + 1;
      double2 s = make_double2(0.0, 0.0);
      for (int k = ju[i] + lane; k <= jlu[i + 1] - 1; k += kWave) s = cadd(s, cmul(to_d2(alu[k]), x[jlu[k] - 1]));
      s = wave_sum(s);
      if (lane == 0) {
        const double2 xi = x[i - 1];
        x[i - 1] = cmul(to_d2(alu[i]), make_double2(xi.x - s.x, xi.y - s.y));
      }
    }
    __syncthreads();
  }
}

template <typename V>
__global__ void __launch_bounds__(1024) k_lusol_flow(int n, const double2* __restrict__ y, double2* __restrict__ x,
                                                     const V* __restrict__ alu, const int* __restrict__ jlu,
                                                     const int* __restrict__ ju, const int* __restrict__ lorder,
                                                     const int* __restrict__ uorder, int* __restrict__ err) {
  extern __shared__ unsigned bits[];  // two bitmaps of ceil(n/32) words
  const int nwords = (n + 31) / 32;
  for (int w = threadIdx.x; w < 2 * nwords; w += blockDim.x) bits[w] = 0u;
  __shared__ int lerr;
  if (threadIdx.x == 0) lerr = 0;
  __syncthreads();
  sweep<V, false>(n, y, x, alu, jlu, ju, lorder, bits, &lerr);
  __syncthreads();  // x of the forward sweep is final before the backward sweep reads it
  sweep<V, true>(n, y, x, alu, jlu, ju, uorder, bits + nwords, &lerr);
  __syncthreads();
  if (threadIdx.x == 0 && lerr) atomicOr(err, 1);
}

}  // namespace

// LDS bytes of the two bitmaps for n rows.
size_t lusol_flow_lds(int n) { return 2 * (size_t)((n + 31) / 32) * sizeof(unsigned); }

int launch_lusol(Handle* h, const double2* y, double2* x) {
  if (h->lusol_flow) {
    const size_t shm = lusol_flow_lds(h->ne);
    if (h->fp64_near)
      hipLaunchKernelGGL(k_lusol_flow<double2>, dim3(1), dim3(1024), shm, h->stream, h->ne, y, x,
                         (const double2*)h->d_alu, h->d_jlu, h->d_ju, h->d_lrows, h->d_urows, h->d_lusol_err);
    else
      hipLaunchKernelGGL(k_lusol_flow<float2>, dim3(1), dim3(1024), shm, h->stream, h->ne, y, x,
                         (const float2*)h->d_alu, h->d_jlu, h->d_ju, h->d_lrows, h->d_urows, h->d_lusol_err);
  } else if (h->fp64_near) {
    hipLaunchKernelGGL(k_lusol<double2>, dim3(1), dim3(1024), 0, h->stream, y, x, (const double2*)h->d_alu,
                       h->d_jlu, h->d_ju, h->d_lptr, h->d_lrows, h->nlev_l, h->d_uptr, h->d_urows, h->nlev_u);
  } else {
    hipLaunchKernelGGL(k_lusol<float2>, dim3(1), dim3(1024), 0, h->stream, y, x, (const float2*)h->d_alu,
                       h->d_jlu, h->d_ju, h->d_lptr, h->d_lrows, h->nlev_l, h->d_uptr, h->d_urows, h->nlev_u);
  }
  LW_HIP(hipGetLastError());
  return LW_OK;
}

int lusol_flow_setup(Handle* h) {
  const size_t shm = lusol_flow_lds(h->ne);
  int lds = 0;
  LW_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
  h->lusol_flow = shm + 64 <= (size_t)lds;
  if (!h->lusol_flow) return LW_OK;  // the level-synchronous kernel serves larger systems
  LW_HIP(hipFuncSetAttribute((const void*)k_lusol_flow<float2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)shm));
  LW_HIP(hipFuncSetAttribute((const void*)k_lusol_flow<double2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)shm));
  return LW_OK;
}

}  // namespace lw
