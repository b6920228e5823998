// sigmaRCS without device allocation: lw_rcs_sigma_noalloc, the variant of
// lw_rcs_sigma for callers that cannot allocate device memory during the call
// (code under stream capture, tight serving loops).  Its scratch memory (the
// angles, the per-chunk partial fields and the sigmas) is one file-scope
// __device__ array of fixed size; each device holds its own copy, whose
// address is looked up once per device when an operator is created on it.
// Calls on one device share that array, so calls that overlap in time on one
// device race; every call writes all the scratch it reads, so nothing needs
// clearing between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "lw_internal.h"

namespace lw {
namespace {

constexpr int kWave = 64;
constexpr int kRcsChunks = 16;  // partial sums per angle (edge ranges)
constexpr int kMaxDevices = 64;

// layout of the scratch array (doubles): theta, phi, sigma, then the partial
// fields part[((a * kRcsChunks + c) * 3 + d) * 2 + {re, im}]
constexpr size_t kOffTheta = 0;
constexpr size_t kOffPhi = kOffTheta + LW_RCS_NOALLOC_MAX_ANGLES;
constexpr size_t kOffSigma = kOffPhi + LW_RCS_NOALLOC_MAX_ANGLES;
constexpr size_t kOffPart = kOffSigma + LW_RCS_NOALLOC_MAX_ANGLES;
constexpr size_t kScratchDoubles = kOffPart + (size_t)LW_RCS_NOALLOC_MAX_ANGLES * kRcsChunks * 3 * 2;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// calc_E (modRCS.f90:60-89) over one range of edges for one observation
// angle: the partial sum of g(r_obs, rm_i) ((m_i.rrm) rrm/|rrm|^2 - m_i).
// Grid (chunk, angle); the observation point r_obs = centre + r_zl r^(theta,
// phi) (calc_RCS, modRCS.f90:109-153) is formed here from the staged angles.
__global__ void __launch_bounds__(256) k_rcs_part(int ne, const double* __restrict__ tb, const int* __restrict__ et,
                                                  const double* __restrict__ el, const double2* __restrict__ cur,
                                                  double cx, double cy, double cz, double rzl, double kfs,
                                                  double* __restrict__ sc) {
  __shared__ double sh[6][256 / kWave];
  const int a = blockIdx.y, c = blockIdx.x;
  const double th = sc[kOffTheta + a], ph = sc[kOffPhi + a];
  const double ro[3] = {cx + (rzl * sin(th) * cos(ph)), cy + (rzl * sin(th) * sin(ph)), cz + (rzl * cos(th))};
  const int per = (ne + kRcsChunks - 1) / kRcsChunks;
  const int i0 = c * per, i1 = min(ne, i0 + per);
  double2 E0 = make_double2(0, 0), E1 = make_double2(0, 0), E2 = make_double2(0, 0);
  for (int i = i0 + (int)threadIdx.x; i < i1; i += blockDim.x) {
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
    const double2 gI = cmul(make_double2(cs / rn, sn / rn), Il);
    const double c0 = dr * rr[0] / rr2 - dm[0], c1 = dr * rr[1] / rr2 - dm[1], c2 = dr * rr[2] / rr2 - dm[2];
    E0.x += gI.x * c0;
    E0.y += gI.y * c0;
    E1.x += gI.x * c1;
    E1.y += gI.y * c1;
    E2.x += gI.x * c2;
    E2.y += gI.y * c2;
  }
  double v[6] = {E0.x, E0.y, E1.x, E1.y, E2.x, E2.y};
  const int wv = threadIdx.x / kWave, ln = threadIdx.x % kWave;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    v[q] = wave_sum(v[q]);
    if (ln == 0) sh[q][wv] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s += sh[threadIdx.x][w];
    sc[kOffPart + ((size_t)a * kRcsChunks + c) * 6 + threadIdx.x] = s;
  }
}

// sigmaRCS (modRCS.f90:91-105) per angle from the partial fields, summed in
// chunk order: E *= j w mu / 4 pi; sigma = 4 pi r_zl^2 |E|^2 / |pol ctte|^2.
__global__ void k_rcs_sigma(int n, double cm, double rzl, double pn, double* __restrict__ sc) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  double e[6] = {0, 0, 0, 0, 0, 0};
  for (int c = 0; c < kRcsChunks; ++c)
    for (int q = 0; q < 6; ++q) e[q] += sc[kOffPart + ((size_t)a * kRcsChunks + c) * 6 + q];
  double e2 = 0.0;
  for (int d = 0; d < 3; ++d) e2 += (e[2 * d] * e[2 * d] + e[2 * d + 1] * e[2 * d + 1]) * (cm * cm);
  sc[kOffSigma + a] = e2 * 4 * Physics::pi * (rzl * rzl) / pn;
}

}  // namespace

// The device address of this device's scratch array (the current device),
// looked up once per device; kept in the handle.
int rcs_static_setup(Handle* h) {
  if (h->device < 0 || h->device >= kMaxDevices) {
    set_error("rcs scratch: device ordinal beyond the supported range");
    return LW_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[h->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("rcs scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return LW_ERR_HIP;
    }
    g_scratch[h->device] = static_cast<double*>(p);
  }
  h->rcs_scratch = g_scratch[h->device];
  return LW_OK;
}

}  // namespace lw

extern "C" {

int lw_rcs_sigma_noalloc(lw_handle* hh, int n, const double* theta, const double* phi, double* sigma) {
  using namespace lw;
  if (!hh || n < 0 || (n > 0 && (!theta || !phi || !sigma))) return LW_ERR_ARG;
  Handle* h = &hh->h;
  if (n > LW_RCS_NOALLOC_MAX_ANGLES) {
    set_error("lw_rcs_sigma_noalloc: more than " + std::to_string(LW_RCS_NOALLOC_MAX_ANGLES) +
              " angles in one call (the size of its static scratch)");
    return LW_ERR_ARG;
  }
  if (!h->rcs_ready) {
    set_error("lw_rcs_sigma_noalloc: call lw_rcs_init first");
    return LW_ERR_STATE;
  }
  if (!h->rcs_scratch) {
    set_error("lw_rcs_sigma_noalloc: no scratch address for this device");
    return LW_ERR_STATE;
  }
  if (n == 0) return LW_OK;
  double* sc = h->rcs_scratch;
  hipStream_t st = h->stream;
  const double w = h->plan.phys.w;
  const double kfs = w * std::sqrt(Physics::mu * Physics::epsi);
  const double cm = w * Physics::mu / (4 * Physics::pi);
  double pn = 0.0;
  for (int c = 0; c < 3; ++c) {
    const double re = h->pol[c].x * h->ctte.x - h->pol[c].y * h->ctte.y;
    const double im = h->pol[c].x * h->ctte.y + h->pol[c].y * h->ctte.x;
    pn += re * re + im * im;
  }
  LW_HIP(hipMemcpyAsync(sc + kOffTheta, theta, sizeof(double) * n, hipMemcpyHostToDevice, st));
  LW_HIP(hipMemcpyAsync(sc + kOffPhi, phi, sizeof(double) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_rcs_part, dim3(kRcsChunks, n), dim3(256), 0, st, h->ne, h->d_tb, h->d_et, h->d_el, h->d_cur,
                     h->rcs_centre[0], h->rcs_centre[1], h->rcs_centre[2], h->rcs_rzl, kfs, sc);
  hipLaunchKernelGGL(k_rcs_sigma, dim3((n + 63) / 64), dim3(64), 0, st, n, cm, h->rcs_rzl, pn, sc);
  LW_HIP(hipGetLastError());
  LW_HIP(hipMemcpyAsync(sigma, sc + kOffSigma, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  LW_HIP(hipStreamSynchronize(st));
  return LW_OK;
}

}  // extern "C"
