// Device-side state of one MLFMA operator (one mesh) and shared helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "lw_plan.h"
#include "lywell.h"

namespace lw {

void set_error(const std::string& m);
int progress_every();

// Per-basis geometry of the radiation pattern (calcRad, modMLFMA.f90:669-699),
// in leaf-slot order: dp = c+ - C, dm = c- - C (triangle centroids relative to
// the leaf-box centre C), rp = c+ - v+, rm = v- - c- (RWG vectors at the
// centroids), and the edge length.
struct RadGeom {
  double dp[3], dm[3], rp[3], rm[3], len;
};

// funcRAD of one basis at one sample (theta and phi components), the
// reference's non-maxPrecision formula with its operation order; used both
// to tabulate the patterns and to regenerate them inside the matvec, with
// contraction off so that both give the same bits.
// The phase exp(j ka khat . d) of one triangle centroid offset d from the
// leaf-box centre, as (cos, sin), with calcRad's operation order.
__device__ inline double2 rad_phase(const double* __restrict__ vk, const double* __restrict__ d, double ka) {
#pragma clang fp contract(off)
  double s, c;
  sincos(ka * ((vk[0] * d[0] + vk[1] * d[1]) + vk[2] * d[2]), &s, &c);
  return make_double2(c, s);
}
// funcRAD of one basis at one sample from the phases of its two triangles
// (pp = plus triangle, pm = minus triangle): both the tabulated patterns and
// the patterns assembled inside the matvec from per-box triangle phases go
// through this function, so both give the same bits.
__device__ inline void rad_from_phase(const RadGeom& g, double2 pp, double2 pm, const double* __restrict__ at,
                                      const double* __restrict__ ap, double2& r0, double2& r1) {
#pragma clang fp contract(off)
  double vx[3], vy[3];
  for (int d = 0; d < 3; ++d) {
    vx[d] = ((pp.x * g.rp[d] + pm.x * g.rm[d]) * g.len) * 0.5;
    vy[d] = ((pp.y * g.rp[d] + pm.y * g.rm[d]) * g.len) * 0.5;
  }
  r0.x = (vx[0] * at[0] + vx[1] * at[1]) + vx[2] * at[2];
  r0.y = (vy[0] * at[0] + vy[1] * at[1]) + vy[2] * at[2];
  r1.x = (vx[0] * ap[0] + vx[1] * ap[1]) + vx[2] * ap[2];
  r1.y = (vy[0] * ap[0] + vy[1] * ap[1]) + vy[2] * ap[2];
}
__device__ inline void rad_eval(const RadGeom& g, const double* __restrict__ vk, const double* __restrict__ at,
                                const double* __restrict__ ap, double ka, double2& r0, double2& r1) {
  rad_from_phase(g, rad_phase(vk, g.dp, ka), rad_phase(vk, g.dm, ka), at, ap, r0, r1);
}

// Samples s = iphi * nth + itheta (theta = (itheta + 1/2) pi / nth, phi =
// (iphi + 1/2) 2 pi / nph, nph = 2 nth).  Sample s < nk/2 and sample
// nk/2 + sample_flip(s, nth) (phi + pi, pi - theta) have opposite directions
// khat, so exp(j ka khat . d) of the second is the conjugate of the first's.
// sample_flip(u) reverses theta within u's phi column (an involution).
__host__ __device__ __forceinline__ int sample_flip(int u, int nth) {
  const int q = u / nth;
  return q * nth + (nth - 1 - (u - q * nth));
}

// Workgroup id w of n -> logical block: the hardware deals workgroups to the
// 8 XCDs round-robin (XCD = w mod 8); this bijection gives each XCD one
// contiguous range of logical blocks, so neighbouring blocks (which share
// operands) share that XCD's L2.
__device__ __forceinline__ int xcd_block(int w, int n) {
  const int q = n / 8, rem = n % 8, x = w % 8, j = w / 8;
  return x * q + min(x, rem) + j;
}

#define LW_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t _e = (call);                                                          \
    if (_e != hipSuccess) {                                                          \
      ::lw::set_error(std::string(#call) + ": " + hipGetErrorString(_e));            \
      return _e == hipErrorOutOfMemory ? LW_ERR_NOMEM : LW_ERR_HIP;                  \
    }                                                                                \
  } while (0)

// One entry of the transposed interpolation stencil (Anterpolar in gather
// form): a weight and the coarse sample it reads, loaded as one 16-byte word.
struct alignas(16) WtEntry {
  double w;
  int r;
  int pad;
};

// Device arrays of one level (levels >= 3).
struct DevLevel {
  int nk = 0, nb = 0, nT = 0;
  double* kap = nullptr;      // (3, nk)
  double* pes = nullptr;      // (nk)
  double* centre = nullptr;   // (3, nb)
  int* parent = nullptr;      // (nb) into level-1
  int* child_ptr = nullptr;   // (nb+1) into level+1 (non-leaf)
  double2* T = nullptr;       // (nk, nT), transfer function of vector t at T[t*nk + k]
  double2* C = nullptr;       // campoC (nk, 2, nb): C[(b*2+p)*nk + k]
  double2* B = nullptr;       // campoB
  double2* A = nullptr;       // campoA
  // interpolation from level+1 to this level (3 <= n < L)
  double* wint = nullptr;     // (9, nk)
  int* iint = nullptr;        // (9, nk) fine sample, 0-based
  // (weight, coarse row) pairs of the transpose for each fine sample, in
  // blocks of 64 samples stored entry-major (entry u of sample s at
  // wt[wt_ptr[s / 64] + 64 u + s % 64]), each block padded with zero weights
  int* wt_ptr = nullptr;      // (ceil(nk_fine / 64) + 1) block offsets
  WtEntry* wt = nullptr;
  // the downward pass through LDS windows (levels whose field does not fit
  // LDS whole): per chunk of 256 fine samples, the cyclic range of this
  // level's phi columns its stencils read (first column, count), and the
  // stencil with rows relative to the window; win_lds = 0 when not in use
  int* win = nullptr;         // (2 * chunks)
  WtEntry* wtw = nullptr;
  int win_lds = 0;
  // exp(+j k khat_i . r_o) for the 8 child offsets r_o = (+-q,+-q,+-q),
  // q = lado/4, octant bits (x,y,z) = (4,2,1), set = positive side (n < L)
  double2* ph = nullptr;      // (8, nk)
  int* oct = nullptr;         // octant of each box within its parent (n > 3)
  long long npairs = 0;
  // translation by parent blocks: per parent at level n-1 a flat source
  // list padded to a multiple of 4 entries (box, offset-code part of the
  // pair), the offset-code part of each box's octant, and
  // Tc[code*nk + k] = T[tr(code)][k] per integer offset code (zero for
  // offsets without a transfer vector)
  int* pc_ptr = nullptr;  // (nb of level n-1, + 1) children of each parent
  int* bcd = nullptr;
  int* ps_ptr = nullptr;
  int2* ps_ent = nullptr;
  long long nps = 0;  // entries of ps_ent
  double2* Tc = nullptr;
};

struct Handle {
  Plan plan;
  int flags = 0;
  bool fp64_near = false;
  int device = 0;
  hipStream_t stream = nullptr;
  // side streams of the matvec (near field; leaf translation; translations
  // of levels L-1..4) and the events that order them with the main stream:
  // fork, near done, leaf translation done, other translations done, and
  // campoC of level n ready (ev_c[n])
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_near = nullptr, ev_ta = nullptr, ev_tb = nullptr;
  std::vector<hipEvent_t> ev_c;
  double2* d_ynear = nullptr;  // near part of the matvec
  std::vector<DevLevel> lv;  // [1..L]
  int L = 0, ne = 0;
  long long nnz = 0;
  // geometry on the device
  double *d_p = nullptr, *d_tb = nullptr, *d_tbs = nullptr, *d_el = nullptr, *d_ec = nullptr;
  int *d_et = nullptr, *d_epo = nullptr;
  // leaf patterns and ordering
  double2* d_rad = nullptr;  // (nk_L, 2, ne) in slot order
  RadGeom* d_rgeo = nullptr;  // (ne) per-basis pattern geometry, slot order
  // the patterns as the matvec uses them: one phase exp(j ka khat . (c_T -
  // C)) per (leaf box, triangle of one of its bases) and leaf sample, and per
  // basis slot the phase rows of its plus and minus triangles
  int npairs = 0;
  double2* d_tph = nullptr;   // (npairs, ph_len) phase rows
  // phase-row length: nk_L / 2 for the factored aggregation and receive
  // (the samples with iphi < nph/2; the others are their antipodes), else nk_L
  int ph_len = 0;
  int2* d_slot_pair = nullptr; // (ne) phase rows of each slot's two triangles
  double *d_ath = nullptr, *d_aph = nullptr;  // (3, nk_L) theta and phi unit vectors of the leaf samples
  int *d_slot_basis = nullptr, *d_basis_slot = nullptr, *d_basis_box = nullptr, *d_box_slot = nullptr;
  // phase rows of each leaf box: [box_row[b], box_row[b+1]); the receive
  // from phase rows needs recv_lds bytes of LDS (0: it does not fit, and
  // the receive streams the tabulated patterns instead)
  int* d_box_row = nullptr;
  // the bases each phase row serves: entries 2 slot + side (0 plus, 1 minus
  // triangle) of row r at [row_ptr[r], row_ptr[r+1])
  int *d_row_ptr = nullptr, *d_row_ent = nullptr;
  int recv_lds = 0;
  // LDS of the factored aggregation (the H vectors of a box's phase rows;
  // used with the receive from phase rows) and of the near product by leaf
  // box (the box's x entries; 0: the rows of some box have different near
  // columns or they do not fit, and the near product runs row by row)
  int agg_lds = 0, near_lds = 0;
  long long near_box_cols = 0;  // sum over leaf boxes of their near-column counts
  // near field CSR (natural order)
  int *d_ia = nullptr, *d_ja = nullptr;
  void* d_z = nullptr;  // float2 or double2 values
  // ILU0 factor (MSR, 1-based indices kept as in the reference)
  bool have_ilu = false;
  void* d_alu = nullptr;  // float2 / double2, alu(k) at index k
  int *d_jlu = nullptr, *d_ju = nullptr;
  int *d_lptr = nullptr, *d_lrows = nullptr, *d_uptr = nullptr, *d_urows = nullptr;
  std::vector<int> h_lptr;  // host copy of the forward sweep's level pointers
  int nlev_l = 0, nlev_u = 0;
  int* d_lusol_err = nullptr;  // set when a triangular solve's wait reached its limit
  bool lusol_flow = false;      // the dataflow triangular solve fits (lw_lusol.hip)
  std::vector<float> h_alu32;
  std::vector<double> h_alu64;
  std::vector<int> h_jlu, h_ju;
  // work vectors (complex(8)[ne])
  double2 *d_x = nullptr, *d_y = nullptr, *d_w[12] = {};
  double* d_part = nullptr;   // reduction partials
  double2* d_scal = nullptr;  // CGS scalars
  int scal_cap = 0;
  // pinned host slots of two CGS iterations (|r|^2, next rho) and their events
  double2* h_cgs = nullptr;
  double2* d_cgs_host = nullptr;  // the same slots' device address
  hipEvent_t ev_cgs[2] = {nullptr, nullptr};
  // RCS state
  double2* d_cur = nullptr;
  double2 pol[3] = {};
  double2 ctte = {};
  bool rcs_ready = false;
  // far-zone distance r_zl and the centre of the edge-centre bounding box
  // (calc_RCS, modRCS.f90:109-153), fixed by the geometry
  double rcs_rzl = 0.0, rcs_centre[3] = {0.0, 0.0, 0.0};
  // this device's static scratch of lw_rcs_sigma_noalloc (lw_rcs_static.hip)
  double* rcs_scratch = nullptr;
  std::vector<void*> allocs;
};

template <typename T>
int dmalloc(Handle* h, T** p, size_t n) {
  void* q = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(&q, n * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return LW_ERR_NOMEM;
  }
  h->allocs.push_back(q);
  *p = reinterpret_cast<T*>(q);
  return LW_OK;
}

template <typename T>
int upload(Handle* h, T** p, const std::vector<T>& v) {
  int rc = dmalloc(h, p, v.size());
  if (rc) return rc;
  // ordered on the handle's stream (a non-blocking stream: null-stream
  // copies would not be ordered with its kernels) and completed before
  // returning, since callers may pass temporaries
  if (!v.empty()) {
    LW_HIP(hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
    LW_HIP(hipStreamSynchronize(h->stream));
  }
  return LW_OK;
}

// Stages of the single-RHS matvec as lw_bench_matvec times them.
constexpr int kMatvecStages = 6;
// Largest LDS request of the receive from phase rows (k_receive_ph).
constexpr int kRecvLdsMax = 128 * 1024;

// setup kernels (lw_setup_kernels.hip)
int launch_setup_rad(Handle* h);
int launch_setup_phases(Handle* h, const int* e_t);
int launch_setup_transfer(Handle* h, int level, const double* d_binom2, const double* d_tu,
                          const int* d_tL, const double* d_th, int hstride);
int launch_setup_near(Handle* h);

// The MSR factor's structure (and value array) on the device, and the level
// sets of both triangular sweeps (lw_device.hip).
int upload_structure(Handle* h);
// lusol0C on the device (lw_lusol.hip): the solve of y into x with the
// handle's factor on its stream, and the set-up of its LDS request.
int launch_lusol(Handle* h, const double2* y, double2* x);
int lusol_flow_setup(Handle* h);
// LW_ERR_STATE (and clears the flag) if a dataflow solve gave up a wait.
int lusol_status(Handle* h);
double sq_threshold(double t);
// ILU(0) of the near field on the device (lw_ilu0.hip); *piv = 0 or the first
// row with a zero pivot.
int ilu0_device(Handle* h, int* piv);

// Looks up this device's address of the static scratch array of
// lw_rcs_sigma_noalloc (once per device) and keeps it in h->rcs_scratch
// (lw_rcs_static.hip).
int rcs_static_setup(Handle* h);
// Parent-block tables of the single-RHS translation (lw_batch.hip). The far
// lists of every level must regroup into parent-neighbour blocks (checked
// against the plan's per-box lists); LW_ERR_STATE otherwise.
int setup_far_parent(Handle* h);
// campoB of level n by parent blocks on stream st (lw_batch.hip).
void launch_translate_parent(Handle* h, int n, hipStream_t st);

// One single-RHS matvec (MatxVec_MLFMA) on device vectors, ordered on the
// handle's stream (lw_device.hip).  With y_near, y receives the far part and
// y_near the near part.  With ev (7 events), the stages run one after the
// other on the handle's stream and ev[s], ev[s+1] bracket stage s
// (aggregation, upward passes, translations, downward passes, near field,
// receive); without, the near field and the translations of levels 4..L run
// on the handle's side streams beside the upward and downward passes.
int run_matvec(Handle* h, const double2* x, double2* y, double2* y_near, hipEvent_t* ev);
// The operator a batch was created on (lw_batch.hip); nullptr for nullptr.
Handle* batch_handle(lw_batch* b);

}  // namespace lw

// The opaque handle of the C ABI.
struct lw_handle {
  lw::Handle h;
};
