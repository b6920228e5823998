// Host-side operator plan: everything the device kernels need that is
// integer/structural work or parity-critical scalar work, computed once per
// mesh on the host (the reference does all of it in inicializar_MLFMA,
// modMLFMA.f90:113-275).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace lw {

struct Geometry {
  int ne = 0, nt = 0, np = 0;
  const double *e_centro = nullptr, *e_long = nullptr, *t_area = nullptr, *t_baric = nullptr,
               *t_baric_sub = nullptr, *p_coord = nullptr, *t_normal = nullptr;
  const int *e_t = nullptr, *e_po = nullptr, *t_p = nullptr, *e_p = nullptr;
};

struct Physics {
  double lambda = 1.0, w = 0.0, kappa = 0.0, vp = 0.0;
  static constexpr double mu = 1.2566370614E-6;     // modGlobalParam.f90:12
  static constexpr double epsi = 8.8541878176E-12;  // modGlobalParam.f90:13
  static constexpr double pi = 3.141592653589793;   // modGlobalParam.f90:17
  void set_lambda(double l);                        // setlambda, :53-61
};

// One octree level (numbered 1..L as in the reference).
struct LevelPlan {
  double lado = 0.0;
  int nmult = 0, nth = 0, nph = 0, nk = 0;
  int nb = 0;
  std::vector<double> centre;   // (3, nb)
  std::vector<int> parent;      // (nb), 0-based index in level-1
  std::vector<int> child_ptr;   // (nb+1) into level+1 boxes (children are contiguous)
  std::vector<std::vector<int>> bases;  // per box, ascending 1-based basis ids
  // sampling (levels >= 3): sample index = (iphi-1)*nth + itheta
  std::vector<double> kap, ath, aph, ang, pes;  // (3,nk) (3,nk) (3,nk) (2,nk) (nk)
  // interaction lists (levels >= 3), CSR over boxes
  std::vector<int> far_ptr, far_box, far_tr;  // 0-based
  // unique transfer vectors in first-seen order
  std::vector<double> trans_u;    // (3, nT) rab / |rab|
  std::vector<int> trans_L;       // (nT) number of multipole terms - 1 (Ltrue)
  std::vector<double> trans_h;    // (2, hstride, nT) h_l(kD), l = 0..Ltrue
  int hstride = 0;
  int nT = 0;
  // interpolation from level+1 samples to this level's samples (3 <= n < L)
  std::vector<double> wint;  // (9, nk)
  std::vector<int> iint;     // (9, nk) 0-based fine sample
  std::vector<int> wt_ptr, wt_row;  // transpose: per fine sample, coarse rows ascending
  std::vector<double> wt_w;
  // the same stencil factored: fine theta rows (pt, lower, upper) and weights
  // per coarse theta index, fine phi columns and weights per coarse phi index
  std::vector<int> sth_idx, sph_idx;  // (3, nth), (3, nph), 0-based
  std::vector<double> sth_w, sph_w;
};

struct Plan {
  Physics phys;
  int L = 0;
  int ne = 0;
  std::vector<LevelPlan> lv;  // lv[1..L]
  // leaf ordering: slot s holds basis slot_basis[s] (0-based); boxes own
  // contiguous slot ranges box_slot[b]..box_slot[b+1]
  std::vector<int> slot_basis, basis_slot, basis_box, box_slot;
  // near CSR in the reference's row order (natural RWG order), 0-based
  std::vector<int> ia;
  std::vector<int> ja;
  long long nnz = 0;
  // binomial table for the Legendre sums: coef[n*(NMAX+1)+i] =
  // (n!/(i!(n-i)!))**2 evaluated as the reference does
  static constexpr int NMAX = 128;
  std::vector<double> binom2;
};

// Builds the plan; returns 0 or a negative LW_ERR_* code with msg filled.
int build_plan(const Geometry& g, double lambda, double ladomax, Plan& plan, std::string& msg);

// flang NORM2 of a 3-vector (one-pass scaled algorithm).
double norm2_3(const double* v);

// Level sets of the forward (L) and backward (U) sweeps of an MSR factor.
void tri_levels(int n, const std::vector<int>& jlu, const std::vector<int>& ju, bool upper,
                std::vector<int>& lev_ptr, std::vector<int>& lev_rows);

}  // namespace lw
