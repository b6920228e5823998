// Monostatic RCS sweep through a batch: the loop of gen_rcs_monostatica
// (main.f90:967-1089) handed to the multi-RHS solver in one call.
//
// The reference walks the n incidence angles phi_a = a (pi / n) one after
// another in the theta = 90 deg plane (propagation d = (-cos phi, -sin phi,
// 0), ajustar_polariz's VV or HH polarisation, main.f90:828-849), solves each
// with CGSold from the previous angle's current times
// exp(-j k (d - d_prev) . e_centro) (main.f90:1017-1022), and stores
// sigmaRCS(pi/2, phi_a) (main.f90:1050).
//
// Here the angles are cut into R chains of consecutive angles, one per batch
// column.  Each column iterates on its own; the host polls the batch every
// few iterations, and a column that has stopped is read out (its sigma taken
// from its current), then restarted on the next angle of its chain from the
// best guess its solved neighbours allow: the learned guess of the two
// nearest solved angles on one side when a network is given, else the
// reference's warm start from one solved neighbour, else zero.  A column
// whose chain is done takes the last pending angle of the longest chain,
// once the neighbours that angle's guess needs are solved, so columns do not
// idle while others still have work.  A sweep over the whole circle wraps
// around: its first and last angles are neighbours.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "lw_internal.h"

namespace {

using cplx = std::complex<double>;
constexpr double kPi = 3.141592653589793;
constexpr int kCheckEvery = 8;  // iterations between polls of the batch

struct Angle {
  double d[3];       // propagation direction
  double pr[3];      // polarisation (real)
  double pc[6];      // polarisation as complex(8) (re, im) triples
  double phi;
};

Angle make_angle(double phi, int pol) {
  Angle a;
  a.phi = phi;
  a.d[0] = -std::cos(phi);
  a.d[1] = -std::sin(phi);
  a.d[2] = 0.0;
  if (pol == LW_POL_VV) {
    a.pr[0] = 0.0, a.pr[1] = 0.0, a.pr[2] = 1.0;
  } else {
    // HH: d x z normalised
    const double hx = a.d[1], hy = -a.d[0];
    const double nh = std::sqrt(hx * hx + hy * hy);
    a.pr[0] = hx / nh, a.pr[1] = hy / nh, a.pr[2] = 0.0;
  }
  for (int k = 0; k < 3; ++k) a.pc[2 * k] = a.pr[k], a.pc[2 * k + 1] = 0.0;
  return a;
}

}  // namespace

extern "C" int lw_sweep_monostatic(lw_handle* hh, lw_batch* bb, lw_mlp* net, int n, double span, int pol,
                                   const double* ctte, double errmax, int maxit, double* sigma, int* iterations,
                                   int* converged, double* currents, int* stats) {
  using lw::set_error;
  if (!hh || !bb || !ctte || !sigma || n < 1 || !(span > 0.0 && span <= 2 * kPi + 1e-12) ||
      (pol != LW_POL_VV && pol != LW_POL_HH) || !(errmax > 0.0) || maxit < 1) {
    set_error("lw_sweep_monostatic: invalid argument");
    return LW_ERR_ARG;
  }
  lw::Handle* h = &hh->h;
  if (lw::batch_handle(bb) != h) {
    set_error("lw_sweep_monostatic: the batch was not created on this handle");
    return LW_ERR_ARG;
  }
  const int R = lw_batch_nrhs(bb);
  const int ne = h->ne;
  const double kappa = h->plan.phys.kappa;
  const bool wrap = std::fabs(span - 2 * kPi) < 1e-9;
  const int need = net ? 2 : 1;  // solved neighbours the best guess uses
  int rc;
#define TRY(x)         \
  do {                 \
    rc = (x);          \
    if (rc) return rc; \
  } while (0)

  std::vector<Angle> ang(n);
  // angphi = (ang - 1) dephi with dephi = pi / n (main.f90:999-1008)
  const double dphi = span / n;
  for (int a = 0; a < n; ++a) ang[a] = make_angle(a * dphi, pol);
  // the edge centres, for the warm start's phase correction
  std::vector<double> ec(3 * (size_t)ne);
  LW_HIP(hipMemcpyAsync(ec.data(), h->d_ec, sizeof(double) * ec.size(), hipMemcpyDeviceToHost, h->stream));
  LW_HIP(hipStreamSynchronize(h->stream));

  // a -> the angle s steps away, or -1 past the ends
  auto nb = [&](int a, int s) {
    const int b = a + s;
    if (wrap) return ((b % n) + n) % n;
    return (b >= 0 && b < n) ? b : -1;
  };
  std::vector<std::vector<cplx>> solved(n);
  auto is_solved = [&](int a) { return a >= 0 && !solved[a].empty(); };
  auto warm = [&](int a, int a1, cplx* g) {
    // x1 exp(-j k (d - d1) . e_centro), per unknown
    const double* d = ang[a].d;
    const double* d1 = ang[a1].d;
    const double dd[3] = {d[0] - d1[0], d[1] - d1[1], d[2] - d1[2]};
    const cplx* x1 = solved[a1].data();
    for (int i = 0; i < ne; ++i) {
      const double* e = &ec[3 * (size_t)i];
      const double t = -kappa * ((dd[0] * e[0] + dd[1] * e[1]) + dd[2] * e[2]);
      g[i] = x1[i] * cplx(std::cos(t), std::sin(t));
    }
  };
  std::vector<cplx> g(ne), b(ne), x(ne);
  std::vector<int> kind(n, 0);
  // the guess for angle a into g (kind[a] = 0: a zero guess, g unused)
  auto guess = [&](int a) -> int {
    kind[a] = 0;
    if (net) {
      for (int s : {-1, 1}) {
        const int a1 = nb(a, s), a2 = nb(a, 2 * s);
        if (is_solved(a1) && is_solved(a2)) {
          TRY(lw_guess_predict(net, hh, 1, ang[a].d, ang[a].pr, ang[a1].d, (const double*)solved[a1].data(),
                               ang[a2].d, (const double*)solved[a2].data(), (double*)g.data()));
          kind[a] = 2;
          return 0;
        }
      }
    }
    for (int s : {-1, 1}) {
      const int a1 = nb(a, s);
      if (is_solved(a1)) {
        warm(a, a1, g.data());
        kind[a] = 1;
        return 0;
      }
    }
    return 0;
  };

  // chains of consecutive angles, lengths differing by at most one
  std::vector<std::deque<int>> chains(R);
  for (int c = 0; c < R; ++c)
    for (int a = (int)((long long)c * n / R); a < (int)((long long)(c + 1) * n / R); ++a) chains[c].push_back(a);
  std::vector<int> work(R, -1);
  for (int c = 0; c < R; ++c)
    if (!chains[c].empty()) work[c] = chains[c].front(), chains[c].pop_front();
  int restarts = 0, steals = 0;

  // every column starts from zero; idle columns repeat the last angle
  {
    std::vector<double> dirs(3 * (size_t)R), pols(6 * (size_t)R);
    for (int c = 0; c < R; ++c) {
      const Angle& A = ang[work[c] >= 0 ? work[c] : n - 1];
      std::memcpy(&dirs[3 * c], A.d, sizeof A.d);
      std::memcpy(&pols[6 * c], A.pc, sizeof A.pc);
    }
    std::vector<cplx> B((size_t)ne * R);
    TRY(lw_batch_plane_wave(bb, dirs.data(), pols.data(), ctte, (double*)B.data()));
    TRY(lw_batch_cgs_start(bb, (const double*)B.data(), nullptr, errmax, 0, maxit));
  }
  auto start = [&](int c, int a) -> int {
    TRY(lw_plane_wave(hh, ang[a].d, ctte, ang[a].pc, (double*)b.data()));
    TRY(guess(a));
    TRY(lw_batch_restart_column(bb, c, (const double*)b.data(), kind[a] ? (const double*)g.data() : nullptr));
    ++restarts;
    return 0;
  };
  // the last pending angle of the longest chain whose last pending angle
  // already has the neighbours its guess needs on the far side
  auto steal = [&]() -> int {
    int best = -1;
    for (int j = 0; j < R; ++j) {
      if (chains[j].empty()) continue;
      const int a = chains[j].back();
      bool ok = true;
      for (int q = 1; q <= need; ++q) ok = ok && is_solved(nb(a, q));
      if (ok && (best < 0 || chains[j].size() > chains[best].size())) best = j;
    }
    if (best < 0) return -1;
    const int a = chains[best].back();
    chains[best].pop_back();
    ++steals;
    return a;
  };

  std::vector<int> st(4 * (size_t)R);
  std::vector<double> rr(R);
  const double half_pi = kPi / 2;
  for (;;) {
    bool any = false;
    for (int c = 0; c < R; ++c) any = any || work[c] >= 0;
    if (!any) break;
    TRY(lw_batch_cgs_advance(bb, kCheckEvery, st.data(), rr.data()));
    for (int c = 0; c < R; ++c) {
      const int a = work[c];
      if (a < 0 || st[4 * c]) continue;  // idle, or still iterating
      TRY(lw_batch_column(bb, c, (double*)x.data()));
      if (iterations) iterations[a] = st[4 * c + 3];
      if (converged) converged[a] = st[4 * c + 2];
      // inicializar_RCS + sigmaRCS(pi/2, phi) (main.f90:1048-1050)
      TRY(lw_rcs_init(hh, (const double*)x.data(), ang[a].pc, ctte));
      TRY(lw_rcs_sigma(hh, 1, &half_pi, &ang[a].phi, &sigma[a]));
      if (currents) std::memcpy(currents + 2 * (size_t)ne * a, x.data(), sizeof(cplx) * ne);
      solved[a] = x;
      work[c] = -1;
      if (!chains[c].empty()) {
        work[c] = chains[c].front();
        chains[c].pop_front();
        TRY(start(c, work[c]));
      }
    }
    for (int c = 0; c < R; ++c) {
      if (work[c] >= 0) continue;
      const int a = steal();
      if (a < 0) break;
      work[c] = a;
      TRY(start(c, a));
    }
  }
  if (stats) {
    stats[0] = restarts;
    stats[1] = steals;
    stats[2] = stats[3] = stats[4] = 0;
    // guesses: from zero, from one neighbour, from two (counted over all angles)
    for (int a = 0; a < n; ++a) ++stats[2 + kind[a]];
  }
#undef TRY
  return LW_OK;
}

// ---------------------------------------------------------------- network file

namespace {

// One stored (uncompressed) member of a zip archive: name -> bytes.
bool zip_member(const std::vector<unsigned char>& z, const std::string& want, const unsigned char** data,
                size_t* size) {
  size_t off = 0;
  auto u16 = [&](size_t o) { return (unsigned)z[o] | ((unsigned)z[o + 1] << 8); };
  auto u32 = [&](size_t o) { return (uint32_t)u16(o) | ((uint32_t)u16(o + 2) << 16); };
  auto u64 = [&](size_t o) { return (uint64_t)u32(o) | ((uint64_t)u32(o + 4) << 32); };
  while (off + 30 <= z.size() && u32(off) == 0x04034b50u) {
    const unsigned flags = u16(off + 6), method = u16(off + 8);
    uint64_t csize = u32(off + 18), usize = u32(off + 22);
    const unsigned nl = u16(off + 26), el = u16(off + 28);
    if (off + 30 + nl + el > z.size()) return false;
    const std::string name((const char*)&z[off + 30], nl);
    // ZIP64 sizes live in the extra field (id 1: uncompressed, compressed)
    for (size_t e = off + 30 + nl; e + 4 <= off + 30 + nl + el;) {
      const unsigned id = u16(e), len = u16(e + 2);
      if (id == 1 && len >= 16 && e + 4 + 16 <= z.size()) {
        if (usize == 0xffffffffu) usize = u64(e + 4);
        if (csize == 0xffffffffu) csize = u64(e + 12);
      }
      e += 4 + len;
    }
    const size_t data_off = off + 30 + nl + el;
    if ((flags & 8) || data_off + csize > z.size()) return false;  // streamed members are not supported
    if (name == want) {
      if (method != 0 || csize != usize) return false;  // stored only
      *data = &z[data_off];
      *size = (size_t)csize;
      return true;
    }
    off = data_off + csize;
  }
  return false;
}

// The payload of a .npy array of dtype `descr` with `count` elements (1-D, C order).
bool npy_payload(const unsigned char* p, size_t n, const char* descr, size_t* count, const unsigned char** data) {
  if (n < 10 || std::memcmp(p, "\x93NUMPY", 6) != 0) return false;
  size_t hl, ho;
  if (p[6] == 1) {
    hl = p[8] | (p[9] << 8);
    ho = 10;
  } else {
    if (n < 12) return false;
    hl = (size_t)p[8] | ((size_t)p[9] << 8) | ((size_t)p[10] << 16) | ((size_t)p[11] << 24);
    ho = 12;
  }
  if (ho + hl > n) return false;
  const std::string hdr((const char*)p + ho, hl);
  if (hdr.find(std::string("'descr': '") + descr + "'") == std::string::npos) return false;
  if (hdr.find("'fortran_order': False") == std::string::npos) return false;
  const size_t s = hdr.find("'shape': (");
  if (s == std::string::npos) return false;
  char* end = nullptr;
  const long long c = std::strtoll(hdr.c_str() + s + 10, &end, 10);
  if (c < 0 || !end || *end != ',') return false;  // 1-D shapes only: "(n,)"
  *count = (size_t)c;
  *data = p + ho + hl;
  return true;
}

}  // namespace

extern "C" int lw_guess_load(const char* path, lw_mlp** out) {
  using lw::set_error;
  if (!path || !out) return LW_ERR_ARG;
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    set_error(std::string("lw_guess_load: cannot open ") + path);
    return LW_ERR_ARG;
  }
  std::vector<unsigned char> z;
  unsigned char buf[1 << 16];
  for (size_t r; (r = std::fread(buf, 1, sizeof buf, f)) > 0;) z.insert(z.end(), buf, buf + r);
  std::fclose(f);
  const unsigned char *pa, *pp, *da, *dp;
  size_t sa, sp, na, np;
  if (!zip_member(z, "arch.npy", &pa, &sa) || !zip_member(z, "params.npy", &pp, &sp) ||
      !npy_payload(pa, sa, "<i8", &na, &da) || !npy_payload(pp, sp, "<f4", &np, &dp) || na != 4 ||
      (size_t)(da - pa) + 8 * na > sa || (size_t)(dp - pp) + 4 * np > sp) {
    set_error("lw_guess_load: not a network file (an uncompressed .npz with arch <i8 (4,) and params <f4 (n,))");
    return LW_ERR_ARG;
  }
  long long arch[4];
  std::memcpy(arch, da, sizeof arch);
  for (long long v : arch)
    if (v < 1 || v > 1 << 20) {
      set_error("lw_guess_load: implausible network shape");
      return LW_ERR_ARG;
    }
  lw_mlp* m = nullptr;
  int rc = lw_mlp_create((int)arch[0], (int)arch[1], (int)arch[2], (int)arch[3], 0, &m);
  if (rc) return rc;
  if ((long long)np != lw_mlp_nparams(m)) {
    lw_mlp_destroy(m);
    set_error("lw_guess_load: the parameter count does not match the network shape");
    return LW_ERR_ARG;
  }
  std::vector<float> p(np);
  std::memcpy(p.data(), dp, 4 * np);
  rc = lw_mlp_set_params(m, p.data());
  if (rc) {
    lw_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return LW_OK;
}
