// The reference's plain-text result files, written the way its Fortran
// writers format them (modGlobalMetod.f90:186-306), so that files from this
// library diff cleanly against the reference's own:
//   guardarVectorI  one I10 per line
//   guardarVectorR  one f16.8 per line
//   guardarMatrizR  each row's ncolumnas values as f16.8, one row per line
//                   (the incremental monostatic table last_rcs_mono_incremental.dat,
//                   main.f90:1056, and the RCS tables)
//   guardarVectorC  real and imaginary parts as two f16.8 per line (the
//                   current I, main.f90:1148)
// F16.8: right-justified in 16 columns with 8 decimals, the value correctly
// rounded (ties to even); a field that does not fit is 16 asterisks; a
// negative value that rounds to zero, and -0.0 itself, keep the minus sign.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "lywell.h"

namespace {

// one f16.8 field into out (16 characters and a terminator)
void f16_8(double v, char* out) {
  if (std::isnan(v)) {
    std::snprintf(out, 17, "%16s", "NaN");
    return;
  }
  if (std::isinf(v)) {
    std::snprintf(out, 17, "%16s", v > 0 ? "Inf" : "-Inf");
    return;
  }
  char buf[64];
  const int n = std::snprintf(buf, sizeof buf, "%16.8f", v);
  if (n < 0 || n > 16) {
    std::memset(out, '*', 16);
    out[16] = 0;
    return;
  }
  std::memcpy(out, buf, 17);
}

// I10: right-justified in 10 columns, asterisks when it does not fit
void i10(int v, char* out) {
  char buf[32];
  const int n = std::snprintf(buf, sizeof buf, "%10d", v);
  if (n < 0 || n > 10) {
    std::memset(out, '*', 10);
    out[10] = 0;
    return;
  }
  std::memcpy(out, buf, 11);
}

}  // namespace

extern "C" {

int lw_dat_format_f16_8(double v, char* out17) {
  if (!out17) return LW_ERR_ARG;
  f16_8(v, out17);
  return LW_OK;
}

int lw_dat_write_vector_c(const char* path, int n, const double* v) {
  if (!path || n < 0 || (n > 0 && !v)) return LW_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return LW_ERR_ARG;
  char a[17], b[17];
  for (int i = 0; i < n; ++i) {
    f16_8(v[2 * (size_t)i], a);
    f16_8(v[2 * (size_t)i + 1], b);
    std::fprintf(f, "%s%s\n", a, b);
  }
  return std::fclose(f) == 0 ? LW_OK : LW_ERR_ARG;
}

int lw_dat_write_matrix_r(const char* path, int nrows, int ncols, const double* m) {
  if (!path || nrows < 0 || ncols < 1 || (nrows > 0 && !m)) return LW_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return LW_ERR_ARG;
  char a[17];
  for (int i = 0; i < nrows; ++i) {
    for (int j = 0; j < ncols; ++j) {
      f16_8(m[(size_t)j * nrows + i], a);  // column-major, as the Fortran array
      std::fputs(a, f);
    }
    std::fputc('\n', f);
  }
  return std::fclose(f) == 0 ? LW_OK : LW_ERR_ARG;
}

int lw_dat_write_vector_r(const char* path, int n, const double* v) {
  return lw_dat_write_matrix_r(path, n, 1, v);
}

int lw_dat_write_vector_i(const char* path, int n, const int* v) {
  if (!path || n < 0 || (n > 0 && !v)) return LW_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return LW_ERR_ARG;
  char a[11];
  for (int i = 0; i < n; ++i) {
    i10(v[i], a);
    std::fprintf(f, "%s\n", a);
  }
  return std::fclose(f) == 0 ? LW_OK : LW_ERR_ARG;
}

}  // extern "C"
