// Host writers of output_solution's files (src/inpout_2d.f90:5934-6423) and
// of write_restart_file (:8037-8084), from records the device assembled
// (imex_snapshot_begin).  The text formats are the reference's edit
// descriptors: '(2e20.12,100(e20.12))' for the .q_2d files and
// '(100ES15.7E2)' for .p_2d, one line per cell in the reference's loop
// order (x fastest), a list-directed blank record after each row of cells
// and two at the end.  Every field has a fixed width, so each line's offset
// in the file is known in advance and threads format and write disjoint
// ranges of lines.  Decimal conversion is std::to_chars (correctly rounded,
// ties to even), the rounding the Fortran runtime applies.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "imex_hip.h"

namespace {

// right-justifies s (length n) into a field of width w at p
inline void put_field(char* p, int w, const char* s, int n) {
  if (n > w) {
    std::memset(p, '*', w);
    return;
  }
  std::memset(p, ' ', w - n);
  std::memcpy(p + w - n, s, n);
}

inline bool put_special(char* p, int w, double x) {
  if (std::isnan(x)) {
    put_field(p, w, "NaN", 3);
    return true;
  }
  if (std::isinf(x)) {
    const char* s = x < 0 ? (w >= 9 ? "-Infinity" : "-Inf") : (w >= 9 ? "Infinity" : "Inf");
    put_field(p, w, s, (int)std::strlen(s));
    return true;
  }
  return false;
}

// mantissa digits and decimal exponent of x in scientific form with `prec`
// digits after the point: buf receives "d.ddd", returns the exponent
inline int sci(double x, int prec, char* mant, int* mlen) {
  char b[64];
  auto r = std::to_chars(b, b + sizeof(b), std::fabs(x), std::chars_format::scientific, prec);
  char* e = std::find(b, r.ptr, 'e');
  *mlen = (int)(e - b);
  std::memcpy(mant, b, *mlen);
  int ex = 0;
  std::from_chars(e + (e[1] == '+' ? 2 : 1), r.ptr, ex);
  return ex;
}

// ESw.dEe with e = 2
inline void put_es(char* p, int w, int d, double x) {
  if (put_special(p, w, x)) return;
  char m[40], s[64];
  int ml;
  const int ex = sci(x, d, m, &ml);
  if (std::abs(ex) > 99) {
    std::memset(p, '*', w);
    return;
  }
  int n = 0;
  if (std::signbit(x)) s[n++] = '-';
  std::memcpy(s + n, m, ml);
  n += ml;
  s[n++] = 'E';
  s[n++] = ex < 0 ? '-' : '+';
  s[n++] = char('0' + std::abs(ex) / 10);
  s[n++] = char('0' + std::abs(ex) % 10);
  put_field(p, w, s, n);
}

// Ew.d: 0.d1..dd with the exponent E+nn, or +nnn beyond 99
inline void put_e(char* p, int w, int d, double x) {
  if (put_special(p, w, x)) return;
  char m[40], s[64];
  int ml, n = 0, ex;
  char dig[32];
  if (x == 0.0) {
    std::memset(dig, '0', d);
    ex = 0;
  } else {
    const int e10 = sci(x, d - 1, m, &ml);
    dig[0] = m[0];
    std::memcpy(dig + 1, m + 2, d - 1);
    ex = e10 + 1;
  }
  if (std::signbit(x)) s[n++] = '-';
  s[n++] = '0';
  s[n++] = '.';
  std::memcpy(s + n, dig, d);
  n += d;
  const int a = std::abs(ex);
  if (a <= 99) {
    s[n++] = 'E';
    s[n++] = ex < 0 ? '-' : '+';
    s[n++] = char('0' + a / 10);
    s[n++] = char('0' + a % 10);
  } else if (a <= 999) {
    s[n++] = ex < 0 ? '-' : '+';
    s[n++] = char('0' + a / 100);
    s[n++] = char('0' + (a / 10) % 10);
    s[n++] = char('0' + a % 10);
  } else {
    std::memset(p, '*', w);
    return;
  }
  if (n > w && s[0] == '0') put_field(p, w, s + 1, n - 1);  // optional leading zero dropped
  else put_field(p, w, s, n);
}

int pwrite_all(int fd, const char* p, size_t n, off_t off) {
  while (n) {
    const ssize_t k = ::pwrite(fd, p, n, off);
    if (k <= 0) return -1;
    p += k;
    n -= (size_t)k;
    off += k;
  }
  return 0;
}

int fwrite_record(FILE* f, const void* data, size_t bytes) {
  if (bytes > 0x7fffffffu) return -2;  // one record marker of 4 bytes
  const int32_t m = (int32_t)bytes;
  if (std::fwrite(&m, 4, 1, f) != 1) return -1;
  if (bytes && std::fwrite(data, 1, bytes, f) != bytes) return -1;
  if (std::fwrite(&m, 4, 1, f) != 1) return -1;
  return 0;
}

}  // namespace

extern "C" int imex_format_value(int fmt, double x, char* out) {
  if (!out) return -2;
  if (fmt == IMEX_FMT_E20_12) {
    put_e(out, 20, 12, x);
    out[20] = 0;
  } else if (fmt == IMEX_FMT_ES15_7) {
    put_es(out, 15, 7, x);
    out[15] = 0;
  } else if (fmt == IMEX_FMT_ES14_7) {
    put_es(out, 14, 7, x);
    out[14] = 0;
  } else {
    return -2;
  }
  return 0;
}

extern "C" int imex_write_records(const char* path, const double* rec, int64_t n_rows, int n_col,
                                  int nx, int fmt, int n_threads) {
  if (!path || (!rec && n_rows) || n_rows < 0 || n_col < 1 || nx < 1 || n_rows % nx) return -2;
  const int W = (fmt == IMEX_FMT_E20_12) ? 20 : (fmt == IMEX_FMT_ES15_7) ? 15 : 0;
  if (!W) return -2;
  const size_t L = (size_t)n_col * W + 1;      // one record line
  const size_t block = (size_t)nx * L + 3;     // a row of cells and its blank record
  const int64_t nk = n_rows / nx;
  const size_t total = (size_t)nk * block + 6;
  const int fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) return -1;
  int nt = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)nt, 64, std::max<int64_t>(nk, 1)}));
  std::vector<int> rc(nt, 0);
  auto work = [&](int t) {
    const int64_t k0 = nk * t / nt, k1 = nk * (t + 1) / nt;
    const int64_t rows_per = std::max<int64_t>(1, (int64_t)((8u << 20) / block));  // ~8 MB chunks
    std::vector<char> buf;
    for (int64_t ka = k0; ka < k1 && !rc[t]; ka += rows_per) {
      const int64_t kb = std::min(k1, ka + rows_per);
      buf.resize((size_t)(kb - ka) * block);
      char* p = buf.data();
      for (int64_t k = ka; k < kb; k++) {
        for (int j = 0; j < nx; j++) {
          const double* r = rec + ((size_t)k * nx + j) * n_col;
          for (int c = 0; c < n_col; c++, p += W) {
            if (W == 20) put_e(p, 20, 12, r[c]);
            else put_es(p, 15, 7, r[c]);
          }
          *p++ = '\n';
        }
        *p++ = ' ';
        *p++ = ' ';
        *p++ = '\n';
      }
      rc[t] = pwrite_all(fd, buf.data(), buf.size(), (off_t)((size_t)ka * block));
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  int r = 0;
  for (int x : rc) r |= x;
  if (!r) r = pwrite_all(fd, "  \n  \n", 6, (off_t)(total - 6));
  if (::close(fd)) r = -1;
  return r ? -1 : 0;
}

extern "C" int imex_write_restart(const char* path, int nx, int ny, int n_vars, int n_solid,
                                  double t, double dt, const double* q_aos, const double* B_cent,
                                  const double* erodible, const double* deposit,
                                  const double* erosion, const double* hmax,
                                  const double* pdynmax, const double* mod_vel_max,
                                  const double* Z) {
  if (!path || !q_aos || !B_cent || !erodible || !deposit || !erosion || !hmax || !pdynmax ||
      !mod_vel_max || nx < 1 || ny < 1)
    return -2;
  FILE* f = std::fopen(path, "wb");
  if (!f) return -1;
  const size_t N = (size_t)nx * ny;
  const int32_t dims[4] = {nx, ny, n_vars, n_solid};
  const double td[2] = {t, dt};
  int r = 0;
  if (!r) r = fwrite_record(f, dims, sizeof(dims));
  if (!r) r = fwrite_record(f, td, sizeof(td));
  if (!r) r = fwrite_record(f, q_aos, 8 * N * n_vars);
  if (!r) r = fwrite_record(f, B_cent, 8 * N);
  if (!r) r = fwrite_record(f, erodible, 8 * N * n_solid);
  if (!r) r = fwrite_record(f, deposit, 8 * N * n_solid);
  if (!r) r = fwrite_record(f, erosion, 8 * N * n_solid);
  if (!r) r = fwrite_record(f, hmax, 8 * N);
  if (!r) r = fwrite_record(f, pdynmax, 8 * N);
  if (!r) r = fwrite_record(f, mod_vel_max, 8 * N);
  if (!r && Z) r = fwrite_record(f, Z, 8 * N);  // stochastic_flag runs (:8078-8080)
  if (std::fclose(f)) r = -1;
  return r;
}
