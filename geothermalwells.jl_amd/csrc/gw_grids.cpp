// gw_grids.cpp -- the grid generators of GeothermalWells.jl (src/grids.jl),
// host-only C entry points of libgeowells_hip (no device code, callable
// without a GPU).
//
// Julia's float ranges -- range(a, b, length=n) and a:s:b over Float64 --
// lift decimal endpoints to exact rationals (base/twiceprecision.jl) and round
// each element once.  Here: each input double is read back as its shortest
// round-trip decimal (std::to_chars), held as an exact fraction of 128-bit
// integers, every element is formed exactly and rounded once to the nearest
// double (ties to even) by long division.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "geowells.h"

namespace {

using i128 = __int128;

struct Frac {
  i128 num;  // value = num / den, den > 0
  i128 den;
};

i128 gcd128(i128 a, i128 b) {
  if (a < 0) a = -a;
  while (b != 0) {
    const i128 t = a % b;
    a = b;
    b = t < 0 ? -t : t;
  }
  return a;
}

// The exact value of x's shortest round-trip decimal representation.
bool decimal_frac(double x, Frac& f) {
  char buf[64];
  const auto res = std::to_chars(buf, buf + sizeof buf - 1, x);
  if (res.ec != std::errc()) return false;
  *res.ptr = 0;
  i128 mant = 0;
  int frac_digits = 0, exp10 = 0, ndig = 0;
  bool neg = false, after_point = false;
  for (const char* c = buf; *c; ++c) {
    if (*c == '-') neg = true;
    else if (*c == '.') after_point = true;
    else if (*c == 'e' || *c == 'E') { exp10 = std::atoi(c + 1); break; }
    else if (*c >= '0' && *c <= '9') {
      mant = mant * 10 + (*c - '0');
      if (after_point) ++frac_digits;
      if (++ndig > 30) return false;
    } else return false;
  }
  int scale = frac_digits - exp10;  // value = mant * 10^-scale
  i128 den = 1;
  for (; scale > 0; --scale) den *= 10;
  for (; scale < 0; ++scale) mant *= 10;
  const i128 g = gcd128(mant, den);
  if (g > 1) { mant /= g; den /= g; }
  f.num = neg ? -mant : mant;
  f.den = den;
  return true;
}

// num / den (den > 0) rounded to the nearest double, ties to even.
double round_div(i128 num, i128 den) {
  if (num == 0) return 0.0;
  const bool neg = num < 0;
  i128 a = neg ? -num : num;
  // scale so that 2^53 <= a / den < 2^54 (a 54-bit quotient: 53 bits and a
  // rounding bit), tracking the power of two
  int e2 = 0;
  auto bits = [](i128 v) { int n = 0; while (v > 0) { v >>= 1; ++n; } return n; };
  int shift = 54 - (bits(a) - bits(den));
  if (bits(a) + shift > 125 || bits(den) - shift > 125)  // far outside the magnitudes of a grid
    return (double)((long double)num / (long double)den);
  i128 d = den;
  if (shift > 0) { a <<= shift; e2 = -shift; }
  else if (shift < 0) { d <<= -shift; e2 = -shift; }
  i128 q = a / d, r = a % d;
  if (q >= ((i128)1 << 54)) {  // one bit too many: halve, keeping the remainder exact
    r = r + (q & 1) * d;
    d <<= 1;
    q >>= 1;
    e2 += 1;
  } else if (q < ((i128)1 << 53)) {  // one bit too few
    a <<= 1;
    e2 -= 1;
    q = a / d;
    r = a % d;
  }
  // q holds 54 bits: drop the last with round-half-even (sticky: r != 0)
  const bool half = (q & 1) != 0;
  const bool sticky = r != 0;
  uint64_t m = (uint64_t)(q >> 1);
  e2 += 1;
  if (half && (sticky || (m & 1))) ++m;
  const double v = std::ldexp((double)m, e2);
  return neg ? -v : v;
}

// lcm of two denominators, false when it would leave the exact range
bool lcm_den(i128 a, i128 b, i128& out) {
  const i128 g = gcd128(a, b);
  const i128 q = a / g;
  if (q > ((i128)1 << 100) / b) return false;
  out = q * b;
  return true;
}

// range(a, b, length = n) for Float64 endpoints.
bool julia_range(double a, double b, int64_t n, std::vector<double>& out) {
  out.clear();
  if (n <= 0) return true;
  if (n == 1) { out.push_back(a); return true; }
  Frac fa, fb;
  if (!decimal_frac(a, fa) || !decimal_frac(b, fb)) return false;
  i128 den;
  if (!lcm_den(fa.den, fb.den, den)) return false;
  const i128 an = fa.num * (den / fa.den), bn = fb.num * (den / fb.den);
  for (int64_t i = 0; i < n; ++i) out.push_back(round_div(an * (n - 1 - i) + bn * i, den * (n - 1)));
  return true;
}

// start:step:stop for Float64 arguments (length floor((stop - start) / step) + 1
// in exact arithmetic).
bool julia_colon(double start, double step, double stop, std::vector<double>& out) {
  out.clear();
  Frac fs, fd, fe;
  if (!decimal_frac(start, fs) || !decimal_frac(step, fd) || !decimal_frac(stop, fe) || fd.num <= 0) return false;
  i128 den;
  if (!lcm_den(fs.den, fd.den, den) || !lcm_den(den, fe.den, den)) return false;
  const i128 sn = fs.num * (den / fs.den), dn = fd.num * (den / fd.den), en = fe.num * (den / fe.den);
  if (en < sn) return true;
  const i128 len = (en - sn) / dn + 1;
  if (len > (i128)100000000) return false;
  for (i128 k = 0; k < len; ++k) out.push_back(round_div(sn + dn * k, den));
  return true;
}

int64_t emit(const std::vector<double>& g, double* out, int64_t cap) {
  if (out)
    for (int64_t q = 0; q < (int64_t)g.size() && q < cap; ++q) out[q] = g[q];
  return (int64_t)g.size();
}

}  // namespace

extern "C" {

int gw_compute_domain(const gw_borehole* bh, int32_t n_bh, double buffer_x, double buffer_y, double buffer_z,
                      double* out) {
  if (!bh || n_bh <= 0 || !out) return GW_EINVAL;
  double xlo = bh[0].xc, xhi = bh[0].xc, ylo = bh[0].yc, yhi = bh[0].yc, hmax = bh[0].h;
  for (int b = 1; b < n_bh; ++b) {
    xlo = std::min(xlo, bh[b].xc); xhi = std::max(xhi, bh[b].xc);
    ylo = std::min(ylo, bh[b].yc); yhi = std::max(yhi, bh[b].yc);
    hmax = std::max(hmax, bh[b].h);
  }
  const double zmax = hmax + buffer_z;
  out[0] = xlo - buffer_x; out[1] = xhi + buffer_x;
  out[2] = ylo - buffer_y; out[3] = yhi + buffer_y;
  out[4] = 0.0 * zmax;     out[5] = zmax;
  return GW_OK;
}

int64_t gw_uniform_gridz(double zmin, double zmax, double dz, const gw_borehole* bh, int32_t n_bh, double* out,
                         int64_t cap) {
  if (!(dz > 0) || n_bh < 0 || (n_bh > 0 && !bh)) return GW_EINVAL;
  std::vector<double> g;
  if (!julia_colon(zmin, dz, zmax, g)) return GW_EINVAL;
  for (int b = 0; b < n_bh; ++b) g.push_back(bh[b].h);
  std::sort(g.begin(), g.end());
  g.erase(std::unique(g.begin(), g.end()), g.end());
  return emit(g, out, cap);
}

int64_t gw_adaptive_grid_1d(double xmin, double xmax, double dx_fine, double growth_factor, double dx_max,
                            const gw_borehole* bh, int32_t n_bh, int32_t direction, double* out, int64_t cap) {
  if (!bh || n_bh <= 0 || (direction != 0 && direction != 1) || !(dx_fine > 0) || !(growth_factor > 1) ||
      !(dx_max > 0))
    return GW_EINVAL;
  // boreholes ordered by their centre along the direction (a stable sort keeps
  // equal centres in input order)
  std::vector<int> order(n_bh);
  for (int b = 0; b < n_bh; ++b) order[b] = b;
  auto centre = [&](int b) { return direction == 0 ? bh[b].xc : bh[b].yc; };
  std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return centre(p) < centre(q); });
  // fine uniform spacing across each borehole's backfill diameter
  std::vector<std::vector<double>> iv(n_bh);
  for (int q = 0; q < n_bh; ++q) {
    const double c = centre(order[q]);
    const double r = 1.0 * bh[order[q]].r_backfill;
    const int64_t n = (int64_t)std::ceil((2 * r) / dx_fine) + 1;
    if (n > 10000000 || !julia_range(c - r, c + r, n, iv[q])) return GW_EINVAL;
  }
  // geometric growth outwards from the first and the last interval
  std::vector<double> head;
  double dx = dx_fine, x = iv[0].front();
  while (x > xmin) {
    dx = std::min(dx * growth_factor, dx_max);
    x = x - dx;
    head.push_back(x > xmin ? x : xmin);
  }
  iv[0].insert(iv[0].begin(), head.rbegin(), head.rend());
  dx = dx_fine;
  x = iv[n_bh - 1].back();
  while (x < xmax) {
    dx = std::min(dx * growth_factor, dx_max);
    x = x + dx;
    iv[n_bh - 1].push_back(x < xmax ? x : xmax);
  }
  // between neighbouring boreholes: grow from both sides, meet in the middle
  for (int q = 0; q + 1 < n_bh; ++q) {
    double xa = iv[q].back(), xb = iv[q + 1].front();
    dx = dx_fine;
    std::vector<double> left, right;
    while (xb - xa > growth_factor * dx) {
      dx = std::min(dx * growth_factor, dx_max);
      xa = xa + dx;
      xb = xb - dx;
      if (xb - xa > dx) {
        left.push_back(xa);
        right.push_back(xb);
      } else {
        left.push_back((xa + xb) / 2);
      }
    }
    iv[q].insert(iv[q].end(), left.begin(), left.end());
    iv[q + 1].insert(iv[q + 1].begin(), right.rbegin(), right.rend());
  }
  std::vector<double> g;
  for (auto& seg : iv) g.insert(g.end(), seg.begin(), seg.end());
  std::sort(g.begin(), g.end());
  g.erase(std::unique(g.begin(), g.end()), g.end());
  return emit(g, out, cap);
}

}  // extern "C"
