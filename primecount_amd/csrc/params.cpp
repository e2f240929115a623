// params.cpp -- exact host-side parameter selection (see params.hpp).
//
// The double-precision expressions below have to be evaluated exactly as the
// reference evaluates them (same operations, same order, no fused multiply-add),
// otherwise y, z and hence D(x) differ.  They are restated from
// src/util.cpp:55-59 (truncate3), 222-263, 276-349, 362-386 and
// src/app/main.cpp:205-246, 435-461.

#include "params.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>

namespace pcg {

namespace {

std::atomic<double> alpha_y_override{-1.0};
std::atomic<double> alpha_z_override{-1.0};
std::atomic<double> alpha_override{-1.0};

/// Keep 3 digits after the decimal point (src/util.cpp:55-59).
double truncate3(double n)
{
  return (double) (int64_t) (n * 1000) / 1000.0;
}

} // namespace

std::string to_string128(i128 v)
{
  if (v == 0) return "0";
  bool neg = v < 0;
  u128 u = neg ? (u128) 0 - (u128) v : (u128) v;
  std::string s;
  while (u > 0) { s.insert(s.begin(), char('0' + (int) (u % 10))); u /= 10; }
  return neg ? "-" + s : s;
}

void set_alpha_y(double a) { alpha_y_override = a < 1.0 ? -1.0 : truncate3(a); }
void set_alpha_z(double a) { alpha_z_override = a < 1.0 ? -1.0 : truncate3(a); }
void set_alpha(double a)   { alpha_override = a < 1.0 ? -1.0 : truncate3(a); }

std::pair<double, double> get_alpha_gourdon(i128 x)
{
  double alpha_y = alpha_y_override;
  double alpha_z = alpha_z_override;
  double x16 = (double) (i128) iroot<6>((u128) x);
  double logx = std::log((double) x);
  double alpha_yz;

  if ((double) x <= 1e11) {
    double a = 0.078173;
    double b = 1;
    alpha_yz = a * logx + b;
  } else {
    double a = 0.007888745747017503;
    double b = -0.7610431057184933;
    double c = 25.74106177417062;
    double d = -288.9823576050215;
    double logx2 = logx * logx;
    double logx3 = logx * logx * logx;
    alpha_yz = a * logx3 + b * logx2 + c * logx + d;
  }

  if (alpha_z < 1) {
    alpha_z = 1.0;
    alpha_z = in_between(1.0, alpha_yz / 5, alpha_z);
  }
  if (alpha_y < 1)
    alpha_y = alpha_yz / alpha_z;

  alpha_y = in_between(1.0, alpha_y, x16);
  alpha_y = truncate3(alpha_y);
  alpha_z = truncate3(alpha_z);
  alpha_y = in_between(1.0, alpha_y, x16);
  double max_alpha_z = std::max(1.0, x16 / alpha_y);
  alpha_z = in_between(1.0, alpha_z, max_alpha_z);
  return {alpha_y, alpha_z};
}

double get_alpha_deleglise_rivat(i128 x)
{
  double alpha = alpha_override;
  double x16 = (double) (i128) iroot<6>((u128) x);

  if (alpha < 1) {
    double logx = std::log((double) x);
    if ((double) x <= 1e9) {
      double a = 0.078173;
      double b = 1;
      alpha = a * logx + b;
    } else {
      double a = 0.00148918;
      double b = -0.0691909;
      double c = 1.00165;
      double d = 0.372253;
      double logx2 = logx * logx;
      double logx3 = logx * logx * logx;
      alpha = a * logx3 + b * logx2 + c * logx + d;
    }
  }

  alpha = in_between(1.0, alpha, x16);
  alpha = truncate3(alpha);
  return in_between(1.0, alpha, x16);
}

int64_t get_x_star_gourdon(i128 x, int64_t y)
{
  y = std::max<int64_t>(y, 1);
  i128 yy = (i128) y * y;
  i128 x_div_yy = ceil_div(x, yy);
  i128 r4 = (i128) iroot<4>((u128) x);
  int64_t x_star = (int64_t) std::max(r4, x_div_yy);
  int64_t sqrt_xy = (int64_t) isqrt((u128) (x / y));
  x_star = std::min(x_star, y);
  x_star = std::min(x_star, sqrt_xy);
  x_star = std::max<int64_t>(x_star, 1);
  return x_star;
}

int64_t get_c(uint64_t y)
{
  // number of primes <= n for n < 20; 8 = number of tabulated primes
  static const uint8_t small_pi[20] = {0, 0, 1, 2, 2, 3, 3, 4, 4, 4, 4, 5, 5, 6, 6, 6, 6, 7, 7, 8};
  return y < 20 ? small_pi[y] : 8;
}

int64_t get_k(i128 x)
{
  return get_c(iroot<4>((u128) x));
}

GourdonParams gourdon_params(i128 x)
{
  GourdonParams p{1, 1, 0, 1, 1.0, 1.0};
  if (x < 1) return p;
  auto alpha = get_alpha_gourdon(x);
  int64_t x13 = (int64_t) iroot<3>((u128) x);
  int64_t sqrtx = (int64_t) isqrt((u128) x);
  int64_t y = (int64_t) (x13 * alpha.first);
  y = std::max(y, x13 + 1);
  y = std::min(y, sqrtx - 1);
  y = std::max<int64_t>(y, 1);
  if (x > ((i128) y << 62))
    throw pcg_error("D(x): x is too large");
  int64_t k = get_k(x);
  int64_t z = (int64_t) (y * alpha.second);
  z = std::max(z, y);
  z = std::min(z, sqrtx - 1);
  z = std::max<int64_t>(z, 1);
  p.y = y; p.z = z; p.k = k;
  p.x_star = get_x_star_gourdon(x, y);
  p.alpha_y = alpha.first; p.alpha_z = alpha.second;
  return p;
}

DRParams deleglise_rivat_params(i128 x)
{
  DRParams p{1, 1, 0, 1.0};
  if (x < 1) return p;
  double alpha = get_alpha_deleglise_rivat(x);
  int64_t y = (int64_t) ((double) (i128) iroot<3>((u128) x) * alpha);
  if (x > ((i128) y << 62))
    throw pcg_error("S2_hard(x): x is too large");
  p.y = y;
  p.z = (int64_t) (x / y);
  p.c = get_c((uint64_t) y);
  p.alpha = alpha;
  return p;
}

} // namespace pcg
