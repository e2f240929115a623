// params.hpp -- exact host-side parameter selection for D(x) and S2_hard(x).
//
// The result of D(x, y, z, k) depends on (y, z, k), and the reference picks
// them in IEEE double from std::log and truncation, so bit-exact D(x) needs
// the same double arithmetic.  The formulas follow
//   get_alpha_gourdon          src/util.cpp:276-349
//   get_x_star_gourdon         src/util.cpp:362-386
//   get_alpha_deleglise_rivat  src/util.cpp:222-263
//   PhiTiny::get_k / get_c     src/PhiTiny.hpp:47-63
//   y, z, k clamps             src/app/main.cpp:205-246 (D), 435-461 (S2_hard)
// This translation unit is compiled with -ffp-contract=off so that no
// multiply-add is fused (the reference's x86-64 build has no FMA either).
#pragma once

#include "pc_math.hpp"

namespace pcg {

struct GourdonParams {
  int64_t y, z, k, x_star;
  double alpha_y, alpha_z;
};

struct DRParams {
  int64_t y, z, c;
  double alpha;
};

/// Tuning-factor overrides, as primecount's set_alpha_y/set_alpha_z/set_alpha
/// (src/util.cpp:103-133): values < 1 mean "use the default formula".
void set_alpha_y(double a);
void set_alpha_z(double a);
void set_alpha(double a);

std::pair<double, double> get_alpha_gourdon(i128 x);
double get_alpha_deleglise_rivat(i128 x);
int64_t get_x_star_gourdon(i128 x, int64_t y);
int64_t get_c(uint64_t y);
int64_t get_k(i128 x);

/// x -> (y, z, k, x_star) as the reference's D(maxint_t x, threads) picks them;
/// throws pcg_error("x is too large") when x > y * 2^62.
GourdonParams gourdon_params(i128 x);

/// x -> (y, z, c) as the reference's S2_hard(maxint_t x, threads) picks them.
DRParams deleglise_rivat_params(i128 x);

} // namespace pcg
