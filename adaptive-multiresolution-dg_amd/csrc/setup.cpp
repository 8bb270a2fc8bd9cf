// Host setup tables; see setup.hpp for the reference locations each part follows.
#include "setup.hpp"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <stdexcept>

namespace adg {

// ---------------------------------------------------------------------------
// Gauss-Legendre nodes/weights on [-1,1].  The literal values are the ones the
// reference tabulates (source/Quad.cpp:31-80); they are kept digit for digit
// because the operator tables must come out bit-identical.
static const std::vector<std::vector<double>> kGLx = {
    {0},
    {-0.5773502691896256, 0.5773502691896256},
    {-0.7745966692414833, 0, 0.7745966692414832},
    {-0.8611363115940527, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526},
    {-0.9061798459386641, -0.538469310105683, 0, 0.5384693101056831, 0.9061798459386639},
    {-0.9324695142031522, -0.6612093864662647, -0.238619186083197, 0.2386191860831969, 0.6612093864662647, 0.9324695142031522},
    {-0.9491079123427585, -0.7415311855993943, -0.4058451513773971, 0, 0.4058451513773971, 0.7415311855993943, 0.9491079123427584},
    {-0.9602898564975365, -0.796666477413627, -0.525532409916329, -0.1834346424956498, 0.1834346424956496, 0.5255324099163292, 0.7966664774136268, 0.9602898564975364},
    {-0.968160239507626, -0.836031107326636, -0.61337143270059, -0.3242534234038094, 0, 0.3242534234038092, 0.6133714327005907, 0.8360311073266359, 0.9681602395076259},
    {-0.9739065285171715, -0.8650633666889844, -0.6794095682990244, -0.4333953941292472, -0.1488743389816311, 0.1488743389816314, 0.4333953941292472, 0.6794095682990243, 0.8650633666889843, 0.9739065285171714}};
static const std::vector<std::vector<double>> kGLw = {
    {2},
    {1, 1},
    {0.5555555555555556, 0.8888888888888895, 0.5555555555555554},
    {0.3478548451374547, 0.6521451548625466, 0.6521451548625458, 0.3478548451374541},
    {0.2369268850561892, 0.4786286704993669, 0.568888888888889, 0.4786286704993672, 0.2369268850561891},
    {0.1713244923791705, 0.3607615730481384, 0.4679139345726904, 0.467913934572691, 0.3607615730481382, 0.1713244923791708},
    {0.1294849661688697, 0.2797053914892765, 0.3818300505051193, 0.4179591836734696, 0.3818300505051192, 0.2797053914892776, 0.1294849661688697},
    {0.101228536290376, 0.2223810344533743, 0.3137066458778873, 0.3626837833783622, 0.3626837833783619, 0.3137066458778869, 0.2223810344533742, 0.1012285362903759},
    {0.08127438836157462, 0.1806481606948576, 0.2606106964029357, 0.3123470770400033, 0.3302393550012602, 0.3123470770400024, 0.2606106964029365, 0.1806481606948582, 0.08127438836157423},
    {0.06667134430868846, 0.14945134915058, 0.2190863625159823, 0.2692667193099961, 0.2955242247147523, 0.2955242247147531, 0.2692667193099947, 0.2190863625159819, 0.1494513491505811, 0.06667134430868851}};

const std::vector<double>& gl_nodes(int n) { return kGLx.at(n - 1); }
const std::vector<double>& gl_weights(int n) { return kGLw.at(n - 1); }

template <class F>
double gl10(F&& f, double tl, double tr) {
  const std::vector<double>& x = kGLx[9];
  const std::vector<double>& w = kGLw[9];
  double integral = 0;
  for (int i = 0; i < 10; ++i) {
    // Quad::linear_map(x, {-1,1}, {tl,tr}) = (x - xl)/(xr - xl)*(tr - tl) + tl
    const double t = (x[i] - (-1.)) / (1. - (-1.)) * (tr - tl) + tl;
    integral += w[i] * f(t);
  }
  integral *= (tr - tl) / 2.;
  return integral;
}

void cell_support(int level, int suppt, double& xl, double& xr) {
  if (level <= 1) { xl = 0.; xr = 1.; return; }
  xl = std::pow(2., -level + 1.) * (suppt - 1.) / 2.;
  xr = std::pow(2., -level + 1.) * (suppt + 1.) / 2.;
}

// ---------------------------------------------------------------------------
// Alpert multiwavelets (AlptBasis).

static double legendre_poly(int p, double x) {
  if (std::abs(x) > 1.) return 0.;
  switch (p) {
    case 0: return 1.;
    case 1: return x;
    case 2: return (3 * x * x - 1.) / 2.;
    case 3: return (5 * x * x * x - 3 * x) / 2.;
    case 4: return (35 * x * x * x * x - 30 * x * x + 3) / 8.;
  }
  throw std::runtime_error("legendre_poly: degree out of range");
}

static double legendre_poly_prime(int p, double x) {
  if (std::abs(x) > 1.) return 0.;
  switch (p) {
    case 0: return 0.;
    case 1: return 1;
    case 2: return 3 * x;
    case 3: return (15 * x * x - 3) / 2.;
    case 4: return (140 * x * x * x - 60 * x) / 8.;
  }
  throw std::runtime_error("legendre_poly_prime: degree out of range");
}

// Alpert's multiwavelets on [-1,1] (value), mirrored for x<0 with parity (p+P+1).
static double alpt_phi0(double x, int P, int p) {
  double sgn = 1.;
  if (x < 0) {
    if ((p + P + 1) % 2 != 0) sgn = -1.;
    return sgn * alpt_phi0(-x, P, p);
  }
  double result = 0.;
  if (P == 0) {
    if (p == 0) result = std::sqrt(0.5);
  } else if (P == 1) {
    if (p == 0) result = std::sqrt(1.5) * (-1. + 2. * x);
    else if (p == 1) result = std::sqrt(0.5) * (-2. + 3. * x);
  } else if (P == 2) {
    if (p == 0) result = std::sqrt(0.5) / 3. * (1. - 24. * x + 30. * x * x);
    else if (p == 1) result = std::sqrt(1.5) * 0.5 * (3. - 16. * x + 15. * x * x);
    else if (p == 2) result = std::sqrt(2.5) / 3. * (4. - 15. * x + 12. * x * x);
  } else if (P == 3) {
    if (p == 0) result = std::sqrt(15. / 34.) * (1. + 4. * x - 30. * x * x + 28. * x * x * x);
    else if (p == 1) result = std::sqrt(1. / 42.) * (-4. + 105. * x - 300. * x * x + 210. * x * x * x);
    else if (p == 2) result = std::sqrt(35. / 34) * 0.5 * (-5. + 48. * x - 105. * x * x + 64. * x * x * x);
    else if (p == 3) result = std::sqrt(5. / 42) * 0.5 * (-16. + 105. * x - 192. * x * x + 105. * x * x * x);
  } else {
    throw std::runtime_error("Alpert basis: degree > 3 is not provided");
  }
  return result;
}

// First derivative; parity of the derivative is opposite to the value's.
static double alpt_phi1(double x, int P, int p) {
  double sgn = 1.;
  if (x < 0) {
    if ((p + P + 1) % 2 == 0) sgn = -1.;
    return sgn * alpt_phi1(-x, P, p);
  }
  double result = 0.;
  if (P == 0) {
    if (p == 0) result = 0.0;
  } else if (P == 1) {
    if (p == 0) result = 2. * std::sqrt(1.5);
    else if (p == 1) result = 3. * std::sqrt(0.5);
  } else if (P == 2) {
    if (p == 0) result = (-24. + 60. * x) / (3. * std::sqrt(2.));
    else if (p == 1) result = (std::sqrt(1.5) * (-16. + 30. * x)) / 2.;
    else if (p == 2) result = (std::sqrt(2.5) * (-15. + 24. * x)) / 3.;
  } else if (P == 3) {
    if (p == 0) result = 0.6642111641550714 * (4. - 60. * x + 84 * x * x);
    else if (p == 1) result = 0.1543033499620919 * (105. - 600. * x + 630. * x * x);
    else if (p == 2) result = 0.5072996561958923 * (48. - 210. * x + 192. * x * x);
    else if (p == 3) result = 0.17251638983558856 * (105. - 384. * x + 315. * x * x);
  } else {
    throw std::runtime_error("Alpert basis derivative: degree > 3 is not provided");
  }
  return result;
}

static double alpt_val0(double x, int n, int j, int p, int P) {
  if (n == 0) return legendre_poly(p, 2 * x - 1) * std::sqrt(2 * p + 1.);
  if (n == 1) {
    if (x >= 0 && x <= 1) return std::sqrt(2.) * alpt_phi0(2 * x - 1, P, p);
    return 0.;
  }
  const int j_odd = (j - 1) / 2;
  const double shift_x = std::pow(2, n - 1) * x - j_odd;
  return std::pow(2, (n - 1) / 2.) * alpt_val0(shift_x, 1, 0, p, P);
}

static double alpt_val1(double x, int n, int j, int p, int P) {
  if (n == 0) return legendre_poly_prime(p, 2 * x - 1) * 2 * std::sqrt(2 * p + 1.);
  if (n == 1) {
    if (x >= 0 && x <= 1) return std::pow(2, 1.5) * alpt_phi1(2 * x - 1, P, p);
    return 0;
  }
  const int j_odd = (j - 1) / 2;
  const double shift_x = std::pow(2, n - 1) * x - j_odd;
  return std::pow(2, (n - 1) * 1.5) * alpt_val1(shift_x, 1, 0, p, P);
}

// ---------------------------------------------------------------------------
// Nested Lagrange interpolation basis, msh_case 1 (LagrBasis::phi).

static double lagr_phi(double xt, int msh, int P, int p) {
  const double x = 2 * xt - 1;  // [0,1] -> [-1,1]
  if (x < -1 || x > 1) return 0.;
  if (P == 1) {
    if (msh == 0) {
      if (p == 0) return -3. / 2. * (x - 1. / 3.);
      if (p == 1) return 3. / 2. * (x + 1. / 3.);
    } else {
      if (p == 0) return (x < 0) ? -3 * (x + 1. / 3.) : 0.;
      if (p == 1) return (x < 0) ? 0. : 3 * (x - 1. / 3.);
    }
  } else if (P == 2) {
    if (msh == 0) {
      if (p == 0) return 3 * (x + 1. / 3.) * (x - 1. / 3.);
      if (p == 1) return -9. / 2. * (x + 2. / 3.) * (x - 1. / 3.);
      if (p == 2) return 3. / 2. * (x + 2. / 3.) * (x + 1. / 3.);
    } else {
      if (p == 0) return (x <= 0) ? 12 * (x + 2. / 3.) * (x + 1. / 3.) : 0.;
      if (p == 1) return (x < 0) ? 0. : 12 * (x - 1. / 3.) * (x - 2. / 3.);
      if (p == 2) return (x < 0) ? 0. : 6 * (x - 1. / 6.) * (x - 1. / 3.);
    }
  } else if (P == 3) {
    if (msh == 0) {
      if (p == 0) return -125. / 48. * (x + 1. / 5.) * (x - 1. / 5.) * (x - 3. / 5.);
      if (p == 1) return 125. / 16. * (x + 3. / 5.) * (x - 1. / 5.) * (x - 3. / 5.);
      if (p == 2) return -125. / 16. * (x + 3. / 5.) * (x + 1. / 5.) * (x - 3. / 5.);
      if (p == 3) return 125. / 48. * (x + 3. / 5.) * (x + 1. / 5.) * (x - 1. / 5.);
    } else {
      if (p == 0) return (x <= 0) ? -125. / 6. * (x + 3. / 5.) * (x + 2. / 5.) * (x + 1. / 5.) : 0.;
      if (p == 1) return (x <= 0) ? -125. / 2. * (x + 4. / 5.) * (x + 3. / 5.) * (x + 1. / 5.) : 0.;
      if (p == 2) return (x < 0) ? 0. : 125. / 2. * (x - 1. / 5.) * (x - 3. / 5.) * (x - 4. / 5.);
      if (p == 3) return (x < 0) ? 0. : 125. / 6. * (x - 1. / 5.) * (x - 2. / 5.) * (x - 3. / 5.);
    }
  } else if (P == 4) {
    if (msh == 0) {
      if (p == 0) return 72. / 5. * (x - 1. / 3.) * (x - 1. / 6.) * (x + 1. / 3.) * (x + 5. / 12.);
      if (p == 1) return -768. / 7. * (x - 1. / 3.) * (x - 1. / 6.) * (x + 1. / 3.) * (x + 2. / 3.);
      if (p == 2) return 108 * (x - 1. / 3.) * (x - 1. / 6.) * (x + 5. / 12.) * (x + 2. / 3.);
      if (p == 3) return -864. / 35. * (x - 1. / 3.) * (x + 1. / 3.) * (x + 5. / 12.) * (x + 2. / 3.);
      if (p == 4) return 12 * (x - 1. / 6.) * (x + 1. / 3.) * (x + 5. / 12.) * (x + 2. / 3.);
    } else {
      if (p == 0) return (x <= 0) ? 1152. / 5. * (x + 1. / 3.) * (x + 5. / 12.) * (x + 2. / 3.) * (x + 17. / 24.) : 0.;
      if (p == 1) return (x <= 0) ? -12288. / 7. * (x + 1. / 3.) * (x + 5. / 12.) * (x + 2. / 3.) * (x + 5. / 6.) : 0.;
      if (p == 2) return (x < 0) ? 0. : -12288. / 7. * (x - 2. / 3.) * (x - 7. / 12.) * (x - 1. / 3.) * (x - 1. / 6.);
      if (p == 3) return (x < 0) ? 0. : -13824. / 35. * (x - 2. / 3.) * (x - 1. / 3.) * (x - 7. / 24.) * (x - 1. / 6.);
      if (p == 4) return (x < 0) ? 0. : 192 * (x - 7. / 12.) * (x - 1. / 3.) * (x - 7. / 24.) * (x - 1. / 6.);
    }
  } else if (P == 5) {
    // source/LagrBasis.cpp:617-681 (msh_case 1, continuous: level-0 points at
    // the cell ends, level-1 points on both sides of the midpoint)
    if (msh == 0) {
      if (p == 0) return -1.0 / 768. * (x - 1.) * (5.0 * x - 3.) * (5.0 * x + 3.) * (5.0 * x + 1.) * (5.0 * x - 1.);
      if (p == 1) return 25.0 / 768. * (x - 1.) * (x + 1.) * (5.0 * x - 3.) * (5.0 * x + 1.) * (5.0 * x - 1.);
      if (p == 2) return -25.0 / 384. * (x - 1.) * (x + 1.) * (5.0 * x - 3.) * (5.0 * x + 3.) * (5.0 * x - 1.);
      if (p == 3) return 25.0 / 384. * (x - 1.) * (x + 1.) * (5.0 * x - 3.) * (5.0 * x + 3.) * (5.0 * x + 1.);
      if (p == 4) return -25.0 / 768. * (x - 1.) * (x + 1.) * (5.0 * x + 3.) * (5.0 * x + 1.) * (5.0 * x - 1.);
      if (p == 5) return 1.0 / 768. * (x + 1.) * (5.0 * x - 3.) * (5.0 * x + 3.) * (5.0 * x + 1.) * (5.0 * x - 1.);
    } else {
      if (p == 0) return (x <= 0) ? 25. / 24. * x * (x + 1.0) * (5.0 * x + 1.) * (5.0 * x + 2.) * (5.0 * x + 3.) : 0.;
      if (p == 1) return (x <= 0) ? 25. / 12. * x * (x + 1.0) * (5.0 * x + 1.) * (5.0 * x + 4.) * (5.0 * x + 3.) : 0.;
      if (p == 2)
        return (x <= 0) ? 1. / 24. * (x + 1.0) * (5.0 * x + 1.) * (5.0 * x + 2.) * (5.0 * x + 3.) * (5.0 * x + 4.) : 0.;
      if (p == 3)
        return (x < 0) ? 0. : -1. / 24. * (x - 1.0) * (5.0 * x - 1.) * (5.0 * x - 2.) * (5.0 * x - 3.) * (5.0 * x - 4.);
      if (p == 4) return (x < 0) ? 0. : -25. / 12. * x * (x - 1.0) * (5.0 * x - 1.) * (5.0 * x - 4.) * (5.0 * x - 3.);
      if (p == 5) return (x < 0) ? 0. : -25. / 24. * x * (x - 1.0) * (5.0 * x - 1.) * (5.0 * x - 2.) * (5.0 * x - 3.);
    }
  }
  throw std::runtime_error("Lagrange basis: only degree 1..5, msh_case 1 are provided");
}

// ---------------------------------------------------------------------------
// Hermite P3 basis, msh_case 1 (HermBasis::phi_P1_L1 / phi_P1_L1_D1,
// source/HermBasis.cpp:347-656): two points per cell, value and slope at each.

static double herm_phi(double x, int msh, int pp) {
  if (x < 0 || x > 1) return 0.;
  const int p = pp % 2, l = pp / 2;
  if (msh == 0) {
    if (p == 0) return l == 0 ? (-1 + x) * (-1 + x) * (1 + 2 * x) : (1 - x) * (1 - x) * x;
    return l == 0 ? -x * x * (-3 + 2 * x) : (-1 + x) * x * x;
  }
  if (p == 0) {
    if (x > 0.5) return 0.;
    return l == 0 ? -4 * x * x * (4 * x - 3) : 2 * x * x * (2 * x - 1);
  }
  if (x < 0.5) return 0.;
  return l == 0 ? 4 * (x - 1) * (x - 1) * (4 * x - 1) : 2 * (x - 1) * (x - 1) * (2 * x - 1);
}

static double herm_phi_d1(double x, int msh, int pp) {
  if (x < 0 || x > 1) return 0.;
  const int p = pp % 2, l = pp / 2;
  if (msh == 0) {
    if (p == 0) return l == 0 ? 6 * x * (-1 + x) : (x - 1) * (3 * x - 1);
    return l == 0 ? -6 * x * (x - 1) : x * (3 * x - 2);
  }
  if (p == 0) {
    if (x > 0.5) return 0.;
    return l == 0 ? 24 * x * (1 - 2 * x) : 4 * x * (3 * x - 1);
  }
  if (x < 0.5) return 0.;
  return l == 0 ? 24 * (x - 1) * (2 * x - 1) : 4 * (x - 1) * (3 * x - 2);
}

// HermBasis::val0 / val1 (source/HermBasis.cpp:157-250): levels >= 2 are the
// level-1 functions shifted and scaled by 2^(-(n-1) * (l - derivative))
static double herm_val(double x, int n, int j, int p, int deriv) {
  if (n <= 1) return deriv == 0 ? herm_phi(x, n, p) : herm_phi_d1(x, n, p);
  const int odd_j = (j - 1) / 2;
  const double xtrans = std::pow(2, n - 1) * x - odd_j;
  const int l = p / 2;
  if (deriv == 0) return herm_phi(xtrans, 1, p) * std::pow(2, -(n - 1) * l);
  return herm_phi_d1(xtrans, 1, p) * std::pow(2, -(n - 1) * (l - 1));
}

static double lagr_val0(double x, int n, int j, int p, int P) {
  if (n <= 1) return lagr_phi(x, n, P, p);
  const int odd_j = (j - 1) / 2;
  const double xtrans = std::pow(2, n - 1) * x - odd_j;
  return lagr_phi(xtrans, 1, P, p);
}

// ---------------------------------------------------------------------------

Basis1D::Basis1D(BasisKind k, int pmax_, int nmax_, int msh_case_)
    : kind(k), pmax(pmax_), msh_case(msh_case_), nmax(nmax_) {
  if (kind == BasisKind::Hermite) {
    // HermBasis::set_interp_msh01 (source/HermBasis.cpp:40-68), P3 only
    if (pmax != 3 || msh_case != 1) throw std::runtime_error("Hermite basis: only degree 3, msh_case 1 is provided");
    msh0 = {-1.0 + ROUND_OFF, 1.0 - ROUND_OFF, -1.0 + ROUND_OFF, 1.0 - ROUND_OFF};
    msh1 = {0.0 - ROUND_OFF, 0.0 + ROUND_OFF, 0.0 - ROUND_OFF, 0.0 + ROUND_OFF};
    for (int i = 0; i < pmax + 1; ++i) {
      msh0[i] = (msh0[i] + 1) / 2.;
      msh1[i] = (msh1[i] + 1) / 2.;
    }
  } else if (kind == BasisKind::Lagrange) {
    if (msh_case != 1) throw std::runtime_error("Lagrange basis: only msh_case 1 is provided");
    // LagrBasis::set_interp_msh01 (source/LagrBasis.cpp:33-154), points on [-1,1]
    if (pmax == 1) { msh0 = {-1. / 3., 1. / 3.}; msh1 = {-2. / 3., 2. / 3.}; }
    else if (pmax == 2) { msh0 = {-2. / 3., -1. / 3., 1. / 3.}; msh1 = {-5. / 6., 1. / 6., 2. / 3.}; }
    else if (pmax == 3) { msh0 = {-3. / 5., -1. / 5., 1. / 5., 3. / 5.}; msh1 = {-4. / 5., -2. / 5., 2. / 5., 4. / 5.}; }
    else if (pmax == 4) {
      msh0 = {-2. / 3., -5. / 12., -1. / 3., 1. / 6., 1. / 3.};
      msh1 = {-5. / 6., -17. / 24., 7. / 24., 7. / 12., 2. / 3.};
    }
    else if (pmax == 5) {
      msh0 = {-1 + ROUND_OFF, -0.6, -0.2, 0.2, 0.6, 1 - ROUND_OFF};
      msh1 = {-0.8, -0.4, 0.0 - ROUND_OFF, 0.0 + ROUND_OFF, 0.4, 0.8};
    }
    else throw std::runtime_error("Lagrange basis: degree out of range");
    for (int i = 0; i < pmax + 1; ++i) {
      msh0[i] = (msh0[i] + 1) / 2.;
      msh1[i] = (msh1[i] + 1) / 2.;
    }
  } else if (pmax < 0 || pmax > 3) {
    throw std::runtime_error("Alpert basis: degree out of range");
  }
}

void Basis1D::decode(int order, int& level, int& suppt, int& p) const {
  const int s = order / (pmax + 1);
  p = order % (pmax + 1);
  level = level_of_slot(s);
  suppt = suppt_of_slot(s);
}

double Basis1D::val(int order, double x, int deriv, int sgn) const {
  int n, j, p;
  decode(order, n, j, p);
  double xlim = x;
  if (sgn == -1) {
    if (std::abs(x) <= ROUND_OFF) return 0.;
    xlim -= ROUND_OFF;
  } else if (sgn == 1) {
    if (std::abs(x - 1) <= ROUND_OFF) return 0.;
    xlim += ROUND_OFF;
  }
  if (kind == BasisKind::Alpert) {
    if (deriv == 0) return alpt_val0(xlim, n, j, p, pmax);
    if (deriv == 1) return alpt_val1(xlim, n, j, p, pmax);
    throw std::runtime_error("Alpert basis: derivative order > 1 is not provided");
  }
  if (kind == BasisKind::Hermite) {
    if (deriv > 1) throw std::runtime_error("Hermite basis: derivative order > 1 is not provided");
    return herm_val(xlim, n, j, p, deriv);
  }
  if (deriv != 0) throw std::runtime_error("Lagrange basis: derivatives are not provided");
  return lagr_val0(xlim, n, j, p, pmax);
}

double Basis1D::point(int order) const {
  int n, j, p;
  decode(order, n, j, p);
  if (n == 0) return msh0[p];
  double xl, xr;
  cell_support(n, j, xl, xr);
  const double h = xr - xl;
  return xl + h * msh1[p];
}

// ---------------------------------------------------------------------------

double product_volume(const Basis1D& U, int ru, const Basis1D& V, int cv, int du, int dv) {
  int nu, ju, pu, nv, jv, pv;
  U.decode(ru, nu, ju, pu);
  V.decode(cv, nv, jv, pv);
  double uxl, uxr, vxl, vxr;
  cell_support(nu, ju, uxl, uxr);
  cell_support(nv, jv, vxl, vxr);
  if ((vxl >= uxr) || (vxr <= uxl)) return 0.;
  double xl = uxl, xr = uxr;
  if (nu <= nv) { xl = vxl; xr = vxr; }
  const double xmid = (xl + xr) / 2.;
  auto f = [&](double x) { return U.val(ru, x, du) * V.val(cv, x, dv); };
  return gl10(f, xl, xmid) + gl10(f, xmid, xr);
}

double product_volume_interv(const Basis1D& B, int ru, int rv, int rw, int du, int dv, int nq) {
  int nu, ju, pu, nv, jv, pv, nw, jw, pw;
  B.decode(ru, nu, ju, pu);
  B.decode(rv, nv, jv, pv);
  B.decode(rw, nw, jw, pw);
  double uxl, uxr, vxl, vxr, wxl, wxr;
  cell_support(nu, ju, uxl, uxr);
  cell_support(nv, jv, vxl, vxr);
  cell_support(nw, jw, wxl, wxr);
  if ((vxl >= uxr) || (vxr <= uxl)) return 0.;
  if ((vxl >= wxr) || (vxr <= wxl)) return 0.;
  if ((uxl >= wxr) || (uxr <= wxl)) return 0.;
  const int max_level = std::max(std::max(nu, nv), nw);
  double xl = uxl, xr = uxr;
  if (nv == max_level) { xl = vxl; xr = vxr; }
  else if (nw == max_level) { xl = wxl; xr = wxr; }
  const double xmid = (xl + xr) / 2.;
  const std::vector<double>& x = kGLx.at(nq - 1);
  const std::vector<double>& w = kGLw.at(nq - 1);
  // Quad::GL_1D with nq points
  auto gl = [&](double tl, double tr) {
    double integral = 0;
    for (int i = 0; i < nq; ++i) {
      const double t = (x[i] - (-1.)) / (1. - (-1.)) * (tr - tl) + tl;
      integral += w[i] * (B.val(ru, t, du) * B.val(rv, t, dv));
    }
    integral *= (tr - tl) / 2.;
    return integral;
  };
  return gl(xl, xmid) + gl(xmid, xr);
}

double product_edge_dis_v_period(const Basis1D& U, int ru, const Basis1D& V, int cv, int slu, int slv) {
  int nv, jv, pv;
  V.decode(cv, nv, jv, pv);
  double vxl, vxr;
  cell_support(nv, jv, vxl, vxr);
  const double dis[3] = {vxl, (vxl + vxr) / 2, vxr};
  double prod_edge = 0.;
  for (int i = 0; i < 3; ++i) {
    if (nv <= 1 && i == 2) continue;  // 0 and 1 are the same point
    double val_u = U.val(ru, dis[i], 0, slu);
    double val_v = V.val(cv, dis[i], 0, slv);
    if (std::abs(dis[i] - 0.) < ROUND_OFF) {
      if (slu == -1) val_u = U.val(ru, 1., 0, slu);
      if (slv == -1) val_v = V.val(cv, 1., 0, slv);
    } else if (std::abs(dis[i] - 1.) < ROUND_OFF) {
      if (slu == 1) val_u = U.val(ru, 0., 0, slu);
      if (slv == 1) val_v = V.val(cv, 0., 0, slv);
    }
    prod_edge += val_u * val_v;
  }
  return prod_edge;
}

// Cells whose closed supports meet (nested, touching or touching across the
// periodic boundary).  Edge products of other pairs are exactly zero because
// every basis function evaluates to exactly 0 outside its support, so only
// these pairs are integrated.
static bool supports_meet(int n1, int j1, int n2, int j2) {
  double a0, a1, b0, b1;
  cell_support(n1, j1, a0, a1);
  cell_support(n2, j2, b0, b1);
  const bool intersect = !((a0 >= b1) || (a1 <= b0));
  const bool adjacent = (std::abs(a0 - b1) < ROUND_OFF) || (std::abs(a1 - b0) < ROUND_OFF);
  const bool periodic = ((std::abs(a0) < ROUND_OFF) && (std::abs(b1 - 1.0) < ROUND_OFF)) ||
                        ((std::abs(b0) < ROUND_OFF) && (std::abs(a1 - 1.0) < ROUND_OFF));
  return intersect || adjacent || periodic;
}

OperatorTables::OperatorTables(const Basis1D& U, const Basis1D& V, bool with_flux)
    : rows(U.size()), cols(V.size()) {
  const size_t n = static_cast<size_t>(rows) * cols;
  u_v.assign(n, 0.);
  u_vx.assign(n, 0.);
  if (with_flux) { ulft_vjp.assign(n, 0.); urgt_vjp.assign(n, 0.); ujp_vjp.assign(n, 0.); }
  const int su = U.pmax + 1, sv = V.pmax + 1;
  for (int r = 0; r < rows; ++r) {
    int nu, ju, pu;
    U.decode(r, nu, ju, pu);
    for (int c = 0; c < cols; ++c) {
      int nv, jv, pv;
      V.decode(c, nv, jv, pv);
      (void)su; (void)sv;
      if (!supports_meet(nu, ju, nv, jv)) continue;
      const size_t idx = static_cast<size_t>(r) * cols + c;
      u_v[idx] = product_volume(U, r, V, c, 0, 0);
      if (V.kind == BasisKind::Lagrange && U.kind == BasisKind::Lagrange) continue;
      u_vx[idx] = product_volume(U, r, V, c, 0, 1);
      if (!with_flux) continue;
      const double ur_vr = product_edge_dis_v_period(U, r, V, c, 1, 1);
      const double ur_vl = product_edge_dis_v_period(U, r, V, c, 1, -1);
      const double ul_vr = product_edge_dis_v_period(U, r, V, c, -1, 1);
      const double ul_vl = product_edge_dis_v_period(U, r, V, c, -1, -1);
      ulft_vjp[idx] = ul_vr - ul_vl;
      urgt_vjp[idx] = ur_vr - ur_vl;
      ujp_vjp[idx] = urgt_vjp[idx] - ulft_vjp[idx];
    }
  }
}

std::vector<double> alpt_at_lagr_points(const Basis1D& alpt, const Basis1D& lagr, int deriv) {
  const int na = alpt.size(), nl = lagr.size();
  std::vector<double> t(static_cast<size_t>(na) * nl, 0.);
  for (int m1 = 0; m1 < nl; ++m1) {
    const double pos = lagr.point(m1);
    for (int m2 = 0; m2 < na; ++m2) {
      int n, j, p;
      alpt.decode(m2, n, j, p);
      double xl, xr;
      cell_support(n, j, xl, xr);
      t[static_cast<size_t>(m2) * nl + m1] = (pos < xl || pos > xr) ? 0. : alpt.val(m2, pos, deriv, 0);
    }
  }
  return t;
}

std::vector<double> alpt_at_herm_points(const Basis1D& alpt, const Basis1D& herm) {
  const int na = alpt.size(), nh = herm.size();
  std::vector<double> t(static_cast<size_t>(na) * nh, 0.);
  for (int m1 = 0; m1 < nh; ++m1) {
    const double pos = herm.point(m1);
    const int m = Basis1D::herm_deriv(m1 % (herm.pmax + 1));
    for (int m2 = 0; m2 < na; ++m2) {
      int n, j, p;
      alpt.decode(m2, n, j, p);
      double xl, xr;
      cell_support(n, j, xl, xr);
      t[static_cast<size_t>(m2) * nh + m1] = (pos < xl || pos > xr) ? 0. : alpt.val(m2, pos, m, 0);
    }
  }
  return t;
}

std::vector<double> lagr_mass_diag(const Basis1D& lagr) {
  std::vector<double> d(lagr.size());
  for (int i = 0; i < lagr.size(); ++i) d[i] = product_volume(lagr, i, lagr, i, 0, 0);
  return d;
}

SurplusTables::SurplusTables(const Basis1D& lagr) {
  p1 = lagr.pmax + 1;
  nslots = 1 << lagr.nmax;
  anc_slot.assign(static_cast<size_t>(nslots) * p1, -1);
  anc_point.assign(static_cast<size_t>(nslots) * p1, -1);
  wt.assign(static_cast<size_t>(nslots) * p1 * p1, 0.);
  for (int s = 1; s < nslots; ++s) {
    const int l = level_of_slot(s), j = suppt_of_slot(s);
    double xleft, xright;
    cell_support(l, j, xleft, xright);
    std::vector<double> p_pos(p1);
    std::vector<int> p_slot(p1), p_num(p1);
    int ic = 0;
    // first P1 coarser-level points strictly inside the support, scanned by
    // level, then support index, then point index
    for (int k = 0; k < l; ++k) {
      if (ic == p1) continue;
      const int p2k = std::max(2, 1 << k);
      for (int i = 1; i < p2k; i += 2) {
        for (int q0 = 0; q0 < p1; ++q0) {
          const int order = slot_of(k, i) * p1 + q0;
          const double pos1 = lagr.point(order);
          if (pos1 > xleft && pos1 < xright) {
            if (ic == p1) throw std::runtime_error("surplus tables: more ancestor points than P+1");
            p_pos[ic] = pos1;
            p_slot[ic] = slot_of(k, i);
            p_num[ic] = q0;
            ++ic;
          }
        }
      }
    }
    if (ic != p1) throw std::runtime_error("surplus tables: fewer ancestor points than P+1");
    std::vector<double> sorted = p_pos;
    std::sort(sorted.begin(), sorted.end());
    for (int c = 0; c < p1; ++c) {
      anc_slot[static_cast<size_t>(s) * p1 + c] = p_slot[c];
      anc_point[static_cast<size_t>(s) * p1 + c] = p_num[c];
    }
    if (lagr.kind == BasisKind::Hermite) {
      // ancestors at the leftmost position take the level-0 left-point
      // degrees of freedom 0, 2, the others 1, 3, in scan order
      const double pl = sorted[0];
      for (int p0 = 0; p0 < p1; ++p0) {
        int i10 = 0, i20 = 1;
        for (int c = 0; c < p1; ++c) {
          int ic0;
          if (std::fabs(p_pos[c] - pl) < 1.0e-14) { ic0 = i10; i10 += 2; }
          else { ic0 = i20; i20 += 2; }
          const int l1 = Basis1D::herm_deriv(p0), l2 = Basis1D::herm_deriv(ic0);
          const double xt = lagr.msh1[p0];
          wt[(static_cast<size_t>(s) * p1 + p0) * p1 + c] =
              -1.0 * std::pow(2, (l - 1) * (l1 - l2)) * lagr.val(ic0, xt, l1, 0);
        }
      }
      continue;
    }
    for (int p0 = 0; p0 < p1; ++p0) {
      for (int c = 0; c < p1; ++c) {
        int ic0 = 0;
        for (int i2 = 0; i2 < p1; ++i2)
          if (std::fabs(p_pos[c] - sorted[i2]) < 1.0e-15) ic0 = i2;
        const double xt = lagr.msh1[p0];
        // level-0 Lagrange basis ic0 evaluated at the level-1 point
        wt[(static_cast<size_t>(s) * p1 + p0) * p1 + c] = -1.0 * lagr.val(ic0, xt, 0, 0);
      }
    }
  }
}

std::vector<double> project_1d(const Basis1D& B, const std::function<double(double)>& func) {
  // linspace(0., 1., 17) accumulates val += 1/16 (include/subs.h:62-73)
  std::vector<double> xint(17);
  {
    const double h = (1. - 0.) / static_cast<double>(17 - 1);
    double v = 0.;
    for (double& x : xint) { x = v; v += h; }
  }
  std::vector<double> out(B.size());
  for (int b = 0; b < B.size(); ++b) {
    int n, j, p;
    B.decode(b, n, j, p);
    auto fx = [&](const double& x) { return B.val(b, x, 0, 0) * func(x); };
    if (n <= 4) {
      double integral = 0.;
      for (int i = 0; i < 16; ++i) integral += gl10(fx, xint[i], xint[i + 1]);
      out[b] = integral;
    } else {
      double xl, xr;
      cell_support(n, j, xl, xr);
      const double xmid = (xl + xr) / 2.;
      out[b] = gl10(fx, xl, xmid) + gl10(fx, xmid, xr);
    }
  }
  return out;
}

std::vector<double> linear_operator_table(const OperatorTables& aa, double coef) {
  const std::vector<double>& flx = coef >= 0 ? aa.ulft_vjp : aa.urgt_vjp;
  std::vector<double> t(aa.u_vx.size());
  for (size_t i = 0; i < t.size(); ++i) {
    double a = coef * aa.u_vx[i];
    double b = coef * flx[i];
    if (!(std::abs(a) >= ROUND_OFF)) a = 0.;
    if (!(std::abs(b) >= ROUND_OFF)) b = 0.;
    t[i] = a + b;
  }
  return t;
}

// ---------------------------------------------------------------------------
// Hash key (source/Hash.cpp:55-167).

static double hk_factorial(double nValue) {
  double result = nValue;
  double result_next;
  double pc = nValue;
  if (nValue < 2) return 1.0;
  do {
    result_next = result * (pc - 1);
    result = result_next;
    pc--;
  } while (pc > 2);
  return result;
}

static double hk_bicoef(double n, double k) {
  if (k == 1) return n;
  return hk_factorial(n) / (hk_factorial(k) * hk_factorial(n - k));
}

HashKey::HashKey() {
  const int Nmax = 100;
  bicoef_.resize(Nmax);
  for (int i = 0; i < Nmax; ++i) {
    bicoef_[i].resize(i + 1);
    for (int j = 0; j <= i; ++j) {
      if (j <= i / 2) bicoef_[i][j] = hk_bicoef(i, j);
      else bicoef_[i][j] = bicoef_[i][i - j];
    }
  }
}

static int hk_pow(int x, int p) {
  if (p == 0) return 1;
  if (p == 1) return x;
  int tmp = hk_pow(x, p / 2);
  if (p % 2 == 0) return tmp * tmp;
  return x * tmp * tmp;
}

const std::vector<int>& HashKey::level_table(int d) const {
  if (d < 1 || d > 4) throw std::runtime_error("hash key table: dimension must be 1..4");
  std::call_once(level_once_[d], [&]() {
    std::vector<int>& tab = level_part_[d];
    tab.assign(static_cast<size_t>(1) << (4 * d), 0);
    std::vector<int> lv(d), ones(d, 1);
    for (size_t c = 0; c < tab.size(); ++c) {
      size_t r = c;
      for (int t = d - 1; t >= 0; --t) { lv[t] = static_cast<int>(r % 16); r /= 16; }
      // j = 1 in every dimension makes ind1 = 0: the key is the level part
      tab[c] = key_direct(lv.data(), ones.data(), d);
    }
  });
  return level_part_[d];
}

int HashKey::key(const int* l, const int* j, int d) const {
  if (d < 1 || d > 4) return key_direct(l, j, d);
  int code = 0;
  for (int t = 0; t < d; ++t) {
    if (l[t] < 0 || l[t] > 15) return key_direct(l, j, d);
    code = code * 16 + l[t];
  }
  const std::vector<int>& tab = level_table(d);
  int ind1 = 0;
  for (int t = 0; t < d; ++t) ind1 = ind1 * hk_pow(2, l[t]) + (j[t] - 1) / 2;
  return ind1 + tab[code];
}

int HashKey::key_direct(const int* l, const int* j, int d) const {
  int ind1 = 0, ind2 = 0, ind3 = 0;
  for (int t = 0; t < d; ++t) ind1 = ind1 * hk_pow(2, l[t]) + (j[t] - 1) / 2;
  int sum = l[0];
  for (int t = 1; t < d; ++t) {
    ind2 -= bicoef_[t + sum][t];  // int -= double: truncates toward zero
    sum += l[t];
    ind2 += bicoef_[t + sum][t];
  }
  ind2 *= hk_pow(2, sum);
  for (int s = 0; s < sum; ++s) ind3 += bicoef_[d - 1 + s][d - 1] * hk_pow(2, s);
  return ind1 + ind2 + ind3;
}

}  // namespace adg
