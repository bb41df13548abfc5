// erk_common.hpp -- device building blocks of the explicit Runge-Kutta path on gfx950.
//
// Tableaux, the PID / constant step-size controllers, end-clipping and the local
// interpolants, written as straight-line device code so that one lane can own one
// trajectory.  Reference behaviour followed (diffrax 0.7.2):
//   tableaux ............. _solver/tsit5.py:18-100, _solver/dopri5.py:11-33, _solver/heun.py:12-17
//   stage algebra ........ _solver/runge_kutta.py:847-1052 (k = control*f, z = y0 + sum a k)
//   Tsit5 dense output ... _solver/tsit5.py:116-156
//   quartic / Hermite .... _local_interpolation.py:50-139, _solver/dopri5.py:36-47
//   PID controller ....... _step_size_controller/pid.py:394-567
//   constant controller .. _step_size_controller/constant.py:30-104
//   _clip_to_end ......... _integrate.py:278-284
// The state type T is double (parity form) or float (mixed form); time, step size,
// the error norm and the controller are always double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dfx {

enum SolverId : int32_t { SOLVER_TSIT5 = 0, SOLVER_DOPRI5 = 1, SOLVER_HEUN = 2 };

// ----------------------------------------------------------------------------------
// Tableaux.  a is stored zero padded; zero entries are skipped at compile time so
// that no 0*k term enters a sum (keeps -0/inf/NaN behaviour of the written formula).
// ----------------------------------------------------------------------------------
struct Tsit5Tab {
  static constexpr int id = SOLVER_TSIT5;
  static constexpr int S = 7;
  static constexpr bool fsal = true, ssal = true;
  static constexpr int order = 5;
  static constexpr double b0 = 0.9646076681806522951816731316512876333711995238157997181903319145764851595234062815396e-1;
  static constexpr double b1 = 1.0 / 100.0;
  static constexpr double b2 = 0.4798896504144995747752495322905965199130404621990332488332634944254542060153074523509;
  static constexpr double b3 = 1.379008574103741893192274821856872770756462643091360525934940067397245698027561293331;
  static constexpr double b4 = -3.290069515436080679901047585711363850115683290894936158531296799594813811049925401677;
  static constexpr double b5 = 2.324710524099773982415355918398765796109060233222962411944060046314465391054716027841;
  static constexpr double a[7][7] = {
      {0, 0, 0, 0, 0, 0, 0},
      {161.0 / 1000.0, 0, 0, 0, 0, 0, 0},
      {-0.8480655492356988544426874250230774675121177393430391537369234245294192976164141156943e-2,
       0.3354806554923569885444268742502307746751211773934303915373692342452941929761641411569, 0, 0, 0, 0, 0},
      {2.897153057105493432130432594192938764924887287701866490314866693455023795137503079289,
       -6.359448489975074843148159912383825625952700647415626703305928850207288721235210244366,
       4.362295432869581411017727318190886861027813359713760212991062156752264926097707165077, 0, 0, 0, 0},
      {5.325864828439256604428877920840511317836476253097040101202360397727981648835607691791,
       -11.74888356406282787774717033978577296188744178259862899288666928009020615663593781589,
       7.495539342889836208304604784564358155658679161518186721010132816213648793440552049753,
       -0.9249506636175524925650207933207191611349983406029535244034750452930469056411389539635e-1, 0, 0, 0},
      {5.861455442946420028659251486982647890394337666164814434818157239052507339770711679748,
       -12.92096931784710929170611868178335939541780751955743459166312250439928519268343184452,
       8.159367898576158643180400794539253485181918321135053305748355423955009222648673734986,
       -0.7158497328140099722453054252582973869127213147363544882721139659546372402303777878835e-1,
       -0.2826905039406838290900305721271224146717633626879770007617876201276764571291579142206e-1, 0, 0},
      {b0, b1, b2, b3, b4, b5, 0}};
  static constexpr double b[7] = {b0, b1, b2, b3, b4, b5, 0.0};
  // b_error entries are the float64 differences b_sol - b_hat, as tsit5.py:75-98 forms them.
  static constexpr double berr[7] = {
      b0 - 0.9468075576583945807478876255758922856117527357724631226139574065785592789071067303271e-1,
      b1 - 0.9183565540343253096776363936645313759813746240984095238905939532922955247253608687270e-2,
      b2 - 0.4877705284247615707855642599631228241516691959761363774365216240304071651579571959813,
      b3 - 1.234297566930478985655109673884237654035539930748192848315425833500484878378061439761,
      b4 + 2.707712349983525454881109975059321670689605166938197378763992255714444407154902012702,
      b5 - 1.866628418170587035753719399566211498666255505244122593996591602841258328965767580089,
      -1.0 / 66.0};
  static constexpr double c[7] = {0.0, 161.0 / 1000.0, 327.0 / 1000.0, 9.0 / 10.0,
                                  0.9800255409045096857298102862870245954942137979563024768854764293221195950761080302604,
                                  1.0, 1.0};
};

struct Dopri5Tab {
  static constexpr int id = SOLVER_DOPRI5;
  static constexpr int S = 7;
  static constexpr bool fsal = true, ssal = true;
  static constexpr int order = 5;
  static constexpr double a[7][7] = {
      {0, 0, 0, 0, 0, 0, 0},
      {1.0 / 5.0, 0, 0, 0, 0, 0, 0},
      {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0, 0},
      {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0, 0},
      {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0, 0},
      {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0, 0},
      {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0}};
  static constexpr double b[7] = {35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0,
                                  -2187.0 / 6784.0, 11.0 / 84.0, 0.0};
  static constexpr double berr[7] = {35.0 / 384.0 - 1951.0 / 21600.0,
                                     0.0,
                                     500.0 / 1113.0 - 22642.0 / 50085.0,
                                     125.0 / 192.0 - 451.0 / 720.0,
                                     -2187.0 / 6784.0 - -12231.0 / 42400.0,
                                     11.0 / 84.0 - 649.0 / 6300.0,
                                     -1.0 / 60.0};
  static constexpr double c[7] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0, 1.0};
  // _Dopri5Interpolation.c_mid (dopri5.py:36-47): each entry is (p/q)/2 in float64.
  static constexpr double cmid[7] = {6025192743.0 / 30085553152.0 / 2.0,
                                     0.0,
                                     51252292925.0 / 65400821598.0 / 2.0,
                                     -2691868925.0 / 45128329728.0 / 2.0,
                                     187940372067.0 / 1594534317056.0 / 2.0,
                                     -1776094331.0 / 19743644256.0 / 2.0,
                                     11237099.0 / 235043384.0 / 2.0};
};

struct HeunTab {
  static constexpr int id = SOLVER_HEUN;
  static constexpr int S = 2;
  static constexpr bool fsal = false, ssal = false;
  static constexpr int order = 2;
  static constexpr double a[2][2] = {{0, 0}, {1.0, 0}};
  static constexpr double b[2] = {0.5, 0.5};
  static constexpr double berr[2] = {0.5, -0.5};
  static constexpr double c[2] = {0.0, 1.0};
};

// ----------------------------------------------------------------------------------
// jnp-style scalar helpers (NaN propagating max/min, as jnp.maximum / jnp.minimum).
// ----------------------------------------------------------------------------------
__host__ __device__ inline double jmax(double a, double b) {
  return (a != a || b != b) ? __builtin_nan("") : (a > b ? a : b);
}
__host__ __device__ inline double jmin(double a, double b) {
  return (a != a || b != b) ? __builtin_nan("") : (a < b ? a : b);
}
__host__ __device__ inline double jclip(double x, double lo, double hi) { return jmin(jmax(x, lo), hi); }
// The same results through the hardware max / min where the sign of a zero cannot differ
// (jmax of two values >= 0; jclip of a value >= 0 between bounds > 0): the NaN case is
// the only one these instructions treat differently, and it is taken first.
__device__ __forceinline__ double jmax_nonneg(double a, double b) {
  return (a != a || b != b) ? __builtin_nan("") : __builtin_fmax(a, b);
}
__device__ __forceinline__ double jclip_pos(double x, double lo, double hi) {
  return (x != x) ? __builtin_nan("") : __builtin_fmin(__builtin_fmax(x, lo), hi);
}

// Next float64 towards -inf (equinox.internal.prevbefore), by integer stepping of the
// IEEE bit pattern.
__host__ __device__ inline double prevbefore(double x) {
  if (x != x) return x;
  if (x == -__builtin_inf()) return x;
  if (x == 0.0) return -4.9406564584124654e-324;
  long long bits;
  __builtin_memcpy(&bits, &x, sizeof(bits));
  bits += (x > 0.0) ? -1 : 1;
  double r;
  __builtin_memcpy(&r, &bits, sizeof(r));
  return r;
}
// Next float64 towards +inf (equinox.internal.nextafter).
__host__ __device__ inline double nextup(double x) { return -prevbefore(-x); }

// ----------------------------------------------------------------------------------
// Correctly rounded x^c for the exponents the controller uses.
//
// PIDController raises 1/scaled_error to (i+p+d)/error_order = 0.2 (the float64 nearest
// 1/5) on every step, and _select_initial_step raises to 1/error_order.  The last bit
// of that power decides the next step size, so it is computed correctly rounded from
// pure IEEE operations (no libm): a fifth root from a float guess and one third-order
// step, a Newton correction on the double-double residual m - r^5, and the factor
// x^(0.2 - 1/5) = 1 + (0.2 - 1/5) ln x.  Any
// correctly rounded pow gives the same bits, on the host and on the device.
// Exponent 0.5 is sqrt (correctly rounded), exponent 1 is the identity, anything else
// goes to the math library.
// ----------------------------------------------------------------------------------
__host__ __device__ inline double bits_to_double(unsigned long long b) {
  double r;
  __builtin_memcpy(&r, &b, sizeof(r));
  return r;
}
__host__ __device__ inline unsigned long long double_to_bits(double x) {
  unsigned long long b;
  __builtin_memcpy(&b, &x, sizeof(b));
  return b;
}

__host__ __device__ inline double cr_pow_fifth(double x_in) {
#pragma clang fp contract(off)
  // Zero, negative, NaN and infinite inputs are resolved by one select at the end and the
  // arithmetic runs on a stand-in for them, so the device code has no branches.
  const bool finite_pos = (x_in > 0.0) && (x_in < __builtin_inf());
  const bool sub = x_in < 2.2250738585072014e-308;  // subnormal: scale by 2^100
  double x = finite_pos ? x_in : 1.0;
  x = sub ? x * 1.2676506002282294e30 : x;
  const int adj = sub ? -20 : 0;
  const unsigned long long bx = double_to_bits(x);
  const int E = (int)((bx >> 52) & 0x7ff) - 1023;
  const int k = (E >= 0) ? E / 5 : -((4 - E) / 5);
  const int j = E - 5 * k;  // 0..4
  const double f = bits_to_double((bx & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  // [1,2)
  const double m = f * (double)(1 << j);  // x = m * 2^(5k), m in [1, 32)
  // w ~ m^(-1/5): a float guess (relative error ~2^-22), then one third-order step
  // w (1 + e/5 + 3e^2/25) with e = 1 - m w^5, whose truncation error (11/125) e^3 ~ 2^-58
  // is below the step's own float64 rounding.
#if defined(__HIP_DEVICE_COMPILE__)
  double w = (double)__builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf((float)m));
#else
  double w = (double)exp2f(-0.2f * log2f((float)m));
#endif
  {
    const double w2 = w * w;
    const double w5 = w2 * w2 * w;
    const double e = 1.0 - m * w5;
    w = w + w * (e * (0.2 + e * 0.12));
  }
  const double w4 = (w * w) * (w * w);
  const double r = m * w4;  // m^(1/5) to ~1 ulp
  // double-double residual m - r^5, Newton correction delta = res / (5 r^4) ~ res * w^4 / 5
  const double h2 = r * r;
  const double l2 = fma(r, r, -h2);
  const double h4 = h2 * h2;
  const double l4 = fma(h2, h2, -h4) + 2.0 * h2 * l2;
  const double h5 = h4 * r;
  const double l5 = fma(h4, r, -h5) + l4 * r;
  const double res = (m - h5) - l5;
  const double delta = res * (w4 * 0.2);
  // ln of the unscaled input: ln(f) + (E + 5*adj) ln 2, ln(f) from the atanh series
  // (f folded into [sqrt(1/2), sqrt(2)) so the truncated series is good to ~1e-9)
  const bool hi = f > 1.4142135623730951;
  const double fl = hi ? f * 0.5 : f;
#if defined(__HIP_DEVICE_COMPILE__)
  // s enters only the correction term (0.2 - 1/5) ln x, 1e-17 of the result, so a
  // reciprocal refined once (relative error ~1e-16) replaces the correctly rounded division
  const double den = fl + 1.0;
  double rd = __builtin_amdgcn_rcp(den);
  rd = fma(rd, fma(-den, rd, 1.0), rd);
  const double s = (fl - 1.0) * rd;
#else
  const double s = (fl - 1.0) / (fl + 1.0);
#endif
  const double s2 = s * s;
  const double lnf = 2.0 * s * (1.0 + s2 * (1.0 / 3.0 + s2 * (0.2 + s2 * (1.0 / 7.0 + s2 * (1.0 / 9.0)))));
  const double lnx = lnf + (double)(E + (hi ? 1 : 0) + 5 * adj) * 0.6931471805599453;
  const double eps = 1.1102230246251565e-17;  // 0.2 (float64) - 1/5
  const double corr = r * (eps * lnx);
  const double out = r + (delta + corr);
  const double scale = bits_to_double((unsigned long long)(1023 + k + adj) << 52);
  const double special = (x_in == 0.0) ? 0.0 : ((x_in > 0.0) ? x_in : __builtin_nan(""));
  return finite_pos ? out * scale : special;
}

// Exponents other than 0.2 / 0.5 / 1 (non-default PID coefficients or error orders) use
// the math library's pow (float64, within an ulp or so of correct rounding).
template <bool Generic = true>
__host__ __device__ inline double pow_ctl(double x, double c) {
  if (c == 0.2) return cr_pow_fifth(x);
  if (c == 1.0) return x;
  if (c == 0.5) {
    if (x == 0.0) return 0.0;
    return (x < 0.0) ? __builtin_nan("") : sqrt(x);
  }
  if constexpr (Generic) return pow(x, c);
  return __builtin_nan("");
}

// rms_norm for D >= 2 with the division by c = sqrt(D) (rounded) written out: the float64
// division a / c is d = div_scale(c), n = div_scale(a), r = rcp(d) refined by two steps
// r += r (1 - d r), q = n r, q + r (n - d q) (div_fmas) and div_fixup.  Here r = rc is
// computed once from c (refined_rcp: the correctly rounded 1/c for D = 2..16), q = a rc,
// and q + rc (a - c q) is the correctly rounded a / c whenever the residual a - c q is
// exact, which holds for a >= 2^-969.  a = sqrt(s) of a float64 s is 0, +inf or in
// [2^-537, 2^512] (s >= 2^-1074 when nonzero), so every reachable a is covered; +inf is
// selected and NaN propagates.  The host build (tests) seeds the reciprocal with a float
// division instead of v_rcp_f64; the two refinement steps give the same r from either.
__host__ __device__ __forceinline__ double refined_rcp(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
#else
  double r = (double)(1.0f / (float)b);
#endif
  r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
  r = __builtin_fma(r, __builtin_fma(-b, r, 1.0), r);
  return r;
}
template <int D>
__host__ __device__ __forceinline__ double rms_norm_rc(const double (&e)[D], double c, double rc) {
  static_assert(D >= 2, "D = 1 is |e| (no division)");
  double s = e[0] * e[0];
#pragma unroll
  for (int j = 1; j < D; ++j) s = s + e[j] * e[j];
  const double a = sqrt(s);
  const double q = a * rc;
  const double out = __builtin_fma(__builtin_fma(-c, q, a), rc, q);
  return (a == __builtin_inf()) ? a : out;
}

// ----------------------------------------------------------------------------------
// Controller parameters (PIDController fields, pid.py:299-311) and state.
// ----------------------------------------------------------------------------------
struct PidParams {
  double rtol, atol;
  double coeff1, coeff2, coeff3;   // (i+p+d)/order, -(p+2d)/order, d/order
  double dtmin, dtmax;             // NaN = None
  double factormin, factormax, safety;
  int force_dtmin, has_dtmin, has_dtmax, pad_;
};

struct PidState {
  double inv, pinv;   // prev_inv_scaled_error, prev_prev_inv_scaled_error
  int at_dtmin;
};

// PIDController.adapt_step_size after the scaled error is known.  Returns keep and
// writes next (t0, t1), the new state and an error code (2 = dt_min_reached).
// The controller kinds:
//   PID_GENERIC  any coefficients (library pow for exponents other than 0.2 / 0.5 / 1);
//   PID_INTEGRAL the integral-only controller (pcoeff = dcoeff = 0, the default) with
//                exponent 0.2, 0.5 or 1 chosen at run time: factor2 = factor3 = 1;
//   PID_FIFTH    the integral-only controller with exponent 0.2 (the default for the
//                fifth-order embedded pairs): the correctly rounded fifth root inline,
//                with no run-time dispatch and no branches.
enum : int { PID_INTEGRAL = 0, PID_GENERIC = 1, PID_FIFTH = 2 };

template <int Ctl>
__device__ __forceinline__ bool pid_adapt(const PidParams& p, double scaled, double tprev,
                                          double tnext, PidState& st, double& ntprev,
                                          double& ntnext, int& code) {
  const double prev_dt = tnext - tprev;
  bool keep = scaled < 1.0;
  if (p.has_dtmin) keep = keep || st.at_dtmin;
  double inv = 1.0 / scaled;
  double f1, f2 = 1.0, f3 = 1.0;
  if constexpr (Ctl == PID_GENERIC) {
    f1 = (p.coeff1 == 0.0) ? 1.0 : pow_ctl<true>(inv, p.coeff1);
    f2 = (p.coeff2 == 0.0) ? 1.0 : pow_ctl<true>(st.inv, p.coeff2);
    f3 = (p.coeff3 == 0.0) ? 1.0 : pow_ctl<true>(st.pinv, p.coeff3);
  } else if constexpr (Ctl == PID_FIFTH) {
    f1 = cr_pow_fifth(inv);
  } else {
    f1 = pow_ctl<false>(inv, p.coeff1);
  }
  const double fmin = keep ? 1.0 : p.factormin;
  const double fmax = keep ? p.factormax : p.safety;
  const double factor = jclip_pos(p.safety * f1 * f2 * f3, fmin, fmax);
  double dt = prev_dt * factor;
  if (inv == 0.0 || __builtin_isinf(inv)) inv = 1.0;
  code = 0;
  if constexpr (Ctl == PID_FIFTH) {
    dt = jmin(dt, p.has_dtmax ? p.dtmax : __builtin_inf());  // jmin(dt, inf) is dt
  } else {
    if (p.has_dtmax) dt = jmin(dt, p.dtmax);
  }
  if (p.has_dtmin) {
    if (!p.force_dtmin && dt < p.dtmin) code = 2;
    if (st.at_dtmin && factor == 1.0) dt = p.dtmin;
    st.at_dtmin = dt <= p.dtmin;
    dt = jmax(dt, p.dtmin);
  }
  ntprev = keep ? tnext : tprev;
  ntnext = ntprev + dt;
  const double new_inv = keep ? inv : st.inv;
  const double new_pinv = keep ? st.inv : st.pinv;
  st.inv = new_inv;
  st.pinv = new_pinv;
  return keep;
}

// Whether the PID_INTEGRAL controller path reproduces these coefficients.
inline bool pid_is_integral(const PidParams& p) {
  return p.coeff2 == 0.0 && p.coeff3 == 0.0 &&
         (p.coeff1 == 0.2 || p.coeff1 == 0.5 || p.coeff1 == 1.0);
}
// Whether the PID_FIFTH controller path reproduces these coefficients.
inline bool pid_is_fifth(const PidParams& p) {
  return p.coeff2 == 0.0 && p.coeff3 == 0.0 && p.coeff1 == 0.2;
}

// optimistix.rms_norm restated: sqrt(sum x^2)/sqrt(n), |x| when n == 1.
template <int D>
__device__ __forceinline__ double rms_norm(const double (&e)[D]) {
  if constexpr (D == 1) {
    return fabs(e[0]);
  } else {
    double s = e[0] * e[0];
#pragma unroll
    for (int j = 1; j < D; ++j) s = s + e[j] * e[j];
    return sqrt(s) / sqrt((double)D);
  }
}

__device__ __forceinline__ double linear_rescale(double t0, double t, double t1) {
  const bool cond = (t0 == t1);
  const double num = cond ? 0.0 : t - t0;
  const double den = cond ? 1.0 : t1 - t0;
  return num / den;
}

// Tsit5 dense-output weights b_i(theta) (tsit5.py:116-156), same factorised form.
__device__ __forceinline__ void tsit5_weights(double th, double (&bb)[7]) {
  const double t2 = th * th;
  bb[0] = -1.0530884977290216 * th * (th - 1.3299890189751412) *
          (t2 - 1.4364028541716351 * th + 0.7139816917074209);
  bb[1] = 0.1017 * t2 * (t2 - 2.1966568338249754 * th + 1.2949852507374631);
  bb[2] = 2.490627285651252793 * t2 * (t2 - 2.38535645472061657 * th + 1.57803468208092486);
  bb[3] = -16.54810288924490272 * (th - 1.21712927295533244) * (th - 0.61620406037800089) * t2;
  bb[4] = 47.37952196281928122 * (th - 1.203071208372362603) * (th - 0.658047292653547382) * t2;
  bb[5] = -34.87065786149660974 * (th - 1.2) * (th - 0.666666666666666667) * t2;
  bb[6] = 2.5 * (th - 1.0) * (th - 0.6) * t2;
}

// d b_i / d theta of the weights above (product rule on the same factors).  The reference
// differentiates the interpolant with jax.jvp (AbstractPath.derivative); this is the same
// derivative in closed form.
__device__ __forceinline__ void tsit5_dweights(double th, double (&db)[7]) {
  const double t2 = th * th;
  {
    const double a = 1.3299890189751412, p = 1.4364028541716351, r = 0.7139816917074209;
    const double q = t2 - p * th + r;
    db[0] = -1.0530884977290216 * ((th - a) * q + th * q + th * (th - a) * (2.0 * th - p));
  }
  {
    const double p = 2.1966568338249754, r = 1.2949852507374631;
    db[1] = 0.1017 * (2.0 * th * (t2 - p * th + r) + t2 * (2.0 * th - p));
  }
  {
    const double p = 2.38535645472061657, r = 1.57803468208092486;
    db[2] = 2.490627285651252793 * (2.0 * th * (t2 - p * th + r) + t2 * (2.0 * th - p));
  }
  auto two_roots = [&](double c, double r1, double r2) {
    // d/dth [c (th - r1)(th - r2) th^2]
    return c * ((th - r2) * t2 + (th - r1) * t2 + 2.0 * th * (th - r1) * (th - r2));
  };
  db[3] = two_roots(-16.54810288924490272, 1.21712927295533244, 0.61620406037800089);
  db[4] = two_roots(47.37952196281928122, 1.203071208372362603, 0.658047292653547382);
  db[5] = two_roots(-34.87065786149660974, 1.2, 0.666666666666666667);
  db[6] = two_roots(2.5, 1.0, 0.6);
}

}  // namespace dfx
