// ALP E-values and bit scores (DIAMOND's finite-size-corrected statistics).
//
// Follows src/alp.rs:60-237 and src/matrix.rs:182-200: the same f64
// operations in the same order, so host results equal the reference's
// bit for bit (both call the platform libm for log/exp). The device build is
// used only as a conservative pre-filter; every E-value that decides or is
// printed is recomputed on the host.
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define DMR_HD __host__ __device__
#else
#define DMR_HD
#endif

namespace dmr {

// isfinite for both host and device builds (false for NaN and +-inf).
DMR_HD inline bool alp_finite(double x) { return fabs(x) <= 1.7976931348623157e308; }

struct AlpParams {
  double lambda, k, ln_k;
  double a, b, alpha, beta, a_j, b_j, alpha_j, beta_j, sigma, tau, vi_thr, vj_thr, c_thr;
};

DMR_HD inline AlpParams alp_make(double lambda, double k, double alpha, double alpha_v, double sigma,
                                 int gap_open, int gap_extend, double u_alpha, double u_alpha_v) {
  const double nat_cut_off = 2.0;
  AlpParams p;
  double g = (double)(gap_open + gap_extend);
  p.lambda = lambda;
  p.k = k;
  p.ln_k = log(k);
  p.a = alpha;
  p.b = 2.0 * g * (u_alpha - alpha);
  p.alpha = alpha_v;
  p.beta = 2.0 * g * (u_alpha_v - alpha_v);
  p.a_j = p.a;
  p.b_j = p.b;
  p.alpha_j = p.alpha;
  p.beta_j = p.beta;
  p.sigma = sigma;
  p.tau = 2.0 * g * (u_alpha_v - sigma);
  double v = nat_cut_off * alpha_v / lambda;
  p.vi_thr = v > 0.0 ? v : 0.0;
  p.vj_thr = p.vi_thr;
  double c = nat_cut_off * sigma / lambda;
  p.c_thr = c > 0.0 ? c : 0.0;
  return p;
}

// GSL Chebyshev evaluation without the error estimate.
DMR_HD inline double alp_cheb(const double* c, int len, double x) {
  double y2 = 2.0 * x, d = 0.0, dd = 0.0;
  for (int j = len - 1; j >= 1; --j) {
    double t = d;
    d = y2 * d - dd + c[j];
    dd = t;
  }
  return x * d - dd + 0.5 * c[0];
}

DMR_HD inline double alp_erfc8_sum(double x) {
  const double P[6] = {2.97886562639399288862, 7.409740605964741794425, 6.1602098531096305440906,
                       5.019049726784267463450058, 1.275366644729965952479585264,
                       0.5641895835477550741253201704};
  const double Q[7] = {3.3690752069827527677, 9.608965327192787870698, 17.08144074746600431571095,
                       12.0489519278551290360340491, 9.396034016235054150430579648,
                       2.260528520767326969591866945, 1.0};
  double num = P[5];
  for (int i = 4; i >= 0; --i) num = x * num + P[i];
  double den = Q[6];
  for (int i = 5; i >= 0; --i) den = x * den + Q[i];
  return num / den;
}

DMR_HD inline double alp_erfc(double x) {
  const double xlt1[20] = {
      1.06073416421769980345174155056,   -0.42582445804381043569204735291,
      0.04955262679620434040357683080,   0.00449293488768382749558001242,
      -0.00129194104658496953494224761,  -0.00001836389292149396270416979,
      0.00002211114704099526291538556,   -5.23337485234257134673693179020e-7,
      -2.78184788833537885382530989578e-7, 1.41158092748813114560316684249e-8,
      2.72571296330561699984539141865e-9, -2.06343904872070629406401492476e-10,
      -2.14273991996785367924201401812e-11, 2.22990255539358204580285098119e-12,
      1.36250074650698280575807934155e-13, -1.95144010922293091898995913038e-14,
      -6.85627169231704599442806370690e-16, 1.44506492869699938239521607493e-16,
      2.45935306460536488037576200030e-18, -9.29599561220523396007359328540e-19};
  const double x15[25] = {
      0.44045832024338111077637466616,   -0.143958836762168335790826895326,
      0.044786499817939267247056666937,  -0.013343124200271211203618353102,
      0.003824682739750469767692372556,  -0.001058699227195126547306482530,
      0.000283859419210073742736310108,  -0.000073906170662206760483959432,
      0.000018725312521489179015872934,  -4.62530981164919445131297264430e-6,
      1.11558657244432857487884006422e-6, -2.63098662650834130067808832725e-7,
      6.07462122724551777372119408710e-8, -1.37460865539865444777251011793e-8,
      3.05157051905475145520096717210e-9, -6.65174789720310713757307724790e-10,
      1.42483346273207784489792999706e-10, -3.00141127395323902092018744545e-11,
      6.22171792645348091472914001250e-12, -1.26994639225668496876152836555e-12,
      2.55385883033257575402681845385e-13, -5.06258237507038698392265499770e-14,
      9.89705409478327321641264227110e-15, -1.90685978789192181051961024995e-15,
      3.50826648032737849245113757340e-16};
  const double x510[20] = {
      1.11684990123545698684297865808,   0.003736240359381998520654927536,
      -0.000916623948045470238763619870, 0.000199094325044940833965078819,
      -0.000040276384918650072591781859, 7.76515264697061049477127605790e-6,
      -1.44464794206689070402099225301e-6, 2.61311930343463958393485241947e-7,
      -4.61833026634844152345304095560e-8, 8.00253111512943601598732144340e-9,
      -1.36291114862793031395712122089e-9, 2.28570483090160869607683087722e-10,
      -3.78022521563251805044056974560e-11, 6.17253683874528285729910462130e-12,
      -9.96019290955316888445830597430e-13, 1.58953143706980770269506726000e-13,
      -2.51045971047162509999527428316e-14, 3.92607828989125810013581287560e-15,
      -6.07970619384160374392535453420e-16, 9.12600607264794717315507477670e-17};
  double ax = fabs(x), v;
  if (ax <= 1.0) v = alp_cheb(xlt1, 20, 2.0 * ax - 1.0);
  else if (ax <= 5.0) v = exp(-x * x) * alp_cheb(x15, 25, 0.5 * (ax - 3.0));
  else if (ax < 10.0) v = exp(-x * x) / ax * alp_cheb(x510, 20, (2.0 * ax - 15.0) / 5.0);
  else v = alp_erfc8_sum(ax) * exp(-x * x);
  return x < 0.0 ? 2.0 - v : v;
}

DMR_HD inline double alp_log_erfc(double x) {
  const double ln2 = 0.6931471805599453;
  const double root6_eps = 2.4607833005759251e-3;
  const double sqrt_pi = 1.772453850905516;
  const double pi = 3.141592653589793;
  if (x < -8.0) return ln2;
  if (x * x < 10.0 * root6_eps) {
    double y = x / sqrt_pi;
    const double c3 = (4.0 - pi) / 3.0, c4 = 2.0 * (1.0 - pi / 3.0);
    const double c5 = -0.001829764677455021, c6 = 0.02629651521057465, c7 = -0.01621575378835404;
    const double c8 = 0.00125993961762116, c9 = 0.00556964649138, c10 = -0.0045563339802;
    const double c11 = 0.0009461589032, c12 = 0.0013200243174, c13 = -0.00142906, c14 = 0.00048204;
    double s = c8 + y * (c9 + y * (c10 + y * (c11 + y * (c12 + y * (c13 + c14 * y)))));
    s = y * (1.0 + y * (1.0 + y * (c3 + y * (c4 + y * (c5 + y * (c6 + y * (c7 + y * s)))))));
    return -2.0 * s;
  }
  if (x > 8.0) return log(alp_erfc8_sum(x)) - x * x;
  return log(alp_erfc(x));
}

DMR_HD inline double alp_ln_or_ninf(double x) { return x > 0.0 ? log(x) : -INFINITY; }

DMR_HD inline double alp_log_sum(double a, double b) {
  double hi = a >= b ? a : b, lo = a >= b ? b : a;
  if (!alp_finite(hi)) return hi;
  if (lo == -INFINITY) return hi;
  return hi + log(1.0 + exp(lo - hi));
}

DMR_HD inline double alp_log_diff(double a, double b) {
  if (a < b) return -INFINITY;
  if (!alp_finite(a)) return a;
  if (b == -INFINITY) return a;
  double t = 1.0 - exp(b - a);
  return t <= 0.0 ? -INFINITY : a + log(t);
}

// One side (subject: m, or query: n) of the ALP area.
DMR_HD inline void alp_side(double len, double a, double b, double alpha, double beta, double thr,
                            double y, double* log_p_side, double* log_p) {
  const double inv_sqrt_2pi = 0.3989422804014327;
  double l_y = len - (a * y + b);
  double v = alpha * y + beta;
  double v_y = thr > v ? thr : v;
  double sq = sqrt(v_y);
  double f = sq == 0.0 ? 1e100 : l_y / sq;
  *log_p = log(0.5) + alp_log_erfc(-sqrt(0.5) * f);
  double log_e = log(inv_sqrt_2pi) - 0.5 * f * f;
  double log_sqrt_e = alp_ln_or_ninf(sq) + log_e;
  *log_p_side = l_y < 0.0 ? alp_log_diff(log_sqrt_e, alp_ln_or_ninf(-l_y) + *log_p)
                          : alp_log_sum(log_sqrt_e, alp_ln_or_ninf(l_y) + *log_p);
}

DMR_HD inline double alp_log_area(const AlpParams& p, double y, double m, double n) {
  double log_p1, log_p_m, log_p2, log_p_n;
  alp_side(m, p.a, p.b, p.alpha, p.beta, p.vi_thr, y, &log_p1, &log_p_m);
  alp_side(n, p.a_j, p.b_j, p.alpha_j, p.beta_j, p.vj_thr, y, &log_p2, &log_p_n);
  double c = p.sigma * y + p.tau;
  double c_y = p.c_thr > c ? p.c_thr : c;
  return alp_log_sum(log_p1 + log_p2, alp_ln_or_ninf(c_y) + log_p_m + log_p_n);
}

// E-value of raw score `raw` for full query/subject lengths and DB letters.
DMR_HD inline double alp_evalue(const AlpParams& p, int raw, uint64_t qlen, uint64_t slen,
                                uint64_t db_letters) {
  double y = (double)raw;
  double m = (double)(slen > 1 ? slen : 1);
  double n = (double)(qlen > 1 ? qlen : 1);
  double la = alp_log_area(p, y, m, n);
  if (!alp_finite(la)) {
    double ex = -p.lambda * y + log(p.k);
    return ex > 700.0 ? INFINITY : (double)qlen * (double)db_letters * exp(ex);
  }
  double log_e = la + log(p.k) - p.lambda * y;
  return log_e > 700.0 ? INFINITY : exp(log_e) * (double)db_letters / m;
}

DMR_HD inline double bit_score(const AlpParams& p, int raw) {
  const double ln2 = 0.6931471805599453;
  return (p.lambda * (double)raw - p.ln_k) / ln2;
}

DMR_HD inline int raw_score(const AlpParams& p, double bits) {
  const double ln2 = 0.6931471805599453;
  return (int)ceil((bits * ln2 + p.ln_k) / p.lambda);
}

}  // namespace dmr
