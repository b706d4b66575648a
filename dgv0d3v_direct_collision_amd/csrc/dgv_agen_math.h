// dgv_agen_math.h -- per-pair values of the collision information operator A
// (host and device code).
//
// For a node pair (xi, xi1) and the Lagrange basis phi_l of the canonical cell K,
// SetA_DGV (dgvtools_mod.f90:377-681) computes
//   Atemp_l = int_0^pi int_0^2pi [phi_l(xi') + phi_l(xi1') - phi_l(xi) - phi_l(xi1)]
//             d eps  sin chi d chi,
//   A_l     = Atemp_l * w(xi1) * w(xi) * |g| / 8          (lines 654-659),
// with xi' = c + r k, xi1' = c - r k, c = (xi+xi1)/2, r = |xi-xi1|/2, k the unit
// vector of the angles (chi, eps) about g (EvalPhiPostCollIntegrand, lines
// 938-975).  The reference integrates with nested adaptive Simpson rules.
//
// Here the same integral is evaluated semi-analytically.  The measure
// sin chi d chi d eps is the surface measure of the unit sphere, so the gain
// term equals 2 * S_l with S_l = int_{S^2} phi_l(c + r k) dk, and the loss term
// is 4 pi (phi_l(xi) + phi_l(xi1)).  S_l is computed with the polar axis along w:
//   S_l = int sin(th) d th L_k(w^(th)) int_{arcs(th)} L_i(u^(eps)) L_j(v^(eps)) d eps,
// w = c_w + r cos(th), rho = r sin(th), u = c_u + rho cos eps, v = c_v + rho sin eps.
//  * The inner integral runs over the arcs of the circle of radius rho that lie
//    in [u0,u1]x[v0,v1]; the integrand is a polynomial in (cos eps, sin eps), so
//    it is integrated exactly with antiderivatives of cos^k sin^n.
//  * The outer integral is split at every polar angle where the arc structure
//    changes (circle tangent to a face line, circle through a corner, the
//    w-faces).  Inside a piece the integrand is analytic in th; at its ends it
//    has at worst square-root behaviour, which the substitution
//    th = ta + (tb-ta)(1-cos s)/2 removes, so Gauss-Legendre in s converges
//    exponentially.
#pragma once
#include <cmath>

#ifndef DGV_HD
#define DGV_HD __host__ __device__
#endif

namespace dgv {

constexpr int kAgenMaxS = 4;
constexpr int kAgenMaxQ = 48;

// Everything the per-pair evaluation needs, precomputed on the host.
struct AgenConsts {
  int su, sv, sw;
  double u0, u1, v0, v1, w0, w1;       // canonical cell
  double kap[3][kAgenMaxS];            // Gauss nodes on [-1,1] per dimension
  double mono[3][kAgenMaxS][kAgenMaxS];// monomial coefficients of L_i(y) = sum_m mono[i][m] y^m
  int nq;                              // Gauss points per smooth piece
  double qs[kAgenMaxQ], qw[kAgenMaxQ]; // Gauss-Legendre on [0, pi] in s
  double Trad;
  int moment_mode;
};

// Lagrange basis function i on nodes kap[0..n-1] (lagrbasfun, dgvtools_mod.f90:31-56).
DGV_HD inline double lagr(int i, double x, const double* kap, int n) {
  double y = 1.0;
  for (int j = 0; j < n; ++j)
    if (j != i) y = y * (kap[j] - x) / (kap[j] - kap[i]);
  return y;
}

// Value of the canonical basis functions at a point (the restored
// EvalLagrBasisFunByNdsDGblzmEZ: zero outside the closed cell).
DGV_HD inline double basis_at(const AgenConsts& C, int iu, int iv, int iw, double u, double v, double w) {
  if (u < C.u0 || C.u1 < u || v < C.v0 || C.v1 < v || w < C.w0 || C.w1 < w) return 0.0;
  const double un = (u - (C.u1 + C.u0) / 2.0) / (C.u1 - C.u0) * 2.0;
  const double vn = (v - (C.v1 + C.v0) / 2.0) / (C.v1 - C.v0) * 2.0;
  const double wn = (w - (C.w1 + C.w0) / 2.0) / (C.w1 - C.w0) * 2.0;
  return lagr(iu, un, C.kap[0], C.su) * lagr(iv, vn, C.kap[1], C.sv) * lagr(iw, wn, C.kap[2], C.sw);
}

// Antiderivatives F_{k,n}(theta) of cos^k sin^n for k, n < kAgenMaxS, by the
// reduction formulas
//   F_{k,n} = -s^{n-1} c^{k+1}/(k+n) + (n-1)/(k+n) F_{k,n-2}     (n >= 2)
//   F_{k,n} =  c^{k-1} s^{n+1}/(k+n) + (k-1)/(k+n) F_{k-2,n}     (k >= 2, n < 2).
DGV_HD inline void antideriv(double th, double c, double s, int ku, int kv, double F[kAgenMaxS][kAgenMaxS]) {
  double cp[kAgenMaxS + 2], sp[kAgenMaxS + 2];
  cp[0] = 1.0;
  sp[0] = 1.0;
  for (int i = 1; i < kAgenMaxS + 2; ++i) { cp[i] = cp[i - 1] * c; sp[i] = sp[i - 1] * s; }
  for (int n = 0; n < kv; ++n)
    for (int k = 0; k < ku; ++k) {
      double v;
      if (n >= 2) {
        v = -sp[n - 1] * cp[k + 1] / (k + n) + (double)(n - 1) / (k + n) * F[k][n - 2];
      } else if (k >= 2) {
        v = cp[k - 1] * sp[n + 1] / (k + n) + (double)(k - 1) / (k + n) * F[k - 2][n];
      } else if (k == 0 && n == 0) {
        v = th;
      } else if (k == 1 && n == 0) {
        v = s;
      } else if (k == 0 && n == 1) {
        v = -c;
      } else {  // k == 1, n == 1
        v = s * s / 2.0;
      }
      F[k][n] = v;
    }
}

DGV_HD inline void sort_small(double* a, int n) {
  for (int i = 1; i < n; ++i) {
    const double x = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > x) { a[j + 1] = a[j]; --j; }
    a[j + 1] = x;
  }
}

// I[k][n] = integral of cos^k sin^n over the arcs of the circle
// (cu + rho cos e, cv + rho sin e) lying inside [u0,u1]x[v0,v1].
template <int SU, int SV>
DGV_HD inline void arc_moments(const AgenConsts& C, double cu, double cv, double rho,
                               double I[kAgenMaxS][kAgenMaxS]) {
  constexpr int ku = SU, kv = SV;
  for (int k = 0; k < ku; ++k)
    for (int n = 0; n < kv; ++n) I[k][n] = 0.0;
  const double twopi = 6.283185307179586476925286766559;
  double ang[9];
  int na = 0;
  const double du[2] = {C.u0 - cu, C.u1 - cu}, dv[2] = {C.v0 - cv, C.v1 - cv};
  for (int e = 0; e < 2; ++e) {
    const double X = du[e] / rho;
    if (X > -1.0 && X < 1.0) {
      const double a = acos(X);
      ang[na++] = a;
      ang[na++] = twopi - a;
    }
    const double Y = dv[e] / rho;
    if (Y > -1.0 && Y < 1.0) {
      double a = asin(Y);
      if (a < 0) a += twopi;
      ang[na++] = a;
      double b = 3.141592653589793238462643383279502884 - asin(Y);
      ang[na++] = b;
    }
  }
  auto inside = [&](double e) {
    const double x = cu + rho * cos(e), y = cv + rho * sin(e);
    return x >= C.u0 && x <= C.u1 && y >= C.v0 && y <= C.v1;
  };
  double Fa[kAgenMaxS][kAgenMaxS], Fb[kAgenMaxS][kAgenMaxS];
  if (na == 0) {
    if (!inside(0.0)) return;
    antideriv(0.0, 1.0, 0.0, ku, kv, Fa);
    antideriv(twopi, 1.0, 0.0, ku, kv, Fb);
    for (int k = 0; k < ku; ++k)
      for (int n = 0; n < kv; ++n) I[k][n] = Fb[k][n] - Fa[k][n];
    return;
  }
  sort_small(ang, na);
  ang[na] = ang[0] + twopi;
  for (int i = 0; i < na; ++i) {
    const double a = ang[i], b = ang[i + 1];
    if (!(b > a)) continue;
    if (!inside(0.5 * (a + b))) continue;
    antideriv(a, cos(a), sin(a), ku, kv, Fa);
    antideriv(b, cos(b), sin(b), ku, kv, Fb);
    for (int k = 0; k < ku; ++k)
      for (int n = 0; n < kv; ++n) I[k][n] += Fb[k][n] - Fa[k][n];
  }
}

// Coefficients c[i][k] of L_i(alpha + beta x) = sum_k c[i][k] x^k.
template <int s>
DGV_HD inline void shifted_coeffs(const double mono[kAgenMaxS][kAgenMaxS], double alpha, double beta,
                                  double c[kAgenMaxS][kAgenMaxS]) {
  // binomial expansion of (alpha + beta x)^m
  double ap[kAgenMaxS], bp[kAgenMaxS];
  ap[0] = 1.0;
  bp[0] = 1.0;
  for (int m = 1; m < s; ++m) { ap[m] = ap[m - 1] * alpha; bp[m] = bp[m - 1] * beta; }
  for (int i = 0; i < s; ++i)
    for (int k = 0; k < s; ++k) {
      double acc = 0.0;
      for (int m = k; m < s; ++m) {
        double binom = 1.0;
        for (int t = 0; t < k; ++t) binom = binom * (m - t) / (t + 1);
        acc += mono[i][m] * binom * ap[m - k];
      }
      c[i][k] = acc * bp[k];
    }
}

// S[l] = int_{S^2} phi_l(c + r k) dk for the su*sv*sw canonical basis functions.
template <int su, int sv, int sw>
DGV_HD inline void sphere_integrals(const AgenConsts& C, const double cen[3], double r, double* S) {
  constexpr int D = su * sv * sw;
  const double pi = 3.141592653589793238462643383279502884;
  for (int l = 0; l < D; ++l) S[l] = 0.0;
  // polar angle theta about the w axis: w = c_w + r cos(theta), rho = r sin(theta)
  // is analytic in theta (no branch points at the poles)
  const double ca = (C.w1 - cen[2]) / r, cb = (C.w0 - cen[2]) / r;
  const double th_lo = ca >= 1.0 ? 0.0 : (ca <= -1.0 ? pi : acos(ca));
  const double th_hi = cb >= 1.0 ? 0.0 : (cb <= -1.0 ? pi : acos(cb));
  if (!(th_hi > th_lo)) return;
  // breakpoints: tangency to the four face lines and passage through the four
  // corners of the cell's u-v rectangle, at rho = d, i.e. theta = asin(d/r)
  // and pi - asin(d/r)
  double bp[20];
  int nb = 0;
  bp[nb++] = th_lo;
  bp[nb++] = th_hi;
  const double du[2] = {C.u0 - cen[0], C.u1 - cen[0]}, dv[2] = {C.v0 - cen[1], C.v1 - cen[1]};
  auto add = [&](double d) {
    if (d < r) {
      const double t = asin(d / r);
      if (t > th_lo && t < th_hi) bp[nb++] = t;
      if (pi - t > th_lo && pi - t < th_hi) bp[nb++] = pi - t;
    }
  };
  for (int e = 0; e < 2; ++e) { add(fabs(du[e])); add(fabs(dv[e])); }
  for (int e = 0; e < 2; ++e)
    for (int f = 0; f < 2; ++f) add(sqrt(du[e] * du[e] + dv[f] * dv[f]));
  sort_small(bp, nb);
  const double hu = 0.5 * (C.u1 - C.u0), hv = 0.5 * (C.v1 - C.v0), hw = 0.5 * (C.w1 - C.w0);
  const double mu = 0.5 * (C.u1 + C.u0), mv = 0.5 * (C.v1 + C.v0), mw = 0.5 * (C.w1 + C.w0);
  const double au = (cen[0] - mu) / hu, av = (cen[1] - mv) / hv;
  double I[kAgenMaxS][kAgenMaxS], cU[kAgenMaxS][kAgenMaxS], cV[kAgenMaxS][kAgenMaxS];
  for (int p = 0; p + 1 < nb; ++p) {
    const double ta = bp[p], tb = bp[p + 1];
    if (!(tb > ta)) continue;
    const double half = 0.5 * (tb - ta);
    for (int q = 0; q < C.nq; ++q) {
      const double sq = C.qs[q];
      const double th = ta + half * (1.0 - cos(sq));
      const double sth = sin(th);
      const double wt = C.qw[q] * half * sin(sq) * sth;   // dt = sin(theta) dtheta
      const double rho = r * sth;
      if (!(rho > 0.0)) continue;
      arc_moments<su, sv>(C, cen[0], cen[1], rho, I);
      shifted_coeffs<su>(C.mono[0], au, rho / hu, cU);
      shifted_coeffs<sv>(C.mono[1], av, rho / hv, cV);
      const double wn = (cen[2] + r * cos(th) - mw) / hw;
      double Lw[kAgenMaxS];
      for (int k = 0; k < sw; ++k) Lw[k] = wt * lagr(k, wn, C.kap[2], sw);
      for (int i = 0; i < su; ++i)
        for (int j = 0; j < sv; ++j) {
          double E = 0.0;
          for (int k = 0; k < su; ++k) {
            double row = 0.0;
            for (int n = 0; n < sv; ++n) row += cV[j][n] * I[k][n];
            E += cU[i][k] * row;
          }
          for (int k = 0; k < sw; ++k) S[(i * sv + j) * sw + k] += Lw[k] * E;
        }
    }
  }
}

// Raw values Atemp_l (before min_sens and the weights) of one pair, or -1 if
// the pair fails SetA_DGV's prefilter (dgvtools_mod.f90:527-530).
template <int SU, int SV, int SW>
DGV_HD inline int pair_raw(const AgenConsts& C, const double x[3], const double y[3], double* raw, double* dsph_out) {
  const double fourpi = 12.566370614359172953850573533118;
  const double d0 = x[0] - y[0], d1 = x[1] - y[1], d2 = x[2] - y[2];
  const double dsph = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  const double s0 = x[0] + y[0] - C.u0 - C.u1, s1 = x[1] + y[1] - C.v0 - C.v1, s2 = x[2] + y[2] - C.w0 - C.w1;
  const double dist2 = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
  const double e0 = C.u1 - C.u0, e1 = C.v1 - C.v0, e2 = C.w1 - C.w0;
  const double dphi = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  *dsph_out = dsph;
  constexpr int D = SU * SV * SW;
  if ((dist2 > dsph + dphi) || (dist2 + dphi < dsph) || (dsph > C.Trad)) return -1;
  const double cen[3] = {0.5 * (x[0] + y[0]), 0.5 * (x[1] + y[1]), 0.5 * (x[2] + y[2])};
  const double r = 0.5 * dsph;
  if (C.moment_mode != 0) {
    // moment test functions psi = 1,u,v,w,|u|^2 over the whole sphere
    // (dgvtools_mod.f90:267-274): closed forms of int psi(c + r k) dk
    double S, px, py;
    switch (C.moment_mode) {
      case 1: S = fourpi; px = 1.0; py = 1.0; break;
      case 2: S = fourpi * cen[0]; px = x[0]; py = y[0]; break;
      case 3: S = fourpi * cen[1]; px = x[1]; py = y[1]; break;
      case 4: S = fourpi * cen[2]; px = x[2]; py = y[2]; break;
      default:
        S = fourpi * (cen[0] * cen[0] + cen[1] * cen[1] + cen[2] * cen[2] + r * r);
        px = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        py = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
        break;
    }
    for (int l = 0; l < D; ++l) raw[l] = 2.0 * S - fourpi * (px + py);
    return 0;
  }
  sphere_integrals<SU, SV, SW>(C, cen, r, raw);
  for (int iu = 0; iu < SU; ++iu)
    for (int iv = 0; iv < SV; ++iv)
      for (int iw = 0; iw < SW; ++iw) {
        const int l = (iu * SV + iv) * SW + iw;
        const double loss = basis_at(C, iu, iv, iw, x[0], x[1], x[2]) + basis_at(C, iu, iv, iw, y[0], y[1], y[2]);
        raw[l] = 2.0 * raw[l] - fourpi * loss;
      }
  return 0;
}

}  // namespace dgv
