/*
 * mmd_oracle.c -- CPU FP64 restatement of jax_qc's McMurchie-Davidson
 * integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP kernels
 * are compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never links,
 * loads or falls back to it.
 *
 * Every routine restates one reference function (cited file:line, paths
 * relative to the reference repository root) with the same loop structure
 * and operation order, so that with boys_mode == ORC_BOYS_REFERENCE the
 * numbers follow the reference's own arithmetic.  boys_mode ==
 * ORC_BOYS_ACCURATE swaps in a uniformly accurate Boys function (series in
 * long double / upward recursion where that recursion is stable) because the
 * reference's upward recursion loses all accuracy for n >= 6 just above
 * t = 0.01 (jax_qc/integrals/boys.py:101-109).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_BOYS_REFERENCE 0
#define ORC_BOYS_ACCURATE 1

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* Boys function                                                       */
/* ------------------------------------------------------------------ */

/* jax_qc/integrals/boys.py:34-49 (_f0_series) */
static double f0_series(double t) {
  return 1.0 - t / 3.0 + t * t / 10.0 - pow(t, 3) / 42.0 + pow(t, 4) / 216.0 -
         pow(t, 5) / 1320.0 + pow(t, 6) / 9360.0 - pow(t, 7) / 75600.0;
}

/* jax_qc/integrals/boys.py:52-55 (_f0_erf) */
static double f0_erf(double t) {
  double safe_t = t > 0 ? t : 1.0;
  return 0.5 * sqrt(M_PI / safe_t) * erf(sqrt(safe_t));
}

/* jax_qc/integrals/boys.py:58-65 (boys_f0) */
static double boys_f0_ref(double t) { return t < 1e-2 ? f0_series(t) : f0_erf(t); }

/* jax_qc/integrals/boys.py:68-83 (_fn_series) */
static double fn_series(int n, double t) {
  double result = 1.0 / (2 * n + 1);
  double sign = -1.0, t_pow = t, fact = 1.0;
  for (int k = 1; k < 8; ++k) {
    fact *= k;
    result = result + sign * t_pow / (fact * (2 * n + 2 * k + 1));
    sign = -sign;
    t_pow = t_pow * t;
  }
  return result;
}

/* jax_qc/integrals/boys.py:86-109 (boys_fn): Taylor series below 0.01,
 * otherwise F_0 from erf followed by the upward recurrence. */
double orc_boys_ref(int n, double t) {
  if (n == 0) return boys_f0_ref(t);
  double safe_t = t > 0 ? t : 1.0;
  double exp_neg_t = exp(-t);
  double fn = boys_f0_ref(t);
  for (int k = 0; k < n; ++k) {
    double next_fn = ((2 * k + 1) * fn - exp_neg_t) / (2.0 * safe_t);
    double next_at_zero = 1.0 / (2.0 * (k + 1) + 1.0);
    fn = t > 0 ? next_fn : next_at_zero;
  }
  return t < 1e-2 ? fn_series(n, t) : fn;
}

/* Accurate F_n(t): for t <= 35 the everywhere-convergent series
 *   F_n(t) = exp(-t) sum_k (2t)^k / ((2n+1)(2n+3)...(2n+2k+1))
 * (all terms positive, evaluated in long double); for t > 35 the upward
 * recurrence from the erf closed form, which is stable while t > n. */
double orc_boys_acc(int n, double td) {
  long double t = td < 0 ? 0.0L : (long double)td;
  if (t <= 35.0L || n > 30) {
    long double term = 1.0L / (2 * n + 1), sum = term;
    for (int k = 1; k < 2000; ++k) {
      term *= 2.0L * t / (long double)(2 * n + 2 * k + 1);
      sum += term;
      if (term < sum * 1e-22L) break;
    }
    return (double)(expl(-t) * sum);
  }
  long double f = 0.5L * sqrtl(3.14159265358979323846264338327950288L / t) * erfl(sqrtl(t));
  long double et = expl(-t);
  for (int k = 0; k < n; ++k) f = ((2 * k + 1) * f - et) / (2.0L * t);
  return (double)f;
}

static double boys(int mode, int n, double t) {
  return mode == ORC_BOYS_REFERENCE ? orc_boys_ref(n, t) : orc_boys_acc(n, t);
}

void orc_boys_array(int mode, int n, const double* t, double* out, int64_t cnt) {
  for (int64_t i = 0; i < cnt; ++i) out[i] = boys(mode, n, t[i]);
}

/* ------------------------------------------------------------------ */
/* Cartesian components and Hermite expansion                          */
/* ------------------------------------------------------------------ */

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

/* jax_qc/integrals/mcmurchie_davidson.py:39-50: lx descending, then ly
 * descending, lz = l - lx - ly. */
static void cart_comps(int l, int (*out)[3]) {
  int k = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      out[k][0] = lx;
      out[k][1] = ly;
      out[k][2] = l - lx - ly;
      ++k;
    }
}

/* E[i][j][t] stored flat with strides (lb+1)*(la+lb+1), (la+lb+1), 1.
 * jax_qc/integrals/mcmurchie_davidson.py:60-108 (hermite_e). */
#define EIDX(i, j, t) (((i) * (lb + 1) + (j)) * (la + lb + 1) + (t))
void orc_hermite_e(int la, int lb, double Xpa, double Xpb, double p, double* E) {
  int T = la + lb + 1;
  memset(E, 0, sizeof(double) * (la + 1) * (lb + 1) * T);
  E[EIDX(0, 0, 0)] = 1.0;
  double inv_2p = 0.5 / p;
  for (int i = 0; i < la; ++i) {
    for (int t = 0; t < i + 1; ++t) {
      double e_tm1 = t - 1 >= 0 ? E[EIDX(i, 0, t - 1)] : 0.0;
      double e_t = E[EIDX(i, 0, t)];
      double e_tp1 = t + 1 <= i + 1 ? E[EIDX(i, 0, t + 1)] : 0.0;
      E[EIDX(i + 1, 0, t)] = inv_2p * e_tm1 + Xpa * e_t + (t + 1) * e_tp1;
    }
    E[EIDX(i + 1, 0, i + 1)] = inv_2p * E[EIDX(i, 0, i)];
  }
  for (int j = 0; j < lb; ++j) {
    for (int i = 0; i < la + 1; ++i) {
      for (int t = 0; t < i + j + 1; ++t) {
        double e_tm1 = t - 1 >= 0 ? E[EIDX(i, j, t - 1)] : 0.0;
        double e_t = E[EIDX(i, j, t)];
        double e_tp1 = t + 1 <= i + j ? E[EIDX(i, j, t + 1)] : 0.0;
        E[EIDX(i, j + 1, t)] = inv_2p * e_tm1 + Xpb * e_t + (t + 1) * e_tp1;
      }
      E[EIDX(i, j + 1, i + j + 1)] = inv_2p * E[EIDX(i, j, i + j)];
    }
  }
}
#undef EIDX

/* R[t][u][v][n] for t<=tmax, u<=umax, v<=vmax, n<=nmax; returned flat with
 * strides (umax+1)(vmax+1)(nmax+1), (vmax+1)(nmax+1), (nmax+1), 1.
 * jax_qc/integrals/mcmurchie_davidson.py:113-172 (hermite_r_aux): the box is
 * grown in the same order (t, then u, then v) through n up to
 * L_max = tmax+umax+vmax+nmax. */
void orc_hermite_r(int mode, int tmax, int umax, int vmax, int nmax, double p, const double* PC,
                   double* out) {
  int Lmax = tmax + umax + vmax + nmax;
  int Nn = Lmax + 1;
  int su = (vmax + 1) * Nn, st = (umax + 1) * su, sv = Nn;
  double* R = (double*)calloc((size_t)(tmax + 1) * st, sizeof(double));
#define RI(t, u, v, n) R[(t) * st + (u) * su + (v) * sv + (n)]
  double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
  double minus_2p = -2.0 * p, pw = 1.0;
  for (int n = 0; n <= Lmax; ++n) {
    RI(0, 0, 0, n) = pw * boys(mode, n, T);
    pw *= minus_2p;
  }
  double X = PC[0], Y = PC[1], Z = PC[2];
  for (int t = 0; t < tmax; ++t)
    for (int n = 0; n < Lmax - t; ++n) {
      double r_tm1 = t - 1 >= 0 ? RI(t - 1, 0, 0, n + 1) : 0.0;
      RI(t + 1, 0, 0, n) = t * r_tm1 + X * RI(t, 0, 0, n + 1);
    }
  for (int u = 0; u < umax; ++u)
    for (int t = 0; t < tmax + 1; ++t)
      for (int n = 0; n < Lmax - t - u; ++n) {
        double r_um1 = u - 1 >= 0 ? RI(t, u - 1, 0, n + 1) : 0.0;
        RI(t, u + 1, 0, n) = u * r_um1 + Y * RI(t, u, 0, n + 1);
      }
  for (int v = 0; v < vmax; ++v)
    for (int u = 0; u < umax + 1; ++u)
      for (int t = 0; t < tmax + 1; ++t)
        for (int n = 0; n < Lmax - t - u - v; ++n) {
          double r_vm1 = v - 1 >= 0 ? RI(t, u, v - 1, n + 1) : 0.0;
          RI(t, u, v + 1, n) = v * r_vm1 + Z * RI(t, u, v, n + 1);
        }
  int No = nmax + 1;
  for (int t = 0; t <= tmax; ++t)
    for (int u = 0; u <= umax; ++u)
      for (int v = 0; v <= vmax; ++v)
        for (int n = 0; n <= nmax; ++n)
          out[((t * (umax + 1) + u) * (vmax + 1) + v) * No + n] = RI(t, u, v, n);
#undef RI
  free(R);
}

/* ------------------------------------------------------------------ */
/* Primitive blocks (jax_qc/integrals/obara_saika.py)                   */
/* ------------------------------------------------------------------ */

/* obara_saika.py:56-61 */
static double k_ab(double a, double b, const double* A, const double* B) {
  double p = a + b, mu = a * b / p;
  double d0 = A[0] - B[0], d1 = A[1] - B[1], d2 = A[2] - B[2];
  return exp(-mu * (d0 * d0 + d1 * d1 + d2 * d2));
}

/* obara_saika.py:64-65 */
static void p_center(double a, double b, const double* A, const double* B, double* P) {
  for (int k = 0; k < 3; ++k) P[k] = (a * A[k] + b * B[k]) / (a + b);
}

/* 1-D overlap S[i][j] = sqrt(pi/p) E[i][j][0] for i<=la+ext, j<=lb+ext
 * (obara_saika.py:70-87). */
static void overlap_1d(int la, int lb, double Xpa, double Xpb, double p, double* S) {
  int T = la + lb + 1;
  double* E = (double*)malloc(sizeof(double) * (la + 1) * (lb + 1) * T);
  orc_hermite_e(la, lb, Xpa, Xpb, p, E);
  double f = sqrt(M_PI / p);
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) S[i * (lb + 1) + j] = f * E[(i * (lb + 1) + j) * T];
  free(E);
}

/* obara_saika.py:90-112 */
void orc_prim_overlap(double a, const double* A, int la, double b, const double* B, int lb,
                      double* out) {
  double p = a + b, P[3];
  p_center(a, b, A, B, P);
  int n1 = (la + 1) * (lb + 1);
  double *Sx = malloc(sizeof(double) * n1), *Sy = malloc(sizeof(double) * n1),
         *Sz = malloc(sizeof(double) * n1);
  overlap_1d(la, lb, P[0] - A[0], P[0] - B[0], p, Sx);
  overlap_1d(la, lb, P[1] - A[1], P[1] - B[1], p, Sy);
  overlap_1d(la, lb, P[2] - A[2], P[2] - B[2], p, Sz);
  double K = k_ab(a, b, A, B);
  int ca[64][3], cb[64][3];
  cart_comps(la, ca);
  cart_comps(lb, cb);
  int na = ncart(la), nb = ncart(lb);
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j)
      out[i * nb + j] = K * Sx[ca[i][0] * (lb + 1) + cb[j][0]] * Sy[ca[i][1] * (lb + 1) + cb[j][1]] *
                        Sz[ca[i][2] * (lb + 1) + cb[j][2]];
  free(Sx);
  free(Sy);
  free(Sz);
}

/* obara_saika.py:115-138 (_kinetic_1d) */
static void kinetic_1d(int la, int lb, double Xpa, double Xpb, double a, double b, double p,
                       double* Tm) {
  int w = lb + 2;
  double* S = malloc(sizeof(double) * (la + 2) * (lb + 2));
  overlap_1d(la + 1, lb + 1, Xpa, Xpb, p, S);
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      double t_val = 4.0 * a * b * S[(i + 1) * w + j + 1];
      if (i > 0 && j > 0) t_val += i * j * S[(i - 1) * w + j - 1];
      if (j > 0) t_val -= 2.0 * a * j * S[(i + 1) * w + j - 1];
      if (i > 0) t_val -= 2.0 * b * i * S[(i - 1) * w + j + 1];
      Tm[i * (lb + 1) + j] = 0.5 * t_val;
    }
  free(S);
}

/* obara_saika.py:141-174 */
void orc_prim_kinetic(double a, const double* A, int la, double b, const double* B, int lb,
                      double* out) {
  double p = a + b, P[3];
  p_center(a, b, A, B, P);
  double K = k_ab(a, b, A, B);
  double X[3][2];
  for (int k = 0; k < 3; ++k) {
    X[k][0] = P[k] - A[k];
    X[k][1] = P[k] - B[k];
  }
  int we = lb + 2, w = lb + 1;
  double* Se[3];
  double* Tk[3];
  for (int k = 0; k < 3; ++k) {
    Se[k] = malloc(sizeof(double) * (la + 2) * (lb + 2));
    Tk[k] = malloc(sizeof(double) * (la + 1) * (lb + 1));
    overlap_1d(la + 1, lb + 1, X[k][0], X[k][1], p, Se[k]);
    kinetic_1d(la, lb, X[k][0], X[k][1], a, b, p, Tk[k]);
  }
  int ca[64][3], cb[64][3];
  cart_comps(la, ca);
  cart_comps(lb, cb);
  int na = ncart(la), nb = ncart(lb);
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j) {
      int ax = ca[i][0], ay = ca[i][1], az = ca[i][2], bx = cb[j][0], by = cb[j][1], bz = cb[j][2];
      double val = Tk[0][ax * w + bx] * Se[1][ay * we + by] * Se[2][az * we + bz] +
                   Se[0][ax * we + bx] * Tk[1][ay * w + by] * Se[2][az * we + bz] +
                   Se[0][ax * we + bx] * Se[1][ay * we + by] * Tk[2][az * w + bz];
      out[i * nb + j] = K * val;
    }
  for (int k = 0; k < 3; ++k) {
    free(Se[k]);
    free(Tk[k]);
  }
}

/* obara_saika.py:177-222 */
void orc_prim_nuclear(int mode, double a, const double* A, int la, double b, const double* B,
                      int lb, int natom, const double* C, const double* Z, double* out) {
  double p = a + b, P[3];
  p_center(a, b, A, B, P);
  double K = k_ab(a, b, A, B);
  int T = la + lb + 1, L = la + lb;
  double *Ex = malloc(sizeof(double) * (la + 1) * (lb + 1) * T),
         *Ey = malloc(sizeof(double) * (la + 1) * (lb + 1) * T),
         *Ez = malloc(sizeof(double) * (la + 1) * (lb + 1) * T);
  orc_hermite_e(la, lb, P[0] - A[0], P[0] - B[0], p, Ex);
  orc_hermite_e(la, lb, P[1] - A[1], P[1] - B[1], p, Ey);
  orc_hermite_e(la, lb, P[2] - A[2], P[2] - B[2], p, Ez);
  int ca[64][3], cb[64][3];
  cart_comps(la, ca);
  cart_comps(lb, cb);
  int na = ncart(la), nb = ncart(lb);
  for (int i = 0; i < na * nb; ++i) out[i] = 0.0;
  double two_pi_over_p = 2.0 * M_PI / p;
  double* R = malloc(sizeof(double) * (L + 1) * (L + 1) * (L + 1));
  for (int c = 0; c < natom; ++c) {
    double PC[3] = {P[0] - C[3 * c], P[1] - C[3 * c + 1], P[2] - C[3 * c + 2]};
    orc_hermite_r(mode, L, L, L, 0, p, PC, R);
    double prefac = -Z[c] * two_pi_over_p * K;
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < nb; ++j) {
        int ax = ca[i][0], ay = ca[i][1], az = ca[i][2], bx = cb[j][0], by = cb[j][1],
            bz = cb[j][2];
        const double* ex = Ex + (ax * (lb + 1) + bx) * T;
        const double* ey = Ey + (ay * (lb + 1) + by) * T;
        const double* ez = Ez + (az * (lb + 1) + bz) * T;
        double acc = 0.0;
        for (int t = 0; t < ax + bx + 1; ++t) {
          if (ex[t] == 0.0) continue;
          for (int u = 0; u < ay + by + 1; ++u) {
            if (ey[u] == 0.0) continue;
            for (int v = 0; v < az + bz + 1; ++v)
              acc += ex[t] * ey[u] * ez[v] * R[(t * (L + 1) + u) * (L + 1) + v];
          }
        }
        out[i * nb + j] += prefac * acc;
      }
  }
  free(R);
  free(Ex);
  free(Ey);
  free(Ez);
}

/* obara_saika.py:227-332 */
void orc_prim_eri(int mode, double a, const double* A, int la, double b, const double* B, int lb,
                  double c, const double* C, int lc, double d, const double* D, int ld,
                  double* out) {
  double p = a + b, q = c + d, rho = p * q / (p + q);
  double P[3], Q[3];
  p_center(a, b, A, B, P);
  p_center(c, d, C, D, Q);
  double Kab = k_ab(a, b, A, B), Kcd = k_ab(c, d, C, D);
  int Tab = la + lb + 1, Tcd = lc + ld + 1;
  double* Eb[3];
  double* Ek[3];
  for (int k = 0; k < 3; ++k) {
    Eb[k] = malloc(sizeof(double) * (la + 1) * (lb + 1) * Tab);
    Ek[k] = malloc(sizeof(double) * (lc + 1) * (ld + 1) * Tcd);
    orc_hermite_e(la, lb, P[k] - A[k], P[k] - B[k], p, Eb[k]);
    orc_hermite_e(lc, ld, Q[k] - C[k], Q[k] - D[k], q, Ek[k]);
  }
  double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
  int L = la + lb + lc + ld, W = L + 1;
  double* R = malloc(sizeof(double) * W * W * W);
  orc_hermite_r(mode, L, L, L, 0, rho, PQ, R);
  double prefac = 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(p + q)) * Kab * Kcd;
  int ca[64][3], cb[64][3], cc[64][3], cd[64][3];
  cart_comps(la, ca);
  cart_comps(lb, cb);
  cart_comps(lc, cc);
  cart_comps(ld, cd);
  int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j) {
      const double* exab = Eb[0] + (ca[i][0] * (lb + 1) + cb[j][0]) * Tab;
      const double* eyab = Eb[1] + (ca[i][1] * (lb + 1) + cb[j][1]) * Tab;
      const double* ezab = Eb[2] + (ca[i][2] * (lb + 1) + cb[j][2]) * Tab;
      int tx = ca[i][0] + cb[j][0], ty = ca[i][1] + cb[j][1], tz = ca[i][2] + cb[j][2];
      for (int k = 0; k < nc; ++k)
        for (int l = 0; l < nd; ++l) {
          const double* excd = Ek[0] + (cc[k][0] * (ld + 1) + cd[l][0]) * Tcd;
          const double* eycd = Ek[1] + (cc[k][1] * (ld + 1) + cd[l][1]) * Tcd;
          const double* ezcd = Ek[2] + (cc[k][2] * (ld + 1) + cd[l][2]) * Tcd;
          int sx = cc[k][0] + cd[l][0], sy = cc[k][1] + cd[l][1], sz = cc[k][2] + cd[l][2];
          double acc = 0.0;
          for (int t = 0; t <= tx; ++t) {
            double e_t = exab[t];
            if (e_t == 0.0) continue;
            for (int u = 0; u <= ty; ++u) {
              double e_u = eyab[u];
              if (e_u == 0.0) continue;
              for (int v = 0; v <= tz; ++v) {
                double e_v = ezab[v];
                if (e_v == 0.0) continue;
                for (int tp = 0; tp <= sx; ++tp) {
                  double f_tp = excd[tp];
                  if (f_tp == 0.0) continue;
                  double sign_t = (tp & 1) ? -1.0 : 1.0;
                  for (int up = 0; up <= sy; ++up) {
                    double f_up = eycd[up];
                    if (f_up == 0.0) continue;
                    double sign_tu = (up & 1) ? -sign_t : sign_t;
                    for (int vp = 0; vp <= sz; ++vp) {
                      double f_vp = ezcd[vp];
                      if (f_vp == 0.0) continue;
                      double sign = (vp & 1) ? -sign_tu : sign_tu;
                      acc += sign * e_t * e_u * e_v * f_tp * f_up * f_vp *
                             R[((t + tp) * W + (u + up)) * W + (v + vp)];
                    }
                  }
                }
              }
            }
          }
          out[((i * nb + j) * nc + k) * nd + l] = prefac * acc;
        }
    }
  free(R);
  for (int k = 0; k < 3; ++k) {
    free(Eb[k]);
    free(Ek[k]);
  }
}

/* ------------------------------------------------------------------ */
/* Basis description passed from Python                                */
/* ------------------------------------------------------------------ */

typedef struct {
  int nshell;
  const int* l;          /* [nshell] */
  const int* nprim;      /* [nshell] */
  const int* prim_off;   /* [nshell] offset into exps/coefs */
  const double* exps;    /* primitive exponents */
  const double* coefs;   /* normalized contraction coefficients */
  const double* centers; /* [nshell*3] Bohr */
} orc_basis;

/* obara_saika.py:394-426 (contracted_eri_block) */
static void contracted_eri(int mode, const orc_basis* bs, int i, int j, int k, int l,
                           double* out) {
  int la = bs->l[i], lb = bs->l[j], lc = bs->l[k], ld = bs->l[l];
  int n = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
  double* blk = malloc(sizeof(double) * n);
  for (int x = 0; x < n; ++x) out[x] = 0.0;
  for (int a = 0; a < bs->nprim[i]; ++a)
    for (int b = 0; b < bs->nprim[j]; ++b)
      for (int c = 0; c < bs->nprim[k]; ++c)
        for (int d = 0; d < bs->nprim[l]; ++d) {
          double w = bs->coefs[bs->prim_off[i] + a] * bs->coefs[bs->prim_off[j] + b] *
                     bs->coefs[bs->prim_off[k] + c] * bs->coefs[bs->prim_off[l] + d];
          orc_prim_eri(mode, bs->exps[bs->prim_off[i] + a], bs->centers + 3 * i, la,
                       bs->exps[bs->prim_off[j] + b], bs->centers + 3 * j, lb,
                       bs->exps[bs->prim_off[k] + c], bs->centers + 3 * k, lc,
                       bs->exps[bs->prim_off[l] + d], bs->centers + 3 * l, ld, blk);
          for (int x = 0; x < n; ++x) out[x] += w * blk[x];
        }
  free(blk);
}

void orc_contracted_eri(int mode, int nshell, const int* l, const int* nprim, const int* prim_off,
                        const double* exps, const double* coefs, const double* centers, int i,
                        int j, int k, int ll, double* out) {
  orc_basis bs = {nshell, l, nprim, prim_off, exps, coefs, centers};
  contracted_eri(mode, &bs, i, j, k, ll, out);
}

/* contracted_eri over a list of nq shell quartets (quartets[4q..4q+3]),
 * threads sharing the list; returns the sum of every Cartesian element (a
 * checksum, so the work cannot be elided).  Used to time the restatement of
 * the reference's quartet loop (jax_qc/integrals/eri.py:89-104) on a sample
 * of a workload's quartets. */
double orc_eri_quartets(int mode, int nshell, const int* l, const int* nprim, const int* prim_off,
                        const double* exps, const double* coefs, const double* centers,
                        int64_t nq, const int* quartets) {
  orc_basis bs = {nshell, l, nprim, prim_off, exps, coefs, centers};
  double total = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
  for (int64_t q = 0; q < nq; ++q) {
    double blk[1296];
    const int* s = quartets + 4 * q;
    contracted_eri(mode, &bs, s[0], s[1], s[2], s[3], blk);
    const int n = ncart(l[s[0]]) * ncart(l[s[1]]) * ncart(l[s[2]]) * ncart(l[s[3]]);
    for (int x = 0; x < n; ++x) total += blk[x];
  }
  return total;
}

/* jax_qc/integrals/spherical.py:34-67: rows follow libcint order (d_xy,
 * d_yz, d_z2, d_xz, d_x2-y2); l = 1 is the identity. */
static void cart_to_sph(int l, double* Cm) {
  if (l == 0) {
    Cm[0] = 1.0;
  } else if (l == 1) {
    for (int i = 0; i < 9; ++i) Cm[i] = 0.0;
    Cm[0] = Cm[4] = Cm[8] = 1.0;
  } else {
    double s3 = sqrt(3.0);
    double M[5][6] = {{0, s3, 0, 0, 0, 0},
                      {0, 0, 0, 0, s3, 0},
                      {-0.5, 0, 0, -0.5, 0, 1.0},
                      {0, 0, s3, 0, 0, 0},
                      {s3 / 2.0, 0, 0, -s3 / 2.0, 0, 0}};
    for (int r = 0; r < 5; ++r)
      for (int c = 0; c < 6; ++c) Cm[r * 6 + c] = M[r][c];
  }
}

static int nfunc(int l, int sph) { return sph ? 2 * l + 1 : ncart(l); }

/* Transform index `axis` of a 4-index block of dims d[4] by C (nout x din). */
static void transform_axis(double* blk, int* d, int axis, const double* Cm, int nout) {
  int pre = 1, post = 1;
  for (int k = 0; k < axis; ++k) pre *= d[k];
  for (int k = axis + 1; k < 4; ++k) post *= d[k];
  int din = d[axis];
  double* tmp = calloc((size_t)pre * nout * post, sizeof(double));
  for (int a = 0; a < pre; ++a)
    for (int o = 0; o < nout; ++o)
      for (int i = 0; i < din; ++i) {
        double c = Cm[o * din + i];
        if (c == 0.0) continue;
        for (int b = 0; b < post; ++b)
          tmp[(a * nout + o) * post + b] += c * blk[(a * din + i) * post + b];
      }
  memcpy(blk, tmp, sizeof(double) * pre * nout * post);
  d[axis] = nout;
  free(tmp);
}

static int shell_offsets(const orc_basis* bs, int sph, int* off) {
  off[0] = 0;
  for (int s = 0; s < bs->nshell; ++s) off[s + 1] = off[s] + nfunc(bs->l[s], sph);
  return off[bs->nshell];
}

/* jax_qc/integrals/eri.py:69-138 (compute_eri_tensor, _scatter_8fold):
 * canonical quartets i>=j, k>=l, ij>=kl, optional spherical transform,
 * scatter to all 8 permutations of a dense (n,n,n,n) tensor.  With shell_atom
 * non-NULL only the canonical quartets that hold a shell on atom `atom` are
 * computed (the rest of the tensor is zero): those are the only elements
 * that move when that atom is displaced, so a finite difference of two such
 * tensors equals the difference of two full rebuilds (the reference's
 * jax_qc/grad/rhf_grad.py:153-209) with the untouched elements' exact zero
 * differences left out. */
static void eri_tensor_impl(int mode, int nshell, const int* l, const int* nprim,
                            const int* prim_off, const double* exps, const double* coefs,
                            const double* centers, int sph, double* eri, const int* shell_atom,
                            int atom) {
  orc_basis bs = {nshell, l, nprim, prim_off, exps, coefs, centers};
  int* off = malloc(sizeof(int) * (nshell + 1));
  int64_t n = shell_offsets(&bs, sph, off);
  memset(eri, 0, sizeof(double) * n * n * n * n);
  /* canonical quartets write disjoint sets of elements, so bra pairs are
   * shared out over threads (the loop order within a pair is unchanged) */
  const int npair = nshell * (nshell + 1) / 2;
#pragma omp parallel for schedule(dynamic, 1)
  for (int ij = npair - 1; ij >= 0; --ij) {
      int i = 0;
      while ((i + 1) * (i + 2) / 2 <= ij) ++i;
      int j = ij - i * (i + 1) / 2;
      double blk[1296];
      double Cm[4][30];
      for (int k = 0; k < nshell; ++k)
        for (int ll = 0; ll <= k; ++ll) {
          int kl = k * (k + 1) / 2 + ll;
          if (ij < kl) continue;
          if (shell_atom && shell_atom[i] != atom && shell_atom[j] != atom &&
              shell_atom[k] != atom && shell_atom[ll] != atom)
            continue;
          contracted_eri(mode, &bs, i, j, k, ll, blk);
          int d[4] = {ncart(l[i]), ncart(l[j]), ncart(l[k]), ncart(l[ll])};
          if (sph) {
            int sh[4] = {i, j, k, ll};
            for (int ax = 0; ax < 4; ++ax) {
              cart_to_sph(l[sh[ax]], Cm[ax]);
              transform_axis(blk, d, ax, Cm[ax], 2 * l[sh[ax]] + 1);
            }
          }
          int o[4] = {off[i], off[j], off[k], off[ll]};
          for (int a = 0; a < d[0]; ++a)
            for (int b = 0; b < d[1]; ++b)
              for (int c = 0; c < d[2]; ++c)
                for (int e = 0; e < d[3]; ++e) {
                  double v = blk[((a * d[1] + b) * d[2] + c) * d[3] + e];
                  int64_t I = o[0] + a, J = o[1] + b, K = o[2] + c, Lx = o[3] + e;
#define PUT(w, x, y, z) eri[(((w) * n + (x)) * n + (y)) * n + (z)] = v
                  PUT(I, J, K, Lx);
                  PUT(J, I, K, Lx);
                  PUT(I, J, Lx, K);
                  PUT(J, I, Lx, K);
                  PUT(K, Lx, I, J);
                  PUT(K, Lx, J, I);
                  PUT(Lx, K, I, J);
                  PUT(Lx, K, J, I);
#undef PUT
                }
        }
  }
  free(off);
}

void orc_eri_tensor(int mode, int nshell, const int* l, const int* nprim, const int* prim_off,
                    const double* exps, const double* coefs, const double* centers, int sph,
                    double* eri) {
  eri_tensor_impl(mode, nshell, l, nprim, prim_off, exps, coefs, centers, sph, eri, NULL, -1);
}

void orc_eri_tensor_atom(int mode, int nshell, const int* l, const int* nprim,
                         const int* prim_off, const double* exps, const double* coefs,
                         const double* centers, int sph, double* eri, const int* shell_atom,
                         int atom) {
  eri_tensor_impl(mode, nshell, l, nprim, prim_off, exps, coefs, centers, sph, eri, shell_atom,
                  atom);
}

/* Shell-pair Schwarz bound Q_ij = sqrt(max_ab (ab|ab)) over Cartesian
 * diagonal components (jax_qc/integrals/screening.py:32-58). */
void orc_schwarz(int mode, int nshell, const int* l, const int* nprim, const int* prim_off,
                 const double* exps, const double* coefs, const double* centers, double* Q) {
  orc_basis bs = {nshell, l, nprim, prim_off, exps, coefs, centers};
  double* blk = malloc(sizeof(double) * 1296);
  for (int i = 0; i < nshell; ++i)
    for (int j = i; j < nshell; ++j) {
      contracted_eri(mode, &bs, i, j, i, j, blk);
      int na = ncart(l[i]), nb = ncart(l[j]);
      double mx = 0.0;
      for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
          double v = blk[((a * nb + b) * na + a) * nb + b];
          if (v > mx) mx = v;
        }
      Q[i * nshell + j] = Q[j * nshell + i] = sqrt(mx > 0 ? mx : 0.0);
    }
  free(blk);
}

/* One-electron matrices: upper triangle of shell blocks then mirrored
 * (overlap.py:67-85, kinetic.py:42-59, nuclear.py:62-81). which: 0=S, 1=T,
 * 2=V. */
void orc_one_electron(int mode, int which, int nshell, const int* l, const int* nprim,
                      const int* prim_off, const double* exps, const double* coefs,
                      const double* centers, int sph, int natom, const double* C, const double* Z,
                      double* M) {
  orc_basis bs = {nshell, l, nprim, prim_off, exps, coefs, centers};
  int* off = malloc(sizeof(int) * (nshell + 1));
  int n = shell_offsets(&bs, sph, off);
  double blk[36], prim[36], tmp[36];
  for (int i = 0; i < nshell; ++i)
    for (int j = i; j < nshell; ++j) {
      int na = ncart(l[i]), nb = ncart(l[j]);
      for (int x = 0; x < na * nb; ++x) blk[x] = 0.0;
      for (int a = 0; a < nprim[i]; ++a)
        for (int b = 0; b < nprim[j]; ++b) {
          double ea = exps[prim_off[i] + a], eb = exps[prim_off[j] + b];
          if (which == 0)
            orc_prim_overlap(ea, centers + 3 * i, l[i], eb, centers + 3 * j, l[j], prim);
          else if (which == 1)
            orc_prim_kinetic(ea, centers + 3 * i, l[i], eb, centers + 3 * j, l[j], prim);
          else
            orc_prim_nuclear(mode, ea, centers + 3 * i, l[i], eb, centers + 3 * j, l[j], natom, C,
                             Z, prim);
          double w = coefs[prim_off[i] + a] * coefs[prim_off[j] + b];
          for (int x = 0; x < na * nb; ++x) blk[x] += w * prim[x];
        }
      int fa = nfunc(l[i], sph), fb = nfunc(l[j], sph);
      if (sph) {
        double Ca[30], Cb[30];
        cart_to_sph(l[i], Ca);
        cart_to_sph(l[j], Cb);
        /* Ca @ blk @ Cb^T (overlap.py:58-64 _shell_block_spherical) */
        for (int r = 0; r < fa; ++r)
          for (int c = 0; c < nb; ++c) {
            double s = 0;
            for (int k = 0; k < na; ++k) s += Ca[r * na + k] * blk[k * nb + c];
            tmp[r * nb + c] = s;
          }
        for (int r = 0; r < fa; ++r)
          for (int c = 0; c < fb; ++c) {
            double s = 0;
            for (int k = 0; k < nb; ++k) s += tmp[r * nb + k] * Cb[c * nb + k];
            blk[r * fb + c] = s;
          }
      }
      for (int r = 0; r < fa; ++r)
        for (int c = 0; c < fb; ++c) {
          M[(off[i] + r) * n + off[j] + c] = blk[r * fb + c];
          M[(off[j] + c) * n + off[i] + r] = blk[r * fb + c];
        }
    }
  free(off);
}
