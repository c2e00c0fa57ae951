/*
 * ndyn_oracle.c -- CPU FP64 restatement of Ndyn's MD step.
 *
 * TEST INFRASTRUCTURE ONLY: this is the parity oracle for the HIP engine.  It is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * never by the product path.  Every routine names the reference lines it restates
 * (paths relative to the Ndyn_OMP source tree).  Built with -O2 -ffp-contract=off so
 * that FP64 expressions round in the order they are written, like the reference
 * built with strict flags.
 *
 * Pinning (which checks hold each routine; see tests/test_oracle.py):
 *   ndo_pair_table      golden tables from the compiled reference + closed forms
 *   ndo_build_half_list golden half lists from the compiled reference (bit-exact,
 *                       entry order included) + brute-force pair sets
 *   ndo_compute0        golden forces/PE/virial from the compiled reference, FCC
 *                       shell sums, the reference logs' PE/pressure, Newton 3
 *   ndo_compute1        golden forces/PE/virial of asymmetric pair matrices from the
 *                       compiled reference; equals ndo_compute0 on symmetric matrices
 *   ndo_integrate1/2,   golden multi-step 'dynamics' runs from the compiled reference
 *   ndo_mk_iso_*        (NVE, NVT, NPT, isoenergetic) + the testT NPT log
 *   ndo_check_list_update  compat: golden runs with ReNeighbor T; SKIN mode is not in
 *                       the reference and is pinned by brute-force list comparison
 *                       after every rebuild (list stays a superset of the cutoff set).
 */
#include "ndyn_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.141592653589793

/* ---------------------------------------------------------------- tables */

/* The damped shifted Coulomb term of MorseC / BuckC / LJC (src/potentials.f90:11,23-26;
 * energy :106-108, force :199-201), alpha = 0.2, ke = 1, for a pair of cutoff rc. */
#define ALPHA_C 0.2
static double coul_v(double r, double rc, double q) {
  double vcon = erfc(ALPHA_C * rc) / rc;
  double fcon = erfc(ALPHA_C * rc) / (rc * rc) +
                2.0 * ALPHA_C / sqrt(PI_D) * exp(-ALPHA_C * ALPHA_C * (rc * rc)) / rc;
  return q * (erfc(ALPHA_C * r) / r - vcon + fcon * (r - rc));
}
static double coul_dv(double r, double rc, double q) {
  double fcon = erfc(ALPHA_C * rc) / (rc * rc) +
                2.0 * ALPHA_C / sqrt(PI_D) * exp(-ALPHA_C * ALPHA_C * (rc * rc)) / rc;
  return q * (erfc(ALPHA_C * r) / (r * r) +
              2.0 * ALPHA_C / sqrt(PI_D) * exp(-(ALPHA_C * ALPHA_C) * (r * r)) / r - fcon);
}

/* pot()  src/potentials.f90:59-152 (Morse :94-110, Buck :112-122, LJ :124-133,
 * spring :88-92, sine/cosine/soft :82-87, John :136-147) */
static double pot_fn(int kind, double r, const double *p, double q) {
  if (kind == NDO_SINE) return sin(r) * sin(r) - 1.0;
  if (kind == NDO_COSINE) return cos(r) * cos(r) - 1.0;
  if (kind == NDO_SOFT) return p[1] * (1.0 + cos(PI_D * r / p[2]));
  if (kind == NDO_JOHNSON) {
    double eps = p[1], r0 = p[2], rc = p[3];
    double A = 2.0 * pow(r0, 3) / pow(rc - r0, 3);
    double B = 3.0 * pow(r0, 2) / pow(rc - r0, 2);
    double t = (r - rc) / r0;
    return r < rc ? -eps * (A * pow(t, 3) + B * pow(t, 2)) : 0.0;
  }
  if (kind == NDO_MORSE_C) return pot_fn(NDO_MORSE, r, p, 0.0) + coul_v(r, p[0], q);
  if (kind == NDO_BUCK_C) return pot_fn(NDO_BUCK, r, p, 0.0) + coul_v(r, p[0], q);
  if (kind == NDO_LJ_C) return pot_fn(NDO_LJ, r, p, 0.0) + coul_v(r, p[0], q);
  if (kind == NDO_SPRING) {
    double rc = p[0], D = p[1], C = p[2];
    double r0 = r / rc * C;
    return D * (r0 * r0 - C * C);
  }
  if (kind == NDO_MORSE || kind == NDO_MORSE_PHYS) {
    double D = p[1], alfa = p[2], r0 = p[3];
    double tmp = r - r0;
    if (r0 <= 0.0) tmp = 0.0;
    double e = exp(-alfa * tmp);
    return D * (e * e - 2.0 * e);
  }
  if (kind == NDO_BUCK) {
    double A = p[1], rho = p[2], C = p[3];
    double r6 = r * r * r * r * r * r;
    double v = A * exp(-r / rho) - C / r6;
    if (rho <= 0.0) v = A - C / r6;
    return v;
  }
  if (kind == NDO_LJ) {
    double eps = p[1], sig = p[2];
    double t = sig / r;
    double t6 = t * t * t * t * t * t;
    return 4.0 * eps * (t6 * t6 - t6);
  }
  return 0.0;
}

/* dpot()  src/potentials.f90:154-247 (Morse :184-204 with the /r at :198,
 * Buck :206-216, LJ :218-227).  Returns -dV/dr except for the reference Morse,
 * which returns -dV/dr / r. */
static double dpot_fn(int kind, double r, const double *p, double q) {
  if (kind == NDO_SINE) return sin(2.0 * r);
  if (kind == NDO_COSINE) return -sin(2.0 * r);
  /* soft's force reads parm(1), parm(2) (src/potentials.f90:176), its energy parm(2), parm(3) */
  if (kind == NDO_SOFT) return PI_D * p[0] * sin(PI_D * r / p[1]) / p[1];
  if (kind == NDO_JOHNSON) {   /* src/potentials.f90:230-241 */
    double eps = p[1], r0 = p[2], rc = p[3];
    double A = 2.0 * pow(r0, 3) / pow(rc - r0, 3);
    double B = 3.0 * pow(r0, 2) / pow(rc - r0, 2);
    double t = (r - rc) / r0;
    return r < rc ? eps / r0 * (3.0 * A * pow(t, 2) + 2.0 * B * t) : 0.0;
  }
  if (kind == NDO_MORSE_C) return dpot_fn(NDO_MORSE, r, p, 0.0) + coul_dv(r, p[0], q);
  if (kind == NDO_BUCK_C) return dpot_fn(NDO_BUCK, r, p, 0.0) + coul_dv(r, p[0], q);
  if (kind == NDO_LJ_C) return dpot_fn(NDO_LJ, r, p, 0.0) + coul_dv(r, p[0], q);
  if (kind == NDO_SPRING) {   /* src/potentials.f90:178-183 */
    double rc = p[0], D = p[1], C = p[2];
    double r0 = r / rc * C;
    return -2.0 * D * r0;
  }
  if (kind == NDO_MORSE || kind == NDO_MORSE_PHYS) {
    double D = p[1], alfa = p[2], r0 = p[3];
    double tmp = r - r0;
    if (r0 <= 0.0) tmp = 0.0;
    double e = exp(-alfa * tmp);
    double d = 2.0 * D * alfa * (e * e - e);
    return kind == NDO_MORSE ? d / r : d;
  }
  if (kind == NDO_BUCK) {
    double A = p[1], rho = p[2], C = p[3];
    double r7 = r * r * r * r * r * r * r;
    double d = A * exp(-r / rho) / rho - 6.0 * C / r7;
    if (rho <= 0.0) d = A - 6.0 * C / r7;
    return d;
  }
  if (kind == NDO_LJ) {
    double eps = p[1], sig = p[2];
    double t = sig / r;
    double t6 = t * t * t * t * t * t;
    return 24.0 * eps * (2.0 * t6 * t6 / r - t6 / r);
  }
  return 0.0;
}

/* buildPairTables  src/potentials.f90:2-56 */
int ndo_pair_table(int kind, const double params[4], double ptrmin, int ptsize,
                   double *P, double *F, double *rc2, double *dr2) {
  return ndo_pair_table_q(kind, params, 0.0, ptrmin, ptsize, P, F, rc2, dr2);
}

/* kind | NDO_ESHIFT (not in the reference): P holds V(r) - V(rc), F unchanged */
int ndo_pair_table_q(int kind, const double params[4], double qprd, double ptrmin, int ptsize,
                     double *P, double *F, double *rc2, double *dr2) {
  int eshift = (kind & NDO_ESHIFT) != 0;
  kind &= ~NDO_ESHIFT;
  if (kind < NDO_MORSE || kind > NDO_LJ_C || ptsize < 2 || (eshift && kind == NDO_SPRING))
    return NDO_ERR_ARG;
  double rcut = params[0];
  double c2 = rcut * rcut;
  double d2 = (c2 - ptrmin) / (double)(ptsize - 1);
  double vc = eshift ? pot_fn(kind, rcut, params, qprd) : 0.0;
  *rc2 = c2;
  *dr2 = d2;
  for (int i = 0; i <= ptsize; ++i) {
    double s = ptrmin + (double)i * d2;
    double r = sqrt(s);
    P[i] = pot_fn(kind, r, params, qprd) - vc;
    F[i] = r * dpot_fn(kind, r, params, qprd);
  }
  return NDO_OK;
}

/* src/nd.f90:84-88; the literals 4./3. and 1.1 are single precision there */
int ndo_neimax_estimate(double rcut, double skin, int np, const double box[3]) {
  double c = (double)(4.0f / 3.0f) * PI_D;
  double rr = (rcut + skin) * (double)1.1f;
  return (int)(c * (rr * rr * rr) * (double)np / (box[0] * box[1] * box[2]));
}

void ndo_cell_grid(const double box[3], double rcut, double skin, int mcell[3]) {
  for (int a = 0; a < 3; ++a) mcell[a] = (int)(box[a] / (rcut + skin));
}

/* ---------------------------------------------------------------- lists */

/* getJcellNP  src/nd.f90:1261-1303: 14-entry half stencil with single-precision fold */
static const int kStencil[14][3] = {
    {0, 0, 0},  {0, 0, 1},  {0, 1, 0},  {0, 1, -1}, {0, 1, 1},
    {1, 0, 0},  {1, 0, -1}, {1, 0, 1},  {1, 1, 0},  {1, 1, -1},
    {1, 1, 1},  {1, -1, 0}, {1, -1, -1}, {1, -1, 1}};

static int fold_cell(int t, int m) {
  float tmp = (float)(t + 1) / (float)(m + 1) - 0.5f;
  return t - m * (int)lroundf(tmp);
}

static double anint(double x) { return round(x); }

static double pair_d2(const double *pi, const double *pj, const double box[3],
                      double rij[3]) {
  double sr0 = pi[0] - pj[0], sr1 = pi[1] - pj[1], sr2 = pi[2] - pj[2];
  sr0 = sr0 - anint(sr0);
  sr1 = sr1 - anint(sr1);
  sr2 = sr2 - anint(sr2);
  rij[0] = box[0] * sr0;
  rij[1] = box[1] * sr1;
  rij[2] = box[2] * sr2;
  double d2 = rij[0] * rij[0];
  d2 = d2 + rij[1] * rij[1];
  d2 = d2 + rij[2] * rij[2];
  return d2;
}

/* buildCellList src/nd.f90:1214-1260 + buildVerletList src/nd.f90:1328-1446 */
int ndo_build_half_list(int np, const double *pos, const int *typ, const double box[3],
                        const int mcell[3], int ntypes, const int *pairmat,
                        const double *rc_pair, double skin, double ptrmin, int neimax,
                        int brute, int *vnl, int *cell_of) {
  const int stride = neimax + 1;
  const int mx = mcell[0], my = mcell[1], mz = mcell[2];
  if (!brute && (mx < 3 || my < 3 || mz < 3)) brute = 1;
  if (brute) {
    for (int i = 0; i < np; ++i) {
      int nj = 0;
      int ti = typ[i] - 1;
      for (int j = i + 1; j < np; ++j) {
        int tj = typ[j] - 1;
        int pij = pairmat[ti + tj * ntypes], pji = pairmat[tj + ti * ntypes];
        int pm = pij > pji ? pij : pji;
        if (pm <= 0) continue;
        double rij[3];
        double d2 = pair_d2(pos + 3 * i, pos + 3 * j, box, rij);
        double rcs = skin + rc_pair[pm - 1];
        if (d2 > rcs * rcs) continue;
        if (nj >= neimax) return NDO_ERR_NEIMAX;
        vnl[stride * i + 1 + nj] = j;
        ++nj;
      }
      vnl[stride * i] = nj;
      if (cell_of) cell_of[i] = 0;
    }
    return NDO_OK;
  }
  const int ncell = mx * my * mz;
  int *head = (int *)malloc(sizeof(int) * ncell);
  int *list = (int *)malloc(sizeof(int) * np);
  for (int c = 0; c < ncell; ++c) head[c] = -1;
  for (int i = 0; i < np; ++i) {
    int ix[3];
    for (int a = 0; a < 3; ++a) {
      double sr = pos[3 * i + a] - anint(pos[3 * i + a]);
      ix[a] = (int)((sr + 0.4999) * (double)mcell[a]);
    }
    int ic = ix[0] + ix[1] * mx + ix[2] * mx * my;
    if (cell_of) cell_of[i] = ic;
    list[i] = head[ic];
    head[ic] = i;
  }
  int err = NDO_OK;
  for (int ic = 0; ic < ncell && !err; ++ic) {
    /* same cell, i < j, and the d2 >= PTrMin test that only this loop has */
    for (int i = head[ic]; i >= 0; i = list[i]) {
      int nj = 0;
      int ti = typ[i] - 1;
      for (int j = head[ic]; j >= 0; j = list[j]) {
        if (i >= j) continue;
        int tj = typ[j] - 1;
        int pij = pairmat[ti + tj * ntypes], pji = pairmat[tj + ti * ntypes];
        int pm = pij > pji ? pij : pji;
        if (pm <= 0) continue;
        double rij[3];
        double d2 = pair_d2(pos + 3 * i, pos + 3 * j, box, rij);
        double rcs = skin + rc_pair[pm - 1];
        if (d2 < ptrmin) continue;
        if (d2 > rcs * rcs) continue;
        if (nj >= neimax) { err = NDO_ERR_NEIMAX; break; }
        vnl[stride * i + 1 + nj] = j;
        ++nj;
      }
      vnl[stride * i] = nj;
      if (err) break;
    }
    if (err) break;
    /* getCellVector(ic, IXYZ, .false.) src/nd.f90:1304-1327 */
    int v0 = ic % mx, v2 = ic / (mx * my), v1 = (ic % (mx * my)) / mx;
    for (int k = 1; k < 14 && !err; ++k) {
      int t0 = fold_cell(v0 + kStencil[k][0], mx);
      int t1 = fold_cell(v1 + kStencil[k][1], my);
      int t2 = fold_cell(v2 + kStencil[k][2], mz);
      int jc = t0 + t1 * mx + t2 * mx * my;
      for (int i = head[ic]; i >= 0; i = list[i]) {
        int nj = vnl[stride * i];
        int ti = typ[i] - 1;
        for (int j = head[jc]; j >= 0; j = list[j]) {
          int tj = typ[j] - 1;
          int pij = pairmat[ti + tj * ntypes], pji = pairmat[tj + ti * ntypes];
          int pm = pij > pji ? pij : pji;
          if (pm <= 0) continue;
          double rij[3];
          double d2 = pair_d2(pos + 3 * i, pos + 3 * j, box, rij);
          double rcs = skin + rc_pair[pm - 1];
          if (d2 > rcs * rcs) continue;
          if (nj >= neimax) { err = NDO_ERR_NEIMAX; break; }
          vnl[stride * i + 1 + nj] = j;
          ++nj;
        }
        vnl[stride * i] = nj;
        if (err) break;
      }
    }
  }
  free(head);
  free(list);
  return err;
}

/* ---------------------------------------------------------------- forces */

/* compute0  src/nd.f90:200-363, serial (one thread: the per-thread replica sum of
 * src/nd.f90:337-343 then adds each force to zero, which is exact). */
void ndo_compute0(int np, const double *pos, const int *typ, const double box[3],
                  const int *vnl, int neimax, int ntypes, const int *pairmat,
                  const double *P, const double *F, const double *rc2, const double *dr2,
                  double ptrmin, int ptsize, int cv, int cu,
                  double *force, double *pe, double virial[6], double *pene) {
  const int stride = neimax + 1;
  const int tab = ptsize + 1;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  double pot = 0.0;
  int npairs_max = ntypes * ntypes;
  float *vdr2 = (float *)malloc(sizeof(float) * (npairs_max + 1));
  for (int p = 0; p < npairs_max + 1; ++p) vdr2[p] = 0.0f;
  /* vPTdr2 = real(1/PTdr2): pair ids present in pairmat */
  for (int a = 0; a < npairs_max; ++a) {
    int p = pairmat[a];
    if (p > 0) vdr2[p] = (float)(1.0 / dr2[p - 1]);
  }
  memset(force, 0, sizeof(double) * 3 * (size_t)np);
  if (cu && pene) memset(pene, 0, sizeof(double) * (size_t)np);
  for (int i = 0; i < np; ++i) {
    const double *pi = pos + 3 * i;
    int ti = typ[i] - 1;
    int n = vnl[stride * i];
    for (int jn = 0; jn < n; ++jn) {
      int j = vnl[stride * i + 1 + jn];
      int pij = pairmat[ti + (typ[j] - 1) * ntypes];
      double rij[3];
      double d2 = pair_d2(pi, pos + 3 * j, box, rij);
      if (d2 > rc2[pij - 1]) continue;
      float rk = (float)(d2 - ptrmin) * vdr2[pij] + 1.0f;
      int k = (int)rk;
      if (k < 1) k = 1;
      double w = (double)rk - (double)k;
      const double *Ft = F + (size_t)tab * (pij - 1);
      double df = w * Ft[k] + (1.0 - w) * Ft[k - 1];
      df = df / d2;
      force[3 * i + 0] = force[3 * i + 0] + rij[0] * df;
      force[3 * i + 1] = force[3 * i + 1] + rij[1] * df;
      force[3 * i + 2] = force[3 * i + 2] + rij[2] * df;
      force[3 * j + 0] = force[3 * j + 0] - rij[0] * df;
      force[3 * j + 1] = force[3 * j + 1] - rij[1] * df;
      force[3 * j + 2] = force[3 * j + 2] - rij[2] * df;
      if (cu) {
        const double *Pt = P + (size_t)tab * (pij - 1);
        double dp = w * Pt[k] + (1.0 - w) * Pt[k - 1];
        pot = pot + dp;
        if (pene) {
          pene[i] += 0.5 * dp;
          pene[j] += 0.5 * dp;
        }
      }
      if (cv) {
        vir[0] = vir[0] + rij[0] * rij[0] * df;
        vir[1] = vir[1] + rij[1] * rij[1] * df;
        vir[2] = vir[2] + rij[2] * rij[2] * df;
        vir[3] = vir[3] + rij[0] * rij[1] * df;
        vir[4] = vir[4] + rij[0] * rij[2] * df;
        vir[5] = vir[5] + rij[1] * rij[2] * df;
      }
    }
  }
  if (cv)
    for (int a = 0; a < 6; ++a) virial[a] = 0.0 + vir[a];
  if (cu) *pe = 0.0 + pot;
  free(vdr2);
}

/* compute1  src/nd.f90:364-530 (asymmetric pair matrices), serial.  For a half-list pair
 * i<j the table is that of max(PairMat(ti,tj), PairMat(tj,ti)) (:445-447); i takes the
 * pair force only if PairMat(tj,ti) > 0 (:474-478), j only if PairMat(ti,tj) > 0
 * (:479-483); energy and virial count only pairs with PairMat(ti,tj) = PairMat(tj,ti)
 * (:485-498).  The reference leaves k unclamped here (:466-467, the guard is commented out);
 * pairs below PTrMin do not occur in valid inputs, and the clamp of compute0 is kept so
 * that such a pair cannot index outside the table.  pene: 1/2 of each counted pair's
 * energy to both atoms (the reference leaves it at zero). */
void ndo_compute1(int np, const double *pos, const int *typ, const double box[3],
                  const int *vnl, int neimax, int ntypes, const int *pairmat,
                  const double *P, const double *F, const double *rc2, const double *dr2,
                  double ptrmin, int ptsize, int cv, int cu,
                  double *force, double *pe, double virial[6], double *pene) {
  const int stride = neimax + 1;
  const int tab = ptsize + 1;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  double pot = 0.0;
  int npairs_max = ntypes * ntypes;
  float *vdr2 = (float *)malloc(sizeof(float) * (npairs_max + 1));
  for (int p = 0; p < npairs_max + 1; ++p) vdr2[p] = 0.0f;
  for (int a = 0; a < npairs_max; ++a) {
    int p = pairmat[a];
    if (p > 0) vdr2[p] = (float)(1.0 / dr2[p - 1]);
  }
  memset(force, 0, sizeof(double) * 3 * (size_t)np);
  if (cu && pene) memset(pene, 0, sizeof(double) * (size_t)np);
  for (int i = 0; i < np; ++i) {
    const double *pi = pos + 3 * i;
    int ti = typ[i] - 1;
    int n = vnl[stride * i];
    for (int jn = 0; jn < n; ++jn) {
      int j = vnl[stride * i + 1 + jn];
      int tj = typ[j] - 1;
      int pij = pairmat[ti + tj * ntypes], pji = pairmat[tj + ti * ntypes];
      int mx = pij > pji ? pij : pji;
      double rij[3];
      double d2 = pair_d2(pi, pos + 3 * j, box, rij);
      if (d2 > rc2[mx - 1]) continue;
      float rk = (float)(d2 - ptrmin) * vdr2[mx] + 1.0f;
      int k = (int)rk;
      if (k < 1) k = 1;
      double w = (double)rk - (double)k;
      const double *Ft = F + (size_t)tab * (mx - 1);
      double df = w * Ft[k] + (1.0 - w) * Ft[k - 1];
      df = df / d2;
      if (pji > 0) {
        force[3 * i + 0] = force[3 * i + 0] + rij[0] * df;
        force[3 * i + 1] = force[3 * i + 1] + rij[1] * df;
        force[3 * i + 2] = force[3 * i + 2] + rij[2] * df;
      }
      if (pij > 0) {
        force[3 * j + 0] = force[3 * j + 0] - rij[0] * df;
        force[3 * j + 1] = force[3 * j + 1] - rij[1] * df;
        force[3 * j + 2] = force[3 * j + 2] - rij[2] * df;
      }
      if (cu && pij == pji) {
        const double *Pt = P + (size_t)tab * (mx - 1);
        double dp = w * Pt[k] + (1.0 - w) * Pt[k - 1];
        pot = pot + dp;
        if (pene) {
          pene[i] += 0.5 * dp;
          pene[j] += 0.5 * dp;
        }
      }
      if (cv && pij == pji) {
        vir[0] = vir[0] + rij[0] * rij[0] * df;
        vir[1] = vir[1] + rij[1] * rij[1] * df;
        vir[2] = vir[2] + rij[2] * rij[2] * df;
        vir[3] = vir[3] + rij[0] * rij[1] * df;
        vir[4] = vir[4] + rij[0] * rij[2] * df;
        vir[5] = vir[5] + rij[1] * rij[2] * df;
      }
    }
  }
  if (cv)
    for (int a = 0; a < 6; ++a) virial[a] = 0.0 + vir[a];
  if (cu) *pe = 0.0 + pot;
  free(vdr2);
}

/* compute2 / compute3 (src/nd.f90:531-933), serial.  The pair id is PairMat(ti,tj) in
 * compute2 (:610) and the larger of both directions in compute3 (:810-812).  A spring pair
 * (:622-660, :827-862): strain against the initial positions, cut at PairParams(3) (no
 * distance cutoff); other pairs: the table as compute0/compute1 (:661-677, :863-882, the
 * index clamped at 1 as in compute0).  Force gating and energy/virial accounting for
 * asymmetric matrices as compute1 (:884-903). */
void ndo_compute23(int asym, int np, const double *pos, const int *typ, const double box[3],
                   const int *vnl, int neimax, int ntypes, const int *pairmat,
                   const double *P, const double *F, const double *rc2, const double *dr2,
                   double ptrmin, int ptsize, const int *spring, const double *pairparams,
                   const double *ipos, const double init_box[3], int cv, int cu,
                   double *force, double *pe, double virial[6], double *pene) {
  const int stride = neimax + 1;
  const int tab = ptsize + 1;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  double pot = 0.0;
  int npairs_max = ntypes * ntypes;
  float *vdr2 = (float *)malloc(sizeof(float) * (npairs_max + 1));
  for (int p = 0; p < npairs_max + 1; ++p) vdr2[p] = 0.0f;
  for (int a = 0; a < npairs_max; ++a) {
    int p = pairmat[a];
    if (p > 0) vdr2[p] = (float)(1.0 / dr2[p - 1]);
  }
  memset(force, 0, sizeof(double) * 3 * (size_t)np);
  if (cu && pene) memset(pene, 0, sizeof(double) * (size_t)np);
  for (int i = 0; i < np; ++i) {
    const double *pi = pos + 3 * i;
    int ti = typ[i] - 1;
    int n = vnl[stride * i];
    for (int jn = 0; jn < n; ++jn) {
      int j = vnl[stride * i + 1 + jn];
      int tj = typ[j] - 1;
      int pij = pairmat[ti + tj * ntypes], pji = pairmat[tj + ti * ntypes];
      int p = asym ? (pij > pji ? pij : pji) : pij;
      const int bal = !asym || pij == pji;
      double rij[3];
      double d2 = pair_d2(pi, pos + 3 * j, box, rij);
      double df, dp = 0.0;
      if (spring[p - 1]) {
        const double *prm = pairparams + 4 * (p - 1);
        double rij0[3];
        double d20 = pair_d2(ipos + 3 * i, ipos + 3 * j, init_box, rij0);
        double r = sqrt(d2);
        double r0 = sqrt(d20);
        double sij = r / r0 - 1.0;
        if (sij > prm[2]) continue;
        double Vi = prm[3];
        df = prm[1] * sij * Vi * Vi;
        if (cu && bal) dp = 0.5 * df * sij * r0;
        df = -df / r;
      } else {
        if (d2 > rc2[p - 1]) continue;
        float rk = (float)(d2 - ptrmin) * vdr2[p] + 1.0f;
        int k = (int)rk;
        if (k < 1) k = 1;
        double w = (double)rk - (double)k;
        const double *Ft = F + (size_t)tab * (p - 1);
        df = w * Ft[k] + (1.0 - w) * Ft[k - 1];
        if (cu && bal) {
          const double *Pt = P + (size_t)tab * (p - 1);
          dp = w * Pt[k] + (1.0 - w) * Pt[k - 1];
        }
        df = df / d2;
      }
      if (!asym || pji > 0) {
        force[3 * i + 0] = force[3 * i + 0] + rij[0] * df;
        force[3 * i + 1] = force[3 * i + 1] + rij[1] * df;
        force[3 * i + 2] = force[3 * i + 2] + rij[2] * df;
      }
      if (!asym || pij > 0) {
        force[3 * j + 0] = force[3 * j + 0] - rij[0] * df;
        force[3 * j + 1] = force[3 * j + 1] - rij[1] * df;
        force[3 * j + 2] = force[3 * j + 2] - rij[2] * df;
      }
      if (cu && bal) {
        pot = pot + dp;
        if (pene) {
          pene[i] += 0.5 * dp;
          pene[j] += 0.5 * dp;
        }
      }
      if (cv && bal) {
        vir[0] = vir[0] + rij[0] * rij[0] * df;
        vir[1] = vir[1] + rij[1] * rij[1] * df;
        vir[2] = vir[2] + rij[2] * rij[2] * df;
        vir[3] = vir[3] + rij[0] * rij[1] * df;
        vir[4] = vir[4] + rij[0] * rij[2] * df;
        vir[5] = vir[5] + rij[1] * rij[2] * df;
      }
    }
  }
  if (cv)
    for (int a = 0; a < 6; ++a) virial[a] = 0.0 + vir[a];
  if (cu) *pe = 0.0 + pot;
  free(vdr2);
}

/* Some PairMat(a,b) != PairMat(b,a): the reference's 'asym' flag (src/read_input.f90:130-133),
 * which selects compute1 in 'dynamics' (src/nd.f90:155-163). */
int ndo_asymmetric(int ntypes, const int *pairmat) {
  for (int a = 0; a < ntypes; ++a)
    for (int b = 0; b < ntypes; ++b)
      if (pairmat[a + b * ntypes] != pairmat[b + a * ntypes]) return 1;
  return 0;
}

/* ---------------------------------------------------------------- integrator */

void ndo_integrate1_hold(int np, double *pos, const double *vel, const double *acc,
                         double *disp, const double box[3], double dt, const int *hold,
                         const int *bc_disp, const int *bc_dir, const double *bc_parm) {
  double li[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
  for (int j = 0; j < np; ++j) {
    double du[3];
    for (int a = 0; a < 3; ++a)
      du[a] = li[a] * (vel[3 * j + a] * dt + 0.5 * acc[3 * j + a] * dt * dt);
    const int g = hold[j];
    if (g != 0 && bc_disp[g - 1])
      for (int a = 0; a < 3; ++a)
        if (bc_dir[3 * (g - 1) + a]) du[a] = bc_parm[3 * (g - 1) + a];
    for (int a = 0; a < 3; ++a) {
      disp[3 * j + a] = disp[3 * j + a] + du[a];
      pos[3 * j + a] = pos[3 * j + a] + du[a];
    }
  }
}

void ndo_integrate1(int np, double *pos, const double *vel, const double *acc,
                    double *disp, const double box[3], double dt) {
  double li[3] = {1.0 / box[0], 1.0 / box[1], 1.0 / box[2]};
  for (int j = 0; j < np; ++j)
    for (int a = 0; a < 3; ++a) {
      double du = li[a] * (vel[3 * j + a] * dt + 0.5 * acc[3 * j + a] * dt * dt);
      disp[3 * j + a] = disp[3 * j + a] + du;
      pos[3 * j + a] = pos[3 * j + a] + du;
    }
}

void ndo_integrate2(int np, double *vel, double *acc, const double *force, const int *typ,
                    const double *mass, double dt, int ck, double *ke) {
  double kin = 0.0;
  for (int j = 0; j < np; ++j) {
    double rmass = 1.0 / mass[typ[j] - 1];
    for (int a = 0; a < 3; ++a)
      vel[3 * j + a] = vel[3 * j + a] + 0.5 * dt * (force[3 * j + a] * rmass + acc[3 * j + a]);
    for (int a = 0; a < 3; ++a) acc[3 * j + a] = force[3 * j + a] * rmass;
    if (ck) {
      double kt = vel[3 * j] * vel[3 * j];
      kt = kt + vel[3 * j + 1] * vel[3 * j + 1];
      kt = kt + vel[3 * j + 2] * vel[3 * j + 2];
      kt = kt * mass[typ[j] - 1];
      kin = kin + kt;
    }
  }
  kin = 0.0 + kin;
  if (ck) kin = kin * 0.5;
  *ke = kin;
}

double ndo_mk_iso_ke(int np, double *vel, double dt, double keparm, double ketarget,
                     double kinetic) {
  double lam = 1.0 + dt / keparm * (ketarget * (double)np / kinetic - 1.0);
  lam = sqrt(lam);
  for (int j = 0; j < 3 * np; ++j) vel[j] = vel[j] * lam;
  return lam;
}

double ndo_mk_iso_e(int np, double *vel, double dt, double eparm, double e0, double pe,
                    double ke) {
  double lam = 1.0 - dt / eparm * (e0 / (pe + ke) - 1.0);
  for (int j = 0; j < 3 * np; ++j) vel[j] = vel[j] * lam;
  return lam;
}

void ndo_mk_iso_p(double box[3], const double virial[6], double dt, double pparm,
                  const double ptarget[3]) {
  double vol = box[0] * box[1] * box[2];
  double press[6], lam[6];
  for (int a = 0; a < 6; ++a) press[a] = virial[a] / vol;
  for (int a = 0; a < 3; ++a) lam[a] = 1.0 - dt / pparm * (ptarget[a] - press[a]);
  lam[3] = dt / pparm * press[3];
  lam[4] = dt / pparm * press[4];
  lam[5] = dt / pparm * press[5];
  box[0] = box[0] * lam[0] + box[1] * lam[5] + box[2] * lam[4];
  box[1] = box[0] * lam[5] + box[1] * lam[1] + box[2] * lam[3];
  box[2] = box[0] * lam[4] + box[1] * lam[3] + box[2] * lam[2];
}

/* CheckListUpdate  src/nd.f90:1447-1499 (top-2 of |disp|^2 with the >= updates) */
int ndo_check_list_update(int np, const double *disp, double skin, const double box[3],
                          const double box_build[3], double rc_max, int mode) {
  double t1 = 0.0, t2 = 0.0;   /* scaled units, as the reference */
  double a1 = 0.0, a2 = 0.0;   /* Angstrom, SKIN mode */
  for (int i = 0; i < np; ++i) {
    const double *d = disp + 3 * i;
    double s = d[0] * d[0];
    s = s + d[1] * d[1];
    s = s + d[2] * d[2];
    if (s >= t1) { t2 = t1; t1 = s; } else if (s >= t2) { t2 = s; }
    double x = d[0] * box[0], y = d[1] * box[1], z = d[2] * box[2];
    double q = x * x + y * y + z * z;
    if (q >= a1) { a2 = a1; a1 = q; } else if (q >= a2) { a2 = q; }
  }
  double dsum = t1 + t2;
  if (mode == NDO_TRIGGER_COMPAT) {
    if (dsum > skin * skin) {
      if (dsum > 1.0) return -NDO_ERR_OUTSIDE;
      return 1;
    }
    return 0;
  }
  if (dsum > 1.0) return -NDO_ERR_OUTSIDE;
  double strain = 0.0;
  for (int a = 0; a < 3; ++a) {
    double e = fabs(box[a] / box_build[a] - 1.0);
    if (e > strain) strain = e;
  }
  double need = sqrt(a1) + sqrt(a2) + (rc_max + skin) * strain;
  return need > skin ? 1 : 0;
}

/* ---------------------------------------------------------------- dynamics */

int ndo_run(const ndo_run_params *rp, int np, int ntypes, int npairs, int ptsize,
            double ptrmin, const double *mass, const int *pairmat, const double *pairparams,
            const double *P, const double *F, const double *rc2, const double *dr2,
            const int *typ, double *pos, double *vel, double *acc, double box[3],
            double *force, double *pene, ndo_thermo *rec, int max_rec, int *nrec,
            int *nupdates) {
  double rcut = 0.0;
  double *rcp = (double *)malloc(sizeof(double) * npairs);
  for (int p = 0; p < npairs; ++p) {
    rcp[p] = pairparams[4 * p];
    if (rcp[p] > rcut) rcut = rcp[p];
  }
  double skin = rp->skin < 0.0 ? rcut * 0.2 : rp->skin;
  int mcell[3];
  ndo_cell_grid(box, rcut, skin, mcell);
  int neimax = rp->neimax > 0 ? rp->neimax : ndo_neimax_estimate(rcut, skin, np, box);
  int *vnl = (int *)malloc(sizeof(int) * (size_t)(neimax + 1) * np);
  double *disp = (double *)calloc((size_t)3 * np, sizeof(double));
  double box_build[3] = {box[0], box[1], box[2]};
  int err = ndo_build_half_list(np, pos, typ, box, mcell, ntypes, pairmat, rcp, skin,
                                ptrmin, neimax, rp->brute, vnl, NULL);
  int nup = 0, nr = 0;
  const int asym = ndo_asymmetric(ntypes, pairmat);
  int *spring = (int *)calloc((size_t)npairs, sizeof(int));
  for (int p = 0; p < npairs && p < 31; ++p) spring[p] = (rp->springs >> p) & 1;
  ndo_run_params *rpw = (ndo_run_params *)rp;   /* ipos / init_box / ipos_set are in-out */
  if (rp->springs && !rp->ipos_set) {
    /* initPositionSet at the start of dynamics (src/nd.f90:103, :1731-1763) */
    memcpy(rpw->ipos, pos, sizeof(double) * 3 * (size_t)np);
    for (int a = 0; a < 3; ++a) rpw->init_box[a] = box[a];
    rpw->ipos_set = 1;
  }
  double pe = 0.0, ke = 0.0, e0 = 0.0, vir[6] = {0, 0, 0, 0, 0, 0};
  int step = rp->step0;
  for (int tt = 1; tt <= rp->nsteps && !err; ++tt) {
    ++step;
    int cv = 0, cu = 0, ck = 0;
    if (step % rp->out_steps == 0 || tt == 1) cv = cu = ck = 1;
    if (rp->isokinetic && step % rp->ke_freq == 0) ck = 1;
    if (rp->isobaric && step % rp->p_freq == 0) cv = 1;
    if (rp->isoenergetic && step % rp->e_freq == 0) cu = ck = 1;
    if (rp->renei && step % rp->renei_freq == 0) {
      int r = ndo_check_list_update(np, disp, skin, box, box_build, rcut, rp->trigger);
      if (r < 0) { err = -r; break; }
      if (r > 0) {
        ++nup;
        if (rp->trigger == NDO_TRIGGER_SKIN) ndo_cell_grid(box, rcut, skin, mcell);
        err = ndo_build_half_list(np, pos, typ, box, mcell, ntypes, pairmat, rcp, skin,
                                  ptrmin, neimax, rp->brute, vnl, NULL);
        memset(disp, 0, sizeof(double) * 3 * (size_t)np);
        for (int a = 0; a < 3; ++a) box_build[a] = box[a];
        if (err) break;
      }
    }
    if (rp->nhold > 0)
      ndo_integrate1_hold(np, pos, vel, acc, disp, box, rp->dt, rp->hold, rp->bc_disp,
                          rp->bc_dir, rp->bc_parm);
    else
      ndo_integrate1(np, pos, vel, acc, disp, box, rp->dt);
    if (rp->springs)
      ndo_compute23(asym, np, pos, typ, box, vnl, neimax, ntypes, pairmat, P, F, rc2, dr2,
                    ptrmin, ptsize, spring, pairparams, rp->ipos, rp->init_box, cv, cu, force,
                    &pe, vir, pene);
    else if (asym)
      ndo_compute1(np, pos, typ, box, vnl, neimax, ntypes, pairmat, P, F, rc2, dr2, ptrmin,
                   ptsize, cv, cu, force, &pe, vir, pene);
    else
      ndo_compute0(np, pos, typ, box, vnl, neimax, ntypes, pairmat, P, F, rc2, dr2, ptrmin,
                   ptsize, cv, cu, force, &pe, vir, pene);
    ndo_integrate2(np, vel, acc, force, typ, mass, rp->dt, ck, &ke);
    if (tt == 1) e0 = pe + ke;
    if (step % rp->out_steps == 0 || tt == 1) {
      if (rec && nr < max_rec) {
        ndo_thermo *t = rec + nr;
        t->step = step;
        t->pe = pe;
        t->ke = ke;
        t->eerr = (pe + ke - e0) / e0;
        t->pressure = (vir[0] + vir[1] + vir[2]) / (box[0] * box[1] * box[2] * 3.0);
        t->volume = box[0] * box[1] * box[2];
        for (int a = 0; a < 6; ++a) t->virial[a] = vir[a];
        for (int a = 0; a < 3; ++a) t->box[a] = box[a];
      }
      ++nr;
    }
    if (rp->isokinetic && step % rp->ke_freq == 0)
      ndo_mk_iso_ke(np, vel, rp->dt, rp->ke_tau, rp->ke_target, ke);
    if (rp->isobaric && step % rp->p_freq == 0)
      ndo_mk_iso_p(box, vir, rp->dt, rp->p_tau, rp->p_target);
    if (rp->isoenergetic && step % rp->e_freq == 0)
      ndo_mk_iso_e(np, vel, rp->dt, rp->e_tau, e0, pe, ke);
    if (rp->box_load)
      for (int a = 0; a < 3; ++a) box[a] = box[a] + rp->box_vel[a] * rp->dt;
  }
  if (nrec) *nrec = nr < max_rec ? nr : max_rec;
  if (nupdates) *nupdates = nup;
  free(rcp);
  free(vnl);
  free(disp);
  free(spring);
  return err;
}
