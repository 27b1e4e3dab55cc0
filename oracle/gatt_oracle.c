/*
 * gatt_oracle.c -- CPU restatement of the reference time-advance path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the parity tests compare
 * the HIP library against; nothing in the product links, loads or calls it.
 *
 * Each routine follows the Fortran it restates statement by statement, in the
 * same evaluation order, so that the only differences to a reference build are
 * the compiler's rounding of identical expressions:
 *   oracle_solve            solveSystem{X,Y,Z}[per]   source/Fortran/derivs.F90:556-948
 *   oracle_apply_filters    applyFilters2D/3D         source/Fortran/derivs.F90:303-552
 *   oracle_nseq             nseq2d / nseq3d           source/Fortran/nseq.F90:10-139, 142-323
 *   forcing (static)        turbulenceGenerator       source/disturbances/turbulenceGenerator.F90:1-44
 *   oracle_apply_bound_cond applyBoundCond            source/Fortran/nseq.F90:327-599
 *   disturbances (static)   lidDrivenMovement, holdInlet
 *                           source/disturbances/lidDrivenMovement.F90:1-19, source/Fortran/holdInlet.F90:3-48
 *   oracle_rk4              RK4_2D / RK4_3D           source/Fortran/timeIntegration.F90:8-156, 338-498
 *   oracle_euler            EULER_2D / EULER_3D       source/Fortran/timeIntegration.F90:158-222, 500-566
 *   oracle_ssprk3           SSPRK3_2D / SSPRK3_3D     source/Fortran/timeIntegration.F90:225-335, 569-685
 *   oracle_time_step        timeStepping select       source/Fortran/main.F90:255-350
 *   oracle_sfd              SFD block of main         source/Fortran/main.F90:381-402
 *   oracle_step             one main-loop iteration   source/Fortran/main.F90:196-462
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no FMA contraction).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/gatt_dns.h"

typedef struct oracle_ctx {
  const gatt_problem_desc* p;
  size_t np;
  /* holdInlet SAVE state, one slot per disturbance */
  int* hold_first;
  double** hold_save; /* [nDisturb][5 * boxsize] */
  /* SFD mean */
  double* mean[5];
  int mean_ready;
  /* scratch */
  double* work[64];
} oracle_ctx;

#define IDX(p, i, j, k) ((size_t)((i) - 1) + (size_t)(p)->nx * ((size_t)((j) - 1) + (size_t)(p)->ny * (size_t)((k) - 1)))

static double* zalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

oracle_ctx* oracle_create(const gatt_problem_desc* p) {
  oracle_ctx* c = (oracle_ctx*)calloc(1, sizeof(oracle_ctx));
  c->p = p;
  c->np = (size_t)p->nx * p->ny * p->nz;
  int nd = p->bc.nDisturb;
  c->hold_first = (int*)calloc(nd > 0 ? nd : 1, sizeof(int));
  c->hold_save = (double**)calloc(nd > 0 ? nd : 1, sizeof(double*));
  for (int d = 0; d < nd; ++d) c->hold_first[d] = 1;
  for (int f = 0; f < 5; ++f) c->mean[f] = zalloc(c->np);
  for (int w = 0; w < 64; ++w) c->work[w] = zalloc(c->np);
  return c;
}

void oracle_destroy(oracle_ctx* c) {
  if (!c) return;
  for (int d = 0; d < c->p->bc.nDisturb; ++d) free(c->hold_save[d]);
  free(c->hold_save);
  free(c->hold_first);
  for (int f = 0; f < 5; ++f) free(c->mean[f]);
  for (int w = 0; w < 64; ++w) free(c->work[w]);
  free(c);
}

/* ------------------------------------------------------------------------- */
/* Line solves: derivs.F90:556-948                                           */
/* ------------------------------------------------------------------------- */

/* element e (1-based) of the line whose other two indices are (a,b) */
static size_t line_addr(const gatt_problem_desc* p, int dir, int a, int b, int e) {
  if (dir == 0) return IDX(p, e, a, b);
  if (dir == 1) return IDX(p, a, e, b);
  return IDX(p, a, b, e);
}

static int modulo(int a, int n) { int r = a % n; return r < 0 ? r + n : r; }

void oracle_solve(const gatt_problem_desc* p, int dir, int kind, const double* F, double* dF) {
  const gatt_lineop* op = kind == GATT_FILTER ? &p->filter[dir] : &p->deriv[dir];
  const int N = op->n, nR = op->nR, per = op->periodic;
  const int NB = N - per, NW = 2 * nR - 1;
  int n1, n2;
  if (dir == 0) { n1 = p->ny; n2 = p->nz; }
  else if (dir == 1) { n1 = p->nx; n2 = p->nz; }
  else { n1 = p->nx; n2 = p->ny; }
  const int32_t* tm = p->types[dir];
#define RR(i, c) op->R[(size_t)((i) - 1) + (size_t)N * ((size_t)((c) - 1) + (size_t)NW * (size_t)(t - 1))]
#define AA(i) op->A[(size_t)((i) - 1) + (size_t)(N - 1) * (size_t)(t - 1)]
#define BB(i) op->B[(size_t)((i) - 1) + (size_t)NB * (size_t)(t - 1)]
#define CC(i) op->C[(size_t)((i) - 1) + (size_t)(N - 1) * (size_t)(t - 1)]
#define DD(i) op->D[(size_t)((i) - 1) + (size_t)N * (size_t)(t - 1)]
#define FL(e) F[line_addr(p, dir, a, b, (e))]
#define DL(e) dF[line_addr(p, dir, a, b, (e))]
  for (int b = 1; b <= n2; ++b)
    for (int a = 1; a <= n1; ++a) {
      const int t = tm[(a - 1) + n1 * (b - 1)];
      /* right-hand side (derivs.F90:583-602) */
      for (int i = 1; i <= nR; ++i) {
        double acc = RR(i, nR) * FL(i);
        for (int l = 1; l <= nR - 1; ++l)
          acc = acc + RR(i, nR + l) * FL(i + l) + RR(i, nR - l) * FL(1 + modulo(i - 1 - l, N));
        DL(i) = acc;
      }
      for (int i = nR + 1; i <= N - nR; ++i) {
        double acc = RR(i, nR) * FL(i);
        for (int l = 1; l <= nR - 1; ++l) acc = acc + RR(i, nR + l) * FL(i + l) + RR(i, nR - l) * FL(i - l);
        DL(i) = acc;
      }
      for (int i = N - nR + 1; i <= N; ++i) {
        double acc = RR(i, nR) * FL(i);
        for (int l = 1; l <= nR - 1; ++l)
          acc = acc + RR(i, nR + l) * FL(1 + modulo(i - 1 + l, N)) + RR(i, nR - l) * FL(i - l);
        DL(i) = acc;
      }
      if (!per) {
        /* Thomas (derivs.F90:605-611) */
        DL(1) = DL(1) * BB(1);
        for (int i = 1; i <= N - 1; ++i) DL(i + 1) = (DL(i + 1) - AA(i) * DL(i)) * BB(i + 1);
        for (int i = N - 1; i >= 1; --i) DL(i) = DL(i) - CC(i) * DL(i + 1);
      } else {
        /* first element from the first row of the inverse (derivs.F90:786-789) */
        DL(1) = DL(1) * DD(1);
        for (int i = 2; i <= N; ++i) DL(1) = DL(1) + DL(i) * DD(i);
        /* known solution into the RHS (792-793) */
        DL(2) = DL(2) - AA(1) * DL(1);
        DL(N) = DL(N) - CC(1) * DL(1);
        /* Thomas on rows 2..N (796-802) */
        DL(2) = DL(2) * BB(1);
        for (int i = 2; i <= N - 1; ++i) DL(i + 1) = (DL(i + 1) - AA(i) * DL(i)) * BB(i);
        for (int i = N - 1; i >= 2; --i) DL(i) = DL(i) - CC(i) * DL(i + 1);
      }
    }
#undef RR
#undef AA
#undef BB
#undef CC
#undef DD
#undef FL
#undef DL
}

/* ------------------------------------------------------------------------- */
/* Filters: derivs.F90:303-552                                               */
/* ------------------------------------------------------------------------- */

void oracle_apply_filters(oracle_ctx* c, double theta, double* const* q) {
  const gatt_problem_desc* p = c->p;
  const size_t np = c->np;
  const int is3d = p->nz > 1;
  const int fields[5] = {GATT_U, GATT_V, GATT_W, GATT_R, GATT_E};
  const int nf = is3d ? 5 : 4;
  int list[5];
  for (int f = 0, m = 0; f < 5; ++f)
    if (is3d || fields[f] != GATT_W) list[m++] = fields[f];
  double* tmp = c->work[0];
  for (int dir = 0; dir < (is3d ? 3 : 2); ++dir) {
    if (!p->filterDir[dir]) continue;
    for (int m = 0; m < nf; ++m) {
      double* F = q[list[m]];
      oracle_solve(p, dir, GATT_FILTER, F, tmp);
      if (theta > 0) {
        for (size_t n = 0; n < np; ++n) F[n] = F[n] + theta * (tmp[n] - F[n]);
      } else {
        memcpy(F, tmp, np * sizeof(double));
      }
    }
  }
}

/* ------------------------------------------------------------------------- */
/* Forcing: turbulenceGenerator.F90:1-44                                     */
/* ------------------------------------------------------------------------- */

static void forcing(const gatt_problem_desc* p, const double* const* q, double* const* dq, int is3d) {
  for (int fi = 0; fi < p->nForcing; ++fi) {
    const gatt_forcing* fc = &p->forcing[fi];
    if (fc->kind == GATT_FORCING_PRESSURE_GRADIENT) {
      /* pressureGradient.F90:13: du = du + dpdx/r on the box */
      const int* bx = fc->box;
      const double dpdx = fc->par[0];
      const double* r = q[fc->var[0]];
      double* du = dq[fc->var[1]];
      for (int k = bx[4]; k <= bx[5]; ++k)
        for (int j = bx[2]; j <= bx[3]; ++j)
          for (int i = bx[0]; i <= bx[1]; ++i) {
            size_t n = IDX(p, i, j, k);
            du[n] = du[n] + dpdx / r[n];
          }
      continue;
    }
    if (fc->kind != GATT_FORCING_TURBULENCE || !is3d) continue;
    const int* bx = fc->box;
    const double eps = fc->par[0], e0 = fc->par[1];
    const double* U = q[GATT_U]; const double* V = q[GATT_V]; const double* W = q[GATT_W];
    const double* E = q[GATT_E];
    double ET = 0, Us = 0, Vs = 0, Ws = 0, Es = 0;
    int N = 0;
    /* Fortran sum() runs in column-major order */
    for (int k = bx[4]; k <= bx[5]; ++k)
      for (int j = bx[2]; j <= bx[3]; ++j)
        for (int i = bx[0]; i <= bx[1]; ++i) {
          size_t n = IDX(p, i, j, k);
          ET += U[n] * U[n] + V[n] * V[n] + W[n] * W[n];
          Us += U[n]; Vs += V[n]; Ws += W[n]; Es += E[n];
          ++N;
        }
    const double gam = 10 * eps;
    const double Nd = (double)N;
    for (int k = bx[4]; k <= bx[5]; ++k)
      for (int j = bx[2]; j <= bx[3]; ++j)
        for (int i = bx[0]; i <= bx[1]; ++i) {
          size_t n = IDX(p, i, j, k);
          dq[GATT_U][n] = dq[GATT_U][n] + 0.25 * eps * Nd / ET * U[n] - gam * Us / Nd;
          dq[GATT_V][n] = dq[GATT_V][n] + 0.25 * eps * Nd / ET * V[n] - gam * Vs / Nd;
          dq[GATT_W][n] = dq[GATT_W][n] + 0.25 * eps * Nd / ET * W[n] - gam * Ws / Nd;
          dq[GATT_E][n] = dq[GATT_E][n] - gam * (Es / Nd - e0);
        }
  }
}

/* ------------------------------------------------------------------------- */
/* Navier-Stokes RHS: nseq.F90:10-323                                        */
/* ------------------------------------------------------------------------- */

void oracle_nseq(oracle_ctx* c, const double* const* q, double* const* dq) {
  const gatt_problem_desc* p = c->p;
  const size_t np = c->np;
  const double Re = p->Re, Ma = p->Ma, Pr = p->Pr, gamma = p->gamma, T0 = p->T0;
  const double *U = q[GATT_U], *V = q[GATT_V], *W = q[GATT_W], *R = q[GATT_R], *E = q[GATT_E];
  double** w = c->work + 1; /* work[0] is used by the filter */
  if (p->nz > 1) {
    double *DUx = w[0], *DUy = w[1], *DUz = w[2], *DVx = w[3], *DVy = w[4], *DVz = w[5];
    double *DWx = w[6], *DWy = w[7], *DWz = w[8], *DRx = w[9], *DRy = w[10], *DRz = w[11];
    double *DEx = w[12], *DEy = w[13], *DEz = w[14];
    double *P = w[15], *DPx = w[16], *DPy = w[17], *DPz = w[18], *mu = w[19];
    double *tauxx = w[20], *tauyy = w[21], *tauzz = w[22], *tauxy = w[23], *tauxz = w[24], *tauyz = w[25];
    double *dtauxxdx = w[26], *dtauyydy = w[27], *dtauzzdz = w[28], *vtx = w[29], *vty = w[30], *vtz = w[31];
    double *dtauxydx = w[32], *dtauxydy = w[33], *dtauxzdx = w[34], *dtauxzdz = w[35];
    double *dtauyzdy = w[36], *dtauyzdz = w[37];
    double *qx = w[38], *qy = w[39], *qz = w[40], *htx = w[41], *hty = w[42], *htz = w[43];
    const double* fld[5] = {U, V, W, R, E};
    double* der[5][3] = {{DUx, DUy, DUz}, {DVx, DVy, DVz}, {DWx, DWy, DWz}, {DRx, DRy, DRz}, {DEx, DEy, DEz}};
    for (int f = 0; f < 5; ++f)
      for (int d = 0; d < 3; ++d) oracle_solve(p, d, GATT_DERIV, fld[f], der[f][d]);
    for (size_t n = 0; n < np; ++n) {
      P[n] = (gamma - 1.0) * R[n] * E[n];
      DPx[n] = (gamma - 1.0) * (R[n] * DEx[n] + DRx[n] * E[n]);
      DPy[n] = (gamma - 1.0) * (R[n] * DEy[n] + DRy[n] * E[n]);
      DPz[n] = (gamma - 1.0) * (R[n] * DEz[n] + DRz[n] * E[n]);
      double mu1 = (1.0 + 110.0 / T0) / (gamma * (gamma - 1.0) * Ma * Ma * E[n] + 110.0 / T0);
      double mu2 = gamma * (gamma - 1.0) * Ma * Ma * E[n];
      mu[n] = mu1 * mu2 * sqrt(mu2);
      double tz = -(2.0 / 3.0 * (DUx[n] + DVy[n] + DWz[n]));
      tauxx[n] = 1.0 / Re * mu[n] * (2.0 * DUx[n] + tz);
      tauyy[n] = 1.0 / Re * mu[n] * (2.0 * DVy[n] + tz);
      tauzz[n] = 1.0 / Re * mu[n] * (2.0 * DWz[n] + tz);
      tauxy[n] = 1.0 / Re * mu[n] * (DUy[n] + DVx[n]);
      tauxz[n] = 1.0 / Re * mu[n] * (DUz[n] + DWx[n]);
      tauyz[n] = 1.0 / Re * mu[n] * (DVz[n] + DWy[n]);
    }
    oracle_solve(p, 0, GATT_DERIV, tauxy, dtauxydx);
    oracle_solve(p, 1, GATT_DERIV, tauxy, dtauxydy);
    oracle_solve(p, 0, GATT_DERIV, tauxz, dtauxzdx);
    oracle_solve(p, 2, GATT_DERIV, tauxz, dtauxzdz);
    oracle_solve(p, 1, GATT_DERIV, tauyz, dtauyzdy);
    oracle_solve(p, 2, GATT_DERIV, tauyz, dtauyzdz);
    oracle_solve(p, 0, GATT_DERIV, tauxx, dtauxxdx);
    oracle_solve(p, 1, GATT_DERIV, tauyy, dtauyydy);
    oracle_solve(p, 2, GATT_DERIV, tauzz, dtauzzdz);
    for (size_t n = 0; n < np; ++n) {
      vtx[n] = dtauxxdx[n] + dtauxydy[n] + dtauxzdz[n];
      vty[n] = dtauyydy[n] + dtauxydx[n] + dtauyzdz[n];
      vtz[n] = dtauzzdz[n] + dtauxzdx[n] + dtauyzdy[n];
      qx[n] = -(gamma / Re / Pr * mu[n] * DEx[n]);
      qy[n] = -(gamma / Re / Pr * mu[n] * DEy[n]);
      qz[n] = -(gamma / Re / Pr * mu[n] * DEz[n]);
    }
    oracle_solve(p, 0, GATT_DERIV, qx, htx);
    oracle_solve(p, 1, GATT_DERIV, qy, hty);
    oracle_solve(p, 2, GATT_DERIV, qz, htz);
    double *dU = dq[GATT_U], *dV = dq[GATT_V], *dW = dq[GATT_W], *dR = dq[GATT_R], *dE = dq[GATT_E];
    for (size_t n = 0; n < np; ++n) {
      double div = DUx[n] + DVy[n] + DWz[n];
      dR[n] = -(R[n] * div) - DRx[n] * U[n] - DRy[n] * V[n] - DRz[n] * W[n];
      dU[n] = -(DUx[n] * U[n] + DUy[n] * V[n] + DUz[n] * W[n]) + (vtx[n] - DPx[n]) / R[n];
      dV[n] = -(DVx[n] * U[n] + DVy[n] * V[n] + DVz[n] * W[n]) + (vty[n] - DPy[n]) / R[n];
      dW[n] = -(DWx[n] * U[n] + DWy[n] * V[n] + DWz[n] * W[n]) + (vtz[n] - DPz[n]) / R[n];
      dE[n] = -(DEx[n] * U[n] + DEy[n] * V[n] + DEz[n] * W[n]) +
              (-(P[n] * div) + tauxx[n] * DUx[n] + tauyy[n] * DVy[n] + tauzz[n] * DWz[n] +
               tauxy[n] * (DVx[n] + DUy[n]) + tauxz[n] * (DWx[n] + DUz[n]) + tauyz[n] * (DWy[n] + DVz[n]) -
               htx[n] - hty[n] - htz[n]) / R[n];
    }
    forcing(p, q, dq, 1);
  } else {
    double *DUx = w[0], *DUy = w[1], *DVx = w[2], *DVy = w[3], *DRx = w[4], *DRy = w[5];
    double *DEx = w[6], *DEy = w[7], *P = w[8], *DPx = w[9], *DPy = w[10], *mu = w[11];
    double *tauxx = w[12], *tauyy = w[13], *tauxy = w[14], *dtauxxdx = w[15], *dtauyydy = w[16];
    double *dtauxydx = w[17], *dtauxydy = w[18], *vtx = w[19], *vty = w[20];
    double *qx = w[21], *qy = w[22], *htx = w[23], *hty = w[24];
    oracle_solve(p, 0, GATT_DERIV, U, DUx); oracle_solve(p, 1, GATT_DERIV, U, DUy);
    oracle_solve(p, 0, GATT_DERIV, V, DVx); oracle_solve(p, 1, GATT_DERIV, V, DVy);
    oracle_solve(p, 0, GATT_DERIV, R, DRx); oracle_solve(p, 1, GATT_DERIV, R, DRy);
    oracle_solve(p, 0, GATT_DERIV, E, DEx); oracle_solve(p, 1, GATT_DERIV, E, DEy);
    for (size_t n = 0; n < np; ++n) {
      P[n] = (gamma - 1.0) * R[n] * E[n];
      DPx[n] = (gamma - 1.0) * (R[n] * DEx[n] + DRx[n] * E[n]);
      DPy[n] = (gamma - 1.0) * (R[n] * DEy[n] + DRy[n] * E[n]);
      double mu1 = (1.0 + 110.0 / T0) / (gamma * (gamma - 1.0) * Ma * Ma * E[n] + 110.0 / T0);
      double mu2 = gamma * (gamma - 1.0) * Ma * Ma * E[n];
      mu[n] = mu1 * mu2 * sqrt(mu2);
      tauxx[n] = 1.0 / Re * mu[n] * (2.0 * DUx[n] - 2.0 / 3.0 * (DUx[n] + DVy[n]));
      tauyy[n] = 1.0 / Re * mu[n] * (2.0 * DVy[n] - 2.0 / 3.0 * (DUx[n] + DVy[n]));
      tauxy[n] = 1.0 / Re * mu[n] * (DUy[n] + DVx[n]);
    }
    oracle_solve(p, 0, GATT_DERIV, tauxy, dtauxydx);
    oracle_solve(p, 1, GATT_DERIV, tauxy, dtauxydy);
    oracle_solve(p, 0, GATT_DERIV, tauxx, dtauxxdx);
    oracle_solve(p, 1, GATT_DERIV, tauyy, dtauyydy);
    for (size_t n = 0; n < np; ++n) {
      vtx[n] = dtauxxdx[n] + dtauxydy[n];
      vty[n] = dtauyydy[n] + dtauxydx[n];
      qx[n] = -(gamma / Re / Pr * mu[n] * DEx[n]);
      qy[n] = -(gamma / Re / Pr * mu[n] * DEy[n]);
    }
    oracle_solve(p, 0, GATT_DERIV, qx, htx);
    oracle_solve(p, 1, GATT_DERIV, qy, hty);
    double *dU = dq[GATT_U], *dV = dq[GATT_V], *dR = dq[GATT_R], *dE = dq[GATT_E];
    for (size_t n = 0; n < np; ++n) {
      double div = DUx[n] + DVy[n];
      dR[n] = -(R[n] * div) - DRx[n] * U[n] - DRy[n] * V[n];
      dU[n] = -(DUx[n] * U[n] + DUy[n] * V[n]) + (vtx[n] - DPx[n]) / R[n];
      dV[n] = -(DVx[n] * U[n] + DVy[n] * V[n]) + (vty[n] - DPy[n]) / R[n];
      dE[n] = -(DEx[n] * U[n] + DEy[n] * V[n]) +
              (-(P[n] * div) + tauxx[n] * DUx[n] + tauyy[n] * DVy[n] + tauxy[n] * (DVx[n] + DUy[n]) - htx[n] -
               hty[n]) / R[n];
    }
    if (dq[GATT_W]) memset(dq[GATT_W], 0, np * sizeof(double));
    forcing(p, q, dq, 0); /* runForcings2D.F90 (nseq.F90:137) */
  }
}

/* ------------------------------------------------------------------------- */
/* Boundary conditions: nseq.F90:327-599                                     */
/* ------------------------------------------------------------------------- */

/* getNeumannIndices (nseq.F90:520-550) */
static void neumann_indices(const int32_t* in, int dir, int L, int ind[6]) {
  for (int i = 0; i < 6; ++i) ind[i] = in[i];
  switch (dir) {
    case 1: ind[1] += L; break;
    case 2: ind[0] -= L; break;
    case 3: ind[3] += L; break;
    case 4: ind[2] -= L; break;
    case 5: ind[5] += L; break;
    case 6: ind[4] -= L; break;
  }
}

/* applyNeumann (nseq.F90:552-599) on the sub-array U(ind) of the full field */
static void apply_neumann(const gatt_problem_desc* p, double* U, const int ind[6], int dir, int L, const double* c) {
  const int ni = ind[1] - ind[0] + 1, nj = ind[3] - ind[2] + 1, nk = ind[5] - ind[4] + 1;
#define SUB(a, b, cc) U[IDX(p, ind[0] + (a) - 1, ind[2] + (b) - 1, ind[4] + (cc) - 1)]
  switch (dir) {
    case 1:
      for (int k = 1; k <= nk; ++k) for (int j = 1; j <= nj; ++j) {
        double v = c[0] * SUB(2, j, k);
        for (int i = 2; i <= L; ++i) v = v + c[i - 1] * SUB(i + 1, j, k);
        SUB(1, j, k) = v;
      }
      break;
    case 2:
      for (int k = 1; k <= nk; ++k) for (int j = 1; j <= nj; ++j) {
        double v = c[0] * SUB(L, j, k);
        for (int i = 2; i <= L; ++i) v = v + c[i - 1] * SUB(L - i + 1, j, k);
        SUB(L + 1, j, k) = v;
      }
      break;
    case 3:
      for (int k = 1; k <= nk; ++k) for (int i = 1; i <= ni; ++i) {
        double v = c[0] * SUB(i, 2, k);
        for (int j = 2; j <= L; ++j) v = v + c[j - 1] * SUB(i, j + 1, k);
        SUB(i, 1, k) = v;
      }
      break;
    case 4:
      for (int k = 1; k <= nk; ++k) for (int i = 1; i <= ni; ++i) {
        double v = c[0] * SUB(i, L, k);
        for (int j = 2; j <= L; ++j) v = v + c[j - 1] * SUB(i, L - j + 1, k);
        SUB(i, L + 1, k) = v;
      }
      break;
    case 5:
      for (int j = 1; j <= nj; ++j) for (int i = 1; i <= ni; ++i) {
        double v = c[0] * SUB(i, j, 2);
        for (int k = 2; k <= L; ++k) v = v + c[k - 1] * SUB(i, j, k + 1);
        SUB(i, j, 1) = v;
      }
      break;
    case 6:
      for (int j = 1; j <= nj; ++j) for (int i = 1; i <= ni; ++i) {
        double v = c[0] * SUB(i, j, L);
        for (int k = 2; k <= L; ++k) v = v + c[k - 1] * SUB(i, j, L - k + 1);
        SUB(i, j, L + 1) = v;
      }
      break;
  }
#undef SUB
}

#define BOXLOOP(bx, body)                                 \
  for (int kk = (bx)[4]; kk <= (bx)[5]; ++kk)             \
    for (int jj = (bx)[2]; jj <= (bx)[3]; ++jj)           \
      for (int ii = (bx)[0]; ii <= (bx)[1]; ++ii) {       \
        size_t n = IDX(p, ii, jj, kk);                    \
        body;                                             \
      }

static void box_of(const gatt_box_list* bl, int i, int out[6]) {
  for (int c = 0; c < 6; ++c) out[c] = bl->box[i + bl->count * c];
}

static double powi18(double x) { return __builtin_powi(x, 18); }

static void disturbances(oracle_ctx* c, double t, double* const* q) {
  const gatt_problem_desc* p = c->p;
  for (int d = 0; d < p->bc.nDisturb; ++d) {
    const gatt_disturbance* ds = &p->bc.disturb[d];
    const int* bx = ds->box;
    if (ds->kind == GATT_DIST_LID_DRIVEN) {
      /* lidDrivenMovement.F90:15-17 on U(box) */
      const int nxb = bx[1] - bx[0] + 1;
      const double* x = ds->x;
      const double U0 = ds->par[0];
      BOXLOOP(bx, {
        int i = ii - bx[0];
        double s = 2 * (x[i] - x[0]) / (x[nxb - 1] - x[0]) - 1;
        double v = 1 - powi18(s);
        q[GATT_U][n] = U0 * (v * v);
      });
    } else if (ds->kind == GATT_DIST_HOLD_INLET) {
      /* holdInlet.F90:18-48 */
      size_t nb = (size_t)(bx[1] - bx[0] + 1) * (bx[3] - bx[2] + 1) * (bx[5] - bx[4] + 1);
      const int hold_r = ds->par[0] > 0.5;
      const int vars[5] = {GATT_U, GATT_V, GATT_W, GATT_E, GATT_R};
      const int nv = hold_r ? 5 : 4;
      if (c->hold_first[d]) {
        c->hold_first[d] = 0;
        c->hold_save[d] = zalloc(5 * nb);
        size_t m = 0;
        BOXLOOP(bx, { for (int v = 0; v < nv; ++v) c->hold_save[d][v * nb + m] = q[vars[v]][n]; ++m; });
      } else {
        size_t m = 0;
        BOXLOOP(bx, { for (int v = 0; v < nv; ++v) q[vars[v]][n] = c->hold_save[d][v * nb + m]; ++m; });
      }
    } else if (ds->kind == GATT_DIST_PERIODIC_SINE) {
      /* periodicSine.F90:13-22; pi and alpha are default (single precision) reals */
      const int nxb = bx[1] - bx[0] + 1;
      const double* x = ds->x;
      const double omega = ds->par[0], amplitude = ds->par[1];
      const float pi = (float)(4.0 * atan(1.0));
      const float alpha = (float)(2.0 * (double)pi / (x[nxb - 1] - x[0]));
      double* u = q[ds->var];
      BOXLOOP(bx, {
        int i = ii - bx[0];
        u[n] = u[n] + amplitude * sin((double)alpha * (x[i] - x[0])) * sin(omega * t);
      });
    } else if (ds->kind == GATT_DIST_PACKET_2D) {
      /* packet_2d.F90:15-39; pi, x0, Tp, Tend, sigmaT, sigmaX are single precision and
       * aux is a single-precision complex.  Every point of an x station receives the
       * same modes: u <- amplitude * (u + sum_m ...), and nothing once t > Tend. */
      const int nxb = bx[1] - bx[0] + 1;
      const double* x = ds->x;
      const double omega = ds->par[0], nmodes = ds->par[1], amplitude = ds->par[2];
      const int nmode = (int)nmodes;
      const float pi = (float)(4.0 * atan(1.0));
      const float x0 = (float)((x[nxb - 1] + x[0]) / 2);
      const float Tend = (float)((2.0 * (double)pi / omega) * 2.5e-1);
      const float Tp = (float)((double)Tend * 5e-1);
      const float sigmaT = (float)((double)Tend * 2.5e-1);
      const float sigmaX = (float)((x[nxb - 1] - x[0]) / 8.0);
      if (t > (double)Tend) continue;
      double* u = q[ds->var];
      BOXLOOP(bx, {
        int i = ii - bx[0];
        double acc = u[n];
        for (int m = 1; m <= nmode; ++m) {
          const double th = (double)m * omega * (t - (double)Tp);
          const float aux_re = (float)cos(th);
          const double dT = t - (double)Tp;
          const double dX = x[i] - (double)x0;
          const double sT = (double)sigmaT;
          const double sX = (double)sigmaX;
          acc = acc + exp(-(dT * dT) / (2.0 * (sT * sT))) * exp(-(dX * dX) / (2.0 * (sX * sX))) * (double)aux_re /
                          nmodes;
        }
        u[n] = amplitude * acc;
      });
    }
  }
}

void oracle_apply_bound_cond(oracle_ctx* c, double t, double* const* q) {
  const gatt_problem_desc* p = c->p;
  const gatt_bc_desc* bc = &p->bc;
  const double gamma = p->gamma;
  double* U = q[GATT_U]; double* V = q[GATT_V]; double* W = q[GATT_W]; double* R = q[GATT_R]; double* E = q[GATT_E];
  double* P = c->work[63];
  const int L = bc->neumannLength, L2 = bc->neumann2Length;
  int bx[6], ind[6];
  /* Dirichlet: U, V, W, E, then P through R (nseq.F90:359-373) */
  double* dfield[4] = {U, V, W, E};
  const int dslot[4] = {GATT_BC_U, GATT_BC_V, GATT_BC_W, GATT_BC_E};
  for (int s = 0; s < 4; ++s) {
    const gatt_box_list* bl = &bc->dir[dslot[s]];
    for (int i = 0; i < bl->count; ++i) {
      box_of(bl, i, bx);
      double v = bl->value[i];
      BOXLOOP(bx, { dfield[s][n] = v; });
    }
  }
  {
    const gatt_box_list* bl = &bc->dir[GATT_BC_P];
    for (int i = 0; i < bl->count; ++i) {
      box_of(bl, i, bx);
      double v = bl->value[i];
      BOXLOOP(bx, { R[n] = v / (gamma - 1.0) / E[n]; });
    }
  }
  /* Neumann: U, V, W, E, then P (377-399) */
  for (int s = 0; s < 4; ++s) {
    const gatt_box_list* bl = &bc->neu[dslot[s]];
    for (int i = 0; i < bl->count; ++i) {
      box_of(bl, i, bx);
      neumann_indices(bx, bl->dir[i], L, ind);
      apply_neumann(p, dfield[s], ind, bl->dir[i], L, bc->neumannCoeffs);
    }
  }
  {
    const gatt_box_list* bl = &bc->neu[GATT_BC_P];
    for (int i = 0; i < bl->count; ++i) {
      box_of(bl, i, bx);
      neumann_indices(bx, bl->dir[i], L, ind);
      BOXLOOP(ind, { P[n] = R[n] * E[n]; });
      apply_neumann(p, P, ind, bl->dir[i], L, bc->neumannCoeffs);
      BOXLOOP(bx, { R[n] = P[n] / E[n]; });
    }
  }
  /* corners (402-488) */
  for (int i = 0; i < bc->nCorners; ++i) {
    int cl[6], cd[3];
    for (int a = 0; a < 6; ++a) cl[a] = bc->cL[i + bc->nCorners * a];
    for (int a = 0; a < 3; ++a) cd[a] = bc->cD[i + bc->nCorners * a];
    const double nsum = (double)(abs(cd[0]) + abs(cd[1]) + abs(cd[2]));
    if (bc->cAdiabatic[i] == 1) {
      BOXLOOP(cl, { E[n] = 0; });
      for (int a = 0; a < 3; ++a) {
        if (cd[a] == 0) continue;
        int d = cd[a] == 1 ? 2 * a + 1 : 2 * a + 2;
        neumann_indices(cl, d, L, ind);
        BOXLOOP(ind, { P[n] = E[n]; });
        apply_neumann(p, P, ind, d, L, bc->neumannCoeffs);
        BOXLOOP(cl, { E[n] = E[n] + P[n]; });
      }
      BOXLOOP(cl, { E[n] = E[n] / nsum; });
    }
    BOXLOOP(cl, { R[n] = 0; });
    for (int a = 0; a < 3; ++a) {
      if (cd[a] == 0) continue;
      int d = cd[a] == 1 ? 2 * a + 1 : 2 * a + 2;
      neumann_indices(cl, d, L, ind);
      BOXLOOP(ind, { P[n] = R[n] * E[n]; });
      apply_neumann(p, P, ind, d, L, bc->neumannCoeffs);
      BOXLOOP(cl, { R[n] = R[n] + P[n]; });
    }
    BOXLOOP(cl, { R[n] = R[n] / E[n] / nsum; });
  }
  /* null second derivative (491-513) */
  for (int s = 0; s < 4; ++s) {
    const gatt_box_list* bl = &bc->sec[dslot[s]];
    for (int i = 0; i < bl->count; ++i) {
      box_of(bl, i, bx);
      neumann_indices(bx, bl->dir[i], L2, ind);
      apply_neumann(p, dfield[s], ind, bl->dir[i], L2, bc->neumann2Coeffs);
    }
  }
  {
    const gatt_box_list* bl = &bc->sec[GATT_BC_P];
    for (int i = 0; i < bl->count; ++i) {
      box_of(bl, i, bx);
      neumann_indices(bx, bl->dir[i], L2, ind);
      BOXLOOP(ind, { P[n] = R[n] * E[n]; });
      apply_neumann(p, P, ind, bl->dir[i], L2, bc->neumann2Coeffs);
      BOXLOOP(bx, { R[n] = P[n] / E[n]; });
    }
  }
  /* disturbances (516) */
  disturbances(c, t, q);
}

/* ------------------------------------------------------------------------- */
/* RK4: timeIntegration.F90:8-156, 338-498                                   */
/* ------------------------------------------------------------------------- */

void oracle_rk4(oracle_ctx* c, double t, double dt, const double* const* q, double* const* qn) {
  const size_t np = c->np;
  const int is3d = c->p->nz > 1;
  double* d[5]; double* delta[5];
  for (int f = 0; f < 5; ++f) { d[f] = c->work[45 + f]; delta[f] = c->work[50 + f]; }
  /* In 2-D the stage BCs act on the old W (RK4_2D passes W, timeIntegration.F90:73). */
  double* stage[5];
  for (int f = 0; f < 5; ++f) stage[f] = qn[f];
  if (!is3d) stage[GATT_W] = (double*)q[GATT_W];
  const int nf = 5;
  oracle_nseq(c, q, d);
  for (int f = 0; f < nf; ++f) {
    if (!is3d && f == GATT_W) continue;
    for (size_t n = 0; n < np; ++n) {
      delta[f][n] = dt / 6.0 * d[f][n];
      stage[f][n] = q[f][n] + dt / 2.0 * d[f][n];
    }
  }
  oracle_apply_bound_cond(c, t + dt / 2.0, stage);
  oracle_nseq(c, (const double* const*)stage, d);
  for (int f = 0; f < nf; ++f) {
    if (!is3d && f == GATT_W) continue;
    for (size_t n = 0; n < np; ++n) {
      delta[f][n] = delta[f][n] + dt / 3.0 * d[f][n];
      stage[f][n] = q[f][n] + dt / 2.0 * d[f][n];
    }
  }
  oracle_apply_bound_cond(c, t + dt / 2.0, stage);
  oracle_nseq(c, (const double* const*)stage, d);
  for (int f = 0; f < nf; ++f) {
    if (!is3d && f == GATT_W) continue;
    for (size_t n = 0; n < np; ++n) {
      delta[f][n] = delta[f][n] + dt / 3.0 * d[f][n];
      stage[f][n] = q[f][n] + dt * d[f][n];
    }
  }
  oracle_apply_bound_cond(c, t + dt, stage);
  oracle_nseq(c, (const double* const*)stage, d);
  for (int f = 0; f < nf; ++f) {
    if (!is3d && f == GATT_W) continue;
    for (size_t n = 0; n < np; ++n) stage[f][n] = q[f][n] + delta[f][n] + dt / 6.0 * d[f][n];
  }
  oracle_apply_bound_cond(c, t + dt, stage);
}

/* EULER: timeIntegration.F90:158-222 (2-D), 500-566 (3-D).  EULER_3D applies the
 * boundary conditions at t, EULER_2D at t + dt; the boundary values here do not
 * depend on time, so both calls give the same state. */
void oracle_euler(oracle_ctx* c, double t, double dt, const double* const* q, double* const* qn) {
  const size_t np = c->np;
  const int is3d = c->p->nz > 1;
  double* d[5];
  for (int f = 0; f < 5; ++f) d[f] = c->work[45 + f];
  double* stage[5];
  for (int f = 0; f < 5; ++f) stage[f] = qn[f];
  if (!is3d) stage[GATT_W] = (double*)q[GATT_W];
  oracle_nseq(c, q, d);
  for (int f = 0; f < 5; ++f) {
    if (!is3d && f == GATT_W) continue;
    for (size_t n = 0; n < np; ++n) stage[f][n] = q[f][n] + dt * d[f][n];
  }
  oracle_apply_bound_cond(c, is3d ? t : t + dt, stage);
}

/* SSPRK3: timeIntegration.F90:225-335 (2-D), 569-685 (3-D).  The running sum k1 + k2
 * is kept so that the third update evaluates (k1 + k2) + 4 k3 in the reference's
 * left-to-right order, and every update is (dt * k) / c as written there. */
void oracle_ssprk3(oracle_ctx* c, double t, double dt, const double* const* q, double* const* qn) {
  const size_t np = c->np;
  const int is3d = c->p->nz > 1;
  double* d[5]; double* ks[5];
  for (int f = 0; f < 5; ++f) { d[f] = c->work[45 + f]; ks[f] = c->work[50 + f]; }
  double* stage[5];
  for (int f = 0; f < 5; ++f) stage[f] = qn[f];
  if (!is3d) stage[GATT_W] = (double*)q[GATT_W];
  oracle_nseq(c, q, d);
  for (int f = 0; f < 5; ++f) {
    if (!is3d && f == GATT_W) continue;
    for (size_t n = 0; n < np; ++n) {
      ks[f][n] = d[f][n];
      stage[f][n] = q[f][n] + dt * d[f][n];
    }
  }
  oracle_apply_bound_cond(c, t + dt, stage);
  oracle_nseq(c, (const double* const*)stage, d);
  for (int f = 0; f < 5; ++f) {
    if (!is3d && f == GATT_W) continue;
    for (size_t n = 0; n < np; ++n) {
      ks[f][n] = ks[f][n] + d[f][n];
      stage[f][n] = q[f][n] + dt * ks[f][n] / 4.0;
    }
  }
  oracle_apply_bound_cond(c, t + dt / 2.0, stage);
  oracle_nseq(c, (const double* const*)stage, d);
  for (int f = 0; f < 5; ++f) {
    if (!is3d && f == GATT_W) continue;
    for (size_t n = 0; n < np; ++n) stage[f][n] = q[f][n] + dt * (ks[f][n] + 4.0 * d[f][n]) / 6.0;
  }
  oracle_apply_bound_cond(c, t + dt, stage);
}

/* The integrator that timeStepping selects (main.F90:255-350): 2 Euler, 3 SSPRK3,
 * anything else RK4. */
void oracle_time_step(oracle_ctx* c, double t, double dt, const double* const* q, double* const* qn) {
  switch (c->p->timeStepping) {
    case 2: oracle_euler(c, t, dt, q, qn); break;
    case 3: oracle_ssprk3(c, t, dt, q, qn); break;
    default: oracle_rk4(c, t, dt, q, qn); break;
  }
}

/* ------------------------------------------------------------------------- */
/* SFD and the main-loop step: main.F90:119-165, 196-462                     */
/* ------------------------------------------------------------------------- */

void oracle_init_mean(oracle_ctx* c, const double* const* q) {
  for (int f = 0; f < 5; ++f) memcpy(c->mean[f], q[f], c->np * sizeof(double));
  c->mean_ready = 1;
}

void oracle_get_mean(oracle_ctx* c, double* const* out) {
  for (int f = 0; f < 5; ++f) memcpy(out[f], c->mean[f], c->np * sizeof(double));
}

void oracle_sfd(oracle_ctx* c, double dt, double* const* qn) {
  const gatt_problem_desc* p = c->p;
  if (p->sfdType <= 0) return;
  const size_t np = c->np;
  const int order[5] = {GATT_U, GATT_V, GATT_R, GATT_E, GATT_W};
  const int nf = p->nz > 1 ? 5 : 4;
  for (int m = 0; m < nf; ++m) {
    int f = order[m];
    double* M = c->mean[f];
    double* Q = qn[f];
    for (size_t n = 0; n < np; ++n) {
      const double X = p->sfdType == 1 ? p->sfdXval : p->sfdX[n];
      if (p->sfdDelta > 0) M[n] = M[n] + dt / p->sfdDelta * (Q[n] - M[n]);
      Q[n] = Q[n] - dt * X * (Q[n] - M[n]);
    }
  }
}

/* One iteration of the time loop with fixed dt; q is replaced by the new state. */
int oracle_step(oracle_ctx* c, double t, double dt, double* const* q, double* umax, double* maxchange) {
  const gatt_problem_desc* p = c->p;
  const size_t np = c->np;
  double* qn[5];
  for (int f = 0; f < 5; ++f) qn[f] = c->work[55 + f];
  if (p->sfdType > 0 && !c->mean_ready) oracle_init_mean(c, (const double* const*)q);
  double um[3] = {0, 0, 0};
  for (size_t n = 0; n < np; ++n) {
    um[0] = fmax(um[0], fabs(q[GATT_U][n]));
    um[1] = fmax(um[1], fabs(q[GATT_V][n]));
    if (p->nz > 1) um[2] = fmax(um[2], fabs(q[GATT_W][n]));
  }
  if (umax) memcpy(umax, um, sizeof(um));
  /* Wnew starts at zero in 2-D and is only touched by boundary conditions (main.F90:184-188) */
  if (p->nz == 1) memset(qn[GATT_W], 0, np * sizeof(double));
  oracle_time_step(c, t, dt, (const double* const*)q, qn);
  double theta = dt / p->filterCharTime;
  oracle_apply_filters(c, theta, qn);
  oracle_sfd(c, dt, qn);
  oracle_apply_bound_cond(c, t + dt, qn);
  double mc[5] = {0, 0, 0, 0, 0};
  int nan = 0;
  for (int f = 0; f < 5; ++f)
    for (size_t n = 0; n < np; ++n) {
      double d = fabs(qn[f][n] - q[f][n]);
      if (d != d) nan = 1;
      if (d > mc[f]) mc[f] = d;
    }
  if (maxchange) memcpy(maxchange, mc, sizeof(mc));
  const int ncopy = 5;
  for (int f = 0; f < ncopy; ++f) {
    if (p->nz == 1 && f == GATT_W) continue;
    memcpy(q[f], qn[f], np * sizeof(double));
  }
  return (nan || (mc[0] + mc[1] + mc[2] + mc[3] + mc[4]) == 0) ? 1 : 0;
}
