/* CPU oracle of the LPM solve path in C (FP64, OpenMP).  TEST INFRASTRUCTURE ONLY: it is used by tests/ as a
 * checker and by the cpu_baseline leg of bench.py as the timed CPU restatement of the reference's own solvers.
 * Nothing in the product links it.
 *
 * Each function follows the reference's Numba kernel of the same role, with the same neighbour table convention
 * (nbr (N, NN) int64, -1 = no bond) and the same loop structure:
 *   orc_apply3  coarse3d/lin3d.py:98-149   (k_shell_sums + k_apply, force form)
 *   orc_cg3     timing/t3d/cgplain.py:11-55 (cg3: unpreconditioned CG from x, ||r|| <= tol ||b||)
 *   orc_pcg3    coarse3d/lin3d.py:464-518  (_pcg: block-Jacobi PCG)
 *   orc_kmv2    timing/t2d/solver2d.py:269-305 (_kmv: energy-gradient form with a (N, 8) scratch)
 *   orc_cg2     timing/t2d/cg2plain.py:11-49 (cg_mf: per-DOF mask)
 * Reductions use OpenMP reduction clauses, so (as with Numba prange) the rounding depends on the thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const double ND3[18][3] = {
    {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
    {-0.70710678118654752440, -0.70710678118654752440, 0}, {0.70710678118654752440, 0.70710678118654752440, 0},
    {-0.70710678118654752440, 0.70710678118654752440, 0}, {0.70710678118654752440, -0.70710678118654752440, 0},
    {-0.70710678118654752440, 0, -0.70710678118654752440}, {0.70710678118654752440, 0, 0.70710678118654752440},
    {-0.70710678118654752440, 0, 0.70710678118654752440}, {0.70710678118654752440, 0, -0.70710678118654752440},
    {0, -0.70710678118654752440, -0.70710678118654752440}, {0, 0.70710678118654752440, 0.70710678118654752440},
    {0, -0.70710678118654752440, 0.70710678118654752440}, {0, 0.70710678118654752440, -0.70710678118654752440}};
static const double ND2[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1},
                                 {-0.70710678118654752440, -0.70710678118654752440},
                                 {0.70710678118654752440, 0.70710678118654752440},
                                 {-0.70710678118654752440, 0.70710678118654752440},
                                 {0.70710678118654752440, -0.70710678118654752440}};

int orc_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* thread count of the following parallel regions (the reference times its CPU solvers at LPMCG_THREADS = 6,
 * timing/t2d/a2common.py:10) */
void orc_set_threads(int n) {
#ifdef _OPENMP
  extern void omp_set_num_threads(int);
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* out = K u (3D).  S: scratch (N, 4) = S0, S1, TS0, TS1. */
void orc_apply3(int64_t N, const int64_t* nbr, const double* u, const double* Kn, const double* Tv, double* out,
                double* S) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    double a0 = 0, a1 = 0, b0 = 0, b1 = 0;
    for (int k = 0; k < 18; ++k) {
      const int64_t q = nbr[i * 18 + k];
      if (q < 0) continue;
      const double dl = ND3[k][0] * (u[3 * i] - u[3 * q]) + ND3[k][1] * (u[3 * i + 1] - u[3 * q + 1]) +
                        ND3[k][2] * (u[3 * i + 2] - u[3 * q + 2]);
      if (k < 6) {
        a0 += dl;
        b0 += Tv[i * 18 + k] * dl;
      } else {
        a1 += dl;
        b1 += Tv[i * 18 + k] * dl;
      }
    }
    S[4 * i + 0] = a0;
    S[4 * i + 1] = a1;
    S[4 * i + 2] = b0;
    S[4 * i + 3] = b1;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    double ox = 0, oy = 0, oz = 0;
    for (int k = 0; k < 18; ++k) {
      const int64_t q = nbr[i * 18 + k];
      if (q < 0) continue;
      const int c = k < 6 ? 0 : 1;
      const double dl = ND3[k][0] * (u[3 * i] - u[3 * q]) + ND3[k][1] * (u[3 * i + 1] - u[3 * q + 1]) +
                        ND3[k][2] * (u[3 * i + 2] - u[3 * q + 2]);
      const double f = 2.0 * Kn[i * 18 + k] * dl + 0.5 * (S[4 * i + 2 + c] + S[4 * q + 2 + c]) +
                       0.5 * Tv[i * 18 + k] * (S[4 * i + c] + S[4 * q + c]);
      ox += f * ND3[k][0];
      oy += f * ND3[k][1];
      oz += f * ND3[k][2];
    }
    out[3 * i] = ox;
    out[3 * i + 1] = oy;
    out[3 * i + 2] = oz;
  }
}

/* Unpreconditioned CG (cg3).  free (N,) uint8; b, x (N, 3) zero on fixed; returns iterations, *relres. */
int orc_cg3(int64_t N, const int64_t* nbr, const double* Kn, const double* Tv, const uint8_t* free_, const double* b,
            double* x, double tol, int maxit, double* relres) {
  double* r = malloc(sizeof(double) * 3 * N);
  double* p = malloc(sizeof(double) * 3 * N);
  double* Ap = malloc(sizeof(double) * 3 * N);
  double* S = malloc(sizeof(double) * 4 * N);
  orc_apply3(N, nbr, x, Kn, Tv, Ap, S);
  double bn = 0, rr = 0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) r[3 * i + d] = free_[i] ? b[3 * i + d] - Ap[3 * i + d] : 0.0;
#pragma omp parallel for reduction(+ : bn) schedule(static)
  for (int64_t i = 0; i < 3 * N; ++i) bn += b[i] * b[i];
  bn = sqrt(bn);
  int it = 0;
  if (bn == 0.0) {
    *relres = 0.0;
    goto done;
  }
  memcpy(p, r, sizeof(double) * 3 * N);
#pragma omp parallel for reduction(+ : rr) schedule(static)
  for (int64_t i = 0; i < 3 * N; ++i) rr += r[i] * r[i];
  for (it = 1; it <= maxit; ++it) {
    orc_apply3(N, nbr, p, Kn, Tv, Ap, S);
    double pAp = 0;
#pragma omp parallel for reduction(+ : pAp) schedule(static)
    for (int64_t i = 0; i < N; ++i)
      if (free_[i]) pAp += p[3 * i] * Ap[3 * i] + p[3 * i + 1] * Ap[3 * i + 1] + p[3 * i + 2] * Ap[3 * i + 2];
    const double alpha = rr / pAp;
    double rn = 0;
#pragma omp parallel for reduction(+ : rn) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      if (!free_[i]) continue;
      for (int d = 0; d < 3; ++d) {
        x[3 * i + d] += alpha * p[3 * i + d];
        r[3 * i + d] -= alpha * Ap[3 * i + d];
      }
      rn += r[3 * i] * r[3 * i] + r[3 * i + 1] * r[3 * i + 1] + r[3 * i + 2] * r[3 * i + 2];
    }
    if (sqrt(rn) <= tol * bn) {
      rr = rn;
      break;
    }
    const double beta = rn / rr;
    rr = rn;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i)
      if (free_[i])
        for (int d = 0; d < 3; ++d) p[3 * i + d] = r[3 * i + d] + beta * p[3 * i + d];
  }
  if (it > maxit) it = maxit;
  *relres = sqrt(rr) / bn;
done:
  free(r);
  free(p);
  free(Ap);
  free(S);
  return it;
}

/* Block-Jacobi PCG (_pcg).  M (N, 3, 3) zero on fixed particles; x starts at the given values. */
int orc_pcg3(int64_t N, const int64_t* nbr, const double* Kn, const double* Tv, const double* M,
             const uint8_t* free_, const double* b, double* x, double tol, int maxit, double* relres) {
  double* r = malloc(sizeof(double) * 3 * N);
  double* z = malloc(sizeof(double) * 3 * N);
  double* p = malloc(sizeof(double) * 3 * N);
  double* Ap = malloc(sizeof(double) * 3 * N);
  double* S = malloc(sizeof(double) * 4 * N);
  int it = 0;
  double rn = 0.0;
  orc_apply3(N, nbr, x, Kn, Tv, Ap, S);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) r[3 * i + d] = free_[i] ? b[3 * i + d] - Ap[3 * i + d] : 0.0;
  double bn = 0;
#pragma omp parallel for reduction(+ : bn) schedule(static)
  for (int64_t i = 0; i < 3 * N; ++i) bn += b[i] * b[i];
  bn = sqrt(bn);
  if (bn == 0.0) {
    *relres = 0.0;
    goto done;
  }
  double rz = 0;
#pragma omp parallel for reduction(+ : rz) schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    for (int d = 0; d < 3; ++d)
      z[3 * i + d] = M[9 * i + 3 * d] * r[3 * i] + M[9 * i + 3 * d + 1] * r[3 * i + 1] + M[9 * i + 3 * d + 2] * r[3 * i + 2];
    rz += r[3 * i] * z[3 * i] + r[3 * i + 1] * z[3 * i + 1] + r[3 * i + 2] * z[3 * i + 2];
  }
  memcpy(p, z, sizeof(double) * 3 * N);
  for (it = 1; it <= maxit; ++it) {
    orc_apply3(N, nbr, p, Kn, Tv, Ap, S);
    double pAp = 0;
#pragma omp parallel for reduction(+ : pAp) schedule(static)
    for (int64_t i = 0; i < N; ++i)
      if (free_[i]) pAp += p[3 * i] * Ap[3 * i] + p[3 * i + 1] * Ap[3 * i + 1] + p[3 * i + 2] * Ap[3 * i + 2];
    const double alpha = rz / pAp;
    rn = 0;
#pragma omp parallel for reduction(+ : rn) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      if (!free_[i]) continue;
      for (int d = 0; d < 3; ++d) {
        x[3 * i + d] += alpha * p[3 * i + d];
        r[3 * i + d] -= alpha * Ap[3 * i + d];
      }
      rn += r[3 * i] * r[3 * i] + r[3 * i + 1] * r[3 * i + 1] + r[3 * i + 2] * r[3 * i + 2];
    }
    rn = sqrt(rn);
    if (rn <= tol * bn) break;
    double rz2 = 0;
#pragma omp parallel for reduction(+ : rz2) schedule(static)
    for (int64_t i = 0; i < N; ++i) {
      for (int d = 0; d < 3; ++d)
        z[3 * i + d] =
            M[9 * i + 3 * d] * r[3 * i] + M[9 * i + 3 * d + 1] * r[3 * i + 1] + M[9 * i + 3 * d + 2] * r[3 * i + 2];
      rz2 += r[3 * i] * z[3 * i] + r[3 * i + 1] * z[3 * i + 1] + r[3 * i + 2] * z[3 * i + 2];
    }
    const double beta = rz2 / rz;
    rz = rz2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 3 * N; ++i) p[i] = z[i] + beta * p[i];
  }
  if (it > maxit) it = maxit;
  *relres = rn / bn;
done:
  free(r);
  free(z);
  free(p);
  free(Ap);
  free(S);
  return it;
}

/* f = K u (2D, energy-gradient form).  g: scratch (N, 8). */
void orc_kmv2(int64_t N, const int64_t* nbr, const double* u, const double* Kn, const double* Tv, double* g,
              double* f) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    double S0 = 0, S1 = 0, T0 = 0, T1 = 0;
    for (int k = 0; k < 8; ++k) {
      const int64_t q = nbr[i * 8 + k];
      if (q < 0) continue;
      const double d = ND2[k][0] * (u[2 * i] - u[2 * q]) + ND2[k][1] * (u[2 * i + 1] - u[2 * q + 1]);
      g[i * 8 + k] = d;
      if (k < 4) {
        S0 += d;
        T0 += Tv[i * 8 + k] * d;
      } else {
        S1 += d;
        T1 += Tv[i * 8 + k] * d;
      }
    }
    for (int k = 0; k < 8; ++k) {
      if (nbr[i * 8 + k] < 0) continue;
      const double d = g[i * 8 + k];
      g[i * 8 + k] = k < 4 ? Kn[i * 8 + k] * d + 0.5 * (Tv[i * 8 + k] * S0 + T0)
                           : Kn[i * 8 + k] * d + 0.5 * (Tv[i * 8 + k] * S1 + T1);
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < N; ++j) {
    double fx = 0, fy = 0;
    for (int k = 0; k < 8; ++k) {
      const int64_t q = nbr[j * 8 + k];
      if (q < 0) continue;
      const double t = g[j * 8 + k] + g[q * 8 + (k ^ 1)];
      fx += t * ND2[k][0];
      fy += t * ND2[k][1];
    }
    f[2 * j] = fx;
    f[2 * j + 1] = fy;
  }
}

/* Unpreconditioned CG with a per-DOF mask (cg_mf).  mask (2N,) 1 free / 0 fixed; returns iterations. */
int orc_cg2(int64_t N, const int64_t* nbr, const double* Kn, const double* Tv, const double* mask, const double* b,
            double* x, double tol, int maxit, double* relres) {
  const int64_t n = 2 * N;
  double* r = malloc(sizeof(double) * n);
  double* p = malloc(sizeof(double) * n);
  double* Ap = malloc(sizeof(double) * n);
  double* g = malloc(sizeof(double) * 8 * N);
  orc_kmv2(N, nbr, x, Kn, Tv, g, Ap);
  double bb = 0, rr = 0;
#pragma omp parallel for reduction(+ : bb, rr) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    r[i] = (b[i] - Ap[i]) * mask[i];
    bb += b[i] * b[i];
    rr += r[i] * r[i];
  }
  const double bn = sqrt(bb);
  int it = 0;
  if (sqrt(rr) <= tol * bn) goto done;
  memcpy(p, r, sizeof(double) * n);
  while (it < maxit) {
    orc_kmv2(N, nbr, p, Kn, Tv, g, Ap);
    double pAp = 0;
#pragma omp parallel for reduction(+ : pAp) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      Ap[i] *= mask[i];
      pAp += p[i] * Ap[i];
    }
    const double alpha = rr / pAp;
    double rn = 0;
#pragma omp parallel for reduction(+ : rn) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
      rn += r[i] * r[i];
    }
    ++it;
    if (sqrt(rn) <= tol * bn) {
      rr = rn;
      break;
    }
    const double beta = rn / rr;
    rr = rn;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
  }
done:
  *relres = bn > 0 ? sqrt(rr) / bn : 0.0;
  free(r);
  free(p);
  free(Ap);
  free(g);
  return it;
}
