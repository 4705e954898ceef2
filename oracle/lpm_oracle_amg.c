/* CPU restatement of the reference's aggregation AMG-PCG (timing/t3d/amg3d.py) in C (FP64, OpenMP).  TEST
 * INFRASTRUCTURE ONLY: the cpu_baseline of bench.py's amg256 workload times it on the host cores; nothing in the
 * product links it.
 *
 *   orc_amg0_restrict / orc_amg0_prolong   _restrict / _prolong_add on the particles (amg3d.py:303-331), the blocks
 *                                           read from shape templates (Gram-Schmidt blocks depend on the shape only)
 *   orc_amg0_probe / orc_amg0_gather       level-1 Galerkin stencil P^T K P by probing with orc_apply3 (27 colours x
 *                                           6 modes) in place of the bond-level _galerkin0 (amg3d.py:151-196)
 *   orc_st_mv                               _st_mv / _sub (amg3d.py:230-253, 373-379)
 *   orc_q_galerkin                          _st_galerkin (amg3d.py:256-300)
 *   orc_q_restrict / orc_q_prolong          _restrict / _prolong_add on the node levels
 *   orc_inv6                                np.linalg.inv(A[:, 13]) (amg3d.py:536)
 *   orc_amg_pcg                             vcycle / _coarse / _vc / pcg (amg3d.py:578-672)
 * Stencils are (n, 27, 6, 6) with d = (dz+1) 9 + (dy+1) 3 + (dx+1) as the reference's.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void orc_apply3(int64_t N, const int64_t* nbr, const double* u, const double* Kn, const double* Tv, double* out,
                double* S);

typedef struct {
  int64_t nx, ny, nz;
  int32_t lo[3], L[3], g0[3], agg;
  const double* tmpl;
  int64_t toff[8];
} orc_amg0;

#define ORC_MAXLEV 12
typedef struct {
  orc_amg0 l0;
  const double* M0; /* (N, 3, 3) particle block-Jacobi inverse, zero on fixed particles */
  double w0;
  int32_t nlev, nu, nu_c, gamma, w_from, pad_;
  int32_t g[ORC_MAXLEV][3];
  const double* A[ORC_MAXLEV];    /* (n_k, 27, 6, 6) */
  const double* Q[ORC_MAXLEV];    /* (n_k, 6, 6) towards level k + 1 */
  const double* Minv[ORC_MAXLEV]; /* (n_k, 6, 6) */
  double w[ORC_MAXLEV];
  const double* dinv; /* dense inverse of the coarsest stencil */
} orc_amg;

static inline int chunk_of(int rel, int agg, int nch) {
  const int c = rel / agg;
  return c < nch - 1 ? c : nch - 1;
}
static inline int chunk_len(int c, int agg, int nch, int L) { return c == nch - 1 ? L - agg * (nch - 1) : agg; }

/* template row of particle (x, y, z); -1 outside the free box; *I the aggregate, cc its coordinates */
static inline const double* p0_row(const orc_amg0* a, int64_t x, int64_t y, int64_t z, int64_t* I, int cc[3]) {
  const int64_t p[3] = {x - a->lo[0], y - a->lo[1], z - a->lo[2]};
  int loc[3], len[3], s = 0;
  for (int k = 0; k < 3; ++k) {
    if (p[k] < 0 || p[k] >= a->L[k]) return NULL;
    cc[k] = chunk_of((int)p[k], a->agg, a->g0[k]);
    loc[k] = (int)p[k] - a->agg * cc[k];
    len[k] = chunk_len(cc[k], a->agg, a->g0[k], a->L[k]);
    if (cc[k] == a->g0[k] - 1) s |= 1 << k;
  }
  *I = ((int64_t)cc[2] * a->g0[1] + cc[1]) * a->g0[0] + cc[0];
  return a->tmpl + a->toff[s] + (int64_t)((loc[2] * len[1] + loc[1]) * len[0] + loc[0]) * 18;
}

void orc_amg0_restrict(const orc_amg0* a, const double* r, double* rc) {
  const int64_t nagg = (int64_t)a->g0[0] * a->g0[1] * a->g0[2];
#pragma omp parallel for schedule(static)
  for (int64_t I = 0; I < nagg; ++I) {
    const int cc[3] = {(int)(I % a->g0[0]), (int)((I / a->g0[0]) % a->g0[1]), (int)(I / ((int64_t)a->g0[0] * a->g0[1]))};
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int s0[3] = {a->lo[0] + a->agg * cc[0], a->lo[1] + a->agg * cc[1], a->lo[2] + a->agg * cc[2]};
    int len[3], sh = 0;
    for (int k = 0; k < 3; ++k) {
      len[k] = chunk_len(cc[k], a->agg, a->g0[k], a->L[k]);
      if (cc[k] == a->g0[k] - 1) sh |= 1 << k;
    }
    int row = 0;
    for (int z = 0; z < len[2]; ++z)
      for (int y = 0; y < len[1]; ++y)
        for (int x = 0; x < len[0]; ++x, ++row) {
          const int64_t p = ((int64_t)(s0[2] + z) * a->ny + (s0[1] + y)) * a->nx + (s0[0] + x);
          const double* pr = a->tmpl + a->toff[sh] + (int64_t)row * 18;
          for (int c = 0; c < 6; ++c) acc[c] += pr[c] * r[3 * p] + pr[6 + c] * r[3 * p + 1] + pr[12 + c] * r[3 * p + 2];
        }
    for (int c = 0; c < 6; ++c) rc[6 * I + c] = acc[c];
  }
}

void orc_amg0_prolong(const orc_amg0* a, const double* e, double* x) {
  const int64_t N = a->nx * a->ny * a->nz;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < N; ++p) {
    int64_t I;
    int cc[3];
    const double* pr = p0_row(a, p % a->nx, (p / a->nx) % a->ny, p / (a->nx * a->ny), &I, cc);
    if (!pr) continue;
    for (int r = 0; r < 3; ++r) {
      double s = 0;
      for (int c = 0; c < 6; ++c) s += pr[6 * r + c] * e[6 * I + c];
      x[3 * p + r] += s;
    }
  }
}

void orc_amg0_probe(const orc_amg0* a, int col, double* v) {
  const int64_t N = a->nx * a->ny * a->nz;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < N; ++p) {
    int64_t I;
    int cc[3];
    const double* pr = p0_row(a, p % a->nx, (p / a->nx) % a->ny, p / (a->nx * a->ny), &I, cc);
    const int on = pr && cc[0] % 3 == col % 3 && cc[1] % 3 == (col / 3) % 3 && cc[2] % 3 == col / 9;
    for (int j = 0; j < 6; ++j)
      for (int r = 0; r < 3; ++r) v[(int64_t)j * N * 3 + 3 * p + r] = on ? pr[6 * r + j] : 0.0;
  }
}

static inline int partner(int ic, int col, int g) {
  for (int o = -1; o <= 1; ++o) {
    const int j = ic + o;
    if (j >= 0 && j < g && j % 3 == col) return j;
  }
  return -1;
}

void orc_amg0_gather(const orc_amg0* a, int col, const double* kv, double* A) {
  const int64_t nagg = (int64_t)a->g0[0] * a->g0[1] * a->g0[2], N = a->nx * a->ny * a->nz;
#pragma omp parallel for schedule(static)
  for (int64_t I = 0; I < nagg; ++I) {
    const int cc[3] = {(int)(I % a->g0[0]), (int)((I / a->g0[0]) % a->g0[1]), (int)(I / ((int64_t)a->g0[0] * a->g0[1]))};
    const int cols[3] = {col % 3, (col / 3) % 3, col / 9};
    int jc[3], ok = 1;
    for (int k = 0; k < 3; ++k) {
      jc[k] = partner(cc[k], cols[k], a->g0[k]);
      if (jc[k] < 0) ok = 0;
    }
    if (!ok) continue;
    const int d = (jc[2] - cc[2] + 1) * 9 + (jc[1] - cc[1] + 1) * 3 + (jc[0] - cc[0] + 1);
    const int s0[3] = {a->lo[0] + a->agg * cc[0], a->lo[1] + a->agg * cc[1], a->lo[2] + a->agg * cc[2]};
    int len[3], sh = 0;
    for (int k = 0; k < 3; ++k) {
      len[k] = chunk_len(cc[k], a->agg, a->g0[k], a->L[k]);
      if (cc[k] == a->g0[k] - 1) sh |= 1 << k;
    }
    for (int j = 0; j < 6; ++j) {
      const double* w = kv + (int64_t)j * N * 3;
      double acc[6] = {0, 0, 0, 0, 0, 0};
      int row = 0;
      for (int z = 0; z < len[2]; ++z)
        for (int y = 0; y < len[1]; ++y)
          for (int x = 0; x < len[0]; ++x, ++row) {
            const int64_t p = ((int64_t)(s0[2] + z) * a->ny + (s0[1] + y)) * a->nx + (s0[0] + x);
            const double* pr = a->tmpl + a->toff[sh] + (int64_t)row * 18;
            for (int r = 0; r < 6; ++r) acc[r] += pr[r] * w[3 * p] + pr[6 + r] * w[3 * p + 1] + pr[12 + r] * w[3 * p + 2];
          }
      for (int r = 0; r < 6; ++r) A[((I * 27 + d) * 6 + r) * 6 + j] = acc[r];
    }
  }
}

void orc_st_mv(const int32_t g[3], const double* A, const double* x, const double* b, double* y) {
  const int64_t n = (int64_t)g[0] * g[1] * g[2];
#pragma omp parallel for schedule(static)
  for (int64_t I = 0; I < n; ++I) {
    const int i0 = (int)(I % g[0]), i1 = (int)((I / g[0]) % g[1]), i2 = (int)(I / ((int64_t)g[0] * g[1]));
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int d = 0; d < 27; ++d) {
      const int j0 = i0 + d % 3 - 1, j1 = i1 + (d / 3) % 3 - 1, j2 = i2 + d / 9 - 1;
      if (j0 < 0 || j0 >= g[0] || j1 < 0 || j1 >= g[1] || j2 < 0 || j2 >= g[2]) continue;
      const int64_t J = ((int64_t)j2 * g[1] + j1) * g[0] + j0;
      const double* Ad = A + (I * 27 + d) * 36;
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) acc[r] += Ad[6 * r + c] * x[6 * J + c];
    }
    for (int r = 0; r < 6; ++r) y[6 * I + r] = b ? b[6 * I + r] - acc[r] : acc[r];
  }
}

static inline void node_agg(const int32_t g[3], const int32_t gc[3], int agg, int64_t n, int64_t* K, int kc[3]) {
  const int c[3] = {(int)(n % g[0]), (int)((n / g[0]) % g[1]), (int)(n / ((int64_t)g[0] * g[1]))};
  for (int k = 0; k < 3; ++k) kc[k] = chunk_of(c[k], agg, gc[k]);
  *K = ((int64_t)kc[2] * gc[1] + kc[1]) * gc[0] + kc[0];
}

void orc_q_restrict(const int32_t g[3], const int32_t gc[3], int agg, const double* Q, const double* r, double* rc) {
  const int64_t nc = (int64_t)gc[0] * gc[1] * gc[2];
#pragma omp parallel for schedule(static)
  for (int64_t K = 0; K < nc; ++K) {
    const int kc[3] = {(int)(K % gc[0]), (int)((K / gc[0]) % gc[1]), (int)(K / ((int64_t)gc[0] * gc[1]))};
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int z = 0; z < chunk_len(kc[2], agg, gc[2], g[2]); ++z)
      for (int y = 0; y < chunk_len(kc[1], agg, gc[1], g[1]); ++y)
        for (int x = 0; x < chunk_len(kc[0], agg, gc[0], g[0]); ++x) {
          const int64_t m = ((int64_t)(agg * kc[2] + z) * g[1] + (agg * kc[1] + y)) * g[0] + (agg * kc[0] + x);
          for (int a_ = 0; a_ < 6; ++a_)
            for (int c = 0; c < 6; ++c) acc[c] += Q[36 * m + 6 * a_ + c] * r[6 * m + a_];
        }
    for (int c = 0; c < 6; ++c) rc[6 * K + c] = acc[c];
  }
}

void orc_q_prolong(const int32_t g[3], const int32_t gc[3], int agg, const double* Q, const double* e, double* x) {
  const int64_t n = (int64_t)g[0] * g[1] * g[2];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int64_t K;
    int kc[3];
    node_agg(g, gc, agg, i, &K, kc);
    for (int a_ = 0; a_ < 6; ++a_) {
      double s = 0;
      for (int c = 0; c < 6; ++c) s += Q[36 * i + 6 * a_ + c] * e[6 * K + c];
      x[6 * i + a_] += s;
    }
  }
}

/* Ac[K, dd] = sum_{n in K, m in J} Q_n^T A_nm Q_m (amg3d.py:256-300) */
void orc_q_galerkin(const int32_t g[3], const int32_t gc[3], int agg, const double* A, const double* Q, double* Ac) {
  const int64_t nc = (int64_t)gc[0] * gc[1] * gc[2];
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t K = 0; K < nc; ++K) {
    const int kc[3] = {(int)(K % gc[0]), (int)((K / gc[0]) % gc[1]), (int)(K / ((int64_t)gc[0] * gc[1]))};
    double* out = Ac + K * 27 * 36;
    memset(out, 0, sizeof(double) * 27 * 36);
    for (int z = 0; z < chunk_len(kc[2], agg, gc[2], g[2]); ++z)
      for (int y = 0; y < chunk_len(kc[1], agg, gc[1], g[1]); ++y)
        for (int x = 0; x < chunk_len(kc[0], agg, gc[0], g[0]); ++x) {
          const int c0 = agg * kc[0] + x, c1 = agg * kc[1] + y, c2 = agg * kc[2] + z;
          const int64_t nn = ((int64_t)c2 * g[1] + c1) * g[0] + c0;
          for (int d = 0; d < 27; ++d) {
            const int j0 = c0 + d % 3 - 1, j1 = c1 + (d / 3) % 3 - 1, j2 = c2 + d / 9 - 1;
            if (j0 < 0 || j0 >= g[0] || j1 < 0 || j1 >= g[1] || j2 < 0 || j2 >= g[2]) continue;
            const int64_t m = ((int64_t)j2 * g[1] + j1) * g[0] + j0;
            const int q0 = chunk_of(j0, agg, gc[0]), q1 = chunk_of(j1, agg, gc[1]), q2 = chunk_of(j2, agg, gc[2]);
            const int dd = (q2 - kc[2] + 1) * 9 + (q1 - kc[1] + 1) * 3 + (q0 - kc[0] + 1);
            const double* Anm = A + (nn * 27 + d) * 36;
            double C[36];
            for (int a_ = 0; a_ < 6; ++a_)
              for (int c = 0; c < 6; ++c) {
                double s = 0;
                for (int b_ = 0; b_ < 6; ++b_) s += Anm[6 * a_ + b_] * Q[36 * m + 6 * b_ + c];
                C[6 * a_ + c] = s;
              }
            for (int r = 0; r < 6; ++r)
              for (int c = 0; c < 6; ++c) {
                double s = 0;
                for (int a_ = 0; a_ < 6; ++a_) s += Q[36 * nn + 6 * a_ + r] * C[6 * a_ + c];
                out[dd * 36 + 6 * r + c] += s;
              }
          }
        }
  }
}

int orc_inv6(int64_t n, const double* A, double* Minv) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int64_t I = 0; I < n; ++I) {
    double m[6][12];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) {
        m[r][c] = A[(I * 27 + 13) * 36 + 6 * r + c];
        m[r][6 + c] = r == c;
      }
    for (int k = 0; k < 6 && !bad; ++k) {
      int pv = k;
      for (int r = k + 1; r < 6; ++r)
        if (fabs(m[r][k]) > fabs(m[pv][k])) pv = r;
      if (!(fabs(m[pv][k]) > 0)) {
        bad = 1;
        break;
      }
      for (int c = 0; c < 12; ++c) {
        const double t = m[k][c];
        m[k][c] = m[pv][c];
        m[pv][c] = t;
      }
      const double ip = 1.0 / m[k][k];
      for (int c = 0; c < 12; ++c) m[k][c] *= ip;
      for (int r = 0; r < 6; ++r) {
        if (r == k) continue;
        const double f = m[r][k];
        for (int c = 0; c < 12; ++c) m[r][c] -= f * m[k][c];
      }
    }
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) Minv[36 * I + 6 * r + c] = m[r][6 + c];
  }
  return bad;
}

/* ---------------------------------------------------------------------------------------------- cycle */
typedef struct {
  const orc_amg* a;
  int64_t N;
  const int64_t* nbr;
  const double* Kn;
  const double* Tv;
  const uint8_t* free_;
  double *r0, *S;
  double *b[ORC_MAXLEV], *e[ORC_MAXLEV], *t[ORC_MAXLEV], *x[ORC_MAXLEV], *r[ORC_MAXLEV];
} cyc;

static int64_t nodes(const int32_t g[3]) { return (int64_t)g[0] * g[1] * g[2]; }

static void bj6(int64_t n, const double* M, const double* b, double w, int add, double* x) {
#pragma omp parallel for schedule(static)
  for (int64_t I = 0; I < n; ++I)
    for (int r = 0; r < 6; ++r) {
      double s = 0;
      for (int c = 0; c < 6; ++c) s += M[36 * I + 6 * r + c] * b[6 * I + c];
      x[6 * I + r] = add ? x[6 * I + r] + w * s : w * s;
    }
}

static void bj3(int64_t N, const double* M, const double* b, double w, int add, double* x) {
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < N; ++p)
    for (int r = 0; r < 3; ++r) {
      const double s = M[9 * p + 3 * r] * b[3 * p] + M[9 * p + 3 * r + 1] * b[3 * p + 1] + M[9 * p + 3 * r + 2] * b[3 * p + 2];
      x[3 * p + r] = add ? x[3 * p + r] + w * s : w * s;
    }
}

/* out = mask (b - K u) */
static void resid0(cyc* c, const double* b, const double* u, double* out) {
  orc_apply3(c->N, c->nbr, u, c->Kn, c->Tv, out, c->S);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < c->N; ++p)
    for (int d = 0; d < 3; ++d) out[3 * p + d] = c->free_[p] ? b[3 * p + d] - out[3 * p + d] : 0.0;
}

static void coarse(cyc* c, int k, const double* rhs, double* ek);

static void vc(cyc* c, int k, const double* rhs, double* out) {
  const orc_amg* a = c->a;
  const int64_t n = nodes(a->g[k]);
  if (k == a->nlev - 1) {
    const int64_t m = 6 * n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
      double s = 0;
      for (int64_t j = 0; j < m; ++j) s += a->dinv[i * m + j] * rhs[j];
      out[i] = s;
    }
    return;
  }
  bj6(n, a->Minv[k], rhs, a->w[k], 0, out);
  for (int it = 0; it < a->nu_c - 1; ++it) {
    orc_st_mv(a->g[k], a->A[k], out, rhs, c->r[k]);
    bj6(n, a->Minv[k], c->r[k], a->w[k], 1, out);
  }
  orc_st_mv(a->g[k], a->A[k], out, rhs, c->r[k]);
  orc_q_restrict(a->g[k], a->g[k + 1], a->l0.agg, a->Q[k], c->r[k], c->b[k + 1]);
  coarse(c, k + 1, c->b[k + 1], c->e[k + 1]);
  orc_q_prolong(a->g[k], a->g[k + 1], a->l0.agg, a->Q[k], c->e[k + 1], out);
  for (int it = 0; it < a->nu_c; ++it) {
    orc_st_mv(a->g[k], a->A[k], out, rhs, c->r[k]);
    bj6(n, a->Minv[k], c->r[k], a->w[k], 1, out);
  }
}

static void coarse(cyc* c, int k, const double* rhs, double* ek) {
  const orc_amg* a = c->a;
  vc(c, k, rhs, ek);
  if (k == a->nlev - 1) return;
  const int extra = (k + 1) >= a->w_from ? a->gamma - 1 : 0;
  const int64_t n6 = 6 * nodes(a->g[k]);
  for (int it = 0; it < extra; ++it) {
    orc_st_mv(a->g[k], a->A[k], ek, rhs, c->t[k]);
    vc(c, k, c->t[k], c->x[k]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n6; ++i) ek[i] += c->x[k][i];
  }
}

static void fine(cyc* c, const double* b, double* z) {
  const orc_amg* a = c->a;
  bj3(c->N, a->M0, b, a->w0, 0, z);
  for (int it = 0; it < a->nu - 1; ++it) {
    resid0(c, b, z, c->r0);
    bj3(c->N, a->M0, c->r0, a->w0, 1, z);
  }
  resid0(c, b, z, c->r0);
  orc_amg0_restrict(&a->l0, c->r0, c->b[0]);
  coarse(c, 0, c->b[0], c->e[0]);
  orc_amg0_prolong(&a->l0, c->e[0], z);
  for (int it = 0; it < a->nu; ++it) {
    resid0(c, b, z, c->r0);
    bj3(c->N, a->M0, c->r0, a->w0, 1, z);
  }
}

static double dot(int64_t n, const double* x, const double* y) {
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
  return s;
}

/* AMG-PCG from x = 0 (amg3d.py:649-672); b (N, 3) zero on fixed particles.  Returns the iterations done. */
int orc_amg_pcg(const orc_amg* a, int64_t N, const int64_t* nbr, const double* Kn, const double* Tv,
                const uint8_t* free_, const double* b, double* x, double tol, int maxit, double* relres) {
  cyc c;
  memset(&c, 0, sizeof c);
  c.a = a;
  c.N = N;
  c.nbr = nbr;
  c.Kn = Kn;
  c.Tv = Tv;
  c.free_ = free_;
  const int64_t n3 = 3 * N;
  c.r0 = malloc(sizeof(double) * n3);
  c.S = malloc(sizeof(double) * 4 * N);
  for (int k = 0; k < a->nlev; ++k) {
    const int64_t n6 = 6 * nodes(a->g[k]);
    c.b[k] = malloc(sizeof(double) * n6);
    c.e[k] = malloc(sizeof(double) * n6);
    c.t[k] = malloc(sizeof(double) * n6);
    c.x[k] = malloc(sizeof(double) * n6);
    c.r[k] = malloc(sizeof(double) * n6);
  }
  double* r = malloc(sizeof(double) * n3);
  double* z = malloc(sizeof(double) * n3);
  double* p = malloc(sizeof(double) * n3);
  double* Ap = malloc(sizeof(double) * n3);
  memset(x, 0, sizeof(double) * n3);
  memcpy(r, b, sizeof(double) * n3);
  const double bn = sqrt(dot(n3, b, b));
  int it = 0;
  double rn = bn;
  if (bn > 0) {
    fine(&c, r, z);
    memcpy(p, z, sizeof(double) * n3);
    double rz = dot(n3, r, z);
    for (it = 1; it <= maxit; ++it) {
      orc_apply3(N, nbr, p, Kn, Tv, Ap, c.S);
#pragma omp parallel for schedule(static)
      for (int64_t q = 0; q < N; ++q)
        if (!free_[q]) Ap[3 * q] = Ap[3 * q + 1] = Ap[3 * q + 2] = 0.0;
      const double alpha = rz / dot(n3, p, Ap);
      double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
      for (int64_t i = 0; i < n3; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * Ap[i];
        s += r[i] * r[i];
      }
      rn = sqrt(s);
      if (rn <= tol * bn) break;
      fine(&c, r, z);
      const double rz2 = dot(n3, r, z);
      const double beta = rz2 / rz;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n3; ++i) p[i] = z[i] + beta * p[i];
      rz = rz2;
    }
    if (it > maxit) it = maxit;
  }
  *relres = bn > 0 ? rn / bn : 0.0;
  free(r);
  free(z);
  free(p);
  free(Ap);
  free(c.r0);
  free(c.S);
  for (int k = 0; k < a->nlev; ++k) {
    free(c.b[k]);
    free(c.e[k]);
    free(c.t[k]);
    free(c.x[k]);
    free(c.r[k]);
  }
  return it;
}

/* rho(M0 K_ff) by power iteration from x = start (amg3d.py:493-505): x = M0 start; repeat: y = M0 K_ff x / |x|,
 * lam = |y|, x = y.  M0 (N, 3, 3) is zero on fixed particles. */
double orc_power_rho0(int64_t N, const int64_t* nbr, const double* Kn, const double* Tv, const double* M0,
                      const uint8_t* free_, const double* start, int its) {
  const int64_t n3 = 3 * N;
  double* x = malloc(sizeof(double) * n3);
  double* y = malloc(sizeof(double) * n3);
  double* S = malloc(sizeof(double) * 4 * N);
  bj3(N, M0, start, 1.0, 0, x);
  double lam = 1.0;
  for (int it = 0; it < its; ++it) {
    const double nx = sqrt(dot(n3, x, x));
    orc_apply3(N, nbr, x, Kn, Tv, y, S);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < N; ++p)
      if (!free_[p]) y[3 * p] = y[3 * p + 1] = y[3 * p + 2] = 0.0;
    bj3(N, M0, y, 1.0 / nx, 0, x);
    lam = sqrt(dot(n3, x, x));
  }
  free(x);
  free(y);
  free(S);
  return lam;
}

/* rho(Minv A) on a stencil level from x = start */
double orc_power_rho_st(const int32_t g[3], const double* A, const double* Minv, const double* start, int its) {
  const int64_t n6 = 6 * nodes(g);
  double* x = malloc(sizeof(double) * n6);
  double* y = malloc(sizeof(double) * n6);
  bj6(n6 / 6, Minv, start, 1.0, 0, x);
  double lam = 1.0;
  for (int it = 0; it < its; ++it) {
    const double nx = sqrt(dot(n6, x, x));
    orc_st_mv(g, A, x, NULL, y);
    bj6(n6 / 6, Minv, y, 1.0 / nx, 0, x);
    lam = sqrt(dot(n6, x, x));
  }
  free(x);
  free(y);
  return lam;
}
