// C ABI of the aggregation AMG preconditioner (declarations and reference citations in include/lpm.h): level
// transfers, probing assembly of the Galerkin stencils, smoothers, the coarsest dense inverse, and the V/W-cycle and
// AMG-PCG drivers (timing/t3d/amg3d.py:508-672).  Fine-level residuals inside the cycle are a plain K u followed by
// one elementwise pass that also applies the smoother; the FP64 residual of the refinement loop is lpm_residual.
#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstddef>
#include <vector>

#include "amg.hpp"
#include "capi_common.hpp"

using namespace lpm;
using lpmc::fail;

namespace {

constexpr int VEC_BLOCKS = 512;  // partial sums of lpm_amg_dot (LPM_AMG_SCRATCH_BYTES = 8 VEC_BLOCKS)
static_assert(LPM_AMG_SCRATCH_BYTES >= 8 * VEC_BLOCKS, "scratch holds the partial sums");
constexpr int PCG_BLOCKS = 2048;
// stencil products y = A x of one system: one thread per (row, node) below kStRowsBelow nodes, the two-pass form
// (every stored block read once, 84 values of workspace per node) from kStTwoPassFrom nodes where the caller has
// that workspace, one thread per node otherwise; lpm_amg_set_stencil_kernel fixes one of them
constexpr int64_t kStRowsBelow = 1 << 17;
constexpr int64_t kStTwoPassFrom = 1 << 17;
std::atomic<int> g_st_mode{LPM_AMG_ST_AUTO};  // blocks of the PCG vector kernels (8 per CU), one partial sum each

inline unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs); }
inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }
inline int64_t nodes(const int g[3]) { return (int64_t)g[0] * g[1] * g[2]; }

// level-0 geometry and grids (enough for the transfer primitives used while the hierarchy is being built)
int check_geo(const lpm_amg* a) {
  if (!a) return fail(LPM_E_BADARG, "amg: null hierarchy");
  if (a->nlev < 1 || a->nlev > LPM_AMG_MAXLEV) return fail(LPM_E_BADARG, "amg: %d stencil levels", a->nlev);
  if (a->agg < 1) return fail(LPM_E_BADARG, "amg: chunk size %d", a->agg);
  for (int k = 0; k < 3; ++k)
    if (a->L[k] < 1 || a->g0[k] < 1 || a->g0[k] != std::max(1, a->L[k] / a->agg))
      return fail(LPM_E_BADARG, "amg: free box and aggregate grid disagree");
  for (int k = 0; k < a->nlev; ++k)
    for (int d = 0; d < 3; ++d) {
      const int nodes_below = k == 0 ? a->L[d] : a->g[k - 1][d];
      const int want = k == 0 ? a->g0[d] : std::max(1, nodes_below / a->agg);
      if (a->g[k][d] != want) return fail(LPM_E_BADARG, "amg: grid of stencil level %d inconsistent", k);
    }
  return LPM_OK;
}

// the complete hierarchy (cycle, PCG, workspace)
int check_amg(const lpm_amg* a) {
  if (int rc = check_geo(a)) return rc;
  for (int k = 0; k < a->nlev; ++k) {
    if (nodes(a->g[k]) < 1 || !a->A[k]) return fail(LPM_E_BADARG, "amg: stencil level %d incomplete", k);
    if (k + 1 < a->nlev && (!a->Q[k] || !a->Minv[k])) return fail(LPM_E_BADARG, "amg: level %d incomplete", k);
  }
  if (!a->dinv) return fail(LPM_E_BADARG, "amg: coarsest inverse missing");
  return LPM_OK;
}

Amg0Geo geo0(const lpm_grid* g, const lpm_amg* a) {
  Amg0Geo r{};
  r.nx = g->nx;
  r.ny = g->ny;
  r.nz = g->nz;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = a->lo[k];
    r.L[k] = a->L[k];
    r.g[k] = a->g0[k];
  }
  r.agg = a->agg;
  for (int s = 0; s < 8; ++s) r.toff[s] = a->toff[s];
  return r;
}

AmgLGeo geoq(const lpm_amg* a, int k) {
  AmgLGeo r{};
  for (int i = 0; i < 3; ++i) {
    r.g[i] = a->g[k][i];
    r.gc[i] = a->g[k + 1][i];
  }
  r.agg = a->agg;
  return r;
}

template <typename T>
int level0(const lpm_grid* g, const lpm_amg* a, int op, int col, const void* in, void* out, hipStream_t st) {
  const Amg0Geo geo = geo0(g, a);
  const T* tm = (const T*)a->tmpl;
  const int64_t N = (int64_t)g->nx * g->ny * g->nz;
  const int64_t nagg = nodes(a->g0);
  switch (op) {
    case LPM_AMG_PROLONG_ADD:
      hipLaunchKernelGGL((amg0_prolong_kernel<T>), dim3(blocks_for(N)), dim3(256), 0, st, geo, tm, (const T*)in, (T*)out);
      break;
    case LPM_AMG_RESTRICT:
      hipLaunchKernelGGL((amg0_restrict_kernel<T>), dim3(blocks_for(nagg)), dim3(256), 0, st, geo, tm, (const T*)in,
                         (T*)out);
      break;
    case LPM_AMG_PROBE:
      hipLaunchKernelGGL((amg0_probe_kernel<T>), dim3(blocks_for(N)), dim3(256), 0, st, geo, tm, col, (T*)out);
      break;
    case LPM_AMG_GATHER:
      hipLaunchKernelGGL((amg0_gather_kernel<T>), dim3(blocks_for(nagg)), dim3(256), 0, st, geo, tm, col,
                         (const T*)in, (T*)out);
      break;
    default:
      return fail(LPM_E_BADARG, "lpm_amg_level0: bad op %d", op);
  }
  LAUNCHCHK();
  return LPM_OK;
}

template <typename T>
int levelq(const lpm_amg* a, int k, int op, int col, const void* in, void* out, hipStream_t st, T* tmp = nullptr) {
  const AmgLGeo geo = geoq(a, k);
  const T* Q = (const T*)a->Q[k];
  const int64_t n = nodes(a->g[k]), nc = nodes(a->g[k + 1]);
  switch (op) {
    case LPM_AMG_PROLONG_ADD:
      hipLaunchKernelGGL((amgq_prolong_kernel<T>), dim3(blocks_for(n)), dim3(256), 0, st, geo, Q, (const T*)in, (T*)out);
      break;
    case LPM_AMG_RESTRICT:
      if (tmp) {
        hipLaunchKernelGGL((amgq_qtr_kernel<T>), dim3(blocks_for(n)), dim3(256), 0, st, (int)n, Q, (const T*)in, tmp);
        hipLaunchKernelGGL((amgq_sum_kernel<T>), dim3(blocks_for(nc)), dim3(256), 0, st, geo, (const T*)tmp, (T*)out);
      } else {
        hipLaunchKernelGGL((amgq_restrict_kernel<T>), dim3(blocks_for(nc)), dim3(256), 0, st, geo, Q, (const T*)in,
                           (T*)out);
      }
      break;
    case LPM_AMG_PROBE:
      hipLaunchKernelGGL((amgq_probe_kernel<T>), dim3(blocks_for(n)), dim3(256), 0, st, geo, Q, col, (T*)out);
      break;
    case LPM_AMG_GATHER:
      hipLaunchKernelGGL((amgq_gather_kernel<T>), dim3(blocks_for(nc)), dim3(256), 0, st, geo, Q, col, (const T*)in,
                         (T*)out);
      break;
    default:
      return fail(LPM_E_BADARG, "lpm_amg_levelq: bad op %d", op);
  }
  LAUNCHCHK();
  return LPM_OK;
}

template <typename T>
int st_mv(const int g[3], int nsys, const void* A, const void* x, const void* b, void* y, hipStream_t st,
          void* two_pass_work = nullptr) {
  const int64_t n = nodes(g);
  if (nsys == 6 && !b) {
    hipLaunchKernelGGL((amg_st_mv_kernel<T, 6>), dim3(blocks_for(n)), dim3(256), 0, st, g[0], g[1], g[2], (const T*)A,
                       (const T*)x, (const T*)nullptr, (T*)y);
  } else {
    const int mode = g_st_mode.load(std::memory_order_relaxed);
    const bool two = two_pass_work && (mode == LPM_AMG_ST_TWO_PASS || (mode == LPM_AMG_ST_AUTO && n >= kStTwoPassFrom));
    const bool rows = !two && (mode == LPM_AMG_ST_ROWS || ((mode == LPM_AMG_ST_AUTO || mode == LPM_AMG_ST_TWO_PASS) &&
                                                           n < kStRowsBelow));
    T* yp = (T*)two_pass_work;
    T* tt = yp + 6 * n;
    for (int s = 0; s < nsys; ++s) {
      const T* xs = (const T*)x + s * n * 6;
      T* ys = (T*)y + s * n * 6;
      if (two) {
        hipLaunchKernelGGL((amg_st_fwd_kernel<T>), dim3(blocks_for(n)), dim3(256), 0, st, g[0], g[1], g[2],
                           (const T*)A, xs, yp, tt);
        hipLaunchKernelGGL((amg_st_bwd_kernel<T>), dim3(blocks_for(n)), dim3(256), 0, st, g[0], g[1], g[2],
                           (const T*)yp, (const T*)tt, (const T*)b, ys);
      } else if (rows) {
        hipLaunchKernelGGL((amg_st_mv_rows_kernel<T>), dim3(blocks_for(6 * n)), dim3(256), 0, st, g[0], g[1], g[2],
                           (const T*)A, xs, (const T*)b, ys);
      } else {
        hipLaunchKernelGGL((amg_st_mv_kernel<T, 1>), dim3(blocks_for(n)), dim3(256), 0, st, g[0], g[1], g[2],
                           (const T*)A, xs, (const T*)b, ys);
      }
    }
  }
  LAUNCHCHK();
  return LPM_OK;
}

template <typename T>
int bj(int64_t n, int k, const void* M, const void* b, double w, int add, void* out, hipStream_t st) {
  if (k == 3)
    hipLaunchKernelGGL((amg_bj3_kernel<T>), dim3(blocks_for(n)), dim3(256), 0, st, n, (const T*)M, (const T*)b, (T)w,
                       add, (T*)out);
  else
    hipLaunchKernelGGL((amg_bj6_kernel<T>), dim3(blocks_for(n)), dim3(256), 0, st, (int)n, (const T*)M, (const T*)b,
                       (T)w, add, (T*)out);
  LAUNCHCHK();
  return LPM_OK;
}

// ---------------------------------------------------------------------------------------------- cycle
template <typename T>
struct Cycle {
  const lpm_grid* g;
  const lpm_amg* a;
  int layout;
  const void* kn;
  const void* tv;
  const uint8_t* mask;
  hipStream_t st;
  // workspace
  T* r0;      // (N, 3) fine residual
  void* rwork;  // lpm_residual partial sums
  T* b[LPM_AMG_MAXLEV];
  T* e[LPM_AMG_MAXLEV];
  T* t[LPM_AMG_MAXLEV];
  T* x[LPM_AMG_MAXLEV];
  T* r[LPM_AMG_MAXLEV];
  T* q[LPM_AMG_MAXLEV];  // restriction stage buffer (n_k, 6)
  void* stw;             // two-pass stencil product: yp (6, n_k) and the transposed products (13, 6, n_k)

  int64_t N() const { return (int64_t)g->nx * g->ny * g->nz; }

  // plain K u into out
  int ku0(const T* u, T* out) {
    const int64_t n3 = N() * 3;
    return lpm_apply(g, 1, 1, sizeof(T) == 4 ? LPM_F32 : LPM_F64, layout, u, lpm_bcast{n3, n3}, kn, tv,
                     lpm_bcast{0, 0}, out, st);
  }

  // out = mask (rhs - K u)
  int resid0(const T* rhs, const T* u, T* out) {
    if (int rc = ku0(u, out)) return rc;
    const int64_t n3 = N() * 3;
    hipLaunchKernelGGL((amg_resid_kernel<T>), dim3(std::min<unsigned>(blocks_for(n3), 8192)), dim3(256), 0, st, n3,
                       rhs, mask, out);
    LAUNCHCHK();
    return LPM_OK;
  }

  // z += w M mask (rhs - K z), with r0 as the K z buffer
  int smooth0(const T* rhs, T* z, double w) {
    if (int rc = ku0(z, r0)) return rc;
    const int64_t N_ = N();
    hipLaunchKernelGGL((amg_resid_bj3_kernel<T>), dim3(blocks_for(N_)), dim3(256), 0, st, N_, (const T*)a->minv0, rhs,
                       mask, (const T*)r0, (T)w, z);
    LAUNCHCHK();
    return LPM_OK;
  }

  // out = approximate solution of A_k out = rhs (one cycle from zero); amg3d.py:590-618
  int vc(int k, const T* rhs, T* out) {
    const int64_t n = nodes(a->g[k]);
    if (k == a->nlev - 1) {
      const int m = (int)(6 * n);
      hipLaunchKernelGGL((amg_dense_mv_kernel<T>), dim3((m + 3) / 4), dim3(256), 0, st, m, (const T*)a->dinv, rhs, out);
      LAUNCHCHK();
      return LPM_OK;
    }
    const double w = a->w[k];
    int rc;
    if ((rc = bj<T>(n, 6, a->Minv[k], rhs, w, 0, out, st))) return rc;
    for (int it = 0; it < a->nu_c - 1; ++it) {
      if ((rc = st_mv<T>(a->g[k], 1, a->A[k], out, rhs, r[k], st, stw))) return rc;
      if ((rc = bj<T>(n, 6, a->Minv[k], r[k], w, 1, out, st))) return rc;
    }
    if ((rc = st_mv<T>(a->g[k], 1, a->A[k], out, rhs, r[k], st, stw))) return rc;
    if ((rc = levelq<T>(a, k, LPM_AMG_RESTRICT, 0, r[k], b[k + 1], st, q[k]))) return rc;
    if ((rc = coarse(k + 1, b[k + 1], e[k + 1]))) return rc;
    if ((rc = levelq<T>(a, k, LPM_AMG_PROLONG_ADD, 0, e[k + 1], out, st))) return rc;
    for (int it = 0; it < a->nu_c; ++it) {
      if ((rc = st_mv<T>(a->g[k], 1, a->A[k], out, rhs, r[k], st, stw))) return rc;
      if ((rc = bj<T>(n, 6, a->Minv[k], r[k], w, 1, out, st))) return rc;
    }
    return LPM_OK;
  }

  // e = gamma cycles on level k from zero (one on the coarsest level or below w_from); amg3d.py:578-588
  int coarse(int k, const T* rhs, T* ek) {
    int rc;
    if ((rc = vc(k, rhs, ek))) return rc;
    if (k == a->nlev - 1) return LPM_OK;
    const int lvl = k + 1;  // the reference numbers the first stencil level 1
    const int extra = lvl >= a->w_from ? a->gamma - 1 : 0;
    const int64_t n6 = nodes(a->g[k]) * 6;
    for (int c = 0; c < extra; ++c) {
      if ((rc = st_mv<T>(a->g[k], 1, a->A[k], ek, rhs, t[k], st, stw))) return rc;
      if ((rc = vc(k, t[k], x[k]))) return rc;
      hipLaunchKernelGGL((amg_add_kernel<T>), dim3(std::min<unsigned>(blocks_for(n6), 4096)), dim3(256), 0, st, n6,
                         (const T*)x[k], ek);
      LAUNCHCHK();
    }
    return LPM_OK;
  }

  // z = B b on the particles; amg3d.py:620-646
  int fine(const T* bb, T* z) {
    const int64_t N_ = N();
    const double w0 = a->w0;
    int rc;
    if ((rc = bj<T>(N_, 3, a->minv0, bb, w0, 0, z, st))) return rc;
    for (int it = 0; it < a->nu - 1; ++it)
      if ((rc = smooth0(bb, z, w0))) return rc;
    if ((rc = resid0(bb, z, r0))) return rc;
    if ((rc = level0<T>(g, a, LPM_AMG_RESTRICT, 0, r0, b[0], st))) return rc;
    if ((rc = coarse(0, b[0], e[0]))) return rc;
    if ((rc = level0<T>(g, a, LPM_AMG_PROLONG_ADD, 0, e[0], z, st))) return rc;
    for (int it = 0; it < a->nu; ++it)
      if ((rc = smooth0(bb, z, w0))) return rc;
    return LPM_OK;
  }
};

struct WorkLayout {
  size_t r0, rwork, lev[LPM_AMG_MAXLEV][6], stw, pcg[4], part, scal, total;
};

template <typename T>
WorkLayout work_layout(const lpm_grid* g, const lpm_amg* a) {
  WorkLayout L{};
  const int64_t N = (int64_t)g->nx * g->ny * g->nz;
  size_t o = 0;
  const size_t vec = align256(sizeof(T) * N * 3);
  L.r0 = o;
  o += vec;
  L.rwork = o;
  o += lpm_residual_workspace(g, 1, sizeof(T) == 4 ? LPM_F32 : LPM_F64);
  for (int k = 0; k < a->nlev; ++k) {
    const size_t v = align256(sizeof(T) * nodes(a->g[k]) * 6);
    for (int j = 0; j < 6; ++j) {
      L.lev[k][j] = o;
      o += v;
    }
  }
  L.stw = o;  // the first stencil level is the largest
  o += align256(sizeof(T) * nodes(a->g[0]) * 6 * 14);
  for (int j = 0; j < 4; ++j) {  // PCG: r, z, p, Ap
    L.pcg[j] = o;
    o += vec;
  }
  L.part = o;
  o += align256(sizeof(double) * PCG_BLOCKS);
  L.scal = o;
  o += align256(sizeof(AmgScal));
  L.total = o;
  return L;
}

template <typename T>
Cycle<T> make_cycle(const lpm_grid* g, const lpm_amg* a, int layout, const void* kn, const void* tv,
                    const uint8_t* mask, void* work, hipStream_t st) {
  const WorkLayout L = work_layout<T>(g, a);
  char* w = (char*)work;
  Cycle<T> c{};
  c.g = g;
  c.a = a;
  c.layout = layout;
  c.kn = kn;
  c.tv = tv;
  c.mask = mask;
  c.st = st;
  c.r0 = (T*)(w + L.r0);
  c.rwork = w + L.rwork;
  for (int k = 0; k < a->nlev; ++k) {
    c.b[k] = (T*)(w + L.lev[k][0]);
    c.e[k] = (T*)(w + L.lev[k][1]);
    c.t[k] = (T*)(w + L.lev[k][2]);
    c.x[k] = (T*)(w + L.lev[k][3]);
    c.r[k] = (T*)(w + L.lev[k][4]);
    c.q[k] = (T*)(w + L.lev[k][5]);
  }
  c.stw = w + L.stw;
  return c;
}

template <typename T>
int dot(int64_t n, const T* x, const T* y, double* part, double* out, hipStream_t st) {
  hipLaunchKernelGGL((amg_dot_kernel<T>), dim3(VEC_BLOCKS), dim3(256), 0, st, n, x, y, part);
  LAUNCHCHK();
  std::vector<double> h(VEC_BLOCKS);
  HIPCHK(hipMemcpyAsync(h.data(), part, sizeof(double) * VEC_BLOCKS, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double s = 0.0;
  for (int i = 0; i < VEC_BLOCKS; ++i) s += h[i];
  *out = s;
  return LPM_OK;
}

// the PCG vector kernels: W values per load where every vector and the mask allow 16-byte loads, one otherwise
template <typename T>
struct PcgVec {
  int64_t n;
  unsigned nblk;
  bool wide;
  double* part;
  AmgScal* sc;
  hipStream_t st;

  template <int W>
  unsigned blocks() const {
    return (unsigned)std::min<int64_t>(PCG_BLOCKS, std::max<int64_t>(1, (n / W + 255) / 256));
  }
  int dot(const T* x, const T* y, int op, double tol = 0.0, int maxit = 0) {
    if (wide)
      hipLaunchKernelGGL((amg_dotv_kernel<T, 16 / sizeof(T)>), dim3(nblk), dim3(256), 0, st, n, x, y, part);
    else
      hipLaunchKernelGGL((amg_dotv_kernel<T, 1>), dim3(nblk), dim3(256), 0, st, n, x, y, part);
    LAUNCHCHK();
    return fin(op, tol, maxit);
  }
  int fin(int op, double tol, int maxit) {
    hipLaunchKernelGGL(amg_fin_kernel, dim3(1), dim3(256), 0, st, (const double*)part, (int)nblk, sc, op, tol, maxit);
    LAUNCHCHK();
    return LPM_OK;
  }
  int step(const T* p, const T* ap, const uint8_t* mask, T* x, T* r, double tol, int maxit) {
    if (wide)
      hipLaunchKernelGGL((amg_cg_stepv_kernel<T, 16 / sizeof(T)>), dim3(nblk), dim3(256), 0, st, n,
                         (const AmgScal*)sc, p, ap, mask, x, r, part);
    else
      hipLaunchKernelGGL((amg_cg_stepv_kernel<T, 1>), dim3(nblk), dim3(256), 0, st, n, (const AmgScal*)sc, p, ap,
                         mask, x, r, part);
    LAUNCHCHK();
    return fin(AF_RN, tol, maxit);
  }
  int xpby(const T* z, T* p) {
    if (wide)
      hipLaunchKernelGGL((amg_xpbyv_kernel<T, 16 / sizeof(T)>), dim3(nblk), dim3(256), 0, st, n, z,
                         (const AmgScal*)sc, p);
    else
      hipLaunchKernelGGL((amg_xpbyv_kernel<T, 1>), dim3(nblk), dim3(256), 0, st, n, z, (const AmgScal*)sc, p);
    LAUNCHCHK();
    return LPM_OK;
  }
};

// the solve's scalars, read back into page-locked memory allocated once per thread
AmgScal* host_scal() {
  thread_local AmgScal* h = nullptr;
  if (!h && hipHostMalloc((void**)&h, sizeof(AmgScal), hipHostMallocDefault) != hipSuccess) h = nullptr;
  return h;
}

// AMG-PCG from x = 0 (amg3d.py:649-672).  The scalars stay on the device: p.Ap, alpha, |r|, beta and the stopping
// test are one-block kernels over fixed-order partial sums, and the vector kernels read them there.  The host waits
// once per iteration, for the 4-byte active flag after the residual update, before it queues the cycle.
template <typename T>
int pcg(const lpm_grid* g, const lpm_amg* a, int layout, const void* kn, const void* tv, const uint8_t* mask,
        const T* b, T* x, double tol, int maxit, void* work, int32_t* iters, double* relres, hipStream_t st) {
  const WorkLayout L = work_layout<T>(g, a);
  Cycle<T> cy = make_cycle<T>(g, a, layout, kn, tv, mask, work, st);
  char* w = (char*)work;
  T* r = (T*)(w + L.pcg[0]);
  T* z = (T*)(w + L.pcg[1]);
  T* p = (T*)(w + L.pcg[2]);
  T* ap = (T*)(w + L.pcg[3]);
  const int64_t N = cy.N(), n3 = N * 3;
  AmgScal* hs = host_scal();
  if (!hs) return fail(LPM_E_HIP, "AMG-PCG: page-locked host memory unavailable");
  constexpr int WV = 16 / sizeof(T);
  PcgVec<T> v{n3, 0, false, (double*)(w + L.part), (AmgScal*)(w + L.scal), st};
  v.wide = ((uintptr_t)x % 16 == 0) && ((uintptr_t)b % 16 == 0) && ((uintptr_t)mask % WV == 0);
  v.nblk = v.wide ? v.template blocks<WV>() : v.template blocks<1>();
  const unsigned eg = std::min<unsigned>(blocks_for(n3), 4096);
  int rc;
  auto poll = [&](size_t bytes) -> int {
    HIPCHK(hipMemcpyAsync(hs, v.sc, bytes, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return LPM_OK;
  };
  hipLaunchKernelGGL((amg_fill_kernel<T>), dim3(eg), dim3(256), 0, st, n3, T(0), x);
  HIPCHK(hipMemcpyAsync(r, b, sizeof(T) * n3, hipMemcpyDeviceToDevice, st));
  if ((rc = v.dot(b, b, AF_BN))) return rc;
  if ((rc = poll(sizeof(AmgScal)))) return rc;
  *iters = 0;
  *relres = 0.0;
  if (hs->nonfinite) return fail(LPM_E_NONFINITE, "AMG-PCG: non-finite right-hand side");
  if (hs->bn == 0.0) return LPM_OK;
  if ((rc = cy.fine(r, z))) return rc;
  HIPCHK(hipMemcpyAsync(p, z, sizeof(T) * n3, hipMemcpyDeviceToDevice, st));
  if ((rc = v.dot(r, z, AF_RZ0))) return rc;
  const lpm_bcast full{n3, n3}, none{0, 0};
  static_assert(offsetof(AmgScal, active) + sizeof(int) <= sizeof(AmgScal), "flag inside the scalars");
  for (int it = 1; it <= maxit; ++it) {
    if ((rc = lpm_apply(g, 1, 1, sizeof(T) == 4 ? LPM_F32 : LPM_F64, layout, p, full, kn, tv, none, ap, st)))
      return rc;
    if ((rc = v.dot(p, ap, AF_PAP))) return rc;
    if ((rc = v.step(p, ap, mask, x, r, tol, maxit))) return rc;
    if ((rc = poll(offsetof(AmgScal, active) + sizeof(int)))) return rc;
    if (!hs->active) break;
    if ((rc = cy.fine(r, z))) return rc;
    if ((rc = v.dot(r, z, AF_RZ))) return rc;
    if ((rc = v.xpby(z, p))) return rc;
  }
  if ((rc = poll(sizeof(AmgScal)))) return rc;
  if (hs->notspd) return fail(LPM_E_NOT_SPD, "AMG-PCG: p.Ap = %g at iteration %d", hs->pap, hs->it + 1);
  if (hs->nonfinite) return fail(LPM_E_NONFINITE, "AMG-PCG: non-finite residual at iteration %d", hs->it);
  *iters = hs->it;
  *relres = hs->rn / hs->bn;
  return LPM_OK;
}

// workspace of the refined solve: the FP32 AMG-PCG workspace, then r64, r32, d32 and the FP64 residual partials
struct RefLayout {
  size_t inner, r64, r32, d32, rwork, part, total;
};

RefLayout ref_layout(const lpm_grid* g, const lpm_amg* a) {
  RefLayout L{};
  const int64_t n3 = (int64_t)g->nx * g->ny * g->nz * 3;
  size_t o = 0;
  L.inner = o;
  o += work_layout<float>(g, a).total;
  L.r64 = o;
  o += align256(8 * n3);
  L.r32 = o;
  o += align256(4 * n3);
  L.d32 = o;
  o += align256(4 * n3);
  L.rwork = o;
  o += lpm_residual_workspace(g, 1, LPM_F64);
  L.part = o;
  o += align256(8 * VEC_BLOCKS);
  L.total = o;
  return L;
}

bool fine_ok(const lpm_grid* g, const lpm_amg* a) {
  if (!g || g->dim != 3) return false;
  const int n[3] = {g->nx, g->ny, g->nz};
  for (int k = 0; k < 3; ++k)
    if (a->lo[k] < 0 || a->lo[k] + a->L[k] > n[k]) return false;
  return true;
}

}  // namespace

extern "C" {

int lpm_amg_level0(const lpm_grid* g, const lpm_amg* a, int32_t dtype, int32_t op, int32_t col, const void* in,
                   void* out, void* stream) {
  if (int rc = check_geo(a)) return rc;
  if (!fine_ok(g, a)) return fail(LPM_E_BADARG, "lpm_amg_level0: free box outside the 3D lattice");
  if (!a->tmpl || !out || (op != LPM_AMG_PROBE && !in) || col < 0 || col > 26)
    return fail(LPM_E_BADARG, "lpm_amg_level0: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == LPM_F64) return level0<double>(g, a, op, col, in, out, st);
  if (dtype == LPM_F32) return level0<float>(g, a, op, col, in, out, st);
  return fail(LPM_E_BADARG, "lpm_amg_level0: bad dtype %d", dtype);
}

int lpm_amg_levelq(const lpm_amg* a, int32_t k, int32_t dtype, int32_t op, int32_t col, const void* in, void* out,
                   void* stream) {
  if (int rc = check_geo(a)) return rc;
  if (k < 0 || k + 1 >= a->nlev || !a->Q[k] || !out || (op != LPM_AMG_PROBE && !in) || col < 0 || col > 26)
    return fail(LPM_E_BADARG, "lpm_amg_levelq: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == LPM_F64) return levelq<double>(a, k, op, col, in, out, st);
  if (dtype == LPM_F32) return levelq<float>(a, k, op, col, in, out, st);
  return fail(LPM_E_BADARG, "lpm_amg_levelq: bad dtype %d", dtype);
}

int lpm_amg_set_stencil_kernel(int32_t mode) {
  if (mode < LPM_AMG_ST_AUTO || mode > LPM_AMG_ST_TWO_PASS)
    return fail(LPM_E_BADARG, "lpm_amg_set_stencil_kernel: mode %d", mode);
  g_st_mode.store(mode);
  return LPM_OK;
}

int lpm_amg_st_mv(int32_t gx, int32_t gy, int32_t gz, int32_t nsys, int32_t dtype, const void* A, const void* x,
                  const void* b, void* y, void* stream) {
  if (gx < 1 || gy < 1 || gz < 1 || nsys < 1 || !A || !x || !y) return fail(LPM_E_BADARG, "lpm_amg_st_mv: bad arguments");
  const int g[3] = {gx, gy, gz};
  hipStream_t st = (hipStream_t)stream;
  if (dtype == LPM_F64) return st_mv<double>(g, nsys, A, x, b, y, st);
  if (dtype == LPM_F32) return st_mv<float>(g, nsys, A, x, b, y, st);
  return fail(LPM_E_BADARG, "lpm_amg_st_mv: bad dtype %d", dtype);
}

int lpm_amg_inv_diag(int32_t n, int32_t dtype, const void* A, void* Minv, void* scratch, void* stream) {
  if (n < 1 || !A || !Minv || !scratch) return fail(LPM_E_BADARG, "lpm_amg_inv_diag: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  int* bad = (int*)scratch;
  hipLaunchKernelGGL((amg_fill_kernel<int>), dim3(1), dim3(64), 0, st, (int64_t)1, 0, bad);
  if (dtype == LPM_F64)
    hipLaunchKernelGGL((amg_inv_diag_kernel<double>), dim3(blocks_for(n)), dim3(256), 0, st, n, (const double*)A,
                       (double*)Minv, bad);
  else if (dtype == LPM_F32)
    hipLaunchKernelGGL((amg_inv_diag_kernel<float>), dim3(blocks_for(n)), dim3(256), 0, st, n, (const float*)A,
                       (float*)Minv, bad);
  else
    return fail(LPM_E_BADARG, "lpm_amg_inv_diag: bad dtype %d", dtype);
  LAUNCHCHK();
  int hb = 0;
  HIPCHK(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (hb) return fail(LPM_E_NOT_SPD, "lpm_amg_inv_diag: singular diagonal block");
  return LPM_OK;
}

int lpm_amg_bj(int64_t n, int32_t k, int32_t dtype, const void* M, const void* b, double w, int32_t add, void* out,
               void* stream) {
  if (n < 1 || (k != 3 && k != 6) || !M || !b || !out || (k == 6 && n >= ((int64_t)1 << 31)))
    return fail(LPM_E_BADARG, "lpm_amg_bj: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == LPM_F64) return bj<double>(n, k, M, b, w, add, out, st);
  if (dtype == LPM_F32) return bj<float>(n, k, M, b, w, add, out, st);
  return fail(LPM_E_BADARG, "lpm_amg_bj: bad dtype %d", dtype);
}

int lpm_amg_coarsest(int32_t gx, int32_t gy, int32_t gz, int32_t dtype, const void* A, void* dinv, void* scratch,
                     void* stream) {
  const int64_t n = (int64_t)gx * gy * gz, m = 6 * n;
  if (n < 1 || !A || !dinv || !scratch || m > 8192)
    return fail(LPM_E_BADARG, "lpm_amg_coarsest: bad arguments (6n = %lld)", (long long)m);
  hipStream_t st = (hipStream_t)stream;
  int* bad = (int*)scratch;
  hipLaunchKernelGGL((amg_fill_kernel<int>), dim3(1), dim3(64), 0, st, (int64_t)1, 0, bad);
  if (dtype == LPM_F64) {
    hipLaunchKernelGGL((amg_dense_kernel<double>), dim3(blocks_for(m * m)), dim3(256), 0, st, gx, gy, gz,
                       (const double*)A, (double*)dinv);
    hipLaunchKernelGGL((amg_spd_inverse_kernel<double>), dim3(1), dim3(1024), 0, st, (int)m, (double*)dinv, bad);
  } else if (dtype == LPM_F32) {
    hipLaunchKernelGGL((amg_dense_kernel<float>), dim3(blocks_for(m * m)), dim3(256), 0, st, gx, gy, gz,
                       (const float*)A, (float*)dinv);
    hipLaunchKernelGGL((amg_spd_inverse_kernel<float>), dim3(1), dim3(1024), 0, st, (int)m, (float*)dinv, bad);
  } else {
    return fail(LPM_E_BADARG, "lpm_amg_coarsest: bad dtype %d", dtype);
  }
  LAUNCHCHK();
  int hb = 0;
  HIPCHK(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (hb) return fail(LPM_E_NOT_SPD, "lpm_amg_coarsest: coarsest operator not positive definite");
  return LPM_OK;
}

int lpm_amg_dot(int64_t n, int32_t dtype, const void* x, const void* y, double* out, void* scratch, void* stream) {
  if (n < 1 || !x || !y || !out || !scratch) return fail(LPM_E_BADARG, "lpm_amg_dot: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  double* part = (double*)scratch;
  if (dtype == LPM_F64) return dot<double>(n, (const double*)x, (const double*)y, part, out, st);
  if (dtype == LPM_F32) return dot<float>(n, (const float*)x, (const float*)y, part, out, st);
  return fail(LPM_E_BADARG, "lpm_amg_dot: bad dtype %d", dtype);
}

size_t lpm_amg_workspace(const lpm_grid* g, const lpm_amg* a, int32_t dtype) {
  if (!g || check_amg(a) || !fine_ok(g, a)) return 0;
  return dtype == LPM_F32 ? work_layout<float>(g, a).total : work_layout<double>(g, a).total;
}

int lpm_amg_vcycle(const lpm_grid* g, const lpm_amg* a, int32_t dtype, int32_t layout, const void* kn, const void* tv,
                   const uint8_t* mask, const void* b, void* z, void* work, void* stream) {
  if (int rc = check_amg(a)) return rc;
  if (!fine_ok(g, a) || !kn || !tv || !b || !z || !work || !a->minv0 || !a->tmpl)
    return fail(LPM_E_BADARG, "lpm_amg_vcycle: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == LPM_F64) return make_cycle<double>(g, a, layout, kn, tv, mask, work, st).fine((const double*)b, (double*)z);
  if (dtype == LPM_F32) return make_cycle<float>(g, a, layout, kn, tv, mask, work, st).fine((const float*)b, (float*)z);
  return fail(LPM_E_BADARG, "lpm_amg_vcycle: bad dtype %d", dtype);
}

int lpm_amg_pcg(const lpm_grid* g, const lpm_amg* a, int32_t dtype, int32_t layout, const void* kn, const void* tv,
                const uint8_t* mask, const void* b, void* x, double tol, int32_t maxit, void* work, int32_t* iters,
                double* relres, void* stream) {
  if (int rc = check_amg(a)) return rc;
  if (!fine_ok(g, a) || !kn || !tv || !mask || !b || !x || !work || !iters || !relres || maxit < 1 || !a->minv0 ||
      !a->tmpl)
    return fail(LPM_E_BADARG, "lpm_amg_pcg: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == LPM_F64)
    return pcg<double>(g, a, layout, kn, tv, mask, (const double*)b, (double*)x, tol, maxit, work, iters, relres, st);
  if (dtype == LPM_F32)
    return pcg<float>(g, a, layout, kn, tv, mask, (const float*)b, (float*)x, tol, maxit, work, iters, relres, st);
  return fail(LPM_E_BADARG, "lpm_amg_pcg: bad dtype %d", dtype);
}

size_t lpm_amg_refined_workspace(const lpm_grid* g, const lpm_amg* a32) {
  if (!g || check_amg(a32) || !fine_ok(g, a32)) return 0;
  return ref_layout(g, a32).total;
}

int lpm_amg_pcg_refined(const lpm_grid* g, const lpm_amg* a32, int32_t layout64, const void* kn64, const void* tv64,
                        int32_t layout32, const void* kn32, const void* tv32, const uint8_t* mask, const void* b64,
                        void* x64, double tol, double inner_tol, int32_t max_outer, int32_t maxit, void* work,
                        int32_t* iters, int32_t* outer, double* relres, void* stream) {
  if (int rc = check_amg(a32)) return rc;
  if (!fine_ok(g, a32) || !kn64 || !tv64 || !kn32 || !tv32 || !mask || !b64 || !x64 || !work || !iters || !outer ||
      !relres || maxit < 1 || max_outer < 1 || !a32->minv0 || !a32->tmpl)
    return fail(LPM_E_BADARG, "lpm_amg_pcg_refined: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const RefLayout L = ref_layout(g, a32);
  char* w = (char*)work;
  double* r64 = (double*)(w + L.r64);
  float* r32 = (float*)(w + L.r32);
  float* d32 = (float*)(w + L.d32);
  double* part = (double*)(w + L.part);
  const int64_t n3 = (int64_t)g->nx * g->ny * g->nz * 3;
  const unsigned eg = std::min<unsigned>(blocks_for(n3), 4096);
  const lpm_bcast full{n3, n3}, none{0, 0};
  hipLaunchKernelGGL((amg_fill_kernel<double>), dim3(eg), dim3(256), 0, st, n3, 0.0, (double*)x64);
  LAUNCHCHK();
  double bn2;
  int rc;
  if ((rc = dot<double>(n3, (const double*)b64, (const double*)b64, part, &bn2, st))) return rc;
  const double bn = std::sqrt(bn2);
  *iters = 0;
  *outer = 0;
  *relres = 0.0;
  if (bn == 0.0) return LPM_OK;
  for (int k = 0;; ++k) {
    // true FP64 residual r = mask (b - K x)
    if ((rc = lpm_residual(g, 1, 1, LPM_F64, layout64, x64, full, kn64, tv64, none, mask, full, b64, full, r64,
                           w + L.rwork, st)))
      return rc;
    double rr;
    if ((rc = dot<double>(n3, r64, r64, part, &rr, st))) return rc;
    const double rel = std::sqrt(rr) / bn;
    if (!std::isfinite(rel)) return fail(LPM_E_NONFINITE, "AMG refinement: non-finite residual");
    *relres = rel;
    *outer = k;
    if (rel <= tol || k >= max_outer) break;
    // FP32 AMG-PCG on the residual, to the tolerance that reaches the target in this step (never looser than 0.5)
    const double t = std::min(0.5, std::max(inner_tol, 0.5 * tol / rel));
    hipLaunchKernelGGL(amg_to32_kernel, dim3(eg), dim3(256), 0, st, n3, (const double*)r64, r32);
    LAUNCHCHK();
    int32_t it = 0;
    double rel_in = 0.0;
    if ((rc = pcg<float>(g, a32, layout32, kn32, tv32, mask, r32, d32, t, maxit, w + L.inner, &it, &rel_in, st)))
      return rc;
    *iters += it;
    hipLaunchKernelGGL(amg_add64_kernel, dim3(eg), dim3(256), 0, st, n3, (const float*)d32, (double*)x64);
    LAUNCHCHK();
  }
  return LPM_OK;
}

}  // extern "C"
