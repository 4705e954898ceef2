// Vector kernels of the batched (P)CG loop.  Every system s = b * nl + l has its own scalars, so systems converge
// and freeze independently (the reference freezes converged systems of a batch, coarse3d/torch_lpm3d.py:123-133, and
// tests ||r|| <= tol ||b|| every iteration in coarse3d/lin3d.py:491-507).  Per-block partial sums are reduced by a
// finalize kernel in a fixed order, so a system's result does not depend on which other systems share the launch.
#pragma once
#include "lpm_common.hpp"

namespace lpm {

template <typename T>
__device__ __forceinline__ bool is_finite(T v) {
  return __builtin_isfinite(v);
}
template <typename T>
__device__ __forceinline__ T nan_of() {
  return (T)__builtin_nan("");
}

// per-system scalars (type T) and integer state
// S_RZN: r.z of the newest residual, written where it is summed (cg_fin_init, cg_pupdate_fused); the fused update
// copies it to S_RZ, the r.z of the iteration's own residual, so no kernel reads a scalar another block of the same
// launch writes.
enum CgScal : int { S_RZ = 0, S_PAP, S_ALPHA, S_BETA, S_RR, S_BN, S_TOL, S_REL, S_RZN, NSCAL };
// I_XUPD: the x += alpha p of the current iteration is still to be applied (set by cg_fin_alpha, consumed by
// cg_pupdate, which folds the x update into its pass over p).
// I_NOTSPD: the system stopped because K (or M) is not positive definite on it: p.Ap <= 0 for a nonzero search
// direction, or r.z < 0.  The reference's dense path raises FloatingPointError there (coarse2d/torch_lpm.py:20-25);
// the host reads the flag at its convergence poll and returns LPM_E_NOT_SPD.
// Snapshots (multi-tolerance solves, timing/t2d/solver2d.py:27, 349-421): I_SNAPK = next tolerance of the descending
// list not yet met; [I_SNAP0, I_SNAP1) = levels first met in the current iteration, whose snapshot is the iterate
// after this iteration's update.
enum CgInt : int { I_IT = 0, I_ACTIVE, I_MAXIT, I_CONV, I_XUPD, I_NOTSPD, I_SNAPK, I_SNAP0, I_SNAP1, NINT };

// Snapshot mode of a solve (nt = 0: off).  tols: descending tolerances (device); snap_it (systems, nt): iteration
// at which each level was first met (-1 = never); snaps (systems, nt, N, D): the iterates.
template <typename T>
struct SnapDev {
  int nt;
  const double* tols;
  int* snap_it;
  T* snaps;
};

// levels k >= k0 met by relative residual rn / bn (bn = 0: all met), recorded with iteration `it`; returns the next
// unmet level
__device__ __forceinline__ int snap_levels(const double* tols, int nt, int k0, double rn, double bn, int* snap_it,
                                           int it) {
  int k = k0;
  while (k < nt && (bn == 0.0 || rn <= tols[k] * bn)) snap_it[k++] = it;
  return k;
}

template <typename T>
struct CgVec {
  int64_t N;  // particles
  int nl;
  T* x;
  T* r;
  T* z;  // null for unpreconditioned CG (z = r)
  T* p;
  T* ap;
  const T* minv;  // symmetric blocks (N, 6 | 3) per system, null = identity
  Bcast bminv;
  const uint8_t* mask;  // per DOF, null = all free
  Bcast bm;
  T* partial;  // (systems, gridDim.x, 3)
  T* scal;     // (systems, NSCAL)
  int* ist;    // (systems, NINT)
};

template <int D, typename T>
__device__ __forceinline__ void apply_minv(const T* __restrict__ m, const T* r, T* z) {
  if constexpr (D == 3) {
    z[0] = m[0] * r[0] + m[5] * r[1] + m[4] * r[2];
    z[1] = m[5] * r[0] + m[1] * r[1] + m[3] * r[2];
    z[2] = m[4] * r[0] + m[3] * r[1] + m[2] * r[2];
  } else {
    z[0] = m[0] * r[0] + m[2] * r[1];
    z[1] = m[2] * r[0] + m[1] * r[1];
  }
}

// r = mask * b (or r given by the residual kernel when `r_ready`), z = M r, p = z; partials (r.z, r.r, b.b).
template <int D, typename T>
__global__ __launch_bounds__(256) void cg_init_kernel(CgVec<T> v, const T* __restrict__ bvec, int r_ready) {
  __shared__ T red[4];
  const int l = blockIdx.y, b = blockIdx.z;
  const int64_t s = (int64_t)b * v.nl + l;
  T rz = T(0), rr = T(0), bb = T(0);
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < v.N; p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = (s * v.N + p) * D;
    T r[3], z[3], m[3];
    for (int a = 0; a < D; ++a) {
      m[a] = v.mask ? (v.mask[v.bm.at(b, l) + D * p + a] ? T(1) : T(0)) : T(1);
      const T bv = m[a] * bvec[o + a];
      bb += bv * bv;
      r[a] = r_ready ? v.r[o + a] : bv;
    }
    if (v.minv) {
      constexpr int NS = D == 3 ? 6 : 3;
      apply_minv<D, T>(v.minv + v.bminv.at(b, l) + NS * p, r, z);
      for (int a = 0; a < D; ++a) z[a] *= m[a];
    } else {
      for (int a = 0; a < D; ++a) z[a] = r[a];
    }
    for (int a = 0; a < D; ++a) {
      if (!r_ready) v.r[o + a] = r[a];
      if (v.z) v.z[o + a] = z[a];
      v.p[o + a] = z[a];
      if (!r_ready) v.x[o + a] = T(0);
      rz += r[a] * z[a];
      rr += r[a] * r[a];
    }
  }
  const T s0 = block_sum(rz, red);
  const T s1 = block_sum(rr, red);
  const T s2 = block_sum(bb, red);
  if (threadIdx.x == 0) {
    T* pp = v.partial + (s * gridDim.x + blockIdx.x) * 3;
    pp[0] = s0;
    pp[1] = s1;
    pp[2] = s2;
  }
}

// sum of column c of the (nblk, stride) partials of one system, fixed order
template <typename T>
__device__ __forceinline__ T sum_partials(const T* __restrict__ pp, int nblk, int stride, int c, T* red) {
  T acc = T(0);
  for (int j = threadIdx.x; j < nblk; j += blockDim.x) acc += pp[(int64_t)j * stride + c];
  return block_sum(acc, red);
}
// the same sum in every thread of the block
template <typename T>
__device__ __forceinline__ T sum_partials_all(const T* __restrict__ pp, int nblk, int stride, int c, T* red,
                                              T* slot) {
  const T v = sum_partials(pp, nblk, stride, c, red);
  if (threadIdx.x == 0) *slot = v;
  __syncthreads();
  const T r = *slot;
  __syncthreads();
  return r;
}

// after cg_init: rz, rr, bn; a system with b = 0 is done at iteration 0 (coarse3d/lin3d.py:479-480).  In snapshot
// mode the levels the initial residual meets are recorded with iteration 0 and the system is done once every level
// is met.
template <typename T>
__global__ __launch_bounds__(256) void cg_fin_init_kernel(T* partial, int nblk, T* scal, int* ist, SnapDev<T> sd) {
  __shared__ T red[4];
  const int s = blockIdx.x;
  const T* pp = partial + (int64_t)s * nblk * 3;
  const T rz = sum_partials(pp, nblk, 3, 0, red);
  const T rr = sum_partials(pp, nblk, 3, 1, red);
  const T bb = sum_partials(pp, nblk, 3, 2, red);
  if (threadIdx.x == 0) {
    T* sc = scal + (int64_t)s * NSCAL;
    int* is = ist + (int64_t)s * NINT;
    const T bn = sqrt(bb);
    sc[S_RZ] = rz;
    sc[S_RZN] = rz;
    sc[S_RR] = rr;
    sc[S_BN] = bn;
    sc[S_ALPHA] = T(0);
    sc[S_BETA] = T(0);
    is[I_IT] = 0;
    // a NaN / Inf in b, M or the mask stops the system with a NaN relative residual (reported as non-finite by the
    // host) instead of passing for a zero right-hand side
    const bool finite = is_finite(bb) && is_finite(rr) && is_finite(rz);
    const bool done = finite && (!(bn > T(0)) || sqrt(rr) <= sc[S_TOL] * bn);
    sc[S_REL] = !finite ? nan_of<T>() : (bn > T(0) ? sqrt(rr) / bn : T(0));
    // r.M r < 0 for the first residual: an indefinite diagonal block of K
    bool done_ = done;
    is[I_SNAPK] = is[I_SNAP0] = is[I_SNAP1] = 0;
    if (sd.nt > 0 && finite) {
      const int k = snap_levels(sd.tols, sd.nt, 0, (double)sqrt(rr), (double)bn, sd.snap_it + (int64_t)s * sd.nt, 0);
      is[I_SNAPK] = is[I_SNAP1] = k;
      done_ = k == sd.nt;
    }
    const bool indef = finite && !done_ && rz < T(0);
    is[I_ACTIVE] = (done_ || indef || !finite) ? 0 : 1;
    is[I_CONV] = done_ ? 1 : 0;
    is[I_NOTSPD] = indef ? 1 : 0;
  }
}

// alpha = rz / pAp from the partials of the masked K p kernel (nblk per system)
template <typename T>
__global__ __launch_bounds__(256) void cg_fin_alpha_kernel(const T* partial, int nblk, T* scal, int* ist) {
  __shared__ T red[4];
  const int s = blockIdx.x;
  int* is = ist + (int64_t)s * NINT;
  const int active = is[I_ACTIVE];
  __syncthreads();
  if (threadIdx.x == 0) is[I_XUPD] = active;
  if (!active) return;
  const T pap = sum_partials(partial + (int64_t)s * nblk, nblk, 1, 0, red);
  if (threadIdx.x == 0) {
    T* sc = scal + (int64_t)s * NSCAL;
    sc[S_PAP] = pap;
    // an active system has r.z > 0, so p = z + beta p != 0 and p.Ap > 0 for SPD K; p.Ap <= 0 means K is not
    // positive definite on the free DOFs: stop the system without an update (a NaN p.Ap is left to the
    // non-finite check of the residual)
    if (pap <= T(0)) {
      sc[S_ALPHA] = T(0);
      is[I_NOTSPD] = 1;
      is[I_ACTIVE] = 0;
      is[I_XUPD] = 0;
    } else {
      sc[S_ALPHA] = sc[S_RZ] / pap;
    }
  }
}

// D components of one site moved by one vector access
template <int D, typename T>
struct Site {
  T c[D];
};
template <int D, typename T>
__device__ __forceinline__ Site<D, T> ld_site(const T* __restrict__ base, int64_t o) {
  return *(const Site<D, T>*)(base + o);
}
template <int D, typename T>
__device__ __forceinline__ void st_site(T* __restrict__ base, int64_t o, const Site<D, T>& v) {
  *(Site<D, T>*)(base + o) = v;
}
constexpr int VU = 4;  // independent sites per thread and pass of the vector kernels (memory-level parallelism)

// r -= alpha mask Ap, z = M r; partials (r.z, r.r).  x += alpha p is applied by cg_pupdate.
template <int D, typename T>
__global__ __launch_bounds__(256) void cg_update_kernel(CgVec<T> v) {
  __shared__ T red[4];
  const int l = blockIdx.y, b = blockIdx.z;
  const int64_t s = (int64_t)b * v.nl + l;
  if (!v.ist[s * NINT + I_ACTIVE]) return;
  constexpr int NS = D == 3 ? 6 : 3;
  const T alpha = v.scal[s * NSCAL + S_ALPHA];
  T* __restrict__ r_ = v.r + s * v.N * D;
  T* __restrict__ z_ = v.z ? v.z + s * v.N * D : nullptr;
  const T* __restrict__ ap_ = v.ap + s * v.N * D;
  const T* __restrict__ mi_ = v.minv ? v.minv + v.bminv.at(b, l) : nullptr;
  const uint8_t* __restrict__ mk_ = v.mask ? v.mask + v.bm.at(b, l) : nullptr;
  T rz = T(0), rr = T(0);
  const int64_t step = (int64_t)gridDim.x * blockDim.x * VU;
  for (int64_t p0 = (int64_t)blockIdx.x * blockDim.x * VU + threadIdx.x; p0 < v.N; p0 += step) {
    Site<D, T> rv[VU], av[VU];
    T m[VU][3];
#pragma unroll
    for (int u = 0; u < VU; ++u) {
      const int64_t p = p0 + (int64_t)u * blockDim.x;
      if (p < v.N) {
        rv[u] = ld_site<D>(r_, p * D);
        av[u] = ld_site<D>(ap_, p * D);
        for (int a = 0; a < D; ++a) m[u][a] = mk_ ? (mk_[D * p + a] ? T(1) : T(0)) : T(1);
      }
    }
#pragma unroll
    for (int u = 0; u < VU; ++u) {
      const int64_t p = p0 + (int64_t)u * blockDim.x;
      if (p < v.N) {
        T r[3], z[3];
        for (int a = 0; a < D; ++a) {
          // the K p kernel leaves fixed DOFs unmasked (p vanishes there, so p.Ap is unaffected); mask here
          r[a] = rv[u].c[a] - alpha * (m[u][a] * av[u].c[a]);
          rv[u].c[a] = r[a];
        }
        st_site<D>(r_, p * D, rv[u]);
        if (mi_) {
          apply_minv<D, T>(mi_ + NS * p, r, z);
          // z (hence p) vanishes on fixed DOFs whatever the preconditioner holds there
          Site<D, T> zv;
          for (int a = 0; a < D; ++a) zv.c[a] = z[a] = m[u][a] * z[a];
          st_site<D>(z_, p * D, zv);
        } else {
          for (int a = 0; a < D; ++a) z[a] = r[a];
        }
        for (int a = 0; a < D; ++a) {
          rz += r[a] * z[a];
          rr += r[a] * r[a];
        }
      }
    }
  }
  const T s0 = block_sum(rz, red);
  const T s1 = block_sum(rr, red);
  if (threadIdx.x == 0) {
    T* pp = v.partial + (s * gridDim.x + blockIdx.x) * 3;
    pp[0] = s0;
    pp[1] = s1;
  }
}

// convergence test and beta (snapshot mode: convergence = every level met)
template <typename T>
__global__ __launch_bounds__(256) void cg_fin_beta_kernel(const T* partial, int nblk, T* scal, int* ist,
                                                          SnapDev<T> sd) {
  __shared__ T red[4];
  const int s = blockIdx.x;
  int* is = ist + (int64_t)s * NINT;
  if (!is[I_ACTIVE]) return;
  const T* pp = partial + (int64_t)s * nblk * 3;
  const T rzn = sum_partials(pp, nblk, 3, 0, red);
  const T rr = sum_partials(pp, nblk, 3, 1, red);
  if (threadIdx.x == 0) {
    T* sc = scal + (int64_t)s * NSCAL;
    const int it = is[I_IT] + 1;
    is[I_IT] = it;
    sc[S_RR] = rr;
    const T rn = sqrt(rr);
    sc[S_REL] = rn / sc[S_BN];
    bool met = rn <= sc[S_TOL] * sc[S_BN];
    if (sd.nt > 0) {
      const int k0 = is[I_SNAPK];
      const int k = is_finite(rn) ? snap_levels(sd.tols, sd.nt, k0, (double)rn, (double)sc[S_BN],
                                                sd.snap_it + (int64_t)s * sd.nt, it)
                                  : k0;
      is[I_SNAP0] = k0;
      is[I_SNAP1] = is[I_SNAPK] = k;
      met = k == sd.nt;
    }
    if (met) {
      is[I_ACTIVE] = 0;
      is[I_CONV] = 1;
    } else if (!is_finite(rn) || !is_finite(rzn)) {
      // NaN / Inf from the coefficients or the preconditioner: stop; the host reports the non-finite residual
      sc[S_REL] = nan_of<T>();
      is[I_ACTIVE] = 0;
    } else if (rzn < T(0)) {
      // r.M r < 0 for r != 0: the preconditioner (the inverse diagonal blocks of K) is indefinite
      is[I_ACTIVE] = 0;
      is[I_NOTSPD] = 1;
    } else if (it >= is[I_MAXIT]) {
      is[I_ACTIVE] = 0;
    } else {
      sc[S_BETA] = sc[S_RZ] != T(0) ? rzn / sc[S_RZ] : T(0);
      sc[S_RZ] = rzn;
    }
  }
}

// x += alpha p (the iteration's update, also in the iteration that converged), then p = z + beta p while the
// system is still active
template <int D, typename T>
__global__ __launch_bounds__(256) void cg_pupdate_kernel(CgVec<T> v, SnapDev<T> sd) {
  const int l = blockIdx.y, b = blockIdx.z;
  const int64_t s = (int64_t)b * v.nl + l;
  if (!v.ist[s * NINT + I_XUPD]) return;
  const bool active = v.ist[s * NINT + I_ACTIVE] != 0;
  // snapshot levels first met in this iteration take the updated iterate
  const int k0 = sd.nt > 0 ? v.ist[s * NINT + I_SNAP0] : 0, k1 = sd.nt > 0 ? v.ist[s * NINT + I_SNAP1] : 0;
  const T alpha = v.scal[s * NSCAL + S_ALPHA];
  const T beta = v.scal[s * NSCAL + S_BETA];
  T* __restrict__ x_ = v.x + s * v.N * D;
  T* __restrict__ p_ = v.p + s * v.N * D;
  const T* __restrict__ z_ = (v.z ? v.z : v.r) + s * v.N * D;
  const int64_t step = (int64_t)gridDim.x * blockDim.x * VU;
  for (int64_t p0 = (int64_t)blockIdx.x * blockDim.x * VU + threadIdx.x; p0 < v.N; p0 += step) {
    Site<D, T> pv[VU], xv[VU], zv[VU];
#pragma unroll
    for (int u = 0; u < VU; ++u) {
      const int64_t q = p0 + (int64_t)u * blockDim.x;
      if (q < v.N) {
        pv[u] = ld_site<D>(p_, q * D);
        xv[u] = ld_site<D>(x_, q * D);
        if (active) zv[u] = ld_site<D>(z_, q * D);
      }
    }
#pragma unroll
    for (int u = 0; u < VU; ++u) {
      const int64_t q = p0 + (int64_t)u * blockDim.x;
      if (q < v.N) {
        for (int a = 0; a < D; ++a) xv[u].c[a] += alpha * pv[u].c[a];
        st_site<D>(x_, q * D, xv[u]);
        for (int k = k0; k < k1; ++k) st_site<D>(sd.snaps + (s * sd.nt + k) * v.N * D, q * D, xv[u]);
        if (active) {
          for (int a = 0; a < D; ++a) pv[u].c[a] = zv[u].c[a] + beta * pv[u].c[a];
          st_site<D>(p_, q * D, pv[u]);
        }
      }
    }
  }
}

// ---- the iteration in three launches (K p, update, p update) instead of five: the scalar steps cg_fin_alpha and
// cg_fin_beta are folded into the vector kernels that consume them.  Every block of a fused kernel sums its
// system's partials of the previous launch itself, in the same fixed order, so all blocks take the same scalars and
// decisions; block 0 records them for later launches and the host.  A block never reads a value another block of
// the same launch writes: I_ACTIVE is read before block 0 may clear it (and a block that reads it cleared has
// p.Ap <= 0 itself), the iteration count is advanced by the update and read by the p update, and the r.z of the new
// residual goes to S_RZN, which the next update copies to S_RZ.  Not used with snapshot levels (cg_fin_beta
// records those).

// cg_fin_alpha + cg_update: r -= alpha mask Ap, z = M r; partials (r.z, r.r) into v.partial (the K p partials are
// kpart, knblk per system).
template <int D, typename T>
__global__ __launch_bounds__(256) void cg_update_fused_kernel(CgVec<T> v, const T* __restrict__ kpart, int knblk) {
  __shared__ T red[4];
  __shared__ T bc;
  const int l = blockIdx.y, b = blockIdx.z;
  const int64_t s = (int64_t)b * v.nl + l;
  int* is = v.ist + s * NINT;
  T* sc = v.scal + s * NSCAL;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!is[I_ACTIVE]) {
    if (lead) is[I_XUPD] = 0;
    return;
  }
  constexpr int NS = D == 3 ? 6 : 3;
  T* __restrict__ r_ = v.r + s * v.N * D;
  T* __restrict__ z_ = v.z ? v.z + s * v.N * D : nullptr;
  const T* __restrict__ ap_ = v.ap + s * v.N * D;
  const T* __restrict__ mi_ = v.minv ? v.minv + v.bminv.at(b, l) : nullptr;
  const uint8_t* __restrict__ mk_ = v.mask ? v.mask + v.bm.at(b, l) : nullptr;
  Site<D, T> rv[VU], av[VU];
  T m[VU][3];
  auto load = [&](int64_t p0) {
#pragma unroll
    for (int u = 0; u < VU; ++u) {
      const int64_t p = p0 + (int64_t)u * blockDim.x;
      if (p < v.N) {
        rv[u] = ld_site<D>(r_, p * D);
        av[u] = ld_site<D>(ap_, p * D);
        for (int a = 0; a < D; ++a) m[u][a] = mk_ ? (mk_[D * p + a] ? T(1) : T(0)) : T(1);
      }
    }
  };
  const int64_t step = (int64_t)gridDim.x * blockDim.x * VU;
  const int64_t first = (int64_t)blockIdx.x * blockDim.x * VU + threadIdx.x;
  // the first group's loads are in flight while the block sums the K p partials
  load(first);
  const T pap = sum_partials_all(kpart + s * knblk, knblk, 1, 0, red, &bc);
  const T rz = sc[S_RZN];
  // an active system has r.z > 0, so p != 0 and p.Ap > 0 for SPD K; p.Ap <= 0: K is not positive definite on the
  // free DOFs, the system stops without an update (a NaN p.Ap is left to the non-finite check of the residual)
  const bool spd = !(pap <= T(0));
  if (lead) {
    sc[S_PAP] = pap;
    sc[S_RZ] = rz;
    sc[S_ALPHA] = spd ? rz / pap : T(0);
    is[I_XUPD] = spd ? 1 : 0;
    if (spd) {
      is[I_IT] = is[I_IT] + 1;
    } else {
      is[I_NOTSPD] = 1;
      is[I_ACTIVE] = 0;
    }
  }
  if (!spd) return;
  const T alpha = rz / pap;
  T rzs = T(0), rr = T(0);
  for (int64_t p0 = first; p0 < v.N; p0 += step) {
    if (p0 != first) load(p0);
#pragma unroll
    for (int u = 0; u < VU; ++u) {
      const int64_t p = p0 + (int64_t)u * blockDim.x;
      if (p < v.N) {
        T r[3], z[3];
        for (int a = 0; a < D; ++a) {
          r[a] = rv[u].c[a] - alpha * (m[u][a] * av[u].c[a]);
          rv[u].c[a] = r[a];
        }
        st_site<D>(r_, p * D, rv[u]);
        if (mi_) {
          apply_minv<D, T>(mi_ + NS * p, r, z);
          Site<D, T> zv;
          for (int a = 0; a < D; ++a) zv.c[a] = z[a] = m[u][a] * z[a];
          st_site<D>(z_, p * D, zv);
        } else {
          for (int a = 0; a < D; ++a) z[a] = r[a];
        }
        for (int a = 0; a < D; ++a) {
          rzs += r[a] * z[a];
          rr += r[a] * r[a];
        }
      }
    }
  }
  const T s0 = block_sum(rzs, red);
  const T s1 = block_sum(rr, red);
  if (threadIdx.x == 0) {
    T* pp = v.partial + (s * gridDim.x + blockIdx.x) * 3;
    pp[0] = s0;
    pp[1] = s1;
  }
}

// cg_fin_beta + cg_pupdate: convergence / failure decision and beta from the update's partials (upart, unblk per
// system), x += alpha p, and p = z + beta p while the system goes on.
template <int D, typename T>
__global__ __launch_bounds__(256) void cg_pupdate_fused_kernel(CgVec<T> v, const T* __restrict__ upart, int unblk) {
  __shared__ T red[4];
  __shared__ T bc;
  const int l = blockIdx.y, b = blockIdx.z;
  const int64_t s = (int64_t)b * v.nl + l;
  int* is = v.ist + s * NINT;
  T* sc = v.scal + s * NSCAL;
  if (!is[I_XUPD]) return;
  T* __restrict__ x_ = v.x + s * v.N * D;
  T* __restrict__ p_ = v.p + s * v.N * D;
  const T* __restrict__ z_ = (v.z ? v.z : v.r) + s * v.N * D;
  Site<D, T> pv[VU], xv[VU], zv[VU];
  auto load_px = [&](int64_t p0) {
#pragma unroll
    for (int u = 0; u < VU; ++u) {
      const int64_t q = p0 + (int64_t)u * blockDim.x;
      if (q < v.N) {
        pv[u] = ld_site<D>(p_, q * D);
        xv[u] = ld_site<D>(x_, q * D);
      }
    }
  };
  const int64_t step = (int64_t)gridDim.x * blockDim.x * VU;
  const int64_t first = (int64_t)blockIdx.x * blockDim.x * VU + threadIdx.x;
  // p and x of the first group are in flight while the block sums the update's partials
  load_px(first);
  const T* pp = upart + s * unblk * 3;
  const T rzn = sum_partials_all(pp, unblk, 3, 0, red, &bc);
  const T rr = sum_partials_all(pp, unblk, 3, 1, red, &bc);
  const int it = is[I_IT];
  const T bn = sc[S_BN], rn = sqrt(rr);
  // the decisions of cg_fin_beta, in its order
  const bool met = rn <= sc[S_TOL] * bn;
  const bool bad = !met && (!is_finite(rn) || !is_finite(rzn));
  const bool indef = !met && !bad && rzn < T(0);
  const bool cap = !met && !bad && !indef && it >= is[I_MAXIT];
  const bool active = !met && !bad && !indef && !cap;
  const T rz = sc[S_RZ];
  const T beta = rz != T(0) ? rzn / rz : T(0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[S_RR] = rr;
    sc[S_REL] = bad ? nan_of<T>() : rn / bn;
    if (met) is[I_CONV] = 1;
    if (indef) is[I_NOTSPD] = 1;
    if (!active) is[I_ACTIVE] = 0;
    if (active) {
      sc[S_BETA] = beta;
      sc[S_RZN] = rzn;
    }
  }
  const T alpha = sc[S_ALPHA];
  for (int64_t p0 = first; p0 < v.N; p0 += step) {
    if (p0 != first) load_px(p0);
#pragma unroll
    for (int u = 0; u < VU; ++u) {
      const int64_t q = p0 + (int64_t)u * blockDim.x;
      if (q < v.N && active) zv[u] = ld_site<D>(z_, q * D);
    }
#pragma unroll
    for (int u = 0; u < VU; ++u) {
      const int64_t q = p0 + (int64_t)u * blockDim.x;
      if (q < v.N) {
        for (int a = 0; a < D; ++a) xv[u].c[a] += alpha * pv[u].c[a];
        st_site<D>(x_, q * D, xv[u]);
        if (active) {
          for (int a = 0; a < D; ++a) pv[u].c[a] = zv[u].c[a] + beta * pv[u].c[a];
          st_site<D>(p_, q * D, pv[u]);
        }
      }
    }
  }
}

// x into the snapshot levels [I_SNAP0, I_SNAP1) of each system (grid (blocks, nl, nb)): the levels met by the
// initial residual, and at the end the levels never met
template <int D, typename T>
__global__ __launch_bounds__(256) void cg_snap_copy_kernel(CgVec<T> v, SnapDev<T> sd) {
  const int l = blockIdx.y, b = blockIdx.z;
  const int64_t s = (int64_t)b * v.nl + l;
  const int k0 = v.ist[s * NINT + I_SNAP0], k1 = v.ist[s * NINT + I_SNAP1];
  if (k1 <= k0) return;
  const T* __restrict__ x_ = v.x + s * v.N * D;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.N * D; i += (int64_t)gridDim.x * blockDim.x) {
    const T xi = x_[i];
    for (int k = k0; k < k1; ++k) sd.snaps[(s * sd.nt + k) * v.N * D + i] = xi;
  }
}
// levels never met: iteration -1, snapshot = the final iterate (timing/t2d/solver2d.py:417-420)
template <typename T>
__global__ void cg_snap_rest_kernel(int* ist, int nsys, SnapDev<T> sd) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  int* is = ist + (int64_t)s * NINT;
  const int k0 = is[I_SNAPK];
  for (int k = k0; k < sd.nt; ++k) sd.snap_it[(int64_t)s * sd.nt + k] = -1;
  is[I_SNAP0] = k0;
  is[I_SNAP1] = sd.nt;
}

// ------------------------------------------------------------------------------------------ refinement helpers
// lo = (TL) hi elementwise (n values)
template <typename TH, typename TL>
__global__ __launch_bounds__(256) void cast_kernel(const TH* __restrict__ hi, TL* __restrict__ lo, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lo[i] = (TL)hi[i];
}
// hi += (TH) lo
template <typename TH, typename TL>
__global__ __launch_bounds__(256) void add_cast_kernel(TH* __restrict__ hi, const TL* __restrict__ lo, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) hi[i] += (TH)lo[i];
}
// sum of squares per system of partials written by the residual kernel (nblk per system) -> out[s]
template <typename T>
__global__ __launch_bounds__(256) void sumsq_fin_kernel(const T* partial, int nblk, double* out) {
  __shared__ T red[4];
  const int s = blockIdx.x;
  const T v = sum_partials(partial + (int64_t)s * nblk, nblk, 1, 0, red);
  if (threadIdx.x == 0) out[s] = (double)v;
}
// ||mask * b||^2 per system (grid (ceil(N/256), nl, nb)); partial layout (systems, gridDim.x)
template <int D, typename T>
__global__ __launch_bounds__(256) void masked_norm2_kernel(const T* __restrict__ bvec, int64_t N, int nl,
                                                           const uint8_t* __restrict__ mask, Bcast bm,
                                                           T* __restrict__ partial) {
  __shared__ T red[4];
  const int l = blockIdx.y, b = blockIdx.z;
  const int64_t s = (int64_t)b * nl + l;
  T acc = T(0);
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
    for (int a = 0; a < D; ++a) {
      const T m = mask ? (mask[bm.at(b, l) + D * p + a] ? T(1) : T(0)) : T(1);
      const T v = m * bvec[(s * N + p) * D + a];
      acc += v * v;
    }
  }
  const T t = block_sum(acc, red);
  if (threadIdx.x == 0) partial[s * gridDim.x + blockIdx.x] = t;
}

}  // namespace lpm
