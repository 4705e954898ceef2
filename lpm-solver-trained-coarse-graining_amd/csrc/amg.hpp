// Kernels of the aggregation AMG preconditioner of the 3D SC lattice (timing/t3d/amg3d.py).
//
// Level 0 is the particle lattice: the free particles form a box (first particle lo, extent L per axis), aggregated
// into chunks of `agg` particles per axis with the remainder merged into the last chunk (amg3d.py:42-48), aggregate
// I = (cz gy + cy) gx + cx.  The tentative prolongator block P_p (3 x 6) of a free particle is the orthonormalised
// rigid-body basis of its aggregate (gmg.py:70-162); Gram-Schmidt removes the translation part of the rotations
// about the lattice centroid, so the block depends only on the aggregate's shape (3 or the merged length per axis,
// at most 8 shapes) and the particle's place in it: the kernels read it from shape templates.
//
// Levels >= 1 are node grids with 6 unknowns per node and a 27-point stencil of 6x6 blocks, offset
// d = (dz+1) 9 + (dy+1) 3 + (dx+1) (amg3d.py:39).  Galerkin operators are symmetric, A[I, d] = A[I + off(d), 26 - d]^T,
// so only the blocks d = 13 ... 26 (the diagonal and the 13 forward offsets) are stored, e = d - 13, structure-of-
// arrays A[(e 36 + r 6 + c) n + I] so that consecutive nodes read consecutive addresses; a block with d < 13 is read
// transposed from the neighbour's e = 13 - d.  Nodes are aggregated like the particles, with per-node blocks
// Q (n, 6, 6) row-major.
//
// Galerkin operators are assembled by probing: with aggregates coloured by (I mod 3) per axis, two aggregates of
// one colour are at least 3 chunks apart, while the operator couples only neighbouring aggregates (the particle
// operator reaches 2 sites, a chunk has at least 3; a stencil reaches 1 node), so P^T A (sum over one colour of
// P e_j) gives, for every aggregate I, column j of its block towards the one aggregate of that colour next to it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lpm {

struct Amg0Geo {
  int nx, ny, nz;
  int lo[3], L[3], g[3];
  int agg;
  int64_t toff[8];  // template of shape s (bit a = last chunk along axis a): (sx sy sz, 3, 6) at tmpl + toff[s]
};

constexpr int ST_BLOCKS = 14;  // stored stencil blocks per node (d = 13 ... 26)

struct AmgLGeo {
  int g[3];   // node grid of this level
  int gc[3];  // aggregate grid (next level)
  int agg;
};

__device__ __forceinline__ int chunk_of(int rel, int agg, int nch) { return min(rel / agg, nch - 1); }
__device__ __forceinline__ int chunk_len(int c, int agg, int nch, int L) { return c == nch - 1 ? L - agg * (nch - 1) : agg; }

// Level-0 coordinates of particle (x, y, z): aggregate, template row pointer; false outside the free box
template <typename T>
__device__ __forceinline__ bool p0_locate(const Amg0Geo& a, const T* __restrict__ tmpl, int x, int y, int z, int& I,
                                          int cc[3], const T*& prow) {
  const int p[3] = {x - a.lo[0], y - a.lo[1], z - a.lo[2]};
  int loc[3], len[3], s = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (p[k] < 0 || p[k] >= a.L[k]) return false;
    cc[k] = chunk_of(p[k], a.agg, a.g[k]);
    loc[k] = p[k] - a.agg * cc[k];
    len[k] = chunk_len(cc[k], a.agg, a.g[k], a.L[k]);
    if (cc[k] == a.g[k] - 1) s |= 1 << k;
  }
  I = (cc[2] * a.g[1] + cc[1]) * a.g[0] + cc[0];
  prow = tmpl + a.toff[s] + (int64_t)((loc[2] * len[1] + loc[1]) * len[0] + loc[0]) * 18;
  return true;
}

// the aggregate of colour col next to aggregate coordinates ic (one axis): -1 if none in [0, g)
__device__ __forceinline__ int colour_partner(int ic, int col, int g) {
#pragma unroll
  for (int o = -1; o <= 1; ++o) {
    const int j = ic + o;
    if (j >= 0 && j < g && j % 3 == col) return j;
  }
  return -1;
}

// x_p += P_p e_{agg(p)} on the free particles (amg3d.py:318-331, level 0)
template <typename T>
__global__ __launch_bounds__(256) void amg0_prolong_kernel(Amg0Geo a, const T* __restrict__ tmpl,
                                                           const T* __restrict__ e, T* __restrict__ xo) {
  const int64_t N = (int64_t)a.nx * a.ny * a.nz;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int x = (int)(p % a.nx), y = (int)((p / a.nx) % a.ny), z = (int)(p / ((int64_t)a.nx * a.ny));
  int I, cc[3];
  const T* pr;
  if (!p0_locate(a, tmpl, x, y, z, I, cc, pr)) return;
  T ev[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) ev[c] = e[(int64_t)I * 6 + c];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < 6; ++c) s += pr[r * 6 + c] * ev[c];
    xo[p * 3 + r] += s;
  }
}

// rc_I = sum_{p in I} P_p^T r_p (amg3d.py:303-315, level 0); one thread per aggregate, members in a fixed order
template <typename T>
__global__ __launch_bounds__(256) void amg0_restrict_kernel(Amg0Geo a, const T* __restrict__ tmpl,
                                                            const T* __restrict__ r, T* __restrict__ rc) {
  const int nagg = a.g[0] * a.g[1] * a.g[2];
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nagg) return;
  const int cc[3] = {I % a.g[0], (I / a.g[0]) % a.g[1], I / (a.g[0] * a.g[1])};
  int s0[3], len[3], sh = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s0[k] = a.lo[k] + a.agg * cc[k];
    len[k] = chunk_len(cc[k], a.agg, a.g[k], a.L[k]);
    if (cc[k] == a.g[k] - 1) sh |= 1 << k;
  }
  const T* t = tmpl + a.toff[sh];
  T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  int row = 0;
  for (int lz = 0; lz < len[2]; ++lz)
    for (int ly = 0; ly < len[1]; ++ly)
      for (int lx = 0; lx < len[0]; ++lx, ++row) {
        const int64_t p = ((int64_t)(s0[2] + lz) * a.ny + (s0[1] + ly)) * a.nx + (s0[0] + lx);
        const T r0 = r[p * 3], r1 = r[p * 3 + 1], r2 = r[p * 3 + 2];
        const T* pr = t + (int64_t)row * 18;
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] += pr[c] * r0 + pr[6 + c] * r1 + pr[12 + c] * r2;
      }
#pragma unroll
  for (int c = 0; c < 6; ++c) rc[(int64_t)I * 6 + c] = acc[c];
}

// probe fields of colour col: v_j(p) = column j of P_p on aggregates of that colour, 0 elsewhere; v (6, N, 3)
template <typename T>
__global__ __launch_bounds__(256) void amg0_probe_kernel(Amg0Geo a, const T* __restrict__ tmpl, int col,
                                                         T* __restrict__ v) {
  const int64_t N = (int64_t)a.nx * a.ny * a.nz;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int x = (int)(p % a.nx), y = (int)((p / a.nx) % a.ny), z = (int)(p / ((int64_t)a.nx * a.ny));
  int I, cc[3];
  const T* pr;
  bool on = p0_locate(a, tmpl, x, y, z, I, cc, pr);
  on = on && cc[0] % 3 == col % 3 && cc[1] % 3 == (col / 3) % 3 && cc[2] % 3 == col / 9;
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int r = 0; r < 3; ++r) v[(int64_t)j * N * 3 + p * 3 + r] = on ? pr[r * 6 + j] : T(0);
}

// column j of the blocks A1[I, d(I, col)] = P_I^T (K v_j) from the probe products kv (6, N, 3) (replaces the
// bond-level Galerkin product of amg3d.py:151-196)
template <typename T>
__global__ __launch_bounds__(256) void amg0_gather_kernel(Amg0Geo a, const T* __restrict__ tmpl, int col,
                                                          const T* __restrict__ kv, T* __restrict__ A) {
  const int nagg = a.g[0] * a.g[1] * a.g[2];
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nagg) return;
  const int64_t N = (int64_t)a.nx * a.ny * a.nz;
  const int cc[3] = {I % a.g[0], (I / a.g[0]) % a.g[1], I / (a.g[0] * a.g[1])};
  const int cols[3] = {col % 3, (col / 3) % 3, col / 9};
  int jc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    jc[k] = colour_partner(cc[k], cols[k], a.g[k]);
    if (jc[k] < 0) return;
  }
  const int d = (jc[2] - cc[2] + 1) * 9 + (jc[1] - cc[1] + 1) * 3 + (jc[0] - cc[0] + 1);
  if (d < 13) return;  // stored as the transpose of the partner's block
  int s0[3], len[3], sh = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s0[k] = a.lo[k] + a.agg * cc[k];
    len[k] = chunk_len(cc[k], a.agg, a.g[k], a.L[k]);
    if (cc[k] == a.g[k] - 1) sh |= 1 << k;
  }
  const T* t = tmpl + a.toff[sh];
  for (int j = 0; j < 6; ++j) {
    const T* w = kv + (int64_t)j * N * 3;
    T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    int row = 0;
    for (int lz = 0; lz < len[2]; ++lz)
      for (int ly = 0; ly < len[1]; ++ly)
        for (int lx = 0; lx < len[0]; ++lx, ++row) {
          const int64_t p = ((int64_t)(s0[2] + lz) * a.ny + (s0[1] + ly)) * a.nx + (s0[0] + lx);
          const T w0 = w[p * 3], w1 = w[p * 3 + 1], w2 = w[p * 3 + 2];
          const T* pr = t + (int64_t)row * 18;
#pragma unroll
          for (int r = 0; r < 6; ++r) acc[r] += pr[r] * w0 + pr[6 + r] * w1 + pr[12 + r] * w2;
        }
#pragma unroll
    for (int r = 0; r < 6; ++r) A[((int64_t)(d - 13) * 36 + r * 6 + j) * nagg + I] = acc[r];
  }
}

// ---------------------------------------------------------------------------------------------- levels >= 1
__device__ __forceinline__ void node_xyz(int n, const int g[3], int c[3]) {
  c[0] = n % g[0];
  c[1] = (n / g[0]) % g[1];
  c[2] = n / (g[0] * g[1]);
}

// y_s = A x_s (b null) or y_s = b - A x_s on a 27-point 6x6 stencil (amg3d.py:230-253) for NS systems sharing A
// (probing); b is used with NS = 1 only.  One thread per node: each stored block is read once for all systems,
// the forward blocks of the node and, transposed, the backward ones from its neighbours.
template <typename T, int NS>
__global__ __launch_bounds__(256) void amg_st_mv_kernel(int gx, int gy, int gz, const T* __restrict__ A,
                                                        const T* __restrict__ x, const T* __restrict__ b,
                                                        T* __restrict__ y) {
  const int n = gx * gy * gz;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  const int g[3] = {gx, gy, gz};
  int c[3];
  node_xyz(I, g, c);
  T acc[NS][6];
#pragma unroll
  for (int s = 0; s < NS; ++s)
#pragma unroll
    for (int r = 0; r < 6; ++r) acc[s][r] = T(0);
  for (int q = 0; q < 27; ++q) {
    const int j0 = c[0] + q % 3 - 1, j1 = c[1] + (q / 3) % 3 - 1, j2 = c[2] + q / 9 - 1;
    if (j0 < 0 || j0 >= gx || j1 < 0 || j1 >= gy || j2 < 0 || j2 >= gz) continue;
    const int J = (j2 * gy + j1) * gx + j0;
    // block (I, J): stored at I (e = q - 13) for q >= 13, else the transpose of J's block e = 13 - q
    const bool fwd = q >= 13;
    const int e = fwd ? q - 13 : 13 - q;
    const int owner = fwd ? I : J;
    T ab[36];
    const T* Ad = A + (int64_t)e * 36 * n + owner;
#pragma unroll
    for (int k = 0; k < 36; ++k) ab[k] = Ad[(int64_t)k * n];
    if (!fwd) {
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int k = r + 1; k < 6; ++k) {
          const T t = ab[r * 6 + k];
          ab[r * 6 + k] = ab[k * 6 + r];
          ab[k * 6 + r] = t;
        }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      T xv[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) xv[k] = x[((int64_t)s * n + J) * 6 + k];
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[s][r] += ab[r * 6 + k] * xv[k];
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    T* ys = y + ((int64_t)s * n + I) * 6;
#pragma unroll
    for (int r = 0; r < 6; ++r) ys[r] = (NS == 1 && b) ? b[(int64_t)I * 6 + r] - acc[s][r] : acc[s][r];
  }
}

// y = A x or y = b - A x as amg_st_mv_kernel with NS = 1, one thread per (row r, node): six times the threads, for
// the small levels where one thread per node leaves the chip mostly idle.  Thread t: r = t / n, node t % n.
template <typename T>
__global__ __launch_bounds__(256) void amg_st_mv_rows_kernel(int gx, int gy, int gz, const T* __restrict__ A,
                                                             const T* __restrict__ x, const T* __restrict__ b,
                                                             T* __restrict__ y) {
  const int n = gx * gy * gz;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * n) return;
  const int r = t / n, I = t - r * n;
  const int g[3] = {gx, gy, gz};
  int c[3];
  node_xyz(I, g, c);
  T acc = T(0);
  for (int q = 0; q < 27; ++q) {
    const int j0 = c[0] + q % 3 - 1, j1 = c[1] + (q / 3) % 3 - 1, j2 = c[2] + q / 9 - 1;
    if (j0 < 0 || j0 >= gx || j1 < 0 || j1 >= gy || j2 < 0 || j2 >= gz) continue;
    const int J = (j2 * gy + j1) * gx + j0;
    const bool fwd = q >= 13;
    const int e = fwd ? q - 13 : 13 - q;
    // row r of block (I, J): entries (r, k) of the stored forward block, or (k, r) of J's stored block
    const T* Ad = A + (int64_t)e * 36 * n + (fwd ? I : J);
    const int o0 = fwd ? r * 6 : r, ok = fwd ? 1 : 6;
    const T* xj = x + (int64_t)J * 6;
#pragma unroll
    for (int k = 0; k < 6; ++k) acc += Ad[(int64_t)(o0 + k * ok) * n] * xj[k];
  }
  y[(int64_t)I * 6 + r] = b ? b[(int64_t)I * 6 + r] - acc : acc;
}

// y = A x or y = b - A x as amg_st_mv_kernel with NS = 1 in two passes that read every stored block once.  Pass 1:
// the forward products of node I (blocks e = 0..13, stored at I) into yp (6, n), and the transposed products
// tt[e - 1] = A_e(I)^T x_I (6, n) for e = 1..13, which belong to the neighbour I + off_e.  Pass 2: y_I = yp_I +
// sum_e tt[e - 1](I - off_e) (or b minus that).
template <typename T>
__global__ __launch_bounds__(256) void amg_st_fwd_kernel(int gx, int gy, int gz, const T* __restrict__ A,
                                                         const T* __restrict__ x, T* __restrict__ yp,
                                                         T* __restrict__ tt) {
  const int n = gx * gy * gz;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  const int g[3] = {gx, gy, gz};
  int c[3];
  node_xyz(I, g, c);
  T xi[6], acc[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    xi[k] = x[(int64_t)I * 6 + k];
    acc[k] = T(0);
  }
  for (int e = 0; e < 14; ++e) {
    const int q = 13 + e;
    const int j0 = c[0] + q % 3 - 1, j1 = c[1] + (q / 3) % 3 - 1, j2 = c[2] + q / 9 - 1;
    if (j0 < 0 || j0 >= gx || j1 < 0 || j1 >= gy || j2 < 0 || j2 >= gz) continue;
    const int J = (j2 * gy + j1) * gx + j0;
    T ab[36], xv[6];
    const T* Ad = A + (int64_t)e * 36 * n + I;
#pragma unroll
    for (int k = 0; k < 36; ++k) ab[k] = Ad[(int64_t)k * n];
#pragma unroll
    for (int k = 0; k < 6; ++k) xv[k] = x[(int64_t)J * 6 + k];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int k = 0; k < 6; ++k) acc[r] += ab[r * 6 + k] * xv[k];
    if (e > 0) {
      T* te = tt + (int64_t)(e - 1) * 6 * n + I;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        T t = T(0);
#pragma unroll
        for (int r = 0; r < 6; ++r) t += ab[r * 6 + k] * xi[r];
        te[(int64_t)k * n] = t;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) yp[(int64_t)r * n + I] = acc[r];
}

template <typename T>
__global__ __launch_bounds__(256) void amg_st_bwd_kernel(int gx, int gy, int gz, const T* __restrict__ yp,
                                                         const T* __restrict__ tt, const T* __restrict__ b,
                                                         T* __restrict__ y) {
  const int n = gx * gy * gz;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  const int g[3] = {gx, gy, gz};
  int c[3];
  node_xyz(I, g, c);
  T acc[6];
#pragma unroll
  for (int r = 0; r < 6; ++r) acc[r] = yp[(int64_t)r * n + I];
  for (int e = 1; e < 14; ++e) {
    const int q = 13 + e;  // the neighbour I - off_e owns the block
    const int j0 = c[0] - (q % 3 - 1), j1 = c[1] - ((q / 3) % 3 - 1), j2 = c[2] - (q / 9 - 1);
    if (j0 < 0 || j0 >= gx || j1 < 0 || j1 >= gy || j2 < 0 || j2 >= gz) continue;
    const T* te = tt + (int64_t)(e - 1) * 6 * n + (j2 * gy + j1) * gx + j0;
#pragma unroll
    for (int r = 0; r < 6; ++r) acc[r] += te[(int64_t)r * n];
  }
#pragma unroll
  for (int r = 0; r < 6; ++r) y[(int64_t)I * 6 + r] = b ? b[(int64_t)I * 6 + r] - acc[r] : acc[r];
}

// members of aggregate K of a node level: first node and lengths per axis
__device__ __forceinline__ void agg_box(const AmgLGeo& l, int K, int s0[3], int len[3], int kc[3]) {
  node_xyz(K, l.gc, kc);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s0[k] = l.agg * kc[k];
    len[k] = chunk_len(kc[k], l.agg, l.gc[k], l.g[k]);
  }
}

// rc_K = sum_{n in K} Q_n^T r_n (amg3d.py:303-315 on a node level)
template <typename T>
__global__ __launch_bounds__(256) void amgq_restrict_kernel(AmgLGeo l, const T* __restrict__ Q, const T* __restrict__ r,
                                                            T* __restrict__ rc) {
  const int nc = l.gc[0] * l.gc[1] * l.gc[2];
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  if (K >= nc) return;
  int s0[3], len[3], kc[3];
  agg_box(l, K, s0, len, kc);
  T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int z = 0; z < len[2]; ++z)
    for (int y = 0; y < len[1]; ++y)
      for (int x = 0; x < len[0]; ++x) {
        const int64_t n = ((int64_t)(s0[2] + z) * l.g[1] + (s0[1] + y)) * l.g[0] + (s0[0] + x);
        const T* q = Q + n * 36;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
          const T rv = r[n * 6 + a];
#pragma unroll
          for (int c = 0; c < 6; ++c) acc[c] += q[a * 6 + c] * rv;
        }
      }
#pragma unroll
  for (int c = 0; c < 6; ++c) rc[(int64_t)K * 6 + c] = acc[c];
}

// first stage of the node-level restriction: tmp_n = Q_n^T r_n, one thread per node (coalesced)
template <typename T>
__global__ __launch_bounds__(256) void amgq_qtr_kernel(int n, const T* __restrict__ Q, const T* __restrict__ r,
                                                       T* __restrict__ tmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T* q = Q + (int64_t)i * 36;
  T rv[6];
#pragma unroll
  for (int a = 0; a < 6; ++a) rv[a] = r[(int64_t)i * 6 + a];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    T s = T(0);
#pragma unroll
    for (int a = 0; a < 6; ++a) s += q[a * 6 + c] * rv[a];
    tmp[(int64_t)i * 6 + c] = s;
  }
}

// second stage: rc_K = sum of tmp over the members of K in a fixed order
template <typename T>
__global__ __launch_bounds__(256) void amgq_sum_kernel(AmgLGeo l, const T* __restrict__ tmp, T* __restrict__ rc) {
  const int nc = l.gc[0] * l.gc[1] * l.gc[2];
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  if (K >= nc) return;
  int s0[3], len[3], kc[3];
  agg_box(l, K, s0, len, kc);
  T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int z = 0; z < len[2]; ++z)
    for (int y = 0; y < len[1]; ++y)
      for (int x = 0; x < len[0]; ++x) {
        const int64_t m = ((int64_t)(s0[2] + z) * l.g[1] + (s0[1] + y)) * l.g[0] + (s0[0] + x);
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] += tmp[m * 6 + c];
      }
#pragma unroll
  for (int c = 0; c < 6; ++c) rc[(int64_t)K * 6 + c] = acc[c];
}

// x_n += Q_n e_{agg(n)} (amg3d.py:318-331 on a node level)
template <typename T>
__global__ __launch_bounds__(256) void amgq_prolong_kernel(AmgLGeo l, const T* __restrict__ Q, const T* __restrict__ e,
                                                           T* __restrict__ xo) {
  const int n = l.g[0] * l.g[1] * l.g[2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c[3];
  node_xyz(i, l.g, c);
  const int K = (chunk_of(c[2], l.agg, l.gc[2]) * l.gc[1] + chunk_of(c[1], l.agg, l.gc[1])) * l.gc[0] +
                chunk_of(c[0], l.agg, l.gc[0]);
  const T* q = Q + (int64_t)i * 36;
  T ev[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) ev[k] = e[(int64_t)K * 6 + k];
#pragma unroll
  for (int a = 0; a < 6; ++a) {
    T s = T(0);
#pragma unroll
    for (int k = 0; k < 6; ++k) s += q[a * 6 + k] * ev[k];
    xo[(int64_t)i * 6 + a] += s;
  }
}

// probe fields of colour col on a node level: v_j(n) = column j of Q_n on aggregates of that colour; v (6, n, 6)
template <typename T>
__global__ __launch_bounds__(256) void amgq_probe_kernel(AmgLGeo l, const T* __restrict__ Q, int col,
                                                         T* __restrict__ v) {
  const int n = l.g[0] * l.g[1] * l.g[2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c[3];
  node_xyz(i, l.g, c);
  const int k0 = chunk_of(c[0], l.agg, l.gc[0]), k1 = chunk_of(c[1], l.agg, l.gc[1]), k2 = chunk_of(c[2], l.agg, l.gc[2]);
  const bool on = k0 % 3 == col % 3 && k1 % 3 == (col / 3) % 3 && k2 % 3 == col / 9;
  const T* q = Q + (int64_t)i * 36;
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int a = 0; a < 6; ++a) v[((int64_t)j * n + i) * 6 + a] = on ? q[a * 6 + j] : T(0);
}

// column j of Ac[K, d(K, col)] = Q_K^T (A v_j) from w (6, n, 6) (replaces amg3d.py:256-300)
template <typename T>
__global__ __launch_bounds__(256) void amgq_gather_kernel(AmgLGeo l, const T* __restrict__ Q, int col,
                                                          const T* __restrict__ w, T* __restrict__ Ac) {
  const int nc = l.gc[0] * l.gc[1] * l.gc[2];
  const int n = l.g[0] * l.g[1] * l.g[2];
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  if (K >= nc) return;
  int s0[3], len[3], kc[3];
  agg_box(l, K, s0, len, kc);
  const int cols[3] = {col % 3, (col / 3) % 3, col / 9};
  int jc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    jc[k] = colour_partner(kc[k], cols[k], l.gc[k]);
    if (jc[k] < 0) return;
  }
  const int d = (jc[2] - kc[2] + 1) * 9 + (jc[1] - kc[1] + 1) * 3 + (jc[0] - kc[0] + 1);
  if (d < 13) return;  // stored as the transpose of the partner's block
  for (int j = 0; j < 6; ++j) {
    const T* wj = w + (int64_t)j * n * 6;
    T acc[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    for (int z = 0; z < len[2]; ++z)
      for (int y = 0; y < len[1]; ++y)
        for (int x = 0; x < len[0]; ++x) {
          const int64_t m = ((int64_t)(s0[2] + z) * l.g[1] + (s0[1] + y)) * l.g[0] + (s0[0] + x);
          const T* q = Q + m * 36;
#pragma unroll
          for (int a = 0; a < 6; ++a) {
            const T wv = wj[m * 6 + a];
#pragma unroll
            for (int r = 0; r < 6; ++r) acc[r] += q[a * 6 + r] * wv;
          }
        }
#pragma unroll
    for (int r = 0; r < 6; ++r) Ac[((int64_t)(d - 13) * 36 + r * 6 + j) * nc + K] = acc[r];
  }
}

// Minv_I = inverse of the diagonal block A[I, 13] (Gauss-Jordan with partial pivoting), SoA (36, n); a singular
// block sets *bad
template <typename T>
__global__ __launch_bounds__(256) void amg_inv_diag_kernel(int n, const T* __restrict__ A, T* __restrict__ Minv,
                                                           int* bad) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  double m[6][12];
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      m[r][c] = (double)A[((int64_t)0 * 36 + r * 6 + c) * n + I];  // e = 0: the diagonal block
      m[r][6 + c] = r == c ? 1.0 : 0.0;
    }
  for (int k = 0; k < 6; ++k) {
    int pv = k;
    for (int r = k + 1; r < 6; ++r)
      if (fabs(m[r][k]) > fabs(m[pv][k])) pv = r;
    if (!(fabs(m[pv][k]) > 0.0)) {
      atomicExch(bad, 1);
      return;
    }
    if (pv != k)
      for (int c = 0; c < 12; ++c) {
        const double tmp = m[k][c];
        m[k][c] = m[pv][c];
        m[pv][c] = tmp;
      }
    const double ip = 1.0 / m[k][k];
    for (int c = 0; c < 12; ++c) m[k][c] *= ip;
    for (int r = 0; r < 6; ++r) {
      if (r == k) continue;
      const double f = m[r][k];
      for (int c = 0; c < 12; ++c) m[r][c] -= f * m[k][c];
    }
  }
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = 0; c < 6; ++c) Minv[(int64_t)(r * 6 + c) * n + I] = (T)m[r][6 + c];
}

// x = w Minv b (add = 0) or x += w Minv b (add = 1), 6x6 SoA blocks (amg3d.py:334-356 on a node level)
template <typename T>
__global__ __launch_bounds__(256) void amg_bj6_kernel(int n, const T* __restrict__ Minv, const T* __restrict__ b, T w,
                                                      int add, T* __restrict__ xo) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n) return;
  T bv[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) bv[k] = b[(int64_t)I * 6 + k];
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    T s = T(0);
#pragma unroll
    for (int k = 0; k < 6; ++k) s += Minv[(int64_t)(r * 6 + k) * n + I] * bv[k];
    xo[(int64_t)I * 6 + r] = add ? xo[(int64_t)I * 6 + r] + w * s : w * s;
  }
}

// particle block Jacobi: x = w M b or x += w M b with the symmetric inverse blocks (N, 6) of lpm_block_jacobi
// (xx yy zz yz xz xy, zero on fixed DOFs) (amg3d.py:334-356 on the particles)
template <typename T>
__global__ __launch_bounds__(256) void amg_bj3_kernel(int64_t N, const T* __restrict__ M, const T* __restrict__ b, T w,
                                                      int add, T* __restrict__ xo) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= N) return;
  const T* m = M + p * 6;
  const T b0 = b[p * 3], b1 = b[p * 3 + 1], b2 = b[p * 3 + 2];
  const T z0 = m[0] * b0 + m[5] * b1 + m[4] * b2;
  const T z1 = m[5] * b0 + m[1] * b1 + m[3] * b2;
  const T z2 = m[4] * b0 + m[3] * b1 + m[2] * b2;
  if (add) {
    xo[p * 3] += w * z0;
    xo[p * 3 + 1] += w * z1;
    xo[p * 3 + 2] += w * z2;
  } else {
    xo[p * 3] = w * z0;
    xo[p * 3 + 1] = w * z1;
    xo[p * 3 + 2] = w * z2;
  }
}

// fine-level residual from a plain K z product kz: r = mask (b - kz), in place over kz (the masked residual of
// _kx_free, amg3d.py:199-229)
template <typename T>
__global__ __launch_bounds__(256) void amg_resid_kernel(int64_t n, const T* __restrict__ b,
                                                        const uint8_t* __restrict__ mask, T* __restrict__ kz) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    kz[i] = mask[i] ? b[i] - kz[i] : T(0);
}

// one fine-level smoothing sweep from a plain K z product: z += w M mask (b - kz) (amg3d.py:640-645)
template <typename T>
__global__ __launch_bounds__(256) void amg_resid_bj3_kernel(int64_t N, const T* __restrict__ M, const T* __restrict__ b,
                                                            const uint8_t* __restrict__ mask, const T* __restrict__ kz,
                                                            T w, T* __restrict__ z) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= N) return;
  const T r0 = mask[p * 3] ? b[p * 3] - kz[p * 3] : T(0);
  const T r1 = mask[p * 3 + 1] ? b[p * 3 + 1] - kz[p * 3 + 1] : T(0);
  const T r2 = mask[p * 3 + 2] ? b[p * 3 + 2] - kz[p * 3 + 2] : T(0);
  const T* m = M + p * 6;
  z[p * 3] += w * (m[0] * r0 + m[5] * r1 + m[4] * r2);
  z[p * 3 + 1] += w * (m[5] * r0 + m[1] * r1 + m[3] * r2);
  z[p * 3 + 2] += w * (m[4] * r0 + m[3] * r1 + m[2] * r2);
}

// dense (6n)^2 matrix of the coarsest stencil (symmetric by its storage; amg3d.py:418-432, 544-546)
template <typename T>
__global__ __launch_bounds__(256) void amg_dense_kernel(int gx, int gy, int gz, const T* __restrict__ A,
                                                        T* __restrict__ D) {
  const int n = gx * gy * gz, m = 6 * n;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)m * m) return;
  const int i = (int)(t / m), j = (int)(t % m);
  const int I = i / 6, r = i % 6, J = j / 6, c = j % 6;
  const int g[3] = {gx, gy, gz};
  int ci[3], cj[3];
  node_xyz(I, g, ci);
  node_xyz(J, g, cj);
  const int dx = cj[0] - ci[0], dy = cj[1] - ci[1], dz = cj[2] - ci[2];
  T v = T(0);
  if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) {
    const int d = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
    v = d >= 13 ? A[((int64_t)(d - 13) * 36 + r * 6 + c) * n + I] : A[((int64_t)(13 - d) * 36 + c * 6 + r) * n + J];
  }
  D[t] = v;
}

// in-place inverse of an SPD matrix (m x m) by Gauss-Jordan elimination without pivoting, one workgroup; a
// non-positive pivot sets *bad and stops (the coarsest level's Cholesky of amg3d.py:544-546 in its inverse form)
template <typename T>
__global__ __launch_bounds__(1024) void amg_spd_inverse_kernel(int m, T* __restrict__ D, int* bad) {
  __shared__ double piv;
  for (int k = 0; k < m; ++k) {
    __syncthreads();
    if (threadIdx.x == 0) piv = (double)D[(int64_t)k * m + k];
    __syncthreads();
    if (!(piv > 0.0)) {
      if (threadIdx.x == 0) *bad = 1;
      return;
    }
    const double ip = 1.0 / piv;
    // row k scaled; D[k][k] becomes 1 / pivot
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
      const int64_t o = (int64_t)k * m + j;
      D[o] = j == k ? (T)ip : (T)((double)D[o] * ip);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      if (i == k) continue;
      const double f = (double)D[(int64_t)i * m + k];
      for (int j = 0; j < m; ++j) {
        const int64_t o = (int64_t)i * m + j;
        if (j == k)
          D[o] = (T)(-f * ip);
        else
          D[o] = (T)((double)D[o] - f * (double)D[(int64_t)k * m + j]);
      }
    }
  }
}

// x = Dinv b for the symmetric dense inverse (m x m, column i taken as row i): one wave per entry of x, the lanes
// over the column (the matrix is small and stays in L2), a fixed order of summation
template <typename T>
__global__ __launch_bounds__(256) void amg_dense_mv_kernel(int m, const T* __restrict__ Dinv, const T* __restrict__ b,
                                                           T* __restrict__ xo) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= m) return;
  T s = T(0);
  for (int j = lane; j < m; j += 64) s += Dinv[(int64_t)j * m + i] * b[j];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) xo[i] = s;
}

// ---------------------------------------------------------------------------------------------- vectors
// per-block partial sums of x.y (fixed order: grid-stride then a block reduction), partial[blockIdx.x]
template <typename T>
__global__ __launch_bounds__(256) void amg_dot_kernel(int64_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                      double* __restrict__ partial) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s += (double)x[i] * (double)y[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// ------------------------------------------------------------------------------------------ AMG-PCG on the device
// Scalars of one AMG-PCG solve kept in device memory (amg3d.py:649-672): the host only polls `active`.
struct AmgScal {
  double bn, rz, pap, alpha, beta, rn;
  int it, active, conv, notspd, nonfinite, pad_;
};
enum AmgFin : int { AF_BN = 0, AF_RZ0, AF_PAP, AF_RN, AF_RZ };

// W consecutive values of T and of the mask bytes (W = 16 / sizeof(T) for 16-byte loads, 1 for unaligned buffers)
template <typename T, int W>
struct VecT {
  T v[W];
};
template <int W>
struct VecM {
  uint8_t v[W];
};

template <typename T>
__device__ inline double block_sum4(double t, double* red) {
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// per-block partial sums of x.y, U vectors of W values in flight per thread; the tail past the last whole vector by
// block 0
template <typename T, int W>
__global__ __launch_bounds__(256) void amg_dotv_kernel(int64_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                       double* __restrict__ partial) {
  __shared__ double red[4];
  constexpr int U = 4;
  using V = VecT<T, W>;
  const int64_t nv = n / W;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s[U] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i0 < nv; i0 += stride * U) {
    V a[U], c[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * stride < nv) {
        a[u] = ((const V*)x)[i0 + u * stride];
        c[u] = ((const V*)y)[i0 + u * stride];
      }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * stride < nv)
#pragma unroll
        for (int j = 0; j < W; ++j) s[u] += (double)a[u].v[j] * (double)c[u].v[j];
  }
  double t = (s[0] + s[1]) + (s[2] + s[3]);
  if (blockIdx.x == 0 && threadIdx.x < n - nv * W) t += (double)x[nv * W + threadIdx.x] * (double)y[nv * W + threadIdx.x];
  t = block_sum4<T>(t, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// the scalar step of the PCG from the fixed-order sum of the partials (one block of 256)
__global__ __launch_bounds__(256) void amg_fin_kernel(const double* __restrict__ partial, int nblk,
                                                      AmgScal* __restrict__ sc, int op, double tol, int maxit) {
  __shared__ double red[4];
  double t = 0.0;
  for (int j = threadIdx.x; j < nblk; j += blockDim.x) t += partial[j];
  const double s = block_sum4<double>(t, red);
  if (threadIdx.x != 0) return;
  if (op == AF_BN) {
    sc->bn = sqrt(s);
    sc->rn = sc->bn;
    sc->it = 0;
    sc->conv = sc->notspd = sc->nonfinite = 0;
    sc->active = sc->bn > 0.0 ? 1 : 0;
    if (!(sc->bn <= 1e300)) {  // non-finite right-hand side
      sc->nonfinite = 1;
      sc->active = 0;
    }
    return;
  }
  if (!sc->active) return;
  if (op == AF_RZ0) {
    sc->rz = s;
  } else if (op == AF_PAP) {
    sc->pap = s;
    if (!(s > 0.0)) {  // K not positive definite on the search direction
      sc->notspd = 1;
      sc->active = 0;
    } else {
      sc->alpha = sc->rz / s;
    }
  } else if (op == AF_RN) {
    sc->it += 1;
    sc->rn = sqrt(s);
    if (!(sc->rn <= 1e300)) {
      sc->nonfinite = 1;
      sc->active = 0;
    } else if (sc->rn <= tol * sc->bn) {
      sc->conv = 1;
      sc->active = 0;
    } else if (sc->it >= maxit) {
      sc->active = 0;
    }
  } else if (op == AF_RZ) {
    sc->beta = s / sc->rz;
    sc->rz = s;
  }
}

// x += alpha p; r -= alpha mask Ap; partial sums of r.r (amg3d.py:399-408 with the free mask of apply0); alpha and
// the active flag from the device scalars
template <typename T, int W>
__global__ __launch_bounds__(256) void amg_cg_stepv_kernel(int64_t n, const AmgScal* __restrict__ sc,
                                                           const T* __restrict__ p, const T* __restrict__ ap,
                                                           const uint8_t* __restrict__ mask, T* __restrict__ x,
                                                           T* __restrict__ r, double* __restrict__ partial) {
  __shared__ double red[4];
  constexpr int U = 2;
  using V = VecT<T, W>;
  using M = VecM<W>;
  const bool active = sc->active != 0;
  const T alpha = (T)sc->alpha;
  const int64_t nv = n / W;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s[U] = {0.0, 0.0};
  if (active) {
    for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i0 < nv; i0 += stride * U) {
      V pv[U], av[U], xv[U], rv[U];
      M mv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = i0 + u * stride;
        if (i < nv) {
          pv[u] = ((const V*)p)[i];
          av[u] = ((const V*)ap)[i];
          xv[u] = ((const V*)x)[i];
          rv[u] = ((const V*)r)[i];
          mv[u] = ((const M*)mask)[i];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t i = i0 + u * stride;
        if (i < nv) {
#pragma unroll
          for (int j = 0; j < W; ++j) {
            xv[u].v[j] += alpha * pv[u].v[j];
            rv[u].v[j] -= mv[u].v[j] ? alpha * av[u].v[j] : T(0);
            s[u] += (double)rv[u].v[j] * (double)rv[u].v[j];
          }
          ((V*)x)[i] = xv[u];
          ((V*)r)[i] = rv[u];
        }
      }
    }
    if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {
      const int64_t i = nv * W + threadIdx.x;
      x[i] += alpha * p[i];
      const T ri = r[i] - (mask[i] ? alpha * ap[i] : T(0));
      r[i] = ri;
      s[0] += (double)ri * (double)ri;
    }
  }
  const double t = block_sum4<T>(s[0] + s[1], red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// p = z + beta p while the solve is active (amg3d.py:411-415)
template <typename T, int W>
__global__ __launch_bounds__(256) void amg_xpbyv_kernel(int64_t n, const T* __restrict__ z,
                                                        const AmgScal* __restrict__ sc, T* __restrict__ p) {
  if (!sc->active) return;
  using V = VecT<T, W>;
  const T beta = (T)sc->beta;
  const int64_t nv = n / W;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
    const V zv = ((const V*)z)[i];
    V pv = ((const V*)p)[i];
#pragma unroll
    for (int j = 0; j < W; ++j) pv.v[j] = zv.v[j] + beta * pv.v[j];
    ((V*)p)[i] = pv;
  }
  if (blockIdx.x == 0 && threadIdx.x < n - nv * W) {
    const int64_t i = nv * W + threadIdx.x;
    p[i] = z[i] + beta * p[i];
  }
}

// out = a - b (elementwise)
template <typename T>
__global__ __launch_bounds__(256) void amg_sub_kernel(int64_t n, const T* __restrict__ a, const T* __restrict__ b,
                                                      T* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}

// y += x (elementwise)
template <typename T>
__global__ __launch_bounds__(256) void amg_add_kernel(int64_t n, const T* __restrict__ x, T* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += x[i];
}

// refinement: r32 = (float) r64, and x64 += (double) d32
__global__ __launch_bounds__(256) void amg_to32_kernel(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}
__global__ __launch_bounds__(256) void amg_add64_kernel(int64_t n, const float* __restrict__ d, double* __restrict__ x) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] += (double)d[i];
}

template <typename T>
__global__ __launch_bounds__(256) void amg_fill_kernel(int64_t n, T v, T* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = v;
}

}  // namespace lpm
