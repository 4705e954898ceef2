// Shared definitions of the LPM solve-path kernels (gfx950, wave64).
//
// Lattices are full boxes.  Particle id p = (z * ny + y) * nx + x (3D) or y * nx + x (2D), x fastest, the same
// numbering as the reference lattices (coarse3d/geom3d.py:61, coarse2d/lin2d.py:35).  Slots come in reverse pairs
// (k, k^1); 3D shell 0 = slots 0-5 (axial), shell 1 = slots 6-17 (face diagonals); 2D shell 0 = slots 0-3,
// shell 1 = slots 4-7 (coarse3d/geom3d.py:16-20, coarse2d/lin2d.py:21-22).  The bond unit vector of slot k is
// n_k = -off_k / |off_k|.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lpm {

// ------------------------------------------------------------------------------------------ slot geometry
// 3D simple cubic, 18 slots
__host__ __device__ constexpr int off3(int k, int a) {
  constexpr int T[18][3] = {{1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                            {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1},
                            {1, 0, -1}, {-1, 0, 1}, {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
  return T[k][a];
}
__host__ __device__ constexpr int shell3(int k) { return k < 6 ? 0 : 1; }
// 2D square, 8 slots
__host__ __device__ constexpr int off2(int k, int a) {
  constexpr int T[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, -1}, {1, -1}, {-1, 1}};
  return T[k][a];
}
__host__ __device__ constexpr int shell2(int k) { return k < 4 ? 0 : 1; }

constexpr double kInvSqrt2 = 0.70710678118654752440;

// component a of n_k (3D): -off / |off|
template <typename T>
__host__ __device__ constexpr T nd3(int k, int a) {
  return off3(k, a) == 0 ? T(0) : (T)((k < 6 ? -1.0 : -kInvSqrt2) * (double)off3(k, a));
}
template <typename T>
__host__ __device__ constexpr T nd2(int k, int a) {
  return off2(k, a) == 0 ? T(0) : (T)((k < 4 ? -1.0 : -kInvSqrt2) * (double)off2(k, a));
}

// ------------------------------------------------------------------------------------------ coefficient layouts
// DIRECTED: (..., N, NN) per directed slot, as the reference tensors Kn/Tv (torch_lpm3d.py:212-227).
// PACKED:   (..., NN/2, P) forward-slot planes (slots 0, 2, 4, ...), valid for symmetric parameters
//           (Kn[i,k] == Kn[q,k^1]): slot k odd at particle i reads forward plane (k-1)/2 at its neighbour q.
//           Each plane holds G zero guard values, the N particles, and at least G more guards: G is one z-plane +
//           one row + one site (2D: one row + one site) rounded up to 64 values and P = N + 2G rounded up to 64,
//           so p + (neighbour offset) is inside the plane for every particle p and slot, bond or not (the kernels
//           read it without a bounds select), and each plane's first particle is 256-byte aligned in FP32.
enum CoefMode : int { COEF_DIRECTED = 0, COEF_PACKED = 1 };

__host__ __device__ inline int64_t packed_guard(int dim, int nx, int ny) {
  const int64_t g = dim == 3 ? (int64_t)nx * ny + nx + 1 : (int64_t)nx + 1;
  return (g + 63) & ~(int64_t)63;
}
// plane stride P of the packed layout for N particles
__host__ __device__ inline int64_t packed_stride(int dim, int nx, int ny, int64_t N) {
  return (N + 2 * packed_guard(dim, nx, ny) + 63) & ~(int64_t)63;
}

// base: first particle of plane 0 (array + G) in PACKED mode; pstride: plane stride P
template <typename C, int MODE, int NN>
__device__ __forceinline__ C coef_at(const C* __restrict__ base, int64_t pstride, int64_t p, int64_t q, int k) {
  if constexpr (MODE == COEF_DIRECTED) {
    return base[p * NN + k];
  } else {
    return (k & 1) ? base[(int64_t)(k >> 1) * pstride + q] : base[(int64_t)(k >> 1) * pstride + p];
  }
}

// ------------------------------------------------------------------------------------------ batching
// A batch of systems is (nb, nl); every array has its own element strides for b and l (0 = broadcast).
struct Bcast {
  int64_t sb, sl;
  __host__ __device__ int64_t at(int b, int l) const { return (int64_t)b * sb + (int64_t)l * sl; }
};

// ------------------------------------------------------------------------------------------ reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide sum with a fixed order (deterministic for a fixed blockDim).  `red` must hold blockDim/64 values.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = T(0);
  if (w == 0) {
    s = lane < nw ? red[lane] : T(0);
    s = wave_sum(s);
  }
  return s;  // valid in wave 0
}

}  // namespace lpm
