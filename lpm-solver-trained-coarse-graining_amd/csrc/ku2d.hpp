// Matrix-free K.u of the 2D square-lattice LPM (8 slots, two shells), fused in one pass.
//
// Same law as the 3D operator (coarse2d/lin2d.py:121-141, timing/t2d/solver2d.py:269-305): the shell sums of the
// tile + 1 halo are formed in LDS from a tile of u with a 2-site halo, then every owned particle sums its 8 bond
// forces.  The tile routine is a device function so that the persistent batched solver reuses it.
#pragma once
#include "ku3d.hpp"
#include "lpm_common.hpp"

namespace lpm {

template <typename T, int TX, int TY>
struct Ku2Smem {
  T u[2][TY + 4][TX + 4];
  T s[4][TY + 2][TX + 2];
};

// K u on the tile [x0, x0+TX) x [y0, y0+TY) of one system; the result for the thread's own site is returned in
// (fx, fy) together with its u; `own` tells whether the thread owns a site of the lattice.  BONDS: bond words of
// removed particles (bit k of particle p = slot k is a bond) restrict the bonds further.
template <typename T, typename C, int MODE, int TX, int TY, bool BONDS = false>
__device__ __forceinline__ void ku2d_tile(const T* __restrict__ u, const C* __restrict__ kn, const C* __restrict__ tv,
                                          int nx, int ny, int64_t N, int x0, int y0, Ku2Smem<T, TX, TY>& sm,
                                          T& fx, T& fy, T& ux, T& uy, bool& own, int64_t& pown,
                                          const uint32_t* __restrict__ bwd = nullptr) {
  // N: plane stride of packed coefficients (unused for directed ones)
  constexpr int HX = TX + 4, HY = TY + 4, SX = TX + 2, SY = TY + 2, NT = TX * TY;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < HX * HY; idx += NT) {
    const int hx = idx % HX, hy = idx / HX;
    const int gx = x0 - 2 + hx, gy = y0 - 2 + hy;
    T v0 = T(0), v1 = T(0);
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) {
      const int64_t p = (int64_t)gy * nx + gx;
      v0 = u[2 * p];
      v1 = u[2 * p + 1];
    }
    sm.u[0][hy][hx] = v0;
    sm.u[1][hy][hx] = v1;
  }
  __syncthreads();
  for (int idx = tid; idx < SX * SY; idx += NT) {
    const int sx = idx % SX, sy = idx / SX;
    const int gx = x0 - 1 + sx, gy = y0 - 1 + sy;
    T s0 = T(0), s1 = T(0), t0 = T(0), t1 = T(0);
    if (gx >= 0 && gx < nx && gy >= 0 && gy < ny) {
      const int64_t p = (int64_t)gy * nx + gx;
      const int hx = sx + 1, hy = sy + 1;
      const T vx = sm.u[0][hy][hx], vy = sm.u[1][hy][hx];
      const uint32_t w = BONDS ? bwd[p] : 0xffu;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int ox = off2(k, 0), oy = off2(k, 1);
        const int qx = gx + ox, qy = gy + oy;
        if (qx < 0 || qx >= nx || qy < 0 || qy >= ny || !((w >> k) & 1u)) continue;
        T dl = T(0);
        if (ox) dl += nd2<T>(k, 0) * (vx - sm.u[0][hy + oy][hx + ox]);
        if (oy) dl += nd2<T>(k, 1) * (vy - sm.u[1][hy + oy][hx + ox]);
        const int64_t q = (int64_t)qy * nx + qx;
        const T t = (T)coef_at<C, MODE, 8>(tv, N, p, q, k);
        if (shell2(k) == 0) {
          s0 += dl;
          t0 += t * dl;
        } else {
          s1 += dl;
          t1 += t * dl;
        }
      }
    }
    sm.s[0][sy][sx] = s0;
    sm.s[1][sy][sx] = s1;
    sm.s[2][sy][sx] = t0;
    sm.s[3][sy][sx] = t1;
  }
  __syncthreads();
  const int tx = tid % TX, ty = tid / TX;
  const int gx = x0 + tx, gy = y0 + ty;
  own = gx < nx && gy < ny;
  fx = fy = T(0);
  ux = uy = T(0);
  pown = 0;
  if (own) {
    const int64_t p = (int64_t)gy * nx + gx;
    pown = p;
    const int hx = tx + 2, hy = ty + 2, sx = tx + 1, sy = ty + 1;
    ux = sm.u[0][hy][hx];
    uy = sm.u[1][hy][hx];
    const T si0 = sm.s[0][sy][sx], si1 = sm.s[1][sy][sx], ti0 = sm.s[2][sy][sx], ti1 = sm.s[3][sy][sx];
    const uint32_t w = BONDS ? bwd[p] : 0xffu;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int ox = off2(k, 0), oy = off2(k, 1);
      const int qx = gx + ox, qy = gy + oy;
      if (qx < 0 || qx >= nx || qy < 0 || qy >= ny || !((w >> k) & 1u)) continue;
      T dl = T(0);
      if (ox) dl += nd2<T>(k, 0) * (ux - sm.u[0][hy + oy][hx + ox]);
      if (oy) dl += nd2<T>(k, 1) * (uy - sm.u[1][hy + oy][hx + ox]);
      const int64_t q = (int64_t)qy * nx + qx;
      const T kk = (T)coef_at<C, MODE, 8>(kn, N, p, q, k);
      const T tt = (T)coef_at<C, MODE, 8>(tv, N, p, q, k);
      T f;
      if (shell2(k) == 0)
        f = T(2) * kk * dl + T(0.5) * (ti0 + sm.s[2][sy + oy][sx + ox]) + T(0.5) * tt * (si0 + sm.s[0][sy + oy][sx + ox]);
      else
        f = T(2) * kk * dl + T(0.5) * (ti1 + sm.s[3][sy + oy][sx + ox]) + T(0.5) * tt * (si1 + sm.s[1][sy + oy][sx + ox]);
      if (ox) fx += f * nd2<T>(k, 0);
      if (oy) fy += f * nd2<T>(k, 1);
    }
  }
}

template <typename T, typename C>
struct Ku2Args {
  int nx, ny;
  int64_t ps;  // plane stride of packed coefficients
  int nl;
  int ntx, nty;
  const T* u;
  Bcast bu;
  const C* kn;
  const C* tv;
  Bcast bc;
  const uint8_t* mask;  // per DOF (N, 2)
  Bcast bm;
  const T* rhs;
  Bcast brhs;
  T* out;
  Bcast bo;
  T* partial;
  const int* active;  // per system flag at active[s * act_stride]; null = all active
  int act_stride;
  const uint32_t* bonds;  // bond words (BONDS kernels)
  Bcast bb;
};

template <typename T, typename C, int MODE, int TX, int TY, int EPI, bool BONDS = false>
__global__ __launch_bounds__(TX* TY) void ku2d_kernel(Ku2Args<T, C> a) {
  __shared__ Ku2Smem<T, TX, TY> sm;
  __shared__ T red[TX * TY / 64];
  const int l = blockIdx.y, b = blockIdx.z;
  const int64_t sys = (int64_t)b * a.nl + l;
  if (a.active && !a.active[sys * a.act_stride]) return;
  const int nblk = gridDim.x;
  const int lid = xcd_remap(blockIdx.x, nblk);
  const int x0 = (lid % a.ntx) * TX, y0 = (lid / a.ntx) * TY;
  T fx, fy, ux, uy;
  bool own;
  int64_t p;
  ku2d_tile<T, C, MODE, TX, TY, BONDS>(a.u + a.bu.at(b, l), a.kn + a.bc.at(b, l), a.tv + a.bc.at(b, l), a.nx, a.ny,
                                       a.ps, x0, y0, sm, fx, fy, ux, uy, own, p,
                                       BONDS ? a.bonds + a.bb.at(b, l) : nullptr);
  T acc = T(0);
  if (own) {
    T* o = a.out + a.bo.at(b, l) + 2 * p;
    if constexpr (EPI == EPI_PLAIN) {
      o[0] = fx;
      o[1] = fy;
    } else {
      if constexpr (EPI == EPI_MASKDOT) {
        acc = ux * fx + uy * fy;  // u vanishes on fixed DOFs: cg_update masks the product
      } else {
        T m0 = T(1), m1 = T(1);
        if (a.mask) {
          const uint8_t* mk = a.mask + a.bm.at(b, l) + 2 * p;
          m0 = mk[0] ? T(1) : T(0);
          m1 = mk[1] ? T(1) : T(0);
        }
        const T* rh = a.rhs + a.brhs.at(b, l) + 2 * p;
        fx = m0 * (rh[0] - fx);
        fy = m1 * (rh[1] - fy);
        acc = fx * fx + fy * fy;
      }
      o[0] = fx;
      o[1] = fy;
    }
  }
  if constexpr (EPI != EPI_PLAIN) {
    const T s = block_sum(acc, red);
    if (threadIdx.x == 0) a.partial[sys * nblk + lid] = s;
  }
}

}  // namespace lpm
