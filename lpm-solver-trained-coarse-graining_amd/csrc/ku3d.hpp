// Matrix-free K.u of the 3D simple-cubic LPM (18 slots, two shells), fused in one pass.
//
// Operator (coarse3d/lin3d.py:98-149, coarse3d/torch_lpm3d.py:63-83):
//   dL_ik = n_k.(u_i - u_q),  S_i^c = sum_{k in c} dL_ik,  TS_i^c = sum_{k in c} Tv_ik dL_ik,
//   F_ik = 2 Kn_ik dL_ik + 1/2 (TS_i^c + TS_q^c) + 1/2 Tv_ik (S_i^c + S_q^c),   (K u)_i = sum_k F_ik n_k.
// F needs the shell sums of the neighbours, so the operator reaches two hops.  A block owns a TX x TY column tile
// and marches along z through a chunk of planes.  Planes of u (tile + 2 halo) sit in an LDS ring of 4 planes and
// the shell sums (tile + 1 halo) in a ring of 3 planes; every u value and coefficient is read from HBM once per
// block (halo re-reads hit L2), and the shell sums never leave LDS.
#pragma once
#include "lpm_common.hpp"

namespace lpm {

enum Ku3Epi : int {
  EPI_PLAIN = 0,    // out = K u
  EPI_MASKDOT = 1,  // out = mask * (K u), partial dot(u, out) per block
  EPI_RESID = 2,    // out = mask * (rhs - K u), partial dot(out, out) per block
};

template <typename T, typename C>
struct Ku3Args {
  int nx, ny, nz;
  int nl;        // systems per outer batch index (grid.y)
  int zc;        // planes per block
  int ntx, nty, nzc;
  const T* u;
  Bcast bu;
  const C* kn;
  const C* tv;
  Bcast bc;
  const uint8_t* mask;  // per DOF (N, 3), 1 = free; may be null (all free)
  Bcast bm;
  const T* rhs;
  Bcast brhs;
  T* out;
  Bcast bo;
  T* partial;         // (systems, blocks per system)
  const int* active;  // per system flag at active[s * act_stride]; null = all active
  int act_stride;
};

// blockIdx.x -> logical tile id, contiguous per XCD (blocks are dispatched round-robin over the 8 XCDs).
__device__ __forceinline__ int xcd_remap(int bid, int nblk) {
  const int q = nblk >> 3, r = nblk & 7;
  const int x = bid & 7, k = bid >> 3;
  return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

template <typename T, typename C, int MODE, int TX, int TY, int EPI>
__global__ __launch_bounds__(TX* TY) void ku3d_kernel(Ku3Args<T, C> a) {
  constexpr int HX = TX + 4, HY = TY + 4;  // u tile with a 2-site halo
  constexpr int SX = TX + 2, SY = TY + 2;  // shell sums with a 1-site halo
  constexpr int NT = TX * TY;
  __shared__ T su[4][3][HY][HX];
  __shared__ T ss[3][4][SY][SX];  // S0, S1, TS0, TS1
  __shared__ T red[NT / 64];

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t N = (int64_t)nx * ny * nz;
  const int nblk = gridDim.x;
  const int lid = xcd_remap(blockIdx.x, nblk);
  const int tix = lid % a.ntx;
  const int tiy = (lid / a.ntx) % a.nty;
  const int tiz = lid / (a.ntx * a.nty);
  const int x0 = tix * TX, y0 = tiy * TY, zb = tiz * a.zc;
  const int ze = min(zb + a.zc, nz);
  const int l = blockIdx.y, b = blockIdx.z;
  if (a.active && !a.active[((int64_t)b * a.nl + l) * a.act_stride]) return;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;

  const T* __restrict__ u = a.u + a.bu.at(b, l);
  const C* __restrict__ kn = a.kn + a.bc.at(b, l);
  const C* __restrict__ tv = a.tv + a.bc.at(b, l);

  auto load_plane = [&](int z) {
    T(*dst)[HY][HX] = su[(z + 8) & 3];
    const bool zin = (z >= 0) && (z < nz);
    for (int idx = tid; idx < HX * HY; idx += NT) {
      const int hx = idx % HX, hy = idx / HX;
      const int gx = x0 - 2 + hx, gy = y0 - 2 + hy;
      T v0 = T(0), v1 = T(0), v2 = T(0);
      if (zin && gx >= 0 && gx < nx && gy >= 0 && gy < ny) {
        const int64_t p = ((int64_t)z * ny + gy) * nx + gx;
        v0 = u[3 * p + 0];
        v1 = u[3 * p + 1];
        v2 = u[3 * p + 2];
      }
      dst[0][hy][hx] = v0;
      dst[1][hy][hx] = v1;
      dst[2][hy][hx] = v2;
    }
  };

  // shell sums of plane z over the tile + 1 halo
  auto shell_plane = [&](int z) {
    T(*dst)[SY][SX] = ss[((z % 3) + 3) % 3];
    const bool zin = (z >= 0) && (z < nz);
    for (int idx = tid; idx < SX * SY; idx += NT) {
      const int sx = idx % SX, sy = idx / SX;
      const int gx = x0 - 1 + sx, gy = y0 - 1 + sy;
      T s0 = T(0), s1 = T(0), t0 = T(0), t1 = T(0);
      if (zin && gx >= 0 && gx < nx && gy >= 0 && gy < ny) {
        const int64_t p = ((int64_t)z * ny + gy) * nx + gx;
        const int hx = sx + 1, hy = sy + 1;
        const T(*up)[HY][HX] = su[(z + 8) & 3];
        const T ux = up[0][hy][hx], uy = up[1][hy][hx], uz = up[2][hy][hx];
#pragma unroll
        for (int k = 0; k < 18; ++k) {
          const int ox = off3(k, 0), oy = off3(k, 1), oz = off3(k, 2);
          const int qx = gx + ox, qy = gy + oy, qz = z + oz;
          if (qx < 0 || qx >= nx || qy < 0 || qy >= ny || qz < 0 || qz >= nz) continue;
          const T(*uq)[HY][HX] = su[(qz + 8) & 3];
          T dl = T(0);
          if (ox) dl += nd3<T>(k, 0) * (ux - uq[0][hy + oy][hx + ox]);
          if (oy) dl += nd3<T>(k, 1) * (uy - uq[1][hy + oy][hx + ox]);
          if (oz) dl += nd3<T>(k, 2) * (uz - uq[2][hy + oy][hx + ox]);
          const int64_t q = ((int64_t)qz * ny + qy) * nx + qx;
          const T t = (T)coef_at<C, MODE, 18>(tv, N, p, q, k);
          if (shell3(k) == 0) {
            s0 += dl;
            t0 += t * dl;
          } else {
            s1 += dl;
            t1 += t * dl;
          }
        }
      }
      dst[0][sy][sx] = s0;
      dst[1][sy][sx] = s1;
      dst[2][sy][sx] = t0;
      dst[3][sy][sx] = t1;
    }
  };

  // prologue: u planes zb-2 .. zb+1, shell sums of planes zb-1 and zb
  load_plane(zb - 2);
  load_plane(zb - 1);
  load_plane(zb);
  __syncthreads();
  shell_plane(zb - 1);
  load_plane(zb + 1);
  __syncthreads();
  shell_plane(zb);

  T acc = T(0);
  for (int z = zb; z < ze; ++z) {
    __syncthreads();
    load_plane(z + 2);
    __syncthreads();
    shell_plane(z + 1);
    __syncthreads();
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx < nx && gy < ny) {
      const int64_t p = ((int64_t)z * ny + gy) * nx + gx;
      const int hx = tx + 2, hy = ty + 2, sx = tx + 1, sy = ty + 1;
      const T(*up)[HY][HX] = su[(z + 8) & 3];
      const T(*sp)[SY][SX] = ss[z % 3];
      const T ux = up[0][hy][hx], uy = up[1][hy][hx], uz = up[2][hy][hx];
      const T si0 = sp[0][sy][sx], si1 = sp[1][sy][sx], ti0 = sp[2][sy][sx], ti1 = sp[3][sy][sx];
      T ox_ = T(0), oy_ = T(0), oz_ = T(0);
#pragma unroll
      for (int k = 0; k < 18; ++k) {
        const int ox = off3(k, 0), oy = off3(k, 1), oz = off3(k, 2);
        const int qx = gx + ox, qy = gy + oy, qz = z + oz;
        if (qx < 0 || qx >= nx || qy < 0 || qy >= ny || qz < 0 || qz >= nz) continue;
        const T(*uq)[HY][HX] = su[(qz + 8) & 3];
        const T(*sq)[SY][SX] = ss[((qz % 3) + 3) % 3];
        T dl = T(0);
        if (ox) dl += nd3<T>(k, 0) * (ux - uq[0][hy + oy][hx + ox]);
        if (oy) dl += nd3<T>(k, 1) * (uy - uq[1][hy + oy][hx + ox]);
        if (oz) dl += nd3<T>(k, 2) * (uz - uq[2][hy + oy][hx + ox]);
        const int64_t q = ((int64_t)qz * ny + qy) * nx + qx;
        const T kk = (T)coef_at<C, MODE, 18>(kn, N, p, q, k);
        const T tt = (T)coef_at<C, MODE, 18>(tv, N, p, q, k);
        T f;
        if (shell3(k) == 0)
          f = T(2) * kk * dl + T(0.5) * (ti0 + sq[2][sy + oy][sx + ox]) + T(0.5) * tt * (si0 + sq[0][sy + oy][sx + ox]);
        else
          f = T(2) * kk * dl + T(0.5) * (ti1 + sq[3][sy + oy][sx + ox]) + T(0.5) * tt * (si1 + sq[1][sy + oy][sx + ox]);
        if (ox) ox_ += f * nd3<T>(k, 0);
        if (oy) oy_ += f * nd3<T>(k, 1);
        if (oz) oz_ += f * nd3<T>(k, 2);
      }
      T* __restrict__ o = a.out + a.bo.at(b, l) + 3 * p;
      if constexpr (EPI == EPI_PLAIN) {
        o[0] = ox_;
        o[1] = oy_;
        o[2] = oz_;
      } else {
        T m0 = T(1), m1 = T(1), m2 = T(1);
        if (a.mask) {
          const uint8_t* mk = a.mask + a.bm.at(b, l) + 3 * p;
          m0 = mk[0] ? T(1) : T(0);
          m1 = mk[1] ? T(1) : T(0);
          m2 = mk[2] ? T(1) : T(0);
        }
        if constexpr (EPI == EPI_MASKDOT) {
          ox_ *= m0;
          oy_ *= m1;
          oz_ *= m2;
          acc += ux * ox_ + uy * oy_ + uz * oz_;
        } else {
          const T* rh = a.rhs + a.brhs.at(b, l) + 3 * p;
          ox_ = m0 * (rh[0] - ox_);
          oy_ = m1 * (rh[1] - oy_);
          oz_ = m2 * (rh[2] - oz_);
          acc += ox_ * ox_ + oy_ * oy_ + oz_ * oz_;
        }
        o[0] = ox_;
        o[1] = oy_;
        o[2] = oz_;
      }
    }
  }
  if constexpr (EPI != EPI_PLAIN) {
    const T s = block_sum(acc, red);
    if (tid == 0) a.partial[((int64_t)b * a.nl + l) * nblk + lid] = s;
  }
}

}  // namespace lpm
