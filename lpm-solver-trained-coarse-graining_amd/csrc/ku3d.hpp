// Matrix-free K.u of the 3D simple-cubic LPM (18 slots, two shells), fused in one pass.
//
// Operator (coarse3d/lin3d.py:98-149, coarse3d/torch_lpm3d.py:63-83):
//   dL_ik = n_k.(u_i - u_q),  S_i^c = sum_{k in c} dL_ik,  TS_i^c = sum_{k in c} Tv_ik dL_ik,
//   F_ik = 2 Kn_ik dL_ik + 1/2 (TS_i^c + TS_q^c) + 1/2 Tv_ik (S_i^c + S_q^c),   (K u)_i = sum_k F_ik n_k.
// F needs the shell sums of the neighbours, so the operator reaches two hops.  A block owns a TX x TY column tile
// and marches along z through a chunk of planes.  Planes of u (tile + 2 halo) sit in an LDS ring of 5 planes and
// the shell sums (tile + 1 halo) in a ring of 3 planes, so the shell sums never leave LDS.
//
// Step z: (1) shell sums of plane z+1 -- every thread for its own column site, keeping that site's 18 bond
// elongations and shell sums in registers for the next step's output, and threads < NR for the ring of halo sites;
// (2) barrier; (3) output of plane z from the registers of the previous step and the neighbours' shell sums in
// LDS; (4) barrier.  Global loads of u and of the coefficients are issued a phase ahead into registers.
//
// Instruction economy (the kernel must stay under the HBM roofline's instruction budget):
//  * each bond elongation is computed once per site;
//  * face-diagonal bonds work with e = +-du_a +- du_b = sqrt2 dL; their shell sums are kept as sqrt2 S^1, sqrt2 TS^1,
//    and G = F / sqrt2 is accumulated, so no 1/sqrt2 multiplies appear;
//  * a missing bond (box face) is dropped by a select on its result; packed coefficient planes carry zero guards, so
//    the neighbour's plane entry is read at p + delta without a bounds select; no branches in the slot loops;
//  * loads use 32-bit byte offsets from a uniform base.
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
  int64_t ps;    // plane stride of packed coefficients (kn / tv point at the first particle of plane 0)
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

// Bond of slot K exists for a site whose face-neighbour flags are (mx, px, my, py, mz, pz).
template <int K>
__device__ __forceinline__ bool slot_ok(bool mx, bool px, bool my, bool py, bool mz, bool pz) {
  constexpr int ox = off3(K, 0), oy = off3(K, 1), oz = off3(K, 2);
  return (ox < 0 ? mx : (ox > 0 ? px : true)) && (oy < 0 ? my : (oy > 0 ? py : true)) &&
         (oz < 0 ? mz : (oz > 0 ? pz : true));
}

template <typename T, typename O>
__device__ __forceinline__ T ld_off(const T* __restrict__ base, O byte_off) {
  return *(const T*)((const char*)base + byte_off);
}

// dL for axial slots; e = sqrt2 dL for diagonal slots (u_q from the LDS plane)
template <typename T, int K, int HY, int HX>
__device__ __forceinline__ T dle3(const T (*uq)[HY][HX], int hy, int hx, T ux, T uy, T uz) {
  constexpr int ox = off3(K, 0), oy = off3(K, 1), oz = off3(K, 2);
  T d = T(0);
  // n_k = -off_k / |off_k|: the sign of each component is -sign(off)
  if constexpr (ox != 0) d += (ox > 0 ? T(-1) : T(1)) * (ux - uq[0][hy + oy][hx + ox]);
  if constexpr (oy != 0) d += (oy > 0 ? T(-1) : T(1)) * (uy - uq[1][hy + oy][hx + ox]);
  if constexpr (oz != 0) d += (oz > 0 ? T(-1) : T(1)) * (uz - uq[2][hy + oy][hx + ox]);
  return d;
}

template <int... Ks>
struct Seq {};
using Slots18 = Seq<0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17>;

template <typename F, int... Ks>
__device__ __forceinline__ void for_slots(Seq<Ks...>, F&& f) {
  (f(std::integral_constant<int, Ks>{}), ...);
}

// the 18 directed coefficients of particle p (in-box); directed: own row; packed: plane entry at p or p + delta
template <typename T, typename C, int MODE>
__device__ __forceinline__ void load18(const C* __restrict__ base, int64_t ps, int p, int nx, int nxy, T (&out)[18]) {
  for_slots(Slots18{}, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    if constexpr (MODE == COEF_DIRECTED) {
      out[K] = (T)ld_off(base, ((unsigned)p * 18u + K) * (unsigned)sizeof(C));  // < 4 GiB (host check)
    } else {
      // offsets from the plane's first guard (G = nxy + nx + 1 before particle 0) are never negative, so the load
      // takes a uniform base and an unsigned 32-bit offset
      const int G = nxy + nx + 1;
      const unsigned q = (unsigned)(p + G + ((K & 1) ? off3(K, 0) + off3(K, 1) * nx + off3(K, 2) * nxy : 0));
      out[K] = (T)ld_off(base + ((K >> 1) * ps - G), q * (unsigned)sizeof(C));
    }
  });
}

template <typename T, typename C, int MODE, int TX, int TY, int EPI>
__global__ __launch_bounds__(TX* TY, sizeof(T) == 4 ? 2 : 1) void ku3d_kernel(Ku3Args<T, C> a) {
  constexpr int HX = TX + 4, HY = TY + 4;  // u tile with a 2-site halo
  constexpr int SX = TX + 2, SY = TY + 2;  // shell sums with a 1-site halo
  constexpr int NT = TX * TY;
  constexpr int NR = 2 * SX + 2 * TY;           // ring of halo sites of the shell sums
  constexpr int NLD = (HX * HY + NT - 1) / NT;  // u sites loaded per thread per plane
  static_assert(NR <= NT, "one ring site per thread at most");
  __shared__ T su[5][3][HY][HX];
  __shared__ T ss[3][4][SY][SX];  // S0, sqrt2 S1, TS0, sqrt2 TS1
  __shared__ T red[NT / 64];

  const int nx = a.nx, ny = a.ny, nz = a.nz, nxy = nx * ny;
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
  const int64_t ps = a.ps;

  const T* __restrict__ u = a.u + a.bu.at(b, l);
  const C* __restrict__ kn = a.kn + a.bc.at(b, l);
  const C* __restrict__ tv = a.tv + a.bc.at(b, l);

  // Site coordinates are clamped into the box for addressing.  Results of halo sites outside the box are never
  // read through an existing bond, so they need no masking.
  auto cl = [](int v, int n) { return min(max(v, 0), n - 1); };

  // own column site
  const int gxo = x0 + tx, gyo = y0 + ty;
  const bool own_xy = gxo < nx && gyo < ny;
  const int cgx = cl(gxo, nx), cgy = cl(gyo, ny);
  const bool omx = cgx > 0, opx = cgx < nx - 1, omy = cgy > 0, opy = cgy < ny - 1;
  const int own_xy_off = cgy * nx + cgx;
  // ring site (threads < NR): shell-sum coordinates (rsx, rsy)
  const bool has_r = tid < NR;
  int rsx, rsy;
  {
    const int r = min(tid, NR - 1);
    if (r < SX) {
      rsx = r;
      rsy = 0;
    } else if (r < 2 * SX) {
      rsx = r - SX;
      rsy = SY - 1;
    } else {
      const int rr = r - 2 * SX;
      rsx = (rr & 1) ? SX - 1 : 0;
      rsy = 1 + (rr >> 1);
    }
  }
  const int rgx = cl(x0 - 1 + rsx, nx), rgy = cl(y0 - 1 + rsy, ny);
  const bool rmx = rgx > 0, rpx = rgx < nx - 1, rmy = rgy > 0, rpy = rgy < ny - 1;
  const int r_xy_off = rgy * nx + rgx;

  T ureg[NLD][3];
  T dC[18], dN[18];  // own elongations of the output plane and of the next shell plane
  T sC[4], sN[4];    // own shell sums, same two planes
  T tvC[18], knC[18], tvN[18], tvR[18];
  // FP32: the next step's coefficients are loaded a full step ahead; FP64 has no registers for a second set and
  // loads them after the output phase
  constexpr bool DEEP = sizeof(T) == 4;
  constexpr int ND = DEEP ? 18 : 1;
  T tvNN[ND], knN[ND], tvRN[ND];

  auto fetch_u = [&](int z) {
    const int zo = cl(z, nz) * nxy;
#pragma unroll
    for (int j = 0; j < NLD; ++j) {
      const int idx = min(tid + j * NT, HX * HY - 1);
      const int hx = idx % HX, hy = idx / HX;
      const int p = zo + cl(y0 - 2 + hy, ny) * nx + cl(x0 - 2 + hx, nx);
      const int o = p * (int)(3 * sizeof(T));
      ureg[j][0] = ld_off(u, o);
      ureg[j][1] = ld_off(u, o + (int)sizeof(T));
      ureg[j][2] = ld_off(u, o + 2 * (int)sizeof(T));
    }
  };
  auto store_u = [&](int z) {
    T(*dst)[HY][HX] = su[(z + 10) % 5];
#pragma unroll
    for (int j = 0; j < NLD; ++j) {
      const int idx = tid + j * NT;
      if (j + 1 < NLD || idx < HX * HY) {
        const int hx = idx % HX, hy = idx / HX;
        dst[0][hy][hx] = ureg[j][0];
        dst[1][hy][hx] = ureg[j][1];
        dst[2][hy][hx] = ureg[j][2];
      }
    }
  };

  // Shell sums of the site at shell coordinates (sx, sy) in plane z, in-plane flags, Tv of the site; optionally
  // keeps the 18 elongations.  Writes S0, sqrt2 S1, TS0, sqrt2 TS1 to the LDS ring and returns them in s.
  auto shell_site = [&](auto keep_tag, int z, int sx, int sy, bool mx, bool px, bool my, bool py,
                        const T (&tvv)[18], T (&keep)[18], T (&s)[4]) {
    constexpr bool KEEP = decltype(keep_tag)::value;
    const int zc_ = cl(z, nz);
    const bool mz = zc_ > 0, pz = zc_ < nz - 1;
    const T(*u0)[HY][HX] = su[(z + 9) % 5];   // plane z-1
    const T(*u1)[HY][HX] = su[(z + 10) % 5];  // plane z
    const T(*u2)[HY][HX] = su[(z + 11) % 5];  // plane z+1
    const int hx = sx + 1, hy = sy + 1;
    const T ux = u1[0][hy][hx], uy = u1[1][hy][hx], uz = u1[2][hy][hx];
    T s0 = T(0), s1 = T(0), t0 = T(0), t1 = T(0);
    for_slots(Slots18{}, [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      constexpr int oz = off3(K, 2);
      const T(*uq)[HY][HX] = oz < 0 ? u0 : (oz > 0 ? u2 : u1);
      T d = dle3<T, K, HY, HX>(uq, hy, hx, ux, uy, uz);
      d = slot_ok<K>(mx, px, my, py, mz, pz) ? d : T(0);
      if constexpr (KEEP) keep[K] = d;
      if constexpr (shell3(K) == 0) {
        s0 += d;
        t0 += tvv[K] * d;
      } else {
        s1 += d;
        t1 += tvv[K] * d;
      }
    });
    T(*dst)[SY][SX] = ss[(z + 9) % 3];
    dst[0][sy][sx] = s0;
    dst[1][sy][sx] = s1;
    dst[2][sy][sx] = t0;
    dst[3][sy][sx] = t1;
    s[0] = s0;
    s[1] = s1;
    s[2] = t0;
    s[3] = t1;
  };
  auto own_p = [&](int z) { return cl(z, nz) * nxy + own_xy_off; };
  auto ring_p = [&](int z) { return cl(z, nz) * nxy + r_xy_off; };

  // ---- prologue: u planes zb-2 .. zb+2 in LDS; shell sums of planes zb-1 and zb
  for (int z = zb - 2; z <= zb + 2; ++z) {
    fetch_u(z);
    store_u(z);
  }
  __syncthreads();
  using Keep = std::integral_constant<bool, true>;
  using Drop = std::integral_constant<bool, false>;
  {
    T tmp[4];
    load18<T, C, MODE>(tv, ps, own_p(zb - 1), nx, nxy, tvN);
    load18<T, C, MODE>(tv, ps, ring_p(zb - 1), nx, nxy, tvR);
    shell_site(Drop{}, zb - 1, tx + 1, ty + 1, omx, opx, omy, opy, tvN, dN, tmp);
    if (has_r) shell_site(Drop{}, zb - 1, rsx, rsy, rmx, rpx, rmy, rpy, tvR, dN, tmp);
    load18<T, C, MODE>(tv, ps, own_p(zb), nx, nxy, tvC);
    load18<T, C, MODE>(tv, ps, ring_p(zb), nx, nxy, tvR);
    shell_site(Keep{}, zb, tx + 1, ty + 1, omx, opx, omy, opy, tvC, dC, sC);
    if (has_r) shell_site(Drop{}, zb, rsx, rsy, rmx, rpx, rmy, rpy, tvR, dN, tmp);
    load18<T, C, MODE>(kn, ps, own_p(zb), nx, nxy, knC);
    load18<T, C, MODE>(tv, ps, own_p(zb + 1), nx, nxy, tvN);
    load18<T, C, MODE>(tv, ps, ring_p(zb + 1), nx, nxy, tvR);
  }
  __syncthreads();

  T acc = T(0);
  const int hx = tx + 2, hy = ty + 2, sx = tx + 1, sy = ty + 1;
  for (int z = zb; z < ze; ++z) {
    fetch_u(z + 3);  // lands during this step
    if constexpr (DEEP) {
      load18<T, C, MODE>(tv, ps, own_p(z + 2), nx, nxy, tvNN);
      load18<T, C, MODE>(kn, ps, own_p(z + 1), nx, nxy, knN);
      load18<T, C, MODE>(tv, ps, ring_p(z + 2), nx, nxy, tvRN);
    }
    {
      T tmp[4];
      if constexpr (!DEEP) load18<T, C, MODE>(tv, ps, own_p(z + 1), nx, nxy, tvN);
      shell_site(Keep{}, z + 1, sx, sy, omx, opx, omy, opy, tvN, dN, sN);
      if constexpr (!DEEP) load18<T, C, MODE>(tv, ps, ring_p(z + 1), nx, nxy, tvR);
      if (has_r) shell_site(Drop{}, z + 1, rsx, rsy, rmx, rpx, rmy, rpy, tvR, dN, tmp);
    }
    __syncthreads();
    if constexpr (!DEEP) {
      load18<T, C, MODE>(kn, ps, own_p(z), nx, nxy, knC);
      load18<T, C, MODE>(tv, ps, own_p(z), nx, nxy, tvC);
    }
    if (own_xy) {
      const int p = (z * ny + gyo) * nx + gxo;
      const bool mx = omx, px = opx, my = omy, py = opy, mz = z > 0, pz = z < nz - 1;
      const T(*s0p)[SY][SX] = ss[(z + 8) % 3];
      const T(*s1p)[SY][SX] = ss[(z + 9) % 3];
      const T(*s2p)[SY][SX] = ss[(z + 10) % 3];
      T fx = T(0), fy = T(0), fz = T(0);
      for_slots(Slots18{}, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int ox = off3(K, 0), oy = off3(K, 1), oz = off3(K, 2);
        const T(*sq)[SY][SX] = oz < 0 ? s0p : (oz > 0 ? s2p : s1p);
        T g;
        if constexpr (shell3(K) == 0) {
          // F = 2 Kn dL + 1/2 (TS_i + TS_q) + 1/2 Tv (S_i + S_q)
          const T t = tvC[K] * (sC[0] + sq[0][sy + oy][sx + ox]) + (sC[2] + sq[2][sy + oy][sx + ox]);
          g = T(2) * knC[K] * dC[K] + T(0.5) * t;
        } else {
          // G = F / sqrt2 = Kn e + 1/4 (TS'_i + TS'_q) + 1/4 Tv (S'_i + S'_q)   (primes: sqrt2-scaled sums)
          const T t = tvC[K] * (sC[1] + sq[1][sy + oy][sx + ox]) + (sC[3] + sq[3][sy + oy][sx + ox]);
          g = knC[K] * dC[K] + T(0.25) * t;
        }
        g = slot_ok<K>(mx, px, my, py, mz, pz) ? g : T(0);
        // (K u)_i += F n_k, n_k = -off_k / |off_k|; for diagonal slots F n_k = -G off_k
        if constexpr (ox != 0) fx += (ox > 0 ? -g : g);
        if constexpr (oy != 0) fy += (oy > 0 ? -g : g);
        if constexpr (oz != 0) fz += (oz > 0 ? -g : g);
      });
      T* __restrict__ o = a.out + a.bo.at(b, l) + 3 * (int64_t)p;
      if constexpr (EPI == EPI_PLAIN) {
        o[0] = fx;
        o[1] = fy;
        o[2] = fz;
      } else {
        const T(*u1)[HY][HX] = su[(z + 10) % 5];
        const T ux = u1[0][hy][hx], uy = u1[1][hy][hx], uz = u1[2][hy][hx];
        T m0 = T(1), m1 = T(1), m2 = T(1);
        if (a.mask) {
          const uint8_t* mk = a.mask + a.bm.at(b, l) + 3 * (int64_t)p;
          m0 = mk[0] ? T(1) : T(0);
          m1 = mk[1] ? T(1) : T(0);
          m2 = mk[2] ? T(1) : T(0);
        }
        if constexpr (EPI == EPI_MASKDOT) {
          fx *= m0;
          fy *= m1;
          fz *= m2;
          acc += ux * fx + uy * fy + uz * fz;
        } else {
          const T* rh = a.rhs + a.brhs.at(b, l) + 3 * (int64_t)p;
          fx = m0 * (rh[0] - fx);
          fy = m1 * (rh[1] - fy);
          fz = m2 * (rh[2] - fz);
          acc += fx * fx + fy * fy + fz * fz;
        }
        o[0] = fx;
        o[1] = fy;
        o[2] = fz;
      }
    }
    // registers of plane z+1 become the output plane; coefficients of the next steps
#pragma unroll
    for (int k = 0; k < 18; ++k) dC[k] = dN[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) sC[k] = sN[k];
    if constexpr (DEEP) {
#pragma unroll
      for (int k = 0; k < 18; ++k) {
        tvC[k] = tvN[k];
        tvN[k] = tvNN[k];
        knC[k] = knN[k];
        tvR[k] = tvRN[k];
      }
    }
    store_u(z + 3);  // into the slot of plane z-2 (no longer read)
    __syncthreads();
  }
  if constexpr (EPI != EPI_PLAIN) {
    const T s = block_sum(acc, red);
    if (tid == 0) a.partial[((int64_t)b * a.nl + l) * nblk + lid] = s;
  }
}

}  // namespace lpm
