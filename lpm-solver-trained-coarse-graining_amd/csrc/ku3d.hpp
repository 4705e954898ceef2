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
#include <utility>

#include "lpm_common.hpp"

namespace lpm {

enum Ku3Epi : int {
  EPI_PLAIN = 0,    // out = K u
  EPI_MASKDOT = 1,  // out = K u, partial dot(u, K u) per block; u vanishes on fixed DOFs inside (P)CG, so the dot
                    // equals dot(u, mask * K u) and the mask is applied where the product is consumed (cg_update)
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
  const uint32_t* bonds;  // bond words (removed particles), per particle: bit k = slot k is a bond; BONDS kernels
  Bcast bb;
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

// the three displacement components of a site, fetched by one vector load
template <typename T>
struct Vec3 {
  T x, y, z;
};

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

// The 18 directed coefficients of a site of one plane.  `zbase` is the plane's first particle (a uniform address,
// directed: row 0 of the plane; packed: plane-0 entry of the plane's first particle) and `xy` the site's in-plane
// index, which stays the same for every plane of a z march: the loads take one loop-invariant lane offset and a
// uniform per-slot base (packed odd slots: the neighbour's forward-plane entry at p + delta, inside the zero guards).
// SKIP: slots (bit K) the caller fills from registers instead (packed layout: slot 5 (-z) is slot 4 of the same
// column one plane below; slots 1 (-x) and 11 (-x, -z) are slots 0 and 10 of the lane to the left).
template <typename T, typename C, int MODE, unsigned SKIP = 0u>
__device__ __forceinline__ void load18(const C* __restrict__ zbase, int64_t ps, unsigned xy, int nx, int nxy,
                                       T (&out)[18]) {
  for_slots(Slots18{}, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    if constexpr ((SKIP >> K) & 1u) {
    } else if constexpr (MODE == COEF_DIRECTED) {
      out[K] = (T)ld_off(zbase, xy * (18u * (unsigned)sizeof(C)) + K * (unsigned)sizeof(C));  // < 4 GiB (host check)
    } else {
      const int64_t sh = (int64_t)(K >> 1) * ps + ((K & 1) ? off3(K, 0) + off3(K, 1) * nx + off3(K, 2) * nxy : 0);
      out[K] = (T)ld_off(zbase + sh, xy * (unsigned)sizeof(C));
    }
  });
}

// v of the lane to the left (wave shift right by one lane); lane 0 takes `lane0`
__device__ __forceinline__ float lane_left(float v, float lane0) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(lane0), __float_as_int(v), 0x138, 0xf, 0xf, false));
}

template <typename T, typename C, int MODE, int TX, int TY, int EPI, bool BONDS = false>
__global__ __launch_bounds__(TX* TY, (sizeof(T) == 4 && TX * TY <= 256) ? 2 : 1) void ku3d_kernel(Ku3Args<T, C> a) {
  constexpr int HX = TX + 4, HY = TY + 4;  // u tile with a 2-site halo
  constexpr int SX = TX + 2, SY = TY + 2;  // shell sums with a 1-site halo
  constexpr int NT = TX * TY;
  constexpr int NR = 2 * SX + 2 * TY;           // ring of halo sites of the shell sums
  constexpr int NLD = (HX * HY + NT - 1) / NT;  // u sites loaded per thread per plane
  static_assert(NR <= NT, "one ring site per thread at most");
  // LDS rings indexed by the plane's offset from zb modulo 6 (u) and 3 (shell sums): with the z loop unrolled by 6
  // every ring slot of a step is a compile-time constant
  __shared__ T su[6][3][HY][HX];
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
  const uint32_t* __restrict__ bwd = BONDS ? a.bonds + a.bb.at(b, l) : nullptr;

  // Site coordinates are clamped into the box for addressing.  Results of halo sites outside the box are never
  // read through an existing bond, so they need no masking.
  auto cl = [](int v, int n) { return min(max(v, 0), n - 1); };

  // own column site
  const int gxo = x0 + tx, gyo = y0 + ty;
  const bool own_xy = gxo < nx && gyo < ny;
  const int cgx = cl(gxo, nx), cgy = cl(gyo, ny);
  const bool omx = cgx > 0, opx = cgx < nx - 1, omy = cgy > 0, opy = cgy < ny - 1;
  const unsigned own_xy_off = (unsigned)(cgy * nx + cgx);
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
  const unsigned r_xy_off = (unsigned)(rgy * nx + rgx);

  // u halo sites of this thread: in-plane byte offsets (loop invariant) and LDS positions
  unsigned uoff[NLD];
#pragma unroll
  for (int j = 0; j < NLD; ++j) {
    const int idx = min(tid + j * NT, HX * HY - 1);
    const int hx = idx % HX, hy = idx / HX;
    uoff[j] = (unsigned)((cl(y0 - 2 + hy, ny) * nx + cl(x0 - 2 + hx, nx)) * 3 * (int)sizeof(T));
  }
  // uniform bases of a plane (clamped into the box)
  constexpr int CROW = MODE == COEF_DIRECTED ? 18 : 1;
  auto kn_z = [&](int z) { return kn + (int64_t)cl(z, nz) * nxy * CROW; };
  auto tv_z = [&](int z) { return tv + (int64_t)cl(z, nz) * nxy * CROW; };

  T ureg[NLD][3];
  // FP32: the coefficients of the next steps are loaded a full step ahead; FP64 has no registers for a second set
  // and loads them after the output phase
  constexpr bool DEEP = sizeof(T) == 4;
  // register banks whose roles rotate with the step (the z loop is unrolled by the rotation period, so the
  // rotation is renaming): own elongations / shell sums of the output plane and of the next plane; own Tv of the
  // output plane, the next plane and (FP32) the plane after; own Kn of the output plane (loaded at the start of
  // its step); ring Tv of the next plane and the plane after
  constexpr int NTV = DEEP ? 3 : 2;
  // packed FP32: the -z coefficients come from the previous plane's registers instead of a load
  constexpr bool REUSE5 = DEEP && MODE == COEF_PACKED;
  // 64-wide tiles: a wave is one row of the tile, so the -x neighbour of a site is the lane to the left
  constexpr bool LANEX = REUSE5 && TX == 64;
  constexpr unsigned SKIPK = LANEX ? ((1u << 1) | (1u << 5) | (1u << 11)) : (REUSE5 ? (1u << 5) : 0u);
  T DD[2][18], SS[2][4], TV[NTV][18], KN[1][18], TR[2][18];
  // bond words (BONDS): own site per plane in the Tv rotation, ring site per plane in the ring-Tv rotation
  uint32_t WO[NTV], WR[2];
  auto word = [&](int z, unsigned xy) -> uint32_t {
    if constexpr (BONDS)
      return ld_off(bwd + (int64_t)cl(z, nz) * nxy, xy * 4u);
    else
      return 0xffffffffu;
  };

  auto fetch_u = [&](int z) {
    const T* uz = u + (int64_t)cl(z, nz) * nxy * 3;
#pragma unroll
    for (int j = 0; j < NLD; ++j) {
      const Vec3<T> v = ld_off((const Vec3<T>*)uz, uoff[j]);
      ureg[j][0] = v.x;
      ureg[j][1] = v.y;
      ureg[j][2] = v.z;
    }
  };
  auto store_u = [&](int slot) {
    T(*dst)[HY][HX] = su[slot];
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

  // Shell sums of the site at shell coordinates (sx, sy) in plane z (R = (z - zb) mod 6), in-plane flags, Tv of
  // the site; optionally keeps the 18 elongations.  Writes S0, sqrt2 S1, TS0, sqrt2 TS1 to the LDS ring and
  // returns them in s.
  auto shell_site = [&](auto keep_tag, auto rel_tag, int z, int sx, int sy, bool mx, bool px, bool my, bool py,
                        const T (&tvv)[18], uint32_t wv, T (&keep)[18], T (&s)[4]) {
    constexpr bool KEEP = decltype(keep_tag)::value;
    constexpr int R = decltype(rel_tag)::value;
    const int zc_ = cl(z, nz);
    const bool mz = zc_ > 0, pz = zc_ < nz - 1;
    const T(*u0)[HY][HX] = su[(R + 5) % 6];  // plane z-1
    const T(*u1)[HY][HX] = su[R];            // plane z
    const T(*u2)[HY][HX] = su[(R + 1) % 6];  // plane z+1
    const int hx = sx + 1, hy = sy + 1;
    const T ux = u1[0][hy][hx], uy = u1[1][hy][hx], uz = u1[2][hy][hx];
    T s0 = T(0), s1 = T(0), t0 = T(0), t1 = T(0);
    for_slots(Slots18{}, [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      constexpr int oz = off3(K, 2);
      const T(*uq)[HY][HX] = oz < 0 ? u0 : (oz > 0 ? u2 : u1);
      T d = dle3<T, K, HY, HX>(uq, hy, hx, ux, uy, uz);
      bool ok = slot_ok<K>(mx, px, my, py, mz, pz);
      if constexpr (BONDS) ok = ok && ((wv >> K) & 1u);
      d = ok ? d : T(0);
      if constexpr (KEEP) keep[K] = d;
      if constexpr (shell3(K) == 0) {
        s0 += d;
        t0 += tvv[K] * d;
      } else {
        s1 += d;
        t1 += tvv[K] * d;
      }
    });
    T(*dst)[SY][SX] = ss[R % 3];
    dst[0][sy][sx] = s0;
    dst[1][sy][sx] = s1;
    dst[2][sy][sx] = t0;
    dst[3][sy][sx] = t1;
    s[0] = s0;
    s[1] = s1;
    s[2] = t0;
    s[3] = t1;
  };

  // entry of forward plane `plane` at the site left of the wave's first site, shifted by `dz` sites (a uniform
  // address: one scalar load per wave)
  const unsigned row0 = __builtin_amdgcn_readfirstlane(own_xy_off);
  auto lane0_entry = [&](const C* zbase, int plane, int dz) -> T {
    return (T)zbase[(int64_t)plane * ps + (int64_t)row0 - 1 + dz];
  };
  T tl1 = T(0), tl11 = T(0), kl1 = T(0), kl11 = T(0), k10 = T(0);

  // ---- prologue: u planes zb-2 .. zb+2 in LDS; shell sums of planes zb-1 and zb
  for (int z = zb - 2; z <= zb + 2; ++z) {
    fetch_u(z);
    store_u((z - zb + 6) % 6);
  }
  __syncthreads();
  using Keep = std::integral_constant<bool, true>;
  using Drop = std::integral_constant<bool, false>;
  {
    T tmp[4];
    load18<T, C, MODE>(tv_z(zb - 1), ps, own_xy_off, nx, nxy, TV[1]);
    load18<T, C, MODE>(tv_z(zb - 1), ps, r_xy_off, nx, nxy, TR[0]);
    WO[1] = word(zb - 1, own_xy_off);
    WR[0] = word(zb - 1, r_xy_off);
    using Rm1 = std::integral_constant<int, 5>;
    using R0 = std::integral_constant<int, 0>;
    shell_site(Drop{}, Rm1{}, zb - 1, tx + 1, ty + 1, omx, opx, omy, opy, TV[1], WO[1], DD[1], tmp);
    if (has_r) shell_site(Drop{}, Rm1{}, zb - 1, rsx, rsy, rmx, rpx, rmy, rpy, TR[0], WR[0], DD[1], tmp);
    load18<T, C, MODE>(tv_z(zb), ps, own_xy_off, nx, nxy, TV[0]);
    load18<T, C, MODE>(tv_z(zb), ps, r_xy_off, nx, nxy, TR[0]);
    WO[0] = word(zb, own_xy_off);
    WR[0] = word(zb, r_xy_off);
    shell_site(Keep{}, R0{}, zb, tx + 1, ty + 1, omx, opx, omy, opy, TV[0], WO[0], DD[0], SS[0]);
    if (has_r) shell_site(Drop{}, R0{}, zb, rsx, rsy, rmx, rpx, rmy, rpy, TR[0], WR[0], DD[1], tmp);
    if constexpr (DEEP) {
      if constexpr (REUSE5) {
        load18<T, C, MODE, SKIPK>(tv_z(zb + 1), ps, own_xy_off, nx, nxy, TV[1]);
        TV[1][5] = TV[0][4];
        if constexpr (LANEX) {
          tl1 = lane0_entry(tv_z(zb + 1), 0, 0);
          tl11 = lane0_entry(tv_z(zb + 1), 5, -nxy);
          KN[0][10] = (T)ld_off(kn_z(zb - 1) + 5 * ps, own_xy_off * (unsigned)sizeof(C));  // slot 10 below zb
        }
        KN[0][4] = (T)ld_off(kn_z(zb - 1) + 2 * ps, own_xy_off * (unsigned)sizeof(C));  // slot 4 below zb
        load18<T, C, MODE, (1u << 5)>(tv_z(zb + 1), ps, r_xy_off, nx, nxy, TR[1]);
        TR[1][5] = TR[0][4];
      } else {
        load18<T, C, MODE>(tv_z(zb + 1), ps, own_xy_off, nx, nxy, TV[1]);
      }
      WO[1] = word(zb + 1, own_xy_off);
      WR[1] = word(zb + 1, r_xy_off);
    }
  }
  __syncthreads();

  T acc = T(0);
  const int hx = tx + 2, hy = ty + 2, sx = tx + 1, sy = ty + 1;
  const int64_t obase = a.bo.at(b, l), mbase = a.bm.at(b, l), rbase = a.brhs.at(b, l);
  const unsigned oxy = (unsigned)(gyo * nx + gxo);

  // one step of the march: shell sums of plane z+1, barrier, output of plane z.  J = step index modulo the
  // rotation period selects the banks.
  auto step = [&](auto jc, int z) __attribute__((always_inline)) {
    constexpr int J = decltype(jc)::value;
    T(&dC)[18] = DD[J % 2];
    T(&dN)[18] = DD[(J + 1) % 2];
    T(&sC)[4] = SS[J % 2];
    T(&sN)[4] = SS[(J + 1) % 2];
    T(&tvC)[18] = TV[J % NTV];
    T(&tvN)[18] = TV[(J + 1) % NTV];
    T(&knC)[18] = KN[0];
    T(&tvR)[18] = TR[(J + 1) % 2];
    uint32_t& wC = WO[J % NTV];
    uint32_t& wN = WO[(J + 1) % NTV];
    uint32_t& wR = WR[(J + 1) % 2];
    fetch_u(z + 3);  // lands during this step
    if constexpr (DEEP) {
      // own Tv a full step ahead (it feeds both phases); the ring's Tv lands during the own shell sums, Kn of the
      // output plane during the shell phase
      if constexpr (REUSE5) {
        const T k4 = knC[4];  // own Kn of plane z-1
        if constexpr (LANEX) {
          // slots 1 and 11 of the Tv plane that becomes the shell plane now: from the lane to the left
          tvN[1] = lane_left(tvN[0], tl1);
          tvN[11] = lane_left(tvC[10], tl11);
          k10 = knC[10];  // own Kn of plane z-1, slot 10
        }
        load18<T, C, MODE, SKIPK>(tv_z(z + 2), ps, own_xy_off, nx, nxy, TV[(J + 2) % NTV]);
        TV[(J + 2) % NTV][5] = TV[(J + 1) % NTV][4];
        if constexpr (LANEX) {
          tl1 = lane0_entry(tv_z(z + 2), 0, 0);
          tl11 = lane0_entry(tv_z(z + 2), 5, -nxy);
          kl1 = lane0_entry(kn_z(z), 0, 0);
          kl11 = lane0_entry(kn_z(z), 5, -nxy);
        }
        load18<T, C, MODE, (1u << 5)>(tv_z(z + 2), ps, r_xy_off, nx, nxy, TR[J % 2]);
        TR[J % 2][5] = tvR[4];
        load18<T, C, MODE, SKIPK>(kn_z(z), ps, own_xy_off, nx, nxy, knC);
        knC[5] = k4;
      } else {
        load18<T, C, MODE>(tv_z(z + 2), ps, own_xy_off, nx, nxy, TV[(J + 2) % NTV]);
        load18<T, C, MODE>(tv_z(z + 1), ps, r_xy_off, nx, nxy, tvR);
        load18<T, C, MODE>(kn_z(z), ps, own_xy_off, nx, nxy, knC);
      }
      WO[(J + 2) % NTV] = word(z + 2, own_xy_off);
      // the non-REUSE5 branch loads the ring's Tv of plane z + 1 into tvR; REUSE5 loads plane z + 2 into TR[J % 2]
      if constexpr (REUSE5)
        WR[J % 2] = word(z + 2, r_xy_off);
      else
        wR = word(z + 1, r_xy_off);
    }
    {
      T tmp[4];
      if constexpr (!DEEP) {
        load18<T, C, MODE>(tv_z(z + 1), ps, own_xy_off, nx, nxy, tvN);
        wN = word(z + 1, own_xy_off);
      }
      using RN = std::integral_constant<int, (J + 1) % 6>;
      shell_site(Keep{}, RN{}, z + 1, sx, sy, omx, opx, omy, opy, tvN, wN, dN, sN);
      if constexpr (!DEEP) {
        load18<T, C, MODE>(tv_z(z + 1), ps, r_xy_off, nx, nxy, tvR);
        wR = word(z + 1, r_xy_off);
      }
      if (has_r) shell_site(Drop{}, RN{}, z + 1, rsx, rsy, rmx, rpx, rmy, rpy, tvR, wR, dN, tmp);
    }
    __syncthreads();
    if constexpr (!DEEP) {
      load18<T, C, MODE>(kn_z(z), ps, own_xy_off, nx, nxy, knC);
      load18<T, C, MODE>(tv_z(z), ps, own_xy_off, nx, nxy, tvC);
    }
    if constexpr (LANEX) {
      knC[1] = lane_left(knC[0], kl1);
      knC[11] = lane_left(k10, kl11);
    }
    if (own_xy) {
      const bool mx = omx, px = opx, my = omy, py = opy, mz = z > 0, pz = z < nz - 1;
      const T(*s0p)[SY][SX] = ss[(J + 2) % 3];  // plane z-1
      const T(*s1p)[SY][SX] = ss[J % 3];        // plane z
      const T(*s2p)[SY][SX] = ss[(J + 1) % 3];  // plane z+1
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
        bool ok = slot_ok<K>(mx, px, my, py, mz, pz);
        if constexpr (BONDS) ok = ok && ((wC >> K) & 1u);
        g = ok ? g : T(0);
        // (K u)_i += F n_k, n_k = -off_k / |off_k|; for diagonal slots F n_k = -G off_k
        if constexpr (ox != 0) fx += (ox > 0 ? -g : g);
        if constexpr (oy != 0) fy += (oy > 0 ? -g : g);
        if constexpr (oz != 0) fz += (oz > 0 ? -g : g);
      });
      const int64_t pz3 = (int64_t)z * nxy * 3;
      T* __restrict__ o = a.out + obase + pz3;
      const unsigned ob = oxy * 3u * (unsigned)sizeof(T);
      if constexpr (EPI == EPI_PLAIN) {
        *(T*)((char*)o + ob) = fx;
        *(T*)((char*)o + ob + sizeof(T)) = fy;
        *(T*)((char*)o + ob + 2 * sizeof(T)) = fz;
      } else {
        const T(*u1)[HY][HX] = su[J % 6];
        const T ux = u1[0][hy][hx], uy = u1[1][hy][hx], uz = u1[2][hy][hx];
        if constexpr (EPI == EPI_MASKDOT) {
          acc += ux * fx + uy * fy + uz * fz;
        } else {
          T m0 = T(1), m1 = T(1), m2 = T(1);
          if (a.mask) {
            const uint8_t* mk = a.mask + mbase + pz3;
            m0 = ld_off(mk, oxy * 3u) ? T(1) : T(0);
            m1 = ld_off(mk, oxy * 3u + 1u) ? T(1) : T(0);
            m2 = ld_off(mk, oxy * 3u + 2u) ? T(1) : T(0);
          }
          const T* rh = a.rhs + rbase + pz3;
          fx = m0 * (ld_off(rh, ob) - fx);
          fy = m1 * (ld_off(rh, ob + (unsigned)sizeof(T)) - fy);
          fz = m2 * (ld_off(rh, ob + 2u * (unsigned)sizeof(T)) - fz);
          acc += fx * fx + fy * fy + fz * fz;
        }
        *(T*)((char*)o + ob) = fx;
        *(T*)((char*)o + ob + sizeof(T)) = fy;
        *(T*)((char*)o + ob + 2 * sizeof(T)) = fz;
      }
    }
    store_u((J + 3) % 6);  // plane z+3, into the slot of plane z-3 (no longer read)
    __syncthreads();
  };

  constexpr int U = 6;  // lcm of the bank counts and of the LDS ring lengths
  for (int z = zb; z < ze; z += U) {
    [&]<int... Js>(std::integer_sequence<int, Js...>) __attribute__((always_inline)) {
      bool go = true;
      ((go = go && (z + Js < ze), go ? (step(std::integral_constant<int, Js>{}, z + Js), 0) : 0), ...);
    }(std::make_integer_sequence<int, U>{});
  }
  if constexpr (EPI != EPI_PLAIN) {
    const T s = block_sum(acc, red);
    if (tid == 0) a.partial[((int64_t)b * a.nl + l) * nblk + lid] = s;
  }
}

}  // namespace lpm
