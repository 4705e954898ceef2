// Lean fused K.u of the 3D simple-cubic LPM for the PCG hot path: FP32, packed symmetric coefficients, full boxes
// without removed particles, column tiles 64 sites wide (one wave per tile row).
//
// Same operator as ku3d_kernel (coarse3d/lin3d.py:98-149, coarse3d/torch_lpm3d.py:63-83).  With e = sqrt2 dL and the
// sqrt2-scaled face-diagonal shell sums S1, TS1 (see ku3d.hpp), F_ik n_k splits into a part that depends on site i
// only and a part that depends on the neighbour q = i + off_k:
//   axial    F n = -off [2 Kn dL + 1/2 TS0_i + 1/2 Tv S0_i]  - off [1/2 TS0_q + 1/2 Tv S0_q]
//   diagonal F n = -off [  Kn e + 1/4 TS1_i + 1/4 Tv S1_i]  - off [1/4 TS1_q + 1/4 Tv S1_q]
// The own part is complete as soon as the site's shell sums are, so it is summed into three registers in the step
// that computes them and the 18 elongations are not kept (the general kernel keeps two banks of them).  That, LDS
// rings of three planes for u and for the shell sums let three blocks of 256 threads stay resident per CU instead of
// two.  Coefficients are read with 32-bit lane offsets from uniform per-plane addresses.
//
// Missing bonds (slots leaving the box) need no per-slot test in the neighbour part: the packed planes hold Kn = Tv
// = 0 on them (lpm_pack_coef writes zeros there) and the shell sums of sites outside the box are written as zero.
// In the own part, an axial elongation of a missing bond vanishes by itself (its neighbour's u is read from the
// clamped halo copy, which is the site itself); the diagonal ones enter S1 through validity factors, and the
// 1/2 TS_i terms of missing slots are removed through sigma_c = sum of off_k over the present slots of shell c
// (zero inside the box).
#pragma once
#include "ku3d.hpp"

namespace lpm {

// waves per SIMD the 64 x 4 lean kernel is compiled for (<= 168 VGPRs); the launch geometry puts as many blocks
// on each CU
constexpr int kLeanWavesPerEU = 3;

typedef __amdgpu_buffer_rsrc_t BufRsrc;

// raw (unstrided) buffer resource over `bytes` bytes for the u planes (one dwordx3 per halo site); loads beyond it
// return zero (gfx9 dword 3)
__device__ __forceinline__ BufRsrc buf_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
// global load at a uniform base plus a uniform and a per-lane 32-bit byte offset (SGPR base, VGPR offset)
__device__ __forceinline__ float gl_ld(const float* __restrict__ base, uint32_t voff, uint32_t soff) {
  const char* b = (const char*)base + soff;
  return *(const float*)(b + voff);
}
typedef unsigned int U32x3 __attribute__((ext_vector_type(3)));
__device__ __forceinline__ Vec3<float> buf_ld3(BufRsrc r, uint32_t voff, uint32_t soff) {
  const U32x3 v = __builtin_amdgcn_raw_buffer_load_b96(r, (int)voff, (int)soff, 0);
  return Vec3<float>{__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z)};
}

// Elongation of slot K from LDS u planes (U0, U1, U2 = planes z-1, z, z+1 of the site): dL for axial slots, e = sqrt2
// dL for diagonal slots, e = off . (u_q - u_i).  For two components a < b with signs sa, sb of off:
// sa u_a + sb u_b = sa (u_a + sa sb u_b), so the site's combination (u_a +- u_b) is shared by every slot that uses it.
template <int K, int HY, int HX>
__device__ __forceinline__ float elong3(const float (*U0)[HY][HX], const float (*U1)[HY][HX],
                                        const float (*U2)[HY][HX], int hy, int hx, float ux, float uy, float uz) {
  constexpr int o[3] = {off3(K, 0), off3(K, 1), off3(K, 2)};
  const float(*Q)[HY][HX] = o[2] < 0 ? U0 : (o[2] > 0 ? U2 : U1);
  const int qy = hy + o[1], qx = hx + o[0];
  const float ui[3] = {ux, uy, uz};
  if constexpr (K < 6) {
    constexpr int a = o[0] != 0 ? 0 : (o[1] != 0 ? 1 : 2);
    const float uq = Q[a][qy][qx];
    return o[a] > 0 ? uq - ui[a] : ui[a] - uq;
  } else {
    constexpr int a = o[0] != 0 ? 0 : 1;
    constexpr int c = o[2] != 0 ? 2 : 1;
    constexpr bool same = (o[a] > 0) == (o[c] > 0);
    const float qa = Q[a][qy][qx], qc = Q[c][qy][qx];
    const float Qq = same ? qa + qc : qa - qc;
    const float Qi = same ? ui[a] + ui[c] : ui[a] - ui[c];
    return o[a] > 0 ? Qq - Qi : Qi - Qq;
  }
}

template <int TY, int EPI>
__global__ __launch_bounds__(64 * TY, TY == 4 ? kLeanWavesPerEU : 2) void ku3l_kernel(Ku3Args<float, float> a) {
  constexpr int TX = 64, NT = TX * TY;
  constexpr int HX = TX + 4, HY = TY + 4;  // u tile with a 2-site halo
  constexpr int SX = TX + 2, SY = TY + 2;  // shell sums with a 1-site halo
  constexpr int NR = 2 * SX + 2 * TY;      // halo ring of the shell sums
  constexpr int NLD = (HX * HY + NT - 1) / NT;
  static_assert(NR <= NT, "one ring site per thread at most");
  static_assert(EPI == EPI_PLAIN || EPI == EPI_MASKDOT, "lean kernel epilogues");
  // rings indexed by the plane's offset from zb modulo 3; the z loop is unrolled by 6 (with the two-bank Tv
  // rotation), so every ring slot of a step is a compile-time constant
  __shared__ float su[3][3][HY][HX];
  __shared__ float2 ss[3][2][SY][SX];  // per shell (S, TS); shell 1 sqrt2-scaled
  __shared__ float red[NT / 64];

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
  const int tx = tid & 63, ty = tid >> 6;

  // packed arrays from their leading guard: every coefficient offset is a non-negative 32-bit byte offset
  const uint32_t ps = (uint32_t)a.ps;
  const uint32_t G = (uint32_t)packed_guard(3, nx, ny);
  const float* __restrict__ knr = a.kn + a.bc.at(b, l) - G;
  const float* __restrict__ tvr = a.tv + a.bc.at(b, l) - G;
  const BufRsrc ru = buf_rsrc(a.u + a.bu.at(b, l), (uint32_t)nxy * (uint32_t)nz * 12u);
  // byte offset of slot K's entry for a site of plane z, relative to the site's in-plane offset
  auto coff = [&](int K, int z) -> uint32_t {
    const int dlt = (K & 1) ? off3(K, 0) + off3(K, 1) * nx + off3(K, 2) * nxy : 0;
    return 4u * ((uint32_t)(K >> 1) * ps + G + (uint32_t)z * (uint32_t)nxy + (uint32_t)dlt);
  };

  auto cl = [](int v, int n) { return min(max(v, 0), n - 1); };
  auto fl = [](bool c) { return c ? 1.0f : 0.0f; };

  // own column site
  const int gxo = x0 + tx, gyo = y0 + ty;
  const bool own_in = gxo < nx && gyo < ny;
  const int cgx = cl(gxo, nx), cgy = cl(gyo, ny);
  const float vxm = fl(cgx > 0), vxp = fl(cgx < nx - 1), vym = fl(cgy > 0), vyp = fl(cgy < ny - 1);
  const uint32_t own_off = (uint32_t)(cgy * nx + cgx) * 4u;
  // ring site (threads < NR) in shell coordinates (rsx, rsy)
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
  const int rgx_ = x0 - 1 + rsx, rgy_ = y0 - 1 + rsy;
  const bool ring_in = rgx_ >= 0 && rgx_ < nx && rgy_ >= 0 && rgy_ < ny;
  const int rgx = cl(rgx_, nx), rgy = cl(rgy_, ny);
  const float rxm = fl(rgx > 0), rxp = fl(rgx < nx - 1), rym = fl(rgy > 0), ryp = fl(rgy < ny - 1);
  const uint32_t r_off = (uint32_t)(rgy * nx + rgx) * 4u;

  // u halo sites of this thread: in-plane byte offsets and LDS positions
  uint32_t uoff[NLD];
#pragma unroll
  for (int j = 0; j < NLD; ++j) {
    const int idx = min(tid + j * NT, HX * HY - 1);
    const int hx = idx % HX, hy = idx / HX;
    uoff[j] = (uint32_t)((cl(y0 - 2 + hy, ny) * nx + cl(x0 - 2 + hx, nx)) * 12);
  }
  Vec3<float> ureg[NLD];
  auto fetch_u = [&](int z) {
    const uint32_t zo = (uint32_t)cl(z, nz) * (uint32_t)nxy * 12u;
#pragma unroll
    for (int j = 0; j < NLD; ++j) ureg[j] = buf_ld3(ru, uoff[j], zo);
  };
  auto store_u = [&](int slot) {
    float(*dst)[HY][HX] = su[slot];
#pragma unroll
    for (int j = 0; j < NLD; ++j) {
      const int idx = tid + j * NT;
      if (j + 1 < NLD || idx < HX * HY) {
        const int hx = idx % HX, hy = idx / HX;
        dst[0][hy][hx] = ureg[j].x;
        dst[1][hy][hx] = ureg[j].y;
        dst[2][hy][hx] = ureg[j].z;
      }
    }
  };

  // the site left of the wave's first site (x0 - 1, same row): its entries come from uniform (scalar) loads
  const uint32_t row0 = __builtin_amdgcn_readfirstlane((uint32_t)(cgy * nx + cgx));
  auto left_entry = [&](const float* __restrict__ raw, int plane, int z) -> float {
    return raw[(uint32_t)plane * ps + G + (uint32_t)z * (uint32_t)nxy + row0 - 1u];
  };

  // Shell sums of one site of plane z.  tv: the site's 18 Tv; validity factors (m = minus, p = plus side) of the
  // in-plane neighbours and of the planes below / above.  Writes (S0, TS0), (S1, TS1) to ring slot R3, zero for a
  // site outside the box.  Returns the elongations in e when KEEP.
  auto shell = [&](auto rel_tag, int hx, int hy, int sx, int sy, bool in, float xm, float xp, float ym, float yp,
                   float zm, float zp, const float (&tv)[18], float (&e)[18], float (&s)[4], float (&uc)[3]) {
    constexpr int R3 = decltype(rel_tag)::value;
    const float(*U0)[HY][HX] = su[(R3 + 2) % 3];
    const float(*U1)[HY][HX] = su[R3];
    const float(*U2)[HY][HX] = su[(R3 + 1) % 3];
    const float ux = U1[0][hy][hx], uy = U1[1][hy][hx], uz = U1[2][hy][hx];
    uc[0] = ux;
    uc[1] = uy;
    uc[2] = uz;
    for_slots(Slots18{}, [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      e[K] = elong3<K, HY, HX>(U0, U1, U2, hy, hx, ux, uy, uz);
    });
    const float s0 = ((e[0] + e[1]) + (e[2] + e[3])) + (e[4] + e[5]);
    float t0 = tv[0] * e[0];
    float t1 = tv[6] * e[6];
#pragma unroll
    for (int k = 1; k < 6; ++k) t0 = fmaf(tv[k], e[k], t0);
#pragma unroll
    for (int k = 7; k < 18; ++k) t1 = fmaf(tv[k], e[k], t1);
    // S1 over the present face-diagonal bonds
    const float sxy = fmaf(xp * yp, e[6], fmaf(xm * ym, e[7], fmaf(xp * ym, e[8], (xm * yp) * e[9])));
    const float sup = fmaf(xp, e[10], fmaf(xm, e[13], fmaf(yp, e[14], ym * e[17])));
    const float sdn = fmaf(xm, e[11], fmaf(xp, e[12], fmaf(ym, e[15], yp * e[16])));
    const float s1 = fmaf(zp, sup, fmaf(zm, sdn, sxy));
    float2(*dst)[SY][SX] = ss[R3];
    dst[0][sy][sx] = in ? make_float2(s0, t0) : make_float2(0.0f, 0.0f);
    dst[1][sy][sx] = in ? make_float2(s1, t1) : make_float2(0.0f, 0.0f);
    s[0] = s0;
    s[1] = s1;
    s[2] = t0;
    s[3] = t1;
  };

  // Own part of (K u)_i from the elongations e, shell sums s and the site's Kn, Tv of the plane (zm, zp: plane
  // validity factors).
  auto own_part = [&](const float (&e)[18], const float (&s)[4], const float (&kn)[18], const float (&tv)[18],
                      float zm, float zp, float (&A)[3]) {
    const float hs0 = 0.5f * s[0];
    // axial pairs (2a, 2a+1): -off [2 Kn dL + 1/2 Tv S0]
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float kd = fmaf(-kn[2 * c], e[2 * c], kn[2 * c + 1] * e[2 * c + 1]);
      A[c] = fmaf(2.0f, kd, hs0 * (tv[2 * c + 1] - tv[2 * c]));
    }
    // face-diagonal pairs (k, k+1), off_{k+1} = -off_k: -off_k [(Kn e)_k - (Kn e)_{k+1} + 1/4 S1 (Tv_k - Tv_{k+1})]
    float W[3] = {0.0f, 0.0f, 0.0f};
    for_slots(Seq<6, 8, 10, 12, 14, 16>{}, [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      const float kd = fmaf(-kn[K + 1], e[K + 1], kn[K] * e[K]);
      const float td = tv[K] - tv[K + 1];
      for_slots(Seq<0, 1, 2>{}, [&](auto cc) {
        constexpr int c = decltype(cc)::value;
        if constexpr (off3(K, c) > 0) {
          A[c] -= kd;
          W[c] += td;
        } else if constexpr (off3(K, c) < 0) {
          A[c] += kd;
          W[c] -= td;
        }
      });
    });
    const float qs1 = -0.25f * s[1];
#pragma unroll
    for (int c = 0; c < 3; ++c) A[c] = fmaf(qs1, W[c], A[c]);
    // -1/2 TS0_i sigma0 - 1/4 TS1_i sigma1 (sigma_c = sum of off_k over the present slots of shell c)
    const float dx = vxp - vxm, dy = vyp - vym, dz = zp - zm;
    const float sgx = vxp + vxm, sgy = vyp + vym, sgz = zp + zm;
    const float h0 = -0.5f * s[2], q1 = -0.25f * s[3];
    A[0] = fmaf(h0, dx, fmaf(q1, dx * (sgy + sgz), A[0]));
    A[1] = fmaf(h0, dy, fmaf(q1, dy * (sgx + sgz), A[1]));
    A[2] = fmaf(h0, dz, fmaf(q1, dz * (sgx + sgy), A[2]));
  };

  // coefficient loads of one site; SKIP: slots the caller fills from registers or neighbouring lanes
  auto load18 = [&](auto skip_tag, const float* __restrict__ raw, uint32_t voff, int z, float (&out)[18]) {
    constexpr unsigned SKIP = decltype(skip_tag)::value;
    for_slots(Slots18{}, [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      if constexpr (!((SKIP >> K) & 1u)) out[K] = gl_ld(raw, voff, coff(K, z));
    });
  };
  using AllSlots = std::integral_constant<unsigned, 0u>;
  // -x (1), -z (5) and (-x, -z) (11) entries: the left lane's slot 0, the site's own slot 4 one plane below and the
  // left lane's slot 10 one plane below
  using Reuse = std::integral_constant<unsigned, (1u << 1) | (1u << 5) | (1u << 11)>;
  using OnlyZ = std::integral_constant<unsigned, (1u << 5)>;

  // registers carried from step to step: own Tv of the output plane (two banks), the own part and displacement of
  // the output plane and of the next plane, Kn slots 4 and 10 of the last shell plane, the ring site's Tv slot 4
  float TV[2][18], A[2][3], UO[2][3];
  float kn4 = 0.0f, kn10 = 0.0f, tr4 = 0.0f;

  // ---- prologue: u planes zb-2 .. zb; shell sums of plane zb-1; plane zb+1; shell sums (+ own part) of plane zb;
  // plane zb+2
  for (int z = zb - 2; z <= zb; ++z) {
    fetch_u(z);
    store_u((z - zb + 6) % 3);
  }
  __syncthreads();
  {
    float e[18], s[4], uc[3], tr[18];
    const int z = zb - 1;
    if (z >= 0) {
      const float zm = fl(z > 0), zp = fl(z < nz - 1);
      load18(AllSlots{}, tvr, own_off, z, TV[1]);
      shell(std::integral_constant<int, 2>{}, tx + 2, ty + 2, tx + 1, ty + 1, own_in, vxm, vxp, vym, vyp, zm, zp,
            TV[1], e, s, uc);
      if (has_r) {
        load18(AllSlots{}, tvr, r_off, z, tr);
        shell(std::integral_constant<int, 2>{}, rsx + 1, rsy + 1, rsx, rsy, ring_in, rxm, rxp, rym, ryp, zm, zp, tr,
              e, s, uc);
      }
    } else {
      // plane -1 is outside the box: zero shell sums
      for (int i = tid; i < 2 * SY * SX; i += NT) (&ss[2][0][0][0])[i] = make_float2(0.0f, 0.0f);
    }
  }
  __syncthreads();
  fetch_u(zb + 1);
  store_u(1);
  __syncthreads();
  {
    float e[18], s[4], kn[18], tr[18];
    const int z = zb;
    const float zm = fl(z > 0), zp = fl(z < nz - 1);
    load18(AllSlots{}, tvr, own_off, z, TV[0]);
    load18(AllSlots{}, knr, own_off, z, kn);
    shell(std::integral_constant<int, 0>{}, tx + 2, ty + 2, tx + 1, ty + 1, own_in, vxm, vxp, vym, vyp, zm, zp, TV[0],
          e, s, UO[0]);
    own_part(e, s, kn, TV[0], zm, zp, A[0]);
    kn4 = kn[4];
    kn10 = kn[10];
    if (has_r) {
      float uc[3];
      load18(AllSlots{}, tvr, r_off, z, tr);
      shell(std::integral_constant<int, 0>{}, rsx + 1, rsy + 1, rsx, rsy, ring_in, rxm, rxp, rym, ryp, zm, zp, tr, e,
            s, uc);
      tr4 = tr[4];
    }
  }
  __syncthreads();
  fetch_u(zb + 2);
  store_u(2);
  __syncthreads();

  float acc = 0.0f;
  const int sx = tx + 1, sy = ty + 1;
  const int64_t obase = a.bo.at(b, l);
  const uint32_t oxy = (uint32_t)(gyo * nx + gxo);

  // one step: shell sums and own part of plane z+1, barrier, output of plane z, u plane z+3 into the ring
  auto step = [&](auto jc, int z) __attribute__((always_inline)) {
    constexpr int J = decltype(jc)::value;
    float(&tvC)[18] = TV[J % 2];
    float(&tvN)[18] = TV[(J + 1) % 2];
    float(&aC)[3] = A[J % 2];
    float(&aN)[3] = A[(J + 1) % 2];
    float(&uC)[3] = UO[J % 2];
    float(&uN)[3] = UO[(J + 1) % 2];
    using RN = std::integral_constant<int, (J + 1) % 3>;
    fetch_u(z + 3);  // lands during the step
    const int zs = z + 1;
    if (zs < nz) {
      const float zm = 1.0f, zp = fl(zs < nz - 1);
      float e[18], s[4], kn[18];
      // the ring site first, so that its Tv registers are free before the own site's loads are issued
      if (has_r) {
        float tr[18], uc[3];
        load18(OnlyZ{}, tvr, r_off, zs, tr);
        tr[5] = tr4;
        shell(RN{}, rsx + 1, rsy + 1, rsx, rsy, ring_in, rxm, rxp, rym, ryp, zm, zp, tr, e, s, uc);
        tr4 = tr[4];
      }
      __builtin_amdgcn_sched_barrier(0);
      // own Tv / Kn of plane zs: -z entries from the previous plane, -x entries from the lane to the left
      load18(Reuse{}, tvr, own_off, zs, tvN);
      load18(Reuse{}, knr, own_off, zs, kn);
      tvN[5] = tvC[4];
      kn[5] = kn4;
      tvN[1] = lane_left(tvN[0], left_entry(tvr, 0, zs));
      tvN[11] = lane_left(tvC[10], left_entry(tvr, 5, zs - 1));
      kn[1] = lane_left(kn[0], left_entry(knr, 0, zs));
      kn[11] = lane_left(kn10, left_entry(knr, 5, zs - 1));
      shell(RN{}, tx + 2, ty + 2, sx, sy, own_in, vxm, vxp, vym, vyp, zm, zp, tvN, e, s, uN);
      own_part(e, s, kn, tvN, zm, zp, aN);
      kn4 = kn[4];
      kn10 = kn[10];
    } else {
      // plane nz is outside the box: zero shell sums
      for (int i = tid; i < 2 * SY * SX; i += NT) (&ss[RN::value][0][0][0])[i] = make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    if (own_in) {
      const float2(*s0p)[SY][SX] = ss[(J + 2) % 3];  // plane z-1
      const float2(*s1p)[SY][SX] = ss[J % 3];        // plane z
      const float2(*s2p)[SY][SX] = ss[(J + 1) % 3];  // plane z+1
      float bx = 0.0f, by = 0.0f, bz = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
      for_slots(Slots18{}, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int ox = off3(K, 0), oy = off3(K, 1), oz = off3(K, 2);
        const float2(*sq)[SY][SX] = oz < 0 ? s0p : (oz > 0 ? s2p : s1p);
        const float2 st = sq[shell3(K)][sy + oy][sx + ox];
        const float v = fmaf(tvC[K], st.x, st.y);
        if constexpr (K < 6) {
          if constexpr (ox != 0) bx += ox > 0 ? -v : v;
          if constexpr (oy != 0) by += oy > 0 ? -v : v;
          if constexpr (oz != 0) bz += oz > 0 ? -v : v;
        } else {
          if constexpr (ox != 0) cx += ox > 0 ? -v : v;
          if constexpr (oy != 0) cy += oy > 0 ? -v : v;
          if constexpr (oz != 0) cz += oz > 0 ? -v : v;
        }
      });
      const float fx = fmaf(0.25f, cx, fmaf(0.5f, bx, aC[0]));
      const float fy = fmaf(0.25f, cy, fmaf(0.5f, by, aC[1]));
      const float fz = fmaf(0.25f, cz, fmaf(0.5f, bz, aC[2]));
      if constexpr (EPI == EPI_MASKDOT) acc = fmaf(uC[0], fx, fmaf(uC[1], fy, fmaf(uC[2], fz, acc)));
      *(Vec3<float>*)(a.out + obase + ((int64_t)z * nxy + oxy) * 3) = Vec3<float>{fx, fy, fz};
    }
    store_u(J % 3);  // plane z+3, into the slot of plane z (read only before the barrier above)
    __syncthreads();
  };

  constexpr int U = 6;
  for (int z = zb; z < ze; z += U) {
    [&]<int... Js>(std::integer_sequence<int, Js...>) __attribute__((always_inline)) {
      bool go = true;
      ((go = go && (z + Js < ze), go ? (step(std::integral_constant<int, Js>{}, z + Js), 0) : 0), ...);
    }(std::make_integer_sequence<int, U>{});
  }
  if constexpr (EPI == EPI_MASKDOT) {
    const float s = block_sum(acc, red);
    if (tid == 0) a.partial[((int64_t)b * a.nl + l) * nblk + lid] = s;
  }
}

}  // namespace lpm
