// Matrix-free K.u of the 3D simple-cubic LPM for small launches: one block per BX x BY x BZ box of sites, no z
// march.  The same operator and arithmetic as ku3d_kernel (coarse3d/lin3d.py:98-149): u of the box + 2 halo sites
// goes to LDS, the shell sums of the box + 1 halo site are formed into LDS (three sites per thread), and each thread
// then forms K u of its own site.  A marching launch over a small lattice with few systems has fewer blocks than the
// chip has CUs, each walking its planes one dependent load / LDS / barrier chain at a time; the box kernel runs one
// such chain per block over about 16 times more blocks, at the price of reading u about 5 times and Tv about 2.5
// times per site through the caches, so it is chosen only for launches the marching kernel cannot spread over the
// chip.  Face-diagonal bonds use e = sqrt2 dL and sqrt2-scaled shell sums, as the marching kernel does.
#pragma once
#include "ku3d.hpp"

namespace lpm {

constexpr int kBoxX = 16, kBoxY = 4, kBoxZ = 4;  // own sites of a block (256 threads)

template <typename T, typename C, int MODE, int EPI, bool BONDS>
__global__ __launch_bounds__(kBoxX* kBoxY* kBoxZ) void ku3b_kernel(Ku3Args<T, C> a) {
  constexpr int BX = kBoxX, BY = kBoxY, BZ = kBoxZ, NT = BX * BY * BZ;
  constexpr int HX = BX + 4, HY = BY + 4, HZ = BZ + 4;  // u with a 2-site halo
  constexpr int SX = BX + 2, SY = BY + 2, SZ = BZ + 2;  // shell sums with a 1-site halo
  constexpr int NU = HX * HY * HZ, NS = SX * SY * SZ;
  constexpr int LU = (NU + NT - 1) / NT, LS = (NS + NT - 1) / NT;
  __shared__ T su[3][HZ][HY][HX];
  __shared__ T ss[4][SZ][SY][SX];  // S0, sqrt2 S1, TS0, sqrt2 TS1
  __shared__ T red[NT / 64];

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t nxy = (int64_t)nx * ny;
  const int bid = blockIdx.x;
  const int tix = bid % a.ntx, tiy = (bid / a.ntx) % a.nty, tiz = bid / (a.ntx * a.nty);
  const int x0 = tix * BX, y0 = tiy * BY, z0 = tiz * BZ;
  const int l = blockIdx.y, b = blockIdx.z;
  if (a.active && !a.active[((int64_t)b * a.nl + l) * a.act_stride]) return;
  const int tid = threadIdx.x;
  const int64_t ps = a.ps;
  const T* __restrict__ u = a.u + a.bu.at(b, l);
  const C* __restrict__ kn = a.kn + a.bc.at(b, l);
  const C* __restrict__ tv = a.tv + a.bc.at(b, l);
  const uint32_t* __restrict__ bwd = BONDS ? a.bonds + a.bb.at(b, l) : nullptr;
  // coordinates are clamped into the box for addressing; results of halo sites outside the box are never read
  // through an existing bond
  auto cl = [](int v, int n) { return min(max(v, 0), n - 1); };
  auto site = [&](int x, int y, int z) { return ((int64_t)z * ny + y) * nx + x; };
  // the 18 coefficients of site p (packed odd slots: the forward plane's entry at the neighbour, inside the guards)
  auto coef18 = [&](const C* __restrict__ base, int64_t p, T (&out)[18]) {
    for_slots(Slots18{}, [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      const int64_t q = (K & 1) ? p + off3(K, 0) + (int64_t)off3(K, 1) * nx + (int64_t)off3(K, 2) * nxy : p;
      out[K] = (T)coef_at<C, MODE, 18>(base, ps, p, q, K);
    });
  };

  // u of the box + 2 halo sites
#pragma unroll
  for (int j = 0; j < LU; ++j) {
    const int i = tid + j * NT;
    if (j + 1 < LU || i < NU) {
      const int hx = i % HX, hy = (i / HX) % HY, hz = i / (HX * HY);
      const int64_t p = site(cl(x0 - 2 + hx, nx), cl(y0 - 2 + hy, ny), cl(z0 - 2 + hz, nz));
      const Vec3<T> v = *(const Vec3<T>*)(u + 3 * p);
      su[0][hz][hy][hx] = v.x;
      su[1][hz][hy][hx] = v.y;
      su[2][hz][hy][hx] = v.z;
    }
  }
  __syncthreads();

  // dL (axial) or e = sqrt2 dL (diagonal) of slot K at the u position (hx, hy, hz)
  auto dle = [&](auto kc, int hx, int hy, int hz, T ux, T uy, T uz) -> T {
    constexpr int K = decltype(kc)::value;
    constexpr int ox = off3(K, 0), oy = off3(K, 1), oz = off3(K, 2);
    T d = T(0);
    if constexpr (ox != 0) d += (ox > 0 ? T(-1) : T(1)) * (ux - su[0][hz + oz][hy + oy][hx + ox]);
    if constexpr (oy != 0) d += (oy > 0 ? T(-1) : T(1)) * (uy - su[1][hz + oz][hy + oy][hx + ox]);
    if constexpr (oz != 0) d += (oz > 0 ? T(-1) : T(1)) * (uz - su[2][hz + oz][hy + oy][hx + ox]);
    return d;
  };

  // shell sums of the box + 1 halo site
#pragma unroll
  for (int j = 0; j < LS; ++j) {
    const int i = tid + j * NT;
    if (j + 1 < LS || i < NS) {
      const int sx = i % SX, sy = (i / SX) % SY, sz = i / (SX * SY);
      const int gx = cl(x0 - 1 + sx, nx), gy = cl(y0 - 1 + sy, ny), gz = cl(z0 - 1 + sz, nz);
      const int64_t p = site(gx, gy, gz);
      T tvv[18];
      coef18(tv, p, tvv);
      const uint32_t wv = BONDS ? bwd[p] : 0xffffffffu;
      const bool mx = gx > 0, px = gx < nx - 1, my = gy > 0, py = gy < ny - 1, mz = gz > 0, pz = gz < nz - 1;
      const int hx = sx + 1, hy = sy + 1, hz = sz + 1;
      const T ux = su[0][hz][hy][hx], uy = su[1][hz][hy][hx], uz = su[2][hz][hy][hx];
      T s0 = T(0), s1 = T(0), t0 = T(0), t1 = T(0);
      for_slots(Slots18{}, [&](auto kc) {
        constexpr int K = decltype(kc)::value;
        T d = dle(kc, hx, hy, hz, ux, uy, uz);
        bool ok = slot_ok<K>(mx, px, my, py, mz, pz);
        if constexpr (BONDS) ok = ok && ((wv >> K) & 1u);
        d = ok ? d : T(0);
        if constexpr (shell3(K) == 0) {
          s0 += d;
          t0 += tvv[K] * d;
        } else {
          s1 += d;
          t1 += tvv[K] * d;
        }
      });
      ss[0][sz][sy][sx] = s0;
      ss[1][sz][sy][sx] = s1;
      ss[2][sz][sy][sx] = t0;
      ss[3][sz][sy][sx] = t1;
    }
  }
  __syncthreads();

  // K u of the own site
  const int tx = tid % BX, ty = (tid / BX) % BY, tz = tid / (BX * BY);
  const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
  T acc = T(0);
  if (gx < nx && gy < ny && gz < nz) {
    const int64_t p = site(gx, gy, gz);
    T knv[18], tvv[18];
    coef18(kn, p, knv);
    coef18(tv, p, tvv);
    const uint32_t wv = BONDS ? bwd[p] : 0xffffffffu;
    const bool mx = gx > 0, px = gx < nx - 1, my = gy > 0, py = gy < ny - 1, mz = gz > 0, pz = gz < nz - 1;
    const int hx = tx + 2, hy = ty + 2, hz = tz + 2, sx = tx + 1, sy = ty + 1, sz = tz + 1;
    const T ux = su[0][hz][hy][hx], uy = su[1][hz][hy][hx], uz = su[2][hz][hy][hx];
    const T si0 = ss[0][sz][sy][sx], si1 = ss[1][sz][sy][sx], ti0 = ss[2][sz][sy][sx], ti1 = ss[3][sz][sy][sx];
    T fx = T(0), fy = T(0), fz = T(0);
    for_slots(Slots18{}, [&](auto kc) {
      constexpr int K = decltype(kc)::value;
      constexpr int ox = off3(K, 0), oy = off3(K, 1), oz = off3(K, 2);
      const T d = dle(kc, hx, hy, hz, ux, uy, uz);
      T g;
      if constexpr (shell3(K) == 0) {
        // F = 2 Kn dL + 1/2 (TS_i + TS_q) + 1/2 Tv (S_i + S_q)
        const T t = tvv[K] * (si0 + ss[0][sz + oz][sy + oy][sx + ox]) + (ti0 + ss[2][sz + oz][sy + oy][sx + ox]);
        g = T(2) * knv[K] * d + T(0.5) * t;
      } else {
        // G = F / sqrt2 = Kn e + 1/4 (TS'_i + TS'_q) + 1/4 Tv (S'_i + S'_q)   (primes: sqrt2-scaled sums)
        const T t = tvv[K] * (si1 + ss[1][sz + oz][sy + oy][sx + ox]) + (ti1 + ss[3][sz + oz][sy + oy][sx + ox]);
        g = knv[K] * d + T(0.25) * t;
      }
      bool ok = slot_ok<K>(mx, px, my, py, mz, pz);
      if constexpr (BONDS) ok = ok && ((wv >> K) & 1u);
      g = ok ? g : T(0);
      // (K u)_i += F n_k, n_k = -off_k / |off_k|; for diagonal slots F n_k = -G off_k
      if constexpr (ox != 0) fx += (ox > 0 ? -g : g);
      if constexpr (oy != 0) fy += (oy > 0 ? -g : g);
      if constexpr (oz != 0) fz += (oz > 0 ? -g : g);
    });
    T* __restrict__ o = a.out + a.bo.at(b, l) + 3 * p;
    if constexpr (EPI == EPI_MASKDOT) {
      acc = ux * fx + uy * fy + uz * fz;
    } else if constexpr (EPI == EPI_RESID) {
      T m0 = T(1), m1 = T(1), m2 = T(1);
      if (a.mask) {
        const uint8_t* mk = a.mask + a.bm.at(b, l) + 3 * p;
        m0 = mk[0] ? T(1) : T(0);
        m1 = mk[1] ? T(1) : T(0);
        m2 = mk[2] ? T(1) : T(0);
      }
      const T* rh = a.rhs + a.brhs.at(b, l) + 3 * p;
      fx = m0 * (rh[0] - fx);
      fy = m1 * (rh[1] - fy);
      fz = m2 * (rh[2] - fz);
      acc = fx * fx + fy * fy + fz * fz;
    }
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
  }
  if constexpr (EPI != EPI_PLAIN) {
    const T s = block_sum(acc, red);
    if (tid == 0) a.partial[((int64_t)b * a.nl + l) * gridDim.x + bid] = s;
  }
}

}  // namespace lpm
