// Serving: many independent 2D solves of different lattice sizes advanced together by one persistent kernel.
//
// Every admitted request owns a slot: a descriptor plus its arrays in a device arena (coefficients, mask, prescribed
// displacement, FP64 solution and right-hand side, FP32 CG vectors).  One workgroup drives one slot through the
// whole mixed-precision solve (coarse2d/torch_lpm.py:216-229 replaced by a matrix-free block-Jacobi PCG):
//   INIT   b = -mask * K u_p (FP64), ||b||, 2x2 block-Jacobi inverse M (free DOFs only);
//   OUTER  r64 = mask * (b - K x) in FP64; stop when ||r64|| <= tol ||b||; else start an FP32 inner PCG on
//          K d = r64 with the relative tolerance min(1/2, max(inner_tol, tol / (2 rel)));
//   INNER  FP32 PCG iterations (shell sums and product in one banded pass through LDS); on inner convergence
//          x += d (FP64) and back to OUTER;
//   DONE   u = u_p + x written over u_p.
// The workgroup runs until its slot is done or the launch's work budget (site-iterations) is spent, then leaves
// its state in the descriptor; the next launch resumes it.  All control values are workgroup-uniform (computed
// from broadcast reductions), so every barrier is reached by every wave.  The arithmetic of a slot does not depend
// on the other slots nor on where launches split the work, so a request's result is bit-identical alone and in a
// mix.  The operator is the force form of lpm::ku (coarse2d/lin2d.py:121-141) with directed FP32 coefficients.
#pragma once
#include "lpm_common.hpp"

namespace lpm {

enum SlotState : int { SLOT_FREE = 0, SLOT_INIT = 1, SLOT_OUTER = 2, SLOT_INNER = 3, SLOT_DONE = 4, SLOT_FAILED = 5 };

struct SlotDesc {
  // request (written at admission)
  int nx, ny;
  float h;
  int maxit;        // total inner iterations allowed
  double tol, inner_tol;
  const float* kn;  // (N, 8) directed
  const float* tv;
  const uint8_t* mask;  // (N, 2) 1 = free
  double* up;           // (N, 2) prescribed displacement; the solution on completion
  double* x;            // (N, 2) FP64 correction
  double* b;            // (N, 2)
  double* w64;          // (N, 4) FP64 shell sums
  float* minv;          // (N, 3) 2x2 symmetric inverse blocks
  float *d, *r, *z, *p, *ap;  // (N, 2)
  float* w32;                 // (N, 4)
  // state (kernel-owned after admission)
  int state;
  int it;        // inner iterations of the current inner solve
  int total;     // inner iterations so far
  int outer;     // refinement steps so far
  double bnorm, rel, thr2;  // ||b||, ||r64|| / ||b||, squared inner stopping threshold
  float rz;
  int pad_;
};

namespace srv {

constexpr int NT_MAX = 1024;  // threads per slot workgroup: 256, 512 or 1024 (blockDim.x), fixed per server
// Fused shell-sum + product pass of the inner iteration: bands of B rows with B + 2 <= BAND_SPT blockDim.x / nx, so
// every thread owns at most BAND_SPT shell sites of a band; LDS holds p of the band rows + 2 each side and the shell
// sums of the band rows + 1 each side.  Lattices wider than BAND_NX_MAX take the two-pass path.
constexpr int BAND_SPT = 3;
constexpr int BAND_NX_MAX = BAND_SPT * NT_MAX / 3;
constexpr int BAND_W_MAX = BAND_SPT * NT_MAX;                // shell sites of a band
constexpr int BAND_P_MAX = BAND_W_MAX + 2 * BAND_NX_MAX;     // p sites of a band

// directed slot geometry of the 2D square lattice
__device__ __forceinline__ int nbr2(int x, int y, int nx, int ny, int k) {
  const int qx = x + off2(k, 0), qy = y + off2(k, 1);
  return (qx < 0 || qx >= nx || qy < 0 || qy >= ny) ? -1 : qy * nx + qx;
}

// sum of two values over the workgroup, broadcast to every thread (fixed order)
template <typename T>
__device__ __forceinline__ void wg_sum2(T& a, T& b, T* red) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[2 * w] = a;
    red[2 * w + 1] = b;
  }
  __syncthreads();
  T sa = T(0), sb = T(0);
  const int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; ++i) {
    sa += red[2 * i];
    sb += red[2 * i + 1];
  }
  a = sa;
  b = sb;
  __syncthreads();
}

// aligned short vectors for the per-site accesses (the arena arrays are 256-byte aligned)
template <typename T, int N>
struct alignas(sizeof(T) * N) Vec {
  T v[N];
};
template <typename T, int N>
__device__ __forceinline__ Vec<T, N> vld(const T* __restrict__ p) {
  return *(const Vec<T, N>*)p;
}

// shell sums (S0, S1, TS0, TS1) of u at every site into w
template <typename T>
__device__ void shell_pass(const SlotDesc& s, const T* __restrict__ u, T* __restrict__ w) {
  const int N = s.nx * s.ny, nx = s.nx, ny = s.ny;
  const float* __restrict__ tvp = s.tv;
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    const int x = p % nx, y = p / nx;
    const Vec<T, 2> ui = vld<T, 2>(u + 2 * p);
    const Vec<float, 4> t0 = vld<float, 4>(tvp + 8 * p), t1 = vld<float, 4>(tvp + 8 * p + 4);
    const float tvk[8] = {t0.v[0], t0.v[1], t0.v[2], t0.v[3], t1.v[0], t1.v[1], t1.v[2], t1.v[3]};
    Vec<T, 4> a;
    a.v[0] = a.v[1] = a.v[2] = a.v[3] = T(0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int q = nbr2(x, y, nx, ny, k);
      if (q < 0) continue;
      const Vec<T, 2> uq = vld<T, 2>(u + 2 * q);
      const T dl = nd2<T>(k, 0) * (ui.v[0] - uq.v[0]) + nd2<T>(k, 1) * (ui.v[1] - uq.v[1]);
      a.v[shell2(k)] += dl;
      a.v[2 + shell2(k)] += (T)tvk[k] * dl;
    }
    *(Vec<T, 4>*)(w + 4 * p) = a;
  }
}

// (K u)_p from u and its shell sums w
template <typename T>
__device__ __forceinline__ void ku_site(const SlotDesc& s, const T* __restrict__ u, const T* __restrict__ w, int p,
                                        T& f0, T& f1) {
  const int nx = s.nx, ny = s.ny;
  const int x = p % nx, y = p / nx;
  const float* __restrict__ knp = s.kn;
  const float* __restrict__ tvp = s.tv;
  const Vec<T, 2> ui = vld<T, 2>(u + 2 * p);
  const Vec<T, 4> wi = vld<T, 4>(w + 4 * p);
  const Vec<float, 4> k0 = vld<float, 4>(knp + 8 * p), k1 = vld<float, 4>(knp + 8 * p + 4);
  const Vec<float, 4> t0 = vld<float, 4>(tvp + 8 * p), t1 = vld<float, 4>(tvp + 8 * p + 4);
  const float knk[8] = {k0.v[0], k0.v[1], k0.v[2], k0.v[3], k1.v[0], k1.v[1], k1.v[2], k1.v[3]};
  const float tvk[8] = {t0.v[0], t0.v[1], t0.v[2], t0.v[3], t1.v[0], t1.v[1], t1.v[2], t1.v[3]};
  f0 = f1 = T(0);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int q = nbr2(x, y, nx, ny, k);
    if (q < 0) continue;
    const int c = shell2(k);
    const Vec<T, 2> uq = vld<T, 2>(u + 2 * q);
    const Vec<T, 4> wq = vld<T, 4>(w + 4 * q);
    const T dl = nd2<T>(k, 0) * (ui.v[0] - uq.v[0]) + nd2<T>(k, 1) * (ui.v[1] - uq.v[1]);
    const T tv = (T)tvk[k];
    const T F = T(2) * (T)knk[k] * dl + T(0.5) * (wi.v[2 + c] + wq.v[2 + c]) + T(0.5) * tv * (wi.v[c] + wq.v[c]);
    f0 += F * nd2<T>(k, 0);
    f1 += F * nd2<T>(k, 1);
  }
}

__device__ __forceinline__ float mfree(const SlotDesc& s, int i) { return s.mask[i] ? 1.0f : 0.0f; }

// INIT: b, ||b||, block-Jacobi inverse
__device__ void phase_init(SlotDesc& s, double* red) {
  const int N = s.nx * s.ny;
  shell_pass<double>(s, s.up, s.w64);
  __syncthreads();
  double bb = 0.0, unused = 0.0;
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    double f0, f1;
    ku_site<double>(s, s.up, s.w64, p, f0, f1);
    const double b0 = -(double)mfree(s, 2 * p) * f0, b1 = -(double)mfree(s, 2 * p + 1) * f1;
    s.b[2 * p] = b0;
    s.b[2 * p + 1] = b1;
    s.x[2 * p] = 0.0;
    s.x[2 * p + 1] = 0.0;
    bb += b0 * b0 + b1 * b1;
    // diagonal block D_p = sum_k (Kn_pk + Kn_qk' + Tv_qk') n n^T + sum_c 1/2 (T_c s_c^T + s_c T_c^T)
    const int x = p % s.nx, y = p / s.nx;
    double m00 = 0, m01 = 0, m11 = 0, T[2][2] = {{0, 0}, {0, 0}}, S[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int q = nbr2(x, y, s.nx, s.ny, k);
      if (q < 0) continue;
      const double n0 = nd2<double>(k, 0), n1 = nd2<double>(k, 1);
      const double cf = (double)s.kn[8 * p + k] + (double)s.kn[8 * q + (k ^ 1)] + (double)s.tv[8 * q + (k ^ 1)];
      m00 += cf * n0 * n0;
      m01 += cf * n0 * n1;
      m11 += cf * n1 * n1;
      const int c = shell2(k);
      const double t = (double)s.tv[8 * p + k];
      T[c][0] += t * n0;
      T[c][1] += t * n1;
      S[c][0] += n0;
      S[c][1] += n1;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      m00 += T[c][0] * S[c][0];
      m11 += T[c][1] * S[c][1];
      m01 += 0.5 * (T[c][0] * S[c][1] + S[c][0] * T[c][1]);
    }
    const bool fr0 = s.mask[2 * p] != 0, fr1 = s.mask[2 * p + 1] != 0;
    float i00 = 0.f, i01 = 0.f, i11 = 0.f;
    if (fr0 && fr1) {
      const double det = m00 * m11 - m01 * m01;
      i00 = (float)(m11 / det);
      i11 = (float)(m00 / det);
      i01 = (float)(-m01 / det);
    } else if (fr0) {
      i00 = (float)(1.0 / m00);
    } else if (fr1) {
      i11 = (float)(1.0 / m11);
    }
    s.minv[3 * p] = i00;
    s.minv[3 * p + 1] = i11;
    s.minv[3 * p + 2] = i01;
  }
  wg_sum2(bb, unused, red);
  s.bnorm = sqrt(bb);
  s.state = s.bnorm > 0.0 ? SLOT_OUTER : SLOT_DONE;
}

// OUTER: true residual; either finish or start an inner solve
__device__ void phase_outer(SlotDesc& s, double* red) {
  const int N = s.nx * s.ny;
  shell_pass<double>(s, s.x, s.w64);
  __syncthreads();
  double rr = 0.0, unused = 0.0;
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    double f0, f1;
    ku_site<double>(s, s.x, s.w64, p, f0, f1);
    const double r0 = (double)mfree(s, 2 * p) * (s.b[2 * p] - f0);
    const double r1 = (double)mfree(s, 2 * p + 1) * (s.b[2 * p + 1] - f1);
    s.r[2 * p] = (float)r0;
    s.r[2 * p + 1] = (float)r1;
    rr += r0 * r0 + r1 * r1;
  }
  wg_sum2(rr, unused, red);
  s.rel = sqrt(rr) / s.bnorm;
  if (s.rel <= s.tol || s.total >= s.maxit) {
    s.state = s.rel <= s.tol ? SLOT_DONE : SLOT_FAILED;
    return;
  }
  const double itol = fmin(0.5, fmax(s.inner_tol, 0.5 * s.tol / s.rel));
  s.thr2 = itol * itol * rr;
  double rz = 0.0;
  unused = 0.0;
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    const float* m = s.minv + 3 * p;
    const float r0 = s.r[2 * p], r1 = s.r[2 * p + 1];
    const float z0 = m[0] * r0 + m[2] * r1, z1 = m[2] * r0 + m[1] * r1;
    s.z[2 * p] = z0;
    s.z[2 * p + 1] = z1;
    s.p[2 * p] = z0;
    s.p[2 * p + 1] = z1;
    s.d[2 * p] = 0.f;
    s.d[2 * p + 1] = 0.f;
    rz += (double)(r0 * z0 + r1 * z1);
  }
  wg_sum2(rz, unused, red);
  s.rz = (float)rz;
  s.it = 0;
  s.outer += 1;
  s.state = SLOT_INNER;
}

// ap = mask (K p) and the partial p.ap in one pass over bands of B rows (the operator of shell_pass + ku_site, same
// arithmetic): p of rows y0 - 2 .. y0 + nb + 1 into LDS, shell sums of rows y0 - 1 .. y0 + nb into LDS by thread t
// for shell sites t + j blockDim.x, which keep their Tv in registers, then the product of the band rows by the
// threads that formed their shell sums.  Every array of the slot is read once per iteration except the halo rows.
__device__ float band_product(const SlotDesc& s, const float* __restrict__ pv, float* __restrict__ apv,
                              Vec<float, 2>* pL, Vec<float, 4>* wL, int B) {
  const int nx = s.nx, ny = s.ny, nt = blockDim.x;
  const float* __restrict__ knp = s.kn;
  const float* __restrict__ tvp = s.tv;
  const uint8_t* __restrict__ mk = s.mask;
  float pap = 0.f;
  for (int y0 = 0; y0 < ny; y0 += B) {
    const int nb = min(B, ny - y0);
    for (int t = threadIdx.x; t < (nb + 4) * nx; t += nt) {
      const int r = t / nx, gy = y0 - 2 + r;
      Vec<float, 2> v;
      v.v[0] = v.v[1] = 0.f;
      if (gy >= 0 && gy < ny) v = vld<float, 2>(pv + 2 * ((size_t)gy * nx + (t - r * nx)));
      pL[t] = v;
    }
    __syncthreads();
    float tvr[BAND_SPT][8];
#pragma unroll
    for (int j = 0; j < BAND_SPT; ++j) {
      const int i = threadIdx.x + j * nt;  // shell site: row i / nx of rows y0 - 1 .., LDS p row i / nx + 1
      const int r = i / nx, x = i - r * nx, gy = y0 - 1 + r;
      if (i < (nb + 2) * nx && gy >= 0 && gy < ny) {
        const int p = gy * nx + x;
        const Vec<float, 4> t0 = vld<float, 4>(tvp + 8 * (size_t)p), t1 = vld<float, 4>(tvp + 8 * (size_t)p + 4);
        tvr[j][0] = t0.v[0], tvr[j][1] = t0.v[1], tvr[j][2] = t0.v[2], tvr[j][3] = t0.v[3];
        tvr[j][4] = t1.v[0], tvr[j][5] = t1.v[1], tvr[j][6] = t1.v[2], tvr[j][7] = t1.v[3];
        const Vec<float, 2> ui = pL[i + nx];
        Vec<float, 4> a;
        a.v[0] = a.v[1] = a.v[2] = a.v[3] = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (nbr2(x, gy, nx, ny, k) < 0) continue;
          const Vec<float, 2> uq = pL[i + nx + off2(k, 1) * nx + off2(k, 0)];
          const float dl = nd2<float>(k, 0) * (ui.v[0] - uq.v[0]) + nd2<float>(k, 1) * (ui.v[1] - uq.v[1]);
          a.v[shell2(k)] += dl;
          a.v[2 + shell2(k)] += tvr[j][k] * dl;
        }
        wL[i] = a;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < BAND_SPT; ++j) {
      const int i = threadIdx.x + j * nt;
      const int r = i / nx, x = i - r * nx, gy = y0 - 1 + r;
      if (i < (nb + 2) * nx && r >= 1 && r <= nb) {
        const int p = gy * nx + x;
        const Vec<float, 4> k0 = vld<float, 4>(knp + 8 * (size_t)p), k1 = vld<float, 4>(knp + 8 * (size_t)p + 4);
        const float knk[8] = {k0.v[0], k0.v[1], k0.v[2], k0.v[3], k1.v[0], k1.v[1], k1.v[2], k1.v[3]};
        const Vec<float, 2> ui = pL[i + nx];
        const Vec<float, 4> wi = wL[i];
        float f0 = 0.f, f1 = 0.f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          if (nbr2(x, gy, nx, ny, k) < 0) continue;
          const int c = shell2(k), dq = off2(k, 1) * nx + off2(k, 0);
          const Vec<float, 2> uq = pL[i + nx + dq];
          const Vec<float, 4> wq = wL[i + dq];
          const float dl = nd2<float>(k, 0) * (ui.v[0] - uq.v[0]) + nd2<float>(k, 1) * (ui.v[1] - uq.v[1]);
          const float F = 2.f * knk[k] * dl + 0.5f * (wi.v[2 + c] + wq.v[2 + c]) + 0.5f * tvr[j][k] * (wi.v[c] + wq.v[c]);
          f0 += F * nd2<float>(k, 0);
          f1 += F * nd2<float>(k, 1);
        }
        Vec<float, 2> o;
        o.v[0] = mk[2 * p] ? f0 : 0.f;
        o.v[1] = mk[2 * p + 1] ? f1 : 0.f;
        *(Vec<float, 2>*)(apv + 2 * (size_t)p) = o;
        pap += ui.v[0] * o.v[0] + ui.v[1] * o.v[1];
      }
    }
    __syncthreads();
  }
  return pap;
}

// rows per band of the fused pass for this slot and workgroup size (0: the lattice is too wide, two passes)
__device__ __forceinline__ int band_rows(int nx) {
  if (nx > BAND_NX_MAX) return 0;
  const int b = BAND_SPT * (int)blockDim.x / nx - 2;
  return b >= 1 ? b : 0;
}

// one inner PCG iteration (FP32); fused = 0 forces the two-pass product
__device__ void inner_iteration(SlotDesc& s, float* red, Vec<float, 2>* pL, Vec<float, 4>* wL, int fused) {
  const int N = s.nx * s.ny;
  float* __restrict__ pv = s.p;
  float* __restrict__ apv = s.ap;
  float* __restrict__ rv = s.r;
  float* __restrict__ zv = s.z;
  float* __restrict__ dv = s.d;
  const float* __restrict__ mi = s.minv;
  const uint8_t* __restrict__ mk = s.mask;
  float pap = 0.f, unused = 0.f;
  const int B = fused ? band_rows(s.nx) : 0;
  if (B > 0) {
    pap = band_product(s, pv, apv, pL, wL, B);
  } else {
    shell_pass<float>(s, pv, s.w32);
    __syncthreads();
    for (int p = threadIdx.x; p < N; p += blockDim.x) {
      float f0, f1;
      ku_site<float>(s, pv, s.w32, p, f0, f1);
      Vec<float, 2> o;
      o.v[0] = mk[2 * p] ? f0 : 0.f;
      o.v[1] = mk[2 * p + 1] ? f1 : 0.f;
      *(Vec<float, 2>*)(apv + 2 * p) = o;
      const Vec<float, 2> pp = vld<float, 2>(pv + 2 * p);
      pap += pp.v[0] * o.v[0] + pp.v[1] * o.v[1];
    }
  }
  wg_sum2(pap, unused, red);
  if (!(pap > 0.f)) {
    s.state = SLOT_FAILED;
    return;
  }
  const float alpha = s.rz / pap;
  float rz = 0.f, rr = 0.f;
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    const int i = 2 * p;
    const Vec<float, 2> pp = vld<float, 2>(pv + i), aa = vld<float, 2>(apv + i), r_ = vld<float, 2>(rv + i);
    Vec<float, 2> dd = vld<float, 2>(dv + i);
    const float m0 = mi[3 * p], m1 = mi[3 * p + 1], m2 = mi[3 * p + 2];
    dd.v[0] += alpha * pp.v[0];
    dd.v[1] += alpha * pp.v[1];
    *(Vec<float, 2>*)(dv + i) = dd;
    Vec<float, 2> r2, z2;
    r2.v[0] = r_.v[0] - alpha * aa.v[0];
    r2.v[1] = r_.v[1] - alpha * aa.v[1];
    *(Vec<float, 2>*)(rv + i) = r2;
    z2.v[0] = m0 * r2.v[0] + m2 * r2.v[1];
    z2.v[1] = m2 * r2.v[0] + m1 * r2.v[1];
    *(Vec<float, 2>*)(zv + i) = z2;
    rz += r2.v[0] * z2.v[0] + r2.v[1] * z2.v[1];
    rr += r2.v[0] * r2.v[0] + r2.v[1] * r2.v[1];
  }
  wg_sum2(rz, rr, red);
  s.it += 1;
  s.total += 1;
  if ((double)rr <= s.thr2 || s.total >= s.maxit) {
    for (int p = threadIdx.x; p < 2 * N; p += blockDim.x) s.x[p] += (double)dv[p];
    __syncthreads();
    s.state = SLOT_OUTER;
    return;
  }
  const float beta = rz / s.rz;
  s.rz = rz;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    Vec<float, 2> pp = vld<float, 2>(pv + 2 * i);
    const Vec<float, 2> z2 = vld<float, 2>(zv + 2 * i);
    pp.v[0] = z2.v[0] + beta * pp.v[0];
    pp.v[1] = z2.v[1] + beta * pp.v[1];
    *(Vec<float, 2>*)(pv + 2 * i) = pp;
  }
  __syncthreads();
}

}  // namespace srv

// One workgroup per slot (blockDim.x a multiple of 64, at most 1024); each advances its slot by at most `budget`
// site-iterations.  Slot descriptors live in
// device memory; the workgroup keeps a private copy (identical in all threads) and thread 0 writes it back.
__global__ __launch_bounds__(srv::NT_MAX) void serve_kernel(SlotDesc* __restrict__ slots, int nslots,
                                                             long long budget, int fused) {
  __shared__ double red64[2 * srv::NT_MAX / 64];
  __shared__ float red32[2 * srv::NT_MAX / 64];
  __shared__ srv::Vec<float, 2> band_p[srv::BAND_P_MAX];
  __shared__ srv::Vec<float, 4> band_w[srv::BAND_W_MAX];
  const int id = blockIdx.x;
  if (id >= nslots) return;
  SlotDesc s = slots[id];
  if (s.state == SLOT_FREE || s.state == SLOT_DONE || s.state == SLOT_FAILED) return;
  const long long N = (long long)s.nx * s.ny;
  long long work = 0;
  while (work < budget) {
    if (s.state == SLOT_INIT) {
      srv::phase_init(s, red64);
      work += 3 * N;
    } else if (s.state == SLOT_OUTER) {
      srv::phase_outer(s, red64);
      work += 3 * N;
    } else if (s.state == SLOT_INNER) {
      srv::inner_iteration(s, red32, band_p, band_w, fused);
      work += N;
    } else {
      break;
    }
  }
  if (s.state == SLOT_DONE) {
    for (int i = threadIdx.x; i < 2 * N; i += blockDim.x) s.up[i] += s.x[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    SlotDesc* o = slots + id;
    o->state = s.state;
    o->it = s.it;
    o->total = s.total;
    o->outer = s.outer;
    o->bnorm = s.bnorm;
    o->rel = s.rel;
    o->thr2 = s.thr2;
    o->rz = s.rz;
  }
}

}  // namespace lpm
