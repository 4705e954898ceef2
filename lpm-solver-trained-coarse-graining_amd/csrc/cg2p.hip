// Single-launch plain FP64 CG of one 2D square-lattice system (timing/t2d/cg2plain.py:11-49): unpreconditioned,
// per-DOF mask, zero start, ||r|| <= tol ||b||.  One cooperative launch runs the whole solve.
//
// Workgroup w owns the rows [w R, min((w + 1) R, ny)) of the lattice.  Its own particles keep x, r, p and their
// directed FP64 Kn, Tv in registers; LDS holds p on the rows + two rows each side, the shell sums on the rows + one
// row each side and the Tv of those two side rows.  An iteration has two grid-wide exchanges, one per reduction
// (p.Ap, r.r).  An exchange is also the barrier: each workgroup publishes its partial sum as epoch-tagged words with
// device-coherent stores once its earlier stores have completed, and every wave polls the words of all workgroups
// and sums the payloads in one fixed order, so every workgroup takes the same decisions from bit-identical scalars.
// No cache writeback or invalidation is needed: every array that crosses workgroups (p_0, r, the words) is read and
// written only with device-coherent accesses.  r is published as epoch-tagged words too (two per double, buffered
// by iteration parity), and the neighbours' halo r is polled in the same loop as the partial sums of r.r, so it
// arrives with that exchange instead of a load round trip after it.  The new p of the two halo rows each side is
// recomputed locally as fma(beta, p_old, r) from that r and the previous p in LDS, which is the owner's own
// arithmetic, so no third exchange is needed before the next K p.  Every wait is bounded: on time-out the solve
// stops with an error.
#include <algorithm>
#include <cmath>

#include "capi_common.hpp"
#include "lpm_common.hpp"

using namespace lpm;
using lpmc::fail;

namespace {

constexpr int kCgpMaxWaves = 16;  // waves of the largest workgroup (1024 threads)
constexpr int kCgpMaxWg = 64;     // one lane of a wave per workgroup in the gather
constexpr int kCgpAutoWg = 48;    // workgroups requested by the automatic launch shape
constexpr long kSpinLimit = 1L << 22;  // polls of one exchange (each with s_sleep, seconds in all) before giving up

enum CgpStatus : int { CGP_CONV = 0, CGP_MAXIT = 1, CGP_NOTSPD = 2, CGP_NONFINITE = 3, CGP_TIMEOUT = 4 };

struct CgpArgs {
  int nx, ny, rows, nwg;
  const double* kn;  // (N, 8) directed
  const double* tv;
  const uint8_t* mask;  // (N, 2), 1 = free
  const double* b;      // (N, 2)
  double* x;            // (N, 2) out
  double* pbuf;         // (N, 2) p_0 (exchanged once)
  uint64_t* rwords;     // (2, N, 2, 2) r by iteration parity, each double as two words (epoch << 32) | half
  uint64_t* slots;      // (2, kCgpMaxWg, 4) published partial sums: (epoch << 32) | half of a double (zero at launch)
  int* abort_flag;      // set by a workgroup whose exchange wait timed out (zero at launch)
  double tol;
  int maxit;
  int* out;  // [iterations, status]
  double* relres;
};

// LDS layout (doubles): p on the rows + 2 each side, shell sums on the rows + 1 each side, Tv of the two rows next to
// the block, two reduction buffers (per-wave partial pairs and per-wave flags)
struct CgpLds {
  int p, w, tv, red, total;
  static constexpr int kRed = 3 * kCgpMaxWaves;
  __host__ __device__ CgpLds(int nx, int R) {
    p = 0;                      // (R + 4, nx, 2)
    w = p + (R + 4) * nx * 2;   // (R + 2, nx, 4)
    tv = w + (R + 2) * nx * 4;  // (2, nx, 8): rows y0 - 1 and y1
    red = tv + 2 * nx * 8;      // (2, kRed)
    total = red + 2 * kRed;
  }
};

// accesses of the exchanged arrays: device-coherent loads and stores, no cache maintenance
__device__ __forceinline__ void xst(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double xld(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide sums of (v0, v1) that also act as a barrier.  The workgroup's sums (fixed order) are published by thread
// 0 as four tagged words after every thread's earlier stores have completed (s_waitcnt); each wave then polls the
// words of all workgroups, lane j those of workgroup j, until every tag equals `epoch`, and sums the payloads over
// the lanes in a fixed order, so every wave of every workgroup obtains the same bits.  Returns false (in every
// thread of the workgroup) when a wait timed out or another workgroup gave up.
template <int NT, bool WAIT, typename Extra>
__device__ bool exchange(const CgpArgs& a, unsigned epoch, double v0, double v1, double* red, double& s0,
                         double& s1, Extra&& extra_poll) {
  constexpr int NW = NT / 64;
  v0 = wave_sum(v0);
  v1 = wave_sum(v1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[2 * wv] = v0;
    red[2 * wv + 1] = v1;
  }
  if constexpr (WAIT) __builtin_amdgcn_s_waitcnt(0);  // this wave's stores of untagged exchanged data completed
  __syncthreads();
  uint64_t* slot = a.slots + (size_t)(epoch & 1u) * kCgpMaxWg * 4;
  if (threadIdx.x == 0) {
    double t0 = 0.0, t1 = 0.0;
    for (int i = 0; i < NW; ++i) {
      t0 += red[2 * i];
      t1 += red[2 * i + 1];
    }
    const uint64_t tag = (uint64_t)epoch << 32;
    const uint64_t b0 = (uint64_t)__double_as_longlong(t0), b1 = (uint64_t)__double_as_longlong(t1);
    uint64_t* mine = slot + (size_t)blockIdx.x * 4;
    __hip_atomic_store(mine + 0, tag | (b0 >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(mine + 1, tag | (b0 & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(mine + 2, tag | (b1 >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(mine + 3, tag | (b1 & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  uint64_t w[4] = {0, 0, 0, 0};
  const bool mine_lane = lane < a.nwg;
  bool ok = true;
  for (long spins = 0;; ++spins) {
    bool have = extra_poll();
    if (mine_lane) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        w[k] = __hip_atomic_load(slot + (size_t)lane * 4 + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        have = have && (unsigned)(w[k] >> 32) == epoch;
      }
    }
    if (__all(have)) break;
    if (spins > kSpinLimit || __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
      ok = false;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  if (!ok && lane == 0) __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  double d0 = 0.0, d1 = 0.0;
  if (mine_lane) {
    d0 = __longlong_as_double((long long)(((w[0] & 0xffffffffull) << 32) | (w[1] & 0xffffffffull)));
    d1 = __longlong_as_double((long long)(((w[2] & 0xffffffffull) << 32) | (w[3] & 0xffffffffull)));
  }
  s0 = wave_sum(d0);
  s1 = wave_sum(d1);
  // every wave of the workgroup takes the same decision
  if (lane == 0) red[2 * kCgpMaxWaves + wv] = ok ? 1.0 : 0.0;
  __syncthreads();
  bool all_ok = true;
  for (int i = 0; i < NW; ++i) all_ok = all_ok && red[2 * kCgpMaxWaves + i] != 0.0;
  __syncthreads();
  return all_ok;
}

__device__ __forceinline__ bool no_extra() { return true; }

// a double as two tagged words and back
__device__ __forceinline__ void put_tagged(uint64_t* w, unsigned epoch, double v) {
  const uint64_t tag = (uint64_t)epoch << 32, b = (uint64_t)__double_as_longlong(v);
  __hip_atomic_store(w, tag | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(w + 1, tag | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double from_tagged(uint64_t hi, uint64_t lo) {
  return __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
}

// one site's slots that stay in the box, as bits
__device__ __forceinline__ unsigned slot_bits(int gx, int gy, int nx, int ny) {
  unsigned v = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int qx = gx + off2(k, 0), qy = gy + off2(k, 1);
    v |= (qx >= 0 && qx < nx && qy >= 0 && qy < ny) ? 1u << k : 0u;
  }
  return v;
}

// shell sums (S0, S1, TS0, TS1) of the site at LDS row lp (of p), column gx: slots leaving the box read the site
// itself with weight zero, so the loop has no branches
__device__ __forceinline__ void shell_site(const double* pL, int nx, int lp, int gx, unsigned vb, const double* tvk,
                                           double w[4]) {
  const double vx = pL[(lp * nx + gx) * 2], vy = pL[(lp * nx + gx) * 2 + 1];
  w[0] = w[1] = w[2] = w[3] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bool in = (vb >> k) & 1u;
    const int q = in ? (lp + off2(k, 1)) * nx + gx + off2(k, 0) : lp * nx + gx;
    const double dl = (in ? 1.0 : 0.0) *
                      (nd2<double>(k, 0) * (vx - pL[2 * q]) + nd2<double>(k, 1) * (vy - pL[2 * q + 1]));
    w[shell2(k)] += dl;
    w[2 + shell2(k)] += tvk[k] * dl;
  }
}

template <int NT, int PPT>
__global__ __launch_bounds__(NT) void cg2p_kernel(CgpArgs a) {
  constexpr int HMAX = (8 * 128 + NT - 1) / NT;  // halo values of r per thread (lattices up to 128 wide)
  extern __shared__ double lds[];
  const int nx = a.nx, ny = a.ny, R = a.rows;
  const CgpLds L(nx, R);
  double* pL = lds + L.p;
  double* wL = lds + L.w;
  double* tvL = lds + L.tv;
  double* red = lds + L.red;
  const int y0 = blockIdx.x * R, y1 = min(ny, y0 + R), S = y1 - y0;
  const int nown = S * nx;
  const int tid = threadIdx.x;
  unsigned epoch = 0;

  // Tv of the rows y0 - 1 and y1 into LDS (their shell sums); Kn, Tv of the own particles in registers
  for (int t = tid; t < 2 * nx * 8; t += NT) {
    const int gy = t < nx * 8 ? y0 - 1 : y1;
    tvL[t] = (gy >= 0 && gy < ny) ? a.tv[(size_t)gy * nx * 8 + t % (nx * 8)] : 0.0;
  }
  double kn[PPT][8], tv[PPT][8];
  double xr[PPT][2], rr_[PPT][2], pr[PPT][2], mk[PPT][2];
  unsigned vb[PPT];
  double bb = 0.0, rr = 0.0;
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int o = tid + j * NT;
    const bool own = o < nown;
    const size_t gp = (size_t)y0 * nx + (own ? o : 0);
    vb[j] = own ? slot_bits(o % nx, y0 + o / nx, nx, ny) : 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      kn[j][k] = own ? a.kn[gp * 8 + k] : 0.0;
      tv[j][k] = own ? a.tv[gp * 8 + k] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const double bv = own ? a.b[2 * gp + c] : 0.0;
      mk[j][c] = (own && a.mask[2 * gp + c]) ? 1.0 : 0.0;
      xr[j][c] = 0.0;
      rr_[j][c] = bv * mk[j][c];
      pr[j][c] = rr_[j][c];
      bb += bv * bv;
      rr += rr_[j][c] * rr_[j][c];
      if (own) {
        xst(a.pbuf + 2 * gp + c, pr[j][c]);
        pL[(2 * nx + o) * 2 + c] = pr[j][c];
      }
    }
  }
  // the sites of rows y0 - 1 and y1 whose shell sums this workgroup forms: thread t < 2 nx, row h = t / nx
  const int ht = tid < 2 * nx ? tid : 0;
  const int hrow = ht / nx, hx = ht % nx, hls = hrow == 0 ? 0 : S + 1, hgy = y0 - 1 + hls;
  const bool hown = tid < 2 * nx && hgy >= 0 && hgy < ny;
  const unsigned hvb = hown ? slot_bits(hx, hgy, nx, ny) : 0u;

  double s0, s1;
  ++epoch;
  bool ok = exchange<NT, true>(a, epoch, bb, rr, red + CgpLds::kRed * (epoch & 1u), s0, s1, [] { return no_extra(); });
  const double bn = sqrt(s0);
  rr = s1;
  int it = 0, status = ok ? CGP_CONV : CGP_TIMEOUT;
  double rlast = rr;  // r.r of the last residual
  bool done = !ok || bn == 0.0 || sqrt(rr) <= a.tol * bn;
  if (!done) {
    // halo rows of p_0
    for (int t = tid; t < 4 * nx; t += NT) {
      const int hr = t / nx, gx = t % nx;
      const int lr = hr < 2 ? hr : S + hr;  // LDS rows 0, 1, S + 2, S + 3
      const int gy = y0 - 2 + lr;
      const size_t gp = (size_t)gy * nx + gx;
      const bool in = gy >= 0 && gy < ny;
      pL[(lr * nx + gx) * 2] = in ? xld(a.pbuf + 2 * gp) : 0.0;
      pL[(lr * nx + gx) * 2 + 1] = in ? xld(a.pbuf + 2 * gp + 1) : 0.0;
    }
    __syncthreads();
  }
  while (!done) {
    // shell sums of p: the own particles by their threads (Tv in registers), rows y0 - 1 and y1 from LDS
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int o = tid + j * NT;
      if (o < nown) {
        double w[4];
        shell_site(pL, nx, o / nx + 2, o % nx, vb[j], tv[j], w);
#pragma unroll
        for (int c = 0; c < 4; ++c) wL[(nx + o) * 4 + c] = w[c];
      }
    }
    if (hown) {
      double w[4];
      shell_site(pL, nx, hls + 1, hx, hvb, tvL + (size_t)(hrow * nx + hx) * 8, w);
#pragma unroll
      for (int c = 0; c < 4; ++c) wL[(hls * nx + hx) * 4 + c] = w[c];
    }
    __syncthreads();
    // A p on the own particles (masked), partial p.Ap
    double ap[PPT][2], pap = 0.0;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int o = tid + j * NT;
      ap[j][0] = ap[j][1] = 0.0;
      if (o < nown) {
        const int gx = o % nx, lp = o / nx + 2, ls = o / nx + 1;
        const double vx = pr[j][0], vy = pr[j][1];
        const double* wi = wL + (size_t)(ls * nx + gx) * 4;
        double f0 = 0.0, f1 = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const bool in = (vb[j] >> k) & 1u;
          const int c = shell2(k);
          const int q = in ? (lp + off2(k, 1)) * nx + gx + off2(k, 0) : lp * nx + gx;
          const int qs = in ? (ls + off2(k, 1)) * nx + gx + off2(k, 0) : ls * nx + gx;
          const double* wq = wL + (size_t)qs * 4;
          const double dl = nd2<double>(k, 0) * (vx - pL[2 * q]) + nd2<double>(k, 1) * (vy - pL[2 * q + 1]);
          const double F = (in ? 1.0 : 0.0) * (2.0 * kn[j][k] * dl + 0.5 * (wi[2 + c] + wq[2 + c]) +
                                               0.5 * tv[j][k] * (wi[c] + wq[c]));
          f0 += F * nd2<double>(k, 0);
          f1 += F * nd2<double>(k, 1);
        }
        ap[j][0] = f0 * mk[j][0];
        ap[j][1] = f1 * mk[j][1];
        pap += pr[j][0] * ap[j][0] + pr[j][1] * ap[j][1];
      }
    }
    ++epoch;
    if (!exchange<NT, false>(a, epoch, pap, 0.0, red + CgpLds::kRed * (epoch & 1u), s0, s1,
                             [] { return no_extra(); })) {
      status = CGP_TIMEOUT;
      break;
    }
    if (!(s0 > 0.0)) {  // K not positive definite on p (or NaN)
      status = s0 == s0 ? CGP_NOTSPD : CGP_NONFINITE;
      break;
    }
    const double alpha = rr / s0;
    double rn = 0.0;
    // r of this iteration, published as tagged words (buffer by iteration parity) with the next exchange's epoch
    const unsigned e2 = epoch + 1;
    uint64_t* rw = a.rwords + (size_t)(it & 1) * nx * ny * 4;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int o = tid + j * NT;
      if (o < nown) {
        const size_t gp = (size_t)y0 * nx + o;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          xr[j][c] += alpha * pr[j][c];
          rr_[j][c] -= alpha * ap[j][c];
          rn += rr_[j][c] * rr_[j][c];
          put_tagged(rw + (gp * 2 + c) * 2, e2, rr_[j][c]);
        }
      }
    }
    // the neighbours' r on the two halo rows each side, polled together with the partial sums of r.r: thread t
    // takes halo values t + i NT (row h / (2 nx), site (h mod 2 nx) / 2, component h mod 2)
    uint64_t hw[HMAX][2];
    auto halo_poll = [&]() -> bool {
      bool have = true;
#pragma unroll
      for (int i = 0; i < HMAX; ++i) {
        const int h = tid + i * NT;
        if (h >= 8 * nx) break;
        const int hr = h / (2 * nx), rem = h - hr * 2 * nx;
        const int gy = y0 - 2 + (hr < 2 ? hr : S + hr);
        if (gy < 0 || gy >= ny) continue;
        const uint64_t* src = rw + (((size_t)gy * nx + (rem >> 1)) * 2 + (rem & 1)) * 2;
        hw[i][0] = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hw[i][1] = __hip_atomic_load(src + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        have = have && (unsigned)(hw[i][0] >> 32) == e2 && (unsigned)(hw[i][1] >> 32) == e2;
      }
      return have;
    };
    ++epoch;
    if (!exchange<NT, false>(a, epoch, rn, 0.0, red + CgpLds::kRed * (epoch & 1u), s0, s1, halo_poll)) {
      status = CGP_TIMEOUT;
      break;
    }
    ++it;
    rlast = s0;
    const double rnorm = sqrt(s0);
    if (!(rnorm <= 1e300)) {
      status = CGP_NONFINITE;
      break;
    }
    if (rnorm <= a.tol * bn) break;
    if (it >= a.maxit) {
      status = CGP_MAXIT;
      break;
    }
    const double beta = s0 / rr;
    rr = s0;
    // p = r + beta p: own particles from registers, halo rows from the neighbours' r and the previous p in LDS
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int o = tid + j * NT;
      if (o < nown)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          pr[j][c] = fma(beta, pr[j][c], rr_[j][c]);
          pL[(2 * nx + o) * 2 + c] = pr[j][c];
        }
    }
#pragma unroll
    for (int i = 0; i < HMAX; ++i) {
      const int h = tid + i * NT;
      if (h >= 8 * nx) break;
      const int hr = h / (2 * nx), rem = h - hr * 2 * nx;
      const int lr = hr < 2 ? hr : S + hr;
      const int gy = y0 - 2 + lr;
      if (gy < 0 || gy >= ny) continue;
      double* ph = pL + (lr * nx + (rem >> 1)) * 2 + (rem & 1);
      *ph = fma(beta, *ph, from_tagged(hw[i][0], hw[i][1]));
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int o = tid + j * NT;
    if (o < nown) {
      const size_t gp = (size_t)y0 * nx + o;
      a.x[2 * gp] = xr[j][0];
      a.x[2 * gp + 1] = xr[j][1];
    }
  }
  if (blockIdx.x == 0 && tid == 0) {
    a.out[0] = it;
    a.out[1] = status;
    *a.relres = bn > 0.0 ? sqrt(rlast) / bn : 0.0;
  }
}

// published words and the abort flag start at zero (no tag of a launch is zero)
__global__ void cgp_zero_kernel(uint64_t* slots, int* abort_flag) {
  for (int i = threadIdx.x; i < 2 * kCgpMaxWg * 4; i += blockDim.x) slots[i] = 0;
  if (threadIdx.x == 0) *abort_flag = 0;
}

// workspace: p_0 (N, 2), tagged r words (2, N, 2, 2), published words, abort flag, result words
struct CgpWork {
  size_t pbuf, rwords, slots, flag, out, rel, total;
  CgpWork(int64_t N) {
    auto al = [](size_t n) { return (n + 255) & ~(size_t)255; };
    size_t o = 0;
    pbuf = o;
    o += al(sizeof(double) * 2 * N);
    rwords = o;
    o += al(sizeof(uint64_t) * 8 * N);
    slots = o;
    o += al(sizeof(uint64_t) * 2 * kCgpMaxWg * 4);
    flag = o;
    o += 256;
    out = o;
    o += 256;
    rel = o;
    o += 256;
    total = o;
  }
};

// rows per workgroup (about `want` workgroups) and the kernel: 256 threads with 1, 2 or 4 particles each (larger
// workgroups run out of registers: x, r, p, Kn and Tv of the own particles live there); false when the rows of one
// workgroup do not fit its threads (the two side rows need 2 nx <= 256) or its LDS
using CgpKernel = void (*)(CgpArgs);
bool cgp_shape(int nx, int ny, int ncu, size_t max_lds, int want, int& rows, int& nwg, int& nt, CgpKernel& fn) {
  want = std::min({want > 0 ? want : kCgpAutoWg, kCgpMaxWg, ncu, ny});
  rows = (ny + want - 1) / want;
  nwg = (ny + rows - 1) / rows;
  const int per = rows * nx;
  nt = 256;
  if (per <= 256)
    fn = cg2p_kernel<256, 1>;
  else if (per <= 512)
    fn = cg2p_kernel<256, 2>;
  else if (per <= 1024)
    fn = cg2p_kernel<256, 4>;
  else
    return false;
  return 2 * nx <= nt && sizeof(double) * (size_t)CgpLds(nx, rows).total <= max_lds;
}

}  // namespace

extern "C" {

size_t lpm_cg2_persistent_workspace(int32_t nx, int32_t ny) {
  if (nx < 1 || ny < 1) return 0;
  return CgpWork((int64_t)nx * ny).total;
}

int lpm_cg2_persistent(int32_t nx, int32_t ny, const double* kn, const double* tv, const uint8_t* mask,
                       const double* b, double* x, double tol, int32_t maxit, int32_t workgroups, void* work,
                       int32_t* iters, double* relres, int32_t* conv, void* stream) {
  if (nx < 1 || ny < 1 || !kn || !tv || !mask || !b || !x || !work || !iters || !relres || !conv || maxit < 1 ||
      !(tol >= 0.0))
    return fail(LPM_E_BADARG, "lpm_cg2_persistent: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  int dev = 0, ncu = 0, max_lds = 0, coop = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  int rows = 0, nwg = 0, nt = 0;
  CgpKernel kfn = nullptr;
  if (!coop || workgroups < 0 || !cgp_shape(nx, ny, ncu, (size_t)max_lds, workgroups, rows, nwg, nt, kfn))
    return fail(LPM_E_UNSUPPORTED, "lpm_cg2_persistent: a %d x %d lattice does not fit one cooperative launch", nx,
                ny);
  const int64_t N = (int64_t)nx * ny;
  const CgpWork W(N);
  char* w = (char*)work;
  CgpArgs a{};
  a.nx = nx;
  a.ny = ny;
  a.rows = rows;
  a.nwg = nwg;
  a.kn = kn;
  a.tv = tv;
  a.mask = mask;
  a.b = b;
  a.x = x;
  a.pbuf = (double*)(w + W.pbuf);
  a.rwords = (uint64_t*)(w + W.rwords);
  a.slots = (uint64_t*)(w + W.slots);
  a.abort_flag = (int*)(w + W.flag);
  a.tol = tol;
  a.maxit = maxit;
  a.out = (int*)(w + W.out);
  a.relres = (double*)(w + W.rel);
  const unsigned lds = (unsigned)(sizeof(double) * CgpLds(nx, rows).total);
  const void* fn = (const void*)kfn;
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);  // a no-op where not needed
  int per_cu = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, nt, lds));
  if (per_cu * ncu < nwg)
    return fail(LPM_E_UNSUPPORTED, "lpm_cg2_persistent: %d workgroups cannot be resident together", nwg);
  hipLaunchKernelGGL(cgp_zero_kernel, dim3(1), dim3(256), 0, st, a.slots, a.abort_flag);
  LAUNCHCHK();
  void* params[] = {&a};
  HIPCHK(hipLaunchCooperativeKernel(fn, dim3(nwg), dim3(nt), params, lds, st));
  int out[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(out, a.out, sizeof(out), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(relres, a.relres, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *iters = out[0];
  *conv = out[1] == CGP_CONV ? 1 : 0;
  if (out[1] == CGP_NOTSPD) return fail(LPM_E_NOT_SPD, "lpm_cg2_persistent: p.Ap <= 0 at iteration %d", out[0] + 1);
  if (out[1] == CGP_NONFINITE) return fail(LPM_E_NONFINITE, "lpm_cg2_persistent: non-finite value at iteration %d",
                                           out[0] + 1);
  if (out[1] == CGP_TIMEOUT) return fail(LPM_E_HIP, "lpm_cg2_persistent: grid barrier timed out");
  return LPM_OK;
}

}  // extern "C"
