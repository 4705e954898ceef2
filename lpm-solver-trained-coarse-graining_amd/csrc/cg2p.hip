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
constexpr int kCgpAutoWg = 64;    // workgroups requested by the automatic launch shape
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
  uint64_t* slots;      // (2, 2, 2, kCgpMaxWg) published partial sums by epoch parity, value and half: (epoch << 32)
                        // | half of the double (zero at launch)
  int* abort_flag;      // set by a workgroup whose exchange wait timed out (zero at launch)
  double tol;
  int maxit;
  int* out;  // [iterations, status]
  double* relres;
};

// LDS layout (doubles), structure of arrays so that lanes reading consecutive sites read consecutive words (no bank
// conflicts): the multiplied vector on the rows + 2 each side as (2, R + 4, nx), shell sums on the rows + 1 each side
// as (4, R + 2, nx), two reduction buffers (per-wave partial pairs, per-wave flags, a broadcast pair)
struct CgpLds {
  int p, w, red, total, vs, ws;
  static constexpr int kRed = 3 * kCgpMaxWaves + 2;
  __host__ __device__ CgpLds(int nx, int R) {
    vs = (R + 4) * nx;  // component stride of the vector
    ws = (R + 2) * nx;  // component stride of the shell sums
    p = 0;
    w = p + 2 * vs;
    red = w + 4 * ws;  // (2, kRed)
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

// a double as two tagged words and back
__device__ __forceinline__ void put_tagged(uint64_t* w, unsigned epoch, double v) {
  const uint64_t tag = (uint64_t)epoch << 32, b = (uint64_t)__double_as_longlong(v);
  __hip_atomic_store(w, tag | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(w + 1, tag | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double from_tagged(uint64_t hi, uint64_t lo) {
  return __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
}

// Grid-wide sums of NV values (v0, v1) that also act as a barrier, in two halves.  xpub: the workgroup's sums
// (fixed order) are published by thread 0, each as two words (epoch << 32) | half, after every thread's earlier
// stores have completed (WAIT: s_waitcnt).  xsum: the first wave polls the words of all workgroups, lane j those of
// workgroup j, until every tag equals `epoch`, and sums the payloads over the lanes in a fixed order, so every
// workgroup obtains the same bits.  The words of one half of one value lie side by side for all workgroups, so one
// wave load reads them (a poll round requests 2 NV runs of 512 bytes, not a line per workgroup and word).  Work
// placed between the halves overlaps the exchange.  xsum returns false (in every thread of the workgroup) when a
// wait timed out or another workgroup gave up.
template <int NT, int NV, bool WAIT>
__device__ void xpub(const CgpArgs& a, unsigned epoch, double v0, double v1, double* red) {
  constexpr int NW = NT / 64;
  v0 = wave_sum(v0);
  if constexpr (NV > 1) v1 = wave_sum(v1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[2 * wv] = v0;
    red[2 * wv + 1] = v1;
  }
  if constexpr (WAIT) __builtin_amdgcn_s_waitcnt(0);  // this wave's stores of untagged exchanged data completed
  __syncthreads();
  uint64_t* slot = a.slots + (size_t)(epoch & 1u) * 4 * kCgpMaxWg + blockIdx.x;
  if (threadIdx.x == 0) {
    double t[2] = {0.0, 0.0};
    for (int i = 0; i < NW; ++i) {
      t[0] += red[2 * i];
      t[1] += red[2 * i + 1];
    }
    const uint64_t tag = (uint64_t)epoch << 32;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const uint64_t b = (uint64_t)__double_as_longlong(t[k]);
      __hip_atomic_store(slot + (2 * k) * kCgpMaxWg, tag | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(slot + (2 * k + 1) * kCgpMaxWg, tag | (b & 0xffffffffull), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Polls until `poll` holds in every lane of every wave of the workgroup (each lane checks the tags of its own
// words).  Bounded: past kSpinLimit polls, or once another workgroup gave up, the abort flag is set and every thread
// of the workgroup returns false.
template <int NT, typename Poll>
__device__ bool wait_words(const CgpArgs& a, double* red, Poll&& poll) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  bool ok = true;
  for (long spins = 0;; ++spins) {
    if (__all(poll())) break;
    // the abort flag is another round trip to memory: read it only every 64th poll, so a poll round waits for
    // one load latency, not two
    if ((spins & 63) == 63 &&
        (spins > kSpinLimit || __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
      ok = false;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  if (!ok && lane == 0) __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // every wave of the workgroup takes the same decision
  if (lane == 0) red[2 * kCgpMaxWaves + wv] = ok ? 1.0 : 0.0;
  __syncthreads();
  bool all_ok = true;
  for (int i = 0; i < NW; ++i) all_ok = all_ok && red[2 * kCgpMaxWaves + i] != 0.0;
  __syncthreads();
  return all_ok;
}

template <int NT, int NV, typename Extra>
__device__ bool xsum(const CgpArgs& a, unsigned epoch, double* red, double& s0, double& s1, Extra&& extra_poll) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t* slot = a.slots + (size_t)(epoch & 1u) * 4 * kCgpMaxWg;
  uint64_t w[4] = {0, 0, 0, 0};
  // the first wave polls the words (lane j those of workgroup j) and passes the sums on through LDS: one wave's
  // polling per workgroup keeps the traffic on the shared words low
  const bool mine_lane = wv == 0 && lane < a.nwg;
  const bool ok = wait_words<NT>(a, red, [&]() -> bool {
    bool have = extra_poll();
    if (mine_lane) {
#pragma unroll
      for (int k = 0; k < 2 * NV; ++k) {
        w[k] = __hip_atomic_load(slot + k * kCgpMaxWg + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        have = have && (unsigned)(w[k] >> 32) == epoch;
      }
    }
    return have;
  });
  if (wv == 0) {
    double d0 = mine_lane ? from_tagged(w[0], w[1]) : 0.0;
    double d1 = mine_lane && NV > 1 ? from_tagged(w[2], w[3]) : 0.0;
    d0 = wave_sum(d0);
    if constexpr (NV > 1) d1 = wave_sum(d1);
    if (lane == 0) {
      red[CgpLds::kRed - 2] = d0;
      red[CgpLds::kRed - 1] = d1;
    }
  }
  __syncthreads();
  s0 = red[CgpLds::kRed - 2];
  s1 = red[CgpLds::kRed - 1];
  return ok;
}

template <int NT, int NV, bool WAIT, typename Extra>
__device__ bool exchange(const CgpArgs& a, unsigned epoch, double v0, double v1, double* red, double& s0,
                         double& s1, Extra&& extra_poll) {
  xpub<NT, NV, WAIT>(a, epoch, v0, v1, red);
  return xsum<NT, NV>(a, epoch, red, s0, s1, extra_poll);
}

__device__ __forceinline__ bool no_extra() { return true; }

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
__device__ __forceinline__ void shell_site(const double* pL, int vs, int nx, int lp, int gx, unsigned vb,
                                           const double (&tvk)[8], double w[4]) {
  const double vx = pL[lp * nx + gx], vy = pL[vs + lp * nx + gx];
  w[0] = w[1] = w[2] = w[3] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bool in = (vb >> k) & 1u;
    const int q = in ? (lp + off2(k, 1)) * nx + gx + off2(k, 0) : lp * nx + gx;
    const double dl = (in ? 1.0 : 0.0) *
                      (nd2<double>(k, 0) * (vx - pL[q]) + nd2<double>(k, 1) * (vy - pL[vs + q]));
    w[shell2(k)] += dl;
    w[2 + shell2(k)] += tvk[k] * dl;
  }
}

// Shell sums of v (LDS rows of the band + two halo rows each side) on the band's own particles and on the side-row
// site of this thread (row y0 - 1 or y1, its Tv in registers), written to wL.  All sums are formed before the stores,
// so the LDS reads of both kinds of site overlap.  The caller synchronises before wL is read.
template <int NT, int PPT>
__device__ __forceinline__ void band_shell(const double* vL, double* wL, const CgpLds& L, int nx, int nown,
                                           const unsigned (&vb)[PPT], const double (&tv)[PPT][8], bool hown,
                                           int hls, int hx, unsigned hvb, const double (&tvh)[8]) {
  double w[PPT][4], wh[4];
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int o = threadIdx.x + j * NT;
    if (o < nown) shell_site(vL, L.vs, nx, o / nx + 2, o % nx, vb[j], tv[j], w[j]);
  }
  if (hown) shell_site(vL, L.vs, nx, hls + 1, hx, hvb, tvh, wh);
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int o = threadIdx.x + j * NT;
    if (o < nown)
#pragma unroll
      for (int c = 0; c < 4; ++c) wL[c * L.ws + nx + o] = w[j][c];
  }
  if (hown)
#pragma unroll
    for (int c = 0; c < 4; ++c) wL[c * L.ws + hls * nx + hx] = wh[c];
}

// av = mask (K v) on the band's own particles from v (registers and LDS) and the shell sums in wL
template <int NT, int PPT>
__device__ __forceinline__ void band_apply(const double* vL, const double* wL, const CgpLds& L, int nx, int nown,
                                           const unsigned (&vb)[PPT], const double (&kn)[PPT][8],
                                           const double (&tv)[PPT][8], const double (&mk)[PPT][2],
                                           const double (&v)[PPT][2], double (&av)[PPT][2]) {
  const int vs = L.vs, ws = L.ws;
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int o = threadIdx.x + j * NT;
    av[j][0] = av[j][1] = 0.0;
    if (o < nown) {
      const int gx = o % nx, lp = o / nx + 2, ls = o / nx + 1;
      const double vx = v[j][0], vy = v[j][1];
      const int si = ls * nx + gx;
      const double wi[4] = {wL[si], wL[ws + si], wL[2 * ws + si], wL[3 * ws + si]};
      double f0 = 0.0, f1 = 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const bool in = (vb[j] >> k) & 1u;
        const int c = shell2(k);
        const int q = in ? (lp + off2(k, 1)) * nx + gx + off2(k, 0) : lp * nx + gx;
        const int qs = in ? (ls + off2(k, 1)) * nx + gx + off2(k, 0) : ls * nx + gx;
        const double wq_c = wL[c * ws + qs], wq_t = wL[(2 + c) * ws + qs];
        const double dl = nd2<double>(k, 0) * (vx - vL[q]) + nd2<double>(k, 1) * (vy - vL[vs + q]);
        const double F = (in ? 1.0 : 0.0) * (2.0 * kn[j][k] * dl + 0.5 * (wi[2 + c] + wq_t) +
                                             0.5 * tv[j][k] * (wi[c] + wq_c));
        f0 += F * nd2<double>(k, 0);
        f1 += F * nd2<double>(k, 1);
      }
      av[j][0] = f0 * mk[j][0];
      av[j][1] = f1 * mk[j][1];
    }
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
  double* red = lds + L.red;
  const int y0 = blockIdx.x * R, y1 = min(ny, y0 + R), S = y1 - y0;
  const int nown = S * nx;
  const int tid = threadIdx.x;
  unsigned epoch = 0;

  // Kn, Tv of the own particles in registers
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
        pL[c * L.vs + 2 * nx + o] = pr[j][c];
      }
    }
  }
  // the sites of rows y0 - 1 and y1 whose shell sums this workgroup forms: thread t < 2 nx, row h = t / nx
  // (threads after the own particles when the workgroup has enough of them, so both kinds run side by side)
  const int hoff = NT >= R * nx + 2 * nx ? R * nx : 0;
  const int hid = tid - hoff;
  const int ht = hid >= 0 && hid < 2 * nx ? hid : 0;
  const int hrow = ht / nx, hx = ht % nx, hls = hrow == 0 ? 0 : S + 1, hgy = y0 - 1 + hls;
  const bool hown = hid >= 0 && hid < 2 * nx && hgy >= 0 && hgy < ny;
  const unsigned hvb = hown ? slot_bits(hx, hgy, nx, ny) : 0u;
  double tvh[8];  // Tv of that site
#pragma unroll
  for (int k = 0; k < 8; ++k) tvh[k] = hown ? a.tv[((size_t)hgy * nx + hx) * 8 + k] : 0.0;

  double s0, s1;
  ++epoch;
  bool ok = exchange<NT, 2, true>(a, epoch, bb, rr, red + CgpLds::kRed * (epoch & 1u), s0, s1, [] { return no_extra(); });
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
      pL[lr * nx + gx] = in ? xld(a.pbuf + 2 * gp) : 0.0;
      pL[L.vs + lr * nx + gx] = in ? xld(a.pbuf + 2 * gp + 1) : 0.0;
    }
    __syncthreads();
  }
  while (!done) {
    // shell sums of p: the own particles by their threads (Tv in registers), rows y0 - 1 and y1 from LDS
    band_shell<NT, PPT>(pL, wL, L, nx, nown, vb, tv, hown, hls, hx, hvb, tvh);
    __syncthreads();
    // A p on the own particles (masked), partial p.Ap
    double ap[PPT][2], pap = 0.0;
    band_apply<NT, PPT>(pL, wL, L, nx, nown, vb, kn, tv, mk, pr, ap);
#pragma unroll
    for (int j = 0; j < PPT; ++j)
      if (tid + j * NT < nown) pap += pr[j][0] * ap[j][0] + pr[j][1] * ap[j][1];
    ++epoch;
    if (!exchange<NT, 1, false>(a, epoch, pap, 0.0, red + CgpLds::kRed * (epoch & 1u), s0, s1,
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
    if (!exchange<NT, 1, false>(a, epoch, rn, 0.0, red + CgpLds::kRed * (epoch & 1u), s0, s1, halo_poll)) {
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
          pL[c * L.vs + 2 * nx + o] = pr[j][c];
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
      double* ph = pL + (rem & 1) * L.vs + lr * nx + (rem >> 1);
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

// Pipelined CG (Ghysels and Vanroose, "Hiding global synchronization latency in the preconditioned Conjugate
// Gradient algorithm", Parallel Computing 40 (2014), Alg. 4 without preconditioner): the same Krylov iterates as the
// classic CG in exact arithmetic with one grid-wide reduction per iteration, (r.r, w.r) with w = A r, whose exchange
// overlaps the product q = A w:
//   gamma = r.r, delta = w.r (published);  q = A w;  (sums arrive)
//   beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)   (first iteration: 0, gamma / delta)
//   z = q + beta z, s = w + beta s, p = r + beta p, x += alpha p, r -= alpha s, w -= alpha z.
// The stopping test ||r|| <= tol ||b|| is taken on the recursively updated r before each update and then confirmed
// on the true residual mask (b - A x) (one more product and exchange), which the recurrences can drift from.  w is
// published each iteration as tagged words (buffered by tag parity) and the neighbours' two halo rows are polled
// before the product.  delta - beta gamma / alpha_old equals p.Ap.  Near the attainable residual that difference
// cancels as the recurrences drift from the quantities they stand for.  Where it is <= 0 the recurrence quantities are
// replaced by their definitions (residual replacement, Cools et al., SIAM J. Sci. Comput. 40 (2018)):
// r = mask (b - A x), w = A r, s = A p, z = A s, and the iteration continues with the same search direction (no
// update counted); so does a true residual that misses the tolerance the recursive one met.  If p.Ap <= 0 right after
// a replacement the iteration restarts from x with the true residual (p = s = z = 0); p.Ap <= 0 on the first update
// after a (re)start, where it is the exact r.Ar, stops with NOT_SPD.
template <int NT, int PPT>
__global__ __launch_bounds__(NT) void cg2pp_kernel(CgpArgs a) {
  constexpr int HMAX = (8 * 128 + NT - 1) / NT;  // halo values per thread (lattices up to 128 wide)
  extern __shared__ double lds[];
  const int nx = a.nx, ny = a.ny, R = a.rows;
  const CgpLds L(nx, R);
  double* vL = lds + L.p;
  double* wL = lds + L.w;
  double* red = lds + L.red;
  const int y0 = blockIdx.x * R, y1 = min(ny, y0 + R), S = y1 - y0;
  const int nown = S * nx;
  const int tid = threadIdx.x;
  const size_t nbuf = (size_t)nx * ny * 4;  // words of one parity buffer
  unsigned epoch = 0;

  // halo rows outside the lattice hold zeros (never read through a bond)
  for (int t = tid; t < 8 * nx; t += NT) {
    const int c = t / (4 * nx), hr = (t % (4 * nx)) / nx, lr = hr < 2 ? hr : S + hr, gy = y0 - 2 + lr;
    if (gy < 0 || gy >= ny) vL[c * L.vs + lr * nx + t % nx] = 0.0;
  }
  double kn[PPT][8], tv[PPT][8], mk[PPT][2];
  double x[PPT][2], r[PPT][2], w[PPT][2], p[PPT][2], sv[PPT][2], z[PPT][2], bm[PPT][2];
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
      bm[j][c] = bv * mk[j][c];
      r[j][c] = bm[j][c];
      x[j][c] = w[j][c] = p[j][c] = sv[j][c] = z[j][c] = 0.0;
      bb += bv * bv;
      rr += r[j][c] * r[j][c];
    }
  }
  // the side-row sites (shell sums only) go to the threads after the own particles when the workgroup has enough
  // of them, so both kinds run side by side
  const int hoff = NT >= R * nx + 2 * nx ? R * nx : 0;
  const int hid = tid - hoff;
  const int ht = hid >= 0 && hid < 2 * nx ? hid : 0;
  const int hrow = ht / nx, hx = ht % nx, hls = hrow == 0 ? 0 : S + 1, hgy = y0 - 1 + hls;
  const bool hown = hid >= 0 && hid < 2 * nx && hgy >= 0 && hgy < ny;
  const unsigned hvb = hown ? slot_bits(hx, hgy, nx, ny) : 0u;
  double tvh[8];  // Tv of that site
#pragma unroll
  for (int k = 0; k < 8; ++k) tvh[k] = hown ? a.tv[((size_t)hgy * nx + hx) * 8 + k] : 0.0;

  // a band vector published as tagged words (buffer tag & 1), the neighbours' halo rows polled into hw and both
  // placed in the LDS rows
  auto publish = [&](unsigned tag, const double (&v)[PPT][2]) {
    uint64_t* buf = a.rwords + (size_t)(tag & 1u) * nbuf;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int o = tid + j * NT;
      if (o < nown) {
        const size_t gp = (size_t)y0 * nx + o;
#pragma unroll
        for (int c = 0; c < 2; ++c) put_tagged(buf + (gp * 2 + c) * 2, tag, v[j][c]);
      }
    }
  };
  uint64_t hw[HMAX][2];
  auto halo_poll = [&](unsigned tag) -> bool {
    const uint64_t* buf = a.rwords + (size_t)(tag & 1u) * nbuf;
    bool have = true;
#pragma unroll
    for (int i = 0; i < HMAX; ++i) {
      const int h = tid + i * NT;
      if (h >= 8 * nx) break;
      const int hr = h / (2 * nx), rem = h - hr * 2 * nx;
      const int gy = y0 - 2 + (hr < 2 ? hr : S + hr);
      if (gy < 0 || gy >= ny) continue;
      const uint64_t* src = buf + (((size_t)gy * nx + (rem >> 1)) * 2 + (rem & 1)) * 2;
      hw[i][0] = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      hw[i][1] = __hip_atomic_load(src + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      have = have && (unsigned)(hw[i][0] >> 32) == tag && (unsigned)(hw[i][1] >> 32) == tag;
    }
    return have;
  };
  auto place = [&](const double (&v)[PPT][2]) {
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
      const int o = tid + j * NT;
      if (o < nown) {
        vL[2 * nx + o] = v[j][0];
        vL[L.vs + 2 * nx + o] = v[j][1];
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
      vL[(rem & 1) * L.vs + lr * nx + (rem >> 1)] = from_tagged(hw[i][0], hw[i][1]);
    }
  };

  // |b|^2 and |r_0|^2
  double s0, s1;
  ++epoch;
  const bool ok = exchange<NT, 2, false>(a, epoch, bb, rr, red + CgpLds::kRed * (epoch & 1u), s0, s1,
                                      [] { return no_extra(); });
  const double bn = sqrt(s0);
  int it = 0, status = ok ? CGP_CONV : CGP_TIMEOUT;
  double rlast = s1;
  bool done = !ok || bn == 0.0 || sqrt(s1) <= a.tol * bn;

  // Each pass of the loop multiplies one band vector: v published under a new tag, the neighbours' halo rows
  // polled, shell sums, product.  The mode says which vector and what the product is for: an iteration (q = A w,
  // with the reduction of gamma, delta overlapping it), the true residual that confirms convergence, or a step of
  // a replacement (A x, then w = A r, s = A p, z = A s).  One product site keeps the kernel's registers in bounds.
  enum : int { M_STEP, M_CONFIRM, M_AX, M_AR, M_AP, M_AS };
  int mode = M_AR;    // w_0 = A r_0 first
  bool fresh = true;  // the next update is the first of a (re)start: beta = 0 and p.Ap = delta = r.Ar
  bool exact = true;  // r, w, s, z are the true quantities (start, replacement, restart), not recurrences
  double gamma_old = 1.0, alpha_old = 1.0;
  unsigned tag = 0;
  while (!done) {
    double* red_e = red + CgpLds::kRed * (epoch & 1u);
    if (mode == M_STEP) {
      double pg = 0.0, pd = 0.0;
#pragma unroll
      for (int j = 0; j < PPT; ++j)
        if (tid + j * NT < nown) {
          pg += r[j][0] * r[j][0] + r[j][1] * r[j][1];
          pd += w[j][0] * r[j][0] + w[j][1] * r[j][1];
        }
      ++epoch;
      red_e = red + CgpLds::kRed * (epoch & 1u);
      xpub<NT, 2, false>(a, epoch, pg, pd, red_e);
    }
    double v[PPT][2], out[PPT][2];
#pragma unroll
    for (int j = 0; j < PPT; ++j)
#pragma unroll
      for (int c = 0; c < 2; ++c)
        v[j][c] = mode == M_STEP ? w[j][c]
                  : mode == M_AR ? r[j][c]
                  : mode == M_AP ? p[j][c]
                  : mode == M_AS ? sv[j][c]
                                 : x[j][c];
    publish(++tag, v);
    if (!wait_words<NT>(a, red_e, [&] { return halo_poll(tag); })) {
      status = CGP_TIMEOUT;
      break;
    }
    place(v);
    __syncthreads();
    band_shell<NT, PPT>(vL, wL, L, nx, nown, vb, tv, hown, hls, hx, hvb, tvh);
    __syncthreads();
    band_apply<NT, PPT>(vL, wL, L, nx, nown, vb, kn, tv, mk, v, out);
    if (mode == M_AR || mode == M_AP || mode == M_AS) {
#pragma unroll
      for (int j = 0; j < PPT; ++j)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          if (mode == M_AR) w[j][c] = out[j][c];
          if (mode == M_AP) sv[j][c] = out[j][c];
          if (mode == M_AS) z[j][c] = out[j][c];
        }
      // after w: with a search direction, its products too (a restart has none)
      if (mode == M_AR && !fresh) {
        mode = M_AP;
      } else if (mode == M_AP) {
        mode = M_AS;
      } else {
        mode = M_STEP;
        exact = true;
      }
      continue;
    }
    if (mode == M_AX || mode == M_CONFIRM) {
      double tr = 0.0;
#pragma unroll
      for (int j = 0; j < PPT; ++j)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          r[j][c] = bm[j][c] - out[j][c];
          tr += r[j][c] * r[j][c];
        }
      if (mode == M_CONFIRM) {
        // the true residual decides
        ++epoch;
        if (!exchange<NT, 1, false>(a, epoch, tr, 0.0, red + CgpLds::kRed * (epoch & 1u), s0, s1,
                                 [] { return no_extra(); })) {
          status = CGP_TIMEOUT;
          break;
        }
        rlast = s0;
        if (sqrt(s0) <= a.tol * bn) break;
        // missed (the recursive residual drifted below the true one): replacement, the iteration goes on along the
        // same search direction (a restart from x would meet the tolerance sooner with a worse error: one
        // steepest-descent-like step takes the residual norm under it while the smooth error stays)
      }
      mode = M_AR;
      continue;
    }
    // M_STEP: q = out
    if (!xsum<NT, 2>(a, epoch, red_e, s0, s1, [] { return no_extra(); })) {
      status = CGP_TIMEOUT;
      break;
    }
    const double gamma = s0, delta = s1;
    rlast = gamma;
    const double rnorm = sqrt(gamma);
    if (!(rnorm <= 1e300) || delta != delta) {
      status = CGP_NONFINITE;
      break;
    }
    if (rnorm <= a.tol * bn) {
      // a recursive r is confirmed by the true one; if that misses the tolerance the quantities are replaced
      if (exact) break;
      mode = M_CONFIRM;
      continue;
    }
    if (it >= a.maxit) {
      status = CGP_MAXIT;
      break;
    }
    const double beta = fresh ? 0.0 : gamma / gamma_old;
    const double pap = fresh ? delta : delta - beta * gamma / alpha_old;
    if (!(pap > 0.0)) {
      if (pap != pap || fresh) {  // non-finite, or K not positive definite on r
        status = pap == pap ? CGP_NOTSPD : CGP_NONFINITE;
        break;
      }
      // replacement; where the true quantities give no descent either, a restart from x
      if (exact) {
        fresh = true;
#pragma unroll
        for (int j = 0; j < PPT; ++j)
#pragma unroll
          for (int c = 0; c < 2; ++c) p[j][c] = sv[j][c] = z[j][c] = 0.0;
      }
      mode = M_AX;
      continue;
    }
    const double alpha = gamma / pap;
#pragma unroll
    for (int j = 0; j < PPT; ++j)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        z[j][c] = fma(beta, z[j][c], out[j][c]);
        sv[j][c] = fma(beta, sv[j][c], w[j][c]);
        p[j][c] = fma(beta, p[j][c], r[j][c]);
        x[j][c] = fma(alpha, p[j][c], x[j][c]);
        r[j][c] = fma(-alpha, sv[j][c], r[j][c]);
        w[j][c] = fma(-alpha, z[j][c], w[j][c]);
      }
    gamma_old = gamma;
    alpha_old = alpha;
    fresh = false;
    exact = false;
    ++it;
  }
#pragma unroll
  for (int j = 0; j < PPT; ++j) {
    const int o = tid + j * NT;
    if (o < nown) {
      const size_t gp = (size_t)y0 * nx + o;
      a.x[2 * gp] = x[j][0];
      a.x[2 * gp + 1] = x[j][1];
    }
  }
  if (blockIdx.x == 0 && tid == 0) {
    a.out[0] = it;
    a.out[1] = status;
    *a.relres = bn > 0.0 ? sqrt(rlast) / bn : 0.0;
  }
}

// published words, tagged vector words and the abort flag start at zero (no tag of a launch is zero), so no word
// left by an earlier launch on the same workspace can match a tag
__global__ void cgp_zero_kernel(uint64_t* slots, uint64_t* rwords, int64_t nrw, int* abort_flag) {
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = i0; i < nrw; i += stride) rwords[i] = 0;
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < 2 * 4 * kCgpMaxWg; i += blockDim.x) slots[i] = 0;
    if (threadIdx.x == 0) *abort_flag = 0;
  }
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
    o += al(sizeof(uint64_t) * 2 * 4 * kCgpMaxWg);
    flag = o;
    o += 256;
    out = o;
    o += 256;
    rel = o;
    o += 256;
    total = o;
  }
};

// rows per workgroup (about `want` workgroups) and the kernel: 256 threads with 1, 2 or 4 particles each, or for the
// classic form with up to 256 own particles 512 threads, the side-row sites on threads of their own (9.06 against
// 9.45 us per iteration at 128^2; larger workgroups run out of registers: the vectors, Kn and Tv of the own
// particles live there); false when the rows of one workgroup do not fit its threads (the two side rows need
// 2 nx <= 256) or its LDS.  The pipelined kernel defaults to the most workgroups too (its reduction is hidden behind
// the product, so smaller bands pay).
using CgpKernel = void (*)(CgpArgs);
bool cgp_shape(bool pipelined, int nx, int ny, int ncu, size_t max_lds, int want, int& rows, int& nwg, int& nt,
               CgpKernel& fn) {
  want = std::min({want > 0 ? want : (pipelined ? kCgpMaxWg : kCgpAutoWg), kCgpMaxWg, ncu, ny});
  rows = (ny + want - 1) / want;
  nwg = (ny + rows - 1) / rows;
  const int per = rows * nx;
  nt = 256;
  if (!pipelined && per <= 256 && per + 2 * nx <= 512) {
    nt = 512;
    fn = cg2p_kernel<512, 1>;
    return sizeof(double) * (size_t)CgpLds(nx, rows).total <= max_lds;
  }
  if (per <= 256)
    fn = pipelined ? cg2pp_kernel<256, 1> : cg2p_kernel<256, 1>;
  else if (per <= 512)
    fn = pipelined ? cg2pp_kernel<256, 2> : cg2p_kernel<256, 2>;
  else if (per <= 1024)
    fn = pipelined ? cg2pp_kernel<256, 4> : cg2p_kernel<256, 4>;
  else
    return false;
  return 2 * nx <= nt && sizeof(double) * (size_t)CgpLds(nx, rows).total <= max_lds;
}

int cg2_persistent(bool pipelined, const char* name, int32_t nx, int32_t ny, const double* kn, const double* tv,
                   const uint8_t* mask, const double* b, double* x, double tol, int32_t maxit, int32_t workgroups,
                   void* work, int32_t* iters, double* relres, int32_t* conv, void* stream) {
  if (nx < 1 || ny < 1 || !kn || !tv || !mask || !b || !x || !work || !iters || !relres || !conv || maxit < 1 ||
      !(tol >= 0.0))
    return fail(LPM_E_BADARG, "%s: bad arguments", name);
  hipStream_t st = (hipStream_t)stream;
  int dev = 0, ncu = 0, max_lds = 0, coop = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIPCHK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  int rows = 0, nwg = 0, nt = 0;
  CgpKernel kfn = nullptr;
  if (!coop || workgroups < 0 ||
      !cgp_shape(pipelined, nx, ny, ncu, (size_t)max_lds, workgroups, rows, nwg, nt, kfn))
    return fail(LPM_E_UNSUPPORTED, "%s: a %d x %d lattice does not fit one cooperative launch", name, nx, ny);
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
  if (per_cu * ncu < nwg) return fail(LPM_E_UNSUPPORTED, "%s: %d workgroups cannot be resident together", name, nwg);
  const int64_t nrw = 8 * N;  // tagged words: 2 parity buffers x N x 2 components x 2 words
  const unsigned zb = (unsigned)std::min<int64_t>(1024, (nrw + 255) / 256);
  hipLaunchKernelGGL(cgp_zero_kernel, dim3(zb), dim3(256), 0, st, a.slots, a.rwords, nrw, a.abort_flag);
  LAUNCHCHK();
  void* params[] = {&a};
  HIPCHK(hipLaunchCooperativeKernel(fn, dim3(nwg), dim3(nt), params, lds, st));
  int out[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(out, a.out, sizeof(out), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(relres, a.relres, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  *iters = out[0];
  *conv = out[1] == CGP_CONV ? 1 : 0;
  if (out[1] == CGP_NOTSPD) return fail(LPM_E_NOT_SPD, "%s: p.Ap <= 0 at iteration %d", name, out[0] + 1);
  if (out[1] == CGP_NONFINITE) return fail(LPM_E_NONFINITE, "%s: non-finite value at iteration %d", name, out[0] + 1);
  if (out[1] == CGP_TIMEOUT) return fail(LPM_E_HIP, "%s: grid barrier timed out", name);
  return LPM_OK;
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
  return cg2_persistent(false, "lpm_cg2_persistent", nx, ny, kn, tv, mask, b, x, tol, maxit, workgroups, work, iters,
                        relres, conv, stream);
}

int lpm_cg2_pipelined(int32_t nx, int32_t ny, const double* kn, const double* tv, const uint8_t* mask,
                      const double* b, double* x, double tol, int32_t maxit, int32_t workgroups, void* work,
                      int32_t* iters, double* relres, int32_t* conv, void* stream) {
  return cg2_persistent(true, "lpm_cg2_pipelined", nx, ny, kn, tv, mask, b, x, tol, maxit, workgroups, work, iters,
                        relres, conv, stream);
}

}  // extern "C"
