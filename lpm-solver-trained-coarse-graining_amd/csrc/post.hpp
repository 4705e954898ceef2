// Per-particle kernels outside the CG loop: shell sums, computeStress / particle energy / bond forces, adjoint
// bond-parameter gradient, block-Jacobi preconditioner, coefficient packing.  They run once per solve, so they use
// a plain two-pass layout (shell sums written to a workspace, then read back for the neighbours).
#pragma once
#include "lpm_common.hpp"

namespace lpm {

template <int DIM>
struct Lat;
template <>
struct Lat<3> {
  static constexpr int D = 3, NN = 18;
  __host__ __device__ static constexpr int off(int k, int a) { return off3(k, a); }
  __host__ __device__ static constexpr int shell(int k) { return shell3(k); }
  template <typename T>
  __host__ __device__ static constexpr T nd(int k, int a) { return nd3<T>(k, a); }
};
template <>
struct Lat<2> {
  static constexpr int D = 2, NN = 8;
  __host__ __device__ static constexpr int off(int k, int a) { return a < 2 ? off2(k, a) : 0; }
  __host__ __device__ static constexpr int shell(int k) { return shell2(k); }
  template <typename T>
  __host__ __device__ static constexpr T nd(int k, int a) { return a < 2 ? nd2<T>(k, a) : T(0); }
};

struct Box {
  int nx, ny, nz;   // nz = 1 in 2D
  int64_t ps;       // plane stride of packed coefficients (packed_stride); unused for directed ones
  // optional bond words (one uint32 per particle and system, bit k = slot k is a bond; removed particles): null =
  // every slot that stays in the box is a bond
  const uint32_t* bonds = nullptr;
  Bcast bb{0, 0};
  __device__ __forceinline__ uint32_t word(int b, int l, int64_t p) const {
    return bonds ? bonds[bb.at(b, l) + p] : 0xffffffffu;
  }
  __device__ __forceinline__ void coords(int64_t p, int& x, int& y, int& z) const {
    x = (int)(p % nx);
    const int64_t r = p / nx;
    y = (int)(r % ny);
    z = (int)(r / ny);
  }
  // neighbour of (x, y, z) in slot k, or -1 outside the box
  template <int DIM>
  __device__ __forceinline__ int64_t nbr(int x, int y, int z, int k) const {
    const int qx = x + Lat<DIM>::off(k, 0), qy = y + Lat<DIM>::off(k, 1), qz = z + Lat<DIM>::off(k, 2);
    if (qx < 0 || qx >= nx || qy < 0 || qy >= ny || qz < 0 || qz >= nz) return -1;
    return ((int64_t)qz * ny + qy) * nx + qx;
  }
  // bond partner of (x, y, z) in slot k given the particle's bond word, or -1
  template <int DIM>
  __device__ __forceinline__ int64_t nbrw(int x, int y, int z, int k, uint32_t w) const {
    return ((w >> k) & 1u) ? nbr<DIM>(x, y, z, k) : -1;
  }
};

// ----------------------------------------------------------------------------------------- shell sums
// ws (systems, N, 4) = S0, S1, TS0, TS1.  Grid (ceil(N/256), nl, nb).
template <int DIM, typename T, typename C, int MODE>
__global__ __launch_bounds__(256) void shell_sums_kernel(Box box, int nl, const T* __restrict__ u, Bcast bu,
                                                         const C* __restrict__ tv, Bcast bc, T* __restrict__ ws) {
  using L = Lat<DIM>;
  const int64_t N = (int64_t)box.nx * box.ny * box.nz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int l = blockIdx.y, b = blockIdx.z;
  const T* uu = u + bu.at(b, l);
  const C* tt = tv + bc.at(b, l);
  int x, y, z;
  box.coords(p, x, y, z);
  const uint32_t bw = box.word(b, l, p);
  T s[2] = {T(0), T(0)}, ts[2] = {T(0), T(0)};
#pragma unroll
  for (int k = 0; k < L::NN; ++k) {
    const int64_t q = box.nbrw<DIM>(x, y, z, k, bw);
    if (q < 0) continue;
    T dl = T(0);
#pragma unroll
    for (int a = 0; a < L::D; ++a)
      if (L::off(k, a)) dl += L::template nd<T>(k, a) * (uu[L::D * p + a] - uu[L::D * q + a]);
    const T t = (T)coef_at<C, MODE, L::NN>(tt, box.ps, p, q, k);
    s[L::shell(k)] += dl;
    ts[L::shell(k)] += t * dl;
  }
  T* w = ws + (((int64_t)b * nl + l) * N + p) * 4;
  w[0] = s[0];
  w[1] = s[1];
  w[2] = ts[0];
  w[3] = ts[1];
}

// ----------------------------------------------------------------------------------------- site outputs
// From u and the shell sums: any of K u (N, D), stress (N, 3|6), energy (N,), dL and F per directed slot (N, NN).
// computeStress (coarse3d/lin3d.py:170-210, coarse2d/lin2d.py:143-154): sigma_i = sum_k of_ik / V * l0_k * F_ik
// n_k n_k^T, of = 1/2 when the particle has all bonds or the opposite bond exists, else 1; Voigt order xx yy zz yz
// xz xy (3D) or xx yy xy (2D).  Energy (coarse3d/lin3d.py:213-236): U_i = 1/2 sum Kn dL^2 + 1/2 sum_c S^c TS^c.
struct PostOut {
  void* ku;
  void* stress;
  void* energy;
  void* dl;
  void* force;
};

template <int DIM, typename T, typename C, int MODE>
__global__ __launch_bounds__(256) void site_post_kernel(Box box, int nl, double h, const T* __restrict__ u, Bcast bu,
                                                        const C* __restrict__ kn, const C* __restrict__ tv, Bcast bc,
                                                        const T* __restrict__ ws, PostOut o) {
  using L = Lat<DIM>;
  constexpr int NN = L::NN, D = L::D, NS = DIM == 3 ? 6 : 3;
  const int64_t N = (int64_t)box.nx * box.ny * box.nz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int l = blockIdx.y, b = blockIdx.z;
  const int64_t sys = (int64_t)b * nl + l;
  const T* uu = u + bu.at(b, l);
  const C* kk = kn + bc.at(b, l);
  const C* tt = tv + bc.at(b, l);
  const T* w = ws + sys * N * 4;
  int x, y, z;
  box.coords(p, x, y, z);
  const uint32_t bw = box.word(b, l, p);
  int nvalid = 0;
#pragma unroll
  for (int k = 0; k < NN; ++k) nvalid += box.nbrw<DIM>(x, y, z, k, bw) >= 0;
  const bool full = nvalid == NN;
  const T si[2] = {w[4 * p + 0], w[4 * p + 1]}, ti[2] = {w[4 * p + 2], w[4 * p + 3]};
  const T vol = (T)(DIM == 3 ? h * h * h : h * h);
  T ku[3] = {T(0), T(0), T(0)}, st[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  T en = T(0);
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    const int64_t q = box.nbrw<DIM>(x, y, z, k, bw);
    T dl = T(0), f = T(0);
    if (q >= 0) {
#pragma unroll
      for (int a = 0; a < D; ++a)
        if (L::off(k, a)) dl += L::template nd<T>(k, a) * (uu[D * p + a] - uu[D * q + a]);
      const int c = L::shell(k);
      const T kv = (T)coef_at<C, MODE, NN>(kk, box.ps, p, q, k);
      const T tv_ = (T)coef_at<C, MODE, NN>(tt, box.ps, p, q, k);
      f = T(2) * kv * dl + T(0.5) * (ti[c] + w[4 * q + 2 + c]) + T(0.5) * tv_ * (si[c] + w[4 * q + c]);
      en += T(0.5) * kv * dl * dl;
#pragma unroll
      for (int a = 0; a < D; ++a) ku[a] += f * L::template nd<T>(k, a);
      const bool opp_ok = box.nbrw<DIM>(x, y, z, k ^ 1, bw) >= 0;
      const T len = (T)(k < (DIM == 3 ? 6 : 4) ? h : h * 1.4142135623730950488);
      const T wgt = ((full || opp_ok) ? T(0.5) : T(1)) / vol * len * f;
      const T n0 = L::template nd<T>(k, 0), n1 = L::template nd<T>(k, 1), n2 = L::template nd<T>(k, 2);
      if constexpr (DIM == 3) {
        st[0] += wgt * n0 * n0;
        st[1] += wgt * n1 * n1;
        st[2] += wgt * n2 * n2;
        st[3] += wgt * n1 * n2;
        st[4] += wgt * n0 * n2;
        st[5] += wgt * n0 * n1;
      } else {
        st[0] += wgt * n0 * n0;
        st[1] += wgt * n1 * n1;
        st[2] += wgt * n0 * n1;
      }
    }
    if (o.dl) ((T*)o.dl)[(sys * N + p) * NN + k] = dl;
    if (o.force) ((T*)o.force)[(sys * N + p) * NN + k] = f;
  }
  en += T(0.5) * (si[0] * ti[0] + si[1] * ti[1]);
  if (o.ku)
    for (int a = 0; a < D; ++a) ((T*)o.ku)[(sys * N + p) * D + a] = ku[a];
  if (o.stress)
    for (int a = 0; a < NS; ++a) ((T*)o.stress)[(sys * N + p) * NS + a] = st[a];
  if (o.energy) ((T*)o.energy)[sys * N + p] = en;
}

// ----------------------------------------------------------------------------------------- adjoint gradient
// For E = sum_s lam_s . (K(Kn, Tv) u_s), the derivative per directed slot (force form of the operator):
//   dE/dKn_ik = 2 dL_ik(u) mu_ik,   dE/dTv_ik = 1/2 mu_ik (S_i^c(u) + S_q^c(u)) + 1/2 dL_ik(u) S_i^c(lam),
// with mu_ik = n_k . lam_i.  This is what autograd of lam . apply(u; Kn, Tv) returns in the reference adjoint
// (coarse3d/torch_lpm3d.py:165-175).  Systems s = cb * spc + j (j < spc) share coefficient batch cb; the sum over
// them is taken in registers in a fixed order.  wsu / wsl hold the shell sums of u and lam.  Output directed
// (ncb, N, NN); grid (ceil(N/256), ncb).
template <int DIM, typename T>
__global__ __launch_bounds__(256) void bond_grad_kernel(Box box, int spc, const T* __restrict__ u,
                                                        const T* __restrict__ lam, const T* __restrict__ wsu,
                                                        const T* __restrict__ wsl, T scale, T* __restrict__ gkn,
                                                        T* __restrict__ gtv) {
  using L = Lat<DIM>;
  constexpr int NN = L::NN, D = L::D;
  const int64_t N = (int64_t)box.nx * box.ny * box.nz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int cb = blockIdx.y;
  int x, y, z;
  box.coords(p, x, y, z);
  T gk[NN], gt[NN];
#pragma unroll
  for (int k = 0; k < NN; ++k) gk[k] = gt[k] = T(0);
  for (int j = 0; j < spc; ++j) {
    const int64_t s = (int64_t)cb * spc + j;
    const T* uu = u + s * N * D;
    const T* ll = lam + s * N * D;
    const T* wu = wsu + s * N * 4;
    const T* wl = wsl + s * N * 4;
    const uint32_t bw = box.word(cb, j, p);
    T lv[3] = {T(0), T(0), T(0)}, uv[3] = {T(0), T(0), T(0)};
    for (int a = 0; a < D; ++a) {
      lv[a] = ll[D * p + a];
      uv[a] = uu[D * p + a];
    }
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      const int64_t q = box.nbrw<DIM>(x, y, z, k, bw);
      if (q < 0) continue;
      T dl = T(0), mu = T(0);
#pragma unroll
      for (int a = 0; a < D; ++a) {
        if (L::off(k, a)) {
          dl += L::template nd<T>(k, a) * (uv[a] - uu[D * q + a]);
          mu += L::template nd<T>(k, a) * lv[a];
        }
      }
      const int c = L::shell(k);
      gk[k] += T(2) * dl * mu;
      gt[k] += T(0.5) * mu * (wu[4 * p + c] + wu[4 * q + c]) + T(0.5) * dl * wl[4 * p + c];
    }
  }
  T* ok = gkn + ((int64_t)cb * N + p) * NN;
  T* ot = gtv + ((int64_t)cb * N + p) * NN;
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    ok[k] = scale * gk[k];
    ot[k] = scale * gt[k];
  }
}

// ----------------------------------------------------------------------------------------- block Jacobi
// D_p = sum_a (Kn_pa + Kn_{q,a^1} + Tv_{q,a^1}) n_a n_a^T + sum_c 1/2 (T_pc s_pc^T + s_pc T_pc^T),
// T_pc = sum_{a in c} Tv_pa n_a, s_pc = sum_{a in c} n_a over existing bonds (coarse3d/lin3d.py:239-269).
// The inverse of the free-DOF sub-block is stored (rows / columns of fixed DOFs zero), so a particle with all
// DOFs fixed gets a zero block as in coarse3d/torch_lpm3d.py:104-106.  Output symmetric: 3D (xx yy zz yz xz xy),
// 2D (xx yy xy), in the solver precision.  `raw` skips the inverse (returns D itself, full (N, D, D) layout).
template <int DIM, typename T, typename C, int MODE>
__global__ __launch_bounds__(256) void block_jacobi_kernel(Box box, int nl, const C* __restrict__ kn,
                                                           const C* __restrict__ tv, Bcast bc,
                                                           const uint8_t* __restrict__ mask, Bcast bm,
                                                           T* __restrict__ out, Bcast bo, int raw) {
  using L = Lat<DIM>;
  constexpr int NN = L::NN, D = L::D;
  const int64_t N = (int64_t)box.nx * box.ny * box.nz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int l = blockIdx.y, b = blockIdx.z;
  const C* kk = kn + bc.at(b, l);
  const C* tt = tv + bc.at(b, l);
  int x, y, z;
  box.coords(p, x, y, z);
  const uint32_t bw = box.word(b, l, p);
  double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  double Tc[2][3] = {{0, 0, 0}, {0, 0, 0}}, sc[2][3] = {{0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    const int64_t q = box.nbrw<DIM>(x, y, z, k, bw);
    if (q < 0) continue;
    const int c = L::shell(k);
    // coefficient of the reverse slot at the neighbour
    double kq, tq;
    if constexpr (MODE == COEF_DIRECTED) {
      kq = (double)kk[q * NN + (k ^ 1)];
      tq = (double)tt[q * NN + (k ^ 1)];
    } else {
      kq = (double)coef_at<C, MODE, NN>(kk, box.ps, p, q, k);  // symmetric: same bond
      tq = (double)coef_at<C, MODE, NN>(tt, box.ps, p, q, k);
    }
    const double kp = (double)coef_at<C, MODE, NN>(kk, box.ps, p, q, k);
    const double tp = (double)coef_at<C, MODE, NN>(tt, box.ps, p, q, k);
    const double cf = kp + kq + tq;
    double n[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) n[a] = L::template nd<double>(k, a);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      Tc[c][a] += tp * n[a];
      sc[c][a] += n[a];
#pragma unroll
      for (int e = 0; e < 3; ++e) M[a][e] += cf * n[a] * n[e];
    }
  }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int e = 0; e < 3; ++e) M[a][e] += 0.5 * (Tc[c][a] * sc[c][e] + sc[c][a] * Tc[c][e]);
  T* o = out + bo.at(b, l);
  if (raw) {
    for (int a = 0; a < D; ++a)
      for (int e = 0; e < D; ++e) o[p * D * D + a * D + e] = (T)M[a][e];
    return;
  }
  bool fr[3] = {true, true, DIM == 3};
  if (mask) {
    const uint8_t* mk = mask + bm.at(b, l) + D * p;
    for (int a = 0; a < D; ++a) fr[a] = mk[a] != 0;
  }
  // identity on fixed DOFs (and the unused third DOF in 2D), invert, zero fixed rows / columns
  for (int a = 0; a < 3; ++a)
    if (!fr[a])
      for (int e = 0; e < 3; ++e) M[a][e] = M[e][a] = (a == e) ? 1.0 : 0.0;
  const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
  const double c01 = -(M[1][0] * M[2][2] - M[1][2] * M[2][0]);
  const double c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  const double det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
  double I[3][3];
  I[0][0] = c00 / det;
  I[1][0] = c01 / det;
  I[2][0] = c02 / det;
  I[0][1] = -(M[0][1] * M[2][2] - M[0][2] * M[2][1]) / det;
  I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
  I[2][1] = -(M[0][0] * M[2][1] - M[0][1] * M[2][0]) / det;
  I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
  I[1][2] = -(M[0][0] * M[1][2] - M[0][2] * M[1][0]) / det;
  I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
  for (int a = 0; a < 3; ++a)
    for (int e = 0; e < 3; ++e)
      if (!fr[a] || !fr[e]) I[a][e] = 0.0;
  if constexpr (DIM == 3) {
    T* w = o + p * 6;
    w[0] = (T)I[0][0];
    w[1] = (T)I[1][1];
    w[2] = (T)I[2][2];
    w[3] = (T)(0.5 * (I[1][2] + I[2][1]));
    w[4] = (T)(0.5 * (I[0][2] + I[2][0]));
    w[5] = (T)(0.5 * (I[0][1] + I[1][0]));
  } else {
    T* w = o + p * 3;
    w[0] = (T)I[0][0];
    w[1] = (T)I[1][1];
    w[2] = (T)(0.5 * (I[0][1] + I[1][0]));
  }
}

// ----------------------------------------------------------------------------------------- packing
// Directed (nbatch, N, NN) -> forward planes (nbatch, NN/2, P) with zero guards, in the solver precision;
// asym[batch] counts the bonds whose two directed slots differ by more than rtol (relative), so the caller can
// keep the directed layout.  Grid (ceil(P/256), nbatch) over the padded plane.
template <int DIM, typename CI, typename CO>
__global__ __launch_bounds__(256) void pack_coef_kernel(Box box, const CI* __restrict__ kn, const CI* __restrict__ tv,
                                                        CO* __restrict__ kp, CO* __restrict__ tp, double rtol,
                                                        unsigned int* __restrict__ asym) {
  using L = Lat<DIM>;
  constexpr int NN = L::NN;
  const int64_t N = (int64_t)box.nx * box.ny * box.nz;
  const int64_t G = packed_guard(DIM, box.nx, box.ny);
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // position in the padded plane
  if (e >= box.ps) return;
  const int b = blockIdx.y;
  const int64_t p = e - G;
  if (p < 0 || p >= N) {
#pragma unroll
    for (int k = 0; k < NN; k += 2) {
      kp[((int64_t)b * (NN / 2) + k / 2) * box.ps + e] = CO(0);
      tp[((int64_t)b * (NN / 2) + k / 2) * box.ps + e] = CO(0);
    }
    return;
  }
  const CI* kk = kn + (int64_t)b * N * NN;
  const CI* tt = tv + (int64_t)b * N * NN;
  int x, y, z;
  box.coords(p, x, y, z);
  unsigned int bad = 0;
#pragma unroll
  for (int k = 0; k < NN; k += 2) {
    const int64_t q = box.nbr<DIM>(x, y, z, k);
    // a slot that leaves the box is no bond: its plane entry is zero (the fused K u kernels rely on it)
    const CI a1 = q >= 0 ? kk[p * NN + k] : CI(0), t1 = q >= 0 ? tt[p * NN + k] : CI(0);
    kp[((int64_t)b * (NN / 2) + k / 2) * box.ps + e] = (CO)a1;
    tp[((int64_t)b * (NN / 2) + k / 2) * box.ps + e] = (CO)t1;
    if (q >= 0) {
      const CI a2 = kk[q * NN + (k ^ 1)], t2 = tt[q * NN + (k ^ 1)];
      const double sa = fabs((double)a1) + fabs((double)a2), st = fabs((double)t1) + fabs((double)t2);
      if (fabs((double)a1 - (double)a2) > rtol * sa || fabs((double)t1 - (double)t2) > rtol * st) ++bad;
    }
  }
  if (bad) atomicAdd(&asym[b], bad);
}

// ----------------------------------------------------------------------------------------- post-processing VJP
// Vector-Jacobian product of (P = K u, sigma = computeStress, U = particle energy) for cotangents (gP, gS, gU):
// E = sum_i gP_i.P_i + gS_i:sigma_i + gU_i U_i.  With the per-slot weight c_ik = gP_i.n_k + w_ik gS_i:(n_k n_k)
// (w_ik the computeStress weight of the slot) the first two terms are sum_ik c_ik F_ik, and
//   dE/dKn_ik = 2 c_ik dL_ik + 1/2 gU_i dL_ik^2
//   dE/dTv_ik = 1/2 c_ik (S_i^c + S_q^c) + 1/2 C_i^c dL_ik + 1/2 gU_i S_i^c dL_ik
//   dE/du_i   = sum_k (W_ik + W_{q,k^1}) n_k,
//   W_jm = 2 Kn_jm c_jm + 1/2 Tv_jm C_j^c + 1/2 D_j^c + gU_j (Kn_jm dL_jm + 1/2 (Tv_jm S_j^c + TS_j^c)),
// C_j^c = sum_{k in c} (c_jk + c_{q,k^1}), D_j^c = sum_{k in c} (c_jk Tv_jk + c_{q,k^1} Tv_{q,k^1}).
// For gS = gU = 0 this is the adjoint of bond_grad_kernel with c = lam.n.  Three kernels: slot weights c, then
// per coefficient batch the parameter gradients and W (summing the systems of the batch in registers), then du.

// computeStress weight of slot k at site (x, y, z) with bond word bw: of / V * l0
template <int DIM, typename T>
__device__ __forceinline__ T stress_weight(const Box& box, int x, int y, int z, int k, double h, uint32_t bw) {
  using L = Lat<DIM>;
  int nvalid = 0;
#pragma unroll
  for (int j = 0; j < L::NN; ++j) nvalid += box.nbrw<DIM>(x, y, z, j, bw) >= 0;
  const bool opp_ok = box.nbrw<DIM>(x, y, z, k ^ 1, bw) >= 0;
  const double vol = DIM == 3 ? h * h * h : h * h;
  const double len = k < (DIM == 3 ? 6 : 4) ? h : h * 1.4142135623730950488;
  return (T)(((nvalid == L::NN || opp_ok) ? 0.5 : 1.0) / vol * len);
}

// c (systems, N, NN) from gP (systems, N, D) and gS (systems, N, 3|6); either may be null.  System s is load
// s % spc of coefficient batch s / spc.
template <int DIM, typename T>
__global__ __launch_bounds__(256) void vjp_weights_kernel(Box box, int spc, double h, const T* __restrict__ gP,
                                                          const T* __restrict__ gS, T* __restrict__ c) {
  using L = Lat<DIM>;
  constexpr int NN = L::NN, D = L::D, NS = DIM == 3 ? 6 : 3;
  const int64_t N = (int64_t)box.nx * box.ny * box.nz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int64_t s = blockIdx.y;
  int x, y, z;
  box.coords(p, x, y, z);
  const uint32_t bw = box.word((int)(s / spc), (int)(s % spc), p);
  T gp[3] = {T(0), T(0), T(0)}, gs[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  if (gP)
    for (int a = 0; a < D; ++a) gp[a] = gP[(s * N + p) * D + a];
  if (gS)
    for (int a = 0; a < NS; ++a) gs[a] = gS[(s * N + p) * NS + a];
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    T v = T(0);
    if (box.nbrw<DIM>(x, y, z, k, bw) >= 0) {
      const T n0 = L::template nd<T>(k, 0), n1 = L::template nd<T>(k, 1), n2 = L::template nd<T>(k, 2);
      v = gp[0] * n0 + gp[1] * n1 + gp[2] * n2;
      T nn;
      if constexpr (DIM == 3)
        nn = gs[0] * n0 * n0 + gs[1] * n1 * n1 + gs[2] * n2 * n2 + gs[3] * n1 * n2 + gs[4] * n0 * n2 + gs[5] * n0 * n1;
      else
        nn = gs[0] * n0 * n0 + gs[1] * n1 * n1 + gs[2] * n0 * n1;
      v += stress_weight<DIM, T>(box, x, y, z, k, h, bw) * nn;
    }
    c[(s * N + p) * NN + k] = v;
  }
}

// parameter gradients (summed over the systems s = cb * spc + j of coefficient batch cb) and the slot weights W.
// ws: shell sums of u (systems, N, 4).  Grid (ceil(N/256), ncb).
template <int DIM, typename T, int MODE>
__global__ __launch_bounds__(256) void vjp_params_kernel(Box box, int spc, const T* __restrict__ u,
                                                         const T* __restrict__ kn, const T* __restrict__ tv,
                                                         int64_t cstride, const T* __restrict__ ws,
                                                         const T* __restrict__ c, const T* __restrict__ gU,
                                                         T* __restrict__ W, T* __restrict__ gkn, T* __restrict__ gtv) {
  using L = Lat<DIM>;
  constexpr int NN = L::NN, D = L::D;
  const int64_t N = (int64_t)box.nx * box.ny * box.nz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int cb = blockIdx.y;
  const T* kk = kn + (int64_t)cb * cstride;
  const T* tt = tv + (int64_t)cb * cstride;
  int x, y, z;
  box.coords(p, x, y, z);
  int64_t q[NN];
  T knv[NN], tvv[NN], tvq[NN];
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    q[k] = box.nbr<DIM>(x, y, z, k);
    const int64_t qq = q[k] >= 0 ? q[k] : p;
    knv[k] = coef_at<T, MODE, NN>(kk, box.ps, p, qq, k);
    tvv[k] = coef_at<T, MODE, NN>(tt, box.ps, p, qq, k);
    // Tv of the reverse slot at the neighbour
    if constexpr (MODE == COEF_DIRECTED)
      tvq[k] = q[k] >= 0 ? tt[q[k] * NN + (k ^ 1)] : T(0);
    else
      tvq[k] = tvv[k];
  }
  T gk[NN], gt[NN];
#pragma unroll
  for (int k = 0; k < NN; ++k) gk[k] = gt[k] = T(0);
  for (int j = 0; j < spc; ++j) {
    const int64_t s = (int64_t)cb * spc + j;
    const T* uu = u + s * N * D;
    const T* w = ws + s * N * 4;
    const T* cs = c + s * N * NN;
    const T gu = gU ? gU[s * N + p] : T(0);
    const uint32_t bw = box.word(cb, j, p);
    T dl[NN], cc[NN], Cc[2] = {T(0), T(0)}, Dc[2] = {T(0), T(0)};
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      dl[k] = T(0);
      cc[k] = cs[p * NN + k];
      if (q[k] < 0 || !((bw >> k) & 1u)) continue;
      for (int a = 0; a < D; ++a)
        if (L::off(k, a)) dl[k] += L::template nd<T>(k, a) * (uu[D * p + a] - uu[D * q[k] + a]);
      const T cq = cs[q[k] * NN + (k ^ 1)];
      const int sh = L::shell(k);
      Cc[sh] += cc[k] + cq;
      Dc[sh] += cc[k] * tvv[k] + cq * tvq[k];
    }
    const T si[2] = {w[4 * p], w[4 * p + 1]}, ti[2] = {w[4 * p + 2], w[4 * p + 3]};
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      T Wk = T(0);
      if (q[k] >= 0 && ((bw >> k) & 1u)) {
        const int sh = L::shell(k);
        const T sq = w[4 * q[k] + sh];
        gk[k] += T(2) * cc[k] * dl[k] + T(0.5) * gu * dl[k] * dl[k];
        gt[k] += T(0.5) * cc[k] * (si[sh] + sq) + T(0.5) * Cc[sh] * dl[k] + T(0.5) * gu * si[sh] * dl[k];
        Wk = T(2) * knv[k] * cc[k] + T(0.5) * tvv[k] * Cc[sh] + T(0.5) * Dc[sh] +
             gu * (knv[k] * dl[k] + T(0.5) * (tvv[k] * si[sh] + ti[sh]));
      }
      W[(s * N + p) * NN + k] = Wk;
    }
  }
  T* ok = gkn + ((int64_t)cb * N + p) * NN;
  T* ot = gtv + ((int64_t)cb * N + p) * NN;
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    ok[k] = gk[k];
    ot[k] = gt[k];
  }
}

// du_i = sum_k (W_ik + W_{q,k^1}) n_k; grid (ceil(N/256), systems), system s = load s % spc of batch s / spc
template <int DIM, typename T>
__global__ __launch_bounds__(256) void vjp_u_kernel(Box box, int spc, const T* __restrict__ W, T* __restrict__ gu) {
  using L = Lat<DIM>;
  constexpr int NN = L::NN, D = L::D;
  const int64_t N = (int64_t)box.nx * box.ny * box.nz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int64_t s = blockIdx.y;
  int x, y, z;
  box.coords(p, x, y, z);
  const T* ws = W + s * N * NN;
  const uint32_t bw = box.word((int)(s / spc), (int)(s % spc), p);
  T g[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int k = 0; k < NN; ++k) {
    const int64_t q = box.nbrw<DIM>(x, y, z, k, bw);
    if (q < 0) continue;
    const T t = ws[p * NN + k] + ws[q * NN + (k ^ 1)];
#pragma unroll
    for (int a = 0; a < D; ++a) g[a] += t * L::template nd<T>(k, a);
  }
  for (int a = 0; a < D; ++a) gu[(s * N + p) * D + a] = g[a];
}

// ----------------------------------------------------------------------------------------- bond words
// Bond words of removed particles (cement3d/torch_v1.py:36-53): bit k of particle p is set iff slot k stays in the
// box and both p and its slot-k neighbour are solid.  solid (nbatch, N) bytes, nonzero = solid; words (nbatch, N).
template <int DIM>
__global__ __launch_bounds__(256) void bond_words_kernel(Box box, const uint8_t* __restrict__ solid,
                                                         uint32_t* __restrict__ words) {
  using L = Lat<DIM>;
  const int64_t N = (int64_t)box.nx * box.ny * box.nz;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint8_t* sd = solid + (int64_t)blockIdx.y * N;
  uint32_t w = 0;
  if (sd[p]) {
    int x, y, z;
    box.coords(p, x, y, z);
#pragma unroll
    for (int k = 0; k < L::NN; ++k) {
      const int64_t q = box.nbr<DIM>(x, y, z, k);
      if (q >= 0 && sd[q]) w |= 1u << k;
    }
  }
  words[(int64_t)blockIdx.y * N + p] = w;
}

}  // namespace lpm
