// Bond-network inference (declarations and reference citations in include/lpm.h): the O_h-invariant 3x3x3
// convolutions of the trunk as class sums staged in LDS followed by a channel mix in registers, the cell means, and
// the bond heads as one fused kernel per direction group: window gather into LDS, the stabiliser-expanded first
// layer on the FP32 matrix cores (v_mfma_f32_16x16x4f32), GELU and the mean over the stabiliser, the two small
// layers, and the Voigt-based scatter of Kn / Tv.
#include <algorithm>
#include <cmath>

#include "capi_common.hpp"

using lpmc::fail;

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu_erf(float v) { return 0.5f * v * (1.0f + erff(v * 0.70710678118654752440f)); }

// ------------------------------------------------------------------------------------------------ trunk layers
// Block = 64 x 4 threads over (x, y), each thread CT_Z = 2 consecutive z voxels (two accumulator sets of COP: 4
// waves per SIMD with 24 channels; four voxels measured slower at 2 waves per SIMD); per input channel the
// (CT_Z+2) x (4+2) x (64+2) zero-padded neighbourhood is staged in LDS (the next channel's values are loaded into
// registers while the current one is mixed).  Per plane a thread forms its in-plane centre c, face f (4 neighbours) and corner k (4
// diagonal neighbours) sums; the 3D classes of voxel z are then
//   centre = c_z, faces = f_z + c_{z-1} + c_{z+1}, edges = k_z + f_{z-1} + f_{z+1}, corners = k_{z-1} + k_{z+1}.
// The weights (cin x COP float4, zero-padded output channels) sit in LDS and are read as broadcasts.
constexpr int CT_X = 64, CT_Y = 4, CT_Z = 2;
constexpr int CT_HALO = (CT_Z + 2) * (CT_Y + 2) * (CT_X + 2);
constexpr int CT_PER = (CT_HALO + 255) / 256;
constexpr int CT_MAXCIN = 32;

template <int COP, bool PROJ>
__global__ __launch_bounds__(256) void iso_conv_kernel(int cin, int cout, int nx, int ny, int nz,
                                                       const float* __restrict__ x, const float* __restrict__ w,
                                                       const float* __restrict__ b, int gelu, int dproj,
                                                       int dstride, const float* __restrict__ pw,
                                                       const float* __restrict__ pb, float* __restrict__ y) {
  __shared__ float tile[CT_Z + 2][CT_Y + 2][CT_X + 2];
  __shared__ f4 wl[CT_MAXCIN * COP];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int x0 = blockIdx.x * CT_X, y0 = blockIdx.y * CT_Y;
  const int zt = (nz + CT_Z - 1) / CT_Z;
  const int bi = blockIdx.z / zt, z0 = (blockIdx.z % zt) * CT_Z;
  const int64_t V = (int64_t)nx * ny * nz;
  const float* xb = x + (int64_t)bi * cin * V;

  for (int i = threadIdx.x; i < cin * COP; i += 256) {
    const int ci = i / COP, co = i % COP;
    f4 v = {0.0f, 0.0f, 0.0f, 0.0f};
    if (co < cout) v = *(const f4*)(w + ((int64_t)co * cin + ci) * 4);
    wl[i] = v;
  }
  // this thread's halo elements: channel-relative offsets (-1 = outside the box, read as zero)
  int hoff[CT_PER];
#pragma unroll
  for (int r = 0; r < CT_PER; ++r) {
    const int idx = threadIdx.x + 256 * r;
    hoff[r] = -1;
    if (idx < CT_HALO) {
      const int xx = idx % (CT_X + 2), q = idx / (CT_X + 2);
      const int yy = q % (CT_Y + 2), zz = q / (CT_Y + 2);
      const int gx = x0 + xx - 1, gy = y0 + yy - 1, gz = z0 + zz - 1;
      if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) hoff[r] = (gz * ny + gy) * nx + gx;
    }
  }
  float hv[CT_PER];
#pragma unroll
  for (int r = 0; r < CT_PER; ++r) hv[r] = hoff[r] >= 0 ? xb[hoff[r]] : 0.0f;

  float acc[CT_Z][COP];
#pragma unroll
  for (int co = 0; co < COP; ++co) {
    const float bias = co < cout ? b[co] : 0.0f;
#pragma unroll
    for (int j = 0; j < CT_Z; ++j) acc[j][co] = bias;
  }
  float* tflat = &tile[0][0][0];

  for (int ci = 0; ci < cin; ++ci) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < CT_PER; ++r) {
      const int idx = threadIdx.x + 256 * r;
      if (idx < CT_HALO) tflat[idx] = hv[r];
    }
    __syncthreads();
    if (ci + 1 < cin) {
      const float* xn = xb + (int64_t)(ci + 1) * V;
#pragma unroll
      for (int r = 0; r < CT_PER; ++r) hv[r] = hoff[r] >= 0 ? xn[hoff[r]] : 0.0f;
    }
    float c[CT_Z + 2], f[CT_Z + 2], k[CT_Z + 2];
#pragma unroll
    for (int p = 0; p < CT_Z + 2; ++p) {
      c[p] = tile[p][ty + 1][tx + 1];
      f[p] = (tile[p][ty + 1][tx] + tile[p][ty + 1][tx + 2]) + (tile[p][ty][tx + 1] + tile[p][ty + 2][tx + 1]);
      k[p] = (tile[p][ty][tx] + tile[p][ty][tx + 2]) + (tile[p][ty + 2][tx] + tile[p][ty + 2][tx + 2]);
    }
    const f4* wc = wl + ci * COP;
#pragma unroll
    for (int co = 0; co < COP; ++co) {
      const f4 wv = wc[co];
#pragma unroll
      for (int j = 0; j < CT_Z; ++j) {
        const float sc = c[j + 1];
        const float sf = f[j + 1] + (c[j] + c[j + 2]);
        const float se = k[j + 1] + (f[j] + f[j + 2]);
        const float sk = k[j] + k[j + 2];
        float a = acc[j][co];
        a = fmaf(wv[0], sc, a);
        a = fmaf(wv[1], sf, a);
        a = fmaf(wv[2], se, a);
        a = fmaf(wv[3], sk, a);
        acc[j][co] = a;
      }
    }
  }

  const int gx = x0 + tx, gy = y0 + ty;
  if (gx >= nx || gy >= ny) return;
#pragma unroll
  for (int j = 0; j < CT_Z; ++j) {
    const int gz = z0 + j;
    if (gz >= nz) break;
    const int64_t vox = ((int64_t)gz * ny + gy) * nx + gx;
    float a[COP];
#pragma unroll
    for (int co = 0; co < COP; ++co) a[co] = gelu ? gelu_erf(acc[j][co]) : acc[j][co];
    if constexpr (PROJ) {
      float* yo = y + ((int64_t)bi * V + vox) * dstride;
      for (int d = 0; d < dstride; ++d) {
        float v = 0.0f;
        if (d < dproj) {
          v = pb[d];
#pragma unroll
          for (int co = 0; co < COP; ++co)
            if (co < cout) v = fmaf(pw[d * cout + co], a[co], v);
        }
        yo[d] = v;
      }
    } else {
#pragma unroll
      for (int co = 0; co < COP; ++co)
        if (co < cout) y[((int64_t)bi * cout + co) * V + vox] = a[co];
    }
  }
}

template <int COP>
int launch_iso(int B, int cin, int cout, int nx, int ny, int nz, const float* x, const float* w, const float* b,
               int gelu, int dproj, int dstride, const float* pw, const float* pb, float* y, hipStream_t st) {
  dim3 grid((nx + CT_X - 1) / CT_X, (ny + CT_Y - 1) / CT_Y, (unsigned)(B * ((nz + CT_Z - 1) / CT_Z)));
  if (dproj > 0)
    iso_conv_kernel<COP, true><<<grid, 256, 0, st>>>(cin, cout, nx, ny, nz, x, w, b, gelu, dproj, dstride, pw, pb, y);
  else
    iso_conv_kernel<COP, false><<<grid, 256, 0, st>>>(cin, cout, nx, ny, nz, x, w, b, gelu, 0, 0, nullptr, nullptr,
                                                       y);
  LAUNCHCHK();
  return LPM_OK;
}

__global__ void cell_pool_kernel(int B, int ncp, int s, int ds, const float* __restrict__ fv, float* __restrict__ fc) {
  const int64_t nce = (int64_t)ncp * ncp * ncp;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)B * nce * ds) return;
  const int d = (int)(idx % ds);
  const int64_t cell = (idx / ds) % nce, bi = idx / (ds * nce);
  const int cx = (int)(cell % ncp), cy = (int)((cell / ncp) % ncp), cz = (int)(cell / ((int64_t)ncp * ncp));
  const int64_t npd = (int64_t)ncp * s;
  const float* src = fv + bi * npd * npd * npd * ds;
  float sum = 0.0f;
  for (int a = 0; a < s; ++a)
    for (int bb = 0; bb < s; ++bb)
      for (int c = 0; c < s; ++c)
        sum += src[(((int64_t)(cz * s + a) * npd + (cy * s + bb)) * npd + (cx * s + c)) * ds + d];
  fc[idx] = sum / (float)(s * s * s);
}

// ------------------------------------------------------------------------------------------------- bond heads
// A block takes HB = 16 consecutive bonds of one group.  The first layer is a GEMM whose rows are the HB * H
// (bond, stabiliser element) pairs, K = nslot * dstride window features and N = hid (padded to NT = 1, 2 or 4
// column tiles of 16): row (b, g) reads window slot inv[g, j] where W1 expects slot j.  Four waves, each RT = H / 4
// row tiles of 16 rows and all NT column tiles.  A window slot is Q = dstride / 4 k-steps of 4; k-step j of lane
// group c = lane / 16 is channel c Q + j, so per slot a lane makes RT x Q LDS reads (A) and NT Q-wide global loads
// (B) for RT x NT x Q MFMAs, and the fragments of two slots are in flight ahead of the MFMAs that use them.
// v_mfma_f32_16x16x4f32 lane layout: A[lane % 16][lane / 16], B[lane / 16][lane % 16], D[4 (lane / 16) + i][lane % 16].
// W1 comes pre-arranged (lpm_bond_w1_layout, [slot][column tile][lane][j]) so that a wave's B load of one column
// tile and slot is one contiguous 64 x Q x 4-byte run.
constexpr int HB = 16;
constexpr int HMAX = 64;

// window slots rounded up to even (the first-layer loop runs two slots per iteration)
__host__ __device__ inline int kpad_of(int nslot) { return nslot + (nslot & 1); }
// Slot pairs of the split-BF16 W1 fragments held in memory: the pair count rounded up to even (the heads loop reads
// pairs two at a time) plus the two zero pairs its two-deep prefetch reads past the end.
__host__ __device__ inline int split_pairs_of(int nslot) {
  const int npair = (nslot + 1) / 2;
  return npair + (npair & 1) + 2;
}

struct HeadArgs {
  int ndir, nv, ncell, nslot, H, D, ds, hid, Q, xs_floats;
  int nc, s, npd, ncp, B;
  int dk[6], dx[6], dy[6], dz[6];
  int64_t nb;
  int64_t cnt[7];
  const int* voff;
  const int* coff;
  const int* inv;
  const float* fv;
  const float* fc;
  const float* w1f;
  const float* b1;
  const float* w2;
  const float* b2;
  const float* w3;
  const float* b3;
  float* out;
  const double* cp;
  double* kn;
  double* tv;
};

struct BondMeta {
  int dir, vbase, cbase, p, q, k;
};

// LDS of a heads block: the window area (xs_floats floats; hidden layers and W2 after the first layer), then the
// inverse stabiliser perms [H][nslot], the window offsets [ndir][nv] and [ndir][ncell], and the bonds' meta [HB]
struct HeadSmem {
  float* xs;
  int* inv_s;
  int* voff_s;
  int* coff_s;
  BondMeta* meta;
};

// bond list decode (direction by direction, x fastest) and the tables into LDS
template <int H>
__device__ __forceinline__ HeadSmem heads_setup(const HeadArgs& a, float* smem, int tid, int64_t b0) {
  HeadSmem sm;
  sm.xs = smem;
  sm.inv_s = (int*)(smem + a.xs_floats);
  sm.voff_s = sm.inv_s + H * a.nslot;
  sm.coff_s = sm.voff_s + a.ndir * a.nv;
  sm.meta = (BondMeta*)(sm.coff_s + a.ndir * a.ncell);
  if (tid < HB) {
    BondMeta m{-1, 0, 0, 0, 0, 0};
    const int64_t gb = b0 + tid;
    if (gb < a.nb) {
      int d = 0;
      while (gb >= a.cnt[d + 1]) ++d;
      const int ox = a.dx[d], oy = a.dy[d], oz = a.dz[d];
      const int ex = a.nc - abs(ox), ey = a.nc - abs(oy);
      const int64_t loc = gb - a.cnt[d];
      const int ix = (int)(loc % ex) + (ox < 0), iy = (int)((loc / ex) % ey) + (oy < 0);
      const int iz = (int)(loc / ((int64_t)ex * ey)) + (oz < 0);
      m.dir = d;
      m.k = a.dk[d];
      m.vbase = (((iz + 2) * a.s) * a.npd + (iy + 2) * a.s) * a.npd + (ix + 2) * a.s;
      m.cbase = ((iz + 2) * a.ncp + (iy + 2)) * a.ncp + (ix + 2);
      m.p = (iz * a.nc + iy) * a.nc + ix;
      m.q = ((iz + oz) * a.nc + (iy + oy)) * a.nc + (ix + ox);
    }
    sm.meta[tid] = m;
  }
  for (int i = tid; i < H * a.nslot; i += 256) sm.inv_s[i] = a.inv[i];
  for (int i = tid; i < a.ndir * a.nv; i += 256) sm.voff_s[i] = a.voff[i];
  for (int i = tid; i < a.ndir * a.ncell; i += 256) sm.coff_s[i] = a.coff[i];
  __syncthreads();
  return sm;
}

// window gather: HB x nslot x (ds / 4) float4 with GB independent loads in flight per thread; store(e, v) puts the
// four channels 4c..4c+3 of (bond, slot) at element e = (bond nslot + slot) ds + 4c of the window area
template <class Store>
__device__ __forceinline__ void heads_gather(const HeadArgs& a, const HeadSmem& sm, int tid, int bi, Store store) {
  constexpr int GB = 8;
  const int nslot = a.nslot, ds = a.ds;
  const int64_t nvox = (int64_t)a.npd * a.npd * a.npd, ncel = (int64_t)a.ncp * a.ncp * a.ncp;
  const float* fvb = a.fv + (int64_t)bi * nvox * ds;
  const float* fcb = a.fc + (int64_t)bi * ncel * ds;
  const int total = HB * nslot * a.Q;
  for (int i0 = 0; i0 < total; i0 += 256 * GB) {
    f4 v[GB];
    int dst[GB];
#pragma unroll
    for (int u = 0; u < GB; ++u) {
      const int idx = i0 + u * 256 + tid;
      v[u] = f4{0.0f, 0.0f, 0.0f, 0.0f};
      dst[u] = -1;
      if (idx < total) {
        const int c = idx % a.Q, r = idx / a.Q;
        const int slot = r % nslot, bl = r / nslot;
        const BondMeta m = sm.meta[bl];
        dst[u] = (bl * nslot + slot) * ds + 4 * c;
        if (m.dir >= 0) {
          const float* src = slot < a.nv ? fvb + (int64_t)(m.vbase + sm.voff_s[m.dir * a.nv + slot]) * ds
                                         : fcb + (int64_t)(m.cbase + sm.coff_s[m.dir * a.ncell + slot - a.nv]) * ds;
          v[u] = *(const f4*)(src + 4 * c);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < GB; ++u)
      if (dst[u] >= 0) store(dst[u], v[u]);
  }
  __syncthreads();
}

// after the first layer: + b1, GELU, mean over the stabiliser, second layer (W2 from LDS), third layer, raw outputs
// and, with cp, Kn / Tv on both slots of every bond.  acc: the wave's RT x NT tiles of first-layer sums.
template <int H, int NT>
__device__ __forceinline__ void heads_epilogue(const HeadArgs& a, const HeadSmem& sm, const f4 (&acc)[H / 4][NT],
                                               int tid, int bi, int64_t b0) {
  constexpr int RT = H / 4;
  const int lane = tid & 63, wave = tid >> 6;
  float* hbuf = sm.xs;                                          // [HB][HMAX]
  float* h2buf = sm.xs + HB * HMAX;                             // [HB][HMAX]
  float* obuf = sm.xs + 2 * HB * HMAX;                          // [HB][2]
  float* w2s = sm.xs + 2 * HB * HMAX + 2 * HB;                  // [HMAX][HMAX]: w2s[m][j] = W2[j][m]
  __syncthreads();  // the window area is reused from here on
  for (int i = tid; i < a.hid * a.hid; i += 256) w2s[(i % a.hid) * HMAX + i / a.hid] = a.w2[i];
  {
    const float invH = 1.0f / (float)H;
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const int col = n * 16 + (lane & 15);
        const float bias = col < a.hid ? a.b1[col] : 0.0f;
        float sum = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) sum += gelu_erf(acc[t][n][i] + bias);
        sum += __shfl_xor(sum, 16);
        if constexpr (H == 16) {
          sum += __shfl_xor(sum, 32);
          if ((lane >> 4) == 0 && col < a.hid) hbuf[(wave * 4 + t) * HMAX + col] = sum * invH;
        } else {
          if ((lane & 16) == 0 && col < a.hid) hbuf[(wave * 4 + t * 2 + (lane >> 5)) * HMAX + col] = sum * invH;
        }
      }
  }
  __syncthreads();
  {
    // 16 threads per bond, each the outputs j0, j0 + 16, j0 + 32, j0 + 48 as independent chains
    constexpr int NQ = HMAX / 16;
    const int bl = tid >> 4, j0 = tid & 15;
    float v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = j0 + 16 * q < a.hid ? a.b2[j0 + 16 * q] : 0.0f;
    for (int m = 0; m < a.hid; ++m) {
      const float hm = hbuf[bl * HMAX + m];
      const float* wm = w2s + m * HMAX + j0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v[q] = fmaf(wm[16 * q], hm, v[q]);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      if (j0 + 16 * q < a.hid) h2buf[bl * HMAX + j0 + 16 * q] = gelu_erf(v[q]);
  }
  __syncthreads();
  if (tid < 2 * HB) {
    const int bl = tid >> 1, c = tid & 1;
    float v = a.b3[c];
    const float* wr = a.w3 + c * a.hid;
    for (int m = 0; m < a.hid; ++m) v = fmaf(wr[m], h2buf[bl * HMAX + m], v);
    obuf[bl * 2 + c] = v;
    if (sm.meta[bl].dir >= 0) a.out[((int64_t)bi * a.nb + b0 + bl) * 2 + c] = v;
  }
  if (!a.cp) return;
  __syncthreads();
  if (tid < HB && sm.meta[tid].dir >= 0) {
    const BondMeta m = sm.meta[tid];
    const int64_t nce = (int64_t)a.nc * a.nc * a.nc;
    const double* cpb = a.cp + (int64_t)bi * nce * 10;
    const double kv = 0.5 * (cpb[(int64_t)m.p * 10 + m.k / 2] + cpb[(int64_t)m.q * 10 + m.k / 2]);
    const double tb = 0.5 * (cpb[(int64_t)m.p * 10 + 9] + cpb[(int64_t)m.q * 10 + 9]);
    const double knv = kv * exp((double)obuf[tid * 2]);
    const double tvv = tb + kv * (double)obuf[tid * 2 + 1];
    double* knb = a.kn + (int64_t)bi * nce * 18;
    double* tvb = a.tv + (int64_t)bi * nce * 18;
    knb[(int64_t)m.p * 18 + m.k] = knv;
    knb[(int64_t)m.q * 18 + (m.k ^ 1)] = knv;
    tvb[(int64_t)m.p * 18 + m.k] = tvv;
    tvb[(int64_t)m.q * 18 + (m.k ^ 1)] = tvv;
  }
}

template <int H, int NT, int Q>
__global__ __launch_bounds__(256) void bond_heads_kernel(HeadArgs a) {
  extern __shared__ __align__(16) float smem[];
  static_assert((H == 8 || H == 16) && (NT == 1 || NT == 2 || NT == 4) && Q >= 1 && Q <= 4, "tile shape");
  constexpr int RT = H / 4;
  const int nslot = a.nslot, ds = a.ds;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int bi = blockIdx.y;
  const int64_t b0 = (int64_t)blockIdx.x * HB;
  const HeadSmem sm = heads_setup<H>(a, smem, tid, b0);
  float* xs = sm.xs;                                            // [HB][nslot][ds]
  const int* inv_s = sm.inv_s;
  heads_gather(a, sm, tid, bi, [&](int e, const f4& v) { *(f4*)(xs + e) = v; });

  f4 acc[RT][NT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[t][n] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  // rows of tile t: wave * RT * 16 + t * 16 + lane % 16 -> bond row / H, stabiliser element row % H; within a
  // window slot, k-step j of lane group c = lane / 16 is channel c Q + j, so a lane's Q values of a slot are
  // contiguous in LDS (A) and in the fragment-ordered W1 (B: one Q-wide load per column tile)
  int xrow[RT], grow[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int row = wave * RT * 16 + t * 16 + (lane & 15);
    xrow[t] = (row / H) * nslot * ds + (lane >> 4) * Q;
    grow[t] = (row % H) * nslot;
  }
  const float* w1l = a.w1f + lane * Q;
  // slots rounded up to even; W1 holds zero fragments for the slots past nslot and two more beyond, and A reads
  // past the last slot repeat it, so the loop has no conditional load
  const int nsl = kpad_of(nslot);
  int lsl = 0;
  auto load = [&](float (&bb)[NT][Q], float (&aa)[RT][Q]) {
    const float* wk = w1l + (int64_t)lsl * NT * 64 * Q;
#pragma unroll
    for (int n = 0; n < NT; ++n)
#pragma unroll
      for (int j = 0; j < Q; ++j) bb[n][j] = wk[n * 64 * Q + j];
    const int sl = lsl < nslot ? lsl : nslot - 1;
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      const int base = xrow[t] + inv_s[grow[t] + sl] * ds;
#pragma unroll
      for (int j = 0; j < Q; ++j) aa[t][j] = xs[base + j];
    }
    ++lsl;
  };
  auto step = [&](const float (&bb)[NT][Q], const float (&aa)[RT][Q]) {
#pragma unroll
    for (int j = 0; j < Q; ++j)
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int n = 0; n < NT; ++n)
          acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x4f32(aa[t][j], bb[n][j], acc[t][n], 0, 0, 0);
  };
  float fb0[NT][Q], fa0[RT][Q], fb1[NT][Q], fa1[RT][Q];
  load(fb0, fa0);
  load(fb1, fa1);
  for (int sl = 0; sl < nsl; sl += 2) {
    step(fb0, fa0);
    load(fb0, fa0);
    __builtin_amdgcn_sched_barrier(0);  // keep each slot's loads between the two halves
    step(fb1, fa1);
    load(fb1, fa1);
    __builtin_amdgcn_sched_barrier(0);
  }
  heads_epilogue<H, NT>(a, sm, acc, tid, bi, b0);
}

// ---- first layer in split BF16 (opt-in): every FP32 operand x = hi + lo with hi = bf16(x), lo = bf16(x - hi), and
// W1 x = Whi xhi + Whi xlo + Wlo xhi accumulated in FP32 by v_mfma_f32_16x16x32_bf16 (lane layout A[lane % 16][8
// (lane / 16) + i], B[8 (lane / 16) + i][lane % 16], D as above).  One instruction covers two window slots of 12
// channels: lane group c reads slot 2 p + (c >> 1), channels 8 (c & 1) + 0..7 (channels 12..15 are zero).  The
// window is stored as hi and lo BF16 planes (the same bytes as FP32); W1 comes split and fragment-ordered
// (lpm_bond_w1_split_layout, [slot pair][column tile][lane] x 8 BF16).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

template <int H, int NT>
__global__ __launch_bounds__(256) void bond_heads_split_kernel(HeadArgs a, const bf16x8* __restrict__ wh,
                                                               const bf16x8* __restrict__ wl) {
  extern __shared__ __align__(16) float smem[];
  static_assert((H == 8 || H == 16) && (NT == 1 || NT == 2 || NT == 4), "tile shape");
  constexpr int RT = H / 4;
  constexpr int DS = 12;
  const int nslot = a.nslot;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int bi = blockIdx.y;
  const int64_t b0 = (int64_t)blockIdx.x * HB;
  const HeadSmem sm = heads_setup<H>(a, smem, tid, b0);
  __bf16* xh = (__bf16*)sm.xs;                                  // [HB][nslot][12] hi
  __bf16* xl = xh + HB * nslot * DS;                            // [HB][nslot][12] lo
  const int* inv_s = sm.inv_s;
  heads_gather(a, sm, tid, bi, [&](int e, const f4& v) {
    bf16x4 h, l;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const __bf16 hb = (__bf16)v[i];
      h[i] = hb;
      l[i] = (__bf16)(v[i] - (float)hb);
    }
    *(bf16x4*)(xh + e) = h;
    *(bf16x4*)(xl + e) = l;
  });

  f4 acc[RT][NT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[t][n] = f4{0.0f, 0.0f, 0.0f, 0.0f};
  const int half = (lane >> 4) & 1, sub = lane >> 5;
  int xrow[RT], grow[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int row = wave * RT * 16 + t * 16 + (lane & 15);
    xrow[t] = (row / H) * nslot * DS + 8 * half;
    grow[t] = (row % H) * nslot;
  }
  const int npair = (nslot + 1) / 2;
  int lp = 0;
  // W1 fragments two slot pairs ahead; window fragments at their step.  The loop makes 2 + 2 ceil(npair / 2)
  // fragment loads, so the last one reads pair npair + (npair & 1) + 1 < split_pairs_of(nslot): W1 holds the pairs
  // rounded up to even plus two zero pairs past the end.
  auto loadb = [&](bf16x8 (&bh)[NT], bf16x8 (&bl)[NT]) {
    const int64_t wo = (int64_t)lp * NT * 64 + lane;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      bh[n] = wh[wo + n * 64];
      bl[n] = wl[wo + n * 64];
    }
    ++lp;
  };
  auto loada = [&](int pr, int t, bf16x8& ah, bf16x8& al) {
    const int slot = 2 * pr + sub;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      ah[i] = (__bf16)0.0f;
      al[i] = (__bf16)0.0f;
    }
    if (slot < nslot) {                                         // an odd last pair has no second slot
      const int base = xrow[t] + inv_s[grow[t] + slot] * DS;
      const bf16x4 h0 = *(const bf16x4*)(xh + base), l0 = *(const bf16x4*)(xl + base);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ah[i] = h0[i];
        al[i] = l0[i];
      }
      if (half == 0) {
        const bf16x4 h1 = *(const bf16x4*)(xh + base + 4), l1 = *(const bf16x4*)(xl + base + 4);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          ah[4 + i] = h1[i];
          al[4 + i] = l1[i];
        }
      }
    }
  };
  auto step = [&](int pr, const bf16x8 (&bh)[NT], const bf16x8 (&bl)[NT]) {
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      bf16x8 ah, al;
      loada(pr, t, ah, al);
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[n], acc[t][n], 0, 0, 0);
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[n], acc[t][n], 0, 0, 0);
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[t][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[n], acc[t][n], 0, 0, 0);
    }
  };
  bf16x8 bh0[NT], bl0[NT], bh1[NT], bl1[NT];
  loadb(bh0, bl0);
  loadb(bh1, bl1);
  for (int pr = 0; pr < npair; pr += 2) {
    step(pr, bh0, bl0);
    loadb(bh0, bl0);
    __builtin_amdgcn_sched_barrier(0);
    if (pr + 1 < npair) step(pr + 1, bh1, bl1);
    loadb(bh1, bl1);
    __builtin_amdgcn_sched_barrier(0);
  }
  heads_epilogue<H, NT>(a, sm, acc, tid, bi, b0);
}

// W1 split into hi / lo BF16 fragments: [slot pair][column tile][lane] x 8, split_pairs_of(nslot) pairs (zero past
// the last slot)
__global__ void w1_split_kernel(int nslot, int NT, int hid, const float* __restrict__ w1, bf16x8* __restrict__ wh,
                                bf16x8* __restrict__ wl) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)split_pairs_of(nslot) * NT * 64) return;
  const int lane = (int)(idx & 63);
  const int64_t r = idx >> 6;
  const int n = (int)(r % NT), pr = (int)(r / NT);
  const int col = n * 16 + (lane & 15);
  const int slot = 2 * pr + (lane >> 5);
  bf16x8 h, l;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int ch = 8 * ((lane >> 4) & 1) + i;
    const float v = (slot < nslot && col < hid && ch < 12) ? w1[(int64_t)col * nslot * 12 + slot * 12 + ch] : 0.0f;
    const __bf16 hb = (__bf16)v;
    h[i] = hb;
    l[i] = (__bf16)(v - (float)hb);
  }
  wh[idx] = h;
  wl[idx] = l;
}

// fragment order [slot][column tile][lane][j]: lane's k-step j of a slot is channel (lane / 16) Q + j
__global__ void w1_layout_kernel(int nslot, int D, int Q, int NT, int hid, const float* __restrict__ w1,
                                 float* __restrict__ w1f) {
  const int64_t total = (int64_t)(kpad_of(nslot) + 2) * NT * 64 * Q;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int j = (int)(idx % Q);
  int64_t r = idx / Q;
  const int lane = (int)(r & 63);
  r >>= 6;
  const int n = (int)(r % NT);
  const int slot = (int)(r / NT);
  const int row = n * 16 + (lane & 15), d = (lane >> 4) * Q + j;
  w1f[idx] = (slot < nslot && row < hid && d < D) ? w1[(int64_t)row * nslot * D + (int64_t)slot * D + d] : 0.0f;
}

// column tiles of the first layer: 1, 2 or 4 (hidden widths above 32 are padded to 64 with zero weights)
inline int col_tiles(int hid) { return hid <= 16 ? 1 : hid <= 32 ? 2 : 4; }

int check_group(const lpm_bond_group* g) {
  if (!g) return fail(LPM_E_BADARG, "bond heads: null group");
  if (!((g->ndir == 3 && g->H == 16) || (g->ndir == 6 && g->H == 8)))
    return fail(LPM_E_BADARG, "bond heads: %d directions with stabiliser order %d", g->ndir, g->H);
  if (g->nv < 1 || g->ncell < 1 || g->nv + g->ncell > 512) return fail(LPM_E_BADARG, "bond heads: window slots");
  if (g->D < 1 || g->dstride < g->D || g->dstride % 4 || g->dstride > 16)
    return fail(LPM_E_BADARG, "bond heads: %d channels with stride %d", g->D, g->dstride);
  if (g->hid < 1 || g->hid > HMAX) return fail(LPM_E_BADARG, "bond heads: hidden width %d (1..%d)", g->hid, HMAX);
  if (!g->voff || !g->coff || !g->inv) return fail(LPM_E_BADARG, "bond heads: null table");
  for (int d = 0; d < g->ndir; ++d) {
    const int k = g->slot[d];
    if (k < 0 || k >= 18 || (k & 1) || (g->H == 16) != (k < 6))
      return fail(LPM_E_BADARG, "bond heads: forward slot %d in a group of stabiliser order %d", k, g->H);
  }
  return LPM_OK;
}

// floats of the LDS area that holds the window features during the GEMM and the hidden layers after it
int head_xs_floats(const lpm_bond_group* g) {
  return std::max(HB * (g->nv + g->ncell) * g->dstride, 2 * HB * HMAX + 2 * HB + HMAX * HMAX);
}

size_t head_lds(const lpm_bond_group* g) {
  return head_xs_floats(g) * sizeof(float) + (size_t)g->H * (g->nv + g->ncell) * sizeof(int) +
         (size_t)g->ndir * (g->nv + g->ncell) * sizeof(int) + HB * sizeof(BondMeta);
}

// slot offsets of the SC lattice (coarse3d/geom3d.py:16-20)
constexpr int SC_OFF[18][3] = {{1, 0, 0},  {-1, 0, 0},  {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                               {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1},
                               {1, 0, -1}, {-1, 0, 1},  {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};

template <int H, int NT, int Q>
int launch_heads(const HeadArgs& a, size_t lds, hipStream_t st) {
  dim3 grid((unsigned)((a.nb + HB - 1) / HB), (unsigned)a.B);
  bond_heads_kernel<H, NT, Q><<<grid, 256, lds, st>>>(a);
  LAUNCHCHK();
  return LPM_OK;
}

template <int H, int NT>
int launch_heads_q(const HeadArgs& a, size_t lds, hipStream_t st) {
  switch (a.Q) {
    case 1: return launch_heads<H, NT, 1>(a, lds, st);
    case 2: return launch_heads<H, NT, 2>(a, lds, st);
    case 3: return launch_heads<H, NT, 3>(a, lds, st);
    default: return launch_heads<H, NT, 4>(a, lds, st);
  }
}

template <int H>
int launch_heads_nt(const HeadArgs& a, size_t lds, hipStream_t st) {
  const int nt = (a.hid + 15) / 16;
  if (nt == 1) return launch_heads_q<H, 1>(a, lds, st);
  if (nt == 2) return launch_heads_q<H, 2>(a, lds, st);
  return launch_heads_q<H, 4>(a, lds, st);
}

// checks and kernel arguments shared by the FP32 and the split-BF16 heads
int head_args(const lpm_bond_group* g, int32_t B, int32_t nc, int32_t s, const float* fv, const float* fc,
              const float* w1, const float* b1, const float* w2, const float* b2, const float* w3, const float* b3,
              float* out, const double* cp, double* kn, double* tv, const char* who, HeadArgs& a, size_t& lds) {
  if (int rc = check_group(g)) return rc;
  if (B < 1 || B > 65535 || nc < 2 || s < 1 || !fv || !fc || !w1 || !b1 || !w2 || !b2 || !w3 || !b3 || !out)
    return fail(LPM_E_BADARG, "%s: bad arguments", who);
  if (cp && (!kn || !tv)) return fail(LPM_E_BADARG, "%s: cp given without kn / tv", who);
  const int64_t npd = (int64_t)(nc + 4) * s;
  if (npd * npd * npd >= ((int64_t)1 << 31)) return fail(LPM_E_BADARG, "%s: padded grid too large", who);
  lds = head_lds(g);
  if (lds > 65536) return fail(LPM_E_BADARG, "%s: window needs %zu bytes of LDS (64 KiB)", who, lds);
  a = HeadArgs{};
  a.ndir = g->ndir;
  a.nv = g->nv;
  a.ncell = g->ncell;
  a.nslot = g->nv + g->ncell;
  a.H = g->H;
  a.D = g->D;
  a.ds = g->dstride;
  a.hid = g->hid;
  a.Q = g->dstride / 4;
  a.nc = nc;
  a.s = s;
  a.npd = (int)npd;
  a.ncp = nc + 4;
  a.B = B;
  a.xs_floats = head_xs_floats(g);
  a.cnt[0] = 0;
  for (int d = 0; d < 6; ++d) {
    const int k = d < g->ndir ? g->slot[d] : 0;
    a.dk[d] = k;
    a.dx[d] = SC_OFF[k][0];
    a.dy[d] = SC_OFF[k][1];
    a.dz[d] = SC_OFF[k][2];
  }
  // bond counts per direction: the cells whose neighbour stays in the box
  for (int d = 0; d < g->ndir; ++d)
    a.cnt[d + 1] = a.cnt[d] + (int64_t)(nc - std::abs(a.dx[d])) * (nc - std::abs(a.dy[d])) * (nc - std::abs(a.dz[d]));
  for (int d = g->ndir + 1; d < 7; ++d) a.cnt[d] = a.cnt[g->ndir];
  a.nb = a.cnt[g->ndir];
  a.voff = g->voff;
  a.coff = g->coff;
  a.inv = g->inv;
  a.fv = fv;
  a.fc = fc;
  a.w1f = w1;
  a.b1 = b1;
  a.w2 = w2;
  a.b2 = b2;
  a.w3 = w3;
  a.b3 = b3;
  a.out = out;
  a.cp = cp;
  a.kn = kn;
  a.tv = tv;
  return LPM_OK;
}

}  // namespace

extern "C" {

int lpm_iso_conv(int32_t B, int32_t cin, int32_t cout, int32_t nx, int32_t ny, int32_t nz, const float* x,
                 const float* w, const float* b, int32_t gelu, int32_t dproj, int32_t dstride, const float* pw,
                 const float* pb, float* y, void* stream) {
  if (B < 1 || cin < 1 || cin > CT_MAXCIN || cout < 1 || cout > 32 || nx < 1 || ny < 1 || nz < 1 || !x || !w || !b ||
      !y)
    return fail(LPM_E_BADARG, "lpm_iso_conv: bad arguments (cin %d, cout %d: 1..32)", cin, cout);
  if ((int64_t)nx * ny * nz >= ((int64_t)1 << 31)) return fail(LPM_E_BADARG, "lpm_iso_conv: grid too large");
  if (dproj < 0 || (dproj > 0 && (dstride < dproj || dstride > 16 || !pw || !pb)))
    return fail(LPM_E_BADARG, "lpm_iso_conv: projection to %d channels with stride %d", dproj, dstride);
  if ((int64_t)B * ((nz + CT_Z - 1) / CT_Z) > 65535) return fail(LPM_E_BADARG, "lpm_iso_conv: grid too deep");
  hipStream_t st = (hipStream_t)stream;
  if (cout <= 8) return launch_iso<8>(B, cin, cout, nx, ny, nz, x, w, b, gelu, dproj, dstride, pw, pb, y, st);
  if (cout <= 16) return launch_iso<16>(B, cin, cout, nx, ny, nz, x, w, b, gelu, dproj, dstride, pw, pb, y, st);
  if (cout <= 24) return launch_iso<24>(B, cin, cout, nx, ny, nz, x, w, b, gelu, dproj, dstride, pw, pb, y, st);
  return launch_iso<32>(B, cin, cout, nx, ny, nz, x, w, b, gelu, dproj, dstride, pw, pb, y, st);
}

int lpm_cell_pool(int32_t B, int32_t ncp, int32_t s, int32_t dstride, const float* fv, float* fc, void* stream) {
  if (B < 1 || ncp < 1 || s < 1 || dstride < 1 || !fv || !fc) return fail(LPM_E_BADARG, "lpm_cell_pool: bad arguments");
  const int64_t n = (int64_t)B * ncp * ncp * ncp * dstride;
  cell_pool_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(B, ncp, s, dstride, fv, fc);
  LAUNCHCHK();
  return LPM_OK;
}

size_t lpm_bond_w1_size(const lpm_bond_group* g) {
  if (!g || g->dstride < 4 || g->hid < 1) return 0;
  return (size_t)(kpad_of(g->nv + g->ncell) + 2) * col_tiles(g->hid) * 64 * (g->dstride / 4) * sizeof(float);
}

int lpm_bond_w1_layout(const lpm_bond_group* g, const float* w1, float* w1f, void* stream) {
  if (int rc = check_group(g)) return rc;
  if (!w1 || !w1f) return fail(LPM_E_BADARG, "lpm_bond_w1_layout: null weights");
  const int nslot = g->nv + g->ncell, Q = g->dstride / 4, NT = col_tiles(g->hid);
  const int64_t n = (int64_t)(kpad_of(nslot) + 2) * NT * 64 * Q;
  w1_layout_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(nslot, g->D, Q, NT, g->hid, w1, w1f);
  LAUNCHCHK();
  return LPM_OK;
}

int lpm_bond_heads(const lpm_bond_group* g, int32_t B, int32_t nc, int32_t s, const float* fv, const float* fc,
                   const float* w1f, const float* b1, const float* w2, const float* b2, const float* w3,
                   const float* b3, float* out, const double* cp, double* kn, double* tv, void* stream) {
  HeadArgs a{};
  size_t lds = 0;
  if (int rc = head_args(g, B, nc, s, fv, fc, w1f, b1, w2, b2, w3, b3, out, cp, kn, tv, "lpm_bond_heads", a, lds))
    return rc;
  hipStream_t st = (hipStream_t)stream;
  return g->H == 16 ? launch_heads_nt<16>(a, lds, st) : launch_heads_nt<8>(a, lds, st);
}

size_t lpm_bond_w1_split_size(const lpm_bond_group* g) {
  if (!g || g->hid < 1) return 0;
  return (size_t)split_pairs_of(g->nv + g->ncell) * col_tiles(g->hid) * 64 * sizeof(bf16x8);
}

int lpm_bond_w1_split_layout(const lpm_bond_group* g, const float* w1, void* w1h, void* w1l, void* stream) {
  if (int rc = check_group(g)) return rc;
  if (g->D != 12 || g->dstride != 12)
    return fail(LPM_E_UNSUPPORTED, "split BF16 heads: %d channels (12 supported)", g->D);
  if (!w1 || !w1h || !w1l) return fail(LPM_E_BADARG, "lpm_bond_w1_split_layout: null weights");
  const int nslot = g->nv + g->ncell, NT = col_tiles(g->hid);
  const int64_t n = (int64_t)split_pairs_of(nslot) * NT * 64;
  w1_split_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(nslot, NT, g->hid, w1, (bf16x8*)w1h,
                                                                                (bf16x8*)w1l);
  LAUNCHCHK();
  return LPM_OK;
}

int lpm_bond_heads_split(const lpm_bond_group* g, int32_t B, int32_t nc, int32_t s, const float* fv, const float* fc,
                         const void* w1h, const void* w1l, const float* b1, const float* w2, const float* b2,
                         const float* w3, const float* b3, float* out, const double* cp, double* kn, double* tv,
                         void* stream) {
  if (g && (g->D != 12 || g->dstride != 12))
    return fail(LPM_E_UNSUPPORTED, "split BF16 heads: %d channels (12 supported)", g->D);
  HeadArgs a{};
  size_t lds = 0;
  if (int rc = head_args(g, B, nc, s, fv, fc, (const float*)w1h, b1, w2, b2, w3, b3, out, cp, kn, tv,
                         "lpm_bond_heads_split", a, lds))
    return rc;
  if (!w1l) return fail(LPM_E_BADARG, "lpm_bond_heads_split: null weights");
  dim3 grid((unsigned)((a.nb + HB - 1) / HB), (unsigned)a.B);
  hipStream_t st = (hipStream_t)stream;
  const int nt = col_tiles(g->hid);
  const bf16x8* wh = (const bf16x8*)w1h;
  const bf16x8* wl = (const bf16x8*)w1l;
  if (g->H == 16) {
    if (nt == 1) bond_heads_split_kernel<16, 1><<<grid, 256, lds, st>>>(a, wh, wl);
    else if (nt == 2) bond_heads_split_kernel<16, 2><<<grid, 256, lds, st>>>(a, wh, wl);
    else bond_heads_split_kernel<16, 4><<<grid, 256, lds, st>>>(a, wh, wl);
  } else {
    if (nt == 1) bond_heads_split_kernel<8, 1><<<grid, 256, lds, st>>>(a, wh, wl);
    else if (nt == 2) bond_heads_split_kernel<8, 2><<<grid, 256, lds, st>>>(a, wh, wl);
    else bond_heads_split_kernel<8, 4><<<grid, 256, lds, st>>>(a, wh, wl);
  }
  LAUNCHCHK();
  return LPM_OK;
}

}  // extern "C"
