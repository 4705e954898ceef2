// C ABI of liblpm.so (declarations and reference citations in include/lpm.h).
#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/lpm.h"
#include "cg.hpp"
#include "ku2d.hpp"
#include "ku3d.hpp"
#include "post.hpp"
#include "serve.hpp"
#include "capi_common.hpp"

using namespace lpm;
using namespace lpmc;

namespace {

thread_local std::string g_err;

}  // namespace

int lpmc::fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {

// Optional per-thread timing of the K u launches inside the (P)CG loop (HIP events on the solver's stream).
struct KuProfile {
  bool on = false;
  double ms = 0.0;
  long long count = 0;
  std::vector<hipEvent_t> ev;
  size_t used = 0;
};
thread_local KuProfile g_prof;

int prof_flush() {
  for (size_t i = 0; i + 1 < g_prof.used; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, g_prof.ev[i], g_prof.ev[i + 1]) == hipSuccess) {
      g_prof.ms += ms;
      ++g_prof.count;
    }
  }
  g_prof.used = 0;
  return 0;
}

inline Bcast bc_of(lpm_bcast b) { return Bcast{b.sb, b.sl}; }
inline int64_t nsites(const lpm_grid* g) { return (int64_t)g->nx * g->ny * (g->dim == 3 ? g->nz : 1); }
inline int64_t guard_of(const lpm_grid* g) { return packed_guard(g->dim, g->nx, g->ny); }
inline int64_t pstride_of(const lpm_grid* g) { return packed_stride(g->dim, g->nx, g->ny, nsites(g)); }
inline Box box_of(const lpm_grid* g) {
  Box b{g->nx, g->ny, g->dim == 3 ? g->nz : 1, pstride_of(g)};
  b.bonds = g->bonds;
  b.bb = Bcast{g->bb.sb, g->bb.sl};
  return b;
}
// kernels address packed planes from their first particle (past the leading guard)
inline const void* cbase(const lpm_grid* g, const void* p, int layout, int dtype) {
  if (layout != LPM_COEF_PACKED || !p) return p;
  return (const char*)p + guard_of(g) * (dtype == LPM_F32 ? 4 : 8);
}

int check_grid(const lpm_grid* g) {
  if (!g) return fail(LPM_E_BADARG, "grid is null");
  if (g->dim != 2 && g->dim != 3) return fail(LPM_E_BADARG, "dim must be 2 or 3 (got %d)", g->dim);
  if (g->nx < 1 || g->ny < 1 || (g->dim == 3 && g->nz < 1))
    return fail(LPM_E_BADARG, "bad lattice %d x %d x %d", g->nx, g->ny, g->nz);
  if (nsites(g) >= (int64_t)1 << 31) return fail(LPM_E_BADARG, "lattice too large");
  return LPM_OK;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// per-thread non-blocking stream (default priority) used only to capture graphs
hipStream_t capture_stream() {
  thread_local hipStream_t cs = nullptr;
  if (!cs && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) cs = nullptr;
  return cs;
}

// blocks per system of the grid-stride vector kernels: at most 1024, so the per-system partial sums stay short
inline int vec_blocks(int64_t N) { return (int)std::min<int64_t>(1024, std::max<int64_t>(1, (N + 255) / 256)); }

// tiling overrides (0 = automatic): x tile width of the 3D kernels and planes per block; g_lean = 1 sends FP32 packed
// K u (plain and masked-dot forms) to the lean kernel.  It is opt-in: at 256^3 it is faster alone (443 vs 477 us)
// but slower inside the PCG iteration (471 vs 444 us), where the vector kernels between launches evict the caches.
int g_tile_wide = 0, g_zc = 0, g_lean = 0;
// 3D box kernel (ku3b.hpp): -1 automatic (launches the marching kernel would run with fewer blocks than CUs), 0
// never, 1 always
int g_box = -1;
int g_serve_fused = 1;  // serving: fused band product of the inner iterations (0: two passes)

// compute units of the current device (cached per device; 256 when the runtime cannot tell)
int cu_count() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cache[dev] <= 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cache[dev] = n;
  }
  return cache[dev];
}

// tile shapes: 3D float 64x4 or 32x8, double 32x4 or 16x8; 2D float 32x8, double 32x4
template <int DIM, typename T>
KuGeom ku_geom(const lpm_grid* g, int nb, int nl, bool bonds, int lean_switch = -1, int box_switch = -2) {
  const int use_lean = lean_switch < 0 ? g_lean : lean_switch;
  const int use_box = box_switch < -1 ? g_box : box_switch;
  KuGeom k{};
  if constexpr (DIM == 3) {
    if (use_box == 1) {
      k.box = 1;
      k.tx = lpm::kBoxX;
      k.ty = lpm::kBoxY;
      k.zc = lpm::kBoxZ;
      k.ntx = (g->nx + lpm::kBoxX - 1) / lpm::kBoxX;
      k.nty = (g->ny + lpm::kBoxY - 1) / lpm::kBoxY;
      k.nzc = (g->nz + lpm::kBoxZ - 1) / lpm::kBoxZ;
      k.grid = dim3(k.ntx * k.nty * k.nzc, nl, nb);
      k.block = dim3(lpm::kBoxX * lpm::kBoxY * lpm::kBoxZ);
      return k;
    }
    const bool wide = g_tile_wide ? g_tile_wide > 0 : g->nx >= 48;
    if (sizeof(T) == 4) {
      k.tx = wide ? 64 : 32;
      // 64 x 8 exists without bond words only
      k.ty = (g_tile_wide == 2 && !bonds) ? 8 : (wide ? 4 : 8);
    } else {
      k.tx = wide ? 32 : 16;
      k.ty = wide ? 4 : 8;
    }
    k.ntx = (g->nx + k.tx - 1) / k.tx;
    k.nty = (g->ny + k.ty - 1) / k.ty;
    const int64_t xy = (int64_t)k.ntx * k.nty * nb * nl;
    // The general 3D tile kernels keep two blocks resident per CU, the lean FP32 kernel (64 x 4 tiles without bond
    // words) kLeanBlocksPerCU (`slots` blocks on the chip).  The z range of a column tile is split into chunks of at
    // least 8 planes so that the launch has about `slots` blocks.  Rounding the chunk count up can leave a short tail round
    // (slots < blocks < 1.5 slots) that costs nearly a whole round; then the count is rounded down to one round.
    // Longer overshoots (1.5 slots and more) are kept: the block scheduler refills freed slots, and they measured
    // faster than one partly filled round.
    const bool lean = sizeof(T) == 4 && use_lean && k.tx == 64 && k.ty == 4 && !bonds;
    if (g_zc > 0) {
      k.zc = g_zc;
    } else {
      const int64_t slots = (lean ? kLeanBlocksPerCU : 2) * (int64_t)cu_count();
      auto zc_for = [&](int64_t chunks) { return std::max<int>(8, (int)((g->nz + chunks - 1) / chunks)); };
      k.zc = zc_for(std::max<int64_t>(1, (slots + xy - 1) / xy));
      const int64_t blocks = xy * ((g->nz + k.zc - 1) / k.zc);
      if (blocks > slots && 2 * blocks < 3 * slots && xy <= slots) k.zc = zc_for(std::max<int64_t>(1, slots / xy));
    }
    k.nzc = (g->nz + k.zc - 1) / k.zc;
    k.grid = dim3(k.ntx * k.nty * k.nzc, nl, nb);
    // automatic: a marching launch with fewer blocks than CUs (small lattices, few systems) leaves the chip idle
    // while each block walks its planes; the box kernel spreads the same sites over about 16 times more blocks.
    // Measured kernel durations (FP32 packed): 24^3 x 16 systems 25.7 -> 14.6 us, 64^3 27.0 -> 20.8 us; with a
    // full marching launch the box kernel is not faster (48^3 x 9: 44.6 against 47.5 us; 128^3: 61 against 116).
    if (use_box < 0 && g_tile_wide == 0 && g_zc == 0 && (int64_t)k.grid.x * nl * nb < cu_count())
      return ku_geom<DIM, T>(g, nb, nl, bonds, lean_switch, 1);
  } else {
    k.tx = 32;
    k.ty = sizeof(T) == 4 ? 8 : 4;
    k.ntx = (g->nx + k.tx - 1) / k.tx;
    k.nty = (g->ny + k.ty - 1) / k.ty;
    k.nzc = 1;
    k.zc = 1;
    k.grid = dim3(k.ntx * k.nty, nl, nb);
  }
  k.block = dim3(k.tx * k.ty);
  return k;
}
template <int DIM, typename T>
KuGeom ku_geom(const lpm_grid* g, int nb, int nl) {
  return ku_geom<DIM, T>(g, nb, nl, g->bonds != nullptr);
}
// blocks per system of the K u launch with or without bond words, general or lean kernel, whichever is larger
// (workspace sizing, independent of the kernel switch at the time of the query)
template <int DIM, typename T>
int ku_blocks_max(const lpm_grid* g, int nsys) {
  unsigned m = 0;
  for (int box = 0; box < 2; ++box)
    for (int lean = 0; lean < 2; ++lean)
      for (int bonds = 0; bonds < 2; ++bonds)
        m = std::max(m, ku_geom<DIM, T>(g, 1, nsys, bonds != 0, lean, box).grid.x);
  return (int)m;
}

// one K u launch with epilogue EPI for either dimension / layout
struct KuCall {
  const void* u;
  Bcast bu;
  const void* kn;
  const void* tv;
  Bcast bc;
  const uint8_t* mask;
  Bcast bm;
  const void* rhs;
  Bcast brhs;
  void* out;
  Bcast bo;
  void* partial;
  const int* active;  // per-system flags with stride NINT (CG state), or null
};

template <int DIM, typename T, int EPI>
int ku_call(const lpm_grid* g, int nb, int nl, int layout, const KuGeom& kg, const KuCall& c, hipStream_t st) {
  // the kernels address a system with 32-bit byte offsets
  const int64_t N = nsites(g);
  if (N * 3 * (int64_t)sizeof(T) >= ((int64_t)1 << 31) ||
      (layout == LPM_COEF_DIRECTED && N * 2 * DIM * 3 * (int64_t)sizeof(T) >= ((int64_t)1 << 32)) ||
      (layout == LPM_COEF_PACKED && pstride_of(g) * (int64_t)sizeof(T) >= ((int64_t)1 << 31)))
    return fail(LPM_E_BADARG, "lattice too large for one system (use the packed layout or split the lattice)");
  if constexpr (DIM == 3) {
    Ku3Args<T, T> a{};
    a.u = (const T*)c.u;
    a.bu = c.bu;
    a.kn = (const T*)cbase(g, c.kn, layout, sizeof(T) == 4 ? LPM_F32 : LPM_F64);
    a.tv = (const T*)cbase(g, c.tv, layout, sizeof(T) == 4 ? LPM_F32 : LPM_F64);
    a.ps = pstride_of(g);
    a.bc = c.bc;
    a.mask = c.mask;
    a.bm = c.bm;
    a.rhs = (const T*)c.rhs;
    a.brhs = c.brhs;
    a.out = (T*)c.out;
    a.bo = c.bo;
    a.partial = (T*)c.partial;
    a.active = c.active;
    a.act_stride = NINT;
    a.bonds = g->bonds;
    a.bb = Bcast{g->bb.sb, g->bb.sl};
    if (kg.box) {
      if (g->bonds)
        return layout == LPM_COEF_PACKED ? launch_ku3b<T, COEF_PACKED, EPI, true>(g, nb, nl, kg, a, st)
                                         : launch_ku3b<T, COEF_DIRECTED, EPI, true>(g, nb, nl, kg, a, st);
      return layout == LPM_COEF_PACKED ? launch_ku3b<T, COEF_PACKED, EPI, false>(g, nb, nl, kg, a, st)
                                       : launch_ku3b<T, COEF_DIRECTED, EPI, false>(g, nb, nl, kg, a, st);
    }
    if (g->bonds)
      return layout == LPM_COEF_PACKED ? launch_ku3<T, COEF_PACKED, EPI, true>(g, nb, nl, kg, a, st)
                                       : launch_ku3<T, COEF_DIRECTED, EPI, true>(g, nb, nl, kg, a, st);
    if constexpr (sizeof(T) == 4 && (EPI == EPI_PLAIN || EPI == EPI_MASKDOT)) {
      // the PCG hot path: lean kernel when the whole packed set and the vectors have 32-bit byte offsets
      if (layout == LPM_COEF_PACKED && g_lean && kg.tx == 64 && kg.ty == 4 &&
          (uint64_t)pstride_of(g) * 9u * 4u < (1ull << 32) && (uint64_t)N * 12u < (1ull << 32))
        return launch_ku3l<EPI>(g, nb, nl, kg, a, st);
    }
    return layout == LPM_COEF_PACKED ? launch_ku3<T, COEF_PACKED, EPI, false>(g, nb, nl, kg, a, st)
                                     : launch_ku3<T, COEF_DIRECTED, EPI, false>(g, nb, nl, kg, a, st);
  } else {
    Ku2Args<T, T> a{};
    a.u = (const T*)c.u;
    a.bu = c.bu;
    a.kn = (const T*)cbase(g, c.kn, layout, sizeof(T) == 4 ? LPM_F32 : LPM_F64);
    a.tv = (const T*)cbase(g, c.tv, layout, sizeof(T) == 4 ? LPM_F32 : LPM_F64);
    a.ps = pstride_of(g);
    a.bc = c.bc;
    a.mask = c.mask;
    a.bm = c.bm;
    a.rhs = (const T*)c.rhs;
    a.brhs = c.brhs;
    a.out = (T*)c.out;
    a.bo = c.bo;
    a.partial = (T*)c.partial;
    a.active = c.active;
    a.act_stride = NINT;
    a.bonds = g->bonds;
    a.bb = Bcast{g->bb.sb, g->bb.sl};
    if (g->bonds)
      return layout == LPM_COEF_PACKED ? launch_ku2<T, COEF_PACKED, EPI, true>(g, nb, nl, kg, a, st)
                                       : launch_ku2<T, COEF_DIRECTED, EPI, true>(g, nb, nl, kg, a, st);
    return layout == LPM_COEF_PACKED ? launch_ku2<T, COEF_PACKED, EPI, false>(g, nb, nl, kg, a, st)
                                     : launch_ku2<T, COEF_DIRECTED, EPI, false>(g, nb, nl, kg, a, st);
  }
}

// ------------------------------------------------------------------------------------------------ PCG
template <typename T>
__global__ void cg_setup_kernel(T* scal, int* ist, int nsys, T tol, int maxit) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  scal[(int64_t)s * NSCAL + S_TOL] = tol;
  ist[(int64_t)s * NINT + I_MAXIT] = maxit;
  ist[(int64_t)s * NINT + I_ACTIVE] = 0;
  ist[(int64_t)s * NINT + I_IT] = 0;
  ist[(int64_t)s * NINT + I_XUPD] = 0;
  ist[(int64_t)s * NINT + I_NOTSPD] = 0;
  ist[(int64_t)s * NINT + I_CONV] = 0;
}
// per-system tolerance from a device array (refinement inner solves)
template <typename T>
__global__ void cg_settol_kernel(T* scal, const double* tol, int nsys) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys) scal[(int64_t)s * NSCAL + S_TOL] = (T)tol[s];
}

struct PcgLayout {
  size_t r, z, p, ap, partial, partial2, scal, ist, tols, snapit, total;
  int maxblk;
};
constexpr int kMaxSnapLevels = 64;  // tolerances of one snapshot solve

template <int DIM, typename T>
PcgLayout pcg_layout(const lpm_grid* g, int nsys) {
  constexpr int D = DIM;
  const int64_t N = nsites(g);
  const int vblk = vec_blocks(N);
  PcgLayout L{};
  L.maxblk = std::max<int>(ku_blocks_max<DIM, T>(g, nsys), vblk);
  size_t o = 0;
  const size_t vec = align256(sizeof(T) * (size_t)nsys * N * D);
  L.r = o;
  o += vec;
  L.z = o;
  o += vec;
  L.p = o;
  o += vec;
  L.ap = o;
  o += vec;
  L.partial = o;
  o += align256(sizeof(T) * (size_t)nsys * L.maxblk * 3);
  L.partial2 = o;  // the update's partials while the K p partials are still read (fused iteration)
  o += align256(sizeof(T) * (size_t)nsys * vblk * 3);
  L.scal = o;
  o += align256(sizeof(T) * (size_t)nsys * NSCAL);
  L.ist = o;
  o += align256(sizeof(int) * (size_t)nsys * NINT);
  L.tols = o;
  o += align256(sizeof(double) * kMaxSnapLevels);
  L.snapit = o;
  o += align256(sizeof(int) * (size_t)nsys * kMaxSnapLevels);
  L.total = o;
  return L;
}

// Core batched (P)CG on device buffers.  When tol_dev is non-null it gives a per-system tolerance.
template <int DIM, typename T>
int pcg_core(const lpm_grid* g, int nb, int nl, int layout, const void* kn, const void* tv, Bcast bc,
             const uint8_t* mask, Bcast bm, const T* minv, Bcast bminv, const T* b, T* x, int use_x0,
             const lpm_cg_opts* o, char* work, const double* tol_dev, int32_t* iters, double* relres, int32_t* conv,
             hipStream_t st, bool no_alloc = false, SnapDev<T> sd = SnapDev<T>{0, nullptr, nullptr, nullptr}) {
  constexpr int D = DIM;
  const int nsys = nb * nl;
  const int64_t N = nsites(g);
  const PcgLayout L = pcg_layout<DIM, T>(g, nsys);
  const KuGeom kg = ku_geom<DIM, T>(g, nb, nl);
  T* r = (T*)(work + L.r);
  T* z = o->precond ? (T*)(work + L.z) : nullptr;
  T* p = (T*)(work + L.p);
  T* ap = (T*)(work + L.ap);
  T* partial = (T*)(work + L.partial);
  T* partial2 = (T*)(work + L.partial2);
  T* scal = (T*)(work + L.scal);
  int* ist = (int*)(work + L.ist);
  const Bcast bvec{(int64_t)nl * N * D, N * D};
  const dim3 vgrid((unsigned)vec_blocks(N), nl, nb);
  const int vblk = (int)vgrid.x;

  CgVec<T> v{};
  v.N = N;
  v.nl = nl;
  v.x = x;
  v.r = r;
  v.z = z;
  v.p = p;
  v.ap = ap;
  v.minv = o->precond ? minv : nullptr;
  v.bminv = bminv;
  v.mask = mask;
  v.bm = bm;
  v.partial = partial;
  v.scal = scal;
  v.ist = ist;

  hipLaunchKernelGGL((cg_setup_kernel<T>), dim3((nsys + 255) / 256), dim3(256), 0, st, scal, ist, nsys, (T)o->tol,
                     o->maxit);
  if (tol_dev) hipLaunchKernelGGL((cg_settol_kernel<T>), dim3((nsys + 255) / 256), dim3(256), 0, st, scal, tol_dev, nsys);
  LAUNCHCHK();
  if (use_x0) {
    KuCall c{x, bvec, kn, tv, bc, mask, bm, b, bvec, r, bvec, partial, nullptr};
    int rc = ku_call<DIM, T, EPI_RESID>(g, nb, nl, layout, kg, c, st);
    if (rc) return rc;
  }
  hipLaunchKernelGGL((cg_init_kernel<D, T>), vgrid, dim3(256), 0, st, v, b, use_x0);
  hipLaunchKernelGGL((cg_fin_init_kernel<T>), dim3(nsys), dim3(256), 0, st, partial, vblk, scal, ist, sd);
  // levels met by the initial residual take the initial iterate
  if (sd.nt > 0) hipLaunchKernelGGL((cg_snap_copy_kernel<D, T>), vgrid, dim3(256), 0, st, v, sd);
  LAUNCHCHK();

  const int check = std::max(1, o->check);
  std::vector<int> host_ist((size_t)nsys * NINT);
  // the captured check block, destroyed on every way out of this function
  struct GraphKeep {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    ~GraphKeep() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
    }
  } gk;
  hipGraphExec_t& gexec = gk.exec;
  hipGraph_t& graph = gk.graph;
  bool capturing = false;
  auto body = [&](hipStream_t sq) -> int {
    KuCall c{p, bvec, kn, tv, bc, mask, bm, nullptr, bvec, ap, bvec, partial, ist + I_ACTIVE};
    // the allocation-free entry point creates no events (profiling) and no graph
    const bool prof = g_prof.on && !capturing && !no_alloc;
    if (prof) {
      while (g_prof.ev.size() < g_prof.used + 2) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        g_prof.ev.push_back(e);
      }
      HIPCHK(hipEventRecord(g_prof.ev[g_prof.used], sq));
    }
    int rc = ku_call<DIM, T, EPI_MASKDOT>(g, nb, nl, layout, kg, c, sq);
    if (rc) return rc;
    if (prof) {
      HIPCHK(hipEventRecord(g_prof.ev[g_prof.used + 1], sq));
      g_prof.used += 2;
    }
    if (sd.nt > 0) {
      hipLaunchKernelGGL((cg_fin_alpha_kernel<T>), dim3(nsys), dim3(256), 0, sq, partial, (int)kg.grid.x, scal,
                         ist);
      hipLaunchKernelGGL((cg_update_kernel<D, T>), vgrid, dim3(256), 0, sq, v);
      hipLaunchKernelGGL((cg_fin_beta_kernel<T>), dim3(nsys), dim3(256), 0, sq, partial, vblk, scal, ist, sd);
      hipLaunchKernelGGL((cg_pupdate_kernel<D, T>), vgrid, dim3(256), 0, sq, v, sd);
    } else {
      // the scalar steps folded into the vector kernels (three launches per iteration)
      CgVec<T> vf = v;
      vf.partial = partial2;
      hipLaunchKernelGGL((cg_update_fused_kernel<D, T>), vgrid, dim3(256), 0, sq, vf, (const T*)partial,
                         (int)kg.grid.x);
      hipLaunchKernelGGL((cg_pupdate_fused_kernel<D, T>), vgrid, dim3(256), 0, sq, vf, (const T*)partial2, vblk);
    }
    LAUNCHCHK();
    return LPM_OK;
  };
  int done_it = 0;
  int rc = LPM_OK;
  while (done_it < o->maxit) {
    const int nrep = std::min(check, o->maxit - done_it);
    if (o->graph && !no_alloc && nrep == check) {
      if (!gexec) {
        // capture on a private stream (the caller's may be the null stream, which cannot capture); the graph is
        // then launched on the caller's stream
        hipStream_t cs = capture_stream();
        if (!cs) return fail(LPM_E_HIP, "graph capture: no capture stream");
        HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        capturing = true;
        for (int i = 0; i < check && rc == LPM_OK; ++i) rc = body(cs);
        capturing = false;
        hipError_t ec = hipStreamEndCapture(cs, &graph);
        if (rc) return rc;
        if (ec != hipSuccess) return fail(LPM_E_HIP, "graph capture: %s", hipGetErrorString(ec));
        HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
      }
      HIPCHK(hipGraphLaunch(gexec, st));
    } else {
      for (int i = 0; i < nrep && rc == LPM_OK; ++i) rc = body(st);
      if (rc) return rc;
    }
    done_it += nrep;
    HIPCHK(hipMemcpyAsync(host_ist.data(), ist, sizeof(int) * nsys * NINT, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    prof_flush();
    bool any = false;
    int notspd = -1;
    for (int s = 0; s < nsys; ++s) {
      any |= host_ist[(size_t)s * NINT + I_ACTIVE] != 0;
      if (notspd < 0 && host_ist[(size_t)s * NINT + I_NOTSPD]) notspd = s;
    }
    if (notspd >= 0) {
      return fail(LPM_E_NOT_SPD, "(P)CG: system %d is not positive definite (p.Ap <= 0 or r.z < 0 at iteration %d)",
                  notspd, host_ist[(size_t)notspd * NINT + I_IT]);
    }
    if (!any) break;
  }
  if (sd.nt > 0) {
    // levels never met: the final iterate, iteration -1
    hipLaunchKernelGGL((cg_snap_rest_kernel<T>), dim3((nsys + 255) / 256), dim3(256), 0, st, ist, nsys, sd);
    hipLaunchKernelGGL((cg_snap_copy_kernel<D, T>), vgrid, dim3(256), 0, st, v, sd);
    LAUNCHCHK();
  }
  std::vector<T> host_scal((size_t)nsys * NSCAL);
  HIPCHK(hipMemcpyAsync(host_scal.data(), scal, sizeof(T) * nsys * NSCAL, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(host_ist.data(), ist, sizeof(int) * nsys * NINT, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  bool bad = false;
  for (int s = 0; s < nsys; ++s) {
    if (host_ist[(size_t)s * NINT + I_NOTSPD])
      return fail(LPM_E_NOT_SPD, "(P)CG: system %d is not positive definite", s);
    if (iters) iters[s] = host_ist[(size_t)s * NINT + I_IT];
    if (relres) relres[s] = (double)host_scal[(size_t)s * NSCAL + S_REL];
    if (conv) conv[s] = host_ist[(size_t)s * NINT + I_CONV];
    if (!std::isfinite((double)host_scal[(size_t)s * NSCAL + S_REL])) bad = true;
  }
  if (bad) return fail(LPM_E_NONFINITE, "non-finite residual in the (P)CG iteration");
  return LPM_OK;
}

}  // namespace

// ================================================================================================ C ABI
extern "C" {

const char* lpm_last_error(void) { return g_err.c_str(); }
int lpm_version(void) { return 1; }

int64_t lpm_packed_len(const lpm_grid* g) { return check_grid(g) ? 0 : pstride_of(g); }

void lpm_set_ku_tiling(int32_t wide, int32_t zc) {
  g_tile_wide = wide;
  g_zc = zc;
}

void lpm_set_ku_kernel(int32_t lean) { g_lean = lean != 0; }

void lpm_set_ku_box(int32_t box) { g_box = box < 0 ? -1 : (box > 0 ? 1 : 0); }

void lpm_profile_ku(int32_t enable) {
  g_prof.on = enable != 0;
  g_prof.ms = 0.0;
  g_prof.count = 0;
  g_prof.used = 0;
}

void lpm_profile_ku_read(double* total_ms, int64_t* launches) {
  if (total_ms) *total_ms = g_prof.ms;
  if (launches) *launches = g_prof.count;
}

int lpm_apply(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* u, lpm_bcast bu,
              const void* kn, const void* tv, lpm_bcast bc, void* out, void* stream) {
  if (int rc = check_grid(g)) return rc;
  if (!u || !kn || !tv || !out || nb < 1 || nl < 1) return fail(LPM_E_BADARG, "lpm_apply: null pointer or empty batch");
  if (nb > 65535 || nl > 65535) return fail(LPM_E_BADARG, "lpm_apply: batch too large");
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = nsites(g);
  const int D = g->dim;
  KuCall c{u, bc_of(bu), kn, tv, bc_of(bc), nullptr, {0, 0}, nullptr, {0, 0}, out,
           {(int64_t)nl * N * D, N * D}, nullptr, nullptr};
  if (g->dim == 3) {
    if (dtype == LPM_F32) return ku_call<3, float, EPI_PLAIN>(g, nb, nl, layout, ku_geom<3, float>(g, nb, nl), c, st);
    if (dtype == LPM_F64) return ku_call<3, double, EPI_PLAIN>(g, nb, nl, layout, ku_geom<3, double>(g, nb, nl), c, st);
  } else {
    if (dtype == LPM_F32) return ku_call<2, float, EPI_PLAIN>(g, nb, nl, layout, ku_geom<2, float>(g, nb, nl), c, st);
    if (dtype == LPM_F64) return ku_call<2, double, EPI_PLAIN>(g, nb, nl, layout, ku_geom<2, double>(g, nb, nl), c, st);
  }
  return fail(LPM_E_BADARG, "lpm_apply: bad dtype %d", dtype);
}

size_t lpm_residual_workspace(const lpm_grid* g, int32_t nsys, int32_t dtype) {
  if (check_grid(g) || nsys < 1) return 0;
  const size_t w = dtype == LPM_F32 ? 4 : 8;
  const int blk = g->dim == 3 ? (dtype == LPM_F32 ? ku_blocks_max<3, float>(g, nsys) : ku_blocks_max<3, double>(g, nsys))
                              : (dtype == LPM_F32 ? ku_blocks_max<2, float>(g, nsys) : ku_blocks_max<2, double>(g, nsys));
  return align256(w * (size_t)nsys * (size_t)blk);
}

int lpm_residual(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* u,
                 lpm_bcast bu, const void* kn, const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm,
                 const void* rhs, lpm_bcast brhs, void* out, void* work, void* stream) {
  if (int rc = check_grid(g)) return rc;
  if (!u || !kn || !tv || !rhs || !out || !work || nb < 1 || nl < 1)
    return fail(LPM_E_BADARG, "lpm_residual: null pointer or empty batch");
  if (nb > 65535 || nl > 65535) return fail(LPM_E_BADARG, "lpm_residual: batch too large");
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = nsites(g);
  const int D = g->dim;
  KuCall c{u, bc_of(bu), kn, tv, bc_of(bc), mask, bc_of(bm), rhs, bc_of(brhs), out,
           {(int64_t)nl * N * D, N * D}, work, nullptr};
  if (g->dim == 3) {
    if (dtype == LPM_F32) return ku_call<3, float, EPI_RESID>(g, nb, nl, layout, ku_geom<3, float>(g, nb, nl), c, st);
    if (dtype == LPM_F64) return ku_call<3, double, EPI_RESID>(g, nb, nl, layout, ku_geom<3, double>(g, nb, nl), c, st);
  } else {
    if (dtype == LPM_F32) return ku_call<2, float, EPI_RESID>(g, nb, nl, layout, ku_geom<2, float>(g, nb, nl), c, st);
    if (dtype == LPM_F64) return ku_call<2, double, EPI_RESID>(g, nb, nl, layout, ku_geom<2, double>(g, nb, nl), c, st);
  }
  return fail(LPM_E_BADARG, "lpm_residual: bad dtype %d", dtype);
}

int lpm_bond_words(const lpm_grid* g, int32_t nbatch, const uint8_t* solid, uint32_t* words, void* stream) {
  if (int rc = check_grid(g)) return rc;
  if (!solid || !words || nbatch < 1 || nbatch > 65535)
    return fail(LPM_E_BADARG, "lpm_bond_words: null pointer or bad batch");
  hipStream_t st = (hipStream_t)stream;
  Box box = box_of(g);
  box.bonds = nullptr;
  const dim3 grid((unsigned)((nsites(g) + 255) / 256), nbatch);
  if (g->dim == 3)
    hipLaunchKernelGGL((bond_words_kernel<3>), grid, dim3(256), 0, st, box, solid, words);
  else
    hipLaunchKernelGGL((bond_words_kernel<2>), grid, dim3(256), 0, st, box, solid, words);
  LAUNCHCHK();
  return LPM_OK;
}

int lpm_pack_coef(const lpm_grid* g, int32_t nbatch, int32_t dtype_in, int32_t dtype_out, const void* kn,
                  const void* tv, void* kn_fwd, void* tv_fwd, double rtol, uint32_t* asym, void* stream) {
  if (int rc = check_grid(g)) return rc;
  if (!kn || !tv || !kn_fwd || !tv_fwd || !asym || nbatch < 1 || nbatch > 65535)
    return fail(LPM_E_BADARG, "lpm_pack_coef: null pointer or bad batch");
  hipStream_t st = (hipStream_t)stream;
  const Box box = box_of(g);
  const dim3 grid((unsigned)((box.ps + 255) / 256), nbatch);
#define PACK(DIM, CI, CO)                                                                                     \
  hipLaunchKernelGGL((pack_coef_kernel<DIM, CI, CO>), grid, dim3(256), 0, st, box, (const CI*)kn, (const CI*)tv, \
                     (CO*)kn_fwd, (CO*)tv_fwd, rtol, asym)
  const int key = (g->dim == 3 ? 100 : 0) + dtype_in * 10 + dtype_out;
  switch (key) {
    case 100: PACK(3, float, float); break;
    case 101: PACK(3, float, double); break;
    case 110: PACK(3, double, float); break;
    case 111: PACK(3, double, double); break;
    case 0: PACK(2, float, float); break;
    case 1: PACK(2, float, double); break;
    case 10: PACK(2, double, float); break;
    case 11: PACK(2, double, double); break;
    default: return fail(LPM_E_BADARG, "lpm_pack_coef: bad dtype");
  }
#undef PACK
  LAUNCHCHK();
  return LPM_OK;
}

int lpm_block_jacobi(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype_c, int32_t dtype_out, int32_t layout,
                     const void* kn, const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm, void* out,
                     int32_t raw, void* stream) {
  if (int rc = check_grid(g)) return rc;
  if (!kn || !tv || !out || nb < 1 || nl < 1 || nb > 65535 || nl > 65535)
    return fail(LPM_E_BADARG, "lpm_block_jacobi: null pointer or bad batch");
  hipStream_t st = (hipStream_t)stream;
  const Box box = box_of(g);
  kn = cbase(g, kn, layout, dtype_c);
  tv = cbase(g, tv, layout, dtype_c);
  const int64_t N = nsites(g);
  const int D = g->dim;
  const int64_t per = raw ? (int64_t)D * D : (D == 3 ? 6 : 3);
  const Bcast bo{(int64_t)nl * N * per, N * per};
  const dim3 grid((unsigned)((N + 255) / 256), nl, nb);
#define BJ(DIM, C, T, MODE)                                                                                        \
  hipLaunchKernelGGL((block_jacobi_kernel<DIM, T, C, MODE>), grid, dim3(256), 0, st, box, nl, (const C*)kn,       \
                     (const C*)tv, bc_of(bc), mask, bc_of(bm), (T*)out, bo, raw)
  const int key = (g->dim == 3 ? 1000 : 0) + dtype_c * 100 + dtype_out * 10 + (layout == LPM_COEF_PACKED);
  switch (key) {
    case 1000: BJ(3, float, float, COEF_DIRECTED); break;
    case 1001: BJ(3, float, float, COEF_PACKED); break;
    case 1010: BJ(3, float, double, COEF_DIRECTED); break;
    case 1011: BJ(3, float, double, COEF_PACKED); break;
    case 1100: BJ(3, double, float, COEF_DIRECTED); break;
    case 1101: BJ(3, double, float, COEF_PACKED); break;
    case 1110: BJ(3, double, double, COEF_DIRECTED); break;
    case 1111: BJ(3, double, double, COEF_PACKED); break;
    case 0: BJ(2, float, float, COEF_DIRECTED); break;
    case 1: BJ(2, float, float, COEF_PACKED); break;
    case 10: BJ(2, float, double, COEF_DIRECTED); break;
    case 11: BJ(2, float, double, COEF_PACKED); break;
    case 100: BJ(2, double, float, COEF_DIRECTED); break;
    case 101: BJ(2, double, float, COEF_PACKED); break;
    case 110: BJ(2, double, double, COEF_DIRECTED); break;
    case 111: BJ(2, double, double, COEF_PACKED); break;
    default: return fail(LPM_E_BADARG, "lpm_block_jacobi: bad dtype");
  }
#undef BJ
  LAUNCHCHK();
  return LPM_OK;
}

int lpm_post(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* u, lpm_bcast bu,
             const void* kn, const void* tv, lpm_bcast bc, void* work, void* ku, void* stress, void* energy, void* dl,
             void* force, void* stream) {
  if (int rc = check_grid(g)) return rc;
  if (!u || !kn || !tv || !work || nb < 1 || nl < 1 || nb > 65535 || nl > 65535)
    return fail(LPM_E_BADARG, "lpm_post: null pointer or bad batch");
  hipStream_t st = (hipStream_t)stream;
  const Box box = box_of(g);
  kn = cbase(g, kn, layout, dtype);
  tv = cbase(g, tv, layout, dtype);
  const int64_t N = nsites(g);
  const dim3 grid((unsigned)((N + 255) / 256), nl, nb);
  PostOut o{ku, stress, energy, dl, force};
#define POST(DIM, T, MODE)                                                                                        \
  do {                                                                                                            \
    hipLaunchKernelGGL((shell_sums_kernel<DIM, T, T, MODE>), grid, dim3(256), 0, st, box, nl, (const T*)u,       \
                       bc_of(bu), (const T*)tv, bc_of(bc), (T*)work);                                            \
    hipLaunchKernelGGL((site_post_kernel<DIM, T, T, MODE>), grid, dim3(256), 0, st, box, nl, g->h, (const T*)u,  \
                       bc_of(bu), (const T*)kn, (const T*)tv, bc_of(bc), (const T*)work, o);                     \
  } while (0)
  const int key = (g->dim == 3 ? 100 : 0) + dtype * 10 + (layout == LPM_COEF_PACKED);
  switch (key) {
    case 100: POST(3, float, COEF_DIRECTED); break;
    case 101: POST(3, float, COEF_PACKED); break;
    case 110: POST(3, double, COEF_DIRECTED); break;
    case 111: POST(3, double, COEF_PACKED); break;
    case 0: POST(2, float, COEF_DIRECTED); break;
    case 1: POST(2, float, COEF_PACKED); break;
    case 10: POST(2, double, COEF_DIRECTED); break;
    case 11: POST(2, double, COEF_PACKED); break;
    default: return fail(LPM_E_BADARG, "lpm_post: bad dtype");
  }
#undef POST
  LAUNCHCHK();
  return LPM_OK;
}

int lpm_bond_grad(const lpm_grid* g, int32_t ncb, int32_t spc, int32_t dtype, int32_t layout, const void* kn,
                  const void* tv, const void* u, const void* lam, double scale, void* work, void* gkn, void* gtv,
                  void* stream) {
  if (int rc = check_grid(g)) return rc;
  if (!kn || !tv || !u || !lam || !work || !gkn || !gtv || ncb < 1 || spc < 1 || ncb > 65535 || spc > 65535)
    return fail(LPM_E_BADARG, "lpm_bond_grad: null pointer or bad batch");
  hipStream_t st = (hipStream_t)stream;
  const Box box = box_of(g);
  const int64_t N = nsites(g);
  const int D = g->dim, NN = D == 3 ? 18 : 8;
  const int64_t cstride = layout == LPM_COEF_PACKED ? box.ps * NN / 2 : N * NN;
  kn = cbase(g, kn, layout, dtype);
  tv = cbase(g, tv, layout, dtype);
  const Bcast bu{(int64_t)spc * N * D, N * D};
  const Bcast bc{cstride, 0};
  const dim3 grid((unsigned)((N + 255) / 256), spc, ncb);
  const dim3 ggrid((unsigned)((N + 255) / 256), ncb);
  const size_t half = (size_t)ncb * spc * N * 4;
#define GRAD(DIM, T, MODE)                                                                                          \
  do {                                                                                                              \
    T* wu = (T*)work;                                                                                               \
    T* wl = wu + half;                                                                                              \
    hipLaunchKernelGGL((shell_sums_kernel<DIM, T, T, MODE>), grid, dim3(256), 0, st, box, spc, (const T*)u, bu,    \
                       (const T*)tv, bc, wu);                                                                       \
    hipLaunchKernelGGL((shell_sums_kernel<DIM, T, T, MODE>), grid, dim3(256), 0, st, box, spc, (const T*)lam, bu,  \
                       (const T*)tv, bc, wl);                                                                       \
    hipLaunchKernelGGL((bond_grad_kernel<DIM, T>), ggrid, dim3(256), 0, st, box, spc, (const T*)u, (const T*)lam,  \
                       (const T*)wu, (const T*)wl, (T)scale, (T*)gkn, (T*)gtv);                                    \
  } while (0)
  const int key = (g->dim == 3 ? 100 : 0) + dtype * 10 + (layout == LPM_COEF_PACKED);
  switch (key) {
    case 100: GRAD(3, float, COEF_DIRECTED); break;
    case 101: GRAD(3, float, COEF_PACKED); break;
    case 110: GRAD(3, double, COEF_DIRECTED); break;
    case 111: GRAD(3, double, COEF_PACKED); break;
    case 0: GRAD(2, float, COEF_DIRECTED); break;
    case 1: GRAD(2, float, COEF_PACKED); break;
    case 10: GRAD(2, double, COEF_DIRECTED); break;
    case 11: GRAD(2, double, COEF_PACKED); break;
    default: return fail(LPM_E_BADARG, "lpm_bond_grad: bad dtype");
  }
#undef GRAD
  LAUNCHCHK();
  (void)NN;
  return LPM_OK;
}

int lpm_post_vjp(const lpm_grid* g, int32_t ncb, int32_t spc, int32_t dtype, int32_t layout, const void* kn,
                 const void* tv, const void* u, const void* gku, const void* gstress, const void* genergy, void* work,
                 void* gu, void* gkn, void* gtv, void* stream) {
  if (int rc = check_grid(g)) return rc;
  if (!kn || !tv || !u || !work || !gu || !gkn || !gtv || ncb < 1 || spc < 1 || ncb > 65535 || spc > 65535 ||
      (int64_t)ncb * spc > 65535)
    return fail(LPM_E_BADARG, "lpm_post_vjp: null pointer or bad batch");
  hipStream_t st = (hipStream_t)stream;
  const Box box = box_of(g);
  const int64_t N = nsites(g);
  const int D = g->dim, NN = D == 3 ? 18 : 8;
  const int64_t cstride = layout == LPM_COEF_PACKED ? box.ps * NN / 2 : N * NN;
  kn = cbase(g, kn, layout, dtype);
  tv = cbase(g, tv, layout, dtype);
  const int nsys = ncb * spc;
  const Bcast bu{(int64_t)spc * N * D, N * D};
  const Bcast bc{cstride, 0};
  const dim3 grid((unsigned)((N + 255) / 256), spc, ncb);
  const dim3 sgrid((unsigned)((N + 255) / 256), nsys);
  const dim3 ggrid((unsigned)((N + 255) / 256), ncb);
#define VJP(DIM, T, MODE)                                                                                           \
  do {                                                                                                              \
    T* ws = (T*)work;                                                                                               \
    T* cw = ws + (size_t)nsys * N * 4;                                                                              \
    T* Ww = cw + (size_t)nsys * N * NN;                                                                             \
    hipLaunchKernelGGL((shell_sums_kernel<DIM, T, T, MODE>), grid, dim3(256), 0, st, box, spc, (const T*)u, bu,    \
                       (const T*)tv, bc, ws);                                                                       \
    hipLaunchKernelGGL((vjp_weights_kernel<DIM, T>), sgrid, dim3(256), 0, st, box, spc, g->h, (const T*)gku,       \
                       (const T*)gstress, cw);                                                                      \
    hipLaunchKernelGGL((vjp_params_kernel<DIM, T, MODE>), ggrid, dim3(256), 0, st, box, spc, (const T*)u,          \
                       (const T*)kn, (const T*)tv, cstride, (const T*)ws, (const T*)cw, (const T*)genergy, Ww,     \
                       (T*)gkn, (T*)gtv);                                                                           \
    hipLaunchKernelGGL((vjp_u_kernel<DIM, T>), sgrid, dim3(256), 0, st, box, spc, (const T*)Ww, (T*)gu);           \
  } while (0)
  const int key = (g->dim == 3 ? 100 : 0) + dtype * 10 + (layout == LPM_COEF_PACKED);
  switch (key) {
    case 100: VJP(3, float, COEF_DIRECTED); break;
    case 101: VJP(3, float, COEF_PACKED); break;
    case 110: VJP(3, double, COEF_DIRECTED); break;
    case 111: VJP(3, double, COEF_PACKED); break;
    case 0: VJP(2, float, COEF_DIRECTED); break;
    case 1: VJP(2, float, COEF_PACKED); break;
    case 10: VJP(2, double, COEF_DIRECTED); break;
    case 11: VJP(2, double, COEF_PACKED); break;
    default: return fail(LPM_E_BADARG, "lpm_post_vjp: bad dtype");
  }
#undef VJP
  LAUNCHCHK();
  return LPM_OK;
}

size_t lpm_pcg_workspace(const lpm_grid* g, int32_t nsys, int32_t dtype) {
  if (check_grid(g) || nsys < 1) return 0;
  if (g->dim == 3) return dtype == LPM_F32 ? pcg_layout<3, float>(g, nsys).total : pcg_layout<3, double>(g, nsys).total;
  return dtype == LPM_F32 ? pcg_layout<2, float>(g, nsys).total : pcg_layout<2, double>(g, nsys).total;
}

static int check_solve_args(const lpm_grid* g, int nb, int nl, const lpm_cg_opts* o, const void* minv) {
  if (int rc = check_grid(g)) return rc;
  if (!o) return fail(LPM_E_BADARG, "options are null");
  if (nb < 1 || nl < 1 || nb > 65535 || nl > 65535) return fail(LPM_E_BADARG, "bad batch %d x %d", nb, nl);
  if (o->precond && !minv) return fail(LPM_E_BADARG, "block-Jacobi PCG needs the preconditioner blocks");
  if (o->maxit < 0) return fail(LPM_E_BADARG, "maxit < 0");
  if (!(o->tol >= 0.0)) return fail(LPM_E_BADARG, "tol must be >= 0");
  return LPM_OK;
}

int lpm_pcg_solve(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* kn,
                  const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm, const void* minv, lpm_bcast bminv,
                  const void* b, void* x, int32_t use_x0, const lpm_cg_opts* opts, void* work, int32_t* iters,
                  double* relres, int32_t* conv, void* stream) {
  if (int rc = check_solve_args(g, nb, nl, opts, minv)) return rc;
  if (!kn || !tv || !b || !x || !work) return fail(LPM_E_BADARG, "lpm_pcg_solve: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const int key = (g->dim == 3 ? 10 : 0) + dtype;
  switch (key) {
    case 10:
      return pcg_core<3, float>(g, nb, nl, layout, kn, tv, bc_of(bc), mask, bc_of(bm), (const float*)minv,
                                bc_of(bminv), (const float*)b, (float*)x, use_x0, opts, (char*)work, nullptr, iters,
                                relres, conv, st);
    case 11:
      return pcg_core<3, double>(g, nb, nl, layout, kn, tv, bc_of(bc), mask, bc_of(bm), (const double*)minv,
                                 bc_of(bminv), (const double*)b, (double*)x, use_x0, opts, (char*)work, nullptr, iters,
                                 relres, conv, st);
    case 0:
      return pcg_core<2, float>(g, nb, nl, layout, kn, tv, bc_of(bc), mask, bc_of(bm), (const float*)minv,
                                bc_of(bminv), (const float*)b, (float*)x, use_x0, opts, (char*)work, nullptr, iters,
                                relres, conv, st);
    case 1:
      return pcg_core<2, double>(g, nb, nl, layout, kn, tv, bc_of(bc), mask, bc_of(bm), (const double*)minv,
                                 bc_of(bminv), (const double*)b, (double*)x, use_x0, opts, (char*)work, nullptr, iters,
                                 relres, conv, st);
    default:
      return fail(LPM_E_BADARG, "lpm_pcg_solve: bad dtype %d", dtype);
  }
}

int lpm_pcg_snapshots(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* kn,
                      const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm, const void* minv,
                      lpm_bcast bminv, const void* b, void* x, int32_t use_x0, const lpm_cg_opts* opts, int32_t ntol,
                      const double* tols, void* snaps, int32_t* snap_iters, void* work, int32_t* iters,
                      double* relres, int32_t* conv, void* stream) {
  if (int rc = check_solve_args(g, nb, nl, opts, minv)) return rc;
  if (!kn || !tv || !b || !x || !work || !tols || !snaps || !snap_iters)
    return fail(LPM_E_BADARG, "lpm_pcg_snapshots: null pointer");
  if (ntol < 1 || ntol > kMaxSnapLevels) return fail(LPM_E_BADARG, "lpm_pcg_snapshots: 1 .. %d tolerances", kMaxSnapLevels);
  for (int k = 0; k < ntol; ++k)
    if (!(tols[k] >= 0.0) || (k > 0 && !(tols[k] <= tols[k - 1])))
      return fail(LPM_E_BADARG, "lpm_pcg_snapshots: tolerances must be non-negative and descending");
  hipStream_t st = (hipStream_t)stream;
  const int nsys = nb * nl;
  int rc = LPM_OK;
  auto run = [&](auto dim_tag, auto t_tag) -> int {
    constexpr int DIM = decltype(dim_tag)::value;
    using T = decltype(t_tag);
    const PcgLayout Lw = pcg_layout<DIM, T>(g, nsys);
    char* w = (char*)work;
    HIPCHK(hipMemcpyAsync(w + Lw.tols, tols, sizeof(double) * ntol, hipMemcpyHostToDevice, st));
    SnapDev<T> sd{ntol, (const double*)(w + Lw.tols), (int*)(w + Lw.snapit), (T*)snaps};
    if (int r = pcg_core<DIM, T>(g, nb, nl, layout, kn, tv, bc_of(bc), mask, bc_of(bm), (const T*)minv, bc_of(bminv),
                                 (const T*)b, (T*)x, use_x0, opts, w, nullptr, iters, relres, conv, st, false, sd))
      return r;
    HIPCHK(hipMemcpyAsync(snap_iters, w + Lw.snapit, sizeof(int32_t) * nsys * ntol, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return LPM_OK;
  };
  if (dtype != LPM_F32 && dtype != LPM_F64) return fail(LPM_E_BADARG, "lpm_pcg_snapshots: bad dtype %d", dtype);
  if (g->dim == 3)
    rc = dtype == LPM_F32 ? run(std::integral_constant<int, 3>{}, float{}) : run(std::integral_constant<int, 3>{}, double{});
  else
    rc = dtype == LPM_F32 ? run(std::integral_constant<int, 2>{}, float{}) : run(std::integral_constant<int, 2>{}, double{});
  return rc;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------ no allocation
// Workspace of lpm_pcg_solve_noalloc: a file-scope device array, one copy per GPU for the life of the process.  In
// host code its name is only a stand-in; the per-device address comes from hipGetSymbolAddress and is the only
// thing host code passes on.
namespace {
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch_ptr[kMaxDevices];

// address of the current device's scratch array (looked up once per device, with that device current)
int scratch_of_current_device(void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(LPM_E_UNSUPPORTED, "device %d: no scratch slot", dev);
  void* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
  if (!p) {
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(LPM_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev].store(p, std::memory_order_release);
  }
  *out = p;
  return LPM_OK;
}
}  // namespace

extern "C" {

size_t lpm_noalloc_scratch_bytes(void) { return kScratchBytes; }

int lpm_pcg_solve_noalloc(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* kn,
                          const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm, const void* minv,
                          lpm_bcast bminv, const void* b, void* x, int32_t use_x0, const lpm_cg_opts* opts,
                          int32_t* iters, double* relres, int32_t* conv, void* stream) {
  if (int rc = check_solve_args(g, nb, nl, opts, minv)) return rc;
  if (!kn || !tv || !b || !x) return fail(LPM_E_BADARG, "lpm_pcg_solve_noalloc: null pointer");
  if (dtype != LPM_F32 && dtype != LPM_F64) return fail(LPM_E_BADARG, "lpm_pcg_solve_noalloc: bad dtype %d", dtype);
  const size_t need = lpm_pcg_workspace(g, nb * nl, dtype);
  if (need == 0 || need > kScratchBytes)
    return fail(LPM_E_BADARG, "lpm_pcg_solve_noalloc: the solve needs %zu bytes of workspace, more than the %zu-byte "
                "static scratch (use lpm_pcg_solve with a caller workspace)", need, kScratchBytes);
  void* work = nullptr;
  if (int rc = scratch_of_current_device(&work)) return rc;
  // every workspace value the solver reads is written by it first (scalars and flags by the setup kernel,
  // vectors by the init kernel, partial sums by the kernels that produce them): nothing to clear
  lpm_cg_opts o = *opts;
  o.graph = 0;
  hipStream_t st = (hipStream_t)stream;
  const int key = (g->dim == 3 ? 10 : 0) + dtype;
  const Bcast bcc = bc_of(bc), bmm = bc_of(bm), bmi = bc_of(bminv);
  switch (key) {
    case 10:
      return pcg_core<3, float>(g, nb, nl, layout, kn, tv, bcc, mask, bmm, (const float*)minv, bmi, (const float*)b,
                                (float*)x, use_x0, &o, (char*)work, nullptr, iters, relres, conv, st, true);
    case 11:
      return pcg_core<3, double>(g, nb, nl, layout, kn, tv, bcc, mask, bmm, (const double*)minv, bmi,
                                 (const double*)b, (double*)x, use_x0, &o, (char*)work, nullptr, iters, relres, conv,
                                 st, true);
    case 0:
      return pcg_core<2, float>(g, nb, nl, layout, kn, tv, bcc, mask, bmm, (const float*)minv, bmi, (const float*)b,
                                (float*)x, use_x0, &o, (char*)work, nullptr, iters, relres, conv, st, true);
    default:
      return pcg_core<2, double>(g, nb, nl, layout, kn, tv, bcc, mask, bmm, (const double*)minv, bmi,
                                 (const double*)b, (double*)x, use_x0, &o, (char*)work, nullptr, iters, relres, conv,
                                 st, true);
  }
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------ refinement
namespace {

__global__ void zero_kernel(double* x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

struct RefLayout {
  size_t inner, r64, r32, d32, part, bn2, rr2, tin, total;
};

template <int DIM>
RefLayout ref_layout(const lpm_grid* g, int nsys) {
  const int64_t N = nsites(g);
  RefLayout L{};
  size_t o = 0;
  L.inner = o;
  o += align256(pcg_layout<DIM, float>(g, nsys).total);
  L.r64 = o;
  o += align256(sizeof(double) * (size_t)nsys * N * DIM);
  L.r32 = o;
  o += align256(sizeof(float) * (size_t)nsys * N * DIM);
  L.d32 = o;
  o += align256(sizeof(float) * (size_t)nsys * N * DIM);
  const int64_t nblk = std::max<int64_t>(ku_blocks_max<DIM, double>(g, nsys), vec_blocks(N));
  L.part = o;
  o += align256(sizeof(double) * (size_t)nsys * nblk);
  L.bn2 = o;
  o += align256(sizeof(double) * nsys);
  L.rr2 = o;
  o += align256(sizeof(double) * nsys);
  L.tin = o;
  o += align256(sizeof(double) * nsys);
  L.total = o;
  return L;
}

template <int DIM>
int refined_core(const lpm_grid* g, int nb, int nl, int layout64, const void* kn64, const void* tv64, Bcast bc64,
                 int layout32, const void* kn32, const void* tv32, Bcast bc32, const uint8_t* mask, Bcast bm,
                 const float* minv32, Bcast bminv, const double* b64, double* x64, int use_x0, const lpm_cg_opts* o,
                 char* work, int32_t* iters, double* relres, int32_t* outer, int32_t* conv, hipStream_t st) {
  constexpr int D = DIM;
  const int nsys = nb * nl;
  const int64_t N = nsites(g);
  const int64_t nvec = (int64_t)nsys * N * D;
  const RefLayout L = ref_layout<DIM>(g, nsys);
  double* r64 = (double*)(work + L.r64);
  float* r32 = (float*)(work + L.r32);
  float* d32 = (float*)(work + L.d32);
  double* part = (double*)(work + L.part);
  double* bn2 = (double*)(work + L.bn2);
  double* rr2 = (double*)(work + L.rr2);
  double* tin = (double*)(work + L.tin);
  const Bcast bvec{(int64_t)nl * N * D, N * D};
  const dim3 vgrid((unsigned)vec_blocks(N), nl, nb);
  const unsigned egrid = (unsigned)((nvec + 255) / 256);
  const KuGeom kg = ku_geom<DIM, double>(g, nb, nl);

  hipLaunchKernelGGL((masked_norm2_kernel<D, double>), vgrid, dim3(256), 0, st, b64, N, nl, mask, bm, part);
  hipLaunchKernelGGL((sumsq_fin_kernel<double>), dim3(nsys), dim3(256), 0, st, part, (int)vgrid.x, bn2);
  if (!use_x0) hipLaunchKernelGGL(zero_kernel, dim3(egrid), dim3(256), 0, st, x64, nvec);
  LAUNCHCHK();
  std::vector<double> hb(nsys), hr(nsys), ht(nsys);
  std::vector<int32_t> it_tot(nsys, 0), it_in(nsys), cv_in(nsys);
  std::vector<double> rel_in(nsys);
  std::vector<int32_t> cv(nsys, 0);
  int nout = 0;
  lpm_cg_opts oi = *o;
  oi.tol = 1.0;
  for (;;) {
    KuCall c{x64, bvec, kn64, tv64, bc64, mask, bm, b64, bvec, r64, bvec, part, nullptr};
    if (int rc = ku_call<DIM, double, EPI_RESID>(g, nb, nl, layout64, kg, c, st)) return rc;
    hipLaunchKernelGGL((sumsq_fin_kernel<double>), dim3(nsys), dim3(256), 0, st, part, (int)kg.grid.x, rr2);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(hb.data(), bn2, sizeof(double) * nsys, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hr.data(), rr2, sizeof(double) * nsys, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    bool any = false;
    for (int s = 0; s < nsys; ++s) {
      const double bn = std::sqrt(hb[s]), rn = std::sqrt(hr[s]);
      if (!std::isfinite(rn)) return fail(LPM_E_NONFINITE, "non-finite FP64 residual");
      relres[s] = bn > 0 ? rn / bn : 0.0;
      cv[s] = (bn == 0.0 || rn <= o->tol * bn) ? 1 : 0;
      // inner tolerance: enough to reach the target in this step, never looser than 0.5
      ht[s] = cv[s] ? 2.0 : std::min(0.5, std::max(o->inner_tol, 0.5 * o->tol / relres[s]));
      any |= !cv[s];
    }
    if (!any || nout >= o->max_outer) break;
    HIPCHK(hipMemcpyAsync(tin, ht.data(), sizeof(double) * nsys, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((cast_kernel<double, float>), dim3(egrid), dim3(256), 0, st, r64, r32, nvec);
    LAUNCHCHK();
    char* iw = work + L.inner;
    if (int rc = pcg_core<DIM, float>(g, nb, nl, layout32, kn32, tv32, bc32, mask, bm, minv32, bminv, r32, d32, 0,
                                      &oi, iw, tin, it_in.data(), rel_in.data(), cv_in.data(), st))
      return rc;
    hipLaunchKernelGGL((add_cast_kernel<double, float>), dim3(egrid), dim3(256), 0, st, x64, d32, nvec);
    LAUNCHCHK();
    for (int s = 0; s < nsys; ++s) it_tot[s] += it_in[s];
    ++nout;
  }
  for (int s = 0; s < nsys; ++s) {
    if (iters) iters[s] = it_tot[s];
    if (conv) conv[s] = cv[s];
  }
  if (outer) *outer = nout;
  return LPM_OK;
}

}  // namespace

extern "C" {

size_t lpm_refined_workspace(const lpm_grid* g, int32_t nsys) {
  if (check_grid(g) || nsys < 1) return 0;
  return g->dim == 3 ? ref_layout<3>(g, nsys).total : ref_layout<2>(g, nsys).total;
}

int lpm_pcg_refined(const lpm_grid* g, int32_t nb, int32_t nl, int32_t layout64, const void* kn64, const void* tv64,
                    lpm_bcast bc64, int32_t layout32, const void* kn32, const void* tv32, lpm_bcast bc32,
                    const uint8_t* mask, lpm_bcast bm, const void* minv32, lpm_bcast bminv, const void* b64,
                    void* x64, int32_t use_x0, const lpm_cg_opts* opts, void* work, int32_t* iters, double* relres,
                    int32_t* outer, int32_t* conv, void* stream) {
  if (int rc = check_solve_args(g, nb, nl, opts, minv32)) return rc;
  if (!kn64 || !tv64 || !kn32 || !tv32 || !b64 || !x64 || !work || !relres)
    return fail(LPM_E_BADARG, "lpm_pcg_refined: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (g->dim == 3)
    return refined_core<3>(g, nb, nl, layout64, kn64, tv64, bc_of(bc64), layout32, kn32, tv32, bc_of(bc32), mask,
                           bc_of(bm), (const float*)minv32, bc_of(bminv), (const double*)b64, (double*)x64, use_x0,
                           opts, (char*)work, iters, relres, outer, conv, st);
  return refined_core<2>(g, nb, nl, layout64, kn64, tv64, bc_of(bc64), layout32, kn32, tv32, bc_of(bc32), mask,
                         bc_of(bm), (const float*)minv32, bc_of(bminv), (const double*)b64, (double*)x64, use_x0, opts,
                         (char*)work, iters, relres, outer, conv, st);
}

}  // extern "C"

// ================================================================================================== serving
namespace {

struct Server {
  int max_slots = 0;
  int threads = 256;  // workgroup size of the slot kernel
  char* arena = nullptr;
  size_t arena_bytes = 0;
  std::map<size_t, size_t> free_blocks;  // offset -> size, coalesced
  SlotDesc* d_slots = nullptr;           // device descriptors
  SlotDesc* h_slots = nullptr;           // pinned host mirror
  std::vector<size_t> off, len;          // arena block of each slot (len 0 = slot free)
  std::map<long long, hipGraphExec_t> graphs;
  hipGraph_t graph_keep[64];
  int ngraph = 0;
  hipStream_t cap = nullptr;  // private stream for graph capture (the caller's may be the null stream)
};

size_t slot_bytes(int nx, int ny) {
  const size_t N = (size_t)nx * ny;
  const size_t sizes[] = {N * 8 * 4, N * 8 * 4, N * 2, N * 2 * 8, N * 2 * 8, N * 2 * 8, N * 4 * 8, N * 3 * 4,
                          N * 2 * 4, N * 2 * 4, N * 2 * 4, N * 2 * 4, N * 2 * 4, N * 4 * 4};
  size_t t = 0;
  for (size_t v : sizes) t += align256(v);
  return t;
}

bool arena_take(Server* s, size_t n, size_t* off) {
  for (auto it = s->free_blocks.begin(); it != s->free_blocks.end(); ++it) {
    if (it->second < n) continue;
    *off = it->first;
    const size_t rest = it->second - n, at = it->first + n;
    s->free_blocks.erase(it);
    if (rest) s->free_blocks[at] = rest;
    return true;
  }
  return false;
}

void arena_give(Server* s, size_t off, size_t n) {
  auto it = s->free_blocks.emplace(off, n).first;
  auto nx = std::next(it);
  if (nx != s->free_blocks.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    s->free_blocks.erase(nx);
  }
  if (it != s->free_blocks.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      s->free_blocks.erase(it);
    }
  }
}

}  // namespace

extern "C" {

size_t lpm_serve_footprint(int32_t nx, int32_t ny) {
  if (nx < 1 || ny < 1) return 0;
  return slot_bytes(nx, ny);
}

int lpm_serve_create(int32_t max_slots, size_t arena_bytes, int32_t threads, void** server) {
  if (!server || max_slots < 1 || max_slots > 65535 || arena_bytes < 256 ||
      (threads != 256 && threads != 512 && threads != 1024))
    return fail(LPM_E_BADARG, "lpm_serve_create: bad arguments");
  Server* s = new Server();
  s->max_slots = max_slots;
  s->threads = threads;
  s->arena_bytes = arena_bytes & ~(size_t)255;
  hipError_t e = hipMalloc((void**)&s->arena, s->arena_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_slots, sizeof(SlotDesc) * max_slots);
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_slots, sizeof(SlotDesc) * max_slots, hipHostMallocDefault);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking);
  if (e == hipSuccess) {
    std::memset(s->h_slots, 0, sizeof(SlotDesc) * max_slots);
    e = hipMemcpy(s->d_slots, s->h_slots, sizeof(SlotDesc) * max_slots, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    if (s->arena) (void)hipFree(s->arena);
    if (s->d_slots) (void)hipFree(s->d_slots);
    if (s->h_slots) (void)hipHostFree(s->h_slots);
    if (s->cap) (void)hipStreamDestroy(s->cap);
    delete s;
    return fail(LPM_E_HIP, "lpm_serve_create: %s", hipGetErrorString(e));
  }
  s->free_blocks[0] = s->arena_bytes;
  s->off.assign(max_slots, 0);
  s->len.assign(max_slots, 0);
  *server = s;
  return LPM_OK;
}

void lpm_serve_destroy(void* server) {
  Server* s = (Server*)server;
  if (!s) return;
  (void)hipDeviceSynchronize();
  for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
  for (int i = 0; i < s->ngraph; ++i) (void)hipGraphDestroy(s->graph_keep[i]);
  (void)hipFree(s->arena);
  (void)hipFree(s->d_slots);
  (void)hipHostFree(s->h_slots);
  (void)hipStreamDestroy(s->cap);
  delete s;
}

int lpm_serve_admit(void* server, int32_t nx, int32_t ny, double h, const float* kn, const float* tv,
                    const uint8_t* mask, const double* up, double tol, double inner_tol, int32_t maxit, void* stream,
                    int32_t* slot) {
  Server* s = (Server*)server;
  if (!s || !kn || !tv || !mask || !up || !slot || nx < 1 || ny < 1 || (int64_t)nx * ny >= (1 << 30) || maxit < 1 ||
      !(tol > 0))
    return fail(LPM_E_BADARG, "lpm_serve_admit: bad arguments");
  int id = -1;
  for (int i = 0; i < s->max_slots; ++i)
    if (s->len[i] == 0) {
      id = i;
      break;
    }
  const size_t need = slot_bytes(nx, ny);
  size_t off = 0;
  if (id < 0 || !arena_take(s, need, &off)) return fail(LPM_E_FULL, "lpm_serve_admit: no free slot or arena space");
  s->off[id] = off;
  s->len[id] = need;
  const size_t N = (size_t)nx * ny;
  char* base = s->arena + off;
  auto take = [&](size_t n) {
    char* p = base;
    base += align256(n);
    return p;
  };
  SlotDesc d;
  std::memset(&d, 0, sizeof d);
  d.nx = nx;
  d.ny = ny;
  d.h = (float)h;
  d.maxit = maxit;
  d.tol = tol;
  d.inner_tol = inner_tol;
  float* dkn = (float*)take(N * 8 * 4);
  float* dtv = (float*)take(N * 8 * 4);
  uint8_t* dmask = (uint8_t*)take(N * 2);
  d.kn = dkn;
  d.tv = dtv;
  d.mask = dmask;
  d.up = (double*)take(N * 2 * 8);
  d.x = (double*)take(N * 2 * 8);
  d.b = (double*)take(N * 2 * 8);
  d.w64 = (double*)take(N * 4 * 8);
  d.minv = (float*)take(N * 3 * 4);
  d.d = (float*)take(N * 2 * 4);
  d.r = (float*)take(N * 2 * 4);
  d.z = (float*)take(N * 2 * 4);
  d.p = (float*)take(N * 2 * 4);
  d.ap = (float*)take(N * 2 * 4);
  d.w32 = (float*)take(N * 4 * 4);
  d.state = SLOT_INIT;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(dkn, kn, N * 8 * 4, hipMemcpyDefault, st));
  HIPCHK(hipMemcpyAsync(dtv, tv, N * 8 * 4, hipMemcpyDefault, st));
  HIPCHK(hipMemcpyAsync(dmask, mask, N * 2, hipMemcpyDefault, st));
  HIPCHK(hipMemcpyAsync(d.up, up, N * 2 * 8, hipMemcpyDefault, st));
  s->h_slots[id] = d;
  HIPCHK(hipMemcpyAsync(s->d_slots + id, s->h_slots + id, sizeof(SlotDesc), hipMemcpyHostToDevice, st));
  *slot = id;
  return LPM_OK;
}

void lpm_serve_set_fused(int32_t fused) { g_serve_fused = fused != 0; }

int lpm_serve_step(void* server, int64_t budget, int32_t use_graph, void* stream) {
  Server* s = (Server*)server;
  if (!s || budget < 1) return fail(LPM_E_BADARG, "lpm_serve_step: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid((unsigned)s->max_slots), block((unsigned)s->threads);
  if (!use_graph) {
    hipLaunchKernelGGL(serve_kernel, grid, block, 0, st, s->d_slots, s->max_slots, (long long)budget,
                       g_serve_fused);
    LAUNCHCHK();
    return LPM_OK;
  }
  const long long key = 2 * (long long)budget + g_serve_fused;
  auto it = s->graphs.find(key);
  if (it == s->graphs.end()) {
    if (s->ngraph >= 64) return fail(LPM_E_BADARG, "lpm_serve_step: too many distinct budgets");
    hipGraph_t gr;
    HIPCHK(hipStreamBeginCapture(s->cap, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(serve_kernel, grid, block, 0, s->cap, s->d_slots, s->max_slots, (long long)budget,
                       g_serve_fused);
    HIPCHK(hipStreamEndCapture(s->cap, &gr));
    s->graph_keep[s->ngraph++] = gr;  // destroyed with the server
    hipGraphExec_t ex;
    HIPCHK(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
    it = s->graphs.emplace(key, ex).first;
  }
  HIPCHK(hipGraphLaunch(it->second, st));
  return LPM_OK;
}

int lpm_serve_poll(void* server, int32_t* state, int32_t* iters, int32_t* outer, double* relres, void* stream) {
  Server* s = (Server*)server;
  if (!s || !state) return fail(LPM_E_BADARG, "lpm_serve_poll: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemcpyAsync(s->h_slots, s->d_slots, sizeof(SlotDesc) * s->max_slots, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int i = 0; i < s->max_slots; ++i) {
    const SlotDesc& d = s->h_slots[i];
    state[i] = s->len[i] ? d.state : SLOT_FREE;
    if (iters) iters[i] = d.total;
    if (outer) outer[i] = d.outer;
    if (relres) relres[i] = d.rel;
  }
  return LPM_OK;
}

int lpm_serve_result(void* server, int32_t slot, double* u, void* stream) {
  Server* s = (Server*)server;
  if (!s || !u || slot < 0 || slot >= s->max_slots || s->len[slot] == 0)
    return fail(LPM_E_BADARG, "lpm_serve_result: bad slot");
  const SlotDesc& d = s->h_slots[slot];
  if (d.state != SLOT_DONE) return fail(LPM_E_BADARG, "lpm_serve_result: slot %d is not done", slot);
  HIPCHK(hipMemcpyAsync(u, d.up, (size_t)d.nx * d.ny * 2 * 8, hipMemcpyDefault, (hipStream_t)stream));
  return LPM_OK;
}

int lpm_serve_release(void* server, int32_t slot) {
  Server* s = (Server*)server;
  if (!s || slot < 0 || slot >= s->max_slots || s->len[slot] == 0)
    return fail(LPM_E_BADARG, "lpm_serve_release: bad slot");
  const int st = s->h_slots[slot].state;
  if (st != SLOT_DONE && st != SLOT_FAILED)
    return fail(LPM_E_BADARG, "lpm_serve_release: slot %d is still running", slot);
  arena_give(s, s->off[slot], s->len[slot]);
  s->len[slot] = 0;
  return LPM_OK;
}

size_t lpm_serve_arena_free(void* server) {
  Server* s = (Server*)server;
  if (!s) return 0;
  size_t t = 0;
  for (auto& kv : s->free_blocks) t += kv.second;
  return t;
}

}  // extern "C"
