/* C ABI of liblpm.so: the MI355X (gfx950) kernels of the LPM elastic solve path.
 *
 * Every buffer is a device pointer owned by the caller (PyTorch tensors' data_ptr(), or hipMalloc from another
 * host).  `stream` is a hipStream_t (the caller's current stream).  Every entry point returns an int status
 * (LPM_OK = 0); lpm_last_error() describes the last failure of the calling thread.  Apart from lpm_serve_create,
 * which allocates a server's arena once, no entry point allocates device memory: solvers take a caller-owned
 * workspace whose size the *_workspace query returns.
 *
 * Lattices are full boxes, particle id p = (z * ny + y) * nx + x; vectors are (systems, N, dim) interleaved,
 * directed coefficients (batch, N, NN) with NN = 18 (3D) or 8 (2D) in the reference slot order; packed
 * coefficients are (batch, NN/2, P) forward-slot planes (valid for symmetric parameters), each plane holding
 * G zero guards, the N particles and at least G trailing zero guards (G = nx ny + nx + 1 in 3D, nx + 1 in 2D,
 * rounded up to 64; P = N + 2G rounded up to 64 = lpm_packed_len), so the kernels read a neighbour's entry without a
 * bounds test and FP32 planes start 256-byte aligned; the entry of a forward slot whose bond leaves the box is zero
 * (lpm_pack_coef writes it so), like the guards.  A batch of systems is
 * (nb, nl) = (specimens, load cases); every operand carries element strides (sb, sl) for the two batch indices,
 * 0 meaning broadcast, which is the reference's (B,1,N,18) against (1|B,L,N,3) convention
 * (coarse3d/torch_lpm3d.py:212-227).
 */
#ifndef LPM_H_
#define LPM_H_
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LPM_OK = 0,
  LPM_E_BADARG = 1,      /* bad shape, dtype or pointer */
  LPM_E_HIP = 2,         /* HIP runtime error */
  LPM_E_NOT_SPD = 3,     /* preconditioner block or curvature not positive */
  LPM_E_NONFINITE = 4,   /* NaN / Inf met in the iteration */
  LPM_E_UNSUPPORTED = 5, /* valid request the library does not implement */
  LPM_E_FULL = 6         /* serving: no free slot or arena space (nothing changed) */
};
enum { LPM_F32 = 0, LPM_F64 = 1 };
enum { LPM_COEF_DIRECTED = 0, LPM_COEF_PACKED = 1 };

typedef struct {
  int64_t sb, sl; /* element strides of the (b, l) batch indices; 0 = broadcast */
} lpm_bcast;

typedef struct {
  int32_t dim;        /* 2 or 3 */
  int32_t nx, ny, nz; /* nz = 1 in 2D */
  double h;           /* lattice spacing */
  /* Optional removed particles (voids), the per-specimen validity of cement3d/torch_v1.py:36-53: NULL, or bond
   * words uint32 (.., N) per system at bonds + b*bb.sb + l*bb.sl (the (b, l) indices of the call's coefficient
   * batch; lpm_bond_grad / lpm_post_vjp: (coefficient batch, load)).  Bit k of particle p set = slot k is a bond;
   * a bond must be set at both ends (bit k at p, bit k^1 at its neighbour), bits of slots leaving the box are
   * ignored.  Every operator, preconditioner, solver and post-processing call then uses only these bonds
   * (shell sums, forces, stress opp-flag rule, block diagonal); particles with no bond must be fixed. */
  const uint32_t* bonds;
  lpm_bcast bb;
} lpm_grid;

typedef struct {
  int32_t precond; /* 0 = none (plain CG), 1 = particle block Jacobi */
  int32_t maxit;   /* iteration cap per system (inner iterations in refined mode) */
  int32_t check;   /* host polls the convergence flags every `check` iterations */
  int32_t graph;   /* 1 = replay the `check` iterations as a captured hipGraph */
  double tol;      /* stop a system at ||r|| <= tol ||b|| */
  double inner_tol;/* refined mode: relative tolerance of every FP32 inner solve */
  int32_t max_outer;/* refined mode: cap on FP64 refinement steps */
  int32_t pad_;
} lpm_cg_opts;

const char* lpm_last_error(void);
int lpm_version(void);

/* Timing of the K u launches inside lpm_pcg_solve / lpm_pcg_refined on the calling thread (HIP events around
 * every launch, graph replays excluded).  No counterpart in the reference: it feeds the roofline measurement. */
void lpm_profile_ku(int32_t enable);
/* Tiling of the 3D K u kernels: wide > 0 selects the 64-wide x tile, < 0 the narrow one, 0 automatic; zc planes
 * per block (0 automatic).  Process-wide; no counterpart in the reference (tuning knob). */
void lpm_set_ku_tiling(int32_t wide, int32_t zc);
/* lean = 1: 3D FP32 K u on packed coefficients without bond words runs the lean kernel (three blocks per CU,
 * own-bond terms summed as the shell sums form, 64-wide tiles); 0 (default): the general kernel.  Process-wide; no
 * counterpart in the reference (tuning / comparison knob). */
void lpm_set_ku_kernel(int32_t lean);
/* box = 1: every 3D K u runs the box kernel (one block per 16 x 4 x 4 sites, no z march); 0: never; -1 (default):
 * automatic, the box kernel where the marching kernel's launch would have fewer blocks than the device has CUs
 * (small lattices with few systems) and no tiling override is set.  Process-wide; no counterpart in the reference
 * (tuning knob). */
void lpm_set_ku_box(int32_t box);
void lpm_profile_ku_read(double* total_ms, int64_t* launches);

/* Bond words from solid flags (nbatch, N) bytes, nonzero = solid: bit k of p set iff slot k stays in the box and
 * p and its neighbour are both solid; words (nbatch, N) uint32.  Replaces SCTV.set_mask (cement3d/torch_v1.py:36-53),
 * which builds the same per-specimen validity as a (B, 1, N, 18) tensor.  g->bonds is not used. */
int lpm_bond_words(const lpm_grid* g, int32_t nbatch, const uint8_t* solid, uint32_t* words, void* stream);

/* out (nb, nl, N, dim) = K u.  Replaces SCT.apply (coarse3d/torch_lpm3d.py:81-83), lin3d.k_apply
 * (coarse3d/lin3d.py:126-149) and the 2D _kmv (timing/t2d/solver2d.py:269-305). */
int lpm_apply(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* u,
              lpm_bcast bu, const void* kn, const void* tv, lpm_bcast bc, void* out, void* stream);

/* out (nb, nl, N, dim) = mask * (rhs - K u): the true residual of the free DOFs (mask per DOF, 1 = free; null =
 * all free).  work: lpm_residual_workspace bytes.  Replaces the fused residual _kx_free(sub=True)
 * (timing/t3d/amg3d.py:199-229) and the FP64 residual of the refinement loop. */
size_t lpm_residual_workspace(const lpm_grid* g, int32_t nsys, int32_t dtype);
int lpm_residual(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* u,
                 lpm_bcast bu, const void* kn, const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm,
                 const void* rhs, lpm_bcast brhs, void* out, void* work, void* stream);

/* Plane length P of the packed layout (N plus the guard bands, rounded up to 64 values). */
int64_t lpm_packed_len(const lpm_grid* g);

/* Directed (nbatch, N, NN) -> packed (nbatch, NN/2, P) forward planes, converting dtype_in -> dtype_out.
 * asym (device, nbatch uint32, zeroed by the caller) counts bonds whose two directed values differ by more than
 * rtol.  Replaces the symmetric duplication of the reference's Scatter (coarse3d/train3d.py:77-97). */
int lpm_pack_coef(const lpm_grid* g, int32_t nbatch, int32_t dtype_in, int32_t dtype_out, const void* kn,
                  const void* tv, void* kn_fwd, void* tv_fwd, double rtol, uint32_t* asym, void* stream);

/* Block-Jacobi preconditioner out (nb, nl, N, 6|3) symmetric (xx yy zz yz xz xy | xx yy xy), inverse of the
 * free-DOF diagonal block, zero rows/cols on fixed DOFs; raw = 1 writes the diagonal blocks D (nb, nl, N, dim, dim)
 * instead.  mask (per DOF, 1 = free) may be null.  Replaces SCT.block_diag / precond (coarse3d/torch_lpm3d.py:94-106),
 * lin3d.k_block_diag / k_block_jacobi (coarse3d/lin3d.py:239-301) and _diag_inv (timing/t2d/solver2d.py:308-346). */
int lpm_block_jacobi(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype_c, int32_t dtype_out, int32_t layout,
                     const void* kn, const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm, void* out,
                     int32_t raw, void* stream);

/* Per-particle outputs from u (any of the output pointers may be null): ku (nb,nl,N,dim), stress (nb,nl,N,6|3),
 * energy (nb,nl,N), dl / force (nb,nl,N,NN).  work: nb*nl*N*4 values of dtype.  Replaces SCT.forces / stress /
 * energy (coarse3d/torch_lpm3d.py:76-92), lin3d.k_bond_forces / k_stress / k_energy (coarse3d/lin3d.py:152-236),
 * Lattice2D.bond_forces / stress / energy_per_particle (coarse2d/lin2d.py:121-163). */
int lpm_post(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* u, lpm_bcast bu,
             const void* kn, const void* tv, lpm_bcast bc, void* work, void* ku, void* stress, void* energy, void* dl,
             void* force, void* stream);

/* Adjoint bond-parameter gradient: gkn, gtv (ncb, N, NN) = scale * d/d(Kn, Tv) sum_s lam_s . K u_s, with systems
 * s = cb * spc + j (j < spc) contiguous in u and lam (ncb*spc, N, dim).  Coefficients are only read for the shell
 * sums' layout check (kn / tv directed or packed, batch ncb).  work: 2 * ncb*spc*N*4 values.  Replaces the
 * autograd part of LinSolve.backward (coarse3d/torch_lpm3d.py:165-175). */
int lpm_bond_grad(const lpm_grid* g, int32_t ncb, int32_t spc, int32_t dtype, int32_t layout, const void* kn,
                  const void* tv, const void* u, const void* lam, double scale, void* work, void* gkn, void* gtv,
                  void* stream);

/* Vector-Jacobian product of lpm_post's (K u, stress, energy): for cotangents gku (ncb*spc, N, dim), gstress
 * (ncb*spc, N, 6|3) and genergy (ncb*spc, N) (any may be NULL = zero) writes gu (ncb*spc, N, dim) and the
 * directed parameter gradients gkn, gtv (ncb, N, NN) summed over the spc systems of each coefficient batch.
 * work: ncb*spc*N*(4 + 2*NN) values.  Replaces the autograd backward through the post-processing of
 * CoarseT.run (coarse3d/torch_lpm3d.py:212-227) and of the 2D CoarseT.stress / energy_density
 * (coarse2d/torch_lpm.py:246-258). */
int lpm_post_vjp(const lpm_grid* g, int32_t ncb, int32_t spc, int32_t dtype, int32_t layout, const void* kn,
                 const void* tv, const void* u, const void* gku, const void* gstress, const void* genergy, void* work,
                 void* gu, void* gkn, void* gtv, void* stream);

/* Batched (P)CG: K_ff x_f = b_f for nb*nl systems, b and x (nb*nl, N, dim) contiguous, x0 = x when use_x0 else 0.
 * minv (symmetric blocks, lpm_block_jacobi layout) is required when opts->precond = 1.  iters / relres / conv are
 * host arrays of nb*nl entries (conv = 1 when ||r|| <= tol ||b||).  Replaces SCT.pcg (coarse3d/torch_lpm3d.py:109-140),
 * lin3d._pcg (coarse3d/lin3d.py:464-518), cgplain.cg3 (timing/t3d/cgplain.py:11-55) and cg2plain.cg_mf
 * (timing/t2d/cg2plain.py:11-49). */
size_t lpm_pcg_workspace(const lpm_grid* g, int32_t nsys, int32_t dtype);
int lpm_pcg_solve(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* kn,
                  const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm, const void* minv, lpm_bcast bminv,
                  const void* b, void* x, int32_t use_x0, const lpm_cg_opts* opts, void* work, int32_t* iters,
                  double* relres, int32_t* conv, void* stream);

/* Multi-tolerance (P)CG: the solve of lpm_pcg_solve, stopped per system once its recursive relative residual meets
 * every tolerance of the descending list tols[0 .. ntol) (ntol <= 64; opts->tol is not used).  snaps (device,
 * (nb*nl, ntol, N, dim), dtype) receives the iterate of the first iteration at which each level is met (iteration 0
 * included) and snap_iters (host, nb*nl*ntol) that iteration; levels never met (maxit) get the final iterate and -1.
 * Replaces pcg_mf / solve_mf with TOLS (timing/t2d/solver2d.py:27, 349-421, 577-613) and the snapshots of pcg_bj
 * (:64-142).  Workspace: lpm_pcg_workspace. */
int lpm_pcg_snapshots(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* kn,
                      const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm, const void* minv,
                      lpm_bcast bminv, const void* b, void* x, int32_t use_x0, const lpm_cg_opts* opts, int32_t ntol,
                      const double* tols, void* snaps, int32_t* snap_iters, void* work, int32_t* iters,
                      double* relres, int32_t* conv, void* stream);

/* Plain FP64 CG of one 2D square-lattice system in one cooperative launch (cg2plain.cg_mf,
 * timing/t2d/cg2plain.py:11-49): kn, tv (nx*ny, 8) directed FP64, mask (N, 2) per-DOF bytes (1 = free), b (N, 2);
 * x (N, 2) receives the solution from a zero start; stops when ||r|| <= tol ||b|| (0 iterations when r0 = mask b
 * already does) or after maxit iterations (conv = 0).  workgroups: 0 = automatic (about 64), else about that many (at
 * most 64), each owning ceil(ny / workgroups) rows; LPM_E_UNSUPPORTED when those rows do not fit one workgroup
 * (1024 particles, nx <= 128, 8 ((6 rows + 16) nx + 100) bytes of LDS) or the workgroups cannot all be resident.
 * LPM_E_NOT_SPD when p.Ap <= 0, LPM_E_NONFINITE on NaN / Inf,
 * LPM_E_HIP when a grid-barrier wait times out.  work: lpm_cg2_persistent_workspace(nx, ny) bytes.  Returns after
 * the solve (synchronises the stream). */
size_t lpm_cg2_persistent_workspace(int32_t nx, int32_t ny);
int lpm_cg2_persistent(int32_t nx, int32_t ny, const double* kn, const double* tv, const uint8_t* mask,
                       const double* b, double* x, double tol, int32_t maxit, int32_t workgroups, void* work,
                       int32_t* iters, double* relres, int32_t* conv, void* stream);

/* The same solve (timing/t2d/cg2plain.py:11-49) in the pipelined form of CG (Ghysels and Vanroose 2014, Alg. 4
 * without preconditioner): one grid-wide reduction per iteration, (r.r, w.r) with w = A r kept by recurrence, whose
 * exchange overlaps the product A w.  Same Krylov iterates as lpm_cg2_persistent in exact arithmetic; rounding
 * differs, so iteration counts may differ by a few.  Arguments, workspace (lpm_cg2_persistent_workspace), limits
 * and errors as lpm_cg2_persistent; workgroups 0 = automatic (up to 64). */
int lpm_cg2_pipelined(int32_t nx, int32_t ny, const double* kn, const double* tv, const uint8_t* mask,
                      const double* b, double* x, double tol, int32_t maxit, int32_t workgroups, void* work,
                      int32_t* iters, double* relres, int32_t* conv, void* stream);

/* lpm_pcg_solve without a workspace argument and without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a file-scope device array of lpm_noalloc_scratch_bytes() (64 MiB) that each GPU the
 * library runs on holds until the process ends; a solve whose lpm_pcg_workspace exceeds it is refused with
 * LPM_E_BADARG.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU (other threads or
 * streams) race.  opts->graph is ignored (no graph is captured) and no profiling events are created. */
size_t lpm_noalloc_scratch_bytes(void);
int lpm_pcg_solve_noalloc(const lpm_grid* g, int32_t nb, int32_t nl, int32_t dtype, int32_t layout, const void* kn,
                          const void* tv, lpm_bcast bc, const uint8_t* mask, lpm_bcast bm, const void* minv,
                          lpm_bcast bminv, const void* b, void* x, int32_t use_x0, const lpm_cg_opts* opts,
                          int32_t* iters, double* relres, int32_t* conv, void* stream);

/* Mixed precision: FP32 (P)CG inner solves inside an FP64 iterative-refinement loop; the true residual
 * r = mask * (b - K x) and the update x += d are FP64.  kn64/tv64 (layout64) define the FP64 operator, kn32/tv32
 * (layout32) the FP32 one; minv32 is the FP32 preconditioner.  relres is the true FP64 relative residual, iters the
 * total inner iterations, outer the refinement steps.  Same problem as lpm_pcg_solve. */
size_t lpm_refined_workspace(const lpm_grid* g, int32_t nsys);
int lpm_pcg_refined(const lpm_grid* g, int32_t nb, int32_t nl, int32_t layout64, const void* kn64, const void* tv64,
                    lpm_bcast bc64, int32_t layout32, const void* kn32, const void* tv32, lpm_bcast bc32,
                    const uint8_t* mask, lpm_bcast bm, const void* minv32, lpm_bcast bminv, const void* b64,
                    void* x64, int32_t use_x0, const lpm_cg_opts* opts, void* work, int32_t* iters, double* relres,
                    int32_t* outer, int32_t* conv, void* stream);

/* ---- Serving: independent 2D square-lattice solves of different sizes, continuously batched.
 * A server owns a device arena and max_slots slots, each advanced by one workgroup of `threads` (256, 512, 1024).  lpm_serve_admit copies one request (kn, tv (N, 8) FP32
 * directed symmetric, mask (N, 2) 1 = free DOF, up (N, 2) FP64 prescribed displacement, zero on free DOFs) into
 * a slot, or returns LPM_E_FULL when no slot or arena space is free.  lpm_serve_step launches (optionally as a
 * replayed hipGraph) one persistent kernel in which every active slot advances its own FP32 block-Jacobi PCG
 * inside an FP64 refinement by at most `budget` site-iterations; requests admitted between steps join the next
 * one.  lpm_serve_poll synchronises the stream and reports per slot: state (0 free, 1-3 running, 4 done,
 * 5 failed), inner iterations, refinement steps and the true FP64 relative residual.  lpm_serve_result copies
 * u = up + x (N, 2) FP64 of a done slot; lpm_serve_release returns the slot and its arena block.  A server is
 * not thread-safe.  The solve replaces CoarseT.solve's dense Cholesky (coarse2d/torch_lpm.py:216-229) for
 * request streams; the slot arena follows the survey's proposed lpm_serving_* interface (admit, step, retire). */
size_t lpm_serve_footprint(int32_t nx, int32_t ny);
int lpm_serve_create(int32_t max_slots, size_t arena_bytes, int32_t threads, void** server);
void lpm_serve_destroy(void* server);
int lpm_serve_admit(void* server, int32_t nx, int32_t ny, double h, const float* kn, const float* tv,
                    const uint8_t* mask, const double* up, double tol, double inner_tol, int32_t maxit, void* stream,
                    int32_t* slot);
int lpm_serve_step(void* server, int64_t budget, int32_t use_graph, void* stream);
/* Tuning switch (process-wide, no counterpart in the reference): 1 (default) runs the inner iterations' shell sums and
 * product as one banded pass through LDS, 0 as two passes through device memory; same operator, the partial sums of
 * p.Ap follow a different order. */
void lpm_serve_set_fused(int32_t fused);
int lpm_serve_poll(void* server, int32_t* state, int32_t* iters, int32_t* outer, double* relres, void* stream);
int lpm_serve_result(void* server, int32_t slot, double* u, void* stream);
int lpm_serve_release(void* server, int32_t slot);
size_t lpm_serve_arena_free(void* server);

/* ---- Aggregation AMG preconditioner of the 3D SC lattice (timing/t3d/amg3d.py, AMG_CFG of timing/t3d/problems.py).
 * Level 0: the free particles (a box: first particle lo, extent L) in chunks of agg particles per axis, remainder
 * merged into the last chunk, g0 aggregates per axis (amg3d.py:42-48, 438-462).  The rigid-body-mode tentative
 * prolongator block of a particle is read from the shape templates tmpl + toff[s] ((sx sy sz, 3, 6) row-major, s =
 * bit a set for the last chunk along axis a), the Gram-Schmidt basis of gmg._tentative (timing/t3d/gmg.py:70-143),
 * which depends on the aggregate's shape only.  Stencil level k (k = 0 is the Galerkin operator of level 0) has
 * the node grid g[k], the symmetric 27-point 6x6 stencil A[k] (offset d = (dz+1) 9 + (dy+1) 3 + (dx+1)) of which the
 * blocks d = 13 ... 26 are stored structure-of-arrays A[((d - 13) 36 + r 6 + c) n + I] (a block d < 13 is the
 * transpose of the neighbour's block 26 - d), and for k < nlev - 1 the per-node blocks Q[k] (n, 6, 6) towards level k + 1,
 * the inverses Minv[k] (SoA (36, n)) of its diagonal blocks and the damping w[k] = omega / rho_k.  dinv is the
 * dense inverse of the coarsest stencil.  All pointers are device memory of the call's dtype. */
#define LPM_AMG_MAXLEV 12
#define LPM_AMG_SCRATCH_BYTES 4096 /* device scratch of lpm_amg_inv_diag / lpm_amg_coarsest / lpm_amg_dot */
typedef struct {
  int32_t lo[3], L[3], g0[3], agg;
  const void* tmpl;
  int64_t toff[8];
  const void* minv0; /* particle block-Jacobi inverse (N, 6) (lpm_block_jacobi layout), zero on fixed DOFs */
  double w0;         /* omega / rho0 */
  int32_t nlev, nu, nu_c, gamma, w_from, pad_;
  int32_t g[LPM_AMG_MAXLEV][3];
  const void* A[LPM_AMG_MAXLEV];
  const void* Q[LPM_AMG_MAXLEV];
  const void* Minv[LPM_AMG_MAXLEV];
  double w[LPM_AMG_MAXLEV];
  const void* dinv;
} lpm_amg;

enum { LPM_AMG_PROLONG_ADD = 0, LPM_AMG_RESTRICT = 1, LPM_AMG_PROBE = 2, LPM_AMG_GATHER = 3 };

/* Level-0 transfer: PROLONG_ADD out (N,3) += P in (nagg,6) (amg3d.py:318-331); RESTRICT out (nagg,6) = P^T in (N,3)
 * (amg3d.py:303-315); PROBE out (6,N,3) = columns of P on the aggregates of colour col (0..26, (I mod 3) per axis);
 * GATHER: column j of out[I, d(I, col)] (stencil A[0] layout, stored blocks only) = P_I^T in_j for in (6,N,3) = K PROBE: with
 * lpm_apply this replaces the bond-level Galerkin product _galerkin0 (amg3d.py:151-196). */
int lpm_amg_level0(const lpm_grid* g, const lpm_amg* a, int32_t dtype, int32_t op, int32_t col, const void* in,
                   void* out, void* stream);
/* The same four operations between stencil levels k and k + 1 with the blocks Q[k] (amg3d.py:256-331). */
int lpm_amg_levelq(const lpm_amg* a, int32_t k, int32_t dtype, int32_t op, int32_t col, const void* in, void* out,
                   void* stream);
/* y = A x (b null) or y = b - A x on a stencil of grid (gx, gy, gz); nsys systems of x and y (n, 6) share A and b
 * (amg3d.py:230-253).  Summation order depends on the stencil kernel (lpm_amg_set_stencil_kernel). */
/* Kernel of the one-system stencil products (process-wide; AUTO by default): ROWS = one thread per (row, node), the
 * small-level form; NODE = one thread per node; TWO_PASS = forward products and stored transposed products, then a
 * gather, reading every stored block once (inside the AMG cycle, whose workspace holds 84 values per node of the
 * first stencil level; lpm_amg_st_mv has no workspace and takes AUTO's other choices).  AUTO: ROWS below 2^17 nodes,
 * TWO_PASS from there (NODE outside the cycle).  A
 * performance switch of this library with no counterpart in the reference (amg3d.py:230-253 is one product). */
enum { LPM_AMG_ST_AUTO = 0, LPM_AMG_ST_ROWS = 1, LPM_AMG_ST_NODE = 2, LPM_AMG_ST_TWO_PASS = 3 };
int lpm_amg_set_stencil_kernel(int32_t mode);
int lpm_amg_st_mv(int32_t gx, int32_t gy, int32_t gz, int32_t nsys, int32_t dtype, const void* A, const void* x,
                  const void* b, void* y, void* stream);
/* Minv (SoA (36, n)) = inverses of the diagonal blocks of the stencil A (np.linalg.inv(A[:, 13]), amg3d.py:536);
 * LPM_E_NOT_SPD on a singular block. */
int lpm_amg_inv_diag(int32_t n, int32_t dtype, const void* A, void* Minv, void* scratch, void* stream);
/* out = w M b (add = 0) or out += w M b: M (n, 6) symmetric particle blocks (k = 3) or SoA 6x6 node blocks
 * (k = 6) (amg3d.py:334-356). */
int lpm_amg_bj(int64_t n, int32_t k, int32_t dtype, const void* M, const void* b, double w, int32_t add, void* out,
               void* stream);
/* dinv ((6n)^2) = inverse of the symmetrised dense matrix of the coarsest stencil (amg3d.py:544-546);
 * LPM_E_NOT_SPD when it is not positive definite. */
int lpm_amg_coarsest(int32_t gx, int32_t gy, int32_t gz, int32_t dtype, const void* A, void* dinv, void* scratch,
                     void* stream);
/* *out = x . y over n values (fixed-order reduction, host result; synchronises the stream). */
int lpm_amg_dot(int64_t n, int32_t dtype, const void* x, const void* y, double* out, void* scratch, void* stream);
/* One preconditioner application z = B b (b (N, 3) zero on fixed DOFs; amg3d.py:620-646 with the W-cycle of
 * amg3d.py:578-618) and the AMG-PCG K_ff x = b from x = 0 with ||r|| <= tol ||b|| tested every iteration
 * (amg3d.py:649-672).  kn / tv (layout) define K; mask per DOF (1 = free).  work: lpm_amg_workspace bytes. */
size_t lpm_amg_workspace(const lpm_grid* g, const lpm_amg* a, int32_t dtype);
int lpm_amg_vcycle(const lpm_grid* g, const lpm_amg* a, int32_t dtype, int32_t layout, const void* kn, const void* tv,
                   const uint8_t* mask, const void* b, void* z, void* work, void* stream);
int lpm_amg_pcg(const lpm_grid* g, const lpm_amg* a, int32_t dtype, int32_t layout, const void* kn, const void* tv,
                const uint8_t* mask, const void* b, void* x, double tol, int32_t maxit, void* work, int32_t* iters,
                double* relres, void* stream);

/* Mixed precision: FP32 AMG-PCG inner solves (hierarchy a32 in FP32, operator kn32 / tv32) inside an FP64
 * iterative refinement in which the true residual r = mask (b - K x) (kn64 / tv64) and the update x += d are FP64;
 * each inner solve stops at min(1/2, max(inner_tol, tol / (2 relres))) of its right-hand side.  iters = total inner
 * iterations, outer = refinement steps, relres = true FP64 relative residual.  Same problem as lpm_amg_pcg; no
 * counterpart in the reference, whose AMG-PCG is FP64 (amg3d.py:649-672). */
size_t lpm_amg_refined_workspace(const lpm_grid* g, const lpm_amg* a32);
int lpm_amg_pcg_refined(const lpm_grid* g, const lpm_amg* a32, int32_t layout64, const void* kn64, const void* tv64,
                        int32_t layout32, const void* kn32, const void* tv32, const uint8_t* mask, const void* b64,
                        void* x64, double tol, double inner_tol, int32_t max_outer, int32_t maxit, void* work,
                        int32_t* iters, int32_t* outer, double* relres, void* stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Bond-network inference: the producer of the coarse Kn / Tv (coarse3d/net3d.py, timing/t3d/infer3d.py).
 * Feature fields are FP32.  Voxel grids are (nz, ny, nx) with x fastest.
 * --------------------------------------------------------------------------------------------------------------- */

/* One O_h-invariant 3x3x3 convolution with zero padding (IsoConv3d.forward, net3d.py:35-49): per output channel o
 * y = b_o + sum_i (w[o,i,0] centre + w[o,i,1] sum of the 6 face + w[o,i,2] sum of the 12 edge + w[o,i,3] sum of the
 * 8 corner neighbours of x_i), then GELU (erf form) when gelu != 0.  x (B, cin, nz, ny, nx), w (cout, cin, 4), b
 * (cout).  With dproj == 0 the output is y (B, cout, nz, ny, nx); with dproj > 0 the 1x1 convolution that ends the
 * trunk (net3d.py:125-126) is applied as well, pw (dproj, cout), pb (dproj), and written voxel-major:
 * y (B, nz*ny*nx, dstride) with channels dproj..dstride-1 set to zero.  cout <= 32, dproj <= dstride <= 16. */
int lpm_iso_conv(int32_t B, int32_t cin, int32_t cout, int32_t nx, int32_t ny, int32_t nz, const float* x,
                 const float* w, const float* b, int32_t gelu, int32_t dproj, int32_t dstride, const float* pw,
                 const float* pb, float* y, void* stream);

/* Cell means of voxel-major features (F.avg_pool3d(Fv, s, s), net3d.py:154): fv (B, (ncp s)^3, dstride) ->
 * fc (B, ncp^3, dstride). */
int lpm_cell_pool(int32_t B, int32_t ncp, int32_t s, int32_t dstride, const float* fv, float* fc, void* stream);

/* One direction group of bond heads (BondNet.forward, net3d.py:151-166, with the bond list of BondWindows,
 * net3d.py:74-118, and the post-processing of infer_fast.py:160-177 / train3d.py:100-108).  Bonds are numbered as
 * the reference lists them: direction by direction, and within a direction the cells I of the nc^3 box whose
 * neighbour I + o is in the box, x fastest.  Window slot j of a bond in direction d reads voxel
 * (I + 2) s + voff[d, j] (j < nv) or cell I + 2 + coff[d, j - nv] (offsets as linear ids in the padded voxel grid of
 * side npd = (nc + 4) s, resp. the padded cell grid of side nc + 4).  inv (H, nslot): slot of the window read for
 * slot j under stabiliser element g, so that z_g = W1 P_g x.  w1f: W1 (hid, nslot * D) rearranged by
 * lpm_bond_w1_layout.  out (B, nb, 2) raw outputs; when cp (B, nc^3, 10) (per-cell 9 direction Kn and Tv) is given,
 * kn / tv (B, nc^3, 18) receive Kn = Kn_V exp(o1), Tv = Tv_V + Kn_V o2 in both directed slots of every bond, with
 * (Kn_V, Tv_V) the two-end means of cp (geom3d.py:214-220); other slots are left untouched. */
typedef struct {
  int32_t ndir;         /* directions of the group (3 axial or 6 face-diagonal) */
  int32_t nv, ncell;    /* voxel slots (2 s^3) and cell slots of the window */
  int32_t H;            /* stabiliser order: 16 (axial) or 8 (diagonal) */
  int32_t D, dstride;   /* feature channels and their stride in fv / fc (a multiple of 4, <= 16) */
  int32_t hid;          /* hidden width, <= 64 */
  int32_t slot[6];      /* forward slot k (even, 0..16) of every direction, in bond order */
  const int32_t* voff;  /* device (ndir, nv); every window must stay inside the padded voxel grid */
  const int32_t* coff;  /* device (ndir, ncell); every window must stay inside the padded cell grid */
  const int32_t* inv;   /* device (H, nv + ncell), entries in [0, nv + ncell): the inverse stabiliser perms */
} lpm_bond_group;
size_t lpm_bond_w1_size(const lpm_bond_group* grp);
int lpm_bond_w1_layout(const lpm_bond_group* grp, const float* w1, float* w1f, void* stream);
int lpm_bond_heads(const lpm_bond_group* grp, int32_t B, int32_t nc, int32_t s, const float* fv, const float* fc,
                   const float* w1f, const float* b1, const float* w2, const float* b2, const float* w3,
                   const float* b3, float* out, const double* cp, double* kn, double* tv, void* stream);

/* Same heads with the first layer in split BF16 (opt-in): every FP32 operand x = hi + lo (hi = bf16(x), lo =
 * bf16(x - hi)) and W1 x ~ Whi xhi + Whi xlo + Wlo xhi accumulated in FP32 on the BF16 matrix cores; D = dstride =
 * 12 (LPM_E_UNSUPPORTED otherwise).  w1h / w1l (lpm_bond_w1_split_size bytes each) from lpm_bond_w1_split_layout.
 * The reference computes this layer in FP32 (net3d.py:160-162); no counterpart for the split. */
size_t lpm_bond_w1_split_size(const lpm_bond_group* grp);
int lpm_bond_w1_split_layout(const lpm_bond_group* grp, const float* w1, void* w1h, void* w1l, void* stream);
int lpm_bond_heads_split(const lpm_bond_group* grp, int32_t B, int32_t nc, int32_t s, const float* fv,
                         const float* fc, const void* w1h, const void* w1l, const float* b1, const float* w2,
                         const float* b2, const float* w3, const float* b3, float* out, const double* cp, double* kn,
                         double* tv, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* LPM_H_ */
