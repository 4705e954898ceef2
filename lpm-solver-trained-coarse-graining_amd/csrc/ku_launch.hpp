// Definitions of the K u launchers (one tile kernel per call, chosen from the launch geometry); included only by
// the ku_*.hip translation units, which instantiate them.
#pragma once
#include "capi_common.hpp"

namespace lpmc {

template <typename T, int MODE, int EPI, bool BONDS>
int launch_ku3(const lpm_grid* g, int nb, int nl, const KuGeom& kg, const lpm::Ku3Args<T, T>& a0, hipStream_t st) {
  lpm::Ku3Args<T, T> a = a0;
  a.nx = g->nx;
  a.ny = g->ny;
  a.nz = g->nz;
  a.nl = nl;
  a.zc = kg.zc;
  a.ntx = kg.ntx;
  a.nty = kg.nty;
  a.nzc = kg.nzc;
  // the tile of the kernel selected below must be the one the geometry was computed for
  const bool shape_ok = sizeof(T) == 4 ? ((kg.tx == 64 && (kg.ty == 4 || (kg.ty == 8 && !BONDS))) ||
                                           (kg.tx == 32 && kg.ty == 8))
                                        : ((kg.tx == 32 && kg.ty == 4) || (kg.tx == 16 && kg.ty == 8));
  if (!shape_ok || kg.block.x != (unsigned)(kg.tx * kg.ty))
    return fail(LPM_E_BADARG, "K u launch: no kernel for a %d x %d tile", kg.tx, kg.ty);
  if constexpr (sizeof(T) == 4) {
    bool done = false;
    if constexpr (!BONDS) {
      if (kg.tx == 64 && kg.ty == 8) {
        hipLaunchKernelGGL((lpm::ku3d_kernel<T, T, MODE, 64, 8, EPI>), kg.grid, kg.block, 0, st, a);
        done = true;
      }
    }
    if (done) {
    } else if (kg.tx == 64)
      hipLaunchKernelGGL((lpm::ku3d_kernel<T, T, MODE, 64, 4, EPI, BONDS>), kg.grid, kg.block, 0, st, a);
    else
      hipLaunchKernelGGL((lpm::ku3d_kernel<T, T, MODE, 32, 8, EPI, BONDS>), kg.grid, kg.block, 0, st, a);
  } else {
    if (kg.tx == 32)
      hipLaunchKernelGGL((lpm::ku3d_kernel<T, T, MODE, 32, 4, EPI, BONDS>), kg.grid, kg.block, 0, st, a);
    else
      hipLaunchKernelGGL((lpm::ku3d_kernel<T, T, MODE, 16, 8, EPI, BONDS>), kg.grid, kg.block, 0, st, a);
  }
  LAUNCHCHK();
  return LPM_OK;
}

template <typename T, int MODE, int EPI, bool BONDS>
int launch_ku3b(const lpm_grid* g, int nb, int nl, const KuGeom& kg, const lpm::Ku3Args<T, T>& a0, hipStream_t st) {
  lpm::Ku3Args<T, T> a = a0;
  a.nx = g->nx;
  a.ny = g->ny;
  a.nz = g->nz;
  a.nl = nl;
  a.zc = kg.zc;
  a.ntx = kg.ntx;
  a.nty = kg.nty;
  a.nzc = kg.nzc;
  constexpr int NT = lpm::kBoxX * lpm::kBoxY * lpm::kBoxZ;
  if (!kg.box || kg.block.x != (unsigned)NT || kg.tx != lpm::kBoxX || kg.ty != lpm::kBoxY || kg.zc != lpm::kBoxZ ||
      kg.grid.x != (unsigned)(kg.ntx * kg.nty * kg.nzc))
    return fail(LPM_E_BADARG, "K u launch: geometry is not the box kernel's");
  hipLaunchKernelGGL((lpm::ku3b_kernel<T, T, MODE, EPI, BONDS>), kg.grid, kg.block, 0, st, a);
  LAUNCHCHK();
  return LPM_OK;
}

template <typename T, int MODE, int EPI, bool BONDS>
int launch_ku2(const lpm_grid* g, int nb, int nl, const KuGeom& kg, const lpm::Ku2Args<T, T>& a0, hipStream_t st) {
  lpm::Ku2Args<T, T> a = a0;
  a.nx = g->nx;
  a.ny = g->ny;
  a.nl = nl;
  a.ntx = kg.ntx;
  a.nty = kg.nty;
  if constexpr (sizeof(T) == 4)
    hipLaunchKernelGGL((lpm::ku2d_kernel<T, T, MODE, 32, 8, EPI, BONDS>), kg.grid, kg.block, 0, st, a);
  else
    hipLaunchKernelGGL((lpm::ku2d_kernel<T, T, MODE, 32, 4, EPI, BONDS>), kg.grid, kg.block, 0, st, a);
  LAUNCHCHK();
  return LPM_OK;
}

}  // namespace lpmc

// explicit instantiation of the 3D launchers of one value type and bond mode, both layouts and all epilogues
#define LPM_INSTANTIATE_KU3(T, B)                                                                                  \
  template int lpmc::launch_ku3<T, lpm::COEF_DIRECTED, lpm::EPI_PLAIN, B>(const lpm_grid*, int, int,            \
                                                                            const lpmc::KuGeom&,                   \
                                                                            const lpm::Ku3Args<T, T>&, hipStream_t); \
  template int lpmc::launch_ku3<T, lpm::COEF_DIRECTED, lpm::EPI_MASKDOT, B>(const lpm_grid*, int, int,          \
                                                                              const lpmc::KuGeom&,                 \
                                                                              const lpm::Ku3Args<T, T>&,           \
                                                                              hipStream_t);                        \
  template int lpmc::launch_ku3<T, lpm::COEF_DIRECTED, lpm::EPI_RESID, B>(const lpm_grid*, int, int,            \
                                                                            const lpmc::KuGeom&,                   \
                                                                            const lpm::Ku3Args<T, T>&, hipStream_t); \
  template int lpmc::launch_ku3<T, lpm::COEF_PACKED, lpm::EPI_PLAIN, B>(const lpm_grid*, int, int,              \
                                                                          const lpmc::KuGeom&,                     \
                                                                          const lpm::Ku3Args<T, T>&, hipStream_t);   \
  template int lpmc::launch_ku3<T, lpm::COEF_PACKED, lpm::EPI_MASKDOT, B>(const lpm_grid*, int, int,            \
                                                                            const lpmc::KuGeom&,                   \
                                                                            const lpm::Ku3Args<T, T>&, hipStream_t); \
  template int lpmc::launch_ku3<T, lpm::COEF_PACKED, lpm::EPI_RESID, B>(const lpm_grid*, int, int,              \
                                                                          const lpmc::KuGeom&,                     \
                                                                          const lpm::Ku3Args<T, T>&, hipStream_t); \
  template int lpmc::launch_ku3b<T, lpm::COEF_DIRECTED, lpm::EPI_PLAIN, B>(const lpm_grid*, int, int,                \
                                                            const lpmc::KuGeom&,                                     \
                                                            const lpm::Ku3Args<T, T>&, hipStream_t);                 \
  template int lpmc::launch_ku3b<T, lpm::COEF_DIRECTED, lpm::EPI_MASKDOT, B>(const lpm_grid*, int, int,              \
                                                            const lpmc::KuGeom&,                                     \
                                                            const lpm::Ku3Args<T, T>&, hipStream_t);                 \
  template int lpmc::launch_ku3b<T, lpm::COEF_DIRECTED, lpm::EPI_RESID, B>(const lpm_grid*, int, int,                \
                                                            const lpmc::KuGeom&,                                     \
                                                            const lpm::Ku3Args<T, T>&, hipStream_t);                 \
  template int lpmc::launch_ku3b<T, lpm::COEF_PACKED, lpm::EPI_PLAIN, B>(const lpm_grid*, int, int,                  \
                                                            const lpmc::KuGeom&,                                     \
                                                            const lpm::Ku3Args<T, T>&, hipStream_t);                 \
  template int lpmc::launch_ku3b<T, lpm::COEF_PACKED, lpm::EPI_MASKDOT, B>(const lpm_grid*, int, int,                \
                                                            const lpmc::KuGeom&,                                     \
                                                            const lpm::Ku3Args<T, T>&, hipStream_t);                 \
  template int lpmc::launch_ku3b<T, lpm::COEF_PACKED, lpm::EPI_RESID, B>(const lpm_grid*, int, int,                  \
                                                            const lpmc::KuGeom&,                                     \
                                                            const lpm::Ku3Args<T, T>&, hipStream_t);

#define LPM_INSTANTIATE_KU2(T, B)                                                                                  \
  template int lpmc::launch_ku2<T, lpm::COEF_DIRECTED, lpm::EPI_PLAIN, B>(const lpm_grid*, int, int,            \
                                                                            const lpmc::KuGeom&,                   \
                                                                            const lpm::Ku2Args<T, T>&, hipStream_t); \
  template int lpmc::launch_ku2<T, lpm::COEF_DIRECTED, lpm::EPI_MASKDOT, B>(const lpm_grid*, int, int,          \
                                                                              const lpmc::KuGeom&,                 \
                                                                              const lpm::Ku2Args<T, T>&,           \
                                                                              hipStream_t);                        \
  template int lpmc::launch_ku2<T, lpm::COEF_DIRECTED, lpm::EPI_RESID, B>(const lpm_grid*, int, int,            \
                                                                            const lpmc::KuGeom&,                   \
                                                                            const lpm::Ku2Args<T, T>&, hipStream_t); \
  template int lpmc::launch_ku2<T, lpm::COEF_PACKED, lpm::EPI_PLAIN, B>(const lpm_grid*, int, int,              \
                                                                          const lpmc::KuGeom&,                     \
                                                                          const lpm::Ku2Args<T, T>&, hipStream_t);   \
  template int lpmc::launch_ku2<T, lpm::COEF_PACKED, lpm::EPI_MASKDOT, B>(const lpm_grid*, int, int,            \
                                                                            const lpmc::KuGeom&,                   \
                                                                            const lpm::Ku2Args<T, T>&, hipStream_t); \
  template int lpmc::launch_ku2<T, lpm::COEF_PACKED, lpm::EPI_RESID, B>(const lpm_grid*, int, int,              \
                                                                          const lpmc::KuGeom&,                     \
                                                                          const lpm::Ku2Args<T, T>&, hipStream_t);
