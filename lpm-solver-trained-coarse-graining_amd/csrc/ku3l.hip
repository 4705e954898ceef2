// Launcher of the lean 3D K u kernel (FP32, packed coefficients, full boxes): a translation unit of its own so the
// library builds in parallel.
#include "capi_common.hpp"
#include "ku3l.hpp"

namespace lpmc {

static_assert(kLeanBlocksPerCU == lpm::kLeanWavesPerEU, "one 4-wave block per wave slot of each SIMD");

template <int EPI>
int launch_ku3l(const lpm_grid* g, int nb, int nl, const KuGeom& kg, const lpm::Ku3Args<float, float>& a0,
                hipStream_t st) {
  lpm::Ku3Args<float, float> a = a0;
  a.nx = g->nx;
  a.ny = g->ny;
  a.nz = g->nz;
  a.nl = nl;
  a.zc = kg.zc;
  a.ntx = kg.ntx;
  a.nty = kg.nty;
  a.nzc = kg.nzc;
  const uint64_t N = (uint64_t)g->nx * g->ny * g->nz;
  if (kg.tx != 64 || kg.ty != 4 || kg.block.x != 256u || g->bonds || (uint64_t)a.ps * 9u * 4u >= (1ull << 32) ||
      N * 12u >= (1ull << 32))
    return fail(LPM_E_BADARG, "lean K u launch: unsupported geometry");
  hipLaunchKernelGGL((lpm::ku3l_kernel<4, EPI>), kg.grid, kg.block, 0, st, a);
  LAUNCHCHK();
  return LPM_OK;
}

template int launch_ku3l<lpm::EPI_PLAIN>(const lpm_grid*, int, int, const KuGeom&, const lpm::Ku3Args<float, float>&,
                                         hipStream_t);
template int launch_ku3l<lpm::EPI_MASKDOT>(const lpm_grid*, int, int, const KuGeom&,
                                           const lpm::Ku3Args<float, float>&, hipStream_t);

}  // namespace lpmc
