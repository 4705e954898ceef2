// Declarations shared by the translation units of liblpm.so: error reporting, the K u launch geometry and the
// K u launchers (instantiated in ku_*.hip so that the tiled kernels compile in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/lpm.h"
#include "ku2d.hpp"
#include "ku3b.hpp"
#include "ku3d.hpp"

namespace lpmc {

// records the message for lpm_last_error() and returns code
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define HIPCHK(expr)                                                                               \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return lpmc::fail(LPM_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK()                                                                                 \
  do {                                                                                              \
    hipError_t e_ = hipGetLastError();                                                              \
    if (e_ != hipSuccess) return lpmc::fail(LPM_E_HIP, "kernel launch: %s", hipGetErrorString(e_)); \
  } while (0)

// K u launch geometry: grid (tiles x z chunks, nl, nb), block = tx * ty threads, zc planes per block; box = 1: the
// 3D box kernel (ku3b.hpp), tiles of kBoxX x kBoxY x kBoxZ sites, nzc = z tiles, zc = kBoxZ
struct KuGeom {
  dim3 grid, block;
  int ntx, nty, nzc, zc, tx, ty, box;
};

template <typename T, int MODE, int EPI, bool BONDS>
int launch_ku3(const lpm_grid* g, int nb, int nl, const KuGeom& kg, const lpm::Ku3Args<T, T>& a0, hipStream_t st);
// lean 3D kernel (ku3l.hpp): FP32, packed coefficients, no bond words, 64 x 4 tiles, epilogues PLAIN / MASKDOT;
// blocks resident per CU (the waves per SIMD it is compiled for)
constexpr int kLeanBlocksPerCU = 3;
template <int EPI>
int launch_ku3l(const lpm_grid* g, int nb, int nl, const KuGeom& kg, const lpm::Ku3Args<float, float>& a0,
                hipStream_t st);
template <typename T, int MODE, int EPI, bool BONDS>
int launch_ku3b(const lpm_grid* g, int nb, int nl, const KuGeom& kg, const lpm::Ku3Args<T, T>& a0, hipStream_t st);
template <typename T, int MODE, int EPI, bool BONDS>
int launch_ku2(const lpm_grid* g, int nb, int nl, const KuGeom& kg, const lpm::Ku2Args<T, T>& a0, hipStream_t st);

}  // namespace lpmc
