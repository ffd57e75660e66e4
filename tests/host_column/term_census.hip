// term_census.hip -- TEST HARNESS ONLY, compiled (never run) by
// tests/test_kernel_resources.py: one gfx950 kernel per term of the column
// code, each loading its inputs from memory, calling the term exactly as the
// forward kernels do, and storing every output, so that the compiler keeps the
// term's whole instruction stream.  The static VALU count of each kernel minus
// that of k_copy (the same loads and stores with no term) is that term's share
// of the forward kernel's instructions.
#include <hip/hip_runtime.h>

#include "../../jax-gcm_amd/csrc/speedy_column.hpp"

using namespace speedy;

constexpr int K = 8;
using CI = ColumnIn<float, K>;
using PA = PhaseA<float, K>;
using PB = PhaseB<float, K>;
struct Args {
  SpeedyParams p;
  SpeedyVertical v;
  Uniforms<float> u;
  uint32_t S, flags;
};

#define CENSUS(name, ...)                                                                             \
  __global__ void k_##name(Args a, const CI* ci, const PA* pa, const PB* pb, PA* oa, PB* ob) {        \
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;                                                 \
    CI c = ci[g];                                                                                      \
    PA A = pa[g];                                                                                      \
    PB B = pb[g];                                                                                      \
    __VA_ARGS__;                                                                                       \
    oa[g] = A;                                                                                         \
    ob[g] = B;                                                                                         \
  }

CENSUS(copy, A.cbmf += c.ps)
CENSUS(zonal, { float z[6]; zonal_column(a.u.zon, c.fsol, c.ozone, z); for (int i = 0; i < 6; ++i) B.tsw[i] = z[i]; })
CENSUS(humidity, humidity<float, K>(c, a.v, A))
CENSUS(convection, { ConvTape<float, K> t; convection<float, K>(c, a.p, a.v, a.u, A, t); })
CENSUS(condensation, condensation<float, K>(c, a.u, A))
CENSUS(clouds, { CloudTape<float, K> t; clouds<float, K>(c, a.p, a.u, A, t); })
CENSUS(albedos, albedos<float, K>(c, a.p, B))
CENSUS(lw_transmissivity, lw_transmissivity<float, K>(c, A, a.p, a.v, a.u, B))
CENSUS(lw_down, { LwdTape<float, K> t; lw_down<float, K>(c, a.p, a.u, B, t); })
CENSUS(vdiff, vdiff<float, K>(c, A, a.p, a.u, B))
CENSUS(shortwave, {
  float fh[4];
  SwTape<float, K> t;
  shortwave<float, K, false>(c, A, a.S, a.p, a.v, a.u, B, fh, t);
  B.flux_h[0] = fh[0];
  B.flux_h[1] = fh[1];
})
CENSUS(surface, {
  float tsfc;
  SfcTape<float> t;
  surface<float, K, false>(c, A, a.p, a.v, a.u, a.flags, B, tsfc, t);
  B.tsfc = tsfc;
  B.rlus2 = t.rlus2;
})
CENSUS(lw_up, {
  const float fh[4] = {B.flux_h[0], B.flux_h[1], 0.0f, 0.0f};
  LwuTape<float, K> t;
  lw_up<float, K, false>(c, A, a.p, a.v, a.u, B, B.rlus2, B.tsfc, fh, t);
})
