// speedy_column.hpp -- SPEEDY column physics for one atmospheric column, held in
// registers (gfx950).  Every routine is templated on the level count KX and
// fully unrolled over levels, so arrays indexed by level stay in VGPRs; data-
// dependent level indices (convective top, cloud top) become select chains.
//
// Each block follows one reference term expression by expression (file:line
// cited), in float32, in the reference's operation order where it matters for
// rounding.  All functions are __host__ __device__ so the same source can be
// compiled for the host in tests of the column routine itself.
#pragma once

#include <math.h>
#include <stdint.h>

#include "speedy_abi.h"

#if defined(__HIPCC__)
#define SPD_HD __host__ __device__ __forceinline__
#else
#define SPD_HD inline
#endif

namespace speedy {

// Math on the column's scalar type: float on the GPU; float or double in the
// host test builds of the same routines.
SPD_HD float mexp(float x) { return expf(x); }
SPD_HD double mexp(double x) { return exp(x); }
SPD_HD float msqrt(float x) { return sqrtf(x); }
SPD_HD double msqrt(double x) { return sqrt(x); }
SPD_HD float mmin(float a, float b) { return fminf(a, b); }
SPD_HD double mmin(double a, double b) { return fmin(a, b); }
SPD_HD float mmax(float a, float b) { return fmaxf(a, b); }
SPD_HD double mmax(double a, double b) { return fmax(a, b); }
SPD_HD double mmin(double a, float b) { return fmin(a, (double)b); }
SPD_HD double mmin(float a, double b) { return fmin((double)a, b); }
SPD_HD double mmax(double a, float b) { return fmax(a, (double)b); }
SPD_HD double mmax(float a, double b) { return fmax((double)a, b); }
SPD_HD float macos(float x) { return acosf(x); }
SPD_HD double macos(double x) { return acos(x); }
SPD_HD float msin(float x) { return sinf(x); }
SPD_HD double msin(double x) { return sin(x); }
SPD_HD float mcos(float x) { return cosf(x); }
SPD_HD double mcos(double x) { return cos(x); }
SPD_HD float masin(float x) { return asinf(x); }
SPD_HD double masin(double x) { return asin(x); }
// Square root where no discrete decision depends on the last bit: on the GPU the
// hardware instruction (v_sqrt_f32, 1 ulp) instead of the correctly rounded
// sequence; the same as msqrt on the host.
SPD_HD float msqrt_1ulp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}
SPD_HD double msqrt_1ulp(double x) { return sqrt(x); }

// Quotient a / b by the IEEE single-precision division sequence without its
// range-scaling and special-case steps: a hardware reciprocal, one Newton step
// on it, then two residual corrections of the quotient (Markstein's method,
// which is also the core of the compiler's own division).  For a zero or
// |a| >= 2^-103, b normal and a normal quotient (every division of the column
// code), the scaling steps are the identity, so the result is bit-identical to
// a / b, in 8 instructions instead of 11.  Only a zero, infinite or NaN
// divisor gives a NaN where a / b would not.
SPD_HD float mdiv(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  float r = __builtin_amdgcn_rcpf(b);
  r = __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
  float q = a * r;
  q = __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
  return __builtin_fmaf(__builtin_fmaf(-b, q, a), r, q);
#else
  return a / b;
#endif
}
SPD_HD double mdiv(double a, double b) { return a / b; }

// Derivatives of max(x, y) and min(x, y) with respect to x, with JAX's rule
// for ties (lax.max/lax.min JVP: half to each operand when x == y).
template <typename R>
SPD_HD R dmax_x(R x, R y) { return x > y ? R(1.0) : (x == y ? R(0.5) : R(0.0)); }
template <typename R>
SPD_HD R dmin_x(R x, R y) { return x < y ? R(1.0) : (x == y ? R(0.5) : R(0.0)); }

// jcm/constants.py:3-5, jcm/physics/speedy/physical_constants.py:8-45.  Python
// double expressions are evaluated in double and rounded once, as JAX does
// when a Python R meets a float32 array.
constexpr double P0 = 1e5;
constexpr double GRAV = 9.81;
constexpr double CP = 1004.0;
constexpr double ALHC = 2501.0;
constexpr double SBC = 5.67e-8;
constexpr double RGAS = 2.0 / 7.0 * 1004.0;
constexpr double RLHC = 1.0 / 2501.0;
constexpr double RCP = 1.0 / 1004.0;
constexpr double TFACT = 2501.0 / 1004.0;
constexpr double PRG = 1e5 / 9.81;
constexpr double RGAS288 = 2.0 / 7.0 * 1004.0 * 288.0;
constexpr double EPS = 1e-9;

template <typename R>
SPD_HD R sq(R x) { return x * x; }
template <typename R>
SPD_HD R pow3(R x) { return x * x * x; }
template <typename R>
SPD_HD R pow4(R x) { R x2 = x * x; return x2 * x2; }

// get_qsat (humidity.py:73-104) for sig > 0 (the sig <= 0 branch never occurs
// on the path: sigma levels and the surface value 1.0 are positive).
template <typename R>
SPD_HD R qsat_f(R ta, R ps, R sig) {
  const bool warm = ta >= R(273.16);
  const R c = warm ? R(17.269) : R(21.875);
  const R tt = warm ? R(35.86) : R(7.66);
  const R e = R(6.108e-3) * mexp(mdiv(c * (ta - R(273.16)), ta - tt));
  return mdiv(R(622.0) * e, sig * ps - R(0.378) * e);
}

// radset (longwave_radiation.py:204-225): (1-epslw) * band fractions.
template <typename R>
SPD_HD void radset4(R temp, R omeps, R fb[4]) {
  const R jt = mmin(mmax(temp, R(200.0)), R(320.0));
  const R f1 = R(0.148) - R(3.0e-6) * sq(jt - R(247.0));
  const R f2 = R(0.356) - R(5.2e-6) * sq(jt - R(282.0));
  const R f3 = R(0.314) + R(1.0e-5) * sq(jt - R(315.0));
  const R f0 = R(1.0) - ((f1 + f2) + f3);
  fb[0] = omeps * f0; fb[1] = omeps * f1; fb[2] = omeps * f2; fb[3] = omeps * f3;
}

// Latitude-dependent part of solar and get_zonal_average_fields
// (shortwave_radiation.py:267-287,446-450) for one column, from the date-only
// uniforms.  out: fsol, ozupp, ozone, zenit, stratz, sqrt(coa).
template <typename R>
struct ZonalUniforms {
  R sol_tdecl, sol_sdecl, sol_cdecl, sol_scale, oz_coz1, zen_c, zen_s;
};
template <typename R>
SPD_HD void zonal_column(const ZonalUniforms<R>& z, R sia, R coa, R out[6]) {
  // clip bounds +-(1-1e-9) are exactly +-1 in float32 (:446)
  const R ch0 = mmin(mmax(-z.sol_tdecl * sia / coa, -R(1.0)), R(1.0));
  const R h0 = macos(ch0);
  // sin(arccos(x)) = sqrt((1 - x)(1 + x)), exactly; (1 - x) and (1 + x) lose
  // nothing near |x| = 1, and a square root is far cheaper than a sine
  const R sh0 = msqrt_1ulp((R(1.0) - ch0) * (R(1.0) + ch0));
  const R topsr = z.sol_scale * (h0 * sia * z.sol_sdecl + sh0 * coa * z.sol_cdecl);
  const R flat2 = R(1.5) * sia * sia - R(0.5);
  const R ozone0 = R(0.4 * 0.020) * (R(1.0) + z.oz_coz1 * sia + R(1.8) * flat2);
  const R zenit = R(1.0) + R(1.0) * sq(R(1.0) - (coa * z.zen_c + sia * z.zen_s));
  out[0] = topsr;
  out[1] = topsr * R(0.5 * 0.020) * zenit;
  out[2] = topsr * ozone0 * zenit;
  out[3] = zenit;
  out[4] = mmax(R(6.0) - topsr, R(0.0));
  out[5] = msqrt_1ulp(coa);
}

// verify_state (physics_interface.py:305-319): where(q < 0, 0, q).
template <typename R>
SPD_HD R verify_q(R q) { return q < R(0.0) ? R(0.0) : q; }

// verify_tendencies (physics_interface.py:321-342) for one level, with the
// unclamped q: where(q + dt*dq >= 0, dq, -q/dt).  Sets `limited`.
template <typename R>
SPD_HD R verify_dq(R q, R dq, R dt, bool& limited) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const R qn = q + dt * dq;
  limited = !(qn >= R(0.0));
  return limited ? -q / dt : dq;
}

// Inputs of one column.
template <typename R, int KX>
struct ColumnIn {
  R T[KX], Q[KX], PHI[KX];
  R ps, u1, v1;                               // u, v at the lowest level only
  R alb0, sice, snowc_am, soilw, stl, sst;    // forcing
  R fmask, orog, phis0;                       // geometry
  R fsol, ozupp, ozone, zenit, stratz, sqcoa; // per-latitude table
};

// Everything phase A (humidity -> convection -> condensation -> clouds) hands on.
template <typename R, int KX>
struct PhaseA {
  R qsat[KX], rh[KX], se[KX];
  R tconv[KX], qconv[KX];     // convection tendencies
  R tlsc[KX], qlsc[KX];       // condensation tendencies
  int iptop_conv, iptop;          // convective top; after condensation (Q4)
  R cbmf, precnv, precls;
  R cloudc, clstr, qcloud, gse;
  int icltop;
};

// Uniform constants that depend only on parameters, levels and the date.  They
// are computed once per call on the host (make_uniforms) in float32, in the
// reference's operation order, and passed to the kernels by value.
template <typename R>
struct Uniforms {
  int32_t kx;
  uint32_t compute_sw;            // set_physics_flags: model_step % nstrad == 0 (speedy_physics.py:26-27)
  R dt_step;                      // DateData.dt_seconds: the time step of verify_tendencies
  R wvi1[SPEEDY_KX_MAX];      // geometry.wvi[:, 1]
  // convection (convection.py:136-142)
  R entr[SPEEDY_KX_MAX], fm0, rdps;
  // large-scale condensation (large_scale_condensation.py:48-63,77)
  R rtlsc, rhref[SPEEDY_KX_MAX], dqmax[SPEEDY_KX_MAX], pfact[SPEEDY_KX_MAX];
  R rhref_d[SPEEDY_KX_MAX];       // d rhref / d rhlsc (0, 1/2 at the tie, 1)
  // clouds (shortwave_radiation.py:348,387)
  R rrcl, rgse;
  // shortwave absorptivity of dry air + aerosol per level (shortwave_radiation.py:82,93)
  R sw_abs1[SPEEDY_KX_MAX];
  // longwave (speedy/forcing.py:33; shortwave_radiation.py:203; longwave_radiation.py:100,156)
  R ablco2, ablco2_d, eps1, omeps, refsfc, corlw;   // ablco2_d = d ablco2 / d co2_scale
  // surface fluxes (surface_flux.py:75-76,101,135-137,149,154,253,311)
  R esbc, ghum0, gtemp0, astab, rdth, rhdrag, cdl, cds, chlcp, chscp;
  // vertical diffusion (vertical_diffusion.py:45-63,99-118)
  R fshcq, fshcse, fvdiq2, drh0, fvdise;
  R rsig[SPEEDY_KX_MAX], rsig1[SPEEDY_KX_MAX], vd_fq[SPEEDY_KX_MAX], vd_drh0[SPEEDY_KX_MAX];
  uint32_t vd_on;                 // bit k: hsg[k+1] > 0.5
  // date-only parts of solar and get_zonal_average_fields (shortwave_radiation.py:250-262,421-443)
  ZonalUniforms<R> zon;
};

template <typename R>
inline void make_uniforms(const SpeedyParams& p, const SpeedyVertical& v, const SpeedyDate& d, Uniforms<R>& u) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const int kx = v.kx;
  u = Uniforms<R>();
  u.kx = kx;
  int32_t ms = d.model_step % 3;
  if (ms < 0) ms += 3;
  u.compute_sw = ms == 0;
  u.dt_step = d.dt_seconds;
  for (int k = 0; k < kx; ++k) u.wvi1[k] = v.wvi[k][1];
  R e[SPEEDY_KX_MAX] = {0};
  R sentr = R(0.0);
  for (int k = 1; k < kx - 1; ++k) {
    e[k] = sq(mmax(R(0.0), v.fsg[k] - R(0.5)));
    sentr = (k == 1) ? e[k] : sentr + e[k];
  }
  const R sc = p.entmax / sentr;
  for (int k = 0; k < kx; ++k) u.entr[k] = e[k] * sc;
  u.fm0 = R(P0) * v.dhs[kx - 1] / (R(GRAV) * p.trcnv * R(3600.0));
  u.rdps = R(2.0) / (R(1.0) - p.psmin);
  u.rtlsc = R(1.0) / (p.trlsc * R(3600.0));
  for (int k = 0; k < kx; ++k) {
    const R sig2 = sq(R(v.fsg[k]));
    const R rharg = p.rhlsc + p.drhlsc * (sig2 - R(1.0));
    u.rhref[k] = mmax(rharg, p.rhblsc);
    u.rhref_d[k] = dmax_x(rharg, R(p.rhblsc));
    u.dqmax[k] = R(10.0) * sig2 * u.rtlsc;
    u.pfact[k] = v.dhs[k] * R(PRG);
    u.sw_abs1[k] = p.absdry + p.absaer * sq(R(v.fsg[k]));
  }
  u.rrcl = R(1.0) / (R(p.rhcl2) - p.rhcl1);
  u.rgse = R(1.0) / (R(p.gse_s1) - p.gse_s0);
  R ab = R(6.0) * mexp((R)p.increase_co2 * R(0.005) * ((R)d.model_year + d.tyear - (R)p.co2_year_ref));
  u.ablco2_d = ab;
  if (p.co2_scale != R(1.0)) ab = ab * p.co2_scale;
  u.ablco2 = ab;
  u.eps1 = p.epslw / (R(v.dhs[0]) + v.dhs[1]);
  u.omeps = R(1.0) - p.epslw;
  u.refsfc = R(1.0) - p.emisfc;
  u.corlw = R(p.epslw) * p.emisfc;
  u.esbc = p.emisfc * R(SBC);
  u.ghum0 = R(1.0) - p.fhum0;
  u.gtemp0 = R(1.0) - p.ftemp0;
  u.astab = p.lscasym ? R(0.5) : R(1.0);
  u.rdth = R(p.fstab) / p.dtheta;
  u.rhdrag = R(1.0) / (R(GRAV) * p.hdrag);
  u.cdl = p.drag_scale != R(1.0) ? R(p.cdl) * p.drag_scale : R(p.cdl);
  u.cds = p.drag_scale != R(1.0) ? R(p.cds) * p.drag_scale : R(p.cds);
  u.chlcp = p.chl * R(CP);
  u.chscp = p.chs * R(CP);
  const int nl1 = kx - 1;
  const R cshc = v.dhs[kx - 1] / R(3600.0);
  const R cvdi = (R(v.hsg[nl1]) - v.hsg[1]) / R((nl1 - 1) * 3600.0);
  u.fshcq = cshc / p.trshc;
  u.fshcse = cshc / (p.trshc * R(CP));
  const R fvdiq = cvdi / p.trvdi;
  u.fvdise = cvdi / (p.trvds * R(CP));
  u.fvdiq2 = fvdiq * v.hsg[nl1];
  u.drh0 = p.rhgrad * (R(v.fsg[kx - 1]) - v.fsg[nl1 - 1]);
  {
    const R asin1 = masin(R(1.0));
    const R pigr = R(2.0) * asin1;
    const R alpha_s = R(2.0) * pigr * d.tyear;
    const R ca1 = mcos(alpha_s), sa1 = msin(alpha_s);
    const R ca2 = ca1 * ca1 - sa1 * sa1;
    const R sa2 = R(2.0) * sa1 * ca1;
    const R ca3 = ca1 * ca2 - sa1 * sa2;
    const R sa3 = sa1 * ca2 + sa2 * ca1;
    const R decl = R(0.006918) - R(0.399912) * ca1 + R(0.070257) * sa1 - R(0.006758) * ca2 + R(0.000907) * sa2 -
                       R(0.002697) * ca3 + R(0.001480) * sa3;
    const R fdis = R(1.000110) + R(0.034221) * ca1 + R(0.001280) * sa1 + R(0.000719) * ca2 + R(0.000077) * sa2;
    u.zon.sol_cdecl = mcos(decl);
    u.zon.sol_sdecl = msin(decl);
    u.zon.sol_tdecl = u.zon.sol_sdecl / u.zon.sol_cdecl;
    u.zon.sol_scale = ((R(4.0) * R(342.0)) / pigr) * fdis;
    const R alpha = R(4.0) * asin1 * (d.tyear + R(10.0 / 365.0));
    u.zon.oz_coz1 = mmax(R(0.0), mcos(alpha));
    const R rzen = -mcos(alpha) * R(23.45) * asin1 / R(90.0);
    u.zon.zen_c = mcos(rzen);
    u.zon.zen_s = msin(rzen);
  }
  u.vd_on = 0;
  for (int k = 0; k < kx; ++k) {
    u.rsig[k] = R(1.0) / v.dhs[k];
    u.rsig1[k] = (k < kx - 1) ? R(1.0) / (R(1.0) - v.hsg[k + 1]) : R(0.0);
    if (k + 1 < kx) {
      u.vd_fq[k] = fvdiq * v.hsg[k + 1];
      u.vd_drh0[k] = p.rhgrad * (R(v.fsg[k + 1]) - v.fsg[k]);
      if (v.hsg[k + 1] > R(0.5)) u.vd_on |= 1u << k;
    }
  }
}

// ---------------------------------------------------------------------------
// Terms.  Each term is a function that computes it from its inputs and records
// in a tape struct the intermediates its adjoint (speedy_adjoint.hpp) needs.  In
// the forward kernels the tapes are dead stores that the compiler removes; the
// adjoint recomputes a term right before reversing it instead of keeping the
// whole forward alive.
// ---------------------------------------------------------------------------

// humidity (humidity.py:40-46) and the dry static energy of convection.py:121.
template <typename R, int KX>
SPD_HD void humidity(const ColumnIn<R, KX>& c, const SpeedyVertical& v, PhaseA<R, KX>& a) {
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    a.qsat[k] = qsat_f(c.T[k], c.ps, R(v.fsg[k]));
    a.rh[k] = mdiv(c.Q[k], a.qsat[k]);
  }
#pragma unroll
  for (int k = 0; k < KX; ++k) a.se[k] = R(CP) * c.T[k] + c.PHI[k];
}

template <typename R, int KX>
struct ConvTape {
  int ktop2;
  bool case1, case2;
  R qthr0, mse0, msthr, qdif, qmax, qb, sb, fpsa, fmass;
  R sbh[KX], qbh[KX], en[KX], FM[KX], FUS[KX], FUQ[KX], FDS[KX], FDQ[KX], delq[KX];
  R fm_t, fus_t, fuq_t, fds_t, fdq_t, qsatb;
  R dfse[KX], dfqa[KX];
};

// convection (convection.py:15-237): diagnose_convection and the fluxes.
template <typename R, int KX>
SPD_HD void convection(const ColumnIn<R, KX>& c, const SpeedyParams& p, const SpeedyVertical& v,
                       const Uniforms<R>& u, PhaseA<R, KX>& a, ConvTape<R, KX>& t) {
  // diagnose_convection (:15-97)
  R mss[KX], mss2[KX];
#pragma unroll
  for (int k = 0; k < KX; ++k) mss[k] = a.se[k] + R(ALHC) * a.qsat[k];
#pragma unroll
  for (int k = 0; k < KX - 1; ++k) mss2[k] = mss[k] + u.wvi1[k] * (mss[k + 1] - mss[k]);
  mss2[KX - 1] = R(0.0);
  const R mse0 = a.se[KX - 1] + R(ALHC) * c.Q[KX - 1];
  const R mse1 = mmin(mse0, a.se[KX - 2] + R(ALHC) * c.Q[KX - 2]);
  const R mss0 = mmax(mse0, mss[KX - 1]);
  int ktop1 = KX, ktop2 = KX;
#pragma unroll
  for (int k = KX - 4; k >= 2; --k) {   // first unstable level from the top wins
    if (mss0 > mss2[k]) ktop1 = k + 1;
    if (mse1 > mss2[k]) ktop2 = k + 1;
  }
  R msthr = mss2[KX - 1];
#pragma unroll
  for (int k = 2; k <= KX - 4; ++k)
    if (ktop2 == k + 1) msthr = mss2[k];
  const R qthr0 = p.rhbl * a.qsat[KX - 1];
  const R qthr1 = p.rhbl * a.qsat[KX - 2];
  const bool lqthr = (c.Q[KX - 1] > qthr0) && (c.Q[KX - 2] > qthr1);
  const bool mask_psa = c.ps > p.psmin;
  const bool case1 = mask_psa && (ktop1 < KX) && (ktop2 < KX);
  const bool case2 = mask_psa && (ktop1 < KX) && !(ktop2 < KX) && lqthr;
  a.iptop_conv = (case1 || case2) ? ktop1 : KX + 1;
  R qdif = R(0.0);
  if (case1) qdif = mmax(c.Q[KX - 1] - qthr0, (mse0 - msthr) * R(RLHC));
  if (case2) qdif = c.Q[KX - 1] - qthr0;
  t.ktop2 = ktop2; t.case1 = case1; t.case2 = case2;
  t.qthr0 = qthr0; t.mse0 = mse0; t.msthr = msthr; t.qdif = qdif;
  // mass fluxes
  const int iptop = a.iptop_conv;
  const bool mask = iptop < KX;
  const R qmax = mmax(R(1.01) * c.Q[KX - 1], a.qsat[KX - 1]);
  R* sbh = t.sbh;
  R* qbh = t.qbh;
  sbh[0] = R(0.0); qbh[0] = R(0.0);
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    sbh[k] = a.se[k - 1] + u.wvi1[k - 1] * (a.se[k] - a.se[k - 1]);
    qbh[k] = c.Q[k - 1] + u.wvi1[k - 1] * (c.Q[k] - c.Q[k - 1]);
  }
  const R sb = sbh[KX - 1];
  const R qb = mmin(qbh[KX - 1], c.Q[KX - 1]);
  const R fpsa = c.ps * mmin(R(1.0), (c.ps - p.psmin) * u.rdps);
  const R fmass = u.fm0 * fpsa * mmin(R(5.0), mdiv(qdif, qmax - qb));
  t.qmax = qmax; t.qb = qb; t.sb = sb; t.fpsa = fpsa; t.fmass = fmass;
  a.cbmf = (mask ? R(1.0) : R(0.0)) * fmass;
  const R fus = fmass * a.se[KX - 1], fuq = fmass * qmax;
  const R fds = fmass * sb, fdq = fmass * qb;
  R* dfse = t.dfse;
  R* dfqa = t.dfqa;
  // entrainment and cumulative upward fluxes from the bottom (:178-191)
  R* en = t.en;
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    const bool lm = (k <= KX - 2) && (k >= iptop);
    en[k] = (lm && k >= 1) ? u.entr[k] * c.ps * a.cbmf : R(0.0);
  }
  R* FM = t.FM;
  R* FUS = t.FUS;
  R* FUQ = t.FUQ;
  {
    R r1 = R(0.0), r2 = R(0.0), r3 = R(0.0);
#pragma unroll
    for (int k = KX - 1; k >= 0; --k) {
      r1 = (k == KX - 1) ? en[k] : r1 + en[k];
      r2 = (k == KX - 1) ? en[k] * a.se[k] : r2 + en[k] * a.se[k];
      r3 = (k == KX - 1) ? en[k] * c.Q[k] : r3 + en[k] * c.Q[k];
      FM[k] = fmass + r1;
      FUS[k] = fus + r2;
      FUQ[k] = fuq + r3;
    }
  }
  R* FDS = t.FDS;
  R* FDQ = t.FDQ;
#pragma unroll
  for (int k = 0; k < KX - 1; ++k) { FDS[k] = FM[k] * sbh[k]; FDQ[k] = FM[k] * qbh[k]; }
  FDS[KX - 1] = fds; FDQ[KX - 1] = fdq;
  dfse[KX - 1] = fds - fus;
  dfqa[KX - 1] = fdq - fuq;
  R sumfsq = R(0.0);
#pragma unroll
  for (int k = 0; k < KX - 1; ++k) {
    const bool lm = (k >= iptop);
    dfse[k] = lm ? ((FUS[k + 1] - FDS[k + 1]) - (FUS[k] - FDS[k])) : R(0.0);
    dfqa[k] = lm ? ((FUQ[k + 1] - FDQ[k + 1]) - (FUQ[k] - FDQ[k])) : R(0.0);
    // secondary moisture flux (:198-202)
    const R delq = lm ? (p.rhil * a.qsat[k] - c.Q[k]) : R(0.0);
    t.delq[k] = delq;
    const R fsq = (delq > R(0.0)) ? p.smf * a.cbmf * delq : R(0.0);
    dfqa[k] = dfqa[k] + fsq;
    sumfsq = (k == 0) ? fsq : sumfsq + fsq;
  }
  dfqa[KX - 1] = dfqa[KX - 1] + (-sumfsq);
  // top layer (:204-219): gather at iptop (0 beyond KX-1), scatter at iptop-1 (Q6)
  R fm_t = R(0.0), fus_t = R(0.0), fuq_t = R(0.0), fds_t = R(0.0), fdq_t = R(0.0), qsatb = R(0.0);
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    if (iptop == k) { fm_t = FM[k]; fus_t = FUS[k]; fuq_t = FUQ[k]; fds_t = FDS[k]; fdq_t = FDQ[k]; }
  }
#pragma unroll
  for (int k = 0; k < KX - 1; ++k)
    if (iptop - 1 == k) qsatb = a.qsat[k] + u.wvi1[k] * (a.qsat[k + 1] - a.qsat[k]);
  t.fm_t = fm_t; t.fus_t = fus_t; t.fuq_t = fuq_t; t.fds_t = fds_t; t.fdq_t = fdq_t; t.qsatb = qsatb;
  a.precnv = mmax(fuq_t - fm_t * qsatb, R(0.0));
  const R top_s = fus_t - fds_t + R(ALHC) * a.precnv;
  const R top_q = fuq_t - fdq_t - a.precnv;
#pragma unroll
  for (int k = 0; k < KX; ++k)
    if (iptop - 1 == k) { dfse[k] = top_s; dfqa[k] = top_q; }
  // flux -> tendency (:226-227)
  const R rps = mdiv(R(1.0), c.ps);
  a.tconv[0] = dfse[0];
  a.qconv[0] = dfqa[0];
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    a.tconv[k] = dfse[k] * rps * v.grdscp[k];
    a.qconv[k] = dfqa[k] * rps * v.grdsig[k];
  }
}

// large-scale condensation (large_scale_condensation.py:14-86)
template <typename R, int KX>
SPD_HD void condensation(const ColumnIn<R, KX>& c, const Uniforms<R>& u, PhaseA<R, KX>& a) {
  const R psa2 = c.ps * c.ps;
  int first = 0;
  R precls = R(0.0);
  a.tlsc[0] = R(0.0);
  a.qlsc[0] = R(0.0);
#pragma unroll
  for (int k = KX - 1; k >= 1; --k) {
    const R dqa = u.rhref[k] * a.qsat[k] - c.Q[k];
    if (!(dqa >= R(0.0))) first = k;         // argmin(dqa[1:] >= 0) + 1 (Q4 when none)
    const bool neg = dqa < R(0.0);
    const R dq = neg ? dqa * u.rtlsc : R(0.0);
    a.qlsc[k] = dq;
    a.tlsc[k] = neg ? R(TFACT) * mmin(-dq, u.dqmax[k] * psa2) : R(0.0);
  }
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    const R t = u.pfact[k] * a.qlsc[k];
    precls = (k == 1) ? t : precls + t;
  }
  a.precls = (R(0.0) - precls) * c.ps;
  const int ipl = first == 0 ? 1 : first;
  a.iptop = ipl < a.iptop_conv ? ipl : a.iptop_conv;
}

template <typename R, int KX>
struct CloudTape {
  R cu[KX];
  uint32_t mk;                    // bit k: cu[k] is rh[k] - rhcl1 (else the level-(KX-2) value)
  bool m0;
  R cmax, pr1, ccraw, cl2, fstab_arg, fstab, clsx, clstr0;
};

// clouds (shortwave_radiation.py:313-404)
template <typename R, int KX>
SPD_HD void clouds(const ColumnIn<R, KX>& c, const SpeedyParams& p, const Uniforms<R>& u, PhaseA<R, KX>& a,
                   CloudTape<R, KX>& t) {
  a.gse = mdiv(a.se[KX - 2] - a.se[KX - 1], c.PHI[KX - 2] - c.PHI[KX - 1]);
  const R rrcl = u.rrcl;
  const bool m0 = a.rh[KX - 2] > p.rhcl1;
  const R cc0 = m0 ? a.rh[KX - 2] - p.rhcl1 : R(0.0);
  const int icl0 = m0 ? KX - 2 : KX;
  R* cu = t.cu;
  bool mk[KX];
  R cmax = R(0.0);
  t.mk = 0u;
#pragma unroll
  for (int k = 2; k <= KX - 3; ++k) {
    const R drh = a.rh[k] - p.rhcl1;
    mk[k] = (drh > cc0) && (c.Q[k] > p.qacl);
    cu[k] = mk[k] ? drh : cc0;
    cmax = (k == 2) ? cu[k] : mmax(cmax, cu[k]);
    if (mk[k]) t.mk |= 1u << k;
  }
  int icl = 0;
#pragma unroll
  for (int k = 2; k <= KX - 3; ++k) {        // Q5: float equality
    const int cand = (cmax == cu[k]) ? (mk[k] ? k : icl0) : icl0;
    icl = (k == 2) ? cand : (cand > icl ? cand : icl);
  }
  const R pr1 = mmin(p.pmaxcl, R(86.4) * (a.precnv + a.precls));
  const R cl2 = mmin(R(1.0), cmax * rrcl);
  const R ccraw = p.wpcl * msqrt_1ulp(mmax(R(EPS), pr1)) + sq(cl2);
  R cloudc = mmin(R(1.0), ccraw);
  if (cloudc != cloudc) cloudc = R(1.0);
  a.cloudc = cloudc;
  a.icltop = a.iptop < icl ? a.iptop : icl;
  a.qcloud = c.Q[KX - 2];
  const R rgse = u.rgse;
  const R fsa = rgse * (a.gse - p.gse_s0);
  const R fstab = mmin(mmax(fsa, R(0.0)), R(1.0));
  const R clsx = p.clsmax - R(1.2) * cloudc;
  R clstr = fstab * mmax(clsx, R(0.0));
  const R clstrl = mmax(clstr, p.clsminl) * a.rh[KX - 1];
  t.m0 = m0; t.cmax = cmax; t.pr1 = pr1; t.ccraw = ccraw; t.cl2 = cl2;
  t.fstab_arg = fsa; t.fstab = fstab; t.clsx = clsx; t.clstr0 = clstr;
  a.clstr = clstr + c.fmask * (clstrl - clstr);
}

template <typename R, int KX>
SPD_HD void phase_a(const ColumnIn<R, KX>& c, const SpeedyParams& p, const SpeedyVertical& v,
                    const Uniforms<R>& u, PhaseA<R, KX>& a) {
  humidity<R, KX>(c, v, a);
  ConvTape<R, KX> ct;
  convection<R, KX>(c, p, v, u, a, ct);
  condensation<R, KX>(c, u, a);
  CloudTape<R, KX> cl;
  clouds<R, KX>(c, p, u, a, cl);
}

// Cloud top only (for the level set S): the decisions of phase A without the
// convective fluxes.  Must agree exactly with phase_a's icltop.
template <typename R, int KX>
SPD_HD int cloud_top(const ColumnIn<R, KX>& c, const SpeedyParams& p, const SpeedyVertical& v,
                     const Uniforms<R>& u) {
  R qsat[KX], rh[KX], se[KX];
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    qsat[k] = qsat_f(c.T[k], c.ps, R(v.fsg[k]));
    rh[k] = mdiv(c.Q[k], qsat[k]);
    se[k] = R(CP) * c.T[k] + c.PHI[k];
  }
  int iptop_conv;
  {
    R mss[KX], mss2[KX];
#pragma unroll
    for (int k = 2; k < KX; ++k) mss[k] = se[k] + R(ALHC) * qsat[k];
#pragma unroll
    for (int k = 2; k <= KX - 4; ++k) mss2[k] = mss[k] + u.wvi1[k] * (mss[k + 1] - mss[k]);
    const R mse0 = se[KX - 1] + R(ALHC) * c.Q[KX - 1];
    const R mse1 = mmin(mse0, se[KX - 2] + R(ALHC) * c.Q[KX - 2]);
    const R mss0 = mmax(mse0, mss[KX - 1]);
    int ktop1 = KX, ktop2 = KX;
#pragma unroll
    for (int k = KX - 4; k >= 2; --k) {
      if (mss0 > mss2[k]) ktop1 = k + 1;
      if (mse1 > mss2[k]) ktop2 = k + 1;
    }
    const R qthr0 = p.rhbl * qsat[KX - 1];
    const R qthr1 = p.rhbl * qsat[KX - 2];
    const bool lqthr = (c.Q[KX - 1] > qthr0) && (c.Q[KX - 2] > qthr1);
    const bool mask_psa = c.ps > p.psmin;
    const bool case1 = mask_psa && (ktop1 < KX) && (ktop2 < KX);
    const bool case2 = mask_psa && (ktop1 < KX) && !(ktop2 < KX) && lqthr;
    iptop_conv = (case1 || case2) ? ktop1 : KX + 1;
  }
  int first = 0;
#pragma unroll
  for (int k = KX - 1; k >= 1; --k) {
    const R dqa = u.rhref[k] * qsat[k] - c.Q[k];
    if (!(dqa >= R(0.0))) first = k;
  }
  const int ipl = first == 0 ? 1 : first;
  const int iptop = ipl < iptop_conv ? ipl : iptop_conv;
  const bool m0 = rh[KX - 2] > p.rhcl1;
  const R cc0 = m0 ? rh[KX - 2] - p.rhcl1 : R(0.0);
  const int icl0 = m0 ? KX - 2 : KX;
  R cu[KX];
  bool mk[KX];
  R cmax = R(0.0);
#pragma unroll
  for (int k = 2; k <= KX - 3; ++k) {
    const R drh = rh[k] - p.rhcl1;
    mk[k] = (drh > cc0) && (c.Q[k] > p.qacl);
    cu[k] = mk[k] ? drh : cc0;
    cmax = (k == 2) ? cu[k] : mmax(cmax, cu[k]);
  }
  int icl = 0;
#pragma unroll
  for (int k = 2; k <= KX - 3; ++k) {
    const int cand = (cmax == cu[k]) ? (mk[k] ? k : icl0) : icl0;
    icl = (k == 2) ? cand : (cand > icl ? cand : icl);
  }
  return iptop < icl ? iptop : icl;
}

// ---------------------------------------------------------------------------
// Phase B.  b1 needs only phase A and the column (albedos, longwave
// transmissivity, longwave down, vertical diffusion); b2 needs the level set S
// and the per-latitude tables (shortwave, surface fluxes, longwave up).
// ---------------------------------------------------------------------------
template <typename R, int KX>
struct PhaseB {
  // b1
  R albsfc, alb_l, alb_s, snowc;
  R tauL[KX][4], st4a0[KX], st4a1[KX], dfl[KX], rlds;
  R tvd[KX], qvd[KX];
  // b2
  R tsw[KX], tlw[KX];
  R du1, dv1, dts1, dqs1;        // surface-flux tendencies at level KX-1
  R evap2;                       // grid-box surface evaporation (budgets)
  R flux_h[2], rlus2;            // shortwave -> longwave and surface -> longwave hand-offs (and tsfc)
  // diagnostics (b2)
  R dfabs_sw[KX], rsds, rsns, ftop_sw, stratc0, stratc1;
  R rlns, ftop_lw, flux_out[4];
  R ustr[3], vstr[3], shf[3], evap[3], rlus[3], hfluxn[2], tsfc, tskin, u0, v0, t0;
};

// Transmissivities use the hardware exponential (v_exp_f32 on x*log2(e)):
// relative error ~6e-8*|x| (< 2e-6 where the result matters), far inside the
// 1e-5 parity budget; decision-bearing exponentials (qsat) stay libm-accurate.
template <typename R>
SPD_HD R exp_tr(R x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(R) == 4) return __expf(x);
#endif
  return mexp(x);
}

// set_forcing albedos (speedy/forcing.py:26-30)
template <typename R, int KX>
SPD_HD void albedos(const ColumnIn<R, KX>& c, const SpeedyParams& p, PhaseB<R, KX>& b) {
  b.snowc = mmin(R(1.0), c.snowc_am);
  b.alb_l = c.alb0 + b.snowc * (p.albsn - c.alb0);
  b.alb_s = p.albsea + c.sice * (R(p.albice) - p.albsea);
  b.albsfc = b.alb_s + c.fmask * (b.alb_l - b.alb_s);
}

// Longwave absorptivities of level k per band (shortwave_radiation.py:177-200).
template <typename R, int KX>
SPD_HD void lw_absorptivity(int k, const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p,
                            const Uniforms<R>& u, R ab[4]) {
  ab[0] = p.ablwin; ab[1] = u.ablco2; ab[2] = p.ablwv1 * c.Q[k]; ab[3] = p.ablwv2 * c.Q[k];
  if (k == 0) { ab[2] = R(0.0); ab[3] = R(0.0); }
  if (k >= 2 && k <= KX - 2) {
    ab[0] = ab[0] + ((k > a.icltop) ? a.cloudc * p.ablcl1 : a.cloudc * p.ablcl2);
    const R ac2 = a.cloudc * p.ablcl2;
    ab[2] = mmax(ab[2], ac2);
    ab[3] = mmax(ab[3], ac2);
  }
}

template <typename R, int KX>
SPD_HD void lw_transmissivity(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p,
                              const SpeedyVertical& v, const Uniforms<R>& u, PhaseB<R, KX>& b) {
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    R ab[4];
    lw_absorptivity<R, KX>(k, c, a, p, u, ab);
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) b.tauL[k][bnd] = exp_tr(-ab[bnd] * c.ps * v.dhs[k]);
  }
}

// One level's longwave quantities, recomputed from the column alone: the band
// transmissivities (shortwave_radiation.py:177-200), the emissions st4a[k][0:2]
// (longwave_radiation.py:48-68) and the radset band fractions of T[k]
// (longwave_radiation.py:216-225).  The same expressions as lw_transmissivity
// and lw_down, so the values are equal to theirs; the adjoint recomputes them per
// level instead of keeping the forward's [KX][4] arrays live.
template <typename R>
struct LwLevel {
  R tau[4], st0, st1, fb[4];
};

// TAU = false: the emissions and band fractions only (the caller has tau).
template <typename R, int KX, bool TAU = true>
SPD_HD void lw_level(int k, const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p,
                     const SpeedyVertical& v, const Uniforms<R>& u, LwLevel<R>& L) {
  if constexpr (TAU) {
    R ab[4];
    lw_absorptivity<R, KX>(k, c, a, p, u, ab);
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) L.tau[bnd] = exp_tr(-ab[bnd] * c.ps * v.dhs[k]);
  }
  auto Th = [&](int j) { return c.T[j] + u.wvi1[j] * (c.T[j + 1] - c.T[j]); };
  if (k == 0) {
    const R s0 = R(0.75) * c.T[0] + R(0.25) * Th(0);
    L.st0 = R(SBC) * pow4(s0);
    L.st1 = R(0.0);
  } else if (k == 1) {
    const R s1 = R(0.50) * c.T[1] + R(0.25) * (Th(0) + Th(1));
    L.st0 = R(SBC) * pow4(s1);
    L.st1 = R(0.0);
  } else {
    const R grad = (k < KX - 1) ? R(0.5) * mmax(Th(k) - Th(k - 1), R(0.0)) : mmax(c.T[KX - 1] - Th(KX - 2), R(0.0));
    const R st3a = R(SBC) * pow3(c.T[k]);
    L.st0 = st3a * c.T[k];
    L.st1 = R(4.0) * st3a * grad;
  }
  radset4(c.T[k], u.omeps, L.fb);
}

template <typename R, int KX>
struct LwdTape {
  R Th[KX], grad[KX], s0, s1;
  R Fd[KX][4];                    // band fluxes leaving level k downward
};

// longwave down (longwave_radiation.py:13-117)
// Half-level temperatures and emissions st4a (longwave_radiation.py:48-68).
template <typename R, int KX>
SPD_HD void lw_emissions(const ColumnIn<R, KX>& c, const Uniforms<R>& u, PhaseB<R, KX>& b, LwdTape<R, KX>& t) {
  R* Th = t.Th;
#pragma unroll
  for (int k = 0; k < KX - 1; ++k) Th[k] = c.T[k] + u.wvi1[k] * (c.T[k + 1] - c.T[k]);
  const R s0 = R(0.75) * c.T[0] + R(0.25) * Th[0];
  const R s1 = R(0.50) * c.T[1] + R(0.25) * (Th[0] + Th[1]);
  t.s0 = s0; t.s1 = s1;
  R* grad = t.grad;
#pragma unroll
  for (int k = 2; k < KX - 1; ++k) grad[k] = R(0.5) * mmax(Th[k] - Th[k - 1], R(0.0));
  grad[KX - 1] = mmax(c.T[KX - 1] - Th[KX - 2], R(0.0));
  b.st4a0[0] = R(SBC) * pow4(s0); b.st4a0[1] = R(SBC) * pow4(s1);
  b.st4a1[0] = R(0.0); b.st4a1[1] = R(0.0);
#pragma unroll
  for (int k = 2; k < KX; ++k) {
    const R st3a = R(SBC) * pow3(c.T[k]);
    b.st4a0[k] = st3a * c.T[k];
    b.st4a1[k] = R(4.0) * st3a * grad[k];
  }
}

template <typename R, int KX>
SPD_HD void lw_down(const ColumnIn<R, KX>& c, const SpeedyParams& p, const Uniforms<R>& u, PhaseB<R, KX>& b,
                    LwdTape<R, KX>& t) {
  lw_emissions<R, KX>(c, u, b, t);
  R Fd[4];
  {
    R fb[4];
    radset4(c.T[0], u.omeps, fb);
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) {
      const R em = R(1.0) - b.tauL[0][bnd];
      Fd[bnd] = bnd < 2 ? em * (fb[bnd] * (b.st4a0[0] + em * b.st4a1[0])) : R(0.0);
      t.Fd[0][bnd] = Fd[bnd];
    }
  }
  R Fprev = ((Fd[0] + Fd[1]) + Fd[2]) + Fd[3];
  b.dfl[0] = -Fprev;
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    R fb[4];
    radset4(c.T[k], u.omeps, fb);
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) {
      const R em = R(1.0) - b.tauL[k][bnd];
      Fd[bnd] = b.tauL[k][bnd] * Fd[bnd] + em * (fb[bnd] * (b.st4a0[k] + em * b.st4a1[k]));
      t.Fd[k][bnd] = Fd[bnd];
    }
    const R Fs = ((Fd[0] + Fd[1]) + Fd[2]) + Fd[3];
    b.dfl[k] = -(Fs - Fprev);
    Fprev = Fs;
  }
  R rlds = ((p.emisfc * Fd[0] + p.emisfc * Fd[1]) + p.emisfc * Fd[2]) + p.emisfc * Fd[3];
  const R corlw = u.corlw * b.st4a0[KX - 1];
  b.dfl[KX - 1] = b.dfl[KX - 1] + (-corlw);
  b.rlds = rlds + corlw;
}

// vertical diffusion (vertical_diffusion.py:10-135)
template <typename R, int KX>
SPD_HD void vdiff(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p, const Uniforms<R>& u,
                  PhaseB<R, KX>& b) {
  const int icnv = KX - a.iptop - 1;
  constexpr int NL1 = KX - 1;
  R tt[KX], qt[KX];
#pragma unroll
  for (int k = 0; k < KX; ++k) { tt[k] = R(0.0); qt[k] = R(0.0); }
  const R dmse = a.se[KX - 1] - a.se[NL1 - 1] + R(ALHC) * (c.Q[KX - 1] - a.qsat[NL1 - 1]);
  const R drh = a.rh[KX - 1] - a.rh[NL1 - 1];
  const R fcnv = (icnv > 0 && dmse >= R(0.0)) ? p.redshc : R(1.0);
  const R rsA = u.rsig[NL1 - 1], rsB = u.rsig[KX - 1];
  if (dmse >= R(0.0)) {
    const R fluxse = fcnv * u.fshcse * dmse;
    tt[NL1 - 1] = fluxse * rsA;
    tt[KX - 1] = -fluxse * rsB;
    if (drh >= R(0.0)) {
      const R fluxq = fcnv * u.fshcq * a.qsat[KX - 1] * drh;
      qt[NL1 - 1] = fluxq * rsA;
      qt[KX - 1] = -fluxq * rsB;
    }
  } else if (dmse < R(0.0) && drh > u.drh0) {
    const R fluxq = u.fvdiq2 * a.qsat[NL1 - 1] * drh;
    qt[NL1 - 1] = fluxq * rsA;
    qt[KX - 1] = -fluxq * rsB;
  }
  // moisture diffusion above the PBL (:99-118)
  R fq[KX];
#pragma unroll
  for (int k = 2; k < KX - 2; ++k) {
    const R d = a.rh[k + 1] - a.rh[k];
    fq[k] = (d >= u.vd_drh0[k] && ((u.vd_on >> k) & 1u)) ? u.vd_fq[k] * a.qsat[k] * d : R(0.0);
  }
#pragma unroll
  for (int k = 2; k < KX - 2; ++k) qt[k] = qt[k] + fq[k] * u.rsig[k];
#pragma unroll
  for (int k = 2; k < KX - 2; ++k) qt[k + 1] = qt[k + 1] + (-fq[k] * u.rsig[k + 1]);
  // super-adiabatic damping (:120-131): all local adds, then the cumulative sum
  R fse[KX];
#pragma unroll
  for (int k = 0; k < NL1; ++k) {
    const R se0 = a.se[k + 1] - p.segrad * (c.PHI[k + 1] - c.PHI[k]);
    fse[k] = (a.se[k] < se0) ? u.fvdise * (se0 - a.se[k]) : R(0.0);
    tt[k] = tt[k] + fse[k] * u.rsig[k];
  }
  R cum = R(0.0);
#pragma unroll
  for (int k = 0; k < NL1; ++k) {
    cum = (k == 0) ? fse[k] * u.rsig1[k] : cum + fse[k] * u.rsig1[k];
    tt[k + 1] = tt[k + 1] + (-cum);
  }
#pragma unroll
  for (int k = 0; k < KX; ++k) { b.tvd[k] = tt[k]; b.qvd[k] = qt[k]; }
}

// Where the longwave transmissivities wait between longwave down (phase b1) and
// longwave up (phase b2): nowhere by default (longwave up recomputes them per
// level); the forward kernels pass a per-column LDS store (put / get / after).
struct NoTauStore {
  static constexpr bool kOn = false;
  template <typename T>
  SPD_HD void put(int, T) const {}
  SPD_HD float get(int) const { return 0.0f; }
  template <typename T>
  SPD_HD NoTauStore after(T) const { return *this; }
};

template <typename R, int KX, class TauStore = NoTauStore>
SPD_HD void phase_b1(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p,
                     const SpeedyVertical& v, const Uniforms<R>& u, PhaseB<R, KX>& b, TauStore ts = TauStore()) {
  albedos<R, KX>(c, p, b);
  lw_transmissivity<R, KX>(c, a, p, v, u, b);
  if constexpr (TauStore::kOn) {
#pragma unroll
    for (int k = 0; k < KX; ++k)
#pragma unroll
      for (int bnd = 0; bnd < 4; ++bnd) ts.put(k * 4 + bnd, b.tauL[k][bnd]);
  }
  LwdTape<R, KX> ld;
  lw_down<R, KX>(c, p, u, b, ld);
  vdiff<R, KX>(c, a, p, u, b);
}

template <typename R, int KX>
struct SwTape {
  R a_cl, psaz, acloud;
  R t22[KX], tau0[KX], tau1[KX], f1[KX], f2[KX], Rf[KX], fu[KX], dfabs[KX];
};

// shortwave (shortwave_radiation.py:27-217); flux_h is the hand-off to the
// longwave (Q2).
template <typename R, int KX, bool DIAG>
SPD_HD void shortwave(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, uint32_t S, const SpeedyParams& p,
                      const SpeedyVertical& v, const Uniforms<R>& u, PhaseB<R, KX>& b, R flux_h[4],
                      SwTape<R, KX>& t) {
  const R ps = c.ps;
  const R rps = mdiv(R(1.0), ps);
  const R a_cl = (a.icltop < KX) ? p.albcl * a.cloudc : R(0.0);
  R* t22 = t.t22;
#pragma unroll
  for (int k = 0; k < KX - 1; ++k) t22[k] = ((S >> k) & 1u) ? a_cl : R(0.0);   // Q1
  t22[KX - 1] = p.albcls * a.clstr;
  const R psaz = ps * c.zenit;
  const R acloud = a.cloudc * mmin(p.abscl1 * a.qcloud, p.abscl2);
  t.a_cl = a_cl; t.psaz = psaz; t.acloud = acloud;
  R* tau0 = t.tau0;
  R* tau1 = t.tau1;
  tau0[0] = exp_tr(-psaz * v.dhs[0] * p.absdry);
#pragma unroll
  for (int k = 1; k < KX - 1; ++k) {
    const R cl = (k >= a.icltop) ? acloud : R(0.0);
    tau0[k] = exp_tr(-psaz * v.dhs[k] * (u.sw_abs1[k] + p.abswv1 * c.Q[k] + cl));
  }
  tau0[KX - 1] = exp_tr(-psaz * v.dhs[KX - 1] * (u.sw_abs1[KX - 1] + p.abswv1 * c.Q[KX - 1]));
  tau1[0] = R(1.0);
#pragma unroll
  for (int k = 1; k < KX; ++k) tau1[k] = exp_tr(-psaz * v.dhs[k] * p.abswv2 * c.Q[k]);
  // downward
  R* dfabs = t.dfabs;
  R* f1 = t.f1;
  R* f2 = t.f2;
  R* Rf = t.Rf;
  const R F0 = c.fsol * R(0.95);
  dfabs[0] = F0;
  f1[0] = tau0[0] * (F0 - c.ozupp * ps);
  dfabs[0] = dfabs[0] + (-f1[0]);
  dfabs[1] = f1[0];
  f1[1] = tau0[1] * (f1[0] - c.ozone * ps);
  dfabs[1] = dfabs[1] + (-f1[1]);
#pragma unroll
  for (int k = 2; k < KX; ++k) f1[k] = f1[k - 1] * tau0[k] * (R(1.0) - t22[k]);
#pragma unroll
  for (int k = 2; k < KX; ++k) {
    dfabs[k] = f1[k - 1] * (R(1.0) - t22[k]) * (R(1.0) - tau0[k]);
    Rf[k] = t22[k] * f1[k - 1];
  }
  Rf[0] = t22[0];
  Rf[1] = t22[1];
  f2[0] = c.fsol * R(0.05);
  R carry = f2[0];
#pragma unroll
  for (int k = 1; k < KX; ++k) { carry = carry * tau1[k]; f2[k] = carry; }
#pragma unroll
  for (int k = 1; k < KX; ++k) dfabs[k] = dfabs[k] + f2[k - 1] * (R(1.0) - tau1[k]);
  // upward
  const R rsds = f1[KX - 1] + f2[KX - 1];
  R U = f1[KX - 1] * b.albsfc;
  const R rsns = rsds - U;
  R* fu = t.fu;
  fu[KX - 1] = U;
#pragma unroll
  for (int k = KX - 1; k >= 1; --k) { U = U * tau0[k] + Rf[k]; fu[k - 1] = U; }
#pragma unroll
  for (int k = 0; k < KX; ++k) dfabs[k] = dfabs[k] + fu[k] * (R(1.0) - tau0[k]);
  const R ftoa = tau0[0] * fu[0] + Rf[0];
  const bool csw = u.compute_sw != 0u;
#pragma unroll
  for (int k = 0; k < KX; ++k) b.tsw[k] = csw ? dfabs[k] * v.grdscp[k] * rps : R(0.0);
  flux_h[0] = ftoa; flux_h[1] = f2[KX - 1]; flux_h[2] = R(0.0); flux_h[3] = R(0.0);   // Q2
  b.rsds = rsds;
  if (DIAG) {
#pragma unroll
    for (int k = 0; k < KX; ++k) b.dfabs_sw[k] = dfabs[k];
    b.rsns = rsns;
    b.ftop_sw = c.fsol - ftoa;
  }
}

template <typename R>
struct SfcTape {
  R u0, v0, Tb, Qb, phi0, dt1, t1_0a, t1_1a, t2_0, t2_1, t1_0, t1_1, t0, wind, den0, tskin0, dd, dthl, den1;
  R forog, cdldv, shf0a, q1_0, qsat00, evx, evap0a, tsk3, drls, rlus0a, hfl0a, clamb, qs1p, qs1, dtskin;
  R ds, dths, den2, q1_1, qsat01;
  R ustr0, vstr0, shf0, evap0, rlus0, ustr1, vstr1, shf1, evap1, rlus1;
  R ustr2, vstr2, shf2, evap2, rlus2;
  bool land, warm;
};

// surface fluxes (surface_flux.py:15-300).  Returns rlus2 and tsfc through t.
template <typename R, int KX, bool DIAG>
SPD_HD void surface(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p, const SpeedyVertical& v,
                    const Uniforms<R>& u, uint32_t flags, PhaseB<R, KX>& b, R& tsfc, SfcTape<R>& t) {
  const SpeedyParams& s = p;
  const R ps = c.ps;
  const R rps = mdiv(R(1.0), ps);
  const R rlds = b.rlds;
  const R u1 = c.u1, v1 = c.v1, Tb = c.T[KX - 1], Qb = c.Q[KX - 1];
  const R u0 = s.fwind0 * u1, v0 = s.fwind0 * v1;
  t.u0 = u0; t.v0 = v0; t.Tb = Tb; t.Qb = Qb;
  R t1_0 = R(0.0), t1_1 = R(0.0), t2_1 = R(0.0), q1_1 = R(0.0);
  R den0 = R(0.0), den2 = R(0.0);
  R ustr0 = R(0.0), vstr0 = R(0.0), shf0 = R(0.0), evap0 = R(0.0), rlus0 = R(0.0), hfl0 = R(0.0), tskin = R(0.0);
  R ustr1 = R(0.0), vstr1 = R(0.0);
  const bool land = flags & SPEEDY_FLAG_LFLUXLAND;
  t.land = land;
  if (land) {
    const R phi0 = c.orog * R(GRAV);
    const R dt1 = v.wvi[KX - 1][1] * (Tb - c.T[KX - 2]);
    t1_0 = Tb + dt1;
    t1_1 = t1_0 - mdiv(phi0 * dt1, R(RGAS288) * v.sigl[KX - 1]);
    t2_1 = Tb + R(RCP) * c.PHI[KX - 1];
    const R t2_0 = t2_1 - R(RCP) * phi0;
    t.phi0 = phi0; t.dt1 = dt1; t.t1_0a = t1_0; t.t1_1a = t1_1; t.t2_0 = t2_0; t.t2_1 = t2_1;
    t.warm = Tb > c.T[KX - 2];
    if (Tb > c.T[KX - 2]) {
      t1_0 = s.ftemp0 * t1_0 + u.gtemp0 * t2_0;
      t1_1 = s.ftemp0 * t1_1 + u.gtemp0 * t2_1;
    } else {
      t1_0 = Tb;
      t1_1 = Tb;
    }
    const R t0 = t1_1 + c.fmask * (t1_0 - t1_1);
    const R wind = msqrt_1ulp(u0 * u0 + v0 * v0 + R(s.vgust) * s.vgust);
    den0 = mdiv(R(P0) * ps, R(RGAS) * t0) * wind;
    tskin = c.stl + s.ctday * c.sqcoa * b.rsds * (R(1.0) - b.alb_l) * ps;
    t.t1_0 = t1_0; t.t1_1 = t1_1; t.t0 = t0; t.wind = wind; t.den0 = den0; t.tskin0 = tskin;
    const R dd = tskin - t2_0;
    const R dthl = (tskin > t2_0) ? mmin(R(s.dtheta), dd) : mmax(-R(s.dtheta), u.astab * dd);
    const R den1 = den0 * (R(1.0) + dthl * u.rdth);
    const R forog = R(1.0) + u.rhdrag * (R(1.0) - exp_tr(-mmax(c.phis0, R(0.0)) * u.rhdrag));
    const R cdldv = u.cdl * den0 * forog;
    t.dd = dd; t.dthl = dthl; t.den1 = den1; t.forog = forog; t.cdldv = cdldv;
    ustr0 = -cdldv * u1;
    vstr0 = -cdldv * v1;
    shf0 = u.chlcp * den1 * (tskin - t1_0);
    t.shf0a = shf0;
    R q1_0 = Qb;
    if (s.fhum0 > R(0.0)) q1_0 = s.fhum0 * (a.rh[KX - 1] * qsat_f(t1_0, ps, R(1.0))) + u.ghum0 * Qb;
    const R qsat00 = qsat_f(tskin, ps, R(1.0));
    const R evx = c.soilw * qsat00 - q1_0;
    evap0 = s.chl * den1 * mmax(R(0.0), evx);
    const R tsk3 = pow3(tskin);
    const R drls = R(4.0) * u.esbc * tsk3;
    rlus0 = u.esbc * tsk3 * tskin;
    hfl0 = b.rsds * (R(1.0) - b.alb_l) + rlds - (rlus0 + shf0 + (R(ALHC) * evap0));
    t.q1_0 = q1_0; t.qsat00 = qsat00; t.evx = evx; t.evap0a = evap0; t.tsk3 = tsk3; t.drls = drls;
    t.rlus0a = rlus0; t.hfl0a = hfl0;
    if (s.lskineb) {
      const R clamb = s.clambda + (b.snowc * (R(s.clambsn) - s.clambda));
      hfl0 = hfl0 - (clamb * (tskin - c.stl));
      const R qs1p = qsat_f(tskin + R(1.0), ps, R(1.0));
      const R qs1 = (evap0 > R(0.0)) ? c.soilw * (qs1p - qsat00) : R(0.0);
      const R dtskin = mdiv(hfl0, clamb + drls + (s.chl * den1 * (R(CP) + (R(ALHC) * qs1))));
      t.clamb = clamb; t.qs1p = qs1p; t.qs1 = qs1; t.dtskin = dtskin;
      tskin = tskin + dtskin;
      shf0 = shf0 + u.chlcp * den1 * dtskin;
      evap0 = evap0 + s.chl * den1 * qs1 * dtskin;
      rlus0 = rlus0 + drls * dtskin;
      hfl0 = clamb * (tskin - c.stl);
    }
    const R ds = c.sst - t2_1;
    const R dths = (c.sst > t2_1) ? mmin(R(s.dtheta), ds) : mmax(-R(s.dtheta), u.astab * ds);
    den2 = den0 * (R(1.0) + dths * u.rdth);
    q1_1 = Qb;
    if (s.fhum0 > R(0.0)) q1_1 = s.fhum0 * (a.rh[KX - 1] * qsat_f(t1_1, ps, R(1.0))) + u.ghum0 * Qb;
    const R cdsdv = u.cds * den2;
    ustr1 = -cdsdv * u1;
    vstr1 = -cdsdv * v1;
    t.ds = ds; t.dths = dths; t.den2 = den2; t.q1_1 = q1_1;
  }
  // sea (:250-261)
  const R shf1 = u.chscp * den2 * (c.sst - t1_1);
  const R qsat01 = qsat_f(c.sst, ps, R(1.0));
  const R evap1 = s.chs * den2 * (qsat01 - q1_1);
  const R rlus1 = u.esbc * pow4(c.sst);
  const R hfl1 = b.rsds * (R(1.0) - b.alb_s) + rlds - rlus1 + shf1 + R(ALHC) * evap1;
  t.qsat01 = qsat01;
  R ustr2 = R(0.0), vstr2 = R(0.0), shf2 = R(0.0), evap2 = R(0.0), t0o = R(0.0), rlus2 = R(0.0);
  tsfc = R(0.0);
  R tskin_o = R(0.0);
  if (land) {
    const R fm = c.fmask;
    ustr2 = ustr1 + fm * (ustr0 - ustr1);
    vstr2 = vstr1 + fm * (vstr0 - vstr1);
    shf2 = shf1 + fm * (shf0 - shf1);
    evap2 = evap1 + fm * (evap0 - evap1);
    rlus2 = rlus1 + fm * (rlus0 - rlus1);
    t0o = t1_1 + fm * (t1_0 - t1_1);
    tsfc = c.sst + fm * (c.stl - c.sst);
    tskin_o = c.sst + fm * (tskin - c.sst);
  }
  t.ustr0 = ustr0; t.vstr0 = vstr0; t.shf0 = shf0; t.evap0 = evap0; t.rlus0 = rlus0;
  t.ustr1 = ustr1; t.vstr1 = vstr1; t.shf1 = shf1; t.evap1 = evap1; t.rlus1 = rlus1;
  t.ustr2 = ustr2; t.vstr2 = vstr2; t.shf2 = shf2; t.evap2 = evap2; t.rlus2 = rlus2;
  b.du1 = ustr2 * rps * v.grdsig[KX - 1];
  b.dv1 = vstr2 * rps * v.grdsig[KX - 1];
  b.dts1 = shf2 * rps * v.grdscp[KX - 1];
  b.dqs1 = evap2 * rps * v.grdsig[KX - 1];
  b.evap2 = evap2;
  if (DIAG) {
    b.ustr[0] = ustr0; b.ustr[1] = ustr1; b.ustr[2] = ustr2;
    b.vstr[0] = vstr0; b.vstr[1] = vstr1; b.vstr[2] = vstr2;
    b.shf[0] = shf0; b.shf[1] = shf1; b.shf[2] = shf2;
    b.evap[0] = evap0; b.evap[1] = evap1; b.evap[2] = evap2;
    b.rlus[0] = rlus0; b.rlus[1] = rlus1; b.rlus[2] = rlus2;
    b.hfluxn[0] = hfl0; b.hfluxn[1] = hfl1;
    b.tsfc = tsfc; b.tskin = tskin_o; b.u0 = u0; b.v0 = v0; b.t0 = t0o;
  }
}

template <typename R, int KX>
struct LwuTape {
  R Fu[KX + 1][4];                // Fu[k]: band fluxes leaving level k upward (Fu[KX]: from the surface)
  R ft[2], dfl[KX];
};

// longwave up (longwave_radiation.py:119-202); dfl enters holding the
// longwave-down flux divergence.  Each level's transmissivities and emissions
// are recomputed (lw_level, equal to lw_transmissivity's and lw_emissions'), so
// the [KX][4] arrays of phase b1 are dead after longwave down and do not stay
// live across the shortwave and the surface fluxes.
template <typename R, int KX, bool DIAG, class TauStore = NoTauStore>
SPD_HD void lw_up(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p, const SpeedyVertical& v,
                  const Uniforms<R>& u, PhaseB<R, KX>& b, R rlus2, R tsfc, const R flux_h[4], LwuTape<R, KX>& t,
                  TauStore ts = TauStore()) {
  // level k's longwave quantities: the transmissivities from the store when
  // there is one (each level read after the previous level's flux exists)
  auto level = [&](int k, R dep, LwLevel<R>& L) {
    if constexpr (TauStore::kOn) {
      lw_level<R, KX, false>(k, c, a, p, v, u, L);
      const auto s = ts.after(dep);
#pragma unroll
      for (int bnd = 0; bnd < 4; ++bnd) L.tau[bnd] = s.get(k * 4 + bnd);
    } else {
      lw_level<R, KX>(k, c, a, p, v, u, L);
    }
  };
  const R ps = c.ps;
  const R rps = mdiv(R(1.0), ps);
  const R stratc0 = c.stratz * ps, stratc1 = u.eps1 * ps;
  const R rlds = b.rlds;
  R fbt[4];
  radset4(tsfc, u.omeps, fbt);
  R Fu[4];
#pragma unroll
  for (int bnd = 0; bnd < 4; ++bnd) { Fu[bnd] = fbt[bnd] * rlus2 + u.refsfc * flux_h[bnd]; t.Fu[KX][bnd] = Fu[bnd]; }
  R dfl[KX];
#pragma unroll
  for (int k = 0; k < KX; ++k) dfl[k] = b.dfl[k];
  dfl[KX - 1] = dfl[KX - 1] + p.epslw * rlus2;
  R Gnext = ((Fu[0] + Fu[1]) + Fu[2]) + Fu[3];
  R st4a0_1 = R(0.0);
#pragma unroll
  for (int k = KX - 1; k >= 1; --k) {
    LwLevel<R> L;
    level(k, Fu[0], L);
    if (k == 1) st4a0_1 = L.st0;
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) {
      const R em = R(1.0) - L.tau[bnd];
      Fu[bnd] = L.tau[bnd] * Fu[bnd] + em * (L.fb[bnd] * (L.st0 - em * L.st1));
      t.Fu[k][bnd] = Fu[bnd];
    }
    const R Gk = ((Fu[0] + Fu[1]) + Fu[2]) + Fu[3];
    dfl[k] = dfl[k] + (Gnext - Gk);
    Gnext = Gk;
  }
  R ft[4];
  LwLevel<R> L0;
  level(0, Fu[0], L0);
#pragma unroll
  for (int bnd = 0; bnd < 2; ++bnd) {
    const R em = R(1.0) - L0.tau[bnd];
    ft[bnd] = L0.tau[bnd] * Fu[bnd] + em * (L0.fb[bnd] * (L0.st0 - em * L0.st1));
  }
  ft[2] = Fu[2];
  ft[3] = Fu[3];
  t.ft[0] = ft[0]; t.ft[1] = ft[1];
  dfl[0] = dfl[0] + ((Fu[0] - ft[0]) + (Fu[1] - ft[1]));
  const R corlw1 = v.dhs[0] * stratc1 * L0.st0 + stratc0;
  const R corlw2 = v.dhs[1] * stratc1 * st4a0_1;
  dfl[0] = dfl[0] + (-corlw1);
  dfl[1] = dfl[1] + (-corlw2);
#pragma unroll
  for (int k = 0; k < KX; ++k) { b.tlw[k] = dfl[k] * v.grdscp[k] * rps; t.dfl[k] = dfl[k]; }
  if (DIAG) {
    b.ftop_lw = (corlw1 + corlw2) + (((ft[0] + ft[1]) + ft[2]) + ft[3]);
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) b.flux_out[bnd] = ft[bnd];
    b.rlns = rlus2 - rlds;
    b.stratc0 = stratc0; b.stratc1 = stratc1;
#pragma unroll
    for (int k = 0; k < KX; ++k) b.dfl[k] = dfl[k];
  }
}

template <typename R, int KX, bool DIAG, class TauStore = NoTauStore>
SPD_HD void phase_b2(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, uint32_t S, const SpeedyParams& p,
                     const SpeedyVertical& v, const Uniforms<R>& u, uint32_t flags, PhaseB<R, KX>& b,
                     TauStore ts = TauStore()) {
  R flux_h[4];
  SwTape<R, KX> sw;
  shortwave<R, KX, DIAG>(c, a, S, p, v, u, b, flux_h, sw);
  R tsfc;
  SfcTape<R> sf;
  surface<R, KX, DIAG>(c, a, p, v, u, flags, b, tsfc, sf);
  b.flux_h[0] = flux_h[0]; b.flux_h[1] = flux_h[1]; b.rlus2 = sf.rlus2; b.tsfc = tsfc;
  LwuTape<R, KX> lu;
  lw_up<R, KX, DIAG, TauStore>(c, a, p, v, u, b, sf.rlus2, tsfc, flux_h, lu, ts);
}

}  // namespace speedy
