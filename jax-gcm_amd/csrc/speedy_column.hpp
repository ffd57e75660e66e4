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

// jcm/constants.py:3-5, jcm/physics/speedy/physical_constants.py:8-45.  Python
// double expressions are evaluated in double and rounded once, as JAX does
// when a Python float meets a float32 array.
constexpr float P0 = 1e5f;
constexpr float GRAV = 9.81f;
constexpr float CP = 1004.0f;
constexpr float ALHC = 2501.0f;
constexpr float SBC = 5.67e-8f;
constexpr float RGAS = (float)(2.0 / 7.0 * 1004.0);
constexpr float RLHC = (float)(1.0 / 2501.0);
constexpr float RCP = (float)(1.0 / 1004.0);
constexpr float TFACT = (float)(2501.0 / 1004.0);
constexpr float PRG = (float)(1e5 / 9.81);
constexpr float RGAS288 = (float)(2.0 / 7.0 * 1004.0 * 288.0);
constexpr float EPS = 1e-9f;

SPD_HD float sq(float x) { return x * x; }
SPD_HD float pow3(float x) { return x * x * x; }
SPD_HD float pow4(float x) { float x2 = x * x; return x2 * x2; }

// get_qsat (humidity.py:73-104) for sig > 0 (the sig <= 0 branch never occurs
// on the path: sigma levels and the surface value 1.0 are positive).
SPD_HD float qsat_f(float ta, float ps, float sig) {
  const bool warm = ta >= 273.16f;
  const float c = warm ? 17.269f : 21.875f;
  const float tt = warm ? 35.86f : 7.66f;
  const float e = 6.108e-3f * expf(c * (ta - 273.16f) / (ta - tt));
  return 622.0f * e / (sig * ps - 0.378f * e);
}

// radset (longwave_radiation.py:204-225): (1-epslw) * band fractions.
SPD_HD void radset4(float temp, float omeps, float fb[4]) {
  const float jt = fminf(fmaxf(temp, 200.0f), 320.0f);
  const float f1 = 0.148f - 3.0e-6f * sq(jt - 247.0f);
  const float f2 = 0.356f - 5.2e-6f * sq(jt - 282.0f);
  const float f3 = 0.314f + 1.0e-5f * sq(jt - 315.0f);
  const float f0 = 1.0f - ((f1 + f2) + f3);
  fb[0] = omeps * f0; fb[1] = omeps * f1; fb[2] = omeps * f2; fb[3] = omeps * f3;
}

// Latitude-dependent part of solar and get_zonal_average_fields
// (shortwave_radiation.py:267-287,446-450) for one column, from the date-only
// uniforms.  out: fsol, ozupp, ozone, zenit, stratz, sqrt(coa).
struct ZonalUniforms {
  float sol_tdecl, sol_sdecl, sol_cdecl, sol_scale, oz_coz1, zen_c, zen_s;
};
SPD_HD void zonal_column(const ZonalUniforms& z, float sia, float coa, float out[6]) {
  // clip bounds +-(1-1e-9) are exactly +-1 in float32 (:446)
  const float ch0 = fminf(fmaxf(-z.sol_tdecl * sia / coa, -1.0f), 1.0f);
  const float h0 = acosf(ch0);
  const float sh0 = sinf(h0);
  const float topsr = z.sol_scale * (h0 * sia * z.sol_sdecl + sh0 * coa * z.sol_cdecl);
  const float flat2 = 1.5f * sia * sia - 0.5f;
  const float ozone0 = (float)(0.4 * 0.020) * (1.0f + z.oz_coz1 * sia + 1.8f * flat2);
  const float zenit = 1.0f + 1.0f * sq(1.0f - (coa * z.zen_c + sia * z.zen_s));
  out[0] = topsr;
  out[1] = topsr * (float)(0.5 * 0.020) * zenit;
  out[2] = topsr * ozone0 * zenit;
  out[3] = zenit;
  out[4] = fmaxf(6.0f - topsr, 0.0f);
  out[5] = sqrtf(coa);
}

// Inputs of one column.
template <int KX>
struct ColumnIn {
  float T[KX], Q[KX], PHI[KX];
  float ps, u1, v1;                               // u, v at the lowest level only
  float alb0, sice, snowc_am, soilw, stl, sst;    // forcing
  float fmask, orog, phis0;                       // geometry
  float fsol, ozupp, ozone, zenit, stratz, sqcoa; // per-latitude table
};

// Everything phase A (humidity -> convection -> condensation -> clouds) hands on.
template <int KX>
struct PhaseA {
  float qsat[KX], rh[KX], se[KX];
  float tconv[KX], qconv[KX];     // convection tendencies
  float tlsc[KX], qlsc[KX];       // condensation tendencies
  int iptop_conv, iptop;          // convective top; after condensation (Q4)
  float cbmf, precnv, precls;
  float cloudc, clstr, qcloud, gse;
  int icltop;
};

// Uniform constants that depend only on parameters, levels and the date.  They
// are computed once per call on the host (make_uniforms) in float32, in the
// reference's operation order, and passed to the kernels by value.
struct Uniforms {
  int32_t kx;
  uint32_t compute_sw;            // set_physics_flags: model_step % nstrad == 0 (speedy_physics.py:26-27)
  float wvi1[SPEEDY_KX_MAX];      // geometry.wvi[:, 1]
  // convection (convection.py:136-142)
  float entr[SPEEDY_KX_MAX], fm0, rdps;
  // large-scale condensation (large_scale_condensation.py:48-63,77)
  float rtlsc, rhref[SPEEDY_KX_MAX], dqmax[SPEEDY_KX_MAX], pfact[SPEEDY_KX_MAX];
  // clouds (shortwave_radiation.py:348,387)
  float rrcl, rgse;
  // shortwave absorptivity of dry air + aerosol per level (shortwave_radiation.py:82,93)
  float sw_abs1[SPEEDY_KX_MAX];
  // longwave (speedy/forcing.py:33; shortwave_radiation.py:203; longwave_radiation.py:100,156)
  float ablco2, eps1, omeps, refsfc, corlw;
  // surface fluxes (surface_flux.py:75-76,101,135-137,149,154,253,311)
  float esbc, ghum0, gtemp0, astab, rdth, rhdrag, cdl, cds, chlcp, chscp;
  // vertical diffusion (vertical_diffusion.py:45-63,99-118)
  float fshcq, fshcse, fvdiq2, drh0, fvdise;
  float rsig[SPEEDY_KX_MAX], rsig1[SPEEDY_KX_MAX], vd_fq[SPEEDY_KX_MAX], vd_drh0[SPEEDY_KX_MAX];
  uint32_t vd_on;                 // bit k: hsg[k+1] > 0.5
  // date-only parts of solar and get_zonal_average_fields (shortwave_radiation.py:250-262,421-443)
  ZonalUniforms zon;
};

inline void make_uniforms(const SpeedyParams& p, const SpeedyVertical& v, const SpeedyDate& d, Uniforms& u) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  const int kx = v.kx;
  u = Uniforms();
  u.kx = kx;
  int32_t ms = d.model_step % 3;
  if (ms < 0) ms += 3;
  u.compute_sw = ms == 0;
  for (int k = 0; k < kx; ++k) u.wvi1[k] = v.wvi[k][1];
  float e[SPEEDY_KX_MAX] = {0};
  float sentr = 0.0f;
  for (int k = 1; k < kx - 1; ++k) {
    e[k] = sq(fmaxf(0.0f, v.fsg[k] - 0.5f));
    sentr = (k == 1) ? e[k] : sentr + e[k];
  }
  const float sc = p.entmax / sentr;
  for (int k = 0; k < kx; ++k) u.entr[k] = e[k] * sc;
  u.fm0 = P0 * v.dhs[kx - 1] / (GRAV * p.trcnv * 3600.0f);
  u.rdps = 2.0f / (1.0f - p.psmin);
  u.rtlsc = 1.0f / (p.trlsc * 3600.0f);
  for (int k = 0; k < kx; ++k) {
    const float sig2 = sq(v.fsg[k]);
    u.rhref[k] = fmaxf(p.rhlsc + p.drhlsc * (sig2 - 1.0f), p.rhblsc);
    u.dqmax[k] = 10.0f * sig2 * u.rtlsc;
    u.pfact[k] = v.dhs[k] * PRG;
    u.sw_abs1[k] = p.absdry + p.absaer * sq(v.fsg[k]);
  }
  u.rrcl = 1.0f / (p.rhcl2 - p.rhcl1);
  u.rgse = 1.0f / (p.gse_s1 - p.gse_s0);
  float ab = 6.0f * expf((float)p.increase_co2 * 0.005f * ((float)d.model_year + d.tyear - (float)p.co2_year_ref));
  if (p.co2_scale != 1.0f) ab = ab * p.co2_scale;
  u.ablco2 = ab;
  u.eps1 = p.epslw / (v.dhs[0] + v.dhs[1]);
  u.omeps = 1.0f - p.epslw;
  u.refsfc = 1.0f - p.emisfc;
  u.corlw = p.epslw * p.emisfc;
  u.esbc = p.emisfc * SBC;
  u.ghum0 = 1.0f - p.fhum0;
  u.gtemp0 = 1.0f - p.ftemp0;
  u.astab = p.lscasym ? 0.5f : 1.0f;
  u.rdth = p.fstab / p.dtheta;
  u.rhdrag = 1.0f / (GRAV * p.hdrag);
  u.cdl = p.drag_scale != 1.0f ? p.cdl * p.drag_scale : p.cdl;
  u.cds = p.drag_scale != 1.0f ? p.cds * p.drag_scale : p.cds;
  u.chlcp = p.chl * CP;
  u.chscp = p.chs * CP;
  const int nl1 = kx - 1;
  const float cshc = v.dhs[kx - 1] / 3600.0f;
  const float cvdi = (v.hsg[nl1] - v.hsg[1]) / (float)((nl1 - 1) * 3600.0);
  u.fshcq = cshc / p.trshc;
  u.fshcse = cshc / (p.trshc * CP);
  const float fvdiq = cvdi / p.trvdi;
  u.fvdise = cvdi / (p.trvds * CP);
  u.fvdiq2 = fvdiq * v.hsg[nl1];
  u.drh0 = p.rhgrad * (v.fsg[kx - 1] - v.fsg[nl1 - 1]);
  {
    const float asin1 = asinf(1.0f);
    const float pigr = 2.0f * asin1;
    const float alpha_s = 2.0f * pigr * d.tyear;
    const float ca1 = cosf(alpha_s), sa1 = sinf(alpha_s);
    const float ca2 = ca1 * ca1 - sa1 * sa1;
    const float sa2 = 2.0f * sa1 * ca1;
    const float ca3 = ca1 * ca2 - sa1 * sa2;
    const float sa3 = sa1 * ca2 + sa2 * ca1;
    const float decl = 0.006918f - 0.399912f * ca1 + 0.070257f * sa1 - 0.006758f * ca2 + 0.000907f * sa2 -
                       0.002697f * ca3 + 0.001480f * sa3;
    const float fdis = 1.000110f + 0.034221f * ca1 + 0.001280f * sa1 + 0.000719f * ca2 + 0.000077f * sa2;
    u.zon.sol_cdecl = cosf(decl);
    u.zon.sol_sdecl = sinf(decl);
    u.zon.sol_tdecl = u.zon.sol_sdecl / u.zon.sol_cdecl;
    u.zon.sol_scale = ((4.0f * 342.0f) / pigr) * fdis;
    const float alpha = 4.0f * asin1 * (d.tyear + (float)(10.0 / 365.0));
    u.zon.oz_coz1 = fmaxf(0.0f, cosf(alpha));
    const float rzen = -cosf(alpha) * 23.45f * asin1 / 90.0f;
    u.zon.zen_c = cosf(rzen);
    u.zon.zen_s = sinf(rzen);
  }
  u.vd_on = 0;
  for (int k = 0; k < kx; ++k) {
    u.rsig[k] = 1.0f / v.dhs[k];
    u.rsig1[k] = (k < kx - 1) ? 1.0f / (1.0f - v.hsg[k + 1]) : 0.0f;
    if (k + 1 < kx) {
      u.vd_fq[k] = fvdiq * v.hsg[k + 1];
      u.vd_drh0[k] = p.rhgrad * (v.fsg[k + 1] - v.fsg[k]);
      if (v.hsg[k + 1] > 0.5f) u.vd_on |= 1u << k;
    }
  }
}

// ---------------------------------------------------------------------------
// Phase A.
// ---------------------------------------------------------------------------
template <int KX>
SPD_HD void phase_a(const ColumnIn<KX>& c, const SpeedyParams& p, const SpeedyVertical& v,
                    const Uniforms& u, PhaseA<KX>& a) {
  // humidity (humidity.py:40-46)
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    a.qsat[k] = qsat_f(c.T[k], c.ps, v.fsg[k]);
    a.rh[k] = c.Q[k] / a.qsat[k];
  }
  // ---- convection (convection.py:99-237)
#pragma unroll
  for (int k = 0; k < KX; ++k) a.se[k] = CP * c.T[k] + c.PHI[k];
  // diagnose_convection (:15-97)
  float mss2[KX];
  {
    float mss[KX];
#pragma unroll
    for (int k = 0; k < KX; ++k) mss[k] = a.se[k] + ALHC * a.qsat[k];
#pragma unroll
    for (int k = 0; k < KX - 1; ++k) mss2[k] = mss[k] + u.wvi1[k] * (mss[k + 1] - mss[k]);
    mss2[KX - 1] = 0.0f;
    const float mse0 = a.se[KX - 1] + ALHC * c.Q[KX - 1];
    const float mse1 = fminf(mse0, a.se[KX - 2] + ALHC * c.Q[KX - 2]);
    const float mss0 = fmaxf(mse0, mss[KX - 1]);
    int ktop1 = KX, ktop2 = KX;
#pragma unroll
    for (int k = KX - 4; k >= 2; --k) {   // first unstable level from the top wins
      if (mss0 > mss2[k]) ktop1 = k + 1;
      if (mse1 > mss2[k]) ktop2 = k + 1;
    }
    float msthr = mss2[KX - 1];
#pragma unroll
    for (int k = 2; k <= KX - 4; ++k)
      if (ktop2 == k + 1) msthr = mss2[k];
    const float qthr0 = p.rhbl * a.qsat[KX - 1];
    const float qthr1 = p.rhbl * a.qsat[KX - 2];
    const bool lqthr = (c.Q[KX - 1] > qthr0) && (c.Q[KX - 2] > qthr1);
    const bool mask_psa = c.ps > p.psmin;
    const bool case1 = mask_psa && (ktop1 < KX) && (ktop2 < KX);
    const bool case2 = mask_psa && (ktop1 < KX) && !(ktop2 < KX) && lqthr;
    a.iptop_conv = (case1 || case2) ? ktop1 : KX + 1;
    float qdif = 0.0f;
    if (case1) qdif = fmaxf(c.Q[KX - 1] - qthr0, (mse0 - msthr) * RLHC);
    if (case2) qdif = c.Q[KX - 1] - qthr0;
    // mass fluxes
    const int iptop = a.iptop_conv;
    const bool mask = iptop < KX;
    const float qmax = fmaxf(1.01f * c.Q[KX - 1], a.qsat[KX - 1]);
    float sbh[KX], qbh[KX];
    sbh[0] = 0.0f; qbh[0] = 0.0f;
#pragma unroll
    for (int k = 1; k < KX; ++k) {
      sbh[k] = a.se[k - 1] + u.wvi1[k - 1] * (a.se[k] - a.se[k - 1]);
      qbh[k] = c.Q[k - 1] + u.wvi1[k - 1] * (c.Q[k] - c.Q[k - 1]);
    }
    const float sb = sbh[KX - 1];
    const float qb = fminf(qbh[KX - 1], c.Q[KX - 1]);
    const float fpsa = c.ps * fminf(1.0f, (c.ps - p.psmin) * u.rdps);
    const float fmass = u.fm0 * fpsa * fminf(5.0f, qdif / (qmax - qb));
    a.cbmf = (mask ? 1.0f : 0.0f) * fmass;
    const float fus = fmass * a.se[KX - 1], fuq = fmass * qmax;
    const float fds = fmass * sb, fdq = fmass * qb;
    float dfse[KX], dfqa[KX];
    // entrainment and cumulative upward fluxes from the bottom (:178-191)
    float en[KX];
#pragma unroll
    for (int k = 0; k < KX; ++k) {
      const bool lm = (k <= KX - 2) && (k >= iptop);
      en[k] = (lm && k >= 1) ? u.entr[k] * c.ps * a.cbmf : 0.0f;
    }
    float FM[KX], FUS[KX], FUQ[KX];
    {
      float r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;
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
    float FDS[KX], FDQ[KX];
#pragma unroll
    for (int k = 0; k < KX - 1; ++k) { FDS[k] = FM[k] * sbh[k]; FDQ[k] = FM[k] * qbh[k]; }
    FDS[KX - 1] = fds; FDQ[KX - 1] = fdq;
    dfse[KX - 1] = fds - fus;
    dfqa[KX - 1] = fdq - fuq;
    float sumfsq = 0.0f;
#pragma unroll
    for (int k = 0; k < KX - 1; ++k) {
      const bool lm = (k >= iptop);
      dfse[k] = lm ? ((FUS[k + 1] - FDS[k + 1]) - (FUS[k] - FDS[k])) : 0.0f;
      dfqa[k] = lm ? ((FUQ[k + 1] - FDQ[k + 1]) - (FUQ[k] - FDQ[k])) : 0.0f;
      // secondary moisture flux (:198-202)
      const float delq = lm ? (p.rhil * a.qsat[k] - c.Q[k]) : 0.0f;
      const float fsq = (delq > 0.0f) ? p.smf * a.cbmf * delq : 0.0f;
      dfqa[k] = dfqa[k] + fsq;
      sumfsq = (k == 0) ? fsq : sumfsq + fsq;
    }
    dfqa[KX - 1] = dfqa[KX - 1] + (-sumfsq);
    // top layer (:204-219): gather at iptop (0 beyond KX-1), scatter at iptop-1 (Q6)
    float fm_t = 0.0f, fus_t = 0.0f, fuq_t = 0.0f, fds_t = 0.0f, fdq_t = 0.0f, qsatb = 0.0f;
#pragma unroll
    for (int k = 0; k < KX; ++k) {
      if (iptop == k) { fm_t = FM[k]; fus_t = FUS[k]; fuq_t = FUQ[k]; fds_t = FDS[k]; fdq_t = FDQ[k]; }
    }
#pragma unroll
    for (int k = 0; k < KX - 1; ++k)
      if (iptop - 1 == k) qsatb = a.qsat[k] + u.wvi1[k] * (a.qsat[k + 1] - a.qsat[k]);
    a.precnv = fmaxf(fuq_t - fm_t * qsatb, 0.0f);
    const float top_s = fus_t - fds_t + ALHC * a.precnv;
    const float top_q = fuq_t - fdq_t - a.precnv;
#pragma unroll
    for (int k = 0; k < KX; ++k)
      if (iptop - 1 == k) { dfse[k] = top_s; dfqa[k] = top_q; }
    // flux -> tendency (:226-227)
    const float rps = 1.0f / c.ps;
    a.tconv[0] = dfse[0];
    a.qconv[0] = dfqa[0];
#pragma unroll
    for (int k = 1; k < KX; ++k) {
      a.tconv[k] = dfse[k] * rps * v.grdscp[k];
      a.qconv[k] = dfqa[k] * rps * v.grdsig[k];
    }
  }
  // ---- large-scale condensation (large_scale_condensation.py:14-86)
  {
    const float psa2 = c.ps * c.ps;
    int first = 0;
    float precls = 0.0f;
    a.tlsc[0] = 0.0f;
    a.qlsc[0] = 0.0f;
#pragma unroll
    for (int k = KX - 1; k >= 1; --k) {
      const float dqa = u.rhref[k] * a.qsat[k] - c.Q[k];
      if (!(dqa >= 0.0f)) first = k;         // argmin(dqa[1:] >= 0) + 1 (Q4 when none)
      const bool neg = dqa < 0.0f;
      const float dq = neg ? dqa * u.rtlsc : 0.0f;
      a.qlsc[k] = dq;
      a.tlsc[k] = neg ? TFACT * fminf(-dq, u.dqmax[k] * psa2) : 0.0f;
    }
#pragma unroll
    for (int k = 1; k < KX; ++k) {
      const float t = u.pfact[k] * a.qlsc[k];
      precls = (k == 1) ? t : precls + t;
    }
    a.precls = (0.0f - precls) * c.ps;
    const int ipl = first == 0 ? 1 : first;
    a.iptop = ipl < a.iptop_conv ? ipl : a.iptop_conv;
  }
  // ---- clouds (shortwave_radiation.py:313-404)
  {
    a.gse = (a.se[KX - 2] - a.se[KX - 1]) / (c.PHI[KX - 2] - c.PHI[KX - 1]);
    const float rrcl = u.rrcl;
    const bool m0 = a.rh[KX - 2] > p.rhcl1;
    const float cc0 = m0 ? a.rh[KX - 2] - p.rhcl1 : 0.0f;
    const int icl0 = m0 ? KX - 2 : KX;
    float cu[KX];
    bool mk[KX];
    float cmax = 0.0f;
#pragma unroll
    for (int k = 2; k <= KX - 3; ++k) {
      const float drh = a.rh[k] - p.rhcl1;
      mk[k] = (drh > cc0) && (c.Q[k] > p.qacl);
      cu[k] = mk[k] ? drh : cc0;
      cmax = (k == 2) ? cu[k] : fmaxf(cmax, cu[k]);
    }
    int icl = 0;
#pragma unroll
    for (int k = 2; k <= KX - 3; ++k) {        // Q5: float equality
      const int cand = (cmax == cu[k]) ? (mk[k] ? k : icl0) : icl0;
      icl = (k == 2) ? cand : (cand > icl ? cand : icl);
    }
    const float pr1 = fminf(p.pmaxcl, 86.4f * (a.precnv + a.precls));
    float cloudc = fminf(1.0f, p.wpcl * sqrtf(fmaxf(EPS, pr1)) + sq(fminf(1.0f, cmax * rrcl)));
    if (cloudc != cloudc) cloudc = 1.0f;
    a.cloudc = cloudc;
    a.icltop = a.iptop < icl ? a.iptop : icl;
    a.qcloud = c.Q[KX - 2];
    const float rgse = u.rgse;
    const float fstab = fminf(fmaxf(rgse * (a.gse - p.gse_s0), 0.0f), 1.0f);
    float clstr = fstab * fmaxf(p.clsmax - 1.2f * cloudc, 0.0f);
    const float clstrl = fmaxf(clstr, p.clsminl) * a.rh[KX - 1];
    a.clstr = clstr + c.fmask * (clstrl - clstr);
  }
}

// Cloud top only (for the level set S): the decisions of phase A without the
// convective fluxes.  Must agree exactly with phase_a's icltop.
template <int KX>
SPD_HD int cloud_top(const ColumnIn<KX>& c, const SpeedyParams& p, const SpeedyVertical& v,
                     const Uniforms& u) {
  float qsat[KX], rh[KX], se[KX];
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    qsat[k] = qsat_f(c.T[k], c.ps, v.fsg[k]);
    rh[k] = c.Q[k] / qsat[k];
    se[k] = CP * c.T[k] + c.PHI[k];
  }
  int iptop_conv;
  {
    float mss[KX], mss2[KX];
#pragma unroll
    for (int k = 2; k < KX; ++k) mss[k] = se[k] + ALHC * qsat[k];
#pragma unroll
    for (int k = 2; k <= KX - 4; ++k) mss2[k] = mss[k] + u.wvi1[k] * (mss[k + 1] - mss[k]);
    const float mse0 = se[KX - 1] + ALHC * c.Q[KX - 1];
    const float mse1 = fminf(mse0, se[KX - 2] + ALHC * c.Q[KX - 2]);
    const float mss0 = fmaxf(mse0, mss[KX - 1]);
    int ktop1 = KX, ktop2 = KX;
#pragma unroll
    for (int k = KX - 4; k >= 2; --k) {
      if (mss0 > mss2[k]) ktop1 = k + 1;
      if (mse1 > mss2[k]) ktop2 = k + 1;
    }
    const float qthr0 = p.rhbl * qsat[KX - 1];
    const float qthr1 = p.rhbl * qsat[KX - 2];
    const bool lqthr = (c.Q[KX - 1] > qthr0) && (c.Q[KX - 2] > qthr1);
    const bool mask_psa = c.ps > p.psmin;
    const bool case1 = mask_psa && (ktop1 < KX) && (ktop2 < KX);
    const bool case2 = mask_psa && (ktop1 < KX) && !(ktop2 < KX) && lqthr;
    iptop_conv = (case1 || case2) ? ktop1 : KX + 1;
  }
  int first = 0;
#pragma unroll
  for (int k = KX - 1; k >= 1; --k) {
    const float dqa = u.rhref[k] * qsat[k] - c.Q[k];
    if (!(dqa >= 0.0f)) first = k;
  }
  const int ipl = first == 0 ? 1 : first;
  const int iptop = ipl < iptop_conv ? ipl : iptop_conv;
  const bool m0 = rh[KX - 2] > p.rhcl1;
  const float cc0 = m0 ? rh[KX - 2] - p.rhcl1 : 0.0f;
  const int icl0 = m0 ? KX - 2 : KX;
  float cu[KX];
  bool mk[KX];
  float cmax = 0.0f;
#pragma unroll
  for (int k = 2; k <= KX - 3; ++k) {
    const float drh = rh[k] - p.rhcl1;
    mk[k] = (drh > cc0) && (c.Q[k] > p.qacl);
    cu[k] = mk[k] ? drh : cc0;
    cmax = (k == 2) ? cu[k] : fmaxf(cmax, cu[k]);
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
template <int KX>
struct PhaseB {
  // b1
  float albsfc, alb_l, alb_s, snowc;
  float tauL[KX][4], st4a0[KX], st4a1[KX], dfl[KX], rlds;
  float tvd[KX], qvd[KX];
  // b2
  float tsw[KX], tlw[KX];
  float du1, dv1, dts1, dqs1;        // surface-flux tendencies at level KX-1
  // diagnostics (b2)
  float dfabs_sw[KX], rsds, rsns, ftop_sw, stratc0, stratc1;
  float rlns, ftop_lw, flux_out[4];
  float ustr[3], vstr[3], shf[3], evap[3], rlus[3], hfluxn[2], tsfc, tskin, u0, v0, t0;
};

// Transmissivities use the hardware exponential (v_exp_f32 on x*log2(e)):
// relative error ~6e-8*|x| (< 2e-6 where the result matters), far inside the
// 1e-5 parity budget; decision-bearing exponentials (qsat) stay libm-accurate.
SPD_HD float exp_tr(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __expf(x);
#else
  return expf(x);
#endif
}

template <int KX>
SPD_HD void phase_b1(const ColumnIn<KX>& c, const PhaseA<KX>& a, const SpeedyParams& p, const SpeedyVertical& v,
                     const Uniforms& u, PhaseB<KX>& b) {
  // ---- set_forcing albedos (speedy/forcing.py:26-30)
  b.snowc = fminf(1.0f, c.snowc_am);
  b.alb_l = c.alb0 + b.snowc * (p.albsn - c.alb0);
  b.alb_s = p.albsea + c.sice * (p.albice - p.albsea);
  b.albsfc = b.alb_s + c.fmask * (b.alb_l - b.alb_s);
  const float ps = c.ps;
  // ---- longwave transmissivity (shortwave_radiation.py:177-200)
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    float ab[4] = {p.ablwin, u.ablco2, p.ablwv1 * c.Q[k], p.ablwv2 * c.Q[k]};
    if (k == 0) { ab[2] = 0.0f; ab[3] = 0.0f; }
    if (k >= 2 && k <= KX - 2) {
      ab[0] = ab[0] + ((k > a.icltop) ? a.cloudc * p.ablcl1 : a.cloudc * p.ablcl2);
      const float ac2 = a.cloudc * p.ablcl2;
      ab[2] = fmaxf(ab[2], ac2);
      ab[3] = fmaxf(ab[3], ac2);
    }
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) b.tauL[k][bnd] = exp_tr(-ab[bnd] * ps * v.dhs[k]);
  }
  // ---- longwave down (longwave_radiation.py:13-117)
  {
    float Th[KX];
#pragma unroll
    for (int k = 0; k < KX - 1; ++k) Th[k] = c.T[k] + u.wvi1[k] * (c.T[k + 1] - c.T[k]);
    const float s0 = 0.75f * c.T[0] + 0.25f * Th[0];
    const float s1 = 0.50f * c.T[1] + 0.25f * (Th[0] + Th[1]);
    float grad[KX];
#pragma unroll
    for (int k = 2; k < KX - 1; ++k) grad[k] = 0.5f * fmaxf(Th[k] - Th[k - 1], 0.0f);
    grad[KX - 1] = fmaxf(c.T[KX - 1] - Th[KX - 2], 0.0f);
    b.st4a0[0] = SBC * pow4(s0); b.st4a0[1] = SBC * pow4(s1);
    b.st4a1[0] = 0.0f; b.st4a1[1] = 0.0f;
#pragma unroll
    for (int k = 2; k < KX; ++k) {
      const float st3a = SBC * pow3(c.T[k]);
      b.st4a0[k] = st3a * c.T[k];
      b.st4a1[k] = 4.0f * st3a * grad[k];
    }
    float Fd[4];
    {
      float fb[4];
      radset4(c.T[0], u.omeps, fb);
#pragma unroll
      for (int bnd = 0; bnd < 4; ++bnd) {
        const float em = 1.0f - b.tauL[0][bnd];
        Fd[bnd] = bnd < 2 ? em * (fb[bnd] * (b.st4a0[0] + em * b.st4a1[0])) : 0.0f;
      }
    }
    float Fprev = ((Fd[0] + Fd[1]) + Fd[2]) + Fd[3];
    b.dfl[0] = -Fprev;
#pragma unroll
    for (int k = 1; k < KX; ++k) {
      float fb[4];
      radset4(c.T[k], u.omeps, fb);
#pragma unroll
      for (int bnd = 0; bnd < 4; ++bnd) {
        const float em = 1.0f - b.tauL[k][bnd];
        Fd[bnd] = b.tauL[k][bnd] * Fd[bnd] + em * (fb[bnd] * (b.st4a0[k] + em * b.st4a1[k]));
      }
      const float Fs = ((Fd[0] + Fd[1]) + Fd[2]) + Fd[3];
      b.dfl[k] = -(Fs - Fprev);
      Fprev = Fs;
    }
    float rlds = ((p.emisfc * Fd[0] + p.emisfc * Fd[1]) + p.emisfc * Fd[2]) + p.emisfc * Fd[3];
    const float corlw = u.corlw * b.st4a0[KX - 1];
    b.dfl[KX - 1] = b.dfl[KX - 1] + (-corlw);
    b.rlds = rlds + corlw;
  }
  // ---- vertical diffusion (vertical_diffusion.py:10-135)
  {
    const int icnv = KX - a.iptop - 1;
    constexpr int NL1 = KX - 1;
    float tt[KX], qt[KX];
#pragma unroll
    for (int k = 0; k < KX; ++k) { tt[k] = 0.0f; qt[k] = 0.0f; }
    const float dmse = a.se[KX - 1] - a.se[NL1 - 1] + ALHC * (c.Q[KX - 1] - a.qsat[NL1 - 1]);
    const float drh = a.rh[KX - 1] - a.rh[NL1 - 1];
    const float fcnv = (icnv > 0 && dmse >= 0.0f) ? p.redshc : 1.0f;
    const float rsA = u.rsig[NL1 - 1], rsB = u.rsig[KX - 1];
    if (dmse >= 0.0f) {
      const float fluxse = fcnv * u.fshcse * dmse;
      tt[NL1 - 1] = fluxse * rsA;
      tt[KX - 1] = -fluxse * rsB;
      if (drh >= 0.0f) {
        const float fluxq = fcnv * u.fshcq * a.qsat[KX - 1] * drh;
        qt[NL1 - 1] = fluxq * rsA;
        qt[KX - 1] = -fluxq * rsB;
      }
    } else if (dmse < 0.0f && drh > u.drh0) {
      const float fluxq = u.fvdiq2 * a.qsat[NL1 - 1] * drh;
      qt[NL1 - 1] = fluxq * rsA;
      qt[KX - 1] = -fluxq * rsB;
    }
    // moisture diffusion above the PBL (:99-118)
    float fq[KX];
#pragma unroll
    for (int k = 2; k < KX - 2; ++k) {
      const float d = a.rh[k + 1] - a.rh[k];
      fq[k] = (d >= u.vd_drh0[k] && ((u.vd_on >> k) & 1u)) ? u.vd_fq[k] * a.qsat[k] * d : 0.0f;
    }
#pragma unroll
    for (int k = 2; k < KX - 2; ++k) qt[k] = qt[k] + fq[k] * u.rsig[k];
#pragma unroll
    for (int k = 2; k < KX - 2; ++k) qt[k + 1] = qt[k + 1] + (-fq[k] * u.rsig[k + 1]);
    // super-adiabatic damping (:120-131): all local adds, then the cumulative sum
    float fse[KX];
#pragma unroll
    for (int k = 0; k < NL1; ++k) {
      const float se0 = a.se[k + 1] - p.segrad * (c.PHI[k + 1] - c.PHI[k]);
      fse[k] = (a.se[k] < se0) ? u.fvdise * (se0 - a.se[k]) : 0.0f;
      tt[k] = tt[k] + fse[k] * u.rsig[k];
    }
    float cum = 0.0f;
#pragma unroll
    for (int k = 0; k < NL1; ++k) {
      cum = (k == 0) ? fse[k] * u.rsig1[k] : cum + fse[k] * u.rsig1[k];
      tt[k + 1] = tt[k + 1] + (-cum);
    }
#pragma unroll
    for (int k = 0; k < KX; ++k) { b.tvd[k] = tt[k]; b.qvd[k] = qt[k]; }
  }
}

template <int KX, bool DIAG>
SPD_HD void phase_b2(const ColumnIn<KX>& c, const PhaseA<KX>& a, uint32_t S, const SpeedyParams& p,
                     const SpeedyVertical& v, const Uniforms& u, uint32_t flags, PhaseB<KX>& b) {
  const float ps = c.ps;
  const float rps = 1.0f / ps;
  // ---- shortwave (shortwave_radiation.py:27-217)
  float flux_h[4];
  {
    const float a_cl = (a.icltop < KX) ? p.albcl * a.cloudc : 0.0f;
    float t22[KX];
#pragma unroll
    for (int k = 0; k < KX - 1; ++k) t22[k] = ((S >> k) & 1u) ? a_cl : 0.0f;   // Q1
    t22[KX - 1] = p.albcls * a.clstr;
    const float psaz = ps * c.zenit;
    const float acloud = a.cloudc * fminf(p.abscl1 * a.qcloud, p.abscl2);
    float tau0[KX], tau1[KX];
    tau0[0] = exp_tr(-psaz * v.dhs[0] * p.absdry);
#pragma unroll
    for (int k = 1; k < KX - 1; ++k) {
      const float cl = (k >= a.icltop) ? acloud : 0.0f;
      tau0[k] = exp_tr(-psaz * v.dhs[k] * (u.sw_abs1[k] + p.abswv1 * c.Q[k] + cl));
    }
    tau0[KX - 1] = exp_tr(-psaz * v.dhs[KX - 1] * (u.sw_abs1[KX - 1] + p.abswv1 * c.Q[KX - 1]));
#pragma unroll
    for (int k = 1; k < KX; ++k) tau1[k] = exp_tr(-psaz * v.dhs[k] * p.abswv2 * c.Q[k]);
    // downward
    float dfabs[KX], f1[KX], f2[KX], R[KX];
    const float F0 = c.fsol * 0.95f;
    dfabs[0] = F0;
    f1[0] = tau0[0] * (F0 - c.ozupp * ps);
    dfabs[0] = dfabs[0] + (-f1[0]);
    dfabs[1] = f1[0];
    f1[1] = tau0[1] * (f1[0] - c.ozone * ps);
    dfabs[1] = dfabs[1] + (-f1[1]);
#pragma unroll
    for (int k = 2; k < KX; ++k) f1[k] = f1[k - 1] * tau0[k] * (1.0f - t22[k]);
#pragma unroll
    for (int k = 2; k < KX; ++k) {
      dfabs[k] = f1[k - 1] * (1.0f - t22[k]) * (1.0f - tau0[k]);
      R[k] = t22[k] * f1[k - 1];
    }
    R[0] = t22[0];
    R[1] = t22[1];
    f2[0] = c.fsol * 0.05f;
    float carry = f2[0];
#pragma unroll
    for (int k = 1; k < KX; ++k) { carry = carry * tau1[k]; f2[k] = carry; }
#pragma unroll
    for (int k = 1; k < KX; ++k) dfabs[k] = dfabs[k] + f2[k - 1] * (1.0f - tau1[k]);
    // upward
    const float rsds = f1[KX - 1] + f2[KX - 1];
    float U = f1[KX - 1] * b.albsfc;
    const float rsns = rsds - U;
    float fu[KX];
    fu[KX - 1] = U;
#pragma unroll
    for (int k = KX - 1; k >= 1; --k) { U = U * tau0[k] + R[k]; fu[k - 1] = U; }
#pragma unroll
    for (int k = 0; k < KX; ++k) dfabs[k] = dfabs[k] + fu[k] * (1.0f - tau0[k]);
    const float ftoa = tau0[0] * fu[0] + R[0];
    const bool csw = u.compute_sw != 0u;
#pragma unroll
    for (int k = 0; k < KX; ++k) b.tsw[k] = csw ? dfabs[k] * v.grdscp[k] * rps : 0.0f;
    flux_h[0] = ftoa; flux_h[1] = f2[KX - 1]; flux_h[2] = 0.0f; flux_h[3] = 0.0f;   // Q2
    b.rsds = rsds;
    if (DIAG) {
#pragma unroll
      for (int k = 0; k < KX; ++k) b.dfabs_sw[k] = dfabs[k];
      b.rsns = rsns;
      b.ftop_sw = c.fsol - ftoa;
    }
  }
  const float stratc0 = c.stratz * ps, stratc1 = u.eps1 * ps;
  const float rlds = b.rlds;
  // ---- surface fluxes (surface_flux.py:15-300)
  float rlus2, tsfc;
  {
    const SpeedyParams& s = p;
    const float u1 = c.u1, v1 = c.v1, Tb = c.T[KX - 1], Qb = c.Q[KX - 1];
    const float u0 = s.fwind0 * u1, v0 = s.fwind0 * v1;
    float t1_0 = 0.0f, t1_1 = 0.0f, t2_1 = 0.0f, q1_1 = 0.0f;
    float den0 = 0.0f, den2 = 0.0f;
    float ustr0 = 0.0f, vstr0 = 0.0f, shf0 = 0.0f, evap0 = 0.0f, rlus0 = 0.0f, hfl0 = 0.0f, tskin = 0.0f;
    float ustr1 = 0.0f, vstr1 = 0.0f;
    const bool land = flags & SPEEDY_FLAG_LFLUXLAND;
    if (land) {
      const float phi0 = c.orog * GRAV;
      const float dt1 = v.wvi[KX - 1][1] * (Tb - c.T[KX - 2]);
      t1_0 = Tb + dt1;
      t1_1 = t1_0 - phi0 * dt1 / (RGAS288 * v.sigl[KX - 1]);
      t2_1 = Tb + RCP * c.PHI[KX - 1];
      const float t2_0 = t2_1 - RCP * phi0;
      if (Tb > c.T[KX - 2]) {
        t1_0 = s.ftemp0 * t1_0 + u.gtemp0 * t2_0;
        t1_1 = s.ftemp0 * t1_1 + u.gtemp0 * t2_1;
      } else {
        t1_0 = Tb;
        t1_1 = Tb;
      }
      const float t0 = t1_1 + c.fmask * (t1_0 - t1_1);
      den0 = (P0 * ps / (RGAS * t0)) * sqrtf(u0 * u0 + v0 * v0 + s.vgust * s.vgust);
      tskin = c.stl + s.ctday * c.sqcoa * b.rsds * (1.0f - b.alb_l) * ps;
      const float dd = tskin - t2_0;
      const float dthl = (tskin > t2_0) ? fminf(s.dtheta, dd) : fmaxf(-s.dtheta, u.astab * dd);
      const float den1 = den0 * (1.0f + dthl * u.rdth);
      const float forog = 1.0f + u.rhdrag * (1.0f - exp_tr(-fmaxf(c.phis0, 0.0f) * u.rhdrag));
      const float cdldv = u.cdl * den0 * forog;
      ustr0 = -cdldv * u1;
      vstr0 = -cdldv * v1;
      shf0 = u.chlcp * den1 * (tskin - t1_0);
      float q1_0 = Qb;
      if (s.fhum0 > 0.0f) q1_0 = s.fhum0 * (a.rh[KX - 1] * qsat_f(t1_0, ps, 1.0f)) + u.ghum0 * Qb;
      const float qsat00 = qsat_f(tskin, ps, 1.0f);
      evap0 = s.chl * den1 * fmaxf(0.0f, c.soilw * qsat00 - q1_0);
      const float tsk3 = pow3(tskin);
      const float drls = 4.0f * u.esbc * tsk3;
      rlus0 = u.esbc * tsk3 * tskin;
      hfl0 = b.rsds * (1.0f - b.alb_l) + rlds - (rlus0 + shf0 + (ALHC * evap0));
      if (s.lskineb) {
        const float clamb = s.clambda + (b.snowc * (s.clambsn - s.clambda));
        hfl0 = hfl0 - (clamb * (tskin - c.stl));
        float qs1 = qsat_f(tskin + 1.0f, ps, 1.0f);
        qs1 = (evap0 > 0.0f) ? c.soilw * (qs1 - qsat00) : 0.0f;
        const float dtskin = hfl0 / (clamb + drls + (s.chl * den1 * (CP + (ALHC * qs1))));
        tskin = tskin + dtskin;
        shf0 = shf0 + u.chlcp * den1 * dtskin;
        evap0 = evap0 + s.chl * den1 * qs1 * dtskin;
        rlus0 = rlus0 + drls * dtskin;
        hfl0 = clamb * (tskin - c.stl);
      }
      const float ds = c.sst - t2_1;
      const float dths = (c.sst > t2_1) ? fminf(s.dtheta, ds) : fmaxf(-s.dtheta, u.astab * ds);
      den2 = den0 * (1.0f + dths * u.rdth);
      q1_1 = Qb;
      if (s.fhum0 > 0.0f) q1_1 = s.fhum0 * (a.rh[KX - 1] * qsat_f(t1_1, ps, 1.0f)) + u.ghum0 * Qb;
      const float cdsdv = u.cds * den2;
      ustr1 = -cdsdv * u1;
      vstr1 = -cdsdv * v1;
    }
    // sea (:250-261)
    const float shf1 = u.chscp * den2 * (c.sst - t1_1);
    const float qsat01 = qsat_f(c.sst, ps, 1.0f);
    const float evap1 = s.chs * den2 * (qsat01 - q1_1);
    const float rlus1 = u.esbc * pow4(c.sst);
    const float hfl1 = b.rsds * (1.0f - b.alb_s) + rlds - rlus1 + shf1 + ALHC * evap1;
    float ustr2 = 0.0f, vstr2 = 0.0f, shf2 = 0.0f, evap2 = 0.0f, t0o = 0.0f;
    rlus2 = 0.0f; tsfc = 0.0f;
    float tskin_o = 0.0f;
    if (land) {
      const float fm = c.fmask;
      ustr2 = ustr1 + fm * (ustr0 - ustr1);
      vstr2 = vstr1 + fm * (vstr0 - vstr1);
      shf2 = shf1 + fm * (shf0 - shf1);
      evap2 = evap1 + fm * (evap0 - evap1);
      rlus2 = rlus1 + fm * (rlus0 - rlus1);
      t0o = t1_1 + fm * (t1_0 - t1_1);
      tsfc = c.sst + fm * (c.stl - c.sst);
      tskin_o = c.sst + fm * (tskin - c.sst);
    }
    b.du1 = ustr2 * rps * v.grdsig[KX - 1];
    b.dv1 = vstr2 * rps * v.grdsig[KX - 1];
    b.dts1 = shf2 * rps * v.grdscp[KX - 1];
    b.dqs1 = evap2 * rps * v.grdsig[KX - 1];
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
  // ---- longwave up (longwave_radiation.py:119-202)
  {
    float fbt[4];
    radset4(tsfc, u.omeps, fbt);
    float Fu[4];
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) Fu[bnd] = fbt[bnd] * rlus2 + u.refsfc * flux_h[bnd];
    float dfl[KX];
#pragma unroll
    for (int k = 0; k < KX; ++k) dfl[k] = b.dfl[k];
    dfl[KX - 1] = dfl[KX - 1] + p.epslw * rlus2;
    float Gnext = ((Fu[0] + Fu[1]) + Fu[2]) + Fu[3];
#pragma unroll
    for (int k = KX - 1; k >= 1; --k) {
      float fb[4];
      radset4(c.T[k], u.omeps, fb);
#pragma unroll
      for (int bnd = 0; bnd < 4; ++bnd) {
        const float em = 1.0f - b.tauL[k][bnd];
        Fu[bnd] = b.tauL[k][bnd] * Fu[bnd] + em * (fb[bnd] * (b.st4a0[k] - em * b.st4a1[k]));
      }
      const float Gk = ((Fu[0] + Fu[1]) + Fu[2]) + Fu[3];
      dfl[k] = dfl[k] + (Gnext - Gk);
      Gnext = Gk;
    }
    float ft[4];
    {
      float fb[4];
      radset4(c.T[0], u.omeps, fb);
#pragma unroll
      for (int bnd = 0; bnd < 2; ++bnd) {
        const float em = 1.0f - b.tauL[0][bnd];
        ft[bnd] = b.tauL[0][bnd] * Fu[bnd] + em * (fb[bnd] * (b.st4a0[0] - em * b.st4a1[0]));
      }
      ft[2] = Fu[2];
      ft[3] = Fu[3];
    }
    dfl[0] = dfl[0] + ((Fu[0] - ft[0]) + (Fu[1] - ft[1]));
    const float corlw1 = v.dhs[0] * stratc1 * b.st4a0[0] + stratc0;
    const float corlw2 = v.dhs[1] * stratc1 * b.st4a0[1];
    dfl[0] = dfl[0] + (-corlw1);
    dfl[1] = dfl[1] + (-corlw2);
#pragma unroll
    for (int k = 0; k < KX; ++k) b.tlw[k] = dfl[k] * v.grdscp[k] * rps;
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
}

}  // namespace speedy
