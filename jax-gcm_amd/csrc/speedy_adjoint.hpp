// speedy_adjoint.hpp -- reverse mode (vector-Jacobian product) of the SPEEDY
// column physics of speedy_column.hpp, written by hand term by term.
//
// It replaces jax.vjp of SpeedyPhysics.compute_tendencies
// (jcm/physics/speedy/speedy_physics.py:88-127) with respect to the state
// fields T, q, u, v, ps and five parameters: convection.rhbl (params.py:24),
// convection.trcnv (:22), condensation.rhlsc (:60), a multiplier on the CO2
// longwave absorptivity (speedy/forcing.py:33) and a multiplier on the surface
// drag coefficients cdl, cds (params.py:191-192).
//
// Conventions follow JAX's differentiation of the reference expressions:
//   * discrete decisions (convective top, cloud top, the level set S, every
//     jnp.where condition) are held fixed; the gradient flows through the
//     selected branch only;
//   * max/min/clip give 1/2 to each operand at a tie (lax.max / lax.min);
//   * a reduction max (jnp.max over levels, clouds) splits the gradient evenly
//     over the tied entries.
// Each term is recomputed (its forward function fills a tape) right before it
// is reversed, so only the phase outputs live across the whole adjoint.
#pragma once

#include "speedy_column.hpp"

namespace speedy {

enum { VJP_RHBL = 0, VJP_TRCNV = 1, VJP_RHLSC = 2, VJP_CO2 = 3, VJP_DRAG = 4, VJP_NPARAM = 5 };

// Floats per column of the longwave tape store used by column_vjp.
template <int KX>
constexpr int vjp_tape_floats() { return 4 * (KX + 1); }

// Cotangents of one column's tendencies.  Only level KX-1 of the u and v
// tendencies is nonzero in the forward (surface stress), so only those matter.
template <typename R, int KX>
struct ColumnCt {
  R du1, dv1;
  R dt[KX], dq[KX];
};

// Adjoint accumulators: the state gradients, the parameter gradients and the
// adjoints of the intermediates that several terms share.
template <typename R, int KX>
struct ColumnAdj {
  R T[KX], Q[KX], ps, u1, v1;
  R prm[VJP_NPARAM];
  R qsat[KX], rh[KX], se[KX];
  R cloudc, clstr, qcloud, precnv, precls, rps;
};

// Materialize values at this point of the program: an empty asm that reads and
// writes them.  Computations feeding them cannot be sunk past it, so a term's
// (or a level's) reversal completes before the next one's recompute starts and
// its intermediates die there.
template <typename R>
SPD_HD void settle(R& x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#else
  (void)x;
#endif
}

// The fences are applied where the kernel then fits in registers (up to eight
// levels: no scratch).  At sixteen levels the accumulators alone are ~100
// values; forcing them into VGPRs at every fence makes the kernel spill around
// the fences, and that build gave wrong forward tendencies in the batched
// forward + VJP kernel on gfx950 (the host build of the same source is right).
// Sixteen-level kernels therefore keep the unfenced form.
template <int KX>
constexpr bool kSettle = KX <= 8;

template <typename R, int KX>
SPD_HD void settle_adj(ColumnAdj<R, KX>& g) {
  if constexpr (!kSettle<KX>) return;
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    settle(g.T[k]); settle(g.Q[k]); settle(g.qsat[k]); settle(g.rh[k]); settle(g.se[k]);
  }
  settle(g.ps); settle(g.u1); settle(g.v1);
#pragma unroll
  for (int i = 0; i < VJP_NPARAM; ++i) settle(g.prm[i]);
  settle(g.cloudc); settle(g.clstr); settle(g.qcloud); settle(g.precnv); settle(g.precls); settle(g.rps);
}

template <typename R, int KX>
SPD_HD void adj_zero(ColumnAdj<R, KX>& g) {
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    g.T[k] = R(0.0); g.Q[k] = R(0.0); g.qsat[k] = R(0.0); g.rh[k] = R(0.0); g.se[k] = R(0.0);
  }
  g.ps = R(0.0); g.u1 = R(0.0); g.v1 = R(0.0);
#pragma unroll
  for (int i = 0; i < VJP_NPARAM; ++i) g.prm[i] = R(0.0);
  g.cloudc = R(0.0); g.clstr = R(0.0); g.qcloud = R(0.0); g.precnv = R(0.0); g.precls = R(0.0);
  g.rps = R(0.0);
}

// verify_tendencies (physics_interface.py:321-342) in reverse: where a level's
// q tendency was limited to -q/dt, its cotangent reaches the unclamped q as
// -1/dt and none of it reaches the physics.  g_qraw receives that part.
template <typename R, int KX>
SPD_HD void verify_ct(uint32_t flags, uint32_t limited, R dt, ColumnCt<R, KX>& ct, R g_qraw[KX]) {
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    g_qraw[k] = R(0.0);
    if ((flags & SPEEDY_FLAG_VERIFY_TENDENCIES) && ((limited >> k) & 1u)) {
      g_qraw[k] = -ct.dq[k] / dt;
      ct.dq[k] = R(0.0);
    }
  }
}

// verify_state in reverse: where(q < 0, 0, q) passes the gradient only where q >= 0.
template <typename R>
SPD_HD R verify_q_grad(R q_raw, R g) { return q_raw < R(0.0) ? R(0.0) : g; }

// d qsat / d ta and d qsat / d ps of qsat_f (humidity.py:95-102).
template <typename R>
SPD_HD void qsat_grad(R ta, R ps, R sig, R& dq_dt, R& dq_dps) {
  const bool warm = ta >= R(273.16);
  const R c = warm ? R(17.269) : R(21.875);
  const R tt = warm ? R(35.86) : R(7.66);
  const R e = R(6.108e-3) * mexp(c * (ta - R(273.16)) / (ta - tt));
  const R D = sig * ps - R(0.378) * e;
  const R rD = R(1.0) / D;
  const R de_dt = e * c * (R(273.16) - tt) / sq(ta - tt);
  dq_dt = R(622.0) * sig * ps * rD * rD * de_dt;
  dq_dps = -R(622.0) * e * sig * rD * rD;
}

// Adjoint of radset4 (longwave_radiation.py:216-225): returns d/d temp of
// sum_b a_fb[b] * fb[b].  jnp.clip(temp, 200, 320) = min(max(temp, 200), 320).
template <typename R>
SPD_HD R radset4_adj(R temp, R omeps, const R a_fb[4]) {
  const R jlo = mmax(temp, R(200.0));
  const R jt = mmin(jlo, R(320.0));
  const R d1 = R(-2.0 * 3.0e-6) * (jt - R(247.0));
  const R d2 = R(-2.0 * 5.2e-6) * (jt - R(282.0));
  const R d3 = R(2.0 * 1.0e-5) * (jt - R(315.0));
  const R d0 = -((d1 + d2) + d3);
  const R a_jt = omeps * (a_fb[0] * d0 + a_fb[1] * d1 + a_fb[2] * d2 + a_fb[3] * d3);
  return a_jt * dmin_x(jlo, R(320.0)) * dmax_x(temp, R(200.0));
}

// Adjoint of level k's longwave transmissivities tauL[k][:] = exp(-ab * ps * dhs)
// (shortwave_radiation.py:177-200) and emissions st4a0[k], st4a1[k]
// (longwave_radiation.py:48-68), applied at once so that no per-level adjoint
// arrays stay live.  Gradients of the half-level temperatures go to a_Th.
template <typename R, int KX>
SPD_HD void lw_level_adj(int k, const R a_tau[4], R a_s0, R a_s1, const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a,
                         const R tau[4], const SpeedyParams& p, const SpeedyVertical& v, const Uniforms<R>& u,
                         ColumnAdj<R, KX>& g, R a_Th[KX]) {
  R ab[4];
  lw_absorptivity<R, KX>(k, c, a, p, u, ab);
#pragma unroll
  for (int bnd = 0; bnd < 4; ++bnd) {
    const R ax = a_tau[bnd] * tau[bnd];
    g.ps += ax * (-(ab[bnd] * v.dhs[k]));
    const R a_ab = ax * (-(c.ps * v.dhs[k]));
    if (bnd == 0) {
      if (k >= 2 && k <= KX - 2) g.cloudc += a_ab * ((k > a.icltop) ? p.ablcl1 : p.ablcl2);
    } else if (bnd == 1) {
      g.prm[VJP_CO2] += a_ab * u.ablco2_d;
    } else if (k >= 1) {
      const R wv = (bnd == 2) ? p.ablwv1 : p.ablwv2;
      if (k >= 2 && k <= KX - 2) {
        const R x = wv * c.Q[k], ac2 = a.cloudc * p.ablcl2;
        g.Q[k] += a_ab * dmax_x(x, ac2) * wv;
        g.cloudc += a_ab * dmax_x(ac2, x) * p.ablcl2;
      } else {
        g.Q[k] += a_ab * wv;
      }
    }
  }
  // emissions; Th[j] = T[j] + wvi1[j] * (T[j+1] - T[j]) as in lw_down
  auto Th = [&](int j) { return c.T[j] + u.wvi1[j] * (c.T[j + 1] - c.T[j]); };
  if (k >= 2) {
    const R T = c.T[k];
    const R st3a = R(SBC) * pow3(T);
    R grad;
    if (k < KX - 1) grad = R(0.5) * mmax(Th(k) - Th(k - 1), R(0.0));
    else grad = mmax(c.T[KX - 1] - Th(KX - 2), R(0.0));
    const R a_st3a = a_s0 * T + a_s1 * R(4.0) * grad;
    g.T[k] += a_s0 * st3a + a_st3a * R(3.0 * SBC) * T * T;
    const R a_grad = a_s1 * R(4.0) * st3a;
    if (k < KX - 1) {
      const R ax = a_grad * R(0.5) * dmax_x(Th(k) - Th(k - 1), R(0.0));
      a_Th[k] += ax;
      a_Th[k - 1] += -ax;
    } else {
      const R ax = a_grad * dmax_x(c.T[KX - 1] - Th(KX - 2), R(0.0));
      g.T[KX - 1] += ax;
      a_Th[KX - 2] += -ax;
    }
  } else if (k == 0) {
    const R s0 = R(0.75) * c.T[0] + R(0.25) * Th(0);
    const R as = a_s0 * R(4.0 * SBC) * pow3(s0);
    g.T[0] += R(0.75) * as;
    a_Th[0] += R(0.25) * as;
  } else {
    const R s1 = R(0.50) * c.T[1] + R(0.25) * (Th(0) + Th(1));
    const R as = a_s0 * R(4.0 * SBC) * pow3(s1);
    g.T[1] += R(0.50) * as;
    a_Th[0] += R(0.25) * as;
    a_Th[1] += R(0.25) * as;
  }
}

// The column as level k's longwave recompute sees it: the values lw_level(k)
// reads (T at k-1..k+1, q at k, ps) are passed through an empty asm that also
// takes `dep`, the adjoint state entering level k, so the recompute cannot be
// hoisted above the reversal of the previous level (which would keep every
// level's recomputed values live at once).
template <typename R, int KX>
SPD_HD ColumnIn<R, KX> column_at_level(const ColumnIn<R, KX>& c, int k, R dep) {
  ColumnIn<R, KX> cc = c;
#if defined(__HIP_DEVICE_COMPILE__)
  for (int j = k - 1; j <= k + 1; ++j)
    if (j >= 0 && j < KX) asm volatile("" : "+v"(cc.T[j]) : "v"(dep));
  asm volatile("" : "+v"(cc.Q[k]) : "v"(dep));
  asm volatile("" : "+v"(cc.ps) : "v"(dep));
#else
  (void)k; (void)dep;
#endif
  return cc;
}

// ---------------------------------------------------------------- longwave up
// Each level's transmissivities, emissions and band fractions are recomputed
// here (lw_level) from the column, not taken from the forward.
template <typename R, int KX, class Store>
SPD_HD void lw_up_adj(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const PhaseB<R, KX>& b,
                      const SpeedyParams& p, const SpeedyVertical& v, const Uniforms<R>& u, const Store& Fu,
                      const R a_dfl[KX], R& a_rlus2, R a_flux_h[2], ColumnAdj<R, KX>& g, R a_Th[KX]) {
  const R ps = c.ps;
  const R stratc1 = u.eps1 * ps;
  LwLevel<R> L0, L1;
  lw_level<R, KX>(0, c, a, p, v, u, L0);
  lw_level<R, KX>(1, c, a, p, v, u, L1);
  // corrections of the two top levels
  const R a_c1 = -a_dfl[0], a_c2 = -a_dfl[1];
  const R a_strat1 = a_c1 * v.dhs[0] * L0.st0 + a_c2 * v.dhs[1] * L1.st0;
  g.ps += a_c1 * c.stratz + a_strat1 * u.eps1;
  // top level: dfl[0] += (F0 - ft0) + (F1 - ft1), F = Fu[1]
  R aF[4];
  {
    R afb[4] = {R(0.0), R(0.0), R(0.0), R(0.0)};
    R atau[4] = {R(0.0), R(0.0), R(0.0), R(0.0)};
    R as0 = a_c1 * v.dhs[0] * stratc1, as1 = R(0.0);
#pragma unroll
    for (int bnd = 0; bnd < 2; ++bnd) {
      const R a_ft = -a_dfl[0];
      const R tau = L0.tau[bnd], em = R(1.0) - tau;
      aF[bnd] = a_dfl[0] + a_ft * tau;
      atau[bnd] = a_ft * (Fu.get(4 + bnd) - L0.fb[bnd] * (L0.st0 - R(2.0) * em * L0.st1));
      afb[bnd] = a_ft * em * (L0.st0 - em * L0.st1);
      as0 += a_ft * em * L0.fb[bnd];
      as1 += -a_ft * em * em * L0.fb[bnd];
    }
    aF[2] = R(0.0); aF[3] = R(0.0);
    g.T[0] += radset4_adj(c.T[0], u.omeps, afb);
    lw_level_adj<R, KX>(0, atau, as0, as1, c, a, L0.tau, p, v, u, g, a_Th);
  }
  // levels 1 .. KX-1 (the forward ran from KX-1 down to 1)
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    LwLevel<R> Lk;
    if (k == 1) Lk = L1;
    else lw_level<R, KX>(k, column_at_level<R, KX>(c, k, aF[0]), a, p, v, u, Lk);
    R afb[4], atau[4];
    R as0 = (k == 1) ? a_c2 * v.dhs[1] * stratc1 : R(0.0), as1 = R(0.0);
    const auto Fk = Fu.after(aF[0]);   // Fu[k+1] read when level k is reversed
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) {
      const R tau = Lk.tau[bnd], em = R(1.0) - tau;
      const R a = aF[bnd] - a_dfl[k];            // dfl[k] += sum Fu[k+1] - sum Fu[k]
      atau[bnd] = a * (Fk.get((k + 1) * 4 + bnd) - Lk.fb[bnd] * (Lk.st0 - R(2.0) * em * Lk.st1));
      afb[bnd] = a * em * (Lk.st0 - em * Lk.st1);
      as0 += a * em * Lk.fb[bnd];
      as1 += -a * em * em * Lk.fb[bnd];
      aF[bnd] = a_dfl[k] + a * tau;
    }
    g.T[k] += radset4_adj(c.T[k], u.omeps, afb);
    lw_level_adj<R, KX>(k, atau, as0, as1, c, a, Lk.tau, p, v, u, g, a_Th);
    settle_adj<R, KX>(g);
    if constexpr (kSettle<KX>) {
#pragma unroll
      for (int j = 0; j < KX; ++j) settle(a_Th[j]);
#pragma unroll
      for (int bnd = 0; bnd < 4; ++bnd) settle(aF[bnd]);
    }
  }
  // surface source: Fu[KX] = fbt * rlus2 + refsfc * flux_h; dfl[KX-1] += epslw * rlus2
  R fbt[4];
  radset4(b.tsfc, u.omeps, fbt);
  a_rlus2 += ((aF[0] * fbt[0] + aF[1] * fbt[1]) + (aF[2] * fbt[2] + aF[3] * fbt[3])) + a_dfl[KX - 1] * p.epslw;
  a_flux_h[0] += aF[0] * u.refsfc;
  a_flux_h[1] += aF[1] * u.refsfc;
}

// ---------------------------------------------------------------- surface fluxes
template <typename R, int KX>
SPD_HD void surface_adj(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const PhaseB<R, KX>& b,
                        const SpeedyParams& s, const SpeedyVertical& v, const Uniforms<R>& u, const SfcTape<R>& t,
                        const ColumnCt<R, KX>& ct, R a_rlus2, R& a_rsds, R& a_rlds, ColumnAdj<R, KX>& g) {
  if (!t.land) return;   // every surface output is zero without lfluxland (Q8)
  const R ps = c.ps;
  const R rps = R(1.0) / ps;
  const R fm = c.fmask;
  const R gs = v.grdsig[KX - 1], gc = v.grdscp[KX - 1];
  const R a_ustr2 = ct.du1 * rps * gs, a_vstr2 = ct.dv1 * rps * gs;
  const R a_shf2 = ct.dt[KX - 1] * rps * gc, a_evap2 = ct.dq[KX - 1] * rps * gs;
  g.rps += ((ct.du1 * t.ustr2 * gs + ct.dv1 * t.vstr2 * gs) + ct.dt[KX - 1] * t.shf2 * gc) + ct.dq[KX - 1] * t.evap2 * gs;
  const R om = R(1.0) - fm;
  R a_ustr0 = fm * a_ustr2, a_ustr1 = om * a_ustr2;
  R a_vstr0 = fm * a_vstr2, a_vstr1 = om * a_vstr2;
  R a_shf0 = fm * a_shf2, a_shf1 = om * a_shf2;
  R a_evap0 = fm * a_evap2, a_evap1 = om * a_evap2;
  R a_rlus0 = fm * a_rlus2;
  R a_Tb = R(0.0), a_Qb = R(0.0), a_t1_0 = R(0.0), a_t1_1 = R(0.0), a_t2_0 = R(0.0), a_t2_1 = R(0.0);
  R a_den0 = R(0.0), a_den1 = R(0.0), a_den2 = R(0.0), a_tskin = R(0.0), a_qsat00 = R(0.0);
  R dqt, dqp;
  // sea
  a_den2 += a_evap1 * s.chs * (t.qsat01 - t.q1_1);
  const R a_qsat01 = a_evap1 * s.chs * t.den2;
  R a_q1_1 = -a_evap1 * s.chs * t.den2;
  qsat_grad(c.sst, ps, R(1.0), dqt, dqp);
  g.ps += a_qsat01 * dqp;
  a_den2 += a_shf1 * u.chscp * (c.sst - t.t1_1);
  a_t1_1 += -a_shf1 * u.chscp * t.den2;
  // land: sea-side stress and humidity
  {
    const R cdsdv = u.cds * t.den2;
    const R a_cdsdv = -(a_ustr1 * c.u1 + a_vstr1 * c.v1);
    g.u1 += -a_ustr1 * cdsdv;
    g.v1 += -a_vstr1 * cdsdv;
    a_den2 += a_cdsdv * u.cds;
    g.prm[VJP_DRAG] += a_cdsdv * t.den2 * s.cds;
  }
  if (s.fhum0 > R(0.0)) {
    const R qs = qsat_f(t.t1_1, ps, R(1.0));
    const R a_qs = a_q1_1 * s.fhum0 * a.rh[KX - 1];
    g.rh[KX - 1] += a_q1_1 * s.fhum0 * qs;
    qsat_grad(t.t1_1, ps, R(1.0), dqt, dqp);
    a_t1_1 += a_qs * dqt;
    g.ps += a_qs * dqp;
    a_Qb += a_q1_1 * u.ghum0;
  } else {
    a_Qb += a_q1_1;
  }
  a_den0 += a_den2 * (R(1.0) + t.dths * u.rdth);
  {
    const R a_dths = a_den2 * t.den0 * u.rdth;
    const R a_ds = (c.sst > t.t2_1) ? a_dths * dmin_x(t.ds, R(s.dtheta))
                                    : a_dths * dmax_x(u.astab * t.ds, -R(s.dtheta)) * u.astab;
    a_t2_1 += -a_ds;
  }
  // skin-temperature energy-balance correction (surface_flux.py:189-221)
  R a_hfl0 = R(0.0), a_drls = R(0.0);
  if (s.lskineb) {
    R a_dtskin = a_shf0 * u.chlcp * t.den1;
    a_den1 += a_shf0 * u.chlcp * t.dtskin;
    a_den1 += a_evap0 * s.chl * t.qs1 * t.dtskin;
    R a_qs1 = a_evap0 * s.chl * t.den1 * t.dtskin;
    a_dtskin += a_evap0 * s.chl * t.den1 * t.qs1;
    a_drls += a_rlus0 * t.dtskin;
    a_dtskin += a_rlus0 * t.drls;
    const R Dn = t.clamb + t.drls + (s.chl * t.den1 * (R(CP) + (R(ALHC) * t.qs1)));
    const R a_hflc = a_dtskin / Dn;
    const R a_Dn = -a_dtskin * t.dtskin / Dn;
    a_drls += a_Dn;
    a_den1 += a_Dn * s.chl * (R(CP) + R(ALHC) * t.qs1);
    a_qs1 += a_Dn * s.chl * t.den1 * R(ALHC);
    if (t.evap0a > R(0.0)) {
      const R a_qs1p = a_qs1 * c.soilw;
      a_qsat00 += -a_qs1 * c.soilw;
      qsat_grad(t.tskin0 + R(1.0), ps, R(1.0), dqt, dqp);
      a_tskin += a_qs1p * dqt;
      g.ps += a_qs1p * dqp;
    }
    a_hfl0 += a_hflc;
    a_tskin += -a_hflc * t.clamb;
  }
  // net surface heat flux hfl0 (surface_flux.py:185)
  a_rsds += a_hfl0 * (R(1.0) - b.alb_l);
  a_rlds += a_hfl0;
  a_rlus0 += -a_hfl0;
  a_shf0 += -a_hfl0;
  a_evap0 += -a_hfl0 * R(ALHC);
  {
    const R a_tsk3 = a_rlus0 * u.esbc * t.tskin0 + a_drls * R(4.0) * u.esbc;
    a_tskin += a_rlus0 * u.esbc * t.tsk3 + a_tsk3 * R(3.0) * t.tskin0 * t.tskin0;
  }
  {
    a_den1 += a_evap0 * s.chl * mmax(R(0.0), t.evx);
    const R a_evx = a_evap0 * s.chl * t.den1 * dmax_x(t.evx, R(0.0));
    a_qsat00 += a_evx * c.soilw;
    const R a_q1_0 = -a_evx;
    qsat_grad(t.tskin0, ps, R(1.0), dqt, dqp);
    a_tskin += a_qsat00 * dqt;
    g.ps += a_qsat00 * dqp;
    if (s.fhum0 > R(0.0)) {
      const R qs = qsat_f(t.t1_0, ps, R(1.0));
      const R a_qs = a_q1_0 * s.fhum0 * a.rh[KX - 1];
      g.rh[KX - 1] += a_q1_0 * s.fhum0 * qs;
      qsat_grad(t.t1_0, ps, R(1.0), dqt, dqp);
      a_t1_0 += a_qs * dqt;
      g.ps += a_qs * dqp;
      a_Qb += a_q1_0 * u.ghum0;
    } else {
      a_Qb += a_q1_0;
    }
  }
  a_den1 += a_shf0 * u.chlcp * (t.tskin0 - t.t1_0);
  a_tskin += a_shf0 * u.chlcp * t.den1;
  a_t1_0 += -a_shf0 * u.chlcp * t.den1;
  {
    const R a_cdldv = -(a_ustr0 * c.u1 + a_vstr0 * c.v1);
    g.u1 += -a_ustr0 * t.cdldv;
    g.v1 += -a_vstr0 * t.cdldv;
    a_den0 += a_cdldv * u.cdl * t.forog;
    g.prm[VJP_DRAG] += a_cdldv * s.cdl * t.den0 * t.forog;
  }
  a_den0 += a_den1 * (R(1.0) + t.dthl * u.rdth);
  {
    const R a_dthl = a_den1 * t.den0 * u.rdth;
    const R a_dd = (t.tskin0 > t.t2_0) ? a_dthl * dmin_x(t.dd, R(s.dtheta))
                                       : a_dthl * dmax_x(u.astab * t.dd, -R(s.dtheta)) * u.astab;
    a_tskin += a_dd;
    a_t2_0 += -a_dd;
  }
  // tskin0 = stl + ctday * sqcoa * rsds * (1 - alb_l) * ps
  {
    const R k0 = s.ctday * c.sqcoa;
    a_rsds += a_tskin * k0 * (R(1.0) - b.alb_l) * ps;
    g.ps += a_tskin * k0 * b.rsds * (R(1.0) - b.alb_l);
  }
  // den0 = (p0 * ps / (rgas * t0)) * wind
  R a_t0;
  {
    const R rho = R(P0) * ps / (R(RGAS) * t.t0);
    const R a_rho = a_den0 * t.wind;
    const R a_wind = a_den0 * rho;
    g.ps += a_rho * (R(P0) / (R(RGAS) * t.t0));
    a_t0 = -a_rho * rho / t.t0;
    g.u1 += a_wind * t.u0 / t.wind * s.fwind0;
    g.v1 += a_wind * t.v0 / t.wind * s.fwind0;
  }
  a_t1_0 += a_t0 * fm;
  a_t1_1 += a_t0 * om;
  R a_t1_0a = R(0.0), a_t1_1a = R(0.0);
  if (t.warm) {
    a_t1_0a = a_t1_0 * s.ftemp0;
    a_t2_0 += a_t1_0 * u.gtemp0;
    a_t1_1a = a_t1_1 * s.ftemp0;
    a_t2_1 += a_t1_1 * u.gtemp0;
  } else {
    a_Tb += a_t1_0 + a_t1_1;
  }
  a_t2_1 += a_t2_0;
  a_Tb += a_t2_1;
  a_t1_0a += a_t1_1a;
  R a_dt1 = -a_t1_1a * t.phi0 / (R(RGAS288) * v.sigl[KX - 1]);
  a_Tb += a_t1_0a;
  a_dt1 += a_t1_0a;
  a_Tb += a_dt1 * v.wvi[KX - 1][1];
  g.T[KX - 2] += -a_dt1 * v.wvi[KX - 1][1];
  g.T[KX - 1] += a_Tb;
  g.Q[KX - 1] += a_Qb;
}

// ---------------------------------------------------------------- shortwave
template <typename R, int KX>
SPD_HD void shortwave_adj(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const PhaseB<R, KX>& b, uint32_t S,
                          const SpeedyParams& p, const SpeedyVertical& v, const Uniforms<R>& u,
                          const SwTape<R, KX>& t, const ColumnCt<R, KX>& ct, const R a_flux_h[2], R a_rsds,
                          ColumnAdj<R, KX>& g) {
  const R ps = c.ps;
  const R rps = R(1.0) / ps;
  const bool csw = u.compute_sw != 0u;
  R a_dfabs[KX];
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    a_dfabs[k] = csw ? ct.dt[k] * v.grdscp[k] * rps : R(0.0);
    if (csw) g.rps += ct.dt[k] * t.dfabs[k] * v.grdscp[k];
  }
  R a_tau0[KX], a_tau1[KX], a_f1[KX], a_f2[KX], a_Rf[KX], a_fu[KX], a_t22[KX];
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    a_tau0[k] = R(0.0); a_tau1[k] = R(0.0); a_f1[k] = R(0.0); a_f2[k] = R(0.0);
    a_Rf[k] = R(0.0); a_fu[k] = R(0.0); a_t22[k] = R(0.0);
  }
  // ftoa = tau0[0] * fu[0] + Rf[0]; flux_h = (ftoa, f2[KX-1])
  a_tau0[0] += a_flux_h[0] * t.fu[0];
  a_fu[0] += a_flux_h[0] * t.tau0[0];
  a_Rf[0] += a_flux_h[0];
  a_f2[KX - 1] += a_flux_h[1];
  // dfabs[k] += fu[k] * (1 - tau0[k])
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    a_fu[k] += a_dfabs[k] * (R(1.0) - t.tau0[k]);
    a_tau0[k] += -a_dfabs[k] * t.fu[k];
  }
  // upward scan fu[k-1] = fu[k] * tau0[k] + Rf[k]
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    a_fu[k] += a_fu[k - 1] * t.tau0[k];
    a_tau0[k] += a_fu[k - 1] * t.fu[k];
    a_Rf[k] += a_fu[k - 1];
  }
  a_f1[KX - 1] += a_fu[KX - 1] * b.albsfc + a_rsds;
  a_f2[KX - 1] += a_rsds;
  // dfabs[k] += f2[k-1] * (1 - tau1[k]); f2[k] = f2[k-1] * tau1[k]
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    a_f2[k - 1] += a_dfabs[k] * (R(1.0) - t.tau1[k]);
    a_tau1[k] += -a_dfabs[k] * t.f2[k - 1];
  }
#pragma unroll
  for (int k = KX - 1; k >= 1; --k) {
    a_f2[k - 1] += a_f2[k] * t.tau1[k];
    a_tau1[k] += a_f2[k] * t.f2[k - 1];
  }
  a_t22[0] += a_Rf[0];
  a_t22[1] += a_Rf[1];
  // downward scan, levels KX-1 .. 2
#pragma unroll
  for (int k = KX - 1; k >= 2; --k) {
    const R om22 = R(1.0) - t.t22[k], omt = R(1.0) - t.tau0[k];
    const R fp = t.f1[k - 1];
    a_f1[k - 1] += a_dfabs[k] * om22 * omt + a_Rf[k] * t.t22[k] + a_f1[k] * t.tau0[k] * om22;
    a_t22[k] += -a_dfabs[k] * fp * omt + a_Rf[k] * fp - a_f1[k] * fp * t.tau0[k];
    a_tau0[k] += -a_dfabs[k] * fp * om22 + a_f1[k] * fp * om22;
  }
  // levels 1 and 0 (ozone absorption)
  a_f1[0] += a_dfabs[1];
  a_f1[1] += -a_dfabs[1];
  a_tau0[1] += a_f1[1] * (t.f1[0] - c.ozone * ps);
  a_f1[0] += a_f1[1] * t.tau0[1];
  g.ps += -a_f1[1] * t.tau0[1] * c.ozone;
  a_f1[0] += -a_dfabs[0];
  const R F0 = c.fsol * R(0.95);
  a_tau0[0] += a_f1[0] * (F0 - c.ozupp * ps);
  g.ps += -a_f1[0] * t.tau0[0] * c.ozupp;
  // transmissivities
  R a_psaz = R(0.0), a_acloud = R(0.0);
  a_psaz += a_tau0[0] * t.tau0[0] * (-(v.dhs[0] * p.absdry));
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    const R cl = (k <= KX - 2 && k >= a.icltop) ? t.acloud : R(0.0);
    const R A = u.sw_abs1[k] + p.abswv1 * c.Q[k] + cl;
    const R ax = a_tau0[k] * t.tau0[k];
    a_psaz += ax * (-(v.dhs[k] * A));
    const R a_A = ax * (-(t.psaz * v.dhs[k]));
    g.Q[k] += a_A * p.abswv1;
    if (k <= KX - 2 && k >= a.icltop) a_acloud += a_A;
    const R bx = a_tau1[k] * t.tau1[k];
    a_psaz += bx * (-(v.dhs[k] * p.abswv2 * c.Q[k]));
    g.Q[k] += bx * (-(t.psaz * v.dhs[k] * p.abswv2));
  }
  {
    const R x = p.abscl1 * a.qcloud;
    g.cloudc += a_acloud * mmin(x, R(p.abscl2));
    g.qcloud += a_acloud * a.cloudc * dmin_x(x, R(p.abscl2)) * p.abscl1;
  }
  g.ps += a_psaz * c.zenit;
  // cloud albedo at the levels of S (Q1) and stratiform albedo at KX-1
  R a_acl = R(0.0);
#pragma unroll
  for (int k = 0; k < KX - 1; ++k)
    if ((S >> k) & 1u) a_acl += a_t22[k];
  g.clstr += a_t22[KX - 1] * p.albcls;
  if (a.icltop < KX) g.cloudc += a_acl * p.albcl;
}

// ---------------------------------------------------------------- longwave down
// As lw_up_adj, each level's longwave quantities are recomputed (lw_level).
template <typename R, int KX, class Store>
SPD_HD void lw_down_adj(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p,
                        const SpeedyVertical& v, const Uniforms<R>& u, const Store& Fd, const R a_dfl[KX], R a_rlds,
                        ColumnAdj<R, KX>& g, R a_Th[KX]) {
  R aF[4];
#pragma unroll
  for (int bnd = 0; bnd < 4; ++bnd) aF[bnd] = a_rlds * p.emisfc;
#pragma unroll
  for (int k = KX - 1; k >= 1; --k) {
    LwLevel<R> Lk;
    lw_level<R, KX>(k, column_at_level<R, KX>(c, k, aF[0]), a, p, v, u, Lk);
    R afb[4], atau[4];
    // rlds = sum_b emisfc * Fd[KX-1][b] + corlw; dfl[KX-1] -= corlw; corlw = corlw_u * st4a0[KX-1]
    R as0 = (k == KX - 1) ? (a_rlds - a_dfl[KX - 1]) * u.corlw : R(0.0), as1 = R(0.0);
    const auto Fk = Fd.after(aF[0]);   // Fd[k-1] read when level k is reversed
#pragma unroll
    for (int bnd = 0; bnd < 4; ++bnd) {
      const R tau = Lk.tau[bnd], em = R(1.0) - tau;
      const R a = aF[bnd] - a_dfl[k];           // dfl[k] = sum Fd[k-1] - sum Fd[k]
      atau[bnd] = a * (Fk.get((k - 1) * 4 + bnd) - Lk.fb[bnd] * (Lk.st0 + R(2.0) * em * Lk.st1));
      afb[bnd] = a * em * (Lk.st0 + em * Lk.st1);
      as0 += a * em * Lk.fb[bnd];
      as1 += a * em * em * Lk.fb[bnd];
      aF[bnd] = a_dfl[k] + a * tau;
    }
    g.T[k] += radset4_adj(c.T[k], u.omeps, afb);
    lw_level_adj<R, KX>(k, atau, as0, as1, c, a, Lk.tau, p, v, u, g, a_Th);
    settle_adj<R, KX>(g);
    if constexpr (kSettle<KX>) {
#pragma unroll
      for (int j = 0; j < KX; ++j) settle(a_Th[j]);
#pragma unroll
      for (int bnd = 0; bnd < 4; ++bnd) settle(aF[bnd]);
    }
  }
  {
    LwLevel<R> L0;
    lw_level<R, KX>(0, column_at_level<R, KX>(c, 0, aF[0]), a, p, v, u, L0);
    R afb[4] = {R(0.0), R(0.0), R(0.0), R(0.0)};
    R atau[4] = {R(0.0), R(0.0), R(0.0), R(0.0)};
    R as0 = R(0.0), as1 = R(0.0);
#pragma unroll
    for (int bnd = 0; bnd < 2; ++bnd) {
      const R tau = L0.tau[bnd], em = R(1.0) - tau;
      const R a = aF[bnd] - a_dfl[0];
      atau[bnd] = -a * L0.fb[bnd] * (L0.st0 + R(2.0) * em * L0.st1);
      afb[bnd] = a * em * (L0.st0 + em * L0.st1);
      as0 += a * em * L0.fb[bnd];
      as1 += a * em * em * L0.fb[bnd];
    }
    g.T[0] += radset4_adj(c.T[0], u.omeps, afb);
    lw_level_adj<R, KX>(0, atau, as0, as1, c, a, L0.tau, p, v, u, g, a_Th);
  }
}

// Half-level temperatures Th[k] = T[k] + wvi1[k] * (T[k+1] - T[k]) back to T.
template <typename R, int KX>
SPD_HD void half_level_adj(const Uniforms<R>& u, const R a_Th[KX], ColumnAdj<R, KX>& g) {
#pragma unroll
  for (int k = 0; k < KX - 1; ++k) {
    g.T[k] += a_Th[k] * (R(1.0) - u.wvi1[k]);
    g.T[k + 1] += a_Th[k] * u.wvi1[k];
  }
}

// ---------------------------------------------------------------- vertical diffusion
template <typename R, int KX>
SPD_HD void vdiff_adj(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p, const Uniforms<R>& u,
                      const ColumnCt<R, KX>& ct, ColumnAdj<R, KX>& g) {
  constexpr int NL1 = KX - 1;
  const int icnv = KX - a.iptop - 1;
  // super-adiabatic damping: tt[k] += fse[k] * rsig[k]; tt[k+1] -= sum_{j<=k} fse[j] * rsig1[j]
  R suf = R(0.0);
#pragma unroll
  for (int k = NL1 - 1; k >= 0; --k) {
    suf = suf + ct.dt[k + 1];
    const R se0 = a.se[k + 1] - p.segrad * (c.PHI[k + 1] - c.PHI[k]);
    if (a.se[k] < se0) {
      const R a_fse = ct.dt[k] * u.rsig[k] - suf * u.rsig1[k];
      const R a_x = a_fse * u.fvdise;
      g.se[k + 1] += a_x;
      g.se[k] += -a_x;
    }
  }
  // moisture diffusion above the PBL
#pragma unroll
  for (int k = 2; k < KX - 2; ++k) {
    const R d = a.rh[k + 1] - a.rh[k];
    if (d >= u.vd_drh0[k] && ((u.vd_on >> k) & 1u)) {
      const R a_fq = ct.dq[k] * u.rsig[k] - ct.dq[k + 1] * u.rsig[k + 1];
      g.qsat[k] += a_fq * u.vd_fq[k] * d;
      const R a_d = a_fq * u.vd_fq[k] * a.qsat[k];
      g.rh[k + 1] += a_d;
      g.rh[k] += -a_d;
    }
  }
  // shallow convection / PBL moisture exchange between the two lowest levels
  const R dmse = a.se[KX - 1] - a.se[NL1 - 1] + R(ALHC) * (c.Q[KX - 1] - a.qsat[NL1 - 1]);
  const R drh = a.rh[KX - 1] - a.rh[NL1 - 1];
  const R fcnv = (icnv > 0 && dmse >= R(0.0)) ? p.redshc : R(1.0);
  const R rsA = u.rsig[NL1 - 1], rsB = u.rsig[KX - 1];
  R a_dmse = R(0.0), a_drh = R(0.0);
  const R a_fq2 = ct.dq[NL1 - 1] * rsA - ct.dq[KX - 1] * rsB;
  if (dmse >= R(0.0)) {
    a_dmse = (ct.dt[NL1 - 1] * rsA - ct.dt[KX - 1] * rsB) * fcnv * u.fshcse;
    if (drh >= R(0.0)) {
      g.qsat[KX - 1] += a_fq2 * fcnv * u.fshcq * drh;
      a_drh = a_fq2 * fcnv * u.fshcq * a.qsat[KX - 1];
    }
  } else if (drh > u.drh0) {
    g.qsat[NL1 - 1] += a_fq2 * u.fvdiq2 * drh;
    a_drh = a_fq2 * u.fvdiq2 * a.qsat[NL1 - 1];
  }
  g.se[KX - 1] += a_dmse;
  g.se[NL1 - 1] += -a_dmse;
  g.Q[KX - 1] += a_dmse * R(ALHC);
  g.qsat[NL1 - 1] += -a_dmse * R(ALHC);
  g.rh[KX - 1] += a_drh;
  g.rh[NL1 - 1] += -a_drh;
}

// ---------------------------------------------------------------- clouds
template <typename R, int KX>
SPD_HD void clouds_adj(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p,
                       const Uniforms<R>& u, const CloudTape<R, KX>& t, ColumnAdj<R, KX>& g) {
  const R fm = c.fmask;
  R a_cl0 = g.clstr * (R(1.0) - fm);
  const R a_cll = g.clstr * fm;
  g.rh[KX - 1] += a_cll * mmax(t.clstr0, R(p.clsminl));
  a_cl0 += a_cll * a.rh[KX - 1] * dmax_x(t.clstr0, R(p.clsminl));
  const R a_fstab = a_cl0 * mmax(t.clsx, R(0.0));
  const R a_clsx = a_cl0 * t.fstab * dmax_x(t.clsx, R(0.0));
  const R a_cc = g.cloudc - R(1.2) * a_clsx;
  const R a_fsa = a_fstab * dmin_x(mmax(t.fstab_arg, R(0.0)), R(1.0)) * dmax_x(t.fstab_arg, R(0.0));
  const R a_gse = a_fsa * u.rgse;
  const R rdphi = R(1.0) / (c.PHI[KX - 2] - c.PHI[KX - 1]);
  g.se[KX - 2] += a_gse * rdphi;
  g.se[KX - 1] += -a_gse * rdphi;
  g.Q[KX - 2] += g.qcloud;
  const R a_raw = a_cc * dmin_x(t.ccraw, R(1.0));
  const R mx = mmax(R(EPS), t.pr1);
  const R a_pr1 = a_raw * p.wpcl * (R(0.5) / msqrt_1ulp(mx)) * dmax_x(t.pr1, R(EPS));
  const R a_sum = a_pr1 * dmin_x(R(86.4) * (a.precnv + a.precls), R(p.pmaxcl)) * R(86.4);
  g.precnv += a_sum;
  g.precls += a_sum;
  const R a_cmax = a_raw * R(2.0) * t.cl2 * dmin_x(t.cmax * u.rrcl, R(1.0)) * u.rrcl;
  // jnp.max over levels 2..KX-3: the gradient is shared evenly by the tied entries
  int nt = 0;
#pragma unroll
  for (int k = 2; k <= KX - 3; ++k) nt += (t.cu[k] == t.cmax) ? 1 : 0;
  const R share = a_cmax / R(nt > 0 ? nt : 1);
  R a_cc0 = R(0.0);
#pragma unroll
  for (int k = 2; k <= KX - 3; ++k) {
    if (t.cu[k] == t.cmax) {
      if ((t.mk >> k) & 1u) g.rh[k] += share;
      else a_cc0 += share;
    }
  }
  if (t.m0) g.rh[KX - 2] += a_cc0;
}

// ---------------------------------------------------------------- condensation
template <typename R, int KX>
SPD_HD void condensation_adj(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const Uniforms<R>& u,
                             const ColumnCt<R, KX>& ct, ColumnAdj<R, KX>& g) {
  const R ps = c.ps;
  const R psa2 = ps * ps;
  R qlsc[KX];
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    const R dqa = u.rhref[k] * a.qsat[k] - c.Q[k];
    qlsc[k] = dqa < R(0.0) ? dqa * u.rtlsc : R(0.0);
  }
  R sum = R(0.0);
#pragma unroll
  for (int k = 1; k < KX; ++k) sum = (k == 1) ? u.pfact[k] * qlsc[k] : sum + u.pfact[k] * qlsc[k];
  g.ps += g.precls * (R(0.0) - sum);
  const R a_sum = -g.precls * ps;
  R a_psa2 = R(0.0);
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    const R dqa = u.rhref[k] * a.qsat[k] - c.Q[k];
    if (dqa < R(0.0)) {
      const R dq = qlsc[k];
      const R cap = u.dqmax[k] * psa2;
      const R a_m = ct.dt[k] * R(TFACT);
      const R a_dq = ct.dq[k] + a_sum * u.pfact[k] - a_m * dmin_x(-dq, cap);
      a_psa2 += a_m * dmin_x(cap, -dq) * u.dqmax[k];
      const R a_dqa = a_dq * u.rtlsc;
      g.qsat[k] += a_dqa * u.rhref[k];
      g.Q[k] += -a_dqa;
      g.prm[VJP_RHLSC] += a_dqa * a.qsat[k] * u.rhref_d[k];
    }
  }
  g.ps += a_psa2 * R(2.0) * ps;
}

// ---------------------------------------------------------------- convection
template <typename R, int KX>
SPD_HD void convection_adj(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyParams& p,
                           const SpeedyVertical& v, const Uniforms<R>& u, const ConvTape<R, KX>& t,
                           const ColumnCt<R, KX>& ct, ColumnAdj<R, KX>& g) {
  const int iptop = a.iptop_conv;
  const R ps = c.ps;
  const R rps = R(1.0) / ps;
  R a_dfse[KX], a_dfqa[KX];
  a_dfse[0] = ct.dt[0];
  a_dfqa[0] = ct.dq[0];
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    a_dfse[k] = ct.dt[k] * rps * v.grdscp[k];
    a_dfqa[k] = ct.dq[k] * rps * v.grdsig[k];
    g.rps += ct.dt[k] * t.dfse[k] * v.grdscp[k] + ct.dq[k] * t.dfqa[k] * v.grdsig[k];
  }
  // top layer: dfse[iptop-1] = top_s, dfqa[iptop-1] = top_q overwrite the fluxes
  R a_top_s = R(0.0), a_top_q = R(0.0);
#pragma unroll
  for (int k = 0; k < KX; ++k)
    if (iptop - 1 == k) {
      a_top_s = a_dfse[k]; a_top_q = a_dfqa[k];
      a_dfse[k] = R(0.0); a_dfqa[k] = R(0.0);
    }
  R a_fm_t = R(0.0), a_fus_t = a_top_s, a_fds_t = -a_top_s, a_fuq_t = a_top_q, a_fdq_t = -a_top_q;
  {
    const R a_pr = g.precnv + R(ALHC) * a_top_s - a_top_q;
    const R P = t.fuq_t - t.fm_t * t.qsatb;
    const R a_P = a_pr * dmax_x(P, R(0.0));
    a_fuq_t += a_P;
    a_fm_t += -a_P * t.qsatb;
    const R a_qsatb = -a_P * t.fm_t;
#pragma unroll
    for (int k = 0; k < KX - 1; ++k)
      if (iptop - 1 == k) {
        g.qsat[k] += a_qsatb * (R(1.0) - u.wvi1[k]);
        g.qsat[k + 1] += a_qsatb * u.wvi1[k];
      }
  }
  R a_FM[KX], a_FUS[KX], a_FUQ[KX], a_FDS[KX], a_FDQ[KX];
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    const bool top = iptop == k;
    a_FM[k] = top ? a_fm_t : R(0.0);
    a_FUS[k] = top ? a_fus_t : R(0.0);
    a_FUQ[k] = top ? a_fuq_t : R(0.0);
    a_FDS[k] = top ? a_fds_t : R(0.0);
    a_FDQ[k] = top ? a_fdq_t : R(0.0);
  }
  // secondary moisture flux and the flux differences
  R a_cbmf = R(0.0);
#pragma unroll
  for (int k = 0; k < KX - 1; ++k) {
    const R a_fsq = a_dfqa[k] - a_dfqa[KX - 1];
    if (t.delq[k] > R(0.0)) {
      a_cbmf += a_fsq * p.smf * t.delq[k];
      const R a_delq = a_fsq * p.smf * a.cbmf;
      g.qsat[k] += a_delq * p.rhil;
      g.Q[k] += -a_delq;
    }
    if (k >= iptop) {
      a_FUS[k + 1] += a_dfse[k]; a_FDS[k + 1] += -a_dfse[k];
      a_FUS[k] += -a_dfse[k]; a_FDS[k] += a_dfse[k];
      a_FUQ[k + 1] += a_dfqa[k]; a_FDQ[k + 1] += -a_dfqa[k];
      a_FUQ[k] += -a_dfqa[k]; a_FDQ[k] += a_dfqa[k];
    }
  }
  R a_fds = a_dfse[KX - 1] + a_FDS[KX - 1], a_fus = -a_dfse[KX - 1];
  R a_fdq = a_dfqa[KX - 1] + a_FDQ[KX - 1], a_fuq = -a_dfqa[KX - 1];
  R a_sbh[KX], a_qbh[KX];
  R a_fmass = R(0.0);
  a_sbh[0] = R(0.0); a_qbh[0] = R(0.0);
#pragma unroll
  for (int k = 0; k < KX - 1; ++k) {
    a_FM[k] += a_FDS[k] * t.sbh[k] + a_FDQ[k] * t.qbh[k];
    a_sbh[k] = a_FDS[k] * t.FM[k];
    a_qbh[k] = a_FDQ[k] * t.FM[k];
  }
  a_sbh[KX - 1] = R(0.0); a_qbh[KX - 1] = R(0.0);
  // cumulative sums from the bottom: FM[k] = fmass + sum_{j>=k} en[j], ...
  {
    R pm = R(0.0), pS = R(0.0), pQ = R(0.0);
#pragma unroll
    for (int j = 0; j < KX; ++j) {
      pm = pm + a_FM[j];
      pS = pS + a_FUS[j];
      pQ = pQ + a_FUQ[j];
      const R a_en = pm + pS * a.se[j] + pQ * c.Q[j];
      g.se[j] += pS * t.en[j];
      g.Q[j] += pQ * t.en[j];
      const bool lm = (j <= KX - 2) && (j >= iptop) && (j >= 1);
      if (lm) {
        g.ps += a_en * u.entr[j] * a.cbmf;
        a_cbmf += a_en * u.entr[j] * ps;
      }
    }
    a_fmass += pm;
    a_fus += pS;
    a_fuq += pQ;
  }
  // fus = fmass*se[KX-1], fuq = fmass*qmax, fds = fmass*sb, fdq = fmass*qb
  a_fmass += ((a_fus * a.se[KX - 1] + a_fuq * t.qmax) + a_fds * t.sb) + a_fdq * t.qb;
  g.se[KX - 1] += a_fus * t.fmass;
  R a_qmax = a_fuq * t.fmass;
  const R a_sb = a_fds * t.fmass;
  R a_qb = a_fdq * t.fmass;
  if (iptop < KX) a_fmass += a_cbmf;
  // fmass = fm0 * fpsa * min(5, qdif / (qmax - qb))
  R a_qdif;
  {
    const R den = t.qmax - t.qb;
    const R ratio = t.qdif / den;
    const R m = mmin(R(5.0), ratio);
    const R a_fpsa = a_fmass * u.fm0 * m;
    const R a_fm0 = a_fmass * t.fpsa * m;
    const R a_ratio = a_fmass * u.fm0 * t.fpsa * dmin_x(ratio, R(5.0));
    a_qdif = a_ratio / den;
    const R a_den = -a_ratio * ratio / den;
    a_qmax += a_den;
    a_qb += -a_den;
    g.prm[VJP_TRCNV] += a_fm0 * (-(u.fm0 / p.trcnv));
    const R y = (ps - p.psmin) * u.rdps;
    g.ps += a_fpsa * mmin(R(1.0), y) + a_fpsa * ps * dmin_x(y, R(1.0)) * u.rdps;
  }
  // qb = min(qbh[KX-1], Q[KX-1]); sb = sbh[KX-1]
  a_qbh[KX - 1] += a_qb * dmin_x(t.qbh[KX - 1], c.Q[KX - 1]);
  g.Q[KX - 1] += a_qb * dmin_x(c.Q[KX - 1], t.qbh[KX - 1]);
  a_sbh[KX - 1] += a_sb;
#pragma unroll
  for (int k = 1; k < KX; ++k) {
    const R w = u.wvi1[k - 1];
    g.se[k - 1] += a_sbh[k] * (R(1.0) - w);
    g.se[k] += a_sbh[k] * w;
    g.Q[k - 1] += a_qbh[k] * (R(1.0) - w);
    g.Q[k] += a_qbh[k] * w;
  }
  // qmax = max(1.01 Q[KX-1], qsat[KX-1])
  {
    const R q1 = R(1.01) * c.Q[KX - 1];
    g.Q[KX - 1] += a_qmax * R(1.01) * dmax_x(q1, a.qsat[KX - 1]);
    g.qsat[KX - 1] += a_qmax * dmax_x(a.qsat[KX - 1], q1);
  }
  // qdif (diagnose_convection)
  R a_qthr0 = R(0.0), a_mse0 = R(0.0), a_msthr = R(0.0);
  if (t.case1) {
    const R A = c.Q[KX - 1] - t.qthr0;
    const R B = (t.mse0 - t.msthr) * R(RLHC);
    const R a_A = a_qdif * dmax_x(A, B), a_B = a_qdif * dmax_x(B, A);
    g.Q[KX - 1] += a_A;
    a_qthr0 = -a_A;
    a_mse0 = a_B * R(RLHC);
    a_msthr = -a_B * R(RLHC);
  } else if (t.case2) {
    g.Q[KX - 1] += a_qdif;
    a_qthr0 = -a_qdif;
  }
  g.prm[VJP_RHBL] += a_qthr0 * a.qsat[KX - 1];
  g.qsat[KX - 1] += a_qthr0 * p.rhbl;
  g.se[KX - 1] += a_mse0;
  g.Q[KX - 1] += a_mse0 * R(ALHC);
#pragma unroll
  for (int k = 2; k <= KX - 4; ++k)
    if (t.ktop2 == k + 1) {   // msthr = mss2[k], mss = se + alhc * qsat
      const R w = u.wvi1[k];
      const R am0 = a_msthr * (R(1.0) - w), am1 = a_msthr * w;
      g.se[k] += am0;
      g.qsat[k] += am0 * R(ALHC);
      g.se[k + 1] += am1;
      g.qsat[k + 1] += am1 * R(ALHC);
    }
}

// humidity, static energy and 1/ps back to the state.
template <typename R, int KX>
SPD_HD void humidity_adj(const ColumnIn<R, KX>& c, const PhaseA<R, KX>& a, const SpeedyVertical& v,
                         ColumnAdj<R, KX>& g) {
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    const R rq = R(1.0) / a.qsat[k];
    g.Q[k] += g.rh[k] * rq;
    const R a_qs = g.qsat[k] - g.rh[k] * a.rh[k] * rq;
    R dqt, dqp;
    qsat_grad(c.T[k], c.ps, R(v.fsg[k]), dqt, dqp);
    g.T[k] += a_qs * dqt + R(CP) * g.se[k];
    g.ps += a_qs * dqp;
  }
  const R rps = R(1.0) / c.ps;
  g.ps += -g.rps * rps * rps;
}

// Per-column storage for the longwave flux tapes between a term's recompute and
// its reversal: a plain array on the host; LDS on the GPU (speedy_kernels.hip).
template <typename R, int N>
struct ArrayStore {
  R v[N];
  SPD_HD void put(int i, R x) { v[i] = x; }
  SPD_HD R get(int i) const { return v[i]; }
  SPD_HD const ArrayStore& after(R) const { return *this; }
};

// A copy of the column whose values the optimizer cannot see through: a term
// recomputed from it is not merged with the same computation elsewhere in the
// kernel (which would keep that term's intermediates live in between).
template <typename R, int KX>
SPD_HD ColumnIn<R, KX> opaque_column(const ColumnIn<R, KX>& c_in) {
  ColumnIn<R, KX> c = c_in;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    asm volatile("" : "+v"(c.T[k]));
    asm volatile("" : "+v"(c.Q[k]));
    asm volatile("" : "+v"(c.PHI[k]));
  }
  asm volatile("" : "+v"(c.ps));
#endif
  return c;
}

// The whole column: given the forward's phase outputs (a: phase A; b: phase B's
// scalars, the longwave-down divergence in b.dfl and the hand-offs flux_h, rlus2,
// tsfc) and the cotangents, accumulate the gradients into g.  The longwave
// transmissivities and emissions are recomputed where the longwave terms are
// reversed, from their own opaque copies of the column, rather than read from b.
template <typename R, int KX, class Store>
SPD_HD void column_vjp(const ColumnIn<R, KX>& c_in, const PhaseA<R, KX>& a, const PhaseB<R, KX>& b, uint32_t S,
                       const SpeedyParams& p, const SpeedyVertical& v, const Uniforms<R>& u, uint32_t flags,
                       const ColumnCt<R, KX>& ct, ColumnAdj<R, KX>& g, Store& tape) {
  adj_zero<R, KX>(g);
  // The terms below recompute the forward from the column; none of the forward's
  // intermediates stay live across the adjoint (they are equal by construction).
  const ColumnIn<R, KX> c = opaque_column<R, KX>(c_in);
  const R rps = R(1.0) / c.ps;
  R a_dfl[KX], a_Th[KX];
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    a_dfl[k] = ct.dt[k] * v.grdscp[k] * rps;
    a_Th[k] = R(0.0);
  }
  R a_rlus2 = R(0.0), a_flux_h[2] = {R(0.0), R(0.0)}, a_rsds = R(0.0), a_rlds = R(0.0);
  {
    const ColumnIn<R, KX> cl = opaque_column<R, KX>(c);
    PhaseB<R, KX> bb;
    LwdTape<R, KX> et;
    lw_transmissivity<R, KX>(cl, a, p, v, u, bb);
    lw_emissions<R, KX>(cl, u, bb, et);
#pragma unroll
    for (int k = 0; k < KX; ++k) bb.dfl[k] = b.dfl[k];
    bb.rlds = b.rlds;
    LwuTape<R, KX> lt;
    const R fh[4] = {b.flux_h[0], b.flux_h[1], R(0.0), R(0.0)};
    lw_up<R, KX, false>(cl, a, p, v, u, bb, b.rlus2, b.tsfc, fh, lt);
#pragma unroll
    for (int k = 0; k < KX; ++k) g.rps += ct.dt[k] * lt.dfl[k] * v.grdscp[k];
#pragma unroll
    for (int k = 1; k <= KX; ++k)
#pragma unroll
      for (int bnd = 0; bnd < 4; ++bnd) tape.put(k * 4 + bnd, lt.Fu[k][bnd]);
  }
  {
    const ColumnIn<R, KX> ca = opaque_column<R, KX>(c);
    lw_up_adj<R, KX>(ca, a, b, p, v, u, tape, a_dfl, a_rlus2, a_flux_h, g, a_Th);
    settle_adj<R, KX>(g);
    if constexpr (kSettle<KX>) { settle(a_rlus2); settle(a_flux_h[0]); settle(a_flux_h[1]); }
  }
  {
    PhaseB<R, KX> bb = b;
    SfcTape<R> st;
    R tsfc;
    surface<R, KX, false>(c, a, p, v, u, flags, bb, tsfc, st);
    surface_adj<R, KX>(c, a, b, p, v, u, st, ct, a_rlus2, a_rsds, a_rlds, g);
    settle_adj<R, KX>(g);
    if constexpr (kSettle<KX>) { settle(a_rsds); settle(a_rlds); }
  }
  {
    PhaseB<R, KX> bb = b;
    SwTape<R, KX> st;
    R fh[4];
    shortwave<R, KX, false>(c, a, S, p, v, u, bb, fh, st);
    shortwave_adj<R, KX>(c, a, b, S, p, v, u, st, ct, a_flux_h, a_rsds, g);
    settle_adj<R, KX>(g);
  }
  {
    const ColumnIn<R, KX> cl = opaque_column<R, KX>(c);
    PhaseB<R, KX> bb;
    LwdTape<R, KX> lt;
    lw_transmissivity<R, KX>(cl, a, p, v, u, bb);
    lw_down<R, KX>(cl, p, u, bb, lt);
#pragma unroll
    for (int k = 0; k < KX - 1; ++k)
#pragma unroll
      for (int bnd = 0; bnd < 4; ++bnd) tape.put(k * 4 + bnd, lt.Fd[k][bnd]);
  }
  {
    const ColumnIn<R, KX> ca = opaque_column<R, KX>(c);
    lw_down_adj<R, KX>(ca, a, p, v, u, tape, a_dfl, a_rlds, g, a_Th);
    settle_adj<R, KX>(g);
  }
  half_level_adj<R, KX>(u, a_Th, g);
  vdiff_adj<R, KX>(c, a, p, u, ct, g);
  settle_adj<R, KX>(g);
  {
    PhaseA<R, KX> aa = a;
    CloudTape<R, KX> t;
    clouds<R, KX>(c, p, u, aa, t);
    clouds_adj<R, KX>(c, a, p, u, t, g);
    settle_adj<R, KX>(g);
  }
  condensation_adj<R, KX>(c, a, u, ct, g);
  settle_adj<R, KX>(g);
  {
    PhaseA<R, KX> aa = a;
    ConvTape<R, KX> t;
    convection<R, KX>(c, p, v, u, aa, t);
    convection_adj<R, KX>(c, a, p, v, u, t, ct, g);
  }
  humidity_adj<R, KX>(c, a, v, g);
}

}  // namespace speedy
