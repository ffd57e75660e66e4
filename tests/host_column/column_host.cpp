// column_host.cpp -- TEST HARNESS ONLY: the kernel's column routine
// (jax-gcm_amd/csrc/speedy_column.hpp) compiled for the host with g++, so the
// CPU test suite can check the device code's logic against the oracle without
// a GPU.  Built twice: without OpenMP for the tests, and with -fopenmp (columns
// split over the host's cores) as the CPU baseline that bench.py times.  It is
// never loaded by the library or the product path.
#include <stdint.h>
#include <string.h>

#include "../../jax-gcm_amd/csrc/speedy_adjoint.hpp"

using namespace speedy;

template <typename R, int KX>
static void run(const R* const* in, R* const* out, const SpeedyVertical* v, const SpeedyParams* p,
                const SpeedyDate* d, int nlon, int nlat, uint32_t flags, uint32_t term_mask, int32_t level_set,
                uint32_t* s_out) {
  const int64_t ncol = (int64_t)nlon * nlat;
  // in: u v t q phi ps alb0 sice snowc soilw stl sst fmask orog phis0 sia coa
  const R *u = in[0], *vv = in[1], *t = in[2], *q = in[3], *phi = in[4], *ps = in[5];
  Uniforms<R> un;
  make_uniforms(*p, *v, *d, un);
  R lat[6][1024];
  for (int j = 0; j < nlat && j < 1024; ++j) {
    R z[6];
    zonal_column(un.zon, in[15][j], in[16][j], z);
    for (int i = 0; i < 6; ++i) lat[i][j] = z[i];
  }
  auto load = [&](int64_t c, ColumnIn<R, KX>& col) {
    for (int k = 0; k < KX; ++k) {
      col.T[k] = t[k * ncol + c]; col.Q[k] = q[k * ncol + c]; col.PHI[k] = phi[k * ncol + c];
      if (flags & SPEEDY_FLAG_VERIFY_STATE) col.Q[k] = verify_q(col.Q[k]);
    }
    col.u1 = u[(KX - 1) * ncol + c]; col.v1 = vv[(KX - 1) * ncol + c]; col.ps = ps[c];
    col.alb0 = in[6][c]; col.sice = in[7][c]; col.snowc_am = in[8][c]; col.soilw = in[9][c];
    col.stl = in[10][c]; col.sst = in[11][c]; col.fmask = in[12][c]; col.orog = in[13][c]; col.phis0 = in[14][c];
    const int j = (int)(c % nlat);
    col.fsol = lat[0][j]; col.ozupp = lat[1][j]; col.ozone = lat[2][j]; col.zenit = lat[3][j];
    col.stratz = lat[4][j]; col.sqcoa = lat[5][j];
  };
  uint32_t S = 0;
#pragma omp parallel for reduction(| : S) schedule(static)
  for (int64_t c = 0; c < ncol; ++c) {
    ColumnIn<R, KX> col;
    load(c, col);
    int icl = cloud_top<R, KX>(col, *p, *v, un);
    icl = icl < 0 ? 0 : (icl > KX - 1 ? KX - 1 : icl);
    S |= 1u << icl;
  }
  *s_out = S;
  if (level_set >= 0) S = (uint32_t)level_set;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < ncol; ++c) {
    ColumnIn<R, KX> col;
    load(c, col);
    PhaseA<R, KX> a;
    phase_a<R, KX>(col, *p, *v, un, a);
    PhaseB<R, KX> b;
    phase_b1<R, KX>(col, a, *p, *v, un, b);
    phase_b2<R, KX, true>(col, a, S, *p, *v, un, flags, b);
    for (int k = 0; k < KX; ++k) {
      R tt = 0, qq = 0, du = 0, dv = 0;
      if (term_mask & (1u << SPEEDY_TERM_CONVECTION)) { tt = tt + a.tconv[k]; qq = qq + a.qconv[k]; }
      if (term_mask & (1u << SPEEDY_TERM_CONDENSATION)) { tt = tt + a.tlsc[k]; qq = qq + a.qlsc[k]; }
      if (term_mask & (1u << SPEEDY_TERM_SHORTWAVE)) tt = tt + b.tsw[k];
      if (k == KX - 1 && (term_mask & (1u << SPEEDY_TERM_SURFACE))) {
        du = du + b.du1; dv = dv + b.dv1; tt = tt + b.dts1; qq = qq + b.dqs1;
      }
      if (term_mask & (1u << SPEEDY_TERM_LW_UP)) tt = tt + b.tlw[k];
      if (term_mask & (1u << SPEEDY_TERM_VDIFF)) { tt = tt + b.tvd[k]; qq = qq + b.qvd[k]; }
      if (flags & SPEEDY_FLAG_VERIFY_TENDENCIES) { bool lim; qq = verify_dq(q[k * ncol + c], qq, un.dt_step, lim); }
      out[0][k * ncol + c] = du; out[1][k * ncol + c] = dv; out[2][k * ncol + c] = tt; out[3][k * ncol + c] = qq;
    }
    // a few diagnostics: iptop, icltop, cbmf, precnv, precls, cloudc, clstr, rsds, rlds, tskin
    R* dg = out[4];
    const R vals[] = {(R)a.iptop, (R)a.icltop, a.cbmf, a.precnv, a.precls, a.cloudc, a.clstr,
                          b.rsds, b.rlds, b.tskin};
    for (int i = 0; i < 10; ++i) dg[i * ncol + c] = vals[i];
  }
}

// Forward + VJP of every column.  ct: ct_du ct_dv ct_dt ct_dq [kx][ncol];
// g: g_u g_v g_t g_q [kx][ncol], g_ps [ncol], g_prm [ncol][5] (per column).
template <typename R, int KX>
static void run_vjp(const R* const* in, const R* const* ct, R* const* out, R* const* g, const SpeedyVertical* v,
                    const SpeedyParams* p, const SpeedyDate* d, int nlon, int nlat, uint32_t flags,
                    int32_t level_set) {
  const int64_t ncol = (int64_t)nlon * nlat;
  const R *u = in[0], *vv = in[1], *t = in[2], *q = in[3], *phi = in[4], *ps = in[5];
  Uniforms<R> un;
  make_uniforms(*p, *v, *d, un);
  auto load = [&](int64_t c, ColumnIn<R, KX>& col) {
    for (int k = 0; k < KX; ++k) {
      col.T[k] = t[k * ncol + c]; col.Q[k] = q[k * ncol + c]; col.PHI[k] = phi[k * ncol + c];
      if (flags & SPEEDY_FLAG_VERIFY_STATE) col.Q[k] = verify_q(col.Q[k]);
    }
    col.u1 = u[(KX - 1) * ncol + c]; col.v1 = vv[(KX - 1) * ncol + c]; col.ps = ps[c];
    col.alb0 = in[6][c]; col.sice = in[7][c]; col.snowc_am = in[8][c]; col.soilw = in[9][c];
    col.stl = in[10][c]; col.sst = in[11][c]; col.fmask = in[12][c]; col.orog = in[13][c]; col.phis0 = in[14][c];
    const int j = (int)(c % nlat);
    R z[6];
    zonal_column(un.zon, in[15][j], in[16][j], z);
    col.fsol = z[0]; col.ozupp = z[1]; col.ozone = z[2]; col.zenit = z[3]; col.stratz = z[4]; col.sqcoa = z[5];
  };
  uint32_t S = 0;
#pragma omp parallel for reduction(| : S) schedule(static)
  for (int64_t c = 0; c < ncol; ++c) {
    ColumnIn<R, KX> col;
    load(c, col);
    int icl = cloud_top<R, KX>(col, *p, *v, un);
    icl = icl < 0 ? 0 : (icl > KX - 1 ? KX - 1 : icl);
    S |= 1u << icl;
  }
  if (level_set >= 0) S = (uint32_t)level_set;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < ncol; ++c) {
    ColumnIn<R, KX> col;
    load(c, col);
    PhaseA<R, KX> a;
    phase_a<R, KX>(col, *p, *v, un, a);
    PhaseB<R, KX> b;
    phase_b1<R, KX>(col, a, *p, *v, un, b);
    phase_b2<R, KX, false>(col, a, S, *p, *v, un, flags, b);
    uint32_t limited = 0;
    for (int k = 0; k < KX; ++k) {
      R tt = 0, qq = 0, du = 0, dv = 0;
      tt = tt + a.tconv[k]; qq = qq + a.qconv[k];
      tt = tt + a.tlsc[k]; qq = qq + a.qlsc[k];
      tt = tt + b.tsw[k];
      if (k == KX - 1) { du = du + b.du1; dv = dv + b.dv1; tt = tt + b.dts1; qq = qq + b.dqs1; }
      tt = tt + b.tlw[k];
      tt = tt + b.tvd[k]; qq = qq + b.qvd[k];
      if (flags & SPEEDY_FLAG_VERIFY_TENDENCIES) {
        bool lim;
        qq = verify_dq(q[k * ncol + c], qq, un.dt_step, lim);
        limited |= (lim ? 1u : 0u) << k;
      }
      out[0][k * ncol + c] = du; out[1][k * ncol + c] = dv; out[2][k * ncol + c] = tt; out[3][k * ncol + c] = qq;
    }
    ColumnCt<R, KX> cc;
    cc.du1 = ct[0][(KX - 1) * ncol + c];
    cc.dv1 = ct[1][(KX - 1) * ncol + c];
    for (int k = 0; k < KX; ++k) { cc.dt[k] = ct[2][k * ncol + c]; cc.dq[k] = ct[3][k * ncol + c]; }
    R g_qraw[KX];
    verify_ct<R, KX>(flags, limited, un.dt_step, cc, g_qraw);
    ColumnAdj<R, KX> ga;
    ArrayStore<R, vjp_tape_floats<KX>()> tape;
    column_vjp<R, KX>(col, a, b, S, *p, *v, un, flags, cc, ga, tape);
    for (int k = 0; k < KX; ++k) {
      if (flags & SPEEDY_FLAG_VERIFY_STATE) ga.Q[k] = verify_q_grad(q[k * ncol + c], ga.Q[k]);
      ga.Q[k] += g_qraw[k];
    }
    for (int k = 0; k < KX; ++k) {
      g[0][k * ncol + c] = k == KX - 1 ? ga.u1 : R(0);
      g[1][k * ncol + c] = k == KX - 1 ? ga.v1 : R(0);
      g[2][k * ncol + c] = ga.T[k];
      g[3][k * ncol + c] = ga.Q[k];
    }
    g[4][c] = ga.ps;
    for (int i = 0; i < VJP_NPARAM; ++i) g[5][c * VJP_NPARAM + i] = ga.prm[i];
  }
}

template <typename R>
static int dispatch_vjp(int kx, const R* const* in, const R* const* ct, R* const* out, R* const* g,
                        const SpeedyVertical* v, const SpeedyParams* p, const SpeedyDate* d, int nlon, int nlat,
                        uint32_t flags, int32_t level_set) {
  switch (kx) {
    case 7: run_vjp<R, 7>(in, ct, out, g, v, p, d, nlon, nlat, flags, level_set); return 0;
    case 8: run_vjp<R, 8>(in, ct, out, g, v, p, d, nlon, nlat, flags, level_set); return 0;
    case 16: run_vjp<R, 16>(in, ct, out, g, v, p, d, nlon, nlat, flags, level_set); return 0;
    default: return 2;
  }
}

extern "C" int column_host_vjp(int kx, const float* const* in, const float* const* ct, float* const* out,
                               float* const* g, const SpeedyVertical* v, const SpeedyParams* p, const SpeedyDate* d,
                               int nlon, int nlat, uint32_t flags, int32_t level_set) {
  return dispatch_vjp<float>(kx, in, ct, out, g, v, p, d, nlon, nlat, flags, level_set);
}

extern "C" int column_host_vjp64(int kx, const double* const* in, const double* const* ct, double* const* out,
                                 double* const* g, const SpeedyVertical* v, const SpeedyParams* p,
                                 const SpeedyDate* d, int nlon, int nlat, uint32_t flags, int32_t level_set) {
  return dispatch_vjp<double>(kx, in, ct, out, g, v, p, d, nlon, nlat, flags, level_set);
}

template <typename R>
static int dispatch(int kx, const R* const* in, R* const* out, const SpeedyVertical* v, const SpeedyParams* p,
                    const SpeedyDate* d, int nlon, int nlat, uint32_t flags, uint32_t term_mask, int32_t level_set,
                    uint32_t* s_out) {
  if (nlat > 1024) return 1;
  switch (kx) {
    case 7: run<R, 7>(in, out, v, p, d, nlon, nlat, flags, term_mask, level_set, s_out); return 0;
    case 8: run<R, 8>(in, out, v, p, d, nlon, nlat, flags, term_mask, level_set, s_out); return 0;
    case 16: run<R, 16>(in, out, v, p, d, nlon, nlat, flags, term_mask, level_set, s_out); return 0;
    default: return 2;
  }
}

// in: u v t q phi ps alb0 sice snowc soilw stl sst fmask orog phis0 sia coa;
// out: du dv dt dq, then 10 diagnostic rows.  float32 and float64 builds.
extern "C" int column_host_fwd(int kx, const float* const* in, float* const* out, const SpeedyVertical* v,
                               const SpeedyParams* p, const SpeedyDate* d, int nlon, int nlat, uint32_t flags,
                               uint32_t term_mask, int32_t level_set, uint32_t* s_out) {
  return dispatch<float>(kx, in, out, v, p, d, nlon, nlat, flags, term_mask, level_set, s_out);
}

extern "C" int column_host_fwd64(int kx, const double* const* in, double* const* out, const SpeedyVertical* v,
                                 const SpeedyParams* p, const SpeedyDate* d, int nlon, int nlat, uint32_t flags,
                                 uint32_t term_mask, int32_t level_set, uint32_t* s_out) {
  return dispatch<double>(kx, in, out, v, p, d, nlon, nlat, flags, term_mask, level_set, s_out);
}
