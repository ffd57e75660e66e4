// speedy_kernels.hip -- gfx950 kernels and the C ABI (include/speedy_abi.h) of the
// SPEEDY column physics.
//
// One call of the physics is two launches on the caller's stream:
//   1. cloudtop_kernel: per-latitude insolation/ozone tables, and the Q1 level
//      set S of every member (an OR over all of the member's columns of
//      1 << clip(icltop, 0, kx-1); shortwave_radiation.py:63-70).  The last
//      block to finish publishes S and clears the accumulators, so no memset
//      is needed between calls.
//   2. fwd_kernel: one thread per column, the whole column held in registers;
//      phase A (humidity, convection, condensation, clouds) is recomputed
//      rather than stored, phase B (shortwave with S, longwave, surface fluxes,
//      vertical diffusion) follows, the term-ordered tendency sum is written.
// Reads/writes are [field][level][column] with consecutive threads on
// consecutive columns, so every access is a coalesced 256-byte wave load.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "speedy_column.hpp"

namespace speedy {

constexpr int kBlock = 256;

struct KArgs {
  SpeedyFields f;
  SpeedyVertical v;
  SpeedyParams p;
  SpeedyDate d;
  int32_t nlon, nlat, nmember, level_set;
  int64_t ncol, ntot;
  uint32_t flags, term_mask;
  uint32_t *counter, *s_accum, *s_final;
  float *lat;   // [6][nlat]
};

// ---------------------------------------------------------------- diagnostics
// Names follow Physics.data_struct_to_dict (physics_interface.py:157-194):
// "<struct>.<field>" and "<struct>.<field>.<channel>" for multi-channel fields.
static const char* kDiag2D[] = {
    "convection.iptop", "convection.cbmf", "convection.precnv", "condensation.precls",
    "shortwave_rad.icltop", "shortwave_rad.cloudc", "shortwave_rad.cloudstr", "shortwave_rad.qcloud",
    "shortwave_rad.gse", "shortwave_rad.rsds", "shortwave_rad.rsns", "shortwave_rad.ftop",
    "shortwave_rad.fsol", "shortwave_rad.ozupp", "shortwave_rad.ozone", "shortwave_rad.zenit",
    "shortwave_rad.stratz", "longwave_rad.ftop", "mod_radcon.albsfc", "mod_radcon.alb_l",
    "mod_radcon.alb_s", "mod_radcon.snowc", "mod_radcon.stratc.0", "mod_radcon.stratc.1",
    "mod_radcon.flux.0", "mod_radcon.flux.1", "mod_radcon.flux.2", "mod_radcon.flux.3",
    "surface_flux.ustr.0", "surface_flux.ustr.1", "surface_flux.ustr.2",
    "surface_flux.vstr.0", "surface_flux.vstr.1", "surface_flux.vstr.2",
    "surface_flux.shf.0", "surface_flux.shf.1", "surface_flux.shf.2",
    "surface_flux.evap.0", "surface_flux.evap.1", "surface_flux.evap.2",
    "surface_flux.rlus.0", "surface_flux.rlus.1", "surface_flux.rlus.2",
    "surface_flux.rlds", "surface_flux.rlns", "surface_flux.hfluxn.0", "surface_flux.hfluxn.1",
    "surface_flux.tsfc", "surface_flux.tskin", "surface_flux.u0", "surface_flux.v0", "surface_flux.t0"};
constexpr int kN2D = sizeof(kDiag2D) / sizeof(kDiag2D[0]);
static const char* kDiag3D[] = {
    "humidity.qsat", "humidity.rh", "convection.se", "shortwave_rad.dfabs", "longwave_rad.dfabs",
    "mod_radcon.st4a.0", "mod_radcon.st4a.1", "mod_radcon.tau2.0", "mod_radcon.tau2.1",
    "mod_radcon.tau2.2", "mod_radcon.tau2.3"};
constexpr int kN3D = sizeof(kDiag3D) / sizeof(kDiag3D[0]);

constexpr int diag_rows(int kx) { return kN2D + kN3D * kx; }

// ---------------------------------------------------------------- loads
template <int KX>
__device__ __forceinline__ void load_column(const KArgs& a, int64_t m, int64_t c, ColumnIn<KX>& col) {
  const int64_t ncol = a.ncol;
  const int64_t base3 = m * KX * ncol + c;
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    col.T[k] = a.f.t[base3 + k * ncol];
    col.Q[k] = a.f.q[base3 + k * ncol];
    col.PHI[k] = a.f.phi[base3 + k * ncol];
  }
  col.u1 = a.f.u[base3 + (KX - 1) * ncol];
  col.v1 = a.f.v[base3 + (KX - 1) * ncol];
  col.ps = a.f.ps[m * ncol + c];
  col.alb0 = a.f.alb0[c];
  col.sice = a.f.sice_am[c];
  col.snowc_am = a.f.snowc_am[c];
  col.soilw = a.f.soilw_am[c];
  col.stl = a.f.stl_am[c];
  col.sst = a.f.sst[c];
  col.fmask = a.f.fmask[c];
  col.orog = a.f.orog[c];
  col.phis0 = a.f.phis0[c];
  const int64_t j = c % a.nlat;
  const int64_t nl = a.nlat;
  col.fsol = a.lat[0 * nl + j];
  col.ozupp = a.lat[1 * nl + j];
  col.ozone = a.lat[2 * nl + j];
  col.zenit = a.lat[3 * nl + j];
  col.stratz = a.lat[4 * nl + j];
  col.sqcoa = a.lat[5 * nl + j];
}

// ---------------------------------------------------------------- kernels
template <int KX>
__global__ __launch_bounds__(kBlock) void cloudtop_kernel(KArgs a) {
  __shared__ uint32_t s_or[kBlock + 1];
  __shared__ int s_last;
  const int tid = threadIdx.x;
  const int64_t gid = (int64_t)blockIdx.x * kBlock + tid;
  if (gid < a.nlat) {
    float z[6];
    zonal_fields(a.d.tyear, a.f.sia[gid], a.f.coa[gid], z);
#pragma unroll
    for (int i = 0; i < 6; ++i) a.lat[i * (int64_t)a.nlat + gid] = z[i];
  }
  if (a.level_set >= 0) return;   // uniform over the grid: S is given
  const int64_t blk0 = (int64_t)blockIdx.x * kBlock;
  const int64_t blk_last = min(blk0 + kBlock, a.ntot) - 1;
  const int64_t m_first = blk0 / a.ncol;
  const int nm = (int)(blk_last / a.ncol - m_first + 1);
  for (int i = tid; i < nm; i += kBlock) s_or[i] = 0u;
  __syncthreads();
  if (gid < a.ntot) {
    const int64_t m = gid / a.ncol;
    const int64_t c = gid - m * a.ncol;
    ColumnIn<KX> col;
    const int64_t ncol = a.ncol;
    const int64_t base3 = m * KX * ncol + c;
#pragma unroll
    for (int k = 1; k < KX; ++k) {
      col.T[k] = a.f.t[base3 + k * ncol];
      col.Q[k] = a.f.q[base3 + k * ncol];
      col.PHI[k] = a.f.phi[base3 + k * ncol];
    }
    col.ps = a.f.ps[m * ncol + c];
    Uniforms<KX> u;
    make_uniforms<KX>(a.p, a.v, a.d, u);
    const int icl = cloud_top<KX>(col, a.p, a.v, u);
    const int lv = icl < 0 ? 0 : (icl > KX - 1 ? KX - 1 : icl);
    atomicOr(&s_or[m - m_first], 1u << lv);
  }
  __syncthreads();
  for (int i = tid; i < nm; i += kBlock)
    if (s_or[i]) atomicOr(&a.s_accum[m_first + i], s_or[i]);
  __threadfence();
  __syncthreads();
  if (tid == 0) {
    const unsigned prev = atomicInc(a.counter, gridDim.x - 1);   // wraps to 0 for the next call
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (s_last) {
    __threadfence();
    for (int m = tid; m < a.nmember; m += kBlock) a.s_final[m] = atomicExch(&a.s_accum[m], 0u);
  }
}

template <int KX, bool DIAG>
__global__ __launch_bounds__(kBlock) void fwd_kernel(KArgs a) {
  const int64_t gid = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (gid >= a.ntot) return;
  const int64_t m = gid / a.ncol;
  const int64_t c = gid - m * a.ncol;
  const int64_t ncol = a.ncol;
  ColumnIn<KX> col;
  load_column<KX>(a, m, c, col);
  Uniforms<KX> u;
  make_uniforms<KX>(a.p, a.v, a.d, u);
  PhaseA<KX> pa;
  phase_a<KX>(col, a.p, a.v, u, pa);
  const uint32_t S = a.level_set >= 0 ? (uint32_t)a.level_set : a.s_final[m];
  PhaseB<KX> pb;
  phase_b<KX, DIAG>(col, pa, S, a.p, a.v, u, a.flags, pb);
  // term-ordered sum (speedy_physics.py:123-125)
  const uint32_t tm = a.term_mask;
  const int64_t base3 = m * KX * ncol + c;
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    float t = 0.0f, q = 0.0f, du = 0.0f, dv = 0.0f;
    if (tm & (1u << SPEEDY_TERM_CONVECTION)) { t = t + pa.tconv[k]; q = q + pa.qconv[k]; }
    if (tm & (1u << SPEEDY_TERM_CONDENSATION)) { t = t + pa.tlsc[k]; q = q + pa.qlsc[k]; }
    if (tm & (1u << SPEEDY_TERM_SHORTWAVE)) t = t + pb.tsw[k];
    if (k == KX - 1 && (tm & (1u << SPEEDY_TERM_SURFACE))) {
      du = du + pb.du1; dv = dv + pb.dv1; t = t + pb.dts1; q = q + pb.dqs1;
    }
    if (tm & (1u << SPEEDY_TERM_LW_UP)) t = t + pb.tlw[k];
    if (tm & (1u << SPEEDY_TERM_VDIFF)) { t = t + pb.tvd[k]; q = q + pb.qvd[k]; }
    a.f.du[base3 + k * ncol] = du;
    a.f.dv[base3 + k * ncol] = dv;
    a.f.dt[base3 + k * ncol] = t;
    a.f.dq[base3 + k * ncol] = q;
  }
  if (DIAG) {
    float* dg = a.f.diag + m * (int64_t)diag_rows(KX) * ncol + c;
    int r = 0;
    auto W = [&](float x) { dg[(int64_t)(r++) * ncol] = x; };
    W((float)pa.iptop); W(pa.cbmf); W(pa.precnv); W(pa.precls);
    W((float)pa.icltop); W(pa.cloudc); W(pa.clstr); W(pa.qcloud); W(pa.gse);
    W(pb.rsds); W(pb.rsns); W(pb.ftop_sw);
    W(col.fsol); W(col.ozupp); W(col.ozone); W(col.zenit); W(col.stratz);
    W(pb.ftop_lw); W(pb.albsfc); W(pb.alb_l); W(pb.alb_s); W(pb.snowc);
    W(pb.stratc0); W(pb.stratc1);
    for (int i = 0; i < 4; ++i) W(pb.flux_out[i]);
    for (int i = 0; i < 3; ++i) W(pb.ustr[i]);
    for (int i = 0; i < 3; ++i) W(pb.vstr[i]);
    for (int i = 0; i < 3; ++i) W(pb.shf[i]);
    for (int i = 0; i < 3; ++i) W(pb.evap[i]);
    for (int i = 0; i < 3; ++i) W(pb.rlus[i]);
    W(pb.rlds); W(pb.rlns); W(pb.hfluxn[0]); W(pb.hfluxn[1]);
    W(pb.tsfc); W(pb.tskin); W(pb.u0); W(pb.v0); W(pb.t0);
#pragma unroll
    for (int k = 0; k < KX; ++k) W(pa.qsat[k]);
#pragma unroll
    for (int k = 0; k < KX; ++k) W(pa.rh[k]);
#pragma unroll
    for (int k = 0; k < KX; ++k) W(pa.se[k]);
#pragma unroll
    for (int k = 0; k < KX; ++k) W(pb.dfabs_sw[k]);
#pragma unroll
    for (int k = 0; k < KX; ++k) W(pb.dfabs_lw[k]);
#pragma unroll
    for (int k = 0; k < KX; ++k) W(pb.st4a0[k]);
#pragma unroll
    for (int k = 0; k < KX; ++k) W(pb.st4a1[k]);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int k = 0; k < KX; ++k) W(pb.tau[k][b]);
  }
}

__global__ void fill_zero_kernel(uint32_t* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------- host side
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// ---------------------------------------------------------------- kernel timers
struct Timer {
  std::mutex mu;
  bool on = false;
  int cap = 0, used = 0;
  std::vector<hipEvent_t> ev;   // begin/end per slot
  std::vector<int> kind;
} g_timer;

int timer_begin(hipStream_t s, int kind) {
  std::lock_guard<std::mutex> lk(g_timer.mu);
  if (!g_timer.on || g_timer.used >= g_timer.cap) return -1;
  const int slot = g_timer.used++;
  g_timer.kind[slot] = kind;
  (void)hipEventRecord(g_timer.ev[2 * slot], s);
  return slot;
}

void timer_end(hipStream_t s, int slot) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_timer.mu);
  (void)hipEventRecord(g_timer.ev[2 * slot + 1], s);
}

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct WsLayout {
  size_t counter, s_accum, s_final, lat, total;
};

WsLayout ws_layout(int32_t nlat, int32_t nmember) {
  WsLayout L;
  L.counter = 0;
  L.s_accum = 256;
  L.s_final = L.s_accum + align256(sizeof(uint32_t) * (size_t)nmember);
  L.lat = L.s_final + align256(sizeof(uint32_t) * (size_t)nmember);
  L.total = L.lat + align256(sizeof(float) * 6 * (size_t)nlat);
  return L;
}

int check_fields(const SpeedyFields* f) {
  const void* req[] = {f->u, f->v, f->t, f->q, f->phi, f->ps, f->alb0, f->sice_am, f->snowc_am,
                       f->soilw_am, f->stl_am, f->sst, f->fmask, f->orog, f->phis0, f->sia, f->coa,
                       f->du, f->dv, f->dt, f->dq};
  for (const void* p : req)
    if (!p) return fail(SPEEDY_ERR_ARG, "speedy: a required field pointer is NULL");
  return SPEEDY_OK;
}

int prepare(const SpeedyFields* fields, const SpeedyVertical* vert, const SpeedyParams* params,
            const SpeedyDate* date, int32_t nlon, int32_t nlat, int32_t nmember, uint32_t flags,
            uint32_t term_mask, int32_t level_set, void* workspace, size_t workspace_bytes, KArgs& a) {
  if (!fields || !vert || !params || !date) return fail(SPEEDY_ERR_ARG, "speedy: NULL struct argument");
  if (nlon <= 0 || nlat <= 0 || nmember <= 0) return fail(SPEEDY_ERR_ARG, "speedy: nlon, nlat, nmember must be > 0");
  if (vert->kx != 7 && vert->kx != 8 && vert->kx != 16)
    return fail(SPEEDY_ERR_KX, "speedy: kx must be 7, 8 or 16 (got " + std::to_string(vert->kx) + ")");
  if (flags & ~(uint32_t)SPEEDY_FLAG_LFLUXLAND) return fail(SPEEDY_ERR_ARG, "speedy: unknown flag bits");
  if (term_mask & ~(uint32_t)SPEEDY_TERMS_ALL) return fail(SPEEDY_ERR_ARG, "speedy: unknown term bits");
  if (level_set >= (1 << vert->kx)) return fail(SPEEDY_ERR_ARG, "speedy: level_set has bits beyond kx");
  if (int rc = check_fields(fields)) return rc;
  const int64_t ncol = (int64_t)nlon * nlat;
  const int64_t ntot = ncol * nmember;
  if (ntot > ((int64_t)1 << 40)) return fail(SPEEDY_ERR_CAPACITY, "speedy: too many columns");
  const WsLayout L = ws_layout(nlat, nmember);
  if (!workspace) return fail(SPEEDY_ERR_WORKSPACE, "speedy: workspace is NULL");
  if (workspace_bytes < L.total)
    return fail(SPEEDY_ERR_WORKSPACE, "speedy: workspace too small (need " + std::to_string(L.total) + " bytes)");
  std::memset(&a, 0, sizeof(a));
  a.f = *fields;
  a.v = *vert;
  a.p = *params;
  a.d = *date;
  a.nlon = nlon; a.nlat = nlat; a.nmember = nmember; a.level_set = level_set;
  a.ncol = ncol; a.ntot = ntot;
  // compute_shortwave = model_step % nstrad == 0 (speedy_physics.py:26-27, nstrad=3)
  int32_t ms = date->model_step % 3;
  if (ms < 0) ms += 3;
  a.flags = flags | ((ms == 0) ? (1u << 16) : 0u);
  a.term_mask = term_mask;
  char* w = static_cast<char*>(workspace);
  a.counter = reinterpret_cast<uint32_t*>(w + L.counter);
  a.s_accum = reinterpret_cast<uint32_t*>(w + L.s_accum);
  a.s_final = reinterpret_cast<uint32_t*>(w + L.s_final);
  a.lat = reinterpret_cast<float*>(w + L.lat);
  return SPEEDY_OK;
}

template <int KX>
int launch_fwd(const KArgs& a, hipStream_t s) {
  const int64_t nblk = (a.ntot + kBlock - 1) / kBlock;
  if (nblk > 0x7fffffff) return fail(SPEEDY_ERR_CAPACITY, "speedy: grid too large");
  int slot = timer_begin(s, 0);
  hipLaunchKernelGGL(cloudtop_kernel<KX>, dim3((unsigned)nblk), dim3(kBlock), 0, s, a);
  timer_end(s, slot);
  slot = timer_begin(s, 1);
  if (a.f.diag)
    hipLaunchKernelGGL((fwd_kernel<KX, true>), dim3((unsigned)nblk), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((fwd_kernel<KX, false>), dim3((unsigned)nblk), dim3(kBlock), 0, s, a);
  timer_end(s, slot);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy: launch failed: ") + hipGetErrorString(e));
  return SPEEDY_OK;
}

}  // namespace speedy

using namespace speedy;

extern "C" {

int speedy_abi_version(void) { return SPEEDY_ABI_VERSION; }

size_t speedy_workspace_bytes(int32_t nlat, int32_t nmember) {
  if (nlat <= 0 || nmember <= 0) return 0;
  return ws_layout(nlat, nmember).total;
}

int speedy_workspace_init(void* workspace, size_t workspace_bytes, void* stream) {
  if (!workspace || workspace_bytes < 256) return fail(SPEEDY_ERR_WORKSPACE, "speedy: bad workspace");
  const int64_t n = (int64_t)(workspace_bytes / 4);
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (uint32_t*)workspace, n);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy: fill failed: ") + hipGetErrorString(e));
  return SPEEDY_OK;
}

int speedy_diag_rows(int32_t kx) {
  if (kx != 7 && kx != 8 && kx != 16) return fail(-SPEEDY_ERR_KX, "speedy: kx must be 7, 8 or 16");
  return diag_rows(kx);
}

int speedy_diag_names(int32_t kx, char* buf, size_t buf_len) {
  if (kx != 7 && kx != 8 && kx != 16) return fail(SPEEDY_ERR_KX, "speedy: kx must be 7, 8 or 16");
  std::string s;
  for (int i = 0; i < kN2D; ++i) { if (!s.empty()) s += ","; s += kDiag2D[i]; }
  for (int i = 0; i < kN3D; ++i) { s += ","; s += kDiag3D[i]; s += "[]"; }
  if (!buf || buf_len < s.size() + 1) return fail(SPEEDY_ERR_ARG, "speedy: name buffer too small");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return SPEEDY_OK;
}

int speedy_fwd(const SpeedyFields* fields, const SpeedyVertical* vert, const SpeedyParams* params,
               const SpeedyDate* date, int32_t nlon, int32_t nlat, int32_t nmember, uint32_t flags,
               uint32_t term_mask, int32_t level_set, void* workspace, size_t workspace_bytes,
               void* stream) {
  KArgs a;
  if (int rc = prepare(fields, vert, params, date, nlon, nlat, nmember, flags, term_mask, level_set,
                       workspace, workspace_bytes, a))
    return rc;
  g_last_error.clear();
  hipStream_t s = (hipStream_t)stream;
  switch (vert->kx) {
    case 7: return launch_fwd<7>(a, s);
    case 8: return launch_fwd<8>(a, s);
    default: return launch_fwd<16>(a, s);
  }
}

int speedy_read_level_sets(const void* workspace, size_t workspace_bytes, int32_t nmember, uint32_t* out,
                           void* stream) {
  if (!workspace || !out || nmember <= 0) return fail(SPEEDY_ERR_ARG, "speedy: bad arguments");
  const WsLayout L = ws_layout(1, nmember);
  if (workspace_bytes < L.lat) return fail(SPEEDY_ERR_WORKSPACE, "speedy: workspace too small");
  hipError_t e = hipMemcpyAsync(out, static_cast<const char*>(workspace) + L.s_final, sizeof(uint32_t) * nmember,
                                hipMemcpyDeviceToHost, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy: copy failed: ") + hipGetErrorString(e));
  return SPEEDY_OK;
}

int speedy_timer_enable(int32_t capacity) {
  if (capacity <= 0 || capacity > (1 << 20)) return fail(SPEEDY_ERR_ARG, "speedy_timer_enable: bad capacity");
  std::lock_guard<std::mutex> lk(g_timer.mu);
  for (hipEvent_t e : g_timer.ev) (void)hipEventDestroy(e);
  g_timer.ev.assign(2 * (size_t)capacity, nullptr);
  g_timer.kind.assign(capacity, 0);
  for (auto& e : g_timer.ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(SPEEDY_ERR_HIP, "speedy_timer_enable: hipEventCreate failed");
  g_timer.cap = capacity;
  g_timer.used = 0;
  g_timer.on = true;
  return SPEEDY_OK;
}

int speedy_timer_disable(void) {
  std::lock_guard<std::mutex> lk(g_timer.mu);
  g_timer.on = false;
  return SPEEDY_OK;
}

int speedy_timer_read(float* ms_per_kind, int32_t* launches_per_kind, int32_t nkinds) {
  if (!ms_per_kind || !launches_per_kind || nkinds <= 0) return fail(SPEEDY_ERR_ARG, "speedy_timer_read: bad args");
  std::lock_guard<std::mutex> lk(g_timer.mu);
  for (int i = 0; i < nkinds; ++i) { ms_per_kind[i] = 0.0f; launches_per_kind[i] = 0; }
  for (int slot = 0; slot < g_timer.used; ++slot) {
    const int k = g_timer.kind[slot];
    float ms = 0.0f;
    if (hipEventSynchronize(g_timer.ev[2 * slot + 1]) != hipSuccess ||
        hipEventElapsedTime(&ms, g_timer.ev[2 * slot], g_timer.ev[2 * slot + 1]) != hipSuccess)
      return fail(SPEEDY_ERR_HIP, "speedy_timer_read: event query failed");
    if (k < nkinds) { ms_per_kind[k] += ms; launches_per_kind[k] += 1; }
  }
  g_timer.used = 0;
  return SPEEDY_OK;
}

const char* speedy_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
