// speedy_kernels.hip -- gfx950 kernels and the C ABI (include/speedy_abi.h) of the
// SPEEDY column physics.
//
// The eleven terms are column-local except for one grid-wide coupling: the
// shortwave cloud albedo is applied at every level of the set
//   S = { clip(icltop[c], 0, kx-1) : every column c of the call }
// (shortwave_radiation.py:63-70, an advanced-index scatter).  S is an OR of
// one bit per column, so a call needs a grid-wide reduction between the cloud
// top (end of phase A) and the shortwave (phase B).  Two paths:
//
//   fused (the grid fits on the GPU at once, e.g. T85 = 128 blocks):
//     one cooperative launch.  Right after the cloud top every block publishes
//     its columns' OR of level bits as one tagged word in its own slot (tag =
//     the call's generation), computes the S-independent work (longwave
//     transmissivity, longwave down, vertical diffusion), then polls only the
//     slots of the blocks covering its member until every tag is current.
//     Block 0 then advances the generation.  No read-modify-write atomics:
//     contended same-address atomics cost microseconds per call on this GPU.
//   two launches (larger grids, members narrower than a block, a given S, and
//     every batch): a level-set kernel, a fixed number of blocks each covering a
//     contiguous chunk of columns, writes per-(chunk, member) ORs; the forward
//     kernel ORs the few chunks of its member, recomputes phase A and does
//     phase B.  Single calls run this as a one-segment batch (seg_desc_kernel
//     writes the descriptor into the workspace).
//
// One thread per column; the column lives in registers.  Accesses are
// [field][level][column] with consecutive lanes on consecutive columns
// (coalesced 256-byte wave loads/stores).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "speedy_adjoint.hpp"

namespace speedy {

constexpr int kBlock = 256;
// Internal flag (above the public SPEEDY_FLAG_* bits): take the one-segment batch
// route even where the fused launch could run (speedy_fwd_noalloc).
constexpr uint32_t kFlagBatchRoute = 1u << 31;
constexpr uint32_t kSpinLimit = 4u << 20;   // barrier wait bound (~seconds): a wave always exits

struct Divisor {          // n / d for n < 2^32 by multiply-high + one correction step
  uint32_t d, m;          // m = floor(2^32 / d) (unused when d == 1)
};

struct KArgs {
  SpeedyFields f;
  SpeedyVertical v;
  SpeedyParams p;
  Uniforms<float> u;
  uint32_t ncol, ntot, nmember, nlat;
  uint32_t nblk;                     // blocks of kBlock columns
  uint32_t ls_blocks, chunk, jmax;   // two-kernel level-set pass: chunks of `chunk` columns
  Divisor dcol, dlat, dchunk;
  int32_t level_set;
  uint32_t flags, term_mask;
  uint32_t *gen, *err, *last_s, *partial;
  // fused forward + VJP only
  SpeedyAdjoint adj;
  float* prm_part;                   // parameter-gradient partials (per block and member slot, or per column)
  double* bud_part;                  // float64 budget partials (same layout, kBudParts doubles each)
};

constexpr int kBudParts = SPEEDY_NBUDGET + 1;   // the budgets and the weight sum

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

// ---------------------------------------------------------------- device helpers
__device__ __forceinline__ uint32_t divq(uint32_t n, const Divisor& dv) {
  if (dv.d == 1u) return n;
  uint32_t q = __umulhi(n, dv.m);
  if (n - q * dv.d >= dv.d) q += 1u;
  return q;
}

// Field access through address-space-1 (global) pointers.  The batch kernels
// read their field pointers from a descriptor in memory, and the compiler emits
// flat instructions for generic pointers; a flat access also counts in the
// scalar-load counter, so every wait for a uniform's scalar load then waits for
// the column's outstanding loads and stores too.  The kernel-argument kernels
// get global instructions either way.
template <typename T>
__device__ __forceinline__ T gld(const T* p, size_t i) {
  return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <typename T>
__device__ __forceinline__ void gst(T* p, size_t i, T v) {
  ((__attribute__((address_space(1))) T*)p)[i] = v;
}

template <int KX, bool EXTRA>
__device__ __forceinline__ void load_state(const KArgs& a, uint32_t m, uint32_t c, ColumnIn<float, KX>& col) {
  const size_t ncol = a.ncol;
  const size_t base3 = (size_t)m * KX * ncol + c;
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    col.T[k] = gld(a.f.t, base3 + k * ncol);
    col.Q[k] = gld(a.f.q, base3 + k * ncol);
    col.PHI[k] = gld(a.f.phi, base3 + k * ncol);
  }
  if (EXTRA && (a.flags & SPEEDY_FLAG_VERIFY_STATE)) {
#pragma unroll
    for (int k = 0; k < KX; ++k) col.Q[k] = verify_q(col.Q[k]);
  }
  col.u1 = gld(a.f.u, base3 + (KX - 1) * ncol);
  col.v1 = gld(a.f.v, base3 + (KX - 1) * ncol);
  col.ps = gld(a.f.ps, (size_t)m * ncol + c);
  col.alb0 = gld(a.f.alb0, c);
  col.sice = gld(a.f.sice_am, c);
  col.snowc_am = gld(a.f.snowc_am, c);
  col.soilw = gld(a.f.soilw_am, c);
  col.stl = gld(a.f.stl_am, c);
  col.sst = gld(a.f.sst, c);
  col.fmask = gld(a.f.fmask, c);
  col.orog = gld(a.f.orog, c);
  col.phis0 = gld(a.f.phis0, c);
}

template <int KX>
__device__ __forceinline__ void load_zonal(const KArgs& a, uint32_t c, ColumnIn<float, KX>& col) {
  const uint32_t j = c - divq(c, a.dlat) * a.nlat;
  float z[6];
  zonal_column(a.u.zon, gld(a.f.sia, j), gld(a.f.coa, j), z);
  col.fsol = z[0]; col.ozupp = z[1]; col.ozone = z[2]; col.zenit = z[3]; col.stratz = z[4]; col.sqcoa = z[5];
}

// u and v tendencies come from the surface stress at the lowest level only
// (surface_flux.py:294-295); the zeros above it are stored at the start of the
// kernel, off the critical path.
template <int KX>
__device__ __forceinline__ void store_uv_zeros(const KArgs& a, uint32_t m, uint32_t c) {
  const size_t ncol = a.ncol;
  const size_t base3 = (size_t)m * KX * ncol + c;
#pragma unroll
  for (int k = 0; k < KX - 1; ++k) {
    gst(a.f.du, base3 + k * ncol, 0.0f);
    gst(a.f.dv, base3 + k * ncol, 0.0f);
  }
}

// Up to eight levels, the term-ordered tendency sums (speedy_physics.py:123-125)
// are formed as early as their terms exist, in the same order of float
// additions: convection and condensation right after phase A; the q tendency of
// every level but the lowest once vertical diffusion (its last term there) is
// known.  The phase-A and vertical-diffusion tendencies then die early instead of
// staying live across phases b1 and b2 (16 fewer registers in the forward, and
// no scratch at two waves per SIMD).  The sixteen-level kernels keep the sums at
// the stores: they spill, and the early form gave wrong 16-level fused forward
// tendencies on gfx950 (the host build of the same order is right).
template <int KX>
constexpr bool kEarlySums = KX <= 8;

template <int KX>
__device__ __forceinline__ void tend_after_a(uint32_t tm, const PhaseA<float, KX>& pa, float at[KX], float aq[KX]) {
  if constexpr (!kEarlySums<KX>) return;
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    float t = 0.0f, q = 0.0f;
    if (tm & (1u << SPEEDY_TERM_CONVECTION)) { t = t + pa.tconv[k]; q = q + pa.qconv[k]; }
    if (tm & (1u << SPEEDY_TERM_CONDENSATION)) { t = t + pa.tlsc[k]; q = q + pa.qlsc[k]; }
    at[k] = t;
    aq[k] = q;
  }
}
template <int KX>
__device__ __forceinline__ void tend_after_b1(uint32_t tm, const PhaseB<float, KX>& pb, float aq[KX]) {
  if constexpr (!kEarlySums<KX>) return;
  if (tm & (1u << SPEEDY_TERM_VDIFF)) {
#pragma unroll
    for (int k = 0; k < KX - 1; ++k) aq[k] = aq[k] + pb.qvd[k];
  }
}

// Returns the levels whose q tendency verify_tendencies limited (bit k).
// at, aq: the partial sums of tend_after_a / tend_after_b1.
template <int KX, bool DIAG, bool EXTRA, bool UV_ALL>
__device__ __forceinline__ uint32_t store_outputs(const KArgs& a, uint32_t m, uint32_t c,
                                                  const ColumnIn<float, KX>& col, const PhaseA<float, KX>& pa,
                                                  const PhaseB<float, KX>& pb, const float at[KX], const float aq[KX],
                                                  double colsum[2]) {
  uint32_t limited = 0u;
  const bool budget = EXTRA && a.f.budget != nullptr;
  const double dpg = (double)col.ps * (double)(P0 / GRAV);   // p0 ps / g; times dhs_k: dp_k / g
  colsum[0] = 0.0;
  colsum[1] = 0.0;
  // term-ordered sum (speedy_physics.py:123-125)
  const uint32_t tm = a.term_mask;
  const size_t ncol = a.ncol;
  const size_t base3 = (size_t)m * KX * ncol + c;
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    float t = 0.0f, q = 0.0f, du = 0.0f, dv = 0.0f;
    if (kEarlySums<KX>) {
      t = at[k];
      q = aq[k];
    } else {
      if (tm & (1u << SPEEDY_TERM_CONVECTION)) { t = t + pa.tconv[k]; q = q + pa.qconv[k]; }
      if (tm & (1u << SPEEDY_TERM_CONDENSATION)) { t = t + pa.tlsc[k]; q = q + pa.qlsc[k]; }
    }
    if (tm & (1u << SPEEDY_TERM_SHORTWAVE)) t = t + pb.tsw[k];
    if (k == KX - 1 && (tm & (1u << SPEEDY_TERM_SURFACE))) {
      du = du + pb.du1; dv = dv + pb.dv1; t = t + pb.dts1; q = q + pb.dqs1;
    }
    if (tm & (1u << SPEEDY_TERM_LW_UP)) t = t + pb.tlw[k];
    if (tm & (1u << SPEEDY_TERM_VDIFF)) {
      t = t + pb.tvd[k];
      if (!kEarlySums<KX> || k == KX - 1) q = q + pb.qvd[k];   // early sums: the levels above took it already
    }
    if (EXTRA && (a.flags & SPEEDY_FLAG_VERIFY_TENDENCIES)) {
      bool lim;
      q = verify_dq(gld(a.f.q, base3 + k * ncol), q, a.u.dt_step, lim);   // the unclamped q (re-read, L2)
      limited |= (lim ? 1u : 0u) << k;
    }
    if (UV_ALL || k == KX - 1) {   // without UV_ALL the levels above were zeroed by store_uv_zeros
      gst(a.f.du, base3 + k * ncol, du);
      gst(a.f.dv, base3 + k * ncol, dv);
    }
    gst(a.f.dt, base3 + k * ncol, t);
    gst(a.f.dq, base3 + k * ncol, q);
    if (budget) {
      const double dp = (double)a.v.dhs[k] * dpg;
      colsum[0] += (double)CP * (double)t * dp;
      colsum[1] += (double)q * dp;
    }
  }
  if (DIAG) {
    float* dg = a.f.diag + (size_t)m * diag_rows(KX) * ncol + c;
    int r = 0;
    auto W = [&](float x) { gst(dg, (size_t)(r++) * ncol, x); };
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
    for (int k = 0; k < KX; ++k) W(pb.dfl[k]);
#pragma unroll
    for (int k = 0; k < KX; ++k) W(pb.st4a0[k]);
#pragma unroll
    for (int k = 0; k < KX; ++k) W(pb.st4a1[k]);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int k = 0; k < KX; ++k) W(pb.tauL[k][b]);
  }
  return limited;
}

// ---------------------------------------------------------------- level set S
// OR over the 64 lanes of a wave (all active).  DPP: six data-parallel-primitive
// ORs within the register file (quad swaps, half-row and row mirrors give every
// lane its row's OR; the row-15 and row-31 broadcasts fold rows into row 3), then
// lane 63 is read: no LDS round trip (the forward kernels).  Otherwise a
// butterfly of lane shuffles: the forward + VJP kernels keep it, because their
// register allocation (several KB of spills per lane at 16 levels) changed with
// another form and gave wrong gradients on gfx950.
template <bool DPP>
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
  if (DPP) {
    v |= __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xF, 0xF, false);    // quad_perm [1,0,3,2]
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xF, 0xF, false);    // quad_perm [2,3,0,1]
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x141, 0xF, 0xF, false);   // row_half_mirror
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x140, 0xF, 0xF, false);   // row_mirror
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xA, 0xF, false);   // row_bcast:15 into rows 1, 3
    v |= __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xC, 0xF, false);   // row_bcast:31 into rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ uint32_t gen_tag(uint32_t g) { return (g & 0x7fffu) | 0x8000u; }



__device__ __forceinline__ uint32_t clip_bit(int icl, int kx) {
  return 1u << (icl < 0 ? 0 : (icl > kx - 1 ? kx - 1 : icl));
}

// Fused path: OR of the level bits of member m over the published words of the
// blocks covering it, polled until every word carries `tag` (one wave).
// Bounded: after kSpinLimit polls it flags the error word and returns.
template <bool DPP>
__device__ __forceinline__ uint32_t collect_fused_t(const KArgs& a, uint32_t tag, uint32_t m) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t b0 = (m * a.ncol) / kBlock, b1 = (m * a.ncol + a.ncol - 1u) / kBlock;
  for (uint32_t spins = 0;; ++spins) {
    uint32_t acc = 0;
    int stale = 0;
    for (uint32_t b = b0 + lane; b <= b1; b += 64u) {
      const uint32_t j = m - divq(b * kBlock, a.dcol);
      const uint32_t w = __hip_atomic_load(&a.partial[2u * b + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      stale |= (w >> 16) != tag;
      acc |= w & 0xffffu;
    }
    if (!__any(stale)) return wave_or<DPP>(acc);
    if (spins >= kSpinLimit) {
      if (lane == 0) __hip_atomic_fetch_or(a.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return wave_or<DPP>(acc);
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// The forward + VJP kernels' form (an outlined call, shuffle OR).
__device__ uint32_t collect_fused(const KArgs& a, uint32_t tag, uint32_t m) { return collect_fused_t<false>(a, tag, m); }

// Two-kernel path: OR over the level-set chunks that cover member m, by one wave
// (the forward + VJP kernels' form).
__device__ uint32_t collect_chunks(const KArgs& a, uint32_t m) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t b0 = divq(m * a.ncol, a.dchunk), b1 = divq(m * a.ncol + a.ncol - 1u, a.dchunk);
  uint32_t acc = 0;
  for (uint32_t b = b0 + lane; b <= b1; b += 64u) acc |= gld(a.partial, b * a.jmax + (m - divq(b * a.chunk, a.dcol)));
  return wave_or<false>(acc);
}

// Two-kernel path: member m's word of level-set chunk b (levelset_batch_kernel output).
__device__ __forceinline__ uint32_t chunk_word(const KArgs& a, uint32_t b, uint32_t m) {
  return gld(a.partial, b * a.jmax + (m - divq(b * a.chunk, a.dcol)));
}

// Two-kernel path, member wider than a block: this thread's share of the OR over
// the chunks covering member m, kChunkLoads words per thread.  Indexes past the
// member's last chunk are clamped to it (OR is idempotent), so the loads carry
// no branches and issue together.  A member spans at most ls_blocks <= 512
// chunks (make_grid), within the kChunkLoads * kBlock = 1024 this reaches.
constexpr uint32_t kChunkLoads = 4;
__device__ __forceinline__ uint32_t chunk_share(const KArgs& a, uint32_t m) {
  const uint32_t b0 = divq(m * a.ncol, a.dchunk), b1 = divq(m * a.ncol + a.ncol - 1u, a.dchunk);
  uint32_t acc = 0u;
#pragma unroll
  for (uint32_t r = 0; r < kChunkLoads; ++r) acc |= chunk_word(a, min(b0 + threadIdx.x + r * kBlock, b1), m);
  return acc;
}

// ---------------------------------------------------------------- kernels
// Two-kernel path, launch 1: every column's cloud top, ORed per (chunk, member).
template <int KX>
__device__ __forceinline__ void levelset_body(const KArgs& a, uint32_t bid, uint32_t* s_or) {
  const uint32_t c_begin = bid * a.chunk;
  const uint32_t c_end = min(c_begin + a.chunk, a.ntot);
  const uint32_t m_first = divq(c_begin, a.dcol);
  const uint32_t nm = divq(c_end - 1u, a.dcol) - m_first + 1u;
  for (uint32_t i = threadIdx.x; i < nm; i += kBlock) s_or[i] = 0u;
  __syncthreads();
  // kLsCols columns per thread per pass: all their loads issue before the first
  // cloud top is computed (at two waves per SIMD one column's loads leave the
  // memory latency exposed).  Columns past the chunk load column c_begin's
  // words (in bounds) and are not ORed.
  constexpr uint32_t kLsCols = 4;
  const size_t ncol = a.ncol;
  for (uint32_t g0 = c_begin + threadIdx.x; g0 < c_end; g0 += kLsCols * kBlock) {
    ColumnIn<float, KX> col[kLsCols];
    uint32_t mcol[kLsCols];
#pragma unroll
    for (uint32_t u = 0; u < kLsCols; ++u) {
      const uint32_t gid = g0 + u * kBlock < c_end ? g0 + u * kBlock : c_begin;
      const uint32_t m = divq(gid, a.dcol);
      const uint32_t c = gid - m * a.ncol;
      const size_t base3 = (size_t)m * KX * ncol + c;
      mcol[u] = m;
#pragma unroll
      for (int k = 1; k < KX; ++k) {
        col[u].T[k] = gld(a.f.t, base3 + k * ncol);
        col[u].Q[k] = gld(a.f.q, base3 + k * ncol);
        col[u].PHI[k] = gld(a.f.phi, base3 + k * ncol);
      }
      col[u].ps = gld(a.f.ps, (size_t)m * ncol + c);
    }
#pragma unroll
    for (uint32_t u = 0; u < kLsCols; ++u) {
      if (a.flags & SPEEDY_FLAG_VERIFY_STATE) {
#pragma unroll
        for (int k = 1; k < KX; ++k) col[u].Q[k] = verify_q(col[u].Q[k]);
      }
      if (g0 + u * kBlock < c_end)
        atomicOr(&s_or[mcol[u] - m_first], clip_bit(cloud_top<float, KX>(col[u], a.p, a.v, a.u), KX));
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nm; i += kBlock) gst(a.partial, bid * a.jmax + i, s_or[i]);
}

// ---------------------------------------------------------------- VJP helpers
// One column's slice of an LDS array laid out [slot][thread] (conflict-free).
struct LdsStore {
  static constexpr bool kOn = true;
  float* base;
  int off = 0;
  __device__ __forceinline__ void put(int i, float x) { base[i * kBlock + off] = x; }
  __device__ __forceinline__ float get(int i) const { return base[i * kBlock + off]; }
  // The same slots, read only after `dep` is computed (the offset passes through
  // an empty asm that takes dep), so the reads are not hoisted above it.
  __device__ __forceinline__ LdsStore after(float dep) const {
    LdsStore t = *this;
    asm volatile("" : "+v"(t.off) : "v"(dep));
    return t;
  }
};

template <int KX>
__device__ __forceinline__ void load_cotangents(const KArgs& a, uint32_t m, uint32_t c, ColumnCt<float, KX>& ct) {
  const size_t ncol = a.ncol;
  const size_t base3 = (size_t)m * KX * ncol + c;
  ct.du1 = gld(a.adj.ct_du, base3 + (KX - 1) * ncol);
  ct.dv1 = gld(a.adj.ct_dv, base3 + (KX - 1) * ncol);
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    ct.dt[k] = gld(a.adj.ct_dt, base3 + k * ncol);
    ct.dq[k] = gld(a.adj.ct_dq, base3 + k * ncol);
  }
}

template <int KX>
__device__ __forceinline__ void store_grads(const KArgs& a, uint32_t m, uint32_t c, const ColumnAdj<float, KX>& g) {
  const size_t ncol = a.ncol;
  const size_t base3 = (size_t)m * KX * ncol + c;
#pragma unroll
  for (int k = 0; k < KX; ++k) {
    gst(a.adj.g_u, base3 + k * ncol, k == KX - 1 ? g.u1 : 0.0f);
    gst(a.adj.g_v, base3 + k * ncol, k == KX - 1 ? g.v1 : 0.0f);
    gst(a.adj.g_t, base3 + k * ncol, g.T[k]);
    gst(a.adj.g_q, base3 + k * ncol, g.Q[k]);
  }
  gst(a.adj.g_ps, (size_t)m * ncol + c, g.ps);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One DPP step of a float sum: x plus the value DPP control CTRL selects (lanes
// outside ROWS keep x).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_add(float x) {
  return x + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, ROWS, 0xF, false));
}

// Float sum over the 64 lanes of a wave (all active) in a fixed order, by the
// DPP steps of wave_or (every lane holds its row's sum after the row mirror; the
// row broadcasts add rows 0-2 into row 3); the total is read from lane 63.
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v = dpp_add<0xB1, 0xF>(v);    // quad_perm [1,0,3,2]
  v = dpp_add<0x4E, 0xF>(v);    // quad_perm [2,3,0,1]
  v = dpp_add<0x141, 0xF>(v);   // row_half_mirror
  v = dpp_add<0x140, 0xF>(v);   // row_mirror
  v = dpp_add<0x142, 0xA>(v);   // row_bcast:15 into rows 1, 3
  v = dpp_add<0x143, 0xC>(v);   // row_bcast:31 into rows 2, 3
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

// Parameter gradients of this block: a fixed-order tree sum per member slot
// (deterministic), written to prm_part[block][slot][5].  Members narrower than
// a block write per-column values instead (prm_part[column][5]).
__device__ __forceinline__ void block_param_partials(const KArgs& a, uint32_t bid, uint32_t gid, bool active,
                                                     uint32_t slot, const float prm[VJP_NPARAM]) {
  if (a.ncol < (uint32_t)kBlock) {
    if (active)
#pragma unroll
      for (int i = 0; i < VJP_NPARAM; ++i) gst(a.prm_part, (size_t)gid * VJP_NPARAM + i, prm[i]);
    return;
  }
  __shared__ float s_red[kBlock / 64][2 * VJP_NPARAM];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < VJP_NPARAM; ++i) {
      const float v = wave_sum((active && slot == (uint32_t)j) ? prm[i] : 0.0f);
      if (lane == 0) s_red[wave][j * VJP_NPARAM + i] = v;
    }
  __syncthreads();
  if (threadIdx.x < 2 * VJP_NPARAM) {
    float acc = s_red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) acc = acc + s_red[w][threadIdx.x];
    gst(a.prm_part, (size_t)bid * 2 * VJP_NPARAM + threadIdx.x, acc);
  }
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Float64 budget partials of this block per member slot (layout as the
// parameter partials, kBudParts doubles): weighted sums, the largest column
// water residual, and the weight sum.
__device__ __forceinline__ void block_budget_partials(const KArgs& a, uint32_t bid, uint32_t gid, bool active,
                                                      uint32_t slot, const double v[kBudParts]) {
  if (a.ncol < (uint32_t)kBlock) {
    if (active)
#pragma unroll
      for (int i = 0; i < kBudParts; ++i) gst(a.bud_part, (size_t)gid * kBudParts + i, v[i]);
    return;
  }
  __shared__ double s_bud[kBlock / 64][2 * kBudParts];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int i = 0; i < kBudParts; ++i) {
      const bool mine = active && slot == (uint32_t)j;
      const double x = i == 5 ? wave_max_d(mine ? v[i] : 0.0) : wave_sum_d(mine ? v[i] : 0.0);
      if (lane == 0) s_bud[wave][j * kBudParts + i] = x;
    }
  __syncthreads();
  if (threadIdx.x < 2 * kBudParts) {
    const bool is_max = (threadIdx.x % kBudParts) == 5;
    double acc = s_bud[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) acc = is_max ? fmax(acc, s_bud[w][threadIdx.x]) : acc + s_bud[w][threadIdx.x];
    gst(a.bud_part, (size_t)bid * 2 * kBudParts + threadIdx.x, acc);
  }
}

// Forward (and, with VJP, the fused forward + VJP).  FUSED: single launch with
// the split grid barrier (all blocks co-resident); otherwise S comes from the
// level-set kernel of the same batch.
// EXTRA: the optional verify steps and float64 budgets are compiled in (their
// flags and pointers are then read at run time); the plain forward omits them.
// The batched forward kernels up to eight levels keep the longwave
// transmissivities in LDS between longwave down and longwave up (4 KX floats a
// column, [slot][thread]) instead of recomputing them per level (T425: 153 fewer
// VALU instructions a wave, ~1 % faster).  The fused kernels recompute: at one
// wave per SIMD the dependent LDS reads in the longwave-up sweep cost more than
// the exponentials (T85 forward ~2 % slower with the store).  The forward + VJP
// kernels' LDS holds the adjoint's tapes, and the sixteen-level kernels keep their
// tested code.
template <bool ON>
__device__ __forceinline__ auto tau_store(float* s) {
  if constexpr (ON)
    return LdsStore{s + threadIdx.x};
  else
    return NoTauStore{};
}

template <int KX, bool DIAG, bool FUSED, bool VJP, bool EXTRA>
__device__ __forceinline__ void fwd_body(const KArgs& a, uint32_t bid) {
  __shared__ uint32_t s_s[2];
  constexpr bool kTau = !VJP && !FUSED && KX <= 8;
  __shared__ float s_tau[kTau ? 4 * KX * kBlock : 1];
  const auto ts = tau_store<kTau>(s_tau);
  const uint32_t gid = bid * kBlock + threadIdx.x;
  const bool active = gid < a.ntot;
  const uint32_t g0 = gid < a.ntot ? gid : a.ntot - 1u;   // inactive lanes shadow a valid column
  const uint32_t m = divq(g0, a.dcol);
  const uint32_t c = g0 - m * a.ncol;
  const uint32_t blk0 = bid * kBlock;
  const uint32_t m_first = divq(blk0, a.dcol);
  const uint32_t nm = divq(min(blk0 + kBlock, a.ntot) - 1u, a.dcol) - m_first + 1u;   // <= 2 when fused
  const bool need_s = a.level_set < 0;
  // The call's generation tag, read first so that its latency overlaps the state
  // loads; wave 0 publishes and gathers, so each of its lanes holds the tag.
  uint32_t tag = 0u;
  if (FUSED && need_s && threadIdx.x < 64u)
    tag = gen_tag(__hip_atomic_load(a.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  // Two-kernel path: the level-set chunk words are loaded first too, so their
  // round trip overlaps the state's.  A member wider than a block: each thread
  // ORs its share of the chunks of the block's (at most two) members; narrower:
  // the <= 2 chunks covering the thread's own member.
  const bool wide = a.ncol >= (uint32_t)kBlock;   // a block spans at most two members
  uint32_t lw0 = 0u, lw1 = 0u;
  if (!FUSED && !VJP && need_s) {
    if (wide) {
      lw0 = chunk_share(a, m_first);
      if (nm > 1u) lw1 = chunk_share(a, m_first + 1u);
    } else {
      const uint32_t b0 = divq(m * a.ncol, a.dchunk), b1 = divq(m * a.ncol + a.ncol - 1u, a.dchunk);
      lw0 = chunk_word(a, b0, m) | chunk_word(a, b1, m);
    }
  }
  ColumnIn<float, KX> col;
  load_state<KX, EXTRA>(a, m, c, col);
  if (!VJP) {
    load_zonal<KX>(a, c, col);   // S-independent; its loads travel with the state's
    if (active) store_uv_zeros<KX>(a, m, c);
  }
  PhaseA<float, KX> pa;
  phase_a<float, KX>(col, a.p, a.v, a.u, pa);
  float tend_t[KX], tend_q[KX];
  tend_after_a<KX>(a.term_mask, pa, tend_t, tend_q);
  if (FUSED && need_s) {
    // publish this block's word(s) for the call's level set: a wave-level OR per
    // member slot (a fused block spans at most two members), then one word per slot
    __shared__ uint32_t s_wor[kBlock / 64][2];
    const uint32_t bit = active ? clip_bit(pa.icltop, KX) : 0u;
    const uint32_t o0 = wave_or<!VJP>(m == m_first ? bit : 0u), o1 = wave_or<!VJP>(m != m_first ? bit : 0u);
    if ((threadIdx.x & 63u) == 0) { s_wor[threadIdx.x >> 6][0] = o0; s_wor[threadIdx.x >> 6][1] = o1; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t j = 0; j < nm; ++j) {
        uint32_t w = 0u;
#pragma unroll
        for (int i = 0; i < kBlock / 64; ++i) w |= s_wor[i][j];
        __hip_atomic_store(&a.partial[2u * bid + j], (tag << 16) | w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  PhaseB<float, KX> pb;
  phase_b1<float, KX>(col, pa, a.p, a.v, a.u, pb, ts);
  tend_after_b1<KX>(a.term_mask, pb, tend_q);
  uint32_t S = (uint32_t)a.level_set;
  if (VJP && need_s) {
    // the forward + VJP kernels: wave 0 gathers S for the block's member(s)
    if (threadIdx.x < 64u) {
      for (uint32_t j = 0; wide && j < nm; ++j) {
        const uint32_t v = FUSED ? collect_fused(a, tag, m_first + j) : collect_chunks(a, m_first + j);
        if (threadIdx.x == 0) s_s[j] = v;
      }
      if (bid == 0) {
        // every member's S for speedy_read_level_sets; in the fused path this also
        // waits for every block's word, i.e. every block has read the generation
        for (uint32_t mm = 0; mm < a.nmember; ++mm) {
          const uint32_t v = FUSED ? collect_fused(a, tag, mm) : collect_chunks(a, mm);
          if (threadIdx.x == 0) gst(a.last_s, mm, v);
        }
        if (FUSED && threadIdx.x == 0) {
          const uint32_t g = __hip_atomic_load(a.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(a.gen, g + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    __syncthreads();
    if (wide) {
      S = s_s[m - m_first];
    } else {   // two-kernel path, members narrower than a block: each spans <= 2 chunks
      S = 0u;
      const uint32_t b0 = divq(m * a.ncol, a.dchunk), b1 = divq(m * a.ncol + a.ncol - 1u, a.dchunk);
      for (uint32_t b = b0; b <= b1; ++b) S |= gld(a.partial, b * a.jmax + (m - divq(b * a.chunk, a.dcol)));
    }
  } else if (FUSED && need_s) {
    if (threadIdx.x < 64u) {   // wave 0 gathers S for the block's member(s)
      for (uint32_t j = 0; j < nm; ++j) {
        const uint32_t v = collect_fused_t<true>(a, tag, m_first + j);
        if (threadIdx.x == 0) s_s[j] = v;
      }
      if (bid == 0) {
        // every member's S for speedy_read_level_sets; this also waits for every
        // block's word, i.e. every block has read the generation
        for (uint32_t mm = 0; mm < a.nmember; ++mm) {
          const uint32_t v = collect_fused_t<true>(a, tag, mm);
          if (threadIdx.x == 0) gst(a.last_s, mm, v);
        }
        if (threadIdx.x == 0) {
          const uint32_t g = __hip_atomic_load(a.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(a.gen, g + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    __syncthreads();
    S = s_s[m - m_first];
  } else if (need_s) {
    if (wide) {   // the block's OR per member slot
      __shared__ uint32_t s_lw[kBlock / 64][2];
      const uint32_t o0 = wave_or<true>(lw0), o1 = wave_or<true>(lw1);
      if ((threadIdx.x & 63u) == 0) { s_lw[threadIdx.x >> 6][0] = o0; s_lw[threadIdx.x >> 6][1] = o1; }
      __syncthreads();
      S = 0u;
#pragma unroll
      for (int i = 0; i < kBlock / 64; ++i) S |= s_lw[i][m - m_first];
    } else {
      S = lw0;
    }
    // every member's S for speedy_read_level_sets, from its first column
    if (active && c == 0u) gst(a.last_s, m, S);
  }
  // The forward + VJP kernels keep the zonal loads and all output stores here:
  // moving them earlier changed their register allocation and, at 16 levels,
  // gave wrong gradients on gfx950 (the kernels spill ~4 KB per lane).
  if (VJP) load_zonal<KX>(a, c, col);
  phase_b2<float, KX, DIAG>(col, pa, S, a.p, a.v, a.u, a.flags, pb, ts);
  uint32_t limited = 0u;
  double colsum[2] = {0.0, 0.0};
  if (active) limited = store_outputs<KX, DIAG, EXTRA, VJP>(a, m, c, col, pa, pb, tend_t, tend_q, colsum);
  if (EXTRA && a.f.budget) {   // float64 column budgets (block-uniform branch)
    const uint32_t j = c - divq(c, a.dlat) * a.nlat;
    const double w = (double)gld(a.f.coa, j);
    const double E = (double)pb.evap2, P = (double)pa.precnv + (double)pa.precls;
    const double res = colsum[1] - (E - P);
    const double v[kBudParts] = {w * colsum[0], w * colsum[1], w * E, w * P, w * res, fabs(res), w};
    block_budget_partials(a, bid, gid, active, m - m_first, v);
  }
  if (VJP && a.adj.g_t != nullptr) {   // batches mix forward-only and VJP segments
    ColumnCt<float, KX> ct;
    load_cotangents<KX>(a, m, c, ct);
    float g_qraw[KX];
    verify_ct<float, KX>(EXTRA ? a.flags : 0u, limited, a.u.dt_step, ct, g_qraw);
    ColumnAdj<float, KX> g;
    // longwave flux tapes in LDS, [slot][thread]: the kernel runs one block per CU
    // (up to eight levels, KX more slots hold g_qraw across the adjoint, where it
    // would otherwise take KX registers)
    constexpr int kQraw = KX <= 8 ? KX : 0;
    __shared__ float s_tape[(vjp_tape_floats<KX>() + kQraw) * kBlock];
    LdsStore tape{s_tape + threadIdx.x};
    if constexpr (kQraw > 0) {
#pragma unroll
      for (int k = 0; k < KX; ++k) tape.put(vjp_tape_floats<KX>() + k, g_qraw[k]);
    }
    column_vjp<float, KX>(col, pa, pb, S, a.p, a.v, a.u, a.flags, ct, g, tape);
    if (EXTRA && (a.flags & SPEEDY_FLAG_VERIFY_STATE)) {
      const size_t base3 = (size_t)m * KX * a.ncol + c;
#pragma unroll
      for (int k = 0; k < KX; ++k) g.Q[k] = verify_q_grad(gld(a.f.q, base3 + k * (size_t)a.ncol), g.Q[k]);
    }
    if constexpr (kQraw > 0) {
      const LdsStore t = tape.after(g.Q[0]);
#pragma unroll
      for (int k = 0; k < KX; ++k) g.Q[k] += t.get(vjp_tape_floats<KX>() + k);
    } else {
#pragma unroll
      for (int k = 0; k < KX; ++k) g.Q[k] += g_qraw[k];
    }
    if (active) store_grads<KX>(a, m, c, g);
    block_param_partials(a, bid, gid, active, m - m_first, g.prm);
  }
}


// Per-member parameter gradients from the partials, one block per member, in a
// fixed order (deterministic).
__device__ __forceinline__ void budget_reduce_body(const KArgs& a, uint32_t m) {
  double acc[kBudParts];
#pragma unroll
  for (int i = 0; i < kBudParts; ++i) acc[i] = 0.0;
  // (unrolled: several iterations' loads in flight; the order of the sums is the loop's)
  if (a.ncol >= (uint32_t)kBlock) {
    const uint32_t b0 = (m * a.ncol) / kBlock, b1 = (m * a.ncol + a.ncol - 1u) / kBlock;
#pragma unroll 4
    for (uint32_t b = b0 + threadIdx.x; b <= b1; b += kBlock) {
      const uint32_t j = m - divq(b * kBlock, a.dcol);
#pragma unroll
      for (int i = 0; i < kBudParts; ++i) {
        const double x = gld(a.bud_part, ((size_t)b * 2 + j) * kBudParts + i);
        acc[i] = i == 5 ? fmax(acc[i], x) : acc[i] + x;
      }
    }
  } else {
#pragma unroll 4
    for (uint32_t cc = threadIdx.x; cc < a.ncol; cc += kBlock)
#pragma unroll
      for (int i = 0; i < kBudParts; ++i) {
        const double x = gld(a.bud_part, ((size_t)m * a.ncol + cc) * kBudParts + i);
        acc[i] = i == 5 ? fmax(acc[i], x) : acc[i] + x;
      }
  }
  __shared__ double s_red[kBlock / 64][kBudParts];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
  for (int i = 0; i < kBudParts; ++i) {
    const double v = i == 5 ? wave_max_d(acc[i]) : wave_sum_d(acc[i]);
    if (lane == 0) s_red[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < kBudParts) {
    const bool is_max = threadIdx.x == 5;
    double t = s_red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) t = is_max ? fmax(t, s_red[w][threadIdx.x]) : t + s_red[w][threadIdx.x];
    s_red[0][threadIdx.x] = t;
  }
  __syncthreads();
  if (threadIdx.x < SPEEDY_NBUDGET) {
    const double wsum = s_red[0][SPEEDY_NBUDGET];
    const double t = s_red[0][threadIdx.x];
    gst(a.f.budget, (size_t)m * SPEEDY_NBUDGET + threadIdx.x, threadIdx.x == 5 ? t : t / wsum);
  }
}

__device__ __forceinline__ void param_reduce_body(const KArgs& a, uint32_t m) {
  float acc[VJP_NPARAM];
#pragma unroll
  for (int i = 0; i < VJP_NPARAM; ++i) acc[i] = 0.0f;
  // (unrolled: several iterations' loads in flight; the order of the sums is the loop's)
  if (a.ncol >= (uint32_t)kBlock) {
    const uint32_t b0 = (m * a.ncol) / kBlock, b1 = (m * a.ncol + a.ncol - 1u) / kBlock;
#pragma unroll 4
    for (uint32_t b = b0 + threadIdx.x; b <= b1; b += kBlock) {
      const uint32_t j = m - divq(b * kBlock, a.dcol);
#pragma unroll
      for (int i = 0; i < VJP_NPARAM; ++i) acc[i] += gld(a.prm_part, ((size_t)b * 2 + j) * VJP_NPARAM + i);
    }
  } else {
#pragma unroll 4
    for (uint32_t cc = threadIdx.x; cc < a.ncol; cc += kBlock)
#pragma unroll
      for (int i = 0; i < VJP_NPARAM; ++i) acc[i] += gld(a.prm_part, ((size_t)m * a.ncol + cc) * VJP_NPARAM + i);
  }
  __shared__ float s_red[kBlock / 64][VJP_NPARAM];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
  for (int i = 0; i < VJP_NPARAM; ++i) {
    const float v = wave_sum_dpp(acc[i]);
    if (lane == 0) s_red[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < VJP_NPARAM) {
    float t = s_red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) t = t + s_red[w][threadIdx.x];
    gst(a.adj.g_params, (size_t)m * VJP_NPARAM + threadIdx.x, t);
  }
}

// Per-member sums of a call: parameter gradients (VJP) and float64 budgets.
__device__ __forceinline__ void member_reduce_body(const KArgs& a, uint32_t m) {
  if (a.adj.g_params) param_reduce_body(a, m);
  if (a.f.budget) {
    __syncthreads();
    budget_reduce_body(a, m);
  }
}

__global__ __launch_bounds__(kBlock) void param_reduce_kernel(KArgs a) { member_reduce_body(a, blockIdx.x); }

template <int KX, bool DIAG, bool FUSED, bool VJP, bool EXTRA>
__global__ __launch_bounds__(kBlock) void fwd_kernel(KArgs a) {
  fwd_body<KX, DIAG, FUSED, VJP, EXTRA>(a, blockIdx.x);
}

// ---------------------------------------------------------------- batches
// A batch descriptor holds one KArgs per segment and, per launch, a table of
// (segment << 20 | block-in-segment) words; every block reads its word once and
// then works exactly as in the single-call kernels.
constexpr uint32_t kSegShift = 20, kSegLocalMask = (1u << kSegShift) - 1u;

template <int KX>
__global__ __launch_bounds__(kBlock) void levelset_batch_kernel(const KArgs* __restrict__ segs,
                                                                const uint32_t* __restrict__ table) {
  extern __shared__ uint32_t s_or[];
  const uint32_t e = __builtin_amdgcn_readfirstlane(table[blockIdx.x]);
  levelset_body<KX>(segs[e >> kSegShift], e & kSegLocalMask, s_or);
}

// Forward-only segments of a batch: the forward without the adjoint's registers.
// The descriptors are read-only and alias nothing the kernel writes
// (__restrict__), so their fields are scalar loads and the field pointers read
// from them are global-memory pointers.
template <int KX>
__global__ __launch_bounds__(kBlock, KX <= 8 ? 2 : 1) void fwd_only_batch_kernel(const KArgs* __restrict__ segs,
                                                                const uint32_t* __restrict__ table) {
  const uint32_t e = __builtin_amdgcn_readfirstlane(table[blockIdx.x]);
  fwd_body<KX, false, false, false, true>(segs[e >> kSegShift], e & kSegLocalMask);
}

// Forward-only segments of a batch in which none needs the verify steps or the
// float64 budgets: the forward without that optional code (4 % faster at T425).
// Both forward-only batch kernels are built for two waves per SIMD at 7 and 8
// levels (at most 256 registers a lane, no scratch): the second wave covers the
// first one's stalls, 128 -> 86 us at T425 x 8 levels.
template <int KX>
__global__ __launch_bounds__(kBlock, KX <= 8 ? 2 : 1) void fwd_plain_batch_kernel(const KArgs* __restrict__ segs,
                                                                 const uint32_t* __restrict__ table) {
  const uint32_t e = __builtin_amdgcn_readfirstlane(table[blockIdx.x]);
  fwd_body<KX, false, false, false, false>(segs[e >> kSegShift], e & kSegLocalMask);
}

// Forward-only segments that also write the PhysicsData diagnostics: single
// calls with diagnostics that the fused launch cannot take.
template <int KX>
__global__ __launch_bounds__(kBlock) void fwd_diag_batch_kernel(const KArgs* __restrict__ segs,
                                                                const uint32_t* __restrict__ table) {
  const uint32_t e = __builtin_amdgcn_readfirstlane(table[blockIdx.x]);
  fwd_body<KX, true, false, false, true>(segs[e >> kSegShift], e & kSegLocalMask);
}

// Forward + VJP segments of a batch.  Built for two waves per SIMD at 7 and 8
// levels (at most 256 registers a lane, 132 B of scratch a lane at 8 levels and
// none at 7, 45 KB of LDS a block): 521 -> 412 us at T425 x 8 levels, against one
// wave per SIMD without scratch.  The fused forward + VJP kernel keeps one wave
// per SIMD and no scratch.
template <int KX>
__global__ __launch_bounds__(kBlock, KX <= 8 ? 2 : 1) void fwd_batch_kernel(const KArgs* segs, const uint32_t* table) {
  const uint32_t e = __builtin_amdgcn_readfirstlane(table[blockIdx.x]);
  fwd_body<KX, false, false, true, true>(segs[e >> kSegShift], e & kSegLocalMask);
}

__global__ __launch_bounds__(kBlock) void param_reduce_batch_kernel(const KArgs* __restrict__ segs,
                                                                    const uint32_t* __restrict__ table) {
  const uint32_t e = __builtin_amdgcn_readfirstlane(table[blockIdx.x]);
  member_reduce_body(segs[e >> kSegShift], e & kSegLocalMask);
}

// The call's arguments as a one-segment batch in the workspace: block 0 copies the
// argument block word by word; every thread writes one table entry (the block-,
// level-set-block- and member-tables in turn, each counting from 0).
__global__ __launch_bounds__(kBlock) void seg_desc_kernel(KArgs a, KArgs* dst, uint32_t* tbl, uint32_t nblk,
                                                          uint32_t nls, uint32_t nred) {
  if (blockIdx.x == 0) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&a);
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
    for (uint32_t w = threadIdx.x; w < (uint32_t)(sizeof(KArgs) / 4); w += kBlock) d[w] = src[w];
  }
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nblk + nls + nred) tbl[i] = i < nblk ? i : (i < nblk + nls ? i - nblk : i - nblk - nls);
}

// acc[e] += diag[e] / steps for e < n (the diagnostics' running mean): four
// elements per thread (16-byte loads and stores when both buffers allow), a
// grid-stride loop, and the scalar tail by the first threads.
__global__ __launch_bounds__(kBlock) void diag_accumulate_kernel(const float* __restrict__ d, float* __restrict__ acc,
                                                                 int64_t n, float steps, int vec) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  const int64_t t0 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t done = 0;
  if (vec) {
    const int64_t n4 = n / 4;
    const float4* d4 = reinterpret_cast<const float4*>(d);
    float4* a4 = reinterpret_cast<float4*>(acc);
    for (int64_t i = t0; i < n4; i += stride) {
      const float4 x = d4[i];
      float4 y = a4[i];
      y.x += x.x / steps; y.y += x.y / steps; y.z += x.z / steps; y.w += x.w / steps;
      a4[i] = y;
    }
    done = n4 * 4;
  }
  for (int64_t i = done + t0; i < n; i += stride) acc[i] += d[i] / steps;
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

// kernel timers
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

// Workspace: word [0] generation, [1] error word; from byte 256 last_s[nmember];
// then the published level-set words (max of the fused path's two per block
// and the two-kernel path's jmax per chunk).
struct Grid {
  uint32_t ncol, ntot, nblk, ls_blocks, chunk, jmax;
};

Grid make_grid(int64_t nlon, int64_t nlat, int64_t nmember) {
  Grid g;
  g.ncol = (uint32_t)(nlon * nlat);
  g.ntot = (uint32_t)(nlon * nlat * nmember);
  g.nblk = (g.ntot + kBlock - 1) / kBlock;
  // At most 512 level-set chunks (~2 blocks per CU): the forward + VJP kernels
  // gather a member's chunk words in one wave, one round trip per 64 chunks, and at
  // T425 512 chunks cost that gather 20 us less than 1024 (fwd+VJP 892 vs 912 us)
  // while the level-set pass itself stays at 25 us (256 chunks: 33 us).
  const uint32_t lsb = g.nblk < 512u ? g.nblk : 512u;
  g.chunk = ((g.ntot + lsb - 1) / lsb + kBlock - 1) / kBlock * kBlock;
  g.ls_blocks = (g.ntot + g.chunk - 1) / g.chunk;
  g.jmax = (g.chunk - 1) / g.ncol + 2;
  return g;
}

struct WsLayout {
  size_t last_s, partial, prm, bud, desc, tbl, total;
};

WsLayout ws_layout(const Grid& g, int32_t nmember) {
  WsLayout L;
  L.last_s = 256;
  L.partial = L.last_s + align256(sizeof(uint32_t) * (size_t)nmember);
  const size_t npart = std::max<size_t>(2 * (size_t)g.nblk, (size_t)g.ls_blocks * g.jmax);
  L.prm = L.partial + align256(sizeof(uint32_t) * npart);
  const size_t nprm = g.ncol >= (uint32_t)kBlock ? (size_t)g.nblk * 2 * VJP_NPARAM : (size_t)g.ntot * VJP_NPARAM;
  L.bud = L.prm + align256(sizeof(float) * nprm);
  const size_t nbud = g.ncol >= (uint32_t)kBlock ? (size_t)g.nblk * 2 * kBudParts : (size_t)g.ntot * kBudParts;
  L.desc = L.bud + align256(sizeof(double) * nbud);
  // one-segment batch descriptor and its block tables, for grids the fused launch cannot take
  L.tbl = L.desc + align256(sizeof(KArgs));
  L.total = L.tbl + align256(sizeof(uint32_t) * ((size_t)g.nblk + g.ls_blocks + (size_t)nmember));
  return L;
}

constexpr uint32_t kMaxJ = 16384;   // LDS words of the level-set kernel's per-member ORs

Divisor make_divisor(uint32_t d) {
  Divisor dv;
  dv.d = d;
  dv.m = d > 1 ? (uint32_t)(((uint64_t)1 << 32) / d) : 0u;
  return dv;
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
  if (flags & ~(uint32_t)SPEEDY_FLAGS_ALL) return fail(SPEEDY_ERR_ARG, "speedy: unknown flag bits");
  if ((flags & SPEEDY_FLAG_VERIFY_TENDENCIES) && !(date->dt_seconds > 0.0f))
    return fail(SPEEDY_ERR_ARG, "speedy: VERIFY_TENDENCIES needs date->dt_seconds > 0");
  if (term_mask & ~(uint32_t)SPEEDY_TERMS_ALL) return fail(SPEEDY_ERR_ARG, "speedy: unknown term bits");
  if (level_set >= (1 << vert->kx)) return fail(SPEEDY_ERR_ARG, "speedy: level_set has bits beyond kx");
  if (int rc = check_fields(fields)) return rc;
  const int64_t ncol = (int64_t)nlon * nlat;
  const int64_t ntot = ncol * nmember;
  if (ntot >= ((int64_t)1 << 31) - kBlock) return fail(SPEEDY_ERR_CAPACITY, "speedy: more than 2^31 columns per call");
  const Grid gr = make_grid(nlon, nlat, nmember);
  if (gr.jmax > kMaxJ) return fail(SPEEDY_ERR_CAPACITY, "speedy: too many members per column chunk");
  if (gr.ls_blocks > kChunkLoads * kBlock)   // chunk_share's reach (make_grid keeps it <= 512)
    return fail(SPEEDY_ERR_CAPACITY, "speedy: too many level-set chunks");
  const WsLayout L = ws_layout(gr, nmember);
  if (!workspace) return fail(SPEEDY_ERR_WORKSPACE, "speedy: workspace is NULL");
  if (workspace_bytes < L.total)
    return fail(SPEEDY_ERR_WORKSPACE, "speedy: workspace too small (need " + std::to_string(L.total) + " bytes)");
  std::memset(&a, 0, sizeof(a));
  a.f = *fields;
  a.v = *vert;
  a.p = *params;
  make_uniforms(*params, *vert, *date, a.u);
  a.ncol = gr.ncol; a.ntot = gr.ntot; a.nmember = (uint32_t)nmember; a.nlat = (uint32_t)nlat;
  a.nblk = gr.nblk; a.ls_blocks = gr.ls_blocks; a.chunk = gr.chunk; a.jmax = gr.jmax;
  a.dcol = make_divisor(gr.ncol);
  a.dlat = make_divisor((uint32_t)nlat);
  a.dchunk = make_divisor(gr.chunk);
  a.level_set = level_set;
  a.flags = flags;
  a.term_mask = term_mask;
  char* w = static_cast<char*>(workspace);
  a.gen = reinterpret_cast<uint32_t*>(w);
  a.err = a.gen + 1;
  a.last_s = reinterpret_cast<uint32_t*>(w + L.last_s);
  a.partial = reinterpret_cast<uint32_t*>(w + L.partial);
  a.prm_part = reinterpret_cast<float*>(w + L.prm);
  a.bud_part = reinterpret_cast<double*>(w + L.bud);
  return SPEEDY_OK;
}

// Make sure the device's per-thread stack limit covers the kernel's private
// segment (register spills of the adjoint kernels: up to ~4 KB per lane at 16
// levels; HIP's default limit is 1 KB).  Called once per kernel before its
// first launch; raises the limit only.
template <typename K>
hipError_t ensure_stack(K kernel) {
  hipFuncAttributes fa;
  hipError_t e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kernel));
  if (e != hipSuccess) return e;
  size_t cur = 0;
  e = hipDeviceGetLimit(&cur, hipLimitStackSize);
  if (e != hipSuccess) return e;
  if ((size_t)fa.localSizeBytes > cur) e = hipDeviceSetLimit(hipLimitStackSize, (size_t)fa.localSizeBytes);
  return e;
}

// The stack limit the batch kernels of one level count need.
template <int KX>
hipError_t batch_stack() {
  hipError_t e = ensure_stack(fwd_batch_kernel<KX>);
  if (e == hipSuccess) e = ensure_stack(fwd_only_batch_kernel<KX>);
  if (e == hipSuccess) e = ensure_stack(fwd_plain_batch_kernel<KX>);
  if (e == hipSuccess) e = ensure_stack(fwd_diag_batch_kernel<KX>);
  return e;
}

// Largest grid that is co-resident on the current device for a kernel.
template <typename K>
int64_t coresident_blocks(K kernel) {
  int dev = 0, ncu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess) return 0;
  return (int64_t)per_cu * ncu;
}

// One-time set-up of a kernel family: the stack limit its kernels need and the
// co-resident grid of its fused kernel.  A function-local static, so the first
// calls from several threads initialise it once (one GPU per process in this
// library's deployments).
struct LaunchInit {
  hipError_t err;
  int64_t cap;
};

template <int KX, bool DIAG, bool VJP, bool EXTRA>
int launch_fwd(const KArgs& a, hipStream_t s) {
  static const LaunchInit init = [] {
    hipError_t se = ensure_stack(fwd_kernel<KX, DIAG, true, VJP, EXTRA>);
    if (se == hipSuccess)
      se = VJP    ? ensure_stack(fwd_batch_kernel<KX>)
           : DIAG ? ensure_stack(fwd_diag_batch_kernel<KX>)
           : EXTRA ? ensure_stack(fwd_only_batch_kernel<KX>) : ensure_stack(fwd_plain_batch_kernel<KX>);
    return LaunchInit{se, se == hipSuccess ? coresident_blocks(fwd_kernel<KX, DIAG, true, VJP, EXTRA>) : 0};
  }();
  if (init.err != hipSuccess)
    return fail(SPEEDY_ERR_HIP, std::string("speedy: stack limit: ") + hipGetErrorString(init.err));
  const int64_t cap = init.cap;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(s, &cs);
  // Sixteen-level forward + VJP calls always take the batch route: at sixteen
  // levels no one-thread-per-column kernel fits the register file, and the fused
  // sixteen-level forward + VJP kernel (3.3 KB of scratch per lane) gave wrong
  // gradients on gfx950 in a case whose host build is right; the batch kernel
  // passes every sixteen-level case (cause of the fused kernel's error not found).
  const bool fused = a.level_set < 0 && (int64_t)a.nblk <= cap && a.ncol >= (uint32_t)kBlock && !(VJP && KX == 16) &&
                     !(a.flags & kFlagBatchRoute);
  const int kind = VJP ? 2 : 1;
  hipError_t e = hipSuccess;
  if (fused) {
    const int slot = timer_begin(s, kind);
    if (cs == hipStreamCaptureStatusNone && !(a.flags & SPEEDY_FLAG_PLAIN_LAUNCH)) {
      void* args[] = {const_cast<KArgs*>(&a)};
      e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&fwd_kernel<KX, DIAG, true, VJP, EXTRA>),
                                     dim3(a.nblk),
                                     dim3(kBlock), args, 0, s);
    } else {
      // Graph capture (or SPEEDY_FLAG_PLAIN_LAUNCH): a plain launch of a grid
      // that fits on the device; the level-set poll is bounded, so the kernel
      // terminates in every case.
      hipLaunchKernelGGL((fwd_kernel<KX, DIAG, true, VJP, EXTRA>), dim3(a.nblk), dim3(kBlock), 0, s, a);
      e = hipGetLastError();
    }
    timer_end(s, slot);
  } else {
    // Grids the fused launch cannot take (beyond one co-resident launch, members
    // narrower than a workgroup, or a given level set) run as a one-segment batch
    // through the batch kernels; the workspace holds the descriptor and tables.
    if (a.nblk > kSegLocalMask || a.ls_blocks > kSegLocalMask)
      return fail(SPEEDY_ERR_CAPACITY, "speedy: more than 2^20 workgroups (2^28 columns) in one call");
    const WsLayout L = ws_layout(make_grid(a.ncol, 1, a.nmember), (int32_t)a.nmember);
    char* w = reinterpret_cast<char*>(a.gen);
    KArgs* d_args = reinterpret_cast<KArgs*>(w + L.desc);
    uint32_t* tbl = reinterpret_cast<uint32_t*>(w + L.tbl);
    const uint32_t nls = a.level_set < 0 ? a.ls_blocks : 0u;
    const uint32_t nred = (VJP || a.f.budget) ? a.nmember : 0u;
    const uint32_t n = a.nblk + nls + nred;
    hipLaunchKernelGGL(seg_desc_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, a, d_args, tbl, a.nblk,
                       nls, nred);
    if (nls) {
      const int slot = timer_begin(s, 0);
      hipLaunchKernelGGL(levelset_batch_kernel<KX>, dim3(nls), dim3(kBlock), a.jmax * sizeof(uint32_t), s,
                         (const KArgs*)d_args, (const uint32_t*)(tbl + a.nblk));
      timer_end(s, slot);
    }
    const int slot = timer_begin(s, kind);
    if (VJP)
      hipLaunchKernelGGL(fwd_batch_kernel<KX>, dim3(a.nblk), dim3(kBlock), 0, s, (const KArgs*)d_args,
                         (const uint32_t*)tbl);
    else if (DIAG)
      hipLaunchKernelGGL(fwd_diag_batch_kernel<KX>, dim3(a.nblk), dim3(kBlock), 0, s, (const KArgs*)d_args,
                         (const uint32_t*)tbl);
    else if (EXTRA)
      hipLaunchKernelGGL(fwd_only_batch_kernel<KX>, dim3(a.nblk), dim3(kBlock), 0, s, (const KArgs*)d_args,
                         (const uint32_t*)tbl);
    else
      hipLaunchKernelGGL(fwd_plain_batch_kernel<KX>, dim3(a.nblk), dim3(kBlock), 0, s, (const KArgs*)d_args,
                         (const uint32_t*)tbl);
    timer_end(s, slot);
    if (nred) {
      const int slot2 = timer_begin(s, 3);
      hipLaunchKernelGGL(param_reduce_batch_kernel, dim3(nred), dim3(kBlock), 0, s, (const KArgs*)d_args,
                         (const uint32_t*)(tbl + a.nblk + nls));
      timer_end(s, slot2);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy: launch failed: ") + hipGetErrorString(e));
    return SPEEDY_OK;
  }
  // Per-member sums in a second small kernel.  (Summing them in the last block to
  // finish instead, after an agent-scope fence per block, cost more than this
  // launch: the fences write back the L2 caches.)
  if (e == hipSuccess && (VJP || a.f.budget)) {
    const int slot = timer_begin(s, 3);
    hipLaunchKernelGGL(param_reduce_kernel, dim3(a.nmember), dim3(kBlock), 0, s, a);
    timer_end(s, slot);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy: launch failed: ") + hipGetErrorString(e));
  return SPEEDY_OK;
}

bool needs_extra(const KArgs& a) {
  return (a.flags & (SPEEDY_FLAG_VERIFY_STATE | SPEEDY_FLAG_VERIFY_TENDENCIES)) || a.f.budget != nullptr;
}

template <int KX>
int launch_fwd_kx(const KArgs& a, hipStream_t s) {
  if (needs_extra(a))
    return a.f.diag ? launch_fwd<KX, true, false, true>(a, s) : launch_fwd<KX, false, false, true>(a, s);
  return a.f.diag ? launch_fwd<KX, true, false, false>(a, s) : launch_fwd<KX, false, false, false>(a, s);
}

// The forward + VJP always runs the instantiation that carries the optional
// verify / budget code (run-time flags).  The instantiation without it gave
// wrong state and parameter gradients at 16 levels on gfx950 (deterministic;
// the same source as a host build, and the instantiation with the optional
// code, agree with the float64 adjoint), so it is not built.
template <int KX>
int launch_fwd_vjp_kx(const KArgs& a, hipStream_t s) {
  return launch_fwd<KX, false, true, true>(a, s);
}

struct BatchLayout {
  size_t blk, ls, red, total;
};

int batch_layout(const SpeedySegment* segs, int32_t nseg, BatchLayout& L) {
  size_t nb = 0, nl = 0, nr = 0;
  for (int32_t i = 0; i < nseg; ++i) {
    const SpeedySegment& sg = segs[i];
    if (sg.nlon <= 0 || sg.nlat <= 0 || sg.nmember <= 0 ||
        (int64_t)sg.nlon * sg.nlat * sg.nmember >= ((int64_t)1 << 31) - kBlock)
      return fail(SPEEDY_ERR_ARG, "speedy_fwd_batch: bad grid in segment " + std::to_string(i));
    const Grid g = make_grid(sg.nlon, sg.nlat, sg.nmember);
    if (g.nblk > kSegLocalMask || g.ls_blocks > kSegLocalMask)
      return fail(SPEEDY_ERR_CAPACITY, "speedy_fwd_batch: segment " + std::to_string(i) + " too large for a batch");
    nb += g.nblk;
    if (sg.level_set < 0) nl += g.ls_blocks;
    if (sg.adj.g_t || sg.fields.budget) nr += (size_t)sg.nmember;
  }
  if (nb > 0x7fffffffu) return fail(SPEEDY_ERR_CAPACITY, "speedy_fwd_batch: too many blocks");
  L.blk = align256(sizeof(KArgs) * (size_t)nseg);
  L.ls = L.blk + align256(sizeof(uint32_t) * nb);
  L.red = L.ls + align256(sizeof(uint32_t) * nl);
  L.total = L.red + align256(sizeof(uint32_t) * (nr ? nr : 1));
  return SPEEDY_OK;
}

}  // namespace speedy

using namespace speedy;

extern "C" {

int speedy_abi_version(void) { return SPEEDY_ABI_VERSION; }

size_t speedy_workspace_bytes(int32_t nlon, int32_t nlat, int32_t nmember) {
  if (nlon <= 0 || nlat <= 0 || nmember <= 0) return 0;
  if ((int64_t)nlon * nlat * nmember >= ((int64_t)1 << 31) - kBlock) return 0;
  return ws_layout(make_grid(nlon, nlat, nmember), nmember).total;
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

int speedy_workspace_status(void* workspace, size_t workspace_bytes, uint32_t* status, void* stream) {
  if (!workspace || workspace_bytes < 256 || !status) return fail(SPEEDY_ERR_ARG, "speedy: bad arguments");
  uint32_t* err = static_cast<uint32_t*>(workspace) + 1;   // word 1: KArgs::err
  hipError_t e = hipMemcpyAsync(status, err, sizeof(uint32_t), hipMemcpyDeviceToHost, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e == hipSuccess && *status) {
    // a failed wait leaves late blocks' level-set words and the generation out of
    // step: the whole workspace starts again from zero
    const int64_t n = (int64_t)(workspace_bytes / 4);
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       static_cast<uint32_t*>(workspace), n);
    e = hipGetLastError();
  }
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy: status read failed: ") + hipGetErrorString(e));
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
    case 7: return launch_fwd_kx<7>(a, s);
    case 8: return launch_fwd_kx<8>(a, s);
    default: return launch_fwd_kx<16>(a, s);
  }
}

int speedy_fwd_vjp(const SpeedyFields* fields, const SpeedyAdjoint* adj, const SpeedyVertical* vert,
                   const SpeedyParams* params, const SpeedyDate* date, int32_t nlon, int32_t nlat, int32_t nmember,
                   uint32_t flags, int32_t level_set, void* workspace, size_t workspace_bytes, void* stream) {
  if (!adj) return fail(SPEEDY_ERR_ARG, "speedy: NULL adjoint struct");
  const void* req[] = {adj->ct_du, adj->ct_dv, adj->ct_dt, adj->ct_dq, adj->g_u, adj->g_v,
                       adj->g_t, adj->g_q, adj->g_ps, adj->g_params};
  for (const void* p : req)
    if (!p) return fail(SPEEDY_ERR_ARG, "speedy: a required adjoint pointer is NULL");
  if (fields && fields->diag) return fail(SPEEDY_ERR_ARG, "speedy: fwd_vjp writes no diagnostics (diag must be NULL)");
  KArgs a;
  if (int rc = prepare(fields, vert, params, date, nlon, nlat, nmember, flags, SPEEDY_TERMS_ALL, level_set,
                       workspace, workspace_bytes, a))
    return rc;
  a.adj = *adj;
  g_last_error.clear();
  hipStream_t s = (hipStream_t)stream;
  switch (vert->kx) {
    case 7: return launch_fwd_vjp_kx<7>(a, s);
    case 8: return launch_fwd_vjp_kx<8>(a, s);
    default: return launch_fwd_vjp_kx<16>(a, s);
  }
}

int speedy_fwd_batch(const SpeedySegment* segs, int32_t nseg, const SpeedyVertical* vert,
                     const SpeedyParams* params, void* desc, size_t desc_bytes, void* host_staging, void* stream) {
  if (!segs || nseg <= 0 || nseg > SPEEDY_BATCH_MAX_SEGMENTS || !vert || !params)
    return fail(SPEEDY_ERR_ARG, "speedy_fwd_batch: bad segment list");
  if (vert->kx != 7 && vert->kx != 8 && vert->kx != 16)
    return fail(SPEEDY_ERR_KX, "speedy_fwd_batch: kx must be 7, 8 or 16");
  BatchLayout L;
  if (int rc = batch_layout(segs, nseg, L)) return rc;
  if (!desc || !host_staging || desc_bytes < L.total)
    return fail(SPEEDY_ERR_WORKSPACE, "speedy_fwd_batch: descriptor buffers missing or too small (need " +
                                          std::to_string(L.total) + " bytes)");
  char* h = static_cast<char*>(host_staging);
  KArgs* ha = reinterpret_cast<KArgs*>(h);
  uint32_t* t_blk = reinterpret_cast<uint32_t*>(h + L.blk);
  uint32_t* t_ls = reinterpret_cast<uint32_t*>(h + L.ls);
  uint32_t* t_red = reinterpret_cast<uint32_t*>(h + L.red);
  uint32_t nb = 0, nl = 0, nr = 0, jmax = 1;
  bool fwd_extra = false;   // a forward-only segment needs the verify steps or budgets
  std::vector<uint8_t> is_vjp((size_t)nseg, 0u);
  for (int32_t i = 0; i < nseg; ++i) {
    const SpeedySegment& sg = segs[i];
    if (sg.fields.diag) return fail(SPEEDY_ERR_ARG, "speedy_fwd_batch: segments write no diagnostics");
    const void* ap[] = {sg.adj.ct_du, sg.adj.ct_dv, sg.adj.ct_dt, sg.adj.ct_dq, sg.adj.g_u,
                        sg.adj.g_v, sg.adj.g_t, sg.adj.g_q, sg.adj.g_ps, sg.adj.g_params};
    int nset = 0;
    for (const void* q : ap) nset += q != nullptr;
    if (nset != 0 && nset != 10)
      return fail(SPEEDY_ERR_ARG, "speedy_fwd_batch: segment " + std::to_string(i) + " sets only some adjoint pointers");
    KArgs a;
    if (int rc = prepare(&sg.fields, vert, params, &sg.date, sg.nlon, sg.nlat, sg.nmember, sg.flags, SPEEDY_TERMS_ALL,
                         sg.level_set, sg.workspace, sg.workspace_bytes, a))
      return rc;
    a.adj = sg.adj;
    std::memcpy(&ha[i], &a, sizeof(KArgs));
    is_vjp[i] = nset != 0;
    if (!nset && needs_extra(a)) fwd_extra = true;
    if (a.level_set < 0) {
      for (uint32_t b = 0; b < a.ls_blocks; ++b) t_ls[nl++] = ((uint32_t)i << kSegShift) | b;
      jmax = std::max(jmax, a.jmax);
    }
    if (nset || sg.fields.budget)
      for (uint32_t m = 0; m < a.nmember; ++m) t_red[nr++] = ((uint32_t)i << kSegShift) | m;
  }
  // block table: the forward-only segments' blocks first, then the VJP segments'
  for (int pass = 0; pass < 2; ++pass)
    for (int32_t i = 0; i < nseg; ++i)
      if (is_vjp[i] == pass)
        for (uint32_t b = 0; b < ha[i].nblk; ++b) t_blk[nb++] = ((uint32_t)i << kSegShift) | b;
  uint32_t nbf = 0;
  for (int32_t i = 0; i < nseg; ++i)
    if (!is_vjp[i]) nbf += ha[i].nblk;
  const uint32_t nbv = nb - nbf;
  g_last_error.clear();
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(desc, host_staging, L.total, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy_fwd_batch: copy failed: ") + hipGetErrorString(e));
  const KArgs* d_args = static_cast<const KArgs*>(desc);
  const char* d = static_cast<const char*>(desc);
  const uint32_t* d_blk = reinterpret_cast<const uint32_t*>(d + L.blk);
  const uint32_t* d_ls = reinterpret_cast<const uint32_t*>(d + L.ls);
  const uint32_t* d_red = reinterpret_cast<const uint32_t*>(d + L.red);
  // The stack limit of each level count's batch kernels, raised once (a
  // function-local static: thread-safe); a failure fails every batch call.
  static const hipError_t stack_err[3] = {
      batch_stack<7>(), batch_stack<8>(), batch_stack<16>()};
  auto launch = [&](auto ls_kernel, auto fwd_only_b, auto fwd_plain_b, auto fwd_kernel_b, int kxi) -> hipError_t {
    if (stack_err[kxi] != hipSuccess) return stack_err[kxi];
    if (nl) {
      const int slot = timer_begin(s, 0);
      hipLaunchKernelGGL(ls_kernel, dim3(nl), dim3(kBlock), jmax * sizeof(uint32_t), s, d_args, d_ls);
      timer_end(s, slot);
    }
    if (nbf) {
      const int slot = timer_begin(s, 1);
      if (fwd_extra)
        hipLaunchKernelGGL(fwd_only_b, dim3(nbf), dim3(kBlock), 0, s, d_args, d_blk);
      else
        hipLaunchKernelGGL(fwd_plain_b, dim3(nbf), dim3(kBlock), 0, s, d_args, d_blk);
      timer_end(s, slot);
    }
    if (nbv) {
      const int slot = timer_begin(s, 2);
      hipLaunchKernelGGL(fwd_kernel_b, dim3(nbv), dim3(kBlock), 0, s, d_args, d_blk + nbf);
      timer_end(s, slot);
    }
    if (nr) {
      const int slot2 = timer_begin(s, 3);
      hipLaunchKernelGGL(param_reduce_batch_kernel, dim3(nr), dim3(kBlock), 0, s, d_args, d_red);
      timer_end(s, slot2);
    }
    return hipSuccess;
  };
  hipError_t le;
  switch (vert->kx) {
    case 7:
      le = launch(levelset_batch_kernel<7>, fwd_only_batch_kernel<7>, fwd_plain_batch_kernel<7>, fwd_batch_kernel<7>, 0);
      break;
    case 8:
      le = launch(levelset_batch_kernel<8>, fwd_only_batch_kernel<8>, fwd_plain_batch_kernel<8>, fwd_batch_kernel<8>, 1);
      break;
    default:
      le = launch(levelset_batch_kernel<16>, fwd_only_batch_kernel<16>, fwd_plain_batch_kernel<16>,
                  fwd_batch_kernel<16>, 2);
      break;
  }
  if (le != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy_fwd_batch: stack limit: ") + hipGetErrorString(le));
  e = hipGetLastError();
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy_fwd_batch: launch failed: ") + hipGetErrorString(e));
  return SPEEDY_OK;
}

size_t speedy_batch_desc_bytes(const SpeedySegment* segs, int32_t nseg) {
  if (!segs || nseg <= 0 || nseg > SPEEDY_BATCH_MAX_SEGMENTS) return 0;
  BatchLayout L;
  if (batch_layout(segs, nseg, L)) return 0;
  return L.total;
}

int speedy_read_level_sets(const void* workspace, size_t workspace_bytes, int32_t nmember, uint32_t* out,
                           void* stream) {
  if (!workspace || !out || nmember <= 0) return fail(SPEEDY_ERR_ARG, "speedy: bad arguments");
  if (workspace_bytes < 256 + align256(sizeof(uint32_t) * (size_t)nmember))
    return fail(SPEEDY_ERR_WORKSPACE, "speedy: workspace too small");
  hipError_t e = hipMemcpyAsync(out, static_cast<const char*>(workspace) + 256, sizeof(uint32_t) * nmember,
                                hipMemcpyDeviceToHost, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy: copy failed: ") + hipGetErrorString(e));
  return SPEEDY_OK;
}

int speedy_diag_accumulate(const float* diag, float* acc, int64_t n, float steps_to_average, void* stream) {
  if (n < 0 || !(steps_to_average > 0.0f)) return fail(SPEEDY_ERR_ARG, "speedy_diag_accumulate: bad n or steps");
  if (n == 0) return SPEEDY_OK;
  if (!diag || !acc) return fail(SPEEDY_ERR_ARG, "speedy_diag_accumulate: NULL buffer");
  const int vec = ((reinterpret_cast<uintptr_t>(diag) | reinterpret_cast<uintptr_t>(acc)) % 16u) == 0u;
  const int64_t work = vec ? (n + 3) / 4 : n;
  const int64_t blocks = std::min<int64_t>((work + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(diag_accumulate_kernel, dim3((uint32_t)blocks), dim3(kBlock), 0, (hipStream_t)stream, diag, acc,
                     n, steps_to_average, vec);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy_diag_accumulate: ") + hipGetErrorString(e));
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

// ---------------------------------------------------------------- allocation-free forward
// speedy_fwd_noalloc: speedy_fwd with its workspace in this file-scope device
// array instead of a caller's buffer, for callers that cannot allocate.  Each GPU
// holds its own copy for the life of the process; its device address is looked
// up once per device (hipGetSymbolAddress) and only that address is passed to
// kernels.  Calls take the one-segment batch route, which reads no word of the
// workspace that the same call has not written first (descriptor, level-set
// words, partials), so the contents left by an earlier call, of any grid, are
// never read and nothing needs clearing.  Calls that overlap on one GPU share
// the array and race.
namespace speedy {
constexpr size_t kNoAllocScratchBytes = (size_t)4 << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};
}  // namespace speedy

extern "C" size_t speedy_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

extern "C" int speedy_fwd_noalloc(const SpeedyFields* fields, const SpeedyVertical* vert, const SpeedyParams* params,
                                  const SpeedyDate* date, int32_t nlon, int32_t nlat, int32_t nmember, uint32_t flags,
                                  uint32_t term_mask, int32_t level_set, void* stream) {
  const size_t need = speedy_workspace_bytes(nlon, nlat, nmember);
  if (need == 0) return fail(SPEEDY_ERR_ARG, "speedy_fwd_noalloc: bad grid");
  if (need > kNoAllocScratchBytes)
    return fail(SPEEDY_ERR_CAPACITY, "speedy_fwd_noalloc: the call needs " + std::to_string(need) +
                                         " bytes of workspace, above the " + std::to_string(kNoAllocScratchBytes) +
                                         "-byte static scratch; use speedy_fwd with a caller workspace");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy_fwd_noalloc: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(SPEEDY_ERR_CAPACITY, "speedy_fwd_noalloc: device index");
  void* ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(SPEEDY_ERR_HIP, std::string("speedy_fwd_noalloc: scratch address: ") + hipGetErrorString(e));
      g_noalloc_ptr[dev] = p;
    }
    ws = g_noalloc_ptr[dev];
  }
  KArgs a;
  if (int rc = prepare(fields, vert, params, date, nlon, nlat, nmember, flags, term_mask, level_set, ws,
                       kNoAllocScratchBytes, a))
    return rc;
  a.flags |= kFlagBatchRoute;
  g_last_error.clear();
  hipStream_t s = (hipStream_t)stream;
  switch (vert->kx) {
    case 7: return launch_fwd_kx<7>(a, s);
    case 8: return launch_fwd_kx<8>(a, s);
    default: return launch_fwd_kx<16>(a, s);
  }
}
