// Dynamics <-> physics transforms for gfx950: spherical-harmonic analysis and
// synthesis on a Gaussian grid as batched fp32 MFMA GEMMs, with the unit
// conversions, the modal vorticity/divergence <-> wind operators and the
// hydrostatic geopotential fused around them.
//
// Reference behaviour: jcm/physics_interface.py:196-303 over dinosaur's
// RealSphericalHarmonics (Grid.T<N>, jcm/utils.py:28-54).
//
// Data flow of one conversion (B stacked fields, Mm = 2*n_m-1 modes):
//   to_nodal:  modal [B][Mm][L] --Legendre--> Y [B][Mm][nlat] --Fourier--> nodal [B][nlon][nlat]
//   to_modal:  nodal [B][nlon][nlat] --Fourier--> Y [B][Mm][nlat] --Legendre--> modal [B][Mm][L]
// Each Legendre arrow is one launch of sht_gemm_kernel<STAGE>: a 64x64 output
// tile per workgroup of four waves, each wave 32x32 as 2x2
// v_mfma_f32_16x16x4f32, K in steps of 32 staged through two LDS stages,
// batched over modes (the basis P[m] is shared by every field).  Each Fourier
// arrow is one launch of fourier_analysis_kernel / fourier_synthesis_kernel: a
// real DFT over longitude factored into two short passes in LDS, batched over
// fields and latitude tiles.  Per-field pre/post operations (1/cos(lat), unit
// scales, reference temperature, log/exp of surface pressure) ride in the
// Fourier stages' loads and stores; the l +- 1 modal operators are one
// elementwise kernel on either side.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "speedy_sht.h"

struct SpeedyShtPlan {
  int32_t n_m, n_l, nlon, nlat, mm;   // mm = 2*n_m-1 modes
  double radius;
  int device;
  int32_t n1, n2;      // nlon = n1 * n2 (the Fourier stages' two passes)
  int32_t tl_a, tl_s;  // latitudes per workgroup of the Fourier analysis / synthesis
  float* P;     // [n_m][nlat][n_l]
  float* w;     // [nlat] Gauss weights * 2 pi / nlon
  float* sec;   // [nlat] 1 / cos(lat)
  float* tw;    // [2][nlon]: cos(2 pi j / nlon), -sin(2 pi j / nlon)
};

namespace sht {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

constexpr int kThreads = 256, kBM = 64, kBN = 64, kBK = 32, kPad = 4;
constexpr int kMaxSeg = 8;

enum Stage { kLegendreAnalysis = 1, kLegendreSynthesis = 2 };
enum : int32_t { kOpLog = 1, kOpExp = 2, kOpSecLat = 4, kOpSubLevel = 8, kOpAddLevel = 16 };

// A run of stacked fields that share a source (Fourier analysis) or a
// destination (Fourier synthesis) and one operation.
struct Seg {
  const float* src;
  float* dst;
  const float* level;   // per-field constant, index (f - first) % period, or null
  int32_t first, count, period, op;
  float mul;
};
struct SegTable {
  Seg s[kMaxSeg];
  int32_t n;
};

struct StageArgs {
  int32_t nlon, nlat, mm, nl, nfield;
  const float* P;
  const float* in;    // LA: Y, LS: modal X
  float* out;         // LA: modal, LS: Y halves
};

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int find_seg(const SegTable& t, int f) {
  int k = 0;
  for (int i = 1; i < t.n; ++i)
    if (f >= t.s[i].first) k = i;
  return k;
}

// C[M x N] = A[M x K] * B[K x N] for one batch entry (blockIdx.y); element
// (r, c) of X at X[r * x_r + c * x_c].
template <int STAGE>
__global__ __launch_bounds__(kThreads) void sht_gemm_kernel(StageArgs a) {
  // blockIdx.y = 2 * mode + parity.  Gaussian latitudes are
  // symmetric (x_{nlat-1-j} = -x_j) and P_l^m(-x) = (-1)^(l+m) P_l^m(x), so each
  // mode splits into the l with l+m even (parity 0) and odd (parity 1), each over
  // the southern half of the latitudes: half the work and half the K of a full
  // contraction.  Analysis reads Y_j +- Y_{nlat-1-j}; synthesis writes the even
  // and odd partial sums E, O for the southern half ([parity][field][mode][h]),
  // which the Fourier synthesis combines as E + O (south) and E - O (north).
  const int b = (int)(blockIdx.y >> 1);
  const int par = (int)(blockIdx.y & 1u);
  const int h = a.nlat >> 1;
  int M, N, K;
  const float* A;
  const float* B;
  float* C;
  long a_m, a_k, b_k, b_n, c_m, c_n = 1;
  const int mode_m = (b + 1) >> 1;    // order m of mode b
  const int lofs = (mode_m + par) & 1;   // first l of this parity (l = lofs + 2c)
  const int npar = (a.nl - lofs + 1) >> 1;
  if (STAGE == kLegendreAnalysis) {
    // Z[:, b, lofs + 2c] [nfield x npar] = (Y_j +- Y_{nlat-1-j})[:, b, :h] [nfield x h] * P_m[:h, lofs + 2c]
    M = a.nfield; N = npar; K = h;
    A = a.in + (size_t)b * a.nlat; a_m = (long)a.mm * a.nlat; a_k = 1;
    B = a.P + (size_t)mode_m * a.nlat * a.nl + lofs; b_k = a.nl; b_n = 2;
    C = a.out + (size_t)b * a.nl + lofs; c_m = (long)a.mm * a.nl; c_n = 2;
  } else {
    // E|O[par][:, b, :h] [nfield x h] = X[:, b, lofs + 2c] [nfield x npar] * P_m[:h, lofs + 2c]^T
    M = a.nfield; N = h; K = npar;
    A = a.in + (size_t)b * a.nl + lofs; a_m = (long)a.mm * a.nl; a_k = 2;
    B = a.P + (size_t)mode_m * a.nlat * a.nl + lofs; b_k = 2; b_n = a.nl;
    C = a.out + (size_t)par * a.nfield * a.mm * h + (size_t)b * h; c_m = (long)a.mm * h;
  }
  const float sgn = par ? -1.0f : 1.0f;
  constexpr bool kBNContig = STAGE != kLegendreSynthesis;
  constexpr int kLoads = kBM * kBK / kThreads;

  const int tiles_n = (N + kBN - 1) / kBN;
  const int tiles_m = (M + kBM - 1) / kBM;
  if ((int)blockIdx.x >= tiles_m * tiles_n) return;    // uniform per workgroup
  const int m0 = (blockIdx.x / tiles_n) * kBM, n0 = (blockIdx.x % tiles_n) * kBN;

  // two LDS stages: the global loads of K-step s+1 are in flight (in registers)
  // while the MFMAs of step s read the other stage; one barrier per step.  Tiles
  // are stored [row][k] (row stride kLdk floats).  Within a step the k order is
  // permuted: lane group g = lane / 16 takes k = 8 g + q at substep q, for A and B
  // alike, so each MFMA still pairs equal k and a lane's eight operands of a step
  // are consecutive: two 16-byte LDS reads per operand row, conflict-free at this
  // stride (16 lanes x 4 banks cover all banks once).
  static_assert(kBK == 32, "the k permutation assumes 32-wide K steps");
  constexpr int kLdk = kBK + kPad;
  __shared__ __align__(16) float As[2][kBM][kLdk];
  __shared__ __align__(16) float Bs[2][kBN][kLdk];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

  float ra[kLoads], rb[kLoads];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < kLoads; ++r) {
      const int e = t + kThreads * r;
      const int mm = e / kBK;
      const int kk = e % kBK;
      // clamped indices (always in bounds), no condition: the loads issue back to
      // back; elements outside the matrix are zeroed when they go to LDS
      const long row = min(m0 + mm, M - 1) * a_m;
      const int ck = min(k0 + kk, K - 1);
      if (STAGE == kLegendreAnalysis)
        ra[r] = A[row + ck] + sgn * A[row + (a.nlat - 1 - ck)];
      else
        ra[r] = A[row + ck * a_k];
    }
#pragma unroll
    for (int r = 0; r < kLoads; ++r) {
      const int e = t + kThreads * r;
      const int nn = kBNContig ? (e % kBN) : (e / kBK);
      const int kk = kBNContig ? (e / kBN) : (e % kBK);
      const int ck = min(k0 + kk, K - 1), cn = min(n0 + nn, N - 1);
      rb[r] = B[ck * b_k + cn * b_n];
    }
  };
  auto store = [&](int st, int k0) {
#pragma unroll
    for (int r = 0; r < kLoads; ++r) {
      const int e = t + kThreads * r;
      const int am = e / kBK;
      const int ak = e % kBK;
      As[st][am][ak] = (m0 + am < M && k0 + ak < K) ? ra[r] : 0.0f;
      const int nn = kBNContig ? (e % kBN) : (e / kBK);
      const int kk = kBNContig ? (e / kBN) : (e % kBK);
      Bs[st][nn][kk] = (k0 + kk < K && n0 + nn < N) ? rb[r] : 0.0f;
    }
  };
  load(0);
  store(0, 0);
  __syncthreads();
  int st = 0;
  for (int k0 = 0; k0 < K; k0 += kBK) {
    const bool more = k0 + kBK < K;
    if (more) load(k0 + kBK);
    // all operands of this step out of LDS first, then the MFMAs
    const int g8 = (lane >> 4) * 8, r16 = lane & 15;
    float4 av[2][2], bv[2][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i][h] = *reinterpret_cast<const float4*>(&As[st][wm + 16 * i + r16][g8 + 4 * h]);
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j][h] = *reinterpret_cast<const float4*>(&Bs[st][wn + 16 * j + r16][g8 + 4 * h]);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i][q >> 2][q & 3], bv[j][q >> 2][q & 3], acc[i][j],
                                                           0, 0, 0);
    // keep the LDS reads ahead of the MFMAs
    __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 32, 0);
    if (more) store(st ^ 1, k0 + kBK);
    __syncthreads();
    st ^= 1;
  }
  // lane holds C[16 i + 4 (lane >> 4) + r][16 j + (lane & 15)] of its wave's 32 x 32
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gn = n0 + wn + 16 * j + (lane & 15);
      if (gn >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        if (gm >= M) continue;
        C[gm * c_m + gn * c_n] = acc[i][j][r];
      }
    }
}

// ------------------------------------------------------------ Fourier stages
// The longitude contraction is a real DFT of length N = nlon truncated to the
// n_m lowest wavenumbers, computed in two short passes (Cooley-Tukey with
// N = n1 * n2, longitude i = i1 + n1 * i2, wavenumber k = k2 + n2 * k1):
//   X_k = sum_i1 w^(i1 k) Z[i1][k mod n2],   Z[i1][k2] = sum_i2 x[i1 + n1 i2] w_n2^(i2 k2),
// with w = exp(-2 pi i / N) and w_n2 = w^n1.  Z of a real line is Hermitian in
// k2, so only k2 <= n2 / 2 is formed.  The modal coefficients are the basis's
// projections: mode 0 = X_0 / sqrt(2 pi), mode 2k-1 = Re X_k / sqrt(pi), mode
// 2k = -Im X_k / sqrt(pi).  Synthesis runs the same passes backwards.  About
// 4x fewer flops than the dense contraction with the basis, and two passes of
// length n1 and n2 instead of one of length nlon.  A workgroup takes one field
// and `tl` latitudes (latitude fastest, as in memory); the line tile, the
// intermediate Z or W and the N roots of unity live in LDS.
struct FourierArgs {
  int32_t nlon, nlat, mm, n_m, n1, n2, tl, nfield;
  const float* twr;     // [nlon] cos(2 pi j / nlon)
  const float* twi;     // [nlon] -sin(2 pi j / nlon)
  const float* w;       // [nlat] Gauss weights * 2 pi / nlon
  const float* sec;     // [nlat]
  const float* in;      // synthesis: E / O halves [2][nfield][mm][nlat / 2]
  float* out;           // analysis: Y [nfield][mm][nlat]
  SegTable seg;
};

constexpr float kRsqrt2Pi = 0.3989422804014327f;   // 1 / sqrt(2 pi)
constexpr float kRsqrtPi = 0.5641895835477563f;    // 1 / sqrt(pi)

__device__ __forceinline__ void load_roots(const FourierArgs& a, float* twr, float* twi) {
  for (int i = threadIdx.x; i < a.nlon; i += kThreads) {
    twr[i] = a.twr[i];
    twi[i] = a.twi[i];
  }
}

// Y_b[mode][lat] = sum_lon F[lon][mode] * prologue(x_b[lon][lat]) for one field b
// (blockIdx.y) and latitudes j0 .. j0 + tl - 1 (blockIdx.x).
__global__ __launch_bounds__(kThreads) void fourier_analysis_generic(FourierArgs a) {
  extern __shared__ float lds[];
  const int N = a.nlon, n1 = a.n1, n2 = a.n2, nh = n2 / 2 + 1, tl = a.tl;
  const int b = blockIdx.y, j0 = blockIdx.x * tl;
  float* twr = lds;
  float* twi = twr + N;
  float* xs = twi + N;            // [N][tl]
  float* zr = xs + N * tl;        // [n1][nh][tl]
  float* zi = zr + n1 * nh * tl;
  const Seg sg = a.seg.s[find_seg(a.seg, b)];
  const float* src = sg.src + (size_t)(b - sg.first) * N * a.nlat;
  const float lvl = sg.level ? sg.level[(b - sg.first) % sg.period] : 0.0f;
  load_roots(a, twr, twi);
  for (int e = threadIdx.x; e < N * tl; e += kThreads) {
    const int lon = e / tl, jj = e - lon * tl, j = j0 + jj;
    float v = 0.0f;
    if (j < a.nlat) {
      v = src[(size_t)lon * a.nlat + j];
      if (sg.op & kOpLog) v = logf(v);
      if (sg.op & kOpSubLevel) v -= lvl;
      v *= sg.mul * a.w[j] * ((sg.op & kOpSecLat) ? a.sec[j] : 1.0f);
    }
    xs[e] = v;
  }
  __syncthreads();
  // pass 1: Z[i1][k2] = sum_i2 x[i1 + n1 i2] w^(n1 i2 k2), k2 <= n2 / 2
  for (int e = threadIdx.x; e < n1 * nh * tl; e += kThreads) {
    const int r = e / tl, jj = e - r * tl, k2 = r % nh, i1 = r / nh;
    float re = 0.0f, im = 0.0f;
    int q = 0;   // (i2 * k2) mod n2
    for (int i2 = 0; i2 < n2; ++i2) {
      const float x = xs[(i1 + n1 * i2) * tl + jj];
      re += x * twr[q * n1];
      im += x * twi[q * n1];
      q += k2;
      if (q >= n2) q -= n2;
    }
    zr[e] = re;
    zi[e] = im;
  }
  __syncthreads();
  // pass 2: X_k = sum_i1 w^(i1 k) Z[i1][k mod n2] for k < n_m, then the modes
  float* Y = a.out + (size_t)b * a.mm * a.nlat;
  for (int e = threadIdx.x; e < a.n_m * tl; e += kThreads) {
    const int k = e / tl, jj = e - k * tl, j = j0 + jj;
    int k2 = k % n2;
    const bool cj = k2 >= nh;          // Z[i1][k2] = conj(Z[i1][n2 - k2])
    if (cj) k2 = n2 - k2;
    const float sgn = cj ? -1.0f : 1.0f;
    float re = 0.0f, im = 0.0f;
    int q = 0;   // (i1 * k) mod N
    const int kq = k % N;
    for (int i1 = 0; i1 < n1; ++i1) {
      const int z = (i1 * nh + k2) * tl + jj;
      const float a_r = zr[z], a_i = sgn * zi[z];
      const float t_r = twr[q], t_i = twi[q];
      re += a_r * t_r - a_i * t_i;
      im += a_r * t_i + a_i * t_r;
      q += kq;
      if (q >= N) q -= N;
    }
    if (j < a.nlat) {
      if (k == 0) {
        Y[j] = re * kRsqrt2Pi;
      } else {
        Y[(size_t)(2 * k - 1) * a.nlat + j] = re * kRsqrtPi;
        Y[(size_t)(2 * k) * a.nlat + j] = -im * kRsqrtPi;
      }
    }
  }
}

// out_b[lon][lat] = epilogue(sum_mode F[lon][mode] * Y_b[mode][lat]), Y_b from the
// Legendre synthesis's even / odd halves: E + O (southern latitudes), E - O
// (northern).
__global__ __launch_bounds__(kThreads) void fourier_synthesis_generic(FourierArgs a) {
  extern __shared__ float lds[];
  const int N = a.nlon, n1 = a.n1, n2 = a.n2, tl = a.tl, h = a.nlat >> 1;
  const int b = blockIdx.y, j0 = blockIdx.x * tl;
  float* twr = lds;
  float* twi = twr + N;
  float* ys = twi + N;             // [mm][tl]
  float* wr = ys + a.mm * tl;      // [n1][n2][tl]
  float* wi = wr + N * tl;
  const Seg sg = a.seg.s[find_seg(a.seg, b)];
  load_roots(a, twr, twi);
  const float* E = a.in + (size_t)b * a.mm * h;
  const size_t odd = (size_t)a.nfield * a.mm * h;
  for (int e = threadIdx.x; e < a.mm * tl; e += kThreads) {
    const int mode = e / tl, jj = e - mode * tl, j = j0 + jj;
    float v = 0.0f;
    if (j < a.nlat) {
      const float* p = E + (size_t)mode * h + (j < h ? j : a.nlat - 1 - j);
      v = j < h ? p[0] + p[odd] : p[0] - p[odd];
    }
    ys[e] = v;
  }
  __syncthreads();
  // pass 1: W[i1][k2] = sum_{k = k2 + n2 k1 < n_m} c_k conj(w^(i1 k)),
  // c_0 = Y_0 / sqrt(2 pi), c_k = (Y_{2k-1} - i Y_{2k}) / sqrt(pi)
  for (int e = threadIdx.x; e < N * tl; e += kThreads) {
    const int r = e / tl, jj = e - r * tl, k2 = r % n2, i1 = r / n2;
    float re = 0.0f, im = 0.0f;
    for (int k = k2; k < a.n_m; k += n2) {
      float c_r, c_i;
      if (k == 0) {
        c_r = ys[jj] * kRsqrt2Pi;
        c_i = 0.0f;
      } else {
        c_r = ys[(2 * k - 1) * tl + jj] * kRsqrtPi;
        c_i = -ys[(2 * k) * tl + jj] * kRsqrtPi;
      }
      const int q = (int)(((long)i1 * k) % N);
      const float t_r = twr[q], t_i = -twi[q];
      re += c_r * t_r - c_i * t_i;
      im += c_r * t_i + c_i * t_r;
    }
    wr[e] = re;
    wi[e] = im;
  }
  __syncthreads();
  // pass 2: x[i1 + n1 i2] = Re sum_k2 W[i1][k2] conj(w^(n1 i2 k2)), then the epilogue
  float* dst = sg.dst + (size_t)(b - sg.first) * N * a.nlat;
  const float lvl = sg.level ? sg.level[(b - sg.first) % sg.period] : 0.0f;
  for (int e = threadIdx.x; e < N * tl; e += kThreads) {
    // latitude fastest, then i2: the lanes of a wave share i1, so their W reads
    // are broadcasts or distinct banks
    const int r = e / tl, jj = e - r * tl, j = j0 + jj;
    const int i2 = r % n2, i1 = r / n2, lon = i1 + n1 * i2;
    float v = 0.0f;
    int q = 0;   // (i2 * k2) mod n2
    for (int k2 = 0; k2 < n2; ++k2) {
      const int z = (i1 * n2 + k2) * tl + jj;
      v += wr[z] * twr[q * n1] + wi[z] * twi[q * n1];
      q += i2;
      if (q >= n2) q -= n2;
    }
    if (j < a.nlat) {
      if (sg.op & kOpSecLat) v *= a.sec[j];
      v = v * sg.mul;
      if (sg.op & kOpAddLevel) v += lvl;
      if (sg.op & kOpExp) v = expf(v);
      dst[(size_t)lon * a.nlat + j] = v;
    }
  }
}

// The same two passes with both factors known at compile time (the nlon of the
// Grid.T<N> table): each pass is a short dense DFT in registers whose roots are
// literals (Roots<n>, evaluated by the compiler), with the line tile read from
// and written to memory directly; only the pass-1 results and the N roots of
// unity for the twiddles w^(i1 k2) go through LDS.
constexpr double kPiD = 3.14159265358979323846;
constexpr double ct_sin(double x) {   // |x| <= pi, Taylor series to 1e-16
  double term = x, sum = x;
  for (int n = 1; n < 24; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double ct_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 24; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
// w_n^q = exp(-2 pi i q / n) = c[q] + i s[q]
template <int n>
struct Roots {
  float c[n], s[n];
  constexpr Roots() : c(), s() {
    for (int q = 0; q < n; ++q) {
      double t = 2.0 * kPiD * q / n;
      if (t > kPiD) t -= 2.0 * kPiD;
      c[q] = (float)ct_cos(t);
      s[q] = (float)-ct_sin(t);
    }
  }
};

template <int N1, int N2>
__global__ __launch_bounds__(kThreads) void fourier_analysis_kernel(FourierArgs a) {
  constexpr int N = N1 * N2, NH = N2 / 2 + 1;
  constexpr Roots<N1> R1{};
  constexpr Roots<N2> R2{};
  extern __shared__ float lds[];
  const int tl = a.tl, b = blockIdx.y, j0 = blockIdx.x * tl;
  float* twr = lds;
  float* twi = twr + N;
  float* zr = twi + N;            // [NH][N1][tl]
  float* zi = zr + NH * N1 * tl;
  const Seg sg = a.seg.s[find_seg(a.seg, b)];
  const float* src = sg.src + (size_t)(b - sg.first) * N * a.nlat;
  const float lvl = sg.level ? sg.level[(b - sg.first) % sg.period] : 0.0f;
  load_roots(a, twr, twi);
  // pass 1: thread (i1, lat): Z[i1][k2] = sum_i2 x[i1 + N1 i2] w_N2^(i2 k2), k2 <= N2 / 2
  for (int e = threadIdx.x; e < N1 * tl; e += kThreads) {
    const int i1 = e / tl, jj = e - i1 * tl, j = j0 + jj;
    const bool in = j < a.nlat;
    const int jc = in ? j : 0;
    const float cs = in ? sg.mul * a.w[jc] * ((sg.op & kOpSecLat) ? a.sec[jc] : 1.0f) : 0.0f;
    float x[N2];
#pragma unroll
    for (int i2 = 0; i2 < N2; ++i2) x[i2] = src[(size_t)(i1 + N1 * i2) * a.nlat + jc];
#pragma unroll
    for (int i2 = 0; i2 < N2; ++i2) {
      float v = x[i2];
      if (sg.op & kOpLog) v = logf(v);
      if (sg.op & kOpSubLevel) v -= lvl;
      x[i2] = in ? v * cs : 0.0f;
    }
#pragma unroll
    for (int k2 = 0; k2 < NH; ++k2) {
      float re = 0.0f, im = 0.0f;
#pragma unroll
      for (int i2 = 0; i2 < N2; ++i2) {
        re += x[i2] * R2.c[(i2 * k2) % N2];
        im += x[i2] * R2.s[(i2 * k2) % N2];
      }
      zr[(k2 * N1 + i1) * tl + jj] = re;
      zi[(k2 * N1 + i1) * tl + jj] = im;
    }
  }
  __syncthreads();
  // pass 2: thread (k2, lat): X_{k2 + N2 k1} = sum_i1 (Z[i1][k2] w^(i1 k2)) w_N1^(i1 k1)
  float* Y = a.out + (size_t)b * a.mm * a.nlat;
  for (int e = threadIdx.x; e < N2 * tl; e += kThreads) {
    const int k2 = e / tl, jj = e - k2 * tl, j = j0 + jj;
    const bool cj = k2 >= NH;        // Z[i1][k2] = conj(Z[i1][N2 - k2])
    const int kz = cj ? N2 - k2 : k2;
    float pr[N1], pi[N1];
#pragma unroll
    for (int i1 = 0; i1 < N1; ++i1) {
      const int z = (kz * N1 + i1) * tl + jj;
      const float a_r = zr[z], a_i = cj ? -zi[z] : zi[z];
      const int q = (i1 * k2) % N;
      const float t_r = twr[q], t_i = twi[q];
      pr[i1] = a_r * t_r - a_i * t_i;
      pi[i1] = a_r * t_i + a_i * t_r;
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      if (N2 * k1 >= a.n_m) break;   // uniform; k < n_m is masked per lane below
      const int k = k2 + N2 * k1;
      float re = 0.0f, im = 0.0f;
#pragma unroll
      for (int i1 = 0; i1 < N1; ++i1) {
        const float c = R1.c[(i1 * k1) % N1], s = R1.s[(i1 * k1) % N1];
        re += pr[i1] * c - pi[i1] * s;
        im += pr[i1] * s + pi[i1] * c;
      }
      if (j < a.nlat && k < a.n_m) {
        if (k == 0) {
          Y[j] = re * kRsqrt2Pi;
        } else {
          Y[(size_t)(2 * k - 1) * a.nlat + j] = re * kRsqrtPi;
          Y[(size_t)(2 * k) * a.nlat + j] = -im * kRsqrtPi;
        }
      }
    }
  }
}

template <int N1, int N2>
__global__ __launch_bounds__(kThreads) void fourier_synthesis_kernel(FourierArgs a) {
  constexpr int N = N1 * N2;
  constexpr Roots<N1> R1{};
  constexpr Roots<N2> R2{};
  extern __shared__ float lds[];
  const int tl = a.tl, h = a.nlat >> 1, b = blockIdx.y, j0 = blockIdx.x * tl;
  float* twr = lds;
  float* twi = twr + N;
  float* wr = twi + N;             // [N1][N2][tl]
  float* wi = wr + N * tl;
  const Seg sg = a.seg.s[find_seg(a.seg, b)];
  load_roots(a, twr, twi);
  __syncthreads();
  const float* E = a.in + (size_t)b * a.mm * h;
  const size_t odd = (size_t)a.nfield * a.mm * h;
  // pass 1: thread (k2, lat): W[i1][k2] = conj(w^(i1 k2)) sum_k1 c_{k2 + N2 k1} conj(w_N1^(i1 k1)),
  // c_0 = Y_0 / sqrt(2 pi), c_k = (Y_{2k-1} - i Y_{2k}) / sqrt(pi); Y from E + O (south), E - O (north)
  for (int e = threadIdx.x; e < N2 * tl; e += kThreads) {
    const int k2 = e / tl, jj = e - k2 * tl, j = j0 + jj;
    const bool in = j < a.nlat;
    const int jh = !in ? 0 : (j < h ? j : a.nlat - 1 - j);
    const float so = j < h ? 1.0f : -1.0f;
    // the wavenumbers k = k2 + N2 k1 < n_m: k1 < ceil(n_m / N2) for every lane
    // (a uniform bound), the last one masked per lane; loads at clamped modes
    // (in bounds) issue back to back, all before any use
    float ec[N1], oc[N1], es[N1], os[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      if (N2 * k1 >= a.n_m) break;
      const int k = k2 + N2 * k1;
      const int kc = k < a.n_m ? k : 0;
      const float* pc = E + (size_t)(kc == 0 ? 0 : 2 * kc - 1) * h + jh;
      const float* ps = E + (size_t)(kc == 0 ? 0 : 2 * kc) * h + jh;
      ec[k1] = pc[0]; oc[k1] = pc[odd]; es[k1] = ps[0]; os[k1] = ps[odd];
    }
    float cr[N1], ci[N1];
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      cr[k1] = 0.0f;
      ci[k1] = 0.0f;
      if (N2 * k1 >= a.n_m) continue;
      const int k = k2 + N2 * k1;
      const bool use = in && k < a.n_m;
      const float y_c = ec[k1] + so * oc[k1], y_s = es[k1] + so * os[k1];
      cr[k1] = use ? y_c * (k == 0 ? kRsqrt2Pi : kRsqrtPi) : 0.0f;
      ci[k1] = use && k > 0 ? -y_s * kRsqrtPi : 0.0f;
    }
    // k1 outermost with the uniform bound: only the ceil(n_m / N2) wavenumbers
    // that exist are summed (6 of 16 at T85), in ascending k1 as before
    float re[N1], im[N1];
#pragma unroll
    for (int i1 = 0; i1 < N1; ++i1) {
      re[i1] = 0.0f;
      im[i1] = 0.0f;
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      if (N2 * k1 >= a.n_m) break;
#pragma unroll
      for (int i1 = 0; i1 < N1; ++i1) {
        const float c = R1.c[(i1 * k1) % N1], s = -R1.s[(i1 * k1) % N1];
        re[i1] += cr[k1] * c - ci[k1] * s;
        im[i1] += cr[k1] * s + ci[k1] * c;
      }
    }
#pragma unroll
    for (int i1 = 0; i1 < N1; ++i1) {
      const int q = (i1 * k2) % N;
      const float t_r = twr[q], t_i = -twi[q];
      wr[(i1 * N2 + k2) * tl + jj] = re[i1] * t_r - im[i1] * t_i;
      wi[(i1 * N2 + k2) * tl + jj] = re[i1] * t_i + im[i1] * t_r;
    }
  }
  __syncthreads();
  // pass 2: thread (i1, lat): x[i1 + N1 i2] = Re sum_k2 W[i1][k2] conj(w_N2^(i2 k2)), then the epilogue
  float* dst = sg.dst + (size_t)(b - sg.first) * N * a.nlat;
  const float lvl = sg.level ? sg.level[(b - sg.first) % sg.period] : 0.0f;
  for (int e = threadIdx.x; e < N1 * tl; e += kThreads) {
    const int i1 = e / tl, jj = e - i1 * tl, j = j0 + jj;
    const int jc = j < a.nlat ? j : 0;
    float pr[N2], pi[N2];
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      pr[k2] = wr[(i1 * N2 + k2) * tl + jj];
      pi[k2] = wi[(i1 * N2 + k2) * tl + jj];
    }
#pragma unroll
    for (int i2 = 0; i2 < N2; ++i2) {
      float v = 0.0f;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) v += pr[k2] * R2.c[(i2 * k2) % N2] + pi[k2] * R2.s[(i2 * k2) % N2];
      if (sg.op & kOpSecLat) v *= a.sec[jc];
      v = v * sg.mul;
      if (sg.op & kOpAddLevel) v += lvl;
      if (sg.op & kOpExp) v = expf(v);
      if (j < a.nlat) dst[(size_t)(i1 + N1 * i2) * a.nlat + j] = v;
    }
  }
}

// eps(m, l) = sqrt((l^2 - m^2) / (4 l^2 - 1)) for m <= l, l >= 1; else 0:
// x P_l^m = eps(m, l+1) P_{l+1}^m + eps(m, l) P_{l-1}^m (unit-norm functions).
__device__ __forceinline__ float eps_ml(int m, int l) {
  if (l <= 0 || m > l) return 0.0f;
  const float fl = (float)l, fm = (float)m;
  return sqrtf((fl * fl - fm * fm) / (4.0f * fl * fl - 1.0f));
}

struct ModalArgs {
  int32_t mm, nl, kx;
  float inv_radius, radius2;
  const float* vor;     // [kx][mm][nl]
  const float* div;
  const float* copy_src[4];   // modal fields copied behind the winds (null: zeros)
  int32_t copy_count[4];
  float* out;           // contiguous [nfield][mm][nl]
  float* vor_out;       // post: destinations of vorticity / divergence
  float* div_out;
  float* copy_dst[4];
  // geopotential rows behind the copies (winds kernel; kx rows when geo_w is set):
  // phi_j = g orography + sum_k W[j][k] (T_var_k + reference_k as a constant)
  const float* tvar;
  const float* orography;
  const float* reference;
  float gravity;
  int32_t geo;
  float geo_w[SPEEDY_SHT_KX_MAX][SPEEDY_SHT_KX_MAX];
};

// Modal coefficient (mode 0, l 0) of the constant 1: 1 / (F_0 P_0^0) = 2 sqrt(pi).
constexpr float kConstCoef = 3.5449077018110318f;

__device__ __forceinline__ float d_dlon(const float* x, int mode, int m, int l, int nl) {
  if (mode == 0) return 0.0f;
  return (mode & 1) ? (float)m * x[(mode + 1) * nl + l] : -(float)m * x[(mode - 1) * nl + l];
}

// Modal (psi, chi) -> (u cos(lat), v cos(lat)) for every level, then copies of
// the remaining modal inputs, into one contiguous stack (vor_div_to_uv_nodal's
// modal half and compute_diagnostic_state's inputs).  blockIdx.y is the row:
// rows below kx are the winds of level blockIdx.y (writing the u and v rows),
// the others write output row blockIdx.y + kx; blockIdx.x covers the row's
// mm x nl coefficients.
__global__ __launch_bounds__(kThreads) void modal_winds_kernel(ModalArgs a) {
  const int per = a.mm * a.nl;
  const int rem = blockIdx.x * kThreads + threadIdx.x;
  if (rem >= per) return;
  int r = blockIdx.y;
  if (r < a.kx) {
    const int k = r, mode = rem / a.nl, l = rem - mode * a.nl;
    const int m = (mode + 1) >> 1;
    const float* z = a.vor + (size_t)k * per;
    const float* d = a.div + (size_t)k * per;
    float u = 0.0f, v = 0.0f;
    if (l < a.nl - 1) {     // the highest total wavenumber is clipped
      auto inv_lap = [&](int ll) { return ll == 0 ? 0.0f : -a.radius2 / ((float)ll * (float)(ll + 1)); };
      // d/dlon of psi, chi at (mode, l)
      const float psi_lon = d_dlon(z, mode, m, l, a.nl) * inv_lap(l);
      const float chi_lon = d_dlon(d, mode, m, l, a.nl) * inv_lap(l);
      // cos(lat) d/dlat at l: (l+2) eps(l+1) x_{l+1} - (l-1) eps(l) x_{l-1}
      float psi_lat = 0.0f, chi_lat = 0.0f;
      if (l + 1 < a.nl) {
        const float c = (float)(l + 2) * eps_ml(m, l + 1) * inv_lap(l + 1);
        psi_lat += c * z[mode * a.nl + l + 1];
        chi_lat += c * d[mode * a.nl + l + 1];
      }
      if (l >= 1) {
        const float c = -(float)(l - 1) * eps_ml(m, l) * inv_lap(l - 1);
        psi_lat += c * z[mode * a.nl + l - 1];
        chi_lat += c * d[mode * a.nl + l - 1];
      }
      u = (chi_lon - psi_lat) * a.inv_radius;
      v = (chi_lat + psi_lon) * a.inv_radius;
    }
    a.out[(size_t)k * per + rem] = u;
    a.out[(size_t)(a.kx + k) * per + rem] = v;
    return;
  }
  float* dst = a.out + (size_t)(blockIdx.y + a.kx) * per + rem;
  r -= a.kx;
  for (int s = 0; s < 4; ++s) {
    if (r < a.copy_count[s]) {
      *dst = a.copy_src[s] ? a.copy_src[s][(size_t)r * per + rem] : 0.0f;
      return;
    }
    r -= a.copy_count[s];
  }
  if (!a.geo || r >= a.kx) return;
  // get_geopotential in modal space: the reference temperature enters the
  // (mode 0, l 0) coefficient only.  Every level's load is issued before the
  // sum (levels past kx re-read the last one and are not summed).
  const int j = r;
  float t[SPEEDY_SHT_KX_MAX];
#pragma unroll
  for (int k = 0; k < SPEEDY_SHT_KX_MAX; ++k) t[k] = a.tvar[(size_t)min(k, a.kx - 1) * per + rem];
  float acc = a.gravity * a.orography[rem];
#pragma unroll
  for (int k = 0; k < SPEEDY_SHT_KX_MAX; ++k) {
    if (k < a.kx) {
      float tk = t[k];
      if (rem == 0) tk += a.reference[k] * kConstCoef;
      acc += a.geo_w[j][k] * tk;
    }
  }
  *dst = acc;
}

// Modal (U, V) = (u, v)/cos(lat) -> vorticity and divergence (curl_cos_lat,
// div_cos_lat, clipped), then the remaining stacked rows to their destinations
// (null source rows: zeros).  blockIdx.y is the row as in modal_winds_kernel:
// rows below kx are the winds of that level, the others copy stacked row
// blockIdx.y + kx.
__global__ __launch_bounds__(kThreads) void modal_vordiv_kernel(ModalArgs a) {
  const int per = a.mm * a.nl;
  const int rem = blockIdx.x * kThreads + threadIdx.x;
  if (rem >= per) return;
  int r = blockIdx.y;
  if (r < a.kx) {
    const int k = r, mode = rem / a.nl, l = rem - mode * a.nl;
    const int m = (mode + 1) >> 1;
    const float* U = a.out + (size_t)k * per;
    const float* V = a.out + (size_t)(a.kx + k) * per;
    float z = 0.0f, d = 0.0f;
    if (l < a.nl - 1) {
      // sec(lat) d/dlat cos^2(lat) at l: l eps(l+1) x_{l+1} - (l+1) eps(l) x_{l-1}
      float u_lat = 0.0f, v_lat = 0.0f;
      if (l + 1 < a.nl) {
        const float c = (float)l * eps_ml(m, l + 1);
        u_lat += c * U[mode * a.nl + l + 1];
        v_lat += c * V[mode * a.nl + l + 1];
      }
      if (l >= 1) {
        const float c = -(float)(l + 1) * eps_ml(m, l);
        u_lat += c * U[mode * a.nl + l - 1];
        v_lat += c * V[mode * a.nl + l - 1];
      }
      z = (d_dlon(V, mode, m, l, a.nl) - u_lat) * a.inv_radius;
      d = (d_dlon(U, mode, m, l, a.nl) + v_lat) * a.inv_radius;
    }
    a.vor_out[(size_t)k * per + rem] = z;
    a.div_out[(size_t)k * per + rem] = d;
    return;
  }
  const float* src = a.out + (size_t)(blockIdx.y + a.kx) * per + rem;
  r -= a.kx;
  for (int s = 0; s < 4; ++s) {
    if (r < a.copy_count[s]) {
      a.copy_dst[s][(size_t)r * per + rem] = a.copy_src[s] ? *src : 0.0f;
      return;
    }
    r -= a.copy_count[s];
  }
}

// ------------------------------------------------------------------ host side
double legendre_p(int n, double x, double* dp) {
  double p0 = 1.0, p1 = x;
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  if (n == 0) {
    *dp = 0.0;
    return 1.0;
  }
  *dp = n * (x * p1 - p0) / (x * x - 1.0);
  return p1;
}

void gauss_nodes(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    // i-th root from the north (cos guess), stored south to north
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      const double p = legendre_p(n, z, &dp);
      const double dz = p / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    legendre_p(n, z, &dp);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

int check_grid(int32_t n_m, int32_t n_l, int32_t nlon, int32_t nlat) {
  if (n_m < 1 || n_l < 2 || nlon < 1 || nlat < 2 || (nlat & 1) || n_m > n_l || 2 * n_m - 1 > nlon || n_l > nlat ||
      (int64_t)n_m * nlat * n_l >= ((int64_t)1 << 31) || (int64_t)nlon * nlat * 64 >= ((int64_t)1 << 31))
    return fail(SPEEDY_ERR_ARG, "speedy_sht: bad grid (need 1 <= n_m <= n_l <= nlat, n_l >= 2, 2*n_m-1 <= nlon, "
                                "nlat even)");
  return SPEEDY_OK;
}

void fill_basis(int32_t n_m, int32_t n_l, int32_t nlon, int32_t nlat, const double* x, float* fourier,
                float* legendre) {
  const int mm = 2 * n_m - 1;
  if (fourier) {
    for (int i = 0; i < nlon; ++i) {
      const double lon = 2.0 * M_PI * i / nlon;
      fourier[(size_t)i * mm] = (float)(1.0 / std::sqrt(2.0 * M_PI));
      for (int m = 1; m < n_m; ++m) {
        fourier[(size_t)i * mm + 2 * m - 1] = (float)(std::cos(m * lon) / std::sqrt(M_PI));
        fourier[(size_t)i * mm + 2 * m] = (float)(std::sin(m * lon) / std::sqrt(M_PI));
      }
    }
  }
  if (legendre) {
    std::vector<double> col(n_l);
    for (int j = 0; j < nlat; ++j) {
      const double xj = x[j], yj = std::sqrt(1.0 - xj * xj);
      double diag = 1.0 / std::sqrt(2.0);
      for (int m = 0; m < n_m; ++m) {
        if (m > 0) diag = -std::sqrt(1.0 + 1.0 / (2.0 * m)) * yj * diag;
        std::fill(col.begin(), col.end(), 0.0);
        col[m] = diag;
        if (m + 1 < n_l) col[m + 1] = std::sqrt(2.0 * m + 3.0) * xj * diag;
        for (int l = m + 2; l < n_l; ++l) {
          const double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
          const double b = std::sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
          col[l] = a * (xj * col[l - 1] - b * col[l - 2]);
        }
        float* dst = legendre + ((size_t)m * nlat + j) * n_l;
        for (int l = 0; l < n_l; ++l) dst[l] = (float)col[l];
      }
    }
  }
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

size_t ws_bytes(const SpeedyShtPlan* p, int32_t nfield) {
  return align256((size_t)nfield * p->mm * p->nlat * 4) + align256((size_t)nfield * p->mm * p->n_l * 4);
}

struct Ws {
  float* y;   // [nfield][mm][nlat]: between the Fourier and Legendre stages
  float* z;   // [nfield][mm][nl]: stacked modal fields
};

Ws carve(const SpeedyShtPlan* p, int32_t nfield, void* ws) {
  char* b = static_cast<char*>(ws);
  Ws r;
  r.y = reinterpret_cast<float*>(b);
  r.z = reinterpret_cast<float*>(b + align256((size_t)nfield * p->mm * p->nlat * 4));
  return r;
}

StageArgs stage_args(const SpeedyShtPlan* p, int32_t nfield) {
  StageArgs a{};
  a.nlon = p->nlon; a.nlat = p->nlat; a.mm = p->mm; a.nl = p->n_l; a.nfield = nfield;
  a.P = p->P;
  return a;
}

FourierArgs fourier_args(const SpeedyShtPlan* p, int32_t nfield, int32_t tl, const SegTable& seg) {
  FourierArgs f{};
  f.nlon = p->nlon; f.nlat = p->nlat; f.mm = p->mm; f.n_m = p->n_m; f.n1 = p->n1; f.n2 = p->n2; f.tl = tl;
  f.nfield = nfield;
  f.twr = p->tw; f.twi = p->tw + p->nlon; f.w = p->w; f.sec = p->sec;
  f.seg = seg;
  return f;
}

// The Fourier kernels of one factor pair nlon = n1 * n2: compiled for the nlon
// of the Grid.T<N> table up to T213 (nlon 640), the generic LDS kernels
// otherwise (at n1 = 32 the register passes would spill to scratch).
typedef void (*FourierKernel)(FourierArgs);
struct FourierPair {
  int n1, n2;
  FourierKernel analysis, synthesis;
};
#define SHT_PAIR(a, b) {a, b, fourier_analysis_kernel<a, b>, fourier_synthesis_kernel<a, b>}
const FourierPair kPairs[] = {SHT_PAIR(8, 8),   SHT_PAIR(8, 12),  SHT_PAIR(8, 16),  SHT_PAIR(16, 16),
                              SHT_PAIR(16, 20), SHT_PAIR(18, 20), SHT_PAIR(16, 32), SHT_PAIR(20, 32)};
#undef SHT_PAIR

const FourierPair* find_pair(int n1, int n2) {
  for (const FourierPair& f : kPairs)
    if (f.n1 == n1 && f.n2 == n2) return &f;
  return nullptr;
}

// LDS bytes of the Fourier kernels (roots, line tile, pass-1 result)
constexpr size_t kFourierLds = 64 * 1024;
size_t analysis_lds(bool fixed, int N, int n1, int n2, int tl) {
  const size_t z = 2 * (size_t)n1 * (n2 / 2 + 1) * tl;
  return 4 * (2 * (size_t)N + z + (fixed ? 0 : (size_t)N * tl));
}
size_t synthesis_lds(bool fixed, int N, int mm, int tl) {
  return 4 * (2 * (size_t)N + 2 * (size_t)N * tl + (fixed ? 0 : (size_t)mm * tl));
}

int tiles(int M, int N) { return ((M + kBM - 1) / kBM) * ((N + kBN - 1) / kBN); }

hipError_t run_analysis(const SpeedyShtPlan* p, int32_t nfield, const SegTable& seg, float* y, float* modal_out,
                        hipStream_t s) {
  FourierArgs f = fourier_args(p, nfield, p->tl_a, seg);
  f.out = y;
  const FourierPair* fp = find_pair(p->n1, p->n2);
  hipLaunchKernelGGL(fp ? fp->analysis : fourier_analysis_generic, dim3((p->nlat + p->tl_a - 1) / p->tl_a, nfield),
                     dim3(kThreads), analysis_lds(fp != nullptr, p->nlon, p->n1, p->n2, p->tl_a), s, f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  StageArgs a = stage_args(p, nfield);
  a.in = y;
  a.out = modal_out;
  hipLaunchKernelGGL(sht_gemm_kernel<kLegendreAnalysis>, dim3(tiles(nfield, (p->n_l + 1) / 2), 2 * p->mm),
                     dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t run_synthesis(const SpeedyShtPlan* p, int32_t nfield, const SegTable& seg, const float* modal_in, float* y,
                         hipStream_t s) {
  StageArgs a = stage_args(p, nfield);
  a.in = modal_in;
  a.out = y;
  hipLaunchKernelGGL(sht_gemm_kernel<kLegendreSynthesis>, dim3(tiles(nfield, p->nlat / 2), 2 * p->mm),
                     dim3(kThreads), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  FourierArgs f = fourier_args(p, nfield, p->tl_s, seg);
  f.in = y;
  const FourierPair* fp = find_pair(p->n1, p->n2);
  hipLaunchKernelGGL(fp ? fp->synthesis : fourier_synthesis_generic, dim3((p->nlat + p->tl_s - 1) / p->tl_s, nfield),
                     dim3(kThreads), synthesis_lds(fp != nullptr, p->nlon, p->mm, p->tl_s), s, f);
  return hipGetLastError();
}

void add_seg(SegTable& t, const float* src, float* dst, int32_t count, int32_t op, float mul,
             const float* level = nullptr, int32_t period = 1) {
  Seg& s = t.s[t.n];
  s.src = src; s.dst = dst; s.level = level; s.period = period; s.op = op; s.mul = mul; s.count = count;
  s.first = t.n == 0 ? 0 : t.s[t.n - 1].first + t.s[t.n - 1].count;
  ++t.n;
}

int check_call(const SpeedyShtPlan* p, int32_t nfield, void* ws, size_t wsb) {
  if (!p) return fail(SPEEDY_ERR_ARG, "speedy_sht: null plan");
  if (nfield < 1 || nfield > 65535) return fail(SPEEDY_ERR_ARG, "speedy_sht: nfield must be in 1..65535");
  if (!ws || wsb < ws_bytes(p, nfield))
    return fail(SPEEDY_ERR_WORKSPACE, "speedy_sht: workspace missing or smaller than speedy_sht_workspace_bytes");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != p->device)
    return fail(SPEEDY_ERR_ARG, "speedy_sht: the plan belongs to another device than the current one");
  return SPEEDY_OK;
}

int check_spec(const SpeedyDynamicsSpec* sp) {
  if (!sp || sp->kx < 1 || sp->kx > SPEEDY_SHT_KX_MAX || !sp->reference_temperature)
    return fail(SPEEDY_ERR_ARG, "speedy_sht: bad dynamics spec (kx 1..16, reference temperature)");
  return SPEEDY_OK;
}

int hip_result(hipError_t e) {
  if (e != hipSuccess) return fail(SPEEDY_ERR_HIP, std::string("speedy_sht: launch failed: ") + hipGetErrorString(e));
  return SPEEDY_OK;
}

int blocks(size_t n) { return (int)((n + kThreads - 1) / kThreads); }

}  // namespace sht

using namespace sht;

extern "C" {

int speedy_sht_host_basis(int32_t n_m, int32_t n_l, int32_t nlon, int32_t nlat, double* sin_lat, double* weights,
                          float* fourier, float* legendre) {
  if (int rc = check_grid(n_m, n_l, nlon, nlat)) return rc;
  if (!sin_lat || !weights) return fail(SPEEDY_ERR_ARG, "speedy_sht_host_basis: null nodes or weights");
  gauss_nodes(nlat, sin_lat, weights);
  fill_basis(n_m, n_l, nlon, nlat, sin_lat, fourier, legendre);
  return SPEEDY_OK;
}

int speedy_sht_create(int32_t n_m, int32_t n_l, int32_t nlon, int32_t nlat, double radius, SpeedyShtPlan** out) {
  if (!out) return fail(SPEEDY_ERR_ARG, "speedy_sht_create: null output");
  *out = nullptr;
  if (int rc = check_grid(n_m, n_l, nlon, nlat)) return rc;
  if (!(radius > 0.0)) return fail(SPEEDY_ERR_ARG, "speedy_sht_create: radius must be positive");
  const int mm = 2 * n_m - 1;
  // nlon = n1 * n2 with n1 the largest divisor <= sqrt(nlon); the widest latitude
  // tile (16, 8, .., 1) whose LDS fits
  int n1 = 1;
  for (int d = 1; (int64_t)d * d <= nlon; ++d)
    if (nlon % d == 0) n1 = d;
  const int n2 = nlon / n1;
  // compiled pairs: a thread per (i1 or k2, latitude) in each pass, so at most
  // kThreads / max(n1, n2) latitudes; generic kernels: 16
  const bool fixed = find_pair(n1, n2) != nullptr;
  int tl_a = fixed ? std::min(16, kThreads / std::max(n1, n2)) : 16, tl_s = tl_a;
  while (tl_a > 1 && analysis_lds(fixed, nlon, n1, n2, tl_a) > kFourierLds) --tl_a;
  while (tl_s > 1 && synthesis_lds(fixed, nlon, mm, tl_s) > kFourierLds) --tl_s;
  if (analysis_lds(fixed, nlon, n1, n2, tl_a) > kFourierLds || synthesis_lds(fixed, nlon, mm, tl_s) > kFourierLds)
    return fail(SPEEDY_ERR_ARG, "speedy_sht_create: nlon too large for the Fourier stages' 64 KiB of LDS");
  std::vector<double> x(nlat), w(nlat);
  gauss_nodes(nlat, x.data(), w.data());
  std::vector<float> P((size_t)n_m * nlat * n_l), wf(nlat), sec(nlat), tw(2 * (size_t)nlon);
  fill_basis(n_m, n_l, nlon, nlat, x.data(), nullptr, P.data());
  for (int j = 0; j < nlat; ++j) {
    wf[j] = (float)(w[j] * 2.0 * M_PI / nlon);
    sec[j] = (float)(1.0 / std::sqrt(1.0 - x[j] * x[j]));
  }
  for (int j = 0; j < nlon; ++j) {
    tw[j] = (float)std::cos(2.0 * M_PI * j / nlon);
    tw[nlon + j] = (float)-std::sin(2.0 * M_PI * j / nlon);
  }
  SpeedyShtPlan* p = new SpeedyShtPlan{n_m, n_l, nlon, nlat, mm, radius, 0, n1, n2, tl_a, tl_s,
                                       nullptr, nullptr, nullptr, nullptr};
  hipError_t e = hipGetDevice(&p->device);
  auto up = [&](float** dst, const std::vector<float>& v) {
    if (e != hipSuccess) return;
    e = hipMalloc(reinterpret_cast<void**>(dst), v.size() * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(*dst, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice);
  };
  up(&p->P, P);
  up(&p->w, wf);
  up(&p->sec, sec);
  up(&p->tw, tw);
  if (e != hipSuccess) {
    speedy_sht_destroy(p);
    return fail(SPEEDY_ERR_HIP, std::string("speedy_sht_create: ") + hipGetErrorString(e));
  }
  *out = p;
  return SPEEDY_OK;
}

int speedy_sht_destroy(SpeedyShtPlan* p) {
  if (!p) return SPEEDY_OK;
  for (float* q : {p->P, p->w, p->sec, p->tw})
    if (q) (void)hipFree(q);
  delete p;
  return SPEEDY_OK;
}

size_t speedy_sht_workspace_bytes(const SpeedyShtPlan* p, int32_t nfield) {
  return p && nfield > 0 ? ws_bytes(p, nfield) : 0;
}

int speedy_sht_to_modal(const SpeedyShtPlan* p, const float* nodal, float* modal, int32_t nfield, void* ws,
                        size_t wsb, void* stream) {
  if (int rc = check_call(p, nfield, ws, wsb)) return rc;
  if (!nodal || !modal) return fail(SPEEDY_ERR_ARG, "speedy_sht_to_modal: null field");
  SegTable seg{};
  add_seg(seg, nodal, nullptr, nfield, 0, 1.0f);
  const Ws w = carve(p, nfield, ws);
  return hip_result(run_analysis(p, nfield, seg, w.y, modal, static_cast<hipStream_t>(stream)));
}

int speedy_sht_to_nodal(const SpeedyShtPlan* p, const float* modal, float* nodal, int32_t nfield, void* ws,
                        size_t wsb, void* stream) {
  if (int rc = check_call(p, nfield, ws, wsb)) return rc;
  if (!nodal || !modal) return fail(SPEEDY_ERR_ARG, "speedy_sht_to_nodal: null field");
  SegTable seg{};
  add_seg(seg, nullptr, nodal, nfield, 0, 1.0f);
  const Ws w = carve(p, nfield, ws);
  return hip_result(run_synthesis(p, nfield, seg, modal, w.y, static_cast<hipStream_t>(stream)));
}

int speedy_dynamics_to_physics(const SpeedyShtPlan* p, const SpeedyDynamicsSpec* spec, const SpeedySpectralState* in,
                               const SpeedyNodalState* out, void* ws, size_t wsb, void* stream) {
  if (int rc = check_spec(spec)) return rc;
  if (!in || !out || !in->vorticity || !in->divergence || !in->temperature_variation || !in->specific_humidity ||
      !in->log_surface_pressure || !spec->orography || !spec->sigma_boundaries || !out->u_wind || !out->v_wind ||
      !out->temperature || !out->specific_humidity || !out->geopotential || !out->normalized_surface_pressure)
    return fail(SPEEDY_ERR_ARG, "speedy_dynamics_to_physics: null field");
  const int kx = spec->kx, nsp = in->nsp;
  if (nsp < 1 || nsp > kx || out->nsp != nsp)
    return fail(SPEEDY_ERR_ARG, "speedy_dynamics_to_physics: nsp must be in 1..kx and equal on both sides");
  const int nfield = 5 * kx + nsp;
  if (int rc = check_call(p, nfield, ws, wsb)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const Ws w = carve(p, nfield, ws);
  // 1. modal (u cos, v cos), the other modal inputs and the modal geopotential,
  //    stacked in z: rows [u kx][v kx][T_var kx][q kx][log ps nsp][phi kx]
  ModalArgs m{};
  m.mm = p->mm; m.nl = p->n_l; m.kx = kx;
  m.inv_radius = (float)(1.0 / p->radius);
  m.radius2 = (float)(p->radius * p->radius);
  m.vor = in->vorticity; m.div = in->divergence;
  m.copy_src[0] = in->temperature_variation; m.copy_count[0] = kx;
  m.copy_src[1] = in->specific_humidity; m.copy_count[1] = kx;
  m.copy_src[2] = in->log_surface_pressure; m.copy_count[2] = nsp;
  m.out = w.z;
  m.tvar = in->temperature_variation;
  m.orography = spec->orography;
  m.reference = spec->reference_temperature;
  m.gravity = (float)spec->gravity;
  m.geo = 1;
  // hydrostatic trapezoid on log(sigma): alpha_k = (ln s_k+1 - ln s_k)/2,
  // alpha_K-1 = -ln s_K-1; W[j][j] = alpha_j, W[j][k>j] = alpha_k + alpha_k-1
  std::vector<double> ls(kx), alpha(kx);
  for (int k = 0; k < kx; ++k) ls[k] = std::log(0.5 * (spec->sigma_boundaries[k] + spec->sigma_boundaries[k + 1]));
  for (int k = 0; k + 1 < kx; ++k) alpha[k] = 0.5 * (ls[k + 1] - ls[k]);
  alpha[kx - 1] = -ls[kx - 1];
  for (int j = 0; j < kx; ++j)
    for (int k = 0; k < kx; ++k)
      m.geo_w[j][k] = (float)(spec->gas_constant * (k < j ? 0.0 : k == j ? alpha[j] : alpha[k] + alpha[k - 1]));
  const size_t per = (size_t)p->mm * p->n_l;
  hipLaunchKernelGGL(modal_winds_kernel, dim3(blocks(per), nfield - kx), dim3(kThreads), 0, s, m);
  if (int rc = hip_result(hipGetLastError())) return rc;
  // 2. synthesis with the nodal post-operations
  SegTable seg{};
  add_seg(seg, nullptr, out->u_wind, kx, kOpSecLat, 1.0f);
  add_seg(seg, nullptr, out->v_wind, kx, kOpSecLat, 1.0f);
  add_seg(seg, nullptr, out->temperature, kx, kOpAddLevel, 1.0f, spec->reference_temperature, kx);
  add_seg(seg, nullptr, out->specific_humidity, kx, 0, 1000.0f);
  add_seg(seg, nullptr, out->normalized_surface_pressure, nsp, kOpExp, 1.0f);
  add_seg(seg, nullptr, out->geopotential, kx, 0, 1.0f);
  return hip_result(run_synthesis(p, nfield, seg, w.z, w.y, s));
}

namespace {
int nodal_to_spectral(const SpeedyShtPlan* p, const SpeedyDynamicsSpec* spec, const SpeedyNodalState* in,
                      const SpeedySpectralState* out, void* ws, size_t wsb, void* stream, bool tendency) {
  const char* who = tendency ? "speedy_physics_tendency_to_dynamics" : "speedy_physics_to_dynamics";
  if (int rc = check_spec(spec)) return rc;
  if (!in || !out || !in->u_wind || !in->v_wind || !in->temperature || !in->specific_humidity ||
      (!tendency && !in->normalized_surface_pressure) || !out->vorticity || !out->divergence ||
      !out->temperature_variation || !out->specific_humidity || !out->log_surface_pressure)
    return fail(SPEEDY_ERR_ARG, std::string(who) + ": null field");
  const int kx = spec->kx, nsp = out->nsp;
  if (nsp < 1 || nsp > kx || (!tendency && in->nsp != nsp))
    return fail(SPEEDY_ERR_ARG, std::string(who) + ": nsp must be in 1..kx and equal on both sides");
  const int nsp_in = tendency ? 0 : nsp;
  const int nfield = 4 * kx + nsp_in;
  if (int rc = check_call(p, nfield, ws, wsb)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const Ws w = carve(p, nfield, ws);
  SegTable seg{};
  add_seg(seg, in->u_wind, nullptr, kx, kOpSecLat, 1.0f);
  add_seg(seg, in->v_wind, nullptr, kx, kOpSecLat, 1.0f);
  if (tendency)
    add_seg(seg, in->temperature, nullptr, kx, 0, 1.0f);
  else
    add_seg(seg, in->temperature, nullptr, kx, kOpSubLevel, 1.0f, spec->reference_temperature, kx);
  add_seg(seg, in->specific_humidity, nullptr, kx, 0, 1e-3f);
  if (!tendency) add_seg(seg, in->normalized_surface_pressure, nullptr, nsp, kOpLog, 1.0f);
  if (int rc = hip_result(run_analysis(p, nfield, seg, w.y, w.z, s))) return rc;
  ModalArgs m{};
  m.mm = p->mm; m.nl = p->n_l; m.kx = kx;
  m.inv_radius = (float)(1.0 / p->radius);
  m.radius2 = (float)(p->radius * p->radius);
  m.out = w.z;
  m.vor_out = out->vorticity; m.div_out = out->divergence;
  m.copy_src[0] = w.z; m.copy_count[0] = kx; m.copy_dst[0] = out->temperature_variation;
  m.copy_src[1] = w.z; m.copy_count[1] = kx; m.copy_dst[1] = out->specific_humidity;
  m.copy_src[2] = tendency ? nullptr : w.z; m.copy_count[2] = nsp; m.copy_dst[2] = out->log_surface_pressure;
  const size_t per = (size_t)p->mm * p->n_l;
  hipLaunchKernelGGL(modal_vordiv_kernel, dim3(blocks(per), 3 * kx + nsp), dim3(kThreads), 0, s, m);
  return hip_result(hipGetLastError());
}
}  // namespace

int speedy_physics_to_dynamics(const SpeedyShtPlan* p, const SpeedyDynamicsSpec* spec, const SpeedyNodalState* in,
                               const SpeedySpectralState* out, void* ws, size_t wsb, void* stream) {
  return nodal_to_spectral(p, spec, in, out, ws, wsb, stream, false);
}

int speedy_physics_tendency_to_dynamics(const SpeedyShtPlan* p, const SpeedyDynamicsSpec* spec,
                                        const SpeedyNodalState* in, const SpeedySpectralState* out, void* ws,
                                        size_t wsb, void* stream) {
  return nodal_to_spectral(p, spec, in, out, ws, wsb, stream, true);
}

const char* speedy_sht_last_error(void) { return g_err.c_str(); }

}  // extern "C"
