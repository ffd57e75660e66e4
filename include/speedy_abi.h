/*
 * speedy_abi.h -- C ABI of the MI355X (gfx950) SPEEDY column-physics library.
 *
 * The library replaces, for the SPEEDY physics suite, the call
 *     SpeedyPhysics.compute_tendencies(state, forcing, geometry, date)
 *       reference: jcm/physics/speedy/speedy_physics.py:88-127
 *       plugin API: jcm/physics_interface.py:137-155 (Physics.compute_tendencies)
 * i.e. the eleven physics terms listed at speedy_physics.py:67-80, summed in
 * term order.  Plain pointers and sizes only: every array pointer is a device
 * pointer (hipMalloc'ed or from a framework), every struct is host memory
 * that is read during the call and may be freed afterwards.
 *
 * Layout (structure of arrays, level-major): a 3-D field is float[kx][ncol],
 * a 2-D field float[ncol], with ncol = nlon*nlat and column c = i*nlat + j
 * (latitude fastest, south to north).  This is the C-order layout of the
 * reference's (kx, ix, il) arrays (jcm/geometry.py:89,130), so framework
 * buffers are consumed without copies.  Members of an ensemble are stacked
 * on a leading axis: state/tendencies float[nmember][kx][ncol],
 * ps float[nmember][ncol]; forcing and geometry are shared by all members.
 *
 * Errors: every entry point returns SPEEDY_OK (0) or a nonzero code and never
 * aborts; speedy_last_error() gives a message for the calling thread.  NaN
 * inputs propagate as in the reference except through min/max, where the
 * GPU's IEEE maxNum/minNum return the non-NaN operand.
 */
#ifndef SPEEDY_ABI_H
#define SPEEDY_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPEEDY_ABI_VERSION 1
#define SPEEDY_KX_MAX 16

enum {
  SPEEDY_OK = 0,
  SPEEDY_ERR_ARG = 1,        /* null pointer, bad size, bad flag */
  SPEEDY_ERR_KX = 2,         /* kx not in {7, 8, 16} */
  SPEEDY_ERR_WORKSPACE = 3,  /* workspace missing or too small */
  SPEEDY_ERR_HIP = 4,        /* a HIP runtime call failed */
  SPEEDY_ERR_CAPACITY = 5    /* request exceeds a stated maximum */
};

/* flags */
#define SPEEDY_FLAG_LFLUXLAND 0x1u   /* ForcingData.lfluxland (forcing.py:17) */
/* The two steps around the physics in get_physical_tendencies
 * (physics_interface.py:371-375), fused into the same kernel:
 * VERIFY_STATE: the physics sees q clamped at 0 (verify_state, :305-319);
 * VERIFY_TENDENCIES: the q tendency is limited so that q + dt*dq >= 0 with the
 * unclamped q and dt = SpeedyDate.dt_seconds (verify_tendencies, :321-342). */
#define SPEEDY_FLAG_VERIFY_STATE 0x2u
#define SPEEDY_FLAG_VERIFY_TENDENCIES 0x4u
/* PLAIN_LAUNCH (no counterpart in the reference): an uncaptured call whose grid
 * fits on the GPU at once makes a plain launch of the single-launch kernel
 * instead of a cooperative one, as calls under graph capture always do.  The
 * cooperative launch guarantees that all workgroups are resident together,
 * which the kernel's grid-wide level-set exchange needs; it costs about 10 us
 * per call on the GPU at T85.  With this flag, co-residency holds only when no
 * other work shares the GPU.  The exchange's wait is bounded, so the kernel
 * always ends, but if it gave up, speedy_workspace_status reports it and the
 * call's results are invalid.  Ignored by speedy_fwd_batch. */
#define SPEEDY_FLAG_PLAIN_LAUNCH 0x8u
#define SPEEDY_FLAGS_ALL 0xFu

/* term_mask bit i selects the tendency of term i (speedy_physics.py:67-80)
 * for the sum; every term always runs, so the data handed between terms are
 * unchanged.  SPEEDY_TERMS_ALL is the reference's composition. */
enum {
  SPEEDY_TERM_FLAGS = 0, SPEEDY_TERM_FORCING, SPEEDY_TERM_HUMIDITY, SPEEDY_TERM_CONVECTION,
  SPEEDY_TERM_CONDENSATION, SPEEDY_TERM_CLOUDS, SPEEDY_TERM_SHORTWAVE, SPEEDY_TERM_LW_DOWN,
  SPEEDY_TERM_SURFACE, SPEEDY_TERM_LW_UP, SPEEDY_TERM_VDIFF, SPEEDY_NTERMS
};
#define SPEEDY_TERMS_ALL ((1u << SPEEDY_NTERMS) - 1u)

/* Parameters.default() (jcm/physics/speedy/params.py:244-254), packed in
 * declaration order, plus two multipliers of this library (co2_scale on
 * ablco2, drag_scale on cdl and cds; 1.0 reproduces the reference). */
typedef struct {
  float psmin, trcnv, rhil, rhbl, entmax, smf;                       /* params.py:11-16 */
  float trlsc, rhlsc, drhlsc, rhblsc;                                /* params.py:51-54 */
  float albcl, albcls, absdry, absaer, abswv1, abswv2, abscl1, abscl2,
        ablwin, ablwv1, ablwv2, ablcl1, ablcl2, rhcl1, rhcl2, qacl, wpcl,
        pmaxcl, clsmax, clsminl, gse_s0, gse_s1;                     /* params.py:72-99 */
  float albsea, albice, albsn, epslw, emisfc;                        /* params.py:134-140 */
  float fwind0, ftemp0, fhum0, cdl, cds, chl, chs, vgust, ctday, dtheta,
        fstab, clambda, clambsn, hdrag;                              /* params.py:157-183 */
  float trshc, trvdi, trvds, redshc, rhgrad, segrad;                 /* params.py:213-218 */
  float co2_scale, drag_scale;                                       /* library multipliers */
  int32_t lscasym, lskineb, increase_co2, co2_year_ref;              /* params.py:180-181,34-35 */
} SpeedyParams;

/* Per-level constants of Geometry (jcm/geometry.py:62-85). */
typedef struct {
  int32_t kx;
  float hsg[SPEEDY_KX_MAX + 1];
  float fsg[SPEEDY_KX_MAX], dhs[SPEEDY_KX_MAX], sigl[SPEEDY_KX_MAX];
  float grdsig[SPEEDY_KX_MAX], grdscp[SPEEDY_KX_MAX];
  float wvi[SPEEDY_KX_MAX][2];
} SpeedyVertical;

/* DateData (jcm/date.py:9-48). */
typedef struct {
  float tyear;
  int32_t model_year;
  int32_t model_step;
  float dt_seconds;
} SpeedyDate;

/* Device pointers of one call.  Inputs are read-only. */
typedef struct {
  /* PhysicsState (physics_interface.py:22-29): [nmember][kx][ncol]; ps [nmember][ncol] */
  const float *u, *v, *t, *q, *phi, *ps;
  /* ForcingData after the day slice (forcing.py:8-17): [ncol] */
  const float *alb0, *sice_am, *snowc_am, *soilw_am, *stl_am, *sst;
  /* Geometry (geometry.py:91-97): fmask, orog, phis0 [ncol]; sia, coa [nlat] */
  const float *fmask, *orog, *phis0, *sia, *coa;
  /* PhysicsTendency (physics_interface.py:89-94): [nmember][kx][ncol] */
  float *du, *dv, *dt, *dq;
  /* optional PhysicsData subset, [nmember][speedy_diag_rows(kx)][ncol], or NULL */
  float *diag;
  /* optional float64 budgets, [nmember][SPEEDY_NBUDGET], or NULL (see below) */
  double *budget;
} SpeedyFields;

/* Float64 column budgets, area-weighted (weight cos(latitude) = coa) global
 * means per member, of the call's tendencies (after any verify step):
 *   0  column energy tendency   sum_k cp dT_k dp_k/g              [W m-2]
 *   1  column water tendency    sum_k dq_k dp_k/g                 [g m-2 s-1]
 *   2  surface evaporation      E                                 [g m-2 s-1]
 *   3  precipitation            P = precnv + precls               [g m-2 s-1]
 *   4  water budget residual    (column water tendency) - (E - P) [g m-2 s-1]
 *   5  largest |water residual| of any column                     [g m-2 s-1]
 * with dp_k/g = dhs_k ps p0 / g.  The per-column sums and the global
 * reduction run in float64 on the GPU (fixed order: deterministic); the
 * physics itself stays float32.  Without lfluxland, E is 0.
 * (Column water conservation: large_scale_condensation.py:77-79,
 * convection.py:181-219, surface_flux.py:297, vertical_diffusion.py:117-118.) */
#define SPEEDY_NBUDGET 6

/* Cotangents and gradients of the fused forward+VJP call. */
typedef struct {
  /* cotangents of the tendencies, [nmember][kx][ncol] */
  const float *ct_du, *ct_dv, *ct_dt, *ct_dq;
  /* gradients w.r.t. the state, [nmember][kx][ncol]; g_ps [nmember][ncol] */
  float *g_u, *g_v, *g_t, *g_q, *g_ps;
  /* gradients w.r.t. (rhbl, trcnv, rhlsc, co2_scale, drag_scale) summed over
   * the columns of each member: float[nmember][5] */
  float *g_params;
} SpeedyAdjoint;

/* Version of this ABI (SPEEDY_ABI_VERSION). */
int speedy_abi_version(void);

/* Bytes of device workspace a call on this grid needs (0 if unsupported).  A
 * workspace may be reused by later calls of the same or a smaller grid. */
size_t speedy_workspace_bytes(int32_t nlon, int32_t nlat, int32_t nmember);

/* Zero a fresh workspace once (a fill kernel on `stream`); later calls leave it
 * ready for the next call. */
int speedy_workspace_init(void *workspace, size_t workspace_bytes, void *stream);

/* Status of a workspace: bit 0 is set if a call's grid barrier gave up
 * waiting (that call's results are invalid).  Reading a set bit clears it and
 * re-zeroes the whole workspace (a fill kernel on `stream`): a failed wait
 * leaves late blocks' level-set words and the generation word out of step, and
 * the next call then starts from a clean workspace. */
int speedy_workspace_status(void *workspace, size_t workspace_bytes, uint32_t *status, void *stream);

/* Rows of the diagnostic output for kx levels, and their names as a
 * comma-separated list ("name" for a 2-D row, "name[k]" rows for 3-D). */
int speedy_diag_rows(int32_t kx);
int speedy_diag_names(int32_t kx, char *buf, size_t buf_len);

/* Forward: PhysicsTendency of SpeedyPhysics.compute_tendencies
 * (speedy_physics.py:88-127) for nmember stacked states.
 * level_set < 0: the Q1 cloud-top level set (shortwave_radiation.py:63-70) is
 * computed per member over all its columns; >= 0: used as given (bit k =
 * level k), for parity tests.  A grid whose workgroups fit on the GPU at once
 * runs as one cooperative launch; other grids (and a given level set) run as a
 * one-segment batch, described in the workspace, in two to four launches.
 * Stream-ordered and safe to capture in a HIP graph; no allocation. */
int speedy_fwd(const SpeedyFields *fields, const SpeedyVertical *vert, const SpeedyParams *params,
               const SpeedyDate *date, int32_t nlon, int32_t nlat, int32_t nmember, uint32_t flags,
               uint32_t term_mask, int32_t level_set, void *workspace, size_t workspace_bytes,
               void *stream);

/* Allocation-free speedy_fwd.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The same forward and
 * the same results as speedy_fwd, with the workspace taken from a static device
 * array of speedy_noalloc_max_bytes() (4 MiB) per GPU, held until the process
 * ends; a grid whose speedy_workspace_bytes exceeds it is rejected with
 * SPEEDY_ERR_CAPACITY.  The caller passes inputs and outputs as for speedy_fwd;
 * no hipMalloc, hipFree or allocating call is made.  Every call on one GPU uses
 * that GPU's array: calls that overlap on one GPU (two streams, two threads)
 * race.  The call always runs as a one-segment batch (two to four launches),
 * which reads nothing an earlier call left in the array. */
size_t speedy_noalloc_max_bytes(void);
int speedy_fwd_noalloc(const SpeedyFields *fields, const SpeedyVertical *vert, const SpeedyParams *params,
                       const SpeedyDate *date, int32_t nlon, int32_t nlat, int32_t nmember, uint32_t flags,
                       uint32_t term_mask, int32_t level_set, void *stream);

/* Fused forward + VJP: the forward tendencies (as speedy_fwd with every term)
 * and the cotangent-weighted gradients w.r.t. T, q, u, v, ps and five
 * parameters.  Replaces jax.vjp of compute_tendencies (speedy_physics.py:88-127)
 * with respect to PhysicsState (physics_interface.py:22-29) and Parameters
 * (params.py:22,24,60 and the two library multipliers).  Discrete decisions
 * (convective and cloud tops, the level set S) are held fixed, as under JAX
 * differentiation.  g_u and g_v are zero below level kx-1 (only the surface
 * stress depends on the winds).  fields->diag must be NULL.  The parameter
 * gradients are summed over each member's columns in a fixed order
 * (deterministic), by a second small kernel on the same stream. */
int speedy_fwd_vjp(const SpeedyFields *fields, const SpeedyAdjoint *adj, const SpeedyVertical *vert,
                   const SpeedyParams *params, const SpeedyDate *date, int32_t nlon, int32_t nlat,
                   int32_t nmember, uint32_t flags, int32_t level_set, void *workspace,
                   size_t workspace_bytes, void *stream);

/* One request of a batched call (speedy_fwd_batch): its own grid, members,
 * fields, date and kind.  All adjoint pointers NULL: forward only; all set:
 * forward + VJP.  The workspace is the segment's own (speedy_workspace_bytes of
 * its grid, initialised once by speedy_workspace_init). */
typedef struct {
  SpeedyFields fields;          /* fields.diag must be NULL */
  SpeedyAdjoint adj;
  SpeedyDate date;
  int32_t nlon, nlat, nmember;
  uint32_t flags;
  int32_t level_set;
  int32_t reserved;
  void *workspace;
  size_t workspace_bytes;
} SpeedySegment;

#define SPEEDY_BATCH_MAX_SEGMENTS 4096

/* Bytes of the device descriptor buffer (and of the host staging buffer) that
 * speedy_fwd_batch needs for these segments (0 if they cannot be batched). */
size_t speedy_batch_desc_bytes(const SpeedySegment *segs, int32_t nseg);

/* Batched forward / forward+VJP of independent requests that share the vertical
 * grid and the parameters (members of one forecast system at mixed resolutions
 * and dates): every segment's result equals its own speedy_fwd / speedy_fwd_vjp
 * call.  No counterpart in the reference (its calls are one grid at a time); it
 * serves continuous batching.  Writes the descriptor to host_staging (pinned
 * host memory of speedy_batch_desc_bytes, not to be reused until this call's
 * copy has completed on `stream`), copies it to `desc` (device memory of the
 * same size) and launches, in stream order: the level sets of every segment,
 * one column kernel over all segments' columns, and the parameter-gradient sums
 * of the VJP segments.  No allocation. */
int speedy_fwd_batch(const SpeedySegment *segs, int32_t nseg, const SpeedyVertical *vert,
                     const SpeedyParams *params, void *desc, size_t desc_bytes, void *host_staging,
                     void *stream);

/* Level set S of the last call's members (bit k = level k), copied to host. */
int speedy_read_level_sets(const void *workspace, size_t workspace_bytes, int32_t nmember,
                           uint32_t *out, void *stream);

/* Running mean of the diagnostics on the device: acc[e] += diag[e] /
 * steps_to_average for e < n, in float32 with an IEEE division, on `stream`.
 * Replaces the tree_map add of DiagnosticsCollector.accumulate_if_physical_step
 * (jcm/model.py:57-64) for the PhysicsData subset the forward writes (diag of
 * SpeedyFields); acc is one output slot of the stacked averages. */
int speedy_diag_accumulate(const float *diag, float *acc, int64_t n, float steps_to_average, void *stream);

/* Kernel timers (no counterpart in the reference; for measurement).  While
 * enabled, every call records HIP events on its stream before and after each
 * of its kernels (up to `capacity` calls, then recording stops).
 * speedy_timer_read synchronizes on the events and returns, per kernel kind
 * (0 = level-set kernel, 1 = forward kernel, 2 = forward+VJP kernel, 3 =
 * parameter-gradient reduction), the summed milliseconds and launch count; it
 * then clears them. */
int speedy_timer_enable(int32_t capacity);
int speedy_timer_disable(void);
int speedy_timer_read(float *ms_per_kind, int32_t *launches_per_kind, int32_t nkinds);

/* Message of the last error on the calling thread ("" if none). */
const char *speedy_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* SPEEDY_ABI_H */
