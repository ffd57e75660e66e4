/*
 * speedy_sht.h -- C ABI of the MI355X (gfx950) dynamics <-> physics transforms.
 *
 * The library replaces the three conversions between the dynamical core's
 * spectral state and the physics' nodal columns:
 *     dynamics_state_to_physics_state        jcm/physics_interface.py:196-234
 *     physics_state_to_dynamics_state        jcm/physics_interface.py:237-273
 *     physics_tendency_to_dynamics_tendency  jcm/physics_interface.py:275-303
 * and the spherical-harmonic transforms under them (dinosaur's
 * Grid.to_modal / Grid.to_nodal with RealSphericalHarmonics, selected at
 * jcm/utils.py:28-54).  Both transform directions are a longitude Fourier
 * contraction (a truncated real DFT in two short passes) and a latitude
 * Legendre contraction (a batched fp32 MFMA GEMM) on gfx950.
 *
 * Conventions (dinosaur RealSphericalHarmonics on a Gaussian grid):
 *   nodal field  float[nlon][nlat], latitude fastest, Gauss-Legendre latitudes
 *                south to north, longitudes 2*pi*i/nlon;
 *   modal field  float[2*n_m-1][n_l]; mode 0 is m=0, mode 2m-1 cos(m lon),
 *                mode 2m sin(m lon); entries with m > l are zero on output
 *                and ignored on input;
 *   Fourier basis 1/sqrt(2 pi), cos(m lon)/sqrt(pi), sin(m lon)/sqrt(pi);
 *   Legendre basis of unit L2 norm on [-1, 1], Condon-Shortley phase;
 *   3-D fields stack levels (or any batch) on a leading axis.
 * Units are SI (dinosaur's SI_SCALE): specific humidity is g/kg on the
 * physics side and kg/kg on the dynamics side.
 *
 * Array pointers are device pointers; structs are host memory read during the
 * call.  Every entry point returns SPEEDY_OK (0) or a nonzero SPEEDY_ERR_*
 * code and never aborts; speedy_sht_last_error() gives a message for the
 * calling thread.
 */
#ifndef SPEEDY_SHT_H
#define SPEEDY_SHT_H

#include <stddef.h>
#include <stdint.h>

#include "speedy_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPEEDY_SHT_KX_MAX 16

typedef struct SpeedyShtPlan SpeedyShtPlan;

/* Host-only: the grid's Gauss-Legendre nodes and weights (float64, south to
 * north; weights sum to 2) and, when not null, the float32 bases the device
 * uses: fourier[nlon][2*n_m-1] and legendre[n_m][nlat][n_l].  No HIP call.
 * Replaces the basis of dinosaur's RealSphericalHarmonics (the nodes come
 * from Newton iteration on P_nlat, the Legendre functions from the
 * three-term recurrence in float64). */
int speedy_sht_host_basis(int32_t n_m, int32_t n_l, int32_t nlon, int32_t nlat, double *sin_lat,
                          double *weights, float *fourier, float *legendre);

/* A plan holds the device bases of one grid (Grid.T<N>: n_m longitude
 * wavenumbers, n_l total wavenumbers, nlon x nlat nodes, sphere radius in m).
 * Creation allocates and uploads them on the current device (synchronous);
 * it requires n_m <= n_l, 2 <= n_l <= nlat, 2*n_m-1 <= nlon and an even nlat
 * (the Legendre stages use the equatorial symmetry of the Gaussian grid), and
 * an nlon whose Fourier stages fit 64 KiB of LDS per workgroup (up to about
 * 3500 longitudes; every Grid.T<N> up to T425 does). */
int speedy_sht_create(int32_t n_m, int32_t n_l, int32_t nlon, int32_t nlat, double radius,
                      SpeedyShtPlan **plan);
int speedy_sht_destroy(SpeedyShtPlan *plan);

/* Device bytes of workspace a call with `nfield` stacked fields needs. */
size_t speedy_sht_workspace_bytes(const SpeedyShtPlan *plan, int32_t nfield);

/* Grid.to_modal: modal[f] = Legendre(weights * Fourier^T nodal[f]) for nfield
 * fields (dinosaur RealSphericalHarmonics.transform). */
int speedy_sht_to_modal(const SpeedyShtPlan *plan, const float *nodal, float *modal, int32_t nfield,
                        void *workspace, size_t workspace_bytes, void *stream);

/* Grid.to_nodal: nodal[f] = Fourier(Legendre^T modal[f])
 * (dinosaur RealSphericalHarmonics.inverse_transform). */
int speedy_sht_to_nodal(const SpeedyShtPlan *plan, const float *modal, float *nodal, int32_t nfield,
                        void *workspace, size_t workspace_bytes, void *stream);

/* What the conversions read from dinosaur's PrimitiveEquations. */
typedef struct {
  int32_t kx;                          /* levels, 1..SPEEDY_SHT_KX_MAX */
  const double *sigma_boundaries;      /* host [kx+1], 0 at the top, 1 at the surface */
  const float *reference_temperature;  /* device [kx], K */
  const float *orography;              /* device modal [2*n_m-1][n_l], m */
  double gravity;                      /* m s-2 (dinosaur: 9.80616) */
  double gas_constant;                 /* J kg-1 K-1 (dinosaur: 1004 * 2/7) */
} SpeedyDynamicsSpec;

/* Modal (spectral) state or tendency, dinosaur primitive_equations.State.
 * vorticity, divergence, temperature_variation, specific_humidity:
 * [kx][2*n_m-1][n_l]; log_surface_pressure: [nsp][2*n_m-1][n_l]. */
typedef struct {
  float *vorticity;
  float *divergence;
  float *temperature_variation;
  float *log_surface_pressure;
  float *specific_humidity;
  int32_t nsp;                         /* fields in log_surface_pressure (1 or kx) */
} SpeedySpectralState;

/* Nodal physics state or tendency: u, v, temperature, specific_humidity,
 * geopotential [kx][nlon][nlat]; normalized_surface_pressure [nsp][nlon][nlat]. */
typedef struct {
  float *u_wind;
  float *v_wind;
  float *temperature;
  float *specific_humidity;
  float *geopotential;
  float *normalized_surface_pressure;
  int32_t nsp;
} SpeedyNodalState;

/* dynamics_state_to_physics_state (physics_interface.py:196-234):
 * u, v from vorticity / divergence (vor_div_to_uv_nodal), temperature =
 * to_nodal(temperature_variation) + reference_temperature, specific humidity
 * in g/kg, geopotential = g * orography + hydrostatic sum of the temperature
 * (get_geopotential), surface pressure = exp(to_nodal(log_surface_pressure)).
 * Reads `in` (nsp log-pressure fields), writes every field of `out`
 * (out->nsp must equal in->nsp). */
int speedy_dynamics_to_physics(const SpeedyShtPlan *plan, const SpeedyDynamicsSpec *spec,
                               const SpeedySpectralState *in, const SpeedyNodalState *out, void *workspace,
                               size_t workspace_bytes, void *stream);

/* physics_state_to_dynamics_state (physics_interface.py:237-273): vorticity
 * and divergence of (u, v) (uv_nodal_to_vor_div_modal), temperature minus
 * the reference, specific humidity in kg/kg, log of surface pressure.  The
 * geopotential of `in` is not read.  out->nsp must equal in->nsp. */
int speedy_physics_to_dynamics(const SpeedyShtPlan *plan, const SpeedyDynamicsSpec *spec,
                               const SpeedyNodalState *in, const SpeedySpectralState *out, void *workspace,
                               size_t workspace_bytes, void *stream);

/* physics_tendency_to_dynamics_tendency (physics_interface.py:275-303):
 * tendencies of u, v, temperature and specific humidity (g/kg/s) to modal
 * vorticity, divergence, temperature and specific humidity (kg/kg/s); the
 * log-surface-pressure tendency is written as zeros (out->nsp fields).
 * in->geopotential and in->normalized_surface_pressure are not read. */
int speedy_physics_tendency_to_dynamics(const SpeedyShtPlan *plan, const SpeedyDynamicsSpec *spec,
                                        const SpeedyNodalState *in, const SpeedySpectralState *out,
                                        void *workspace, size_t workspace_bytes, void *stream);

/* Message of the calling thread's last failed call ("" if none). */
const char *speedy_sht_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SPEEDY_SHT_H */
