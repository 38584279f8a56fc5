/*
 * teuk_oracle.h -- CPU FP64 restatement of teuk-kokkos's linear per-stage
 * Teukolsky operator and RK4 stepper.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (teuk_kokkos_amd/) never links, loads or calls it.
 *
 * It is a nodal restatement that follows the reference's own structure
 * (nodal (mode, field=13, radial, theta) state, per-mode analyze/diag/
 * synthesize angular Laplacian, three-pass radial RHS, per-field band
 * projection, four-stage classical RK4), so it is independent of the modal
 * formulation used on the GPU.  Each function cites the reference file:line
 * it follows.
 *
 * Parity pins: see tests/test_oracle_*.py (SBP identities, exactness and
 * orders, dissipation energy identity and spectral radius, coefficient point
 * oracle, Gauss-Legendre moments, harmonic closed forms, orthonormality,
 * Galerkin residual, band invariance, RK4 temporal order, QNM values).
 * Complex division uses the libstdc++ algorithm; the reference uses
 * Kokkos::complex division, whose bitwise rounding is "parity unpinned".
 */
#ifndef TEUK_ORACLE_H
#define TEUK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Radial scheme ids: 0 = d4-2, 1 = d8-4, 2 = d10-5. */
/* Reduction ids: 0 = free_damped, 1 = stage_constrained. */

typedef struct OracleParams {
  double mass, spin, length;      /* Kerr M, a, compactification L        */
  int nr, ntheta;                 /* radial points, Gauss-Legendre nodes  */
  int ell_max_first, ell_max_second;
  int n_modes;                    /* stored signed modes (sorted)         */
  int modes[64];
  int n_parents;                  /* first-order modes                    */
  int parents[64];
  int n_targets;                  /* second-order modes                   */
  int targets[64];
  double reduction_damping;       /* gamma                                */
  double dissipation;             /* epsilon                              */
  int reduction;                  /* 0 free_damped, 1 stage_constrained   */
  int scheme;                     /* 0 d4-2, 1 d8-4, 2 d10-5              */
  int threads;                    /* OpenMP threads (<=0: runtime default)*/
  /* Second-order source policy (spatial_pipeline.hpp:35-62): 0 disabled
   * (linear only), 1 constraint-aware, 2 unrestricted. */
  int source_mode;
  double source_start_time;
  double constraint_tolerance;
  int required_passes;
} OracleParams;

/* SourceActivationState (source_activation.hpp:17-22). */
typedef struct OracleActivation {
  int active;
  double activation_time;
  int consecutive_passes;
  double last_eligibility_time;
} OracleActivation;

/* SourceConstraintNorms (source_activation.hpp:24-37); families in the
 * order psi3 Bianchi, psi2 Bianchi, h_ll reality. */
typedef struct OracleConstraintNorms {
  double absolute_maximum[3];
  double normalized_maximum[3];
  double weighted_rms[3];
  double normalized_weighted[3];
  double controlling_normalized;
  int all_finite;
} OracleConstraintNorms;

typedef struct OraclePipeline OraclePipeline;

/* Returns NULL on invalid parameters (message via oracle_last_error). */
OraclePipeline* oracle_create(const OracleParams* params);
void oracle_destroy(OraclePipeline* p);
const char* oracle_last_error(void);

/* State is interleaved complex128 (re, im) in (mode, field=13, radial,
 * theta) order: 2 * M * 13 * nr * ntheta doubles. */
long long oracle_value_count(const OraclePipeline* p);
void oracle_get_state(const OraclePipeline* p, double* out);
void oracle_set_state(OraclePipeline* p, const double* in);

/* Gaussian initial data (pipeline_initial_data.hpp:211-340).  center and
 * width are absolute compactified radii.  Seeds are already expanded (sharp
 * partners included) and sorted by (m, ell).  Returns the Galerkin residual
 * or a negative value on failure. */
double oracle_init_gaussian(OraclePipeline* p, double center, double width,
                            int compact_support, int n_seeds, const int* ells,
                            const int* ms, const double* amp_re,
                            const double* amp_im);

/* One spatial RHS (spatial_pipeline.hpp:458-515, Disabled policy). */
void oracle_evaluate_rhs(OraclePipeline* p, const double* stage, double* out);
/* Classical RK4 steps of the owned state (device_rk4.hpp:84-128), split at
 * the source start time with the activation latch considered on accepted
 * states (spatial_pipeline.hpp:590-620).  Returns 0, or -1 on error. */
int oracle_step(OraclePipeline* p, double time, double dt);
int oracle_advance(OraclePipeline* p, int steps, double time0, double dt);

/* Activation latch and the last constraint sample. */
void oracle_get_activation(const OraclePipeline* p, OracleActivation* out);
void oracle_set_activation(OraclePipeline* p, const OracleActivation* in);
void oracle_constraint_norms(const OraclePipeline* p, OracleConstraintNorms* out);
/* Consider activation on the owned state at `time`; returns the latch. */
int oracle_consider_activation(OraclePipeline* p, double time);
/* Latch seen by each RK stage of the last step (count returned). */
int oracle_last_stage_sources(const OraclePipeline* p, int* flags);
/* Intermediates of the last RHS evaluation (M x nr x ntheta complex, or
 * M x 2 x nr x ntheta for the D, T pairs): 0 forcing, 1 source_over_r3,
 * 2 projected inner (D, T), 3 projected inner tangent, 4 constraint
 * residuals (M x 3 x ...), 5 constraint terms (M x 6 x ...), 6 radial D of
 * the stage reconstruction fields G..U (M x 7 x ...), 7 the stage chain's
 * angular inputs Eth1F, Eth2G, Eth2C, Eth2Pi, Eth'1B#, Eth'2C# (M x 6 x ...),
 * 8 the ordered-pair sums D, T before projection (M x 2 x ...).
 * Returns the
 * complex value count, or -1. */
long long oracle_intermediate(const OraclePipeline* p, int which, double* out);

/* Modal endpoint coefficients of FirstPsi (waveform_output.hpp:40-160) for
 * one stored mode index: out[2*n_ell] at scri (radial 0) and horizon
 * (radial nr-1); returns n_ell or -1. */
int oracle_endpoint_modal(const OraclePipeline* p, int mode_index,
                          double* scri_out, double* horizon_out);

/* ---- primitives used by the known-answer tests ---- */
void oracle_gauss_legendre(int n, double* x, double* w);
double oracle_swsh_theta(int ell, int m, int spin, double theta);
/* S is ntheta x n_ell row-major, A is n_ell x ntheta row-major. */
int oracle_transform_matrices(int spin, int m, int ell_max, int ntheta,
                              double* S, double* A);
/* out: C_T, K, H_R, Re G, Im G, Re q, Im q, Re psi, Im psi. */
void oracle_teukolsky_coefficients(double mass, double spin, double length,
                                   int spin_weight, int m, double radius,
                                   double theta, double* out);
double oracle_sbp_derivative_entry(int scheme, int n, int row, int column);
double oracle_sbp_norm_weight(int scheme, int n, int index);
int oracle_sbp_minimum_points(int scheme);
/* complex lines (interleaved), length n */
void oracle_sbp_derivative(int scheme, int n, double h, const double* in,
                           double* out);
void oracle_sbp_dissipation(int scheme, int n, double h, double strength,
                            const double* in, double* out);
double oracle_dissipation_spectral_radius_bound(int scheme);
int oracle_dissipation_step_admissible(int scheme, double strength,
                                       double step_over_spacing);
/* maximum |coordinate velocity| over the grid (solver_driver.hpp:36-57) */
double oracle_maximum_coordinate_speed(double mass, double spin,
                                       double length, int nr, int ntheta);
/* dense complex partial-pivot solve (dense_solve.hpp:20-72); returns the
 * relative residual or -1 on a singular matrix. */
double oracle_dense_solve(int n, const double* matrix, const double* rhs,
                          double* solution);

/* ---- second-order point functions used by the known-answer tests ---- */
/* out: mu0, tau0, pi0, psi20, rho0, epsilon0 as (re, im) pairs (12). */
void oracle_background(double mass, double spin, double length, double radius,
                       double cos_theta, double sin_theta, double* out);
/* Ordered-pair source with Jet1 tangents.  bg: mu0, tau0, pi0 (re, im);
 * fields: 19 jets (value re, im, tangent re, im) in OrderedPairFields
 * order; derivs: 11 jets in OrderedPairDerivatives order; out: D then T
 * jets (8 doubles). */
void oracle_pair_source(double radius, const double* bg, const double* fields,
                        const double* derivs, double* out);
/* eth_n (prime = 0) or eth'_n (prime = 1) of one theta row of a (spin,
 * boost, m, ell_max) band on ntheta nodes at radius R: x, dt, out are
 * complex rows of length ntheta.  Returns 0 or -1. */
int oracle_ghp_row(int spin, int boost, int m, int ell_max, int ntheta, int prime,
                   double kerr_spin, double length, double radius, const double* x,
                   const double* dt, double* out);
/* Point forms used by the pipeline (complex values as (re, im) pairs).
 * eth_n (prime = 0) / eth'_n (prime = 1) of X at one point from X, its time
 * derivative and its angular part `ang` (raised / lowered synthesis). */
void oracle_eth_point(int prime, int spin, int boost, double kerr_spin, double length,
                      double radius, double cos_theta, double sin_theta, const double* x,
                      const double* dt, const double* ang, double* out);
/* R_X of Delta_n X = R_X for G, Lambda, H, B, Pi, C, U (out: 7 values);
 * bg: mu0, tau0, pi0; in: F, G, Lambda, H, B, Pi, C, U, Pi#, B#, C#,
 * eth_1 F, eth_2 G, eth_2 C, eth_2 Pi, eth'_1 B#, eth'_2 C# (17 values). */
void oracle_recon_rhs(double radius, const double* bg, const double* in, double* out);
/* S / R^3 from in: Delta_3 D, D, eth'_3 T, T; bg: mu0, tau0, pi0. */
void oracle_outer_source_point(double radius, const double* bg, const double* in, double* out);
/* 2 (L^4 + a^2 R^2 cos^2) * source. */
void oracle_coordinate_forcing(double length, double kerr_spin, double radius, double cos_theta,
                               const double* source, double* out);
/* Constraint terms (left, right) x (psi3 Bianchi, psi2 Bianchi, h_ll
 * reality); bg: the six background values of oracle_background; in: F, G,
 * H, Lambda, Pi, B, C, U, U#, thorn_1 F, thorn_2 G, eth'_2 G, eth'_3 H. */
void oracle_constraint_terms(double radius, const double* bg, const double* in, double* out);
/* thorn_n at one point (ghp.hpp:24-58). */
void oracle_thorn(const double* value, const double* dt, const double* dr, int n, int spin,
                  int boost, int m, double radius, double cos_theta, double mass,
                  double kerr_spin, double length, const double* epsilon0, double* out);
/* (spin, boost, falloff) of F, G, H, Lambda, Pi, B, C, U, D, T (30 ints):
 * the weights the pipeline's coordinators, Delta_n, eth_n and thorn_n use. */
void oracle_field_weights(int* out);
/* Delta_n and its inverse for the time derivative (complex in/out). */
void oracle_delta_n(const double* value, const double* dt, const double* dr, int n,
                    double radius, double mass, double length, double* out);
void oracle_recon_dt(const double* value, const double* dr, const double* rhs, int n,
                     double radius, double mass, double length, double* out);

#ifdef __cplusplus
}
#endif
#endif
