/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A from-scratch FP64 C++ restatement of the TrixiAtmo.jl / Trixi.jl DGSEM
 * right-hand side that the HIP library accelerates.  Only the test suite,
 * the smoke check and the CPU-baseline leg of bench.py load this library;
 * the product library never links or calls it.
 *
 * Every function names the reference file:line it restates.  Functions whose
 * per-call behaviour cannot be pinned by a reference fixture are pinned by the
 * closed forms and invariants listed in the tests (free-stream preservation,
 * conservation, entropy conservation, SBP identities, exactness on
 * polynomials, 4th-order convergence of the Runge-Kutta scheme).
 *
 * Layouts are Julia column-major (first index fastest), as Trixi's
 * `wrap_array` produces:
 *   u, du           [v, i, j, k, e]         v fastest
 *   contravariant   [c, d, i, j, k, e]      c = Cartesian component, d = direction
 *   inverse_jac     [i, j, k, e]
 *   node_coords     [c, i, j, k, e]
 * Interface / boundary arrays follow Trixi's P4estInterfaceContainer and
 * P4estBoundaryContainer with 0-based element ids and node-index codes
 *   0 = :begin, 1 = :end, 2 = :i_forward, 3 = :i_backward,
 *   4 = :j_forward, 5 = :j_backward.
 */
#ifndef ATMO_ORACLE_H
#define ATMO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- basis (Trixi LobattoLegendreBasis, restated from Kopriva 2009) ---- */
void ora_lgl_nodes_weights(int n, double* nodes, double* weights);
void ora_derivative_matrix(int n, const double* nodes, double* D); /* D[i*n+j] */
void ora_derivative_split(int n, const double* nodes, const double* weights,
                          double* Dsplit);

/* ---- mean values (Ranocha, Schlottke-Lakemper, Chan, Rueda-Ramirez,
 *      Winters, Hindenlang, Gassner 2023, ACM TOMS; imported by
 *      src/TrixiAtmo.jl:37) ---- */
double ora_ln_mean(double x, double y);
double ora_inv_ln_mean(double x, double y);

/* ---- dry Euler, energy form, with gravity (6 variables) ---- */
enum {
  ORA_VOL_RANOCHA = 0,
  ORA_VOL_KENNEDY_GRUBER = 1,
  ORA_VOL_SHIMA = 2,
  ORA_VOL_CENTRAL = 3
};
enum { ORA_NC_ZERO = 0, ORA_NC_WARUSZEWSKI = 1, ORA_NC_SOUZA = 2 };
enum {
  ORA_SURF_LAX_FRIEDRICHS = 0,
  ORA_SURF_RANOCHA = 1,
  ORA_SURF_KENNEDY_GRUBER = 2,
  ORA_SURF_LMARS = 3,
  ORA_SURF_KG_PLUS_LLF = 4,      /* FluxPlusDissipation(flux_kennedy_gruber, DissipationLocalLaxFriedrichs()) */
  ORA_SURF_RANOCHA_PLUS_LLF = 5  /* FluxPlusDissipation(flux_ranocha, DissipationLocalLaxFriedrichs()) */
};

typedef struct {
  int n;          /* nodes per direction (polydeg + 1) */
  int n_elements;
  double c_p, c_v, gravity;
  int volume_flux, volume_noncons, surface_flux, surface_noncons;
  double lmars_c; /* speed of sound estimate for FluxLMARS */
  double coriolis_rate; /* > 0: source_terms_coriolis with this rotation rate */
} ora_e3_params;

void ora_e3_flux(const ora_e3_params* p, const double* u, const double* normal,
                 double* f);
void ora_e3_volume_flux(const ora_e3_params* p, int which, const double* ull,
                        const double* urr, const double* normal, double* f);
void ora_e3_noncons_flux(const ora_e3_params* p, int which, const double* ull,
                         const double* urr, const double* normal, double* f);
void ora_e3_surface_flux(const ora_e3_params* p, int which, const double* ull,
                         const double* urr, const double* normal, double* f);
double ora_e3_max_abs_speed(const ora_e3_params* p, const double* ull,
                            const double* urr, const double* normal);
void ora_e3_cons2entropy(const ora_e3_params* p, const double* u, double* w);
double ora_e3_entropy(const ora_e3_params* p, const double* u);

typedef struct {
  int n_interfaces;
  const int* neighbor_ids;  /* [2, I] */
  const int* node_indices;  /* [3, 2, I] codes */
  int n_boundaries;
  const int* bnd_element;   /* [B] */
  const int* bnd_node_indices; /* [3, B] */
  const int* bnd_kind;      /* [B], 0 = slip wall */
} ora_connectivity;

/* full rhs!: du = -(1/J) (volume + surface) */
int ora_e3_rhs(const ora_e3_params* p, const ora_connectivity* c,
               const double* contravariant, const double* inverse_jacobian,
               const double* u, double* du);

/* sum over all contributions of |contribution| for every entry of du: the
 * scale that bounds the rounding error of the tendency */
int ora_e3_rhs_magnitude(const ora_e3_params* p, const ora_connectivity* c,
                         const double* contravariant, const double* inverse_jacobian,
                         const double* u, double* du_abs);

/* CarpenterKennedy2N54 steps with fixed dt (OrdinaryDiffEq 2N scheme) */
int ora_e3_ck2n54(const ora_e3_params* p, const ora_connectivity* c,
                  const double* contravariant, const double* inverse_jacobian,
                  double* u, double dt, int n_steps);

/* StepsizeCallback max_dt for P4estMesh{3} (without the CFL factor) */
/* operation counts of one rhs! (the same code run with an operation-counting
 * scalar, single-threaded; du receives the same values as ora_e3_rhs):
 * counts[4] = additions and subtractions, multiplications, divisions,
 * sqrt/log/exp/pow */
int ora_e3_count_flops(const ora_e3_params* p, const ora_connectivity* c, const double* cv,
                       const double* invJ, const double* u, double* du, long long* counts);
double ora_e3_max_dt(const ora_e3_params* p, const double* contravariant,
                     const double* inverse_jacobian, const double* u);

/* integrals: out[0] = mass, out[1] = entropy, out[2] = dS/dt = int w(u).du,
 * out[3..8] = integral of every conserved variable */
void ora_e3_integrals(const ora_e3_params* p, const double* inverse_jacobian,
                      const double* u, const double* du, double* out);

/* CarpenterKennedy2N54 coefficients A[5], B[5], c[5] */
void ora_ck2n54_coefficients(double* A, double* B, double* c);

/* ---- covariant shallow water (SplitCovariantShallowWaterEquations2D) on
 *      2D P4est meshes of the sphere; state (h, h v^1, h v^2), 26 auxiliary
 *      variables per node, node coordinates [3, n, n, E] ---- */
typedef struct {
  int n, n_elements;
  double gravity, rotation_rate;
  int volume_flux;     /* 0 flux_ec, 1 flux_central, 2 VolumeIntegralWeakForm() of
                          CovariantShallowWaterEquations2D (the standard form) */
  int surface_flux;    /* 0 FluxPlusDissipation(flux_ec, DissipationLocalLaxFriedrichs()),
                          1 flux_lax_friedrichs, 2 flux_ec, 3 flux_lax_friedrichs of the
                          standard form */
  int surface_noncons; /* 0 flux_nonconservative_surface_simplified, 1 flux_nonconservative_ec
                          (ignored by the standard form) */
  int source_terms;    /* 1: source_terms_geometric_coriolis, 2: the same of the standard form */
} ora_swe_params;

typedef struct {
  int n_interfaces;
  const int* neighbor_ids; /* [2, I] */
  const int* node_indices; /* [2, 2, I] 2D codes: 0 begin, 1 end, 2 i_forward, 3 i_backward */
} ora_connectivity2d;

int ora_swe_rhs(const ora_swe_params* p, const ora_connectivity2d* c, const double* aux,
                const double* xyz, const double* u, double* du);
/* sum of |contribution| for every entry of du (rounding-error scale) */
int ora_swe_rhs_magnitude(const ora_swe_params* p, const ora_connectivity2d* c, const double* aux,
                          const double* xyz, const double* u, double* du_abs);
double ora_swe_max_dt(const ora_swe_params* p, const double* aux, const double* u);
int ora_swe_ck2n54(const ora_swe_params* p, const ora_connectivity2d* c, const double* aux,
                   const double* xyz, double* u, double dt, int n_steps);
void ora_swe_integrals(const ora_swe_params* p, const double* aux, const double* u,
                       const double* du, double* out);
/* cons2prim_and_vorticity of SaveSolutionCallback: out[6][nodes] = (H, global
 * velocity v1..v3, h_s, relative vorticity), variable-major */
void ora_swe_prim_vorticity(const ora_swe_params* p, const double* aux, const double* u,
                            double* out);

/* ---- compressible rainy Euler (CompressibleRainyEulerEquations2D) on a
 *      uniform Cartesian StructuredMesh{2}; state [9, n, n, E] = (rho_dry,
 *      rho_moist, rho_rain, rho_v1, rho_v2, energy, rho_vapour, rho_cloud, T),
 *      elements numbered x fastest ---- */
typedef struct {
  int n, nx, ny;
  double x0, y0, x1, y1;
  int periodic_y;      /* 0: boundary_condition_simple_slip_wall at y = y0, y1 */
  int volume_integral; /* 0 VolumeIntegralWeakForm, 1 flux differencing with flux_ec_rain */
  int source;          /* 0 none, 1 source_terms_rainy, 2 convergence-test source,
                          3 source_terms_no_phase_change */
  int surface_flux;    /* 0 flux_lax_friedrichs, 1 flux_ec_rain */
} ora_rainy_params;

void ora_rainy_saturation_vapour_pressure(double T, double* svp, double* dsvp);
void ora_rainy_residual(const double* u, const double* guess, double* r);
void ora_rainy_residual_jacobian(const double* u, const double* guess, double* J);
int ora_rainy_newton_node(double* u, const double* guess, double tol, int max_iter);
void ora_rainy_flux(const double* u, const double* normal, double* f);
void ora_rainy_flux_ec(const double* ul, const double* ur, const double* normal, double* f);
void ora_rainy_flux_llf(const double* ul, const double* ur, const double* normal, double* f);
double ora_rainy_max_abs_speed_naive(const double* ul, const double* ur, const double* normal);
void ora_rainy_source(const double* u, double* s);
void ora_rainy_node_coordinates(const ora_rainy_params* p, double* xy);
int ora_rainy_rhs(const ora_rainy_params* p, const double* u, double t, double* du);
double ora_rainy_max_dt(const ora_rainy_params* p, const double* u);
int ora_rainy_newton(const ora_rainy_params* p, double* u, double tol, int max_iter,
                     int* iterations);
int ora_rainy_ck2n54(const ora_rainy_params* p, double* u, double t, double dt, int n_steps,
                     double tol, int max_iter);
void ora_rainy_cons2entropy(const double* u, double* w);
void ora_rainy_integrals(const ora_rainy_params* p, const double* u, const double* du, double* out);

/* ---- the MLP seed of the saturation solve (3 -> 32 -> 32 -> 2, tanh):
 *      FP64 loss and gradient for FP32 parameters and samples ---- */
void ora_mlp_loss_grad(const float* params, const float* norm, const float* x, const float* y,
                       long n, double* loss, double* grad);
void ora_mlp_forward(const float* params, const float* norm, const float* x, long n, double* y);

#ifdef __cplusplus
}
#endif
#endif
