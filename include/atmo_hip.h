/*
 * atmo_hip.h -- C ABI of libatmo_hip.so, an MI355X (gfx950) implementation of
 * the DGSEM semidiscretization `rhs!` used by TrixiAtmo.jl on P4est meshes.
 *
 * All entry points are `extern "C"`, take plain pointers and sizes, and return
 * an int status (ATMO_OK = 0, negative on error; `atmo_last_error` gives the
 * message).  Arrays use Julia's column-major layout (first index fastest), so
 * a Julia caller passes `pointer(u_ode)` etc. without copies or transposes.
 * One handle owns its device buffers and one HIP stream; calls on one handle
 * must be serialised by the caller (as OrdinaryDiffEq does for `rhs!`).
 * Device-pointer entry points (`*_device`) take pointers into device memory
 * that the caller owns.
 */
#ifndef ATMO_HIP_H
#define ATMO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define ATMO_OK 0
#define ATMO_ERR_INVALID (-1)     /* bad argument / unsupported combination */
#define ATMO_ERR_DEVICE (-2)      /* HIP runtime error */
#define ATMO_ERR_NONFINITE (-3)   /* non-finite or non-physical state found */
#define ATMO_ERR_NEWTON (-4)      /* saturation Newton: too many failures */
#define ATMO_ERR_MESH (-5)        /* mesh construction failed */

/* -------------------------------------------------------- enumerations */
/* equation systems */
#define ATMO_EQ_EULER_ENERGY_GRAVITY_3D 1 /* CompressibleEulerEnergyEquationsWithGravity3D */
#define ATMO_EQ_COVARIANT_SWE_2D 2        /* SplitCovariantShallowWaterEquations2D */
#define ATMO_EQ_RAINY_EULER_2D 3          /* CompressibleRainyEulerEquations2D */

/* volume (symmetric two-point) fluxes */
#define ATMO_VOL_RANOCHA 0        /* flux_ranocha */
#define ATMO_VOL_KENNEDY_GRUBER 1 /* flux_kennedy_gruber */
#define ATMO_VOL_SHIMA 2          /* flux_shima_etal */
#define ATMO_VOL_CENTRAL 3        /* flux_central */

/* nonconservative gravity fluxes */
#define ATMO_NC_ZERO 0            /* flux_zero */
#define ATMO_NC_WARUSZEWSKI 1     /* flux_nonconservative_waruszewski_etal */
#define ATMO_NC_SOUZA 2           /* flux_nonconservative_souza_etal */

/* surface fluxes */
#define ATMO_SURF_LAX_FRIEDRICHS 0 /* flux_lax_friedrichs (max_abs_speed) */
#define ATMO_SURF_RANOCHA 1        /* flux_ranocha (entropy conservative) */
#define ATMO_SURF_KENNEDY_GRUBER 2 /* flux_kennedy_gruber */
#define ATMO_SURF_LMARS 3          /* FluxLMARS(c) */
#define ATMO_SURF_KG_PLUS_LLF 4    /* FluxPlusDissipation(flux_kennedy_gruber, DissipationLocalLaxFriedrichs()) */
#define ATMO_SURF_RANOCHA_PLUS_LLF 5 /* FluxPlusDissipation(flux_ranocha, DissipationLocalLaxFriedrichs()) */

/* arithmetic of the two-point flux evaluation and volume sums */
#define ATMO_PREC_FP64 0          /* everything in FP64 */
#define ATMO_PREC_FP32_KAHAN 1    /* FP32 fluxes, Kahan-compensated FP32 sums */
#define ATMO_PREC_FP32 2          /* FP32 fluxes, plain FP32 sums */
#define ATMO_PREC_FP32_GEOMETRY 3 /* metric terms (3D: contravariant vectors;
                                     covariant form: metric terms and face maps)
                                     stored in FP32, all arithmetic FP64 */

/* node-local source terms (added after the Jacobian, as calc_sources!) */
#define ATMO_SRC_NONE 0
#define ATMO_SRC_CORIOLIS 1       /* source_terms_coriolis: -2 Omega e_z x rho v */

/* boundary conditions (per boundary face) */
#define ATMO_BC_SLIP_WALL 0       /* boundary_condition_slip_wall */

/* node-index codes of Trixi's P4estInterfaceContainer.node_indices */
#define ATMO_IDX_BEGIN 0
#define ATMO_IDX_END 1
#define ATMO_IDX_I_FORWARD 2
#define ATMO_IDX_I_BACKWARD 3
#define ATMO_IDX_J_FORWARD 4
#define ATMO_IDX_J_BACKWARD 5

/* ------------------------------------------------------------ basis */
/* LobattoLegendreBasis(polydeg): nodes[n], weights[n], D[n*n] and
 * Dsplit[n*n] (row-major, D[i*n+j] = d l_j/dx (x_i)); any pointer may be NULL.
 * Replaces Trixi's LobattoLegendreBasis as constructed by DGSEM(polydeg=...)
 * (examples/euler/dry_air/global_circulation/elixir_energy_baroclinic_instability.jl:27). */
int atmo_lgl_basis(int n, double* nodes, double* weights, double* D, double* Dsplit);

/* ------------------------------------------------------------ meshes */
typedef struct atmo_mesh3d atmo_mesh3d;

/* P4estMesh(trees_per_dimension; polydeg, coordinates_min, coordinates_max,
 * periodicity) with initial_refinement_level = 0 (one element per tree). */
int atmo_mesh3d_cartesian(int polydeg, const int32_t trees_per_dimension[3],
                          const double coordinates_min[3],
                          const double coordinates_max[3],
                          const int32_t periodicity[3], atmo_mesh3d** mesh);

/* P4estMeshCubedSphereTopography(trees_per_face_dimension, layers,
 * inner_radius, thickness; polydeg) with the perfect-sphere topography
 * (src/meshes/p4est_cubed_sphere_topography.jl:114-163). */
int atmo_mesh3d_cubed_sphere_shell(int polydeg, int trees_per_face_dimension,
                                   int layers, double inner_radius,
                                   double thickness, atmo_mesh3d** mesh);

/* sizes of a built mesh */
int atmo_mesh3d_sizes(const atmo_mesh3d* mesh, int32_t* n_elements,
                      int32_t* n_interfaces, int32_t* n_boundaries);

/* Borrowed pointers into the mesh (valid until atmo_mesh3d_destroy):
 *   node_coordinates [3,n,n,n,E], contravariant_vectors [3,3,n,n,n,E],
 *   inverse_jacobian [n,n,n,E], interface_neighbor_ids [2,I] (0-based),
 *   interface_node_indices [3,2,I], boundary_element [B],
 *   boundary_node_indices [3,B], boundary_kind [B] (0 = :inside, 1 = :outside,
 *   2 = other).  These are the arrays of Trixi's P4estElementContainer,
 *   P4estInterfaceContainer and P4estBoundaryContainer. */
int atmo_mesh3d_arrays(const atmo_mesh3d* mesh, const double** node_coordinates,
                       const double** contravariant_vectors,
                       const double** inverse_jacobian,
                       const int32_t** interface_neighbor_ids,
                       const int32_t** interface_node_indices,
                       const int32_t** boundary_element,
                       const int32_t** boundary_node_indices,
                       const int32_t** boundary_kind);
void atmo_mesh3d_destroy(atmo_mesh3d* mesh);

/* The face connectivity of a P4estMesh{3} built on the device from its node
 * coordinates alone (p4est's connectivity and Trixi's
 * init_interfaces! / init_boundaries! [ext];
 * src/meshes/p4est_cubed_sphere_topography.jl:271-594 for the shell): faces
 * matched by their quantised centres (radix sort), the primary side the
 * element face that sorts first, the secondary side's node-index codes the
 * orientation under which its nodes coincide; lists in the order of
 * atmo_mesh3d_arrays, equal to them entry for entry.  period[3] gives the
 * period of each axis (0 or NULL: not periodic).  node_coordinates
 * [3,n,n,n,E] and the outputs are device pointers sized for the maxima
 * (interfaces <= 3E: neighbor ids [2,3E], node indices [3,2,3E]; boundaries
 * <= 6E); the counts come back in host memory.  ATMO_ERR_MESH when a face
 * matches more than one other or no orientation fits.  Complete on return. */
int atmo_mesh3d_connectivity_device(int polydeg, int32_t n_elements,
                                    const double* node_coordinates, const double* period,
                                    int32_t* interface_neighbor_ids,
                                    int32_t* interface_node_indices, int32_t* boundary_element,
                                    int32_t* boundary_node_indices, int32_t* boundary_kind,
                                    int32_t* n_interfaces, int32_t* n_boundaries);

/* 2D cubed sphere for the covariant form */
typedef struct atmo_mesh2d atmo_mesh2d;

/* P4estMeshCubedSphere2D(trees_per_face_dimension, radius; polydeg,
 * element_local_mapping = true) with the 26 auxiliary variables per node of
 * init_auxiliary_node_variables! (GlobalCartesianCoordinates; Christoffel
 * symbols from the exact metric derivatives, or from the collocation
 * derivative when `christoffel_collocation` is nonzero); bottom topography
 * (aux row 20) is zero and may be set by the caller
 * (src/meshes/p4est_cubed_sphere.jl:47-79,
 * src/solvers/dgsem_p4est/containers_2d_manifold_in_3d_covariant.jl:232-489). */
int atmo_mesh2d_cubed_sphere(int polydeg, int trees_per_face_dimension, double radius,
                             int32_t christoffel_collocation, atmo_mesh2d** mesh);
int atmo_mesh2d_sizes(const atmo_mesh2d* mesh, int32_t* n_elements, int32_t* n_interfaces);
/* node_coordinates [3,n,n,E], aux_node_vars [26,n,n,E],
 * interface_neighbor_ids [2,I], interface_node_indices [2,2,I] (2D codes
 * begin/end/i_forward/i_backward) */
int atmo_mesh2d_arrays(const atmo_mesh2d* mesh, const double** node_coordinates,
                       const double** aux_node_vars, const int32_t** interface_neighbor_ids,
                       const int32_t** interface_node_indices);
void atmo_mesh2d_destroy(atmo_mesh2d* mesh);

/* The edge connectivity of P4estMeshCubedSphere2D built on the device from
 * its node coordinates [3,n,n,E] (src/meshes/p4est_cubed_sphere.jl:82-299
 * with p4est's connectivity): equal entry for entry to atmo_mesh2d_arrays'
 * interface lists (interfaces = 2E on the closed sphere: neighbor ids
 * [2,2E], node indices [2,2,2E], device pointers); ATMO_ERR_MESH when an
 * edge has no partner. Complete on return. */
int atmo_mesh2d_connectivity_device(int polydeg, int32_t n_elements,
                                    const double* node_coordinates,
                                    int32_t* interface_neighbor_ids,
                                    int32_t* interface_node_indices, int32_t* n_interfaces);

/* ---------------------------------------------- semidiscretization */
typedef struct atmo_handle atmo_handle;

typedef struct {
  int32_t equations;        /* ATMO_EQ_* */
  int32_t polydeg;          /* 1..7 */
  int32_t n_elements;
  int32_t volume_flux;      /* ATMO_VOL_* */
  int32_t volume_noncons;   /* ATMO_NC_* */
  int32_t surface_flux;     /* ATMO_SURF_* */
  int32_t surface_noncons;  /* ATMO_NC_* */
  int32_t precision;        /* ATMO_PREC_* */
  double c_p, c_v, gravity; /* CompressibleEulerEnergyEquationsWithGravity3D(; c_p, c_v, gravity) */
  double lmars_c;           /* FluxLMARS speed of sound */
  int32_t source_terms;     /* ATMO_SRC_* */
  double rotation_rate;     /* Omega of source_terms_coriolis */
  /* Mixed-precision paths only: hydrostatic background split off in FP64,
   * p_bg = background_p0 exp(-phi / (R background_T0)); <= 0 disables. */
  double background_p0, background_T0;
} atmo_config;

typedef struct {
  /* metric terms, or NULL (both) to compute them on the device from
   * node_coordinates */
  const double* contravariant_vectors;   /* [3,3,n,n,n,E] */
  const double* inverse_jacobian;        /* [n,n,n,E] */
  int32_t n_interfaces;
  const int32_t* interface_neighbor_ids; /* [2,I] 0-based element ids */
  const int32_t* interface_node_indices; /* [3,2,I] ATMO_IDX_* codes */
  int32_t n_boundaries;
  const int32_t* boundary_element;       /* [B] */
  const int32_t* boundary_node_indices;  /* [3,B] */
  const int32_t* boundary_condition;     /* [B] ATMO_BC_* */
  const double* node_coordinates;        /* [3,n,n,n,E]; read when the metric terms are NULL
                                          * or n_interfaces < 0 */
  /* n_interfaces < 0: the interface and boundary lists are built on the
   * device from node_coordinates (atmo_mesh3d_connectivity_device, the same
   * lists as atmo_mesh3d_arrays) and the lists above are not read; each
   * boundary takes the condition of its kind (0 = :inside, 1 = :outside,
   * 2 = other faces) */
  const double* period;                  /* [3] period per axis, 0 = not periodic (NULL: none) */
  int32_t boundary_condition_by_kind[3]; /* ATMO_BC_* */
} atmo_geometry3d;

/* Metric terms of P4estMesh{3} elements from their node coordinates
 * [3,n,n,n,E] (Trixi init_elements! with MetricTermsInvariantCurl,
 * src/semidiscretization/metric_terms.jl:25 [ext]): the Jacobian matrix by
 * the collocation derivative, the contravariant vectors [3,3,n,n,n,E] in
 * Kopriva's invariant curl form and the inverse Jacobian [n,n,n,E].
 * atmo_metric_terms3d takes host arrays; atmo_metric_terms3d_device device
 * pointers (one workgroup per element, complete on return); both give the
 * same bits. */
int atmo_metric_terms3d(int polydeg, int32_t n_elements, const double* node_coordinates,
                        double* contravariant_vectors, double* inverse_jacobian);
int atmo_metric_terms3d_device(int polydeg, int32_t n_elements, const double* node_coordinates,
                               double* contravariant_vectors, double* inverse_jacobian);

/* covariant shallow water: fluxes.  SplitCovariantShallowWaterEquations2D
 * takes ATMO_SWE_VOL_EC / _CENTRAL with ATMO_SWE_SURF_EC_PLUS_LLF /
 * _LAX_FRIEDRICHS / _EC and source_terms 0 or 1; the standard
 * CovariantShallowWaterEquations2D (covariant_shallow_water.jl) takes
 * ATMO_SWE_VOL_WEAK_FORM with ATMO_SWE_SURF_LAX_FRIEDRICHS_STANDARD and
 * source_terms 0 or 2 (surface_noncons is then ignored). */
#define ATMO_SWE_VOL_EC 0                  /* flux_ec */
#define ATMO_SWE_VOL_CENTRAL 1             /* flux_central */
#define ATMO_SWE_VOL_WEAK_FORM 2           /* VolumeIntegralWeakForm() of the standard form */
#define ATMO_SWE_SURF_EC_PLUS_LLF 0        /* FluxPlusDissipation(flux_ec, DissipationLocalLaxFriedrichs()) */
#define ATMO_SWE_SURF_LAX_FRIEDRICHS 1     /* flux_lax_friedrichs */
#define ATMO_SWE_SURF_EC 2                 /* flux_ec */
#define ATMO_SWE_SURF_LAX_FRIEDRICHS_STANDARD 3 /* flux_lax_friedrichs of the standard form */
#define ATMO_SWE_NC_SURFACE_SIMPLIFIED 0   /* flux_nonconservative_surface_simplified */
#define ATMO_SWE_NC_EC 1                   /* flux_nonconservative_ec */

typedef struct {
  int32_t polydeg;          /* 1..7 */
  int32_t n_elements;
  int32_t volume_flux;      /* ATMO_SWE_VOL_*; the volume nonconservative term is flux_nonconservative_ec */
  int32_t surface_flux;     /* ATMO_SWE_SURF_* */
  int32_t surface_noncons;  /* ATMO_SWE_NC_* */
  int32_t source_terms;     /* 1: source_terms_geometric_coriolis, 2: the same of the standard form */
  int32_t precision;        /* ATMO_PREC_FP64 or ATMO_PREC_FP32_GEOMETRY */
  double gravity, rotation_rate; /* (Split)CovariantShallowWaterEquations2D(gravity, rotation_rate) */
} atmo_swe_config;

typedef struct {
  /* [26,n,n,E], or NULL: the auxiliary variables are then computed on the
   * device from node_coordinates (as atmo_covariant_aux_device does, with the
   * two fields at the end of this struct) */
  const double* aux_node_vars;
  const double* node_coordinates;        /* [3,n,n,E] */
  int32_t n_interfaces;
  const int32_t* interface_neighbor_ids; /* [2,I] */
  const int32_t* interface_node_indices; /* [2,2,I] */
  /* ensemble: members own contiguous element ranges
   * [member_first_element[m], member_first_element[m+1]); NULL: one member */
  int32_t n_members;
  const int32_t* member_first_element;   /* [n_members + 1] */
  /* read only when aux_node_vars is NULL */
  int32_t christoffel_collocation;       /* 0: exact metric derivatives, 1: collocation */
  const double* bottom_topography;       /* [n,n,E] host array (aux row 20), or NULL for zero */
  /* n_interfaces < 0: the interface lists are built on the device from
   * node_coordinates (atmo_mesh2d_connectivity_device; one member, or each
   * ensemble member's elements by themselves) and the lists above are not
   * read */
} atmo_geometry2d;

/* init_auxiliary_node_variables! on the device
 * (src/solvers/dgsem_p4est/containers_2d_manifold_in_3d_covariant.jl:232-489,
 * GlobalCartesianCoordinates, element_local_mapping = true): the 26
 * auxiliary variables per node [26,n,n,E] of P4estMeshCubedSphere2D elements
 * from their node coordinates [3,n,n,E] (the corners of each element and the
 * radius of the first node, as the reference reads them), with the
 * Christoffel symbols from the exact metric derivatives (forward-mode dual
 * numbers) or, when christoffel_collocation is nonzero, the collocation
 * derivative; bottom_topography [n,n,E] (may be NULL: zero) fills row 20.
 * Device pointers; complete on return.  Every value equals that of
 * atmo_mesh2d_cubed_sphere bit for bit except the exact Christoffel symbols,
 * which agree to rounding (the host builder differentiates by complex step). */
int atmo_covariant_aux_device(int polydeg, int32_t n_elements, const double* node_coordinates,
                              int32_t christoffel_collocation, const double* bottom_topography,
                              double* aux_node_vars);

/* SemidiscretizationHyperbolic for the covariant form on P4estMesh{2}
 * (src/semidiscretization/semidiscretization_hyperbolic_2d_manifold_in_3d.jl:78-111,
 * cache of src/solvers/dgsem_p4est/dg.jl:5-32); an ensemble of independent
 * members is one handle whose stage launches cover all members.  The kernels
 * recompute G^-1 (and, in FP64, J) from G: aux rows 12 and 16-18 must equal
 * sqrt(det G) and G^-1 of rows 13-15 to 1e-10 relative (as the reference's
 * auxiliary variables do), else ATMO_ERR_INVALID. */
int atmo_swe_create(const atmo_swe_config* config, const atmo_geometry2d* geometry,
                    atmo_handle** handle);

/* Autotuner of the stage kernels (no counterpart in the reference): times
 * every workgroup shape on the resident state (as an RK stage) and keeps the
 * fastest; `ms` (may be NULL, else room for ATMO_SWE_VARIANTS values)
 * receives the time of each shape.  Covariant form: 1, 2 or 4 waves per
 * workgroup, each with a full or a persistent grid (ATMO_SWE_VARIANTS
 * shapes).  3D Euler: one or two elements per workgroup (variants 0 and 1;
 * the second is built for polydeg 3 and 4 of the direct volume form, elsewhere
 * variant 0 is the only shape and its time is reported as 0).  atmo_step and
 * atmo_solve tune by themselves on their first call; every shape gives
 * bit-identical results. */
#define ATMO_SWE_VARIANTS 6
int atmo_autotune(atmo_handle* handle, int32_t* variant, double* ms);
/* pin the workgroup shape of the stage kernel (covariant:
 * 0..ATMO_SWE_VARIANTS-1; 3D: 0 or 1 elements-per-workgroup minus one) */
int atmo_set_kernel_variant(atmo_handle* handle, int32_t variant);

/* ensemble variants: per-member max_dt (dt[n_members]) and RK steps with
 * per-member time steps (two steps at a time replayed from a captured
 * hipGraph) */
int atmo_max_dt_members(atmo_handle* handle, double cfl, double* dt);
int atmo_step_members(atmo_handle* handle, const double* dt, int32_t n_steps);

/* ------------------------------------------------- rainy Euler in 2D */
#define ATMO_RAINY_VOL_WEAK_FORM 0        /* VolumeIntegralWeakForm() */
#define ATMO_RAINY_VOL_FLUX_EC_RAIN 1     /* VolumeIntegralFluxDifferencing(flux_ec_rain) */
#define ATMO_RAINY_SURF_LAX_FRIEDRICHS 0  /* flux_lax_friedrichs (max_abs_speed_naive) */
#define ATMO_RAINY_SURF_FLUX_EC_RAIN 1    /* flux_ec_rain */
#define ATMO_RAINY_SRC_NONE 0
#define ATMO_RAINY_SRC_RAINY 1            /* source_terms_rainy */
#define ATMO_RAINY_SRC_CONVERGENCE_TEST 2 /* source_terms_convergence_test_rainy */
#define ATMO_RAINY_SRC_NO_PHASE_CHANGE 3  /* source_terms_no_phase_change */

typedef struct {
  int32_t polydeg;                /* 1..7 */
  int32_t cells_x, cells_y;       /* StructuredMesh cells_per_dimension */
  double x_min, y_min, x_max, y_max;
  int32_t periodic_y;             /* 0: boundary_condition_simple_slip_wall at y_min, y_max */
  int32_t volume_integral;        /* ATMO_RAINY_VOL_* */
  int32_t surface_flux;           /* ATMO_RAINY_SURF_* */
  int32_t source_terms;           /* ATMO_RAINY_SRC_* */
  double newton_tolerance;        /* NonlinearSolveDG(; tolerance) */
  int32_t newton_max_iterations;  /* NonlinearSolveDG(; max_iterations); 0: no stage limiter */
} atmo_rainy_config;

/* SemidiscretizationHyperbolic(StructuredMesh(cells, min, max; periodicity),
 * CompressibleRainyEulerEquations2D(), initial_condition, DGSEM(...);
 * source_terms, boundary_conditions) on a uniform Cartesian mesh (periodic
 * in x) with NonlinearSolveDG(saturation_residual,
 * saturation_residual_jacobian, SVector(7, 8, 9)) as the stage limiter of
 * atmo_step (src/callbacks_stage/nonlinear_solve_dg.jl:34-45).  State
 * [9, n, n, E], elements numbered x fastest.  atmo_integrals writes the
 * integral of each of the nine variables (out[0..8]). */
int atmo_rainy_create(const atmo_rainy_config* config, atmo_handle** handle);

/* time of the resident state (used by time-dependent sources and as the
 * start of atmo_solve); atmo_step advances it */
int atmo_set_time(atmo_handle* handle, double t);

typedef struct {
  int64_t nodes;
  int64_t updates;         /* Newton updates summed over nodes */
  int64_t max_updates;     /* largest number of updates of one node */
  int64_t failures;        /* nodes that missed the tolerance (last iterate kept) */
  int64_t fallbacks;       /* seeded solves redone from the default guess */
} atmo_newton_stats;

/* the saturation solve of nonlinear_solve_dg_2d! (nonlinear_solve_dg2d.jl:6-45)
 * on the resident state: guess u[7:9] (seeded = 0), the attached MLP's
 * (T, q_v) prediction (seeded = 1) or the closed-form guess (seeded = 2: all
 * moisture as vapour with T from the energy row, and where that is
 * supersaturated one saturation adjustment at that temperature), a seeded
 * solve that fails being redone from u[7:9]; iterations[node] (device
 * pointer, may be NULL) receives the update count (-1: tolerance missed). */
int atmo_saturation_newton(atmo_handle* handle, int32_t seeded, int32_t* iterations_device,
                           atmo_newton_stats* stats);
/* Newton statistics accumulated by atmo_step's stage limiter since the last
 * call (enable with atmo_enable_newton_stats) */
int atmo_enable_newton_stats(atmo_handle* handle, int32_t enable);
int atmo_newton_stats_read(atmo_handle* handle, atmo_newton_stats* stats);

/* Entry points that take device pointers (x_device, iterations_device, ...)
 * run on the library's own stream and read them at once: the caller finishes
 * whatever writes them on other streams first.  They synchronise the
 * library's stream before returning, so their results are complete. */

/* features of the resident state for the saturation seed (device pointers):
 * x[3 * node] = (rho, internal energy density, rho_moist + rho_rain),
 * y[2 * node] = (T, rho_vapour / rho) */
int atmo_rainy_features(atmo_handle* handle, float* x_device, float* y_device);

/* -------------------------------- MLP seed of the saturation solve
 * (no counterpart in the reference): 3 -> 32 -> 32 -> 2, tanh, FP32,
 * parameters W1[32][3], b1[32], W2[32][32], b2[32], W3[2][32], b3[2] (1250),
 * inputs and outputs normalised as (v - mean) / scale. */
typedef struct atmo_mlp atmo_mlp;
#define ATMO_MLP_PARAMS 1250
int atmo_mlp_create(uint64_t seed, atmo_mlp** mlp);
void atmo_mlp_destroy(atmo_mlp* mlp);
/* norm = in_mean[3], in_scale[3], out_mean[2], out_scale[2] */
int atmo_mlp_set_normalization(atmo_mlp* mlp, const float* norm);
int atmo_mlp_set_params(atmo_mlp* mlp, const float* params);
int atmo_mlp_get_params(atmo_mlp* mlp, float* params);
/* loss = mean over samples of the squared error of the normalised outputs
 * (summed over the two outputs) and its gradient (host arrays), for samples
 * x[3 n], y[2 n] in device memory */
int atmo_mlp_loss_grad(atmo_mlp* mlp, const float* x_device, const float* y_device, int64_t n,
                       float* loss, float* grad);
/* Adam (lr, beta1 0.9, beta2 0.999, eps 1e-8) over mini-batches of
 * `batch` samples taken in order from shuffled index lists held by the
 * caller: one call runs `n_batches` steps on x[perm[b * batch + i]] */
int atmo_mlp_train(atmo_mlp* mlp, const float* x_device, const float* y_device,
                   const int32_t* perm_device, int64_t n, int32_t batch, int32_t n_batches,
                   float lr, float* last_loss);
/* y[2 n] = network(x[3 n]) in physical units (device pointers) */
int atmo_mlp_forward(atmo_mlp* mlp, const float* x_device, float* y_device, int64_t n);
/* the network that seeds atmo_saturation_newton (NULL detaches) */
int atmo_rainy_attach_mlp(atmo_handle* handle, atmo_mlp* mlp);
/* seed the stage limiter of atmo_step: enable = 1 with the attached
 * network, 2 with the closed-form guess of atmo_saturation_newton, 0 not
 * (with the same fallback to u[7:9] per node) */
int atmo_rainy_seed_stage_limiter(atmo_handle* handle, int32_t enable);

/* SemidiscretizationHyperbolic(mesh, equations, initial_condition, solver;
 * boundary_conditions) for P4estMesh{3}: copies geometry and connectivity to
 * the device; replaces Trixi.create_cache (src/solvers/dgsem_p4est/dg.jl:5-32). */
int atmo_create(const atmo_config* config, const atmo_geometry3d* geometry,
                atmo_handle** handle);
void atmo_destroy(atmo_handle* handle);
const char* atmo_last_error(const atmo_handle* handle); /* NULL handle: global */

/* number of doubles in one state vector (n_vars * n^3 * n_elements) */
int64_t atmo_state_length(const atmo_handle* handle);

/* Trixi.rhs!(du_ode, u_ode, semi, t) with host arrays in Trixi layout
 * [v,i,j,k,e]; replaces rhs_hyperbolic!
 * (src/solvers/dgsem_p4est/dg_2d_manifold_in_3d_covariant.jl:8-71 for the
 * covariant 2D form; the 3D P4est form of Trixi [ext]). */
int atmo_rhs(atmo_handle* handle, double* du, const double* u, double t);
/* same with device pointers (no copies) */
int atmo_rhs_device(atmo_handle* handle, double* du, const double* u, double t);

/* resident state: upload / download the handle's u */
int atmo_set_state(atmo_handle* handle, const double* u_host);
int atmo_get_state(atmo_handle* handle, double* u_host);

/* n_steps CarpenterKennedy2N54 steps of size dt on the resident state, every
 * stage one fused launch (OrdinaryDiffEq low-storage 2N scheme as used by
 * examples/shallow_water/covariant/elixir_barotropic_instability.jl:79). */
int atmo_step(atmo_handle* handle, double dt, int32_t n_steps);

/* StepsizeCallback max_dt on the resident state: dt = cfl * 2 / (n * max
 * scaled speed) (src/callbacks_step/stepsize_dg2d.jl:5-43 for the 2D form). */
int atmo_max_dt(atmo_handle* handle, double cfl, double* dt);

/* solve(ode, CarpenterKennedy2N54(williamson_condition = false); callback =
 * StepsizeCallback(cfl)) from the handle's time (atmo_set_time; atmo_step
 * advances it) to t_end, on the device: before every step
 * dt = min(atmo_max_dt, t_end - t), computed on the device from the resident
 * state, and the five stages read it there
 * (examples/shallow_water/covariant/elixir_barotropic_instability.jl:79-80
 * with the StepsizeCallback of :67).  The same steps and state as a host loop
 * of atmo_max_dt + atmo_step; the host waits for the device only every
 * `check_every` (>= 1) steps.  Ensemble members each take their own step
 * sizes.  steps (may be NULL) receives the steps taken per member
 * (the config's n_members entries, 1 for other handles), t_reached (may be NULL) the
 * time reached (the smallest over members), which becomes the handle's time.
 * ATMO_ERR_NONFINITE when a wave speed is not finite (the state is left at
 * the last finite step). */
int atmo_solve(atmo_handle* handle, double t_end, double cfl, int32_t check_every,
               int64_t* steps, double* t_reached);

/* Allocation-free max_dt for the 3D Euler equations.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight time loops, harnesses that forbid hipMalloc).  It
 * computes what atmo_max_dt computes on a handle -- dt = cfl * 2 / (n * max
 * over elements of the summed maximum scaled speeds) -- for a state u
 * [6,n,n,n,E], contravariant vectors [3,3,n,n,n,E] and inverse Jacobian
 * [n,n,n,E] in device memory owned by the caller (config: equations,
 * polydeg, n_elements, c_p, c_v; the other fields are not read), and writes
 * dt to device memory (*dt_device; NaN when a wave speed is not finite,
 * where atmo_max_dt returns ATMO_ERR_NONFINITE).  Its kernels keep their
 * per-element partials in a fixed scratch array of the library (one per GPU,
 * held until the process ends), so n_elements is at most
 * ATMO_NOALLOC_MAX_ELEMENTS (ATMO_ERR_INVALID beyond).  It calls no hipMalloc
 * or hipFree, enqueues its kernels on `stream` (a hipStream_t; NULL: the
 * null stream) and returns without waiting for them.  Calls on one GPU share
 * that GPU's scratch array: calls that overlap on one GPU race, so the
 * caller serialises them. */
#define ATMO_NOALLOC_MAX_ELEMENTS 262144
int atmo_max_dt3d_noalloc(const atmo_config* config, const double* u_device,
                          const double* contravariant_vectors_device,
                          const double* inverse_jacobian_device, double cfl,
                          double* dt_device, void* stream);

/* Integrals on the resident state (analysis_covariant.jl:53-105 analogues):
 * out[0] mass, out[1] entropy, out[2] dS/dt = int w(u).du of the last
 * computed rhs (or of a fresh one when `fresh_rhs`), out[3..] integral of each
 * conserved variable (3D Euler: out[0..8]; covariant shallow water: out[0..2]). */
int atmo_integrals(atmo_handle* handle, int32_t fresh_rhs, double* out);

/* SaveSolutionCallback's solution_variables of the resident state, computed
 * on the device: kind 0 cons2cons, kind 1 cons2prim (3D Euler,
 * compressible_euler_energy_with_gravity_3d.jl:568-578; rainy Euler,
 * compressible_rainy_euler_2d.jl:114-129) or cons2prim_and_vorticity
 * (covariant shallow water, save_solution_covariant.jl:65-117: H, the global
 * velocity, h_s, the relative vorticity).  out (host) receives
 * [n_vars][nodes], each variable contiguous in Trixi's node order (the
 * variables_<v> datasets of the solution file); n_vars (may be NULL) the
 * number of variables: 6 (3D), 3 or 6 (covariant), 9 (rainy). */
int atmo_solution_variables(atmo_handle* handle, int32_t kind, double* out, int32_t* n_vars);

/* wait for all work of the handle's stream */
int atmo_synchronize(atmo_handle* handle);

/* Kernel timing: when enabled, every launch is bracketed by HIP events on the
 * handle's stream; atmo_kernel_times returns per kernel class the number of
 * launches and the total milliseconds (classes: 0 fused stage/rhs kernel,
 * 1 reductions). */
int atmo_enable_timing(atmo_handle* handle, int32_t enable);
int atmo_kernel_times(atmo_handle* handle, int32_t kernel_class,
                      int64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif
