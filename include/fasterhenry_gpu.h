/*
 * fasterhenry_gpu.h -- C ABI of the MI355X (gfx950) extraction hot path.
 *
 * A drop-in for the solve paths of fasterhenry 0.2.0.  The reference is
 * a safe Rust library without an FFI; each entry below names the reference
 * item it replaces (paths relative to the reference repository root).  The
 * ABI is plain C: fixed-width integers, doubles and pointers, caller-owned
 * inputs and outputs, no C++ or HIP types.  A Rust caller binds it with
 * `extern "C"` and `#[repr(C)]` mirrors of the structs.
 *
 * Threading: one fh_system per thread at a time; distinct systems may be
 * used concurrently (each owns its HIP stream).  Results are deterministic:
 * no reduction on the device depends on scheduling order.
 */
#ifndef FASTERHENRY_GPU_H
#define FASTERHENRY_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* DENSE_PATH_MAX_FILAMENTS (fasterhenry/src/solve.rs:133). */
#define FH_DENSE_PATH_MAX_FILAMENTS 10000u

/* Status codes.  They mirror SolveError (fasterhenry/src/solve.rs:567-637),
 * KernelError (inductance/mod.rs:93-115), MeshError (mesh.rs:99-148) and
 * GeometryError (geometry.rs).  fh_last_error() and fh_error_detail() give
 * the message and the indices. */
typedef enum fh_status {
  FH_OK = 0,
  FH_E_INVALID_ARGUMENT = 1, /* null pointer, bad size or option */
  FH_E_GEOMETRY = 2,         /* GeometryError: index0 = node or segment */
  FH_E_DISCRETIZE = 3,       /* SolveError::Discretize: index0 = segment */
  FH_E_KERNEL = 4,           /* SolveError::Kernel: index0, index1 = filament pair */
  FH_E_MESH = 5,             /* SolveError::Mesh: code = which MeshError */
  FH_E_INVALID_GRADING = 6,  /* SolveError::InvalidGrading: index0 = segment */
  FH_E_INVALID_FREQUENCY = 7,/* SolveError::InvalidFrequency: index0, value */
  FH_E_SINGULAR = 8,         /* SolveError::Singular: index0 = frequency index, value = Hz */
  FH_E_NOT_CONVERGED = 9,    /* index0 = frequency index; refinement missed 1e-10
                                (index1 = steps, value = Hz) or GMRES missed its
                                tolerance (index1 = port, value = relative residual,
                                code = iterations) */
  FH_E_UNSUPPORTED = 10,     /* a path this build does not provide */
  FH_E_NO_MEMORY = 11,       /* host or device allocation failed */
  FH_E_HIP = 12,             /* HIP runtime error */
  FH_E_PFFT = 13,            /* PfftError (pfft/mod.rs:236-251): code 1 = InvalidParameter,
                                2 = GridTooLarge (index0 = padded points, index1 = limit) */
  FH_E_INVALID_EXCITATION = 14, /* SolveError::InvalidExcitation: index0 = ports expected,
                                   index1 = entries given */
  FH_E_DECK = 15               /* deck ParseError (fasterhenry-cli/src/inp.rs:682-702):
                                  index0 = line (0: the deck as a whole) */
} fh_status;

/* SolverChoice (solve.rs:169-189) and Solver (solve.rs:143-158). */
typedef enum fh_solver_choice { FH_AUTO = 0, FH_DENSE = 1, FH_ITERATIVE = 2 } fh_solver_choice;
typedef enum fh_solver { FH_SOLVER_DENSE = 0, FH_SOLVER_ITERATIVE = 1 } fh_solver;

/* Arithmetic of the per-frequency factorization.  FH_FP64_PLAIN: the whole
 * solve in FP64.  FH_FP32_REFINED: the factorization in FP32 on the matrix
 * cores, FP64 iterative refinement of the solution to a 1e-10 relative
 * Schur-complement change, FP64 everywhere else. */
typedef enum fh_precision { FH_FP64_PLAIN = 0, FH_FP32_REFINED = 1 } fh_precision;

/* PfftParams (pfft/mod.rs:183-231).  auto_grid = 1: GridSpacing::Auto with
 * cells_per_filament; 0: GridSpacing::Fixed(spacing) in metres.
 * near_field_radius in grid spacings (>= interpolation_order + 1).
 * fft_precision: arithmetic of the far-field FFT convolution, 0 = FP64 (the
 * reference's), 1 = FP32; projection, near field and sums stay FP64. */
typedef struct fh_pfft_params {
  int32_t auto_grid;
  int32_t interpolation_order;
  double cells_per_filament;
  double spacing;
  double near_field_radius;
  int32_t fft_precision;
  int32_t reserved;
} fh_pfft_params;

/* GmresParams (gmres.rs:29-60). */
typedef struct fh_gmres_params {
  double tolerance;
  uint64_t restart;
  uint64_t max_iterations;
} fh_gmres_params;

/* IterativeParams (iterative.rs:98-105). */
typedef struct fh_iterative_params {
  fh_pfft_params pfft;
  fh_gmres_params gmres;
} fh_iterative_params;

/* PfftStats (pfft/mod.rs:255-270). */
typedef struct fh_pfft_stats {
  uint64_t filaments;
  double grid_spacing;
  uint64_t grid_dims[3];
  uint64_t fft_dims[3];
  uint64_t projection_entries;
  uint64_t near_entries;
  uint64_t memory_bytes; /* device memory held by the operator */
  /* the convolution used: 0 = FP64 pruned line FFTs, 1 = FP32 pruned line
   * FFTs, 2 = FP64 3-D FFT by global-memory Stockham passes (a padded axis
   * too long for one line in LDS) */
  int32_t fft_kind;
  /* lines per workgroup the set-up autotuner chose for the axis-1 forward,
   * axis-0 and axis-1 inverse line passes (0 when not tuned) */
  int32_t line_group[3];
} fh_pfft_stats;

/* SolverChoice::resolve (solve.rs:179-189). */
fh_solver fh_resolve(fh_solver_choice choice, size_t n_filaments);

/* Geometry.  SegmentDef (geometry.rs:289-330): nodes a, b; cross-section
 * width x height; conductivity sigma; optional width direction. */
typedef struct fh_segment {
  uint32_t a, b;
  double width, height, sigma;
  double width_dir[3];
  uint32_t has_width_dir;
  uint32_t reserved;
} fh_segment;

/* A segment's resolved grid: Discretization::resolve (solve.rs:468-510)
 * gives one per segment.  graded = 0 cuts with discretize (filament.rs:320),
 * graded = 1 with discretize_graded_per_axis (filament.rs:478). */
typedef struct fh_grid {
  uint32_t nw, nh;
  double width_ratio, height_ratio;
  uint32_t graded;
  uint32_t reserved;
} fh_grid;

/* Port (mesh.rs:68-94). */
typedef struct fh_port {
  uint32_t positive, negative;
} fh_port;

/* A problem: nodes (xyz triples), segments with one grid each, ports. */
typedef struct fh_problem {
  const double *nodes;
  size_t n_nodes;
  const fh_segment *segments;
  const fh_grid *grids;
  size_t n_segments;
  const fh_port *ports;
  size_t n_ports;
} fh_problem;

/* A filament (filament.rs:34-140): centreline, frame, cross-section. */
typedef struct fh_filament {
  double start[3], end[3];
  double length;
  double dir[3], wdir[3], hdir[3];
  double width, height, sigma;
} fh_filament;

/* Counts (result.rs:110-125) plus timing (result.rs:129-140) of a sweep. */
typedef struct fh_sweep_info {
  size_t nodes, segments, filaments, meshes, internal_meshes, ports;
  double inductance_s; /* fill of L */
  double assembly_s;   /* discretization + fill + congruence */
  double solve_s;      /* the sweep, from frequencies in to Z out */
  /* FH_FP32_REFINED: worst refinement steps and final relative Schur change */
  int32_t refine_steps_max;
  /* arithmetic of the trailing updates that produced the result: 0 FP64,
   * 1 FP32 matrix cores, 2 FP32 with split-bf16 matrix-core products */
  int32_t factor_arith;
  double refine_change_max;
  /* instrumentation (fh_system_set_profiling): time and launch count of
   * the dominant kernel (the trailing update over each outer block), HIP
   * events on the system's stream */
  double update_kernel_ms;
  int64_t update_kernel_launches;
  double update_kernel_flops;      /* algorithmic (FP64/FP32-equivalent) flops */
  double update_kernel_unit_flops; /* flops executed on the matrix-core units */
  /* the path that produced the result (fh_solver) */
  int32_t solver;
  /* iterative path: most operator products of one port's GMRES solve, their
   * total, the worst relative residual reached; with profiling, the time in
   * operator products (each on the real and imaginary branch currents) */
  int32_t gmres_iterations_max;
  int64_t gmres_iterations;
  double gmres_residual_max;
  int64_t operator_applies;
  double operator_apply_ms;
  /* refined form: the last refinement stage any frequency needed (0-4, see
   * fh_system_set_refinement); -1 for the plain form and the iterative path */
  int32_t refine_stage;
} fh_sweep_info;

typedef struct fh_system fh_system;

/* MeshSystem::assemble (solve.rs:680-766) or, when the choice resolves to
 * the iterative path, IterativeSystem::assemble (iterative.rs:155-238) with
 * the default parameters (solve.rs:931-938).  Dense: discretize, build the
 * loop basis, fill L on the GPU, reduce to mesh blocks on the GPU; the
 * blocks stay resident in device memory until fh_system_destroy.
 * Iterative: the pFFT operator, port columns and Jacobi diagonals. */
fh_status fh_system_create(const fh_problem *problem, fh_solver_choice choice,
                           fh_precision precision, fh_system **out);

/* IterativeParams::default (iterative.rs:98-105): PfftParams::default
 * (pfft/mod.rs:213-231) and GmresParams::default (gmres.rs:48-60). */
void fh_iterative_default_params(fh_iterative_params *out);

/* IterativeSystem::assemble (iterative.rs:155-238) with explicit parameters.
 * FH_E_PFFT if the operator cannot be built from params->pfft. */
fh_status fh_system_create_iterative(const fh_problem *problem, const fh_iterative_params *params,
                                     fh_system **out);

/* IterativeSystem::operator_stats (iterative.rs:255-258); FH_E_UNSUPPORTED
 * for a dense system. */
fh_status fh_system_operator_stats(const fh_system *system, fh_pfft_stats *out);

/* MeshSystem::sweep (solve.rs:881-894, 1039-1079) or IterativeSystem::sweep
 * (iterative.rs:387-404): z_out receives
 * K x p x p complex values, row-major, each as {re, im}.  f = 0 is solved
 * in real arithmetic with Im Z exactly 0 (solve.rs:809-812).  The first
 * failing frequency in list order is reported.  info may be NULL. */
fh_status fh_sweep(fh_system *system, const double *freqs_hz, size_t K, double *z_out,
                   fh_sweep_info *info);

/* MeshSystem::impedance (solve.rs:806-830) at one frequency. */
fh_status fh_impedance(fh_system *system, double freq_hz, double *z_out);

/* Changes the factorization arithmetic of an assembled dense system (no
 * effect on an iterative one). */
fh_status fh_system_set_precision(fh_system *system, fh_precision precision);

/* Brackets each launch of the dominant kernel with HIP events on the
 * system's stream; totals are reported in fh_sweep_info. */
fh_status fh_system_set_profiling(fh_system *system, int enable);

/* Counts (result.rs:110-125): nodes, segments, filaments, meshes,
 * internal meshes, ports. */
fh_status fh_system_counts(const fh_system *system, size_t counts[6]);

/* Copies of the resident data, for parity checks.  Any pointer may be NULL.
 * filaments: n; L: n x n row-major; mesh blocks row-major: l_ee m x m,
 * l_ep m x p, l_pp p x p, r_ee m x m, r_ep m x p, r_pp p x p. */
fh_status fh_system_filaments(const fh_system *system, fh_filament *out);
/* Scheduling of refined dense sweeps (no counterpart in the reference):
 * a chunk of frequencies is factorized as `parts` (1 to 4) groups on as
 * many streams, side by side, so that the latency-bound kernels of one
 * group overlap the bulk updates of the others.  Default 2.  Results do not
 * depend on it beyond the refinement's rounding (each frequency is refined
 * to the same acceptance bound). */
fh_status fh_system_set_sweep_parts(fh_system *system, int32_t parts);

/* Refinement stages of the FP32_REFINED form (no counterpart in the
 * reference, whose LU fails only on an exactly zero pivot, dense.rs:43-84):
 * each frequency goes through the allowed stages in this order until its Z
 * is accepted; 0 split-bf16 factorization with classical refinement, 1 the
 * same factorization under GMRES-IR (GMRES on the preconditioned correction
 * equation, FP64 residuals), 2 the FP32 factorization with classical
 * refinement, 3 the FP32 factorization under GMRES-IR, 4 a plain FP64
 * L D L^T solve of the frequencies still left (sweeps only; fh_grad and
 * fh_currents need x and stop at stage 3).  A frequency no allowed stage
 * solves is FH_E_NOT_CONVERGED.  Default FH_REFINE_ALL; other masks are for
 * tests and studies.  fh_sweep_info.refine_stage reports the last stage any
 * frequency needed. */
#define FH_REFINE_SPLIT_CLASSICAL 1u
#define FH_REFINE_SPLIT_GMRES 2u
#define FH_REFINE_FP32_CLASSICAL 4u
#define FH_REFINE_FP32_GMRES 8u
#define FH_REFINE_FP64 16u
#define FH_REFINE_ALL 31u
fh_status fh_system_set_refinement(fh_system *system, uint32_t stages);

/* Schedule of a dense system's sweep (no counterpart in the reference): the
 * outer block of the factorization (columns; a multiple of 64 from 64 to
 * 1024; default 128 FP64, 256 refined) and the sweep parts (refined form).
 * Setting one fixes it for the system; otherwise a schedule the tuning cache
 * holds for the system's size class (precision, internal loops and
 * frequencies rounded up to powers of two) is used.  Results do not depend
 * on the schedule beyond the refinement's rounding. */
typedef struct fh_schedule {
  int32_t outer_block;
  int32_t sweep_parts;
} fh_schedule;
fh_status fh_system_get_schedule(const fh_system *system, fh_schedule *out);
fh_status fh_system_set_schedule(fh_system *system, const fh_schedule *schedule);
/* Autotuner: sweeps `freqs` once per candidate schedule (outer blocks 128,
 * 256, 512 x parts 1-3 refined; 64, 128, 256 FP64), keeps the fastest for the
 * system and stores it in the tuning cache; chosen and best_ms (either may
 * be NULL) receive it and its sweep time.  A system with no dense
 * factorization has nothing to tune. */
fh_status fh_system_autotune(fh_system *system, const double *freqs, size_t K, fh_schedule *chosen,
                             double *best_ms);
/* Graph replay of repeated refined sweeps (no counterpart in the reference):
 * from the second sweep of a chunk with the same frequency count, each
 * sweep part's factorization and first two refinement steps, with their
 * convergence test on the device, run as one captured HIP graph; frequencies
 * that need more steps continue as before.  The results are the same bits
 * either way.  Default on; 0 turns it off (and frees the graphs). */
fh_status fh_system_set_graphs(fh_system *system, int32_t enable);
/* The tuning cache to and from a JSON file (version 1). */
fh_status fh_tune_save(const char *path);
fh_status fh_tune_load(const char *path);

/* Dense systems only (FH_E_UNSUPPORTED for an iterative one). */
fh_status fh_system_inductance(const fh_system *system, double *L_out);
fh_status fh_system_blocks(const fh_system *system, double *r_ee, double *l_ee, double *r_ep,
                           double *l_ep, double *r_pp, double *l_pp);

void fh_system_destroy(fh_system *system);

/* solve() (solve.rs:931-938): create + sweep + destroy. */
fh_status fh_solve(const fh_problem *problem, fh_solver_choice choice, fh_precision precision,
                   const double *freqs_hz, size_t K, double *z_out, fh_sweep_info *info);

/* A batch of independent problems swept over the same K frequencies -- the
 * work of one solve() per problem (solve.rs:931-938), scheduled together:
 * systems are assembled and swept concurrently on the device by
 * `concurrency` workers (0: the library's choice), largest problems first,
 * each on its own stream.  Each
 * problem's result is identical to its own
 * fh_solve.  z_out: problem b's K x p_b x p_b {re, im} values start at
 * 2 K sum_{a<b} p_a^2 doubles.  infos (may be NULL): one fh_sweep_info per
 * problem.  On failure the first failing problem in batch order is reported
 * (its status and detail) and *failed (may be NULL) receives its index. */
fh_status fh_solve_batch(const fh_problem *problems, size_t count, fh_solver_choice choice,
                         fh_precision precision, const double *freqs_hz, size_t K, double *z_out,
                         fh_sweep_info *infos, int32_t concurrency, size_t *failed);

/* Geometry parameters for gradients: how the problem moves along each of
 * n_params directions.  d_nodes: n_params x n_nodes x 3 (d node xyz / dp);
 * d_widths, d_heights: n_params x n_segments (d segment width / dp, d height
 * / dp), either may be NULL (no change).  Filament counts and grading
 * ratios are held at the problem's (frozen discretization). */
typedef struct fh_param_map {
  size_t n_params;
  const double *d_nodes;
  const double *d_widths;
  const double *d_heights;
} fh_param_map;

/* PowerBalance (currents.rs:103-111), watts, peak-phasor convention. */
typedef struct fh_power_balance {
  double dissipated_w;      /* sum of 1/2 |I|^2 R over filaments */
  double input_w;           /* 1/2 Re(I_p^H V_p) */
  double relative_residual; /* |dissipated - input| / max(dissipated, input), 0 if both 0 */
} fh_power_balance;

/* MeshSystem::currents / IterativeSystem::currents (solve.rs:835-868,
 * iterative.rs:329-380, solve_currents solve.rs:1110-1149,
 * assemble_solution currents.rs:190-274): for one frequency and the port
 * current phasors `excitation` (n_excitation = p complex {re, im}), the
 * internal loop currents -Zee^-1 Zep I_p mapped to filaments.  Outputs, each
 * may be NULL: filament currents (n complex), |J| = |I| / (w h) (n), power
 * 1/2 |I|^2 R (n); per segment the summed current (complex) and power, and
 * its conductor (connected component, numbered by first segment); the
 * number of conductors and, per conductor, the filament of largest |J|
 * (capacity n_segments; the first of equals); the port voltages
 * Zpp I_p + Zep^T I_e (p complex); the power balance.  The dense form
 * solves with the refined factorization refined to working accuracy; the
 * iterative form with GMRES.  Errors: FH_E_INVALID_FREQUENCY,
 * FH_E_INVALID_EXCITATION (wrong length or a non-finite entry),
 * FH_E_SINGULAR, FH_E_NOT_CONVERGED. */
fh_status fh_currents(fh_system *system, double frequency_hz, const double *excitation,
                      size_t n_excitation, double *filament_current, double *current_density,
                      double *power_w, double *segment_current, double *segment_power_w,
                      int32_t *conductor_of_segment, size_t *n_conductors, int64_t *peak_filament,
                      double *port_voltage, fh_power_balance *balance);

/* Writers of a sweep (K frequencies, z K x p x p {re, im} row-major, as
 * fh_sweep returns it).  fh_zc_mat: write_zc_mat
 * (fasterhenry-cli/src/mat.rs:31-95), the MAT v4 bytes of `freqs` (1 x K)
 * and `Zc_1 .. Zc_K` (complex p x p); *size receives the byte count, out
 * (may be NULL) the bytes.  fh_spice_subckt: write_spice_subckt
 * (fasterhenry-cli/src/spice.rs:64-126) at frequency `index`, the text
 * with its terminating NUL in out (may be NULL), *length its length.
 * fh_subckt_name_from_path: subckt_name_from_path (spice.rs:37-60). */
fh_status fh_zc_mat(const double *freqs, size_t K, const double *z, size_t p, uint8_t *out,
                    size_t capacity, size_t *size);
fh_status fh_spice_subckt(const double *freqs, size_t K, const double *z, size_t p, size_t index,
                          const char *name, char *out, size_t capacity, size_t *length);
fh_status fh_subckt_name_from_path(const char *path, char *out, size_t capacity, size_t *length);

/* Gradients of the port impedance.  No counterpart in the reference (it has
 * no derivatives); the north star's adjoint: for each frequency k, parameter
 * q and port pair (a, b), dR_dp = Re dZ_ab/dp_q and dL_dp = Im dZ_ab/dp_q / w
 * (0 at f = 0), both K x n_params x p x p row-major; z_out (may be NULL)
 * receives Z (K x p x p {re, im}) in the stationary form i_a^T (R + jw L) i_b
 * of the refined currents: fh_sweep's Z to within its tolerance, without the
 * Schur complement's cancellation.  Z is stationary in the loop currents, so
 * dZ_ab = i_a^T (dR + jw dL) i_b with i_a the filament currents of port a's
 * solution: one refined solve per frequency, then the derivative kernels of
 * the partial inductances (methods, quadrature orders and subdivisions held
 * at the nominal geometry) contracted with the currents on the device.
 * Dense systems only. */
fh_status fh_grad(fh_system *system, const fh_param_map *map, const double *freqs_hz, size_t K,
                  double *z_out, double *dR_dp, double *dL_dp);

/* Host level: d(filaments)/dp along one direction (d_nodes n_nodes x 3,
 * d_widths / d_heights n_segments or NULL): out (n filaments) receives
 * d(start), d(end), d(length), d(dir), d(wdir), d(hdir), d(width),
 * d(height) of each filament, sigma 0.  For checking fh_grad's geometry
 * derivatives against finite differences of fh_discretize. */
fh_status fh_discretize_tangent(const fh_problem *problem, const double *d_nodes,
                                const double *d_widths, const double *d_heights, fh_filament *out,
                                size_t capacity);

/* Kernel level, gradients: nominal partial inductances and their
 * derivatives along filament tangents, pair by pair.  No counterpart in the
 * reference.  tangents: n_dirs x n filaments (as fh_discretize_tangent
 * writes them); pairs: count x {f, g, direction}.  value_out[k] is L_fg
 * (the same bits as fh_partial_inductance's entry), deriv_out[k] its
 * derivative along the direction with the pair's method, quadrature orders
 * and aligned-quadrature subdivision held at the nominal geometry. */
fh_status fh_partial_inductance_derivative(const fh_filament *filaments, size_t n,
                                           const fh_filament *tangents, size_t n_dirs,
                                           const uint32_t *pairs, size_t count, double *value_out,
                                           double *deriv_out);

/* PairMask (inductance/batch.rs:37-110): the group of each of n filaments
 * and a groups x groups keep table (row-major; the caller symmetrizes it and
 * sets the diagonal).  A pair is evaluated when either filament has no
 * entry (index >= n) or a negative group, when the table index
 * a * groups + b is past the table, or when the entry is nonzero; otherwise
 * its L entry is exactly 0 and no kernel runs for it. */
typedef struct fh_pair_mask {
  const int32_t *group;
  size_t n;
  size_t groups;
  const uint8_t *coupled;
} fh_pair_mask;

/* A coupling resolved against the segments (Coupling, coupling.rs:173-353):
 * the group index of every segment and the groups x groups keep table.
 * coupled = NULL is Coupling::all_pairs. */
typedef struct fh_coupling {
  const int32_t *segment_group;
  size_t n_segments;
  size_t groups;
  const uint8_t *coupled;
} fh_coupling;

/* MeshSystem::assemble_with_coupling (solve.rs:703-766): as
 * fh_system_create, with the partial inductances of uncoupled groups
 * truncated (never evaluated).  A truncating coupling stays dense under
 * FH_AUTO and returns FH_E_UNSUPPORTED under FH_ITERATIVE
 * (fasterhenry-cli/src/lib.rs:89-107).  The coupling's names, validation and
 * truncation warnings are resolved by the caller (the Python mirror's
 * Coupling). */
fh_status fh_system_create_coupled(const fh_problem *problem, fh_solver_choice choice,
                                   fh_precision precision, const fh_coupling *coupling,
                                   fh_system **out);

/* partial_inductance_matrix_masked (inductance/batch.rs:305-312): mask may
 * be NULL (every pair); methods (may be NULL) holds -1 for truncated pairs. */
fh_status fh_partial_inductance_masked(const fh_filament *filaments, size_t n,
                                       const fh_pair_mask *mask, double *L_out, int8_t *methods);

/* The FastHenry `.inp` deck (fasterhenry-cli/src/inp.rs, parse :3447-3625),
 * conductor subset: .title, .units (mandatory), .default, N nodes, E
 * segments (w h nwinc nhinc rw rh wx wy wz sigma|rho group), .external,
 * .freq, .equiv, .couples, .end; ground planes (G, .hole, .contact) return
 * FH_E_DECK.  fh_deck_problem points into the deck (valid until
 * fh_deck_destroy): the nodes segments or ports reference, one grid per
 * segment, the ports.  fh_deck_info: the sweep (Hz), the discretization
 * (0 Uniform, 1 PerSegment, 2 PerSegmentGraded: discretization_of,
 * inp.rs:4733-4765) and the number of coupling groups.  fh_deck_coupling:
 * the resolved `.couples` truncation for fh_system_create_coupled
 * (coupled = NULL: all pairs).  Names: port labels (an explicit label as
 * written, else "<+>/<->" lowercased), group names, the title (NULL if
 * none). */
typedef struct fh_deck fh_deck;
fh_status fh_deck_parse(const char *text, size_t length, fh_deck **out);
fh_status fh_deck_problem(const fh_deck *deck, fh_problem *out);
fh_status fh_deck_info(const fh_deck *deck, size_t *n_frequencies, const double **frequencies,
                       int32_t *discretization, size_t *n_groups);
fh_status fh_deck_coupling(const fh_deck *deck, fh_coupling *out);
const char *fh_deck_title(const fh_deck *deck);
const char *fh_deck_port_name(const fh_deck *deck, size_t port);
const char *fh_deck_group_name(const fh_deck *deck, size_t group);
void fh_deck_destroy(fh_deck *deck);

/* Kernel level.  partial_inductance_matrix (inductance/batch.rs:246-253):
 * L_out (n x n, row-major, host) for host filaments.  methods (may be NULL)
 * receives the Method of inductance/mod.rs:130-152 per entry (0 orthogonal,
 * 1 bar closed form, 2 aligned quadrature, 3 point quadrature, 4 sampled). */
fh_status fh_partial_inductance(const fh_filament *filaments, size_t n, double *L_out,
                                int8_t *methods);

/* The fill of fh_partial_inductance without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  d_filaments (n records), d_L_out
 * (n x n doubles, row-major) and d_methods_out (n x n bytes, may be NULL)
 * are device pointers the caller owns; stream is a hipStream_t (NULL: the
 * null stream).  The deferred-pair queue lives in a fixed scratch array of
 * each GPU (64 MB, held until the process ends), sized for
 * FH_NOALLOC_MAX_FILAMENTS filaments; a larger n returns
 * FH_E_INVALID_ARGUMENT.  Every call on a GPU shares that array, so calls
 * that overlap on one GPU race: serialize them.  The values and methods are
 * fh_partial_inductance's, bit for bit; the call synchronizes the stream
 * before it returns (to report a kernel error). */
#define FH_NOALLOC_MAX_FILAMENTS 4000
fh_status fh_partial_inductance_noalloc(const fh_filament *d_filaments, size_t n, double *d_L_out,
                                        int8_t *d_methods_out, void *stream);

/* discretize_geometry (solve.rs:947-975) on the host: n_out receives the
 * filament count; out (capacity entries) may be NULL to query the count. */
fh_status fh_discretize(const fh_problem *problem, fh_filament *out, size_t capacity,
                        size_t *n_out);

/* The precorrected-FFT operator on its own: PfftOperator (pfft/mod.rs:270-528). */
typedef struct fh_pfft fh_pfft;

/* PfftParams::default (pfft/mod.rs:213-231). */
void fh_pfft_default_params(fh_pfft_params *out);

/* PfftOperator::new (pfft/mod.rs:332-445) for host filaments.  FH_E_PFFT
 * (code 1 InvalidParameter, 2 GridTooLarge), FH_E_KERNEL for a near pair
 * whose exact partial inductance fails (index0, index1). */
fh_status fh_pfft_create(const fh_filament *filaments, size_t n, const fh_pfft_params *params,
                         fh_pfft **out);

/* PfftOperator::apply (pfft/mod.rs:482-528): y = L x for nvec host vectors
 * of n values each, stored one after another. */
fh_status fh_pfft_apply(fh_pfft *op, const double *x, double *y, size_t nvec);

/* PfftOperator::stats (pfft/mod.rs:255-270). */
fh_status fh_pfft_get_stats(const fh_pfft *op, fh_pfft_stats *out);

/* Copies of the operator's sparse data, for parity checks; any pointer may
 * be NULL.  Projection CSR (proj_off n + 1, proj_idx / proj_w
 * projection_entries, grid points ascending within a row) and near-field
 * corrections (near_off n + 1, near_col / near_val near_entries, full
 * rows: partners j >= i in ascending order, mirrored). */
fh_status fh_pfft_data(const fh_pfft *op, int64_t *proj_off, uint32_t *proj_idx, double *proj_w,
                       int64_t *near_off, uint32_t *near_col, double *near_val);

void fh_pfft_destroy(fh_pfft *op);

/* gmres (gmres.rs:91-206) with the identity preconditioner on a dense
 * complex matrix: a is n x n column-major {re, im}, b and x (initial guess
 * in, solution out) n {re, im}.  complex_arith = 0 runs the real form on the
 * real parts (imaginary parts must be zero).  Outcome (gmres.rs:63-73):
 * iterations, relative residual, converged. */
fh_status fh_gmres_dense(size_t n, const double *a, const double *b, double *x,
                         const fh_gmres_params *params, int32_t complex_arith,
                         uint64_t *iterations, double *relative_residual, int32_t *converged);

/* Device-side log and atan of the fill kernels, for bit-level checks. */
fh_status fh_device_math(const double *x, size_t n, double *log_out, double *atan_out);

/* Message of the last error on the calling thread ("" when none). */
const char *fh_last_error(void);
/* Indices and value attached to the last error: index0, index1, value, code. */
void fh_error_detail(int64_t *index0, int64_t *index1, double *value, int32_t *code);

/* Library version string. */
const char *fh_version(void);

#ifdef __cplusplus
}
#endif
#endif
