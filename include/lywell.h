/*
 * lywell: MI355X (gfx950) implementation of the MLFMAxwell hot path.
 *
 * C ABI called by the Fortran modules modMLFMA / modIterativo / modRCS
 * (lywell_amd/fortran/) through ISO_C_BINDING, and by any other host code.
 * Plain pointers and sizes only.  Conventions follow the reference's Fortran
 * arrays so a Fortran caller passes its arrays unchanged:
 *   - geometry arrays are column-major: (3,n) means xyz contiguous per item;
 *   - connectivity holds 1-based indices;
 *   - complex(8) vectors are interleaved (re, im) doubles, complex(4) vectors
 *     interleaved floats.
 * Every entry point returns a status code (LW_OK = 0) unless stated otherwise.
 * The reference has no error returns; the codes replace its STOP statements
 * and silent exits.  One handle per mesh; a handle is not reentrant.
 */
#ifndef LYWELL_H
#define LYWELL_H

#ifdef __cplusplus
extern "C" {
#endif

#define LW_OK 0
#define LW_ERR_ARG (-1)      /* invalid argument or size */
#define LW_ERR_HIP (-2)      /* a HIP runtime call failed */
#define LW_ERR_OCTREE (-3)   /* nivel_max <= 3: the reference STOPs (modMLFMA.f90:195-198) */
#define LW_ERR_NOMEM (-4)    /* device allocation failed */
#define LW_ERR_STATE (-5)    /* call order violated (e.g. lusol before ilu0) */
#define LW_ERR_PIVOT (-6)    /* ilu0 zero pivot; the row is returned in *ierr */
#define LW_ERR_NODEV (-7)    /* no HIP device visible */

/* Flags for lw_mlfma_init. */
#define LW_NEAR_FP64 1 /* store near field and ILU0 factor in complex(8)
                          (the reference's dp_near = dp_prec = 8 variant);
                          default keeps complex(4) storage as shipped */
#define LW_PATTERN_ORDER 2 /* aggregate and receive pattern by pattern, in the
                              reference's order over the bases (the tabulated
                              radiation patterns, receive and near field in one
                              kernel) instead of factored through the per-box
                              triangle phase rows; the default falls back to
                              this where a box's rows do not fit in LDS */

typedef struct lw_handle lw_handle;

/* Last error message for this thread (static storage). */
const char *lw_last_error(void);
/* Number of HIP devices visible (0 when none). */
int lw_device_count(void);

/* ------------------------------------------------------------------ mesh
 * Caller-side geometry helpers (the reference does this in host code:
 * modReadStl.f90:1378-1442, modRWG.f90:8-174, 270-328).  No GPU use. */

/* Icosphere sizes for a subdivision level: nt = 20*4^s, np = nt/2 + 2. */
int lw_mesh_icosphere_size(int subdiv, int *np, int *nt);
/* Icosphere scaled to semi-axes (ax, ay, az); t_p(3,nt) 1-based, outward
 * unit normals t_normal(3,nt) (may be NULL). */
int lw_mesh_icosphere(int subdiv, double ax, double ay, double az, double *p_coord,
                      int *t_p, double *t_normal);
/* Radial perturbation by real spherical harmonics: r *= 1 + sum c_i Y_{l_i m_i}. */
int lw_mesh_perturb_sh(int np, double *p_coord, int ncoef, const int *deg, const int *ord,
                       const double *coef);
/* ASCII STL (obtener_p_t, modReadStl.f90:1378-1533): sizes after merging
 * duplicate nodes, then the merged nodes p_coord(3,np), triangles t_p(3,nt)
 * (1-based) and facet normals t_normal(3,nt) in the order the reference's
 * reader and limpiarnodos produce them. */
int lw_stl_size(const char *path, int *np, int *nt);
int lw_stl_read(const char *path, double *p_coord, int *t_p, double *t_normal);
/* Writes an ASCII STL whose coordinates the reference's parser (s_to_r8)
 * reads back exactly (fixed point, 12 decimals, |x| < 1000).  t_normal may be
 * NULL (normals from the vertex order). */
int lw_stl_write(const char *path, int np, const double *p_coord, int nt, const int *t_p,
                 const double *t_normal);
/* parametros_rwg (modRWG.f90:133-174): outputs sized for 3*nt edges.
 * Returns the number of RWG edges (>0) or a negative status. */
int lw_rwg_build(int np, const double *p_coord, int nt, const int *t_p, int *e_p, int *e_t,
                 int *e_po, double *e_long, double *e_centro);
/* hallar_geometria_rwg (modRWG.f90:270-328). */
int lw_tri_geometry(int np, const double *p_coord, int nt, const int *t_p, double *t_area,
                    double *t_baric, double *t_baric_sub);

/* ------------------------------------------------------------------ result files
 * The reference's text writers (modGlobalMetod.f90:186-306), same layout and
 * F16.8 / I10 fields, so outputs diff against the reference's own files.
 * Host code, no GPU use.  Fields that do not fit are asterisks; NaN and
 * infinities are right-justified NaN / Inf / -Inf, as the flang build of the
 * reference writes them (ifort writes Infinity). */
/* guardarVectorC: v complex(8)[n] interleaved; one "re im" line per entry. */
int lw_dat_write_vector_c(const char *path, int n, const double *v);
/* guardarMatrizR: m(nrows, ncols) column-major (Fortran order); one line per row. */
int lw_dat_write_matrix_r(const char *path, int nrows, int ncols, const double *m);
/* guardarVectorR and guardarVectorI: one value per line. */
int lw_dat_write_vector_r(const char *path, int n, const double *v);
int lw_dat_write_vector_i(const char *path, int n, const int *v);
/* One F16.8 field of v into out17 (16 characters + terminator). */
int lw_dat_format_f16_8(double v, char *out17);

/* ------------------------------------------------------------------ modMLFMA */

/* inicializar_MLFMA (modMLFMA.f90:113-275).  Arguments mirror the
 * reference's dummy arguments plus lambda and conf%criterio_ladomax, which
 * the reference reads from module globals (modGlobalParam.f90:20-43).
 * Builds the octree and operators and uploads them to the current device.
 * LW_ERR_STATE if a level's far lists do not regroup into parent-neighbour
 * blocks (the translation runs per parent block; the octree built here
 * always regroups, the check guards the plan). */
int lw_mlfma_init(double lambda, double ladomax, int flags, int num_e, const double *e_centro,
                  const int *e_t, const int *e_po, const double *e_long, int num_t,
                  const double *t_area, const double *t_baric, const double *t_baric_sub,
                  int num_p, const double *p_coord, const int *t_p, const int *e_p,
                  const double *t_normal, lw_handle **out);
/* destruir_MLFMA (modMLFMA.f90:85-110). */
int lw_mlfma_destroy(lw_handle *h);

/* Octree shape: levels 1..nlevels as in the reference. */
int lw_mlfma_nlevels(lw_handle *h);
int lw_mlfma_level_info(lw_handle *h, int level, int *nboxes, int *nmult, int *nth, int *nph,
                        int *ntransf, long long *npairs);

/* MatxVec_MLFMA (modMLFMA.f90:1901-1946): y = Z x, host buffers
 * (complex(8)[num_e] each, interleaved). */
int lw_matvec(lw_handle *h, const double *x, double *y);
/* Same product on device buffers on the handle's stream; no host sync. */
int lw_matvec_dev(lw_handle *h, const void *x_dev, void *y_dev);
/* Far-field part only / near-field part only (per-stage checks). */
int lw_matvec_parts(lw_handle *h, const double *x, double *y_far, double *y_near);

/* n_esparcida_MLFMA and the CSR arrays Zesparcida_MLFMA / ia_MLFMA /
 * ja_MLFMA (modMLFMA.f90:30-32), copied to host.  z is complex(4)
 * interleaved (float[2*nnz]) unless the handle was built with LW_NEAR_FP64,
 * in which case use lw_near_csr64. */
long long lw_near_nnz(lw_handle *h);
int lw_near_csr(lw_handle *h, int *ia, int *ja, float *z);
int lw_near_csr64(lw_handle *h, int *ia, int *ja, double *z);

/* ------------------------------------------------------------------ modIterativo */

/* crear_precond ILU0 branch (main.f90:884-903) -> ilu0C (modIterativo.f90:1805)
 * on the handle's near CSR, factorised on the device (level-scheduled, in the
 * reference's elimination order and arithmetic); the MSR factor stays on the
 * device and a host copy is kept for lw_ilu0_get.
 * *ierr = 0 or the zero-pivot row (then LW_ERR_PIVOT is returned). */
int lw_ilu0(lw_handle *h, int *ierr);
/* Host copy of the MSR factor: alu(nnz+1) complex(4) (complex(8) with
 * LW_NEAR_FP64: pass alu64), jlu(nnz+1), ju(num_e), 1-based. */
int lw_ilu0_get(lw_handle *h, float *alu, double *alu64, int *jlu, int *ju);
/* Install an MSR factor held by the caller (the arrays ilu0C fills, same
 * layout as lw_ilu0_get) as the handle's preconditioner: for callers that
 * build or edit the factor themselves.  alu is complex(4) unless the handle
 * stores complex(8) (then pass alu64). */
int lw_ilu0_set(lw_handle *h, const float *alu, const double *alu64, const int *jlu, const int *ju);
/* ilu0C (modIterativo.f90:1805-1976) on any n x n CSR matrix held by the
 * caller: ia(n+1), ja(nnz) 1-based, every row with ascending columns and its
 * diagonal; a(nnz) and the factor alu(nnz+1) are complex(4) interleaved, or
 * complex(8) with LW_NEAR_FP64 (the reference's dp_near = dp_prec = 8).
 * Outputs the MSR arrays alu, jlu(nnz+1), ju(n) as ilu0C leaves them,
 * factorised on the device like lw_ilu0.  *ierr = 0, or the first row with a
 * zero pivot (then LW_ERR_PIVOT is returned and the factor is not usable).
 * Needs a device; no handle. */
int lw_ilu0_csr(int n, const int *ia, const int *ja, const void *a, int flags, void *alu, int *jlu, int *ju,
                int *ierr);
/* lusol0C (modIterativo.f90:504-560) with the handle's factor, host buffers. */
int lw_lusol0(lw_handle *h, const double *y, double *x);

/* CGSold (modIterativo.f90:331-493) with MatxVec_MLFMA as the operator.
 * use_prec: apply the ILU0 factor (lw_ilu0 must have succeeded).
 * maxit: iteration cap (the reference has none).  Outputs: x, iteration
 * count, recurrence relative residual ||r||/||b||, converged flag. */
int lw_cgs(lw_handle *h, const double *b, const double *guess, double errmax, int use_prec,
           int maxit, double *x, int *iters, double *relres, int *converged);

/* ------------------------------------------------------------------ excitation, modRCS */

/* generar_onda (main.f90:1488-1525): b for unit direction dir(3), complex
 * amplitude ctte (2 doubles) and complex polarisation pol (6 doubles). */
int lw_plane_wave(lw_handle *h, const double *dir, const double *ctte, const double *pol,
                  double *b);
/* inicializar_RCS (modRCS.f90:23-57): current vector and wave parameters. */
int lw_rcs_init(lw_handle *h, const double *current, const double *pol, const double *ctte);
/* sigmaRCS (modRCS.f90:91-105) at n angle pairs (radians) -> sigma in m^2. */
int lw_rcs_sigma(lw_handle *h, int n, const double *theta, const double *phi, double *sigma);

/* sigmaRCS at n angle pairs like lw_rcs_sigma, allocating no device memory:
 * no counterpart in the reference; for callers that cannot allocate during
 * the call (stream capture, serving loops).  Its scratch (angles, partial
 * fields, sigmas) is a fixed static device array per GPU, so
 * n <= LW_RCS_NOALLOC_MAX_ANGLES (else LW_ERR_ARG), and calls that overlap in
 * time on one GPU race.  Host buffers, synchronous like lw_rcs_sigma. */
#define LW_RCS_NOALLOC_MAX_ANGLES 4096
int lw_rcs_sigma_noalloc(lw_handle *h, int n, const double *theta, const double *phi, double *sigma);

/* ------------------------------------------------------------------ measurement */

/* Time `reps` device-resident matvecs (after one warm-up) with HIP events on
 * the handle's stream.  *ms receives the mean time of a matvec as the solver
 * runs it (the near field and the translations of levels 4..L on side
 * streams beside the rest).  stage_ms (may be NULL) receives the mean time
 * of each stage when the stages run one after the other: [0] aggregation,
 * [1] interpolation up, [2] translation, [3] anterpolation down, [4] near
 * field, [5] receive (0 in [4] where the receive and the near field are one
 * kernel). */
int lw_bench_matvec(lw_handle *h, int reps, double *ms, double *stage_ms);
/* Time `reps` ILU0 triangular solves on device-resident vectors. */
int lw_bench_lusol(lw_handle *h, int reps, double *ms);
/* Algorithmic HBM bytes of one matvec in this library's layout (the sum of
 * lw_matvec_stage_bytes): the phase rows (one exp(j k khat.(c_T - C)) row per
 * leaf box and triangle of its bases) and RWG vectors, read by the
 * aggregation and by the receive, campo arrays 6x, transfer tables once, near
 * CSR, x and y; and of the aggregation and of the near field + receive.
 * (The reference's layout reads the pattern array twice instead.) */
int lw_matvec_bytes(lw_handle *h, double *total, double *aggregation, double *receive_near);
/* Mean time in ms of a device-to-device copy of `bytes` bytes (a multiple of
 * 16) by the library's copy kernel on the handle's stream: the measured
 * bandwidth reference of the roofline figures. */
int lw_bench_copy(lw_handle *h, long long bytes, int reps, double *ms);
/* Algorithmic bytes of each matvec stage (6 values), in the order of
 * lw_bench_matvec's stage_ms: aggregation, upward passes, translations,
 * downward passes, near field, receive. */
int lw_matvec_stage_bytes(lw_handle *h, double *stage_bytes);

/* ---- multi-RHS form ------------------------------------------------------
 * nrhs right-hand sides of one mesh through one operator pass.  The far field
 * (aggregation, interpolation, translation, disaggregation, receive) runs in
 * FP32 (complex(4) campo arrays, FP32 accumulation) or, with
 * LW_BATCH_FAR_FP64, in FP64; the near field, the ILU0 solves and the CGS
 * vectors and scalars are FP64 in both.
 * Blocks are complex(8) arrays of shape (nrhs, ne) in Fortran order
 * (x(r,i) = right-hand side r, unknown i: nrhs values contiguous per unknown).
 * A batch shares the handle's stream and operators; one batch per handle at a
 * time.  No reference counterpart: the reference solves one RHS per CGSold
 * call (modIterativo.f90:331-493) from a loop over incidence angles. */
typedef struct lw_batch lw_batch;

/* Flags for lw_batch_create. */
#define LW_BATCH_FAR_FP64 1 /* far field in complex(8) (default: complex(4)) */
#define LW_BATCH_FP64_RESIDUAL 2 /* complex(4) far field in the iteration, judged by
                                    true residuals b - Z x with the complex(8)
                                    far field (a second operator on the handle):
                                    a column counts as converged only when its
                                    true residual meets errmax (otherwise CGSold
                                    starts afresh from its iterate), and with
                                    lw_batch_set_replacement the recurrence
                                    residual of every active column is replaced
                                    by its true residual periodically */

/* nrhs in {64, 128}. */
int lw_batch_create(lw_handle *h, int nrhs, int flags, lw_batch **out);
int lw_batch_destroy(lw_batch *b);
/* The block width, or LW_ERR_ARG. */
int lw_batch_nrhs(lw_batch *b);
/* LW_BATCH_FP64_RESIDUAL batches: replace the recurrence residual (and
 * CGSold's rho) of every active column by its true FP64 residual every
 * `period` iterations (checked every 8 iterations; 0 = never, the default).
 * Reliable updating in the sense of van der Vorst and Ye (2000). */
int lw_batch_set_replacement(lw_batch *b, int period);
/* Number of FP64 true-residual passes since the batch was created. */
long long lw_batch_true_residual_passes(lw_batch *b);
/* Y = Z X column by column (MatxVec_MLFMA per column, FP32 far field). */
int lw_batch_matvec(lw_batch *b, const double *x, double *y);
/* lusol0C per column with the handle's ILU0 factor. */
int lw_batch_lusol0(lw_batch *b, const double *y, double *x);
/* CGSold per column: each column stops at its own relative residual errmax
 * (its later updates are masked); iters/relres/converged are per column. */
int lw_batch_cgs(lw_batch *b, const double *rhs, const double *guess /* NULL: zero */, double errmax,
                 int use_prec, int maxit, double *x, int *iters, double *relres, int *converged);
/* Persistent form of lw_batch_cgs for sweeps whose columns run independent
 * chains of problems: lw_batch_cgs_start sets every column up (initial
 * residuals); lw_batch_cgs_advance runs up to `iterations` iterations (fewer
 * once no column is active) and returns per column state(4,nrhs) = (active,
 * first, converged, iterations) and relres(nrhs); lw_batch_column copies out
 * the current iterate of one column; lw_batch_restart_column gives column col
 * a new right-hand side and guess (NULL: zero), its initial residual taken
 * with the single-RHS FP64 operator, while the other columns keep iterating
 * in later calls. */
int lw_batch_cgs_start(lw_batch *b, const double *rhs, const double *guess, double errmax, int use_prec,
                       int maxit);
int lw_batch_cgs_advance(lw_batch *b, int iterations, int *state, double *relres);
int lw_batch_column(lw_batch *b, int col, double *x);
int lw_batch_restart_column(lw_batch *b, int col, const double *rhs, const double *guess);
/* generar_onda per column: dirs (3,nrhs) real, pols (3,nrhs) complex(8). */
int lw_batch_plane_wave(lw_batch *b, const double *dirs, const double *pols, const double *ctte, double *rhs);
/* sigmaRCS per column at one (theta(r), phi(r)) each: the monostatic sweep. */
int lw_batch_rcs(lw_batch *b, const double *cur, const double *pols, const double *ctte, const double *theta,
                 const double *phi, double *sigma);
/* Mean batched matvec time and one batched lusol0 time in ms (HIP events). */
int lw_bench_batch(lw_batch *b, int reps, double *matvec_ms, double *lusol_ms);
/* Tile shapes of the batched operator, chosen by timing each candidate on
 * this batch (HIP events, `reps` launches each): per level, the theta block
 * of the upward pass (4, 8 or 16 coarse rows per wave) and, for the fused
 * translation + downward pass, the sources per pipelined group and columns
 * per lane ((4,2), (2,2), (4,1) or (2,1)), whether the column chunks run
 * outermost, whether the transfer row is read from LDS, whether the
 * downward pass runs in a kernel of its own, and the parent ranges over the
 * XCDs (1, 2 or 8).  Every choice gives the same results.  ms (optional, 2):
 * the batched matvec time before and after; config (optional,
 * 3 (lw_mlfma_nlevels + 1)): per level n, (theta block, 10000 separate +
 * 1000 LDS + 100 chunks-outermost + 10 sources + columns, ranges), zeros
 * where a level has no such pass.  It overwrites the batch's
 * iterate: call it before a solve (a solve in progress must be started
 * again).  No reference counterpart. */
int lw_batch_tune(lw_batch *b, int reps, double *ms, int *config);
/* Sets the tile shapes explicitly, in lw_batch_tune's config layout (levels
 * 0-2 ignored; theta block 0 at level 3): pins a tuned configuration, or
 * compares shapes.  LW_ERR_ARG for a shape outside the lists above. */
int lw_batch_set_tiles(lw_batch *b, const int *config);

/* ---- learned initial guess for monostatic sweeps ---------------------------
 * No reference counterpart except the warm start it generalises: the
 * reference starts each angle of a sweep from the previous angle's current
 * times exp(-j k (d - d_prev) . e_centro) (main.f90:1017-1022).  Here a
 * pointwise MLP (in -> hidden x depth (tanh) -> out; FP32 master weights,
 * bfloat16 GEMM operands, FP32 accumulation on MFMA) maps 19 features per
 * unknown to 4 outputs o, and the guess for a new angle is
 *   (1 + o0 + j o1) xt1 + (o2 + j o3) (xt1 - xt2),
 * xt1 and xt2 being the two nearest solved currents with that phase
 * correction (o = 0 is the reference's warm start).  Parameters are laid out
 * layer by layer as W (out x in, row-major) then b. */
typedef struct lw_mlp lw_mlp;

int lw_mlp_create(int in_dim, int hidden, int depth, int out_dim, unsigned long long seed, lw_mlp **out);
int lw_mlp_destroy(lw_mlp *m);
long long lw_mlp_nparams(lw_mlp *m);
int lw_mlp_get_params(lw_mlp *m, float *p);
int lw_mlp_set_params(lw_mlp *m, const float *p);
/* y(rows, out) = MLP(x(rows, in)), host row-major buffers. */
int lw_mlp_forward(lw_mlp *m, const float *x, int rows, float *y);
/* Gradients of sum(dy .* MLP(x)) with respect to every parameter (grad,
 * nparams) and to x (dx, rows x in); either may be NULL. */
int lw_mlp_backward(lw_mlp *m, const float *x, int rows, const float *dy, float *grad, float *dx);
/* The guess for ncol new angles on the handle's mesh: dir(3,ncol) unit
 * propagation directions, pol(3,ncol) real polarisations, dir1/dir2 the
 * directions of the solved neighbours and x1/x2 their currents (complex(8),
 * (ncol, num_e) in Fortran order as in lw_batch_*).  guess: same layout. */
int lw_guess_predict(lw_mlp *m, lw_handle *h, int ncol, const double *dir, const double *pol,
                     const double *dir1, const double *x1, const double *dir2, const double *x2,
                     double *guess);
/* Appends training rows (one per unknown and column) with the converged
 * currents xtrue as targets; lw_guess_train then runs Adam steps on random
 * minibatches of batch_rows rows (loss per step into loss[steps] if not
 * NULL: mean relative squared error of the guess). */
int lw_guess_add_samples(lw_mlp *m, lw_handle *h, int ncol, const double *dir, const double *pol,
                         const double *dir1, const double *x1, const double *dir2, const double *x2,
                         const double *xtrue);
long long lw_guess_nsamples(lw_mlp *m);
int lw_guess_train(lw_mlp *m, int steps, int batch_rows, float lr, unsigned long long seed, float *loss);
/* A trained network from its file: an uncompressed .npz holding arch (int64
 * [in, hidden, depth, out]) and params (float32, the layout above), as
 * lywell_amd/guess_train.py writes it (lywell_amd/guess_mlp.npz). */
int lw_guess_load(const char *path, lw_mlp **out);

/* ---- monostatic sweep ------------------------------------------------------
 * gen_rcs_monostatica (main.f90:967-1089) in one call: n incidence angles
 * phi_a = a (span / n) (a = 0..n-1; the reference's span is pi) in the
 * theta = 90 deg plane, propagation (-cos phi, -sin phi, 0), polarisation as
 * ajustar_polariz sets it (main.f90:828-849: VV = z, HH = d x z / |d x z|),
 * each solved by CGSold (no preconditioner) to errmax and sigma[a] =
 * sigmaRCS(pi/2, phi_a) (main.f90:1048-1050; linear, as sigmaRCS returns it).
 * The angles run as chains of consecutive angles through the batch b (created
 * on h), each column restarted as soon as it stops, from the learned guess of
 * two solved neighbours when net is given, else the reference's warm start
 * from one (main.f90:1017-1022), else zero; a span of 2 pi wraps around.
 * iterations, converged (per angle), currents (complex(8), (num_e, n) in
 * Fortran order) and stats (restarts, steals, angles started from zero / one /
 * two solved neighbours) are optional. */
#define LW_POL_VV 0
#define LW_POL_HH 1
int lw_sweep_monostatic(lw_handle *h, lw_batch *b, lw_mlp *net, int n, double span, int pol, const double *ctte,
                        double errmax, int maxit, double *sigma, int *iterations, int *converged, double *currents,
                        int *stats);

#ifdef __cplusplus
}
#endif
#endif
