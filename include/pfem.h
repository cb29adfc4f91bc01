/*
 * pfem.h -- C-ABI of the MI355X-native Newton-step hot path (libpfem.so).
 *
 * Replaces two seams of the reference package PlasticityFEM (lagrangian/src):
 *   assemble! / assemble_threaded!     Assembly.jl:238-257, :340-354
 *   _linear_solve!                     Solver.jl:235-290
 * and offers the driver around them (newton! Solver.jl:317-389, solve!
 * Solver.jl:419-475) as fused entry points.  Plain pointers and sizes only; no
 * HIP or torch types cross this boundary.  All device memory (mesh, BSR K,
 * R, U, dU, Gauss state, AMG hierarchy, workspaces) is owned by the context
 * and stays resident between calls; host arrays are borrowed for the call.
 *
 * Conventions: indices 0-based (the reference is 1-based; a binding converts
 * once).  U, R are node-major (dof = 3*node + comp, Mesh.jl:111).  Gauss state
 * arrays are 6 x ngp column-major (st[6*gp + k], gp = 8*e + g, Mesh.jl:124-138).
 * Every call is synchronous on return and returns a pfem_status.
 * A context is not thread-safe; distinct contexts may be used concurrently.
 */
#ifndef PFEM_H
#define PFEM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pfem_status {
    PFEM_OK = 0,
    PFEM_ERR_INVERTED = 1,      /* ElementInversionError(e, g, J): pfem_fail filled (FiniteStrain.jl:42-52) */
    PFEM_ERR_NOT_CONVERGED = 2, /* Krylov not converged after the fallback ladder (Solver.jl:270-286) */
    PFEM_ERR_NON_SPD = 3,       /* breakdown: p'Ap <= 0 or r'z <= 0 with no fallback left */
    PFEM_ERR_UNSUPPORTED = 4,   /* material / element branch not implemented */
    PFEM_ERR_HIP = 5,           /* HIP runtime error: code in pfem_fail.hip */
    PFEM_ERR_BAD_ARG = 6,       /* invalid argument (guards of Materials.jl:90-103 included) */
    PFEM_ERR_NO_DEVICE = 7      /* no GPU visible */
} pfem_status;

typedef enum pfem_element_kind {
    PFEM_HEX8_FINITE = 0,       /* Hex8Finite      (FiniteStrain.jl:64) */
    PFEM_HEX8_FINITE_FBAR = 1,  /* Hex8FiniteFbar  (FiniteStrain.jl:65) */
    PFEM_HEX8_SMALL = 2         /* Hex8Small       (FiniteStrain.jl:63, Elements.jl:323-360) */
} pfem_element_kind;

/* J2Material inputs (Materials.jl:82-108).  ssat = sy0 and delta = 0 select the
 * linear isotropic law; Hkin > 0 adds linear kinematic hardening (with finite
 * strain the back stress is stored in the reference configuration and pushed
 * forward by the polar rotation of F, FiniteStrain.jl:172-346, which makes the
 * tangent non-symmetric).  Saturation together with Hkin > 0 is rejected
 * (Materials.jl:98-101). */
typedef struct pfem_material {
    double E, nu, sy0, Hiso, Hkin, ssat, delta;
} pfem_material;

/* Mesh and element description (Mesh.jl:17-22, Model.jl:63-75). */
typedef struct pfem_mesh_desc {
    int64_t nnodes;
    int64_t nelem;
    const double *nodes;        /* 3 x nnodes (x,y,z per node) */
    const int64_t *conn;        /* 8 x nelem, VTK hex order, 0-based */
    const int32_t *material_id; /* per element index into materials, or NULL (all 0) */
    int32_t nmaterials;
    const pfem_material *materials;
    int32_t element_kind;       /* pfem_element_kind */
} pfem_mesh_desc;

typedef enum pfem_precision {
    PFEM_PREC_FP64 = 0,         /* plain FP64 PCG + SA-AMG (the drop-in) */
    PFEM_PREC_FP32_IR = 1       /* FP32 copies of every AMG level and of the fine K; with fp32_krylov the
                                   Krylov iteration runs in FP32 inside FP64 iterative refinement, else
                                   FP64 flexible PCG refined by FP32 V-cycle corrections; flexible GCR
                                   with the FP32 cycle after a breakdown; the FP64 ladder last */
} pfem_precision;

typedef enum pfem_smoother {
    PFEM_SMOOTH_JACOBI = 0,     /* damped block-Jacobi, one sweep (Solver.jl:138 Jacobi(2/3; iter=1)) */
    PFEM_SMOOTH_CHEBYSHEV = 1,  /* Chebyshev polynomial in D^-1 A (block diagonal D) on [lmax/30, lmax] */
    PFEM_SMOOTH_CHEBYSHEV4 = 2  /* fourth-kind Chebyshev smoother in D^-1 A (Lottes 2022), no lower bound */
} pfem_smoother;

/* AMG rebuild policy of pfem_newton (pfem_options.amg_rebuild).  The reference
 * rebuilds once per load step (Solver.jl:374); the others trade set-up time
 * against Krylov iterations. */
typedef enum pfem_amg_rebuild {
    PFEM_AMG_PER_LOAD_STEP = 0,     /* full set-up at the first Newton iteration of a load step (the reference) */
    PFEM_AMG_EVERY_ITERATION = 1,   /* full set-up from every Newton iteration's tangent */
    PFEM_AMG_FIRST_TWO = 2,         /* at the first and second Newton iteration of a load step */
    PFEM_AMG_FINE_SMOOTHER = 3,     /* at the first, then only the fine-level smoother from each tangent */
    PFEM_AMG_KEEP_P_IN_STEP = 4,    /* at the first, then all but the prolongators from each tangent */
    PFEM_AMG_KEEP_P = 5,            /* prolongators from the first set-up, the rest from every tangent */
    PFEM_AMG_KEEP_P_UNTIL_FALLBACK = 6, /* as 5, the prolongators rebuilt after a solve that needed a fallback
                                           (the library default) */
    PFEM_AMG_REUSE_COARSE = 7           /* prolongators from the first set-up (as 5); at every Newton iteration
                                           except the second the levels below the finest are those built at
                                           the same iteration of an earlier load step (kept aside when built),
                                           and only the fine-level smoother comes from the new tangent; a kept
                                           set-up is redone when a solve on it needs more Krylov iterations
                                           than the solve right after it was built (by max(1, 1/16)), and all
                                           are dropped after a solve that needed a fallback */
} pfem_amg_rebuild;

/* Linear-solve options (LinearSolveState, Solver.jl:86-111) and the switches of the
 * device driver.  pfem_default_options fills every field. */
typedef struct pfem_options {
    int32_t precision;          /* pfem_precision */
    int32_t smoother;           /* pfem_smoother */
    int32_t smoother_degree;    /* sweeps (Jacobi) or polynomial degree (Chebyshev); <=0 => default */
    int32_t cg_itmax;           /* inner Krylov cap (default 200) */
    double eta_min, eta_max;    /* Eisenstat-Walker clamp (defaults 1e-8, 0.1) */
    double ew_gamma, ew_alpha;  /* defaults 0.9, 1.5 */
    int32_t coarse_max;         /* coarsest-level size target in DOF (default 1500) */
    int32_t max_levels;         /* AMG level cap (default 10) */
    int32_t verbose;            /* print per-iteration lines to stderr */
    int32_t amg_rebuild;        /* pfem_amg_rebuild (default PFEM_AMG_KEEP_P_UNTIL_FALLBACK) */
    int32_t amg_wcycle;         /* 1 = W-cycle (levels below the finest visited twice), 0 = V-cycle (default) */
    int32_t krylov_fallback;    /* after a CG breakdown or stall: 0 = flexible GCR(16) with the same AMG cycle
                                   (default), 1 = BiCGStab */
    int32_t gcr_directions;     /* directions kept by the learned preconditioner's flexible GCR (<= 0: 16) */
    int32_t fp16_fine;          /* FP32-IR: 1 = the V-cycle's fine-level smoother products use an FP16 copy of
                                   S K S (S = diag|K_ii|^-1/2); 0 = the FP32 copy of K (default) */
    int32_t line_search;        /* pfem_newton backtracking on ||R|| with at most this many step halvings
                                   (0: the reference's full Newton step, Solver.jl:383; default) */
    int32_t rho_iters_cold;     /* power-iteration steps of the rho(D^-1 A) estimate behind the Chebyshev */
    int32_t rho_iters_warm;     /* bounds: first set-up / warm-started rebuilds (<= 0: 25 / 4) */
    int32_t host_build;         /* 1 = build the pattern, slots and colours on the host in pfem_create */
    int32_t vcycle_eager;       /* 1 = launch the V-cycle's kernels one by one instead of replaying a HIP graph */
    int32_t prolong_rowwise;    /* 1 = fine-level prolongation with one thread per row instead of over chunks
                                   of consecutive rows */
    int32_t fp32_krylov;        /* FP32-IR: 1 = the Krylov iteration itself runs in FP32 (FP32 K p and FP32
                                   vectors) inside FP64 iterative refinement (FP64 residual b - K x once per
                                   refinement sweep); 0 = FP64 flexible CG whose preconditioner is the FP32
                                   V-cycle (default) */
    double newton_diverge;      /* pfem_newton stops a load step early (not converged) when, from the third
                                   iteration on, ||R_k|| > newton_diverge * min_{j<k} ||R_j||; 0 = never (the
                                   reference runs maxiter iterations; default).  For load-step cut-back. */
    int32_t assembly_chunk;     /* > 0: pipelined assembly -- the elements of each colour in chunks of at most
                                   this many; the Gauss-point kernel of chunk k+1 runs beside the scatter of
                                   chunk k on a second stream, and the packets live in two chunk-sized
                                   buffers instead of one per Gauss point of the mesh.  0: the Gauss-point
                                   kernel over the whole mesh, then the scatter colour by colour. */
    int32_t smoother_unfused;   /* 1 = the Chebyshev steps (fine and coarse levels) as a product launch followed
                                   by a vector launch; 0 = one launch that does both (default; same results bit
                                   for bit) */
    int32_t coarse_degree;      /* Chebyshev degree / Jacobi sweeps on the levels below the finest (<= 0: the
                                   fine level's) */
    int32_t krylov_sync;        /* 1 = PCG / flexible CG read each iteration's scalars on the host before they
                                   launch the next iteration; 0 = the next search direction and its product
                                   are launched first and the stopping test overlaps them (default; the same
                                   results bit for bit) */
    int32_t split_operator;     /* FP32-IR: 1 = flexible CG's product K p reads the FP32 copy of K (the V-cycle's)
                                   plus a BF16 copy of K - fl32(K): FP64 arithmetic on entries that agree with
                                   K to 2^-33 relative, 54 instead of 72 bytes of values per block (the
                                   default of pfem_default_options); 0 = the FP64 K */
    int32_t reserved[3];        /* zero */
} pfem_options;

typedef struct pfem_fail {
    int64_t e;                  /* element (0-based) */
    int32_t g;                  /* Gauss point 0..7 */
    double J;                   /* offending det F */
    int32_t hip;                /* hipError_t for PFEM_ERR_HIP */
} pfem_fail;

typedef struct pfem_ctx pfem_ctx;

/* Fills `o` with the reference defaults (Solver.jl:103-111). */
void pfem_default_options(pfem_options *o);

/* box_mesh(lx, ly, lz, nx, ny, nz) (Mesh.jl:30-59) generated on the device and copied
 * to nodes (3 x nnodes) and conn (8 x nelem, 0-based), with the seeded jitter of the
 * interior nodes: each coordinate moves by (2u - 1) * jitter * h, u =
 * splitmix64(splitmix64(seed) ^ (3 node + comp)) >> 11 scaled by 2^-53 (jitter = 0:
 * the reference's box_mesh exactly). */
int pfem_box_mesh(double lx, double ly, double lz, int64_t nx, int64_t ny, int64_t nz, double jitter,
                  uint64_t seed, double *nodes, int64_t *conn);

/* Model(mesh, material; element) set-up (Model.jl:63-75): uploads the mesh,
 * builds the block pattern (Assembly.jl:62-162, as 3x3 BSR) and the scatter slots
 * on the device, colours the elements (the parity colours of Assembly.jl:172-188
 * for box numbering, on the device; greedy first-fit in element order otherwise,
 * on the host), and resets the Gauss state (Mesh.jl:132-138).  The pattern is
 * built on the host instead when opts->host_build is set or when a node is shared
 * by more than 32 elements. */
int pfem_create(const pfem_mesh_desc *mesh, const pfem_options *opts, pfem_ctx **out);
void pfem_destroy(pfem_ctx *ctx);

/* Dirichlet / Neumann data (Model.jl:102-145, frozen as in _freeze_bcs
 * Solver.jl:192-218: duplicates resolved last-wins).  ramp[i] != 0 scales by
 * the load factor. */
int pfem_set_dirichlet(pfem_ctx *ctx, int64_t n, const int64_t *dofs, const double *values,
                       const uint8_t *ramp);
int pfem_set_neumann(pfem_ctx *ctx, int64_t n, const int64_t *dofs, const double *values);

/* reset! (Model.jl:157-162): U = 0, Gauss state undeformed. */
int pfem_reset(pfem_ctx *ctx);

int pfem_set_displacement(pfem_ctx *ctx, const double *U);
int pfem_get_displacement(pfem_ctx *ctx, double *U);
/* Committed Gauss state (state_committed, GaussState, Mesh.jl:124-163); any pointer
 * may be NULL.  epsp, beta, cpinv, sig: 6 x ngp; abar: ngp.  beta is the back stress
 * (for finite strain in the reference configuration); it is stored only for the
 * small-strain element or when a material has Hkin > 0 -- otherwise it is
 * identically zero (set ignores it, get returns zeros).  sig is the committed
 * stress (Kirchhoff for finite strain, Cauchy for small strain), written by
 * pfem_commit only (Solver.jl:470-471); it is zero after pfem_reset. */
int pfem_set_state(pfem_ctx *ctx, const double *epsp, const double *beta, const double *abar,
                   const double *cpinv);
int pfem_get_state(pfem_ctx *ctx, double *epsp, double *beta, double *abar, double *cpinv, double *sig);

/* The stress of the last assembly (the trial state's sigma that assemble! overwrites,
 * Assembly.jl:238-257): 6 x ngp, Kirchhoff for finite strain. */
int pfem_get_trial_stress(pfem_ctx *ctx, double *sig);

/* gauss_stress (Model.jl:188-208): the committed stress as Cauchy stress, 6 x ngp.
 * Finite strain: tau / det F with the un-barred F = I + grad U at each Gauss point
 * and the device-resident U (1/det F replaced by 1 where det F <= 0); small strain:
 * the committed stress itself.  Computed on the device. */
int pfem_gauss_stress(pfem_ctx *ctx, double *sig);

/* U <- U at the last pfem_commit (zero after pfem_reset or pfem_create): returns the
 * model to its committed state after a load step that failed, for the load-step
 * bisection of lagrangian/docs/SCALING.md:268-274.  The committed history is never
 * touched by a trial assembly, so U is all that has to be restored. */
int pfem_rollback(pfem_ctx *ctx);

/* assemble! (Assembly.jl:238-257): K (device) and R = F_int overwritten from U
 * (host array, or NULL to use the device-resident U) and the committed
 * history; the trial stress overwritten; commit != 0 also writes the new history
 * and the committed stress.  R_out may be NULL.  Returns PFEM_ERR_INVERTED with
 * `fail` filled on det F <= 0; K is then marked invalid (pfem_get_K_* values,
 * pfem_impose_dirichlet* and pfem_linear_solve refuse it with PFEM_ERR_BAD_ARG)
 * until an assembly succeeds.  Before the first assembly K is the zero matrix on
 * the pattern. */
int pfem_assemble(pfem_ctx *ctx, const double *U, int commit, double *R_out, pfem_fail *fail);

/* R -= lambda * Fext (Solver.jl:326-327, :349). */
int pfem_subtract_external(pfem_ctx *ctx, double lambda);

/* Both impose_dirichlet! calls of newton! (Solver.jl:354-355,
 * BoundaryConditions.jl:62-109) on the device K and R at load factor lambda:
 * the ramped set, then the fixed set. */
int pfem_impose_dirichlet(pfem_ctx *ctx, double lambda);

typedef enum pfem_bc_set {
    PFEM_BC_RAMPED = 0,         /* dir_ramp of _freeze_bcs (Solver.jl:192-218): g = lambda * value */
    PFEM_BC_FIXED = 1           /* dir_fix: g = value */
} pfem_bc_set;

/* One impose_dirichlet!(K, R, bc, lambda, U) call (BoundaryConditions.jl:62-109) for one
 * of the two frozen sets registered by pfem_set_dirichlet: delta g = g - U[d];
 * R[r] += K[r,d] delta g on rows r outside the set; the set's rows and columns of K
 * zeroed with a unit diagonal; R[d] = -delta g.  U (ndof, or NULL: the device U)
 * is uploaded first.  R (ndof, or NULL) is in/out on the host: when given, it
 * replaces the device R before the imposition and receives the result, so a host
 * driver that keeps R itself (newton! with R -= Fext on the host) can call it
 * unchanged. */
int pfem_impose_dirichlet_set(pfem_ctx *ctx, int set, double lambda, const double *U, double *R);

/* ||R||_2 (Solver.jl:358). */
int pfem_residual_norm(pfem_ctx *ctx, double *out);

/* _linear_solve!(dU, K, b, ls, rtol, rebuild_pc) (Solver.jl:235-290).
 * b: host right-hand side (ndof), or NULL for b = -R from the device R.  The solution
 * stays on the device (pfem_update_displacement, pfem_get_dU) and is also copied to
 * the host array dU when it is not NULL.  niter = Krylov iterations (all attempts);
 * solver_used: 0 PCG, 1 PCG after AMG refresh, 2 BiCGStab fallback (or the FP64
 * ladder after FP32-IR failed), 3 FP32-IR flexible PCG, 4 FP32-IR with BiCGStab,
 * 5 flexible GCR with the learned preconditioner, 6 the same finished by AMG,
 * 7 FP32-IR finished by flexible GCR(16) with the FP32 AMG cycle after a CG breakdown,
 * 8 FP64 finished by flexible GCR(16) with the AMG cycle after PCG failed, 9 FP32-IR with
 * the Krylov iteration in FP32 inside FP64 refinement (fp32_krylov), 10 FP32-IR flexible
 * PCG with the network's correction of the FP32 V-cycle (pfem_attach_net mode 5), 11 the
 * FP32 Krylov iteration (9) finished by FP64 flexible PCG with the FP32 cycle after its
 * recursion broke down at its rounding level.  The GCR
 * fallbacks (7, 8) are the default; krylov_fallback = 1 selects BiCGStab (2, 4).
 * true_relres = ||b - K dU|| / ||b||.
 * A solve that ends with ||b - K dU|| > ||b|| returns dU = 0 (not applied); its status
 * is then PFEM_ERR_NOT_CONVERGED or PFEM_ERR_NON_SPD.  Both statuses leave a usable dU
 * otherwise (the reference's fallbacks always produce a step, Solver.jl:257-286). */
int pfem_linear_solve(pfem_ctx *ctx, const double *b, double rtol, int rebuild_pc, int *niter,
                      int *solver_used, double *true_relres, double *dU);

/* U += dU (Solver.jl:383). */
int pfem_update_displacement(pfem_ctx *ctx);

/* Copy the device K out as BSR (rowptr nb+1, colind nnzb, vals 9*nnzb
 * row-major blocks) or as the reference's CSC (Assembly.jl:99-162).  Any
 * pointer NULL => only sizes returned. */
int pfem_get_K_bsr(pfem_ctx *ctx, int64_t *nnzb, int32_t *rowptr, int32_t *colind, double *vals);
int pfem_get_K_csc(pfem_ctx *ctx, int64_t *nnz, int64_t *colptr, int64_t *rowval, double *nzval);
int pfem_get_R(pfem_ctx *ctx, double *R);
int pfem_get_dU(pfem_ctx *ctx, double *dU);
/* Element colouring: colour[e], returns count in *ncolors. */
int pfem_get_colors(pfem_ctx *ctx, int32_t *colour, int32_t *ncolors);

typedef struct pfem_newton_result {
    int32_t converged;
    int32_t iters;              /* Newton iterations (length of the residual history) */
    int32_t krylov_iters[64];   /* per Newton iteration that solved (iters - converged of them) */
    int32_t solver_used[64];
    double residuals[64];
    double t_assemble, t_bc, t_setup, t_solve, t_total; /* seconds */
    double eta[64];             /* Eisenstat-Walker forcing term = Krylov rtol per iteration (Solver.jl:371) */
    int32_t solves;             /* linear solves done (= iters - converged unless stopped early) */
    int32_t discarded;          /* solves whose correction was discarded (||b - K dU|| > ||b||) */
    int32_t stopped;            /* 0 ran to convergence or maxiter; 1 stopped by newton_diverge; 2 stopped
                                   after two discarded solves in a row (no usable correction) */
} pfem_newton_result;

/* One load step of newton! at load factor lambda (Solver.jl:317-389): at each iteration
 * assemble, R -= lambda Fext, impose the ramped then the fixed Dirichlet set, test
 * ||R|| <= tol * max(||R_1||, ||Fext||, 1); otherwise solve with the Eisenstat-Walker
 * tolerance and U += dU -- also at iteration maxiter, as the reference does, so a
 * step that does not converge leaves U where the reference leaves it.  A solve whose
 * correction is discarded makes the next iteration rebuild the whole preconditioner;
 * a second one in a row ends the step (not converged, stopped = 2). */
int pfem_newton(pfem_ctx *ctx, double lambda, double tol, int maxiter,
                pfem_newton_result *res, pfem_fail *fail);

/* Commit pass of solve! (Solver.jl:470-471): the history (eps_p, abar, Cp^-1) and
 * the stress are updated from the current U.  solve! discards the K and R this pass
 * re-assembles, so the device K and R are left as they were. */
int pfem_commit(pfem_ctx *ctx, pfem_fail *fail);

/* Element-level entry for kernel checks (element_force_tangent_finite!
 * Elements.jl:405-494, element_force_tangent! :323-360): for nelem independent
 * elements with node coordinates Xe (24 per element, node-major) and
 * displacements ue (24), and per-GP history (epsp, beta 6x8, abar 8, cpinv 6x8
 * per element; beta NULL = zero back stress, cpinv may be NULL for the
 * small-strain kind), compute Fe (24), Ke (576, column-major 24x24) and sig
 * (6x8) on the device with the production per-GP and tangent-contraction
 * kernels.  commit != 0 updates the history arrays given.  material applies to
 * all elements. */
int pfem_element_batch(int32_t element_kind, const pfem_material *mat, int64_t nelem,
                       const double *Xe, const double *ue, double *epsp, double *beta, double *abar,
                       double *cpinv, int commit, double *Fe, double *Ke, double *sig,
                       pfem_fail *fail);

/* pfem_element_batch without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * material-point loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results as pfem_element_batch, limited to nelem <= PFEM_BATCH_NOALLOC_MAX
 * (PFEM_ERR_BAD_ARG above).  The device copies of inputs and outputs and the
 * per-Gauss-point packets live in one file-scope device array of about 32 MB that each
 * GPU the library runs on holds until the process ends; the call runs on the null
 * stream and is synchronous.  Calls that overlap on one GPU (several host threads) share
 * the array and race. */
#define PFEM_BATCH_NOALLOC_MAX 2048
int pfem_element_batch_noalloc(int32_t element_kind, const pfem_material *mat, int64_t nelem,
                               const double *Xe, const double *ue, double *epsp, double *beta,
                               double *abar, double *cpinv, int commit, double *Fe, double *Ke,
                               double *sig, pfem_fail *fail);

typedef struct pfem_stats {
    int64_t nnodes, nelem, ndof, nnzb, ncolors;
    int32_t amg_levels;
    int64_t amg_ndof[12], amg_nnzb[12];
    double amg_operator_complexity;
    double t_assemble, t_setup, t_solve;    /* accumulated seconds */
    double t_gp_kernel, t_scatter_kernel;   /* accumulated device seconds (events) */
    double t_spmv;                           /* last pfem_bench_spmv result (s per SpMV) */
    double t_pcg_spmv;                       /* accumulated device time of the fine FP64 SpMV in PCG (s) */
    int64_t n_pcg_spmv;                      /* number of those launches */
    double t_setup_rho, t_setup_galerkin, t_setup_coarse;  /* AMG numeric set-up phases (s) */
    int64_t n_assemble, n_setup, n_solve, n_krylov;
    int64_t device_bytes;                    /* device memory held by the context */
    double t_create_pattern;                 /* pfem_create: pattern, scatter slots and colours (s) */
    int32_t built_on_device;                 /* bit 0: pattern + slots built on the device, bit 1: colours */
    double t_krylov_spmv32;                  /* accumulated device time of the FP32 Krylov SpMV (fp32_krylov) */
    int64_t n_krylov_spmv32;                 /* number of those launches */
} pfem_stats;

int pfem_get_stats(pfem_ctx *ctx, pfem_stats *st);

/* Dense inverse of the n x n row-major A on the device with the coarsest-level AMG
 * solver (Gauss-Jordan with partial pivoting, blocked panels when they fit in LDS;
 * an all-zero row is decoupled with a unit diagonal).  No counterpart in the
 * reference; exported so the coarse solve can be checked on its own. */
int pfem_dense_inverse(int64_t n, const double *A, double *inv);

/* Achievable device peaks for the rooflines: the best of `reps` launches of a
 * STREAM-style copy of `copy_bytes` (>= 1 MiB; read + write bytes counted) and of a
 * read-only stream over the same bytes, in GB/s (read_gbs may be NULL), and of an
 * FP64 FMA loop (8 independent chains per lane) in TFLOP/s.  Allocates 2 x copy_bytes
 * of device memory for the call.  No counterpart in the reference. */
int pfem_device_peaks(int64_t copy_bytes, int reps, double *copy_gbs, double *read_gbs, double *fp64_tflops);

/* Time `reps` fine-level SpMVs y = K x (FP64 if prec == 0, FP32 if 1, FP64 from the
 * FP32 + BF16 split copy of K if 2) on the context's stream with events; returns mean
 * seconds per SpMV.  prec 1 and 2 make their copies from the current K. */
int pfem_bench_spmv(pfem_ctx *ctx, int prec, int reps, double *sec_per_spmv);

/* y = K x for host vectors of ndof entries with the current device K after
 * assembly and Dirichlet imposition: prec 0 the FP64 K, 1 its FP32 copy (FP32 x and
 * products, widened), 2 the FP32 + BF16 split copy (FP64 arithmetic).  prec 1 and 2 make
 * their copies from the current K.  No counterpart in the reference (operator checks). */
int pfem_apply_K(pfem_ctx *ctx, int prec, const double *x, double *y);

/* Time `reps` applications of the current AMG cycle (prec 0: FP64, 1: FP32, which needs
 * PFEM_PREC_FP32_IR) and `reps` fused fine-level Chebyshev smoother steps (product
 * t = K d with the update r -= t, d' = c1 d + c2 D^-1 r, x += d') on the context's
 * stream with events, after one untimed application.  Returns seconds per cycle and per
 * step and their algorithmic bytes (every launch's operands read once, outputs written
 * once).  Needs a built preconditioner (a pfem_linear_solve or pfem_newton before);
 * overwrites the cycle's work vectors.  No counterpart in the reference (measurement). */
int pfem_bench_vcycle(pfem_ctx *ctx, int prec, int reps, double *sec_per_cycle, double *bytes_per_cycle,
                      double *sec_per_smoother_step, double *bytes_per_smoother_step);

/* ---------------------------------------------------------------------------
 * Learned preconditioner.  No counterpart in the reference: the network and
 * its training are additions beside the drop-in path.  A bias-free 3D
 * convolutional U-net (4 levels, channels[0..3] per level, ReLU, additive skip
 * connections) on the structured node grid of a box mesh (node index
 * x + nx*(y + ny*z)); BF16 weights and activations on MFMA with FP32
 * accumulation; hand-written forward, backward and Adam kernels.  It maps the
 * 3-component nodal residual to a smooth correction n; the preconditioner
 * z = g n + D^-1 (r - g K n) (g = (Kn)'r / (Kn)'(Kn), D the 3x3 block diagonal of K)
 * replaces the V-cycle inside FP64 flexible CG, and AMG takes over from the
 * current iterate when the iteration stalls.  Trained on (residual, V-cycle correction) samples that
 * the solver logs from its own AMG-preconditioned solves, with the
 * scale-invariant loss 1 - cos(N(r), z).
 * ------------------------------------------------------------------------- */
typedef struct pfem_net pfem_net;

typedef struct pfem_net_desc {
    int32_t nx, ny, nz;         /* node grid (x fastest) */
    int32_t channels[4];        /* per level, each 16, 32 or 64 (default 16, 32, 64, 64) */
    int64_t capacity;           /* sample store: (residual, correction) pairs kept on the device */
    uint64_t seed;              /* weight initialisation (He normal) */
} pfem_net_desc;

int pfem_net_create(const pfem_net_desc *desc, pfem_net **out);
void pfem_net_destroy(pfem_net *net);
int64_t pfem_net_num_params(const pfem_net *net);
/* Kernel of the 3x3x3 layers: 1 (default) LDS-tiled implicit GEMM on the 16x16x32 BF16
 * MFMA, 0 one wave per 16 voxels on the 16x16x16 MFMA (kept for comparison). */
int pfem_net_set_conv_kernel(pfem_net *net, int kind);
/* FP32 master weights, layer by layer, each [tap][c_in][c_out]. */
int pfem_net_get_params(pfem_net *net, float *w);
int pfem_net_set_params(pfem_net *net, const float *w);   /* also resets Adam */
/* out = N(in) for host 3-channel fields (3 * nx*ny*nz floats, node-major). */
int pfem_net_forward(pfem_net *net, const float *in, float *out);
/* loss = 1 - cos(N(in), target) and its gradient w.r.t. the weights (nparams
 * floats, may be NULL); inputs are rounded to BF16 like logged samples. */
int pfem_net_loss_grad(pfem_net *net, const float *in, const float *target, double *loss, float *grad);
/* Append one sample (host data) to the store (ring buffer). */
int pfem_net_add_pair(pfem_net *net, const float *r, const float *z);
int pfem_net_num_pairs(const pfem_net *net, int64_t *count);
/* `steps` Adam steps (beta 0.9/0.999, eps 1e-8), each over `batch` samples drawn
 * from the store; mean_loss = mean loss over the samples seen. */
int pfem_net_train(pfem_net *net, int steps, int batch, double lr, uint64_t seed, double *mean_loss);
/* mode 0: detach; 1: log (r, V-cycle(r)) samples from every `stride`-th
 * preconditioner application of pfem_linear_solve / pfem_newton into the
 * net's store; 2: precondition pfem_linear_solve with the network (flexible
 * CG, AMG fallback); 3: the same preconditioner without its network
 * (block-Jacobi flexible CG, AMG fallback; net may be NULL), the baseline that
 * shows what the network adds; 4 (FP32-IR only): log error samples from every
 * `stride`-th FP32 V-cycle application of the FP32-IR flexible PCG -- input the
 * cycle's output z_k = V(r_k), target the error it left, x* - x_k - z_k, both divided
 * by max |z_k|, kept only from solves that converge; 5 (FP32-IR only): precondition the
 * FP32-IR flexible PCG with the V-cycle and the network's correction of its output,
 * z = V(r) + c N(V(r) / c), c = max |V(r)| (the plain cycle finishes a solve the
 * corrected one does not).  The net's grid must hold exactly the context's nodes. */
int pfem_attach_net(pfem_ctx *ctx, pfem_net *net, int mode, int stride);
/* Training loss: 0 (default) 1 - cos(N(x), t) (scale-invariant, for mode 1/2 samples),
 * 1 ||N(x) - t||^2 / ||t||^2 (for the error samples of mode 4, whose scale matters). */
int pfem_net_set_loss(pfem_net *net, int kind);

/* Human-readable text for a status. */
const char *pfem_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
