/*
 * e3d.h — C ABI of the MI355X-native Euler3D explicit-solver hot path (libe3d_hip.so).
 *
 * Drop-in boundary for the reference's in-process solver seam. Each entry point names the
 * reference interface it replaces (paths are in the Euler3D source tree):
 *
 *   EulerSolver(FlowField&, MPIHandler&, Mesh<MeshPartition>&, SimConfig&, Metrics&)
 *       include/solver/EulerSolver.hpp:19-23          -> e3d_create
 *   EulerSolver::Run()            src/solver/EulerSolver.cpp:60-167   -> e3d_run / e3d_iterate
 *     (TIME_INTEGRATION 0/1/2: EulerExplicit / EulerImplicit / RungeKutta, EulerSolver.cpp:419-472;
 *      RESIDUAL_SMOOTHING: smoothResiduals, EulerSolver.cpp:483-518)
 *   EulerSolver::updateBC()       src/solver/EulerSolver.cpp:201-251  -> e3d_update_bc
 *   Solver::Roe / Solver::Fc      src/solver/ConvectiveFluxesSchemes.cpp:10-187 -> e3d_face_flux
 *   EulerSolver::computeResidual  src/solver/EulerSolver.cpp:303-359  -> e3d_residual
 *   TimeStepMethod2               src/solver/TimeIntegration.cpp:8-40 -> e3d_timestep
 *   AeroCoefficients::Update      src/solver/AeroCoefficients.cpp:38-52 -> e3d_aero
 *   FlowField set/get             include/solver/FlowField.hpp:29-115 -> e3d_set_state / e3d_get_state
 *   SimConfig(file)               src/parser/SimConfig.cpp:15-236      -> e3d_config_read
 *   SU2MeshParser / MeshPartition src/parser/SU2MeshParser.cpp:12-240,
 *                                 src/parser/MeshPartition.cpp:15-112  -> e3d_mesh_read
 *   SU2Writer                     src/partition/SU2Writer.cpp:17-123   -> e3d_mesh_write
 *   MeshPartition x N ranks + MPIHandler  MeshPartition.cpp:15-112,
 *                                 MPIHandler.cpp:16-313               -> e3d_mesh_read_partitions
 *   Partition::Write (E3D_PRE)    src/partition/Partition.cpp:357-460  -> e3d_mesh_partition_write
 *   Post(config).Write()          src/post/post.cpp:18-133             -> e3d_post_tecplot
 *   Mesh::solveConnectivity + Metrics  include/mesh/Mesh.hpp:36-105,
 *                                 src/mesh/Metrics.cpp:9-56            -> e3d_mesh_preprocess
 *
 * Conventions. Every function returns 0 (E3D_OK) or a positive error code; the message of the
 * last failure on the calling thread is e3d_last_error(). Host arrays are caller-owned and copied
 * in or out; device memory is owned by the opaque context. One host thread per context; contexts
 * are independent, several may share a device. Cell-indexed arrays follow the reference's
 * FlowField numbering: interior cells 0..n_elem-1, then ghost cells n_elem..n_elem+n_ghost-1
 * (ghost k sits on boundary face n_face_int + k). Per-face and per-cell 5-vectors are
 * component-minor: out[5*i + k], k = (rho, rho u, rho v, rho w, rho H) as in ResidualVar.
 * All arithmetic is IEEE FP64.
 */
#ifndef E3D_H
#define E3D_H

#ifdef __cplusplus
extern "C" {
#endif

#define E3D_OK 0
#define E3D_ERR_ARG 1
#define E3D_ERR_IO 2
#define E3D_ERR_MESH 3
#define E3D_ERR_DEVICE 4
#define E3D_ERR_UNSUPPORTED 5
#define E3D_ERR_BUSY 6 /* batch: no free case slot or the HBM budget is exhausted; retry after a removal */
#define E3D_ERR_NUMERIC 7 /* a NaN residual in the residual smoothing (the reference exits the process), or a
                           * batched case whose residual became NaN or infinite */

#define E3D_SCHEME_ROE 0
#define E3D_SCHEME_AUSM 1

typedef struct e3d_mesh e3d_mesh;
typedef struct e3d_ctx e3d_ctx;

/* Solver parameters, the fields of SimConfig that the hot path reads
 * (include/parser/SimConfig.hpp; keys parsed in src/parser/SimConfig.cpp:49-175). */
typedef struct e3d_params {
	double mach;              /* SPEED_VALUE (SPEED_OPTION 0) or derived from m/s (SimConfig.cpp:176-182) */
	double aoa_deg;           /* AOA */
	double gamma;             /* GAMMA */
	double cfl;               /* CFL */
	double sref, cref;        /* SREF, CREF */
	double ref_point[3];      /* MESH_REF_POINT_X/Y/Z */
	int orient_cl[2];         /* (axis, sign) of MESH_ORIENTATION_CL (SimConfig.cpp:185-207) */
	int orient_cd[2];         /* (axis, sign) of MESH_ORIENTATION_CD */
	int orient_cm[2];         /* derived, SimConfig.cpp:231 */
	int scheme;               /* SCHEME: 0 Roe, 1 AUSM */
	int time_integration;     /* TIME_INTEGRATION: only 2 (RK5) is on this path */
	int residual_smoothing;   /* RESIDUAL_SMOOTHING: must be 0 on this path */
	int max_iter;             /* MAX_ITER */
	int sampling;             /* SAMPLING */
	double min_residual;      /* MIN_RESIDUAL */
	double min_aerocoeff_error; /* MIN_AEROCOEFF_ERROR */
} e3d_params;

typedef struct e3d_mesh_info {
	int n_node, n_elem, n_face, n_face_int, n_ghost, n_wall, n_far, n_sym;
} e3d_mesh_info;

const char *e3d_last_error(void);
const char *e3d_version(void);
/* HIP devices visible to this process (0 without a device or driver). No counterpart in the
 * reference, which runs on the CPU. */
int e3d_device_count(int *n);

/* ---- configuration (.e3d) ---------------------------------------------------------------- */
int e3d_params_default(e3d_params *out);
/* Parses an .e3d file. mesh_path receives the first partition file (PARTITION_FILES pattern with
 * '#' -> 0), out_dir the directory of OUTPUT_FILE. */
int e3d_config_read(const char *path, e3d_params *out, char *mesh_path, int mesh_path_len, char *out_dir,
                    int out_dir_len);
/* SimConfig::getPartitionedMeshFiles (SimConfig.cpp:66-87): partition file `index` of the .e3d
 * file into buf (may be NULL) and the partition count into *n_part (may be NULL). */
int e3d_config_partition(const char *path, int index, char *buf, int len, int *n_part);
/* INITIAL_MESH of the .e3d file (SimConfig::getInitialMeshFile, parsed at SimConfig.cpp:64-65) */
int e3d_config_initial_mesh(const char *path, char *buf, int len);

/* ---- meshes ------------------------------------------------------------------------------- */
/* SU2 (.su2) or partition (.par, require_par = 1: the NINTERNAL_MARKER block must be present) */
int e3d_mesh_read(const char *path, int require_par, e3d_mesh **out);
/* SU2Writer layout; precision = digits after the point of the %e coordinates; par_block adds
 * "NINTERNAL_MARKER= 0" so the file is a one-partition .par */
int e3d_mesh_write(const e3d_mesh *m, const char *path, int precision, int par_block);

/* ---- several partitions ------------------------------------------------------------------- */
/* The n_part partition files one multi-rank run reads, rank r reading paths[r]
 * (Mesh<MeshPartition>, MeshPartition.cpp:15-112, src/Solver.cpp:37-49), as ONE mesh: after
 * e3d_mesh_preprocess every partition is numbered as the reference numbers it on its rank, and
 * the composite's cells, nodes, ghosts and wall faces are the partitions' in rank order. A context
 * created from it runs the reference's multi-rank solve on one device: interface faces read the
 * cell that the reference's ghost exchange delivers (MPIHandler.cpp:16-267), the RMS sums all
 * partitions, force and moment coefficients are summed per partition in rank order
 * (MPIHandler.cpp:268-313), and residual smoothing stops per partition. */
int e3d_mesh_read_partitions(int n_part, const char *const *paths, e3d_mesh **out);
typedef struct e3d_partition_info {
	int elem0, n_elem;       /* interior cells of the partition in the composite */
	int node0, n_node;       /* its nodes */
	int face_int0, n_face_int; /* its interior faces */
	int ghost0, n_ghost;     /* its ghosts: composite ghost n_elem_total + ghost0 + k, boundary face
	                          * n_face_int_total + ghost0 + k */
	int wall0, n_wall;       /* its entries of the wall lists (wallCell, wallFace, ...) */
} e3d_partition_info;
/* 1 for a mesh read on its own */
int e3d_mesh_partition_count(const e3d_mesh *m, int *n);
int e3d_mesh_partition_info(const e3d_mesh *m, int r, e3d_partition_info *out);
/* E3D_PRE for n_part > 1 (Partition::Write, Partition.cpp:357-460): writes the partition files of
 * a preprocessed single mesh, paths[r] for partition r, coordinates with `precision` digits (the
 * reference writes 10). elem2part gives each cell's partition; NULL uses this library's recursive
 * coordinate bisection (the reference calls METIS, which its sources do not include). */
int e3d_mesh_partition_write(const e3d_mesh *m, int n_part, const int *elem2part, const char *const *paths,
                             int precision);
/* The recursive-coordinate-bisection assignment e3d_mesh_partition_write uses by default. */
int e3d_mesh_partition_assign(const e3d_mesh *m, int n_part, int *elem2part);
/* Post::Post(config) + Post::Write()  src/post/post.cpp:18-133 (the E3D_POST executable,
 * src/postprocessor.cpp:17-40): writes the Tecplot ASCII file OUTPUT_FILE of the .e3d file, one
 * FEBRICK zone per listed partition, with the cell values read from
 * <directory of OUTPUT_FILE>/<partition file name>.sol. Host only: no device is used.
 * Missing or unreadable inputs give E3D_ERR_IO (the reference exits the process). */
int e3d_post_tecplot(const char *config_path);
/* The same writer over explicit paths: n_part partition meshes and the .sol file of each. */
int e3d_post_tecplot_files(const char *out_path, int n_part, const char *const *par_paths,
                           const char *const *sol_paths);
/* Synthetic meshes. tags6: six comma-separated tags for the x-, x+, y-, y+, z-, z+ sides. */
int e3d_mesh_gen_box(int nx, int ny, int nz, double lx, double ly, double lz, const char *tags6, e3d_mesh **out);
int e3d_mesh_gen_naca(int ni, int nj, int nk, double thickness, double camber, double camber_pos,
                      double wall_spacing, double farfield_radius, double dz, e3d_mesh **out);
int e3d_mesh_gen_bump(int nx, int ny, int nz, double jitter, unsigned long long seed, int permute, e3d_mesh **out);
int e3d_mesh_gen_wing(int ni, int nj, int nk, double thickness, double sweep_deg, double taper, double semispan,
                      double farfield, e3d_mesh **out);
/* Connectivity, ghost cells and FP64 metrics, numbered as the reference numbers them. */
int e3d_mesh_preprocess(e3d_mesh *m, double cref);
int e3d_mesh_get_info(const e3d_mesh *m, e3d_mesh_info *out);
/* Copies a named array out (capacity in elements; *n receives the full length). Names:
 * int32: face2element, face2nodeStart, face2node, element2faceStart, element2face,
 *        element2element, wallGhost, wallCell, wallFace, farGhost, farCell, farFace, symGhost,
 *        symCell, symFace, ghostKind, elemVtk, elemStart, elemNodes, mpiGhost, mpiCell, partStart,
 *        and of several partitions read together the interface exchange (MPIHandler::
 *        updateFlowField, one message per sending/receiving partition pair): haloStart (message
 *        offsets), haloFrom, haloTo (its partitions), haloSrc (sent cells), haloDst (receiving
 *        ghosts), and deviceOrder (the interior cells in the order the solver's default plan
 *        lays them out on the device: reference cell id of each device position)
 * f64:   faceNormals, faceNormalsUnit, faceSurfaces, faceCenters, cellVolumes, cellCentroids, xyz */
int e3d_mesh_get_i32(const e3d_mesh *m, const char *name, int *out, long long capacity, long long *n);
int e3d_mesh_get_f64(const e3d_mesh *m, const char *name, double *out, long long capacity, long long *n);
void e3d_mesh_free(e3d_mesh *m);

/* ---- device solver context ---------------------------------------------------------------- */
/* Uploads a preprocessed mesh to HIP device `device` and initialises the freestream state
 * (FlowField::Initialize, src/solver/FlowField.cpp:68-101). */
int e3d_create(const e3d_mesh *m, const e3d_params *p, int device, e3d_ctx **out);
/* fields = {rho, u, v, w, p, E, H, M}, each n_elem + n_ghost values. The explicit Euler step's
 * conservative state (FlowField's rho u, rho v, rho w, rho E) is set to u*rho, v*rho, w*rho, E*rho. */
int e3d_set_state(e3d_ctx *c, const double *const fields[8]);
int e3d_get_state(e3d_ctx *c, double *const fields[8]);
int e3d_update_bc(e3d_ctx *c);
/* per-face flux of the current state as computeResidual dispatches it; F unscaled, FS = F*S
 * (either may be NULL), 5*n_face values each */
int e3d_face_flux(e3d_ctx *c, double *F, double *FS);
/* residual of the current state (no BC update), 5*n_elem values */
int e3d_residual(e3d_ctx *c, double *R);
int e3d_timestep(e3d_ctx *c, double *dt);
/* n iterations of Run's body: updateBC, computeResidual, RungeKutta (5 stages). rms_sum (n values
 * or NULL) receives sum over interior cells of R_rho^2 of each iteration's last stage
 * (EulerSolver.cpp:96); RMS = sqrt(rms_sum / n_elem). */
int e3d_iterate(e3d_ctx *c, int n, double *rms_sum);
/* e3d_iterate without any device allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The per-iteration sums go through a static device array of 65536 doubles
 * (512 KiB per device, held until the process ends) instead of a buffer allocated per call, so n
 * is at most 65536 (E3D_ERR_ARG beyond). Every context on a device shares that array: calls that
 * overlap on one device (from several threads) race. Results are bitwise those of e3d_iterate. */
int e3d_iterate_noalloc(e3d_ctx *c, int n, double *rms_sum);
/* force and moment coefficients of the current state (AeroCoefficients::Update) */
int e3d_aero(e3d_ctx *c, double cf[3], double cm[3]);
/* cp of every wall face, in wall-list order, as the last force update left it (AeroCoefficients::
 * GetCp, include/solver/AeroCoefficients.hpp:53, written by EulerSolver::PrintCp,
 * src/solver/EulerSolver.cpp:282-300): inside e3d_run that is the last sampled iteration, which
 * need not be the last iteration. Zeros before any update. *n receives the wall face count;
 * cp may be NULL to ask for it. */
int e3d_wall_cp(e3d_ctx *c, double *cp, long long capacity, long long *n);
/* Run loop with MIN_RESIDUAL, SAMPLING and coefficient convergence (EulerSolver.cpp:60-167).
 * rows: up to max_rows rows of 8 doubles (iter, rms, Cx, Cy, Cz, CMx, CMy, CMz) as residuals.dat
 * holds them. coeffs = CL, CD, CM after the loop. */
int e3d_run(e3d_ctx *c, double *rows, int max_rows, int *n_rows, int *iters, double coeffs[3], double cf[3],
            double cm[3]);
/* Exact FP64 checkpoint of a context's solution (every cell's primitives including ghosts, the
 * explicit Euler step's conservative state, and whether the stored interior Mach numbers are
 * current), in reference cell numbering. Reading it into a context of the same mesh and
 * parameters continues bit for bit as if the run had not been interrupted. No counterpart in the
 * reference (restart files are not part of Euler3D). */
int e3d_checkpoint_write(e3d_ctx *c, const char *path);
int e3d_checkpoint_read(e3d_ctx *c, const char *path);
/* device time of the last e3d_iterate / e3d_run call in milliseconds (HIP events) */
int e3d_last_elapsed_ms(e3d_ctx *c, double *ms);
/* Per-launch timing of the dominant (per-stage residual + update) kernel: when enabled, every
 * launch of it during e3d_iterate / e3d_run is bracketed by HIP events on the context's stream.
 * e3d_kernel_timing reports the mean duration of those launches, their count and the kernel's
 * name, then clears the record. No counterpart in the reference (measurement hook). */
int e3d_set_kernel_timing(e3d_ctx *c, int enable);
int e3d_kernel_timing(e3d_ctx *c, double *avg_ms, int *launches, char *name, int name_len);

/* Self-check of the short FP64 sequences the fused stage uses in place of the library's square
 * root and division (no operand scaling): on device `device`, n pseudo-random inputs from `seed`
 * (square roots of positive values 2^-700 .. 2^700; reciprocals and quotients of operands
 * 2^-400 .. 2^400 of either sign, an eighth of the numerators +-0) are compared bit for bit with
 * the library results. mismatches[0..2] receive the counts for the square root, the reciprocal and
 * the quotient. No counterpart in the reference (a check of this library's arithmetic). */
int e3d_math_selfcheck(int device, long long n, unsigned long long seed, unsigned long long mismatches[3]);
/* One launch of a plain copy kernel over `bytes` (a positive multiple of 16) on device `device`:
 * every byte read once and written once, so its memory-counter readings scale the stage kernels'.
 * *ms (may be NULL) receives its duration. No counterpart in the reference (measurement hook). */
int e3d_copy_probe(int device, long long bytes, double *ms);
void e3d_destroy(e3d_ctx *c);

/* ---- mixed-precision multi-case engine (batched and served solves) ------------------------
 * No counterpart in the reference, which runs one case per process (src/Solver.cpp:19-73). Many
 * cases of any sizes stay resident on one device and advance together: one kernel launch per RK
 * stage covers the blocks of every resident case (one workgroup per block by default), and one iteration (5 stages plus
 * the per-case RMS and convergence test) is replayed from a captured HIP graph. Face fluxes,
 * ghosts and local time steps are FP32 (perturbation form, on FP32 copies of the state and the
 * face geometry); the conservative state, the RK update, the Kahan-compensated residual
 * accumulation, the residual norms and the force integration are FP64. A case's arithmetic does
 * not depend on the other resident cases, so its result is bitwise the same alone and in a mix.
 * Cases are admitted while others run, against the HBM budget and the slot count, and retired
 * without draining the batch. A converged case stops advancing on the device by itself. */
typedef struct e3d_batch e3d_batch;

typedef struct e3d_batch_opts {
	double hbm_budget;   /* bytes the resident cases may hold; <= 0: 90 % of the free device memory */
	int max_cases;       /* case slots */
	int wg_per_cu;       /* 0: one workgroup per (case, block) pair; else persistent workgroups per
	                        compute unit (launch shape; results unchanged) */
	int xcds;            /* XCD count of the L2-aware block order, 1 = dispatch order */
	int nontemporal;     /* non-temporal loads for read-once arrays */
	int use_graph;       /* replay iterations from a captured HIP graph */
} e3d_batch_opts;

/* how a batched case finished (the `done` value of e3d_batch_status / e3d_batch_poll) */
#define E3D_CASE_RUNNING 0
#define E3D_CASE_CONVERGED 1 /* residual drop or MIN_RESIDUAL reached */
#define E3D_CASE_CAPPED 2    /* iteration cap reached first */
#define E3D_CASE_FAILED 3    /* residual NaN or infinite: retired, the other cases are unaffected */
#define E3D_CASE_STALLED 4   /* residual stopped falling at the FP32 flux floor before the drop (stall
                                test of e3d_case_opts); finish on the FP64 path: e3d_batch_get_state,
                                e3d_create, e3d_set_state, e3d_iterate */

typedef struct e3d_case_opts {
	double drop;         /* converged when rms < drop * rms of the first iteration (0: off) */
	int max_iter;        /* iteration cap */
	int hist_cap;        /* residual RMS history entries kept on the device */
	int stall_window;    /* stall test, 0: off. Iterations are cut into windows of stall_window; a
	                        case retires as E3D_CASE_STALLED when a window's lowest RMS is not below
	                        stall_ratio x the previous window's lowest and is below stall_level x the
	                        first RMS */
	double stall_ratio;  /* (0, 1] */
	double stall_level;  /* > 0 */
} e3d_case_opts;

int e3d_batch_opts_default(e3d_batch_opts *out);
int e3d_batch_create(int device, const e3d_batch_opts *opts, e3d_batch **out);
/* device bytes a case of this preprocessed mesh holds while resident */
int e3d_batch_case_bytes(const e3d_mesh *m, const e3d_case_opts *co, double *bytes);
/* Admits a case (freestream start, FlowField::Initialize) and returns its id. The mesh may be
 * freed afterwards. E3D_ERR_BUSY when no slot is free or the case does not fit in what the
 * resident cases leave of the budget (retry after a removal); E3D_ERR_ARG when it needs more than
 * the whole budget. The
 * upload overlaps the iterations already queued; the case starts with the next e3d_batch_run. May
 * be called from other threads while e3d_batch_run / e3d_batch_poll run (a batch is otherwise
 * used from one thread). */
int e3d_batch_add(e3d_batch *b, const e3d_mesh *m, const e3d_params *p, const e3d_case_opts *co, int *case_id);
/* Queues n iterations of every resident, unfinished case and returns without waiting. */
int e3d_batch_run(e3d_batch *b, int n);
/* Waits for the queued iterations. */
int e3d_batch_wait(e3d_batch *b);
/* Waits, then reports a case's progress: done (E3D_CASE_*), iterations, RMS of the first and of
 * the last iteration. A failed case fills the outputs and returns E3D_ERR_NUMERIC. */
int e3d_batch_status(e3d_batch *b, int case_id, int *done, int *iters, double *rms0, double *rms_last);
/* Waits, then lists the resident cases (up to cap) with their progress (done = E3D_CASE_*); *n
 * receives their count. */
int e3d_batch_poll(e3d_batch *b, int *ids, int *done, int *iters, double *rms_last, int cap, int *n);
/* Residual RMS history (one value per iteration, at most cap) and the force and moment
 * coefficients of the current state: coeffs = CL, CD, CM (AeroCoefficients::Update,
 * src/solver/AeroCoefficients.cpp:38-52, FP64). A failed case fills the outputs and returns
 * E3D_ERR_NUMERIC. */
int e3d_batch_result(e3d_batch *b, int case_id, double *hist, int cap, int *n_hist, double coeffs[3], double cf[3],
                     double cm[3]);
/* FP64 state of a case in the FlowField numbering as e3d_get_state gives it: interior cells from
 * the FP64 conservative state, ghost cells from their FP32 copies */
int e3d_batch_get_state(e3d_batch *b, int case_id, double *const fields[8]);
/* Retires a case and releases its device memory (stream-ordered; no device-wide wait). */
int e3d_batch_remove(e3d_batch *b, int case_id);
/* Changes the launch shape (0 or persistent workgroups per CU, XCD order, non-temporal loads) of the
 * next iterations; results are bitwise the same for every shape. Used by the autotuner. */
int e3d_batch_set_tuning(e3d_batch *b, int wg_per_cu, int xcds, int nontemporal);
/* resident cases, bytes they hold, budget */
int e3d_batch_info(e3d_batch *b, int *n_cases, double *bytes, double *budget);
/* Per-launch timing of the stage kernel over the next e3d_batch_run calls (HIP events around
 * each stage when graphs are off, around each iteration otherwise); same contract as
 * e3d_kernel_timing. */
int e3d_batch_set_timing(e3d_batch *b, int enable);
int e3d_batch_timing(e3d_batch *b, double *avg_ms, int *launches, int *per_iteration);
void e3d_batch_destroy(e3d_batch *b);

#ifdef __cplusplus
}
#endif
#endif
