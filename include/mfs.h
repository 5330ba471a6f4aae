/* mfs.h - C ABI of the MI355X (gfx950) FEM solve path.
 *
 * One shared library, libmfs.so, exports every function declared here.  All
 * entry points return int status (MFS_OK = 0, negative = error); the message
 * of the last error on the calling thread is returned by mfs_last_error().
 * Host pointers passed in are borrowed and copied before the call returns.
 * Device objects (matrices, vectors, solvers, batches, networks) are owned by
 * the context that created them; all work of a context runs in order on the
 * context's single HIP stream.  A context is not re-entrant: use one per host
 * thread.
 *
 * Each declaration names the reference interface it replaces (file:line in
 * AnastasiaKutishcheva/MultiscaleFEMSolver3D).  Sparse patterns use the
 * reference's CSSD row order: row = down ++ diag ++ up, every list sorted
 * ascending (Library/CSSD_Matrix.h:10-17), i.e. a sorted CSR with the
 * diagonal present in every row.
 */
#ifndef MFS_H
#define MFS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
  MFS_OK = 0,
  MFS_ERR_ARG = -1,           /* bad argument / shape mismatch            */
  MFS_ERR_OOM = -2,           /* device or host allocation failed         */
  MFS_ERR_NOT_CONVERGED = -3, /* iteration limit reached                  */
  MFS_ERR_NONFINITE = -4,     /* NaN/Inf met in a residual or input       */
  MFS_ERR_HIP = -5,           /* HIP runtime error                        */
  MFS_ERR_UNSUPPORTED = -6,   /* valid request this build does not serve  */
  MFS_ERR_NO_DEVICE = -7      /* no gfx950 device visible                  */
};

/* precisions / formats / methods */
enum { MFS_F64 = 0, MFS_F32 = 1 };
enum { MFS_FMT_CSR = 0, MFS_FMT_BSR3 = 1 };
enum { MFS_PCG_JACOBI = 0, MFS_CG = 1, MFS_GMRES = 2, MFS_MINRES = 3, MFS_BICGSTAB = 4 };
enum { MFS_PREC_F64 = 0, MFS_PREC_MIXED = 1 };
enum { MFS_PROBLEM_SCALAR = 0, MFS_PROBLEM_ELASTIC = 1 };
/* Elastic element block convention.  REFERENCE: block(I,J) = V B_J^T D B_I,
 * exactly what FEM/FEM_Grid.h:2125-2191 assembles (the transpose of the
 * textbook block; identical rows for nodes off the Neumann/free boundary).
 * TEXTBOOK: block(I,J) = V B_I^T D B_J. */
enum { MFS_ELEM_REFERENCE = 0, MFS_ELEM_TEXTBOOK = 1 };
/* MsFEM boundary data of the local problems: the trilinear coarse hat times
 * e_d, or the reference's edge -> face -> volume hierarchy
 * (MsFEM/MsFEM_Grid.h:5026-5506) driven by e_d. */
enum { MFS_TRACE_Q1 = 0, MFS_TRACE_REFERENCE = 1 };
/* MsFEM coarse functions: the 24 bases phi_{c,d} themselves, or the
 * reference's component-wise use of one vector basis per coarse node,
 * Psi_{c,a} = e_a (sum_d phi_{c,d})_a (MsFEM/MsFEM_Grid.h:3872-3903). */
enum { MFS_BASIS_PER_DIRECTION = 0, MFS_BASIS_COMPONENTWISE = 1 };

typedef struct mfs_ctx_s* mfs_ctx;
typedef struct mfs_vec_s* mfs_vec;
typedef struct mfs_mat_s* mfs_mat;
typedef struct mfs_solver_s* mfs_solver;
typedef struct mfs_asm_s* mfs_asm;
typedef struct mfs_batch_s* mfs_batch;
typedef struct mfs_mlp_s* mfs_mlp;

/* ---- errors / context ------------------------------------------------- */
/* No reference counterpart (the reference prints and continues,
 * CSSD_Matrix.h:359-362); text of the last error on this thread. */
const char* mfs_last_error(void);
/* Library version string. */
const char* mfs_version(void);
/* Create a context on HIP device `device`: one stream, one allocator pool.
 * Replaces the implicit global state of the reference (math::NUM_THREADS,
 * Library/Math.h:19). */
int mfs_ctx_create(int device, mfs_ctx* out);
int mfs_ctx_destroy(mfs_ctx ctx);
/* Wait until all work queued on the context's stream has finished. */
int mfs_ctx_sync(mfs_ctx ctx);
/* Device name and compute-unit count of the context's device. */
int mfs_ctx_info(mfs_ctx ctx, char* name, int name_len, int* cu_count);

/* ---- device vectors --------------------------------------------------- */
/* Replaces std::vector<double> X, F of CSSD_Matrix (CSSD_Matrix.h:15). */
int mfs_vec_create(mfs_ctx ctx, int64_t n, int prec, mfs_vec* out);
int mfs_vec_destroy(mfs_vec v);
int mfs_vec_size(mfs_vec v, int64_t* n, int* prec);
/* host f64 <-> device (converted to the vector's precision) */
int mfs_vec_upload(mfs_vec v, const double* host);
int mfs_vec_download(mfs_vec v, double* host);
int mfs_vec_fill(mfs_vec v, double value);
int mfs_vec_copy(mfs_vec dst, mfs_vec src); /* converts precision if needed */

/* ---- sparse matrices -------------------------------------------------- */
/* Scalar CSSD matrix given as sorted CSR (row = down ++ diag ++ up).
 * Replaces CSSD_Matrix<double,double>::Initialization + value fill
 * (CSSD_Matrix.h:79-99).  rowptr[n] = nnz, int32 indices. */
int mfs_mat_create_csr(mfs_ctx ctx, int64_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                       mfs_mat* out);
/* Block 3x3 CSSD matrix (CSSD_Matrix<Tensor2Rank3D, Point<double>>) as BSR3:
 * nb block rows, browptr/bcol sorted, val = 9 doubles per block, row-major
 * (val[9k + 3a + b] = block.val[a][b]).  The scalar view is the expansion of
 * math::MakeCopyMatrix_A_into_B (Library/Math.h:1963-2054): scalar row 3i+a,
 * column 3j+b. */
int mfs_mat_create_bsr3(mfs_ctx ctx, int64_t nb, const int32_t* browptr, const int32_t* bcol, const double* val,
                        mfs_mat* out);
/* Replace the values, keeping the pattern (same layout as at creation). */
int mfs_mat_update_values(mfs_mat A, const double* val);
int mfs_mat_destroy(mfs_mat A);
/* n = scalar rows, nnz = scalar nonzeros, format = MFS_FMT_*; block counts
 * (nb, nnzb) equal n, nnz for CSR. */
int mfs_mat_info(mfs_mat A, int64_t* n, int64_t* nnz, int* format, int64_t* nb, int64_t* nnzb);
/* Export the scalar CSR (pattern + FP64 values) in the reference's order:
 * for BSR3 this is exactly the row/column order MakeCopyMatrix_A_into_B
 * produces (Math.h:1963-2054).  Arrays sized n+1 / nnz / nnz. */
int mfs_mat_export_csr(mfs_mat A, int32_t* rowptr, int32_t* col, double* val);

/* ---- kernels ----------------------------------------------------------- */
/* y = A x.  prec selects the matrix copy and vector type (MFS_F64 / MFS_F32;
 * x, y must have that precision).  Replaces
 * CSSD_Matrix::MultiplicationMatrixVector (CSSD_Matrix.h:395-423). */
int mfs_spmv(mfs_mat A, int prec, mfs_vec x, mfs_vec y);
/* result = x . y  (Math.h:1633-1654 MakeInnerProduct) */
int mfs_dot(mfs_vec x, mfs_vec y, double* result);
/* y = a x + b y  (Math.h:1601-1616 MakeSummVectors) */
int mfs_axpby(double a, mfs_vec x, double b, mfs_vec y);
/* Average device time (ms, HIP events on the context stream) of `reps`
 * back-to-back SpMV launches; measurement helper, no reference counterpart. */
int mfs_spmv_time(mfs_mat A, int prec, mfs_vec x, mfs_vec y, int reps, double* avg_ms);

/* ---- solvers ----------------------------------------------------------- */
typedef struct {
  int method;       /* MFS_PCG_JACOBI | MFS_CG | MFS_GMRES | MFS_MINRES |
                       MFS_BICGSTAB                                            */
  int precision;    /* MFS_PREC_F64 | MFS_PREC_MIXED                          */
  double tol;       /* stop when ||b - A x||_2 / ||b||_2 <= tol (FP64, true)  */
  int maxit;        /* total inner iterations cap                             */
  int restart;      /* GMRES restart length m (default 30)                    */
  double inner_tol; /* MIXED: relative tolerance of each FP32 inner solve     */
  int check_every;  /* iterations per device-resident chunk (0 = auto)        */
} mfs_solve_desc;

typedef struct {
  int iterations;        /* total (inner) iterations                         */
  int outer_steps;       /* FP64 refinement steps (MIXED), else 1            */
  double true_relres;    /* ||b - A x|| / ||b|| in FP64 at exit              */
  double recursive_relres; /* solver's own residual estimate at exit         */
  double seconds;        /* wall time of the solve (first residual to exit)  */
  int status;            /* MFS_OK or MFS_ERR_NOT_CONVERGED / NONFINITE      */
} mfs_solve_report;

/* Fill desc with defaults: Jacobi-PCG, FP64, tol 1e-8, maxit 100000,
 * restart 30, inner_tol 1e-4. */
void mfs_solve_desc_default(mfs_solve_desc* desc);
/* Plan a solve on A: allocates work vectors and captures the iteration
 * chunk into HIP graphs.  MFS_MINRES (Jacobi-preconditioned, symmetric
 * indefinite A) serves the reference element convention's elastic operator,
 * which is indefinite for heterogeneous or free-surface problems.
 * MFS_BICGSTAB (right Jacobi preconditioning) serves non-symmetric
 * systems.  Replaces the solver family of CSSD_Matrix (MSG :1496-1663,
 * BiCG_Diag :1181-1300, MSG_PreconditioningSSOR :1302-1495, GMRES
 * :1774-1989, BiCG_Stab :730-853 / BiCG_Stab_Preconditioning :1058-1163)
 * together with the drivers' continuation loops (FEM/FEM.h:626-695,
 * :2399-2485). */
int mfs_solver_create(mfs_mat A, const mfs_solve_desc* desc, mfs_solver* out);
int mfs_solver_destroy(mfs_solver S);
/* Solve A x = b; x holds the initial guess on entry (FP64 vectors). */
int mfs_solver_run(mfs_solver S, mfs_vec b, mfs_vec x, mfs_solve_report* report);
/* Time the iteration kernels of a CG-method solver in its inner precision
 * (FP32 for MFS_PREC_MIXED): kernel 0 = direction update fused with the
 * SpMV (q = A (z + beta p)), kernel 1 = the x/r/z update.  The kernel is
 * launched
 * reps times back to back on the context stream and timed with HIP events;
 * *bytes receives its algorithmic HBM bytes per launch.  Overwrites the
 * solver's work vectors (a later run starts afresh).  No reference
 * counterpart: measures the kernels replacing the inner loop of
 * CSSD_Matrix::MSG / BiCG_Diag (Library/CSSD_Matrix.h:1496-1663). */
int mfs_solver_kernel_time(mfs_solver S, int kernel, int reps, double* avg_ms, double* bytes);
/* Solve k <= 8 systems A x_j = b_j that share A (FP64 vectors; x_j holds
 * the initial guess on entry) as one iteration stream: each iteration reads
 * A once for all systems (SELL SpMM over interleaved vectors), per-system
 * alpha, beta and stop tests, a converged system is frozen while the others
 * continue.  desc->method: MFS_PCG_JACOBI or MFS_CG; precision MFS_PREC_F64
 * or MFS_PREC_MIXED (FP64 refinement per system around the FP32
 * iteration).  reports[j] receives system j's outcome (seconds: the whole
 * call).  Replaces the homogenisation driver's six sequential solves with
 * the same matrix (Mechanical/MechanicalMain.cpp:243-420, one
 * CSSD_Matrix::MSG_PreconditioningSSOR per STAGE). */
int mfs_solve_multi(mfs_mat A, const mfs_solve_desc* desc, int k, mfs_vec* b, mfs_vec* x,
                    mfs_solve_report* reports);
/* One-shot convenience: create + run + destroy. */
int mfs_solve(mfs_mat A, const mfs_solve_desc* desc, mfs_vec b, mfs_vec x, mfs_solve_report* report);
/* mfs_solve without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code around
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The work vectors, iteration state and reduction partials live in a fixed
 * 64 MiB scratch array of the library that each device holds until the
 * process ends; kernels are launched directly and the host waits on the
 * context stream after each iteration chunk (the same kernels and order as
 * mfs_solve, hence the same iterates).  Methods MFS_PCG_JACOBI and MFS_CG,
 * precisions MFS_PREC_F64 and MFS_PREC_MIXED.  Systems whose layout does
 * not fit the scratch (about 0.9 M unknowns in FP64, 1.3 M mixed) return
 * MFS_ERR_ARG with the size needed.  Every call on one device shares that
 * device's scratch: calls that overlap on one device (two host threads,
 * two contexts) race and must be serialised by the caller. */
int mfs_solve_noalloc(mfs_mat A, const mfs_solve_desc* desc, mfs_vec b, mfs_vec x, mfs_solve_report* report);

/* ---- assembly ---------------------------------------------------------- */
typedef struct {
  int problem;            /* MFS_PROBLEM_SCALAR | MFS_PROBLEM_ELASTIC         */
  int64_t n_nodes;
  int64_t n_tets;
  const double* xyz;      /* 3 per node                                       */
  const int32_t* tets;    /* 4 node ids per tet                               */
  const int32_t* material;/* material id per tet (renumbering), may be NULL   */
  int renumber;           /* 1: FEM.h:1990-2081 node renumbering (elastic)   */
  int convention;         /* MFS_ELEM_REFERENCE | MFS_ELEM_TEXTBOOK           */
  int64_t n_fixed;        /* Dirichlet nodes, in the drivers' tag-list order  */
  const int32_t* fixed_nodes;
  const double* fixed_values; /* 1 (scalar) or 3 (elastic) values per node    */
  int64_t n_tris;         /* Neumann triangles (elastic)                      */
  const int32_t* tris;    /* 3 node ids per triangle                          */
  const double* traction; /* 3 per triangle                                   */
} mfs_fem_desc;

/* Host setup (renumbering FEM.h:1990-2081, portrait FEM_Grid.h:3236-3287 /
 * :665-716, Neumann loads FEM_Grid.h:2622-2661) and device upload.  Creates
 * the matrix (CSR for scalar, BSR3 for elastic) and the right-hand side. */
int mfs_assemble_create(mfs_ctx ctx, const mfs_fem_desc* desc, mfs_asm* out);
/* GPU assembly: batched per-tet geometry, deterministic gather of element
 * blocks in ascending tet order (the summation order of SummPartOfMatrix,
 * CSSD_Matrix.h:307-363, called in element order, FEM.h:2146-2155), then
 * Dirichlet elimination (FEM.h:505-619).  coef: sigma (scalar) or E
 * (elastic) per tet; nu used for elastic. */
int mfs_assemble_run(mfs_asm S, const double* coef, double nu);
int mfs_assemble_destroy(mfs_asm S);
/* Borrowed handles (owned by S): system matrix, right-hand side b, and the
 * initial vector x0 (Dirichlet values on fixed DOFs, 0 elsewhere). */
int mfs_assemble_matrix(mfs_asm S, mfs_mat* A);
int mfs_assemble_rhs(mfs_asm S, mfs_vec* b);
int mfs_assemble_x0(mfs_asm S, mfs_vec* x0);
/* original node id -> DOF node id (identity when renumber == 0) */
int mfs_assemble_node_map(mfs_asm S, int32_t* map);
/* Average device time (ms) of mfs_assemble_run's kernels over reps runs. */
int mfs_assemble_time(mfs_asm S, int reps, double* avg_ms);

/* ---- host setup utilities (CPU only, no device needed) -------------------
 * Synthetic inputs: the reference has no mesh or field generator; meshes it
 * reads come from Gmsh files (Library/Math.h:327-520). */
/* Node / tet counts of a structured nx*ny*nz Freudenthal box. */
int mfs_mesh_box_counts(int nx, int ny, int nz, int64_t* n_nodes, int64_t* n_tets);
/* Nodes i + (nx+1)(j + (ny+1)k); 6 tets per hex (axis permutations xyz,
 * xzy, yxz, yzx, zxy, zyx), hexes x fastest. */
int mfs_mesh_box(int nx, int ny, int nz, double x0, double y0, double z0, double hx, double hy, double hz,
                 double* xyz, int32_t* tets);
/* Nodes on face f (0:x-,1:x+,2:y-,3:y+,4:z-,5:z+) ascending; returns count
 * via *count (pass NULL nodes to count only). */
int mfs_mesh_box_face_nodes(int nx, int ny, int nz, int face, int32_t* nodes, int64_t* count);
/* Boundary triangles of face f (the tet faces lying on it). */
int mfs_mesh_box_face_tris(int nx, int ny, int nz, int face, int32_t* tris, int64_t* count);
/* Per-tet fields from a counter-based generator (splitmix64). */
int mfs_field_inclusions(int64_t n_tets, const int32_t* tets, const double* xyz, int count, double rmin,
                         double rmax, uint64_t seed, double v_bg, double v_in, double* out);
int mfs_field_lognormal(int64_t n_tets, uint64_t seed, uint64_t stream, int64_t index0, double mu, double sigma,
                        double* out);
int mfs_field_grf(int64_t n_nodes, int64_t n_tets, const int32_t* tets, const double* xyz, int modes, double ell,
                  uint64_t seed, double v_lo, double v_hi, double* out, int32_t* material);
/* Mechanical renumbering (FEM.h:1990-2081) on the host. */
int mfs_renumber(int64_t n_nodes, int64_t n_tets, const int32_t* tets, const int32_t* material, int64_t n_fixed,
                 const int32_t* fixed_nodes, int32_t* map);
/* Block portrait (FEM_Grid.h:3236-3287) as sorted CSR including the
 * diagonal.  Two calls: rowptr only (col NULL), then col. */
int mfs_portrait(int64_t n_rows, int64_t n_tets, const int32_t* tets, int32_t* rowptr, int32_t* col);
/* Write a Gmsh 2.2 ASCII mesh in the layout SimpleGrid::ReadFromMSH reads
 * (Library/Math.h:327-520): nodes, then boundary triangles (type 2) with
 * their tag and parent tet, then tets (type 4) with the material in the
 * physical-tag field (the reader's domain id). */
int mfs_write_gmsh(const char* path, int64_t n_nodes, const double* xyz, int64_t n_tets, const int32_t* tets,
                   const int32_t* material, int64_t n_tris, const int32_t* tris, const int32_t* tri_tags);
/* Property file of the Mechanical driver, param_for_solver_%d.txt as
 * MechanicalMain.cpp:859-1201 reads it: log flag, threads, mesh format (10:
 * ReadFromMSH, 9: ReadFromMSH_v2208), mesh path, critical residual, the
 * Dirichlet rows (tag digit; per component a value, or '*' when free), the
 * Neumann rows (tag, value, direction; direction 0 0 0 makes the value a
 * normal pressure), the materials (E, nu per domain) and the result
 * directory.  Tags are single digits in this format. */
typedef struct {
  int tag;
  int fixed[3];     /* 1: component prescribed, 0: free                     */
  double value[3];
} mfs_dirichlet_spec;
typedef struct {
  int tag;
  double value;
  double direction[3];
} mfs_neumann_spec;
typedef struct {
  double E, nu;
} mfs_material_spec;
int mfs_write_solver_params(const char* path, int log_flag, int threads, int mesh_format, const char* mesh_path,
                            double critical_residual, int n_dirichlet, const mfs_dirichlet_spec* dirichlet,
                            int n_neumann, const mfs_neumann_spec* neumann, int n_materials,
                            const mfs_material_spec* materials, const char* result_dir);
/* Read a Gmsh 2.2 ASCII mesh with SimpleGrid::ReadFromMSH's rules
 * (Library/Math.h:327-520): 9-integer type-2 rows are boundary triangles
 * (tag = 4th field; parent tet = 5th field in the "modified" layout, whose
 * header line holds 4 numbers), 10-integer type-4 rows are tets (material =
 * 4th field).  Ids become 0-based.  Two calls: with NULL arrays it returns
 * the counts, then it fills the arrays (sizes must match).  Parent tets
 * absent from the file come from the node-tet adjacency; the reference
 * searches all tets and rewrites the file (:455-511), this reader never
 * writes. */
int mfs_read_gmsh(const char* path, int64_t* n_nodes, double* xyz, int64_t* n_tets, int32_t* tets,
                  int32_t* material, int64_t* n_tris, int32_t* tris, int32_t* tri_tags, int32_t* tri_parent);

/* ---- MsFEM batched local problems ----------------------------------------
 * Replaces Grid_forMech_Poly_Order1::CreateFunctionalSpaces'
 * per-coarse-cell local solves (MsFEM/MsFEM_Grid.h:4795-5760, the volume
 * problems via FEM/FEM.h:2969-3795) and the coarse element matrices
 * (MsFEM_Grid.h:3837-3975, here the exact Galerkin Phi^T K Phi). */
typedef struct {
  int n_cells;
  const int32_t* cell_n;     /* fine hexes per edge of each cell (>= 2)      */
  const double* cell_origin; /* 3 per cell                                   */
  double cell_h;             /* coarse cell edge length                      */
  const double* E;           /* per fine tet, cells concatenated in order     */
  double nu;
  int convention;            /* MFS_ELEM_*                                    */
  int trace;                 /* MFS_TRACE_* (0: Q1 hats)                      */
  int coarse_basis;          /* MFS_BASIS_* (0: per direction)                */
  double trace_tol;          /* reference traces: largest accepted face-solve */
                             /* residual ||r||/||b|| (0: 1e-11)               */
} mfs_batch_desc;

typedef struct {
  int precision;   /* MFS_PREC_F64 (plain) | MFS_PREC_MIXED                   */
  double tol;      /* per local problem: ||r||/||b|| (FP64, true)             */
  int maxit;       /* inner iteration cap per problem                         */
  double inner_tol;/* MIXED: FP32 inner relative tolerance                    */
  int use_graph;   /* capture the persistent launch in a hipGraph             */
  int wg_size;     /* workgroup size (0 = autotuned; 64..256, 512 for
                      variant 3)                                             */
  int use_guess;   /* start from mfs_batch_set_guess values                   */
  int variant;     /* kernel variant: 0 = autotuned/default; rows carried
                      per thread: FP32 inner 1 -> 2 rows, 2 -> 3 rows;
                      FP64 inner 1 -> 1 row, 2 -> 2 rows; with r, d, q
                      kept in registers: 3 -> 1 row (wg_size up to 512),
                      4 -> 2 rows (these need nint <= rows * wg_size for
                      every cell, else MFS_ERR_ARG)                         */
  int fallback_wg_size;  /* MIXED: shape of the FP64 pass over the problems  */
  int fallback_variant;  /* the FP32 launch could not solve (0 = the FP64
                            autotuned shape, else as wg_size / variant)      */
} mfs_batch_solve_desc;

typedef struct {
  double seconds;       /* device time of the whole batch solve               */
  int max_iterations;   /* worst problem (inner iterations)                   */
  double mean_iterations;
  double max_true_relres;
  int n_not_converged;
  int n_fallback;       /* guesses rejected (worse than zero guess)           */
  int n_precision_fallback; /* problems the FP32 inner solve could not finish,
                               re-solved in FP64 (a second persistent launch
                               queued behind the first, reading the list of
                               failed tasks the first one writes)              */
  double kernel_seconds; /* device time of the first persistent launch alone
                            (seconds: both launches)                          */
} mfs_batch_report;

int mfs_batch_create(mfs_ctx ctx, const mfs_batch_desc* desc, mfs_batch* out);
int mfs_batch_destroy(mfs_batch B);
/* Sizes: interior DOFs of cell c, total fine DOFs of cell c. */
int mfs_batch_cell_dofs(mfs_batch B, int cell, int64_t* n_interior, int64_t* n_total);
/* Assemble every cell matrix on the GPU (batched element + gather).  In
 * the reference-trace mode this first runs the batched 1-D edge solves
 * (MsFEM_Grid.h:5026-5175 via FEM/FEM.h:3796-4365) and 2-D face solves
 * (MsFEM_Grid.h:5176-5325 via FEM/FEM.h:4367-5178) that give the boundary
 * data of the volume problems. */
int mfs_batch_assemble(mfs_batch B);
/* Reference-trace mode, after mfs_batch_assemble: largest relative
 * residual ||b - K x|| / ||b|| of the face solves, the number of face
 * problems (12 per face) above trace_tol, and the device time of the edge +
 * face launches (ms).  No reference counterpart (the reference prints its
 * best residual per solve, FEM/FEM.h:4343-4351).  Any pointer may be NULL. */
int mfs_batch_trace_status(mfs_batch B, double* max_face_relres, int* n_face_failed, double* trace_ms);
/* Solve the 24 local problems of every cell in one persistent launch. */
int mfs_batch_solve(mfs_batch B, const mfs_batch_solve_desc* desc, mfs_batch_report* report);
/* Initial guesses for the interior DOFs of every (cell, basis), FP32 device
 * layout [cell][interior dof][24], uploaded from host f64 in that order. */
int mfs_batch_set_guess(mfs_batch B, const double* guess);
/* Download basis of one cell: 24 x n_total (full fine vectors, boundary
 * traces included), basis index 3*coarse_node + direction. */
int mfs_batch_get_basis(mfs_batch B, int cell, double* basis);
/* Coarse element matrices Kc = Phi^T K Phi, 24x24 per cell, all cells. */
int mfs_batch_coarse_matrices(mfs_batch B, double* Kc);
/* Per-problem inner iteration counts, 24 per cell. */
int mfs_batch_iterations(mfs_batch B, int32_t* iters);
/* Per-problem status of the last solve (24 per cell, any pointer may be
 * NULL): inner iterations, true FP64 relative residual, flags (bit 0: not
 * converged, bit 1: initial guess rejected in favour of the zero guess). */
int mfs_batch_status(mfs_batch B, int32_t* iters, double* relres, int32_t* flags);
/* Learned-guess glue (all cells must share one subdivision n; device data):
 * features X[cell][tet] = (ln E - mu) / sigma (FP32, ncells*6n^3);
 * targets T[cell][(v*3+a)*24+q] = basis - hat interpolant (FP32, interior);
 * guess: x = hat interpolant + G (G NULL: the hat interpolant alone). */
int mfs_batch_features(mfs_batch B, double mu, double sigma, mfs_vec X);
int mfs_batch_targets(mfs_batch B, mfs_vec T);
int mfs_batch_guess_from(mfs_batch B, mfs_vec G);
/* Coarse system of the MsFEM solve (MsFEM/MsFEM.h:571-1062 coarse
 * portrait, scatter, elimination): the cells must tile a box on a grid of
 * spacing cell_h; Kc from mfs_batch_coarse_matrices (host), load = 3 per
 * coarse node; fixed coarse nodes with 3 values each.  Creates a BSR3 matrix
 * and the right-hand side / initial vector (owned by the caller). */
int mfs_msfem_coarse_system(mfs_batch B, const double* Kc, const double* load, int64_t n_fixed,
                            const int32_t* fixed_nodes, const double* fixed_values, mfs_mat* A, mfs_vec* b,
                            mfs_vec* x0);
/* Coarse load Phi^T f of a uniform traction t[3] on the top face of the box. */
int mfs_msfem_top_traction(mfs_batch B, const double* t, double* load);
/* Fine displacement of one cell, u = sum Phi u_coarse (MechanicalMain.cpp:
 * 2176-2181), full fine vector 3*(n+1)^3. */
int mfs_msfem_downscale(mfs_batch B, const double* u_coarse, int cell, double* u_fine);
/* MsFEM-vs-FEM error of MechanicalMain.cpp:2112-2221: per component a,
 * sqrt(sum_tets V (u_ms - u_fem)_a^2 (centroid)) / sqrt(sum_tets V
 * u_fem,a^2) (denominator 1 when below 1e-10).  All cells share one n;
 * u_fem is the fine FEM solution on the whole box's fine grid
 * ((ncx*n+1) x (ncy*n+1) x (ncz*n+1) nodes, x fastest, 3 per node). */
int mfs_msfem_error(mfs_batch B, const double* u_coarse, const double* u_fem, double* err3);
/* Autotuner for mfs_batch_solve in desc->precision: times the whole batched
 * solve for every kernel variant (rows carried per thread), workgroup size
 * (64..256, up to 512 for variant 3) and workgroups per CU (as many as the
 * kernel's registers and LDS
 * allow, or one fewer), keeps the fastest for later solves of this
 * precision and returns it as *best_wg = wg*100 + per_cu*10 + variant
 * (0-based; desc->variant takes it plus one) with
 * its time in *best_ms.  No reference counterpart (the reference solves
 * cells one at a time, MsFEM/MsFEM_Grid.h:4795-5760). */
int mfs_batch_autotune(mfs_batch B, const mfs_batch_solve_desc* desc, int* best_wg, double* best_ms);

/* ---- learned initial guess (BF16 MFMA MLP) --------------------------------
 * No reference counterpart: the network predicts local bases as CG initial
 * guesses; the zero-guess solve stays the fallback. */
typedef struct {
  int n_layers;          /* number of linear layers                           */
  const int32_t* dims;   /* n_layers + 1 sizes                                */
  uint64_t seed;
  double lr;             /* Adam learning rate                                */
  double beta1, beta2, eps;
} mfs_mlp_desc;

int mfs_mlp_create(mfs_ctx ctx, const mfs_mlp_desc* desc, mfs_mlp* out);
int mfs_mlp_destroy(mfs_mlp M);
/* Parameters as FP32 master copies: layer l weights [out][in] then bias. */
int mfs_mlp_get_params(mfs_mlp M, int layer, float* W, float* b);
int mfs_mlp_set_params(mfs_mlp M, int layer, const float* W, const float* b);
/* Forward for a batch: X [B][dims0] f32 host, Y [B][dimsL] f32 host. */
int mfs_mlp_forward(mfs_mlp M, int batch, const float* X, float* Y);
/* Loss 0.5*mean((Y-T)^2) and gradients (no update); grads as FP32 host
 * arrays per layer (W then b), for checks. */
int mfs_mlp_loss_grad(mfs_mlp M, int batch, const float* X, const float* T, double* loss, float* const* gW,
                      float* const* gb);
/* Train on device-resident data: X [N][dims0], T [N][dimsL] (host f32,
 * uploaded once), `epochs` passes of minibatches of `batch`; returns final
 * mean loss. */
int mfs_mlp_train(mfs_mlp M, int64_t n_samples, const float* X, const float* T, int batch, int epochs,
                  double* final_loss);
/* Device-resident variants: X [n][dims0], T [n][dimsL], Y [n][dimsL] FP32
 * vectors; minibatches of `batch`, shuffled with `seed`. */
int mfs_mlp_train_vec(mfs_mlp M, int64_t n_samples, mfs_vec X, mfs_vec T, int batch, int epochs, uint64_t seed,
                      double* final_loss);
int mfs_mlp_forward_vec(mfs_mlp M, int64_t n_samples, mfs_vec X, mfs_vec Y, int batch);
/* Time the BF16 MFMA GEMM kernel alone on an M x N x K product (A row-major
 * [M][K], B [N][K], FP32 output; split-K when the output has few tiles),
 * reps launches back to back: average ms and TFLOP/s (2MNK per launch).
 * No reference counterpart (the learned guess is specific to this build). */
int mfs_gemm_bf16_time(mfs_ctx ctx, int M, int N, int K, int reps, double* avg_ms, double* tflops);
/* Measured VALU multiply-add rate (TFLOP/s, 2 flops per FMA) of the
 * device in MFS_F32 (plain, unpacked instructions, as every kernel of this
 * library is compiled) or MFS_F64.  No reference counterpart: a
 * measurement that explains the mixed-precision speedups of the solvers. */
int mfs_valu_fma_rate(mfs_ctx ctx, int dtype, double* tflops);
/* Average device time (ms) and achieved TFLOP/s of the forward+backward GEMMs
 * for one minibatch of `batch`. */
int mfs_mlp_time_step(mfs_mlp M, int batch, int reps, double* avg_ms, double* tflops);

#ifdef __cplusplus
}
#endif
#endif /* MFS_H */
