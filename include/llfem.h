/*
 * libllfem -- MI355X (gfx950) operator-assembly library for LowLevelFEM.jl.
 *
 * C ABI only: plain pointers, sizes and POD structs; no C++ or torch types.
 * Every host array is owned by the caller; the library never frees caller
 * memory.  Device memory is owned by the context / mesh handles.  Functions
 * return LLFEM_OK (0) or a negative LLFEM_E* code; llfem_last_error() then
 * holds a message for the calling thread (a Julia shim rethrows it with
 * error(msg), keeping the reference's messages where it has one).
 *
 * Indices at this boundary are int64.  `one_based` selects Julia's 1-based
 * node tags / DoF numbers (dof = pdim*(tag-1)+c, src/linear.jl:309-312);
 * 0 selects 0-based.  Output CSC arrays use the same base.
 *
 * Each entry point names the reference function it replaces.  Reference
 * citations are to LowLevelFEM.jl 1.13.31.
 */
#ifndef LLFEM_H
#define LLFEM_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define LLFEM_API __attribute__((visibility("default")))
#else
#define LLFEM_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct llfem_ctx llfem_ctx;
typedef struct llfem_mesh llfem_mesh;

enum {
  LLFEM_OK = 0,
  LLFEM_EINVAL = -1,   /* bad argument / size mismatch (src/multifield.jl:2784-2792) */
  LLFEM_EHIP = -2,     /* HIP runtime failure */
  LLFEM_ENOMEM = -3,   /* device allocation failed or HBM budget exceeded */
  LLFEM_EGEOM = -4,    /* singular geometry at an integration point (src/multifield.jl:1737-1739) */
  LLFEM_EPATTERN = -5, /* row missing from the CSC pattern (src/multifield.jl:2282) */
  LLFEM_ENOCONV = -6,  /* iterative solver hit maxiter */
  LLFEM_EUNSUP = -7    /* element type / option not supported */
};

enum { LLFEM_OP_STIFFNESS = 0, LLFEM_OP_MASS = 1 };

/* Arithmetic of the element integration.  The global CSC summation is FP64
 * in every mode. */
enum {
  LLFEM_FP64 = 0,               /* FP64 throughout (the parity path) */
  LLFEM_FP32_INT_FP64_SUM = 1,  /* FP32 element integration, FP64 summation */
  LLFEM_BF16X3 = 2              /* high-order B'DB as split BF16 MFMA, FP32 accumulate */
};

enum {
  LLFEM_PCG_JACOBI_FP64 = 0,    /* Jacobi-PCG in FP64 */
  LLFEM_PCG_MIXED_REFINE = 1,   /* FP32 Jacobi-PCG inner solves + FP64 residual refinement */
  LLFEM_PCG_MIXED_RELIABLE = 2, /* FP32 block-Jacobi PCG with FP64 reliable updates (the true
                                   FP64 residual replaces the recursive one every 10x reduction;
                                   converged on ||b - K x|| <= reltol ||b||); llfem_solve only */
  LLFEM_PCG_F32OP_RELIABLE = 3, /* as 2 with FP64 Krylov vectors: only the operator is FP32 */
  LLFEM_PCG_AMG_RELIABLE = 4    /* as 2 with an aggregation-AMG V-cycle as the preconditioner
                                   (rigid-body-mode coarse spaces, Chebyshev smoothing); the
                                   hierarchy is built once per mesh and support and kept */
};

/* One operator request.  The tables are what gmsh hands the reference:
 * getIntegrationPoints / getBasisFunctions(..., "Lagrange"|"GradLagrange")
 * in gmsh's flattened layout (src/linear.jl:235-239). */
typedef struct {
  int32_t op;              /* LLFEM_OP_* */
  int32_t q;               /* number of integration points */
  const double* weights;   /* [q] */
  const double* N_ref;     /* [q][npe]      (mass, traction) */
  const double* dN_ref;    /* [q][npe][3]   (stiffness, geometry) */
  const double* D;         /* 6x6 row-major in SymGrad row order
                              [xx,yy,zz,xy,xz,yz] (src/multifield.jl:635-657),
                              or NULL for Hooke's law from E, nu
                              (constitutive_matrix, src/linear.jl:154-165) */
  double E, nu, rho;
  int32_t lumped;          /* mass: row-sum diagonal (src/linear.jl:1999-2001) */
  int32_t precision;       /* LLFEM_FP64 | LLFEM_FP32_INT_FP64_SUM | LLFEM_BF16X3 */
  const double* element_scale; /* [E] or NULL: element e's integrand is multiplied by
                              element_scale[e] (finite, >= 0).  A problem with several
                              material groups (src/linear.jl:190-330 loops over
                              problem.material) is assembled as one call per group
                              with 1 on the group's elements and 0 elsewhere, and
                              the results added in group order -- the order in which
                              the reference's sparse() combines the groups' triplets.
                              llfem_assemble only (batch and server: NULL). */
  int32_t accumulate;      /* llfem_assemble: 0 replaces the mesh's resident values of
                              this operator, 1 adds this call's values to them (the
                              first material group with 0, the others with 1: the mesh
                              then holds the sum, in group order).  A resident operator
                              assembled with element_scale counts as complete -- usable
                              by llfem_solve / llfem_mesh_values -- once every element
                              has had a non-zero factor in some call since the last
                              accumulate = 0 call.  Batch and server: 0. */
} llfem_op_desc;

typedef struct {
  double kernel_ms;        /* device time of the assembly kernels (HIP events, library stream) */
  double main_kernel_ms;   /* device time of the dominant kernel alone */
  double total_ms;         /* host wall time of the call */
  int64_t nnz;             /* structural non-zeros written */
  int64_t elements;
  double alg_bytes;        /* algorithmic HBM bytes of the call */
  int64_t iterations;      /* solver: total (inner) CG iterations */
  int64_t refinements;     /* solver: FP64 refinement steps */
  double relres;           /* solver: ||(K u - f)_free|| / ||f_free|| */
  double alg_flops;        /* element-integration flops of the call (see llfem_assemble) */
  double setup_ms;         /* solver: preconditioner set-up wall time (AMG: hierarchy build when
                              not cached, plus the per-solve coarse operators) */
  int64_t rollbacks;       /* solver (reliable-update methods): updates whose FP32 correction
                              was not finite, so the previous FP64 solution was kept and the
                              recursion restarted from it */
  int64_t restarts;        /* solver (AMG): reliable updates whose preconditioned residual
                              was recomputed by a V-cycle because the rescaled one was no
                              longer consistent (r.z <= 0 or relative drift > 1/4) */
} llfem_stats;

/* ---- context ---------------------------------------------------------- */
/* One context per host thread; calls on a context are serialised on its
 * HIP stream.  hbm_budget_bytes = 0 means no budget. */
LLFEM_API int llfem_ctx_create(int device, size_t hbm_budget_bytes, llfem_ctx** out);
LLFEM_API int llfem_ctx_destroy(llfem_ctx* ctx);
LLFEM_API int llfem_ctx_synchronize(llfem_ctx* ctx);
LLFEM_API size_t llfem_ctx_bytes_in_use(const llfem_ctx* ctx);
LLFEM_API const char* llfem_last_error(void);

/* ---- mesh --------------------------------------------------------------
 * Replaces the gmsh queries + pattern construction of the reference:
 * dense_node_coordinates (src/multifield.jl:1612-1648), getElementsByType,
 * and build_csc_pattern (src/multifield.jl:2083-2250).  Uploads coordinates
 * (xyz: [N][3], node-tag order) and connectivity ([E][npe], element-major),
 * then builds on the device, once per mesh: the structural CSC pattern
 * (node adjacency expanded to pdim x pdim blocks, columns node-major /
 * component-minor, rows sorted) and the element->nnz map that replaces
 * find_csc_position (src/multifield.jl:2260-2283). */
LLFEM_API int llfem_mesh_create(llfem_ctx* ctx, const double* xyz, int64_t N, int32_t elem_type,
                      int32_t npe, const int64_t* conn, int64_t E, int32_t pdim,
                      int32_t one_based, llfem_mesh** out);
LLFEM_API int llfem_mesh_destroy(llfem_mesh* mesh);

/* Structural pattern of an operator (LLFEM_OP_STIFFNESS: the K pattern of
 * build_csc_pattern; LLFEM_OP_MASS: node adjacency x I_pdim, the pattern
 * massMatrix returns after dropzeros!, src/linear.jl:2002).  Two-phase:
 * pass colptr = rowval = NULL to query nnz. */
LLFEM_API int llfem_mesh_pattern(llfem_mesh* mesh, int32_t op, int32_t one_based, int64_t* nnz,
                       int64_t* colptr, int64_t* rowval);

/* ---- assembly ----------------------------------------------------------
 * LLFEM_OP_STIFFNESS replaces stiffnessMatrix(problem) (src/linear.jl:70-87,
 * stiffnessMatrixSolid :190-330) and the DSL
 * ∫(SymGrad(Pu)⋅D⋅SymGrad(Pu); csc_matrix=K) path (src/multifield.jl:5295,
 * assemble_operator :2630-3280).  LLFEM_OP_MASS replaces
 * massMatrix(problem; lumped) (src/linear.jl:1843-2004).
 * The values stay resident on the device (reused by llfem_solve); when
 * nzval_out != NULL they are also copied to the host, in the order of
 * llfem_mesh_pattern(op) (for lumped mass: the pdim*N diagonal).
 * Stiffness of hex125 (and hex64 with LLFEM_BF16X3, or LLFEM_HIGHORDER=1 in
 * the environment) runs as a per-element Gram product on the matrix cores
 * (FP64 MFMA, or BF16x3 with FP32 accumulation); those need positive
 * quadrature weights and q <= 128.  FP32_INT_FP64_SUM on that path
 * computes in FP64. */
LLFEM_API int llfem_assemble(llfem_mesh* mesh, const llfem_op_desc* desc, double* nzval_out,
                   llfem_stats* stats);

/* ---- load vector -------------------------------------------------------
 * Replaces the traction branch of loadVector(problem, [load(name, fx,fy,fz)])
 * (src/linear.jl:3088-3434): f_a += sum_q N_a t Ja w_q over the boundary
 * elements face_conn [F][nf] of a surface, curve or point group, with
 * Ja = |t1 x t2| for surface elements, |t1| for curve (line) elements and 1
 * for point elements (src/linear.jl:3378-3388); f_out is [pdim*N],
 * zero-initialised here.  face_type is gmsh's element type code. */
LLFEM_API int llfem_load_traction(llfem_mesh* mesh, int32_t face_type, int32_t nf, const int64_t* face_conn,
                        int64_t F, int32_t one_based, int32_t q, const double* weights,
                        const double* N_ref, const double* dN_ref, const double traction[3],
                        double* f_out);

/* Physical coordinates of the face Gauss points, x_q = sum_a N_a(xi_q) x_a
 * (the x, y, z the reference evaluates a function-valued load at,
 * src/linear.jl:3278-3280): xyz_out is [F][q][3].  With
 * llfem_load_traction_gp this replaces loadVector with fx/fy/fz given as
 * functions of (x, y, z) (_loadvec_helper, src/linear.jl:3288-3299): the
 * caller evaluates its functions at these points. */
LLFEM_API int llfem_face_points(llfem_mesh* mesh, int32_t face_type, int32_t nf, const int64_t* face_conn,
                                int64_t F, int32_t one_based, int32_t q, const double* N_ref,
                                double* xyz_out);

/* Traction given per face Gauss point, t_gp [F][q][3]:
 * f_a += sum_q N_a t_q Ja w_q (src/linear.jl:3378-3413), Ja as for
 * llfem_load_traction, summed per node in ascending face order; f_out is
 * [pdim*N], zero-initialised here. */
LLFEM_API int llfem_load_traction_gp(llfem_mesh* mesh, int32_t face_type, int32_t nf, const int64_t* face_conn,
                                     int64_t F, int32_t one_based, int32_t q, const double* weights,
                                     const double* N_ref, const double* dN_ref, const double* t_gp,
                                     double* f_out);

/* Body force on volume elements of the mesh: replaces loadVector(problem,
 * [load(volume group, fx, fy, fz)]) ("Volume -> body force",
 * src/linear.jl:3054-3068, integrated by :3240-3413 with Ja = det J) and the
 * DSL linear form  ∫(Pu ⋅ g; Ω)  (assemble_linear, src/multifield.jl:3496-3748).
 * f_a += sum_q N_a(xi_q) g(x_q) det J_q w_q over the elements elems[0..ne)
 * (0-based indices into the mesh's elements; elems = NULL with ne = 0: all
 * elements), summed per node in ascending element order.  g is the constant
 * force [3] when g_gp is NULL, else g_gp [ne][q][3] gives it per Gauss point
 * (a function of (x, y, z) evaluated by the caller at llfem_volume_points).
 * f_out is [3N], zero-initialised here. */
LLFEM_API int llfem_load_body(llfem_mesh* mesh, int32_t q, const double* weights, const double* N_ref,
                              const double* dN_ref, const int64_t* elems, int64_t ne, const double g[3],
                              const double* g_gp, double* f_out);

/* Physical coordinates of the volume Gauss points of elems[0..ne) (NULL, 0:
 * all elements): xyz_out is [ne][q][3], x_q = sum_a N_a(xi_q) x_a. */
LLFEM_API int llfem_volume_points(llfem_mesh* mesh, int32_t q, const double* N_ref, const int64_t* elems,
                                  int64_t ne, double* xyz_out);

/* Pressure p on a face set of a 3D solid: replaces loadVector(problem,
 * [load(name, p = p)]) (src/linear.jl:3196-3224).  As the reference does it:
 * the unit normal n = t1 x t2 / |t1 x t2| of each face at each of its nodes
 * (normalVector, src/general.jl:4848-4940; dN_nodes [nf][nf][3] are the face
 * basis gradients at the face's own reference nodes), the nodal traction
 * t_b = (sum over the faces at node b of (-n) p_b) / (number of such faces)
 * (elementsToNodes, src/general.jl:5519-5640), interpolated at the Gauss
 * points, t(x_q) = sum_b N_b t_b (_loadvec_helper, src/linear.jl:3033-3044),
 * and f_a += sum_q N_a t(x_q) |t1 x t2| w_q.  p_b = p_nodal[b] ([N], the
 * caller's function at the node coordinates) or the constant p when p_nodal
 * is NULL.  f_out is [3N], zero-initialised here. */
LLFEM_API int llfem_load_pressure(llfem_mesh* mesh, int32_t face_type, int32_t nf, const int64_t* face_conn,
                                  int64_t F, int32_t one_based, int32_t q, const double* weights,
                                  const double* N_ref, const double* dN_ref, const double* dN_nodes, double p,
                                  const double* p_nodal, double* f_out);

/* ---- solve -------------------------------------------------------------
 * Replaces solveDisplacement(K, f; support, iterative=true, reltol, maxiter,
 * preconditioner) (src/linear.jl:3711-3743) on the stiffness most recently
 * assembled on `mesh`: u_fixed on the constrained DoFs, rhs = f - K u_fixed
 * on the free DoFs, K_ff u_free = rhs by PCG (method: LLFEM_PCG_*; the
 * reference's `preconditioner` keyword, passed to cg as Pl at :3736, maps
 * to the method: Jacobi, or LLFEM_PCG_AMG_RELIABLE for the aggregation-AMG
 * V-cycle whose hierarchy is kept on the mesh per support; stats->setup_ms
 * reports its set-up).  fixed[] lists the constrained DoFs (constrainedDoFs,
 * src/general.jl:5289-5322) and u_fixed[] their values (NULL = zeros).
 * u_out is [pdim*N]. */
LLFEM_API int llfem_solve(llfem_mesh* mesh, const double* f, const int64_t* fixed, int64_t nfixed,
                int32_t one_based, const double* u_fixed, int32_t method, double reltol,
                int64_t maxiter, double* u_out, llfem_stats* stats);

/* Same solve for a caller-supplied symmetric CSC matrix (the SystemMatrix
 * the reference passes in): ndof x ndof, colptr [ndof+1], rowval/nzval [nnz]. */
LLFEM_API int llfem_solve_csc(llfem_ctx* ctx, int64_t ndof, const int64_t* colptr, const int64_t* rowval,
                    const double* nzval, int32_t one_based, const double* f,
                    const int64_t* fixed, int64_t nfixed, const double* u_fixed, int32_t method,
                    double reltol, int64_t maxiter, double* u_out, llfem_stats* stats);

/* Allocation-free variant of llfem_solve_csc with LLFEM_PCG_JACOBI_FP64.  It
 * has no counterpart in the reference: it is for callers that cannot
 * allocate device memory during the call (tight serving loops, code that
 * forbids hipMalloc).  Every argument is DEVICE memory the caller owns:
 * zero-based colptr [ndof+1] (int64), rowval [nnz] (int32), nzval [nnz],
 * f [ndof], fixed [ndof] (1 = constrained DoF), u_fixed [ndof] (values on the
 * constrained DoFs; NULL = zeros), u_out [ndof].  The solver's work vectors
 * live in one static device array per GPU (64 MiB, held until the process
 * ends), so ndof is limited to about 1,000,000; larger systems return
 * LLFEM_EINVAL with the limit in llfem_last_error.  Index errors give the
 * same LLFEM_EINVAL as llfem_solve_csc.  Calls on one GPU share that array:
 * within a process they are serialised by a lock; calls from different
 * processes on one GPU each have their own copy.  Results are bit-identical
 * to llfem_solve_csc(..., LLFEM_PCG_JACOBI_FP64, ...) on the same system. */
LLFEM_API int llfem_solve_csc_noalloc(llfem_ctx* ctx, int64_t ndof, const int64_t* colptr,
                                      const int32_t* rowval, const double* nzval, const double* f,
                                      const uint8_t* fixed, const double* u_fixed, double reltol,
                                      int64_t maxiter, double* u_out, llfem_stats* stats);

/* ---- batched assembly ----------------------------------------------------
 * Replaces a host loop of stiffnessMatrix(problem) / massMatrix(problem;
 * lumped) (src/linear.jl:70-87, 1843-2004) over independent problems: all
 * (mesh, operator) pairs are integrated by ONE persistent panel launch
 * (wavefronts take panel work from a device queue), after one batched
 * geometry launch.  meshes[i] / descs[i] for i < n; a mesh may appear twice
 * (stiffness and mass).  All meshes must share one context.  Values stay on
 * the device (llfem_mesh_values copies them out).  Pairs the persistent
 * kernel does not cover (general D, FP32/BF16 variants, npe not in
 * {4,6,8,10}, two-byte element->nnz maps, order >= 3 hexahedra) are
 * assembled by llfem_assemble inside the call.  stats (optional) sums over
 * the batch: kernel_ms = geometry + panel device time, main_kernel_ms = the
 * persistent launch, alg_bytes = sum of 24N + 4 n E + 8 nnz per pair. */
LLFEM_API int llfem_assemble_batch(int32_t n, llfem_mesh* const* meshes, const llfem_op_desc* descs,
                                   llfem_stats* stats);

/* Copy the device-resident values of the last assembly of `op` on `mesh`
 * (order of llfem_mesh_pattern(op); lumped=1: the pdim*N diagonal). */
LLFEM_API int llfem_mesh_values(llfem_mesh* mesh, int32_t op, int32_t lumped, double* nzval_out);

/* ---- autotuning ------------------------------------------------------------
 * Times the panel assembly of `desc` on `sample` over a grid of launch
 * shapes (launch-width multiplier of the record planner x waves per block),
 * then, for the fastest shape, over the number of streams the node-class
 * launches of one assembly are spread over (1-3) x round kernels on/off;
 * keeps the fastest for the sample's element type in the context (meshes
 * created afterwards, and llfem_assemble / llfem_assemble_batch, use it) and
 * re-plans the sample.  Values do not depend on the choice.  No reference
 * counterpart: the reference's per-element loops have no launch shape. */
typedef struct {
  int32_t elem_type;       /* gmsh element type code tuned */
  int32_t mult;            /* chosen launch-width multiplier */
  int32_t waves_per_block; /* chosen waves per block */
  int32_t candidates;      /* configurations timed */
  double best_ms;          /* device time of the chosen configuration */
  double default_ms;       /* device time of the default configuration */
  int32_t streams;         /* chosen stream count for the class launches */
  int32_t round_kernels;   /* 1: step-unrolled round kernels where they exist */
} llfem_tune_result;

LLFEM_API int llfem_autotune(llfem_mesh* sample, const llfem_op_desc* desc, llfem_tune_result* out);

/* ---- serving -------------------------------------------------------------
 * A device-resident loop for assemble-and-solve requests, each the
 * reference's  K = stiffnessMatrix(problem); f = loadVector(problem,
 * [load(face, fx, fy, fz)]); u = solveDisplacement(K, f, support = [...])
 * (src/linear.jl:70-87, 3088-3434, 3711-3743) on its own mesh and material.
 * A worker thread (own context on `device`) admits requests first come first
 * served while fewer than max_active run and their allocations fit the HBM
 * budget (0 = no budget), and advances all admitted solves together, one
 * launch per kernel for the whole batch (continuous batching).  Meshes with
 * their CSC pattern and element->nnz map, face-load weights and
 * constrained-DoF masks stay on the device for later requests with the same
 * content (or the same non-zero mesh_key).  A request's result does not
 * depend on the other requests in flight.
 * Every array a request points to must stay valid until llfem_server_wait
 * returns for it; u_out is written before the request completes. */
typedef struct llfem_server llfem_server;

typedef struct {
  const double* xyz;          /* [N][3] */
  int64_t N;
  int32_t elem_type, npe;     /* gmsh volume element type */
  const int64_t* conn;        /* [E][npe], 0-based */
  int64_t E;
  llfem_op_desc stiffness;    /* volume rule, E, nu (op is taken as stiffness) */
  int32_t face_type, nf;      /* traction face set (gmsh face type), 0-based [F][nf] */
  const int64_t* face_conn;
  int64_t F;
  int32_t face_q;             /* face rule: q, weights [q], N [q][nf], dN [q][nf][3] */
  const double* face_weights;
  const double* face_N_ref;
  const double* face_dN_ref;
  double traction[3];         /* constant traction on the face set */
  const int64_t* fixed;       /* constrained DoFs (0-based), held at 0 */
  int64_t nfixed;
  int32_t method;             /* LLFEM_PCG_MIXED_RELIABLE, _F32OP_RELIABLE or _AMG_RELIABLE */
  double reltol;              /* ||b - K x|| <= reltol ||b|| */
  int64_t maxiter;
  uint64_t mesh_key;          /* 0: key the mesh cache by content */
  double* u_out;              /* [3N] */
} llfem_request;

typedef struct {
  double latency_ms;          /* submit -> result written */
  double queue_ms;            /* submit -> admitted */
  double service_ms;          /* admitted -> result written */
  int64_t iterations;         /* PCG iterations */
  int64_t updates;            /* FP64 reliable updates */
  double relres;              /* ||(K u - f)_free|| / ||f_free|| */
  int32_t mesh_cache_hit;
  int32_t status;             /* LLFEM_OK or the request's error code */
  int64_t rollbacks;          /* as llfem_stats.rollbacks */
  int64_t restarts;           /* as llfem_stats.restarts */
} llfem_request_stats;

LLFEM_API int llfem_server_create(int device, size_t hbm_budget_bytes, int32_t max_active, llfem_server** out);
LLFEM_API int llfem_server_submit(llfem_server* server, const llfem_request* request, int64_t* id);
/* Blocks until request `id` completes; returns its status (message in
 * llfem_last_error on failure).  Each id is waited for once. */
LLFEM_API int llfem_server_wait(llfem_server* server, int64_t id, llfem_request_stats* stats);
/* Finishes the requests already submitted, then frees everything. */
LLFEM_API int llfem_server_destroy(llfem_server* server);

#ifdef __cplusplus
}
#endif
#endif /* LLFEM_H */
