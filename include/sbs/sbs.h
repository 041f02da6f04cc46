/* sbs.h -- C ABI of the MI355X (gfx950) projective-dynamics library.
 *
 * The boundary is the reference's PD solver and linear-solver interface,
 * flattened to plain pointers and sizes so any host language can bind it:
 *
 *   PdSolver(int threadsPerBlock, const SolverData<float>&)
 *        src/simulation/solver/projective/pdSolver.h:18      -> sbs_pd_create
 *   PdSolver::SetGlobalSolver(SolverType)    pdSolver.h:20   -> sbs_pd_set_global_solver
 *   PdSolver::Update(SolverData&, const SolverParams&)
 *        pdSolver.h:21, pdSolver.cu:210-232                  -> sbs_pd_update
 *   PdSolver::Reset()                         pdSolver.h:22  -> sbs_pd_reset
 *   Solver::SetPerf(bool)        src/simulation/solver/solver.h:18 -> sbs_pd_set_perf
 *   Solver::GetPerformanceData() solver.h:19                 -> sbs_pd_perf_count / sbs_pd_perf_entry
 *   LinearSolver<T>::Solve(N, d_b, d_x, ...) with the prefactored PD matrix
 *        src/simulation/solver/linear/linear.h:62, pdSolver.cu:188 -> sbs_pd_solve
 *
 * Data: SbsSolverData mirrors SolverData<Scalar> (src/def.h:22-56) -- raw
 * device pointers, vectors as interleaved xyz (glm::vec3 AoS), tets as
 * u32[4*N_t].  SbsSolverParams mirrors SolverParams<Scalar> (src/def.h:83-100)
 * plus the settings of the PCG global step.  Scalars are float or double as
 * chosen at creation (the reference's PD path is float; double gives the FP64
 * form of the same step).
 *
 * Errors: every entry point returns 0 on success or a negative SBS_E_* code
 * and records a message readable with sbs_last_error() (thread-local).  The
 * reference prints and continues (linear.h:16-49) or exits; this library never
 * exits the process.
 *
 * Threads and streams: a handle is used by one host thread at a time.  Work
 * is queued on the handle's HIP stream (default: the null stream; see
 * sbs_pd_set_stream); sbs_pd_update returns without waiting for the GPU.
 */
#ifndef SBS_SBS_H
#define SBS_SBS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBS_OK 0
#define SBS_E_INVALID (-1)
#define SBS_E_HIP (-2)
#define SBS_E_NOT_READY (-3)
#define SBS_E_FACTOR (-4)
#define SBS_E_OOM (-5)
#define SBS_E_INTERNAL (-6)

/* PdSolver::SolverType (pdSolver.h:14-17) plus the Jacobi-PCG global step. */
#define SBS_SOLVER_JACOBI 0
#define SBS_SOLVER_CUSOLVER_CHOLESKY 1
#define SBS_SOLVER_EIGEN_CHOLESKY 2
#define SBS_SOLVER_PCG 3
/* Per-body dense inverses applied with MFMA + FP64 iterative refinement: the
 * CholeskyDnLinearSolver path (src/simulation/solver/linear/cholesky.cu:26-70,
 * 161-166) for scenes of small bodies (contiguous vertex ranges). */
#define SBS_SOLVER_DENSE_CHOLESKY 4
/* Batched bodies: every connected body of the scene (the bodies the
 * reference's DataLoader concatenates, dataLoader.cu:292-378) is solved on its
 * own -- its own Jacobi-PCG (alpha, beta and stop test per body,
 * pcgJacobi.cu:89-173) + Chebyshev -- with one workgroup per body running the
 * whole step in one persistent launch per body-size class; the classes'
 * workgroup sizes come from an autotuner run at the first batched update
 * (SbsSolverParams.batch_autotune).  Results equal one PdSolver per body with
 * SBS_SOLVER_PCG and pcg_chebyshev.  moreDBC is not taken. */
#define SBS_SOLVER_BATCHED_PCG 5

/* Mirrors SolverData<Scalar> (src/def.h:22-56); pointers are device memory
 * holding float (precision 32) or double (precision 64). */
typedef struct SbsSolverData {
    int num_verts;
    int num_tets;
    int num_dbc;
    void* X;              /* 3*num_verts: positions x_n, written by update   */
    void* V;              /* 3*num_verts: velocities, written by update      */
    void* X0;             /* 3*num_verts: rest positions                      */
    void* XTilde;         /* 3*num_verts: end-of-step positions               */
    void* DBCX;           /* 3*num_verts: pin targets (set from X0 at prepare);
                           * a dragged vertex (moreDBC > 0) is held at its
                           * DBCX entry, which the caller sets between updates */
    const uint32_t* Tet;  /* 4*num_tets, 16-byte aligned                      */
    const void* mass;     /* num_verts                                        */
    const void* mu;       /* num_tets                                         */
    const void* DBC;      /* num_verts: 1 for pinned vertices, else 0         */
    const void* moreDBC;  /* num_verts or NULL: mouse-drag weights (0 = none):
                           * c = (m + w)/h^2, s_n = b/w = DBCX, V = 0       */
    void* V0;             /* num_tets or NULL: rest volumes, written at create */
    /* Surface for collision handling (used when handle_collision is set):   */
    int num_tris;
    const uint32_t* Tri;        /* 3*num_tris                                   */
    const uint32_t* TriFathers; /* num_tris: body of each triangle (dev_TriFathers) */
    void* tIs;                  /* num_verts: 1, or 0.5 on hit vertices (dev_tIs) */
    float* Normals;             /* 3*num_verts: contact normals (dev_Normals)   */
    struct SbsFixed* fixed_bodies; /* NULL or a fixed-body set (SolverData::pFixedBodies) */
} SbsSolverData;

/* Mirrors SolverParams<Scalar> (src/def.h:83-100) for the PD step. */
typedef struct SbsSolverParams {
    int num_iterations;   /* local/global iterations per step                 */
    double dt;
    double gravity;       /* magnitude along -y                               */
    double tol;           /* early exit of the direct/PCG loop on the iterate */
    double rho;           /* Chebyshev spectral-radius estimate               */
    int handle_collision; /* CCD between bodies after the solve + CCD response */
    double pcg_tol;       /* PCG: ||D^-1 r_k|| <= pcg_tol * ||D^-1 r_0||       */
    double pcg_floor;     /*  ... or <= pcg_floor * ||D^-1 b||                 */
    int pcg_max_iter;     /* PCG iteration limit per solve                     */
    int pcg_chebyshev;    /* apply the Chebyshev recurrence after each solve   */
    int pcg_adaptive;     /* 1: learned launch cap (no host sync); 0: exact    */
    int dense_precision;  /* dense solver: 0 FP32 MFMA, 1 BF16 MFMA            */
    double dense_tol;     /* dense solver: max|D^-1 r| <= tol max|D^-1 b|      */
    int dense_max_iter;   /* dense solver: FP64 refinement corrections per solve */
    int pcg_halo;         /* PCG kernels: 1 halo-staged (LDS), 0 per-row gathers */
    double mu_n;          /* fixed-body response: normal restitution (muN)  */
    double mu_t;          /* fixed-body response: friction (muT)            */
    int use_graph;        /* 1: on a non-null stream, capture the step's launches once into a
                             HIP graph and replay it while launch sequence, arrays and
                             parameters stay the same (Jacobi, or adaptive PCG with learned
                             caps; no collision handling, perf or kernel timing).  On MI355X
                             a replay costs what the eager launches cost (64^3 cube: 248 vs
                             256 steps/s; 8^3: 660 vs 654), so 0 is the usual choice */
    int batch_autotune;   /* SBS_SOLVER_BATCHED_PCG: 1 times the workgroup sizes of each
                             body-size class at the first batched update (state unchanged);
                             0 keeps one thread per 16 tets of the class's largest body */
    int polar_rotation;   /* FP32 local step: 0 takes each element's rotation from the TR1690
                             SVD as the reference does (computeLocal, pdUtil.cu:112-125);
                             1 from Newton's polar iteration R <- (R + R^-T)/2, 5 steps,
                             falling back to the SVD for elements it does not settle
                             (det F <= 0, singular values outside ~[0.27, 3.7]): about a
                             third of the SVD's arithmetic, and the exact polar factor
                             where the SVD's 4 Jacobi sweeps stop short of it (strongly
                             sheared elements, up to ~1e-3).  FP64 always uses the SVD. */
} SbsSolverParams;

typedef struct SbsPd SbsPd;

/* Step statistics of the last completed update (see sbs_pd_stats). */
typedef struct SbsPdStats {
    int pcg_iterations_last_step; /* PCG iterations summed over the last step */
    int pcg_cap;                  /* largest learned PCG launch cap (iterations) */
    long long pcg_cap_hits;       /* solves stopped (pcg_max_iter or launch cap) before the tolerance */
    long long nnz_offdiag;        /* off-diagonal nonzeros of the scalar matrix */
    long long pcg_iterations_total; /* PCG iterations over all completed steps  */
    long long pcg_solves_total;     /* PCG solves over all completed steps      */
    int halo_groups;                /* workgroups of the halo-staged PCG (0: mesh not tiled) */
    long long graph_captures;       /* steps captured into a HIP graph          */
    long long graph_replays;        /* steps run by replaying a captured graph  */
    long long partial_entries;      /* (chunk, vertex) pairs of the local step's partial forces */
} SbsPdStats;

const char* sbs_last_error(void);
int sbs_version(void);

/* PdSolver(threadsPerBlock, solverData): computes Dm^-1 and V0 (FEMSolver ctor). */
int sbs_pd_create(int precision, int threads_per_block, const SbsSolverData* data, SbsPd** out);
int sbs_pd_destroy(SbsPd* pd);
int sbs_pd_set_global_solver(SbsPd* pd, int solver_type);
/* One PdSolver::Update: prepares on the first call, then one time step. */
int sbs_pd_update(SbsPd* pd, SbsSolverData* data, const SbsSolverParams* params);
int sbs_pd_reset(SbsPd* pd);
int sbs_pd_set_perf(SbsPd* pd, int on);
int sbs_pd_perf_count(SbsPd* pd, int* count);
int sbs_pd_perf_entry(SbsPd* pd, int index, const char** name, double* ms);
int sbs_pd_set_stream(SbsPd* pd, void* hip_stream);
/* Waits for queued work of the handle and returns its statistics. */
int sbs_pd_stats(SbsPd* pd, SbsPdStats* out);

/* Kernel timing with HIP events on the handle's stream: set a bit mask of
 * SBS_KERNEL_* ids, run updates, then read launches and summed milliseconds
 * (reading waits for the stream).  Setting the mask clears the sums. */
#define SBS_KERNEL_LOCAL_STEP 0        /* computeLocal: one launch per iteration    */
#define SBS_KERNEL_PCG_SOLVE 1         /* a whole PCG global solve (all its launches) */
#define SBS_KERNEL_JACOBI_CHEBYSHEV 2
#define SBS_KERNEL_DIRECT_SOLVE 3
#define SBS_KERNEL_RHS_GATHER 4
#define SBS_KERNEL_PCG_ITERATION 5     /* one launch per PCG iteration (+ stop probe) */
#define SBS_KERNEL_COUNT 6
int sbs_pd_set_kernel_timing(SbsPd* pd, unsigned mask);
int sbs_pd_kernel_time(SbsPd* pd, int kernel_id, const char** name, int* launches, double* ms);

/* Dense solver of the handle (after an update with SBS_SOLVER_DENSE_CHOLESKY):
 * number of bodies, largest body, bytes of the inverse fragments, and the last
 * solve's refinement corrections (-1-max_iter if not converged) and residual
 * ratio.  Waits for the handle's stream. */
int sbs_pd_dense_info(SbsPd* pd, int* bodies, int* max_body, long long* inverse_bytes, int* last_iterations,
                      double* last_ratio);

/* LinearSolver::Solve(3N, d_b, d_x[, d_guess]) against the prepared PD matrix
 * (direct when a Cholesky solver type is set, else PCG); synchronous. */
int sbs_pd_solve(SbsPd* pd, const void* d_b, void* d_x, const void* d_guess, const SbsSolverParams* params,
                 int* iterations);

/* computeLocal + computeDBCLocal alone on an iterate q (device arrays of
 * 3*num_verts): b = c*sn_old + sum of per-tet projections, pins overwritten. */
int sbs_pd_local_step(SbsPd* pd, const void* d_q, const void* d_sn_old, int residual_form, double dt, void* d_b);

/* Prepared system copied to host buffers (any pointer may be NULL):
 * c, diag(A_s), diag(L_s) (num_verts); V0, w=|V0|mu (num_tets); Dm^-1
 * (9*num_tets, row-major per tet); CSR of A_s (rowptr num_verts+1, col/val nnz).
 * sbs_pd_system_nnz gives nnz (diagonal included). */
int sbs_pd_system_nnz(SbsPd* pd, long long* nnz);
int sbs_pd_export_system(SbsPd* pd, void* c, void* diagA, void* diagL, void* V0, void* w, void* dminv, int* rowptr,
                         int* col, void* val);

/* CollisionDetection<Scalar> (src/collision/bvh.h:44-89): LBVH broad phase,
 * vertex-face / edge-edge CCD narrow phase and hit resolution over surface
 * triangles.  The triangle arrays are bound at creation (device, caller-owned,
 * read at every call); positions are passed per call.
 *
 *   DetectCollision(numVerts, numTris, Tri, X, XTilde, TriFathers, tI, nors,
 *                   ignoreSelf)              bvh.h:48, bvh.cu:171-181 -> sbs_collision_detect
 *   ComputeMinStepSize(...)                  bvh.h:50, bvh.cu:183-190 -> sbs_collision_min_step
 *   GetNumQueries / GetQueries               bvh.h:57-63     -> sbs_collision_num_queries / _get_queries
 *
 * tI: num_verts scalars (1, or 0.5 on the vertices of a hit); normals:
 * 3*num_verts floats, written on hit vertices.  Where the reference leaves
 * the result to thread scheduling, the order is fixed: equal times of impact
 * keep the query first in (v1, v2, v3) order, and a vertex touched by several
 * hits takes the normal of the last one in the sorted query list. */
typedef struct SbsCollision SbsCollision;
/* Batched solver of the handle (after an update with SBS_SOLVER_BATCHED_PCG):
 * bodies and size classes; per class info = {min tets, max tets, bodies,
 * threads per workgroup} and the tuner's (threads, ms per step) pairs; per
 * body 6 ints {vertices, tets, PCG iterations, solves, most iterations of one
 * solve, solves stopped by pcg_max_iter} accumulated over the updates. */
int sbs_pd_batch_info(SbsPd* pd, int* num_bodies, int* num_classes);
int sbs_pd_batch_class(SbsPd* pd, int cls, int* info, double* tuned_ms, int max_tried, int* num_tried);
int sbs_pd_batch_bodies(SbsPd* pd, int* stats);

int sbs_collision_create(int precision, int num_verts, int num_tris, const uint32_t* d_tri,
                         const uint32_t* d_tri_fathers, SbsCollision** out);
int sbs_collision_destroy(SbsCollision* c);
int sbs_collision_set_stream(SbsCollision* c, void* hip_stream);
int sbs_collision_detect(SbsCollision* c, const void* d_X, const void* d_XTilde, int ignore_self, void* d_tI,
                         float* d_normals);
int sbs_collision_min_step(SbsCollision* c, const void* d_X, const void* d_XTilde, int ignore_self, double* toi);
/* Queries of the last call: sorted, unique, canonical (type 1 = vertex v0 vs
 * triangle v1 v2 v3; type 2 = edge v0 v1 vs edge v2 v3), 5 u32 each, with
 * their time of impact and normal (host buffers; any may be NULL). */
int sbs_collision_num_queries(SbsCollision* c, long long* n);
int sbs_collision_get_queries(SbsCollision* c, uint32_t* queries, double* toi, double* normals);
int sbs_collision_stats(SbsCollision* c, long long* candidate_pairs, int* hits);
/* SetBuildType / GetBuildType (bvh.h:67-68): BVH<Scalar>::BuildType Serial (0),
 * Atomic (1, the default), Cooperative (2) -- how the internal boxes are
 * fitted (bvh.cu:118-135).  Serial fits one tree level per launch until the
 * root is done; Atomic and Cooperative run the one-launch bottom-up climb.
 * Every type gives the same boxes and so the same queries. */
int sbs_collision_set_build_type(SbsCollision* c, int build_type);
int sbs_collision_get_build_type(SbsCollision* c, int* build_type);
/* UpdateQueries(numVerts, numTris, Tri, X, TriFathers, dhat) (bvh.h:52,
 * broadphase.cu:574-587), precision 64 (the FP64 IPC solver is its caller):
 * the queries of the triangle pairs whose boxes at X, grown by 2 dhat, overlap,
 * with the ignore-self flag of the last detect / min_step call (false before
 * any, as the reference's member), kept while the squared distance is
 * <= dhat, sorted by (squared distance, distance type), ties by vertices.
 * get_proximity copies them (5 u32 each), their distance types (the order
 * of distance_type.cuh's DistanceType) and squared distances. */
int sbs_collision_update_queries(SbsCollision* c, const void* d_X, double dhat, long long* num_queries);
int sbs_collision_get_proximity(SbsCollision* c, uint32_t* queries, int* distance_types, double* distances);
/* CCDKernel (src/simulation/collisionUtil.cu:50-70): vertices with tI < 1 get
 * V = -((XTilde - X).n) n and keep X; the others take X = XTilde. */
int sbs_ccd_response(int precision, int num_verts, void* d_X, const void* d_XTilde, void* d_V, const void* d_tI,
                     const float* d_normals);

/* LinearSolver<T>::Solve(int N, T* d_b, T* d_x, T* A, int nz, int* rowIdx,
 * int* colIdx, T* d_guess) (src/simulation/solver/linear/linear.h:62) with
 * the PCGJacobiSolver algorithm (pcgJacobi.h:7-10, pcgJacobi.cu:89-173), the
 * linear solver IPCSolver uses (ipc.h:16) and calls with its Hessian's COO
 * triplets (ipc.cu:239):
 *   PCGJacobiSolver(N, max_iter = 2000, tolerance = 1e-5) -> sbs_linear_create
 *   Solve(N, d_b, d_x, A, nz, rowIdx, colIdx, d_guess)    -> sbs_linear_solve
 * The triplets (device arrays) may be unsorted and repeat an entry: as
 * sort_coo (linear.cu:7-40) the repeats are summed.  x starts at d_guess (or
 * 0); each iteration stops first when ||r||_2 < tolerance (absolute), then
 * when |r.z| < 1e-15, else takes the Jacobi-preconditioned CG step; at most
 * max_iter iterations, whose count is returned.  Indices outside [0, N) are
 * rejected.  sbs_linear_stats reports the entries left after summing. */
typedef struct SbsLinear SbsLinear;
int sbs_linear_create(int precision, int max_iter, double tolerance, SbsLinear** out);
int sbs_linear_destroy(SbsLinear* h);
int sbs_linear_set_stream(SbsLinear* h, void* hip_stream);
int sbs_linear_solve(SbsLinear* h, int n, void* d_b, void* d_x, const void* d_A, int nz, const int* d_rowIdx,
                     const int* d_colIdx, const void* d_guess, int* iterations);
int sbs_linear_stats(SbsLinear* h, long long* nnz);
/* The same solve without any device allocation: no counterpart in the
 * reference; for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Its
 * workspace is a fixed 32 MiB array per GPU, held until the process ends
 * (about N = 100,000 unknowns with 400,000 triplets in FP64); larger systems
 * are rejected with SBS_E_INVALID.  Every call on one GPU uses that GPU's
 * array, so calls that overlap in time on one GPU race; each call
 * reinitialises what it reads.  Results equal sbs_linear_solve's. */
int sbs_linear_solve_noalloc(SbsLinear* h, int n, void* d_b, void* d_x, const void* d_A, int nz,
                             const int* d_rowIdx, const int* d_colIdx, const void* d_guess, int* iterations);

/* Fixed (kinematic) bodies: FixedBody (src/collision/rigid/rigid.h:14-28) as
 * Sphere (sphere.h), Plane (plane.h) or Cylinder (cylinder.h), held as
 * FixedBodyData (src/simulation/fixedBodyData.h:10-30):
 *   FixedBodyData(int threadsPerBlock, const std::vector<FixedBody*>&)
 *        fixedBodyData.h:15, fixedBodyData.cu:8-61    -> sbs_fixed_create
 *   FixedBodyData::HandleCollisions(X, V, numVerts, muT, muN)
 *        fixedBodyData.h:21, fixedBodyData.cu:136-148 -> sbs_fixed_handle_collisions
 * A body is its type (BodyType order, rigid.h:8-12), its model matrix
 * (FixedBody::m_model, glm column-major: model[4*col + row]) and its radius
 * (Sphere::m_radius; Cylinder::m_radius, the x scale; unused by planes).
 * Derived in float at creation: the plane normal m_floorUp (row 1 of the
 * inverse of the model's 3x3 part, normalised; plane.cpp:9) and the cylinder
 * axis normalize(model * (0,1,0,0)) (fixedBodyData.cu:121).  The response
 * visits spheres, planes, then cylinders, each in creation order
 * (fixedBodyData.cu:139-147).  Attach a set to a PD solve through
 * SbsSolverData.fixed_bodies (SolverData::pFixedBodies, def.h:43): the
 * update then applies it to XTilde and V after X is taken (pdSolver.cu:228-231),
 * with SbsSolverParams.mu_t / mu_n. */
#define SBS_FIXED_SPHERE 0
#define SBS_FIXED_PLANE 1
#define SBS_FIXED_CYLINDER 2
typedef struct SbsFixedBody {
    int type;          /* SBS_FIXED_*                                        */
    float model[16];   /* model matrix, column-major                         */
    float radius;      /* sphere / cylinder radius                           */
} SbsFixedBody;
typedef struct SbsFixed SbsFixed;
int sbs_fixed_create(const SbsFixedBody* bodies, int num_bodies, SbsFixed** out);
int sbs_fixed_destroy(SbsFixed* f);
/* X, V: 3*num_verts device scalars of the precision (32 or 64), updated in
 * place on the stream (NULL: the null stream). */
int sbs_fixed_handle_collisions(SbsFixed* f, int precision, void* d_X, void* d_V, int num_verts, double mu_t,
                                double mu_n, void* hip_stream);
/* Derived parameters, 8 floats per body in creation order: p[3] (translation),
 * u[3] (plane normal / cylinder axis, else 0), radius, 0. */
int sbs_fixed_params(SbsFixed* f, float* out);

/* ------------------------------------------------------- scene loading --
 * DataLoader (src/simulation/dataLoader.h:10-37) for one soft body from
 * TetGen files, on the host:
 *   CollectData(nodeFile, eleFile, faceFile, pos, scale, rot, centralize,
 *               startIndex, attrib)  dataLoader.h:17, dataLoader.cu:204-237 -> sbs_mesh_load_tetgen
 * with loadNodeFile (:130-170: coordinates read as float, optional centring
 * with a y/z swap), loadEleFaceFile (:39-127: the .face triangles, or when
 * there is none the tet faces seen an odd number of times, ascending by
 * sorted vertex triple, with the winding of their last occurrence) and
 * CollectEdges (:172-199), then the model transform T(pos) S(scale) Rx Ry Rz
 * (rotations in degrees) of TransformVertices (utilities.cu:57-64) in the
 * chosen precision.  Indices come back 0-based.  Comment lines ('#') are
 * skipped; an unreadable node or ele file, a malformed record or a vertex
 * index outside the node list is an error (the reference prints and goes
 * on); a face file that is NULL, empty or does not open means "extract". */
typedef struct SbsTetgenOptions {
    int start_index;      /* the files' first vertex index (0 or 1)        */
    int centralize;       /* subtract the centroid, then swap y and z      */
    float pos[3];
    float scale[3];
    float rot[3];         /* degrees about x, then y, then z               */
    int precision;        /* 32: float arithmetic, 64: double              */
} SbsTetgenOptions;
typedef struct SbsMesh SbsMesh;
int sbs_mesh_load_tetgen(const char* node_file, const char* ele_file, const char* face_file,
                         const SbsTetgenOptions* opt, SbsMesh** out);
int sbs_mesh_counts(const SbsMesh* m, int* num_verts, int* num_tets, int* num_tris, int* num_edges,
                    int* tris_from_file);
/* Host buffers of the sizes sbs_mesh_counts gives (3*nv, 4*nt, 3*ntri, 2*nedge); any may be NULL. */
int sbs_mesh_copy(const SbsMesh* m, double* X, uint32_t* tet, uint32_t* tri, uint32_t* edges);
int sbs_mesh_destroy(SbsMesh* m);

/* ---------------------------------------------------------------- IPC --
 * The FP64 incremental-potential-contact solver of the reference:
 *   IPCSolver(int threadsPerBlock, const SolverData<double>&)
 *        src/simulation/solver/IPC/ipc.h:19, ipc.cu:101-117 -> sbs_ipc_create
 *   IPCSolver::Update(SolverData&, const SolverParams&)
 *        ipc.h:21, ipc.cu:128-135 (SolverStep :142-228)    -> sbs_ipc_update
 *   IPCSolver::Reset()            ipc.h:22, ipc.cu:137-141 -> sbs_ipc_reset
 *   IPCSolver::SetLinearSolver(SolverType) ipc.h:23        -> sbs_ipc_set_linear_solver
 * Newton on 1/2|x - x~|_M^2 + h^2 (gravity + corotated elasticity + barrier on
 * the squared distances of the contact queries), CCD-limited line search,
 * stop when |p|_inf / dt < tol.  All arrays are device memory of doubles
 * (positions and velocities interleaved xyz), indices u32. */
typedef struct SbsIpcData {
    int num_verts;
    int num_tets;
    int num_tris;
    double* X;                  /* 3*num_verts: positions, updated by a step     */
    double* V;                  /* 3*num_verts: velocities, updated by a step    */
    double* XTilde;             /* 3*num_verts: x~ of the step (written)         */
    const uint32_t* Tet;        /* 4*num_tets                                    */
    const uint32_t* Tri;        /* 3*num_tris surface triangles                  */
    const uint32_t* TriFathers; /* num_tris: body of each triangle               */
    const double* mass;         /* num_verts                                     */
    const double* mu;           /* num_tets                                      */
    const double* lambda;       /* num_tets                                      */
    const double* DBC;          /* num_verts: nonzero = fixed vertex (DBCIdx)    */
    struct SbsFixed* fixed_bodies; /* NULL or the bodies of the implicit barrier
                                      (SolverData::pFixedBodies, ImplicitBarrierEnergy) */
} SbsIpcData;

/* The IPC fields of SolverParams<double> (src/def.h:83-100) plus the linear solve. */
typedef struct SbsIpcParams {
    double dt;
    double damp;          /* x~ = x + dt v damp                                 */
    double dhat;          /* barrier distance (queries kept while d^2 <= dhat)  */
    double tol;           /* Newton stop: |p|_inf / dt < tol                    */
    int max_iterations;   /* Newton iterations per step (failure beyond)        */
    double lin_tol;       /* Jacobi-PCG: |r| <= lin_tol |g|                     */
    int lin_max_iter;
} SbsIpcParams;

typedef struct SbsIpcStats {
    int newton_iterations;    /* of the last step                              */
    int line_search_halvings; /* of the last step                              */
    int linear_iterations;    /* PCG iterations of the last step's solves
                                 (newton_iterations + 1 solves)                */
    int failed;               /* latched: a step failed (reset clears it)      */
    long long queries;        /* contact queries kept at the last update       */
    double kappa;             /* barrier stiffness of the last step            */
    double alpha0;            /* first line-search step of the last step       */
    double energy;            /* incremental potential at the end of the step  */
} SbsIpcStats;

/* IPCSolver::SolverType (ipc.h:14-17). */
#define SBS_IPC_CUSOLVER_CHOLESKY 0
#define SBS_IPC_ICPCG 1
#define SBS_IPC_PCG_JACOBI 2
#define SBS_IPC_JACOBI 3

typedef struct SbsIpc SbsIpc;
int sbs_ipc_create(const SbsIpcData* data, SbsIpc** out);
int sbs_ipc_destroy(SbsIpc* h);
/* One SolverStep; *converged = 0 when the step failed (then, like the
 * reference's latched `failed`, later updates do nothing until a reset). */
int sbs_ipc_update(SbsIpc* h, SbsIpcData* data, const SbsIpcParams* params, int* converged);
/* The elastic term of the IPC energy: 0 CorotatedEnergy (what IPEnergy's
 * constructor builds, ip.cu:22), 1 NeoHookean08Energy (neohookean08.cu:99-206,
 * the one-line swap there).  Default 0. */
int sbs_ipc_set_elastic(SbsIpc* h, int model);
int sbs_ipc_reset(SbsIpc* h);
/* Linear solver of the Newton system.  Every type that converges to the
 * solution (0, 1, 2) runs Jacobi-PCG to lin_tol; 3 (plain Jacobi iteration)
 * is rejected. */
int sbs_ipc_set_linear_solver(SbsIpc* h, int solver_type);
int sbs_ipc_set_stream(SbsIpc* h, void* hip_stream);
int sbs_ipc_stats(SbsIpc* h, SbsIpcStats* out);
/* Energy (and gradient, host buffer of 3*num_verts, may be NULL) at the host
 * positions X with the queries recomputed there, with the current x~ and kappa
 * (set by the last step or by sbs_ipc_set_state, host x~). */
int sbs_ipc_energy(SbsIpc* h, const double* X, double* energy, double* grad);
int sbs_ipc_set_state(SbsIpc* h, const double* XTilde, double kappa);

/* Device memory and copies for callers without their own allocator. */
int sbs_malloc(void** ptr, size_t bytes);
int sbs_free(void* ptr);
int sbs_memcpy_htod(void* dst, const void* src, size_t bytes);
int sbs_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int sbs_memcpy_dtod(void* dst, const void* src, size_t bytes);
int sbs_memset(void* dst, int value, size_t bytes);
int sbs_device_synchronize(void);
int sbs_device_count(int* count);
/* A non-blocking HIP stream at the default priority, for callers without one
 * (a PD solver replays its step as a captured HIP graph only on a non-null
 * stream, see SbsSolverParams.use_graph). */
int sbs_stream_create(void** stream);
int sbs_stream_destroy(void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SBS_SBS_H */
