// Projective-dynamics engine for gfx950: state, prepared system and the step.
//
// Reference behaviour this engine reproduces:
//   PdSolver::SolverPrepare   src/simulation/solver/projective/pdSolver.cu:40-137
//   PdSolver::SolverStep      src/simulation/solver/projective/pdSolver.cu:141-208
//   PdSolver::Update          src/simulation/solver/projective/pdSolver.cu:210-232
//   FEMSolver ctor            src/simulation/solver/femSolver.cu:7-17
//
// Layout in HBM (one body or many bodies concatenated, like the reference):
//   * vertex vectors are interleaved xyz (3*N_v scalars), the reference's
//     glm::vec3 AoS layout, so a gather of one neighbour is one 12-byte run;
//   * tets are reordered internally along a Morton curve of their rest
//     centroids and grouped into chunks of kChunkTets; a chunk stores the
//     sorted list of the vertices it touches, and each tet its 4 vertices as
//     u16 indices into that list; per-tet constants are SoA in that order
//     (Dm^-1 as 9 planes, w = |V0| mu as one) so every plane is read coalesced;
//   * the local step writes one partial force per (chunk, vertex) pair
//     instead of scattering with global atomics; the consumers of the
//     right-hand side gather a vertex's partials through a vertex->entry
//     list and add c*s_n there;
//   * the system is the scalar matrix A_s = L_s + diag(c) (the reference's
//     3N x 3N matrix is A_s (x) I_3): its diagonal is one N_v array and the
//     off-diagonal part is sliced ELL with 64-row slices (one slice per
//     wavefront, column-major inside the slice).
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "common.hpp"

namespace sbs {

template <class T> class BandCholesky;
template <class T> class DenseRefinedSolver;
template <class T> class CollisionDetector;
template <class T> class BatchEngine;

enum class GlobalSolver : int {
    Jacobi = 0, CuSolverCholesky = 1, EigenCholesky = 2, PCG = 3, DenseCholesky = 4,
    BatchedPCG = 5  // every connected body with its own Jacobi-PCG, one workgroup per body
};

// Mirrors the fields of SolverParams<Scalar> (src/def.h:83-100) used by the
// PD step, plus the settings of the PCG global step.
struct PdParams {
    int num_iterations = 1;
    double dt = 0.001;
    double gravity = 9.8;
    double tol = 1e-2;
    double rho = 0.9992;
    int handle_collision = 0;
    double pcg_tol = 1e-6;      // residual reduction ||D^-1 r_k|| <= pcg_tol ||D^-1 r_0||
    double pcg_floor = 1e-8;    // ... or <= pcg_floor ||D^-1 b||
    int pcg_max_iter = 100;
    int pcg_chebyshev = 0;      // apply the Chebyshev recurrence after each PCG solve
    int pcg_adaptive = 1;       // 1: learned launch caps; 0: every solve checked by the host (exact)
    int dense_precision = 0;    // DenseCholesky: 0 FP32 MFMA, 1 BF16 MFMA inverse fragments
    double dense_tol = 1e-10;   // DenseCholesky: max|D^-1 r| <= tol max|D^-1 b| (FP64 refinement)
    int dense_max_iter = 8;     // DenseCholesky: refinement corrections per solve
    int pcg_halo = 1;           // 1: halo-staged PCG kernels (when the tiles could be built); 0: row gathers
    int use_graph = 0;          // replay a captured HIP graph of the step when its launch sequence is fixed
    int batch_autotune = 1;     // BatchedPCG: time the workgroup sizes of each body-size class at prepare
    int polar_rotation = 0;     // FP32 local step: 0 TR1690 SVD (as the reference), 1 Newton polar iteration
    double mu_n = 0.5;          // fixed-body response: normal restitution (SolverParams::muN)
    double mu_t = 0.5;          // fixed-body response: friction (SolverParams::muT)
};

class FixedBodies;

// Caller-owned device arrays: the reference's SolverData<Scalar> (src/def.h:22-56).
template <class T> struct PdData {
    int num_verts = 0, num_tets = 0, num_dbc = 0;
    T* X = nullptr;        // 3*N_v
    T* V = nullptr;        // 3*N_v
    T* X0 = nullptr;       // 3*N_v, rest positions (Dm^-1, V0 and DBCX source)
    T* XTilde = nullptr;   // 3*N_v
    T* DBCX = nullptr;     // 3*N_v, pin targets (set to X0 at prepare)
    const uint32_t* Tet = nullptr;  // 4*N_t
    const T* mass = nullptr;        // N_v
    const T* mu = nullptr;          // N_t
    const T* DBC = nullptr;         // N_v, 0/1 pin flags
    const T* moreDBC = nullptr;     // N_v, extra soft pins (may be null: all zero)
    T* V0 = nullptr;       // N_t, filled at construction when non-null
    int num_tris = 0;                   // surface, for collision handling
    const uint32_t* Tri = nullptr;      // 3*num_tris
    const uint32_t* TriFathers = nullptr;  // num_tris: body of each triangle
    T* tIs = nullptr;                   // N_v: 1, or 0.5 on hit vertices
    float* Normals = nullptr;           // 3*N_v contact normals
    const FixedBodies* fixed = nullptr; // SolverData::pFixedBodies (def.h:43); null: none
};

struct PerfEntry {
    std::string name;
    double ms = 0.0;
};

// Kernels that can be timed with HIP events on the engine's stream.
enum KernelId : int {
    kKLocal = 0,
    kKPcg,
    kKJacobi,
    kKDirect,
    kKRhs,
    kKPcgIter,  // PCG iteration launches (the stop probe and launches after convergence included), timed per run
    kKCount
};
const char* kernel_name(int id);

struct KernelTime {
    int launches = 0;
    double ms = 0.0;
};

#ifndef SBS_CHUNK_TETS
#define SBS_CHUNK_TETS 256
#endif
constexpr int kChunkTets = SBS_CHUNK_TETS;  // tets per local-step chunk (one workgroup)
constexpr int kMaxSolveLog = 64;  // per-step log of PCG iterations, by solve index
constexpr int kCounterInts = 6 + kMaxSolveLog;  // step counters copied to the host

template <class T> class PdEngine {
public:
    PdEngine(int threads_per_block, const PdData<T>& data);
    ~PdEngine();

    void set_global_solver(GlobalSolver s) { solver_ = s; }
    GlobalSolver global_solver() const { return solver_; }
    void set_perf(bool on) { perf_ = on; }
    void set_stream(hipStream_t s) {
        if (s != stream_) drop_graph();
        stream_ = s;
    }
    long long graph_captures() const { return graph_captures_; }
    long long graph_replays() const { return graph_replays_; }
    size_t partial_entries() const { return nentries_; }
    hipStream_t stream() const { return stream_; }
    const std::vector<PerfEntry>& perf() const { return perf_data_; }
    void reset();

    // One PdSolver::Update (prepare on first call, then the step).
    void update(PdData<T>& data, const PdParams& params);
    void prepare(PdData<T>& data, const PdParams& params);

    // Local step alone on a given iterate, for kernel-level tests:
    // b = c*sn_old + sum_t w_t P_t Dm^-T G^T, then the pinned overwrite.
    void local_only(const T* q, const T* sn_old, bool residual_form, double dt, T* b_out);

    // Scalar linear solve A_s x = b for 3 interleaved RHS with the prepared
    // matrix (direct when factored, else PCG); exposes the Solve(b, x) call.
    int solve(const T* b, T* x, const PdParams& params, bool use_guess);

    // Prepared system, for tests (caller tet order).
    struct Export {
        std::vector<T> c, diagA, diagL, V0, w, dminv;  // dminv row-major per tet
        std::vector<int> rowptr, col;
        std::vector<T> val;
    };
    Export export_system() const;

    int last_iterations() const { return last_iters_; }
    int last_pcg_iterations() const { return last_pcg_iters_; }
    int num_verts() const { return nv_; }
    int num_tets() const { return nt_; }
    size_t nnz_offdiag() const { return nnz_off_; }
    bool prepared() const { return ready_; }
    int pcg_cap() const;  // largest learned PCG launch cap (iterations)
    long long pcg_cap_hits() const { return unconverged_total_; }
    long long pcg_iterations_total() const { return pcg_iters_total_; }
    long long pcg_solves_total() const { return pcg_solves_total_; }
    int halo_groups() const { return halo_groups_; }  // workgroups of the halo tiles (0: not built)
    // Waits for queued work and folds the last step's counters in.
    void sync_counters();
    void set_kernel_timing(unsigned mask);
    KernelTime kernel_time(int id);
    // DenseCholesky: bodies, largest body, bytes of the inverse fragments, and
    // the last solve's corrections and residual ratio (waits for the stream).
    void dense_info(int* bodies, int* max_body, long long* bytes, int* iterations, double* ratio);
    // BatchedPCG: the engine (null before the first batched update)
    BatchEngine<T>* batch() const { return batch_.get(); }

private:
    void build_chunks(const std::vector<double>& X0h);
    void build_structure();
    void build_halo(const std::vector<int>& rowstart, const std::vector<uint64_t>& keys,
                    const std::vector<int>& sptr);
    void assemble(const PdParams& params);
    void launch_local(const T* q, bool residual_form, const int* skip);
    // the local step that also finishes the previous PCG solve (its stop test
    // and the Chebyshev recurrence from sn_in/prev_in into sn_out/prev_out)
    void launch_local_fold(const T* x_int, const T* sn_in, const T* prev_in, T* sn_out, T* prev_out, T omega,
                           int close_L, const PdParams& params);
    size_t local_lds_bytes() const;
    // sn/prev: the Chebyshev state the finish updates (null: sn_, prev_);
    // defer_finish: launch no finish (the next local step does it)
    int launch_pcg(int gathered, const T* b, const T* x0, T* x, bool cheb, T omega, const PdParams& params,
                   const int* skip, int solve_index, bool exact, T* sn = nullptr, T* prev = nullptr,
                   bool defer_finish = false);
    void reset_caps_if_changed(const PdParams& p);
    void enqueue_step(PdData<T>& data, const PdParams& params);
    std::vector<uint64_t> step_key(const PdData<T>& d, const PdParams& p) const;
    void drop_graph();
    bool graph_caps_stale(int n);
    static int cap_margin(int used, int moved, int steady);
    bool is_direct() const;
    void direct_factor(const PdParams& params);
    void direct_solve(const T* b, T* x, const PdParams& params, bool warm = false);
    void tick(int id, bool start, int launches = 1);  // a stop covers `launches` launches
    void flush_timing();

    int tpb_ = 256;
    int nv_ = 0, nt_ = 0;
    PdData<T> data_{};
    GlobalSolver solver_ = GlobalSolver::Jacobi;
    bool polar_ = false;  // local step rotation: Newton polar iteration (FP32) instead of the SVD
    bool ready_ = false;
    bool system_built_ = false;
    bool perf_ = false;
    hipStream_t stream_ = nullptr;
    std::vector<PerfEntry> perf_data_;

    // tets in internal (Morton chunk) order
    std::vector<uint32_t> host_tets_;   // caller order, 4*N_t
    std::vector<int> tet_perm_;         // internal position -> caller tet index
    DevBuf<int> tet_perm_d_;
    DevBuf<uint2> tet_lidx_;            // 4 x u16 chunk-local vertex indices
    int nchunks_ = 0, chunk_max_verts_ = 0;
    DevBuf<int> chunk_vptr_;            // nchunks+1: first entry of each chunk
    DevBuf<uint32_t> chunk_verts_;      // global vertex of each entry
    DevBuf<uint32_t> chunk_rows_;       // internal (halo) row of each entry's vertex, bit 31: its first entry
    DevBuf<int> row_inc_ptr_;           // N_v+1: each row's tet incidences (assembly)
    DevBuf<uint32_t> row_inc_;          // internal tet << 2 | corner, ascending tet per row
    std::vector<uint32_t> host_cverts_;
    std::vector<int> host_vptr_;             // vertex -> its partials [host_vptr_[v], host_vptr_[v+1])
    std::vector<unsigned char> host_first_;  // entry is the first of its vertex (vertex-major order)
    DevBuf<T> sn_alt_, prev_alt_;       // second Chebyshev buffers of the folded finish
    DevBuf<int> vent_ptr_;              // N_v+1: entries of each vertex
    DevBuf<uint32_t> entry_pos_;        // entry -> slot of its partial in vertex-major order
    size_t nentries_ = 0;
    DevBuf<T> partial_f_;               // 3 scalars per entry: local-step partial forces
    DevBuf<uint16_t> inc_codes_;        // per chunk 4*kChunkTets: (tet << 2 | corner) grouped by vertex
    DevBuf<uint16_t> inc_end_;          // per entry: end of its incidences in the chunk's list
    DevBuf<T> dminv_;                   // 9 planes of N_t
    DevBuf<T> w_;                       // |V0| * mu
    DevBuf<T> v0_;                      // V0
    // per vertex
    DevBuf<T> c_, diagL_, diagA_, dinv_;
    DevBuf<T> sn_, sn_old_, prev_, next_, rhs_;
    // matrix: sliced ELL (64-row slices) of the off-diagonal part
    int nslices_ = 0;
    size_t nnz_off_ = 0;
    DevBuf<int> slice_ptr_, slice_w_, rowlen_;
    DevBuf<uint32_t> ecol_;
    DevBuf<T> eval_;
    // PCG workspace and control
    DevBuf<T> cg_rec_[2], cg_p_;        // per-row PCG records (r, D^-1, w, s), double-buffered
    // halo tiles of the halo-staged PCG kernels (rows in Morton order)
    std::vector<uint32_t> vertex_key_;  // Morton code of each vertex's rest position
    int halo_groups_ = 0;
    size_t halo_entries_ = 0;
    DevBuf<int> h_grp_, h_src_;
    DevBuf<uint32_t> h_halo_, h_halo_v_, h_perm_;
    DevBuf<uint2> h_vent_;                   // internal row -> its vertex's partial range
    DevBuf<uint16_t> h_col_;
    DevBuf<T> h_val_, cg_x_;
    DevBuf<T> cg_plane_[2];             // halo PCG: D^-1-scaled r, w, s planes, double-buffered
    DevBuf<double> slots_;              // address-striped reduction slots
    DevBuf<unsigned char> ctl_;         // device control block
    void* host_ctl_ = nullptr;          // pinned copy of the control block (host checks)
    // learned PCG launch caps per solve index, for one set of PCG settings
    // per solve index: launch cap, last count, its last change, how many
    // observations in a row repeated it
    std::vector<int> cap_, seen_, moved_, steady_;
    int cap_max_iter_ = 0, cap_cheb_ = -1;
    double cap_tol_ = -1, cap_floor_ = -1, cap_dt_ = -1;
    long long cap_epoch_ = 0;
    long long step_caps_[2] = {-1, -1};  // cap epoch of each counter slot
    // direct solver (band Cholesky factor, GPU triangular solves) for small bodies
    std::unique_ptr<BandCholesky<T>> direct_;
    std::unique_ptr<DenseRefinedSolver<T>> dense_;   // per-body dense inverses (DenseCholesky)
    std::unique_ptr<BatchEngine<T>> batch_;          // per-body persistent solve (BatchedPCG)
    void build_batch(const PdParams& params);
    int dense_prec_ = -1;
    // collision handling after the step (built on first use)
    std::unique_ptr<CollisionDetector<T>> collision_;
    int last_iters_ = 0, last_pcg_iters_ = 0;
    // pinned counters of the last two steps (slot = step parity), kCounterInts
    // each: pcg iterations, unconverged solves, max per solve, solves, last, pad, log
    int* host_counters_ = nullptr;
    hipEvent_t step_event_[2] = {nullptr, nullptr};
    bool step_pending_[2] = {false, false};
    long long steps_ = 0;
    void fold_slot(int slot);
    long long unconverged_total_ = 0;
    long long pcg_iters_total_ = 0, pcg_solves_total_ = 0;
    // kernel timing
    unsigned timing_mask_ = 0;
    hipGraphExec_t graph_exec_ = nullptr;  // the captured step
    std::vector<uint64_t> graph_key_;
    std::vector<int> graph_caps_;          // PCG launch caps the graph was captured with
    int graph_slack_steps_ = 0;            // consecutive steps its caps sat well above the learned ones
    long long graph_captures_ = 0, graph_replays_ = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tev_[kKCount];
    std::vector<int> tev_n_[kKCount];  // launches between each start/stop pair
    size_t tev_used_[kKCount] = {};
    KernelTime ktime_[kKCount];
};

}  // namespace sbs
