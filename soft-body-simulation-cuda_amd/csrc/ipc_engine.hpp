// FP64 incremental-potential-contact (IPC) time step on gfx950.
//
// Reference: IPCSolver (src/simulation/solver/IPC/ipc.h:11-40, ipc.cu:101-258)
// with IPEnergy (src/simulation/energy/ip.cu): Newton on
//   E(x) = 1/2 |x - x~|_M^2 + h^2 (gravity + corotated elasticity + barrier)
// over the squared distances of the contact queries (UpdateQueries), with a
// line search started at 0.95 min(1, 0.9 toi) (CCD over x -> x - p) and halved
// until E does not increase, until |p|_inf / h < tol.
//
// Fixed bodies (ImplicitBarrierEnergy, implicitBarrier.cu) add, per vertex
// and body within dhat, area * dhat * 50 (s - 1) ln s (s = d/dhat, d the signed
// distance to the body) with its gradient and its made-PD 3x3 Hessian, and
// bound the first step of the line search (InitStepSize, :352-407).
//
// Layout: per-tet Dm^-1 (row-major) and V0; per Newton iteration the 12x12
// Hessian of every tet and of every active query (row-major, the query's or
// tet's 4 vertices x 3 coordinates) and the gradient; the linear system
// H p = g is solved by Jacobi-PCG applied element by element (no assembled
// matrix), with fixed degrees of freedom as rows and columns of the identity.
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "collision.hpp"
#include "common.hpp"

namespace sbs {

struct IpcData {
    int num_verts = 0, num_tets = 0, num_tris = 0;
    double* X = nullptr;        // 3*N_v positions (updated by a step)
    double* V = nullptr;        // 3*N_v velocities (updated by a step)
    double* XTilde = nullptr;   // 3*N_v x~ of the step (written)
    const uint32_t* Tet = nullptr;
    const uint32_t* Tri = nullptr;
    const uint32_t* TriFathers = nullptr;
    const double* mass = nullptr;    // N_v
    const double* mu = nullptr;      // N_t
    const double* lambda = nullptr;  // N_t
    const double* DBC = nullptr;     // N_v: nonzero = fixed vertex (DBCIdx of the reference)
    const FixedBodies* fixed_bodies = nullptr;  // SolverData::pFixedBodies: the implicit barrier's bodies
};

enum : int { kElasticCorotated = 0, kElasticNeoHookean = 1 };

// UpdateQueries (broadphase.cu:574-587) on a detector's triangles: the kept
// queries (d <= dhat) on the host, 5 u32 each, with their distance types and
// squared distances, sorted by (d, type).  Returns their number.
long long update_queries_on(CollisionDetector<double>& cd, const double* X, double dhat, bool ignore_self,
                            std::vector<uint32_t>& q5, std::vector<int>& dtype, std::vector<double>& dist);

struct IpcParams {
    double dt = 0.01;
    double damp = 0.999;
    double dhat = 1e-3;
    double tol = 5e-2;        // Newton stop: |p|_inf / dt < tol
    int max_iterations = 100; // Newton iterations per step
    double lin_tol = 1e-13;   // PCG: |r| <= lin_tol |g|
    int lin_max_iter = 20000;
};

struct IpcStats {
    int newton_iterations = 0, line_search_halvings = 0, linear_iterations = 0;
    long long queries = 0, active_queries = 0;
    double kappa = 0.0, alpha0 = 1.0, energy = 0.0;
    int failed = 0;
};

class IpcEngine {
public:
    explicit IpcEngine(const IpcData& data);
    ~IpcEngine();
    // IPCSolver::Update: one SolverStep unless a previous step failed
    // (the latched `failed`); returns false when this step failed.
    bool update(IpcData& data, const IpcParams& params);
    void reset() { failed_ = false; }
    // The elastic term of IPEnergy: CorotatedEnergy (what ip.cu:22 builds) or
    // NeoHookean08Energy (neohookean08.cu, the one-line swap of ip.cu:22).
    void set_elastic(int model) { elastic_ = model; }
    int elastic_model() const { return elastic_; }
    const IpcStats& stats() const { return stats_; }
    void set_stream(hipStream_t s);
    // Energy and gradient at X with the current x~ and kappa, the queries
    // recomputed at X (tests).
    double energy_at(const double* X, double* grad_host);
    void set_state(const double* XTilde, double kappa);

private:
    bool step(IpcData& data, const IpcParams& p);
    void update_queries(const double* X);
    double energy(const double* X);
    void gradient(const double* X, bool hessian, double coef_barrier_kappa);
    void update_kappa(const double* X);
    bool search_direction(const double* X);
    void build_rows(const std::vector<uint32_t>& tet);
    int solve();
    void assemble_contact();
    double reduce_sum(const double* parts, int n);
    double max_abs(const double* v, size_t n);
    double body_step(const double* X, const double* p);
    void body_gradient(const double* X, double coef, bool hessian);
    void elastic(const double* X, double coef, bool hessian);

    int nv_ = 0, nt_ = 0, ntri_ = 0;
    IpcData d_{};
    IpcParams p_{};
    int elastic_ = kElasticCorotated;
    hipStream_t stream_ = nullptr;
    bool failed_ = false;
    bool ignore_self_ = false;  // the detector's flag: UpdateQueries uses the last value set (bvh.h:82)
    double kappa_ = 100.0;
    IpcStats stats_{};
    std::unique_ptr<CollisionDetector<double>> col_;
    DevBuf<double> dm_, v0_;                // Dm^-1 (9 per tet, row-major), V0
    DevBuf<unsigned char> fixed_;
    DevBuf<double> xn_, xtmp_, p_dir_, g_, diag_;
    DevBuf<double> ht_, hq_;                // 144 per tet / per query
    // the tets' Hessian assembled per Newton iteration into 3x3 blocks per
    // (vertex, vertex) pair: rows in vertex order, blocks in column order
    long long nnzb_ = 0;
    DevBuf<long long> row_ptr_, contrib_ptr_;
    DevBuf<int> row_col_, row_of_;
    DevBuf<double> row_val_;                // 9 per block, row-major
    DevBuf<uint32_t> contrib_;              // tet << 4 | a << 2 | b per summed entry
    DevBuf<int> qtype_;                     // distance type per candidate query
    DevBuf<double> qd_;                     // squared distance per candidate query
    size_t qcap_ = 0;
    long long nq_ = 0;
    DevBuf<double> r_, z_, q_, p2_;         // PCG vectors (p double-buffered in x_tmp and p2)
    // the active queries' Hessians assembled per Newton iteration into 3x3
    // blocks per (vertex, vertex) pair: rows crow, columns cbcol, values cval
    size_t ecap_ = 0, ctmp_bytes_ = 0;
    DevBuf<unsigned long long> ckey_, ckey_s_;
    DevBuf<uint32_t> ccode_, ccode_s_;
    DevBuf<int> cflag_, cpos_, chead_, cbrow_, cbcol_, crow_, ccount_;
    DevBuf<double> cval_;
    DevBuf<unsigned char> ctmp_;
    DevBuf<double> parts_, out_;            // reduction partials, results
    DevBuf<double> area_;                   // contact area per vertex (the implicit barrier's weight)
    DevBuf<double> hv_;                     // 9 per vertex: the implicit barrier's Hessian blocks
    bool hv_live_ = false;                  // hv_ holds this Newton iteration's blocks
    double* host_ = nullptr;                // pinned scalars
    DevBuf<int> pcg_ctl_;                   // PCG control: [finished, completed iterations]
    DevBuf<double> pcg_parts_;              // PCG block partials of r.z and r.r
    DevBuf<double> prod_parts_;             // PCG block partials of p.q
    int* host_ctl_ = nullptr;               // its pinned copy
    int pcg_last_ = 0;                      // iterations of the previous solve (when to check first)
};

}  // namespace sbs
