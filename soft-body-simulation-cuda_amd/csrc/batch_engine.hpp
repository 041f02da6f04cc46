// Batched PD step for many independent bodies on gfx950: every body advances
// with its own Jacobi-PCG global solve (its own alpha, beta and stop test,
// pcgJacobi.cu:134-135 applied per body) inside one persistent launch per
// body-size class, one workgroup per body running the whole step: the
// prologue, the local/global iterations (computeLocal, the right-hand side,
// the PCG solve, the Chebyshev recurrence) and updateVelPos
// (pdSolver.cu:141-208).  The classes launch on their own streams, so a
// captured step is one graph with a branch per class.
//
// Bodies are the connected components of the tet mesh (the reference's
// DataLoader concatenates bodies, dataLoader.cu:292-378; they stay
// independent because no tet spans two).  Internally the vertices are
// renumbered body by body; caller arrays are read and written through that
// permutation at the step's ends.
//
// Precision: each body's vectors are held relative to a local origin (the rest
// position of its first vertex), so FP32 rounds offsets of the body's size
// rather than of its distance from the world origin; dot products accumulate
// in FP64.
//
// Tile choice: the workgroup size of each class (64..1024 threads) comes from
// an autotuner that times one step of the class with every candidate on a
// copy of the state at prepare.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace sbs {

struct BatchBody {
    int verts = 0, tets = 0, iterations = 0, solves = 0, max_per_solve = 0, unconverged = 0;
};

struct BatchClass {
    int min_tets = 0, max_tets = 0, bodies = 0, threads = 0;  // threads per workgroup chosen by the tuner
    int max_verts = 0;                                        // vertices of the class's largest body
    std::vector<std::pair<int, double>> tried;               // (threads, ms of one step) the tuner measured
};

// Host inputs of the batched system, all in the caller's vertex and tet order.
template <class T> struct BatchSystem {
    int nv = 0, nt = 0;
    std::vector<uint32_t> tets;       // 4 per tet
    std::vector<T> dminv, w;          // 9 (row-major) and 1 per tet
    std::vector<T> c0, diagA, mass, dbc;  // per vertex
    std::vector<T> x0;                // 3 per vertex: rest positions (the bodies' local origins)
    std::vector<int> rowptr, col;     // CSR of the scalar matrix, diagonal included
    std::vector<T> val;
};

struct BatchParams {
    int iterations = 10;
    double dt = 1.0 / 60.0, gravity = 9.8, rho = 0.9992;
    double pcg_tol = 1e-6, pcg_floor = 1e-8;
    int pcg_max_iter = 200;
    bool write_x = true;  // false when the collision response decides X
    bool polar = false;   // FP32 local step: rotation by Newton's polar iteration (SVD fallback)
};

template <class T> class BatchEngine {
public:
    explicit BatchEngine(const BatchSystem<T>& sys);
    ~BatchEngine();
    // One step of every body on `stream` (and the class streams joined back
    // into it); X, V, XTilde and DBCX are the caller's arrays.
    void step(T* X, T* V, T* XTilde, const T* DBCX, const BatchParams& p, hipStream_t stream);
    // Chooses each class's workgroup size by timing one step per candidate on
    // copies of X, V, XTilde (the caller's state is not changed).
    void autotune(const T* X, const T* V, const T* DBCX, const BatchParams& p, hipStream_t stream);
    int num_bodies() const { return nb_; }
    const std::vector<BatchClass>& classes() const { return cls_; }
    // Per-body PCG statistics accumulated since the last reset (waits for the GPU).
    std::vector<BatchBody> body_stats();
    void reset_stats();

private:
    void launch_class(int c, T* X, T* V, T* XTilde, const T* DBCX, const BatchParams& p, hipStream_t s, int threads);

    int nv_ = 0, nt_ = 0, nb_ = 0;
    std::vector<int> vptr_h_, tptr_h_;
    size_t lds_p_limit_ = 0;  // largest search direction (bytes) a workgroup keeps in LDS
    DevBuf<int> vptr_, tptr_, eptr_, ew_, body_list_;
    DevBuf<uint32_t> perm_, ecol_, inc_;
    DevBuf<int> inc_ptr_;
    DevBuf<uint4> tets_;
    DevBuf<T> dminv_, w_, eval_, diagA_, c0_, mass_, dbc_, origin_;
    DevBuf<T> sn_, sn_old_, prev_, x_, r_, p_, q_, force_;
    DevBuf<int> stats_;
    std::vector<BatchClass> cls_;
    std::vector<int> cls_first_;      // offset of each class's bodies in body_list_
    std::vector<hipStream_t> streams_;
    std::vector<hipEvent_t> events_;  // fork + one join per class
};

}  // namespace sbs
