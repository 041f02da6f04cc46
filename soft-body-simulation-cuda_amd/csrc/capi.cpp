// extern "C" entry points declared in include/sbs/sbs.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/sbs/sbs.h"
#include "collision.hpp"
#include "ipc_engine.hpp"
#include "linear_solver.hpp"
#include "direct_solver.hpp"
#include "pd_engine.hpp"
#include "batch_engine.hpp"
#include "scene_loader.hpp"

namespace sbs {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

}  // namespace sbs

struct SbsPd {
    int precision = 32;
    std::unique_ptr<sbs::PdEngine<float>> f32;
    std::unique_ptr<sbs::PdEngine<double>> f64;
};

struct SbsIpc {
    std::unique_ptr<sbs::IpcEngine> e;
};

struct SbsMesh {
    sbs::BodyMesh m;
};

struct SbsFixed {
    std::unique_ptr<sbs::FixedBodies> f;
};

struct SbsCollision {
    int precision = 32;
    std::unique_ptr<sbs::CollisionDetector<float>> f32;
    std::unique_ptr<sbs::CollisionDetector<double>> f64;
    // the ignoreSelfCollision member: false until DetectCollision or
    // ComputeMinStepSize sets it, and UpdateQueries uses its last value (bvh.h:82)
    bool ignore_self = false;
    std::vector<uint32_t> prox_q;  // the last UpdateQueries result
    std::vector<int> prox_type;
    std::vector<double> prox_d;
};

struct SbsLinear {
    int precision = 64;
    hipStream_t stream = nullptr;
    std::unique_ptr<sbs::CooPcgSolver<float>> f32;
    std::unique_ptr<sbs::CooPcgSolver<double>> f64;
};

namespace {

template <class F> int guarded(F&& f) {
    try {
        f();
        sbs::set_last_error("");
        return SBS_OK;
    } catch (const sbs::Error& e) {
        sbs::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        sbs::set_last_error("host allocation failed");
        return SBS_E_OOM;
    } catch (const std::exception& e) {
        sbs::set_last_error(e.what());
        return SBS_E_INTERNAL;
    }
}

template <class T> sbs::PdData<T> to_data(const SbsSolverData* d) {
    sbs::PdData<T> o;
    o.num_verts = d->num_verts;
    o.num_tets = d->num_tets;
    o.num_dbc = d->num_dbc;
    o.X = static_cast<T*>(d->X);
    o.V = static_cast<T*>(d->V);
    o.X0 = static_cast<T*>(d->X0);
    o.XTilde = static_cast<T*>(d->XTilde);
    o.DBCX = static_cast<T*>(d->DBCX);
    o.Tet = d->Tet;
    o.mass = static_cast<const T*>(d->mass);
    o.mu = static_cast<const T*>(d->mu);
    o.DBC = static_cast<const T*>(d->DBC);
    o.moreDBC = static_cast<const T*>(d->moreDBC);
    o.V0 = static_cast<T*>(d->V0);
    o.num_tris = d->num_tris;
    o.Tri = d->Tri;
    o.TriFathers = d->TriFathers;
    o.tIs = static_cast<T*>(d->tIs);
    o.Normals = d->Normals;
    o.fixed = d->fixed_bodies ? d->fixed_bodies->f.get() : nullptr;
    return o;
}

sbs::PdParams to_params(const SbsSolverParams* p) {
    sbs::PdParams o;
    o.num_iterations = p->num_iterations;
    o.dt = p->dt;
    o.gravity = p->gravity;
    o.tol = p->tol;
    o.rho = p->rho;
    o.handle_collision = p->handle_collision;
    o.pcg_tol = p->pcg_tol;
    o.pcg_floor = p->pcg_floor;
    o.pcg_max_iter = p->pcg_max_iter;
    o.pcg_chebyshev = p->pcg_chebyshev;
    o.pcg_adaptive = p->pcg_adaptive;
    o.dense_precision = p->dense_precision;
    o.dense_tol = p->dense_tol;
    o.dense_max_iter = p->dense_max_iter;
    o.pcg_halo = p->pcg_halo;
    o.mu_n = p->mu_n;
    o.mu_t = p->mu_t;
    o.use_graph = p->use_graph;
    o.batch_autotune = p->batch_autotune;
    o.polar_rotation = p->polar_rotation;
    return o;
}

template <class F> void with_detector(SbsCollision* c, F&& f) {
    SBS_REQUIRE(c != nullptr, "null SbsCollision handle");
    if (c->precision == 64) f(*c->f64);
    else f(*c->f32);
}

// Calls f with the engine of the handle's precision.
template <class F> void with_engine(SbsPd* pd, F&& f) {
    SBS_REQUIRE(pd != nullptr, "null SbsPd handle");
    if (pd->precision == 64) f(*pd->f64);
    else f(*pd->f32);
}

}  // namespace

extern "C" {

const char* sbs_last_error(void) { return sbs::last_error(); }
int sbs_version(void) { return 1; }

int sbs_pd_create(int precision, int threads_per_block, const SbsSolverData* data, SbsPd** out) {
    return guarded([&] {
        SBS_REQUIRE(out != nullptr && data != nullptr, "sbs_pd_create: null argument");
        SBS_REQUIRE(precision == 32 || precision == 64, "sbs_pd_create: precision must be 32 or 64");
        SBS_REQUIRE(threads_per_block > 0, "sbs_pd_create: threads_per_block must be positive");
        *out = nullptr;
        auto h = std::make_unique<SbsPd>();
        h->precision = precision;
        if (precision == 64) h->f64 = std::make_unique<sbs::PdEngine<double>>(threads_per_block, to_data<double>(data));
        else h->f32 = std::make_unique<sbs::PdEngine<float>>(threads_per_block, to_data<float>(data));
        *out = h.release();
    });
}

int sbs_pd_destroy(SbsPd* pd) {
    return guarded([&] { delete pd; });
}

int sbs_pd_set_global_solver(SbsPd* pd, int solver_type) {
    return guarded([&] {
        SBS_REQUIRE(solver_type >= 0 && solver_type <= 5, "sbs_pd_set_global_solver: unknown solver type");
        with_engine(pd, [&](auto& e) { e.set_global_solver(static_cast<sbs::GlobalSolver>(solver_type)); });
    });
}

int sbs_pd_update(SbsPd* pd, SbsSolverData* data, const SbsSolverParams* params) {
    return guarded([&] {
        SBS_REQUIRE(data != nullptr && params != nullptr, "sbs_pd_update: null argument");
        const sbs::PdParams p = to_params(params);
        if (pd && pd->precision == 64) {
            auto d = to_data<double>(data);
            pd->f64->update(d, p);
        } else {
            SBS_REQUIRE(pd != nullptr, "null SbsPd handle");
            auto d = to_data<float>(data);
            pd->f32->update(d, p);
        }
    });
}

int sbs_pd_reset(SbsPd* pd) {
    return guarded([&] { with_engine(pd, [&](auto& e) { e.reset(); }); });
}

int sbs_pd_set_perf(SbsPd* pd, int on) {
    return guarded([&] { with_engine(pd, [&](auto& e) { e.set_perf(on != 0); }); });
}

int sbs_pd_perf_count(SbsPd* pd, int* count) {
    return guarded([&] {
        SBS_REQUIRE(count != nullptr, "null count");
        with_engine(pd, [&](auto& e) { *count = (int)e.perf().size(); });
    });
}

int sbs_pd_perf_entry(SbsPd* pd, int index, const char** name, double* ms) {
    return guarded([&] {
        with_engine(pd, [&](auto& e) {
            SBS_REQUIRE(index >= 0 && index < (int)e.perf().size(), "perf index out of range");
            if (name) *name = e.perf()[index].name.c_str();
            if (ms) *ms = e.perf()[index].ms;
        });
    });
}

int sbs_pd_set_stream(SbsPd* pd, void* hip_stream) {
    return guarded([&] { with_engine(pd, [&](auto& e) { e.set_stream(static_cast<hipStream_t>(hip_stream)); }); });
}

int sbs_pd_stats(SbsPd* pd, SbsPdStats* out) {
    return guarded([&] {
        SBS_REQUIRE(out != nullptr, "null stats");
        with_engine(pd, [&](auto& e) {
            e.sync_counters();
            out->pcg_iterations_last_step = e.last_pcg_iterations();
            out->pcg_cap = e.pcg_cap();
            out->pcg_cap_hits = e.pcg_cap_hits();
            out->nnz_offdiag = (long long)e.nnz_offdiag();
            out->pcg_iterations_total = e.pcg_iterations_total();
            out->pcg_solves_total = e.pcg_solves_total();
            out->halo_groups = e.halo_groups();
            out->graph_captures = e.graph_captures();
            out->graph_replays = e.graph_replays();
            out->partial_entries = (long long)e.partial_entries();
        });
    });
}

int sbs_pd_set_kernel_timing(SbsPd* pd, unsigned mask) {
    return guarded([&] { with_engine(pd, [&](auto& e) { e.set_kernel_timing(mask); }); });
}

int sbs_pd_kernel_time(SbsPd* pd, int kernel_id, const char** name, int* launches, double* ms) {
    return guarded([&] {
        with_engine(pd, [&](auto& e) {
            const sbs::KernelTime t = e.kernel_time(kernel_id);
            if (name) *name = sbs::kernel_name(kernel_id);
            if (launches) *launches = t.launches;
            if (ms) *ms = t.ms;
        });
    });
}

int sbs_pd_solve(SbsPd* pd, const void* d_b, void* d_x, const void* d_guess, const SbsSolverParams* params,
                 int* iterations) {
    return guarded([&] {
        SBS_REQUIRE(d_b && d_x && params, "sbs_pd_solve: null argument");
        const sbs::PdParams p = to_params(params);
        with_engine(pd, [&](auto& e) {
            using T = std::remove_reference_t<decltype(e)>;
            using S = std::conditional_t<std::is_same_v<T, sbs::PdEngine<double>>, double, float>;
            const size_t bytes = 3 * (size_t)e.num_verts() * sizeof(S);
            if (d_guess && d_guess != d_x)
                SBS_HIP(hipMemcpyAsync(d_x, d_guess, bytes, hipMemcpyDeviceToDevice, e.stream()));
            const int it = e.solve(static_cast<const S*>(d_b), static_cast<S*>(d_x), p, d_guess != nullptr);
            if (iterations) *iterations = it;
        });
    });
}

int sbs_pd_local_step(SbsPd* pd, const void* d_q, const void* d_sn_old, int residual_form, double dt, void* d_b) {
    return guarded([&] {
        SBS_REQUIRE(d_q && d_sn_old && d_b, "sbs_pd_local_step: null argument");
        with_engine(pd, [&](auto& e) {
            using T = std::remove_reference_t<decltype(e)>;
            using S = std::conditional_t<std::is_same_v<T, sbs::PdEngine<double>>, double, float>;
            e.local_only(static_cast<const S*>(d_q), static_cast<const S*>(d_sn_old), residual_form != 0, dt,
                         static_cast<S*>(d_b));
            SBS_HIP(hipStreamSynchronize(e.stream()));
        });
    });
}

int sbs_pd_system_nnz(SbsPd* pd, long long* nnz) {
    return guarded([&] {
        SBS_REQUIRE(nnz != nullptr, "null nnz");
        with_engine(pd, [&](auto& e) {
            SBS_REQUIRE(e.prepared(), "system not prepared (call sbs_pd_update first)");
            *nnz = (long long)e.nnz_offdiag() + e.num_verts();
        });
    });
}

int sbs_pd_export_system(SbsPd* pd, void* c, void* diagA, void* diagL, void* V0, void* w, void* dminv, int* rowptr,
                         int* col, void* val) {
    return guarded([&] {
        with_engine(pd, [&](auto& e) {
            auto ex = e.export_system();
            using S = typename decltype(ex.c)::value_type;
            auto put = [](void* dst, const std::vector<S>& src) {
                if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(S));
            };
            put(V0, ex.V0);
            put(w, ex.w);
            put(dminv, ex.dminv);  // row-major per tet, caller tet order
            if (e.prepared()) {
                put(c, ex.c);
                put(diagA, ex.diagA);
                put(diagL, ex.diagL);
                if (rowptr) std::memcpy(rowptr, ex.rowptr.data(), ex.rowptr.size() * sizeof(int));
                if (col) std::memcpy(col, ex.col.data(), ex.col.size() * sizeof(int));
                put(val, ex.val);
            } else {
                SBS_REQUIRE(!(c || diagA || diagL || rowptr || col || val), "system not prepared (call sbs_pd_update first)");
            }
        });
    });
}

int sbs_malloc(void** ptr, size_t bytes) {
    return guarded([&] {
        SBS_REQUIRE(ptr != nullptr, "null pointer");
        *ptr = nullptr;
        SBS_HIP(hipMalloc(ptr, bytes));
    });
}
int sbs_free(void* ptr) {
    return guarded([&] { if (ptr) SBS_HIP(hipFree(ptr)); });
}
int sbs_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return guarded([&] { SBS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int sbs_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return guarded([&] { SBS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int sbs_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    return guarded([&] { SBS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); });
}
int sbs_memset(void* dst, int value, size_t bytes) {
    return guarded([&] { SBS_HIP(hipMemset(dst, value, bytes)); });
}
int sbs_device_synchronize(void) {
    return guarded([&] { SBS_HIP(hipDeviceSynchronize()); });
}
int sbs_stream_create(void** stream) {
    return guarded([&] {
        SBS_REQUIRE(stream != nullptr, "sbs_stream_create: null argument");
        hipStream_t s = nullptr;
        SBS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
    });
}
int sbs_stream_destroy(void* stream) {
    return guarded([&] { if (stream) SBS_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream))); });
}
int sbs_pd_dense_info(SbsPd* pd, int* bodies, int* max_body, long long* inverse_bytes, int* last_iterations,
                      double* last_ratio) {
    return guarded(
        [&] { with_engine(pd, [&](auto& e) { e.dense_info(bodies, max_body, inverse_bytes, last_iterations, last_ratio); }); });
}

int sbs_pd_batch_info(SbsPd* pd, int* num_bodies, int* num_classes) {
    return guarded([&] {
        with_engine(pd, [&](auto& e) {
            auto* b = e.batch();
            SBS_REQUIRE(b != nullptr, "sbs_pd_batch_info: no batched update has run on this handle");
            if (num_bodies) *num_bodies = b->num_bodies();
            if (num_classes) *num_classes = (int)b->classes().size();
        });
    });
}

int sbs_pd_batch_class(SbsPd* pd, int cls, int* info, double* tuned_ms, int max_tried, int* num_tried) {
    return guarded([&] {
        with_engine(pd, [&](auto& e) {
            auto* b = e.batch();
            SBS_REQUIRE(b != nullptr, "sbs_pd_batch_class: no batched update has run on this handle");
            SBS_REQUIRE(cls >= 0 && cls < (int)b->classes().size(), "sbs_pd_batch_class: class out of range");
            const auto& c = b->classes()[cls];
            if (info) {
                info[0] = c.min_tets;
                info[1] = c.max_tets;
                info[2] = c.bodies;
                info[3] = c.threads;
            }
            const int n = (int)c.tried.size();
            if (num_tried) *num_tried = n;
            for (int k = 0; tuned_ms && k < std::min(n, max_tried); ++k) {
                tuned_ms[2 * k] = c.tried[k].first;
                tuned_ms[2 * k + 1] = c.tried[k].second;
            }
        });
    });
}

int sbs_pd_batch_bodies(SbsPd* pd, int* stats) {
    return guarded([&] {
        SBS_REQUIRE(stats != nullptr, "sbs_pd_batch_bodies: null output");
        with_engine(pd, [&](auto& e) {
            auto* b = e.batch();
            SBS_REQUIRE(b != nullptr, "sbs_pd_batch_bodies: no batched update has run on this handle");
            const auto v = b->body_stats();
            for (size_t i = 0; i < v.size(); ++i) {
                int* o = stats + 6 * i;
                o[0] = v[i].verts; o[1] = v[i].tets; o[2] = v[i].iterations;
                o[3] = v[i].solves; o[4] = v[i].max_per_solve; o[5] = v[i].unconverged;
            }
        });
    });
}

int sbs_collision_create(int precision, int num_verts, int num_tris, const uint32_t* d_tri,
                         const uint32_t* d_tri_fathers, SbsCollision** out) {
    return guarded([&] {
        SBS_REQUIRE(out != nullptr, "sbs_collision_create: null argument");
        SBS_REQUIRE(precision == 32 || precision == 64, "sbs_collision_create: precision must be 32 or 64");
        *out = nullptr;
        auto h = std::make_unique<SbsCollision>();
        h->precision = precision;
        if (precision == 64)
            h->f64 = std::make_unique<sbs::CollisionDetector<double>>(num_verts, num_tris, d_tri, d_tri_fathers, nullptr);
        else
            h->f32 = std::make_unique<sbs::CollisionDetector<float>>(num_verts, num_tris, d_tri, d_tri_fathers, nullptr);
        *out = h.release();
    });
}

int sbs_collision_destroy(SbsCollision* c) {
    return guarded([&] { delete c; });
}

int sbs_collision_set_stream(SbsCollision* c, void* hip_stream) {
    return guarded([&] { with_detector(c, [&](auto& d) { d.set_stream(static_cast<hipStream_t>(hip_stream)); }); });
}

int sbs_collision_detect(SbsCollision* c, const void* d_X, const void* d_XTilde, int ignore_self, void* d_tI,
                         float* d_normals) {
    return guarded([&] {
        if (c) c->ignore_self = ignore_self != 0;
        with_detector(c, [&](auto& d) {
            using T = std::remove_reference_t<decltype(d)>;
            using S = std::conditional_t<std::is_same_v<T, sbs::CollisionDetector<double>>, double, float>;
            d.detect(static_cast<const S*>(d_X), static_cast<const S*>(d_XTilde), ignore_self != 0,
                     static_cast<S*>(d_tI), d_normals);
        });
    });
}

int sbs_collision_min_step(SbsCollision* c, const void* d_X, const void* d_XTilde, int ignore_self, double* toi) {
    return guarded([&] {
        SBS_REQUIRE(toi != nullptr, "sbs_collision_min_step: null output");
        if (c) c->ignore_self = ignore_self != 0;
        with_detector(c, [&](auto& d) {
            using T = std::remove_reference_t<decltype(d)>;
            using S = std::conditional_t<std::is_same_v<T, sbs::CollisionDetector<double>>, double, float>;
            *toi = d.min_step(static_cast<const S*>(d_X), static_cast<const S*>(d_XTilde), ignore_self != 0);
        });
    });
}

int sbs_collision_num_queries(SbsCollision* c, long long* n) {
    return guarded([&] {
        SBS_REQUIRE(n != nullptr, "null output");
        with_detector(c, [&](auto& d) { *n = d.num_queries(); });
    });
}

int sbs_collision_get_queries(SbsCollision* c, uint32_t* queries, double* toi, double* normals) {
    return guarded([&] { with_detector(c, [&](auto& d) { d.get_queries(queries, toi, normals); }); });
}

int sbs_collision_set_build_type(SbsCollision* c, int build_type) {
    return guarded([&] { with_detector(c, [&](auto& d) { d.set_build_type(build_type); }); });
}

int sbs_collision_get_build_type(SbsCollision* c, int* build_type) {
    return guarded([&] {
        SBS_REQUIRE(build_type != nullptr, "sbs_collision_get_build_type: null output");
        with_detector(c, [&](auto& d) { *build_type = d.build_type(); });
    });
}

int sbs_collision_update_queries(SbsCollision* c, const void* d_X, double dhat, long long* num_queries) {
    return guarded([&] {
        SBS_REQUIRE(c != nullptr, "null SbsCollision handle");
        SBS_REQUIRE(c->precision == 64, "sbs_collision_update_queries: the reference calls UpdateQueries from the "
                                        "FP64 IPC solver only; create the detector with precision 64");
        SBS_REQUIRE(d_X != nullptr, "sbs_collision_update_queries: null X");
        SBS_REQUIRE(dhat > 0, "sbs_collision_update_queries: dhat must be positive");
        const long long n = sbs::update_queries_on(*c->f64, static_cast<const double*>(d_X), dhat, c->ignore_self,
                                                   c->prox_q, c->prox_type, c->prox_d);
        if (num_queries) *num_queries = n;
    });
}

int sbs_collision_get_proximity(SbsCollision* c, uint32_t* queries, int* distance_types, double* distances) {
    return guarded([&] {
        SBS_REQUIRE(c != nullptr, "null SbsCollision handle");
        const size_t n = c->prox_d.size();
        if (queries) std::copy(c->prox_q.begin(), c->prox_q.end(), queries);
        if (distance_types) std::copy(c->prox_type.begin(), c->prox_type.begin() + n, distance_types);
        if (distances) std::copy(c->prox_d.begin(), c->prox_d.begin() + n, distances);
    });
}

int sbs_linear_create(int precision, int max_iter, double tolerance, SbsLinear** out) {
    return guarded([&] {
        SBS_REQUIRE(out != nullptr, "sbs_linear_create: null output");
        SBS_REQUIRE(precision == 32 || precision == 64, "sbs_linear_create: precision must be 32 or 64");
        SBS_REQUIRE(max_iter >= 0 && tolerance >= 0, "sbs_linear_create: negative limit");
        auto h = std::make_unique<SbsLinear>();
        h->precision = precision;
        if (precision == 64) h->f64 = std::make_unique<sbs::CooPcgSolver<double>>(max_iter, tolerance);
        else h->f32 = std::make_unique<sbs::CooPcgSolver<float>>(max_iter, tolerance);
        *out = h.release();
    });
}

int sbs_linear_destroy(SbsLinear* h) {
    return guarded([&] { delete h; });
}

int sbs_linear_set_stream(SbsLinear* h, void* hip_stream) {
    return guarded([&] {
        SBS_REQUIRE(h != nullptr, "null SbsLinear handle");
        h->stream = static_cast<hipStream_t>(hip_stream);
    });
}

int sbs_linear_solve(SbsLinear* h, int n, void* d_b, void* d_x, const void* d_A, int nz, const int* d_rowIdx,
                     const int* d_colIdx, const void* d_guess, int* iterations) {
    return guarded([&] {
        SBS_REQUIRE(h != nullptr, "null SbsLinear handle");
        int it = 0;
        if (h->precision == 64)
            it = h->f64->solve(n, static_cast<const double*>(d_b), static_cast<double*>(d_x),
                               static_cast<const double*>(d_A), nz, d_rowIdx, d_colIdx,
                               static_cast<const double*>(d_guess), h->stream);
        else
            it = h->f32->solve(n, static_cast<const float*>(d_b), static_cast<float*>(d_x),
                               static_cast<const float*>(d_A), nz, d_rowIdx, d_colIdx,
                               static_cast<const float*>(d_guess), h->stream);
        if (iterations) *iterations = it;
    });
}

int sbs_linear_solve_noalloc(SbsLinear* h, int n, void* d_b, void* d_x, const void* d_A, int nz,
                             const int* d_rowIdx, const int* d_colIdx, const void* d_guess, int* iterations) {
    return guarded([&] {
        SBS_REQUIRE(h != nullptr, "null SbsLinear handle");
        int it = 0;
        if (h->precision == 64)
            it = h->f64->solve_no_alloc(n, static_cast<const double*>(d_b), static_cast<double*>(d_x),
                                        static_cast<const double*>(d_A), nz, d_rowIdx, d_colIdx,
                                        static_cast<const double*>(d_guess), h->stream);
        else
            it = h->f32->solve_no_alloc(n, static_cast<const float*>(d_b), static_cast<float*>(d_x),
                                        static_cast<const float*>(d_A), nz, d_rowIdx, d_colIdx,
                                        static_cast<const float*>(d_guess), h->stream);
        if (iterations) *iterations = it;
    });
}

int sbs_linear_stats(SbsLinear* h, long long* nnz) {
    return guarded([&] {
        SBS_REQUIRE(h != nullptr, "null SbsLinear handle");
        if (nnz) *nnz = h->precision == 64 ? h->f64->last_nnz() : h->f32->last_nnz();
    });
}

int sbs_collision_stats(SbsCollision* c, long long* candidate_pairs, int* hits) {
    return guarded([&] {
        with_detector(c, [&](auto& d) {
            if (candidate_pairs) *candidate_pairs = d.num_pairs();
            if (hits) *hits = d.num_hits();
        });
    });
}

int sbs_ccd_response(int precision, int num_verts, void* d_X, const void* d_XTilde, void* d_V, const void* d_tI,
                     const float* d_normals) {
    return guarded([&] {
        SBS_REQUIRE(num_verts >= 0, "sbs_ccd_response: negative size");
        SBS_REQUIRE(num_verts == 0 || (d_X && d_XTilde && d_V && d_tI && d_normals), "sbs_ccd_response: null array");
        if (num_verts == 0) return;
        if (precision == 64)
            sbs::ccd_response<double>(num_verts, static_cast<double*>(d_X), static_cast<const double*>(d_XTilde),
                                      static_cast<double*>(d_V), static_cast<const double*>(d_tI), d_normals, nullptr);
        else
            sbs::ccd_response<float>(num_verts, static_cast<float*>(d_X), static_cast<const float*>(d_XTilde),
                                     static_cast<float*>(d_V), static_cast<const float*>(d_tI), d_normals, nullptr);
        SBS_HIP(hipStreamSynchronize(nullptr));
    });
}

int sbs_device_count(int* count) {
    return guarded([&] {
        SBS_REQUIRE(count != nullptr, "null count");
        *count = 0;
        SBS_HIP(hipGetDeviceCount(count));
    });
}

int sbs_mesh_load_tetgen(const char* node_file, const char* ele_file, const char* face_file,
                         const SbsTetgenOptions* opt, SbsMesh** out) {
    return guarded([&] {
        SBS_REQUIRE(out != nullptr && node_file != nullptr && ele_file != nullptr && opt != nullptr,
                    "sbs_mesh_load_tetgen: null argument");
        *out = nullptr;
        sbs::TetgenOptions o;
        o.start_index = opt->start_index;
        o.centralize = opt->centralize != 0;
        for (int k = 0; k < 3; ++k) {
            o.pos[k] = opt->pos[k];
            o.scale[k] = opt->scale[k];
            o.rot[k] = opt->rot[k];
        }
        o.precision = opt->precision;
        auto h = std::make_unique<SbsMesh>();
        try {
            h->m = sbs::load_tetgen(node_file, ele_file, face_file ? face_file : "", o);
        } catch (const std::invalid_argument& e) {
            throw sbs::Error(SBS_E_INVALID, std::string("sbs_mesh_load_tetgen: ") + e.what());
        }
        *out = h.release();
    });
}

int sbs_mesh_counts(const SbsMesh* m, int* num_verts, int* num_tets, int* num_tris, int* num_edges,
                    int* tris_from_file) {
    return guarded([&] {
        SBS_REQUIRE(m != nullptr, "null SbsMesh handle");
        if (num_verts) *num_verts = (int)(m->m.X.size() / 3);
        if (num_tets) *num_tets = (int)(m->m.tet.size() / 4);
        if (num_tris) *num_tris = (int)(m->m.tri.size() / 3);
        if (num_edges) *num_edges = (int)(m->m.edges.size() / 2);
        if (tris_from_file) *tris_from_file = m->m.tri_from_file ? 1 : 0;
    });
}

int sbs_mesh_copy(const SbsMesh* m, double* X, uint32_t* tet, uint32_t* tri, uint32_t* edges) {
    return guarded([&] {
        SBS_REQUIRE(m != nullptr, "null SbsMesh handle");
        if (X) std::copy(m->m.X.begin(), m->m.X.end(), X);
        if (tet) std::copy(m->m.tet.begin(), m->m.tet.end(), tet);
        if (tri) std::copy(m->m.tri.begin(), m->m.tri.end(), tri);
        if (edges) std::copy(m->m.edges.begin(), m->m.edges.end(), edges);
    });
}

int sbs_mesh_destroy(SbsMesh* m) {
    return guarded([&] { delete m; });
}

int sbs_fixed_create(const SbsFixedBody* bodies, int num_bodies, SbsFixed** out) {
    return guarded([&] {
        SBS_REQUIRE(out != nullptr, "sbs_fixed_create: null argument");
        SBS_REQUIRE(num_bodies >= 0 && (num_bodies == 0 || bodies != nullptr), "sbs_fixed_create: bad body list");
        *out = nullptr;
        std::vector<sbs::FixedBodySpec> specs(num_bodies);
        for (int k = 0; k < num_bodies; ++k) {
            specs[k].type = bodies[k].type;
            std::memcpy(specs[k].model, bodies[k].model, sizeof(specs[k].model));
            specs[k].radius = bodies[k].radius;
        }
        auto h = std::make_unique<SbsFixed>();
        h->f = std::make_unique<sbs::FixedBodies>(specs);
        *out = h.release();
    });
}

int sbs_fixed_destroy(SbsFixed* f) {
    return guarded([&] { delete f; });
}

int sbs_fixed_handle_collisions(SbsFixed* f, int precision, void* d_X, void* d_V, int num_verts, double mu_t,
                                double mu_n, void* hip_stream) {
    return guarded([&] {
        SBS_REQUIRE(f != nullptr, "null SbsFixed handle");
        SBS_REQUIRE(precision == 32 || precision == 64, "sbs_fixed_handle_collisions: precision must be 32 or 64");
        const hipStream_t s = static_cast<hipStream_t>(hip_stream);
        if (precision == 64)
            f->f->handle(static_cast<double*>(d_X), static_cast<double*>(d_V), num_verts, mu_t, mu_n, s);
        else
            f->f->handle(static_cast<float*>(d_X), static_cast<float*>(d_V), num_verts, (float)mu_t, (float)mu_n, s);
    });
}

int sbs_fixed_params(SbsFixed* f, float* out) {
    return guarded([&] {
        SBS_REQUIRE(f != nullptr && out != nullptr, "sbs_fixed_params: null argument");
        const auto& p = f->f->params();
        std::copy(p.begin(), p.end(), out);
    });
}

}  // extern "C"

namespace {
sbs::IpcData to_ipc(const SbsIpcData* d) {
    sbs::IpcData o;
    o.num_verts = d->num_verts;
    o.num_tets = d->num_tets;
    o.num_tris = d->num_tris;
    o.X = d->X;
    o.V = d->V;
    o.XTilde = d->XTilde;
    o.Tet = d->Tet;
    o.Tri = d->Tri;
    o.TriFathers = d->TriFathers;
    o.mass = d->mass;
    o.mu = d->mu;
    o.lambda = d->lambda;
    o.DBC = d->DBC;
    o.fixed_bodies = d->fixed_bodies ? d->fixed_bodies->f.get() : nullptr;
    return o;
}
}  // namespace

int sbs_ipc_create(const SbsIpcData* data, SbsIpc** out) {
    return guarded([&] {
        SBS_REQUIRE(data != nullptr && out != nullptr, "sbs_ipc_create: null argument");
        *out = nullptr;
        auto h = std::make_unique<SbsIpc>();
        h->e = std::make_unique<sbs::IpcEngine>(to_ipc(data));
        *out = h.release();
    });
}

int sbs_ipc_destroy(SbsIpc* h) {
    return guarded([&] { delete h; });
}

int sbs_ipc_update(SbsIpc* h, SbsIpcData* data, const SbsIpcParams* params, int* converged) {
    return guarded([&] {
        SBS_REQUIRE(h && data && params, "sbs_ipc_update: null argument");
        sbs::IpcParams p;
        p.dt = params->dt;
        p.damp = params->damp;
        p.dhat = params->dhat;
        p.tol = params->tol;
        p.max_iterations = params->max_iterations;
        p.lin_tol = params->lin_tol;
        p.lin_max_iter = params->lin_max_iter;
        sbs::IpcData d = to_ipc(data);
        const bool ok = h->e->update(d, p);
        if (converged) *converged = ok ? 1 : 0;
    });
}

int sbs_ipc_set_elastic(SbsIpc* h, int model) {
    return guarded([&] {
        SBS_REQUIRE(h, "sbs_ipc_set_elastic: null handle");
        SBS_REQUIRE(model == sbs::kElasticCorotated || model == sbs::kElasticNeoHookean,
                    "sbs_ipc_set_elastic: model must be 0 (corotated) or 1 (Neo-Hookean)");
        h->e->set_elastic(model);
    });
}

int sbs_ipc_reset(SbsIpc* h) {
    return guarded([&] {
        SBS_REQUIRE(h, "sbs_ipc_reset: null handle");
        h->e->reset();
    });
}

int sbs_ipc_set_linear_solver(SbsIpc* h, int solver_type) {
    return guarded([&] {
        SBS_REQUIRE(h, "sbs_ipc_set_linear_solver: null handle");
        SBS_REQUIRE(solver_type >= 0 && solver_type <= 2,
                    "sbs_ipc_set_linear_solver: types 0..2 run Jacobi-PCG; the plain Jacobi iteration is not provided");
    });
}

int sbs_ipc_set_stream(SbsIpc* h, void* hip_stream) {
    return guarded([&] {
        SBS_REQUIRE(h, "sbs_ipc_set_stream: null handle");
        h->e->set_stream(static_cast<hipStream_t>(hip_stream));
    });
}

int sbs_ipc_stats(SbsIpc* h, SbsIpcStats* out) {
    return guarded([&] {
        SBS_REQUIRE(h && out, "sbs_ipc_stats: null argument");
        const sbs::IpcStats& s = h->e->stats();
        out->newton_iterations = s.newton_iterations;
        out->line_search_halvings = s.line_search_halvings;
        out->linear_iterations = s.linear_iterations;
        out->failed = s.failed;
        out->queries = s.queries;
        out->kappa = s.kappa;
        out->alpha0 = s.alpha0;
        out->energy = s.energy;
    });
}

int sbs_ipc_energy(SbsIpc* h, const double* X, double* energy, double* grad) {
    return guarded([&] {
        SBS_REQUIRE(h && X && energy, "sbs_ipc_energy: null argument");
        *energy = h->e->energy_at(X, grad);
    });
}

int sbs_ipc_set_state(SbsIpc* h, const double* XTilde, double kappa) {
    return guarded([&] {
        SBS_REQUIRE(h && XTilde, "sbs_ipc_set_state: null argument");
        h->e->set_state(XTilde, kappa);
    });
}
