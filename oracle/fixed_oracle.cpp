// CPU oracle for the fixed (kinematic) bodies of the PD step.
//
// TEST INFRASTRUCTURE ONLY.  Plain C++17, used by tests/ as the checker; never
// linked into, loaded by, or called from the product library.  Built with
// -ffp-contract=off, so the arithmetic of the response is the IEEE sequence
// written here and the GPU kernel (compiled to the same sequence) is compared
// with it bit for bit when both start from the same body parameters.
//
// What it restates (reference file:line):
//   * body parameters: the translation column of FixedBody::m_model
//     (fixedBodyData.cu:71,95,120); Plane::m_floorUp = normalize(vec3(
//     transpose(inverse(model)) * (0,1,0,0))) (plane.cpp:9, rigid.cpp:5);
//     the cylinder axis normalize(model * (0,1,0,0)) (fixedBodyData.cu:121);
//     radii (sphere.cpp:9, cylinder.cpp:9);
//   * HandleCollisions (fixedBodyData.cu:136-148): spheres (:91-109), then
//     planes (:68-88), then cylinders (:112-134), bodies in creation order,
//     each seeing the vertex as the previous body left it; the shared
//     response v = -muN v_N + max(1 - muT (1 + muN) |v_N| / |v_T|, 0) v_T
//     (0 when |v_T| = 0).
//
// The plane normal is computed here by solving A^T w = e_y with partial
// pivoting in double (w = row 1 of A^-1), then rounded to float and
// normalised in float, an independent route to the library's float
// adjugate; the two agree to float rounding (compared with a tolerance),
// and the bit-exact comparison of the response uses one set of parameters.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

enum { kSphere = 0, kPlane = 1, kCylinder = 2 };

// glm::normalize in float: v * (1 / sqrt(dot(v, v))), dot summed left to right
void normalize_f(const float* v, int n, float* out) {
    float d = 0.f;
    for (int i = 0; i < n; ++i) d = d + v[i] * v[i];
    const float s = 1.f / std::sqrt(d);
    for (int i = 0; i < n; ++i) out[i] = v[i] * s;
}

// row 1 of the inverse of the 3x3 part of a column-major 4x4 (solve A^T w = e_y)
bool plane_normal(const float* m, float* up) {
    double M[3][4];
    for (int r = 0; r < 3; ++r) {          // row r of A^T = column r of A
        for (int c = 0; c < 3; ++c) M[r][c] = m[4 * r + c];
        M[r][3] = r == 1 ? 1.0 : 0.0;
    }
    for (int k = 0; k < 3; ++k) {
        int piv = k;
        for (int r = k + 1; r < 3; ++r)
            if (std::fabs(M[r][k]) > std::fabs(M[piv][k])) piv = r;
        if (M[piv][k] == 0.0) return false;
        for (int c = 0; c < 4; ++c) std::swap(M[k][c], M[piv][c]);
        for (int r = 0; r < 3; ++r) {
            if (r == k) continue;
            const double f = M[r][k] / M[k][k];
            for (int c = k; c < 4; ++c) M[r][c] -= f * M[k][c];
        }
    }
    const float w[3] = {(float)(M[0][3] / M[0][0]), (float)(M[1][3] / M[1][1]), (float)(M[2][3] / M[2][2])};
    normalize_f(w, 3, up);
    return true;
}

template <class T> struct Vec {
    T v[3];
};
template <class T> T dot3(const T* a, const T* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <class T> T len3(const T* a) { return std::sqrt(dot3(a, a)); }

template <class T> void respond(T* vel, const T* n, T mu_t, T mu_n) {
    const T s = dot3(vel, n);
    T vn[3], vt[3];
    for (int k = 0; k < 3; ++k) vn[k] = s * n[k];
    for (int k = 0; k < 3; ++k) vt[k] = vel[k] - vn[k];
    const T mag = len3(vt);
    T a = 0;
    if (mag != (T)0) {
        a = (T)1 - mu_t * ((T)1 + mu_n) * len3(vn) / mag;
        if (a < (T)0) a = 0;
    }
    for (int k = 0; k < 3; ++k) vel[k] = (-mu_n) * vn[k] + a * vt[k];
}

template <class T>
void handle(int nv, T* X, T* V, const int* type, const float* prm, int nb, T mu_t, T mu_n) {
    for (int i = 0; i < nv; ++i) {
        T* x = X + 3 * (size_t)i;
        T* vel = V + 3 * (size_t)i;
        for (int pass : {kSphere, kPlane, kCylinder}) {
            for (int j = 0; j < nb; ++j) {
                if (type[j] != pass) continue;
                const float* q = prm + 8 * (size_t)j;
                const T p[3] = {(T)q[0], (T)q[1], (T)q[2]};
                const T u[3] = {(T)q[3], (T)q[4], (T)q[5]};
                const T r = (T)q[6];
                if (pass == kPlane) {
                    const T w[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
                    const T d = dot3(w, u);
                    if (d < (T)0 && dot3(vel, u) < (T)0) {
                        for (int k = 0; k < 3; ++k) x[k] = x[k] - d * u[k];
                        respond(vel, u, mu_t, mu_n);
                    }
                    continue;
                }
                T n[3];
                const T w[3] = {x[0] - p[0], x[1] - p[1], x[2] - p[2]};
                if (pass == kSphere) {
                    for (int k = 0; k < 3; ++k) n[k] = w[k];
                } else {  // (I - a a^T) w, row k: sum_c (delta_kc - a_k a_c) w_c
                    for (int k = 0; k < 3; ++k) {
                        T acc = 0;
                        for (int c = 0; c < 3; ++c) {
                            const T m = (k == c ? (T)1 : (T)0) - u[k] * u[c];
                            acc = c == 0 ? m * w[c] : acc + m * w[c];
                        }
                        n[k] = acc;
                    }
                }
                const T d = len3(n);
                if (d < r) {
                    const T s = (T)1 / std::sqrt(dot3(n, n));
                    T nn[3];
                    for (int k = 0; k < 3; ++k) nn[k] = s * n[k];
                    for (int k = 0; k < 3; ++k) x[k] = x[k] + (r - d) * nn[k];
                    respond(vel, nn, mu_t, mu_n);
                }
            }
        }
    }
}

}  // namespace

extern "C" {

// 8 floats per body: translation, plane normal / cylinder axis (else 0),
// radius, 0.  Returns 0, or -1 for an unknown type or a singular plane matrix.
int orc_fixed_params(int n, const int* type, const float* models, const float* radius, float* out) {
    for (int j = 0; j < n; ++j) {
        const float* m = models + 16 * (size_t)j;
        float* q = out + 8 * (size_t)j;
        for (int k = 0; k < 8; ++k) q[k] = 0.f;
        q[0] = m[12];
        q[1] = m[13];
        q[2] = m[14];
        q[6] = radius[j];
        if (type[j] == kPlane) {
            if (!plane_normal(m, q + 3)) return -1;
        } else if (type[j] == kCylinder) {
            float a[4];
            normalize_f(m + 4, 4, a);
            q[3] = a[0];
            q[4] = a[1];
            q[5] = a[2];
        } else if (type[j] != kSphere) {
            return -1;
        }
    }
    return 0;
}

void orc_fixed_handle(int precision, int nv, void* X, void* V, int n, const int* type, const float* params,
                      double mu_t, double mu_n) {
    if (precision == 64)
        handle<double>(nv, static_cast<double*>(X), static_cast<double*>(V), type, params, n, mu_t, mu_n);
    else
        handle<float>(nv, static_cast<float*>(X), static_cast<float*>(V), type, params, n, (float)mu_t,
                      (float)mu_n);
}

}  // extern "C"
