// Device vector math with the reference kernels' operation order (Warp builtins as used in
// src/motion_engine/contact_engine_gpu*.py). Compiled with -ffp-contract=off and IEEE div/sqrt,
// so every expression below rounds exactly as written:
//   dot(a,b)   = a.x*b.x + a.y*b.y + a.z*b.z (left to right)
//   cross(a,b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   M*v        = row i: (M[i][0]*v.x + M[i][1]*v.y) + M[i][2]*v.z
//   M*N        = element ij: ((0 + M[i][0]*N[0][j]) + M[i][1]*N[1][j]) + M[i][2]*N[2][j]
//   quat_rotate(q,x) = x*(2w*w-1) + qv*(2*(qv.x)) + cross(qv,x)*w*2; quat_to_matrix = its columns
//   quat normalize = q * (1/|q|); vector normalize = v / |v|
// Those are the uncontracted forms of the kernels the reference builds with fuse_fp=False (the
// manifold selection and the chunk sweeps of its unfused engines). Every other kernel it builds with
// Warp's default FMA contraction, and the c-prefixed forms below write its fusions out with fmaf:
//   a*b + c*d -> fma(a, b, c*d); s + a*b and a*b + s -> fma(a, b, s); s - a*b -> fma(-a, b, s);
//   a*b - c*d -> fma(a, b, -(c*d)); a product by the literal 1 folds away first; an add of two
//   non-products stays an add. So cdot(a,b) = fma(a.z,b.z, fma(a.x,b.x, a.y*b.y)), a row of M*v is
//   fma(M[i][2],v.z, fma(M[i][0],v.x, M[i][1]*v.y)), and ij of M*N is fma(M[i][2],N[2][j],
//   fma(M[i][1],N[1][j], fma(M[i][0],N[0][j], 0))).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace me {

struct f3 { float x, y, z; };
struct f4 { float x, y, z, w; };
struct m3 { float m[3][3]; };

__host__ __device__ __forceinline__ f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__host__ __device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__host__ __device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__host__ __device__ __forceinline__ f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
__host__ __device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__host__ __device__ __forceinline__ f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__host__ __device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__host__ __device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__host__ __device__ __forceinline__ float length(f3 a) { return sqrtf(dot(a, a)); }
__host__ __device__ __forceinline__ f3 normalize(f3 a) {
    float l = length(a);
    if (l > 0.0f) return a / l;
    return mk(0.0f, 0.0f, 0.0f);
}
// max(a, b) = a > b ? a : b (explicit: same signed-zero and NaN behaviour as the oracle)
__host__ __device__ __forceinline__ float wmax(float a, float b) { return a > b ? a : b; }

__host__ __device__ __forceinline__ f3 mv(const m3 &M, f3 b) {
    return mk(M.m[0][0] * b.x + M.m[0][1] * b.y + M.m[0][2] * b.z,
              M.m[1][0] * b.x + M.m[1][1] * b.y + M.m[1][2] * b.z,
              M.m[2][0] * b.x + M.m[2][1] * b.y + M.m[2][2] * b.z);
}
__host__ __device__ __forceinline__ m3 mm(const m3 &A, const m3 &B) {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float s = 0.0f;
#pragma unroll
            for (int k = 0; k < 3; ++k) s += A.m[i][k] * B.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
__host__ __device__ __forceinline__ m3 transpose(const m3 &A) {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r.m[i][j] = A.m[j][i];
    return r;
}
__host__ __device__ __forceinline__ f3 col(const m3 &R, int i) { return mk(R.m[0][i], R.m[1][i], R.m[2][i]); }

__host__ __device__ __forceinline__ f3 quat_rotate(f4 q, f3 x) {
    float c = 2.0f * q.w * q.w - 1.0f;
    float d = 2.0f * (q.x * x.x + q.y * x.y + q.z * x.z);
    f3 qv = mk(q.x, q.y, q.z);
    return (x * c + qv * d) + (cross(qv, x) * q.w) * 2.0f;
}
__host__ __device__ __forceinline__ m3 quat_to_matrix(f4 q) {
    f3 c1 = quat_rotate(q, mk(1.0f, 0.0f, 0.0f));
    f3 c2 = quat_rotate(q, mk(0.0f, 1.0f, 0.0f));
    f3 c3 = quat_rotate(q, mk(0.0f, 0.0f, 1.0f));
    m3 R;
    R.m[0][0] = c1.x; R.m[1][0] = c1.y; R.m[2][0] = c1.z;
    R.m[0][1] = c2.x; R.m[1][1] = c2.y; R.m[2][1] = c2.z;
    R.m[0][2] = c3.x; R.m[1][2] = c3.y; R.m[2][2] = c3.z;
    return R;
}
__host__ __device__ __forceinline__ f4 qmul(f4 a, f4 b) {
    f4 r;
    r.x = a.w * b.x + b.w * a.x + a.y * b.z - b.y * a.z;
    r.y = a.w * b.y + b.w * a.y + a.z * b.x - b.z * a.x;
    r.z = a.w * b.z + b.w * a.z + a.x * b.y - b.x * a.y;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}
__host__ __device__ __forceinline__ f4 qnormalize(f4 q) {
    float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    f4 r;
    if (l > 0.0f) {
        float inv = 1.0f / l;
        r.x = q.x * inv; r.y = q.y * inv; r.z = q.z * inv; r.w = q.w * inv;
    } else {
        r.x = 0.0f; r.y = 0.0f; r.z = 0.0f; r.w = 1.0f;
    }
    return r;
}

// ── contracted forms (kernels the reference builds with FMA contraction) ──
__host__ __device__ __forceinline__ float cdot(f3 a, f3 b) { return fmaf(a.z, b.z, fmaf(a.x, b.x, a.y * b.y)); }
__host__ __device__ __forceinline__ f3 ccross(f3 a, f3 b) {
    return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}
__host__ __device__ __forceinline__ float clength(f3 a) { return sqrtf(cdot(a, a)); }
// s + a*u, per component
__host__ __device__ __forceinline__ f3 cmadd(float a, f3 u, f3 s) {
    return mk(fmaf(a, u.x, s.x), fmaf(a, u.y, s.y), fmaf(a, u.z, s.z));
}
__host__ __device__ __forceinline__ f3 cnormalize(f3 a) {
    float l = clength(a);
    if (l > 0.0f) return a / l;
    return mk(0.0f, 0.0f, 0.0f);
}
__host__ __device__ __forceinline__ f3 cmv(const m3 &M, f3 b) {
    return mk(fmaf(M.m[0][2], b.z, fmaf(M.m[0][0], b.x, M.m[0][1] * b.y)),
              fmaf(M.m[1][2], b.z, fmaf(M.m[1][0], b.x, M.m[1][1] * b.y)),
              fmaf(M.m[2][2], b.z, fmaf(M.m[2][0], b.x, M.m[2][1] * b.y)));
}
__host__ __device__ __forceinline__ m3 cmm(const m3 &A, const m3 &B) {
    m3 r;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float s = fmaf(A.m[i][0], B.m[0][j], 0.0f);
            s = fmaf(A.m[i][1], B.m[1][j], s);
            r.m[i][j] = fmaf(A.m[i][2], B.m[2][j], s);
        }
    return r;
}
// normalize(aa - dot(aa, n)*n) with aa = e_x, or e_y when |n.x| >= 0.9
__host__ __device__ __forceinline__ f3 cfixed_t1(f3 n) {
    f3 aa = mk(1.0f, 0.0f, 0.0f);
    if (fabsf(n.x) >= 0.9f) aa = mk(0.0f, 1.0f, 0.0f);
    return cnormalize(cmadd(-cdot(aa, n), n, aa));
}
// quat_rotate: c = fma(2w, w, -1), d = 2*cdot(qv, x); x*c + qv*d + (cross(qv,x)*w)*2
__host__ __device__ __forceinline__ f3 cquat_rotate(f4 q, f3 x) {
    float c = fmaf(2.0f * q.w, q.w, -1.0f);
    float d = 2.0f * fmaf(q.z, x.z, fmaf(q.x, x.x, q.y * x.y));
    f3 cr = ccross(mk(q.x, q.y, q.z), x);
    return mk(fmaf(cr.x * q.w, 2.0f, fmaf(x.x, c, q.x * d)), fmaf(cr.y * q.w, 2.0f, fmaf(x.y, c, q.y * d)),
              fmaf(cr.z * q.w, 2.0f, fmaf(x.z, c, q.z * d)));
}
// quat_to_matrix: column k is quat_rotate(q, e_k) with the unit factors folded (d = 2*q_k, the
// diagonal fma(q_k, d, c))
__host__ __device__ __forceinline__ m3 cquat_to_matrix(f4 q) {
    float c = fmaf(2.0f * q.w, q.w, -1.0f);
    float qq[3] = {q.x, q.y, q.z};
    m3 R;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float d = 2.0f * qq[k];
        f3 cr = cross(mk(q.x, q.y, q.z), mk(k == 0 ? 1.0f : 0.0f, k == 1 ? 1.0f : 0.0f, k == 2 ? 1.0f : 0.0f));
        float cj[3] = {cr.x, cr.y, cr.z};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float inner = j == k ? fmaf(qq[j], d, c) : fmaf(qq[j], d, 0.0f * c);
            R.m[j][k] = fmaf(cj[j] * q.w, 2.0f, inner);
        }
    }
    return R;
}
__host__ __device__ __forceinline__ f4 cqmul(f4 a, f4 b) {
    f4 r;
    r.x = fmaf(-b.y, a.z, fmaf(a.y, b.z, fmaf(a.w, b.x, b.w * a.x)));
    r.y = fmaf(-b.z, a.x, fmaf(a.z, b.x, fmaf(a.w, b.y, b.w * a.y)));
    r.z = fmaf(-b.x, a.y, fmaf(a.x, b.y, fmaf(a.w, b.z, b.w * a.z)));
    r.w = fmaf(-a.z, b.z, fmaf(-a.y, b.y, fmaf(a.w, b.w, -(a.x * b.x))));
    return r;
}
__host__ __device__ __forceinline__ f4 cqnormalize(f4 q) {
    float l = sqrtf(fmaf(q.w, q.w, fmaf(q.z, q.z, fmaf(q.x, q.x, q.y * q.y))));
    f4 r;
    if (l > 0.0f) {
        float inv = 1.0f / l;
        r.x = q.x * inv; r.y = q.y * inv; r.z = q.z * inv; r.w = q.w * inv;
    } else {
        r.x = 0.0f; r.y = 0.0f; r.z = 0.0f; r.w = 1.0f;
    }
    return r;
}

// SoA loads / stores
__device__ __forceinline__ f3 ld3(const float *p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void st3(float *p, int i, f3 a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }
__device__ __forceinline__ f4 ld4(const float *p, int i) {
    f4 r; r.x = p[4 * i]; r.y = p[4 * i + 1]; r.z = p[4 * i + 2]; r.w = p[4 * i + 3]; return r;
}
__device__ __forceinline__ m3 ldm(const float *p, int i) {
    m3 r;
#pragma unroll
    for (int k = 0; k < 9; ++k) r.m[k / 3][k % 3] = p[9 * i + k];
    return r;
}
__device__ __forceinline__ void stm(float *p, int i, const m3 &M) {
#pragma unroll
    for (int k = 0; k < 9; ++k) p[9 * i + k] = M.m[k / 3][k % 3];
}

}  // namespace me
