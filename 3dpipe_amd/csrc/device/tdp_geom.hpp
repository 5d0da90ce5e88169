// Device geometry for the refinement path: the triangle-triangle distance of
// the reference (PQP-derived TriDist) evaluated exactly in a chosen precision,
// and the voxel-pair bounds. Compiled with -ffp-contract=off so that the
// float instantiation reproduces the reference's FP32 arithmetic bit for bit
// (same operation order as src/geometry/computing_cpu.cpp:60-439 and
// src/include/geometry.h:24-98; IEEE division and square root).
// The TriDist algorithm (SegPoints, the edge-edge slab tests, the vertex-face
// tests) is from PQP, Copyright 1999 The University of North Carolina at
// Chapel Hill (E. Larsen et al.), as the reference's computing_cpu.cpp:1-37
// states; the steps are restated here, the code is this project's.
#pragma once

#include <hip/hip_runtime.h>

namespace tdp {

template <typename R>
struct Tri {
    R v[9];  // v0.xyz v1.xyz v2.xyz
};

template <typename R>
__device__ __forceinline__ R dsqrt(R x);
template <>
__device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_sqrtf(x); }
template <>
__device__ __forceinline__ double dsqrt<double>(double x) { return __builtin_sqrt(x); }

// Arithmetic policies for tri_dist: Exact keeps IEEE division and square
// root (the reference's results bit for bit); Fast uses the hardware
// reciprocal and square root (about 1 ulp each), for FP32 estimates whose
// error is bounded separately.
struct Exact {
    template <typename R>
    static __device__ __forceinline__ R div(R a, R b) { return a / b; }
    template <typename R>
    static __device__ __forceinline__ R sqrt(R a) { return dsqrt<R>(a); }
};
struct Fast {
    static __device__ __forceinline__ float div(float a, float b) {
        return a * __builtin_amdgcn_rcpf(b);
    }
    static __device__ __forceinline__ float sqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
};

#define TDP_SUB(r, a, b) { (r)[0] = (a)[0] - (b)[0]; (r)[1] = (a)[1] - (b)[1]; (r)[2] = (a)[2] - (b)[2]; }
#define TDP_ADD(r, a, b) { (r)[0] = (a)[0] + (b)[0]; (r)[1] = (a)[1] + (b)[1]; (r)[2] = (a)[2] + (b)[2]; }
#define TDP_ADDS(r, a, b, s) { (r)[0] = (a)[0] + (b)[0] * (s); (r)[1] = (a)[1] + (b)[1] * (s); (r)[2] = (a)[2] + (b)[2] * (s); }
#define TDP_CROSS(r, a, b) { (r)[0] = (a)[1] * (b)[2] - (a)[2] * (b)[1]; (r)[1] = (a)[2] * (b)[0] - (a)[0] * (b)[2]; (r)[2] = (a)[0] * (b)[1] - (a)[1] * (b)[0]; }
#define TDP_DOT(a, b) ((a)[0] * (b)[0] + (a)[1] * (b)[1] + (a)[2] * (b)[2])
#define TDP_COPY(r, a) { (r)[0] = (a)[0]; (r)[1] = (a)[1]; (r)[2] = (a)[2]; }

// closest points on segments P + tA and Q + uB; only the separating vector
// VEC and the closest points X, Y are produced (SegPoints)
template <typename R, typename Ar = Exact>
__device__ __forceinline__ void seg_points(R VEC[3], R X[3], R Y[3], const R *P, const R *A,
                                           const R *Q, const R *B) {
    R T[3], TMP[3];
    TDP_SUB(T, Q, P);
    const R AA = TDP_DOT(A, A), BB = TDP_DOT(B, B), AB = TDP_DOT(A, B);
    const R AT = TDP_DOT(A, T), BT = TDP_DOT(B, T);
    R t, u;
    const R denom = AA * BB - AB * AB;
    if (denom == 0) t = 0;
    else t = Ar::div(AT * BB - BT * AB, denom);
    if (BB == 0) u = 0;
    else u = Ar::div(t * AB - BT, BB);
    if (u <= 0) {
        TDP_COPY(Y, Q);
        if (AA == 0) t = 0; else t = Ar::div(AT, AA);
        if (t <= 0) { TDP_COPY(X, P); TDP_SUB(VEC, Q, P); }
        else if (t >= 1) { TDP_ADD(X, P, A); TDP_SUB(VEC, Q, X); }
        else { TDP_ADDS(X, P, A, t); TDP_CROSS(TMP, T, A); TDP_CROSS(VEC, A, TMP); }
    } else if (u >= 1) {
        TDP_ADD(Y, Q, B);
        if (AA == 0) t = 0; else t = Ar::div(AB + AT, AA);
        if (t <= 0) { TDP_COPY(X, P); TDP_SUB(VEC, Y, P); }
        else if (t >= 1) { TDP_ADD(X, P, A); TDP_SUB(VEC, Y, X); }
        else {
            TDP_ADDS(X, P, A, t);
            TDP_SUB(T, Y, P);
            TDP_CROSS(TMP, T, A);
            TDP_CROSS(VEC, A, TMP);
        }
    } else {
        TDP_ADDS(Y, Q, B, u);
        if (t <= 0) { TDP_COPY(X, P); TDP_CROSS(TMP, T, B); TDP_CROSS(VEC, B, TMP); }
        else if (t >= 1) {
            TDP_ADD(X, P, A);
            TDP_SUB(T, Q, X);
            TDP_CROSS(TMP, T, B);
            TDP_CROSS(VEC, B, TMP);
        } else {
            TDP_ADDS(X, P, A, t);
            TDP_CROSS(VEC, A, B);
            if (TDP_DOT(VEC, T) < 0) { VEC[0] = VEC[0] * (R)-1; VEC[1] = VEC[1] * (R)-1; VEC[2] = VEC[2] * (R)-1; }
        }
    }
}

// vertex-face case for the face of A against the vertices of B; returns -1
// when it does not apply
template <typename R, typename Ar = Exact>
__device__ __forceinline__ R face_case(const R *A, const R Av[3][3], const R *B, bool &disjoint,
                                       bool a_first) {
    R N[3], V[3], Z[3], Bp[3];
    TDP_CROSS(N, Av[0], Av[1]);
    const R Nl = TDP_DOT(N, N);
    if ((double)Nl > 1e-15) {
        TDP_SUB(V, A, B);     Bp[0] = TDP_DOT(V, N);
        TDP_SUB(V, A, B + 3); Bp[1] = TDP_DOT(V, N);
        TDP_SUB(V, A, B + 6); Bp[2] = TDP_DOT(V, N);
        int point = -1;
        if ((Bp[0] > 0) && (Bp[1] > 0) && (Bp[2] > 0)) {
            point = (Bp[0] < Bp[1]) ? 0 : 1;
            if (Bp[2] < Bp[point]) point = 2;
        } else if ((Bp[0] < 0) && (Bp[1] < 0) && (Bp[2] < 0)) {
            point = (Bp[0] > Bp[1]) ? 0 : 1;
            if (Bp[2] > Bp[point]) point = 2;
        }
        if (point >= 0) {
            disjoint = true;
            const R *Bv = B + 3 * point;
            TDP_SUB(V, Bv, A);
            TDP_CROSS(Z, N, Av[0]);
            if (TDP_DOT(V, Z) > 0) {
                TDP_SUB(V, Bv, A + 3);
                TDP_CROSS(Z, N, Av[1]);
                if (TDP_DOT(V, Z) > 0) {
                    TDP_SUB(V, Bv, A + 6);
                    TDP_CROSS(Z, N, Av[2]);
                    if (TDP_DOT(V, Z) > 0) {
                        R P[3], Q[3];
                        const R s = Ar::div(Bp[point], Nl);
                        if (a_first) { TDP_ADDS(P, Bv, N, s); TDP_COPY(Q, Bv); }
                        else { TDP_COPY(P, Bv); TDP_ADDS(Q, Bv, N, s); }
                        const R d2 = (P[0] - Q[0]) * (P[0] - Q[0]) + (P[1] - Q[1]) * (P[1] - Q[1]) +
                                     (P[2] - Q[2]) * (P[2] - Q[2]);
                        return Ar::sqrt(d2);
                    }
                }
            }
        }
    }
    return (R)-1;
}

// TriDist: 9 edge-edge cases with the slab test, then the 6 vertex-face
// cases, then "disjoint ? sqrt(min squared edge distance) : 0"
//
// cert (optional) tells whether the value carries a certificate: the slab
// test of an edge pair or the inside test of a vertex-face case. Without one
// the result is the unverified fallback (smallest edge distance when some test
// showed the triangles disjoint, else 0), which is where FP32 and FP64 may
// take different branches.
template <typename R, typename Ar = Exact>
__device__ __forceinline__ R tri_dist(const R *S, const R *T, bool *cert = nullptr) {
    R Sv[3][3], Tv[3][3];
    TDP_SUB(Sv[0], S + 3, S); TDP_SUB(Sv[1], S + 6, S + 3); TDP_SUB(Sv[2], S, S + 6);
    TDP_SUB(Tv[0], T + 3, T); TDP_SUB(Tv[1], T + 6, T + 3); TDP_SUB(Tv[2], T, T + 6);
    bool disjoint = false;
    R mindd = (R)__builtin_inff();
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            R VEC[3], P[3], Q[3], V[3], Z[3];
            seg_points<R, Ar>(VEC, P, Q, S + i * 3, Sv[i], T + j * 3, Tv[j]);
            TDP_SUB(V, Q, P);
            const R dd = TDP_DOT(V, V);
            if (dd <= mindd) {
                mindd = dd;
                TDP_SUB(Z, (S + ((i + 2) % 3) * 3), P);
                R a = TDP_DOT(Z, VEC);
                TDP_SUB(Z, (T + ((j + 2) % 3) * 3), Q);
                R b = TDP_DOT(Z, VEC);
                if ((a <= 0) && (b >= 0)) {
                    if (cert) *cert = true;
                    return Ar::sqrt(mindd);
                }
                const R p = TDP_DOT(V, VEC);
                if (a < 0) a = 0;
                if (b > 0) b = 0;
                if ((p - a + b) > 0) disjoint = true;
            }
        }
    }
    if (cert) *cert = true;
    R r = face_case<R, Ar>(S, Sv, T, disjoint, true);
    if (r != (R)-1) return r;
    r = face_case<R, Ar>(T, Tv, S, disjoint, false);
    if (r != (R)-1) return r;
    if (cert) *cert = false;
    if (disjoint) return Ar::sqrt(mindd);
    return (R)0;
}

// voxel-pair bounds: [box min distance, core-core distance]
// (voxel_dist_core, computing_gpu.cu:1606-1631); boxes are 9 floats
// squared box gap s and squared core distance c of two voxel records
// {low, high, core}, given in R; voxel_bound is (sqrt s, sqrt c). The square
// root being monotone, min sqrt(s) over many pairs is sqrt(min s), bit for
// bit. Per axis the reference adds t2^2 when t2 < 0, else t1^2 when t1 > 0
// (t1 <= t2 always: the extents are ordered). Here the gap g = max(-t2, t1,
// 0) is squared and always added: g is -t2 when t2 < 0, t1 when t1 > 0 and
// +-0 otherwise, and (-t2)^2 = t2^2 while +0 added to s (>= 0) leaves it
// unchanged, so the bits are the same, with no branch (a NaN operand of max
// gives the other, as a false comparison skips its term).
template <typename R, typename T>
__device__ __forceinline__ void voxel_sq(const T *a, const T *b, R &s, R &c) {
    s = 0;
    c = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const R t1 = (R)a[i] - (R)b[i + 3];
        const R t2 = (R)a[i + 3] - (R)b[i];
        const R g = __builtin_fmax(__builtin_fmax(-t2, t1), (R)0);
        s += g * g;
        const R dc = (R)a[i + 6] - (R)b[i + 6];
        c += dc * dc;
    }
}

template <typename R>
__device__ __forceinline__ void voxel_bound(const float *a, const float *b, R &mn, R &mx) {
    R s, c;
    voxel_sq<R>(a, b, s, c);
    mn = dsqrt<R>(s);
    mx = dsqrt<R>(c);
}

}  // namespace tdp
