// kwh_adjoint.h -- elementwise kernels of the reverse-time (adjoint) loop.
//
// The adjoint of the lossless, linear step (p0 problem, pressure recorded) is, with s_a / q_a the
// staggered / regular PML along a, r_a = dt / rho0_a (staggered), m = dt rho0, D+-_a the
// staggered derivative operators (whose transposes are -D-+_a), S the sensor sampling:
//   lp^n   = S^T g^n + sum_a D-_a[r_a s_a lu_a^{n+1}]
//   lr_a^n = c^2 lp^n + q_a^2 lr_a^{n+1}
//   lu_a^n = s_a^2 lu_a^{n+1} + D+_a[m q_a lr_a^n]
//   dL/dc += 2 lp^n p^n / c
// and at n = 0 (the p0 override, KSpaceFirstOrder.swift:724-733)
//   dL/dp0 = lp^0 + sum_a q_a^2 lr_a^1 / (d c^2) - sum_a D-_a[(r_a / 2) s_a^2 lu_a^1]
//   dL/dc += -2 p0 / (d c^3) sum_a q_a^2 lr_a^1.
// The spectral operators reuse the forward's transform passes; these kernels are the diagonal
// parts between them. Every kernel is grid-stride over the N cells.
#pragma once
#include "kwh_kernels.h"

namespace kwh {

// Index of cell i along internal axis ia.
__device__ __forceinline__ int cell_axis_index(const Geo& g, int64_t i, int ia) {
    if (ia == 2) return (int)(i % g.m2);
    const int64_t row = i / g.m2;
    return ia == 1 ? (int)(row % g.m1) : (int)(row / g.m1);
}

#define KWH_GRID_LOOP(i, n) \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// w = scale * r_a * s_a^spow * lu   (spow 1: the divergence input; 2: the t = 0 term)
template <typename T>
__global__ void adj_prediv_kernel(T* w, const T* lu, SMap<T> r, const T* s, int ia, int spow, T scale, Geo g) {
    KWH_GRID_LOOP(i, g.N) {
        const T sv = s[cell_axis_index(g, i, ia)];
        const T f = spow == 2 ? sv * sv : sv;
        w[i] = scale * r.at(i) * f * lu[i];
    }
}

// lr = c^2 lp + q^2 lr (in place); w = m q lr, m = dt rho0
template <typename T>
__global__ void adj_rho_kernel(T* lr, T* w, const T* lp, SMap<T> c0, SMap<T> rho0, T dt, const T* q, int ia, Geo g) {
    KWH_GRID_LOOP(i, g.N) {
        const T qv = q[cell_axis_index(g, i, ia)];
        const T c = c0.at(i);
        const T v = c * c * lp[i] + qv * qv * lr[i];
        lr[i] = v;
        w[i] = dt * rho0.at(i) * qv * v;
    }
}

// lu = s^2 lu (the D+ term is accumulated onto it afterwards)
template <typename T>
__global__ void adj_u_decay_kernel(T* lu, const T* s, int ia, Geo g) {
    KWH_GRID_LOOP(i, g.N) {
        const T sv = s[cell_axis_index(g, i, ia)];
        lu[i] *= sv * sv;
    }
}

// dL/dc += 2 lp p / c
template <typename T>
__global__ void adj_grad_c_kernel(T* gc, const T* lp, const T* p, SMap<T> c0, int64_t n) {
    KWH_GRID_LOOP(i, n) { gc[i] += T(2) * lp[i] * p[i] / c0.at(i); }
}

// lp += S^T g for one time step: g[k] at sensor point k, spread over its corners (binary
// masks: one corner, weight 1). Atomic: Cartesian points may share corners.
template <typename T>
__global__ void adj_scatter_kernel(T* lp, const T* g, const int64_t* idx, const T* w, int nc, int64_t nsens) {
    KWH_GRID_LOOP(s, nsens * nc) {
        const int64_t k = s / nc;
        const T v = w ? g[k] * w[s] : g[k];
        atomicAdd(&lp[idx[s]], v);
    }
}

// dL/dp0 = lp0 - t0 + lr0 / (d c^2), dL/dc += -2 p0 lr0 / (d c^3), lr0 = sum_a q_a^2 lr_a^1.
template <typename T>
struct AdjFinalArgs {
    T* gp0;
    T* gc;
    const T* lp0;
    const T* t0;
    const T* lr[3];
    const T* q[3];
    const T* p0;
    SMap<T> c0;
    int dim;
};

template <typename T>
__global__ void adj_final_kernel(AdjFinalArgs<T> a, Geo g) {
    KWH_GRID_LOOP(i, g.N) {
        T lr0 = 0;
        for (int ax = 0; ax < a.dim; ++ax) {
            const T qv = a.q[ax][cell_axis_index(g, i, ax + g.off)];
            lr0 += qv * qv * a.lr[ax][i];
        }
        const T c = a.c0.at(i);
        const T dc2 = T(a.dim) * c * c;
        a.gp0[i] = a.lp0[i] - a.t0[i] + lr0 / dc2;
        a.gc[i] += T(-2) * a.p0[i] * lr0 / (dc2 * c);
    }
}

// p = c^2 sum_a rho_a (rebuilds the pressure of a restored checkpoint)
template <typename T>
__global__ void eos_linear_kernel(T* p, const T* r0, const T* r1, const T* r2, int dim, SMap<T> c0, int64_t n) {
    KWH_GRID_LOOP(i, n) {
        T s = r0[i];
        if (dim > 1) s += r1[i];
        if (dim > 2) s += r2[i];
        const T c = c0.at(i);
        p[i] = c * c * s;
    }
}

#undef KWH_GRID_LOOP

}  // namespace kwh
