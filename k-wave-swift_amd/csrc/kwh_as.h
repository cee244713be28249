// kwh_as.h -- elementwise kernels of the axisymmetric solver (kspaceFirstOrderAS,
// Sources/KWave/Solver/KSpaceFirstOrderAS.swift:17-309).
//
// The physical fields are [nx][ny] (x axial, y radial, y = 0 on the axis). Every spectral
// operation mirrors its input into a 4x radially expanded grid [nx][4 ny] with the variable's
// symmetry and uses ordinary transforms there (WSWA for p and ux, HAHS for uy, HSHA for uy / r);
// the kernels below build those expansions, read the expanded results in place (the first ny
// columns, no crop copies), and carry the updates of the time loop (:237-295).
#pragma once
#include "kwh_kernels.h"

namespace kwh {

enum AsMirror { AS_WSWA = 0, AS_HAHS = 1, AS_HSHA = 2 };

// E[x][Y], Y in [0, 4 ny), from w[x][y] (times invr[y] when given) (:201-214):
//   WSWA [w, 0, -w[ny-1..1], -w, 0, w[ny-1..1]]; HAHS [w, flip w, -w, -flip w];
//   HSHA [w, -flip w, -w, flip w].
template <typename T>
__global__ void as_mirror_kernel(const T* __restrict__ w, const T* __restrict__ invr, T* __restrict__ E, int nx,
                                 int ny, int kind) {
    const int ne = 4 * ny;
    const int64_t total = (int64_t)nx * ne;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i / ne);
        const int Y = (int)(i - (int64_t)x * ne);
        const int q = Y / ny, r = Y - q * ny;
        int y;
        T sgn;
        if (kind == AS_WSWA) {
            // quarters [0, ny): +w[Y]; [ny, 2ny): 0 at ny, -w[2ny - Y]; [2ny, 3ny): -w[Y - 2ny];
            // [3ny, 4ny): 0 at 3ny, +w[4ny - Y]
            if (q == 0) { y = r; sgn = T(1); }
            else if (q == 2) { y = r; sgn = T(-1); }
            else if (r == 0) { y = -1; sgn = T(0); }
            else { y = ny - r; sgn = q == 1 ? T(-1) : T(1); }
        } else {
            const bool flip = q & 1;
            y = flip ? ny - 1 - r : r;
            if (kind == AS_HAHS) sgn = q < 2 ? T(1) : T(-1);
            else sgn = (q == 0 || q == 3) ? T(1) : T(-1);
        }
        T v = T(0);
        if (y >= 0) {
            v = w[(int64_t)x * ny + y];
            if (invr) v *= invr[y];
            v *= sgn;
        }
        E[i] = v;
    }
}

// Velocity update from the pressure gradients (:237-241): expanded dp/dx and dp/dy read at
// [x][y < ny].
template <typename T>
__global__ void as_velocity_kernel(T* ux, T* uy, const T* __restrict__ gx, const T* __restrict__ gy, SMap<T> dtrx,
                                   SMap<T> dtry, const T* __restrict__ pmlx_sg, const T* __restrict__ pmly_sg, int nx,
                                   int ny) {
    const int64_t n = (int64_t)nx * ny;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i / ny), y = (int)(i - (int64_t)x * ny);
        const int64_t e = (int64_t)x * 4 * ny + y;
        const T qx = pmlx_sg[x], qy = pmly_sg[y];
        ux[i] = qx * (qx * ux[i] - dtrx.at(i) * gx[e]);
        uy[i] = qy * (qy * uy[i] - dtry.at(i) * gy[e]);
    }
}

// Mass conservation (:253-261): linear dt rho0 div, or nonlinear dt (2 sum rho + rho0) div with
// the previous step's densities.
template <typename T>
__global__ void as_density_kernel(T* rhox, T* rhoy, const T* __restrict__ dux, const T* __restrict__ duy,
                                  SMap<T> rho0, const T* __restrict__ pmlx, const T* __restrict__ pmly, T dt,
                                  int nonlinear, int nx, int ny) {
    const int64_t n = (int64_t)nx * ny;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i / ny), y = (int)(i - (int64_t)x * ny);
        const int64_t e = (int64_t)x * 4 * ny + y;
        const T rx = rhox[i], ry = rhoy[i];
        const T r0 = rho0.at(i);
        const T f = nonlinear ? dt * (T(2) * (rx + ry) + r0) : dt * r0;
        const T qx = pmlx[x], qy = pmly[y];
        rhox[i] = qx * (qx * rx - f * dux[e]);
        rhoy[i] = qy * (qy * ry - f * duy[e]);
    }
}

// Equation of state (:282-287): p = c0^2 (rho [+ tau rho0 div u] [+ B/A rho^2 / (2 rho0)]).
template <typename T>
__global__ void as_eos_kernel(T* p, const T* __restrict__ rhox, const T* __restrict__ rhoy, const T* __restrict__ dux,
                              const T* __restrict__ duy, SMap<T> c0, SMap<T> rho0, SMap<T> tau, SMap<T> bona,
                              int stokes, int nonlinear, int nx, int ny) {
    const int64_t n = (int64_t)nx * ny;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i / ny), y = (int)(i - (int64_t)x * ny);
        const int64_t e = (int64_t)x * 4 * ny + y;
        const T rt = rhox[i] + rhoy[i];
        T eos = rt;
        if (stokes) eos = eos + tau.at(i) * rho0.at(i) * (dux[e] + duy[e]);
        if (nonlinear) eos = eos + bona.at(i) * (rt * rt) / (T(2) * rho0.at(i));
        const T c = c0.at(i);
        p[i] = c * c * eos;
    }
}

// f0 += E[x][y] (and f1 += E[x][y] when given): a k-space corrected source field (:176-178).
template <typename T>
__global__ void as_add_crop_kernel(T* f0, T* f1, const T* __restrict__ E, int nx, int ny) {
    const int64_t n = (int64_t)nx * ny;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i / ny), y = (int)(i - (int64_t)x * ny);
        const T v = E[(int64_t)x * 4 * ny + y];
        f0[i] += v;
        if (f1) f1[i] += v;
    }
}

// Point sources at step t: dst[index] = v (dirichlet) or dst[index] += v (additive without
// correction), v = signal[row][t] * base * factor; dst1 (optional) gets the same (:264-276).
template <typename T>
__global__ void as_scatter_kernel(T* dst0, T* dst1, const int64_t* __restrict__ index, const T* __restrict__ signal,
                                  const T* __restrict__ base, int rows, int len, int n, int t, T factor, int add) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n || t >= len) return;
    const T v = signal[(int64_t)(rows == 1 ? 0 : e) * len + t] * base[e] * factor;
    const int64_t c = index[e];
    dst0[c] = add ? dst0[c] + v : v;
    if (dst1) dst1[c] = add ? dst1[c] + v : v;
}

// t = 0 override with p0 (:290-297): p = p0, rho_a = p0 / (2 c0^2), u_a = dt/rho_a / 2 * grad p0.
template <typename T>
__global__ void as_init_kernel(const T* __restrict__ p0, T* p, T* rhox, T* rhoy, T* ux, T* uy, const T* __restrict__ gx,
                               const T* __restrict__ gy, SMap<T> c0, SMap<T> dtrx, SMap<T> dtry, int nx, int ny) {
    const int64_t n = (int64_t)nx * ny;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i / ny), y = (int)(i - (int64_t)x * ny);
        const int64_t e = (int64_t)x * 4 * ny + y;
        const T f = p0[i];
        const T c = c0.at(i);
        p[i] = f;
        const T r = f / (T(2) * (c * c));
        rhox[i] = r;
        rhoy[i] = r;
        ux[i] = (dtrx.at(i) / T(2)) * gx[e];
        uy[i] = (dtry.at(i) / T(2)) * gy[e];
    }
}

}  // namespace kwh
