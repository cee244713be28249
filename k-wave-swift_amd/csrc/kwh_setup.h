// kwh_setup.h -- one-time operator precompute kernels, float64 arithmetic on the device.
//
// The reference precomputes kappa / source kappa (KSpaceFirstOrder.swift:581-583), the
// fractional-Laplacian absorption operators (:812-816, :851-882), the staggered densities
// (:799-806) and the smoothing window (Filters.swift:203-347) in float32; here they are
// evaluated in float64 and rounded once to the working type.
#pragma once
#include "kwh_kernels.h"

namespace kwh {

struct KGeo {
    int m0, m1, m2, mzh;
    int64_t NH;
    const double* kv0;   // wavenumbers of internal axes (FFT order, full length)
    const double* kv1;
    const double* kv2;
};

__device__ __forceinline__ void kspace_index(const KGeo& g, int64_t h, int& i0, int& i1, int& i2) {
    i2 = (int)(h % g.mzh);
    const int64_t r = h / g.mzh;
    i1 = (int)(r % g.m1);
    i0 = (int)(r / g.m1);
}

// kappa = sinc(c_ref dt |k| / 2), source kappa = cos(...), nabla1 = |k|^(y-2), nabla2 = |k|^(y-1)
// (DC -> 0). Null outputs are skipped.
template <typename T>
__global__ void kspace_tables_kernel(KGeo g, double carg, double pw1, double pw2, T* kappa, T* skappa, T* nab1,
                                     T* nab2) {
    for (int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; h < g.NH; h += (int64_t)gridDim.x * blockDim.x) {
        int i0, i1, i2;
        kspace_index(g, h, i0, i1, i2);
        const double k = sqrt(g.kv0[i0] * g.kv0[i0] + g.kv1[i1] * g.kv1[i1] + g.kv2[i2] * g.kv2[i2]);
        const double a = k * carg;
        if (kappa) kappa[h] = T(a == 0.0 ? 1.0 : sin(a) / a);
        if (skappa) skappa[h] = T(cos(a));
        if (nab1) nab1[h] = T(k == 0.0 ? 0.0 : pow(k, pw1));
        if (nab2) nab2[h] = T(k == 0.0 ? 0.0 : pow(k, pw2));
    }
}

// Smoothing window (Filters.swift:252-347), Hermitian part over the half spectrum.
// place2d selects the reference's 2-D placement flat[q] = w[(q + (n+1)/2) % n]; otherwise the
// 1-D / 3-D placement w[s] -> (s + (n+1)/2) % n.
struct WinGeo {
    int dim;
    int n[3];          // physical sizes
    int off;           // internal axis = physical + off
    int m1, mzh;
    int64_t NH;
    const double* win; // symmetric Blackman of length Lw
    int Lw;
    double radius;
    int place2d;
};

__device__ __forceinline__ double win_axis_coord(const WinGeo& w, int a, int s) {
    const int n = w.n[a];
    const int count = (n % 2) ? n : n + 1;
    if (count <= 1) return 0.0;
    return -w.radius + 2.0 * w.radius * (double)s / (double)(count - 1);
}

__device__ double win_value(const WinGeo& w, const int q[3]) {
    // q: FFT-order index per physical axis -> source index s per axis -> radial interp
    double r2 = 0.0;
    for (int a = 0; a < w.dim; ++a) {
        const int n = w.n[a];
        const int sh = (n + 1) / 2;
        const int s = w.place2d ? (q[a] + sh) % n : ((q[a] - sh) % n + n) % n;
        const double x = win_axis_coord(w, a, s);
        r2 += x * x;
    }
    const double r = sqrt(r2);
    const double pos = fmin(r, w.radius) + w.radius;
    int i0 = (int)floor(pos);
    if (i0 > w.Lw - 2) i0 = w.Lw - 2;
    if (i0 < 0) i0 = 0;
    const double fr = pos - i0;
    const double hi = (w.Lw >= 2) ? w.win[i0 + 1] : w.win[0];
    return fabs(w.win[i0] * (1 - fr) + hi * fr);
}

template <typename T>
__global__ void window_kernel(WinGeo w, T* out) {
    for (int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; h < w.NH; h += (int64_t)gridDim.x * blockDim.x) {
        const int i2 = (int)(h % w.mzh);
        const int64_t r = h / w.mzh;
        const int iint[3] = {(int)(r / w.m1), (int)(r % w.m1), i2};
        int q[3], qm[3];
        for (int a = 0; a < w.dim; ++a) {
            q[a] = iint[a + w.off];
            qm[a] = (w.n[a] - q[a]) % w.n[a];
        }
        out[h] = T(0.5 * (win_value(w, q) + win_value(w, qm)));
    }
}

// 2-D directivity weight (Directivity.swift:54-71), Hermitian part.
struct DirGeo {
    int nx, ny, mzh;
    int64_t NH;
    const double* kx;
    const double* ky;
    double ct, st;     // cos / sin of the angle
    double half_size;  // element size / 2
    int gradient;
};

__device__ __forceinline__ double dir_weight(const DirGeo& g, int ix, int iy) {
    const double kx = g.kx[ix], ky = g.ky[iy];
    if (g.gradient) {
        const double kn = g.st * ky + g.ct * kx;
        const double km = sqrt(kx * kx + ky * ky);
        return km == 0.0 ? 0.0 : kn / km;
    }
    const double x = (g.ct * ky - g.st * kx) * g.half_size;
    if (x == 0.0) return 1.0;
    const double px = 3.14159265358979323846 * x;
    return sin(px) / px;
}

template <typename T>
__global__ void directivity_kernel(DirGeo g, T* out) {
    for (int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; h < g.NH; h += (int64_t)gridDim.x * blockDim.x) {
        const int iy = (int)(h % g.mzh);
        const int ix = (int)(h / g.mzh);
        const int mx = (g.nx - ix) % g.nx, my = (g.ny - iy) % g.ny;
        out[h] = T(0.5 * (dir_weight(g, ix, iy) + dir_weight(g, mx, my)));
    }
}

// dt / rho0 staggered +d/2 along internal axis ax (neighbour mean, last cell replicated).
template <typename T>
__global__ void stagger_kernel(const T* rho0, T* out, int m0, int m1, int m2, int ax, double dt) {
    const int64_t N = (int64_t)m0 * m1 * m2;
    const int64_t st = ax == 2 ? 1 : (ax == 1 ? m2 : (int64_t)m1 * m2);
    const int len = ax == 2 ? m2 : (ax == 1 ? m1 : m0);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const int pos = (int)((i / st) % len);
        const double a = rho0[i];
        const double r = pos + 1 < len ? 0.5 * (a + (double)rho0[i + st]) : a;
        out[i] = T(dt / r);
    }
}

// tau = -2 a c^(y-1) (or -2 a c for Stokes), eta = 2 a c^y tan(pi y / 2) (makeAbsorption :863-878).
template <typename T>
__global__ void absorption_kernel(const T* c0, T cs, const T* alpha, T as, int64_t N, double db2np, double y,
                                  int stokes, double tanfac, T* tau, T* eta) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const double c = c0 ? (double)c0[i] : (double)cs;
        const double a = (alpha ? (double)alpha[i] : (double)as) * db2np;
        if (tau) tau[i] = T(stokes ? -2.0 * a * c : -2.0 * a * pow(c, y - 1.0));
        if (eta) eta[i] = T(2.0 * a * pow(c, y) * tanfac);
    }
}

// Delayed copies of one drive row (a transducer transmit): rows[k][t] = scale[k] *
// sig[t - delay[k]] for 0 <= t - delay[k] < len, else 0; rows is [n][len_eff].
template <typename T>
__global__ void delayed_rows_kernel(const float* sig, int len, const int* delay, const float* scale, int64_t n,
                                    int len_eff, T* rows) {
    const int64_t total = n * (int64_t)len_eff;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = i / len_eff;
        const int s = (int)(i - k * len_eff) - delay[k];
        const double a = scale ? (double)scale[k] : 1.0;
        rows[i] = (s >= 0 && s < len) ? T(a * (double)sig[s]) : T(0);
    }
}

template <typename T>
__global__ void gather_values_kernel(const T* f, const int64_t* idx, int64_t n, double* out) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x)
        out[s] = (double)f[idx[s]];
}

}  // namespace kwh
