// kwh_kernels.h -- HIP kernels of the k-space first-order time loop (gfx950).
//
// Data layout in HBM. Every grid is handled as internal 3-D [m0][m1][m2] with m2 contiguous:
// 1D [nx] -> [1][1][nx], 2D [nx,ny] -> [1][nx][ny], 3D [nx,ny,nz] -> [nx][ny][nz] (C-order, the
// reference's MLXArray layout). Real fields are m0*m1*m2 values. Spectra are half spectra along
// m2: [m0][m1][mzh] complex, mzh = m2/2+1 (R2C along the contiguous axis, then complex passes).
//
// Per time step the reference computes (KSpaceFirstOrder.swift:669-702)
//   u_a  = pml_sg_a * (pml_sg_a * u_a - dt/rho_a * Re ifft(ddxPos_a * kappa * fft p))
//   du_a = Re ifft(ddxNeg_a * kappa * fft u_a)
//   rho_a = pml_a * (pml_a * rho_a - dt*rho0 * du_a * nl)
//   p = c^2 (sum rho + tau Re ifft(nab1 fft(rho0 sum du)) - eta Re ifft(nab2 fft(sum rho)) + B/A ...)
// with C2C transforms and Re(). Here every transform is R2C/C2R. Re(ifft(op * fft f)) equals
// C2R(op_H * R2C f) where op_H is the Hermitian part of op; for the single-axis operators used
// op_H == op except at an even-length Nyquist bin, where op_H = Re(op) (built on the host).
//
// Kernels:
//   strided_pass   complex FFT along internal axis 0 or 1, in place, several fields per launch
//   middle_pass    forward FFT along the outermost axis, per-output k-space operator, inverse
//                  FFT, written to output spectra (fuses the two outermost passes of a
//                  transform pair and the operator multiply)
//   z_* kernels    R2C / C2R along the contiguous axis with the elementwise update fused:
//                  two real rows form one complex transform (a + i b)
#pragma once
#include "kwh_fft.h"

namespace kwh {

struct Geo {
    int m0, m1, m2, mzh;
    int64_t N, NH;
    int nrows;     // m0 * m1
    int dim;
    int off;       // 3 - dim: internal axis of physical axis a is a + off
};

// Per-internal-axis tables for one working type T.
template <typename T> struct AxisTabs {
    FftPlan plan[3];
    const cx<T>* tw[3];
    const cx<T>* ddxp[3];   // i k e^{+ik d/2}, Hermitian part at Nyquist
    const cx<T>* ddxn[3];   // i k e^{-ik d/2}
    const cx<T>* coll[3];   // e^{-ik d/2}
    const T* pml[3];        // regular-grid PML (length m_axis; 1 on dummy axes)
    const T* pmlsg[3];      // staggered PML
};

// XCD-aware tile order: consecutive tiles share one XCD's L2 (8 XCDs, round-robin dispatch).
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
    const int nx = 8;
    if (nblocks < 2 * nx) return bid;
    const int per = nblocks / nx;
    const int full = per * nx;
    if (bid >= full) return bid;
    return (bid % nx) * per + bid / nx;
}

// Block coordinates of one simulation's share of a launch. A single launch reads them from the
// hardware and applies the XCD-aware tile order to x; a launch batched over several simulations
// (kwh_batch.h) hands each its own block range, already in XCD order.
struct BIdx {
    int x, nx, y;
    int remap;
    __device__ __forceinline__ int tile() const { return remap ? xcd_tile(x, nx) : x; }
    // XCD-aware order with each XCD's run of tiles walked from its end (single launches only)
    __device__ __forceinline__ int tile_rev() const {
        if (!remap) return x;
        const int t = xcd_tile(x, nx), per = nx / 8;
        return (nx < 16 || t >= per * 8) ? t : (t / per) * per + (per - 1 - t % per);
    }
};
__device__ __forceinline__ BIdx bidx_here() { return BIdx{(int)blockIdx.x, (int)gridDim.x, (int)blockIdx.y, 1}; }

template <typename T> __device__ __forceinline__ T& comp(cx<T>& v, int part) {
    return reinterpret_cast<T*>(&v)[part];
}

// ---------------------------------------------------------------------------------------------
// Strided passes
// ---------------------------------------------------------------------------------------------

struct StridedArgs {
    void* fields[8];
    int nfields;
    int axis;        // internal axis 0 or 1
    int L;
    int64_t stride;  // element stride along the axis (complex elements)
    int64_t inner;   // columns are c -> (c / inner) * L * inner + c % inner
    int64_t ncols;
    int lg_ntr;
    int pitch;
    int walk_rev;    // register strided pass: walk each XCD's run of tiles from its end
    // register inverse pass with two outputs from field dual - 1 (dual = 0: none): that field becomes
    // IFFT(d1[n] X) in place (n: index along the pass axis) and dual_out gets IFFT(d2[col] X)
    // (col: the column's index within `inner`); both tables are cx<T>
    int dual;
    void* dual_out;
    const void* d1;
    const void* d2;
};

template <typename T>
__device__ __forceinline__ void tile_colbase(int64_t* colbase, int64_t c0, int ntr, int64_t ncols, int64_t inner,
                                             int L, int tid) {
    if (tid < ntr) {
        const int64_t c = c0 + tid;
        colbase[tid] = c < ncols ? (c / inner) * (int64_t)L * inner + c % inner : -1;
    }
}

template <typename T, bool INV>
__device__ __forceinline__ void strided_pass_body(const BIdx& bi, const StridedArgs& a, const FftPlan& pl, const cx<T>* __restrict__ tw) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int ntr = 1 << a.lg_ntr;
    const int tid = threadIdx.x, nthr = blockDim.x;
    int64_t* colbase = reinterpret_cast<int64_t*>(smem);
    cx<T>* b0 = reinterpret_cast<cx<T>*>(smem + 64 * sizeof(int64_t));
    cx<T>* b1 = b0 + (size_t)a.L * a.pitch;
    const int tile = bi.tile();
    const int64_t c0 = (int64_t)tile * ntr;
    tile_colbase<T>(colbase, c0, ntr, a.ncols, a.inner, a.L, tid);
    __syncthreads();
    cx<T>* data = reinterpret_cast<cx<T>*>(a.fields[bi.y]);
    GlobAcc<T> g{data, colbase, a.stride};
    if (pl.nstages >= 2) {
        fft_run<T, INV>(pl, g, g, b0, b1, a.pitch, a.lg_ntr, tw, tid, nthr);
    } else {
        LdsAcc<T> B{b1, a.pitch};
        fft_run<T, INV>(pl, g, B, b0, b1, a.pitch, a.lg_ntr, tw, tid, nthr);
        __syncthreads();
        for (int i = tid; i < (a.L << a.lg_ntr); i += nthr) g.st(i & (ntr - 1), i >> a.lg_ntr, B.ld(i & (ntr - 1), i >> a.lg_ntr));
    }
}
template <typename T, bool INV>
__global__ void __launch_bounds__(256) strided_pass(StridedArgs a, FftPlan pl, const cx<T>* __restrict__ tw) {
    strided_pass_body<T, INV>(bidx_here(), a, pl, tw);
}

// ---------------------------------------------------------------------------------------------
// Middle pass: forward along the middle axis, k-space operator, inverse, store
// ---------------------------------------------------------------------------------------------

enum OpKind { OP_DDXP_KAPPA = 0, OP_DDXN_KAPPA = 1, OP_COLL = 2, OP_TAB = 3, OP_DDXP = 4 };

struct MidOp {
    int kind;
    int axis;      // internal axis of the 1-D operator
    int in;        // input field index
    void* out;     // output spectrum
    const void* table;  // OP_TAB: real table over the half spectrum; kappa for *_KAPPA
};

struct MidArgs {
    const void* in[3];
    int nin;
    MidOp op[8];
    int nop;
    int axis;      // middle axis (0 or 1)
    int L;
    int64_t stride, inner, ncols;
    int lg_ntr, pitch;
    double scale;  // 1/N normalisation folded into the operator
    // input 0 is zero outside the middle-axis slabs [slab_lo, slab_lo + slab_n) (slab_n = 0: dense
    // input): the register middle pass then forms its forward transform directly from those slabs
    int slab_lo, slab_n;
};

template <typename T>
__device__ __forceinline__ cx<T> mid_op_value(const MidOp& op, const AxisTabs<T>& tab, const Geo& g, int mid_axis,
                                              int n, int64_t col, T scale) {
    // internal indices of spectrum element (column col, index n along the middle axis)
    int i0, i1, i2;
    int64_t h;
    if (mid_axis == 0) {
        i0 = n;
        i1 = (int)(col / g.mzh);
        i2 = (int)(col % g.mzh);
        h = (int64_t)n * g.m1 * g.mzh + col;
    } else {
        i0 = (int)(col / g.mzh);
        i2 = (int)(col % g.mzh);
        i1 = n;
        h = ((int64_t)i0 * g.m1 + n) * g.mzh + i2;
    }
    const int idx = op.axis == 0 ? i0 : (op.axis == 1 ? i1 : i2);
    cx<T> v;
    switch (op.kind) {
        case OP_DDXP_KAPPA: {
            const T k = reinterpret_cast<const T*>(op.table)[h];
            v = tab.ddxp[op.axis][idx] * (k * scale);
            break;
        }
        case OP_DDXN_KAPPA: {
            const T k = reinterpret_cast<const T*>(op.table)[h];
            v = tab.ddxn[op.axis][idx] * (k * scale);
            break;
        }
        case OP_COLL: {   // with a table: the shift times that table (kappa)
            const T k = op.table ? reinterpret_cast<const T*>(op.table)[h] : T(1);
            v = tab.coll[op.axis][idx] * (k * scale);
            break;
        }
        case OP_DDXP: v = tab.ddxp[op.axis][idx] * scale; break;
        default: {
            const T k = reinterpret_cast<const T*>(op.table)[h];
            v = cx<T>{k * scale, T(0)};
            break;
        }
    }
    return v;
}

template <typename T>
__device__ __forceinline__ void middle_pass_body(const BIdx& bi, const MidArgs& a, const AxisTabs<T>& tab, const Geo& g) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int ntr = 1 << a.lg_ntr;
    const int tid = threadIdx.x, nthr = blockDim.x;
    int64_t* colbase = reinterpret_cast<int64_t*>(smem);
    cx<T>* b0 = reinterpret_cast<cx<T>*>(smem + 64 * sizeof(int64_t));
    const size_t bl = (size_t)(a.L > 0 ? a.L : 1) * a.pitch;
    cx<T>* b1 = b0 + bl;
    cx<T>* X = b1 + bl;
    const int tile = bi.tile();
    const int64_t c0 = (int64_t)tile * ntr;
    tile_colbase<T>(colbase, c0, ntr, a.ncols, a.inner, a.L, tid);
    __syncthreads();
    const FftPlan& pl = tab.plan[a.axis];
    const cx<T>* tw = tab.tw[a.axis];
    const T scale = T(a.scale);
    const int tot = a.L << a.lg_ntr;
    LdsAcc<T> LX{X, a.pitch}, L0{b0, a.pitch};
    for (int f = 0; f < a.nin; ++f) {
        GlobAcc<T> gin{const_cast<cx<T>*>(reinterpret_cast<const cx<T>*>(a.in[f])), colbase, a.stride};
        if (pl.nstages == 0) {
            for (int i = tid; i < tot; i += nthr) LX.st(i & (ntr - 1), i >> a.lg_ntr, gin.ld(i & (ntr - 1), i >> a.lg_ntr));
        } else {
            fft_run<T, false>(pl, gin, LX, b0, b1, a.pitch, a.lg_ntr, tw, tid, nthr);
        }
        __syncthreads();
        for (int o = 0; o < a.nop; ++o) {
            const MidOp op = a.op[o];
            if (op.in != f) continue;
            GlobAcc<T> gout{reinterpret_cast<cx<T>*>(op.out), colbase, a.stride};
            if (pl.nstages == 0) {
                for (int i = tid; i < tot; i += nthr) {
                    const int w = i & (ntr - 1), n = i >> a.lg_ntr;
                    if (colbase[w] < 0) continue;
                    gout.st(w, n, cmul(LX.ld(w, n), mid_op_value<T>(op, tab, g, a.axis, n, c0 + w, scale)));
                }
                continue;
            }
            for (int i = tid; i < tot; i += nthr) {
                const int w = i & (ntr - 1), n = i >> a.lg_ntr;
                cx<T> v = LX.ld(w, n);
                if (colbase[w] >= 0) v = cmul(v, mid_op_value<T>(op, tab, g, a.axis, n, c0 + w, scale));
                L0.st(w, n, v);
            }
            __syncthreads();
            if (pl.nstages >= 1) {
                // stage 0 reads b0 (first) and writes b1; later stages ping-pong; the last writes global
                fft_run<T, true>(pl, L0, gout, b0, b1, a.pitch, a.lg_ntr, tw, tid, nthr);
            }
            __syncthreads();
        }
    }
}
template <typename T>
__global__ void __launch_bounds__(256) middle_pass(MidArgs a, AxisTabs<T> tab, Geo g) {
    middle_pass_body<T>(bidx_here(), a, tab, g);
}

// ---------------------------------------------------------------------------------------------
// Contiguous-axis (z) helpers: two real rows (2q, 2q+1) form one complex transform a + i b
// ---------------------------------------------------------------------------------------------

struct ZTile {
    int q0;       // first row pair of the tile
    int ntr;      // pairs per tile (power of two)
    int lg_ntr;
    int pitch;    // LDS pitch (>= ntr)
};

// Build the full Hermitian-extended spectrum Z = A + iB of rows (2q, 2q+1) in LDS buffer Z.
template <typename T>
__device__ void z_load_spectrum(const cx<T>* __restrict__ spec, const Geo& g, const ZTile& zt, cx<T>* Z, int tid,
                                int nthr) {
    const int L = g.m2, mzh = g.mzh;
    for (int i = tid; i < zt.ntr * mzh; i += nthr) {
        const int w = i / mzh, k = i - w * mzh;
        const int ra = 2 * (zt.q0 + w), rb = ra + 1;
        cx<T> A = ra < g.nrows ? spec[(int64_t)ra * mzh + k] : cx<T>{T(0), T(0)};
        cx<T> B = rb < g.nrows ? spec[(int64_t)rb * mzh + k] : cx<T>{T(0), T(0)};
        if (k == 0 || 2 * k == L) {
            Z[k * zt.pitch + w] = cx<T>{A.x, B.x};
        } else {
            Z[k * zt.pitch + w] = cx<T>{A.x - B.y, A.y + B.x};
            Z[(L - k) * zt.pitch + w] = cx<T>{A.x + B.y, B.x - A.y};
        }
    }
}

// Split Z = FFT(a + i b) into the half spectra of rows 2q (a) and 2q+1 (b) and store them.
template <typename T>
__device__ void z_store_spectrum(cx<T>* __restrict__ spec, const Geo& g, const ZTile& zt, const cx<T>* Z, int tid,
                                 int nthr) {
    const int L = g.m2, mzh = g.mzh;
    for (int i = tid; i < zt.ntr * mzh; i += nthr) {
        const int w = i / mzh, k = i - w * mzh;
        const int ra = 2 * (zt.q0 + w), rb = ra + 1;
        const cx<T> zk = Z[k * zt.pitch + w];
        const int km = k == 0 ? 0 : L - k;
        const cx<T> zm = Z[km * zt.pitch + w];
        if (ra < g.nrows) spec[(int64_t)ra * mzh + k] = cx<T>{T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
        if (rb < g.nrows) spec[(int64_t)rb * mzh + k] = cx<T>{T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
    }
}

// Inverse along z of spectrum `spec` for this tile: result (a in .x, b in .y) in the returned
// LDS buffer. Caller must __syncthreads() before reading it.
template <typename T>
__device__ cx<T>* z_inverse(const cx<T>* spec, const Geo& g, const ZTile& zt, const FftPlan& pl, const cx<T>* tw,
                            cx<T>* b0, cx<T>* b1, int tid, int nthr) {
    z_load_spectrum<T>(spec, g, zt, b0, tid, nthr);
    __syncthreads();
    if (pl.nstages == 0) return b0;
    cx<T>* out = ((pl.nstages - 1) & 1) ? b0 : b1;
    LdsAcc<T> A{b0, zt.pitch}, O{out, zt.pitch};
    fft_run<T, true>(pl, A, O, b0, b1, zt.pitch, zt.lg_ntr, tw, tid, nthr);
    return out;
}

// Forward along z of the real rows held in LDS buffer `src` (a in .x, b in .y), stored to spec.
template <typename T>
__device__ void z_forward(cx<T>* spec, const Geo& g, const ZTile& zt, const FftPlan& pl, const cx<T>* tw, cx<T>* src,
                          cx<T>* other, int tid, int nthr) {
    __syncthreads();
    // Odd row count: the last pair has no b row. Its slot holds what the kernel left there (the
    // rounding residue of an inverse, at the scale of the field before the update), and the
    // packed transform's rounding is relative to max(|a|, |b|): clear it so that row a keeps its
    // own precision. The condition is uniform over the workgroup.
    if (g.nrows & 1) {
        const int lastq = (g.nrows - 1) / 2;
        if (lastq >= zt.q0 && lastq < zt.q0 + zt.ntr) {
            for (int n = tid; n < g.m2; n += nthr) src[n * zt.pitch + (lastq - zt.q0)].y = T(0);
            __syncthreads();
        }
    }
    cx<T>* res = src;
    if (pl.nstages > 0) {
        // ping-pong between src and other; result lands in the buffer the last stage writes
        res = ((pl.nstages - 1) & 1) ? src : other;
        LdsAcc<T> A{src, zt.pitch}, O{res, zt.pitch};
        fft_run<T, false>(pl, A, O, src, other, zt.pitch, zt.lg_ntr, tw, tid, nthr);
    }
    __syncthreads();
    z_store_spectrum<T>(spec, g, zt, res, tid, nthr);
}

// Cell loop helper: calls f(cell, row, n, slot_ref) for every valid cell of the tile.
#define KWH_CELL_LOOP(zt, g, tid, nthr, BODY)                                      \
    for (int _i = (tid); _i < 2 * (zt).ntr * (g).m2; _i += (nthr)) {              \
        const int _lr = _i / (g).m2;                                              \
        const int n = _i - _lr * (g).m2;                                          \
        const int row = 2 * (zt).q0 + _lr;                                        \
        if (row >= (g).nrows) continue;                                           \
        const int w = _lr >> 1, part = _lr & 1;                                   \
        const int64_t cell = (int64_t)row * (g).m2 + n;                           \
        BODY                                                                      \
    }

// Position of a cell along internal axis ax.
__device__ __forceinline__ int axis_pos(const Geo& g, int row, int n, int ax) {
    return ax == 2 ? n : (ax == 1 ? row % g.m1 : row / g.m1);
}

// Scalar-or-map accessor. `dp` always points at device memory: the map itself (mask = ~0) or a
// buffer holding the scalar (mask = 0), so at() is a branch-free load. A scalar's buffer holds
// kScalarRep copies: lane(off, j) returns a per-lane pointer whose [J*m] elements (j + J*m <
// kScalarRep, the longest register-resident row) are the map's row elements or the scalar, so the
// register-resident kernels address every element as base + immediate. `p` / `s` keep the
// host-side view (p == nullptr: scalar).
constexpr int kScalarRep = 1024;
template <typename T> struct SMap {
    const T* p;
    T s;
    const T* dp;
    int64_t mask;
    __device__ __forceinline__ T at(int64_t i) const { return dp[i & mask]; }
    __device__ __forceinline__ const T* lane(int64_t off, int j) const { return p ? p + off : dp + j; }
};

// Sparse sources of one field: rowptr over internal rows, flat indices, signal table.
template <typename T> struct SparseSrc {
    const int* rowptr;      // [nrows+1], null = none
    const int64_t* index;   // flat cell indices (ascending)
    const T* signal;        // [rows][len]
    const T* base;          // per-point base scale
    int rows, len;
    int mode;               // 0 dirichlet, 2 additive-no-correction
    T factor;               // per-axis scale factor
};

template <typename T>
__device__ void apply_sparse(const SparseSrc<T>& s, const Geo& g, const ZTile& zt, cx<T>* Z, int t, int tid,
                             int nthr) {
    if (!s.rowptr || t >= s.len) return;
    const int r0 = 2 * zt.q0;
    const int r1 = min(g.nrows, r0 + 2 * zt.ntr);
    const int e0 = s.rowptr[r0], e1 = s.rowptr[r1];
    for (int e = e0 + tid; e < e1; e += nthr) {
        const int64_t cell = s.index[e];
        const int row = (int)(cell / g.m2);
        const int n = (int)(cell - (int64_t)row * g.m2);
        const int lr = row - r0;
        const T v = s.signal[(int64_t)(s.rows == 1 ? 0 : e) * s.len + t] * s.base[e] * s.factor;
        T& slot = comp(Z[n * zt.pitch + (lr >> 1)], lr & 1);
        slot = s.mode == 0 ? v : slot + v;
    }
}

__device__ __forceinline__ bool sparse_in_tile(const int* rowptr, const Geo& g, const ZTile& zt) {
    if (!rowptr) return false;
    const int r0 = 2 * zt.q0;
    const int r1 = min(g.nrows, r0 + 2 * zt.ntr);
    return rowptr[r1] > rowptr[r0];
}

// ---------------------------------------------------------------------------------------------
// z kernels
// ---------------------------------------------------------------------------------------------

// Plain forward R2C along z of a real field (times `scale`) into `spec`.
template <typename T>
__device__ __forceinline__ void z_forward_real_body(const BIdx& bi, const T* __restrict__ src, cx<T>* spec, T scale, const Geo& g,
                                                      const ZTile& zt0, const AxisTabs<T>& tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, nthr = blockDim.x;
    ZTile zt = zt0;
    zt.q0 = bi.tile() * zt.ntr;
    cx<T>* b0 = reinterpret_cast<cx<T>*>(smem);
    cx<T>* b1 = b0 + (size_t)g.m2 * zt.pitch;
    KWH_CELL_LOOP(zt, g, tid, nthr, { comp(b0[n * zt.pitch + w], part) = src[cell] * scale; })
    z_forward<T>(spec, g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);   // clears the missing b row
}
template <typename T>
__global__ void __launch_bounds__(256) z_forward_real(const T* __restrict__ src, cx<T>* spec, T scale, Geo g,
                                                      ZTile zt0, AxisTabs<T> tab) {
    z_forward_real_body<T>(bidx_here(), src, spec, scale, g, zt0, tab);
}

// Plain inverse C2R along z: dst = scale * C2R(spec) (accumulate adds to dst).
template <typename T>
__device__ __forceinline__ void z_inverse_real_body(const BIdx& bi, const cx<T>* spec, T* __restrict__ dst, T scale, int accumulate,
                                                      const Geo& g, const ZTile& zt0, const AxisTabs<T>& tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, nthr = blockDim.x;
    ZTile zt = zt0;
    zt.q0 = bi.tile() * zt.ntr;
    cx<T>* b0 = reinterpret_cast<cx<T>*>(smem);
    cx<T>* b1 = b0 + (size_t)g.m2 * zt.pitch;
    cx<T>* R = z_inverse<T>(spec, g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
    __syncthreads();
    KWH_CELL_LOOP(zt, g, tid, nthr, {
        const T v = comp(R[n * zt.pitch + w], part) * scale;
        dst[cell] = accumulate ? dst[cell] + v : v;
    })
}
template <typename T>
__global__ void __launch_bounds__(256) z_inverse_real(const cx<T>* spec, T* __restrict__ dst, T scale, int accumulate,
                                                      Geo g, ZTile zt0, AxisTabs<T> tab) {
    z_inverse_real_body<T>(bidx_here(), spec, dst, scale, accumulate, g, zt0, tab);
}

// Velocity update: u_a = pml_sg_a (pml_sg_a u_a - dt/rho_a * g_a) [+ sources]; then U_a = R2C_z(u_a)
// written over the gradient spectrum (KSpaceFirstOrder.swift:669-678, sources :705-711).
template <typename T> struct VelArgs {
    cx<T>* G[3];           // in: gradient spectra after the outer inverse passes; out: R2C_z(u_a)
    // plane kernels: when gdual is set, axis a reads Gin[a] instead of G[a], and axes 1 and 2 apply
    // d_y (per row) and d_z (per column) of the +dx/2 gradient to the loaded y-z spectrum
    const cx<T>* Gin[3];
    int gdual;
    // plane kernels: the k-space-corrected source of axis corr_axis as its y-z spectrum (after the
    // x pass), inverted in the velocity kernel itself instead of read as a real field (null: none)
    const cx<T>* corr_spec;
    int corr_axis;
    T* u[3];
    SMap<T> dtr[3];        // dt / rho0 staggered along a
    SMap<T> corr[3];       // additive k-space-corrected source contribution (zero scalar when absent)
    SparseSrc<T> sp[3];
    const int* ctr;        // device step counters {t, record slot}
};

template <typename T>
__device__ __forceinline__ void z_velocity_body(const BIdx& bi, const VelArgs<T>& a, const Geo& g, const ZTile& zt0, const AxisTabs<T>& tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, nthr = blockDim.x;
    ZTile zt = zt0;
    zt.q0 = bi.tile() * zt.ntr;
    cx<T>* b0 = reinterpret_cast<cx<T>*>(smem);
    cx<T>* b1 = b0 + (size_t)g.m2 * zt.pitch;
    for (int ax = 0; ax < g.dim; ++ax) {
        const int ia = ax + g.off;
        cx<T>* R = z_inverse<T>(a.G[ax], g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
        cx<T>* other = R == b0 ? b1 : b0;
        __syncthreads();
        const bool sparse = sparse_in_tile(a.sp[ax].rowptr, g, zt);
        T* u = a.u[ax];
        const T* pmls = tab.pmlsg[ia];
        const SMap<T> corr = a.corr[ax];
        const SMap<T> dtr = a.dtr[ax];
        KWH_CELL_LOOP(zt, g, tid, nthr, {
            T& s = comp(R[n * zt.pitch + w], part);
            const T pm = pmls[axis_pos(g, row, n, ia)];
            T v = pm * (pm * u[cell] - dtr.at(cell) * s);
            v += corr.at(cell);
            s = v;
            if (!sparse) u[cell] = v;
        })
        if (sparse) {
            __syncthreads();
            apply_sparse<T>(a.sp[ax], g, zt, R, a.ctr[0], tid, nthr);
            __syncthreads();
            KWH_CELL_LOOP(zt, g, tid, nthr, { u[cell] = comp(R[n * zt.pitch + w], part); })
        }
        z_forward<T>(a.G[ax], g, zt, tab.plan[2], tab.tw[2], R, other, tid, nthr);
        __syncthreads();
    }
}
template <typename T>
__global__ void __launch_bounds__(256) z_velocity(VelArgs<T> a, Geo g, ZTile zt0, AxisTabs<T> tab) {
    z_velocity_body<T>(bidx_here(), a, g, zt0, tab);
}

// Density update + EOS (KSpaceFirstOrder.swift:680-702, sources :716-720).
// mode 0: p = c^2 (rhoT [+ stokes] [+ B/A]) written, and P = R2C_z(p) stored in outP.
// mode 1: power-law absorption follows: stores rhoT, F1 = R2C_z(rho0 divU) [, F2 = R2C_z(rhoT)].
template <typename T> struct RhoArgs {
    const cx<T>* D[3];     // divergence spectra after outer inverse passes
    T* rho[3];
    SMap<T> rho0;
    SMap<T> c0;
    SMap<T> bona;          // nonlinearity, active if nonlinear != 0
    SMap<T> tau;           // Stokes absorption coefficient (mode 0 with stokes != 0)
    SMap<T> corr;          // additive corrected p-source field (base scale); zero scalar when absent
    T corr_factor[3];      // 1/d_a
    SparseSrc<T> sp[3];
    T* p;                  // mode 0 output
    cx<T>* outP;           // mode 0 spectrum output
    T* rhoT;               // mode 1 output
    cx<T>* F1;             // mode 1 outputs
    cx<T>* F2;
    T dt;
    int mode, nonlinear, stokes, dispersion;
    const int* ctr;        // device step counters {t, record slot}
};

template <typename T>
__device__ __forceinline__ void z_density_body(const BIdx& bi, const RhoArgs<T>& a, const Geo& g, const ZTile& zt0, const AxisTabs<T>& tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, nthr = blockDim.x;
    ZTile zt = zt0;
    zt.q0 = bi.tile() * zt.ntr;
    const size_t bl = (size_t)g.m2 * zt.pitch;
    cx<T>* b0 = reinterpret_cast<cx<T>*>(smem);
    cx<T>* b1 = b0 + bl;
    // LDS: b0, b1, RT = new total density, DIV = sum of du (Stokes / absorption inputs only),
    // RP = previous total density (nonlinear only); a/b rows in .x/.y
    const bool need_div = a.mode != 0 || a.stokes;
    cx<T>* RT = b1 + bl;
    cx<T>* DIV = need_div ? RT + bl : nullptr;
    cx<T>* RP = RT + (need_div ? 2 : 1) * bl;
    if (a.nonlinear) {
        KWH_CELL_LOOP(zt, g, tid, nthr, {
            T s = 0;
            for (int ax = 0; ax < g.dim; ++ax) s += a.rho[ax][cell];
            comp(RP[n * zt.pitch + w], part) = s;
        })
    }
    for (int ax = 0; ax < g.dim; ++ax) {
        const int ia = ax + g.off;
        cx<T>* R = z_inverse<T>(a.D[ax], g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
        __syncthreads();
        const bool sparse = sparse_in_tile(a.sp[ax].rowptr, g, zt);
        T* rho = a.rho[ax];
        const T* pm_ = tab.pml[ia];
        const T cf = a.corr_factor[ax];
        KWH_CELL_LOOP(zt, g, tid, nthr, {
            T& s = comp(R[n * zt.pitch + w], part);
            const T du = s;
            const T r0 = a.rho0.at(cell);
            T mass = a.dt * r0 * du;
            if (a.nonlinear) mass *= (T(2) * comp(RP[n * zt.pitch + w], part) + r0) / r0;
            const T pm = pm_[axis_pos(g, row, n, ia)];
            T v = pm * (pm * rho[cell] - mass);
            v += a.corr.at(cell) * cf;
            s = v;
            if (!sparse) rho[cell] = v;
            if (need_div) {
                T& dv = comp(DIV[n * zt.pitch + w], part);
                dv = ax == 0 ? du : dv + du;
            }
            T& rt = comp(RT[n * zt.pitch + w], part);
            if (!sparse) rt = ax == 0 ? v : rt + v;
        })
        if (sparse) {
            __syncthreads();
            apply_sparse<T>(a.sp[ax], g, zt, R, a.ctr[0], tid, nthr);
            __syncthreads();
            KWH_CELL_LOOP(zt, g, tid, nthr, {
                const T v = comp(R[n * zt.pitch + w], part);
                rho[cell] = v;
                T& rt = comp(RT[n * zt.pitch + w], part);
                rt = ax == 0 ? v : rt + v;
            })
        }
        __syncthreads();
    }
    if (a.mode == 0) {
        KWH_CELL_LOOP(zt, g, tid, nthr, {
            const T rt = comp(RT[n * zt.pitch + w], part);
            T e = rt;
            const T r0 = a.rho0.at(cell);
            if (a.stokes) e += a.tau.at(cell) * r0 * comp(DIV[n * zt.pitch + w], part);
            if (a.nonlinear) e += a.bona.at(cell) * (rt * rt) / (T(2) * r0);
            const T c = a.c0.at(cell);
            const T p = c * c * e;
            a.p[cell] = p;
            comp(b0[n * zt.pitch + w], part) = p;
        })
        z_forward<T>(a.outP, g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
    } else {
        KWH_CELL_LOOP(zt, g, tid, nthr, {
            const T rt = comp(RT[n * zt.pitch + w], part);
            a.rhoT[cell] = rt;
            comp(b0[n * zt.pitch + w], part) = a.rho0.at(cell) * comp(DIV[n * zt.pitch + w], part);
        })
        z_forward<T>(a.F1, g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
        if (a.dispersion) {
            __syncthreads();
            KWH_CELL_LOOP(zt, g, tid, nthr, { comp(b0[n * zt.pitch + w], part) = comp(RT[n * zt.pitch + w], part); })
            z_forward<T>(a.F2, g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
        }
    }
}
template <typename T>
__global__ void __launch_bounds__(256) z_density(RhoArgs<T> a, Geo g, ZTile zt0, AxisTabs<T> tab) {
    z_density_body<T>(bidx_here(), a, g, zt0, tab);
}

// EOS with power-law absorption/dispersion (KSpaceFirstOrder.swift:836-845, :694-702):
// p = c^2 (rhoT + tau * A1 - eta * A2 + B/A rhoT^2 / (2 rho0)); P = R2C_z(p).
// Adjoint row kernels (reverse sweep of the lossless linear loop; the recursion is documented in
// kwh_adjoint.h): a diagonal update fused into a contiguous-axis transform.
//   forward mode 0: w = scale r s^spow lu;  mode 1: lr = c^2 lp + s^2 lr (stored), w = dt rho0 s lr;
//   then spec = R2C_z(w).
//   inverse mode 0: dst = C2R_z(spec); 1: dst += C2R_z(spec); 2: dst = s^2 dst + C2R_z(spec).
// s is the PML profile along internal axis ia (staggered for the velocity-side terms).
template <typename T> struct AdjRowArgs {
    const T* lu;
    T* lr;
    const T* lp;
    SMap<T> r;      // dt / rho0 staggered along the axis
    SMap<T> c0, rho0;
    const T* s;
    int ia;
    int spow;
    T scale, dt;
    cx<T>* spec;
};

template <typename T> struct EosArgs {
    const cx<T>* A1;
    const cx<T>* A2;       // null without dispersion
    const T* rhoT;
    SMap<T> c0, rho0, tau, eta, bona;
    int nonlinear;
    // 1: tau = tk c^te and eta = ek c^ee from the cell's sound speed (scalar alphaCoeff over a
    // sound-speed map), instead of reading precomputed tau / eta maps
    int tau_eta_from_c;
    T tk, te, ek, ee;
    T* p;
    cx<T>* outP;
};

// Absorption coefficients of one cell (makeAbsorption, KSpaceFirstOrder.swift:855-878).
template <typename T>
__device__ __forceinline__ void eos_tau_eta(const EosArgs<T>& a, int64_t cell, T c, T& tau, T& eta) {
    if (a.tau_eta_from_c) {
        const T l = log2(c);
        tau = a.tk * exp2(a.te * l);
        eta = a.ek * exp2(a.ee * l);
    } else {
        tau = a.tau.at(cell);
        eta = a.eta.at(cell);
    }
}

template <typename T>
__device__ __forceinline__ void z_eos_body(const BIdx& bi, const EosArgs<T>& a, const Geo& g, const ZTile& zt0, const AxisTabs<T>& tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, nthr = blockDim.x;
    ZTile zt = zt0;
    zt.q0 = bi.tile() * zt.ntr;
    const size_t bl = (size_t)g.m2 * zt.pitch;
    cx<T>* b0 = reinterpret_cast<cx<T>*>(smem);
    cx<T>* b1 = b0 + bl;
    cx<T>* E = b1 + bl;
    cx<T>* R = z_inverse<T>(a.A1, g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
    __syncthreads();
    KWH_CELL_LOOP(zt, g, tid, nthr, {
        T tau = 0;
        T eta = 0;
        eos_tau_eta(a, cell, a.c0.at(cell), tau, eta);
        comp(E[n * zt.pitch + w], part) = a.rhoT[cell] + tau * comp(R[n * zt.pitch + w], part);
    })
    __syncthreads();
    if (a.A2) {
        R = z_inverse<T>(a.A2, g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
        __syncthreads();
        KWH_CELL_LOOP(zt, g, tid, nthr, {
            T tau = 0;
            T eta = 0;
            eos_tau_eta(a, cell, a.c0.at(cell), tau, eta);
            comp(E[n * zt.pitch + w], part) -= eta * comp(R[n * zt.pitch + w], part);
        })
        __syncthreads();
    }
    KWH_CELL_LOOP(zt, g, tid, nthr, {
        T e = comp(E[n * zt.pitch + w], part);
        if (a.nonlinear) {
            const T rt = a.rhoT[cell];
            e += a.bona.at(cell) * (rt * rt) / (T(2) * a.rho0.at(cell));
        }
        const T c = a.c0.at(cell);
        const T p = c * c * e;
        a.p[cell] = p;
        comp(b0[n * zt.pitch + w], part) = p;
    })
    z_forward<T>(a.outP, g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
}
template <typename T>
__global__ void __launch_bounds__(256) z_eos(EosArgs<T> a, Geo g, ZTile zt0, AxisTabs<T> tab) {
    z_eos_body<T>(bidx_here(), a, g, zt0, tab);
}

// t = 0 initial-pressure override (KSpaceFirstOrder.swift:724-733):
// p = p0; rho_a = p0 / (dim c^2); u_a = dt/rho_a / 2 * Re ifft(ddxPos_a kappa fft p0); P = R2C_z(p0).
template <typename T> struct InitArgs {
    const cx<T>* G[3];
    T* u[3];
    T* rho[3];
    SMap<T> dtr[3];
    SMap<T> c0;
    const T* p0;
    T* p;
    cx<T>* outP;
};

template <typename T>
__device__ __forceinline__ void z_init_body(const BIdx& bi, const InitArgs<T>& a, const Geo& g, const ZTile& zt0, const AxisTabs<T>& tab) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x, nthr = blockDim.x;
    ZTile zt = zt0;
    zt.q0 = bi.tile() * zt.ntr;
    cx<T>* b0 = reinterpret_cast<cx<T>*>(smem);
    cx<T>* b1 = b0 + (size_t)g.m2 * zt.pitch;
    for (int ax = 0; ax < g.dim; ++ax) {
        cx<T>* R = z_inverse<T>(a.G[ax], g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);
        __syncthreads();
        T* u = a.u[ax];
        const SMap<T> dtr = a.dtr[ax];
        KWH_CELL_LOOP(zt, g, tid, nthr, { u[cell] = dtr.at(cell) * T(0.5) * comp(R[n * zt.pitch + w], part); })
        __syncthreads();
    }
    const T nd = T(g.dim);
    KWH_CELL_LOOP(zt, g, tid, nthr, {
        const T v = a.p0[cell];
        const T c = a.c0.at(cell);
        const T r = v / (nd * (c * c));
        for (int ax = 0; ax < g.dim; ++ax) a.rho[ax][cell] = r;
        a.p[cell] = v;
        comp(b0[n * zt.pitch + w], part) = v;
    })
    z_forward<T>(a.outP, g, zt, tab.plan[2], tab.tw[2], b0, b1, tid, nthr);   // clears the missing b row
}
template <typename T>
__global__ void __launch_bounds__(256) z_init(InitArgs<T> a, Geo g, ZTile zt0, AxisTabs<T> tab) {
    z_init_body<T>(bidx_here(), a, g, zt0, tab);
}

// ---------------------------------------------------------------------------------------------
// Elementwise / utility kernels
// ---------------------------------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ void fill_kernel_body(const BIdx& bi, T* p, T v, int64_t n) {
    for (int64_t i = bi.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)bi.nx * blockDim.x)
        p[i] = v;
}
template <typename T>
__global__ void fill_kernel(T* p, T v, int64_t n) {
    fill_kernel_body<T>(bidx_here(), p, v, n);
}

template <typename S, typename D>
__global__ void convert_kernel(const S* s, D* d, D scale, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = D(s[i]) * scale;
}

// max |x| -> atomicMax on the bit pattern of a non-negative double.
template <typename T>
__global__ void maxabs_kernel(const T* x, int64_t n, unsigned long long* out) {
    __shared__ double red[256];
    double m = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs((double)x[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

// Scatter scaled source values into a zeroed real field (k-space-corrected additive sources).
template <typename T>
__device__ __forceinline__ void scatter_kernel_body(const BIdx& bi, T* dst, const int64_t* index, const T* signal, const T* base, int rows, int len, int n,
                               const int* ctr, T factor) {
    const int e = bi.x * blockDim.x + threadIdx.x;
    const int t = ctr[0];
    if (e >= n || t >= len) return;
    dst[index[e]] = signal[(int64_t)(rows == 1 ? 0 : e) * len + t] * base[e] * factor;
}
template <typename T>
__global__ void scatter_kernel(T* dst, const int64_t* index, const T* signal, const T* base, int rows, int len, int n,
                               const int* ctr, T factor) {
    scatter_kernel_body<T>(bidx_here(), dst, index, signal, base, rows, len, n, ctr, factor);
}

// End of a time step: t += 1 and, when the step recorded, slot += 1. Kernels read both from
// here, so a captured step graph replays unchanged.
static __global__ void set_counters_kernel(int* ctr, int t, int slot) {
    if (threadIdx.x == 0) {
        ctr[0] = t;
        ctr[1] = slot;
    }
}

__device__ __forceinline__ void advance_kernel_body(const BIdx& bi, int* ctr, int recorded) {
    if (threadIdx.x == 0) {
        ctr[0] += 1;
        ctr[1] += recorded;
    }
}
static __global__ void advance_kernel(int* ctr, int recorded) {
    advance_kernel_body(bidx_here(), ctr, recorded);
}

// Sensor gather + streaming aggregates (Recording.swift:37-45, :139-156).
template <typename T> struct GatherArgs {
    const T* field;
    const int64_t* idx;    // [n * ncorner]
    const T* weight;       // null for binary masks
    int ncorner;
    int64_t n;
    T* series;             // [nt][n] or null
    T* vmax;
    T* vmin;
    double* sum2;
    const int* ctr;        // device step counters {t, record slot}; slot 0 initialises max/min
};

// Up to four fields (p, ux, uy, uz) sampled by one launch: block y selects the field.
template <typename T> struct GatherSet {
    GatherArgs<T> f[4];
};

// Running max / min that keep a NaN once one has been seen (the semantics of a max / min over the
// whole series, as numpy's), so a blow-up shows in pMax / pMin / uMax instead of leaving them at
// their last finite values.
template <typename T>
__device__ __forceinline__ T nan_max(T m, T v) { return (v > m || v != v) ? v : m; }
template <typename T>
__device__ __forceinline__ T nan_min(T m, T v) { return (v < m || v != v) ? v : m; }

template <typename T>
__device__ __forceinline__ void gather_kernel_body(const BIdx& bi, const GatherSet<T>& set) {
    const GatherArgs<T>& a = set.f[bi.y];
    const int slot = a.ctr[1];
    const bool first = slot == 0;
    for (int64_t s = bi.x * (int64_t)blockDim.x + threadIdx.x; s < a.n; s += (int64_t)bi.nx * blockDim.x) {
        T v;
        if (a.weight) {
            v = T(0);
            for (int c = 0; c < a.ncorner; ++c) v += a.field[a.idx[s * a.ncorner + c]] * a.weight[s * a.ncorner + c];
        } else {
            v = a.field[a.idx[s]];
        }
        if (a.series) a.series[(int64_t)slot * a.n + s] = v;
        if (a.vmax) a.vmax[s] = first ? v : nan_max(a.vmax[s], v);
        if (a.vmin) a.vmin[s] = first ? v : nan_min(a.vmin[s], v);
        if (a.sum2) a.sum2[s] = (first ? 0.0 : a.sum2[s]) + (double)v * (double)v;
    }
}
template <typename T>
__global__ void gather_kernel(GatherSet<T> set) {
    gather_kernel_body<T>(bidx_here(), set);
}

// Gather through output positions (a subset of sensor rows, e.g. one directivity angle).
template <typename T>
__device__ __forceinline__ void gather_pos_kernel_body(const BIdx& bi, const T* field, const int64_t* idx, const int64_t* pos, int64_t n, int64_t nsens,
                                  T* series, T* vmax, T* vmin, double* sum2, const int* ctr) {
    const int t = ctr[1];
    const bool first = t == 0;
    for (int64_t s = bi.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)bi.nx * blockDim.x) {
        const T v = field[idx[s]];
        const int64_t q = pos[s];
        if (series) series[(int64_t)t * nsens + q] = v;
        if (vmax) vmax[q] = first ? v : nan_max(vmax[q], v);
        if (vmin) vmin[q] = first ? v : nan_min(vmin[q], v);
        if (sum2) sum2[q] = (first ? 0.0 : sum2[q]) + (double)v * (double)v;
    }
}
template <typename T>
__global__ void gather_pos_kernel(const T* field, const int64_t* idx, const int64_t* pos, int64_t n, int64_t nsens,
                                  T* series, T* vmax, T* vmin, double* sum2, const int* ctr) {
    gather_pos_kernel_body<T>(bidx_here(), field, idx, pos, n, nsens, series, vmax, vmin, sum2, ctr);
}

// Non-finite check of a field after step `step`: flag[0] = min(flag[0], step) if any value is a
// NaN or an infinity (flag starts at INT_MAX).
template <typename T>
__global__ void nonfinite_kernel(const T* x, int64_t n, int step, int* flag) {
    bool bad = false;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        bad |= !isfinite(x[i]);
    if (bad) atomicMin(flag, step);
}

// pRms / uRms = sqrt(sum2 / nt).
template <typename T>
__global__ void rms_kernel(const double* sum2, T* out, int64_t n, double inv_nt) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x)
        out[s] = T(sqrt(sum2[s] * inv_nt));
}

// Receive combination (KWaveTransducer.swift:263-289) over a time-major series [nt][n]:
// out[e][t] = scale * sum_{k in element e} series[t + delay[k]][k], t + delay[k] < nt.
// rowptr/rows list the sensor points of each element (CSR).
template <typename T>
__global__ void element_traces_kernel(const T* series, int64_t n, int nt, const int* rowptr, const int* rows,
                                      const int* delay, int nel, double scale, T* out) {
    const int64_t total = (int64_t)nel * nt;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int e = (int)(i / nt);
        const int t = (int)(i - (int64_t)e * nt);
        double acc = 0.0;
        for (int q = rowptr[e]; q < rowptr[e + 1]; ++q) {
            const int k = rows[q];
            const int64_t ts = (int64_t)t + delay[k];
            if (ts < nt) acc += (double)series[ts * n + k];
        }
        out[i] = T(acc * scale);
    }
}

// Time reversal (TimeReversal.swift:33-50): reverse every source row in time, in place ...
template <typename T>
__global__ void reverse_rows_kernel(T* sig, int rows, int len) {
    const int64_t half = (int64_t)rows * (len / 2);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < half; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / (len / 2);
        const int t = (int)(i - r * (len / 2));
        T* row = sig + r * len;
        const T a = row[t];
        row[t] = row[len - 1 - t];
        row[len - 1 - t] = a;
    }
}

// ... and the estimate max(0, factor * pFinal).
template <typename T>
__global__ void scaled_positive_kernel(const T* p, T* out, T factor, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T v = factor * p[i];
        out[i] = v > T(0) ? v : T(0);
    }
}

// Intensity average (Recording.swift:113-124) from stored series [nt][n]:
// I = (1/nt) sum_t p[t] * Re(ifft(e^{i w/2} fft(u)))[t] = (1/nt^2) Re sum_k conj(P_k) S_k U_k.
// One thread per sensor point, direct DFT in float64 (cost n * nt^2).
template <typename T>
__device__ double intensity_point(const T* p, const T* u, int64_t n, int nt, int64_t s) {
    const double two_pi = 6.283185307179586476925286766559;
    double acc = 0.0;
    const int half = nt / 2;
    for (int k = 0; k < nt; ++k) {
        double pr = 0, pi = 0, ur = 0, ui = 0;
        for (int t = 0; t < nt; ++t) {
            const double ang = -two_pi * (double)((int64_t)k * t % nt) / nt;
            double sn, cs;
            sincos(ang, &sn, &cs);
            const double pv = (double)p[(int64_t)t * n + s], uv = (double)u[(int64_t)t * n + s];
            pr += pv * cs; pi += pv * sn;
            ur += uv * cs; ui += uv * sn;
        }
        const int kk = k <= half ? k : k - nt;
        const double th = two_pi * kk / nt * 0.5;
        double ssn, scs;
        sincos(th, &ssn, &scs);
        // S U
        const double sur = scs * ur - ssn * ui, sui = scs * ui + ssn * ur;
        // Re(conj(P) * SU)
        acc += pr * sur + pi * sui;
    }
    return acc / ((double)nt * (double)nt);
}

template <typename T>
__global__ void intensity_kernel(const T* p, const T* u, T* out, int64_t n, int nt) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x)
        out[s] = T(intensity_point(p, u, n, nt, s));
}

// The same for sensors [s0, s0 + cnt) into out[0, cnt).
template <typename T>
__global__ void intensity_range_kernel(const T* p, const T* u, T* out, int64_t n, int nt, int64_t s0, int64_t cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = T(intensity_point(p, u, n, nt, s0 + i));
}

// Recorded series, time-major [nt][n], to sensor-major rows [cnt][nt] for sensors [s0, s0 + cnt)
// (the [nSensor, nt] layout of SimulationOutput).
template <typename T>
__global__ void series_rows_kernel(const T* series, int64_t n, int nt, int64_t s0, int64_t cnt, T* out) {
    const int64_t total = cnt * nt;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / nt;
        const int t = (int)(i - r * nt);
        out[i] = series[(int64_t)t * n + s0 + r];
    }
}

}  // namespace kwh
