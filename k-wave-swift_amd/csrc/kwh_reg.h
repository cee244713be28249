// kwh_reg.h -- register-resident kernels for axis lengths L = J*E with J | E
// (J, E) = (4, 8): 32, (8, 8): 64, (4, 24): 96, (8, 16): 128, (4, 40): 160, (8, 24): 192,
// (16, 16): 256, (16, 32): 512, (32, 32): 1024.
//
// A length-L transform is a two-step Cooley-Tukey split L = J * E. Lane j (0..J-1) of a
// transform holds elements n = j + J*m (m = 0..E-1) in registers:
//   step 1: DFT_E over m in registers -> Y[n1 = j][k2], k2 = 0..E-1, written to LDS;
//   step 2: lane j takes the columns k2 = j + J*a (a = 0..E/J-1): reads Y[n1][k2] for n1 =
//           0..J-1, twiddles W_L^(n1*k2), DFT_J over n1 -> X[k2 + E*k1] = element
//           j + J*(a + (E/J)*k1) of the lane-natural pattern.
// Input and output of a transform are therefore both in the lane-natural pattern n = j + J*m, so
// an inverse, an elementwise update and a forward chain without extra LDS traffic, and the
// outputs of a forward feed the next inverse directly (middle pass). One LDS transpose per
// transform; every lane busy in both steps. For J = E (64, 256) this is two radix-R stages.
#pragma once
#include "kwh_kernels.h"
#include "kwh_xlane.h"

// Occupancy floors (waves per SIMD) of the fused row kernels with E <= 16 in FP32: without them
// the register allocator spends the whole register file on load hoisting and drops a wave per
// SIMD. The density kernels that also accumulate rho0 * div u stay at two. Longer rows (E > 16)
// hold more registers per lane and carry no floor (at E = 32 a floor of two costs the EOS and the
// density kernels scratch spills and measured no faster).
#ifndef KWH_DENSITY_WAVES
#define KWH_DENSITY_WAVES 3
#endif
#ifndef KWH_VELOCITY_WAVES
#define KWH_VELOCITY_WAVES 4
#endif
#ifndef KWH_MIDDLE_WAVES
#define KWH_MIDDLE_WAVES 3
#endif
#ifndef KWH_EOS_WAVES
#define KWH_EOS_WAVES 3
#endif
// Loads issued together with a kernel's first spectrum load, so their latency overlaps the
// transform instead of following it (E <= 16 FP32): the EOS kernel's sum rho and c0 (it then runs
// at KWH_EOS_PRELOAD_WAVES), the FFT-absorption density kernel's rho_a, the middle pass's operator
// table. Measured at 256^3: EOS 102 -> 95 us, density 176 -> 167 us, middle passes -1..2 us.
#ifndef KWH_EOS_PRELOAD
#define KWH_EOS_PRELOAD 1
#endif
#ifndef KWH_EOS_PRELOAD_WAVES
#define KWH_EOS_PRELOAD_WAVES 2
#endif
#ifndef KWH_DENSITY_PRELOAD
#define KWH_DENSITY_PRELOAD 1
#endif
#ifndef KWH_MIDDLE_KAPPA_EARLY
#define KWH_MIDDLE_KAPPA_EARLY 1
#endif
#ifndef KWH_DENSITY_HOLD_RHO0
#define KWH_DENSITY_HOLD_RHO0 1
#endif
// Element loops versioned on run-time switches that are uniform over a launch (operator kind of a
// middle pass; the EOS kernel's tau / eta source, dispersion and B/A), so no element carries a
// branch.
#ifndef KWH_MIDDLE_VERSIONED
#define KWH_MIDDLE_VERSIONED 1
#endif
#ifndef KWH_EOS_VERSIONED
#define KWH_EOS_VERSIONED 1
#endif

// Largest J whose middle pass uses the cross-lane (DPP) inner radix by default: none, as it
// measured slower than the LDS transpose at every length (KWH_MIDDLE_XLANE=1 opts in).
#ifndef KWH_MIDDLE_XLANE_MAXJ
#define KWH_MIDDLE_XLANE_MAXJ 0
#endif

namespace kwh {

template <typename T, int E>
constexpr int reg_waves(int w) { return (sizeof(T) == 4 && E <= 16) ? w : 1; }

// Twiddles W_L^(n1*k2) of step 2, in an LDS table laid out [n1][k2] (L elements at the start of
// the workgroup's LDS): lane j reads slot (j + J*a) + E*n1, a base + immediate address, and the
// table costs no registers between transforms. The inverse uses the conjugates. The table is
// filled cooperatively at kernel start; the barrier inside the first transform orders the fill
// before any read.
template <typename T, int J, int E>
struct LaneTw {
    static constexpr int L = J * E;
    const cx<T>* p;
    // tw: W_(S*L)^i; S = 2 for the half-row transforms, which read W_L from the row's W_2L table
    template <int S = 1>
    __device__ __forceinline__ void load(cx<T>* tab_lds, const cx<T>* __restrict__ tw, int j) {
        for (int i = threadIdx.x; i < L; i += blockDim.x) tab_lds[i] = tw[S * ((i % E) * (i / E))];
        p = tab_lds + j;
    }
    __device__ __forceinline__ cx<T> w(int a, int n1) const { return p[J * a + E * n1]; }
};

// Step 2 of a transform: lane j reads Y[n1][j + J*a] at rd[n1*RS + J*a*CS] (RS, CS: the LDS
// strides of n1 and of one k2), twiddles, DFT_J over n1, and keeps X[k2 + E*k1] in
// v[a + (E/J)*k1].
template <typename T, int J, int E, int RS, int CS, bool INV>
__device__ __forceinline__ void reg_step2(cx<T>* v, const cx<T>* rd, const LaneTw<T, J, E>& tw) {
    constexpr int Q = E / J;
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        cx<T> s[J];
#pragma unroll
        for (int n1 = 0; n1 < J; ++n1) {
            cx<T> x = rd[n1 * RS + J * a * CS];
            if (n1 > 0) x = cmul(x, INV ? conjv(tw.w(a, n1)) : tw.w(a, n1));
            s[n1] = x;
        }
        dftN<T, J, INV>(s);
#pragma unroll
        for (int k1 = 0; k1 < J; ++k1) v[a + Q * k1] = s[k1];
    }
}

// ------------------------------------------------------------------------------------------
// Contiguous-axis transforms of a row pair (a + i b), lanes j of transform w in a workgroup.
// LDS per transform: LP = L + E + 1 complex.
// ------------------------------------------------------------------------------------------
template <typename T, int J, int E>
struct RegRow {
    static_assert(E % J == 0 && E % 2 == 0, "lanes per transform must divide the (even) elements per lane");
    static constexpr int L = J * E;
    static constexpr int LP = L + E + 1;
    // padded LDS slot of element n is n + n / J; for the access pattern used (n = j + J*m):
    // j + (J+1)*m. Step 1 writes Y[n1][k2] at n1*(E+1) + k2.
    __device__ static __forceinline__ int pidx(int n) { return n + n / J; }

    // Two-step transform of v (lane-natural in, lane-natural out). lds: this transform's slab.
    template <bool INV>
    __device__ static __forceinline__ void transform(cx<T>* v, cx<T>* lds, int j, const LaneTw<T, J, E>& tw) {
        dftN<T, E, INV>(v);
        cx<T>* wr = lds + j * (E + 1);
#pragma unroll
        for (int r = 0; r < E; ++r) wr[r] = v[r];
        __syncthreads();
        reg_step2<T, J, E, E + 1, 1, INV>(v, lds + j, tw);
        __syncthreads();
    }

    // Load the Hermitian-extended Z = A + iB of rows (ra, rb) in the lane-natural pattern.
    // Elements n = j + J m < L/2 read k = n from base (j); n >= L/2 read k = L - n from base
    // (L - j) with negative compile-time offsets, so every load is base + immediate.
    template <bool FULL = false>
    __device__ static __forceinline__ void load_spectrum(const cx<T>* __restrict__ spec, int ra, int rb, int nrows,
                                                         int mzh, int j, cx<T>* v) {
        const bool va = FULL || ra < nrows, vb = FULL || rb < nrows;
        const T ma = va ? T(1) : T(0), mb = vb ? T(1) : T(0);
        const cx<T>* sa = spec + (int64_t)(va ? ra : 0) * mzh;
        const cx<T>* sb = spec + (int64_t)(vb ? rb : 0) * mzh;
        const cx<T>* la = sa + j;
        const cx<T>* lb = sb + j;
        const cx<T>* ha = sa + (L - j);
        const cx<T>* hb = sb + (L - j);
#pragma unroll
        for (int m = 0; m < E / 2; ++m) {
            const cx<T> A = la[J * m] * ma;
            const cx<T> B = lb[J * m] * mb;
            v[m] = cx<T>{A.x - B.y, A.y + B.x};
        }
#pragma unroll
        for (int m = E / 2; m < E; ++m) {
            const cx<T> A = ha[-J * m] * ma;
            const cx<T> B = hb[-J * m] * mb;
            v[m] = cx<T>{A.x + B.y, B.x - A.y};
        }
        if (j == 0) {   // DC (n = 0) and Nyquist (n = L/2) bins are real parts only
            v[0] = cx<T>{sa[0].x * ma, sb[0].x * mb};
            v[E / 2] = cx<T>{sa[L / 2].x * ma, sb[L / 2].x * mb};
        }
    }

    // Forward of v (a in .x, b in .y), split into half spectra, stored to rows (ra, rb).
    template <bool FULL = false>
    __device__ static __forceinline__ void forward_store(cx<T>* spec, int ra, int rb, int nrows, int mzh, cx<T>* lds,
                                                         int j, const LaneTw<T, J, E>& tw, cx<T>* v) {
        transform<false>(v, lds, j, tw);
        cx<T>* wr = lds + j;
#pragma unroll
        for (int m = 0; m < E; ++m) wr[(J + 1) * m] = v[m];
        __syncthreads();
        const bool va = FULL || ra < nrows, vb = FULL || rb < nrows;
        // k = j + J m: slot j + (J+1) m; mirror L - k: slot (L - j + E - [j>0]) - (J+1) m (k > 0)
        const cx<T>* pk = lds + j;
        const cx<T>* pm = lds + (L - j + E - (j > 0 ? 1 : 0));
        cx<T>* oa = spec + (int64_t)(va ? ra : 0) * mzh + j;
        cx<T>* ob = spec + (int64_t)(vb ? rb : 0) * mzh + j;
#pragma unroll
        for (int m = 0; m <= E / 2; ++m) {
            const int k = j + J * m;
            if (k > L / 2) break;
            const cx<T> zk = pk[(J + 1) * m];
            const cx<T> zm = k == 0 ? zk : pm[-(J + 1) * m];
            if (va) oa[J * m] = cx<T>{T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
            if (vb) ob[J * m] = cx<T>{T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
        }
        __syncthreads();
    }
};

// ------------------------------------------------------------------------------------------
// One real row of length 2L per length-L complex transform (even/odd split): z[n] = x[2n] +
// i x[2n+1], Z = DFT_L(z); E[k] = (Z[k] + conj Z[L-k]) / 2 and O[k] = (Z[k] - conj Z[L-k]) / 2i
// are the DFTs of the even and odd samples, and X[k] = E[k] + W_2L^k O[k] (k = 0..L; E and O are
// L-periodic). Against a row pair as one length-2L transform this holds half the elements per
// lane: length-512 rows run as 256-point transforms (16 x 16) instead of 512-point ones (16 x 32).
// tw2: W_2L^i (i < 2L), the row's own twiddle table.
// ------------------------------------------------------------------------------------------
template <typename T, int J, int E>
struct RegHalf {
    using R = RegRow<T, J, E>;
    static constexpr int L = J * E;

    // Z'[k] = X[k] + conj X[L-k] + i W_2L^-k (X[k] - conj X[L-k]) = 2 (E[k] + i O[k]), k = j + J m;
    // its unnormalised inverse is 2L (x[2n] + i x[2n+1]), the unnormalised C2R of X with the DC
    // and Nyquist bins taken as real.
    template <bool FULL = false>
    __device__ static __forceinline__ void load_spectrum(const cx<T>* __restrict__ spec, int row, int nrows, int mzh,
                                                         int j, const cx<T>* __restrict__ tw2, cx<T>* v) {
        const bool ok = FULL || row < nrows;
        const T mk = ok ? T(1) : T(0);
        const cx<T>* s = spec + (int64_t)(ok ? row : 0) * mzh;
        const cx<T>* lo = s + j;
        const cx<T>* hi = s + (L - j);
        const cx<T>* w = tw2 + j;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const cx<T> A = lo[J * m], B = hi[-J * m], c = w[J * m];
            const T sx = A.x + B.x, sy = A.y - B.y;       // A + conj B
            const T dx = A.x - B.x, dy = A.y + B.y;       // A - conj B
            const T fx = dx * c.x + dy * c.y, fy = dy * c.x - dx * c.y;   // (A - conj B) conj(c)
            v[m] = cx<T>{(sx - fy) * mk, (sy + fx) * mk};
        }
        if (j == 0) {
            const T x0 = s[0].x * mk, xn = s[L].x * mk;
            v[0] = cx<T>{x0 + xn, x0 - xn};
        }
    }

    // Forward of v = z (lane-natural), then X[k] for k = j + J m and, from lane 0, X[L].
    template <bool FULL = false>
    __device__ static __forceinline__ void forward_store(cx<T>* spec, int row, int nrows, int mzh, cx<T>* lds, int j,
                                                         const LaneTw<T, J, E>& tw, const cx<T>* __restrict__ tw2,
                                                         cx<T>* v) {
        R::template transform<false>(v, lds, j, tw);
        cx<T>* wr = lds + j;
#pragma unroll
        for (int m = 0; m < E; ++m) wr[(J + 1) * m] = v[m];
        __syncthreads();
        const bool ok = FULL || row < nrows;
        // slot of Z[L - k], k = j + J m > 0: (L - j + E - [j > 0]) - (J + 1) m
        const cx<T>* pm = lds + (L - j + E - (j > 0 ? 1 : 0));
        cx<T>* o = spec + (int64_t)(ok ? row : 0) * mzh + j;
        const cx<T>* w = tw2 + j;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const int k = j + J * m;
            const cx<T> A = v[m];
            const cx<T> B = k == 0 ? A : pm[-(J + 1) * m];
            const T ex = T(0.5) * (A.x + B.x), ey = T(0.5) * (A.y - B.y);   // E[k]
            const T ox = T(0.5) * (A.y + B.y), oy = T(0.5) * (B.x - A.x);   // O[k]
            const cx<T> c = w[J * m];
            if (ok) o[J * m] = cx<T>{ex + c.x * ox - c.y * oy, ey + c.x * oy + c.y * ox};
        }
        if (j == 0 && ok) o[L] = cx<T>{v[0].x - v[0].y, T(0)};   // X[L] = E[0] - O[0]
        __syncthreads();
    }
};

// Rows of one lane group in the fused row kernels (velocity, density, EOS).
//   Row-pair form (HALF = false): rows 2q and 2q + 1 are the real and imaginary parts of one
//   length-L transform; element m of lane j is sample n = j + J m of each row.
//   Half form: one row q of length 2L; element m of lane j holds samples 2n (component a) and
//   2n + 1 (component b), n = j + J m.
// P lane groups per workgroup; rows outside the grid are clamped to row 0 and masked (va / vb).
template <int J, int E, bool HALF>
struct ZRows {
    static constexpr int L = J * E;
    static constexpr int RL = HALF ? 2 * L : L;   // samples per row
    static constexpr int SJ = HALF ? 2 * J : J;   // sample stride between a lane's elements
    int w, j;
    int qa, qb;          // rows of components a and b
    bool va, vb;
    int64_t oa, ob;      // sample offsets of element 0 (rows clamped)
    int na, nb;          // position of element 0 along the row
    int row0, nrow;      // first row and number of rows of the workgroup's tile
    template <bool FULL>
    __device__ __forceinline__ void init(int P, const BIdx& bi, int nrows) {
        const int tid = threadIdx.x;
        w = tid / J;
        j = tid - w * J;
        const int tile = bi.tile();
        if constexpr (HALF) {
            qa = qb = tile * P + w;
            row0 = tile * P;
            nrow = P;
            na = 2 * j;
            nb = 2 * j + 1;
        } else {
            qa = 2 * (tile * P + w);
            qb = qa + 1;
            row0 = 2 * tile * P;
            nrow = 2 * P;
            na = nb = j;
        }
        va = FULL || qa < nrows;
        vb = FULL || qb < nrows;
        oa = (int64_t)(va ? qa : 0) * RL + na;
        ob = (int64_t)(vb ? qb : 0) * RL + nb;
    }
};
template <bool HALF>
__device__ __forceinline__ bool zrows_tile_full(const BIdx& bi, int P, int nrows) {
    return ((int64_t)bi.tile() + 1) * P * (HALF ? 1 : 2) <= nrows;
}
__device__ __forceinline__ bool rows_have_sparse(const int* rowptr, const Geo& g, int row0, int nrow) {
    if (!rowptr) return false;
    return rowptr[min(g.nrows, row0 + nrow)] > rowptr[row0];
}
template <typename T, int J, int E, bool HALF, bool FULL>
__device__ __forceinline__ void zrows_load(const cx<T>* spec, const ZRows<J, E, HALF>& ps, const Geo& g,
                                           const AxisTabs<T>& tab, cx<T>* v) {
    if constexpr (HALF) RegHalf<T, J, E>::template load_spectrum<FULL>(spec, ps.qa, g.nrows, g.mzh, ps.j, tab.tw[2], v);
    else RegRow<T, J, E>::template load_spectrum<FULL>(spec, ps.qa, ps.qb, g.nrows, g.mzh, ps.j, v);
}
template <typename T, int J, int E, bool HALF, bool FULL>
__device__ __forceinline__ void zrows_store(cx<T>* spec, const ZRows<J, E, HALF>& ps, const Geo& g, cx<T>* lds,
                                            const LaneTw<T, J, E>& lt, const AxisTabs<T>& tab, cx<T>* v) {
    if constexpr (HALF)
        RegHalf<T, J, E>::template forward_store<FULL>(spec, ps.qa, g.nrows, g.mzh, lds, ps.j, lt, tab.tw[2], v);
    else RegRow<T, J, E>::template forward_store<FULL>(spec, ps.qa, ps.qb, g.nrows, g.mzh, lds, ps.j, lt, v);
}

// Row pair handled by a lane group: P transforms per workgroup, J lanes each.
template <int J>
struct RegZPos {
    int w, j, q, ra, rb;
    __device__ RegZPos(int P, const BIdx& bi) {
        const int tid = threadIdx.x;
        w = tid / J;
        j = tid - w * J;
        const int tile = bi.tile();
        q = tile * P + w;
        ra = 2 * q;
        rb = ra + 1;
    }
};

// Every row of the workgroup's tile (P row pairs) is inside the grid: the row kernels are then
// instantiated with FULL = true, so their per-element row predicates fold away and no store carries
// an exec-mask branch. Uniform over the workgroup.
__device__ __forceinline__ bool reg_tile_full(const BIdx& bi, int P, int nrows) {
    return ((int64_t)bi.tile() + 1) * P * 2 <= nrows;
}
#ifndef KWH_ROW_FULL_TILES
#define KWH_ROW_FULL_TILES 1
#endif

// Sparse sources on register-resident rows: go through LDS (natural order) when the tile has any.
// Half-form rows (HALF): row lr of the tile is lane group lr, sample n in slot n / 2, component n % 2.
template <typename T, int J, int E, bool HALF = false>
__device__ void reg_apply_sparse(const SparseSrc<T>& s, const Geo& g, int row0, int nrow_tile, cx<T>* lds_all,
                                 int w, int j, int t, cx<T>* v) {
    using S = RegRow<T, J, E>;
    cx<T>* lds = lds_all + w * S::LP;
#pragma unroll
    for (int m = 0; m < E; ++m) lds[S::pidx(j + J * m)] = v[m];
    __syncthreads();
    if (t < s.len) {
        const int r1 = min(g.nrows, row0 + nrow_tile);
        const int e0 = s.rowptr[row0], e1 = s.rowptr[r1];
        for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
            const int64_t cell = s.index[e];
            const int row = (int)(cell / g.m2);
            const int n = (int)(cell - (int64_t)row * g.m2);
            const int lr = row - row0;
            const T val = s.signal[(int64_t)(s.rows == 1 ? 0 : e) * s.len + t] * s.base[e] * s.factor;
            T& slot = HALF ? comp(lds_all[lr * S::LP + S::pidx(n >> 1)], n & 1)
                           : comp(lds_all[(lr >> 1) * S::LP + S::pidx(n)], lr & 1);
            slot = s.mode == 0 ? val : slot + val;
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = lds[S::pidx(j + J * m)];
    __syncthreads();
}

template <typename T, int J, int E, bool FULL>
__device__ __forceinline__ void reg_z_forward_real_impl(const BIdx& bi, const T* __restrict__ src, cx<T>* spec,
                                                        T scale, const Geo& g, int P, const AxisTabs<T>& tab) {
    using S = RegRow<T, J, E>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds = ldsb + S::L;
    RegZPos<J> ps(P, bi);
    LaneTw<T, J, E> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    const bool va = FULL || ps.ra < g.nrows, vb = FULL || ps.rb < g.nrows;
    cx<T> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = ps.j + J * m;
        v[m].x = va ? src[(int64_t)ps.ra * S::L + n] * scale : T(0);
        v[m].y = vb ? src[(int64_t)ps.rb * S::L + n] * scale : T(0);
    }
    S::template forward_store<FULL>(spec, ps.ra, ps.rb, g.nrows, g.mzh, lds + ps.w * S::LP, ps.j, lt, v);
}
template <typename T, int J, int E>
__device__ __forceinline__ void reg_z_forward_real_body(const BIdx& bi, const T* __restrict__ src, cx<T>* spec,
                                                        T scale, const Geo& g, int P, const AxisTabs<T>& tab) {
    if (KWH_ROW_FULL_TILES && reg_tile_full(bi, P, g.nrows))
        reg_z_forward_real_impl<T, J, E, true>(bi, src, spec, scale, g, P, tab);
    else
        reg_z_forward_real_impl<T, J, E, false>(bi, src, spec, scale, g, P, tab);
}
template <typename T, int J, int E>
__global__ void __launch_bounds__(256) reg_z_forward_real(const T* __restrict__ src, cx<T>* spec, T scale, Geo g,
                                                          int P, AxisTabs<T> tab) {
    reg_z_forward_real_body<T, J, E>(bidx_here(), src, spec, scale, g, P, tab);
}

template <typename T, int J, int E, bool FULL, bool ACC>
__device__ __forceinline__ void reg_z_inverse_real_impl(const BIdx& bi, const cx<T>* spec, T* __restrict__ dst,
                                                        T scale, const Geo& g, int P, const AxisTabs<T>& tab) {
    using S = RegRow<T, J, E>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds = ldsb + S::L;
    RegZPos<J> ps(P, bi);
    LaneTw<T, J, E> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    cx<T> v[E];
    S::template load_spectrum<FULL>(spec, ps.ra, ps.rb, g.nrows, g.mzh, ps.j, v);
    S::template transform<true>(v, lds + ps.w * S::LP, ps.j, lt);
    const bool va = FULL || ps.ra < g.nrows, vb = FULL || ps.rb < g.nrows;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = ps.j + J * m;
        if (va) {
            const int64_t c = (int64_t)ps.ra * S::L + n;
            dst[c] = ACC ? dst[c] + v[m].x * scale : v[m].x * scale;
        }
        if (vb) {
            const int64_t c = (int64_t)ps.rb * S::L + n;
            dst[c] = ACC ? dst[c] + v[m].y * scale : v[m].y * scale;
        }
    }
}
template <typename T, int J, int E>
__device__ __forceinline__ void reg_z_inverse_real_body(const BIdx& bi, const cx<T>* spec, T* __restrict__ dst,
                                                        T scale, int accumulate, const Geo& g, int P,
                                                        const AxisTabs<T>& tab) {
    const bool full = KWH_ROW_FULL_TILES && reg_tile_full(bi, P, g.nrows);
    if (accumulate) {
        if (full) reg_z_inverse_real_impl<T, J, E, true, true>(bi, spec, dst, scale, g, P, tab);
        else reg_z_inverse_real_impl<T, J, E, false, true>(bi, spec, dst, scale, g, P, tab);
    } else {
        if (full) reg_z_inverse_real_impl<T, J, E, true, false>(bi, spec, dst, scale, g, P, tab);
        else reg_z_inverse_real_impl<T, J, E, false, false>(bi, spec, dst, scale, g, P, tab);
    }
}
template <typename T, int J, int E>
__global__ void __launch_bounds__(256) reg_z_inverse_real(const cx<T>* spec, T* __restrict__ dst, T scale,
                                                          int accumulate, Geo g, int P, AxisTabs<T> tab) {
    reg_z_inverse_real_body<T, J, E>(bidx_here(), spec, dst, scale, accumulate, g, P, tab);
}

// Collocated velocity along the contiguous axis for a list of rows: dst[row] =
// C2R_z(coll_z * Z[row]) / L, Z = R2C_z(u) of the updated component (as the velocity kernel
// leaves it). The shift e^{-i k dz/2} only involves the contiguous axis, so the full 3-D
// transform of the reference (KSpaceFirstOrder.swift:735-745) reduces to this 1-D one on the
// rows the sensor reads. Rows are paired in list order; an odd tail pairs with nothing.
template <typename T, int J, int E>
__device__ __forceinline__ void reg_z_colloc_rows_body(const BIdx& bi, const cx<T>* spec, T* __restrict__ dst,
                                                       const int* __restrict__ rows, int nlist, const Geo& g, int P,
                                                       const AxisTabs<T>& tab) {
    using S = RegRow<T, J, E>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    RegZPos<J> ps(P, bi);
    LaneTw<T, J, E> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    const bool va = ps.ra < nlist, vb = ps.rb < nlist;
    const int ra = va ? rows[ps.ra] : g.nrows;   // g.nrows: masked by load_spectrum
    const int rb = vb ? rows[ps.rb] : g.nrows;
    cx<T> v[E];
    S::load_spectrum(spec, ra, rb, g.nrows, g.mzh, ps.j, v);
    const cx<T>* co = tab.coll[2] + ps.j;
    const T scale = T(1) / T(S::L);
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const cx<T> c = co[J * m];
        v[m] = cmul(v[m], cx<T>{c.x * scale, c.y * scale});
    }
    S::template transform<true>(v, ldsb + S::L + ps.w * S::LP, ps.j, lt);
#pragma unroll
    for (int m = 0; m < E; ++m) {
        const int n = ps.j + J * m;
        if (va) dst[(int64_t)ra * S::L + n] = v[m].x;
        if (vb) dst[(int64_t)rb * S::L + n] = v[m].y;
    }
}
template <typename T, int J, int E>
__global__ void __launch_bounds__(256) reg_z_colloc_rows(const cx<T>* spec, T* __restrict__ dst,
                                                         const int* __restrict__ rows, int nlist, Geo g, int P,
                                                         AxisTabs<T> tab) {
    reg_z_colloc_rows_body<T, J, E>(bidx_here(), spec, dst, rows, nlist, g, P, tab);
}

// Scheduling fence: keeps the compiler from hoisting the loads of later element groups above
// earlier ones (bounds live registers; other waves hide the latency).
#define KWH_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
constexpr int kGroup = 8;      // density update
#ifndef KWH_VEL_GROUP
#define KWH_VEL_GROUP 4
#endif
constexpr int kVelGroup = KWH_VEL_GROUP;   // velocity update (keeps it within the four-wave register budget)


// Velocity update, per axis: C2R_z(G_a) -> u_a = pml_sg (pml_sg u_a - dt/rho_a g) [+ sources] -> R2C_z.
// HALF: one row of length 2L per lane group (ZRows).
template <typename T, int J, int E, bool CORR, bool FULL, bool HALF = false>
__device__ __forceinline__ void reg_z_velocity_impl(const BIdx& bi, const VelArgs<T>& a, const Geo& g, int P,
                                                    const AxisTabs<T>& tab) {
    using S = RegRow<T, J, E>;
    using Z = ZRows<J, E, HALF>;
    constexpr int SJ = Z::SJ;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds_all = ldsb + S::L;
    Z ps;
    ps.template init<FULL>(P, bi, g.nrows);
    LaneTw<T, J, E> lt;
    lt.template load<HALF ? 2 : 1>(ldsb, tab.tw[2], ps.j);
    cx<T>* lds = lds_all + ps.w * S::LP;
    const bool va = ps.va, vb = ps.vb;
    const int ra = va ? ps.qa : 0, rb = vb ? ps.qb : 0;
    const int64_t oa = ps.oa, ob = ps.ob;
    for (int ax = 0; ax < g.dim; ++ax) {
        const int ia = ax + g.off;
        cx<T> v[E];
        T* u = a.u[ax];
        const SMap<T> dtr = a.dtr[ax];
        const SMap<T> corr = a.corr[ax];
        const T* pm = tab.pmlsg[ia];
        const T* ua = u + oa;
        const T* ub = u + ob;
        const T* dta = dtr.lane(oa, ps.na);
        const T* dtb = dtr.lane(ob, ps.nb);
        const T* cra = corr.lane(oa, ps.na);
        const T* crb = corr.lane(ob, ps.nb);
        zrows_load<T, J, E, HALF, FULL>(a.G[ax], ps, g, tab, v);
        S::template transform<true>(v, lds, ps.j, lt);
        KWH_SCHED_FENCE();
        auto upd = [&](int m, T qa, T qb) {
            T xa = qa * (qa * ua[SJ * m] - dta[SJ * m] * v[m].x);
            T xb = qb * (qb * ub[SJ * m] - dtb[SJ * m] * v[m].y);
            if constexpr (CORR) {
                xa += cra[SJ * m];
                xb += crb[SJ * m];
            }
            v[m].x = va ? xa : T(0);
            v[m].y = vb ? xb : T(0);
        };
        if (ia == 2) {   // staggered PML along z varies along the row (shared by a row pair)
            const T* pza = pm + ps.na;
            const T* pzb = pm + ps.nb;
#pragma unroll
            for (int g0 = 0; g0 < E; g0 += kVelGroup) {
#pragma unroll
                for (int m = g0; m < g0 + kVelGroup && m < E; ++m) {
                    if constexpr (HALF) {
                        upd(m, pza[SJ * m], pzb[SJ * m]);
                    } else {
                        const T q = pza[SJ * m];
                        upd(m, q, q);
                    }
                }
                KWH_SCHED_FENCE();
            }
        } else {
            const T qa = pm[axis_pos(g, ra, 0, ia)], qb = pm[axis_pos(g, rb, 0, ia)];
#pragma unroll
            for (int g0 = 0; g0 < E; g0 += kVelGroup) {
#pragma unroll
                for (int m = g0; m < g0 + kVelGroup && m < E; ++m) upd(m, qa, qb);
                KWH_SCHED_FENCE();
            }
        }

        if (rows_have_sparse(a.sp[ax].rowptr, g, ps.row0, ps.nrow))
            reg_apply_sparse<T, J, E, HALF>(a.sp[ax], g, ps.row0, ps.nrow, lds_all, ps.w, ps.j, a.ctr[0], v);
        {
            T* wa = u + oa;
            T* wb = u + ob;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                if (va) wa[SJ * m] = v[m].x;
                if (vb) wb[SJ * m] = v[m].y;
            }
        }
        zrows_store<T, J, E, HALF, FULL>(a.G[ax], ps, g, lds, lt, tab, v);
    }
}
template <typename T, int J, int E, bool CORR>
__device__ __forceinline__ void reg_z_velocity_body(const BIdx& bi, const VelArgs<T>& a, const Geo& g, int P, const AxisTabs<T>& tab) {
    if (KWH_ROW_FULL_TILES && reg_tile_full(bi, P, g.nrows))
        reg_z_velocity_impl<T, J, E, CORR, true>(bi, a, g, P, tab);
    else
        reg_z_velocity_impl<T, J, E, CORR, false>(bi, a, g, P, tab);
}
template <typename T, int J, int E, bool CORR>
__device__ __forceinline__ void reg_z_velocity_half_body(const BIdx& bi, const VelArgs<T>& a, const Geo& g, int P,
                                                         const AxisTabs<T>& tab) {
    if (KWH_ROW_FULL_TILES && zrows_tile_full<true>(bi, P, g.nrows))
        reg_z_velocity_impl<T, J, E, CORR, true, true>(bi, a, g, P, tab);
    else
        reg_z_velocity_impl<T, J, E, CORR, false, true>(bi, a, g, P, tab);
}
template <typename T, int J, int E, bool CORR>
__global__ void __launch_bounds__(256, (reg_waves<T, E>(KWH_VELOCITY_WAVES)))
    reg_z_velocity_half(VelArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    reg_z_velocity_half_body<T, J, E, CORR>(bidx_here(), a, g, P, tab);
}

template <typename T, int J, int E, bool CORR>
__global__ void __launch_bounds__(256, (reg_waves<T, E>(KWH_VELOCITY_WAVES)))
    reg_z_velocity(VelArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    reg_z_velocity_body<T, J, E, CORR>(bidx_here(), a, g, P, tab);
}

// Density update, per axis: C2R_z(D_a) -> rho_a = pml (pml rho_a - dt rho0 [nl] du_a) [+ sources],
// followed by the EOS (MODE 0), the EOS with Stokes absorption (MODE 2) or the inputs of the
// FFT absorption term (MODE 1: rhoT, F1 = R2C_z(rho0 sum_a du_a), F2 = R2C_z(rhoT)). NL: B/A.
// The total density and, for MODE 1/2, rho0 * sum_a du_a are accumulated in registers while the
// axes are processed, so each divergence spectrum is read once. Length-512 rows run in the half
// form (RegHalf, E = 16 per lane) and keep everything in the kernel. As row pairs (E = 32) in MODE 1
// they keep only rho0 * sum_a du_a: rhoT and F2 then come from reg_z_rhot, which re-reads the
// stored rho_a (d reads of a real field) and frees E accumulator registers per lane. Measured on
// 3-D power-law grids: E = 32 density 571 -> 300 + 144 us (64 x 512 x 512), half form 190 us; at
// E = 24 and 40 the split is no faster (115 -> 79 + 38, 245 -> 189 + 83 us), so those keep the
// fused kernel.
template <typename T, int E, int MODE>
constexpr bool density_rhot_split() { return MODE == 1 && E == 32 && sizeof(T) == 4; }
template <typename T, int J, int E, bool NL, int MODE, bool CORR>
constexpr int density_waves() {
    return (NL || CORR) ? 1 : reg_waves<T, E>(MODE == 0 ? KWH_DENSITY_WAVES : 2);
}
template <typename T, int J, int E, bool NL, int MODE, bool CORR, bool FULL, bool HALF = false>
__device__ __forceinline__ void reg_z_density_impl(const BIdx& bi, const RhoArgs<T>& a, const Geo& g, int P,
                                                   const AxisTabs<T>& tab) {
    using S = RegRow<T, J, E>;
    using Z = ZRows<J, E, HALF>;
    constexpr int SJ = Z::SJ;
    constexpr bool ACC = MODE != 0;   // accumulate rho0 * sum_a du_a
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds_all = ldsb + S::L;
    Z ps;
    ps.template init<FULL>(P, bi, g.nrows);
    LaneTw<T, J, E> lt;
    lt.template load<HALF ? 2 : 1>(ldsb, tab.tw[2], ps.j);
    cx<T>* lds = lds_all + ps.w * S::LP;
    const bool va = ps.va, vb = ps.vb;
    const int ra = va ? ps.qa : 0, rb = vb ? ps.qb : 0;
    const int64_t oa = ps.oa, ob = ps.ob;
    const T* r0pa = a.rho0.lane(oa, ps.na);
    const T* r0pb = a.rho0.lane(ob, ps.nb);
    // rho0 of the row pair held in registers across the axes (one HBM read instead of one per axis)
    constexpr bool HOLD = KWH_DENSITY_HOLD_RHO0 && E <= 16 && sizeof(T) == 4;
    cx<T> r0h[HOLD ? E : 1];
    if constexpr (HOLD) {
#pragma unroll
        for (int m = 0; m < E; ++m) r0h[m] = cx<T>{r0pa[SJ * m], r0pb[SJ * m]};
    }
    constexpr bool RT = !density_rhot_split<T, E, MODE>();
    cx<T> rt[RT ? E : 1];
    cx<T> sd[ACC ? E : 1];
    cx<T> rp[NL ? E : 1];
    if constexpr (NL) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            T sa = 0, sb = 0;
            for (int ax = 0; ax < g.dim; ++ax) {
                sa += a.rho[ax][oa + SJ * m];
                sb += a.rho[ax][ob + SJ * m];
            }
            rp[m] = cx<T>{sa, sb};
        }
    }
    // rho_a of the row pair loaded together with the divergence spectrum (KWH_DENSITY_PRELOAD)
    constexpr bool PRE = KWH_DENSITY_PRELOAD && !NL && !CORR && MODE == 1 && E <= 16 && sizeof(T) == 4;
    for (int ax = 0; ax < g.dim; ++ax) {
        const int ia = ax + g.off;
        cx<T> v[E];
        T* rho = a.rho[ax];
        const T* rha = rho + oa;
        const T* rhb = rho + ob;
        zrows_load<T, J, E, HALF, FULL>(a.D[ax], ps, g, tab, v);
        cx<T> rpre[PRE ? E : 1];
        if constexpr (PRE) {
#pragma unroll
            for (int m = 0; m < E; ++m) rpre[m] = cx<T>{rha[SJ * m], rhb[SJ * m]};
        }
        S::template transform<true>(v, lds, ps.j, lt);
        KWH_SCHED_FENCE();
        const T* pm = tab.pml[ia];
        const T cf = a.corr_factor[ax];
        const T* cra = a.corr.lane(oa, ps.na);
        const T* crb = a.corr.lane(ob, ps.nb);
        auto upd = [&](int m, T qa, T qb) {
            const T r0a = HOLD ? r0h[m].x : r0pa[SJ * m];
            const T r0b = HOLD ? r0h[m].y : r0pb[SJ * m];
            const T da = r0a * v[m].x, db = r0b * v[m].y;
            if constexpr (ACC) {
                if (ax == 0) sd[m] = cx<T>{da, db};
                else { sd[m].x += da; sd[m].y += db; }
            }
            T ma = a.dt * da, mb = a.dt * db;
            if constexpr (NL) {
                ma *= (T(2) * rp[m].x + r0a) / r0a;
                mb *= (T(2) * rp[m].y + r0b) / r0b;
            }
            T xa = qa * (qa * (PRE ? rpre[m].x : rha[SJ * m]) - ma);
            T xb = qb * (qb * (PRE ? rpre[m].y : rhb[SJ * m]) - mb);
            if constexpr (CORR) {
                xa += cra[SJ * m] * cf;
                xb += crb[SJ * m] * cf;
            }
            v[m] = cx<T>{xa, xb};
        };
        if (ia == 2) {
            const T* pza = pm + ps.na;
            const T* pzb = pm + ps.nb;
#pragma unroll
            for (int g0 = 0; g0 < E; g0 += kGroup) {
#pragma unroll
                for (int m = g0; m < g0 + kGroup && m < E; ++m) {
                    if constexpr (HALF) {
                        upd(m, pza[SJ * m], pzb[SJ * m]);
                    } else {
                        const T q = pza[SJ * m];
                        upd(m, q, q);
                    }
                }
                KWH_SCHED_FENCE();
            }
        } else {
            const T qa = pm[axis_pos(g, ra, 0, ia)], qb = pm[axis_pos(g, rb, 0, ia)];
#pragma unroll
            for (int g0 = 0; g0 < E; g0 += kGroup) {
#pragma unroll
                for (int m = g0; m < g0 + kGroup && m < E; ++m) upd(m, qa, qb);
                KWH_SCHED_FENCE();
            }
        }
        if (rows_have_sparse(a.sp[ax].rowptr, g, ps.row0, ps.nrow))
            reg_apply_sparse<T, J, E, HALF>(a.sp[ax], g, ps.row0, ps.nrow, lds_all, ps.w, ps.j, a.ctr[0], v);
        T* wa = rho + oa;
        T* wb = rho + ob;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            if (va) wa[SJ * m] = v[m].x;
            if (vb) wb[SJ * m] = v[m].y;
            if constexpr (RT) {
                if (ax == 0) rt[m] = v[m];
                else { rt[m].x += v[m].x; rt[m].y += v[m].y; }
            }
        }
    }
    cx<T> v[E];
    if constexpr (MODE != 1) {
        const T* tpa = a.tau.lane(oa, ps.na);
        const T* tpb = a.tau.lane(ob, ps.nb);
        const T* bpa = a.bona.lane(oa, ps.na);
        const T* bpb = a.bona.lane(ob, ps.nb);
        const T* cpa = a.c0.lane(oa, ps.na);
        const T* cpb = a.c0.lane(ob, ps.nb);
#pragma unroll
        for (int m = 0; m < E; ++m) {
            T ea = rt[m].x, eb = rt[m].y;
            const T r0a = HOLD ? r0h[m].x : r0pa[SJ * m];
            const T r0b = HOLD ? r0h[m].y : r0pb[SJ * m];
            if constexpr (MODE == 2) {
                ea += tpa[SJ * m] * sd[m].x;
                eb += tpb[SJ * m] * sd[m].y;
            }
            if constexpr (NL) {
                ea += bpa[SJ * m] * (rt[m].x * rt[m].x) / (T(2) * r0a);
                eb += bpb[SJ * m] * (rt[m].y * rt[m].y) / (T(2) * r0b);
            }
            const T cA = cpa[SJ * m];
            const T cB = cpb[SJ * m];
            const T pA = cA * cA * ea, pB = cB * cB * eb;
            if (va) a.p[oa + SJ * m] = pA;
            if (vb) a.p[ob + SJ * m] = pB;
            v[m] = cx<T>{va ? pA : T(0), vb ? pB : T(0)};
        }
        zrows_store<T, J, E, HALF, FULL>(a.outP, ps, g, lds, lt, tab, v);
    } else {
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cx<T>{va ? sd[m].x : T(0), vb ? sd[m].y : T(0)};
        zrows_store<T, J, E, HALF, FULL>(a.F1, ps, g, lds, lt, tab, v);
        if constexpr (RT) {
#pragma unroll
            for (int m = 0; m < E; ++m) {
                if (va) a.rhoT[oa + SJ * m] = rt[m].x;
                if (vb) a.rhoT[ob + SJ * m] = rt[m].y;
                rt[m] = cx<T>{va ? rt[m].x : T(0), vb ? rt[m].y : T(0)};
            }
            if (a.dispersion) zrows_store<T, J, E, HALF, FULL>(a.F2, ps, g, lds, lt, tab, rt);
        }
    }
}

// rhoT = sum_a rho_a of a row pair (summed in axis order, as the density kernel does) and, with
// dispersion, F2 = R2C_z(rhoT): the part of the FFT-absorption density kernel that long rows leave
// to a pass of their own (density_rhot_split).
template <typename T, int J, int E, bool FULL>
__device__ __forceinline__ void reg_z_rhot_impl(const BIdx& bi, const RhoArgs<T>& a, const Geo& g, int P,
                                                const AxisTabs<T>& tab) {
    using S = RegRow<T, J, E>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    RegZPos<J> ps(P, bi);
    LaneTw<T, J, E> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    cx<T>* lds = ldsb + S::L + ps.w * S::LP;
    const bool va = FULL || ps.ra < g.nrows, vb = FULL || ps.rb < g.nrows;
    const int64_t oa = (int64_t)(va ? ps.ra : 0) * S::L + ps.j, ob = (int64_t)(vb ? ps.rb : 0) * S::L + ps.j;
    cx<T> rt[E];
#pragma unroll
    for (int m = 0; m < E; ++m) rt[m] = cx<T>{a.rho[0][oa + J * m], a.rho[0][ob + J * m]};
    for (int ax = 1; ax < g.dim; ++ax) {
        const T* rho = a.rho[ax];
#pragma unroll
        for (int m = 0; m < E; ++m) {
            rt[m].x += rho[oa + J * m];
            rt[m].y += rho[ob + J * m];
        }
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
        if (va) a.rhoT[oa + J * m] = rt[m].x;
        if (vb) a.rhoT[ob + J * m] = rt[m].y;
        rt[m] = cx<T>{va ? rt[m].x : T(0), vb ? rt[m].y : T(0)};
    }
    if (a.dispersion) S::template forward_store<FULL>(a.F2, ps.ra, ps.rb, g.nrows, g.mzh, lds, ps.j, lt, rt);
}
template <typename T, int J, int E>
__device__ __forceinline__ void reg_z_rhot_body(const BIdx& bi, const RhoArgs<T>& a, const Geo& g, int P,
                                                const AxisTabs<T>& tab) {
    if (KWH_ROW_FULL_TILES && reg_tile_full(bi, P, g.nrows)) reg_z_rhot_impl<T, J, E, true>(bi, a, g, P, tab);
    else reg_z_rhot_impl<T, J, E, false>(bi, a, g, P, tab);
}
template <typename T, int J, int E>
__global__ void __launch_bounds__(256) reg_z_rhot(RhoArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    reg_z_rhot_body<T, J, E>(bidx_here(), a, g, P, tab);
}
template <typename T, int J, int E, bool NL, int MODE, bool CORR>
__device__ __forceinline__ void reg_z_density_body(const BIdx& bi, const RhoArgs<T>& a, const Geo& g, int P, const AxisTabs<T>& tab) {
    if (KWH_ROW_FULL_TILES && reg_tile_full(bi, P, g.nrows))
        reg_z_density_impl<T, J, E, NL, MODE, CORR, true>(bi, a, g, P, tab);
    else
        reg_z_density_impl<T, J, E, NL, MODE, CORR, false>(bi, a, g, P, tab);
}

template <typename T, int J, int E, bool NL, int MODE, bool CORR>
__global__ void __launch_bounds__(256, (density_waves<T, J, E, NL, MODE, CORR>()))
    reg_z_density(RhoArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    reg_z_density_body<T, J, E, NL, MODE, CORR>(bidx_here(), a, g, P, tab);
}
template <typename T, int J, int E, bool NL, int MODE, bool CORR>
__device__ __forceinline__ void reg_z_density_half_body(const BIdx& bi, const RhoArgs<T>& a, const Geo& g, int P,
                                                        const AxisTabs<T>& tab) {
    if (KWH_ROW_FULL_TILES && zrows_tile_full<true>(bi, P, g.nrows))
        reg_z_density_impl<T, J, E, NL, MODE, CORR, true, true>(bi, a, g, P, tab);
    else
        reg_z_density_impl<T, J, E, NL, MODE, CORR, false, true>(bi, a, g, P, tab);
}
template <typename T, int J, int E, bool NL, int MODE, bool CORR>
__global__ void __launch_bounds__(256, (density_waves<T, J, E, NL, MODE, CORR>()))
    reg_z_density_half(RhoArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    reg_z_density_half_body<T, J, E, NL, MODE, CORR>(bidx_here(), a, g, P, tab);
}

// EOS with FFT absorption: p = c^2 (rhoT + tau C2R_z(A1) - eta C2R_z(A2) [+ B/A rhoT^2 / (2 rho0)]),
// P = R2C_z(p). Both spectra are loaded before the first transform (one memory phase).
template <typename T, int E>
constexpr int eos_waves() {
    return reg_waves<T, E>(KWH_EOS_PRELOAD && E <= 16 && sizeof(T) == 4 ? KWH_EOS_PRELOAD_WAVES : KWH_EOS_WAVES);
}
template <typename T, int J, int E, bool FULL, bool HALF = false>
__device__ __forceinline__ void reg_z_eos_impl(const BIdx& bi, const EosArgs<T>& a, const Geo& g, int P,
                                               const AxisTabs<T>& tab) {
    using S = RegRow<T, J, E>;
    using Z = ZRows<J, E, HALF>;
    constexpr int SJ = Z::SJ;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    Z ps;
    ps.template init<FULL>(P, bi, g.nrows);
    LaneTw<T, J, E> lt;
    lt.template load<HALF ? 2 : 1>(ldsb, tab.tw[2], ps.j);
    cx<T>* lds = ldsb + S::L + ps.w * S::LP;
    const bool va = ps.va, vb = ps.vb;
    cx<T> e[E], v[E];
    const int64_t oa = ps.oa, ob = ps.ob;
    const T* cpa = a.c0.lane(oa, ps.na);
    const T* cpb = a.c0.lane(ob, ps.nb);
    zrows_load<T, J, E, HALF, FULL>(a.A1, ps, g, tab, e);
    if (a.A2) zrows_load<T, J, E, HALF, FULL>(a.A2, ps, g, tab, v);
    constexpr bool PRE = KWH_EOS_PRELOAD && E <= 16 && sizeof(T) == 4;
    cx<T> rtp[PRE ? E : 1], cp[PRE ? E : 1];   // sum rho and c0 of the row pair, loaded with the spectra
    if constexpr (PRE) {
#pragma unroll
        for (int m = 0; m < E; ++m) {
            rtp[m] = cx<T>{a.rhoT[oa + SJ * m], a.rhoT[ob + SJ * m]};
            cp[m] = cx<T>{cpa[SJ * m], cpb[SJ * m]};
        }
    }
    S::template transform<true>(e, lds, ps.j, lt);
    if (a.A2) S::template transform<true>(v, lds, ps.j, lt);
    KWH_SCHED_FENCE();
    const T* tpa = a.tau.lane(oa, ps.na);
    const T* tpb = a.tau.lane(ob, ps.nb);
    const T* epa = a.eta.lane(oa, ps.na);
    const T* epb = a.eta.lane(ob, ps.nb);
    const T* bpa = a.bona.lane(oa, ps.na);
    const T* bpb = a.bona.lane(ob, ps.nb);
    const T* rpa = a.rho0.lane(oa, ps.na);
    const T* rpb = a.rho0.lane(ob, ps.nb);
    // The element loop, instantiated per combination of the three run-time switches (uniform over
    // the launch) so that its body carries no per-element branch.
    auto elements = [&](auto from_c, auto has_a2, auto nl) {
        constexpr bool FROMC = decltype(from_c)::value, A2 = decltype(has_a2)::value, NLIN = decltype(nl)::value;
#pragma unroll
        for (int g0 = 0; g0 < E; g0 += kGroup) {
#pragma unroll
            for (int m = g0; m < g0 + kGroup && m < E; ++m) {
                const int64_t ca = oa + SJ * m, cb = ob + SJ * m;
                const T rta = PRE ? rtp[m].x : a.rhoT[ca], rtb = PRE ? rtp[m].y : a.rhoT[cb];
                const T cA = PRE ? cp[m].x : cpa[SJ * m], cB = PRE ? cp[m].y : cpb[SJ * m];
                T tauA, etaA = T(0), tauB, etaB = T(0);
                if constexpr (FROMC) {
                    const T la = log2(cA), lb = log2(cB);
                    tauA = a.tk * exp2(a.te * la);
                    tauB = a.tk * exp2(a.te * lb);
                    if constexpr (A2) {
                        etaA = a.ek * exp2(a.ee * la);
                        etaB = a.ek * exp2(a.ee * lb);
                    }
                } else {
                    tauA = tpa[SJ * m];
                    tauB = tpb[SJ * m];
                    if constexpr (A2) {
                        etaA = epa[SJ * m];
                        etaB = epb[SJ * m];
                    }
                }
                T ea = rta + tauA * e[m].x;
                T eb = rtb + tauB * e[m].y;
                if constexpr (A2) {
                    ea -= etaA * v[m].x;
                    eb -= etaB * v[m].y;
                }
                if constexpr (NLIN) {
                    ea += bpa[SJ * m] * (rta * rta) / (T(2) * rpa[SJ * m]);
                    eb += bpb[SJ * m] * (rtb * rtb) / (T(2) * rpb[SJ * m]);
                }
                const T pA = cA * cA * ea, pB = cB * cB * eb;
                if (va) a.p[ca] = pA;
                if (vb) a.p[cb] = pB;
                e[m] = cx<T>{va ? pA : T(0), vb ? pB : T(0)};
            }
            KWH_SCHED_FENCE();
        }
    };
    using Y = std::true_type;
    using N = std::false_type;
#if KWH_EOS_VERSIONED
    const bool has_a2 = a.A2 != nullptr;
    if (a.tau_eta_from_c) {
        if (has_a2) { if (a.nonlinear) elements(Y{}, Y{}, Y{}); else elements(Y{}, Y{}, N{}); }
        else { if (a.nonlinear) elements(Y{}, N{}, Y{}); else elements(Y{}, N{}, N{}); }
    } else {
        if (has_a2) { if (a.nonlinear) elements(N{}, Y{}, Y{}); else elements(N{}, Y{}, N{}); }
        else { if (a.nonlinear) elements(N{}, N{}, Y{}); else elements(N{}, N{}, N{}); }
    }
#else
    (void)elements;
    {
#pragma unroll
        for (int g0 = 0; g0 < E; g0 += kGroup) {
#pragma unroll
            for (int m = g0; m < g0 + kGroup && m < E; ++m) {
                const int64_t ca = oa + SJ * m, cb = ob + SJ * m;
                const T rta = PRE ? rtp[m].x : a.rhoT[ca], rtb = PRE ? rtp[m].y : a.rhoT[cb];
                const T cA = PRE ? cp[m].x : cpa[SJ * m], cB = PRE ? cp[m].y : cpb[SJ * m];
                T tauA, etaA, tauB, etaB;
                if (a.tau_eta_from_c) {
                    const T la = log2(cA), lb = log2(cB);
                    tauA = a.tk * exp2(a.te * la);
                    tauB = a.tk * exp2(a.te * lb);
                    etaA = a.ek * exp2(a.ee * la);
                    etaB = a.ek * exp2(a.ee * lb);
                } else {
                    tauA = tpa[SJ * m];
                    tauB = tpb[SJ * m];
                    etaA = epa[SJ * m];
                    etaB = epb[SJ * m];
                }
                T ea = rta + tauA * e[m].x;
                T eb = rtb + tauB * e[m].y;
                if (a.A2) {
                    ea -= etaA * v[m].x;
                    eb -= etaB * v[m].y;
                }
                if (a.nonlinear) {
                    ea += bpa[SJ * m] * (rta * rta) / (T(2) * rpa[SJ * m]);
                    eb += bpb[SJ * m] * (rtb * rtb) / (T(2) * rpb[SJ * m]);
                }
                const T pA = cA * cA * ea, pB = cB * cB * eb;
                if (va) a.p[ca] = pA;
                if (vb) a.p[cb] = pB;
                e[m] = cx<T>{va ? pA : T(0), vb ? pB : T(0)};
            }
            KWH_SCHED_FENCE();
        }
    }
#endif
    zrows_store<T, J, E, HALF, FULL>(a.outP, ps, g, lds, lt, tab, e);
}
template <typename T, int J, int E>
__device__ __forceinline__ void reg_z_eos_body(const BIdx& bi, const EosArgs<T>& a, const Geo& g, int P, const AxisTabs<T>& tab) {
    if (KWH_ROW_FULL_TILES && reg_tile_full(bi, P, g.nrows))
        reg_z_eos_impl<T, J, E, true>(bi, a, g, P, tab);
    else
        reg_z_eos_impl<T, J, E, false>(bi, a, g, P, tab);
}

template <typename T, int J, int E>
__global__ void __launch_bounds__(256, (eos_waves<T, E>()))
    reg_z_eos(EosArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    reg_z_eos_body<T, J, E>(bidx_here(), a, g, P, tab);
}
template <typename T, int J, int E>
__device__ __forceinline__ void reg_z_eos_half_body(const BIdx& bi, const EosArgs<T>& a, const Geo& g, int P,
                                                    const AxisTabs<T>& tab) {
    if (KWH_ROW_FULL_TILES && zrows_tile_full<true>(bi, P, g.nrows))
        reg_z_eos_impl<T, J, E, true, true>(bi, a, g, P, tab);
    else
        reg_z_eos_impl<T, J, E, false, true>(bi, a, g, P, tab);
}
template <typename T, int J, int E>
__global__ void __launch_bounds__(256, (eos_waves<T, E>()))
    reg_z_eos_half(EosArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    reg_z_eos_half_body<T, J, E>(bidx_here(), a, g, P, tab);
}

// Adjoint forward row kernel (AdjRowArgs, kwh_kernels.h): the diagonal update of the reverse sweep
// computed on the row pair in registers, then R2C_z into a.spec. Instantiated per launch-uniform
// case (PML along the row or across it, s or s^2, full tile), so no element carries a branch.
template <typename T, int J, int E, int MODE, bool FULL, bool ZAX, bool SQ>
__device__ __forceinline__ void reg_adj_fwd_impl(const BIdx& bi, const AdjRowArgs<T>& a, const Geo& g, int P,
                                                 const AxisTabs<T>& tab) {
    using S = RegRow<T, J, E>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    RegZPos<J> ps(P, bi);
    LaneTw<T, J, E> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    const bool va = FULL || ps.ra < g.nrows, vb = FULL || ps.rb < g.nrows;
    const int ra = va ? ps.ra : 0, rb = vb ? ps.rb : 0;
    const int64_t oa = (int64_t)ra * S::L + ps.j, ob = (int64_t)rb * S::L + ps.j;
    const T* sz = a.s + ps.j;
    const T sa0 = ZAX ? T(1) : a.s[axis_pos(g, ra, 0, a.ia)];
    const T sb0 = ZAX ? T(1) : a.s[axis_pos(g, rb, 0, a.ia)];
    cx<T> v[E];
    if constexpr (MODE == 0) {
        const T* rpa = a.r.lane(oa, ps.j);
        const T* rpb = a.r.lane(ob, ps.j);
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const T qa = ZAX ? sz[J * m] : sa0, qb = ZAX ? sz[J * m] : sb0;
            const T fa = SQ ? qa * qa : qa, fb = SQ ? qb * qb : qb;
            const T xa = a.scale * rpa[J * m] * fa * a.lu[oa + J * m];
            const T xb = a.scale * rpb[J * m] * fb * a.lu[ob + J * m];
            v[m] = cx<T>{va ? xa : T(0), vb ? xb : T(0)};
        }
    } else {
        const T* cpa = a.c0.lane(oa, ps.j);
        const T* cpb = a.c0.lane(ob, ps.j);
        const T* mpa = a.rho0.lane(oa, ps.j);
        const T* mpb = a.rho0.lane(ob, ps.j);
#pragma unroll
        for (int m = 0; m < E; ++m) {
            const T qa = ZAX ? sz[J * m] : sa0, qb = ZAX ? sz[J * m] : sb0;
            const T ca = cpa[J * m], cb = cpb[J * m];
            const T la = ca * ca * a.lp[oa + J * m] + qa * qa * a.lr[oa + J * m];
            const T lb = cb * cb * a.lp[ob + J * m] + qb * qb * a.lr[ob + J * m];
            if (va) a.lr[oa + J * m] = la;
            if (vb) a.lr[ob + J * m] = lb;
            v[m] = cx<T>{va ? a.dt * mpa[J * m] * qa * la : T(0), vb ? a.dt * mpb[J * m] * qb * lb : T(0)};
        }
    }
    S::template forward_store<FULL>(a.spec, ps.ra, ps.rb, g.nrows, g.mzh, ldsb + S::L + ps.w * S::LP, ps.j, lt, v);
}
template <typename T, int J, int E, int MODE>
__device__ __forceinline__ void reg_adj_fwd_body(const BIdx& bi, const AdjRowArgs<T>& a, const Geo& g, int P,
                                                 const AxisTabs<T>& tab) {
    const bool full = KWH_ROW_FULL_TILES && reg_tile_full(bi, P, g.nrows);
    const bool zax = a.ia == 2, sq = MODE == 0 && a.spow == 2;
    auto go = [&](auto F, auto Z, auto Q) {
        reg_adj_fwd_impl<T, J, E, MODE, decltype(F)::value, decltype(Z)::value, decltype(Q)::value>(bi, a, g, P, tab);
    };
    using Y = std::true_type;
    using N = std::false_type;
    if (full) {
        if (zax) { if (sq) go(Y{}, Y{}, Y{}); else go(Y{}, Y{}, N{}); }
        else { if (sq) go(Y{}, N{}, Y{}); else go(Y{}, N{}, N{}); }
    } else {
        if (zax) { if (sq) go(N{}, Y{}, Y{}); else go(N{}, Y{}, N{}); }
        else { if (sq) go(N{}, N{}, Y{}); else go(N{}, N{}, N{}); }
    }
}
template <typename T, int J, int E, int MODE>
__global__ void __launch_bounds__(256) reg_adj_fwd(AdjRowArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    reg_adj_fwd_body<T, J, E, MODE>(bidx_here(), a, g, P, tab);
}

// Adjoint inverse row kernel: C2R_z(spec) stored (mode 0), accumulated (1) or added to the
// decayed field s^2 dst (2). Instantiated per mode, PML orientation and full tile.
template <typename T, int J, int E, int MODE, bool FULL, bool ZAX>
__device__ __forceinline__ void reg_adj_inv_impl(const BIdx& bi, const cx<T>* spec, T* dst, const T* s, int ia,
                                                 const Geo& g, int P, const AxisTabs<T>& tab) {
    using S = RegRow<T, J, E>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    RegZPos<J> ps(P, bi);
    LaneTw<T, J, E> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    cx<T> v[E];
    S::template load_spectrum<FULL>(spec, ps.ra, ps.rb, g.nrows, g.mzh, ps.j, v);
    S::template transform<true>(v, ldsb + S::L + ps.w * S::LP, ps.j, lt);
    const bool va = FULL || ps.ra < g.nrows, vb = FULL || ps.rb < g.nrows;
    const int ra = va ? ps.ra : 0, rb = vb ? ps.rb : 0;
    const T sa0 = MODE == 2 && !ZAX ? s[axis_pos(g, ra, 0, ia)] : T(1);
    const T sb0 = MODE == 2 && !ZAX ? s[axis_pos(g, rb, 0, ia)] : T(1);
    const T* sz = s + ps.j;
    T* da = dst + (int64_t)ra * S::L + ps.j;
    T* db = dst + (int64_t)rb * S::L + ps.j;
#pragma unroll
    for (int m = 0; m < E; ++m) {
        if constexpr (MODE == 0) {
            if (va) da[J * m] = v[m].x;
            if (vb) db[J * m] = v[m].y;
        } else if constexpr (MODE == 1) {
            if (va) da[J * m] = da[J * m] + v[m].x;
            if (vb) db[J * m] = db[J * m] + v[m].y;
        } else {
            const T qa = ZAX ? sz[J * m] : sa0, qb = ZAX ? sz[J * m] : sb0;
            if (va) da[J * m] = da[J * m] * (qa * qa) + v[m].x;
            if (vb) db[J * m] = db[J * m] * (qb * qb) + v[m].y;
        }
    }
}
template <typename T, int J, int E>
__device__ __forceinline__ void reg_adj_inv_body(const BIdx& bi, const cx<T>* spec, T* dst, const T* s, int ia,
                                                 int mode, const Geo& g, int P, const AxisTabs<T>& tab) {
    const bool full = KWH_ROW_FULL_TILES && reg_tile_full(bi, P, g.nrows);
    const bool zax = ia == 2;
    auto go = [&](auto M, auto F, auto Z) {
        reg_adj_inv_impl<T, J, E, decltype(M)::value, decltype(F)::value, decltype(Z)::value>(bi, spec, dst, s, ia, g,
                                                                                              P, tab);
    };
    using Y = std::true_type;
    using N = std::false_type;
    using M0 = std::integral_constant<int, 0>;
    using M1 = std::integral_constant<int, 1>;
    using M2 = std::integral_constant<int, 2>;
    if (mode == 0) { if (full) go(M0{}, Y{}, N{}); else go(M0{}, N{}, N{}); }
    else if (mode == 1) { if (full) go(M1{}, Y{}, N{}); else go(M1{}, N{}, N{}); }
    else if (full) { if (zax) go(M2{}, Y{}, Y{}); else go(M2{}, Y{}, N{}); }
    else { if (zax) go(M2{}, N{}, Y{}); else go(M2{}, N{}, N{}); }
}
template <typename T, int J, int E>
__global__ void __launch_bounds__(256) reg_adj_inv(const cx<T>* spec, T* dst, const T* s, int ia, int mode, Geo g,
                                                   int P, AxisTabs<T> tab) {
    reg_adj_inv_body<T, J, E>(bidx_here(), spec, dst, s, ia, mode, g, P, tab);
}

// ------------------------------------------------------------------------------------------
// Strided / middle passes over a tile of NTR contiguous columns: lane (w, j), w fastest.
// LDS: element (n1, k2) of step 1 of column w at (n1*E + k2)*NTR + w + n1 (pad one slot per E
// rows), after the L-element twiddle table.
// ------------------------------------------------------------------------------------------
template <typename T, int J, int E, int NTR>
struct RegCol {
    static constexpr int L = J * E;
    static constexpr size_t lds_elems() { return (size_t)L * NTR + J + L; }
    //   step 1 write (n1 = j, k2 = r) -> (j*E*NTR + w + j) + r*NTR
    //   step 2 read  (n1, k2 = j + J*a) -> (j*NTR + w) + n1*(E*NTR + 1) + J*a*NTR
    template <bool INV>
    __device__ static __forceinline__ void transform(cx<T>* v, cx<T>* lds, int j, int w, const LaneTw<T, J, E>& tw) {
        dftN<T, E, INV>(v);
        cx<T>* wr = lds + (j * E * NTR + w + j);
#pragma unroll
        for (int r = 0; r < E; ++r) wr[r * NTR] = v[r];
        __syncthreads();
        reg_step2<T, J, E, E * NTR + 1, NTR, INV>(v, lds + (j * NTR + w), tw);
        __syncthreads();
    }
};

// Largest tile width (columns per workgroup) of the column kernels: 256 lanes in FP32 (128 in
// FP64), halved until the LDS slab and twiddle table fit 64 KiB.
template <typename T, int J, int E>
constexpr int reg_ntr_max() {
    int n = (sizeof(T) == 8 ? 128 : 256) / J;
    while (n > 1 && ((size_t)J * E * n + J + J * E) * sizeof(cx<T>) > 65536) n /= 2;
    return n;
}

template <typename T, int J, int E, int NTR, bool INV, bool DUAL = false>
__device__ __forceinline__ void reg_strided_body(const BIdx& bi, const StridedArgs& a, const cx<T>* __restrict__ tw) {
    using S = RegCol<T, J, E, NTR>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds = ldsb + S::L;
    const int w = threadIdx.x % NTR, j = threadIdx.x / NTR;
    LaneTw<T, J, E> lt;
    lt.load(ldsb, tw, j);
    const int tile = a.walk_rev ? bi.tile_rev() : bi.tile();
    const int64_t c = (int64_t)tile * NTR + w;
    const bool valid = c < a.ncols;
    const int64_t cc = valid ? c : 0;
    const int64_t base = (cc / a.inner) * (int64_t)S::L * a.inner + cc % a.inner;
    cx<T>* data = reinterpret_cast<cx<T>*>(a.fields[bi.y]) + base + (int64_t)j * a.stride;
    const int64_t step = (int64_t)J * a.stride;
    cx<T> v[E];
    {
        const cx<T>* q = data;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            v[m] = *q;
            q += step;
        }
    }
    if constexpr (DUAL) {
        if (bi.y == a.dual - 1) {   // uniform over the workgroup
            cx<T> v2[E];
            const cx<T> d2 = reinterpret_cast<const cx<T>*>(a.d2)[cc % a.inner];
            const cx<T>* d1 = reinterpret_cast<const cx<T>*>(a.d1) + j;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                v2[m] = cmul(v[m], d2);
                v[m] = cmul(v[m], d1[J * m]);
            }
            S::template transform<INV>(v2, lds, j, w, lt);
            if (valid) {
                cx<T>* q = reinterpret_cast<cx<T>*>(a.dual_out) + base + (int64_t)j * a.stride;
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    *q = v2[m];
                    q += step;
                }
            }
        }
    }
    S::template transform<INV>(v, lds, j, w, lt);
    if (valid) {
        cx<T>* q = data;
#pragma unroll
        for (int m = 0; m < E; ++m) {
            *q = v[m];
            q += step;
        }
    }
}
template <typename T, int J, int E, int NTR, bool INV, bool DUAL = false>
__global__ void __launch_bounds__(256) reg_strided(StridedArgs a, const cx<T>* __restrict__ tw) {
    reg_strided_body<T, J, E, NTR, INV, DUAL>(bidx_here(), a, tw);
}

// Middle pass: forward along the middle axis, operators, inverse, store. Every table-using
// operator of one input reads the same real k-space table (checked on the host), loaded once.
// XL = false: lane (w, j) with w fastest, the inner radix through an LDS transpose (RegCol);
// spectra between the transforms in the lane-natural pattern k = j + J*m.
// XL = true: lane (w, j) with j fastest, the inner radix across the lanes of a DPP row (XLane);
// spectra between the transforms in the block pattern k = m + E*brev(j). LDS: the twiddle table.
template <typename T, int J, int E, int NTR, bool XL = false>
__device__ __forceinline__ void reg_middle_body(const BIdx& bi, const MidArgs& a, const AxisTabs<T>& tab, const Geo& g) {
    using S = RegCol<T, J, E, NTR>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds = ldsb + S::L;
    const int w = XL ? (int)threadIdx.x / J : (int)threadIdx.x % NTR;
    const int j = XL ? (int)threadIdx.x % J : (int)threadIdx.x / NTR;
    LaneTw<T, J, E> lt;
    struct NoXLane {};
    std::conditional_t<XL, XLane<T, J, E>, NoXLane> xl;   // the cross-lane form exists for J <= 16 only
    // spectrum element held in slot m: k = kb + kst * m
    int kb = j;
    if constexpr (XL) {
        xl.init(ldsb, tab.tw[a.axis], j);
        __syncthreads();
        kb = xl.block_start();
    } else {
        lt.load(ldsb, tab.tw[a.axis], j);
    }
    constexpr int kst = XL ? 1 : J;
    const int tile = bi.tile();
    const int64_t c = (int64_t)tile * NTR + w;
    const bool valid = c < a.ncols;
    const int64_t cc = valid ? c : 0;
    const int64_t base = (cc / a.inner) * (int64_t)S::L * a.inner + cc % a.inner + (int64_t)j * a.stride;
    const int64_t step = (int64_t)J * a.stride;
    const T scale = T(a.scale);
    int i0c, i1c, i2c;
    if (a.axis == 0) { i0c = 0; i1c = (int)(cc / g.mzh); i2c = (int)(cc % g.mzh); }
    else { i0c = (int)(cc / g.mzh); i1c = 0; i2c = (int)(cc % g.mzh); }
    const int64_t hstride = a.axis == 0 ? (int64_t)g.m1 * g.mzh : (int64_t)g.mzh;
    const int64_t hbase = (a.axis == 0 ? cc : (int64_t)i0c * g.m1 * g.mzh + i2c) + (int64_t)kb * hstride;
    const int64_t hstep = (int64_t)kst * hstride;
    T kap[E];
    const T* kap_src = nullptr;   // table currently held in kap (inputs sharing one table load it once)
    for (int f = 0; f < a.nin; ++f) {
        cx<T> X[E];
        if (!(f == 0 && a.slab_n > 0)) {
            const cx<T>* q = reinterpret_cast<const cx<T>*>(a.in[f]) + base;
#pragma unroll
            for (int m = 0; m < E; ++m) {
                X[m] = *q;
                q += step;
            }
        }
        const T* ktab = nullptr;
        for (int o = 0; o < a.nop; ++o)
            if (a.op[o].in == f && a.op[o].table) { ktab = reinterpret_cast<const T*>(a.op[o].table); break; }
        // the operator table (kappa or |k|^p) is loaded with the input when KWH_MIDDLE_KAPPA_EARLY
        // is set, else after the forward transform
        auto load_table = [&] {
            if (ktab && ktab != kap_src) {
                const T* q = ktab + hbase;
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    kap[m] = *q * scale;
                    q += hstep;
                }
                kap_src = ktab;
            }
        };
        if (f == 0 && a.slab_n > 0) {
            // input nonzero only in slabs x0 = slab_lo + s: X[k] = sum_s x[x0] W_L^(x0 k), formed
            // directly (slab_n loads and complex products per element instead of E loads and a
            // transform); k = kb + kst m is the element slot m holds after a forward transform
            load_table();
            const cx<T>* col = reinterpret_cast<const cx<T>*>(a.in[0]) + (base - (int64_t)j * a.stride);
            const cx<T>* twl = tab.tw[a.axis];
#pragma unroll
            for (int m = 0; m < E; ++m) X[m] = cx<T>{T(0), T(0)};
            for (int sl = 0; sl < a.slab_n; ++sl) {
                const int x0 = a.slab_lo + sl;
                // W_L^(x0 k) = W_L^(x0 kb) W_L^(x0 kst m): a per-lane factor times one that is
                // uniform over the wave (scalar loads)
                const cx<T> xs = cmul(col[(int64_t)x0 * a.stride], twl[(x0 * kb) % S::L]);
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    const cx<T> w = twl[(x0 * kst * m) % S::L];
                    X[m].x += xs.x * w.x - xs.y * w.y;
                    X[m].y += xs.x * w.y + xs.y * w.x;
                }
            }
        } else {
            if (KWH_MIDDLE_KAPPA_EARLY) load_table();
            if constexpr (XL) xl.template to_block<false>(X);
            else S::template transform<false>(X, lds, j, w, lt);
            if (!KWH_MIDDLE_KAPPA_EARLY) load_table();
        }
        for (int o = 0; o < a.nop; ++o) {
            const MidOp op = a.op[o];
            if (op.in != f) continue;
            const bool use_kappa = op.table != nullptr;
            const bool use_vec = op.kind != OP_TAB;
            const cx<T>* tab1 = op.kind == OP_DDXP_KAPPA || op.kind == OP_DDXP ? tab.ddxp[op.axis]
                              : op.kind == OP_DDXN_KAPPA ? tab.ddxn[op.axis] : tab.coll[op.axis];
            const bool vec_on_mid = op.axis == a.axis;
            const int fixed_idx = op.axis == 0 ? i0c : (op.axis == 1 ? i1c : i2c);
            const cx<T> fixed_vec = use_vec && !vec_on_mid ? tab1[fixed_idx] : cx<T>{T(1), T(0)};
            const cx<T>* tj = tab1 + kb;
            cx<T> v[E];
#if KWH_MIDDLE_VERSIONED
            // The operator's kind is uniform over the workgroup: choose the element loop once, so
            // the loop body is straight-line (its table loads issue together, no per-element branch).
            if (use_vec && vec_on_mid) {
                cx<T> mv[E];
#pragma unroll
                for (int m = 0; m < E; ++m) mv[m] = tj[kst * m];
                if (use_kappa) {
#pragma unroll
                    for (int m = 0; m < E; ++m) v[m] = cmul(X[m], mv[m] * kap[m]);
                } else {
#pragma unroll
                    for (int m = 0; m < E; ++m) v[m] = cmul(X[m], mv[m] * scale);
                }
            } else if (use_kappa) {
#pragma unroll
                for (int m = 0; m < E; ++m) v[m] = cmul(X[m], fixed_vec * kap[m]);
            } else {
                const cx<T> f = fixed_vec * scale;
#pragma unroll
                for (int m = 0; m < E; ++m) v[m] = cmul(X[m], f);
            }
#else
#pragma unroll
            for (int m = 0; m < E; ++m) {
                cx<T> mv = use_vec ? (vec_on_mid ? tj[kst * m] : fixed_vec) : cx<T>{T(1), T(0)};
                const T s = use_kappa ? kap[m] : scale;
                v[m] = cmul(X[m], cx<T>{mv.x * s, mv.y * s});
            }
#endif
            if constexpr (XL) xl.template to_cyclic<true>(v);
            else S::template transform<true>(v, lds, j, w, lt);
            if (valid) {
                cx<T>* q = reinterpret_cast<cx<T>*>(op.out) + base;
#pragma unroll
                for (int m = 0; m < E; ++m) {
                    *q = v[m];
                    q += step;
                }
            }
        }
    }
}
template <typename T, int J, int E, int NTR, bool XL = false>
__global__ void __launch_bounds__(256, (reg_waves<T, E>(KWH_MIDDLE_WAVES)))
    reg_middle(MidArgs a, AxisTabs<T> tab, Geo g) {
    reg_middle_body<T, J, E, NTR, XL>(bidx_here(), a, tab, g);
}

}  // namespace kwh
