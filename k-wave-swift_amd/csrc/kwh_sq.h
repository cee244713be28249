// kwh_sq.h -- register-resident kernels for axis lengths L = R*R (R = 16: 256, R = 8: 64).
//
// A length-L transform is two radix-R stages. Lane j (0..R-1) of a transform holds elements
// n = j + R*r (r = 0..R-1) in registers:
//   stage A (Stockham Ns = 1): DFT_R over r, result to LDS at n' = j*R + r;
//   stage B (Ns = R): read n = j + R*r, twiddle W_L^(j*r), DFT_R -> element j + R*r in registers.
// Input and output of a transform are therefore both in the lane-natural pattern n = j + R*r, so
// an inverse, an elementwise update and a forward chain without extra LDS traffic, and the
// outputs of a forward feed the next inverse directly (middle pass). One LDS transpose per
// transform; every lane busy in both stages.
#pragma once
#include "kwh_kernels.h"

// Occupancy floors (waves per SIMD) of the fused row kernels: without them the register
// allocator spends the whole register file on load hoisting and drops a wave per SIMD. The
// density kernels that also accumulate rho0 * div u stay at two.
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

namespace kwh {

template <typename T, int R, bool INV>
__device__ __forceinline__ void dftR(cx<T>* v) {
    dft<T, R, INV>(v);
}

// ------------------------------------------------------------------------------------------
// Contiguous-axis transforms of a row pair (a + i b), lanes j of transform w in a workgroup.
// LDS per transform: L + R complex (one pad slot per R elements).
// ------------------------------------------------------------------------------------------
// Twiddles W_L^(j*r) of the second stage, in an LDS table laid out [r][j] (R*R elements at the
// start of the workgroup's LDS): lane j reads slot j + R*r, a base + immediate address, and the
// table costs no registers between transforms. The inverse uses the conjugates. The table is
// filled cooperatively at kernel start; the barrier inside the first transform orders the fill
// before any read.
template <typename T, int R>
struct LaneTw {
    const cx<T>* p;
    __device__ __forceinline__ void load(cx<T>* tab_lds, const cx<T>* __restrict__ tw, int j) {
        for (int i = threadIdx.x; i < R * R; i += blockDim.x) tab_lds[i] = tw[(i % R) * (i / R)];
        p = tab_lds + j;
    }
    __device__ __forceinline__ cx<T> w(int r) const { return p[R * r]; }
};

template <typename T, int R>
struct SqRow {
    static constexpr int L = R * R;
    static constexpr int LP = L + R + 1;
    // padded LDS slot of element n is n + n / R; for the two access patterns used:
    //   n = j*R + r -> j*(R+1) + r        n = j + R*r -> j + (R+1)*r
    __device__ static __forceinline__ int pidx(int n) { return n + n / R; }

    // Two-stage transform of v (lane-natural in, lane-natural out). lds: this transform's slab.
    template <bool INV>
    __device__ static __forceinline__ void transform(cx<T>* v, cx<T>* lds, int j, const LaneTw<T, R>& tw) {
        dftR<T, R, INV>(v);
        cx<T>* wr = lds + j * (R + 1);
#pragma unroll
        for (int r = 0; r < R; ++r) wr[r] = v[r];
        __syncthreads();
        const cx<T>* rd = lds + j;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            cx<T> x = rd[(R + 1) * r];
            if (r > 0) x = cmul(x, INV ? conjv(tw.w(r)) : tw.w(r));
            v[r] = x;
        }
        dftR<T, R, INV>(v);
        __syncthreads();
    }

    // Load the Hermitian-extended Z = A + iB of rows (ra, rb) in the lane-natural pattern.
    // Elements n = j + R r < L/2 read k = n from base (j); n >= L/2 read k = L - n from base
    // (L - j) with negative compile-time offsets, so every load is base + immediate.
    __device__ static __forceinline__ void load_spectrum(const cx<T>* __restrict__ spec, int ra, int rb, int nrows,
                                                         int mzh, int j, cx<T>* v) {
        const T ma = ra < nrows ? T(1) : T(0), mb = rb < nrows ? T(1) : T(0);
        const cx<T>* sa = spec + (int64_t)(ra < nrows ? ra : 0) * mzh;
        const cx<T>* sb = spec + (int64_t)(rb < nrows ? rb : 0) * mzh;
        const cx<T>* la = sa + j;
        const cx<T>* lb = sb + j;
        const cx<T>* ha = sa + (L - j);
        const cx<T>* hb = sb + (L - j);
#pragma unroll
        for (int r = 0; r < R / 2; ++r) {
            const cx<T> A = la[R * r] * ma;
            const cx<T> B = lb[R * r] * mb;
            v[r] = cx<T>{A.x - B.y, A.y + B.x};
        }
#pragma unroll
        for (int r = R / 2; r < R; ++r) {
            const cx<T> A = ha[-R * r] * ma;
            const cx<T> B = hb[-R * r] * mb;
            v[r] = cx<T>{A.x + B.y, B.x - A.y};
        }
        if (j == 0) {   // DC (n = 0) and Nyquist (n = L/2) bins are real parts only
            v[0] = cx<T>{sa[0].x * ma, sb[0].x * mb};
            v[R / 2] = cx<T>{sa[L / 2].x * ma, sb[L / 2].x * mb};
        }
    }

    // Sum of the Hermitian-extended spectra of nf fields (rows ra, rb): the spectrum of the sum.
    __device__ static __forceinline__ void load_spectrum_sum(const cx<T>* const* specs, int nf, int ra, int rb,
                                                             int nrows, int mzh, int j, cx<T>* v) {
        load_spectrum(specs[0], ra, rb, nrows, mzh, j, v);
        for (int f = 1; f < nf; ++f) {
            cx<T> t[R];
            load_spectrum(specs[f], ra, rb, nrows, mzh, j, t);
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = v[r] + t[r];
        }
    }

    // Forward of v (a in .x, b in .y), split into half spectra, stored to rows (ra, rb).
    __device__ static __forceinline__ void forward_store(cx<T>* spec, int ra, int rb, int nrows, int mzh, cx<T>* lds,
                                                         int j, const LaneTw<T, R>& tw, cx<T>* v) {
        transform<false>(v, lds, j, tw);
        cx<T>* wr = lds + j;
#pragma unroll
        for (int r = 0; r < R; ++r) wr[(R + 1) * r] = v[r];
        __syncthreads();
        const bool va = ra < nrows, vb = rb < nrows;
        // k = j + R m: slot j + (R+1) m; mirror L - k: slot (L - j + R - [j>0]) - (R+1) m (k > 0)
        const cx<T>* pk = lds + j;
        const cx<T>* pm = lds + (L - j + R - (j > 0 ? 1 : 0));
        cx<T>* oa = spec + (int64_t)(va ? ra : 0) * mzh + j;
        cx<T>* ob = spec + (int64_t)(vb ? rb : 0) * mzh + j;
#pragma unroll
        for (int m = 0; m <= R / 2; ++m) {
            const int k = j + R * m;
            if (k > L / 2) break;
            const cx<T> zk = pk[(R + 1) * m];
            const cx<T> zm = k == 0 ? zk : pm[-(R + 1) * m];
            if (va) oa[R * m] = cx<T>{T(0.5) * (zk.x + zm.x), T(0.5) * (zk.y - zm.y)};
            if (vb) ob[R * m] = cx<T>{T(0.5) * (zk.y + zm.y), T(0.5) * (zm.x - zk.x)};
        }
        __syncthreads();
    }
};

// Row pair handled by a lane group: P transforms per workgroup, R lanes each.
template <int R>
struct SqZPos {
    int w, j, q, ra, rb;
    __device__ SqZPos(int P, const BIdx& bi) {
        const int tid = threadIdx.x;
        w = tid / R;
        j = tid - w * R;
        const int tile = bi.tile();
        q = tile * P + w;
        ra = 2 * q;
        rb = ra + 1;
    }
};

// Sparse sources on register-resident rows: go through LDS (natural order) when the tile has any.
template <typename T, int R>
__device__ void sq_apply_sparse(const SparseSrc<T>& s, const Geo& g, int row0, int nrow_tile, cx<T>* lds_all,
                                int w, int j, int t, cx<T>* v) {
    using S = SqRow<T, R>;
    cx<T>* lds = lds_all + w * S::LP;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[S::pidx(j + R * r)] = v[r];
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
            T& slot = comp(lds_all[(lr >> 1) * S::LP + S::pidx(n)], lr & 1);
            slot = s.mode == 0 ? val : slot + val;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = lds[S::pidx(j + R * r)];
    __syncthreads();
}

template <typename T, int R>
__device__ __forceinline__ void sq_z_forward_real_body(const BIdx& bi, const T* __restrict__ src, cx<T>* spec, T scale, const Geo& g,
                                                         int P, const AxisTabs<T>& tab) {
    using S = SqRow<T, R>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds = ldsb + R * R;
    SqZPos<R> ps(P, bi);
    LaneTw<T, R> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    const bool va = ps.ra < g.nrows, vb = ps.rb < g.nrows;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int n = ps.j + R * r;
        v[r].x = va ? src[(int64_t)ps.ra * S::L + n] * scale : T(0);
        v[r].y = vb ? src[(int64_t)ps.rb * S::L + n] * scale : T(0);
    }
    S::forward_store(spec, ps.ra, ps.rb, g.nrows, g.mzh, lds + ps.w * S::LP, ps.j, lt, v);
}
template <typename T, int R>
__global__ void __launch_bounds__(256) sq_z_forward_real(const T* __restrict__ src, cx<T>* spec, T scale, Geo g,
                                                         int P, AxisTabs<T> tab) {
    sq_z_forward_real_body<T, R>(bidx_here(), src, spec, scale, g, P, tab);
}

template <typename T, int R>
__device__ __forceinline__ void sq_z_inverse_real_body(const BIdx& bi, const cx<T>* spec, T* __restrict__ dst, T scale,
                                                         int accumulate, const Geo& g, int P, const AxisTabs<T>& tab) {
    using S = SqRow<T, R>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds = ldsb + R * R;
    SqZPos<R> ps(P, bi);
    LaneTw<T, R> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    cx<T> v[R];
    S::load_spectrum(spec, ps.ra, ps.rb, g.nrows, g.mzh, ps.j, v);
    S::template transform<true>(v, lds + ps.w * S::LP, ps.j, lt);
    const bool va = ps.ra < g.nrows, vb = ps.rb < g.nrows;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int n = ps.j + R * r;
        if (va) {
            const int64_t c = (int64_t)ps.ra * S::L + n;
            dst[c] = accumulate ? dst[c] + v[r].x * scale : v[r].x * scale;
        }
        if (vb) {
            const int64_t c = (int64_t)ps.rb * S::L + n;
            dst[c] = accumulate ? dst[c] + v[r].y * scale : v[r].y * scale;
        }
    }
}
template <typename T, int R>
__global__ void __launch_bounds__(256) sq_z_inverse_real(const cx<T>* spec, T* __restrict__ dst, T scale,
                                                         int accumulate, Geo g, int P, AxisTabs<T> tab) {
    sq_z_inverse_real_body<T, R>(bidx_here(), spec, dst, scale, accumulate, g, P, tab);
}

// Collocated velocity along the contiguous axis for a list of rows: dst[row] =
// C2R_z(coll_z * Z[row]) / L, Z = R2C_z(u) of the updated component (as the velocity kernel
// leaves it). The shift e^{-i k dz/2} only involves the contiguous axis, so the full 3-D
// transform of the reference (KSpaceFirstOrder.swift:735-745) reduces to this 1-D one on the
// rows the sensor reads. Rows are paired in list order; an odd tail pairs with nothing.
template <typename T, int R>
__device__ __forceinline__ void sq_z_colloc_rows_body(const BIdx& bi, const cx<T>* spec, T* __restrict__ dst,
                                                        const int* __restrict__ rows, int nlist, const Geo& g, int P,
                                                        const AxisTabs<T>& tab) {
    using S = SqRow<T, R>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    SqZPos<R> ps(P, bi);
    LaneTw<T, R> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    const bool va = ps.ra < nlist, vb = ps.rb < nlist;
    const int ra = va ? rows[ps.ra] : g.nrows;   // g.nrows: masked by load_spectrum
    const int rb = vb ? rows[ps.rb] : g.nrows;
    cx<T> v[R];
    S::load_spectrum(spec, ra, rb, g.nrows, g.mzh, ps.j, v);
    const cx<T>* co = tab.coll[2] + ps.j;
    const T scale = T(1) / T(S::L);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const cx<T> c = co[R * r];
        v[r] = cmul(v[r], cx<T>{c.x * scale, c.y * scale});
    }
    S::template transform<true>(v, ldsb + R * R + ps.w * S::LP, ps.j, lt);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int n = ps.j + R * r;
        if (va) dst[(int64_t)ra * S::L + n] = v[r].x;
        if (vb) dst[(int64_t)rb * S::L + n] = v[r].y;
    }
}
template <typename T, int R>
__global__ void __launch_bounds__(256) sq_z_colloc_rows(const cx<T>* spec, T* __restrict__ dst,
                                                        const int* __restrict__ rows, int nlist, Geo g, int P,
                                                        AxisTabs<T> tab) {
    sq_z_colloc_rows_body<T, R>(bidx_here(), spec, dst, rows, nlist, g, P, tab);
}

// Scheduling fence: keeps the compiler from hoisting the loads of later element groups above
// earlier ones (bounds live registers; other waves hide the latency).
#define KWH_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
constexpr int kGroup = 8;      // density update
constexpr int kVelGroup = 4;   // velocity update (keeps it within the three-wave register budget)

// Velocity update, per axis: C2R_z(G_a) -> u_a = pml_sg (pml_sg u_a - dt/rho_a g) [+ sources] -> R2C_z.
template <typename T, int R, bool CORR>
__device__ __forceinline__ void sq_z_velocity_body(const BIdx& bi, const VelArgs<T>& a, const Geo& g, int P, const AxisTabs<T>& tab) {
    using S = SqRow<T, R>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds_all = ldsb + R * R;
    SqZPos<R> ps(P, bi);
    LaneTw<T, R> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    cx<T>* lds = lds_all + ps.w * S::LP;
    const bool va = ps.ra < g.nrows, vb = ps.rb < g.nrows;
    const int ra = va ? ps.ra : 0, rb = vb ? ps.rb : 0;
    const int64_t oa = (int64_t)ra * S::L + ps.j, ob = (int64_t)rb * S::L + ps.j;
    for (int ax = 0; ax < g.dim; ++ax) {
        const int ia = ax + g.off;
        cx<T> v[R];
        S::load_spectrum(a.G[ax], ps.ra, ps.rb, g.nrows, g.mzh, ps.j, v);
        S::template transform<true>(v, lds, ps.j, lt);
        KWH_SCHED_FENCE();
        T* u = a.u[ax];
        const SMap<T> dtr = a.dtr[ax];
        const SMap<T> corr = a.corr[ax];
        const T* pm = tab.pmlsg[ia];
        const T* ua = u + oa;
        const T* ub = u + ob;
        const T* dta = dtr.lane(oa, ps.j);
        const T* dtb = dtr.lane(ob, ps.j);
        const T* cra = corr.lane(oa, ps.j);
        const T* crb = corr.lane(ob, ps.j);
        auto upd = [&](int r, T qa, T qb) {
            T xa = qa * (qa * ua[R * r] - dta[R * r] * v[r].x);
            T xb = qb * (qb * ub[R * r] - dtb[R * r] * v[r].y);
            if constexpr (CORR) {
                xa += cra[R * r];
                xb += crb[R * r];
            }
            v[r].x = va ? xa : T(0);
            v[r].y = vb ? xb : T(0);
        };
        if (ia == 2) {   // staggered PML along z varies with n and is shared by both rows
            const T* pz = pm + ps.j;
#pragma unroll
            for (int g0 = 0; g0 < R; g0 += kVelGroup) {
#pragma unroll
                for (int r = g0; r < g0 + kVelGroup && r < R; ++r) {
                    const T q = pz[R * r];
                    upd(r, q, q);
                }
                KWH_SCHED_FENCE();
            }
        } else {
            const T qa = pm[axis_pos(g, ra, 0, ia)], qb = pm[axis_pos(g, rb, 0, ia)];
#pragma unroll
            for (int g0 = 0; g0 < R; g0 += kVelGroup) {
#pragma unroll
                for (int r = g0; r < g0 + kVelGroup && r < R; ++r) upd(r, qa, qb);
                KWH_SCHED_FENCE();
            }
        }
        if (sparse_in_tile(a.sp[ax].rowptr, g, ZTile{(int)(bi.tile() * P), P, 0, 0}))
            sq_apply_sparse<T, R>(a.sp[ax], g, 2 * bi.tile() * P, 2 * P, lds_all, ps.w, ps.j,
                                  a.ctr[0], v);
        {
            T* wa = u + oa;
            T* wb = u + ob;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (va) wa[R * r] = v[r].x;
                if (vb) wb[R * r] = v[r].y;
            }
        }
        S::forward_store(a.G[ax], ps.ra, ps.rb, g.nrows, g.mzh, lds, ps.j, lt, v);
    }
}
template <typename T, int R, bool CORR>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? KWH_VELOCITY_WAVES : 1) sq_z_velocity(VelArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    sq_z_velocity_body<T, R, CORR>(bidx_here(), a, g, P, tab);
}

// Density update, per axis: C2R_z(D_a) -> rho_a = pml (pml rho_a - dt rho0 [nl] du_a) [+ sources],
// followed by the EOS (MODE 0), the EOS with Stokes absorption (MODE 2) or the inputs of the
// FFT absorption term (MODE 1: rhoT, F1 = R2C_z(rho0 sum_a du_a), F2 = R2C_z(rhoT)). NL: B/A.
// The total density and, for MODE 1/2, rho0 * sum_a du_a are accumulated in registers while the
// axes are processed, so each divergence spectrum is read once.
template <typename T, int R, bool NL, int MODE, bool CORR>
__device__ __forceinline__ void sq_z_density_body(const BIdx& bi, const RhoArgs<T>& a, const Geo& g, int P, const AxisTabs<T>& tab) {
    using S = SqRow<T, R>;
    constexpr bool ACC = MODE != 0;   // accumulate rho0 * sum_a du_a
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds_all = ldsb + R * R;
    SqZPos<R> ps(P, bi);
    LaneTw<T, R> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    cx<T>* lds = lds_all + ps.w * S::LP;
    const bool va = ps.ra < g.nrows, vb = ps.rb < g.nrows;
    const int ra = va ? ps.ra : 0, rb = vb ? ps.rb : 0;
    const int row0 = 2 * bi.tile() * P;
    const int64_t oa = (int64_t)ra * S::L + ps.j, ob = (int64_t)rb * S::L + ps.j;
    const T* r0pa = a.rho0.lane(oa, ps.j);
    const T* r0pb = a.rho0.lane(ob, ps.j);
    cx<T> rt[R];
    cx<T> sd[ACC ? R : 1];
    cx<T> rp[NL ? R : 1];
    if constexpr (NL) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            T sa = 0, sb = 0;
            for (int ax = 0; ax < g.dim; ++ax) {
                sa += a.rho[ax][oa + R * r];
                sb += a.rho[ax][ob + R * r];
            }
            rp[r] = cx<T>{sa, sb};
        }
    }
    for (int ax = 0; ax < g.dim; ++ax) {
        const int ia = ax + g.off;
        cx<T> v[R];
        S::load_spectrum(a.D[ax], ps.ra, ps.rb, g.nrows, g.mzh, ps.j, v);
        S::template transform<true>(v, lds, ps.j, lt);
        KWH_SCHED_FENCE();
        T* rho = a.rho[ax];
        const T* pm = tab.pml[ia];
        const T cf = a.corr_factor[ax];
        const T* rha = rho + oa;
        const T* rhb = rho + ob;
        const T* cra = a.corr.lane(oa, ps.j);
        const T* crb = a.corr.lane(ob, ps.j);
        auto upd = [&](int r, T qa, T qb) {
            const T r0a = r0pa[R * r];
            const T r0b = r0pb[R * r];
            const T da = r0a * v[r].x, db = r0b * v[r].y;
            if constexpr (ACC) {
                if (ax == 0) sd[r] = cx<T>{da, db};
                else { sd[r].x += da; sd[r].y += db; }
            }
            T ma = a.dt * da, mb = a.dt * db;
            if constexpr (NL) {
                ma *= (T(2) * rp[r].x + r0a) / r0a;
                mb *= (T(2) * rp[r].y + r0b) / r0b;
            }
            T xa = qa * (qa * rha[R * r] - ma);
            T xb = qb * (qb * rhb[R * r] - mb);
            if constexpr (CORR) {
                xa += cra[R * r] * cf;
                xb += crb[R * r] * cf;
            }
            v[r] = cx<T>{xa, xb};
        };
        if (ia == 2) {
            const T* pz = pm + ps.j;
#pragma unroll
            for (int g0 = 0; g0 < R; g0 += kGroup) {
#pragma unroll
                for (int r = g0; r < g0 + kGroup && r < R; ++r) {
                    const T q = pz[R * r];
                    upd(r, q, q);
                }
                KWH_SCHED_FENCE();
            }
        } else {
            const T qa = pm[axis_pos(g, ra, 0, ia)], qb = pm[axis_pos(g, rb, 0, ia)];
#pragma unroll
            for (int g0 = 0; g0 < R; g0 += kGroup) {
#pragma unroll
                for (int r = g0; r < g0 + kGroup && r < R; ++r) upd(r, qa, qb);
                KWH_SCHED_FENCE();
            }
        }
        if (sparse_in_tile(a.sp[ax].rowptr, g, ZTile{row0 / 2, P, 0, 0}))
            sq_apply_sparse<T, R>(a.sp[ax], g, row0, 2 * P, lds_all, ps.w, ps.j, a.ctr[0], v);
        T* wa = rho + oa;
        T* wb = rho + ob;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (va) wa[R * r] = v[r].x;
            if (vb) wb[R * r] = v[r].y;
            if (ax == 0) rt[r] = v[r];
            else { rt[r].x += v[r].x; rt[r].y += v[r].y; }
        }
    }
    cx<T> v[R];
    if constexpr (MODE != 1) {
        const T* tpa = a.tau.lane(oa, ps.j);
        const T* tpb = a.tau.lane(ob, ps.j);
        const T* bpa = a.bona.lane(oa, ps.j);
        const T* bpb = a.bona.lane(ob, ps.j);
        const T* cpa = a.c0.lane(oa, ps.j);
        const T* cpb = a.c0.lane(ob, ps.j);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            T ea = rt[r].x, eb = rt[r].y;
            const T r0a = r0pa[R * r];
            const T r0b = r0pb[R * r];
            if constexpr (MODE == 2) {
                ea += tpa[R * r] * sd[r].x;
                eb += tpb[R * r] * sd[r].y;
            }
            if constexpr (NL) {
                ea += bpa[R * r] * (rt[r].x * rt[r].x) / (T(2) * r0a);
                eb += bpb[R * r] * (rt[r].y * rt[r].y) / (T(2) * r0b);
            }
            const T cA = cpa[R * r];
            const T cB = cpb[R * r];
            const T pA = cA * cA * ea, pB = cB * cB * eb;
            if (va) a.p[oa + R * r] = pA;
            if (vb) a.p[ob + R * r] = pB;
            v[r] = cx<T>{va ? pA : T(0), vb ? pB : T(0)};
        }
        S::forward_store(a.outP, ps.ra, ps.rb, g.nrows, g.mzh, lds, ps.j, lt, v);
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = cx<T>{va ? sd[r].x : T(0), vb ? sd[r].y : T(0)};
        S::forward_store(a.F1, ps.ra, ps.rb, g.nrows, g.mzh, lds, ps.j, lt, v);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (va) a.rhoT[oa + R * r] = rt[r].x;
            if (vb) a.rhoT[ob + R * r] = rt[r].y;
            rt[r] = cx<T>{va ? rt[r].x : T(0), vb ? rt[r].y : T(0)};
        }
        if (a.dispersion) S::forward_store(a.F2, ps.ra, ps.rb, g.nrows, g.mzh, lds, ps.j, lt, rt);
    }
}
template <typename T, int R, bool NL, int MODE, bool CORR>
__global__ void __launch_bounds__(256, (NL || CORR || sizeof(T) != 4) ? 1 : (MODE == 0 ? KWH_DENSITY_WAVES : 2)) sq_z_density(RhoArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    sq_z_density_body<T, R, NL, MODE, CORR>(bidx_here(), a, g, P, tab);
}

// EOS with FFT absorption: p = c^2 (rhoT + tau C2R_z(A1) - eta C2R_z(A2) [+ B/A rhoT^2 / (2 rho0)]),
// P = R2C_z(p). Both spectra are loaded before the first transform (one memory phase).
template <typename T, int R>
__device__ __forceinline__ void sq_z_eos_body(const BIdx& bi, const EosArgs<T>& a, const Geo& g, int P, const AxisTabs<T>& tab) {
    using S = SqRow<T, R>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    SqZPos<R> ps(P, bi);
    LaneTw<T, R> lt;
    lt.load(ldsb, tab.tw[2], ps.j);
    cx<T>* lds = ldsb + R * R + ps.w * S::LP;
    const bool va = ps.ra < g.nrows, vb = ps.rb < g.nrows;
    const int ra = va ? ps.ra : 0, rb = vb ? ps.rb : 0;
    cx<T> e[R], v[R];
    S::load_spectrum(a.A1, ps.ra, ps.rb, g.nrows, g.mzh, ps.j, e);
    if (a.A2) S::load_spectrum(a.A2, ps.ra, ps.rb, g.nrows, g.mzh, ps.j, v);
    S::template transform<true>(e, lds, ps.j, lt);
    if (a.A2) S::template transform<true>(v, lds, ps.j, lt);
    KWH_SCHED_FENCE();
    const int64_t oa = (int64_t)ra * S::L + ps.j, ob = (int64_t)rb * S::L + ps.j;
    const T* cpa = a.c0.lane(oa, ps.j);
    const T* cpb = a.c0.lane(ob, ps.j);
    const T* tpa = a.tau.lane(oa, ps.j);
    const T* tpb = a.tau.lane(ob, ps.j);
    const T* epa = a.eta.lane(oa, ps.j);
    const T* epb = a.eta.lane(ob, ps.j);
    const T* bpa = a.bona.lane(oa, ps.j);
    const T* bpb = a.bona.lane(ob, ps.j);
    const T* rpa = a.rho0.lane(oa, ps.j);
    const T* rpb = a.rho0.lane(ob, ps.j);
#pragma unroll
    for (int g0 = 0; g0 < R; g0 += kGroup) {
#pragma unroll
        for (int r = g0; r < g0 + kGroup && r < R; ++r) {
            const int64_t ca = oa + R * r, cb = ob + R * r;
            const T rta = a.rhoT[ca], rtb = a.rhoT[cb];
            const T cA = cpa[R * r], cB = cpb[R * r];
            T tauA, etaA, tauB, etaB;
            if (a.tau_eta_from_c) {
                const T la = log2(cA), lb = log2(cB);
                tauA = a.tk * exp2(a.te * la);
                tauB = a.tk * exp2(a.te * lb);
                etaA = a.ek * exp2(a.ee * la);
                etaB = a.ek * exp2(a.ee * lb);
            } else {
                tauA = tpa[R * r];
                tauB = tpb[R * r];
                etaA = epa[R * r];
                etaB = epb[R * r];
            }
            T ea = rta + tauA * e[r].x;
            T eb = rtb + tauB * e[r].y;
            if (a.A2) {
                ea -= etaA * v[r].x;
                eb -= etaB * v[r].y;
            }
            if (a.nonlinear) {
                ea += bpa[R * r] * (rta * rta) / (T(2) * rpa[R * r]);
                eb += bpb[R * r] * (rtb * rtb) / (T(2) * rpb[R * r]);
            }
            const T pA = cA * cA * ea, pB = cB * cB * eb;
            if (va) a.p[ca] = pA;
            if (vb) a.p[cb] = pB;
            e[r] = cx<T>{va ? pA : T(0), vb ? pB : T(0)};
        }
        KWH_SCHED_FENCE();
    }
    S::forward_store(a.outP, ps.ra, ps.rb, g.nrows, g.mzh, lds, ps.j, lt, e);
}
template <typename T, int R>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? KWH_EOS_WAVES : 1) sq_z_eos(EosArgs<T> a, Geo g, int P, AxisTabs<T> tab) {
    sq_z_eos_body<T, R>(bidx_here(), a, g, P, tab);
}

// ------------------------------------------------------------------------------------------
// Strided / middle passes over a tile of ntr contiguous columns: lane (w, j), w fastest.
// LDS: element n of column w at n * ntr + w + n / R (pad one slot per R rows).
// ------------------------------------------------------------------------------------------
template <typename T, int R, int NTR>
struct SqCol {
    static constexpr int L = R * R;
    // element n of column w at n*NTR + w + n/R:
    //   stage A write n = j*R + r -> (j*R*NTR + w + j) + r*NTR
    //   stage B read  n = j + R*r -> (j*NTR + w) + r*(R*NTR + 1)
    template <bool INV>
    __device__ static __forceinline__ void transform(cx<T>* v, cx<T>* lds, int j, int w, const LaneTw<T, R>& tw) {
        dftR<T, R, INV>(v);
        cx<T>* wr = lds + (j * R * NTR + w + j);
#pragma unroll
        for (int r = 0; r < R; ++r) wr[r * NTR] = v[r];
        __syncthreads();
        const cx<T>* rd = lds + (j * NTR + w);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            cx<T> x = rd[r * (R * NTR + 1)];
            if (r > 0) x = cmul(x, INV ? conjv(tw.w(r)) : tw.w(r));
            v[r] = x;
        }
        dftR<T, R, INV>(v);
        __syncthreads();
    }
};

template <typename T, int R, int NTR, bool INV>
__device__ __forceinline__ void sq_strided_body(const BIdx& bi, const StridedArgs& a, const cx<T>* __restrict__ tw) {
    using S = SqCol<T, R, NTR>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds = ldsb + R * R;
    const int w = threadIdx.x % NTR, j = threadIdx.x / NTR;
    LaneTw<T, R> lt;
    lt.load(ldsb, tw, j);
    const int tile = bi.tile();
    const int64_t c = (int64_t)tile * NTR + w;
    const bool valid = c < a.ncols;
    const int64_t cc = valid ? c : 0;
    const int64_t base = (cc / a.inner) * (int64_t)S::L * a.inner + cc % a.inner;
    cx<T>* data = reinterpret_cast<cx<T>*>(a.fields[bi.y]) + base + (int64_t)j * a.stride;
    const int64_t step = (int64_t)R * a.stride;
    cx<T> v[R];
    {
        const cx<T>* q = data;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[r] = *q;
            q += step;
        }
    }
    S::template transform<INV>(v, lds, j, w, lt);
    if (valid) {
        cx<T>* q = data;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            *q = v[r];
            q += step;
        }
    }
}
template <typename T, int R, int NTR, bool INV>
__global__ void __launch_bounds__(256) sq_strided(StridedArgs a, const cx<T>* __restrict__ tw) {
    sq_strided_body<T, R, NTR, INV>(bidx_here(), a, tw);
}

// Middle pass: forward along the middle axis, operators, inverse, store. Every table-using
// operator of one input reads the same real k-space table (checked on the host), loaded once.
template <typename T, int R, int NTR>
__device__ __forceinline__ void sq_middle_body(const BIdx& bi, const MidArgs& a, const AxisTabs<T>& tab, const Geo& g) {
    using S = SqCol<T, R, NTR>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* const ldsb = reinterpret_cast<cx<T>*>(smem);
    cx<T>* lds = ldsb + R * R;
    const int w = threadIdx.x % NTR, j = threadIdx.x / NTR;
    LaneTw<T, R> lt;
    lt.load(ldsb, tab.tw[a.axis], j);
    const int tile = bi.tile();
    const int64_t c = (int64_t)tile * NTR + w;
    const bool valid = c < a.ncols;
    const int64_t cc = valid ? c : 0;
    const int64_t base = (cc / a.inner) * (int64_t)S::L * a.inner + cc % a.inner + (int64_t)j * a.stride;
    const int64_t step = (int64_t)R * a.stride;
    const T scale = T(a.scale);
    int i0c, i1c, i2c;
    if (a.axis == 0) { i0c = 0; i1c = (int)(cc / g.mzh); i2c = (int)(cc % g.mzh); }
    else { i0c = (int)(cc / g.mzh); i1c = 0; i2c = (int)(cc % g.mzh); }
    const int64_t hstride = a.axis == 0 ? (int64_t)g.m1 * g.mzh : (int64_t)g.mzh;
    const int64_t hbase = (a.axis == 0 ? cc : (int64_t)i0c * g.m1 * g.mzh + i2c) + (int64_t)j * hstride;
    const int64_t hstep = (int64_t)R * hstride;
    T kap[R];
    const T* kap_src = nullptr;   // table currently held in kap (inputs sharing one table load it once)
    for (int f = 0; f < a.nin; ++f) {
        cx<T> X[R];
        {
            const cx<T>* q = reinterpret_cast<const cx<T>*>(a.in[f]) + base;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                X[r] = *q;
                q += step;
            }
        }
        S::template transform<false>(X, lds, j, w, lt);
        const T* ktab = nullptr;
        for (int o = 0; o < a.nop; ++o)
            if (a.op[o].in == f && a.op[o].table) { ktab = reinterpret_cast<const T*>(a.op[o].table); break; }
        if (ktab && ktab != kap_src) {
            const T* q = ktab + hbase;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                kap[r] = *q * scale;
                q += hstep;
            }
            kap_src = ktab;
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
            const cx<T>* tj = tab1 + j;
            cx<T> v[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> m = use_vec ? (vec_on_mid ? tj[R * r] : fixed_vec) : cx<T>{T(1), T(0)};
                const T s = use_kappa ? kap[r] : scale;
                v[r] = cmul(X[r], cx<T>{m.x * s, m.y * s});
            }
            S::template transform<true>(v, lds, j, w, lt);
            if (valid) {
                cx<T>* q = reinterpret_cast<cx<T>*>(op.out) + base;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    *q = v[r];
                    q += step;
                }
            }
        }
    }
}
template <typename T, int R, int NTR>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? KWH_MIDDLE_WAVES : 1) sq_middle(MidArgs a, AxisTabs<T> tab, Geo g) {
    sq_middle_body<T, R, NTR>(bidx_here(), a, tab, g);
}

}  // namespace kwh
