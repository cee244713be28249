// kwh_plane.h -- y-z plane kernels: one workgroup transforms a whole (y, z) plane of a 3-D field
// in LDS, so the strided (y) pass and the contiguous (z) pass of a step are one HBM round trip.
//
// The step of the reference (KSpaceFirstOrder.swift:669-692) transforms every field along all
// three axes. The split passes (kwh_reg.h) read and write each field once per axis: z row
// kernels, y column kernels and x middle kernels. When a plane [m1][m2/2+1] of a half spectrum
// fits in LDS (64 KiB at 128 x 128, 36 KiB at 96 x 96 in float32), a plane kernel does the y and
// z transforms of one x plane together: load the plane (contiguous, coalesced), inverse along y
// and z, elementwise update, forward along z and y, store. Only the x pass stays separate.
//
// Layout in LDS: complex slots pl[r * PITCH + c], r < LY rows, c < M = LZ / 2 columns.
//  * Spectrum view: slot (ky, kz) for 0 < kz < M; slot (ky, 0) packs the kz = 0 and kz = M
//    (Nyquist) columns as G0 + i GN. Both columns are transforms along y of real sequences
//    (the z spectrum of real rows is real at DC and Nyquist), so one complex column transform
//    carries both, and every row has exactly M complex slots.
//  * Real view: row r holds its LZ real values x[2m] + i x[2m+1] in slot m, the half-length
//    complex sequence of the real transform identity (Cooley-Tukey split of the row into even
//    and odd samples), so the spectrum and the real plane share one buffer.
// Row transforms are length-M complex DFTs with the even/odd split before (inverse) or after
// (forward); column transforms are length-LY complex DFTs. Each is a two-factor split L = R1 R2
// done in place by all threads of the workgroup (stage A: DFT_R2 over n2 per n1, twiddles, back
// to the same slots; stage B: DFT_R1 over n1 per k2, natural-order output after a barrier).
// Lanes are mapped transform-fastest: consecutive lanes touch consecutive columns (column
// stages) or rows PITCH apart (row stages, PITCH odd so rows spread over the LDS banks).
// All transforms are unnormalised; the x pass applies 1/N, as with the split passes.
#pragma once
#include "kwh_kernels.h"

namespace kwh {

// threadIdx.x behind an empty asm statement: the lane-dependent addresses of each phase are then
// computed where the phase runs instead of once for the whole kernel, where the compiler would
// keep every phase's addresses live across all of them (and spill).
__device__ __forceinline__ int plane_tid() {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Split L = R1 * R2 of the in-place transforms (n = n1 + R1 n2).
constexpr int plane_r1(int L) {
    return L == 128 ? 8 : L == 96 ? 12 : L == 64 ? 8 : L == 48 ? 8 : L == 32 ? 4 : L == 24 ? 4 : L == 16 ? 4 : 0;
}

template <typename T, int LY, int LZ>
struct PlaneShape {
    static constexpr int M = LZ / 2;                  // complex slots per row
    static constexpr int MZH = M + 1;                 // half-spectrum row length in HBM
    static constexpr int PITCH = MZH;                 // LDS row pitch = HBM row: planes copy verbatim
    static constexpr int R1Y = plane_r1(LY), R2Y = LY / R1Y;
    static constexpr int R1M = plane_r1(M), R2M = M / R1M;
    static constexpr int TWY = LY * PITCH;            // W_LY^(n1 k2) laid out [n1][k2] (LY entries)
    static constexpr int TWM = TWY + LY;              // W_M^(n1 k2) laid out [n1][k2] (M entries)
    static constexpr int TWS = TWM + M;               // W_LZ^k, k <= M / 2 (M entries reserved)
    static constexpr size_t kLds = (size_t)(TWS + M) * sizeof(cx<T>);
    static_assert(LZ % 4 == 0 && R1Y > 0 && R1M > 0, "unsupported plane shape");
};

// Occupancy floor (waves per SIMD) of a plane kernel with NT lanes: as many workgroups per CU as
// the 160 KiB of LDS hold. Without it the register allocator hoists the element loops' loads
// across the whole plane share and drops to one wave per SIMD.
template <typename T, int LY, int LZ, int NT>
constexpr int plane_waves() {
    constexpr size_t lds_cu = 160 * 1024;
    constexpr int wg = (int)(lds_cu / PlaneShape<T, LY, LZ>::kLds);
    constexpr int w = wg * (NT / 64) / 4;
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// Twiddle tables into LDS (from the per-axis tables tw[a][m] = W_La^m). Laid out [n1][k2] so that
// a task's twiddles are base + immediate addresses. Ordered before first use by the barrier that
// follows the first plane load.
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_twiddles(cx<T>* lds, const AxisTabs<T>& tab) {
    using P = PlaneShape<T, LY, LZ>;
    for (int i = threadIdx.x; i < LY + 2 * P::M; i += NT) {
        cx<T> w;
        if (i < LY) {
            w = tab.tw[1][((i / P::R2Y) * (i % P::R2Y)) % LY];
        } else if (i < LY + P::M) {
            const int j = i - LY;
            w = tab.tw[2][(2 * (j / P::R2M) * (j % P::R2M)) % LZ];
        } else {
            w = tab.tw[2][i - LY - P::M];
        }
        lds[P::TWY + i] = w;
    }
}

// In-place DFTs of NTR transforms of length L = R1 * R2 held in LDS: element n of transform t
// at buf[t * TS + n * ES]. Lanes are transform-fastest (lane = t + NTR * sub-task); NT is a
// multiple of NTR, so a lane keeps its transform across iterations. tw[n1 * R2 + k2] =
// W_L^(n1 k2) (conjugated for the inverse). Ends with a barrier.
template <typename T, int L, int NTR, int TS, int ES, int NT, bool INV>
__device__ __forceinline__ void plane_dft(cx<T>* buf, const cx<T>* tw) {
    constexpr int R1 = plane_r1(L), R2 = L / R1;
    static_assert(R1 > 0 && R1 * R2 == L, "no split for this length");
    static_assert(NT % NTR == 0, "workgroup size must be a multiple of the transform count");
    constexpr int SPI = NT / NTR;   // sub-tasks per iteration
    {
        const int tid = plane_tid();
        const int t = tid % NTR, s0 = tid / NTR;
        constexpr int IA = (R1 + SPI - 1) / SPI;
#pragma unroll
        for (int it = 0; it < IA; ++it) {
            const int n1 = s0 + it * SPI;
            if (R1 % SPI == 0 || n1 < R1) {
                cx<T>* b = buf + t * TS + n1 * ES;
                const cx<T>* w = tw + n1 * R2;
                cx<T> v[R2];
#pragma unroll
                for (int n2 = 0; n2 < R2; ++n2) v[n2] = b[n2 * R1 * ES];
                dftN<T, R2, INV>(v);
#pragma unroll
                for (int k2 = 1; k2 < R2; ++k2) v[k2] = cmul(v[k2], INV ? conjv(w[k2]) : w[k2]);
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) b[k2 * R1 * ES] = v[k2];
            }
        }
    }
    __syncthreads();
    {
        const int tid = plane_tid();
        const int t = tid % NTR, s0 = tid / NTR;
        constexpr int IB = (R2 + SPI - 1) / SPI;
        cx<T> s[IB][R1];
#pragma unroll
        for (int it = 0; it < IB; ++it) {
            const int k2 = s0 + it * SPI;
            if (R2 % SPI == 0 || k2 < R2) {
                const cx<T>* b = buf + t * TS + k2 * R1 * ES;
#pragma unroll
                for (int n1 = 0; n1 < R1; ++n1) s[it][n1] = b[n1 * ES];
                dftN<T, R1, INV>(s[it]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IB; ++it) {
            const int k2 = s0 + it * SPI;
            if (R2 % SPI == 0 || k2 < R2) {
                cx<T>* b = buf + t * TS + k2 * ES;
#pragma unroll
                for (int k1 = 0; k1 < R1; ++k1) b[k1 * R2 * ES] = s[it][k1];
            }
        }
    }
    __syncthreads();
}

template <typename T, int LY, int LZ, int NT, bool INV>
__device__ __forceinline__ void plane_cols(cx<T>* pl) {
    using P = PlaneShape<T, LY, LZ>;
    plane_dft<T, LY, P::M, 1, P::PITCH, NT, INV>(pl, pl + P::TWY);
}
template <typename T, int LY, int LZ, int NT, bool INV>
__device__ __forceinline__ void plane_rows(cx<T>* pl) {
    using P = PlaneShape<T, LY, LZ>;
    plane_dft<T, P::M, LY, P::PITCH, 1, NT, INV>(pl, pl + P::TWM);
}

// Task (r, k), k = 0 .. M/2, of the row splits: lanes row-fastest (NT a multiple of LY).
template <int LY, int M, int NT, class F>
__device__ __forceinline__ void plane_split_tasks(F&& f) {
    static_assert(NT % LY == 0, "workgroup size must be a multiple of the rows");
    constexpr int SPI = NT / LY, H = M / 2 + 1, IQ = (H + SPI - 1) / SPI;
    const int tid = plane_tid();
    const int r = tid % LY, k0 = tid / LY;
#pragma unroll
    for (int it = 0; it < IQ; ++it) {
        const int k = k0 + it * SPI;
        if (H % SPI == 0 || k < H) f(r, k);
    }
}

// Inverse real transform of every row, before plane_rows<INV = true>: the packed half spectrum X
// (slot 0 = X[0] + i X[M]) becomes Z[k] = E2[k] + i O2[k] with E2 = X[k] + conj X[M-k] and
// O2 = (X[k] - conj X[M-k]) W_LZ^-k, whose length-M inverse DFT is x[2m] + i x[2m+1] (LZ times the
// normalised inverse). Task (r, k) owns slots k and M - k. Ends with a barrier.
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_c2r_split(cx<T>* pl) {
    using P = PlaneShape<T, LY, LZ>;
    constexpr int M = P::M;
    const cx<T>* tws = pl + P::TWS;
    plane_split_tasks<LY, M, NT>([&](int r, int k) {
        cx<T>* row = pl + r * P::PITCH;
        if (k == 0) {
            const cx<T> a = row[0];
            row[0] = cx<T>{a.x + a.y, a.x - a.y};
        } else {
            const cx<T> a = row[k], b = row[M - k];
            const cx<T> e2{a.x + b.x, a.y - b.y};
            const cx<T> o2 = cmul(cx<T>{a.x - b.x, a.y + b.y}, conjv(tws[k]));
            row[k] = cx<T>{e2.x - o2.y, e2.y + o2.x};
            if (k != M - k) row[M - k] = cx<T>{e2.x + o2.y, o2.x - e2.y};
        }
    });
    __syncthreads();
}

// Forward real transform of every row, after plane_rows<INV = false> of x[2m] + i x[2m+1]:
// X[k] = E + W_LZ^k O, X[M-k] = conj(E - W_LZ^k O) with E = (Z[k] + conj Z[M-k]) / 2,
// O = (Z[k] - conj Z[M-k]) / 2i; slot 0 = X[0] + i X[M]. Ends with a barrier.
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_r2c_split(cx<T>* pl) {
    using P = PlaneShape<T, LY, LZ>;
    constexpr int M = P::M;
    const cx<T>* tws = pl + P::TWS;
    plane_split_tasks<LY, M, NT>([&](int r, int k) {
        cx<T>* row = pl + r * P::PITCH;
        if (k == 0) {
            const cx<T> a = row[0];
            row[0] = cx<T>{a.x + a.y, a.x - a.y};
        } else {
            const cx<T> a = row[k], b = row[M - k];
            const cx<T> e{T(0.5) * (a.x + b.x), T(0.5) * (a.y - b.y)};
            const cx<T> o{T(0.5) * (a.y + b.y), T(0.5) * (b.x - a.x)};
            const cx<T> wo = cmul(tws[k], o);
            row[k] = e + wo;
            if (k != M - k) row[M - k] = conjv(e - wo);
        }
    });
    __syncthreads();
}

// Two complex values: the unit of the verbatim plane copies (16 B in float32), moved with the
// nontemporal hint (gld / gst, kwh_fft.h).
template <typename T> struct alignas(4 * sizeof(T)) CPair { cx<T> a, b; };
template <typename T> __device__ __forceinline__ CPair<T> gldp(const CPair<T>* p) {
    typedef T v4 __attribute__((ext_vector_type(4)));
    const v4 v = gld(reinterpret_cast<const v4*>(p));
    return CPair<T>{{v.x, v.y}, {v.z, v.w}};
}
template <typename T> __device__ __forceinline__ void gstp(CPair<T>* p, const CPair<T>& v) {
    typedef T v4 __attribute__((ext_vector_type(4)));
    gst(reinterpret_cast<v4*>(p), v4{v.a.x, v.a.y, v.b.x, v.b.y});
}

// HBM half-spectrum plane [LY][M + 1] -> LDS (verbatim, the LDS rows have the HBM row length),
// then slot 0 of every row packs columns 0 and M as G0 + i GN. The caller places a barrier
// before the slots are read.
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_load_spec(cx<T>* pl, const cx<T>* __restrict__ src,
                                                const cx<T>* __restrict__ op = nullptr, bool op_rows = false) {
    using P = PlaneShape<T, LY, LZ>;
    constexpr int NP = LY * P::MZH / 2, IP = (NP + NT - 1) / NT;
    static_assert((LY * P::MZH) % 2 == 0, "plane copies move complex pairs");
    {
        // every load issued before the first LDS write (clamped index instead of a branch, so the
        // loads are not serialised behind per-iteration waits)
        const int tid = plane_tid();
        const CPair<T>* s = reinterpret_cast<const CPair<T>*>(src);
        CPair<T>* d = reinterpret_cast<CPair<T>*>(pl);
        CPair<T> v[IP];
#pragma unroll
        for (int it = 0; it < IP; ++it) v[it] = gldp(s + min(tid + it * NT, NP - 1));
#pragma unroll
        for (int it = 0; it < IP; ++it) {
            const int i = tid + it * NT;
            if (NP % NT == 0 || i < NP) d[i] = v[it];
        }
    }
    __syncthreads();
    if (op) {   // uniform: an operator per row (ky) or per column (kz) of the unpacked spectrum
        for (int i = threadIdx.x; i < LY * P::MZH; i += NT) {
            const int r = i / P::MZH, k = i - r * P::MZH;
            cx<T>& e = pl[r * P::PITCH + k];
            e = cmul(e, op[op_rows ? r : k]);
        }
        __syncthreads();
    }
    for (int r = threadIdx.x; r < LY; r += NT) {
        const cx<T> g0 = pl[r * P::PITCH], gn = pl[r * P::PITCH + P::M];
        pl[r * P::PITCH] = cx<T>{g0.x - gn.y, g0.y + gn.x};
    }
}

// Slot 0 of the rows -> columns 0 and M, G0 = (C[ky] + conj C[-ky]) / 2 and
// GN = (C[ky] - conj C[-ky]) / 2i (rows ky and LY - ky handled together), then the plane to HBM
// (verbatim). The caller places a barrier before the plane is written again.
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_store_spec(cx<T>* pl, cx<T>* __restrict__ dst) {
    using P = PlaneShape<T, LY, LZ>;
    constexpr int M = P::M, NP = LY * P::MZH / 2, IP = (NP + NT - 1) / NT;
    for (int r = threadIdx.x; r <= LY / 2; r += NT) {
        const int rm = r == 0 ? 0 : LY - r;
        const cx<T> c = pl[r * P::PITCH], m = pl[rm * P::PITCH];
        pl[r * P::PITCH] = cx<T>{T(0.5) * (c.x + m.x), T(0.5) * (c.y - m.y)};
        pl[r * P::PITCH + M] = cx<T>{T(0.5) * (c.y + m.y), T(0.5) * (m.x - c.x)};
        if (rm != r) {
            pl[rm * P::PITCH] = cx<T>{T(0.5) * (m.x + c.x), T(0.5) * (m.y - c.y)};
            pl[rm * P::PITCH + M] = cx<T>{T(0.5) * (m.y + c.y), T(0.5) * (c.x - m.x)};
        }
    }
    __syncthreads();
    const int tid = plane_tid();
    const CPair<T>* s = reinterpret_cast<const CPair<T>*>(pl);
    CPair<T>* d = reinterpret_cast<CPair<T>*>(dst);
    CPair<T> v[IP];
#pragma unroll
    for (int it = 0; it < IP; ++it) v[it] = s[min(tid + it * NT, NP - 1)];
#pragma unroll
    for (int it = 0; it < IP; ++it) {
        const int i = tid + it * NT;
        if (NP % NT == 0 || i < NP) gstp(d + i, v[it]);
    }
}

// Inverse transform of the spectrum plane into the real view (ends with a barrier).
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_inverse(cx<T>* pl, const cx<T>* __restrict__ src,
                                              const cx<T>* __restrict__ op = nullptr, bool op_rows = false) {
    plane_load_spec<T, LY, LZ, NT>(pl, src, op, op_rows);
    __syncthreads();
    plane_cols<T, LY, LZ, NT, true>(pl);
    plane_c2r_split<T, LY, LZ, NT>(pl);
    plane_rows<T, LY, LZ, NT, true>(pl);
}

// Forward transform of the real view, stored as a half-spectrum plane. The caller places a
// barrier between the last write of the real view and this call, and another before the plane
// is written again.
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_forward(cx<T>* pl, cx<T>* __restrict__ dst) {
    plane_rows<T, LY, LZ, NT, false>(pl);
    plane_r2c_split<T, LY, LZ, NT>(pl);
    plane_cols<T, LY, LZ, NT, false>(pl);
    plane_store_spec<T, LY, LZ, NT>(pl, dst);
}

// Two adjacent cells (2m, 2m+1) of a real row: the element of the real view's slot m.
template <typename T> struct alignas(2 * sizeof(T)) Pair { T a, b; };
template <typename T>
__device__ __forceinline__ Pair<T> ld2(const T* p) { return *reinterpret_cast<const Pair<T>*>(p); }
template <typename T>
__device__ __forceinline__ void st2(T* p, T a, T b) { *reinterpret_cast<Pair<T>*>(p) = Pair<T>{a, b}; }
template <typename T>
__device__ __forceinline__ Pair<T> ld2m(const SMap<T>& m, int64_t c) {
    return m.p ? ld2(m.p + c) : Pair<T>{m.s, m.s};
}

// Element loop over the real view: f(iteration, row r, column pair m, cell offset in plane);
// slot r * M + m of iteration `it` is threadIdx.x + it * NT. Iterations run in groups of G
// between scheduling fences, so the loads of later groups are not hoisted above earlier ones
// (bounded live registers; the other waves of the CU hide the latency).
template <typename T, int LY, int LZ, int NT, int G = 2, class F>
__device__ __forceinline__ void plane_elements(F&& f) {
    using P = PlaneShape<T, LY, LZ>;
    constexpr int NS = LY * P::M;
    static_assert(NS % NT == 0 && NT % P::M == 0, "workgroup size must divide the plane and be a multiple of a row");
    const int tid = plane_tid();
    const int m = tid % P::M, r0 = tid / P::M;
#pragma unroll
    for (int it = 0; it < NS / NT; ++it) {
        const int r = r0 + it * (NT / P::M);
        f(it, r, m, r * LZ + 2 * m);
        if ((it + 1) % G == 0) __builtin_amdgcn_sched_barrier(0);
    }
}

// Element loop in groups of G iterations whose global loads (load(it, r, m, c) -> V) are all issued
// before the group's updates and stores (body(it, r, m, c, V)): a store to one array may alias a later
// load of another as far as the compiler knows, so interleaved code waits out a full memory latency
// per iteration.
template <typename T, int LY, int LZ, int NT, int G, class LD, class F>
__device__ __forceinline__ void plane_elements_lf(LD&& load, F&& body) {
    using P = PlaneShape<T, LY, LZ>;
    constexpr int NS = LY * P::M, NI = NS / NT;
    static_assert(NS % NT == 0 && NT % P::M == 0 && NI % G == 0, "workgroup size must divide the plane");
    const int tid = plane_tid();
    const int m = tid % P::M, r0 = tid / P::M;
    using V = decltype(load(0, 0, 0, 0));
#pragma unroll
    for (int g0 = 0; g0 < NI; g0 += G) {
        V v[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int r = r0 + (g0 + k) * (NT / P::M);
            v[k] = load(g0 + k, r, m, r * LZ + 2 * m);
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int r = r0 + (g0 + k) * (NT / P::M);
            body(g0 + k, r, m, r * LZ + 2 * m, v[k]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}
template <int LY, int M, int NT> constexpr int plane_group() {
    constexpr int NI = LY * M / NT;
    return NI % 4 == 0 ? 4 : (NI % 2 == 0 ? 2 : 1);
}

// Sparse sources of rows [x LY, (x+1) LY) into the real view (dirichlet or additive).
template <typename T, int LY, int LZ, int NT>
__device__ void plane_apply_sparse(const SparseSrc<T>& s, int x, cx<T>* pl, int t) {
    using P = PlaneShape<T, LY, LZ>;
    if (t >= s.len) return;
    const int r0 = x * LY;
    const int e0 = s.rowptr[r0], e1 = s.rowptr[r0 + LY];
    T* f = reinterpret_cast<T*>(pl);
    for (int e = e0 + (int)threadIdx.x; e < e1; e += NT) {
        const int64_t cell = s.index[e];
        const int row = (int)(cell / LZ);
        const int n = (int)(cell - (int64_t)row * LZ);
        const T val = s.signal[(int64_t)(s.rows == 1 ? 0 : e) * s.len + t] * s.base[e] * s.factor;
        T& slot = f[(row - r0) * 2 * P::PITCH + n];
        slot = s.mode == 0 ? val : slot + val;
    }
}
__device__ __forceinline__ bool sparse_in_plane(const int* rowptr, int x, int ly) {
    return rowptr && rowptr[(x + 1) * ly] > rowptr[x * ly];
}

// ---------------------------------------------------------------------------------------------
// Kernels. One workgroup per x plane; grid.x = number of planes (block coordinate = plane).
// ---------------------------------------------------------------------------------------------

// spec = R2C_yz(scale * src), the plane-wise counterpart of the z pass + y pass.
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_fwd_real_body(const BIdx& bi, const T* __restrict__ src, cx<T>* spec,
                                                    T scale, const Geo& g, const AxisTabs<T>& tab) {
    using P = PlaneShape<T, LY, LZ>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* pl = reinterpret_cast<cx<T>*>(smem);
    plane_twiddles<T, LY, LZ, NT>(pl, tab);
    const int x = bi.x;
    const T* s = src + (int64_t)x * LY * LZ;
    plane_elements<T, LY, LZ, NT>([&](int, int r, int m, int c) {
        const Pair<T> v = ld2(s + c);
        pl[r * P::PITCH + m] = cx<T>{v.a * scale, v.b * scale};
    });
    __syncthreads();
    plane_forward<T, LY, LZ, NT>(pl, spec + (int64_t)x * LY * P::MZH);
    (void)g;
}
template <typename T, int LY, int LZ, int NT>
__global__ void __launch_bounds__(NT, (plane_waves<T, LY, LZ, NT>())) plane_fwd_real(const T* __restrict__ src, cx<T>* spec, T scale, Geo g,
                                                     AxisTabs<T> tab) {
    plane_fwd_real_body<T, LY, LZ, NT>(bidx_here(), src, spec, scale, g, tab);
}

// dst (= or +=) scale * C2R_yz(spec).
template <typename T, int LY, int LZ, int NT, bool ACC>
__device__ __forceinline__ void plane_inv_real_impl(const BIdx& bi, const cx<T>* spec, T* __restrict__ dst,
                                                    T scale, const AxisTabs<T>& tab) {
    using P = PlaneShape<T, LY, LZ>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* pl = reinterpret_cast<cx<T>*>(smem);
    plane_twiddles<T, LY, LZ, NT>(pl, tab);
    const int x = bi.x;
    plane_inverse<T, LY, LZ, NT>(pl, spec + (int64_t)x * LY * P::MZH);
    T* d = dst + (int64_t)x * LY * LZ;
    plane_elements<T, LY, LZ, NT>([&](int, int r, int m, int c) {
        const cx<T> v = pl[r * P::PITCH + m];
        if (ACC) {
            const Pair<T> o = ld2(d + c);
            st2(d + c, o.a + v.x * scale, o.b + v.y * scale);
        } else {
            st2(d + c, v.x * scale, v.y * scale);
        }
    });
}
template <typename T, int LY, int LZ, int NT>
__device__ __forceinline__ void plane_inv_real_body(const BIdx& bi, const cx<T>* spec, T* __restrict__ dst,
                                                    T scale, int accumulate, const Geo& g, const AxisTabs<T>& tab) {
    if (accumulate) plane_inv_real_impl<T, LY, LZ, NT, true>(bi, spec, dst, scale, tab);
    else plane_inv_real_impl<T, LY, LZ, NT, false>(bi, spec, dst, scale, tab);
    (void)g;
}
template <typename T, int LY, int LZ, int NT>
__global__ void __launch_bounds__(NT, (plane_waves<T, LY, LZ, NT>())) plane_inv_real(const cx<T>* spec, T* __restrict__ dst, T scale,
                                                     int accumulate, Geo g, AxisTabs<T> tab) {
    plane_inv_real_body<T, LY, LZ, NT>(bidx_here(), spec, dst, scale, accumulate, g, tab);
}

// Velocity update of plane x, per axis a (KSpaceFirstOrder.swift:669-678, sources :705-711):
// g_a = C2R_yz(G_a); u_a = pml_sg_a (pml_sg_a u_a - dt/rho_a g_a) [+ corrected source]
// [+ sparse sources]; G_a = R2C_yz(u_a) for the divergence's x pass.
// CORR: 0 no corrected source, 1 the corrected source as a real field (VelArgs::corr), 2 as the y-z
// spectrum of axis corr_axis (VelArgs::corr_spec), inverted here and held in registers.
template <typename T, int LY, int LZ, int NT, int CORR>
__device__ __forceinline__ void plane_velocity_body(const BIdx& bi, const VelArgs<T>& a, const Geo& g,
                                                    const AxisTabs<T>& tab) {
    using P = PlaneShape<T, LY, LZ>;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* pl = reinterpret_cast<cx<T>*>(smem);
    plane_twiddles<T, LY, LZ, NT>(pl, tab);
    const int x = bi.x;
    const int64_t rb = (int64_t)x * LY * LZ, sb = (int64_t)x * LY * P::MZH;
    // corrected source given as a spectrum: its plane inverted first, the real values kept in
    // registers in the element order of the update loop
    constexpr int NI = LY * P::M / NT;
    cx<T> crr[CORR == 2 ? NI : 1];
    if constexpr (CORR == 2) {
        plane_inverse<T, LY, LZ, NT>(pl, a.corr_spec + sb);
        plane_elements<T, LY, LZ, NT>([&](int it, int r, int m, int) { crr[it] = pl[r * P::PITCH + m]; });
        __syncthreads();
    }
    sfor<0, 3>([&](auto cax) {
        constexpr int ax = decltype(cax)::value;
        if (a.gdual && ax > 0)
            plane_inverse<T, LY, LZ, NT>(pl, a.Gin[ax] + sb, tab.ddxp[ax], ax == 1);
        else
            plane_inverse<T, LY, LZ, NT>(pl, (a.gdual ? a.Gin[ax] : a.G[ax]) + sb);
        T* u = a.u[ax] + rb;
        const SMap<T> dtr = a.dtr[ax], corr = a.corr[ax];
        const T* pm = tab.pmlsg[ax];
        const bool sp = sparse_in_plane(a.sp[ax].rowptr, x, LY);
        const T qx = pm[ax == 0 ? x : 0];
        struct VL { Pair<T> uu, dt, cr; };
        plane_elements_lf<T, LY, LZ, NT, plane_group<LY, P::M, NT>()>([&](int, int, int, int c) {
            VL l;
            l.uu = ld2(u + c);
            l.dt = ld2m(dtr, rb + c);
            if constexpr (CORR == 1) l.cr = ld2m(corr, rb + c);
            return l;
        }, [&](int it, int r, int m, int c, const VL& l) {
            const cx<T> gv = pl[r * P::PITCH + m];
            const Pair<T> uu = l.uu, dt = l.dt;
            T q0, q1;
            if constexpr (ax == 0) { q0 = qx; q1 = qx; }
            else if constexpr (ax == 1) { q0 = pm[r]; q1 = q0; }
            else { const Pair<T> qq = ld2(pm + 2 * m); q0 = qq.a; q1 = qq.b; }
            T v0 = q0 * (q0 * uu.a - dt.a * gv.x);
            T v1 = q1 * (q1 * uu.b - dt.b * gv.y);
            if constexpr (CORR == 1) {
                v0 += l.cr.a;
                v1 += l.cr.b;
            } else if constexpr (CORR == 2) {
                if (a.corr_axis == ax) {
                    v0 += crr[it].x;
                    v1 += crr[it].y;
                }
            }
            pl[r * P::PITCH + m] = cx<T>{v0, v1};
            if (!sp) st2(u + c, v0, v1);
        });
        __syncthreads();
        if (sp) {
            plane_apply_sparse<T, LY, LZ, NT>(a.sp[ax], x, pl, a.ctr[0]);
            __syncthreads();
            plane_elements<T, LY, LZ, NT>([&](int, int r, int m, int c) {
                const cx<T> v = pl[r * P::PITCH + m];
                st2(u + c, v.x, v.y);
            });
            __syncthreads();
        }
        plane_forward<T, LY, LZ, NT>(pl, a.G[ax] + sb);
        __syncthreads();
    });
    (void)g;
}
template <typename T, int LY, int LZ, int NT, int CORR>
__global__ void __launch_bounds__(NT, (plane_waves<T, LY, LZ, NT>())) plane_velocity(VelArgs<T> a, Geo g, AxisTabs<T> tab) {
    plane_velocity_body<T, LY, LZ, NT, CORR>(bidx_here(), a, g, tab);
}

// Density update of plane x (KSpaceFirstOrder.swift:680-702): per axis a, du_a = C2R_yz(D_a);
// rho_a = pml_a (pml_a rho_a - dt rho0 [B/A factor] du_a) [+ sources]. MODE 0 (lossless) / 2
// (Stokes): then the EOS p = c0^2 (sum_a rho_a [+ tau rho0 sum_a du_a] [+ B/A (sum rho)^2 /
// (2 rho0)]) and P = R2C_yz(p) for the next step's x pass. MODE 1 (FFT power-law absorption): the
// inputs of the absorption term instead, rhoT = sum_a rho_a, F1 = R2C_yz(rho0 sum_a du_a) and, with
// dispersion, F2 = R2C_yz(rhoT) (the EOS follows in plane_eos). The sums are held in registers
// (one slot pair per element of the thread's share of the plane).
template <typename T, int LY, int LZ, int NT, bool NL, int MODE, bool CORR>
__device__ __forceinline__ void plane_density_body(const BIdx& bi, const RhoArgs<T>& a, const Geo& g,
                                                   const AxisTabs<T>& tab) {
    using P = PlaneShape<T, LY, LZ>;
    static_assert(MODE >= 0 && MODE <= 2, "plane density kernel: lossless, FFT absorption or Stokes");
    constexpr int NV = LY * P::M / NT;
    // rho0 kept in registers across the axes when the occupancy floor leaves room (>= 128 registers)
    constexpr bool HOLD = !NL && MODE == 0 && plane_waves<T, LY, LZ, NT>() <= 4;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* pl = reinterpret_cast<cx<T>*>(smem);
    plane_twiddles<T, LY, LZ, NT>(pl, tab);
    const int x = bi.x;
    const int64_t rb = (int64_t)x * LY * LZ, sb = (int64_t)x * LY * P::MZH;
    cx<T> rt[NV];
    cx<T> sd[MODE != 0 ? NV : 1];
    cx<T> rp[NL ? NV : 1];
    cx<T> r0h[HOLD ? NV : 1];
    if constexpr (NL) {
        plane_elements<T, LY, LZ, NT>([&](int it, int, int, int c) {
            T s0 = 0, s1 = 0;
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const Pair<T> v = ld2(a.rho[ax] + rb + c);
                s0 += v.a;
                s1 += v.b;
            }
            rp[it] = cx<T>{s0, s1};
        });
    }
    if constexpr (HOLD) {
        plane_elements<T, LY, LZ, NT>([&](int it, int, int, int c) {
            const Pair<T> v = ld2m(a.rho0, rb + c);
            r0h[it] = cx<T>{v.a, v.b};
        });
    }
    sfor<0, 3>([&](auto cax) {
        constexpr int ax = decltype(cax)::value;
        plane_inverse<T, LY, LZ, NT>(pl, a.D[ax] + sb);
        T* rho = a.rho[ax] + rb;
        const T* pm = tab.pml[ax];
        const T cf = a.corr_factor[ax];
        const bool sp = sparse_in_plane(a.sp[ax].rowptr, x, LY);
        const T qx = pm[ax == 0 ? x : 0];
        struct DL { Pair<T> r0, ro, cr; };
        plane_elements_lf<T, LY, LZ, NT, plane_group<LY, P::M, NT>()>([&](int it, int, int, int c) {
            DL l;
            if constexpr (HOLD) l.r0 = Pair<T>{r0h[it].x, r0h[it].y};
            else l.r0 = ld2m(a.rho0, rb + c);
            l.ro = ld2(rho + c);
            if constexpr (CORR) l.cr = ld2m(a.corr, rb + c);
            return l;
        }, [&](int it, int r, int m, int c, const DL& l) {
            const cx<T> du = pl[r * P::PITCH + m];
            const Pair<T> r0 = l.r0, ro = l.ro;
            T q0, q1;
            if constexpr (ax == 0) { q0 = qx; q1 = qx; }
            else if constexpr (ax == 1) { q0 = pm[r]; q1 = q0; }
            else { const Pair<T> qq = ld2(pm + 2 * m); q0 = qq.a; q1 = qq.b; }
            const T d0 = r0.a * du.x, d1 = r0.b * du.y;
            if constexpr (MODE != 0) {
                if (ax == 0) sd[it] = cx<T>{d0, d1};
                else { sd[it].x += d0; sd[it].y += d1; }
            }
            T m0 = a.dt * d0, m1 = a.dt * d1;
            if constexpr (NL) {
                m0 *= (T(2) * rp[it].x + r0.a) / r0.a;
                m1 *= (T(2) * rp[it].y + r0.b) / r0.b;
            }
            T v0 = q0 * (q0 * ro.a - m0);
            T v1 = q1 * (q1 * ro.b - m1);
            if constexpr (CORR) {
                v0 += l.cr.a * cf;
                v1 += l.cr.b * cf;
            }
            if (sp) {
                pl[r * P::PITCH + m] = cx<T>{v0, v1};
            } else {
                st2(rho + c, v0, v1);
                if (ax == 0) rt[it] = cx<T>{v0, v1};
                else { rt[it].x += v0; rt[it].y += v1; }
            }
        });
        __syncthreads();
        if (sp) {
            plane_apply_sparse<T, LY, LZ, NT>(a.sp[ax], x, pl, a.ctr[0]);
            __syncthreads();
            plane_elements<T, LY, LZ, NT>([&](int it, int r, int m, int c) {
                const cx<T> v = pl[r * P::PITCH + m];
                st2(rho + c, v.x, v.y);
                if (ax == 0) rt[it] = v;
                else { rt[it].x += v.x; rt[it].y += v.y; }
            });
            __syncthreads();
        }
    });
    if constexpr (MODE == 1) {
        // rhoT, F1 = R2C_yz(rho0 sum du), F2 = R2C_yz(rhoT)
        T* rT = a.rhoT + rb;
        plane_elements<T, LY, LZ, NT>([&](int it, int r, int m, int c) {
            st2(rT + c, rt[it].x, rt[it].y);
            pl[r * P::PITCH + m] = sd[it];
        });
        __syncthreads();
        plane_forward<T, LY, LZ, NT>(pl, a.F1 + sb);
        if (a.dispersion) {
            __syncthreads();
            plane_elements<T, LY, LZ, NT>([&](int it, int r, int m, int) { pl[r * P::PITCH + m] = rt[it]; });
            __syncthreads();
            plane_forward<T, LY, LZ, NT>(pl, a.F2 + sb);
        }
        (void)g;
        return;
    }
    // EOS into the real view, then P = R2C_yz(p)
    T* p = a.p + rb;
    struct EL { Pair<T> r0, tu, bn, cc; };
    plane_elements_lf<T, LY, LZ, NT, plane_group<LY, P::M, NT>()>([&](int it, int, int, int c) {
        EL l;
        l.r0 = Pair<T>{T(1), T(1)};
        if constexpr (NL) {
            if constexpr (HOLD) l.r0 = Pair<T>{r0h[it].x, r0h[it].y};
            else l.r0 = ld2m(a.rho0, rb + c);
            l.bn = ld2m(a.bona, rb + c);
        }
        if constexpr (MODE == 2) l.tu = ld2m(a.tau, rb + c);
        l.cc = ld2m(a.c0, rb + c);
        return l;
    }, [&](int it, int r, int m, int c, const EL& l) {
        T e0 = rt[it].x, e1 = rt[it].y;
        const Pair<T> r0 = l.r0;
        if constexpr (MODE == 2) {
            e0 += l.tu.a * sd[it].x;
            e1 += l.tu.b * sd[it].y;
        }
        if constexpr (NL) {
            e0 += l.bn.a * (rt[it].x * rt[it].x) / (T(2) * r0.a);
            e1 += l.bn.b * (rt[it].y * rt[it].y) / (T(2) * r0.b);
        }
        const Pair<T> cc = l.cc;
        const T p0 = cc.a * cc.a * e0, p1 = cc.b * cc.b * e1;
        st2(p + c, p0, p1);
        pl[r * P::PITCH + m] = cx<T>{p0, p1};
    });
    __syncthreads();
    plane_forward<T, LY, LZ, NT>(pl, a.outP + sb);
    (void)g;
}
template <typename T, int LY, int LZ, int NT, bool NL, int MODE, bool CORR>
__global__ void __launch_bounds__(NT, (plane_waves<T, LY, LZ, NT>())) plane_density(RhoArgs<T> a, Geo g, AxisTabs<T> tab) {
    plane_density_body<T, LY, LZ, NT, NL, MODE, CORR>(bidx_here(), a, g, tab);
}

// EOS with FFT power-law absorption on plane x (KSpaceFirstOrder.swift:693-702): A1, A2 are the
// x-pass outputs |k|^(y-2) F1 and |k|^(y-1) F2; e = rhoT + tau C2R_yz(A1) [- eta C2R_yz(A2)]
// [+ B/A rhoT^2 / (2 rho0)], p = c0^2 e, P = R2C_yz(p). tau, eta from the cell's sound speed
// (tau_eta_from_c) or from their maps. The partial sum is held in registers while the second
// spectrum is transformed. Same arithmetic, in the same order, as reg_z_eos.
template <typename T, int LY, int LZ, int NT, bool NL>
__device__ __forceinline__ void plane_eos_body(const BIdx& bi, const EosArgs<T>& a, const Geo& g,
                                               const AxisTabs<T>& tab) {
    using P = PlaneShape<T, LY, LZ>;
    constexpr int NV = LY * P::M / NT;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* pl = reinterpret_cast<cx<T>*>(smem);
    plane_twiddles<T, LY, LZ, NT>(pl, tab);
    const int x = bi.x;
    const int64_t rb = (int64_t)x * LY * LZ, sb = (int64_t)x * LY * P::MZH;
    const T* rT = a.rhoT + rb;
    cx<T> e[NV];
    plane_inverse<T, LY, LZ, NT>(pl, a.A1 + sb);
    {
        struct L1 { Pair<T> rt, c, tau; };
        plane_elements_lf<T, LY, LZ, NT, plane_group<LY, P::M, NT>()>([&](int, int, int, int c) {
            L1 l;
            l.rt = ld2(rT + c);
            l.c = ld2m(a.c0, rb + c);
            if (!a.tau_eta_from_c) l.tau = ld2m(a.tau, rb + c);
            return l;
        }, [&](int it, int r, int m, int, const L1& l) {
            const cx<T> a1 = pl[r * P::PITCH + m];
            T tA, tB;
            if (a.tau_eta_from_c) {
                tA = a.tk * exp2(a.te * log2(l.c.a));
                tB = a.tk * exp2(a.te * log2(l.c.b));
            } else {
                tA = l.tau.a;
                tB = l.tau.b;
            }
            e[it] = cx<T>{l.rt.a + tA * a1.x, l.rt.b + tB * a1.y};
        });
    }
    __syncthreads();
    if (a.A2) {
        plane_inverse<T, LY, LZ, NT>(pl, a.A2 + sb);
        struct L2 { Pair<T> c, eta; };
        plane_elements_lf<T, LY, LZ, NT, plane_group<LY, P::M, NT>()>([&](int, int, int, int c) {
            L2 l;
            if (a.tau_eta_from_c) l.c = ld2m(a.c0, rb + c);
            else l.eta = ld2m(a.eta, rb + c);
            return l;
        }, [&](int it, int r, int m, int, const L2& l) {
            const cx<T> a2 = pl[r * P::PITCH + m];
            T hA, hB;
            if (a.tau_eta_from_c) {
                hA = a.ek * exp2(a.ee * log2(l.c.a));
                hB = a.ek * exp2(a.ee * log2(l.c.b));
            } else {
                hA = l.eta.a;
                hB = l.eta.b;
            }
            e[it].x -= hA * a2.x;
            e[it].y -= hB * a2.y;
        });
        __syncthreads();
    }
    T* p = a.p + rb;
    struct L3 { Pair<T> c, rt, bn, r0; };
    plane_elements_lf<T, LY, LZ, NT, plane_group<LY, P::M, NT>()>([&](int, int, int, int c) {
        L3 l;
        l.c = ld2m(a.c0, rb + c);
        if constexpr (NL) {
            l.rt = ld2(rT + c);
            l.bn = ld2m(a.bona, rb + c);
            l.r0 = ld2m(a.rho0, rb + c);
        }
        return l;
    }, [&](int it, int r, int m, int c, const L3& l) {
        T ea = e[it].x, eb = e[it].y;
        if constexpr (NL) {
            ea += l.bn.a * (l.rt.a * l.rt.a) / (T(2) * l.r0.a);
            eb += l.bn.b * (l.rt.b * l.rt.b) / (T(2) * l.r0.b);
        }
        const T pA = l.c.a * l.c.a * ea, pB = l.c.b * l.c.b * eb;
        st2(p + c, pA, pB);
        pl[r * P::PITCH + m] = cx<T>{pA, pB};
    });
    __syncthreads();
    plane_forward<T, LY, LZ, NT>(pl, a.outP + sb);
    (void)g;
}
template <typename T, int LY, int LZ, int NT, bool NL>
__global__ void __launch_bounds__(NT, (plane_waves<T, LY, LZ, NT>())) plane_eos(EosArgs<T> a, Geo g, AxisTabs<T> tab) {
    plane_eos_body<T, LY, LZ, NT, NL>(bidx_here(), a, g, tab);
}

}  // namespace kwh
