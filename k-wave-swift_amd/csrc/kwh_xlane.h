// kwh_xlane.h -- length L = J * E transforms whose inner radix J runs across lanes (J = 4, 8, 16).
//
// The J lanes of one transform are adjacent lanes of one DPP row (lane j = lane id % J) and
// exchange values with DPP moves; no LDS round trip and no workgroup barrier inside a transform.
// With n = j + J*m and k = k2 + E*k1:
//   X[k2 + E*k1] = sum_j W_J^(j*k1) W_L^(j*k2) sum_m x[j + J*m] W_E^(m*k2).
// Two data layouts across the lanes of a transform:
//   cyclic: lane j holds elements j + J*m (m = 0..E-1), the layout of the loads and stores;
//   block:  lane j holds elements m + E*brev(j), a contiguous run of E.
// to_block (cyclic -> block): DFT_E in registers, twiddle W_L^(j*m), radix-2 decimation in
// frequency over the lanes (natural lane order in, bit-reversed out). to_cyclic (block ->
// cyclic): the mirror image, radix-2 decimation in time over the lanes, twiddle, DFT_E. A
// forward and an inverse therefore chain through an elementwise step in the block layout.
#pragma once
#include "kwh_fft.h"

namespace kwh {

// DPP controls: quad_perm [1,0,3,2] (lane ^ 1), quad_perm [2,3,0,1] (lane ^ 2), quad_perm
// [3,2,1,0], row_half_mirror (lane ^ 7 within eight) and row_ror:8 (lane ^ 8 within sixteen).
// lane ^ 4 = quad reverse of the half-row mirror.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppQuadRev = 0x1B, kDppHalfMirror = 0x141, kDppRor8 = 0x128;

template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_v(float x) {
    return __int_as_float(dpp_i<CTRL>(__float_as_int(x)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_v(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)dpp_i<CTRL>((int)(unsigned)b);
    const unsigned hi = (unsigned)dpp_i<CTRL>((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Value of lane (lane ^ H) within the DPP row.
template <int H, typename V>
__device__ __forceinline__ V lane_xor(V x) {
    static_assert(H == 1 || H == 2 || H == 4 || H == 8, "lane distance");
    if constexpr (H == 1) return dpp_v<kDppXor1>(x);
    else if constexpr (H == 2) return dpp_v<kDppXor2>(x);
    else if constexpr (H == 4) return dpp_v<kDppQuadRev>(dpp_v<kDppHalfMirror>(x));
    else return dpp_v<kDppRor8>(x);
}
template <int H, typename T>
__device__ __forceinline__ cx<T> lane_xor(cx<T> a) {
    return cx<T>{lane_xor<H>(a.x), lane_xor<H>(a.y)};
}

__host__ __device__ constexpr int brev_bits(int x, int bits) {
    int r = 0;
    for (int b = 0; b < bits; ++b) r |= ((x >> b) & 1) << (bits - 1 - b);
    return r;
}
__host__ __device__ constexpr int ilog2_c(int x) { return x <= 1 ? 0 : 1 + ilog2_c(x / 2); }

template <typename T, int J, int E>
struct XLane {
    static_assert(J == 4 || J == 8 || J == 16, "lanes per transform");
    static constexpr int L = J * E;
    static constexpr int LB = ilog2_c(J);
    // Per-lane constants: W_L^(j*m) table row (LDS, [n1][k2] layout), radix-2 stage twiddles
    // W_(2h)^(j mod h) on the lower half of each lane pair (1 on the upper half), stage signs.
    const cx<T>* twrow;
    cx<T> ws[LB];    // stage h = 2^s
    T sg[LB];        // +1 on the upper lane of a pair (j & h == 0), -1 on the lower
    int j;

    // tab_lds: L slots; tw: W_L^i, i < L (forward sign). Caller barriers before first use.
    __device__ __forceinline__ void init(cx<T>* tab_lds, const cx<T>* __restrict__ tw, int lane) {
        for (int i = threadIdx.x; i < L; i += blockDim.x) tab_lds[i] = tw[(i % E) * (i / E)];
        j = lane;
        twrow = tab_lds + j * E;
#pragma unroll
        for (int s = 0; s < LB; ++s) {
            const int h = 1 << s;
            const bool lower = (j & h) != 0;
            sg[s] = lower ? T(-1) : T(1);
            ws[s] = lower ? tw[(j & (h - 1)) * (L / (2 * h))] : cx<T>{T(1), T(0)};
        }
    }
    __device__ __forceinline__ int block_start() const { return E * brev_bits(j, LB); }

    template <bool INV>
    __device__ __forceinline__ void twiddle(cx<T>* v) const {
#pragma unroll
        for (int m = 1; m < E; ++m) {
            const cx<T> w = twrow[m];
            v[m] = cmul(v[m], INV ? conjv(w) : w);
        }
    }

    // cyclic -> block
    template <bool INV>
    __device__ __forceinline__ void to_block(cx<T>* v) const {
        dftN<T, E, INV>(v);
        twiddle<INV>(v);
        sfor<0, LB>([&](auto cs) {
            constexpr int s = LB - 1 - decltype(cs)::value;   // h = J/2 .. 1
            constexpr int h = 1 << s;
            const T g = sg[s];
            const cx<T> w = INV ? conjv(ws[s]) : ws[s];
#pragma unroll
            for (int m = 0; m < E; ++m) {
                const cx<T> p = lane_xor<h>(v[m]);
                cx<T> r{g * v[m].x + p.x, g * v[m].y + p.y};
                if constexpr (h > 1) r = cmul(r, w);
                v[m] = r;
            }
        });
    }

    // block -> cyclic
    template <bool INV>
    __device__ __forceinline__ void to_cyclic(cx<T>* v) const {
        sfor<0, LB>([&](auto cs) {
            constexpr int s = decltype(cs)::value;   // h = 1 .. J/2
            constexpr int h = 1 << s;
            const T g = sg[s];
            const cx<T> w = INV ? conjv(ws[s]) : ws[s];
#pragma unroll
            for (int m = 0; m < E; ++m) {
                cx<T> u = v[m];
                if constexpr (h > 1) u = cmul(u, w);
                const cx<T> p = lane_xor<h>(u);
                v[m] = cx<T>{g * u.x + p.x, g * u.y + p.y};
            }
        });
        twiddle<INV>(v);
        dftN<T, E, INV>(v);
    }
};

}  // namespace kwh
