// kwh_fft.h -- device-side mixed-radix Stockham FFT for batches of 1-D transforms.
//
// Replaces the MLXFFT complex transforms the reference calls inside its time loop
// (Sources/KWave/Solver/KSpaceFirstOrder.swift:270-296, 432-461, 647-686). Design:
//  * A workgroup owns `ntr` transforms (ntr a power of two) of length L. In LDS an element n of
//    transform w lives at buf[n * pitch + w], so lanes with consecutive w touch consecutive
//    addresses (conflict-free) and consecutive columns of a strided global tile.
//  * Each stage of radix R runs L/R butterflies per transform; a butterfly reads R elements,
//    applies the Stockham twiddles, does a register DFT_R and writes R elements. Stages ping-pong
//    between two buffers; the first stage may read global memory directly and the last stage may
//    write global memory directly (accessor functors), which removes staging copies for the
//    strided passes.
//  * Radices 2, 3, 4, 5, 8, 16 have unrolled register DFTs; any other prime factor runs a
//    direct-DFT stage (one output per lane iteration), so every length is accepted.
//  * Twiddles come from a per-length table tw[m] = exp(-2*pi*i*m/L) computed in float64 on the
//    host and rounded once to the working type.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace kwh {

template <typename T> struct cx { T x, y; };

template <typename T> __device__ __forceinline__ cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T> __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ cx<T> conjv(cx<T> a) { return {a.x, -a.y}; }
// multiply by +i / -i
template <typename T> __device__ __forceinline__ cx<T> mul_pi(cx<T> a) { return {-a.y, a.x}; }
template <typename T> __device__ __forceinline__ cx<T> mul_mi(cx<T> a) { return {a.y, -a.x}; }

constexpr int kMaxStages = 24;

// One transform length's plan. radix[] multiplies to L; tw points at the length-L table.
struct FftPlan {
    int L;
    int nstages;
    int radix[kMaxStages];
};

// Loads and stores with the nontemporal hint, for whole-line streaming copies (each lane moves 16
// contiguous bytes, every line read or written once per pass). Measured on a 1 GiB device copy:
// 5.46-5.77 TB/s plain, 5.93-6.31 TB/s with nontemporal loads and stores. Not for the register
// kernels' 8-byte lane accesses, which the hint made 1.7x slower at 256^3 (their lines are completed
// by other waves).
#ifndef KWH_STREAM_HINTS
#define KWH_STREAM_HINTS 1
#endif
template <typename U> __device__ __forceinline__ U gld(const U* p) {
    if constexpr (KWH_STREAM_HINTS) return __builtin_nontemporal_load(p);
    else return *p;
}
template <typename U> __device__ __forceinline__ void gst(U* p, U v) {
    if constexpr (KWH_STREAM_HINTS) __builtin_nontemporal_store(v, p);
    else *p = v;
}
template <typename T> __device__ __forceinline__ cx<T> gld(const cx<T>* p) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    if constexpr (KWH_STREAM_HINTS) {
        const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
        return cx<T>{v.x, v.y};
    } else {
        return *p;
    }
}
template <typename T> __device__ __forceinline__ void gst(cx<T>* p, cx<T> v) {
    typedef T v2 __attribute__((ext_vector_type(2)));
    if constexpr (KWH_STREAM_HINTS) {
        const v2 w = {v.x, v.y};
        __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
    } else {
        *p = v;
    }
}

// cos / sin of 2*pi*m/16, m = 0..15 (exact decimal expansions of the standard constants).
#define KWH_C16_1 0.92387953251128675613
#define KWH_S16_1 0.38268343236508977173
#define KWH_R2_2 0.70710678118654752440

template <typename T, bool INV>
__device__ __forceinline__ cx<T> w16(int m) {
    // exp(-2*pi*i*m/16) (forward) or its conjugate (inverse), for compile-time m after unrolling
    constexpr double C[16] = {1.0, KWH_C16_1, KWH_R2_2, KWH_S16_1, 0.0, -KWH_S16_1, -KWH_R2_2, -KWH_C16_1,
                              -1.0, -KWH_C16_1, -KWH_R2_2, -KWH_S16_1, 0.0, KWH_S16_1, KWH_R2_2, KWH_C16_1};
    constexpr double S[16] = {0.0, KWH_S16_1, KWH_R2_2, KWH_C16_1, 1.0, KWH_C16_1, KWH_R2_2, KWH_S16_1,
                              0.0, -KWH_S16_1, -KWH_R2_2, -KWH_C16_1, -1.0, -KWH_C16_1, -KWH_R2_2, -KWH_S16_1};
    const int mm = m & 15;
    return cx<T>{T(C[mm]), INV ? T(S[mm]) : T(-S[mm])};
}

template <typename T, bool INV>
__device__ __forceinline__ void dft2(cx<T>* v) {
    cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cx<T>& v0, cx<T>& v1, cx<T>& v2, cx<T>& v3) {
    cx<T> a = v0 + v2, b = v0 - v2, c = v1 + v3, d = v1 - v3;
    cx<T> jd = INV ? mul_pi(d) : mul_mi(d);
    v0 = a + c;
    v2 = a - c;
    v1 = b + jd;
    v3 = b - jd;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(cx<T>* v) {
    const T s = T(0.86602540378443864676);  // sqrt(3)/2
    cx<T> t1 = v[1] + v[2];
    cx<T> t2 = v[0] - t1 * T(0.5);
    cx<T> t3 = (v[1] - v[2]) * s;
    cx<T> jt = INV ? mul_pi(t3) : mul_mi(t3);
    v[0] = v[0] + t1;
    v[1] = t2 + jt;
    v[2] = t2 - jt;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft5(cx<T>* v) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
    cx<T> a1 = v[1] + v[4], a2 = v[2] + v[3], b1 = v[1] - v[4], b2 = v[2] - v[3];
    cx<T> x0 = v[0];
    cx<T> r1 = x0 + a1 * c1 + a2 * c2;
    cx<T> r2 = x0 + a1 * c2 + a2 * c1;
    cx<T> i1 = b1 * s1 + b2 * s2;
    cx<T> i2 = b1 * s2 - b2 * s1;
    cx<T> j1 = INV ? mul_pi(i1) : mul_mi(i1);
    cx<T> j2 = INV ? mul_pi(i2) : mul_mi(i2);
    v[0] = x0 + a1 + a2;
    v[1] = r1 + j1;
    v[4] = r1 - j1;
    v[2] = r2 + j2;
    v[3] = r2 - j2;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cx<T>* v) {
    // n = n1 + 2*n2: DFT4 over n2 for n1 = 0,1; twiddle W8^(n1*k2); DFT2 over n1 -> X[k2 + 4*k1]
    cx<T> a0 = v[0], a1 = v[2], a2 = v[4], a3 = v[6];
    cx<T> b0 = v[1], b1 = v[3], b2 = v[5], b3 = v[7];
    dft4<T, INV>(a0, a1, a2, a3);
    dft4<T, INV>(b0, b1, b2, b3);
    b1 = cmul(b1, w16<T, INV>(2));
    b2 = cmul(b2, w16<T, INV>(4));
    b3 = cmul(b3, w16<T, INV>(6));
    v[0] = a0 + b0; v[4] = a0 - b0;
    v[1] = a1 + b1; v[5] = a1 - b1;
    v[2] = a2 + b2; v[6] = a2 - b2;
    v[3] = a3 + b3; v[7] = a3 - b3;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft16(cx<T>* v) {
    // n = n1 + 4*n2: DFT4 over n2 per n1; twiddle W16^(n1*k2); DFT4 over n1 -> X[k2 + 4*k1]
    cx<T> y[16];  // y[n1*4 + k2]
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        cx<T> a = v[n1], b = v[n1 + 4], c = v[n1 + 8], d = v[n1 + 12];
        dft4<T, INV>(a, b, c, d);
        y[n1 * 4 + 0] = a;
        y[n1 * 4 + 1] = b;
        y[n1 * 4 + 2] = c;
        y[n1 * 4 + 3] = d;
    }
#pragma unroll
    for (int n1 = 1; n1 < 4; ++n1)
#pragma unroll
        for (int k2 = 1; k2 < 4; ++k2) y[n1 * 4 + k2] = cmul(y[n1 * 4 + k2], w16<T, INV>(n1 * k2));
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
        cx<T> a = y[k2], b = y[4 + k2], c = y[8 + k2], d = y[12 + k2];
        dft4<T, INV>(a, b, c, d);
        v[k2] = a;
        v[k2 + 4] = b;
        v[k2 + 8] = c;
        v[k2 + 12] = d;
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(cx<T>* v) {
    if constexpr (R == 2) dft2<T, INV>(v);
    else if constexpr (R == 3) dft3<T, INV>(v);
    else if constexpr (R == 4) dft4<T, INV>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 5) dft5<T, INV>(v);
    else if constexpr (R == 8) dft8<T, INV>(v);
    else if constexpr (R == 16) dft16<T, INV>(v);
}

// LDS accessor: element n of transform w at buf[n * pitch + w].
template <typename T> struct LdsAcc {
    cx<T>* buf;
    int pitch;
    __device__ __forceinline__ cx<T> ld(int w, int n) const { return buf[n * pitch + w]; }
    __device__ __forceinline__ void st(int w, int n, cx<T> v) const { buf[n * pitch + w] = v; }
};

// Strided global accessor over a tile of columns: column w starts at base + colbase(w).
template <typename T> struct GlobAcc {
    cx<T>* data;
    const int64_t* colbase;  // per-w offsets (in LDS), -1 = column absent
    int64_t stride;
    __device__ __forceinline__ cx<T> ld(int w, int n) const {
        int64_t b = colbase[w];
        return b >= 0 ? data[b + (int64_t)n * stride] : cx<T>{T(0), T(0)};
    }
    __device__ __forceinline__ void st(int w, int n, cx<T> v) const {
        int64_t b = colbase[w];
        if (b >= 0) data[b + (int64_t)n * stride] = v;
    }
};

// One Stockham stage of fixed radix R (natural order in, natural order out).
template <typename T, int R, bool INV, class Src, class Dst>
__device__ __forceinline__ void stage_fixed(const Src& src, const Dst& dst, int L, int Ns, int lg_ntr,
                                            const cx<T>* __restrict__ tw, int tid, int nthr) {
    const int nbf = L / R;
    const int total = nbf << lg_ntr;
    const int ntr_m1 = (1 << lg_ntr) - 1;
    const int step = L / (Ns * R);
    for (int i = tid; i < total; i += nthr) {
        const int w = i & ntr_m1;
        const int j = i >> lg_ntr;
        const int jm = j % Ns;
        cx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src.ld(w, j + r * nbf);
        if (Ns > 1) {
            const int m = jm * step;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cx<T> t = tw[m * r];
                if (INV) t.y = -t.y;
                v[r] = cmul(v[r], t);
            }
        }
        dft<T, R, INV>(v);
        const int idxD = (j - jm) * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) dst.st(w, idxD + r * Ns, v[r]);
    }
}

// Direct-DFT stage for any radix (used for prime factors other than 2, 3, 5).
template <typename T, bool INV, class Src, class Dst>
__device__ __forceinline__ void stage_generic(const Src& src, const Dst& dst, int L, int R, int Ns, int lg_ntr,
                                              const cx<T>* __restrict__ tw, int tid, int nthr) {
    const int nbf = L / R;
    const int total = (nbf * R) << lg_ntr;
    const int ntr_m1 = (1 << lg_ntr) - 1;
    const int step = L / (Ns * R);
    for (int i = tid; i < total; i += nthr) {
        const int w = i & ntr_m1;
        const int q = i >> lg_ntr;
        const int j = q % nbf;
        const int r = q / nbf;
        const int jm = j % Ns;
        cx<T> acc{T(0), T(0)};
        for (int m = 0; m < R; ++m) {
            const int e = (jm * step * m + m * r * nbf) % L;
            cx<T> t = tw[e];
            if (INV) t.y = -t.y;
            acc = acc + cmul(src.ld(w, j + m * nbf), t);
        }
        dst.st(w, (j - jm) * R + jm + r * Ns, acc);
    }
}

template <typename T, bool INV, class Src, class Dst>
__device__ __forceinline__ void run_stage(int R, const Src& src, const Dst& dst, int L, int Ns, int lg_ntr,
                                          const cx<T>* tw, int tid, int nthr) {
    switch (R) {
        case 16: stage_fixed<T, 16, INV>(src, dst, L, Ns, lg_ntr, tw, tid, nthr); break;
        case 8: stage_fixed<T, 8, INV>(src, dst, L, Ns, lg_ntr, tw, tid, nthr); break;
        case 4: stage_fixed<T, 4, INV>(src, dst, L, Ns, lg_ntr, tw, tid, nthr); break;
        case 2: stage_fixed<T, 2, INV>(src, dst, L, Ns, lg_ntr, tw, tid, nthr); break;
        case 3: stage_fixed<T, 3, INV>(src, dst, L, Ns, lg_ntr, tw, tid, nthr); break;
        case 5: stage_fixed<T, 5, INV>(src, dst, L, Ns, lg_ntr, tw, tid, nthr); break;
        default: stage_generic<T, INV>(src, dst, L, R, Ns, lg_ntr, tw, tid, nthr); break;
    }
}

// Run all stages. Input is read through `first` (stage 0), the output is written through `last`
// (final stage). Intermediate stages ping-pong between LDS buffers b0 / b1 (pitch = ntr).
// Requires plan.nstages >= 1. All lanes of the workgroup must call this.
template <typename T, bool INV, class Src, class Dst>
__device__ void fft_run(const FftPlan& pl, const Src& first, const Dst& last, cx<T>* b0, cx<T>* b1, int pitch,
                        int lg_ntr, const cx<T>* tw, int tid, int nthr) {
    const int L = pl.L;
    int Ns = 1;
    LdsAcc<T> A{b0, pitch}, B{b1, pitch};
    const int ns = pl.nstages;
    for (int s = 0; s < ns; ++s) {
        const int R = pl.radix[s];
        const bool isFirst = s == 0, isLast = s == ns - 1;
        const LdsAcc<T>& lsrc = (s & 1) ? B : A;  // stage s>0 reads the buffer stage s-1 wrote
        const LdsAcc<T>& ldst = (s & 1) ? A : B;
        if (isFirst && isLast) run_stage<T, INV>(R, first, last, L, Ns, lg_ntr, tw, tid, nthr);
        else if (isFirst) run_stage<T, INV>(R, first, B, L, Ns, lg_ntr, tw, tid, nthr);
        else if (isLast) run_stage<T, INV>(R, (s & 1) ? B : A, last, L, Ns, lg_ntr, tw, tid, nthr);
        else run_stage<T, INV>(R, lsrc, ldst, L, Ns, lg_ntr, tw, tid, nthr);
        (void)lsrc;
        (void)ldst;
        if (!isLast) __syncthreads();
        Ns *= R;
    }
}

// ---------------------------------------------------------------------------------------------
// Register DFTs of composite length with compile-time twiddles (register-resident kernels)
// ---------------------------------------------------------------------------------------------

// cos and sin of 2*pi*e/N, evaluated at compile time: reduce to the nearest quadrant,
// 2*pi*e/N = q*pi/2 + psi with |psi| <= pi/4, Taylor series of sin / cos of psi to the 25th power
// (truncation error below 1e-25), then rotate by q quarter turns exactly. Multiples of pi/2 give
// exact 0 and +-1.
struct CosSin { double c, s; };
constexpr CosSin ct_cis(long e, long n) {
    long m = e % n;
    if (m < 0) m += n;
    const long q = (8 * m + n) / (2 * n);              // nearest quadrant, 0..4
    const double psi = 1.57079632679489661923132169163975144 * (double)(4 * m - q * n) / (double)n;
    const double x2 = psi * psi;
    double ts = psi, ss = psi, tc = 1.0, sc = 1.0;
    for (int k = 1; k <= 12; ++k) {
        ts *= -x2 / (double)((2 * k) * (2 * k + 1));
        ss += ts;
        tc *= -x2 / (double)((2 * k - 1) * (2 * k));
        sc += tc;
    }
    switch (q & 3) {
        case 0: return CosSin{sc, ss};
        case 1: return CosSin{-ss, sc};
        case 2: return CosSin{-sc, -ss};
        default: return CosSin{ss, -sc};
    }
}

// W_N^EXP = exp(-2 pi i EXP / N) (forward) or its conjugate (inverse), a compile-time constant.
template <typename T, int N, int EXP, bool INV>
__device__ __forceinline__ cx<T> twk() {
    constexpr CosSin v = ct_cis(EXP, N);
    return cx<T>{T(v.c), INV ? T(v.s) : T(-v.s)};
}

// Compile-time loop: f(integral_constant<int, i>) for i = B .. E-1.
template <int B, int E, class F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

// Factor of a composite length N = N1 * N2 (N2 = 3 or 5 when they divide N, else 4 or 8).
template <int N> struct CtSplit {
    static constexpr int N2 = N % 3 == 0 ? 3 : (N % 5 == 0 ? 5 : (N >= 64 ? 8 : 4));
    static constexpr int N1 = N / N2;
};

template <typename T, int N, bool INV> __device__ __forceinline__ void dftN(cx<T>* v);

// Cooley-Tukey in registers, natural order in and out: with n = n1 + N1 n2 and k = k2 + N2 k1,
// X[k] = sum_n1 W_N^(n1 k2) W_N1^(n1 k1) sum_n2 x[n] W_N2^(n2 k2).
template <typename T, int N1, int N2, bool INV>
__device__ __forceinline__ void dft_ct(cx<T>* v) {
    constexpr int N = N1 * N2;
    cx<T> y[N];
    sfor<0, N1>([&](auto c1) {
        constexpr int n1 = decltype(c1)::value;
        cx<T> t[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) t[n2] = v[n1 + N1 * n2];
        dftN<T, N2, INV>(t);
        sfor<0, N2>([&](auto c2) {
            constexpr int k2 = decltype(c2)::value;
            if constexpr ((n1 * k2) % N == 0) y[n1 * N2 + k2] = t[k2];
            else y[n1 * N2 + k2] = cmul(t[k2], twk<T, N, n1 * k2, INV>());
        });
    });
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
        cx<T> s[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) s[n1] = y[n1 * N2 + k2];
        dftN<T, N1, INV>(s);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) v[k2 + N2 * k1] = s[k1];
    }
}

// DFT of any length whose prime factors are 2, 3 and 5, in registers.
template <typename T, int N, bool INV>
__device__ __forceinline__ void dftN(cx<T>* v) {
    if constexpr (N == 1) {
    } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5 || N == 8 || N == 16) {
        dft<T, N, INV>(v);
    } else {
        dft_ct<T, CtSplit<N>::N1, CtSplit<N>::N2, INV>(v);
    }
}

}  // namespace kwh
