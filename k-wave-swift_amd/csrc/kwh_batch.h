// kwh_batch.h -- one launch per pass for a batch of independent simulations.
//
// Every kernel of the time step is a __device__ body taking its block coordinates (BIdx) plus a
// __global__ wrapper that reads them from the hardware. A batched launch runs the same body for
// many simulations at once: the host records each simulation's launch (its arguments and grid)
// into a BatchPack, concatenates the packs of all simulations at one position of the step, and
// launches batched_kernel once over the sum of their grids. Each workgroup looks up which
// simulation owns it (one table read), takes that simulation's arguments from its pack and runs
// the body with the simulation-local block coordinates. Simulations of uneven sizes therefore
// share every launch of a step, and the whole step of the batch is a short fixed list of
// launches that a hipGraph replays.
//
// The packs are read through the constant address space: they are written once by the host and
// never change while a kernel runs, and loading pointers from constant memory lets the compiler
// keep the bodies' loads and stores in the global (not flat) address space.
#pragma once
#include <tuple>

#include "kwh_kernels.h"

namespace kwh {

// A kernel body F with its single-launch wrapper K. W: minimum waves per SIMD of the batched
// variant (the wrapper's own launch bound). FLEX: the body accepts any workgroup size (loops
// over blockDim), so a batched launch may run it with the largest size of its simulations.
// NTH: the largest workgroup size the body is launched with (the batched variant's bound).
template <auto F, auto K, int W = 1, bool FLEX = true, int NTH = 256>
struct Body {
    static constexpr int kWaves = W;
    static constexpr int kThreads = NTH;
    static constexpr bool kFlex = FLEX;
    static constexpr auto kernel = K;
    template <class... A>
    __device__ static void run(const BIdx& b, const A&... a) { F(b, a...); }
};

// One simulation's launch: first global block, grid x extent, the launch arguments.
template <class... A>
struct BatchPack {
    int start;
    int gx;
    int pad0, pad1;
    std::tuple<A...> args;
};

template <class B, class... A>
__global__ void __launch_bounds__(B::kThreads, B::kWaves) batched_kernel(const BatchPack<A...>* packs, const int* owner) {
    using CPack = const __attribute__((address_space(4))) BatchPack<A...>;
    using CInt = const __attribute__((address_space(4))) int;
    const int gb = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int s = ((CInt*)owner)[gb];
    const BatchPack<A...>& pk = *(const BatchPack<A...>*)&((CPack*)packs)[s];
    const int lb = gb - pk.start;
    const BIdx bi{lb % pk.gx, pk.gx, lb / pk.gx, 0};
    std::apply([&](const A&... a) { B::run(bi, a...); }, pk.args);
}

// Bodies of the step kernels.
template <typename T, bool INV> using StridedB = Body<&strided_pass_body<T, INV>, &strided_pass<T, INV>>;
template <typename T> using MiddleB = Body<&middle_pass_body<T>, &middle_pass<T>>;
template <typename T> using ZForwardRealB = Body<&z_forward_real_body<T>, &z_forward_real<T>>;
template <typename T> using ZInverseRealB = Body<&z_inverse_real_body<T>, &z_inverse_real<T>>;
template <typename T> using ZVelocityB = Body<&z_velocity_body<T>, &z_velocity<T>>;
template <typename T> using ZDensityB = Body<&z_density_body<T>, &z_density<T>>;
template <typename T> using ZEosB = Body<&z_eos_body<T>, &z_eos<T>>;
template <typename T> using ZInitB = Body<&z_init_body<T>, &z_init<T>>;
template <typename T> using FillB = Body<&fill_kernel_body<T>, &fill_kernel<T>>;
template <typename T> using ScatterB = Body<&scatter_kernel_body<T>, &scatter_kernel<T>>;
using AdvanceB = Body<&advance_kernel_body, &advance_kernel>;
template <typename T> using GatherB = Body<&gather_kernel_body<T>, &gather_kernel<T>>;
template <typename T> using GatherPosB = Body<&gather_pos_kernel_body<T>, &gather_pos_kernel<T>>;

}  // namespace kwh
