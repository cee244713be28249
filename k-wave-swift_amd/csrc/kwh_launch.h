// kwh_launch.h -- how a simulation issues the launches of its time step.
//
// Every launch of a step goes through an Emit: launched on the simulation's stream, or, while a
// batched run records one simulation's step, appended to the recording as a LaunchRec (the
// batched kernel of the body, the simulation's own grid / workgroup / LDS and its packed
// arguments) so that launches of many simulations can be concatenated (kwh_batch.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <tuple>
#include <vector>

#include "kwh_batch.h"

namespace kwh {

struct LaunchRec {
    const void* fn;
    unsigned gx, gy, block;
    size_t lds;
    bool flex;                          // body runs with any workgroup size
    std::vector<unsigned char> pack;
};

template <class B, class... A>
LaunchRec make_rec(dim3 grid, dim3 block, size_t lds, const A&... a) {
    using P = BatchPack<A...>;
    P pk{0, (int)grid.x, 0, 0, std::tuple<A...>(a...)};
    LaunchRec r{(const void*)&batched_kernel<B, A...>, grid.x, grid.y, block.x, lds, B::kFlex, {}};
    r.pack.resize(sizeof(P));
    std::memcpy(r.pack.data(), &pk, sizeof(P));
    return r;
}

struct Emit {
    hipStream_t stream;
    std::vector<LaunchRec>* recording;
    template <class B, class... A>
    void go(dim3 grid, dim3 block, size_t lds, const A&... a) const {
        if (recording) {
            recording->push_back(make_rec<B>(grid, block, lds, a...));
            return;
        }
        hipLaunchKernelGGL(B::kernel, grid, block, lds, stream, a...);
    }
};

}  // namespace kwh
