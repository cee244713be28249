// kwh_planeops_impl.h -- builds the PlaneOps table of one plane shape for one working type.
// Included only by the per-shape translation units kwh_plane_<m1>x<m2>.hip.
#pragma once
#include "kwh_plane.h"
#include "kwh_planeops.h"

namespace kwh {

template <typename T, int LY, int LZ, int NT>
struct PlaneImpl {
    using S = PlaneShape<T, LY, LZ>;
    static constexpr int W = plane_waves<T, LY, LZ, NT>();
    using FwdB = Body<&plane_fwd_real_body<T, LY, LZ, NT>, &plane_fwd_real<T, LY, LZ, NT>, W, false, NT>;
    using InvB = Body<&plane_inv_real_body<T, LY, LZ, NT>, &plane_inv_real<T, LY, LZ, NT>, W, false, NT>;
    template <int CORR>
    using VelB = Body<&plane_velocity_body<T, LY, LZ, NT, CORR>, &plane_velocity<T, LY, LZ, NT, CORR>, W, false, NT>;
    // the velocity kernel that inverts a corrected source's spectrum itself: 128 x 128 float32 only
    static constexpr bool kCorrFuse = LY * LZ >= 128 * 128 && sizeof(T) == 4;
    template <bool NL, int MODE, bool CORR>
    using DenB = Body<&plane_density_body<T, LY, LZ, NT, NL, MODE, CORR>, &plane_density<T, LY, LZ, NT, NL, MODE, CORR>,
                      W, false, NT>;
    template <bool NL>
    using EosB = Body<&plane_eos_body<T, LY, LZ, NT, NL>, &plane_eos<T, LY, LZ, NT, NL>, W, false, NT>;

    static hipError_t set_lds(const void* fn) {
        return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S::kLds);
    }
    template <bool NL, int MODE>
    static hipError_t set_density() {
        hipError_t e = set_lds((const void*)&plane_density<T, LY, LZ, NT, NL, MODE, false>);
        return e != hipSuccess ? e : set_lds((const void*)&plane_density<T, LY, LZ, NT, NL, MODE, true>);
    }
    static hipError_t prepare() {
        if (S::kLds <= 65536) return hipSuccess;
        const void* fns[] = {(const void*)&plane_fwd_real<T, LY, LZ, NT>, (const void*)&plane_inv_real<T, LY, LZ, NT>,
                             (const void*)&plane_velocity<T, LY, LZ, NT, 0>,
                             (const void*)&plane_velocity<T, LY, LZ, NT, 1>};
        for (const void* f : fns)
            if (hipError_t e = set_lds(f); e != hipSuccess) return e;
        if constexpr (kCorrFuse) {
            if (hipError_t e = set_lds((const void*)&plane_velocity<T, LY, LZ, NT, 2>); e != hipSuccess) return e;
        }
        hipError_t e = set_density<false, 0>();
        if (e == hipSuccess) e = set_density<true, 0>();
        if (e == hipSuccess) e = set_density<false, 1>();
        if (e == hipSuccess) e = set_density<true, 1>();
        if (e == hipSuccess) e = set_density<false, 2>();
        if (e == hipSuccess) e = set_density<true, 2>();
        if (e == hipSuccess) e = set_lds((const void*)&plane_eos<T, LY, LZ, NT, false>);
        if (e == hipSuccess) e = set_lds((const void*)&plane_eos<T, LY, LZ, NT, true>);
        return e;
    }

    static void fwd_real(const Emit& em, int planes, const T* src, cx<T>* spec, T scale, const Geo& g,
                         const AxisTabs<T>& tab) {
        em.go<FwdB>(dim3(planes), dim3(NT), S::kLds, src, spec, scale, g, tab);
    }
    static void inv_real(const Emit& em, int planes, const cx<T>* spec, T* dst, T scale, int accumulate, const Geo& g,
                         const AxisTabs<T>& tab) {
        em.go<InvB>(dim3(planes), dim3(NT), S::kLds, spec, dst, scale, accumulate, g, tab);
    }
    static void velocity(const Emit& em, int planes, const VelArgs<T>& va, const Geo& g, const AxisTabs<T>& tab,
                         bool corr) {
        if constexpr (kCorrFuse) {
            if (va.corr_spec) {
                em.go<VelB<2>>(dim3(planes), dim3(NT), S::kLds, va, g, tab);
                return;
            }
        }
        if (corr) em.go<VelB<1>>(dim3(planes), dim3(NT), S::kLds, va, g, tab);
        else em.go<VelB<0>>(dim3(planes), dim3(NT), S::kLds, va, g, tab);
    }
    template <bool NL, int MODE>
    static void density_m(const Emit& em, int planes, const RhoArgs<T>& ra, const Geo& g, const AxisTabs<T>& tab,
                          bool corr) {
        if (corr) em.go<DenB<NL, MODE, true>>(dim3(planes), dim3(NT), S::kLds, ra, g, tab);
        else em.go<DenB<NL, MODE, false>>(dim3(planes), dim3(NT), S::kLds, ra, g, tab);
    }
    static void density(const Emit& em, int planes, const RhoArgs<T>& ra, const Geo& g, const AxisTabs<T>& tab,
                        bool nonlinear, int mode, bool corr) {
        if (mode == 2) {
            if (nonlinear) density_m<true, 2>(em, planes, ra, g, tab, corr);
            else density_m<false, 2>(em, planes, ra, g, tab, corr);
        } else if (mode == 1) {
            if (nonlinear) density_m<true, 1>(em, planes, ra, g, tab, corr);
            else density_m<false, 1>(em, planes, ra, g, tab, corr);
        } else {
            if (nonlinear) density_m<true, 0>(em, planes, ra, g, tab, corr);
            else density_m<false, 0>(em, planes, ra, g, tab, corr);
        }
    }

    static void eos(const Emit& em, int planes, const EosArgs<T>& ea, const Geo& g, const AxisTabs<T>& tab) {
        if (ea.nonlinear) em.go<EosB<true>>(dim3(planes), dim3(NT), S::kLds, ea, g, tab);
        else em.go<EosB<false>>(dim3(planes), dim3(NT), S::kLds, ea, g, tab);
    }

    static const PlaneOps<T>* ops() {
        static const PlaneOps<T> o{LY, LZ, NT, S::kLds, kCorrFuse, &prepare, &fwd_real, &inv_real,
                                   &velocity, &density, &eos};
        return &o;
    }
};

}  // namespace kwh
