// kwh_regops_impl.h -- builds the RegOps table of one register-resident length (L = J * E) for
// one working type. Included only by the per-length translation units kwh_reg_<L>.hip.
#pragma once
#include "kwh_reg.h"
#include "kwh_regops.h"
#include <cstdlib>

namespace kwh {

// Launchers of the half-form row kernels (one real row of length 2 J E per lane group).
template <typename T, int J, int E>
struct RegHalfImpl {
    template <bool CORR>
    using VelB = Body<&reg_z_velocity_half_body<T, J, E, CORR>, &reg_z_velocity_half<T, J, E, CORR>,
                      reg_waves<T, E>(KWH_VELOCITY_WAVES), false>;
    template <bool NL, int MODE, bool CORR>
    using DenB = Body<&reg_z_density_half_body<T, J, E, NL, MODE, CORR>, &reg_z_density_half<T, J, E, NL, MODE, CORR>,
                      density_waves<T, J, E, NL, MODE, CORR>(), false>;
    static void z_velocity(const Emit& em, dim3 gr, dim3 bl, size_t lds, const VelArgs<T>& va, const Geo& g, int P,
                           const AxisTabs<T>& tab, bool corr) {
        if (corr) em.go<VelB<true>>(gr, bl, lds, va, g, P, tab);
        else em.go<VelB<false>>(gr, bl, lds, va, g, P, tab);
    }
    template <bool NL>
    static void density_nl(const Emit& em, dim3 gr, dim3 bl, size_t lds, const RhoArgs<T>& ra, const Geo& g, int P,
                           const AxisTabs<T>& tab, int mode, bool corr) {
        if (mode == 1) {
            if (corr) em.go<DenB<NL, 1, true>>(gr, bl, lds, ra, g, P, tab);
            else em.go<DenB<NL, 1, false>>(gr, bl, lds, ra, g, P, tab);
        } else if (mode == 2) {
            if (corr) em.go<DenB<NL, 2, true>>(gr, bl, lds, ra, g, P, tab);
            else em.go<DenB<NL, 2, false>>(gr, bl, lds, ra, g, P, tab);
        } else {
            if (corr) em.go<DenB<NL, 0, true>>(gr, bl, lds, ra, g, P, tab);
            else em.go<DenB<NL, 0, false>>(gr, bl, lds, ra, g, P, tab);
        }
    }
    static void z_density(const Emit& em, dim3 gr, dim3 bl, size_t lds, const RhoArgs<T>& ra, const Geo& g, int P,
                          const AxisTabs<T>& tab, bool nonlinear, int mode, bool corr) {
        if (nonlinear) density_nl<true>(em, gr, bl, lds, ra, g, P, tab, mode, corr);
        else density_nl<false>(em, gr, bl, lds, ra, g, P, tab, mode, corr);
    }
    static void z_eos(const Emit& em, dim3 gr, dim3 bl, size_t lds, const EosArgs<T>& ea, const Geo& g, int P,
                      const AxisTabs<T>& tab) {
        em.go<Body<&reg_z_eos_half_body<T, J, E>, &reg_z_eos_half<T, J, E>, eos_waves<T, E>(), false>>(
            gr, bl, lds, ea, g, P, tab);
    }
};

template <typename T, int J, int E>
struct RegImpl {
    static_assert(J * E <= kScalarRep, "a scalar map's buffer must cover a row (SMap::lane)");
    static constexpr int NTR = reg_ntr_max<T, J, E>();
    static_assert(NTR >= 2, "column tile too narrow");

    template <bool CORR>
    using VelB = Body<&reg_z_velocity_body<T, J, E, CORR>, &reg_z_velocity<T, J, E, CORR>,
                      reg_waves<T, E>(KWH_VELOCITY_WAVES), false>;
    template <bool NL, int MODE, bool CORR>
    using DenB = Body<&reg_z_density_body<T, J, E, NL, MODE, CORR>, &reg_z_density<T, J, E, NL, MODE, CORR>,
                      density_waves<T, J, E, NL, MODE, CORR>(), false>;
    template <int N, bool INV>
    using StrB = Body<&reg_strided_body<T, J, E, N, INV>, &reg_strided<T, J, E, N, INV>, 1, false>;
    template <int N>
    using StrDualB = Body<&reg_strided_body<T, J, E, N, true, true>, &reg_strided<T, J, E, N, true, true>, 1, false>;
    template <int N>
    using MidB = Body<&reg_middle_body<T, J, E, N>, &reg_middle<T, J, E, N>, reg_waves<T, E>(KWH_MIDDLE_WAVES), false>;
    template <int N>
    using MidXB = Body<&reg_middle_body<T, J, E, N, true>, &reg_middle<T, J, E, N, true>,
                       reg_waves<T, E>(KWH_MIDDLE_WAVES), false>;
    // Middle pass with the inner radix across lanes (DPP) for J up to KWH_MIDDLE_XLANE_MAXJ;
    // KWH_MIDDLE_XLANE=0/1 in the environment overrides.
    static bool middle_xlane() {
        static const bool on = [] {
            const char* e = std::getenv("KWH_MIDDLE_XLANE");
            return e && *e ? std::atoi(e) != 0 : J <= KWH_MIDDLE_XLANE_MAXJ;
        }();
        return on;
    }

    static void z_forward_real(const Emit& em, dim3 gr, dim3 bl, size_t lds, const T* src, cx<T>* spec, T scale,
                               const Geo& g, int P, const AxisTabs<T>& tab) {
        em.go<Body<&reg_z_forward_real_body<T, J, E>, &reg_z_forward_real<T, J, E>, 1, false>>(gr, bl, lds, src, spec,
                                                                                               scale, g, P, tab);
    }
    static void z_inverse_real(const Emit& em, dim3 gr, dim3 bl, size_t lds, const cx<T>* spec, T* dst, T scale,
                               int accumulate, const Geo& g, int P, const AxisTabs<T>& tab) {
        em.go<Body<&reg_z_inverse_real_body<T, J, E>, &reg_z_inverse_real<T, J, E>, 1, false>>(
            gr, bl, lds, spec, dst, scale, accumulate, g, P, tab);
    }
    static void z_colloc_rows(const Emit& em, dim3 gr, dim3 bl, size_t lds, const cx<T>* spec, T* dst,
                              const int* rows, int nlist, const Geo& g, int P, const AxisTabs<T>& tab) {
        em.go<Body<&reg_z_colloc_rows_body<T, J, E>, &reg_z_colloc_rows<T, J, E>, 1, false>>(
            gr, bl, lds, spec, dst, rows, nlist, g, P, tab);
    }
    static void z_velocity(const Emit& em, dim3 gr, dim3 bl, size_t lds, const VelArgs<T>& va, const Geo& g, int P,
                           const AxisTabs<T>& tab, bool corr) {
        if (corr) em.go<VelB<true>>(gr, bl, lds, va, g, P, tab);
        else em.go<VelB<false>>(gr, bl, lds, va, g, P, tab);
    }
    template <bool NL>
    static void density_nl(const Emit& em, dim3 gr, dim3 bl, size_t lds, const RhoArgs<T>& ra, const Geo& g, int P,
                           const AxisTabs<T>& tab, int mode, bool corr) {
        if (mode == 1) {
            if (corr) em.go<DenB<NL, 1, true>>(gr, bl, lds, ra, g, P, tab);
            else em.go<DenB<NL, 1, false>>(gr, bl, lds, ra, g, P, tab);
        } else if (mode == 2) {
            if (corr) em.go<DenB<NL, 2, true>>(gr, bl, lds, ra, g, P, tab);
            else em.go<DenB<NL, 2, false>>(gr, bl, lds, ra, g, P, tab);
        } else {
            if (corr) em.go<DenB<NL, 0, true>>(gr, bl, lds, ra, g, P, tab);
            else em.go<DenB<NL, 0, false>>(gr, bl, lds, ra, g, P, tab);
        }
    }
    static void z_density(const Emit& em, dim3 gr, dim3 bl, size_t lds, const RhoArgs<T>& ra, const Geo& g, int P,
                          const AxisTabs<T>& tab, bool nonlinear, int mode, bool corr) {
        if (nonlinear) density_nl<true>(em, gr, bl, lds, ra, g, P, tab, mode, corr);
        else density_nl<false>(em, gr, bl, lds, ra, g, P, tab, mode, corr);
    }
    static void z_eos(const Emit& em, dim3 gr, dim3 bl, size_t lds, const EosArgs<T>& ea, const Geo& g, int P,
                      const AxisTabs<T>& tab) {
        em.go<Body<&reg_z_eos_body<T, J, E>, &reg_z_eos<T, J, E>, eos_waves<T, E>(), false>>(
            gr, bl, lds, ea, g, P, tab);
    }
    static void z_rhot(const Emit& em, dim3 gr, dim3 bl, size_t lds, const RhoArgs<T>& ra, const Geo& g, int P,
                       const AxisTabs<T>& tab) {
        em.go<Body<&reg_z_rhot_body<T, J, E>, &reg_z_rhot<T, J, E>, 1, false>>(gr, bl, lds, ra, g, P, tab);
    }
    static void adj_fwd(const Emit& em, dim3 gr, dim3 bl, size_t lds, const AdjRowArgs<T>& a, const Geo& g, int P,
                        const AxisTabs<T>& tab, int mode) {
        if (mode == 0) em.go<Body<&reg_adj_fwd_body<T, J, E, 0>, &reg_adj_fwd<T, J, E, 0>, 1, false>>(gr, bl, lds, a, g, P, tab);
        else em.go<Body<&reg_adj_fwd_body<T, J, E, 1>, &reg_adj_fwd<T, J, E, 1>, 1, false>>(gr, bl, lds, a, g, P, tab);
    }
    static void adj_inv(const Emit& em, dim3 gr, dim3 bl, size_t lds, const cx<T>* spec, T* dst, const T* s, int ia,
                        int mode, const Geo& g, int P, const AxisTabs<T>& tab) {
        em.go<Body<&reg_adj_inv_body<T, J, E>, &reg_adj_inv<T, J, E>, 1, false>>(gr, bl, lds, spec, dst, s, ia, mode, g,
                                                                                 P, tab);
    }
    static void strided(const Emit& em, dim3 gr, dim3 bl, size_t lds, const StridedArgs& a, const cx<T>* tw, bool inv,
                        int ntr) {
        if (inv && a.dual > 0) {
            if (ntr == NTR) em.go<StrDualB<NTR>>(gr, bl, lds, a, tw);
            else em.go<StrDualB<NTR / 2>>(gr, bl, lds, a, tw);
            return;
        }
        if (ntr == NTR) {
            if (inv) em.go<StrB<NTR, true>>(gr, bl, lds, a, tw);
            else em.go<StrB<NTR, false>>(gr, bl, lds, a, tw);
        } else {
            if (inv) em.go<StrB<NTR / 2, true>>(gr, bl, lds, a, tw);
            else em.go<StrB<NTR / 2, false>>(gr, bl, lds, a, tw);
        }
    }
    static void middle(const Emit& em, dim3 gr, dim3 bl, size_t lds, const MidArgs& a, const AxisTabs<T>& tab,
                       const Geo& g, int ntr) {
        if constexpr (J <= 16) {   // the cross-lane network covers J = 4, 8, 16 (one DPP row)
            if (middle_xlane()) {
                const size_t lx = (size_t)J * E * sizeof(cx<T>);   // twiddle table only
                if (ntr == NTR) em.go<MidXB<NTR>>(gr, bl, lx, a, tab, g);
                else em.go<MidXB<NTR / 2>>(gr, bl, lx, a, tab, g);
                return;
            }
        }
        if (ntr == NTR) em.go<MidB<NTR>>(gr, bl, lds, a, tab, g);
        else em.go<MidB<NTR / 2>>(gr, bl, lds, a, tab, g);
    }

    // Length-512 FP32 rows also run in the half form (256-point transforms, 16 x 16): E = 32 row
    // pairs hold 64 complex registers per lane for the two transforms of the EOS kernel.
    static constexpr bool kHalf = sizeof(T) == 4 && J == 16 && E == 32;
    using H = RegHalfImpl<T, J, (kHalf ? E / 2 : E)>;
    static constexpr decltype(RegOps<T>::z_velocity) half_vel() {
        if constexpr (kHalf) return &H::z_velocity;
        else return nullptr;
    }
    static constexpr decltype(RegOps<T>::z_density) half_den() {
        if constexpr (kHalf) return &H::z_density;
        else return nullptr;
    }
    static constexpr decltype(RegOps<T>::z_eos) half_eos() {
        if constexpr (kHalf) return &H::z_eos;
        else return nullptr;
    }

    static const RegOps<T>* ops() {
        static const RegOps<T> o{J, E, NTR, &z_forward_real, &z_inverse_real, &z_colloc_rows, &z_velocity,
                                 &z_density, &z_eos, density_rhot_split<T, E, 1>(), &z_rhot, &adj_fwd, &adj_inv,
                                 &strided, &middle, kHalf ? J : 0, kHalf ? E / 2 : 0, half_vel(), half_den(),
                                 half_eos()};
        return &o;
    }
};

}  // namespace kwh
