// kwh_regops.h -- launchers of the register-resident kernels for one axis length.
//
// The kernels of each supported length L = J*E (kwh_reg.h) are compiled in a translation unit of
// their own (kwh_reg_<L>.hip), which fills a RegOps table of launchers; the host code of the
// time step looks the table up by length (reg_ops) and falls back to the generic Stockham
// kernels (kwh_kernels.h) for lengths without one. Every launcher issues its launch through an
// Emit, so single and batched runs use the same table.
#pragma once
#include "kwh_kernels.h"
#include "kwh_launch.h"

namespace kwh {

template <typename T>
struct RegOps {
    int J, E;        // lanes per transform, elements per lane (L = J * E)
    int ntr_max;     // widest column tile of the strided / middle kernels (ntr_max / 2 also built)
    // contiguous-axis row kernels: P row pairs per workgroup (P * J lanes)
    void (*z_forward_real)(const Emit&, dim3, dim3, size_t, const T*, cx<T>*, T, const Geo&, int, const AxisTabs<T>&);
    void (*z_inverse_real)(const Emit&, dim3, dim3, size_t, const cx<T>*, T*, T, int, const Geo&, int,
                           const AxisTabs<T>&);
    void (*z_colloc_rows)(const Emit&, dim3, dim3, size_t, const cx<T>*, T*, const int*, int, const Geo&, int,
                          const AxisTabs<T>&);
    void (*z_velocity)(const Emit&, dim3, dim3, size_t, const VelArgs<T>&, const Geo&, int, const AxisTabs<T>&,
                       bool corr);
    // mode 0: EOS; 1: FFT absorption inputs; 2: EOS with Stokes absorption
    void (*z_density)(const Emit&, dim3, dim3, size_t, const RhoArgs<T>&, const Geo&, int, const AxisTabs<T>&,
                      bool nonlinear, int mode, bool corr);
    void (*z_eos)(const Emit&, dim3, dim3, size_t, const EosArgs<T>&, const Geo&, int, const AxisTabs<T>&);
    // rhoT = sum_a rho_a and F2 = R2C_z(rhoT), after z_density mode 1 when rhot_split is set
    bool rhot_split;
    void (*z_rhot)(const Emit&, dim3, dim3, size_t, const RhoArgs<T>&, const Geo&, int, const AxisTabs<T>&);
    // adjoint row kernels (AdjRowArgs): forward mode 0 / 1, inverse mode 0 / 1 / 2
    void (*adj_fwd)(const Emit&, dim3, dim3, size_t, const AdjRowArgs<T>&, const Geo&, int, const AxisTabs<T>&,
                    int mode);
    void (*adj_inv)(const Emit&, dim3, dim3, size_t, const cx<T>*, T*, const T*, int ia, int mode, const Geo&, int,
                    const AxisTabs<T>&);
    // column kernels over tiles of ntr columns (ntr_max or ntr_max / 2), ntr * J lanes
    void (*strided)(const Emit&, dim3, dim3, size_t, const StridedArgs&, const cx<T>*, bool inv, int ntr);
    void (*middle)(const Emit&, dim3, dim3, size_t, const MidArgs&, const AxisTabs<T>&, const Geo&, int ntr);
    // Rows of this length as one real row per length-L/2 complex transform (even/odd split, half_J
    // lanes x half_E elements, P rows per workgroup) for the velocity, density and EOS kernels;
    // half_J = 0: none. The density kernel then also forms rhoT and F2 (no z_rhot pass).
    int half_J, half_E;
    void (*z_velocity_h)(const Emit&, dim3, dim3, size_t, const VelArgs<T>&, const Geo&, int, const AxisTabs<T>&,
                         bool corr);
    void (*z_density_h)(const Emit&, dim3, dim3, size_t, const RhoArgs<T>&, const Geo&, int, const AxisTabs<T>&,
                        bool nonlinear, int mode, bool corr);
    void (*z_eos_h)(const Emit&, dim3, dim3, size_t, const EosArgs<T>&, const Geo&, int, const AxisTabs<T>&);
};

// Launchers for axis length L, or nullptr when no register-resident kernel covers it.
const RegOps<float>* reg_ops_f32(int L);
const RegOps<double>* reg_ops_f64(int L);
template <typename T> inline const RegOps<T>* reg_ops(int L);
template <> inline const RegOps<float>* reg_ops<float>(int L) { return reg_ops_f32(L); }
template <> inline const RegOps<double>* reg_ops<double>(int L) { return reg_ops_f64(L); }

}  // namespace kwh
