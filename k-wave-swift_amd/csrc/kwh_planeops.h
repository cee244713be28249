// kwh_planeops.h -- launchers of the y-z plane kernels (kwh_plane.h) for one plane shape.
//
// Each supported plane shape [m1][m2] is compiled in a translation unit of its own
// (kwh_plane_<m1>x<m2>.hip) that fills a PlaneOps table; the host looks it up by shape
// (plane_ops) and uses the split passes for shapes without one. Launchers issue through an
// Emit, so single and batched runs share the table. Grid: one workgroup per x plane.
#pragma once
#include "kwh_kernels.h"
#include "kwh_launch.h"

namespace kwh {

template <typename T>
struct PlaneOps {
    int LY, LZ, NT;      // plane shape, workgroup size
    size_t lds;          // dynamic LDS per workgroup
    // the velocity kernel inverts a corrected source's spectrum itself (VelArgs::corr_spec): faster
    // than a separate plane inverse at 128 x 128 (batched velocity +47 us vs the inverse's 72 us),
    // slower at 96 x 96 (+75 vs 60 us)
    bool corr_fuse;
    // Raise the kernels' dynamic-LDS limit above the 64 KiB default (once, outside graph capture).
    hipError_t (*prepare)();
    void (*fwd_real)(const Emit&, int planes, const T*, cx<T>*, T, const Geo&, const AxisTabs<T>&);
    void (*inv_real)(const Emit&, int planes, const cx<T>*, T*, T, int accumulate, const Geo&, const AxisTabs<T>&);
    void (*velocity)(const Emit&, int planes, const VelArgs<T>&, const Geo&, const AxisTabs<T>&, bool corr);
    // mode 0: lossless EOS; 1: inputs of the FFT absorption term; 2: EOS with Stokes absorption
    void (*density)(const Emit&, int planes, const RhoArgs<T>&, const Geo&, const AxisTabs<T>&, bool nonlinear,
                    int mode, bool corr);
    // EOS with FFT power-law absorption
    void (*eos)(const Emit&, int planes, const EosArgs<T>&, const Geo&, const AxisTabs<T>&);
};

// Launchers for plane shape m1 x m2, or nullptr when no plane kernel covers it.
const PlaneOps<float>* plane_ops_f32(int m1, int m2);
const PlaneOps<double>* plane_ops_f64(int m1, int m2);
template <typename T> inline const PlaneOps<T>* plane_ops(int m1, int m2);
template <> inline const PlaneOps<float>* plane_ops<float>(int m1, int m2) { return plane_ops_f32(m1, m2); }
template <> inline const PlaneOps<double>* plane_ops<double>(int m1, int m2) { return plane_ops_f64(m1, m2); }

}  // namespace kwh
