// kwh_plane_table.hip -- plane shape -> y-z plane kernel launchers (kwh_planeops.h).
#include "kwh_planeops.h"

namespace kwh {

const PlaneOps<float>* plane_ops_f32_128x128();
const PlaneOps<double>* plane_ops_f64_128x128();
const PlaneOps<float>* plane_ops_f32_96x96();
const PlaneOps<double>* plane_ops_f64_96x96();
const PlaneOps<float>* plane_ops_f32_64x64();
const PlaneOps<double>* plane_ops_f64_64x64();
const PlaneOps<float>* plane_ops_f32_32x32();
const PlaneOps<double>* plane_ops_f64_32x32();

const PlaneOps<float>* plane_ops_f32(int m1, int m2) {
    if (m1 == 128 && m2 == 128) return plane_ops_f32_128x128();
    if (m1 == 96 && m2 == 96) return plane_ops_f32_96x96();
    if (m1 == 64 && m2 == 64) return plane_ops_f32_64x64();
    if (m1 == 32 && m2 == 32) return plane_ops_f32_32x32();
    return nullptr;
}

const PlaneOps<double>* plane_ops_f64(int m1, int m2) {
    if (m1 == 128 && m2 == 128) return plane_ops_f64_128x128();
    if (m1 == 96 && m2 == 96) return plane_ops_f64_96x96();
    if (m1 == 64 && m2 == 64) return plane_ops_f64_64x64();
    if (m1 == 32 && m2 == 32) return plane_ops_f64_32x32();
    return nullptr;
}

}  // namespace kwh
