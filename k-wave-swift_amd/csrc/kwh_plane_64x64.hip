// kwh_plane_64x64.hip -- y-z plane kernels for 64 x 64 planes (256 lanes per plane).
#include "kwh_planeops_impl.h"

namespace kwh {
const PlaneOps<float>* plane_ops_f32_64x64() { return PlaneImpl<float, 64, 64, 256>::ops(); }
const PlaneOps<double>* plane_ops_f64_64x64() { return PlaneImpl<double, 64, 64, 256>::ops(); }
}  // namespace kwh
