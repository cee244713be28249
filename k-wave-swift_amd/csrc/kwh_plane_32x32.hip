// kwh_plane_32x32.hip -- y-z plane kernels for 32 x 32 planes (256 lanes per plane).
#include "kwh_planeops_impl.h"

namespace kwh {
const PlaneOps<float>* plane_ops_f32_32x32() { return PlaneImpl<float, 32, 32, 256>::ops(); }
const PlaneOps<double>* plane_ops_f64_32x32() { return PlaneImpl<double, 32, 32, 256>::ops(); }
}  // namespace kwh
