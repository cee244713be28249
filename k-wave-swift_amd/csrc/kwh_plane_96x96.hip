// kwh_plane_96x96.hip -- y-z plane kernels for 96 x 96 planes (384 lanes per plane).
#include "kwh_planeops_impl.h"

namespace kwh {
const PlaneOps<float>* plane_ops_f32_96x96() { return PlaneImpl<float, 96, 96, 384>::ops(); }
const PlaneOps<double>* plane_ops_f64_96x96() { return PlaneImpl<double, 96, 96, 384>::ops(); }
}  // namespace kwh
