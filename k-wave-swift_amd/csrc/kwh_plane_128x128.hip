// kwh_plane_128x128.hip -- y-z plane kernels for 128 x 128 planes (1024 lanes per plane).
#include "kwh_planeops_impl.h"

namespace kwh {
const PlaneOps<float>* plane_ops_f32_128x128() { return PlaneImpl<float, 128, 128, 1024>::ops(); }
const PlaneOps<double>* plane_ops_f64_128x128() { return PlaneImpl<double, 128, 128, 1024>::ops(); }
}  // namespace kwh
