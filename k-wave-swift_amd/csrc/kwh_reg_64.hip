// kwh_reg_64.hip -- register-resident kernels for axis length 64 (J = 8 lanes x E = 8 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_64() { return RegImpl<float, 8, 8>::ops(); }
const RegOps<double>* reg_ops_f64_64() { return RegImpl<double, 8, 8>::ops(); }
}  // namespace kwh
