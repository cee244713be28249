// kwh_reg_256.hip -- register-resident kernels for axis length 256 (J = 16 lanes x E = 16 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_256() { return RegImpl<float, 16, 16>::ops(); }
const RegOps<double>* reg_ops_f64_256() { return RegImpl<double, 16, 16>::ops(); }
}  // namespace kwh
