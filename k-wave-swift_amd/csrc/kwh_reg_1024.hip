// kwh_reg_1024.hip -- register-resident kernels for axis length 1024 (J = 32 lanes x E = 32 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_1024() { return RegImpl<float, 32, 32>::ops(); }
}  // namespace kwh
