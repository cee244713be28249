// kwh_reg_96.hip -- register-resident kernels for axis length 96 (J = 4 lanes x E = 24 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_96() { return RegImpl<float, 4, 24>::ops(); }
}  // namespace kwh
