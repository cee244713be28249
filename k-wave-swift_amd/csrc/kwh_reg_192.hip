// kwh_reg_192.hip -- register-resident kernels for axis length 192 (J = 8 lanes x E = 24 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_192() { return RegImpl<float, 8, 24>::ops(); }
}  // namespace kwh
