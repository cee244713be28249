// kwh_reg_32.hip -- register-resident kernels for axis length 32 (J = 4 lanes x E = 8 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_32() { return RegImpl<float, 4, 8>::ops(); }
}  // namespace kwh
