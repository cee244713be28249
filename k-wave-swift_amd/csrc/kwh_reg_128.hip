// kwh_reg_128.hip -- register-resident kernels for axis length 128 (J = 8 lanes x E = 16 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_128() { return RegImpl<float, 8, 16>::ops(); }
}  // namespace kwh
