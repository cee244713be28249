// kwh_reg_512.hip -- register-resident kernels for axis length 512 (J = 16 lanes x E = 32 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_512() { return RegImpl<float, 16, 32>::ops(); }
}  // namespace kwh
