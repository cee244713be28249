// kwh_reg_160.hip -- register-resident kernels for axis length 160 (J = 4 lanes x E = 40 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_160() { return RegImpl<float, 4, 40>::ops(); }
}  // namespace kwh
