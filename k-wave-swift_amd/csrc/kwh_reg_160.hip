// kwh_reg_160.hip -- register-resident kernels for axis length 160 (J = 4 lanes x E = 40 elements;
// the middle pass also with 8 lanes x 20 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_160() {
    static const RegOps<float> o = with_middle_t<RegImpl<float, 4, 40>, float, 8, 20>();
    return &o;
}
}  // namespace kwh
