// kwh_reg_96.hip -- register-resident kernels for axis length 96 (J = 4 lanes x E = 24 elements;
// the middle pass also with 8 lanes x 12 elements).
#include "kwh_regops_impl.h"

namespace kwh {
const RegOps<float>* reg_ops_f32_96() {
    static const RegOps<float> o = with_middle_t<RegImpl<float, 4, 24>, float, 8, 12>();
    return &o;
}
}  // namespace kwh
