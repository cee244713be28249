// kwh_reg_table.hip -- axis length -> register-resident launchers (kwh_regops.h).
#include "kwh_regops.h"

namespace kwh {

const RegOps<float>* reg_ops_f32_32();
const RegOps<float>* reg_ops_f32_64();
const RegOps<float>* reg_ops_f32_96();
const RegOps<float>* reg_ops_f32_128();
const RegOps<float>* reg_ops_f32_160();
const RegOps<float>* reg_ops_f32_192();
const RegOps<float>* reg_ops_f32_256();
const RegOps<float>* reg_ops_f32_512();
const RegOps<float>* reg_ops_f32_1024();
const RegOps<double>* reg_ops_f64_64();
const RegOps<double>* reg_ops_f64_256();

const RegOps<float>* reg_ops_f32(int L) {
    switch (L) {
        case 32: return reg_ops_f32_32();
        case 64: return reg_ops_f32_64();
        case 96: return reg_ops_f32_96();
        case 128: return reg_ops_f32_128();
        case 160: return reg_ops_f32_160();
        case 192: return reg_ops_f32_192();
        case 256: return reg_ops_f32_256();
        case 512: return reg_ops_f32_512();
        case 1024: return reg_ops_f32_1024();
        default: return nullptr;
    }
}

const RegOps<double>* reg_ops_f64(int L) {
    switch (L) {
        case 64: return reg_ops_f64_64();
        case 256: return reg_ops_f64_256();
        default: return nullptr;
    }
}

}  // namespace kwh
