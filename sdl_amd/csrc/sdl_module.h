// The module detId layout of the reference (Module.cc:159-358, Module.h:40-230),
// decoded on the host: the one place this library reads detId bits.
//
//   barrel (subdet 5):  subdet[27:25] layer[22:20] side[19:18] rod[16:10] module[8:2] lower/upper[0]
//   endcap (subdet 4):  subdet[27:25] side[24:23] layer[20:18] ring[15:12] module[8:2] lower/upper[0]
//
// A module is "inverted" when its lower sensor carries the odd detId bit 0 = 0
// (alternate modules are mounted the other way round); isLower and the partner
// sensor's detId follow from it.
#pragma once

#include <cstdint>

namespace sdl {

enum : uint8_t { MOD_SUBDET_ENDCAP = 4, MOD_SUBDET_BARREL = 5 };
enum : uint8_t { MOD_SIDE_NEGZ = 1, MOD_SIDE_POSZ = 2, MOD_SIDE_CENTER = 3 };
enum : uint8_t { MOD_TYPE_PS = 0, MOD_TYPE_2S = 1 };          // Module::ModuleType
enum : uint8_t { MOD_LAYER_PIXEL = 0, MOD_LAYER_STRIP = 1 };  // Module::ModuleLayerType

struct ModuleId {
    uint32_t detId, partner;
    uint16_t subdet, side, layer, rod, ring, module;
    bool isLower, isInverted;
    uint8_t type, layerType;
};

inline ModuleId decodeDetId(uint32_t d) {
    ModuleId m{};
    m.detId = d;
    m.subdet = (d >> 25) & 7;
    m.module = (d >> 2) & 127;
    if (m.subdet == MOD_SUBDET_ENDCAP) {
        m.side = (d >> 23) & 3;
        m.layer = (d >> 18) & 7;
        m.ring = (d >> 12) & 15;
    } else if (m.subdet == MOD_SUBDET_BARREL) {
        m.side = (d >> 18) & 3;
        m.layer = (d >> 20) & 7;
        m.rod = (d >> 10) & 127;
    }
    // Module::parseIsInverted (Module.cc:239-302): by side and layer, the odd or
    // the even module numbers are inverted; categories the reference cannot
    // place read false
    const bool odd = m.module % 2 == 1;
    if (m.subdet == MOD_SUBDET_ENDCAP) {
        m.isInverted = m.side == MOD_SIDE_NEGZ ? odd : (m.side == MOD_SIDE_POSZ ? !odd : false);
    } else if (m.subdet == MOD_SUBDET_BARREL) {
        if (m.side == MOD_SIDE_CENTER) m.isInverted = m.layer <= 3 ? odd : !odd;
        else if (m.side == MOD_SIDE_NEGZ || m.side == MOD_SIDE_POSZ)
            m.isInverted = m.layer <= 2 ? odd : (m.layer == 3 ? !odd : false);
    }
    // Module::parseIsLower and parsePartnerDetId (Module.cc:234-237, 304-310)
    m.isLower = m.isInverted ? !(d & 1) : (d & 1);
    m.partner = m.isLower != m.isInverted ? d + 1 : d - 1;
    // Module::parseModuleType, parseModuleLayerType (Module.cc:312-358)
    if (m.subdet == MOD_SUBDET_BARREL) m.type = m.layer <= 3 ? MOD_TYPE_PS : MOD_TYPE_2S;
    else if (m.layer <= 2) m.type = m.ring <= 10 ? MOD_TYPE_PS : MOD_TYPE_2S;
    else m.type = m.ring <= 7 ? MOD_TYPE_PS : MOD_TYPE_2S;
    if (m.type == MOD_TYPE_2S) m.layerType = MOD_LAYER_STRIP;
    else m.layerType = (m.isLower != m.isInverted) ? MOD_LAYER_PIXEL : MOD_LAYER_STRIP;
    return m;
}

}  // namespace sdl
