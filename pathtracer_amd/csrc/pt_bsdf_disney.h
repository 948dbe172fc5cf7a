// pt_bsdf_disney.h -- DisneyPrincipledBSDF and HairMarschner on the device.
#pragma once
#include "pt_math.h"

namespace pt {

PT_D bool disney_scatter(const pt_material_desc &m, f3 wo, ShadePoint &h, float uc, float u0, float u1,
                         Scatter &s) {
  return false;
}
PT_D f3 disney_f(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) { return s3(0.f); }
PT_D float disney_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) { return 0.f; }
PT_D bool hair_scatter(const pt_material_desc &m, f3 wo, ShadePoint &h, float uc, float u0, float u1,
                       Scatter &s) {
  return false;
}
PT_D f3 hair_f(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) { return s3(0.f); }
PT_D float hair_pdf(const pt_material_desc &m, f3 wo, f3 wi, const ShadePoint &h) { return 0.f; }

}  // namespace pt
