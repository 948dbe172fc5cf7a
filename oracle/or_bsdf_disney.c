/*
 * or_bsdf_disney.c -- oracle restatement of DisneyPrincipledBSDF and
 * HairMarschner.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <string.h>
#include "oracle.h"

int or_disney_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc,
                      float u0, float u1, or_srec *s) {
  (void)m; (void)ray_d; (void)h; (void)uc; (void)u0; (void)u1; (void)s;
  return 0;
}
v3 or_disney_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  (void)m; (void)wo; (void)wi; (void)h;
  return vs(0.f);
}
float or_disney_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  (void)m; (void)wo; (void)wi; (void)h;
  return 0.f;
}
int or_hair_scatter(const pt_material_desc *m, v3 ray_d, const or_hit *h, float uc,
                    float u0, float u1, or_srec *s) {
  (void)m; (void)ray_d; (void)h; (void)uc; (void)u0; (void)u1; (void)s;
  return 0;
}
v3 or_hair_f(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  (void)m; (void)wo; (void)wi; (void)h;
  return vs(0.f);
}
float or_hair_pdf(const pt_material_desc *m, v3 wo, v3 wi, const or_hit *h) {
  (void)m; (void)wo; (void)wi; (void)h;
  return 0.f;
}
