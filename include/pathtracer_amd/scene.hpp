// scene.hpp -- source-compatible scene API over the C ABI of pathtracer_amd.h.
//
// The reference renderer's scenes are C++ functions that build a SceneInfo
// out of std::make_shared<Sphere>(...), Lambertian, Conductor, Translate,
// Rotate<Axis::Y>, loadOBJMesh(...) and so on (e.g. scenes/scenes_hdri.hpp)
// and call SceneInfo::render (render/renderManager.hpp:51-158).  This header
// declares the same class names and constructor signatures.  The objects only
// record parameters; SceneInfo::render lowers them onto pt_scene_* calls and
// renders on the GPU.  A scene function written for the reference compiles
// against this header unchanged when it uses the classes below.
//
// Lowering rules (the reference's run-time behaviour, resolved once at scene
// build time):
//   * Translate / Rotate<Axis> / Scale are flattened into world-space
//     geometry.  Rotate<Axis::A>(obj, angle) maps an object point p to the
//     world point R_A(angle)^T p (objectTransform.cpp:93-213 rotates the ray
//     by R_A); Scale(obj, s) maps p to s*p (objectTransform.hpp:121-146);
//     Translate(obj, o) to p + o (objectTransform.hpp:43-60).  A rect that a
//     rotation leaves non-axis-aligned becomes two triangles.
//   * HitRecord::is_inside: every Translate/Rotate/Scale re-runs set_normal
//     on a ray-opposed normal, which forces it true; FlipFace toggles it
//     (objectTransform.hpp:21).  The chain is folded from the innermost
//     wrapper outwards into PT_PRIM_FORCE_INSIDE / PT_PRIM_FLIP_FACE.
//   * Light discovery (lightDiscovery.hpp:5-81) only looks at top-level
//     objects whose underlying shape is an AARectangle or a Sphere; emissive
//     primitives anywhere else (inside a Box or a nested ObjectList) are
//     lowered with PT_PRIM_NOT_LIGHT.  The reference builds the light of a
//     wrapped top-level emitter from the untransformed shape; here the light
//     follows the transformed geometry.
//   * Materials are lowered once each, in order of first use.
//
// Supported: Lambertian (colour or texture), DiffuseEmitter (colour), Conductor,
// MFDielectric, ThinDielectric, Phong, DisneyPrincipledBSDF, HairMarschner,
// KajiyaKay; ConstantTexture, ImageTexture, CheckerTexture (on Lambertian and
// the Disney base colour); Sphere, AARectangle<Axis> (XY/XZ/YZRectangle), Box,
// TriangleMesh (with MeshData normals), loadOBJMesh, ObjectList, FlipFace,
// Translate, Rotate<Axis>, Scale; PerspectiveCamera; EnvironmentLight from
// .hdr or .pfm files.  NoiseTexture and textured emitters are not.
#pragma once

#include <array>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <optional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <variant>
#include <vector>

#include "../pathtracer_amd.h"
#include "image_io.hpp"

// ------------------------------------------------------------------ gl::vec
namespace gl {

struct vec2 {
  float e[2] = {0.f, 0.f};
  vec2() = default;
  explicit vec2(float s) : e{s, s} {}
  vec2(float a, float b) : e{a, b} {}
  float u() const { return e[0]; }
  float v() const { return e[1]; }
  float x() const { return e[0]; }
  float y() const { return e[1]; }
  float &operator[](int i) { return e[i]; }
  float operator[](int i) const { return e[i]; }
};

struct vec3 {
  float e[3] = {0.f, 0.f, 0.f};
  vec3() = default;
  vec3(float s) : e{s, s, s} {}  // implicit, as gl::vec3(0.7f) and colour constants use it
  vec3(float a, float b, float c) : e{a, b, c} {}
  float x() const { return e[0]; }
  float y() const { return e[1]; }
  float z() const { return e[2]; }
  float r() const { return e[0]; }
  float g() const { return e[1]; }
  float b() const { return e[2]; }
  float &operator[](int i) { return e[i]; }
  float operator[](int i) const { return e[i]; }
  vec3 operator-() const { return vec3(-e[0], -e[1], -e[2]); }
  vec3 &operator+=(const vec3 &o) { e[0] += o.e[0]; e[1] += o.e[1]; e[2] += o.e[2]; return *this; }
  vec3 &operator*=(float s) { e[0] *= s; e[1] *= s; e[2] *= s; return *this; }
  float length() const { return std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]); }
  vec3 normalize() const { float l = length(); return vec3(e[0] / l, e[1] / l, e[2] / l); }
};
inline vec3 operator+(const vec3 &a, const vec3 &b) { return vec3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline vec3 operator-(const vec3 &a, const vec3 &b) { return vec3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline vec3 operator*(const vec3 &a, const vec3 &b) { return vec3(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
inline vec3 operator*(const vec3 &a, float s) { return vec3(a[0] * s, a[1] * s, a[2] * s); }
inline vec3 operator*(float s, const vec3 &a) { return a * s; }
inline vec3 operator/(const vec3 &a, float s) { return vec3(a[0] / s, a[1] / s, a[2] / s); }
inline float dot(const vec3 &a, const vec3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline vec3 cross(const vec3 &a, const vec3 &b) {
  return vec3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline vec3 normalize(const vec3 &a) { return a.normalize(); }
inline float to_radian(float deg) { return deg * 3.14159265358979323846f / 180.f; }

}  // namespace gl

enum class Axis { X = 0, Y = 1, Z = 2 };

// ------------------------------------------------------------------ lowering
namespace ptamd {

// What a scene lowers to: the C ABI calls, in order.  SceneInfo::render uses
// the ABI sink below; tests can record the calls instead.
struct SceneSink {
  virtual ~SceneSink() = default;
  virtual uint32_t material(const pt_material_desc &d) = 0;
  virtual void mesh(const std::vector<float> &pos, const std::vector<int32_t> &idx, const std::vector<float> *nrm,
                    const std::vector<int32_t> *nidx, uint32_t mat, uint32_t flags) = 0;
  virtual void sphere(const float c[3], float r, uint32_t mat, uint32_t flags) = 0;
  virtual void rect(int axis, float k, float d0min, float d0max, float d1min, float d1max, uint32_t mat,
                    uint32_t flags) = 0;
  virtual void curve(const float cps[12], float w0, float w1, uint32_t type, uint32_t mat, uint32_t flags) {
    (void)cps, (void)w0, (void)w1, (void)type, (void)mat, (void)flags;
    throw std::runtime_error("this scene sink does not take curves");
  }
  // Textures return their id; mesh_uv gives the last mesh added its texture
  // coordinates (2 floats per uv, 3 uv indices per triangle).
  virtual uint32_t texture_image(const std::vector<float> &rgb, uint32_t w, uint32_t h) {
    (void)rgb, (void)w, (void)h;
    throw std::runtime_error("this scene sink does not take textures");
  }
  virtual uint32_t texture_checker(const float c1[3], const float c2[3], float scale) {
    (void)c1, (void)c2, (void)scale;
    throw std::runtime_error("this scene sink does not take textures");
  }
  virtual void mesh_uv(const std::vector<float> &uv, const std::vector<int32_t> &uvidx) { (void)uv, (void)uvidx; }
};

inline void check(pt_status st, const char *what) {
  if (st != PT_OK) throw std::runtime_error(std::string(what) + ": " + pt_last_error());
}

struct AbiSink : SceneSink {
  pt_scene *s;
  uint32_t meshes = 0;
  explicit AbiSink(pt_scene *scene) : s(scene) {}
  uint32_t material(const pt_material_desc &d) override {
    uint32_t id = 0;
    check(pt_scene_add_material(s, &d, &id), "pt_scene_add_material");
    return id;
  }
  void mesh(const std::vector<float> &pos, const std::vector<int32_t> &idx, const std::vector<float> *nrm,
            const std::vector<int32_t> *nidx, uint32_t mat, uint32_t flags) override {
    const uint32_t nv = (uint32_t)(pos.size() / 3), nt = (uint32_t)(idx.size() / 3);
    if (nrm)
      check(pt_scene_add_mesh_normals(s, pos.data(), nv, idx.data(), nt, nrm->data(), (uint32_t)(nrm->size() / 3),
                                      nidx->data(), mat, flags),
            "pt_scene_add_mesh_normals");
    else
      check(pt_scene_add_mesh(s, pos.data(), nv, idx.data(), nt, mat, flags), "pt_scene_add_mesh");
    ++meshes;
  }
  uint32_t texture_image(const std::vector<float> &rgb, uint32_t w, uint32_t h) override {
    uint32_t id = 0;
    check(pt_scene_add_texture_image(s, rgb.data(), w, h, &id), "pt_scene_add_texture_image");
    return id;
  }
  uint32_t texture_checker(const float c1[3], const float c2[3], float scale) override {
    uint32_t id = 0;
    check(pt_scene_add_texture_checker(s, c1, c2, scale, &id), "pt_scene_add_texture_checker");
    return id;
  }
  void mesh_uv(const std::vector<float> &uv, const std::vector<int32_t> &uvidx) override {
    if (meshes == 0) throw std::runtime_error("mesh_uv before any mesh");
    check(pt_scene_set_mesh_uv(s, meshes - 1, uv.data(), (uint32_t)(uv.size() / 2), uvidx.data(),
                               (uint32_t)(uvidx.size() / 3)),
          "pt_scene_set_mesh_uv");
  }
  void sphere(const float c[3], float r, uint32_t mat, uint32_t flags) override {
    check(pt_scene_add_sphere(s, c, r, mat, flags), "pt_scene_add_sphere");
  }
  void rect(int axis, float k, float d0min, float d0max, float d1min, float d1max, uint32_t mat,
            uint32_t flags) override {
    check(pt_scene_add_rect(s, axis, k, d0min, d0max, d1min, d1max, mat, flags), "pt_scene_add_rect");
  }
  void curve(const float cps[12], float w0, float w1, uint32_t type, uint32_t mat, uint32_t flags) override {
    check(pt_scene_add_curve(s, cps, w0, w1, type, mat, flags), "pt_scene_add_curve");
  }
};

// Object-to-world map p -> M p + t (rotations and uniform scale only).
struct Xform {
  float m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  float t[3] = {0, 0, 0};
  float scale = 1.f;
  bool identity_rotation = true;
  gl::vec3 point(const gl::vec3 &p) const {
    gl::vec3 r;
    for (int i = 0; i < 3; ++i) r[i] = m[i][0] * p[0] + m[i][1] * p[1] + m[i][2] * p[2] + t[i];
    return r;
  }
  gl::vec3 dir(const gl::vec3 &n) const {  // normals: rotation part only (uniform scale)
    gl::vec3 r;
    for (int i = 0; i < 3; ++i) r[i] = (m[i][0] * n[0] + m[i][1] * n[1] + m[i][2] * n[2]) / scale;
    return r;
  }
  // this o inner: apply `inner` first
  Xform then_outer_of(const Xform &inner) const {
    Xform r;
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) r.m[i][j] = m[i][0] * inner.m[0][j] + m[i][1] * inner.m[1][j] + m[i][2] * inner.m[2][j];
      r.t[i] = m[i][0] * inner.t[0] + m[i][1] * inner.t[1] + m[i][2] * inner.t[2] + t[i];
    }
    r.scale = scale * inner.scale;
    r.identity_rotation = identity_rotation && inner.identity_rotation;
    return r;
  }
};

// Wrapper chain from the outermost object inwards: 'F' FlipFace, 'T' transform.
struct Chain {
  std::string w;
  uint32_t flags() const {
    std::optional<bool> forced;
    bool flip = false;
    for (auto it = w.rbegin(); it != w.rend(); ++it) {  // innermost first
      if (*it == 'T') {
        forced = true;
        flip = false;
      } else if (forced) {
        forced = !*forced;
      } else {
        flip = !flip;
      }
    }
    if (forced) return *forced ? (uint32_t)PT_PRIM_FORCE_INSIDE : (uint32_t)(PT_PRIM_FORCE_INSIDE | PT_PRIM_FLIP_FACE);
    return flip ? (uint32_t)PT_PRIM_FLIP_FACE : 0u;
  }
};

struct Lowering;
}  // namespace ptamd

// ------------------------------------------------------------------ textures
// texture/texture.hpp:20-94.  A texture either is one colour everywhere
// (ConstantTexture) or lowers to a pt_scene texture that materials reference
// by id.  LERP_MODE is BILINEAR throughout, as every caller in the
// reference leaves it.
class Texture2D {
 public:
  virtual ~Texture2D() = default;
  virtual std::optional<gl::vec3> constant() const { return std::nullopt; }
  virtual uint32_t lower(ptamd::SceneSink &sink) const = 0;
};
class ConstantTexture : public Texture2D {
 public:
  ConstantTexture() = default;
  explicit ConstantTexture(const gl::vec3 &c) : c(c) {}
  std::optional<gl::vec3> constant() const override { return c; }
  uint32_t lower(ptamd::SceneSink &sink) const override {
    const float rgb[3] = {c[0], c[1], c[2]};
    return sink.texture_checker(rgb, rgb, 1.f);
  }
  gl::vec3 c;
};
using SolidColor = ConstantTexture;
// ImageTexture(file, flip_y, isHDR) (texture.cpp:3-91).  Files are read as
// image_io::read_texture reads them (.hdr/.pic, .pfm, .ppm); the in-memory
// form takes texels in the pt_scene_add_texture_image layout.
class ImageTexture : public Texture2D {
 public:
  ImageTexture(const std::string &filename, bool flip_y = false, bool isHDR = false) {
    ptamd::image_io::Image img = ptamd::image_io::read_texture(filename, flip_y, isHDR);
    width = img.width;
    height = img.height;
    rgb = std::move(img.rgb);
  }
  ImageTexture(std::vector<float> texels, uint32_t w, uint32_t h) : rgb(std::move(texels)), width(w), height(h) {}
  uint32_t lower(ptamd::SceneSink &sink) const override { return sink.texture_image(rgb, width, height); }
  std::vector<float> rgb;
  uint32_t width = 0, height = 0;
};
// texture.cpp:151-183: color1 where sin(u scale) sin(v scale) < 0.
class CheckerTexture : public Texture2D {
 public:
  explicit CheckerTexture(float scale = 10.f) : c1(0.f), c2(1.f), scale(scale) {}
  CheckerTexture(const gl::vec3 &color1, const gl::vec3 &color2, float scale = 10.f)
      : c1(color1), c2(color2), scale(scale) {}
  uint32_t lower(ptamd::SceneSink &sink) const override {
    const float a[3] = {c1[0], c1[1], c1[2]}, b[3] = {c2[0], c2[1], c2[2]};
    return sink.texture_checker(a, b, scale);
  }
  gl::vec3 c1, c2;
  float scale;
};
// texture.cpp:185-193 (fractal Perlin noise): declared so that scenes that
// name it compile; it has no GPU lowering and a scene using it is refused.
class NoiseTexture : public Texture2D {
 public:
  NoiseTexture() = default;
  NoiseTexture(float scale, int fractal = 1) : scale(scale), fractal(fractal) {}
  uint32_t lower(ptamd::SceneSink &) const override {
    throw std::runtime_error("NoiseTexture has no GPU lowering: use ImageTexture, CheckerTexture or a colour");
  }
  float scale = 1.f;
  int fractal = 1;
};
using ColorVariant = std::variant<gl::vec3, std::shared_ptr<Texture2D>>;

namespace ptamd {
// A colour argument as (constant colour, texture): the texture is null when
// the colour is the same everywhere.
inline std::pair<gl::vec3, std::shared_ptr<Texture2D>> color_or_texture(const ColorVariant &v) {
  if (std::holds_alternative<gl::vec3>(v)) return {std::get<gl::vec3>(v), nullptr};
  const auto &t = std::get<std::shared_ptr<Texture2D>>(v);
  if (!t) throw std::runtime_error("null texture");
  if (auto c = t->constant()) return {*c, nullptr};
  return {gl::vec3(0.f), t};
}
inline gl::vec3 color_of(const ColorVariant &v, const char *who) {
  auto ct = color_or_texture(v);
  if (ct.second) throw std::runtime_error(std::string(who) + " takes constant colours only");
  return ct.first;
}
}  // namespace ptamd

// ------------------------------------------------------------------ materials
class Material {
 public:
  virtual ~Material() = default;
  virtual pt_material_desc desc() const = 0;
  virtual bool is_emitter() const { return false; }
  // the texture that replaces the material's colour, if any
  virtual std::shared_ptr<Texture2D> texture() const { return nullptr; }

 protected:
  static pt_material_desc make(uint32_t kind, std::initializer_list<float> p) {
    pt_material_desc d{};
    d.kind = kind;
    int i = 0;
    for (float x : p) d.p[i++] = x;
    return d;
  }
};

// material.hpp:79-137
class Lambertian : public Material {
 public:
  Lambertian(const ColorVariant &a) { std::tie(albedo, tex) = ptamd::color_or_texture(a); }
  pt_material_desc desc() const override { return make(PT_MAT_LAMBERTIAN, {albedo[0], albedo[1], albedo[2]}); }
  std::shared_ptr<Texture2D> texture() const override { return tex; }
  gl::vec3 albedo;
  std::shared_ptr<Texture2D> tex;
};
// material.hpp:467-506
class DiffuseEmitter : public Material {
 public:
  DiffuseEmitter(const ColorVariant &a, float intensity = 1.0f) : color(ptamd::color_of(a, "DiffuseEmitter")), intensity(intensity) {}
  pt_material_desc desc() const override {
    return make(PT_MAT_EMITTER, {color[0], color[1], color[2], intensity});
  }
  bool is_emitter() const override { return true; }
  gl::vec3 color;
  float intensity;
};
// conductor.hpp:4-145
class Conductor : public Material {
 public:
  Conductor(const gl::vec3 &eta, const gl::vec3 &k, float alpha_x, float alpha_y)
      : eta(eta), k(k), ax(alpha_x), ay(alpha_y) {}
  pt_material_desc desc() const override {
    return make(PT_MAT_CONDUCTOR, {eta[0], eta[1], eta[2], k[0], k[1], k[2], ax, ay});
  }
  gl::vec3 eta, k;
  float ax, ay;
};
// dielectric.hpp:5-167
class MFDielectric : public Material {
 public:
  MFDielectric(float eta, float alpha_x, float alpha_y) : eta(eta), ax(alpha_x), ay(alpha_y) {}
  pt_material_desc desc() const override { return make(PT_MAT_MFDIELECTRIC, {eta, ax, ay}); }
  float eta, ax, ay;
};
// thinDielectric.hpp:3-70
class ThinDielectric : public Material {
 public:
  ThinDielectric(float eta, bool use_split_ray = false) : eta(eta), split(use_split_ray) {}
  pt_material_desc desc() const override { return make(PT_MAT_THINDIELECTRIC, {eta, split ? 1.f : 0.f}); }
  float eta;
  bool split;
};
// material.hpp:303-465
class Phong : public Material {
 public:
  Phong(const gl::vec3 &diffuse, const gl::vec3 &specular, const gl::vec3 &ambient, float shininess)
      : diffuse(diffuse), specular(specular), ambient(ambient), shininess(shininess) {}
  pt_material_desc desc() const override {
    return make(PT_MAT_PHONG, {diffuse[0], diffuse[1], diffuse[2], specular[0], specular[1], specular[2], ambient[0],
                               ambient[1], ambient[2], shininess});
  }
  gl::vec3 diffuse, specular, ambient;
  float shininess;
};
// disneyPrincipledBSDF.hpp:34-63 (constructor argument order)
class DisneyPrincipledBSDF : public Material {
 public:
  DisneyPrincipledBSDF(const gl::vec3 &baseColor, float specular_transmission, float metallic, float subsurface,
                       float specular, float roughness, float specular_tint, float anisotropic, float sheen,
                       float sheen_tint, float clearcoat, float clearcoat_gloss, float eta)
      : p{baseColor[0], baseColor[1], baseColor[2], specular_transmission, metallic, subsurface, specular,
          roughness, specular_tint, anisotropic, sheen, sheen_tint, clearcoat, clearcoat_gloss, eta} {}
  // beyond the reference's constructor: a textured base colour (as its
  // Disney lobes take one, disneyDiffuse.hpp:11-19, disneyMetal.hpp:12-21)
  DisneyPrincipledBSDF(std::shared_ptr<Texture2D> baseColor, float specular_transmission, float metallic,
                       float subsurface, float specular, float roughness, float specular_tint, float anisotropic,
                       float sheen, float sheen_tint, float clearcoat, float clearcoat_gloss, float eta)
      : p{0.f, 0.f, 0.f, specular_transmission, metallic, subsurface, specular,
          roughness, specular_tint, anisotropic, sheen, sheen_tint, clearcoat, clearcoat_gloss, eta} {
    gl::vec3 c;
    std::tie(c, tex) = ptamd::color_or_texture(baseColor);
    for (int i = 0; i < 3; ++i) p[i] = c[i];
  }
  pt_material_desc desc() const override {
    pt_material_desc d{};
    d.kind = PT_MAT_DISNEY;
    for (int i = 0; i < 15; ++i) d.p[i] = p[i];
    return d;
  }
  std::shared_ptr<Texture2D> texture() const override { return tex; }
  float p[15];
  std::shared_ptr<Texture2D> tex;
};
// hairMarsch.hpp:7-9
class HairMarschner : public Material {
 public:
  HairMarschner(const gl::vec3 &sigma_a, float h, float eta, float beta_m, float beta_n, float alpha)
      : sigma_a(sigma_a), h(h), eta(eta), beta_m(beta_m), beta_n(beta_n), alpha(alpha) {}
  pt_material_desc desc() const override {
    return make(PT_MAT_HAIR, {sigma_a[0], sigma_a[1], sigma_a[2], h, eta, beta_m, beta_n, alpha});
  }
  gl::vec3 sigma_a;
  float h, eta, beta_m, beta_n, alpha;
};
// No class in the reference (README.md:74 names the model): diffuse +
// specular around the tangent, exponent p.
class KajiyaKay : public Material {
 public:
  KajiyaKay(const gl::vec3 &diffuse, const gl::vec3 &specular, float exponent)
      : diffuse(diffuse), specular(specular), exponent(exponent) {}
  pt_material_desc desc() const override {
    return make(PT_MAT_KAJIYA_KAY, {diffuse[0], diffuse[1], diffuse[2], specular[0], specular[1], specular[2],
                                    exponent});
  }
  gl::vec3 diffuse, specular;
  float exponent;
};

// ------------------------------------------------------------------ lowering state
namespace ptamd {
struct Lowering {
  SceneSink &sink;
  std::map<const Material *, uint32_t> ids;
  std::map<const Texture2D *, uint32_t> tex_ids;
  explicit Lowering(SceneSink &s) : sink(s) {}
  uint32_t tex(const std::shared_ptr<Texture2D> &t) {
    auto it = tex_ids.find(t.get());
    if (it != tex_ids.end()) return it->second;
    const uint32_t id = t->lower(sink);
    tex_ids[t.get()] = id;
    return id;
  }
  uint32_t mat(const std::shared_ptr<Material> &m) {
    if (!m) throw std::runtime_error("object without a material");
    auto it = ids.find(m.get());
    if (it != ids.end()) return it->second;
    pt_material_desc d = m->desc();
    if (auto t = m->texture()) d.flags = PT_MAT_TEXTURE(tex(t));
    uint32_t id = sink.material(d);
    ids[m.get()] = id;
    return id;
  }
};
}  // namespace ptamd

// ------------------------------------------------------------------ geometry
class Hittable {
 public:
  virtual ~Hittable() = default;
  // Lowers this object with the object-to-world map X and wrapper chain C.
  // `light_ok`: an emissive rect/sphere reached here may be discovered as a light.
  virtual void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool light_ok) const = 0;
  virtual std::shared_ptr<Material> get_material() const { return nullptr; }
};

namespace ptamd {
inline uint32_t light_flags(const std::shared_ptr<Material> &m, bool light_ok) {
  return (m && m->is_emitter() && !light_ok) ? (uint32_t)PT_PRIM_NOT_LIGHT : 0u;
}
}  // namespace ptamd

// primitive.hpp:81-203
class Sphere : public Hittable {
 public:
  Sphere(const gl::vec3 &center, float radius) : center(center), radius(radius) {}
  Sphere(const gl::vec3 &center, float radius, std::shared_ptr<Material> material)
      : center(center), radius(radius), material(std::move(material)) {}
  void setMaterial(std::shared_ptr<Material> m) { material = std::move(m); }
  std::shared_ptr<Material> get_material() const override { return material; }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool light_ok) const override {
    if (!(X.scale > 0.f)) throw std::runtime_error("Scale must be positive");
    gl::vec3 c = X.point(center);
    float cc[3] = {c[0], c[1], c[2]};
    L.sink.sphere(cc, radius * X.scale, L.mat(material), C.flags() | ptamd::light_flags(material, light_ok));
  }
  gl::vec3 center;
  float radius;
  std::shared_ptr<Material> material;
};

// Curve (primitives/curve.hpp:6-140): a cubic Bezier hit as 8 cylinder
// segments; every type intersects as a cylinder, as in the reference.
enum class CurveType { Flat, Cylinder, Ribbon };
class Curve : public Hittable {
 public:
  Curve(const std::array<gl::vec3, 4> &cps, float width0, float width1, CurveType type = CurveType::Flat,
        std::shared_ptr<Material> material = nullptr)
      : cps(cps), width0(width0), width1(width1), type(type), material(std::move(material)) {}
  std::shared_ptr<Material> get_material() const override { return material; }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool light_ok) const override {
    (void)light_ok;  // curves are never lights ("for now, we don't treat them as emitters", curve.hpp:133)
    if (!(X.scale > 0.f)) throw std::runtime_error("Scale must be positive");
    float p[12];
    for (int i = 0; i < 4; ++i) {
      gl::vec3 q = X.point(cps[i]);
      p[3 * i] = q[0];
      p[3 * i + 1] = q[1];
      p[3 * i + 2] = q[2];
    }
    L.sink.curve(p, width0 * X.scale, width1 * X.scale, (uint32_t)type, L.mat(material), C.flags());
  }
  std::array<gl::vec3, 4> cps;
  float width0, width1;
  CurveType type;
  std::shared_ptr<Material> material;
};

// primitive.hpp:205-273; d0/d1 are the two other axes in increasing order
template <Axis axis>
class AARectangle : public Hittable {
 public:
  AARectangle(float k, float d0_min, float d0_max, float d1_min, float d1_max)
      : _d0_min(d0_min), _d0_max(d0_max), _d1_min(d1_min), _d1_max(d1_max), _k(k) {}
  AARectangle(float k, float d0_min, float d0_max, float d1_min, float d1_max, std::shared_ptr<Material> material)
      : _d0_min(d0_min), _d0_max(d0_max), _d1_min(d1_min), _d1_max(d1_max), _k(k), material(std::move(material)) {}
  std::shared_ptr<Material> get_material() const override { return material; }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool light_ok) const override {
    if (!(X.scale > 0.f)) throw std::runtime_error("Scale must be positive");
    const int a = (int)axis, b0 = (a == 0) ? 1 : 0, b1 = (a == 2) ? 1 : 2;
    const uint32_t mat = L.mat(material);
    const uint32_t flags = C.flags() | ptamd::light_flags(material, light_ok);
    if (X.identity_rotation) {
      const float s = X.scale;
      L.sink.rect(a, _k * s + X.t[a], _d0_min * s + X.t[b0], _d0_max * s + X.t[b0], _d1_min * s + X.t[b1],
                  _d1_max * s + X.t[b1], mat, flags);
      return;
    }
    // a rotated rect: two triangles over the same corners
    std::vector<float> pos;
    const float d0[2] = {_d0_min, _d0_max}, d1[2] = {_d1_min, _d1_max};
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        gl::vec3 p;
        p[a] = _k;
        p[b0] = d0[i];
        p[b1] = d1[j];
        gl::vec3 w = X.point(p);
        pos.insert(pos.end(), {w[0], w[1], w[2]});
      }
    std::vector<int32_t> idx = {0, 1, 3, 0, 3, 2};
    L.sink.mesh(pos, idx, nullptr, nullptr, mat, flags | (uint32_t)PT_PRIM_NOT_LIGHT);
    // the rect's own texture coordinates (primitive.cpp:55-57), affine over
    // the corners, so that interpolating them gives the rect's (u, v)
    if (material && material->texture()) L.sink.mesh_uv({0.f, 0.f, 1.f, 0.f, 0.f, 1.f, 1.f, 1.f}, idx);
  }
  float _d0_min, _d0_max, _d1_min, _d1_max, _k;
  std::shared_ptr<Material> material;
};
using XYRectangle = AARectangle<Axis::Z>;
using XZRectangle = AARectangle<Axis::Y>;
using YZRectangle = AARectangle<Axis::X>;

// mesh_io/meshData.hpp:6-13
struct MeshData {
  std::vector<gl::vec3> positions;
  std::vector<gl::vec3> normals;
  std::vector<gl::vec2> uvs;
  std::vector<std::array<int, 3>> indices;
  std::vector<std::array<int, 3>> normalIndices;
  std::vector<std::array<int, 3>> uvIndices;
};

// primitives/triangleMesh.cpp:5-56.  Shading normals are used when every
// face has three valid normal indices.
class TriangleMesh : public Hittable {
 public:
  TriangleMesh(MeshData &&data, std::shared_ptr<Material> mat) : mesh(std::move(data)), material(std::move(mat)) {}
  std::shared_ptr<Material> get_material() const override { return material; }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool) const override {
    if (!(X.scale > 0.f)) throw std::runtime_error("Scale must be positive");
    std::vector<float> pos;
    pos.reserve(mesh.positions.size() * 3);
    for (const auto &p : mesh.positions) {
      gl::vec3 w = X.point(p);
      pos.insert(pos.end(), {w[0], w[1], w[2]});
    }
    std::vector<int32_t> idx;
    idx.reserve(mesh.indices.size() * 3);
    for (const auto &f : mesh.indices) idx.insert(idx.end(), {f[0], f[1], f[2]});
    bool smooth = !mesh.normals.empty() && mesh.normalIndices.size() == mesh.indices.size();
    if (smooth)
      for (const auto &f : mesh.normalIndices)
        for (int c = 0; c < 3; ++c)
          if (f[c] < 0 || (size_t)f[c] >= mesh.normals.size()) smooth = false;
    const uint32_t mat = L.mat(material);
    // meshes are never discovered as lights (lightDiscovery.hpp:62-66)
    const uint32_t flags = C.flags() | (uint32_t)PT_PRIM_NOT_LIGHT;
    if (!smooth) {
      L.sink.mesh(pos, idx, nullptr, nullptr, mat, flags);
    } else {
      std::vector<float> nrm;
      for (const auto &n : mesh.normals) {
        gl::vec3 w = X.dir(n);
        nrm.insert(nrm.end(), {w[0], w[1], w[2]});
      }
      std::vector<int32_t> nidx;
      for (const auto &f : mesh.normalIndices) nidx.insert(nidx.end(), {f[0], f[1], f[2]});
      L.sink.mesh(pos, idx, &nrm, &nidx, mat, flags);
    }
    // texture coordinates when every face has three valid uv indices
    // (meshBVH.cpp:133-138 interpolates them); else the planar fallback
    bool has_uv = !mesh.uvs.empty() && mesh.uvIndices.size() == mesh.indices.size();
    if (has_uv)
      for (const auto &f : mesh.uvIndices)
        for (int c = 0; c < 3; ++c)
          if (f[c] < 0 || (size_t)f[c] >= mesh.uvs.size()) has_uv = false;
    if (!has_uv) {
      if ((!X.identity_rotation || X.scale != 1.f) && material && material->texture()) planar_uv(L);
      return;
    }
    std::vector<float> uv;
    uv.reserve(mesh.uvs.size() * 2);
    for (const auto &t : mesh.uvs) uv.insert(uv.end(), {t[0], t[1]});
    std::vector<int32_t> uvidx;
    uvidx.reserve(mesh.uvIndices.size() * 3);
    for (const auto &f : mesh.uvIndices) uvidx.insert(uvidx.end(), {f[0], f[1], f[2]});
    L.sink.mesh_uv(uv, uvidx);
  }
  // The planar fallback is taken in object space, over the object's box
  // (triangleMesh.cpp:10-18, 38-51).  Without rotation or scale the
  // library's world-space fallback is the same map; under them, the object
  // space coordinates of each face's corners are passed as the mesh's uvs:
  // the fallback is affine over a face whose normal picks the axes, so
  // interpolating them reproduces it (the face normal stands in for the
  // interpolated shading normal of a smooth mesh).
  void planar_uv(ptamd::Lowering &L) const {
    gl::vec3 mn(INFINITY), mx(-INFINITY);
    for (const auto &v : mesh.positions)
      for (int c = 0; c < 3; ++c) {
        mn[c] = std::fmin(mn[c], v[c]);
        mx[c] = std::fmax(mx[c], v[c]);
      }
    for (int c = 0; c < 3; ++c) {
      mn[c] -= 1e-3f;  // gl::epsilon (constants.hpp:7)
      mx[c] += 1e-3f;
    }
    std::vector<float> uv;
    std::vector<int32_t> uvidx;
    uv.reserve(mesh.indices.size() * 6);
    uvidx.reserve(mesh.indices.size() * 3);
    for (const auto &f : mesh.indices) {
      const gl::vec3 &p0 = mesh.positions[f[0]], &p1 = mesh.positions[f[1]], &p2 = mesh.positions[f[2]];
      gl::vec3 e1 = p1 - p0, e2 = p2 - p0, n;
      n[0] = std::fabs(e1[1] * e2[2] - e1[2] * e2[1]);
      n[1] = std::fabs(e1[2] * e2[0] - e1[0] * e2[2]);
      n[2] = std::fabs(e1[0] * e2[1] - e1[1] * e2[0]);
      const int au = (n[0] > n[1] && n[0] > n[2]) ? 1 : 0;
      const int av = au == 1 ? 2 : (n[1] > n[2] ? 2 : 1);
      for (int c = 0; c < 3; ++c) {
        const gl::vec3 &q = mesh.positions[f[c]];
        uvidx.push_back((int32_t)(uv.size() / 2));
        uv.push_back((q[au] - mn[au]) / (mx[au] - mn[au]));
        uv.push_back((q[av] - mn[av]) / (mx[av] - mn[av]));
      }
    }
    L.sink.mesh_uv(uv, uvidx);
  }
  MeshData mesh;
  std::shared_ptr<Material> material;
};

// base/objectList.hpp:9-105
class ObjectList : public Hittable {
 public:
  ObjectList() = default;
  void addObject(std::shared_ptr<Hittable> object) { objects.push_back(std::move(object)); }
  void clear() { objects.clear(); }
  std::vector<std::shared_ptr<Hittable>> getLists() const { return objects; }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool) const override {
    for (const auto &o : objects) o->lower(L, X, C, false);  // nested lists hold no lights
  }
  std::vector<std::shared_ptr<Hittable>> objects;
};

// primitives/box.hpp:5-53: six rects with the box's material
class Box : public Hittable {
 public:
  Box(gl::vec3 min_xyz, gl::vec3 max_xyz, std::shared_ptr<Material> material = nullptr)
      : min_xyz(min_xyz), max_xyz(max_xyz), material(std::move(material)) {}
  std::shared_ptr<Material> get_material() const override { return material; }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool) const override {
    const float x0 = min_xyz[0], x1 = max_xyz[0], y0 = min_xyz[1], y1 = max_xyz[1], z0 = min_xyz[2],
                z1 = max_xyz[2];
    YZRectangle(x1, y0, y1, z0, z1, material).lower(L, X, C, false);
    YZRectangle(x0, y0, y1, z0, z1, material).lower(L, X, C, false);
    XZRectangle(y1, x0, x1, z0, z1, material).lower(L, X, C, false);
    XZRectangle(y0, x0, x1, z0, z1, material).lower(L, X, C, false);
    XYRectangle(z1, x0, x1, y0, y1, material).lower(L, X, C, false);
    XYRectangle(z0, x0, x1, y0, y1, material).lower(L, X, C, false);
  }
  gl::vec3 min_xyz, max_xyz;
  std::shared_ptr<Material> material;
};

// utils/objectTransform.hpp:5-181
class FlipFace : public Hittable {
 public:
  FlipFace(std::shared_ptr<Hittable> object) : object(std::move(object)) {}
  std::shared_ptr<Material> get_material() const override { return object->get_material(); }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool light_ok) const override {
    ptamd::Chain c2{C.w + "F"};
    object->lower(L, X, c2, light_ok);
  }
  std::shared_ptr<Hittable> object;
};
class Translate : public Hittable {
 public:
  Translate(std::shared_ptr<Hittable> object, const gl::vec3 &offset) : object(std::move(object)), offset(offset) {}
  std::shared_ptr<Material> get_material() const override { return object->get_material(); }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool light_ok) const override {
    ptamd::Xform inner;
    for (int i = 0; i < 3; ++i) inner.t[i] = offset[i];
    object->lower(L, X.then_outer_of(inner), ptamd::Chain{C.w + "T"}, light_ok);
  }
  std::shared_ptr<Hittable> object;
  gl::vec3 offset;
};
template <Axis axis>
class Rotate : public Hittable {
 public:
  Rotate(std::shared_ptr<Hittable> object, float angle)
      : object(std::move(object)), angle(angle), sin_theta(std::sin(angle)), cos_theta(std::cos(angle)) {}
  std::shared_ptr<Material> get_material() const override { return object->get_material(); }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool light_ok) const override {
    // the reference rotates rays by R (objectTransform.cpp:93-213): for Y,
    // x' = c x - s z, z' = s x + c z; geometry therefore maps by R^T
    const float c = cos_theta, s = sin_theta;
    int i = 0, j = 0;
    if (axis == Axis::X) { i = 1; j = 2; }
    if (axis == Axis::Y) { i = 0; j = 2; }
    if (axis == Axis::Z) { i = 0; j = 1; }
    ptamd::Xform inner;
    inner.m[i][i] = c;
    inner.m[i][j] = s;
    inner.m[j][i] = -s;
    inner.m[j][j] = c;
    inner.identity_rotation = (s == 0.f && c == 1.f);
    object->lower(L, X.then_outer_of(inner), ptamd::Chain{C.w + "T"}, light_ok);
  }
  std::shared_ptr<Hittable> object;
  float angle, sin_theta, cos_theta;
};
class Scale : public Hittable {
 public:
  Scale(std::shared_ptr<Hittable> obj, float s) : object(std::move(obj)), scale(s) {}
  std::shared_ptr<Material> get_material() const override { return object->get_material(); }
  void lower(ptamd::Lowering &L, const ptamd::Xform &X, const ptamd::Chain &C, bool light_ok) const override {
    ptamd::Xform inner;
    for (int i = 0; i < 3; ++i) inner.m[i][i] = scale;
    inner.scale = scale;
    object->lower(L, X.then_outer_of(inner), ptamd::Chain{C.w + "T"}, light_ok);
  }
  std::shared_ptr<Hittable> object;
  float scale;
};

// mesh_io/meshLoader.cpp:7-58: positions, normals, uvs and triangular faces
// (any other face size is an error) with per-corner indices.
inline std::shared_ptr<TriangleMesh> loadOBJMesh(const std::string &path, std::shared_ptr<Material> mat) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  MeshData d;
  std::string line;
  auto fix = [](long i, size_t n) -> int {  // 1-based, negative = relative to the end
    if (i > 0) return (int)(i - 1);
    if (i < 0) return (int)((long)n + i);
    return -1;
  };
  size_t lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    std::istringstream is(line);
    std::string tag;
    if (!(is >> tag) || tag[0] == '#') continue;
    if (tag == "v") {
      float x = 0, y = 0, z = 0;
      is >> x >> y >> z;
      d.positions.emplace_back(x, y, z);
    } else if (tag == "vn") {
      float x = 0, y = 0, z = 0;
      is >> x >> y >> z;
      d.normals.emplace_back(x, y, z);
    } else if (tag == "vt") {
      float u = 0, v = 0;
      is >> u >> v;
      d.uvs.emplace_back(u, v);
    } else if (tag == "f") {
      std::array<int, 3> vi{}, ti{}, ni{};
      std::string tok;
      int k = 0;
      while (is >> tok) {
        if (k == 3) throw std::runtime_error(path + ":" + std::to_string(lineno) + ": non-triangle face in OBJ");
        long a = 0, b = 0, c = 0;
        size_t s1 = tok.find('/');
        a = std::stol(tok.substr(0, s1));
        if (s1 != std::string::npos) {
          size_t s2 = tok.find('/', s1 + 1);
          std::string t = tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
          if (!t.empty()) b = std::stol(t);
          if (s2 != std::string::npos && s2 + 1 < tok.size()) c = std::stol(tok.substr(s2 + 1));
        }
        vi[k] = fix(a, d.positions.size());
        ti[k] = fix(b, d.uvs.size());
        ni[k] = fix(c, d.normals.size());
        ++k;
      }
      if (k != 3) throw std::runtime_error(path + ":" + std::to_string(lineno) + ": non-triangle face in OBJ");
      for (int c = 0; c < 3; ++c)
        if (vi[c] < 0 || (size_t)vi[c] >= d.positions.size())
          throw std::runtime_error(path + ":" + std::to_string(lineno) + ": vertex index out of range");
      d.indices.push_back(vi);
      d.uvIndices.push_back(ti);
      d.normalIndices.push_back(ni);
    }
  }
  return std::make_shared<TriangleMesh>(std::move(d), std::move(mat));
}

// ------------------------------------------------------------------ camera, environment
// base/camera.hpp:16-35, camera.cpp:7-56.  fov in radians; `aperture` with
// the reference's meaning (lens half-width focalDist/aperture/2).
class PerspectiveCamera {
 public:
  PerspectiveCamera(float fov, float aspect, float focalDist, float aperture, gl::vec3 up, gl::vec3 front,
                    gl::vec3 position)
      : fov(fov), aspect(aspect), focalDist(focalDist), aperture(aperture), up(up), front(front), position(position),
        position1(position) {}
  pt_camera_desc desc() const {
    pt_camera_desc c{};
    c.fov = fov;
    c.aspect = aspect;
    c.focal_dist = focalDist;
    c.aperture = aperture;
    for (int i = 0; i < 3; ++i) {
      c.up[i] = up[i];
      c.front[i] = front[i];
      c.pos[i] = position[i];
      c.pos1[i] = position1[i];
    }
    c.shutter = shutter;
    c.normalize_dir = normalize_directions ? 1u : 0u;
    return c;
  }
  float fov, aspect, focalDist, aperture;
  gl::vec3 up, front, position;
  // beyond the reference: motion blur (position at shutter close) and the
  // unit-length camera directions that keep t in one unit for every shape
  gl::vec3 position1;
  float shutter = 0.f;
  bool normalize_directions = true;
};

// light/envLight.hpp:4-92 with HDRITexture (texture.hpp:97-132)
class EnvironmentLight {
 public:
  EnvironmentLight(const std::string &hdri_filepath, float intensity_scale = 1.0f) : scale(intensity_scale) {
    ptamd::image_io::Image img = ptamd::image_io::read_env(hdri_filepath);
    width = img.width;
    height = img.height;
    rgb = std::move(img.rgb);
  }
  // an in-memory map in the pt_scene_set_env layout
  EnvironmentLight(std::vector<float> texels, uint32_t w, uint32_t h, float intensity_scale = 1.0f)
      : rgb(std::move(texels)), width(w), height(h), scale(intensity_scale) {}
  std::vector<float> rgb;
  uint32_t width = 0, height = 0;
  float scale = 1.f;
};

// ------------------------------------------------------------------ SceneInfo
// render/renderManager.hpp:35-158.  The render settings the reference takes
// from config.hpp (MAX_RAY_DEPTH 10, 2x2 light samples, GAMMA 2.0, the
// maxdepth_nee integrator, the Halton sampler) are fields here.
// Medium / HomogeneousMedium(sigma_a, sigma_s, sigma_scale, Le, Le_scale, g)
// (medium/homogeneous.hpp:31-68): the scene-filling medium the volume
// estimator (PT_EST_VOLUME, the reference's USE_MAXDEPTH_VOLUME) renders.
class Medium {
 public:
  virtual ~Medium() = default;
  gl::vec3 sigma_a, sigma_s, Le;
  float g = 0.f;
};
class HomogeneousMedium : public Medium {
 public:
  HomogeneousMedium(gl::vec3 sa, gl::vec3 ss, float sigma_scale, gl::vec3 le, float le_scale, float g_) {
    sigma_a = sa * sigma_scale;
    sigma_s = ss * sigma_scale;
    Le = le * le_scale;
    g = g_;
  }
};

struct SceneInfo {
  std::shared_ptr<PerspectiveCamera> camera = nullptr;
  std::shared_ptr<Medium> global_medium = nullptr;  // renderManager.hpp:48
  ObjectList objects;
  std::shared_ptr<EnvironmentLight> environment_light = nullptr;
  gl::vec3 bg_color = gl::vec3(0.7f, 0.8f, 1.0f);
  bool use_bvh = true;
  unsigned _width = 800;
  unsigned _height = 800;
  unsigned spp_x = 2;
  unsigned spp_y = 2;
  float _gamma = 1.0f;
  // config.hpp settings
  unsigned max_depth = 10;
  unsigned light_samples = 4;
  float write_gamma = 2.0f;
  uint32_t sampler = PT_SAMPLER_HALTON_OWEN;
  uint32_t estimator = PT_EST_REFERENCE_NEE;
  uint32_t rr_depth = 0;
  uint32_t seed = 0;

  SceneInfo() = default;

  // Lowers the scene into `sink` (materials, then objects in order).
  void lower(ptamd::SceneSink &sink) const {
    ptamd::Lowering L(sink);
    for (const auto &o : objects.objects) o->lower(L, ptamd::Xform{}, ptamd::Chain{}, true);
  }

  pt_render_cfg render_cfg() const {
    pt_render_cfg c{};
    c.width = _width;
    c.height = _height;
    c.spp_x = spp_x;
    c.spp_y = spp_y;
    c.max_depth = max_depth;
    c.sampler = sampler;
    c.estimator = estimator;
    c.rr_depth = rr_depth;
    c.light_samples = light_samples;
    c.seed = seed;
    for (int i = 0; i < 3; ++i) c.bg[i] = bg_color[i];
    return c;
  }

  // Builds the device scene.  The caller owns the returned handle.
  pt_scene *build() const {
    if (!camera) throw std::runtime_error("Camera is not initialized!");  // renderManager.hpp:58-62
    if (objects.objects.empty()) throw std::runtime_error("No objects in the scene!");
    pt_scene *s = nullptr;
    ptamd::check(pt_scene_create(&s), "pt_scene_create");
    try {
      ptamd::AbiSink sink(s);
      lower(sink);
      if (environment_light)
        ptamd::check(pt_scene_set_env(s, environment_light->rgb.data(), environment_light->width,
                                      environment_light->height, environment_light->scale),
                     "pt_scene_set_env");
      pt_camera_desc cd = camera->desc();
      ptamd::check(pt_scene_set_camera(s, &cd), "pt_scene_set_camera");
      if (global_medium) {
        const float sa[3] = {global_medium->sigma_a[0], global_medium->sigma_a[1], global_medium->sigma_a[2]};
        const float ss[3] = {global_medium->sigma_s[0], global_medium->sigma_s[1], global_medium->sigma_s[2]};
        const float le[3] = {global_medium->Le[0], global_medium->Le[1], global_medium->Le[2]};
        ptamd::check(pt_scene_set_medium(s, sa, ss, 1.f, le, 1.f, global_medium->g), "pt_scene_set_medium");
      }
      ptamd::check(pt_scene_commit(s, nullptr), "pt_scene_commit");
    } catch (...) {
      pt_scene_destroy(s);
      throw;
    }
    return s;
  }

  // Renders the frame and writes it as PNG (gamma `write_gamma`, vertical
  // flip, FrameBuffer::writeToFile); a path ending in .pfm gets the FP32
  // film instead.  Returns the film (height x width x 3, row 0 at the bottom).
  std::vector<float> render(const std::string &out_path = "./output.png", bool show_progress = true) {
    pt_scene *s = build();
    std::vector<float> film((size_t)_width * _height * 3);
    try {
      pt_render_cfg cfg = render_cfg();
      pt_stats st{};
      ptamd::check(pt_render(s, &cfg, &st), "pt_render");
      ptamd::check(pt_film_read(s, film.data()), "pt_film_read");
      if (show_progress)
        std::printf("rendered %ux%u, %u spp in %.1f ms (%llu path samples)\n", _width, _height, spp_x * spp_y,
                    st.ms_total, (unsigned long long)st.paths);
    } catch (...) {
      pt_scene_destroy(s);
      throw;
    }
    pt_scene_destroy(s);
    if (!out_path.empty()) {
      if (out_path.size() > 4 && out_path.substr(out_path.size() - 4) == ".pfm")
        ptamd::image_io::write_pfm(out_path, film.data(), _width, _height);
      else
        ptamd::image_io::write_png(out_path, film.data(), _width, _height, write_gamma);
    }
    return film;
  }
};
