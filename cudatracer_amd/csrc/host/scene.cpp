#include "scene.h"

#include <cmath>
#include <cstring>

namespace ct {

Material Material::light(vec3 emission) {
    Material mat;
    mat.refractIdx = 1.0f;
    mat.diffuseTex = -1;
    mat.normalTex = -1;
    mat.type = Light;
    mat.albedo = vmax(emission, vec3(0.01f));
    mat.absorption = vec3(0.0f);
    return mat;
}

Material Material::lambertian(vec3 albedo, int diffuseTex, int normalTex, int maskTex, int displaceTex,
                              float refractionIdx) {
    Material mat;
    mat.albedo = albedo;
    mat.refractIdx = refractionIdx;
    mat.diffuseTex = diffuseTex;
    mat.normalTex = normalTex;
    mat.maskTex = maskTex;
    mat.displaceTex = displaceTex;
    mat.type = Lambertian;
    mat.roughness = 0.0f;
    return mat;
}

Material Material::specular(vec3 albedo, float refractionIdx) {
    Material mat;
    mat.albedo = albedo;
    mat.refractIdx = refractionIdx;
    mat.type = Specular;
    mat.roughness = 0.0001f;
    return mat;
}

Material Material::fresnel(vec3 albedo, float refractIdx, vec3 absorption, float roughness, int diffuseTex,
                           int normalTex) {
    Material mat;
    mat.albedo = albedo;
    mat.refractIdx = refractIdx;
    mat.diffuseTex = diffuseTex;
    mat.normalTex = normalTex;
    mat.type = Fresnel;
    mat.absorption = absorption;
    mat.roughness = roughness;
    return mat;
}

Material Material::microfacet(vec3 albedo, MaterialType t, float roughness) {
    Material mat;
    mat.albedo = albedo;
    mat.type = t;
    mat.roughness = roughness;
    return mat;
}

ct_material Material::to_abi() const {
    ct_material m;
    std::memset(&m, 0, sizeof(m));
    m.albedo[0] = albedo.x;
    m.albedo[1] = albedo.y;
    m.albedo[2] = albedo.z;
    m.refract_idx = refractIdx;
    m.diffuse_tex = diffuseTex;
    m.normal_tex = normalTex;
    m.mask_tex = maskTex;
    m.displace_tex = displaceTex;
    m.type = (int32_t)type;
    m.roughness = roughness;
    m.absorption[0] = absorption.x;
    m.absorption[1] = absorption.y;
    m.absorption[2] = absorption.z;
    return m;
}

Material Material::from_abi(const ct_material& m) {
    Material r;
    r.albedo = vec3(m.albedo[0], m.albedo[1], m.albedo[2]);
    r.refractIdx = m.refract_idx;
    r.diffuseTex = m.diffuse_tex;
    r.normalTex = m.normal_tex;
    r.maskTex = m.mask_tex;
    r.displaceTex = m.displace_tex;
    r.type = (MaterialType)m.type;
    r.roughness = m.roughness;
    r.absorption = vec3(m.absorption[0], m.absorption[1], m.absorption[2]);
    return r;
}

AABB triangle_bounds(vec3 p0, vec3 p1, vec3 p2) {
    const vec3 mn = vmin(p0, vmin(p1, p2));
    const vec3 mx = vmax(p0, vmax(p1, p2));
    const vec3 lo = mn - 0.001f;
    const vec3 hi = mx + 0.01f;
    AABB b;
    b.bmin[0] = lo.x; b.bmin[1] = lo.y; b.bmin[2] = lo.z;
    b.bmax[0] = hi.x; b.bmax[1] = hi.y; b.bmax[2] = hi.z;
    return b;
}

TriangleList::TriangleList() {
    // TriangleList.cpp:8-13: "Have at least 1 material".  The reference passes
    // 1.0f as the second argument, the int diffuse texture index, so its
    // material 0 samples texture 1 (out of range without two textures);
    // here material 0 is the untextured white Lambert it was meant to be.
    addMaterial(Material::lambertian(vec3(1.0f)));
}

void TriangleList::addTriangle(vec3 p0, vec3 p1, vec3 p2, vec3 n0, vec3 n1, vec3 n2, unsigned int matIdx, vec2 t0,
                               vec2 t1, vec2 t2) {
    const uint32_t idx0 = (uint32_t)m_Vertices.size();
    m_Vertices.push_back(p0);
    m_Vertices.push_back(p1);
    m_Vertices.push_back(p2);
    m_Normals.push_back(normalize(n0));
    m_Normals.push_back(normalize(n1));
    m_Normals.push_back(normalize(n2));
    m_CenterNormals.push_back((n0 + n1 + n2) / 3.0f);
    m_TexCoords.push_back(t0);
    m_TexCoords.push_back(t1);
    m_TexCoords.push_back(t2);
    m_Indices.emplace_back(idx0, idx0 + 1, idx0 + 2);
    m_MaterialIdxs.push_back(matIdx);
    m_AABBs.push_back(triangle_bounds(p0, p1, p2));
    // Extension: the reference registers lights only in loadModel
    // (TriangleList.cpp:136-137), which leaves procedural lights invisible
    // to next-event estimation.
    if (matIdx < m_Materials.size() && m_Materials[matIdx].type == Light)
        m_LightIndices.push_back((unsigned int)(m_Indices.size() - 1));
}

unsigned int TriangleList::addMaterial(Material mat, bool keepRoughness) {
    const unsigned int m = (unsigned int)m_Materials.size();
    if (!keepRoughness) mat.roughness = 1.0f;
    m_Materials.push_back(mat);
    return m;
}

Camera::Camera(int width, int height, float fov, vec3 pos)
    : m_Position(pos), m_Fov(fov), m_Width((float)width), m_Height((float)height),
      m_InvWidth(1.f / (float)width), m_InvHeight(1.f / (float)height) {
    m_Forward = vec3(0, 0, -1);
    m_AspectRatio = m_Width / m_Height;
    // glm::radians(x) = x * 0.01745329251994329576923690768489f
    m_PlaneDistance = std::tan(fov * 0.5f * 0.01745329251994329576923690768489f);
}

void Camera::setDimensions(int width, int height) {
    m_Width = (float)width;
    m_Height = (float)height;
    m_InvWidth = 1.0f / (float)width;
    m_InvHeight = 1.0f / (float)height;
    m_AspectRatio = (float)width / (float)height;
}

void Camera::lookAt(vec3 target) { m_Forward = normalize(target - m_Position); }

ct_camera Camera::to_abi() const {
    ct_camera c;
    c.position[0] = m_Position.x; c.position[1] = m_Position.y; c.position[2] = m_Position.z;
    c.forward[0] = m_Forward.x; c.forward[1] = m_Forward.y; c.forward[2] = m_Forward.z;
    c.up[0] = m_Up.x; c.up[1] = m_Up.y; c.up[2] = m_Up.z;
    c.plane_distance = m_PlaneDistance;
    return c;
}

}  // namespace ct
