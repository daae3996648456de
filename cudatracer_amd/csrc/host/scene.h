// Host scene classes with the reference's names and semantics:
//   Material      src/Core/Material.cuh:29-97
//   TriangleList  src/Core/TriangleList.h:23-84, TriangleList.cpp:8-61
//   Camera        src/Core/Camera.h:12-93, Camera.cpp:9-37
// The device side never sees these types: ct_scene_upload takes a
// ct_scene_view of their arrays.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../../include/cudatracer.h"
#include "ct_vec.h"

namespace ct {

enum MaterialType {
    None = CT_MAT_NONE,
    Light = CT_MAT_LIGHT,
    Specular = CT_MAT_SPECULAR,
    Fresnel = CT_MAT_FRESNEL,
    Lambertian = CT_MAT_LAMBERTIAN,
    Beckmann = CT_MAT_BECKMANN,
    GGX = CT_MAT_GGX,
    Trowbridge = CT_MAT_TROWBRIDGE,
    FresnelBeckmann = CT_MAT_FRESNEL_BECKMANN,
    FresnelGGX = CT_MAT_FRESNEL_GGX,
    FresnelTrowbridge = CT_MAT_FRESNEL_TROWBRIDGE,
};

struct Material {
    vec3 albedo;  // emission for lights
    float refractIdx = 1.0f;
    int diffuseTex = -1, normalTex = -1, maskTex = -1, displaceTex = -1;
    MaterialType type = Lambertian;
    float roughness = 0.0f;
    vec3 absorption = vec3(0.0f);

    bool is_light() const { return type == Light; }

    // Material::light, Material.cuh:50-60 (emission clamped to >= 0.01)
    static Material light(vec3 emission);
    // Material::lambertian, Material.cuh:62-74
    static Material lambertian(vec3 albedo, int diffuseTex = -1, int normalTex = -1, int maskTex = -1,
                               int displaceTex = -1, float refractionIdx = 1.0f);
    // Material::specular, Material.cuh:76-84
    static Material specular(vec3 albedo, float refractionIdx);
    // Material::fresnel, Material.cuh:86-97
    static Material fresnel(vec3 albedo, float refractIdx, vec3 absorption = vec3(0.0f), float roughness = 0.0f,
                            int diffuseTex = -1, int normalTex = -1);
    // Microfacet reflectors (types GGX / Beckmann); the reference has the
    // enum values but no factory.
    static Material microfacet(vec3 albedo, MaterialType t, float roughness);

    ct_material to_abi() const;
    static Material from_abi(const ct_material& m);
};

// core::Surface's float texel buffer (src/Utils/Surface.cpp:33-100): RGBA,
// row 0 at the bottom of the image (FreeImage scan-line order).
struct Texture {
    uint32_t width = 0, height = 0;
    std::vector<float> rgba;
};

// Axis-aligned box in the padded form of triangle::getBounds
// (Triangle.cuh:59-65): [min - 0.001, max + 0.01].
struct AABB {
    float bmin[3] = {1e34f, 1e34f, 1e34f};
    float bmax[3] = {-1e34f, -1e34f, -1e34f};
};

class TriangleList {
public:
    TriangleList();  // adds the default material 0 (white Lambert)

    void addTriangle(vec3 p0, vec3 p1, vec3 p2, vec3 n0, vec3 n1, vec3 n2, unsigned int matIdx,
                     vec2 t0 = vec2(), vec2 t1 = vec2(), vec2 t2 = vec2());
    // Forces roughness = 1 like the reference (TriangleList.cpp:56);
    // keepRoughness = true is the extension the config scenes use.
    unsigned int addMaterial(Material mat, bool keepRoughness = false);

    // TriangleList::loadModel (TriangleList.h:49, TriangleList.cpp:63-146)
    // for Wavefront OBJ files (host/objload.cpp).  mat: column-major 4x4
    // (glm layout), nullptr = identity.  material < 0: material 0 (MTL files
    // are not read).  Returns false with a message in *error on failure,
    // leaving the list unchanged.
    bool loadModel(const std::string& path, float scale = 1.0f, const float* mat = nullptr, int material = -1,
                   bool normalize = true, std::string* error = nullptr);

    // TriangleList::loadTexture (TriangleList.cpp:377-391) for binary PPM (P6)
    // and PFM files (host/texload.cpp); a path loaded before keeps its index.
    // Returns the texture index, or -1 with a message in *error.
    int loadTexture(const std::string& path, std::string* error = nullptr);
    // The same for a texture already in memory (width * height RGBA floats).
    int addTexture(uint32_t width, uint32_t height, const float* rgba);

    size_t getPrimCount() const { return m_Indices.size(); }

    std::vector<vec3> m_Vertices;
    std::vector<vec3> m_Normals;
    std::vector<vec3> m_CenterNormals;
    std::vector<vec2> m_TexCoords;
    std::vector<unsigned int> m_MaterialIdxs;
    std::vector<unsigned int> m_LightIndices;
    std::vector<Material> m_Materials;
    std::vector<AABB> m_AABBs;
    std::vector<uvec3> m_Indices;
    std::vector<Texture> m_Textures;
    std::map<std::string, int> m_LoadedTextures;
};

class Camera {
public:
    Camera() = default;
    // Camera.cpp:9-17: forward (0,0,-1), plane distance tan(radians(fov/2)).
    Camera(int width, int height, float fov, vec3 pos = vec3(0.f, 0.f, 1.f));

    void setDimensions(int width, int height);
    void lookAt(vec3 target);
    ct_camera to_abi() const;

    vec3 m_Up = vec3(0.0f, 1.0f, 0.0f);
    vec3 m_Forward = vec3(0.0f, 0.0f, -1.0f);
    vec3 m_Position = vec3(0.0f, 0.0f, 1.0f);
    float m_Fov = 70.0f;
    float m_PlaneDistance = 0.0f;
    float m_Width = 1, m_Height = 1;
    float m_InvWidth = 1, m_InvHeight = 1;
    float m_AspectRatio = 1;
};

// Vertex AABB with the reference padding (Triangle.cuh:59-65).
AABB triangle_bounds(vec3 p0, vec3 p1, vec3 p2);

}  // namespace ct
