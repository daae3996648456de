// Seeded procedural scenes for the benchmark configurations (C1 box+sphere,
// C2 atrium, C3 displaced sphere).  The reference builds its scene in main.cpp
// (main.cpp:42-74) from model files; these generators replace that step with
// deterministic geometry of the stated sizes.
#pragma once

#include <cstdint>

#include "../../../include/cudatracer.h"
#include "scene.h"

namespace ct {

struct SceneConfig {
    int config = 0;
    uint64_t seed = 0;
    vec3 cam_pos, cam_target;
    float cam_fov = 70.0f;
    int sky_enabled = 0;
    vec3 sky;
    int sky_texture = -1;  // SceneData::skyboxTexture
    ct_frame_opts opts;   // the configuration's depth / spp / accel / epsilons
    int width = 256, height = 256;
};

// splitmix64 stream (Steele, Lea & Flood 2014).
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    // uniform in [0, 1) with 24 random bits
    float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
};

// Generate config 1, 2 or 3 into an empty TriangleList.  Returns false for an
// unknown config.
bool generate_scene(int config, uint64_t seed, TriangleList& list, SceneConfig& cfg);

}  // namespace ct
