// Host-side sanitizer driver for the scene loaders, generators and builders
// (built with g++ -fsanitize=address,undefined by tests/test_host_sanitizers.py).
//   host_driver gen <config>      generate, build, view and export the wide tree
//   host_driver load <files...>   load each file as a model (.obj) or texture
// Loaders must refuse malformed files with an error code, never read or write
// out of bounds.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cudatracer.h"

// The frame defaults live in the HIP runtime, which this host-only build
// leaves out; the generators only need a value to fill in.
extern "C" void ct_frame_opts_default(ct_frame_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->max_depth = 5;
    o->spp = 1;
}

static int gen(int config) {
    ct_host_scene* s = nullptr;
    if (ct_host_scene_generate(config, 7, &s) != 0) return 3;
    int r = ct_host_scene_build(s);
    ct_scene_view v;
    r |= ct_host_scene_view(s, &v);
    uint32_t nn = 0, bound = 0;
    r |= ct_build_wide_nodes(&v, nullptr, 0, &nn, &bound);
    std::vector<uint32_t> w((size_t)nn * 32);
    r |= ct_build_wide_nodes(&v, w.data(), nn, &nn, &bound);
    std::printf("config %d: rc %d, %u wide nodes, stack bound %u\n", config, r, nn, bound);
    ct_host_scene_destroy(s);
    return r ? 4 : 0;
}

int main(int argc, char** argv) {
    if (argc >= 3 && std::string(argv[1]) == "gen") return gen(std::atoi(argv[2]));
    if (argc < 2 || std::string(argv[1]) != "load") return 2;
    int ok = 0, refused = 0;
    for (int i = 2; i < argc; ++i) {
        const std::string p = argv[i];
        ct_host_scene* s = nullptr;
        if (ct_host_scene_create(&s) != 0) return 3;
        int r;
        if (p.size() > 4 && p.compare(p.size() - 4, 4, ".obj") == 0) {
            r = ct_host_scene_load_model(s, p.c_str(), 1.0f, nullptr, -1, i & 1);
            if (r == 0 && ct_host_scene_build(s) == 0) {
                ct_scene_view v;
                (void)ct_host_scene_view(s, &v);
            }
        } else {
            r = ct_host_scene_load_texture(s, p.c_str());
        }
        (r == 0 ? ok : refused)++;
        ct_host_scene_destroy(s);
    }
    std::printf("loaded %d refused %d\n", ok, refused);
    return 0;
}
