// psyref — TEST INFRASTRUCTURE ONLY (parity oracle / CPU baseline).
//
// Renderer: restates the light tree (src/accel/light_tree.rs), light
// selection (src/scene/mod.rs:48-121, src/scene/assembly.rs:46-141), the
// camera (src/camera.rs), the light-path state machine and bucket render
// loop (src/renderer.rs:194-692) and the XYZ film accumulation.
#pragma once

#include <map>
#include <vector>

#include "psyref_color.h"
#include "psyref_scene.h"
#include "psyref_shading.h"

namespace psyref {

// light_tree.rs: a binary SAH tree over light instances collapsed to arity 8.
struct LightTree {
    struct Node {
        bool leaf = true;
        std::vector<BBox> bounds;
        float energy = 0;
        size_t light_index = 0;
        std::vector<Node> children;
    };
    bool has_root = false;
    Node root;
    size_t depth = 0;

    // objects: keys permuted in place; info(key) -> (bounds, energy)
    void build(uint32_t* objects, size_t n, const std::function<BoundsSlice(uint32_t)>& bounds_of,
               const std::function<float(uint32_t)>& energy_of);
    // light_tree.rs:135-201.  Returns false when there is no light.
    bool select(Vec3 inc, Vec3 pos, Vec3 nor, Vec3 nor_g, const Closure& sc, float time, float n, size_t& index,
                float& prob, float& whittled_n) const;
    float approximate_energy() const { return has_root ? root.energy : 0.0f; }
};

struct World {
    Color background;
    std::vector<DistantDiskLight> lights;
};

// Camera (camera.rs:12-19): time samples of each lens parameter, each list
// interpolated over the shutter on its own.
struct CameraDesc {
    std::vector<Transform> transforms;         // camera to world
    std::vector<float> fovs{1.0f};             // radians
    std::vector<float> aperture_radii{0.0f};
    std::vector<float> focus_distances{1.0f};
};

// The camera after Camera::new's checks, with linear fovs.
struct Camera {
    std::vector<Transform> transforms;
    std::vector<float> tfovs, apertures, focuses;
};
Camera camera_new(const CameraDesc& d);
Ray generate_ray(const Camera& c, float x, float y, float time, float wavelength, float u, float v);

struct RenderSettings {
    uint32_t width = 64, height = 64, spp = 1, max_bounces = 2, seed = 0;
};

// The light accel of one assembly (assembly.rs:297-338): a light tree over
// its light instances and the instances of sub-assemblies that hold lights,
// in light-tree order.
struct AssemblyLights {
    LightTree tree;
    std::vector<Instance> instances;
};

// A scene ready to render: the root assembly with the light trees of it and
// its sub-assemblies, shaders, world and camera.
struct RenderScene {
    Assembly* root = nullptr;
    std::vector<Closure> shaders;  // SimpleSurfaceShader -> closure (shading/mod.rs:42-61)
    World world;
    CameraDesc camera;
    std::map<const Assembly*, AssemblyLights> lights;

    void build_lights();
    const AssemblyLights& lights_of(const Assembly* a) const { return lights.at(a); }
};

// Renders into xyz (width*height*3).  Returns the summed time of the trace
// calls and the ray count.
void render(const RenderScene& scene, const RenderSettings& rs, float* xyz, uint32_t threads, double& trace_seconds,
            uint64_t& rays);

// The per-path sample streams, exposed for tests.
float probit(float n, float width);

}  // namespace psyref
