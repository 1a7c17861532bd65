// flatten.h — Scene -> rtg_scene_desc (the librtg boundary format).
#pragma once
#include "Scene.h"
#include "../../include/rtg.h"

// Flattened view of a Scene.  Triangles and texels are not copied: parts and
// textures point into the Scene's own storage, so the view is valid while the
// Scene is alive and unmodified.
struct SceneFlat {
    std::vector<rtg_object> objects;
    std::vector<rtg_part> parts;
    std::vector<rtg_material> materials;
    std::vector<rtg_texture> textures;
    std::vector<rtg_light> lights;
    rtg_scene_desc desc{};
};

struct SceneFlattener {
    static void flatten(const Scene &scene, SceneFlat &out);
};

// 64-bit content hash of everything a render reads from `desc`.
uint64_t rtg_desc_fingerprint(const rtg_scene_desc &desc);
