// Scene.cpp — the scene container of the drop-in API: material library,
// object and light creation, and the edits the viewer makes (reference
// Scene.cpp:6-301).  Headless: nothing here touches GL.
//
// The float steps that produce render inputs are the reference's (AddLight's
// component split and luminance scale, ChgLightComp's renormalisation, the
// poses of new objects); the bookkeeping around them is this file's own.
#include "Scene.h"

namespace {

// The built-in material library (Scene.cpp:12-66): name, ambient, diffuse,
// specular, shininess, reflectance, transmittance, IOR.  Rows in the
// reference's order; MtlLiby[0] dresses new cubes, MtlLiby[1] new spheres and
// ball planes.
struct LibraryRow {
    const char *name;
    float ka[3], kd[3], ks[3];
    float shiness, reflect, refract, rfr;
};
const LibraryRow kLibrary[] = {
    {"brass", {0.329412f, 0.223529f, 0.027451f}, {0.780392f, 0.568627f, 0.113725f},
     {0.992157f, 0.941176f, 0.807843f}, 27.8974f, 0.0f, 0.0f, 1.0f},
    {"bas-sphere", {0.1f, 0.1f, 0.1f}, {0.1f, 0.5f, 0.8f}, {1.0f, 1.0f, 1.0f}, 100.0f, 0.35f, 0.0f, 1.0f},
    {"mirror", {0.1f, 0.1f, 0.1f}, {0.1f, 0.1f, 0.1f}, {1.0f, 1.0f, 1.0f}, 127.0f, 0.95f, 0.0f, 1.0f},
    {"green grass", {0.1f, 0.4f, 0.1f}, {0.1f, 0.5f, 0.1f}, {1.0f, 1.0f, 1.0f}, 127.0f, 0.55f, 0.0f, 1.0f},
    {"grass", {0.1f, 0.1f, 0.1f}, {0.9f, 0.9f, 0.9f}, {1.0f, 1.0f, 1.0f}, 127.0f, 0.15f, 0.75f, 1.5f},
    {"wall", {0.2f, 0.2f, 0.2f}, {0.7f, 0.7f, 0.7f}, {0.5f, 0.5f, 0.5f}, 5.0f, 0.15f, 0.0f, 1.0f},
};

Material from_row(const LibraryRow &r) {
    Material m;
    m.name = r.name;
    m.SetMtl(MY_MODEL_AMBIENT, r.ka[0], r.ka[1], r.ka[2]);
    m.SetMtl(MY_MODEL_DIFFUSE, r.kd[0], r.kd[1], r.kd[2]);
    m.SetMtl(MY_MODEL_SPECULAR, r.ks[0], r.ks[1], r.ks[2]);
    m.SetMtl(MY_MODEL_SHINESS, r.shiness);
    m.reflect = r.reflect, m.refract = r.refract, m.rfr = r.rfr;
    return m;
}

constexpr size_t kMaxLights = 8;  // AddLight refuses a ninth (Scene.cpp:85-86)
constexpr uint8_t kFull = 0xff;

template <class T>
T *slot(vector<T> &items, uint8_t num) {
    return num < items.size() ? &items[num] : nullptr;
}

// Channel c (0 x/r, 1 y/g, 2 z/b) of a Vertex.
float &lane(Vertex &v, int c) { return c == 0 ? v.x : c == 1 ? v.y : v.z; }

}  // namespace

Scene::Scene() {
    EnvLight = Vertex(0.2f, 0.2f, 0.2f, 1.0f);
    for (const LibraryRow &r : kLibrary) MtlLiby.push_back(from_row(r));
}

Scene::~Scene() {
    for (DrawObject *o : Objects) delete o;
}

uint8_t Scene::adopt(DrawObject *obj) {
    Objects.push_back(obj);
    return (uint8_t)(Objects.size() - 1);
}

// The light's ambient / diffuse / specular weights are the components of
// `comp` over their sum (a reciprocal multiply, as Vertex / float is), grey
// per component; then the attenuation, and the luminance atte.alpha scales
// all three.
uint8_t Scene::AddLight(const uint8_t type, const Vertex &comp, const Vertex &atte) {
    if (Lights.size() >= kMaxLights) return kFull;
    Light light(type);
    const Vertex share = comp / (comp.x + comp.y + comp.z);
    light.SetProperty(MY_MODEL_AMBIENT, share.x, share.x, share.x);
    light.SetProperty(MY_MODEL_DIFFUSE, share.y, share.y, share.y);
    light.SetProperty(MY_MODEL_SPECULAR, share.z, share.z, share.z);
    light.SetProperty(MY_MODEL_ATTENUATION, atte.x, atte.y, atte.z);
    light.SetLumi(atte.alpha);
    Lights.push_back(light);
    return (uint8_t)(Lights.size() - 1);
}

// New spheres rest on the ground (centre at height r), cubes likewise (at
// half their edge); planes get a plain material with reflectance 0.6.
uint8_t Scene::AddSphere(const float radius) {
    Sphere *s = new Sphere(radius);
    s->position = Vertex(0.0f, radius, 0.0f);
    s->SetMtl(MtlLiby[1]);
    return adopt(s);
}

uint8_t Scene::AddCube(const float len) {
    Box *b = new Box(len);
    b->position = Vertex(0.0f, len / 2, 0.0f);
    b->SetMtl(MtlLiby[0]);
    return adopt(b);
}

uint8_t Scene::AddModel(const wstring &objname, const wstring &mtlname, uint8_t code) {
    Model *m = new Model();
    m->loadOBJ(objname, mtlname, code);
    return adopt(m);
}

uint8_t Scene::AddModel(Model *model) { return adopt(model); }

uint8_t Scene::AddPlane() {
    Plane *p = new Plane();
    Material plain;
    plain.reflect = 0.6f;
    p->SetMtl(plain);
    return adopt(p);
}

uint8_t Scene::AddBallPlane(const float radius) {
    BallPlane *b = new BallPlane(radius);
    b->SetMtl(MtlLiby[1]);
    return adopt(b);
}

// MY_LIGHT_COMPENT: weight the ambient, diffuse and specular parts by v.x,
// v.y, v.z, then rescale each colour channel so that the three parts sum to
// the light's luminance again.  MY_LIGHT_LUMI: multiply the luminance by v.w.
bool Scene::ChgLightComp(const uint8_t type, const uint8_t num, const Vertex &v) {
    Light *light = slot(Lights, num);
    if (light == nullptr) return false;
    if (type == MY_LIGHT_LUMI) {
        light->SetLumi(light->attenuation.alpha * v.alpha);
        return true;
    }
    if (type != MY_LIGHT_COMPENT) return false;
    Vertex parts[3] = {light->ambient * v.x, light->diffuse * v.y, light->specular * v.z};
    for (int c = 0; c < 3; ++c) {
        const float scale = (lane(parts[0], c) + lane(parts[1], c) + lane(parts[2], c)) / light->attenuation.alpha;
        for (Vertex &p : parts) lane(p, c) /= scale;
    }
    light->SetProperty(MY_MODEL_AMBIENT, parts[0].x, parts[0].y, parts[0].z);
    light->SetProperty(MY_MODEL_DIFFUSE, parts[1].x, parts[1].y, parts[1].z);
    light->SetProperty(MY_MODEL_SPECULAR, parts[2].x, parts[2].y, parts[2].z);
    return true;
}

// A plane dressed in the library's "wall" material loses its checker (plain
// white texture).
bool Scene::ChgMtl(const uint8_t num, const Material &mtl) {
    DrawObject **obj = slot(Objects, num);
    if (obj == nullptr) return false;
    (*obj)->SetMtl(mtl);
    if (Plane *p = dynamic_cast<Plane *>(*obj); p != nullptr && mtl.name == "wall") p->setTex(Texture(false));
    return true;
}

// Recolour the diffuse term (every material of a model): with clr.w < 0.5 the
// new hue keeps the old colour's total r + g + b, otherwise clr replaces it.
bool Scene::ChgMtl(const uint8_t num, const Normal &clr) {
    DrawObject **obj = slot(Objects, num);
    if (obj == nullptr) return false;
    const bool keep_brightness = clr.w < 0.5f;
    auto recolour = [&](Vertex &diffuse) {
        diffuse = keep_brightness ? clr * (diffuse.r + diffuse.g + diffuse.b) : Vertex(clr);
    };
    if ((*obj)->type == MY_OBJECT_MODEL)
        for (Material &m : static_cast<Model *>(*obj)->mtls) recolour(m.diffuse);
    else
        recolour((*obj)->mtl.diffuse);
    return true;
}

bool Scene::Delete(const uint8_t type, const uint8_t num) {
    if (type == MY_MODEL_OBJECT) {
        if (num >= Objects.size()) return false;
        delete Objects[num];
        Objects.erase(Objects.begin() + num);
    } else if (type == MY_MODEL_LIGHT) {
        if (num >= Lights.size()) return false;
        Lights.erase(Lights.begin() + num);
    }
    return true;
}

// Objects: planes and ball planes turn (rotate by the angle deltas of v),
// everything else translates by v.  Lights orbit (Light::move).
bool Scene::MovePos(const uint8_t type, const uint8_t num, const Vertex &v) {
    if (type == MY_MODEL_LIGHT) {
        Light *light = slot(Lights, num);
        if (light != nullptr) light->move(v.x, v.y, v.z);
        return light != nullptr;
    }
    if (type != MY_MODEL_OBJECT) return false;
    DrawObject **obj = slot(Objects, num);
    if (obj == nullptr) return false;
    switch ((*obj)->type) {
    case MY_OBJECT_PLANE:
        static_cast<Plane *>(*obj)->rotate(v);
        break;
    case MY_OBJECT_BALLPLANE:
        static_cast<BallPlane *>(*obj)->rotate(v);
        break;
    default:
        (*obj)->position += v;
    }
    return true;
}

// Show or hide (or, with MY_MODEL_SWITCH in `type`, toggle) a light or an
// object.  A light reports success; an object reports whether it was shown
// before.
bool Scene::Switch(const uint8_t type, const uint8_t num, const bool isShow) {
    const bool toggle = (type & MY_MODEL_SWITCH) != 0;
    auto next = [&](bool cur) { return toggle ? !cur : isShow; };
    switch (type & 0x7f) {
    case MY_MODEL_LIGHT: {
        Light *light = slot(Lights, num);
        if (light == nullptr) return false;
        light->bLight = next(light->bLight);
        return true;
    }
    case MY_MODEL_OBJECT: {
        DrawObject **obj = slot(Objects, num);
        if (obj == nullptr) return false;
        const bool was = (*obj)->bShow;
        (*obj)->bShow = next(was);
        return was;
    }
    default:
        return false;
    }
}
