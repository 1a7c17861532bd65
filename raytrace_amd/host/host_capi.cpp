// host_capi.cpp — C ABI over the headless Scene / RayTracer API (rtg_host.h).
#include "RayTracer.h"
#include "flatten.h"
#include "scenegen.h"
#include "../../include/rtg_host.h"

#include <chrono>
#include <codecvt>
#include <locale>

struct rtg_host_scene_s {
    std::unique_ptr<Scene> scene;
    SceneFlat flat;
};

namespace {
std::wstring widen(const char *s) {
    std::wstring w;
    if (s == nullptr) return w;
    std::mbstate_t st{};
    const char *p = s;
    size_t n = std::strlen(s);
    while (n > 0) {
        wchar_t c;
        size_t k = std::mbrtowc(&c, p, n, &st);
        if (k == 0 || k == (size_t)-1 || k == (size_t)-2) {
            c = (unsigned char)*p;
            k = 1;
        }
        w.push_back(c);
        p += k;
        n -= k;
    }
    return w;
}
bool valid(rtg_host_scene_t h) { return h != nullptr && h->scene != nullptr; }
}  // namespace

extern "C" {

int rtg_host_scene_config(int config, const rtg_gen_opts *opts, rtg_host_scene_t *out, rtg_gen_info *info) {
    if (out == nullptr) return RTG_E_INVALID;
    rtg_gen_opts o{};
    o.n_spheres = -1;
    o.n_lights = -1;
    o.max_level = -1;
    if (opts != nullptr) o = *opts;
    rtg_gen_info gi{};
    Scene *s = rtg_make_config_scene(config, o, gi);
    if (s == nullptr) return RTG_E_INVALID;
    auto *h = new rtg_host_scene_s();
    h->scene.reset(s);
    *out = h;
    if (info != nullptr) *info = gi;
    return RTG_OK;
}

int rtg_host_gen_requests(uint64_t seed, int n, double rate, rtg_gen_request *out) {
    if (n < 0 || (n > 0 && out == nullptr) || !(rate >= 0)) return RTG_E_INVALID;
    rtg_make_requests(seed, n, rate, out);
    return RTG_OK;
}

int rtg_host_scene_new(rtg_host_scene_t *out) {
    if (out == nullptr) return RTG_E_INVALID;
    auto *h = new rtg_host_scene_s();
    h->scene.reset(new Scene());
    *out = h;
    return RTG_OK;
}

int rtg_host_scene_destroy(rtg_host_scene_t h) {
    delete h;
    return RTG_OK;
}

const rtg_scene_desc *rtg_host_scene_desc(rtg_host_scene_t h) {
    if (!valid(h)) return nullptr;
    SceneFlattener::flatten(*h->scene, h->flat);
    return &h->flat.desc;
}

int rtg_host_add_sphere(rtg_host_scene_t h, float r) { return valid(h) ? h->scene->AddSphere(r) : RTG_E_INVALID; }
int rtg_host_add_plane(rtg_host_scene_t h) { return valid(h) ? h->scene->AddPlane() : RTG_E_INVALID; }
int rtg_host_add_cube(rtg_host_scene_t h, float len) { return valid(h) ? h->scene->AddCube(len) : RTG_E_INVALID; }
int rtg_host_add_ballplane(rtg_host_scene_t h, float r) {
    return valid(h) ? h->scene->AddBallPlane(r) : RTG_E_INVALID;
}
int rtg_host_add_model_obj(rtg_host_scene_t h, const char *obj_path, const char *mtl_path) {
    if (!valid(h)) return RTG_E_INVALID;
    return h->scene->AddModel(widen(obj_path), widen(mtl_path));
}
int rtg_host_add_light(rtg_host_scene_t h, int type, const float comp[3], const float atte[4]) {
    if (!valid(h) || comp == nullptr) return RTG_E_INVALID;
    Vertex a = atte != nullptr ? Vertex(atte[0], atte[1], atte[2], atte[3]) : Vertex(1, 0, 0, 1);
    return h->scene->AddLight((uint8_t)type, Vertex(comp[0], comp[1], comp[2]), a);
}
int rtg_host_move_pos(rtg_host_scene_t h, int type, int num, const float v[3]) {
    if (!valid(h) || v == nullptr) return RTG_E_INVALID;
    return h->scene->MovePos((uint8_t)type, (uint8_t)num, Vertex(v[0], v[1], v[2])) ? RTG_OK : RTG_E_INVALID;
}
int rtg_host_switch(rtg_host_scene_t h, int type, int num, int show) {
    if (!valid(h)) return RTG_E_INVALID;
    return h->scene->Switch((uint8_t)type, (uint8_t)num, show != 0) ? 1 : 0;
}
int rtg_host_chg_light_comp(rtg_host_scene_t h, int type, int num, const float v[4]) {
    if (!valid(h) || v == nullptr) return RTG_E_INVALID;
    return h->scene->ChgLightComp((uint8_t)type, (uint8_t)num, Vertex(v[0], v[1], v[2], v[3])) ? RTG_OK
                                                                                             : RTG_E_INVALID;
}
int rtg_host_recolour(rtg_host_scene_t h, int num, const float clr[4]) {
    if (!valid(h) || clr == nullptr) return RTG_E_INVALID;
    return h->scene->ChgMtl((uint8_t)num, Normal(clr[0], clr[1], clr[2], clr[3])) ? RTG_OK : RTG_E_INVALID;
}
int rtg_host_chg_mtl_library(rtg_host_scene_t h, int num, int lib_index) {
    if (!valid(h) || lib_index < 0 || lib_index >= (int)h->scene->MtlLiby.size()) return RTG_E_INVALID;
    return h->scene->ChgMtl((uint8_t)num, h->scene->MtlLiby[lib_index]) ? RTG_OK : RTG_E_INVALID;
}
int rtg_host_delete(rtg_host_scene_t h, int type, int num) {
    if (!valid(h)) return RTG_E_INVALID;
    return h->scene->Delete((uint8_t)type, (uint8_t)num) ? RTG_OK : RTG_E_INVALID;
}
int rtg_host_light(rtg_host_scene_t h, int num, rtg_light *out) {
    if (!valid(h) || out == nullptr || num < 0 || num >= (int)h->scene->Lights.size()) return RTG_E_INVALID;
    const Light &l = h->scene->Lights[num];
    auto v4 = [](const Vertex &v) { return rtg_vec4{v.x, v.y, v.z, v.w}; };
    *out = rtg_light{};
    out->type = l.type, out->on = l.bLight;
    out->position = v4(l.position), out->ambient = v4(l.ambient), out->diffuse = v4(l.diffuse);
    out->specular = v4(l.specular), out->attenuation = v4(l.attenuation);
    return RTG_OK;
}
int rtg_host_set_material(rtg_host_scene_t h, int num, const rtg_material *m) {
    if (!valid(h) || m == nullptr) return RTG_E_INVALID;
    Material mt;
    mt.ambient = Vertex(m->ambient.x, m->ambient.y, m->ambient.z, m->ambient.w);
    mt.diffuse = Vertex(m->diffuse.x, m->diffuse.y, m->diffuse.z, m->diffuse.w);
    mt.specular = Vertex(m->specular.x, m->specular.y, m->specular.z, m->specular.w);
    mt.emission = Vertex(m->emission.x, m->emission.y, m->emission.z, m->emission.w);
    mt.shiness = m->shiness;
    mt.reflect = m->reflect;
    mt.refract = m->refract;
    mt.rfr = m->rfr;
    return h->scene->ChgMtl((uint8_t)num, mt) ? RTG_OK : RTG_E_INVALID;
}
int rtg_host_set_position(rtg_host_scene_t h, int num, const float p[3]) {
    if (!valid(h) || p == nullptr || num < 0 || (size_t)num >= h->scene->Objects.size()) return RTG_E_INVALID;
    h->scene->Objects[num]->position = Vertex(p[0], p[1], p[2]);
    return RTG_OK;
}
int rtg_host_set_light_position(rtg_host_scene_t h, int num, const float p[4]) {
    if (!valid(h) || p == nullptr || num < 0 || (size_t)num >= h->scene->Lights.size()) return RTG_E_INVALID;
    h->scene->Lights[num].SetProperty(MY_MODEL_POSITION, p[0], p[1], p[2], p[3]);
    return RTG_OK;
}
int rtg_host_set_env_light(rtg_host_scene_t h, const float e[4]) {
    if (!valid(h) || e == nullptr) return RTG_E_INVALID;
    h->scene->EnvLight = Vertex(e[0], e[1], e[2], e[3]);
    return RTG_OK;
}
int rtg_host_resize(rtg_host_scene_t h, int w, int hh) {
    if (!valid(h) || w <= 0 || hh <= 0) return RTG_E_INVALID;
    h->scene->cam.resize(w, hh);
    return RTG_OK;
}

int rtg_host_raytrace(rtg_host_scene_t h, int type, int max_level, int frames, uint8_t *out, double *use_time,
                      uint64_t *uploads) {
    if (!valid(h) || out == nullptr || frames < 1) return RTG_E_INVALID;
    RayTracer rt(*h->scene);
    rt.maxLevel = (uint8_t)max_level;
    for (int f = 0; f < frames; ++f) {
        rt.start((uint8_t)type, 4);
        while (!rt.isFinish) std::this_thread::sleep_for(std::chrono::milliseconds(1));
        if (rt.lastStatus != RTG_OK) break;
    }
    memcpy(out, rt.output, 2048 * 2048 * 3);
    if (use_time != nullptr) *use_time = rt.useTime;
    if (uploads != nullptr) *uploads = rt.uploads;
    return rt.lastStatus;
}

int rtg_host_intersect(rtg_host_scene_t h, int num, const rtg_host_ray *ray, int64_t prev_obj, float prev_distance,
                       float min, rtg_host_hit *out) {
    if (!valid(h) || ray == nullptr || out == nullptr || num < 0 || num >= (int)h->scene->Objects.size())
        return RTG_E_INVALID;
    DrawObject *obj = h->scene->Objects[num];
    obj->RTPrepare();
    auto vtx = [](const float *a) { return Vertex(a[0], a[1], a[2], a[3]); };
    const float *d = ray->direction;
    Ray r(vtx(ray->origin), Normal(d[0], d[1], d[2], d[3]), (uint8_t)ray->type);  // taken as given, not renormalised
    r.isInside = (uint8_t)ray->inside;
    HitRes prev(prev_distance);
    prev.obj = (intptr_t)prev_obj;
    const HitRes res = obj->intersect(r, prev, min);
    memset(out, 0, sizeof(*out));
    const float *p = &res.position.x, *n = &res.normal.x;
    for (int k = 0; k < 4; ++k) out->position[k] = p[k], out->normal[k] = n[k];
    out->u = res.tcoord.u, out->v = res.tcoord.v;
    out->distance = res.distance, out->rfr = res.rfr;
    out->obj = (int64_t)res.obj;
    out->object = (int64_t)(intptr_t)obj;
    out->inside = res.isInside;
    out->hit = !(res.obj == prev.obj && res.distance == prev.distance);
    out->has_tex = res.tex != nullptr;
    return RTG_OK;
}

int rtg_host_raytrace_stop(rtg_host_scene_t h, int type, int max_level, double stop_after_s, uint8_t *out,
                           int *finished_before_stop) {
    if (!valid(h) || out == nullptr || !(stop_after_s >= 0.0)) return RTG_E_INVALID;
    RayTracer rt(*h->scene);
    rt.maxLevel = (uint8_t)max_level;
    rt.start((uint8_t)type, 4);
    std::this_thread::sleep_for(std::chrono::duration<double>(stop_after_s));
    if (finished_before_stop != nullptr) *finished_before_stop = rt.isFinish ? 1 : 0;
    rt.stop();
    while (!rt.isFinish) std::this_thread::sleep_for(std::chrono::milliseconds(1));
    memcpy(out, rt.output, 2048 * 2048 * 3);
    return rt.lastStatus;
}

}  // extern "C"
