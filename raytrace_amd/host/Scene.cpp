// Scene.cpp — scene construction and mutation, with the reference's rules
// (Scene.cpp:6-301): material library, light normalisation, object poses.
#include "Scene.h"

Scene::Scene() {
    EnvLight = Vertex(0.2f, 0.2f, 0.2f, 1.0f);
    // Material library (Scene.cpp:12-66).
    Material mtl;
    mtl.name = "brass";
    mtl.SetMtl(MY_MODEL_AMBIENT, 0.329412f, 0.223529f, 0.027451f);
    mtl.SetMtl(MY_MODEL_DIFFUSE, 0.780392f, 0.568627f, 0.113725f);
    mtl.SetMtl(MY_MODEL_SPECULAR, 0.992157f, 0.941176f, 0.807843f);
    mtl.SetMtl(MY_MODEL_SHINESS, 27.8974f);
    MtlLiby.push_back(mtl);

    mtl = Material();
    mtl.name = "bas-sphere";
    mtl.SetMtl(MY_MODEL_AMBIENT, 0.1f, 0.1f, 0.1f);
    mtl.SetMtl(MY_MODEL_DIFFUSE, 0.1f, 0.5f, 0.8f);
    mtl.SetMtl(MY_MODEL_SPECULAR, 1.0f, 1.0f, 1.0f);
    mtl.SetMtl(MY_MODEL_SHINESS, 100);
    mtl.reflect = 0.35f;
    MtlLiby.push_back(mtl);

    mtl = Material();
    mtl.name = "mirror";
    mtl.SetMtl(MY_MODEL_AMBIENT, 0.1f, 0.1f, 0.1f);
    mtl.SetMtl(MY_MODEL_DIFFUSE, 0.1f, 0.1f, 0.1f);
    mtl.SetMtl(MY_MODEL_SPECULAR, 1.0f, 1.0f, 1.0f);
    mtl.SetMtl(MY_MODEL_SHINESS, 127);
    mtl.reflect = 0.95f;
    MtlLiby.push_back(mtl);

    mtl = Material();
    mtl.name = "green grass";
    mtl.SetMtl(MY_MODEL_AMBIENT, 0.1f, 0.4f, 0.1f);
    mtl.SetMtl(MY_MODEL_DIFFUSE, 0.1f, 0.5f, 0.1f);
    mtl.SetMtl(MY_MODEL_SPECULAR, 1.0f, 1.0f, 1.0f);
    mtl.SetMtl(MY_MODEL_SHINESS, 127);
    mtl.reflect = 0.55f;
    MtlLiby.push_back(mtl);

    mtl = Material();
    mtl.name = "grass";
    mtl.SetMtl(MY_MODEL_AMBIENT, 0.1f, 0.1f, 0.1f);
    mtl.SetMtl(MY_MODEL_DIFFUSE, 0.9f, 0.9f, 0.9f);
    mtl.SetMtl(MY_MODEL_SPECULAR, 1.0f, 1.0f, 1.0f);
    mtl.SetMtl(MY_MODEL_SHINESS, 127);
    mtl.reflect = 0.15f;
    mtl.refract = 0.75f;
    mtl.rfr = 1.5f;
    MtlLiby.push_back(mtl);

    mtl = Material();
    mtl.name = "wall";
    mtl.SetMtl(MY_MODEL_AMBIENT, 0.2f, 0.2f, 0.2f);
    mtl.SetMtl(MY_MODEL_DIFFUSE, 0.7f, 0.7f, 0.7f);
    mtl.SetMtl(MY_MODEL_SPECULAR, 0.5f, 0.5f, 0.5f);
    mtl.SetMtl(MY_MODEL_SHINESS, 5);
    mtl.reflect = 0.15f;
    mtl.refract = 0;
    MtlLiby.push_back(mtl);
}

Scene::~Scene() {
    for (auto &obj : Objects) delete obj;
}

// Light components are normalised to sum 1, then scaled by atte.alpha
// through SetLumi (Scene.cpp:83-97).  At most 8 lights, as in the reference.
uint8_t Scene::AddLight(const uint8_t type, const Vertex &comp, const Vertex &atte) {
    if (Lights.size() == 8) return 0xff;
    Light light(type);
    float sum = comp.x + comp.y + comp.z;
    Vertex nc = comp / sum;
    light.SetProperty(MY_MODEL_AMBIENT, nc.x, nc.x, nc.x);
    light.SetProperty(MY_MODEL_DIFFUSE, nc.y, nc.y, nc.y);
    light.SetProperty(MY_MODEL_SPECULAR, nc.z, nc.z, nc.z);
    light.SetProperty(MY_MODEL_ATTENUATION, atte.x, atte.y, atte.z);
    light.SetLumi(atte.alpha);
    Lights.push_back(light);
    return (uint8_t)(Lights.size() - 1);
}

uint8_t Scene::AddSphere(const float radius) {
    Sphere *sphere = new Sphere(radius);
    sphere->position = Vertex(0.0f, radius, 0.0f);
    sphere->SetMtl(MtlLiby[1]);
    Objects.push_back(sphere);
    return (uint8_t)(Objects.size() - 1);
}

uint8_t Scene::AddCube(const float len) {
    Box *box = new Box(len);
    box->position = Vertex(0.0f, len / 2, 0.0f);
    box->SetMtl(MtlLiby[0]);
    Objects.push_back(box);
    return (uint8_t)(Objects.size() - 1);
}

uint8_t Scene::AddModel(const wstring &objname, const wstring &mtlname, uint8_t code) {
    Model *model = new Model();
    model->loadOBJ(objname, mtlname, code);
    Objects.push_back(model);
    return (uint8_t)(Objects.size() - 1);
}

uint8_t Scene::AddModel(Model *model) {
    Objects.push_back(model);
    return (uint8_t)(Objects.size() - 1);
}

uint8_t Scene::AddPlane() {
    Plane *plane = new Plane();
    Material mtl;
    mtl.reflect = 0.6f;
    plane->SetMtl(mtl);
    Objects.push_back(plane);
    return (uint8_t)(Objects.size() - 1);
}

uint8_t Scene::AddBallPlane(const float radius) {
    BallPlane *ballp = new BallPlane(radius);
    ballp->SetMtl(MtlLiby[1]);
    Objects.push_back(ballp);
    return (uint8_t)(Objects.size() - 1);
}

bool Scene::ChgLightComp(const uint8_t type, const uint8_t num, const Vertex &v) {
    if (num >= Lights.size()) return false;
    Light &light = Lights[num];
    Vertex ta, td, ts;
    float tmp;
    switch (type) {
    case MY_LIGHT_COMPENT:
        ta = light.ambient * v.x;
        td = light.diffuse * v.y;
        ts = light.specular * v.z;
        tmp = (ta.x + td.x + ts.x) / light.attenuation.alpha;
        ta.x /= tmp, ts.x /= tmp, td.x /= tmp;
        tmp = (ta.y + td.y + ts.y) / light.attenuation.alpha;
        ta.y /= tmp, ts.y /= tmp, td.y /= tmp;
        tmp = (ta.z + td.z + ts.z) / light.attenuation.alpha;
        ta.z /= tmp, ts.z /= tmp, td.z /= tmp;
        light.SetProperty(MY_MODEL_AMBIENT, ta.x, ta.y, ta.z);
        light.SetProperty(MY_MODEL_DIFFUSE, td.x, td.y, td.z);
        light.SetProperty(MY_MODEL_SPECULAR, ts.x, ts.y, ts.z);
        break;
    case MY_LIGHT_LUMI:
        tmp = light.attenuation.alpha * v.alpha;
        light.SetLumi(tmp);
        break;
    default:
        return false;
    }
    return true;
}

bool Scene::ChgMtl(const uint8_t num, const Material &mtl) {
    if (num >= Objects.size()) return false;
    Objects[num]->SetMtl(mtl);
    Plane *p = dynamic_cast<Plane *>(Objects[num]);
    if (mtl.name == "wall" && p != nullptr) p->setTex(Texture(false));
    return true;
}

bool Scene::ChgMtl(const uint8_t num, const Normal &clr) {
    if (num >= Objects.size()) return false;
    auto chgclr = [clr](Vertex &mclr) {
        if (clr.w < 0.5f) {
            float sum = mclr.r + mclr.g + mclr.b;
            mclr = clr * sum;
        } else
            mclr = clr;
    };
    if (Objects[num]->type == MY_OBJECT_MODEL) {
        Model &m = dynamic_cast<Model &>(*Objects[num]);
        for (Material &mtl : m.mtls) chgclr(mtl.diffuse);
    } else
        chgclr(Objects[num]->mtl.diffuse);
    return true;
}

bool Scene::Delete(const uint8_t type, const uint8_t num) {
    switch (type) {
    case MY_MODEL_OBJECT:
        if (num >= Objects.size()) return false;
        delete Objects[num];
        Objects.erase(Objects.begin() + num);
        break;
    case MY_MODEL_LIGHT:
        if (num >= Lights.size()) return false;
        Lights.erase(Lights.begin() + num);
    }
    return true;
}

bool Scene::MovePos(const uint8_t type, const uint8_t num, const Vertex &v) {
    switch (type) {
    case MY_MODEL_OBJECT:
        if (num >= Objects.size()) return false;
        if (Objects[num]->type == MY_OBJECT_PLANE)
            dynamic_cast<Plane &>(*Objects[num]).rotate(v);
        else if (Objects[num]->type == MY_OBJECT_BALLPLANE)
            dynamic_cast<BallPlane &>(*Objects[num]).rotate(v);
        else
            Objects[num]->position += v;
        break;
    case MY_MODEL_LIGHT:
        if (num >= Lights.size()) return false;
        Lights[num].move(v.x, v.y, v.z);
        break;
    default:
        return false;
    }
    return true;
}

bool Scene::Switch(const uint8_t type, const uint8_t num, const bool isShow) {
    bool old, isSwitch = type & MY_MODEL_SWITCH;
    switch (type & 0x7f) {
    case MY_MODEL_LIGHT:
        if (num >= Lights.size()) return false;
        Lights[num].bLight = isSwitch ? !Lights[num].bLight : isShow;
        return true;
    case MY_MODEL_OBJECT:
        if (num >= Objects.size()) return false;
        old = Objects[num]->bShow;
        Objects[num]->bShow = isSwitch ? !old : isShow;
        return old;
    }
    return false;
}
