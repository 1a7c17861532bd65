// 3DElement.cpp — scalar implementation of the drop-in math records.
//
// Each operation reproduces the rounding of the reference's live SSE/AVX
// variant (rely.h:5-15 defines SSE2..AVX2 but never AVX_):
//   dot   : DPPS 0x71 -> (x*x' + y*y') + (z*z' + 0)       (3DElement.cpp:206-214)
//   cross : per lane l - r with l = (y z', z x', x y')     (3DElement.cpp:190-197)
//   v / f : v * (1/f)                                      (3DElement.cpp:137-158)
//   Normal(v): v / sqrt(dot(v, v)), true division          (3DElement.cpp:218-238)
// Compiled with -ffp-contract=off so no product is fused into an add.
#include "3DElement.h"

const char MY_OBJECT_NAME[][10] = {"ERROR", "sphere", "cube", "model", "plane", "ballplane"};
const char MY_LIGHT_NAME[][10] = {"ERROR", "parallel", "point", "spot"};

// mod (3DElement.cpp:4-10): remainder through modf of the float quotient.
float mod(const float &l, const float &r) {
    float t;
    std::modf(l / r, &t);
    float e = t * r;
    return l - e;
}

// Spherical -> cartesian in double, rounded to float on store (3DElement.cpp:12-17).
void Coord_sph2car(float &angy, float &angz, const float dis, Vertex &v) {
    v.z = (float)(dis * std::sin(angy * PI / 180) * std::cos(angz * PI / 180.0));
    v.x = (float)(dis * std::sin(angy * PI / 180) * std::sin(angz * PI / 180));
    v.y = (float)(dis * std::cos(angy * PI / 180));
}

// Variant with the angz >= 180 fold (3DElement.cpp:19-31); updates angy/angz.
void Coord_sph2car2(float &angy, float &angz, const float dis, Vertex &v) {
    bool fix = false;
    if (angz >= 180) {
        angz = mod(angz, 180.0f);
        angy = mod(360 - angy, 360.0f);
        fix = true;
    }
    if ((double)angy < 1e-6) angy = 360;
    v.z = (float)(dis * std::sin(angy * PI / 180) * std::cos(angz * PI / 180.0));
    v.x = (float)(dis * std::sin(angy * PI / 180) * std::sin(angz * PI / 180));
    if (fix && (double)mod(angy, 180.0f) < 1e-6) {
        v.z *= -1;
        v.x *= -1;
    }
    v.y = (float)(dis * std::cos(angy * PI / 180));
}

void Coord_car2sph(const Vertex &v, float &angy, float &angz, float &dis) {
    dis = v.length();
    angy = (float)(std::acos(v.y / dis) * 180 / PI);
    angz = (float)(std::atan2(v.x, v.z) * 180 / PI);
}

Coord2D Coord2D::operator+(const Coord2D &c) const { return Coord2D(u + c.u, v + c.v); }
Coord2D Coord2D::operator*(const float &n) const { return Coord2D(u * n, v * n); }

Vertex::Vertex() : x(0), y(0), z(0), w(0) {}

float Vertex::operator&(const Vertex &v) const {
    float xy = x * v.x + y * v.y;
    float zw = z * v.z + 0.0f;
    return xy + zw;
}
float Vertex::length_sqr() const { return *this & *this; }
float Vertex::length() const { return std::sqrt(length_sqr()); }
Vertex Vertex::muladd(const float &n, const Vertex &v) const {
    return Vertex(std::fma(x, n, v.x), std::fma(y, n, v.y), std::fma(z, n, v.z), std::fma(w, n, v.w));
}
Vertex Vertex::mixmul(const Vertex &v) const { return Vertex(x * v.x, y * v.y, z * v.z, w * v.w); }
Vertex Vertex::operator+(const Vertex &v) const { return Vertex(x + v.x, y + v.y, z + v.z, w + v.w); }
Vertex &Vertex::operator+=(const Vertex &v) { return *this = *this + v; }
Vertex Vertex::operator-(const Vertex &v) const { return Vertex(x - v.x, y - v.y, z - v.z, w - v.w); }
Vertex &Vertex::operator-=(const Vertex &v) { return *this = *this - v; }
Vertex Vertex::operator/(const float &n) const {
    float rec = 1 / n;
    return *this * rec;
}
Vertex &Vertex::operator/=(const float &n) { return *this = *this / n; }
Vertex Vertex::operator*(const float &n) const { return Vertex(x * n, y * n, z * n, w * n); }
Vertex &Vertex::operator*=(const float &n) { return *this = *this * n; }
Vertex Vertex::operator*(const Vertex &v) const {
    float a = y * v.z - z * v.y;
    float b = z * v.x - x * v.z;
    float c = x * v.y - y * v.x;
    float d = w * v.w - w * v.w;
    return Vertex(a, b, c, d);
}

Normal::Normal(const Vertex &v) {
    float s = std::sqrt(v & v);
    x = v.x / s;
    y = v.y / s;
    z = v.z / s;
    w = v.w / s;
}

Texture::Texture(bool check) {
    w = 4, h = 4;
    data = new uint8_t[48];
    if (check) {
        name = "check";
        // 4x4 grey checker in 2x2 blocks (3DElement.cpp:242-256)
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                int begin = (4 * a + b) * 3;
                uint8_t c = (a & 0x2) == (b & 0x2) ? 0xff : 0x7f;
                data[begin] = data[begin + 1] = data[begin + 2] = c;
            }
    } else {
        name = "empty";
        memset(data, 0xff, 48);
    }
}
Texture::Texture(const string &iname, const int16_t iw, const int16_t ih, const uint8_t *img) {
    name = iname;
    w = iw, h = ih;
    int32_t size = w * h * 3;
    data = new uint8_t[size];
    memcpy(data, img, size);
}
Texture::~Texture() { delete[] data; }
Texture::Texture(const Texture &t) {
    name = t.name;
    w = t.w, h = t.h;
    if (t.data != nullptr) {
        data = new uint8_t[w * h * 3];
        memcpy(data, t.data, w * h * 3);
    }
}
Texture::Texture(Texture &&t) {
    name = std::move(t.name);
    w = t.w, h = t.h;
    data = t.data;
    t.data = nullptr;
}
Texture &Texture::operator=(const Texture &t) {
    if (this == &t) return *this;
    name = t.name;
    w = t.w, h = t.h;
    delete[] data;
    data = nullptr;
    if (t.data != nullptr) {
        data = new uint8_t[w * h * 3];
        memcpy(data, t.data, w * h * 3);
    }
    return *this;
}

// Material defaults (3DElement.cpp:340-348).
Material::Material() {
    name = "simple";
    SetMtl(MY_MODEL_AMBIENT | MY_MODEL_DIFFUSE, 0.588f, 0.588f, 0.588f);
    SetMtl(MY_MODEL_EMISSION | MY_MODEL_SPECULAR, 0.0f, 0.0f, 0.0f);
    SetMtl(MY_MODEL_SHINESS, 10.0f);
    reflect = refract = 0.0f;
    rfr = 1.0f;
}
Material::~Material() {}
void Material::SetMtl(const uint8_t prop, const float r, const float g, const float b, const float a) {
    SetMtl(prop, Vertex(r, g, b, a));
    SetMtl(prop, a);
}
void Material::SetMtl(const uint8_t prop, const Vertex &v) {
    if (prop & MY_MODEL_AMBIENT) ambient = v;
    if (prop & MY_MODEL_DIFFUSE) diffuse = v;
    if (prop & MY_MODEL_EMISSION) emission = v;
    if (prop & MY_MODEL_SPECULAR) specular = v;
}
void Material::SetMtl(const uint8_t prop, const float val) {
    if (prop & MY_MODEL_SHINESS) shiness = val;
}

Triangle::Triangle() {}
Triangle::Triangle(const Vertex &va, const Vertex &vb, const Vertex &vc) {
    points[0] = va, points[1] = vb, points[2] = vc;
}
Triangle::Triangle(const Vertex &va, const Normal &na, const Vertex &vb, const Normal &nb, const Vertex &vc,
                   const Normal &nc) {
    points[0] = va, points[1] = vb, points[2] = vc;
    norms[0] = na, norms[1] = nb, norms[2] = nc;
}
Triangle::Triangle(const Vertex &va, const Normal &na, const Coord2D &ta, const Vertex &vb, const Normal &nb,
                   const Coord2D &tb, const Vertex &vc, const Normal &nc, const Coord2D &tc) {
    points[0] = va, points[1] = vb, points[2] = vc;
    norms[0] = na, norms[1] = nb, norms[2] = nc;
    tcoords[0] = ta, tcoords[1] = tb, tcoords[2] = tc;
}

Color::Color(const bool white) {
    r = g = b = white ? 1.0f : 0.0f;
    alpha = 1e20f;
}
Color::Color(const Vertex &v) {
    r = v.x, g = v.y, b = v.z;
    alpha = 1e20f;
}
Color::Color(const Normal &n) {
    r = (float)(0.5 * (n.x + 1));
    g = (float)(0.5 * (n.y + 1));
    b = (float)(0.5 * (n.z + 1));
}
// Nearest texel after wrapping uv into [0,1) (3DElement.cpp:430-450).
Color::Color(const Texture *tex, const Coord2D &coord) {
    if (tex == nullptr) {
        b = g = r = 1.0f;
        return;
    }
    float empty;
    float nu = std::modf(coord.u, &empty), nv = std::modf(coord.v, &empty);
    if (nu < 0) nu += 1;
    if (nv < 0) nv += 1;
    int16_t px = (int16_t)(nu * tex->w), py = (int16_t)(nv * tex->h);
    int32_t offset = (py * tex->w + px) * 3;
    b = tex->data[offset] / 255.0f;
    g = tex->data[offset + 1] / 255.0f;
    r = tex->data[offset + 2] / 255.0f;
}
void Color::set(const float depth, const float mindepth, const float maxdepth) {
    if (depth <= mindepth)
        r = 1.0f, g = b = 0.0f;
    else if (depth >= maxdepth)
        r = g = b = 0.0f;
    else {
        float after = std::log(depth), mx = std::log(maxdepth);
        r = g = b = (mx - after) / mx;
    }
}
static inline uint8_t quant8(float c) {
    if (c > 1.0f) return 255;
    if (c < 0.0f) return 0;
    float s = c * 255;
    return s == s ? (uint8_t)s : 0;
}
void Color::put(uint8_t *addr) {
    addr[0] = quant8(r);
    addr[1] = quant8(g);
    addr[2] = quant8(b);
}
void Color::get(uint8_t *addr) {
    r = addr[0] / 255.0f, g = addr[1] / 255.0f, b = addr[2] / 255.0f;
}

HitRes::HitRes(bool b) { distance = b ? 1e8f : 1e20f; }
bool HitRes::operator<(const HitRes &right) { return distance < right.distance; }
HitRes::operator bool() { return distance < 1e8f; }

// Light (3DElement.cpp:504-569).
Light::Light(const uint8_t type) {
    this->type = type;
    bLight = true;
    rangy = 90, rangz = 0, rdis = 16;
    coang = 0, exponent = 0;
    move(0, 0, 0);
    SetProperty(MY_MODEL_AMBIENT, 0.05f, 0.05f, 0.05f);
    SetProperty(MY_MODEL_DIFFUSE | MY_MODEL_SPECULAR, 1.0f, 1.0f, 1.0f);
    SetProperty(MY_MODEL_ATTENUATION, 1.0f, 0.0f, 0.0f);
    switch (type) {
    case MY_LIGHT_PARALLEL: position.alpha = 0.0f; break;
    case MY_LIGHT_POINT: position.alpha = 1.0f; break;
    case MY_LIGHT_SPOT: position.alpha = 1.0f; break;
    }
}
bool Light::turn() { return bLight = !bLight; }
void Light::move(const float dangy, const float dangz, const float ddis) {
    rdis += ddis;
    if (rdis < 2)
        rdis = 2;
    else if (rdis > 64)
        rdis = 64;
    angy = rangy = mod(360 + rangy + dangy, 360.0f);
    angz = rangz = mod(360 + rangz + dangz, 360.0f);
    dis = rdis;
    Coord_sph2car2(angy, angz, dis, position);
}
void Light::SetProperty(const int16_t prop, const float r, const float g, const float b, const float a) {
    Vertex set(r, g, b, a);
    if (prop & MY_MODEL_AMBIENT) ambient = set;
    if (prop & MY_MODEL_DIFFUSE) diffuse = set;
    if (prop & MY_MODEL_SPECULAR) specular = set;
    if (prop & MY_MODEL_ATTENUATION) attenuation = set;
    if (prop & MY_MODEL_POSITION) position = set;
}
void Light::SetLumi(const float lum) {
    float ext = lum / attenuation.alpha;
    attenuation.alpha = lum;
    ambient *= ext;
    diffuse *= ext;
    specular *= ext;
}

// Camera (3DElement.cpp:573-635).
Camera::Camera(GLint w, GLint h) {
    width = w, height = h;
    aspect = (float)w / h;
    fovy = 45.0f, zNear = 1.0f, zFar = 100.0f;
    position = Vertex(0, 4, 15);
    u = Vertex(1, 0, 0);
    v = Vertex(0, 1, 0);
    n = Vertex(0, 0, -1);
}
void Camera::move(const float &x, const float &y, const float &z) {
    position += u * x;
    position += v * y;
    position += n * z;
}
void Camera::yaw(const float angz) {
    float oangy = (float)(std::acos(n.y / 1) * 180 / PI), oangz = (float)(std::atan2(n.x, n.z) * 180 / PI);
    oangz -= angz;
    Coord_sph2car(oangy, oangz, 1, n);
    oangy = (float)(std::acos(u.y / 1) * 180 / PI);
    oangz = (float)(std::atan2(u.x, u.z) * 180 / PI);
    oangz -= angz;
    Coord_sph2car(oangy, oangz, 1, u);
}
void Camera::pitch(float angy) {
    float oangy = (float)(std::acos(n.y / 1) * 180 / PI), oangz = (float)(std::atan2(n.x, n.z) * 180 / PI);
    if (oangy - angy < 1.0) angy = (float)(oangy - 1.0);
    if (oangy - angy > 179.0) angy = (float)(oangy - 179.0);
    oangy -= angy;
    Coord_sph2car(oangy, oangz, 1, n);
    v = u * n;
}
void Camera::resize(GLint w, GLint h) {
    width = w, height = h;
    aspect = (float)w / h;
}
