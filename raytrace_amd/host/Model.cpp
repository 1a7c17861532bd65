// Model.cpp — mesh loading and model-space preparation for the drop-in API.
//
// Parsing and post-processing follow the reference's rules so that a model
// loaded here has the same parts, borders, materials and rescale as in the
// reference (Model.cpp:14-400); ray intersection itself is done on the device.
#include "Model.h"

#include <cwchar>

namespace {

std::string narrow(const wstring &w) {
    std::string s;
    s.reserve(w.size());
    std::mbstate_t st{};
    char buf[16];
    for (wchar_t c : w) {
        size_t n = std::wcrtomb(buf, c, &st);
        if (n == (size_t)-1) {
            s.push_back('?');
            continue;
        }
        s.append(buf, n);
    }
    return s;
}

// Line tokenizer with the reference's conventions (Model.cpp:870-910): up to
// five whitespace-separated tokens, "****" appended to the first so that
// short keywords compare on three characters, tokens of earlier lines kept
// in `data` when a line has fewer.
class Loader {
    FILE *fp;
    char tmpdat[8][256];
public:
    explicit Loader(const wstring &fname) { fp = std::fopen(narrow(fname).c_str(), "r"); }
    ~Loader() {
        if (fp != nullptr) std::fclose(fp);
    }
    bool ok() const { return fp != nullptr; }
    int read(string data[]) {
        if (fp == nullptr) return INT8_MIN;
        if (std::fgets(tmpdat[0], 256, fp) == nullptr) return INT8_MIN;
        int a = std::sscanf(tmpdat[0], "%255s%255s%255s%255s%255s", tmpdat[1], tmpdat[2], tmpdat[3], tmpdat[4],
                            tmpdat[5]);
        std::strncat(tmpdat[1], "****", 255 - std::strlen(tmpdat[1]));
        for (int b = 0; b < a; ++b) data[b] = string(tmpdat[b + 1]);
        return a;
    }
    static int parseInt(const string &in, int out[]) {
        int a = std::sscanf(in.c_str(), "%d/%d/%d/%d", &out[0], &out[1], &out[2], &out[3]);
        if (a < 2) a = std::sscanf(in.c_str(), "%d//%d/%d", &out[0], &out[2], &out[3]);
        return a;
    }
};

constexpr int32_t key3(char a, char b, char c) { return a * 65536 + b * 256 + c; }

int32_t parseStr(string str) {
    if (str.length() < 3) str += "****";
    return str[2] + 256 * str[1] + 65536 * str[0];
}

void border(Vertex &mn, Vertex &mx, const Vertex &v) {
    mn = Vertex(std::min(mn.x, v.x), std::min(mn.y, v.y), std::min(mn.z, v.z), std::min(mn.w, v.w));
    mx = Vertex(std::max(mx.x, v.x), std::max(mx.y, v.y), std::max(mx.z, v.z), std::max(mx.w, v.w));
}

template <typename T>
const T &at_or_default(const vector<T> &v, int i) {
    static const T empty{};
    return (i >= 0 && (size_t)i < v.size()) ? v[i] : empty;
}

}  // namespace

Model::~Model() { reset(); }

void Model::reset() {
    VerMin = VerMax = Vertex();
    mtl_tex.clear();
    part_mtl.clear();
    texs.clear();
    mtls.clear();
    parts.clear();
    octclparts.clear();
    borders.clear();
    bboxs.clear();
    vers.clear();
    nors.clear();
    txcs.clear();
}

// Whole-model border and rescale to a largest extent of 8 (Model.cpp:174-198).
// Points are multiplied by the reciprocal (Vertex::operator/=).
void Model::rescale(bool apply) {
    VerMin = Vertex(1000, 1000, 1000);
    VerMax = Vertex(-1000, -1000, -1000);
    for (size_t a = 0; a + 1 < borders.size(); a += 2) {
        border(VerMin, VerMax, borders[a]);
        border(VerMin, VerMax, borders[a + 1]);
    }
    if (!apply) return;
    Vertex Dif = VerMax - VerMin;
    float mdif = Dif.x > Dif.y ? Dif.x : Dif.y;
    mdif = (float)(Dif.z > mdif ? Dif.z / 8.0 : mdif / 8.0);
    VerMax /= mdif, VerMin /= mdif;
    for (Vertex &v : borders) v /= mdif;
    for (vector<Triangle> &vt : parts)
        for (Triangle &t : vt)
            for (Vertex &v : t.points) v /= mdif;
}

int32_t Model::loadMesh(const vector<MeshPart> &mesh, const vector<Material> &materials, bool rescaleTo8) {
    reset();
    mtls.push_back(Material());
    mtl_tex.push_back(-1);
    for (const Material &m : materials) {
        mtls.push_back(m);
        mtl_tex.push_back(-1);
    }
    for (const MeshPart &mp : mesh)
        if (mp.tris.size() > 32767) return -1;
    vers.push_back(Vertex());
    nors.push_back(Normal());
    txcs.push_back(Coord2D());
    for (const MeshPart &mp : mesh) {
        Vertex mn(1000, 1000, 1000), mx(-1000, -1000, -1000);
        for (const Triangle &t : mp.tris)
            for (const Vertex &p : t.points) border(mn, mx, p);
        parts.push_back(mp.tris);
        borders.push_back(mn);
        borders.push_back(mx);
        part_mtl.push_back(mp.mtl);
    }
    rescale(rescaleTo8);
    return (int32_t)parts.size();
}

int32_t Model::loadobj(const wstring &fname, const uint8_t) {
    Loader ldr(fname);
    if (!ldr.ok()) return -1;
    string ele[5];
    int ti[16] = {0};
    bool bFirstO = true;
    vector<Triangle> tris;
    Vertex v;
    Normal n;
    Coord2D tc;
    vers.push_back(v);
    nors.push_back(n);
    txcs.push_back(tc);
    VerMin = Vertex(1000, 1000, 1000);
    VerMax = Vertex(-1000, -1000, -1000);
    while (true) {
        int num = ldr.read(ele);
        if (num == INT8_MIN) break;
        if (num == -1) continue;
        switch (parseStr(ele[0])) {
        case key3('v', '*', '*'):
            v = Vertex((float)atof(ele[1].c_str()), (float)atof(ele[2].c_str()), (float)atof(ele[3].c_str()));
            vers.push_back(v);
            break;
        case key3('v', 'n', '*'):
            n = Normal((float)atof(ele[1].c_str()), (float)atof(ele[2].c_str()), (float)atof(ele[3].c_str()));
            nors.push_back(n);
            break;
        case key3('v', 't', '*'):
            tc = Coord2D((float)atof(ele[1].c_str()), (float)atof(ele[2].c_str()));
            txcs.push_back(tc);
            break;
        case key3('f', '*', '*'): {
            // A quad is split along its 0-2 diagonal (Model.cpp:88-117).
            const int corners = num > 4 ? 4 : 3;
            int pn = 0;
            for (int c = 0; c < corners; ++c) pn += Loader::parseInt(ele[1 + c], &ti[3 * c]);
            const bool textured = pn >= 3 * corners;
            auto V = [&](int c) { return at_or_default(vers, ti[3 * c]); };
            auto N = [&](int c) { return at_or_default(nors, ti[3 * c + 2]); };
            auto T = [&](int c) { return textured ? at_or_default(txcs, ti[3 * c + 1]) : Coord2D(); };
            tris.push_back(Triangle(V(0), N(0), T(0), V(1), N(1), T(1), V(2), N(2), T(2)));
            if (corners == 4) tris.push_back(Triangle(V(0), N(0), T(0), V(2), N(2), T(2), V(3), N(3), T(3)));
            for (int c = 0; c < corners; ++c) border(VerMin, VerMax, V(c));
            break;
        }
        case key3('u', 's', 'e'): {
            // Each `usemtl` after the first closes a part (Model.cpp:144-161).
            if (!bFirstO) {
                parts.push_back(std::move(tris));
                tris.clear();
                borders.push_back(VerMin);
                borders.push_back(VerMax);
            }
            VerMin = Vertex(1000, 1000, 1000);
            VerMax = Vertex(-1000, -1000, -1000);
            bFirstO = false;
            int8_t a = (int8_t)mtls.size();
            while (--a > 0)
                if (mtls[a].name == ele[1]) break;
            part_mtl.push_back(a);
            break;
        }
        }
    }
    if (bFirstO) part_mtl.push_back(0);
    parts.push_back(std::move(tris));
    borders.push_back(VerMin);
    borders.push_back(VerMax);
    rescale();
    return (int32_t)parts.size();
}

int32_t Model::loadmtl(const wstring &fname, const uint8_t code) {
    Loader ldr(fname);
    string ele[5];
    int8_t cur_tex_id = -1;
    bool bFirstM = true;
    Material mtl;
    mtls.push_back(mtl);
    mtl_tex.push_back(cur_tex_id);
    while (true) {
        int num = ldr.read(ele);
        if (num == INT8_MIN) break;
        if (num == -1) continue;
        switch (parseStr(ele[0])) {
        case key3('n', 'e', 'w'):
            // `newmtl` does not reset the running material (Model.cpp:229-239).
            if (!bFirstM) {
                mtls.push_back(mtl);
                mtl_tex.push_back(cur_tex_id);
            }
            mtl.name = ele[1];
            cur_tex_id = -1;
            bFirstM = false;
            break;
        case key3('K', 'a', '*'):
            mtl.SetMtl(MY_MODEL_AMBIENT, (float)atof(ele[1].c_str()), (float)atof(ele[2].c_str()),
                       (float)atof(ele[3].c_str()));
            break;
        case key3('K', 'd', '*'):
            mtl.SetMtl(MY_MODEL_DIFFUSE, (float)atof(ele[1].c_str()), (float)atof(ele[2].c_str()),
                       (float)atof(ele[3].c_str()));
            break;
        case key3('K', 's', '*'):
            mtl.SetMtl(MY_MODEL_SPECULAR, (float)atof(ele[1].c_str()), (float)atof(ele[2].c_str()),
                       (float)atof(ele[3].c_str()));
            break;
        case key3('K', 'e', '*'):
            mtl.SetMtl(MY_MODEL_EMISSION, (float)atof(ele[1].c_str()), (float)atof(ele[2].c_str()),
                       (float)atof(ele[3].c_str()));
            break;
        case key3('N', 's', '*'):
            mtl.SetMtl(MY_MODEL_SHINESS, 0, 0, 0, (float)atof(ele[1].c_str()));
            break;
        case key3('m', 'a', 'p'):
            if (ele[0] == "map_Kd****") {
                // The texture file is <basename>.bmp (Model.cpp:256-273).
                size_t loff = ele[1].find_last_of('\\'), roff = ele[1].find_last_of('.');
                ele[1] = ele[1].substr(loff + 1, roff - loff) + "bmp";
                int8_t a = (int8_t)texs.size();
                while (--a >= 0)
                    if (texs[a].name == ele[1]) {
                        cur_tex_id = a;
                        break;
                    }
                if (a < 0) {
                    loadtex(ele[1], code);
                    cur_tex_id = (int8_t)(texs.size() - 1);
                }
            }
            break;
        }
    }
    mtls.push_back(mtl);
    mtl_tex.push_back(cur_tex_id);
    return 0;
}

// 24-bit BMP texels, stored bottom-up in B,G,R order as read (Model.cpp:283-312).
int32_t Model::loadtex(const string &texname, const uint8_t) {
    FILE *fp = std::fopen(texname.c_str(), "rb");
    if (fp == nullptr) return -1;
    uint8_t fh[14], ih[40];
    if (std::fread(fh, 1, 14, fp) != 14 || fh[0] != 'B' || fh[1] != 'M' || std::fread(ih, 1, 40, fp) != 40) {
        std::fclose(fp);
        return -1;
    }
    auto u32 = [](const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; };
    uint32_t off = u32(fh + 10);
    int32_t width = (int32_t)u32(ih + 4), height = (int32_t)u32(ih + 8);
    int32_t size = (int32_t)u32(ih + 20);
    if (size == 0) size = width * height * 3;
    if (width <= 0 || height <= 0 || width > 32767 || height > 32767 || size <= 0) {
        std::fclose(fp);
        return -1;
    }
    std::vector<uint8_t> image((size_t)std::max<int64_t>(size, (int64_t)width * height * 3), 0);
    std::fseek(fp, off, SEEK_SET);
    size_t got = std::fread(image.data(), 1, (size_t)size, fp);
    (void)got;
    texs.push_back(Texture(texname, (int16_t)width, (int16_t)height, image.data()));
    std::fclose(fp);
    return 0;
}

int32_t Model::loadOBJ(const wstring &obj, const wstring &mtl, const uint8_t code) {
    this->objname = obj, this->mtlname = mtl;
    reset();
    loadmtl(mtl, code);
    if (loadobj(obj, code) < 0) return -1;
    return 1;
}

// Y-up -> Z-up turn (Model.cpp:351-393), including the reference's handling of
// the border boxes (z of min and max swapped) and of VerMin/VerMax (not swapped).
void Model::zRotate() {
    for (Vertex &v : vers) {
        std::swap(v.y, v.z);
        v.z *= -1;
    }
    for (Normal &n : nors) {
        std::swap(n.y, n.z);
        n.z *= -1;
    }
    for (vector<Triangle> &part : parts)
        for (Triangle &t : part)
            for (int a = 0; a < 3; ++a) {
                std::swap(t.points[a].y, t.points[a].z);
                std::swap(t.norms[a].y, t.norms[a].z);
                t.points[a].z *= -1;
                t.norms[a].z *= -1;
            }
    for (size_t a = 0; a + 1 < borders.size(); a += 2) {
        Vertex &va = borders[a], &vb = borders[a + 1];
        std::swap(va.y, va.z);
        std::swap(vb.y, vb.z);
        va.z *= -1;
        vb.z *= -1;
        std::swap(va.z, vb.z);
    }
    std::swap(VerMin.y, VerMin.z);
    VerMin.z *= -1;
    std::swap(VerMax.y, VerMax.z);
    VerMax.z *= -1;
}

void Model::SetMtl(const Material &m) {
    for (auto &old : mtls) old = m;
}
