// scenegen.cpp — seeded synthetic render workloads, built through the Scene API.
//
// Random numbers: splitmix64, u = (x >> 40) * 2^-24 in [0, 1); uniforms are
// computed in double and rounded to float.  Draw order is fixed per workload
// and listed at each generator.  Every mesh enters through Model::loadMesh,
// i.e. with the reference's per-part borders and extent-8 rescale
// (Model.cpp:174-198) unless a workload says otherwise.
#include "scenegen.h"

#include <array>
#include <map>

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u() { return (double)(next() >> 40) * (1.0 / 16777216.0); }
    float uni(double a, double b) { return (float)(a + (b - a) * u()); }
    // integer uniform on [lo, hi]
    int range(int lo, int hi) { return lo + (int)(((next() >> 40) * (uint64_t)(hi - lo + 1)) >> 24); }
};

uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// 3D value noise: hashed lattice values in [-1, 1], smoothstep-weighted
// trilinear interpolation.
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
    uint64_t h = mix64(seed ^ ((uint64_t)x * 0x8CB92BA72F3D8DD7ull) ^ ((uint64_t)y * 0x9E3779B97F4A7C15ull) ^
                       ((uint64_t)z * 0xC2B2AE3D27D4EB4Full));
    return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
double vnoise(double x, double y, double z, uint64_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    double tx = x - fx, ty = y - fy, tz = z - fz;
    tx = tx * tx * (3 - 2 * tx);
    ty = ty * ty * (3 - 2 * ty);
    tz = tz * tz * (3 - 2 * tz);
    double c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    double x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    double x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    double y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}
// 3 octaves at f, 2f, 4f with amplitudes 1, 1/2, 1/4, normalised to [-1, 1].
double fbm3(double x, double y, double z, double f, uint64_t seed) {
    double s = vnoise(x * f, y * f, z * f, seed) + 0.5 * vnoise(x * 2 * f, y * 2 * f, z * 2 * f, seed + 1) +
               0.25 * vnoise(x * 4 * f, y * 4 * f, z * 4 * f, seed + 2);
    return s / 1.75;
}

struct D3 {
    double x, y, z;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double len(D3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
D3 unit(D3 a) { return a * (1.0 / len(a)); }

// Indexed mesh -> reference Triangles with area-weighted vertex normals.
struct IndexedMesh {
    std::vector<D3> pos;
    std::vector<std::array<uint32_t, 3>> tri;
    std::vector<int> part_of;  // part index per triangle
    int n_parts = 1;
};

std::vector<MeshPart> to_parts(const IndexedMesh &m, int8_t mtl) {
    std::vector<Vertex> fp(m.pos.size());
    for (size_t i = 0; i < m.pos.size(); ++i)
        fp[i] = Vertex((float)m.pos[i].x, (float)m.pos[i].y, (float)m.pos[i].z);
    std::vector<D3> acc(m.pos.size(), D3{0, 0, 0});
    for (const auto &t : m.tri) {
        D3 p0{fp[t[0]].x, fp[t[0]].y, fp[t[0]].z}, p1{fp[t[1]].x, fp[t[1]].y, fp[t[1]].z},
            p2{fp[t[2]].x, fp[t[2]].y, fp[t[2]].z};
        D3 n = cross(p1 - p0, p2 - p0);
        for (int k = 0; k < 3; ++k) acc[t[k]] = acc[t[k]] + n;
    }
    std::vector<Normal> nrm(m.pos.size());
    for (size_t i = 0; i < acc.size(); ++i) {
        double l = len(acc[i]);
        D3 u = l > 0 ? acc[i] * (1.0 / l) : D3{0, 1, 0};
        nrm[i] = Normal((float)u.x, (float)u.y, (float)u.z);
    }
    std::vector<MeshPart> parts(m.n_parts);
    for (auto &p : parts) p.mtl = mtl;
    for (size_t i = 0; i < m.tri.size(); ++i) {
        const auto &t = m.tri[i];
        parts[m.part_of[i]].tris.push_back(Triangle(fp[t[0]], nrm[t[0]], fp[t[1]], nrm[t[1]], fp[t[2]], nrm[t[2]]));
    }
    return parts;
}

// Geodesic icosphere of frequency F (20*F*F triangles, one part per
// icosahedron face), vertices displaced radially to r = 1 + amp * fbm3(p).
IndexedMesh icosphere(int F, double amp, double freq, uint64_t seed) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    const D3 base[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                         {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                              {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                              {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    D3 V[12];
    for (int i = 0; i < 12; ++i) V[i] = unit(base[i]);
    std::map<std::pair<int, int>, int> edge_id;
    for (auto &f : faces)
        for (int e = 0; e < 3; ++e) {
            int a = f[e], b = f[(e + 1) % 3];
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            if (!edge_id.count(key)) {
                int id = (int)edge_id.size();
                edge_id[key] = id;
            }
        }
    const int n_edge_pts = F - 1, n_int = (F - 1) * (F - 2) / 2;
    IndexedMesh m;
    m.n_parts = 20;
    m.pos.resize(12 + 30 * n_edge_pts + 20 * n_int);
    auto displace = [&](D3 p) {
        D3 u = unit(p);
        double r = 1.0 + amp * fbm3(u.x, u.y, u.z, freq, seed);
        return u * r;
    };
    for (int i = 0; i < 12; ++i) m.pos[i] = displace(V[i]);
    for (auto &kv : edge_id) {
        int a = kv.first.first, b = kv.first.second;
        for (int k = 1; k < F; ++k)
            m.pos[12 + kv.second * n_edge_pts + (k - 1)] =
                displace(V[a] * ((double)(F - k) / F) + V[b] * ((double)k / F));
    }
    auto int_index = [&](int i, int j) {  // i >= 1, j >= 1, i + j <= F - 1
        // rows of constant i hold F-1-i interior points
        int before = 0;
        for (int r = 1; r < i; ++r) before += F - 1 - r;
        return before + (j - 1);
    };
    std::vector<int> row_start(F + 1, 0);
    for (int i = 1; i < F; ++i) row_start[i] = int_index(i, 1);
    for (int f = 0; f < 20; ++f) {
        const D3 &A = V[faces[f][0]], &B = V[faces[f][1]], &C = V[faces[f][2]];
        for (int i = 1; i < F; ++i)
            for (int j = 1; i + j < F; ++j)
                m.pos[12 + 30 * n_edge_pts + f * n_int + row_start[i] + (j - 1)] =
                    displace(A * ((double)(F - i - j) / F) + B * ((double)i / F) + C * ((double)j / F));
    }
    auto edge_vertex = [&](int u, int v, int k) {  // k steps from u towards v, 0 < k < F
        int a = std::min(u, v), b = std::max(u, v);
        int kk = u == a ? k : F - k;
        return 12 + edge_id[std::make_pair(a, b)] * n_edge_pts + (kk - 1);
    };
    for (int f = 0; f < 20; ++f) {
        const int a = faces[f][0], b = faces[f][1], c = faces[f][2];
        auto vid = [&](int i, int j) -> uint32_t {
            if (i == 0 && j == 0) return a;
            if (i == F && j == 0) return b;
            if (i == 0 && j == F) return c;
            if (j == 0) return edge_vertex(a, b, i);
            if (i == 0) return edge_vertex(a, c, j);
            if (i + j == F) return edge_vertex(b, c, j);
            return 12 + 30 * n_edge_pts + f * n_int + row_start[i] + (j - 1);
        };
        for (int i = 0; i < F; ++i)
            for (int j = 0; i + j < F; ++j) {
                m.tri.push_back({vid(i, j), vid(i + 1, j), vid(i, j + 1)});
                m.part_of.push_back(f);
                if (i + j < F - 1) {
                    m.tri.push_back({vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)});
                    m.part_of.push_back(f);
                }
            }
    }
    return m;
}

// Height field over [-sx, sx] x [-sz, sz]: nx * nz vertices, 2 (nx-1)(nz-1)
// triangles, parts of whole quad rows holding at most 32767 triangles.
IndexedMesh heightfield(int nx, int nz, double sx, double sz, double amp, double freq, uint64_t seed) {
    IndexedMesh m;
    m.pos.resize((size_t)nx * nz);
    for (int j = 0; j < nz; ++j)
        for (int i = 0; i < nx; ++i) {
            double x = -sx + 2 * sx * i / (nx - 1), z = -sz + 2 * sz * j / (nz - 1);
            m.pos[(size_t)j * nx + i] = D3{x, amp * fbm3(x, 0.37, z, freq, seed), z};
        }
    const int per_row = 2 * (nx - 1);
    const int rows_per_part = std::max(1, 32767 / per_row);
    m.n_parts = (nz - 1 + rows_per_part - 1) / rows_per_part;
    for (int j = 0; j + 1 < nz; ++j)
        for (int i = 0; i + 1 < nx; ++i) {
            uint32_t v00 = j * nx + i, v10 = v00 + 1, v01 = v00 + nx, v11 = v01 + 1;
            m.tri.push_back({v00, v01, v10});
            m.tri.push_back({v10, v01, v11});
            m.part_of.push_back(j / rows_per_part);
            m.part_of.push_back(j / rows_per_part);
        }
    return m;
}

// UV sphere with `rings` x `sectors` quads (2 triangles each, degenerate at the poles).
IndexedMesh uvsphere(int rings, int sectors, double radius) {
    IndexedMesh m;
    for (int r = 0; r <= rings; ++r)
        for (int s = 0; s <= sectors; ++s) {
            double th = M_PI * r / rings, ph = 2 * M_PI * s / sectors;
            m.pos.push_back(D3{radius * std::sin(th) * std::cos(ph), radius * std::cos(th),
                               radius * std::sin(th) * std::sin(ph)});
        }
    for (int r = 0; r < rings; ++r)
        for (int s = 0; s < sectors; ++s) {
            uint32_t a = r * (sectors + 1) + s, b = a + 1, c = a + sectors + 1, d = c + 1;
            m.tri.push_back({a, b, c});
            m.tri.push_back({b, d, c});
            m.part_of.push_back(0);
            m.part_of.push_back(0);
        }
    return m;
}

// Random triangle soup in [-1, 1]^3, parts of at most 32767 triangles.
IndexedMesh soup(int64_t T, SplitMix64 &rng) {
    IndexedMesh m;
    const double s = 1.2 / std::cbrt((double)T);
    const int64_t per_part = 32767;
    m.n_parts = (int)((T + per_part - 1) / per_part);
    for (int64_t i = 0; i < T; ++i) {
        D3 c{2 * rng.u() - 1, 2 * rng.u() - 1, 2 * rng.u() - 1};
        uint32_t base = (uint32_t)m.pos.size();
        for (int k = 0; k < 3; ++k) m.pos.push_back(c + D3{s * (2 * rng.u() - 1), s * (2 * rng.u() - 1), s * (2 * rng.u() - 1)});
        m.tri.push_back({base, base + 1, base + 2});
        m.part_of.push_back((int)(i / per_part));
    }
    return m;
}

// Phong material of the workloads: kd ~ U[0,1]^3, ka = 0.2 kd, ks ~ U[0,1] grey.
Material random_material(SplitMix64 &rng, float reflect) {
    Material m;
    float kd[3] = {rng.uni(0, 1), rng.uni(0, 1), rng.uni(0, 1)};
    float ks = rng.uni(0, 1);
    m.SetMtl(MY_MODEL_DIFFUSE, kd[0], kd[1], kd[2]);
    m.SetMtl(MY_MODEL_AMBIENT, 0.2f * kd[0], 0.2f * kd[1], 0.2f * kd[2]);
    m.SetMtl(MY_MODEL_SPECULAR, ks, ks, ks);
    m.SetMtl(MY_MODEL_SHINESS, 32.0f);
    m.reflect = reflect;
    m.refract = 0.0f;
    return m;
}

// Point light as in main.cpp:168, placed at U[-8,8] x U[4,12] x U[-8,8].
void add_point_light(Scene &s, SplitMix64 &rng, bool bypass_cap) {
    float x = rng.uni(-8, 8), y = rng.uni(4, 12), z = rng.uni(-8, 8);
    if (!bypass_cap) {
        uint8_t li = s.AddLight(MY_LIGHT_POINT, Vertex(0.15f, 0.55f, 0.3f), Vertex(0.0f, 0.0f, 1.0f, 256));
        if (li == 0xff) return;
        s.Lights[li].SetProperty(MY_MODEL_POSITION, x, y, z, 1.0f);
        return;
    }
    // Past the reference's cap of 8 lights: build the light exactly as
    // AddLight does and append it to Scene::Lights directly.
    Light light(MY_LIGHT_POINT);
    Vertex comp(0.15f, 0.55f, 0.3f);
    float sum = comp.x + comp.y + comp.z;
    Vertex nc = comp / sum;
    light.SetProperty(MY_MODEL_AMBIENT, nc.x, nc.x, nc.x);
    light.SetProperty(MY_MODEL_DIFFUSE, nc.y, nc.y, nc.y);
    light.SetProperty(MY_MODEL_SPECULAR, nc.z, nc.z, nc.z);
    light.SetProperty(MY_MODEL_ATTENUATION, 0.0f, 0.0f, 1.0f);
    light.SetLumi(256);
    light.SetProperty(MY_MODEL_POSITION, x, y, z, 1.0f);
    s.Lights.push_back(light);
}

// Sphere via Scene::AddSphere, then placed at c and given a random material.
void add_random_sphere(Scene &s, SplitMix64 &rng, double lo, double hi, float reflect) {
    float r = rng.uni(0.3, 1.5);
    uint8_t idx = s.AddSphere(r);
    float cx = rng.uni(lo, hi), cy = rng.uni(lo, hi), cz = rng.uni(lo, hi);
    s.Objects[idx]->position = Vertex(cx, cy, cz);
    s.Objects[idx]->SetMtl(random_material(rng, reflect));
}

int pick(int v, int dflt) { return v > 0 ? v : dflt; }

}  // namespace

// Workload 1 (seed 1): 640x480, EnvLight 0.05, plane (reflect 0.3), 8 spheres
// (draws per sphere: r, cx, cy, cz, kd.rgb, ks), 2 point lights (x, y, z),
// maxLevel 3.
static Scene *gen_config1(const rtg_gen_opts &o, rtg_gen_info &info) {
    SplitMix64 rng(o.seed ? o.seed : 1);
    Scene *s = new Scene();
    info.width = pick(o.width, 640);
    info.height = pick(o.height, 480);
    info.max_level = o.max_level >= 0 ? o.max_level : 3;
    info.spp = 1;
    s->cam.resize(info.width, info.height);
    s->EnvLight = Vertex(0.05f, 0.05f, 0.05f, 1.0f);
    uint8_t pl = s->AddPlane();
    s->Objects[pl]->mtl.reflect = 0.3f;
    const int ns = o.n_spheres >= 0 ? o.n_spheres : 8;
    for (int k = 0; k < ns; ++k) add_random_sphere(*s, rng, -5, 5, 0.3f);
    const int nl = o.n_lights >= 0 ? o.n_lights : 2;
    for (int k = 0; k < nl; ++k) add_point_light(*s, rng, k >= 8);
    return s;
}

// Workload 2 (seed 2): 1920x1080, maxLevel 4.  Object 0: geodesic icosphere
// of frequency F = 100 (20 parts x 10,000 triangles), radially displaced by
// 0.1 * fbm3 (base frequency 4, noise seed derived from the workload seed),
// area-weighted normals, extent-8 rescale; its material is drawn first
// (kd.rgb, ks).  Then 64 spheres with centres in U[-8,8]^3 and 4 point lights.
static Scene *gen_config2(const rtg_gen_opts &o, rtg_gen_info &info) {
    const uint64_t seed = o.seed ? o.seed : 2;
    SplitMix64 rng(seed);
    Scene *s = new Scene();
    info.width = pick(o.width, 1920);
    info.height = pick(o.height, 1080);
    info.max_level = o.max_level >= 0 ? o.max_level : 4;
    info.spp = 1;
    s->cam.resize(info.width, info.height);
    s->EnvLight = Vertex(0.05f, 0.05f, 0.05f, 1.0f);
    Material mm = random_material(rng, 0.3f);
    const int F = pick(o.detail, 100);
    IndexedMesh ico = icosphere(F, 0.1, 4.0, mix64(seed * 0x1234567ull + 99));
    Model *model = new Model();
    model->loadMesh(to_parts(ico, 1), {mm});
    s->AddModel(model);
    const int ns = o.n_spheres >= 0 ? o.n_spheres : 64;
    for (int k = 0; k < ns; ++k) add_random_sphere(*s, rng, -8, 8, 0.3f);
    const int nl = o.n_lights >= 0 ? o.n_lights : 4;
    for (int k = 0; k < nl; ++k) add_point_light(*s, rng, k >= 8);
    return s;
}

// Workloads 3/4 (seed 3): resident scene `index` of 8.  Per scene (rng seeded
// with mix(seed, index)): T = round(10^(3 + 3u)) triangles (or `detail`), an
// even index is a triangle soup, an odd one a noise height field; then its
// material, U{1..8} point lights and a maxLevel in U{2..6}.
static Scene *gen_config3(const rtg_gen_opts &o, rtg_gen_info &info) {
    const uint64_t seed = o.seed ? o.seed : 3;
    SplitMix64 rng(mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)o.index));
    Scene *s = new Scene();
    info.width = pick(o.width, 256);
    info.height = pick(o.height, 256);
    info.spp = 1;
    s->cam.resize(info.width, info.height);
    s->EnvLight = Vertex(0.05f, 0.05f, 0.05f, 1.0f);
    int64_t T = (int64_t)std::llround(std::pow(10.0, 3.0 + 3.0 * rng.u()));
    if (o.detail > 0) T = o.detail;
    IndexedMesh mesh;
    if ((o.index & 1) == 0) {
        mesh = soup(T, rng);
    } else {
        int n = std::max(2, (int)std::llround(std::sqrt((double)T / 2.0)) + 1);
        mesh = heightfield(n, n, 1.0, 1.0, 0.25, 2.0, mix64(seed + 17 * (uint64_t)o.index));
    }
    Material mm = random_material(rng, 0.3f);
    Model *model = new Model();
    model->loadMesh(to_parts(mesh, 1), {mm});
    s->AddModel(model);
    const int nl = o.n_lights >= 0 ? o.n_lights : rng.range(1, 8);
    const int depth = rng.range(2, 6);
    info.max_level = o.max_level >= 0 ? o.max_level : depth;
    for (int k = 0; k < nl; ++k) add_point_light(*s, rng, k >= 8);
    return s;
}

// Workload 5 (seed 5): 3840x2160, 4 spp, maxLevel 8.  Height-field terrain of
// 2048x1024 vertices (4,188,162 triangles, 128 parts) over 80 x 70 units,
// then 256 models of a 32x32-quad UV sphere (2,048 triangles) with radius
// U[0.4,1.2] at U[-20,20] x U[0.5,3] x U[-40,5]; every second sphere refracts
// (refract 0.5, rfr U[1.3,1.7]).  Both kinds keep model units (no extent-8
// rescale).  16 point lights, past the reference's cap of 8.
static Scene *gen_config5(const rtg_gen_opts &o, rtg_gen_info &info) {
    const uint64_t seed = o.seed ? o.seed : 5;
    SplitMix64 rng(seed);
    Scene *s = new Scene();
    info.width = pick(o.width, 3840);
    info.height = pick(o.height, 2160);
    info.max_level = o.max_level >= 0 ? o.max_level : 8;
    info.spp = 4;
    s->cam.resize(info.width, info.height);
    s->EnvLight = Vertex(0.05f, 0.05f, 0.05f, 1.0f);
    const int div = pick(o.detail, 1);
    IndexedMesh terrain = heightfield(2048 / div, 1024 / div, 40.0, 35.0, 3.0, 0.15, mix64(seed + 5));
    for (D3 &p : terrain.pos) p.z -= 25.0;
    Material tm = random_material(rng, 0.1f);
    Model *tmodel = new Model();
    tmodel->loadMesh(to_parts(terrain, 1), {tm}, false);
    s->AddModel(tmodel);
    const int ns = o.n_spheres >= 0 ? o.n_spheres : 256;
    for (int k = 0; k < ns; ++k) {
        double r = 0.4 + 0.8 * rng.u();
        IndexedMesh sph = uvsphere(32, 32, r);
        Material m = random_material(rng, 0.3f);
        if (k % 2 == 1) {
            m.refract = 0.5f;
            m.rfr = rng.uni(1.3, 1.7);
        }
        Model *sm = new Model();
        sm->loadMesh(to_parts(sph, 1), {m}, false);
        sm->position = Vertex(rng.uni(-20, 20), rng.uni(0.5, 3.0), rng.uni(-40, 5));
        s->AddModel(sm);
    }
    const int nl = o.n_lights >= 0 ? o.n_lights : 16;
    for (int k = 0; k < nl; ++k) add_point_light(*s, rng, k >= 8);
    return s;
}

// Request stream of workloads 3/4: scene U{0..7}, 64 round(16^u) x
// 64 round(12^u) pixels, Poisson arrivals at `rate` (exponential gaps).
void rtg_make_requests(uint64_t seed, int n, double rate, rtg_gen_request *out) {
    SplitMix64 rng(mix64(seed ? seed : 3));
    double t = 0.0;
    for (int k = 0; k < n; ++k) {
        rtg_gen_request &q = out[k];
        q.scene = std::min(7, (int)(rng.u() * 8.0));
        q.width = 64 * (int)std::llround(std::pow(16.0, rng.u()));
        q.height = 64 * (int)std::llround(std::pow(12.0, rng.u()));
        q.pad = 0;
        if (rate > 0) t += -std::log(1.0 - rng.u()) / rate;
        q.arrival_s = rate > 0 ? t : 0.0;
    }
}

Scene *rtg_make_config_scene(int config, const rtg_gen_opts &opts, rtg_gen_info &info) {
    Scene *s = nullptr;
    switch (config) {
    case 1: s = gen_config1(opts, info); break;
    case 2: s = gen_config2(opts, info); break;
    case 3:
    case 4: s = gen_config3(opts, info); break;
    case 5: s = gen_config5(opts, info); break;
    default: return nullptr;
    }
    int64_t nt = 0;
    for (DrawObject *o : s->Objects)
        if (o->type == MY_OBJECT_MODEL)
            for (auto &p : static_cast<Model *>(o)->parts) nt += (int64_t)p.size();
    info.n_tris = nt;
    info.n_objects = (int32_t)s->Objects.size();
    info.n_lights = (int32_t)s->Lights.size();
    return s;
}
