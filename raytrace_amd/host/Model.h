// Model.h — headless triangle-mesh object of the drop-in API (reference
// Model.h:6-50).  Holds the same per-part triangle lists, border boxes and
// material tables as the reference.  The renderer builds a BVH on the device
// (librtg) and does not call RTPrepare; RTPrepare's octant bins
// (Model.cpp:402-480) serve intersect() on the host (intersect.cpp).
#pragma once
#include "3DElement.h"

struct SceneFlattener;

// One part of an in-memory mesh: model-space triangles and the index of its
// material in the Model's `mtls` table.
struct MeshPart {
    std::vector<Triangle> tris;
    int8_t mtl = 0;
};

class Model : public DrawObject {
private:
    Vertex VerMin, VerMax, BorderMin, BorderMax;
    friend struct SceneFlattener;
    int32_t loadobj(const wstring &objname, const uint8_t code);
    int32_t loadmtl(const wstring &mtlname, const uint8_t code);
    int32_t loadtex(const string &texname, const uint8_t code);
    void rescale(bool apply = true);
public:
    vector<Vertex> vers;
    vector<Normal> nors;
    vector<Coord2D> txcs;
    vector<vector<Triangle>> parts;
    vector<vector<clTri>> octclparts;  // RTPrepare: 8 bins per part, each ending in a sentinel
    vector<Vertex> borders;
    vector<Vertex> bboxs;              // RTPrepare: part border boxes + position
    vector<Material> mtls;
    vector<Texture> texs;
    vector<int8_t> part_mtl, mtl_tex;
    wstring objname, mtlname;

    Model(GLuint num = 0) : DrawObject(num) { type = MY_OBJECT_MODEL; }
    ~Model() override;
    void reset();
    // OBJ + MTL (+ BMP textures) from disk, with the reference's parsing rules
    // (Model.cpp:14-312).  Returns 1 (the reference's value) or -1 if the OBJ
    // file cannot be opened.
    int32_t loadOBJ(const wstring &objname, const wstring &mtlname, const uint8_t code = 0x0);
    // In-memory entry with the same post-processing as loadobj: per-part
    // borders and the rescale of the whole model to a largest extent of 8
    // (Model.cpp:174-198); `rescaleTo8 = false` keeps model units (the bounds
    // are still computed).  `materials` become mtls[1..]; mtls[0] is the
    // default Material as after loadmtl (Model.cpp:215-217).  Returns the
    // number of parts, or -1 if a part exceeds 32767 triangles.
    int32_t loadMesh(const vector<MeshPart> &mesh, const vector<Material> &materials, bool rescaleTo8 = true);
    void zRotate();
    virtual void SetMtl(const Material &mtl) override;
    void GLPrepare() override {}
    void RTPrepare() override;
    HitRes intersect(const Ray &ray, const HitRes &hr, const float min = 0) override;
    const Vertex &getVerMin() const { return VerMin; }
    const Vertex &getVerMax() const { return VerMax; }
};
