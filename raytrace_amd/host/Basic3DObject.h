// Basic3DObject.h — headless Sphere / Box / Plane / BallPlane of the drop-in API
// (reference Basic3DObject.h:5-68).  The renderer intersects them on the
// device (librtg); intersect() answers one ray on the host (intersect.cpp).
#pragma once
#include "3DElement.h"

struct SceneFlattener;

class Sphere : public DrawObject {
private:
    float radius, radius_sqr;
    friend struct SceneFlattener;
public:
    Sphere(const float r = 1.0f, GLuint lnum = 0);
    ~Sphere() override {}
    float getRadius() const { return radius; }
    void GLPrepare() override {}
    HitRes intersect(const Ray &ray, const HitRes &hr, const float min = 0) override;
};

class Box : public DrawObject {
private:
    float width, height, length;
    Vertex min, max;
    friend struct SceneFlattener;
public:
    Box(const float len = 2.0, GLuint lnum = 0);
    Box(const float l, const float w, const float h, GLuint lnum = 0);
    void GLPrepare() override {}
    HitRes intersect(const Ray &ray, const HitRes &hr, const float dmin = 0) override;
};

class Plane : public DrawObject {
private:
    Vertex ang;
    Normal axisx, axisy;
    Texture tex;
    GLuint texList = 0;
    friend struct SceneFlattener;
public:
    Normal normal;
    Plane(GLuint lnum = 0);
    void rotate(const Vertex &v);
    void setTex(const Texture &tex) { this->tex = tex; }
    const Normal &getAxisX() const { return axisx; }
    const Normal &getAxisY() const { return axisy; }
    void GLPrepare() override {}
    HitRes intersect(const Ray &ray, const HitRes &hr, const float min = 0) override;
};

class BallPlane : public DrawObject {
private:
    Vertex ang;
    Normal axisx, axisy;
    float radius, radius_sqr;
    friend struct SceneFlattener;
public:
    Normal normal;
    BallPlane(const float r = 0.3f, GLuint lnum = 0);
    ~BallPlane() override {}
    void rotate(const Vertex &v);
    void GLPrepare() override {}
    HitRes intersect(const Ray &ray, const HitRes &hr, const float min = 0) override;
};
