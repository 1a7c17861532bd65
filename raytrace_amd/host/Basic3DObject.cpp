// Basic3DObject.cpp — scene-side state of the built-in primitives.
#include "Basic3DObject.h"

Sphere::Sphere(const float r, GLuint lnum) : DrawObject(lnum) {
    type = MY_OBJECT_SPHERE;
    radius = r;
    radius_sqr = r * r;  // Basic3DObject.cpp:88-89
}

Box::Box(const float len, GLuint lnum) : DrawObject(lnum) {
    type = MY_OBJECT_CUBE;
    width = height = length = len;
    float l = len / 2;
    max = Vertex(l, l, l);
    min = max * -1;
}

// As in the reference (Basic3DObject.cpp:234-240) this constructor leaves
// `min` at the origin and negates `max`.
Box::Box(const float l, const float w, const float h, GLuint lnum) : DrawObject(lnum) {
    type = MY_OBJECT_CUBE;
    length = l, width = w, height = h;
    max = Vertex(l / 2, w / 2, h / 2);
    max = max * -1;
}

Plane::Plane(GLuint lnum) : DrawObject(lnum) {
    type = MY_OBJECT_PLANE;
    tex = Texture(true);
    rotate(Vertex(0, 36, 0));
}

// Orientation from spherical angles (Basic3DObject.cpp:312-330): the plane's
// position is dis along the spherical direction, its normal points back to the
// origin, and axisy/axisx span the plane.
void Plane::rotate(const Vertex &v) {
    bool fix = false;
    ang.x = mod(360 + ang.x - v.y * 5, 360.0f);
    ang.y = mod(360 + ang.y - v.x * 5, 360.0f);
    ang.z += v.z;
    if (ang.z < 0.0f) ang.z = 0.0f;
    if (std::fabs(ang.z) < 1e-5f) fix = true, ang.z = 1;
    Coord_sph2car2(ang.x, ang.y, ang.z, position);
    normal = Normal(position * -1);
    if (fix) ang.z = 0, position = Vertex();
    float tangy = mod(90 + ang.x, 360.0f);
    Coord_sph2car2(tangy, ang.y, 1, axisy);
    axisx = axisy * normal;
}

BallPlane::BallPlane(const float r, GLuint lnum) : DrawObject(lnum) {
    type = MY_OBJECT_BALLPLANE;
    radius = r;
    radius_sqr = r * r;
    rotate(Vertex(0, 36, 0));
}

void BallPlane::rotate(const Vertex &v) {
    bool fix = false;
    ang.x = mod(360 + ang.x - v.y * 5, 360.0f);
    ang.y = mod(360 + ang.y - v.x * 5, 360.0f);
    ang.z += v.z;
    if (ang.z < 0.0f) ang.z = 0.0f;
    if (std::fabs(ang.z) < 1e-5f) fix = true, ang.z = 1;
    Coord_sph2car2(ang.x, ang.y, ang.z, position);
    normal = Normal(position * -1);
    if (fix) ang.z = 0, position = Vertex();
    float tangy = mod(90 + ang.x, 360.0f);
    Coord_sph2car2(tangy, ang.y, 1, axisy);
    axisx = axisy * normal;
}
