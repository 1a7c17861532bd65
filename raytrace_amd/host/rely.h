// rely.h — headless platform header for the drop-in Scene / RayTracer API.
//
// Mirrors the role of the reference's rely.h (rely.h:1-51) without Win32,
// GLEW or freeglut: GL handle types become plain integers, and the SIMD
// feature cascade is replaced by scalar code that reproduces the live
// variants' operation order (3DElement.cpp:56-238).  All host code of this
// package is compiled with -ffp-contract=off.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <string>
#include <vector>
#include <algorithm>
#include <memory>
#include <thread>
#include <atomic>

#define PI 3.1415926535897932384
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_PI_2
#define M_PI_2 1.57079632679489661923
#endif

typedef unsigned int GLuint;
typedef int GLint;
typedef unsigned short GLushort;

using namespace std;
