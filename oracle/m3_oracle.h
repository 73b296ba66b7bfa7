// Restated leaf functions of the Maul3D step, plain C17 (test
// infrastructure only: the product never includes or links this).
// Each function names the reference routine it restates.  The full-step
// checker is the reference engine built from its own sources into
// oracle/_ref; these restatements pin the arithmetic of the leaf routines
// independently of both libraries.
#ifndef M3_ORACLE_H
#define M3_ORACLE_H

#include <stdint.h>

typedef struct { float x, y, z; } orVec3;

// FNV-1a 64 (src/core.c:46-54).
uint64_t or_fnv1a(uint64_t h, const void* bytes, int32_t count);

// Tangent basis from the axis of smallest |n_i| (src/manifold.c:19-39).
void or_tangent_basis(const float n[3], float t1[3], float t2[3]);

// Minimax atan2 (src/math_functions.c:14-49).
float or_atan2(float y, float x);

// Bhaskara I cosine/sine, renormalized (src/math_functions.c:54-96).
void or_cos_sin(float radians, float* c, float* s);

// Soft constraint coefficients (src/solver.c:30-41): out = bias rate,
// mass scale, impulse scale.
void or_make_soft(float hertz, float zeta, float h, float out[3]);

// Sphere-sphere manifold fields (src/manifold.c:41-62): returns the point
// count; fills normal, anchors and separation.
int or_collide_spheres(const float d[3], float rA, float rB, float normal[3], float anchorA[3],
                       float anchorB[3], float* separation);

// Plane-sphere manifold fields (src/manifold.c:64-85).
int or_collide_plane_sphere(const float n[3], float dist, float radius, float normal[3],
                            float anchorB[3], float* separation);

// Unit-density box mass and centroid inertia diagonal (src/hull.c:61-70).
void or_box_mass(float hx, float hy, float hz, float* mass, float inertiaDiag[3]);

// Sphere and capsule mass and centroid inertia (src/shape.c:114-161);
// inertia is the full 3x3, column major.
void or_sphere_mass(float density, float radius, float* mass, float inertia[9]);
void or_capsule_mass(float density, const float p1[3], const float p2[3], float radius,
                     float* mass, float com[3], float inertia[9]);

#endif
