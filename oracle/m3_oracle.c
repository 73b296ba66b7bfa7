// Restated leaf functions (see m3_oracle.h).  Test infrastructure: compiled
// by oracle/Makefile into oracle/libm3oracle.so and loaded only by tests.
// Built under the reference FP law (-ffp-contract=off, no fast math) so the
// float results are the reference's bits; every expression keeps the
// reference's evaluation order, which is the arithmetic being pinned.
#include "m3_oracle.h"

#include <math.h>
#include <string.h>

#define OR_PI 3.14159265358979323846f

uint64_t or_fnv1a(uint64_t h, const void* bytes, int32_t count)
{
    const unsigned char* p = (const unsigned char*)bytes;
    for (int32_t i = 0; i < count; ++i)
    {
        h ^= (uint64_t)p[i];
        h *= 0x100000001B3ull;
    }
    return h;
}

static float Abs(float a) { return a < 0.0f ? -a : a; }

static void Cross(const float a[3], const float b[3], float out[3])
{
    float x = a[1] * b[2] - a[2] * b[1];
    float y = a[2] * b[0] - a[0] * b[2];
    float z = a[0] * b[1] - a[1] * b[0];
    out[0] = x;
    out[1] = y;
    out[2] = z;
}

static void Normalize(const float v[3], float out[3])
{
    float len = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len < 1.19209290e-7f)
    {
        out[0] = 0.0f;
        out[1] = 1.0f;
        out[2] = 0.0f;
        return;
    }
    float inv = 1.0f / len;
    out[0] = inv * v[0];
    out[1] = inv * v[1];
    out[2] = inv * v[2];
}

void or_tangent_basis(const float n[3], float t1[3], float t2[3])
{
    float ax = Abs(n[0]);
    float ay = Abs(n[1]);
    float az = Abs(n[2]);
    float axis[3] = {0.0f, 0.0f, 0.0f};
    if (ax <= ay && ax <= az)
    {
        axis[0] = 1.0f;
    }
    else if (ay <= az)
    {
        axis[1] = 1.0f;
    }
    else
    {
        axis[2] = 1.0f;
    }
    float c[3];
    Cross(n, axis, c);
    Normalize(c, t1);
    Cross(n, t1, t2);
}

float or_atan2(float y, float x)
{
    if (x == 0.0f && y == 0.0f)
    {
        return 0.0f;
    }
    float ax = Abs(x);
    float ay = Abs(y);
    float mx = ay > ax ? ay : ax;
    float mn = ay < ax ? ay : ax;
    float a = mn / mx;
    float s = a * a;
    float c = s * a;
    float q = s * s;
    float r = 0.024840285f * q + 0.18681418f;
    float t = -0.094097948f * q - 0.33213072f;
    r = r * s + t;
    r = r * c + a;
    if (ay > ax)
    {
        r = 1.57079637f - r;
    }
    if (x < 0.0f)
    {
        r = 3.14159274f - r;
    }
    if (y < 0.0f)
    {
        r = -r;
    }
    return r;
}

void or_cos_sin(float radians, float* cOut, float* sOut)
{
    float x = remainderf(radians, 2.0f * OR_PI);
    float pi2 = OR_PI * OR_PI;
    float c;
    if (x < -0.5f * OR_PI)
    {
        float y = x + OR_PI;
        float y2 = y * y;
        c = -(pi2 - 4.0f * y2) / (pi2 + y2);
    }
    else if (x > 0.5f * OR_PI)
    {
        float y = x - OR_PI;
        float y2 = y * y;
        c = -(pi2 - 4.0f * y2) / (pi2 + y2);
    }
    else
    {
        float y2 = x * x;
        c = (pi2 - 4.0f * y2) / (pi2 + y2);
    }
    float s;
    if (x < 0.0f)
    {
        float y = x + OR_PI;
        s = -16.0f * y * (OR_PI - y) / (5.0f * pi2 - 4.0f * y * (OR_PI - y));
    }
    else
    {
        s = 16.0f * x * (OR_PI - x) / (5.0f * pi2 - 4.0f * x * (OR_PI - x));
    }
    float mag = sqrtf(s * s + c * c);
    float invMag = mag > 0.0f ? 1.0f / mag : 0.0f;
    *cOut = c * invMag;
    *sOut = s * invMag;
}

void or_make_soft(float hertz, float zeta, float h, float out[3])
{
    if (hertz == 0.0f)
    {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    float omega = 2.0f * OR_PI * hertz;
    float a1 = 2.0f * zeta + h * omega;
    float a2 = h * omega * a1;
    float a3 = 1.0f / (1.0f + a2);
    out[0] = omega / a1;
    out[1] = a2 * a3;
    out[2] = a3;
}

int or_collide_spheres(const float d[3], float rA, float rB, float normal[3], float anchorA[3],
                       float anchorB[3], float* separation)
{
    float distance = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float sep = distance - rA - rB;
    if (sep > 4.0f * 0.005f)
    {
        return 0;
    }
    Normalize(d, normal);
    for (int k = 0; k < 3; ++k)
    {
        anchorA[k] = rA * normal[k];
        anchorB[k] = -rB * normal[k];
    }
    *separation = sep;
    return 1;
}

int or_collide_plane_sphere(const float n[3], float dist, float radius, float normal[3],
                            float anchorB[3], float* separation)
{
    float sep = dist - radius;
    if (sep > 4.0f * 0.005f)
    {
        return 0;
    }
    for (int k = 0; k < 3; ++k)
    {
        normal[k] = n[k];
        anchorB[k] = -radius * n[k];
    }
    *separation = sep;
    return 1;
}

void or_box_mass(float hx, float hy, float hz, float* mass, float inertiaDiag[3])
{
    float m = 8.0f * hx * hy * hz;
    *mass = m;
    inertiaDiag[0] = (m / 3.0f) * (hy * hy + hz * hz);
    inertiaDiag[1] = (m / 3.0f) * (hx * hx + hz * hz);
    inertiaDiag[2] = (m / 3.0f) * (hx * hx + hy * hy);
}

void or_sphere_mass(float density, float radius, float* mass, float inertia[9])
{
    float m = density * (4.0f / 3.0f) * OR_PI * radius * radius * radius;
    float ic = 0.4f * m * radius * radius;
    memset(inertia, 0, 9 * sizeof(float));
    inertia[0] = ic;
    inertia[4] = ic;
    inertia[8] = ic;
    *mass = m;
}

void or_capsule_mass(float density, const float p1[3], const float p2[3], float r, float* mass,
                     float com[3], float inertia[9])
{
    float axis[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    float length = sqrtf(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    float u[3] = {1.0f, 0.0f, 0.0f};
    if (length > 0.0f)
    {
        float inv = 1.0f / length;
        u[0] = inv * axis[0];
        u[1] = inv * axis[1];
        u[2] = inv * axis[2];
    }
    float mCyl = density * OR_PI * r * r * length;
    float mSph = density * (4.0f / 3.0f) * OR_PI * r * r * r;
    float axial = 0.5f * mCyl * r * r + 0.4f * mSph * r * r;
    float perp = mCyl * (length * length / 12.0f + 0.25f * r * r) +
                 mSph * (0.4f * r * r + 0.25f * length * length + 0.375f * length * r);
    *mass = mCyl + mSph;
    for (int k = 0; k < 3; ++k)
    {
        com[k] = 0.5f * (p1[k] + p2[k]);
    }
    float d = axial - perp;
    // column major: inertia[3*col + row]
    inertia[0] = perp + d * u[0] * u[0];
    inertia[1] = d * u[0] * u[1];
    inertia[2] = d * u[0] * u[2];
    inertia[3] = d * u[0] * u[1];
    inertia[4] = perp + d * u[1] * u[1];
    inertia[5] = d * u[1] * u[2];
    inertia[6] = d * u[0] * u[2];
    inertia[7] = d * u[1] * u[2];
    inertia[8] = perp + d * u[2] * u[2];
}
