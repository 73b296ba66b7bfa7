// Host-side hull construction shared by the world (box and point-cloud
// hulls).  Internal to the library.
#ifndef M3_HULL_H
#define M3_HULL_H

#include "m3_core.h"

#ifdef __cplusplus
extern "C" {
#endif

#define M3_HULL_MAX_INPUT 256 // points accepted by the point-cloud builder

// Half-edge adjacency (twins at 2k, 2k+1) and the vertex centroid from the
// face loops (src/hull.c:77-137).
void m3HostHullHalfEdges(m3HullData* hull);

// Convex hull of 4..256 points (QuickHull with face merging) plus its
// unit-density mass properties; 0 for a degenerate cloud or one over the
// hull caps (src/quickhull.c:1110-1216).
int m3HostComputeHull(const m3Vec3* points, int32_t count, m3HullData* out);

#ifdef __cplusplus
}
#endif

#endif
