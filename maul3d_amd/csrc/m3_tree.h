// The host mirror of the reference's broadphase tree (src/dynamic_tree.h,
// src/dynamic_tree.c).  The device step never builds a tree: its pair list
// is a sweep over per-shape leaf boxes (the tree's leaves).  The tree itself
// is needed for two things only, and both need its exact node layout:
//
//  * snapshot bytes (src/snapshot.c:244-247 walk the node array and the
//    proxy ids; the header carries the root and the free list), and
//  * the order in which the continuous pass meets its tree candidates
//    (src/solver.c:3402, a depth-first query whose order is a function of
//    the tree's shape).
//
// The host replays the reference's operation history on the same node
// array: insert at shape creation (src/shape.c:592-621), remove and
// reinsert of every shape whose box escaped its leaf, in shape order, each
// step (src/broad_phase.c:505-539), and the canonical rebuild
// (src/dynamic_tree.c:439-564).  The node record and every branch decision
// are the reference's, so the node array is byte-identical.
#ifndef M3_TREE_H
#define M3_TREE_H

#include <stdbool.h>
#include <stdint.h>

#define M3_TREE_NULL (-1)

typedef struct m3TreeNode // src/dynamic_tree.h:18-28
{
    double lo[3];
    double hi[3];
    int32_t parent; // next free node while on the free chain
    int32_t child1;
    int32_t child2;
    int32_t userData; // shape slot of a leaf
    int32_t height;   // -1 free, 0 leaf
    int32_t pad;
} m3TreeNode;

_Static_assert(sizeof(m3TreeNode) == 72, "m3TreeNode must be 72 padding-free bytes");

typedef struct m3Tree
{
    m3TreeNode* nodes;
    int32_t capacity;
    int32_t root;
    int32_t freeList;
} m3Tree;

// A tree of `capacity` free nodes chained in ascending order; false when the
// host is out of memory.
bool m3TreeCreate(m3Tree* tree, int32_t capacity);
void m3TreeDestroy(m3Tree* tree);
// Puts every node back on an ascending free chain (the empty tree).
void m3TreeReset(m3Tree* tree);

// Leaf insertion by the surface-area descent with AVL repair; returns the
// leaf node, or M3_TREE_NULL when the node pool is exhausted.
int32_t m3TreeInsert(m3Tree* tree, const double lo[3], const double hi[3], int32_t userData);
void m3TreeRemove(m3Tree* tree, int32_t leaf);

// Top-down rebuild from leaves given in canonical (shape) order; outNodes
// receives each input's leaf node.  False when 2 * count - 1 nodes do not
// fit.
bool m3TreeRebuild(m3Tree* tree, const double (*los)[3], const double (*his)[3],
                   const int32_t* userDatas, int32_t count, int32_t* outNodes);

// Depth-first leaf order of a query that overlaps every node (child1 before
// child2, the order of src/dynamic_tree.c:367-406): rank[userData] = the
// leaf's position; every other entry of rank[0..rankCount) is set to -1.
// Any query visits the leaves it reaches in this order, because a pruned
// subtree only removes leaves and the stack never overflows when the full
// walk does not.  Returns the deepest stack the walk used.
int32_t m3TreeLeafRanks(const m3Tree* tree, int32_t* rank, int32_t rankCount);

// Height of the root (0 for a single leaf, -1 for an empty tree).
int32_t m3TreeHeight(const m3Tree* tree);

// True when the node array (as a restore brings it in) is a well-formed
// tree: root and free chain in range, every reachable node reached once,
// internal nodes with two in-range children pointing back and heights one
// above the taller child, leaves of height 0, free nodes of height -1, and
// the reachable nodes plus the free chain covering the pool exactly.
bool m3TreeValid(const m3Tree* tree);

#endif // M3_TREE_H
