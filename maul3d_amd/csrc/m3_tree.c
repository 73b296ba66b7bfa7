// Host mirror of the reference's broadphase tree (header: m3_tree.h).
//
// The node array must come out byte-identical to the reference's after the
// same operation history, so the allocation discipline (LIFO free chain, the
// order nodes are taken and returned), the surface-area descent
// (src/dynamic_tree.c:221-307) with its strict comparisons, the splice, the
// AVL rotations (src/dynamic_tree.c:90-200, including which grandchild moves
// where) and the removal (src/dynamic_tree.c:324-358) follow the reference
// step for step.  Bounds arithmetic is double with compare-and-select min and
// max, as there.
#include "m3_tree.h"

#include <stdlib.h>
#include <string.h>

static double Area(const double lo[3], const double hi[3])
{
    const double ex = hi[0] - lo[0];
    const double ey = hi[1] - lo[1];
    const double ez = hi[2] - lo[2];
    return 2.0 * (ex * ey + ey * ez + ez * ex);
}

static void Merge(const double aLo[3], const double aHi[3], const double bLo[3], const double bHi[3],
                  double lo[3], double hi[3])
{
    for (int k = 0; k < 3; ++k)
    {
        lo[k] = aLo[k] < bLo[k] ? aLo[k] : bLo[k];
        hi[k] = aHi[k] > bHi[k] ? aHi[k] : bHi[k];
    }
}

static int32_t MaxI(int32_t a, int32_t b) { return a > b ? a : b; }

static void ChainFree(m3Tree* t)
{
    for (int32_t i = 0; i < t->capacity; ++i)
    {
        t->nodes[i].parent = i + 1 < t->capacity ? i + 1 : M3_TREE_NULL;
        t->nodes[i].height = -1;
    }
    t->freeList = t->capacity > 0 ? 0 : M3_TREE_NULL;
    t->root = M3_TREE_NULL;
}

bool m3TreeCreate(m3Tree* t, int32_t capacity)
{
    memset(t, 0, sizeof(*t));
    t->nodes = (m3TreeNode*)calloc((size_t)(capacity > 0 ? capacity : 1), sizeof(m3TreeNode));
    if (t->nodes == NULL)
    {
        return false;
    }
    t->capacity = capacity;
    ChainFree(t);
    return true;
}

void m3TreeDestroy(m3Tree* t)
{
    free(t->nodes);
    memset(t, 0, sizeof(*t));
    t->root = M3_TREE_NULL;
    t->freeList = M3_TREE_NULL;
}

void m3TreeReset(m3Tree* t)
{
    memset(t->nodes, 0, (size_t)t->capacity * sizeof(m3TreeNode));
    ChainFree(t);
}

// Takes the head of the free chain (src/dynamic_tree.c:62-78).
static int32_t Take(m3Tree* t)
{
    const int32_t id = t->freeList;
    if (id == M3_TREE_NULL)
    {
        return M3_TREE_NULL;
    }
    m3TreeNode* n = &t->nodes[id];
    t->freeList = n->parent;
    n->parent = M3_TREE_NULL;
    n->child1 = M3_TREE_NULL;
    n->child2 = M3_TREE_NULL;
    n->height = 0;
    n->userData = -1;
    n->pad = 0;
    return id;
}

// Wipes a node and pushes it on the free chain (src/dynamic_tree.c:80-87).
static void Give(m3Tree* t, int32_t id)
{
    m3TreeNode* n = &t->nodes[id];
    memset(n, 0, sizeof(*n));
    n->parent = t->freeList;
    n->height = -1;
    t->freeList = id;
}

// Points the parent of `old` (or the root) at `repl`.
static void Relink(m3Tree* t, int32_t parent, int32_t old, int32_t repl)
{
    if (parent == M3_TREE_NULL)
    {
        t->root = repl;
    }
    else if (t->nodes[parent].child1 == old)
    {
        t->nodes[parent].child1 = repl;
    }
    else
    {
        t->nodes[parent].child2 = repl;
    }
}

// One rotation at node a when its children's heights differ by more than one
// (src/dynamic_tree.c:90-200): the taller child `up` takes a's place, a
// becomes up's child1, and of up's two children the taller stays with up as
// its child2 while the other replaces `up` under a.  Returns the subtree's
// new root.
static int32_t Rotate(m3Tree* t, int32_t ia)
{
    m3TreeNode* nodes = t->nodes;
    m3TreeNode* a = &nodes[ia];
    if (a->height < 2 || a->child1 == M3_TREE_NULL)
    {
        return ia;
    }
    const int32_t ib = a->child1;
    const int32_t ic = a->child2;
    const int32_t balance = nodes[ic].height - nodes[ib].height;
    if (balance >= -1 && balance <= 1)
    {
        return ia;
    }
    // up: the taller child; stay: the shorter child; slot: which of a's
    // child fields `up` occupied (2 = child2 when c rotates up).
    const bool cUp = balance > 1;
    const int32_t iu = cUp ? ic : ib;
    const int32_t is = cUp ? ib : ic;
    m3TreeNode* u = &nodes[iu];
    m3TreeNode* s = &nodes[is];
    const int32_t i1 = u->child1;
    const int32_t i2 = u->child2;
    u->child1 = ia;
    u->parent = a->parent;
    a->parent = iu;
    Relink(t, u->parent, ia, iu);
    // The taller grandchild stays under u; ties keep child2 (the reference's
    // strict '>' on the first grandchild's height).
    const bool firstTaller = nodes[i1].height > nodes[i2].height;
    const int32_t keep = firstTaller ? i1 : i2;
    const int32_t move = firstTaller ? i2 : i1;
    u->child2 = keep;
    if (cUp)
    {
        a->child2 = move;
    }
    else
    {
        a->child1 = move;
    }
    nodes[move].parent = ia;
    Merge(s->lo, s->hi, nodes[move].lo, nodes[move].hi, a->lo, a->hi);
    Merge(a->lo, a->hi, nodes[keep].lo, nodes[keep].hi, u->lo, u->hi);
    a->height = 1 + MaxI(s->height, nodes[move].height);
    u->height = 1 + MaxI(a->height, nodes[keep].height);
    return iu;
}

// Rebalance and refit from `index` to the root (src/dynamic_tree.c:202-219).
static void Refit(m3Tree* t, int32_t index)
{
    m3TreeNode* nodes = t->nodes;
    while (index != M3_TREE_NULL)
    {
        index = Rotate(t, index);
        m3TreeNode* n = &nodes[index];
        const m3TreeNode* c1 = &nodes[n->child1];
        const m3TreeNode* c2 = &nodes[n->child2];
        n->height = 1 + MaxI(c1->height, c2->height);
        Merge(c1->lo, c1->hi, c2->lo, c2->hi, n->lo, n->hi);
        index = n->parent;
    }
}

// Cost of descending into child c when the new leaf's box is (lo, hi)
// (src/dynamic_tree.c:257-279): a leaf costs the merged area, an internal
// node the merged area minus its own, both plus the inherited growth.
static double ChildCost(const m3TreeNode* c, const double lo[3], const double hi[3], double inherit)
{
    double mlo[3];
    double mhi[3];
    Merge(c->lo, c->hi, lo, hi, mlo, mhi);
    if (c->height == 0)
    {
        return Area(mlo, mhi) + inherit;
    }
    return Area(mlo, mhi) - Area(c->lo, c->hi) + inherit;
}

int32_t m3TreeInsert(m3Tree* t, const double lo[3], const double hi[3], int32_t userData)
{
    const int32_t leaf = Take(t);
    if (leaf == M3_TREE_NULL)
    {
        return M3_TREE_NULL;
    }
    m3TreeNode* nodes = t->nodes;
    memcpy(nodes[leaf].lo, lo, sizeof(double) * 3);
    memcpy(nodes[leaf].hi, hi, sizeof(double) * 3);
    nodes[leaf].userData = userData;
    nodes[leaf].height = 0;
    if (t->root == M3_TREE_NULL)
    {
        t->root = leaf;
        return leaf;
    }
    // Surface-area descent: stop where a new parent here is strictly
    // cheaper than both descents; otherwise take the cheaper child (child2
    // on a tie).
    int32_t at = t->root;
    while (nodes[at].height > 0)
    {
        double mlo[3];
        double mhi[3];
        Merge(nodes[at].lo, nodes[at].hi, lo, hi, mlo, mhi);
        const double merged = Area(mlo, mhi);
        const double here = 2.0 * merged;
        const double inherit = 2.0 * (merged - Area(nodes[at].lo, nodes[at].hi));
        const int32_t c1 = nodes[at].child1;
        const int32_t c2 = nodes[at].child2;
        const double cost1 = ChildCost(&nodes[c1], lo, hi, inherit);
        const double cost2 = ChildCost(&nodes[c2], lo, hi, inherit);
        if (here < cost1 && here < cost2)
        {
            break;
        }
        at = cost1 < cost2 ? c1 : c2;
    }
    // A new parent spliced above the chosen sibling: sibling child1, leaf
    // child2 (src/dynamic_tree.c:283-319).
    const int32_t sibling = at;
    const int32_t above = nodes[sibling].parent;
    const int32_t parent = Take(t);
    if (parent == M3_TREE_NULL)
    {
        Give(t, leaf);
        return M3_TREE_NULL;
    }
    nodes[parent].parent = above;
    nodes[parent].userData = -1;
    Merge(lo, hi, nodes[sibling].lo, nodes[sibling].hi, nodes[parent].lo, nodes[parent].hi);
    nodes[parent].height = nodes[sibling].height + 1;
    Relink(t, above, sibling, parent);
    nodes[parent].child1 = sibling;
    nodes[parent].child2 = leaf;
    nodes[sibling].parent = parent;
    nodes[leaf].parent = parent;
    Refit(t, parent);
    return leaf;
}

void m3TreeRemove(m3Tree* t, int32_t leaf)
{
    m3TreeNode* nodes = t->nodes;
    if (leaf == t->root)
    {
        t->root = M3_TREE_NULL;
        Give(t, leaf);
        return;
    }
    const int32_t parent = nodes[leaf].parent;
    const int32_t grand = nodes[parent].parent;
    const int32_t sibling = nodes[parent].child1 == leaf ? nodes[parent].child2 : nodes[parent].child1;
    if (grand != M3_TREE_NULL)
    {
        Relink(t, grand, parent, sibling);
        nodes[sibling].parent = grand;
        Give(t, parent);
        Refit(t, grand);
    }
    else
    {
        t->root = sibling;
        nodes[sibling].parent = M3_TREE_NULL;
        Give(t, parent);
    }
    Give(t, leaf);
}

// ---------------------------------------------------------------------------
// Canonical rebuild (src/dynamic_tree.c:411-564): median split on the longest
// centroid axis, stable insertion sort (ties keep the lower input first),
// nodes taken from a fresh ascending chain in recursion order.
// ---------------------------------------------------------------------------

typedef struct Build
{
    const double (*los)[3];
    const double (*his)[3];
    const int32_t* userDatas;
    int32_t* order;
    int32_t* outNodes;
} Build;

static double Centre(const Build* b, int32_t input, int axis)
{
    return 0.5 * (b->los[input][axis] + b->his[input][axis]);
}

static void SortRange(Build* b, int32_t first, int32_t end, int axis)
{
    for (int32_t i = first + 1; i < end; ++i)
    {
        const int32_t key = b->order[i];
        const double c = Centre(b, key, axis);
        int32_t j = i - 1;
        while (j >= first)
        {
            const double cj = Centre(b, b->order[j], axis);
            if (cj < c || (cj == c && b->order[j] < key))
            {
                break;
            }
            b->order[j + 1] = b->order[j];
            j -= 1;
        }
        b->order[j + 1] = key;
    }
}

static int32_t BuildRange(m3Tree* t, Build* b, int32_t first, int32_t end)
{
    const int32_t id = Take(t);
    if (id == M3_TREE_NULL)
    {
        return M3_TREE_NULL;
    }
    if (end - first == 1)
    {
        const int32_t input = b->order[first];
        m3TreeNode* n = &t->nodes[id];
        memcpy(n->lo, b->los[input], sizeof(double) * 3);
        memcpy(n->hi, b->his[input], sizeof(double) * 3);
        n->userData = b->userDatas[input];
        n->height = 0;
        b->outNodes[input] = id;
        return id;
    }
    double clo[3];
    double chi[3];
    for (int k = 0; k < 3; ++k)
    {
        clo[k] = Centre(b, b->order[first], k);
        chi[k] = clo[k];
    }
    for (int32_t i = first + 1; i < end; ++i)
    {
        for (int k = 0; k < 3; ++k)
        {
            const double c = Centre(b, b->order[i], k);
            clo[k] = c < clo[k] ? c : clo[k];
            chi[k] = c > chi[k] ? c : chi[k];
        }
    }
    int axis = 0;
    if (chi[1] - clo[1] > chi[0] - clo[0])
    {
        axis = 1;
    }
    if (chi[2] - clo[2] > chi[axis] - clo[axis])
    {
        axis = 2;
    }
    SortRange(b, first, end, axis);
    const int32_t mid = first + (end - first) / 2;
    const int32_t c1 = BuildRange(t, b, first, mid);
    const int32_t c2 = BuildRange(t, b, mid, end);
    m3TreeNode* n = &t->nodes[id];
    n->child1 = c1;
    n->child2 = c2;
    t->nodes[c1].parent = id;
    t->nodes[c2].parent = id;
    Merge(t->nodes[c1].lo, t->nodes[c1].hi, t->nodes[c2].lo, t->nodes[c2].hi, n->lo, n->hi);
    n->height = 1 + MaxI(t->nodes[c1].height, t->nodes[c2].height);
    return id;
}

bool m3TreeRebuild(m3Tree* t, const double (*los)[3], const double (*his)[3],
                   const int32_t* userDatas, int32_t count, int32_t* outNodes)
{
    if (count < 0 || 2 * count - 1 > t->capacity)
    {
        return false;
    }
    // Every node back on an ascending chain with canonical contents.
    for (int32_t i = 0; i < t->capacity; ++i)
    {
        m3TreeNode* n = &t->nodes[i];
        memset(n, 0, sizeof(*n));
        n->parent = i + 1 < t->capacity ? i + 1 : M3_TREE_NULL;
        n->height = -1;
        n->child1 = M3_TREE_NULL;
        n->child2 = M3_TREE_NULL;
    }
    t->freeList = t->capacity > 0 ? 0 : M3_TREE_NULL;
    t->root = M3_TREE_NULL;
    if (count == 0)
    {
        return true;
    }
    int32_t* order = (int32_t*)malloc((size_t)count * sizeof(int32_t));
    if (order == NULL)
    {
        return false;
    }
    for (int32_t i = 0; i < count; ++i)
    {
        order[i] = i;
    }
    Build b = {los, his, userDatas, order, outNodes};
    t->root = BuildRange(t, &b, 0, count);
    free(order);
    return t->root != M3_TREE_NULL;
}

int32_t m3TreeLeafRanks(const m3Tree* t, int32_t* rank, int32_t rankCount)
{
    for (int32_t i = 0; i < rankCount; ++i)
    {
        rank[i] = -1;
    }
    if (t->root == M3_TREE_NULL)
    {
        return 0;
    }
    // The reference's walk: an explicit stack, child2 pushed before child1.
    int32_t stack[256];
    int32_t top = 0;
    int32_t deepest = 0;
    int32_t next = 0;
    stack[top++] = t->root;
    while (top > 0)
    {
        const m3TreeNode* n = &t->nodes[stack[--top]];
        if (n->height == 0)
        {
            if (n->userData >= 0 && n->userData < rankCount)
            {
                rank[n->userData] = next;
            }
            next += 1;
            continue;
        }
        if (top + 2 > 256)
        {
            return 257; // deeper than any walk of the reference can go
        }
        stack[top++] = n->child2;
        stack[top++] = n->child1;
        deepest = top > deepest ? top : deepest;
    }
    return deepest;
}

int32_t m3TreeHeight(const m3Tree* t)
{
    return t->root == M3_TREE_NULL ? -1 : t->nodes[t->root].height;
}

bool m3TreeValid(const m3Tree* t)
{
    const int32_t n = t->capacity;
    if (t->root < M3_TREE_NULL || t->root >= n || t->freeList < M3_TREE_NULL || t->freeList >= n)
    {
        return false;
    }
    uint8_t* seen = (uint8_t*)calloc((size_t)(n > 0 ? n : 1), 1);
    int32_t* stack = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    bool ok = seen != NULL && stack != NULL;
    int32_t top = 0;
    int32_t visited = 0;
    if (ok && t->root != M3_TREE_NULL)
    {
        if (t->nodes[t->root].parent != M3_TREE_NULL)
        {
            ok = false;
        }
        stack[top++] = t->root;
    }
    while (ok && top > 0)
    {
        const int32_t id = stack[--top];
        if (seen[id] != 0)
        {
            ok = false;
            break;
        }
        seen[id] = 1;
        visited += 1;
        const m3TreeNode* node = &t->nodes[id];
        if (node->height == 0)
        {
            continue;
        }
        const int32_t c1 = node->child1;
        const int32_t c2 = node->child2;
        if (node->height < 0 || c1 < 0 || c1 >= n || c2 < 0 || c2 >= n || c1 == c2 ||
            t->nodes[c1].parent != id || t->nodes[c2].parent != id ||
            node->height != 1 + MaxI(t->nodes[c1].height, t->nodes[c2].height) || top + 2 > n)
        {
            ok = false;
            break;
        }
        stack[top++] = c1;
        stack[top++] = c2;
    }
    for (int32_t id = t->freeList; ok && id != M3_TREE_NULL; id = t->nodes[id].parent)
    {
        if (id < 0 || id >= n || seen[id] != 0 || t->nodes[id].height != -1)
        {
            ok = false;
            break;
        }
        seen[id] = 1;
        visited += 1;
    }
    ok = ok && visited == n;
    free(seen);
    free(stack);
    return ok;
}
