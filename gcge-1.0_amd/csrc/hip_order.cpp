// hip_order.cpp -- a locality ordering of a sparse matrix's rows, computed
// on the host from the graph of its pattern alone (no coordinates).
//
// The SpMM kernels read, for every row block of 256 rows, the X rows its
// entries reference.  In the natural order of a 3-D grid those are the
// block's own rows plus whole neighbouring planes (about 3.3 times the
// block); in an order where each block is a compact piece of the grid they
// are the block plus its surface (about 1 times).  The order is built by
// recursive halving: a part of the graph is listed in breadth-first order
// from a vertex found at its periphery (the last vertex of a first
// breadth-first sweep), its first half and its second half become the next
// parts, and parts of at most `leaf` rows keep their breadth-first order.
// The halves of a graph region are two compact sub-regions next to each
// other, so neighbouring row blocks hold neighbouring pieces at every scale.
// Vertices a part's sweep does not reach (disconnected pieces) follow in
// their previous order.
#include <algorithm>
#include <vector>

#include "gcge_app_hip.h"

namespace {

struct Part {
    int start, len;
};

// BFS over the vertices of one part (tag[v] == t) from `root`, writing the
// visit order to out; returns the number of vertices reached.
int bfs(int root, int t, const int *rowptr, const int *col, const std::vector<int> &tag, std::vector<int> &seen,
        int stamp, int *out)
{
    int head = 0, tail = 0;
    out[tail++] = root;
    seen[root] = stamp;
    while (head < tail) {
        const int v = out[head++];
        for (int e = rowptr[v]; e < rowptr[v + 1]; e++) {
            const int c = col[e];
            if (tag[c] == t && seen[c] != stamp) {
                seen[c] = stamp;
                out[tail++] = c;
            }
        }
    }
    return tail;
}

}  // namespace

extern "C" GCGE_INT HIP_MatLocalityOrder(GCGE_INT n, const GCGE_INT *rowptr, const GCGE_INT *col, GCGE_INT leaf,
                                         GCGE_INT *perm)
{
    if (n <= 0 || !rowptr || !col || !perm) return 1;
    for (int i = 0; i < n; i++)
        if (rowptr[i + 1] < rowptr[i]) return 2;
    for (int e = 0; e < rowptr[n]; e++)
        if (col[e] < 0 || col[e] >= n) return 3;
    if (leaf < 1) leaf = 512;
    std::vector<int> order(n), tag(n, 0), seen(n, 0), tmp(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::vector<Part> stack{{0, n}};
    int next_tag = 0, stamp = 0;
    while (!stack.empty()) {
        const Part p = stack.back();
        stack.pop_back();
        if (p.len <= leaf) continue;
        const int t = ++next_tag;
        for (int i = p.start; i < p.start + p.len; i++) tag[order[i]] = t;
        // a peripheral vertex: the last one a sweep from the part's first vertex reaches
        int reached = bfs(order[p.start], t, rowptr, col, tag, seen, ++stamp, tmp.data());
        const int root = tmp[reached - 1];
        reached = bfs(root, t, rowptr, col, tag, seen, ++stamp, tmp.data());
        // unreached vertices of the part follow in their previous order
        for (int i = p.start; i < p.start + p.len && reached < p.len; i++)
            if (seen[order[i]] != stamp) {
                seen[order[i]] = stamp;
                tmp[reached++] = order[i];
            }
        std::copy(tmp.begin(), tmp.begin() + p.len, order.begin() + p.start);
        const int h = p.len / 2;
        stack.push_back({p.start + h, p.len - h});
        stack.push_back({p.start, h});
    }
    std::copy(order.begin(), order.end(), perm);
    return 0;
}
