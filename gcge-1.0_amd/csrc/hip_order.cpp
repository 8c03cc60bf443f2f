// hip_order.cpp -- a locality ordering of a sparse matrix's rows, computed
// on the host from the pattern alone (no coordinates are passed in).
//
// The SpMM kernels read, for every row block of 256 rows, the X rows its
// entries reference.  In the natural order of a 3-D grid those are the
// block's own rows plus whole neighbouring planes (about 2.4 times the
// block); in an order where each block is a compact piece of the grid they
// are the block plus its surface (about 0.9 times).
//
// Two methods, tried in this order:
// 1. A box grid numbered lexicographically (x fastest) is recognized from
//    the pattern: the frequent positive column offsets of a sample of rows
//    propose the x extent nx (and the plane size nx*ny); a candidate is
//    accepted when at least 97 % of all stored entries join rows whose grid
//    coordinates differ by at most 2 in every direction.  The rows are then
//    listed tile by tile (8 x 8 x 8 in 3-D, 16 x 16 in 2-D, clipped at the
//    box's faces), tiles in lexicographic order, rows within a tile
//    lexicographically.
// 2. Otherwise recursive halving: a part of the graph is listed in
//    breadth-first order from a vertex found at its periphery (the last
//    vertex of a first breadth-first sweep), its halves (split at a multiple
//    of `leaf` rows) become the next parts, and parts of at most `leaf` rows
//    keep their breadth-first order.  The halves of a graph region are two
//    compact sub-regions next to each other, so neighbouring row blocks hold
//    neighbouring pieces at every scale.  Vertices a part's sweep does not
//    reach (disconnected pieces) follow in their previous order.
// Either way the result is only a permutation: any order gives the same
// products up to the summation order of reductions over rows.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <map>
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

// Fraction of stored entries joining rows at most kReach apart in every
// grid direction, for a box nx * ny * nz numbered x fastest.
constexpr int kReach = 2;

double box_fit(int n, const int *rowptr, const int *col, int64_t nx, int64_t ny)
{
    const int64_t plane = nx * ny;
    int64_t good = 0;
    for (int r = 0; r < n; r++) {
        const int64_t x = r % nx, y = (r / nx) % ny, z = r / plane;
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
            const int64_t c = col[e];
            const int64_t dx = c % nx - x, dy = (c / nx) % ny - y, dz = c / plane - z;
            good += (std::llabs(dx) <= kReach && std::llabs(dy) <= kReach && std::llabs(dz) <= kReach);
        }
    }
    return rowptr[n] > 0 ? (double)good / (double)rowptr[n] : 0.0;
}

// Recognizes a lexicographically numbered box grid; fills dims = {nx, ny, nz}
// and returns its dimension (3 or 2), or 0.
int detect_box(int n, const int *rowptr, const int *col, int64_t dims[3])
{
    if (n < 64 || rowptr[n] == 0) return 0;
    // offsets col - row > 1 over a sample of about 65,536 rows
    const int step = std::max(1, n / 65536);
    std::map<int64_t, int64_t> count;
    int64_t rows = 0;
    for (int r = 0; r < n; r += step, rows++)
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
            const int64_t d = (int64_t)col[e] - r;
            if (d > 1) count[d]++;
        }
    std::vector<int64_t> freq;  // offsets present in at least 20 % of sampled rows
    for (const auto &kv : count)
        if (kv.second * 5 >= rows) freq.push_back(kv.first);
    if (freq.empty() || freq.size() > 32) return 0;
    int tried = 0;
    // 3-D: an x extent a and a plane size p, both frequent offsets
    for (int64_t a : freq)
        for (int64_t p : freq) {
            if (p <= a || p % a || n % p || p / a < 2 || n / p < 2 || tried >= 8) continue;
            tried++;
            if (box_fit(n, rowptr, col, a, p / a) >= 0.97) {
                dims[0] = a;
                dims[1] = p / a;
                dims[2] = n / p;
                return 3;
            }
        }
    for (int64_t a : freq) {
        if (n % a || n / a < 2 || tried >= 12) continue;
        tried++;
        if (box_fit(n, rowptr, col, a, n / a) >= 0.97) {
            dims[0] = a;
            dims[1] = n / a;
            dims[2] = 1;
            return 2;
        }
    }
    return 0;
}

// Tile-by-tile order of a box grid (tiles tx * ty * tz, clipped at the faces).
void tile_order(const int64_t dims[3], int tx, int ty, int tz, int *perm)
{
    const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
    int64_t pos = 0;
    for (int64_t z0 = 0; z0 < nz; z0 += tz)
        for (int64_t y0 = 0; y0 < ny; y0 += ty)
            for (int64_t x0 = 0; x0 < nx; x0 += tx)
                for (int64_t z = z0; z < std::min(nz, z0 + tz); z++)
                    for (int64_t y = y0; y < std::min(ny, y0 + ty); y++)
                        for (int64_t x = x0; x < std::min(nx, x0 + tx); x++)
                            perm[pos++] = (int)((z * ny + y) * nx + x);
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
    int64_t dims[3];
    const int nd = detect_box(n, rowptr, col, dims);
    if (nd == 3) {
        tile_order(dims, 8, 8, 8, perm);
        return 0;
    }
    if (nd == 2) {
        tile_order(dims, 16, 16, 1, perm);
        return 0;
    }
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
        // split at a multiple of the leaf size, so that leaves (and with
        // leaf = 256 the SpMM's row blocks) coincide
        int h = ((p.len / 2 + leaf / 2) / leaf) * leaf;
        if (h <= 0) h = leaf;
        if (h >= p.len) h = p.len / 2;
        stack.push_back({p.start + h, p.len - h});
        stack.push_back({p.start, h});
    }
    std::copy(order.begin(), order.end(), perm);
    return 0;
}
