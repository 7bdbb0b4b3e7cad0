/*
 * oracle.cpp -- CPU restatement of uAstar's CPU solvers (test infrastructure only;
 * see oracle.h for the rules on who may use it and for the reference line map).
 *
 * Costs are carried exactly as (orth, diag) step counts, value = orth + diag*sqrt(2),
 * and compared with integer arithmetic (cmp_pair), so the oracle shares no cost
 * encoding with the product's fixed-point keys.
 */
#include "oracle.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <queue>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* utils.hpp:119-122 */
const float SQRT2F = 1.4142135623731f;
const int DX[8] = {1, -1, 0, 0, 1, 1, -1, -1};
const int DY[8] = {0, 0, 1, -1, 1, -1, 1, -1};

/* sign of (a1 + b1*sqrt2) - (a2 + b2*sqrt2), exact */
int cmp_pair(int64_t a1, int64_t b1, int64_t a2, int64_t b2) {
    __int128 x = (__int128)a1 - a2, y = (__int128)b1 - b2;
    if (x >= 0 && y >= 0) return (x || y) ? 1 : 0;
    if (x <= 0 && y <= 0) return (x || y) ? -1 : 0;
    if (x > 0) return (x * x > 2 * y * y) ? 1 : -1;   /* y < 0 */
    return (2 * y * y > x * x) ? 1 : -1;              /* x < 0, y > 0 */
}

struct Pair {
    int64_t a, b;
};
inline bool pair_less(const Pair &p, const Pair &q) { return cmp_pair(p.a, p.b, q.a, q.b) < 0; }

/* octile heuristic as an exact pair (CPU-solver.cpp:105-112 uses min*SQRT2 + |dx-dy|) */
inline Pair octile(int x, int y, int ex, int ey) {
    int dx = std::abs(x - ex), dy = std::abs(y - ey);
    return Pair{std::abs(dx - dy), std::min(dx, dy)};
}

}  // namespace

/* ------------------------------------------------------------------ generators */

uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint64_t orc_cell_hash(uint64_t seed, uint64_t id) {
    return orc_mix64(orc_mix64(seed) + id * 0x9E3779B97F4A7C15ULL);
}

void orc_grid_bernoulli(int H, int W, uint64_t seed, uint64_t thr, int sx, int sy, int ex, int ey,
                        uint8_t *bytes) {
    for (int64_t x = 0; x < H; ++x)
        for (int64_t y = 0; y < W; ++y) {
            uint64_t id = (uint64_t)(x * W + y);
            bytes[id] = ((orc_cell_hash(seed, id) >> 32) < thr) ? 0 : 0xFF;
        }
    bytes[(int64_t)sx * W + sy] = 0xFF;
    bytes[(int64_t)ex * W + ey] = 0xFF;
}

/* pathway/input/zigzag.cpp:14-40 */
void orc_grid_zigzag(int H, int W, int gap, uint8_t *bytes, int *sx, int *sy, int *ex, int *ey) {
    *sx = 0;
    *sy = W / 2;
    *ex = H - 1;
    *ey = W / 2;
    bool left = false;
    uint8_t *buf = bytes;
    for (int i = 0; i < H; ++i) {
        if (i && i % gap == 0) {
            if (left) *buf++ = 0xFF;
            for (int j = 0; j < W - 1; ++j) *buf++ = 0;
            if (!left) *buf++ = 0xFF;
            left = !left;
        } else {
            for (int j = 0; j < W; ++j) *buf++ = 0xFF;
        }
    }
}

int orc_puzzle_solvable(int n, const uint8_t *tiles) {
    int nn = n * n;
    std::vector<int> p(nn);
    int blank = -1;
    for (int i = 0; i < nn; ++i) {
        p[i] = tiles[i] ? tiles[i] - 1 : nn - 1;  /* goal index of the item in cell i */
        if (!tiles[i]) blank = i;
    }
    std::vector<char> seen(nn, 0);
    int cycles = 0;
    for (int i = 0; i < nn; ++i) {
        if (seen[i]) continue;
        ++cycles;
        for (int j = i; !seen[j]; j = p[j]) seen[j] = 1;
    }
    int perm_parity = (nn - cycles) & 1;
    int man = (n - 1 - blank / n) + (n - 1 - blank % n);
    return perm_parity == (man & 1);
}

void orc_puzzle_random(int n, uint64_t seed, uint8_t *tiles) {
    int nn = n * n;
    for (int i = 0; i < nn; ++i) tiles[i] = (uint8_t)i;
    uint64_t s = seed;
    for (int i = nn - 1; i >= 1; --i) {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t r = orc_mix64(s);
        int j = (int)(r % (uint64_t)(i + 1));
        std::swap(tiles[i], tiles[j]);
    }
    if (!orc_puzzle_solvable(n, tiles)) {
        int a = -1, b = -1;
        for (int i = 0; i < nn; ++i)
            if (tiles[i]) {
                if (a < 0) a = i;
                else if (b < 0) b = i;
            }
        std::swap(tiles[a], tiles[b]);
    }
}

/* ------------------------------------------------------------------ grid A* */

namespace {

struct GNode {
    int id;
    int64_t a, b;
    float fdist;
    GNode *prev;
};

int grid_ref_containers(const uint8_t *g, int H, int W, int sx, int sy, int ex, int ey, bool f32,
                        orc_grid_out *out, int32_t *path, int64_t cap) {
    /* CPU-solver.cpp:29-44 (initialize) and :46-90 (solve) */
    struct Item {
        Pair f;
        float ff;
        GNode *node;
    };
    auto worse = [f32](const Item &p, const Item &q) {
        if (f32) return p.ff > q.ff;
        int c = cmp_pair(p.f.a, p.f.b, q.f.a, q.f.b);
        if (c) return c > 0;
        return p.node->id > q.node->id;
    };
    std::priority_queue<Item, std::vector<Item>, decltype(worse)> open(worse);
    std::unordered_set<int> closed;
    std::unordered_map<int, GNode *> global;
    std::vector<GNode *> owned;
    auto fval = [&](GNode *nd) {
        int x = nd->id / W, y = nd->id % W;
        Pair h = octile(x, y, ex, ey);
        Item it;
        it.f = Pair{nd->a + h.a, nd->b + h.b};
        int dx = std::abs(x - ex), dy = std::abs(y - ey);
        it.ff = nd->fdist + std::min(dx, dy) * SQRT2F + std::abs(dx - dy);  /* CPU-solver.cpp:111 */
        it.node = nd;
        return it;
    };
    int target = ex * W + ey;
    GNode *start = new GNode{sx * W + sy, 0, 0, 0.0f, nullptr};
    owned.push_back(start);
    global[start->id] = start;
    open.push(fval(start));
    out->inserted = 1;
    GNode *goal = nullptr;
    while (!open.empty()) {
        GNode *node = nullptr;
        while (!open.empty()) { /* B14: the reference calls top() without this guard */
            node = open.top().node;
            open.pop();
            if (!closed.count(node->id)) break;
            node = nullptr;
        }
        if (!node) break;
        closed.insert(node->id);
        out->expanded++;
        if (node->id == target) {
            goal = node;
            break;
        }
        int x = node->id / W, y = node->id % W;
        for (int i = 0; i < 8; ++i) {
            if (~g[node->id] & (1 << i)) continue;
            int nx = x + DX[i], ny = y + DY[i];
            if (!(0 <= nx && nx < H && 0 <= ny && ny < W)) continue;
            int nid = nx * W + ny;
            out->generated++;
            int64_t na = node->a + (i < 4), nb = node->b + (i >= 4);
            float nf = node->fdist + (i < 4 ? 1.0f : SQRT2F);
            auto it = global.find(nid);
            if (it == global.end()) {
                GNode *nn = new GNode{nid, na, nb, nf, node};
                owned.push_back(nn);
                global[nid] = nn;
                open.push(fval(nn));
                out->inserted++;
            } else {
                GNode *on = it->second;
                bool better = f32 ? (nf < on->fdist) : (cmp_pair(na, nb, on->a, on->b) < 0);
                if (better) {
                    on->a = na;
                    on->b = nb;
                    on->fdist = nf;
                    on->prev = node;
                    open.push(fval(on));
                    out->inserted++;
                }
            }
        }
    }
    if (goal) {
        out->found = 1;
        out->orth = goal->a;
        out->diag = goal->b;
        out->cost_f32 = f32 ? goal->fdist : (float)goal->a + (float)goal->b * SQRT2F;
        std::vector<int> ids;
        for (GNode *p = goal; p; p = p->prev) ids.push_back(p->id);
        std::reverse(ids.begin(), ids.end());
        out->path_len = (int64_t)ids.size();
        if (path && (int64_t)ids.size() <= cap) std::copy(ids.begin(), ids.end(), path);
    }
    for (GNode *p : owned) delete p;
    return 0;
}

/* Same A* on dense arrays: best pair per cell, closed bit, parent direction. */
int grid_dense(const uint8_t *g, int H, int W, int sx, int sy, int ex, int ey, bool dijkstra,
               orc_grid_out *out, int32_t *path, int64_t cap) {
    int64_t n = (int64_t)H * W;
    const int32_t INF = INT32_MAX;
    std::vector<int32_t> ba(n, INF), bb(n, INF);
    std::vector<uint8_t> closed(n, 0), pdir(n, 0xFF);
    struct Item {
        int64_t fa, fb;
        int64_t id;
    };
    auto worse = [](const Item &p, const Item &q) {
        int c = cmp_pair(p.fa, p.fb, q.fa, q.fb);
        if (c) return c > 0;
        return p.id > q.id;
    };
    std::priority_queue<Item, std::vector<Item>, decltype(worse)> open(worse);
    auto push = [&](int64_t id) {
        int x = (int)(id / W), y = (int)(id % W);
        Pair h = dijkstra ? Pair{0, 0} : octile(x, y, ex, ey);
        open.push(Item{ba[id] + h.a, bb[id] + h.b, id});
        out->inserted++;
    };
    int64_t s = (int64_t)sx * W + sy, t = (int64_t)ex * W + ey;
    ba[s] = 0;
    bb[s] = 0;
    push(s);
    bool found = false;
    while (!open.empty()) {
        Item it = open.top();
        open.pop();
        int64_t id = it.id;
        if (closed[id]) continue;
        closed[id] = 1;
        out->expanded++;
        if (id == t) {
            found = true;
            break;
        }
        int x = (int)(id / W), y = (int)(id % W);
        for (int i = 0; i < 8; ++i) {
            if (~g[id] & (1 << i)) continue;
            int nx = x + DX[i], ny = y + DY[i];
            if (!(0 <= nx && nx < H && 0 <= ny && ny < W)) continue;
            int64_t nid = (int64_t)nx * W + ny;
            out->generated++;
            int64_t na = ba[id] + (i < 4), nb = bb[id] + (i >= 4);
            if (ba[nid] == INF || cmp_pair(na, nb, ba[nid], bb[nid]) < 0) {
                ba[nid] = (int32_t)na;
                bb[nid] = (int32_t)nb;
                pdir[nid] = (uint8_t)i;
                if (!closed[nid]) push(nid);
            }
        }
    }
    if (found) {
        out->found = 1;
        out->orth = ba[t];
        out->diag = bb[t];
        out->cost_f32 = (float)ba[t] + (float)bb[t] * SQRT2F;
        std::vector<int32_t> ids;
        int64_t c = t;
        while (true) {
            ids.push_back((int32_t)c);
            if (c == s) break;
            int d = pdir[c];
            int x = (int)(c / W) - DX[d], y = (int)(c % W) - DY[d];
            c = (int64_t)x * W + y;
        }
        std::reverse(ids.begin(), ids.end());
        out->path_len = (int64_t)ids.size();
        if (path && (int64_t)ids.size() <= cap) std::copy(ids.begin(), ids.end(), path);
    }
    return 0;
}

}  // namespace

int orc_grid_astar(const uint8_t *graph, int H, int W, int sx, int sy, int ex, int ey, int variant,
                   orc_grid_out *out, int32_t *path, int64_t path_cap) {
    if (!graph || !out || H <= 0 || W <= 0) return -1;
    if (sx < 0 || sx >= H || sy < 0 || sy >= W || ex < 0 || ex >= H || ey < 0 || ey >= W) return -1;
    std::memset(out, 0, sizeof(*out));
    switch (variant) {
    case 0: return grid_ref_containers(graph, H, W, sx, sy, ex, ey, false, out, path, path_cap);
    case 1: return grid_dense(graph, H, W, sx, sy, ex, ey, false, out, path, path_cap);
    case 2: return grid_ref_containers(graph, H, W, sx, sy, ex, ey, true, out, path, path_cap);
    case 3: return grid_dense(graph, H, W, sx, sy, ex, ey, true, out, path, path_cap);
    default: return -1;
    }
}

/* ------------------------------------------------------------------ pattern databases */

namespace {

struct Pattern {
    int n, nn, k;
    std::vector<int> tracked;
    std::vector<int> map;        /* tile number -> index in tracked (database.cpp:12-16) */
    std::vector<uint64_t> mult;  /* mult[i] = prod_{m=i}^{k-1} (nn-m); mult[k] = 1 */
    uint64_t size;
    Pattern(int n_, const int *t, int k_) : n(n_), nn(n_ * n_), k(k_), tracked(t, t + k_) {
        map.assign(nn + 1, -1);
        mult.assign(k + 1, 1);
        size = 1;
        for (int i = 0; i < k; ++i) {
            map[tracked[i]] = i;
            size *= (uint64_t)(nn - i);
        }
        for (int i = k - 1; i >= 0; --i) mult[i] = mult[i + 1] * (uint64_t)(nn - i);
    }
    /* database.cpp:140-163, with #{i<j: idx_i < idx_j} taken by popcount */
    uint32_t rank(const int *pos) const {
        uint32_t mask = 0;
        uint64_t code = 0;
        for (int j = 0; j < k; ++j) {
            int p = pos[j];
            int v = p - __builtin_popcount(mask & ((1u << p) - 1));
            code += (uint64_t)v * mult[j + 1];
            mask |= 1u << p;
        }
        return (uint32_t)code;
    }
    /* database.cpp:165-199 */
    void unrank(uint32_t code, int *pos) const {
        uint32_t used = 0;
        uint64_t c = code;
        for (int j = 0; j < k; ++j) {
            int v = (int)(c / mult[j + 1]);
            c %= mult[j + 1];
            int cnt = 0;
            while (used >> cnt & 1) ++cnt;
            for (int s = 0; s < v; ++s) {
                ++cnt;
                while (used >> cnt & 1) ++cnt;
            }
            pos[j] = cnt;
            used |= 1u << cnt;
        }
    }
    void positions(const uint8_t *conf, int *pos) const {
        for (int i = 0; i < nn; ++i)
            if (conf[i] && map[conf[i]] >= 0) pos[map[conf[i]]] = i;
    }
};

/* database.cpp:37-84: queue BFS from the goal pattern (tile v on cell v-1) */
void pdb_bfs_queue(const Pattern &P, uint8_t *out) {
    std::memset(out, 0xFF, P.size);
    std::vector<uint64_t> visited((P.size + 63) / 64, 0);
    std::queue<uint32_t> q;
    int pos[25];
    for (int j = 0; j < P.k; ++j) pos[j] = P.tracked[j] - 1;
    uint32_t c0 = P.rank(pos);
    q.push(c0);
    out[c0] = 0;
    visited[c0 >> 6] |= 1ULL << (c0 & 63);
    while (!q.empty()) {
        uint32_t code = q.front();
        q.pop();
        P.unrank(code, pos);
        uint32_t occ = 0;
        for (int j = 0; j < P.k; ++j) occ |= 1u << pos[j];
        /* database.cpp:59-79 iterates cells in row-major order and the 4 moves in DX/DY order */
        int order[25];
        for (int j = 0; j < P.k; ++j) order[j] = j;
        std::sort(order, order + P.k, [&](int a, int b) { return pos[a] < pos[b]; });
        for (int oj = 0; oj < P.k; ++oj) {
            int j = order[oj];
            int x = pos[j] / P.n, y = pos[j] % P.n;
            for (int d = 0; d < 4; ++d) {
                int nx = x + DX[d], ny = y + DY[d];
                if (!(0 <= nx && nx < P.n && 0 <= ny && ny < P.n)) continue;
                int np = nx * P.n + ny;
                if (occ >> np & 1) continue;
                int save = pos[j];
                pos[j] = np;
                uint32_t nc = P.rank(pos);
                pos[j] = save;
                if (!(visited[nc >> 6] >> (nc & 63) & 1)) {
                    visited[nc >> 6] |= 1ULL << (nc & 63);
                    q.push(nc);
                    out[nc] = (uint8_t)(out[code] + 1);
                }
            }
        }
    }
}

/* Level-synchronous BFS: same distances, parallel over the frontier scan. */
void pdb_bfs_levels(const Pattern &P, uint8_t *out) {
    std::memset(out, 0xFF, P.size);
    int pos0[25];
    for (int j = 0; j < P.k; ++j) pos0[j] = P.tracked[j] - 1;
    out[P.rank(pos0)] = 0;
    for (int d = 0; d < 254; ++d) {
        long long added = 0;
        int64_t sz = (int64_t)P.size;
#pragma omp parallel for schedule(dynamic, 65536) reduction(+ : added)
        for (int64_t c = 0; c < sz; ++c) {
            if (__atomic_load_n(&out[c], __ATOMIC_RELAXED) != d) continue;
            int pos[25];
            P.unrank((uint32_t)c, pos);
            uint32_t occ = 0;
            for (int j = 0; j < P.k; ++j) occ |= 1u << pos[j];
            for (int j = 0; j < P.k; ++j) {
                int x = pos[j] / P.n, y = pos[j] % P.n;
                for (int dd = 0; dd < 4; ++dd) {
                    int nx = x + DX[dd], ny = y + DY[dd];
                    if (!(0 <= nx && nx < P.n && 0 <= ny && ny < P.n)) continue;
                    int np = nx * P.n + ny;
                    if (occ >> np & 1) continue;
                    int save = pos[j];
                    pos[j] = np;
                    uint32_t nc = P.rank(pos);
                    pos[j] = save;
                    if (__atomic_load_n(&out[nc], __ATOMIC_RELAXED) == 0xFF) {
                        __atomic_store_n(&out[nc], (uint8_t)(d + 1), __ATOMIC_RELAXED);
                        ++added;
                    }
                }
            }
        }
        if (!added) break;
    }
}

/* puzzle/CPU-solver.hpp:44-104 */
int partitions(int n, std::vector<std::vector<int>> &t) {
    t.clear();
    if (n == 3) {
        t.push_back({1, 2, 3, 4, 5, 6, 7, 8});
    } else if (n == 4) {
        t.push_back({1, 2, 3, 4, 5, 6, 9, 13});
        t.push_back({7, 8, 10, 11, 12, 14, 15});
    } else if (n == 5) {
        t.push_back({3, 4, 5, 10, 15, 20});
        t.push_back({2, 1, 6, 11, 16, 21});
        t.push_back({7, 8, 9, 12, 17, 22});
        t.push_back({13, 14, 18, 19, 23, 24});
    } else {
        return -1;
    }
    return (int)t.size();
}

}  // namespace

uint64_t orc_pdb_size(int n, int k) {
    uint64_t s = 1;
    for (int i = 0; i < k; ++i) s *= (uint64_t)(n * n - i);
    return s;
}

uint32_t orc_pdb_encode(int n, const int *tracked, int k, const uint8_t *conf) {
    Pattern P(n, tracked, k);
    int pos[25];
    P.positions(conf, pos);
    return P.rank(pos);
}

void orc_pdb_decode(int n, const int *tracked, int k, uint32_t code, uint8_t *conf) {
    Pattern P(n, tracked, k);
    int pos[25];
    P.unrank(code, pos);
    std::memset(conf, 0, (size_t)n * n);
    for (int j = 0; j < k; ++j) conf[pos[j]] = (uint8_t)tracked[j];
}

int orc_pdb_build(int n, const int *tracked, int k, int method, uint8_t *out) {
    if (n < 2 || n > 5 || k < 1 || k > n * n - 1) return -1;
    Pattern P(n, tracked, k);
    if (method == 0) pdb_bfs_queue(P, out);
    else pdb_bfs_levels(P, out);
    return 0;
}

int orc_partitions(int n, int *counts, int *tiles) {
    std::vector<std::vector<int>> t;
    int c = partitions(n, t);
    if (c < 0) return c;
    for (int i = 0; i < c; ++i) {
        counts[i] = (int)t[i].size();
        for (size_t j = 0; j < t[i].size(); ++j) tiles[i * 8 + (int)j] = t[i][j];
    }
    return c;
}

/* ------------------------------------------------------------------ puzzle search */

namespace {

typedef unsigned __int128 State;  /* 5 bits per cell, cell i at bits 5i..5i+4 */

struct PuzzleCtx {
    int n, nn;
    std::vector<Pattern> pats;
    std::vector<uint64_t> offs;
    std::vector<int> pat_of, idx_of;  /* tile -> (pattern, index): CPU-solver.hpp:108-125 */
    const uint8_t *pdb;
    PuzzleCtx(int n_, const uint8_t *pdb_) : n(n_), nn(n_ * n_), pdb(pdb_) {
        std::vector<std::vector<int>> t;
        partitions(n, t);
        pat_of.assign(nn, -1);
        idx_of.assign(nn, -1);
        uint64_t off = 0;
        for (size_t i = 0; i < t.size(); ++i) {
            pats.emplace_back(n, t[i].data(), (int)t[i].size());
            offs.push_back(off);
            off += pats.back().size;
            for (size_t j = 0; j < t[i].size(); ++j) {
                pat_of[t[i][j]] = (int)i;
                idx_of[t[i][j]] = (int)j;
            }
        }
    }
    /* CPU-solver.hpp:287-314 */
    int h(const uint8_t *conf) const {
        int pos[4][25];
        for (int i = 0; i < nn; ++i)
            if (conf[i]) pos[pat_of[conf[i]]][idx_of[conf[i]]] = i;
        int r = 0;
        for (size_t p = 0; p < pats.size(); ++p) r += pdb[offs[p] + pats[p].rank(pos[p])];
        return r;
    }
};

inline State pack(const uint8_t *c, int nn) {
    State s = 0;
    for (int i = 0; i < nn; ++i) s |= (State)c[i] << (5 * i);
    return s;
}
inline void unpack(State s, uint8_t *c, int nn) {
    for (int i = 0; i < nn; ++i) c[i] = (uint8_t)((s >> (5 * i)) & 31);
}
struct StateHash {
    size_t operator()(const State &s) const {
        uint64_t lo = (uint64_t)s, hi = (uint64_t)(s >> 64);
        return (size_t)orc_mix64(lo ^ orc_mix64(hi));
    }
};

int blank_of(const uint8_t *c, int nn) {
    for (int i = 0; i < nn; ++i)
        if (!c[i]) return i;
    return -1;
}

/* CPU-solver.hpp:154-260 */
int puzzle_astar(const PuzzleCtx &P, const uint8_t *tiles, int64_t limit, int64_t *expanded,
                 std::vector<int> &moves) {
    struct Node {
        State ps;
        Node *prev;
        int f, g;
    };
    struct HeapItem {
        int f;
        Node *node;
        bool operator>(const HeapItem &r) const { return f > r.f; }
    };
    std::priority_queue<HeapItem, std::vector<HeapItem>, std::greater<HeapItem>> open;
    std::unordered_map<State, Node *, StateHash> closed;
    std::vector<Node *> owned;
    int n = P.n, nn = P.nn;
    uint8_t conf[25], goal[25];
    for (int i = 0; i < nn; ++i) goal[i] = (uint8_t)(i + 1);
    goal[nn - 1] = 0;
    State target = pack(goal, nn);
    Node *root = new Node{pack(tiles, nn), nullptr, P.h(tiles), 0};
    owned.push_back(root);
    open.push(HeapItem{root->f, root});
    closed[root->ps] = root;
    Node *best = nullptr;
    int64_t exp = 0;
    int rc = -1;
    while (!open.empty()) {
        HeapItem now{0, nullptr};
        while (!open.empty()) {
            now = open.top();
            open.pop();
            if (closed[now.node->ps]->f == now.f) break;  /* CPU-solver.hpp:162-165 */
            now.node = nullptr;
        }
        if (!now.node) break;
        Node *node = now.node;
        ++exp;
        if (node->ps == target) {
            best = node;
            break;
        }
        if (limit > 0 && exp > limit) {
            rc = -3;
            break;
        }
        unpack(node->ps, conf, nn);
        int b = blank_of(conf, nn);
        int x = b / n, y = b % n;
        for (int i = 0; i < 4; ++i) {
            int nx = x + DX[i], ny = y + DY[i];
            if (!(0 <= nx && nx < n && 0 <= ny && ny < n)) continue;
            std::swap(conf[b], conf[nx * n + ny]);
            State ns = pack(conf, nn);
            int g = node->g + 1, f = g + P.h(conf);
            auto it = closed.find(ns);
            if (it == closed.end()) {
                Node *nd = new Node{ns, node, f, g};
                owned.push_back(nd);
                closed[ns] = nd;
                open.push(HeapItem{f, nd});
            } else if (it->second->f > f) {  /* CPU-solver.hpp:216 */
                *it->second = Node{ns, node, f, g};
                open.push(HeapItem{f, it->second});
            }
            std::swap(conf[b], conf[nx * n + ny]);
        }
    }
    *expanded = exp;
    if (best) {
        rc = best->f;
        moves.clear();
        uint8_t cc[25], pc[25];
        for (Node *c = best; c->prev; c = c->prev) {
            unpack(c->ps, cc, nn);
            unpack(c->prev->ps, pc, nn);
            int cb = blank_of(cc, nn), pb = blank_of(pc, nn);
            for (int i = 0; i < 4; ++i)
                if (pb / n + DX[i] == cb / n && pb % n + DY[i] == cb % n) moves.push_back(i);
        }
        std::reverse(moves.begin(), moves.end());
    }
    for (Node *p : owned) delete p;
    return rc;
}

/* IDA* (cross-check): heuristic = PDB (use_pdb) or Manhattan distance */
struct Ida {
    const PuzzleCtx *P;
    bool use_pdb;
    int n, nn;
    uint8_t conf[25];
    int64_t nodes = 0, limit = 0;
    std::vector<int> path;
    int hman() const {
        int s = 0;
        for (int i = 0; i < nn; ++i)
            if (conf[i]) {
                int t = conf[i] - 1;
                s += std::abs(i / n - t / n) + std::abs(i % n - t % n);
            }
        return s;
    }
    int h() const { return use_pdb ? P->h(conf) : hman(); }
    /* returns -1 found, else the min f above bound; -2 limit */
    int dfs(int b, int g, int bound, int last) {
        int f = g + h();
        if (f > bound) return f;
        bool goal = true;
        for (int i = 0; i < nn - 1 && goal; ++i) goal = conf[i] == i + 1;
        if (goal) return -1;
        if (limit > 0 && ++nodes > limit) return -2;
        int best = 1 << 30;
        int x = b / n, y = b % n;
        for (int i = 0; i < 4; ++i) {
            if (last >= 0 && (i ^ 1) == last) continue; /* do not undo the previous move */
            int nx = x + DX[i], ny = y + DY[i];
            if (!(0 <= nx && nx < n && 0 <= ny && ny < n)) continue;
            int nb = nx * n + ny;
            std::swap(conf[b], conf[nb]);
            path.push_back(i);
            int r = dfs(nb, g + 1, bound, i);
            if (r == -1) return -1;
            path.pop_back();
            std::swap(conf[b], conf[nb]);
            if (r == -2) return -2;
            best = std::min(best, r);
        }
        return best;
    }
};

}  // namespace

int orc_puzzle_h(int n, const uint8_t *tiles, const uint8_t *pdb) {
    if (n < 3 || n > 5 || !pdb) return -2;
    PuzzleCtx P(n, pdb);
    return P.h(tiles);
}

int orc_puzzle_astar(int n, const uint8_t *tiles, const uint8_t *pdb, int variant, int64_t node_limit,
                     int64_t *expanded, int32_t *moves, int64_t moves_cap, int64_t *moves_len) {
    if (n < 3 || n > 5 || !tiles) return -2;
    if (variant != 2 && !pdb) return -2;
    int nn = n * n;
    {
        std::vector<int> seen(nn, 0);
        for (int i = 0; i < nn; ++i) {
            if (tiles[i] >= nn || seen[tiles[i]]) return -2;
            seen[tiles[i]] = 1;
        }
    }
    int64_t exp = 0;
    std::vector<int> mv;
    int rc;
    if (variant == 0) {
        PuzzleCtx P(n, pdb);
        rc = puzzle_astar(P, tiles, node_limit, &exp, mv);
    } else {
        if (!orc_puzzle_solvable(n, tiles)) {
            rc = -1;
        } else {
            PuzzleCtx *P = variant == 1 ? new PuzzleCtx(n, pdb) : nullptr;
            Ida I;
            I.P = P;
            I.use_pdb = variant == 1;
            I.n = n;
            I.nn = nn;
            I.limit = node_limit;
            std::memcpy(I.conf, tiles, nn);
            int b = blank_of(I.conf, nn);
            int bound = I.h();
            rc = -3;
            while (true) {
                int r = I.dfs(b, 0, bound, -1);
                if (r == -1) {
                    rc = (int)I.path.size();
                    mv = I.path;
                    break;
                }
                if (r == -2 || r >= (1 << 30)) break;
                bound = r;
            }
            exp = I.nodes;
            delete P;
        }
    }
    if (expanded) *expanded = exp;
    if (rc >= 0) {
        if (moves_len) *moves_len = (int64_t)mv.size();
        if (moves && (int64_t)mv.size() <= moves_cap) std::copy(mv.begin(), mv.end(), moves);
    }
    return rc;
}
