// cpu_solvers.cpp -- sequential CPU A* for the command line (grid and sliding-tile
// puzzle) and the puzzle's pattern databases with uAstar's cache files.
#include "cpu_solvers.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <deque>
#include <queue>
#include <unordered_map>

namespace cli {

namespace {

// Moves 0..7: four orthogonal, then four diagonal (uAstar's utils.hpp:120-121 order).
const int DX[8] = {1, -1, 0, 0, 1, 1, -1, -1};
const int DY[8] = {0, 0, 1, -1, 1, -1, 1, -1};

// Exact cost keys: a + b*sqrt(2) as a * 2^36 + b * floor(sqrt(2) * 2^36); the key order
// equals the order of the exact values for costs below 180000.
constexpr uint64_t KONE = 1ull << 36;
constexpr uint64_t KSQRT2 = 97184015999ull;

uint64_t octile(int x, int y, int ex, int ey) {
    int64_t dx = std::abs(x - ex), dy = std::abs(y - ey);
    int64_t d = std::min(dx, dy);
    return (uint64_t)(std::max(dx, dy) - d) * KONE + (uint64_t)d * KSQRT2;
}

}  // namespace

double GridCost::value() const { return (double)orth + (double)diag * std::sqrt(2.0); }

bool cpu_grid_astar(const std::vector<uint8_t> &graph, int H, int W, int sx, int sy, int ex, int ey,
                    GridCost *cost, std::vector<int> *path, int64_t *expanded) {
    const size_t N = (size_t)H * W;
    int64_t n_exp = 0;
    std::vector<uint64_t> g(N, ~0ull);
    std::vector<uint8_t> via(N, 0xFF);
    using Item = std::pair<uint64_t, uint32_t>;  // (f key, cell)
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> open;
    const uint32_t s = (uint32_t)((size_t)sx * W + sy), t = (uint32_t)((size_t)ex * W + ey);
    g[s] = 0;
    open.push({octile(sx, sy, ex, ey), s});
    bool found = false;
    while (!open.empty()) {
        auto [f, u] = open.top();
        open.pop();
        int x = (int)(u / W), y = (int)(u % W);
        if (f != g[u] + octile(x, y, ex, ey)) continue;  // a better entry was pushed
        if (u == t) {
            found = true;
            break;
        }
        ++n_exp;
        for (int d = 0; d < 8; ++d) {
            if (!(graph[u] & (1u << d))) continue;
            int nx = x + DX[d], ny = y + DY[d];
            if (nx < 0 || nx >= H || ny < 0 || ny >= W) continue;
            uint32_t v = (uint32_t)((size_t)nx * W + ny);
            uint64_t ng = g[u] + (d < 4 ? KONE : KSQRT2);
            if (ng >= g[v]) continue;
            g[v] = ng;
            via[v] = (uint8_t)d;
            open.push({ng + octile(nx, ny, ex, ey), v});
        }
    }
    if (expanded) *expanded = n_exp;
    if (!found) return false;
    std::vector<int> rev;
    int64_t orth = 0, diag = 0;
    for (uint32_t v = t;;) {
        rev.push_back((int)v);
        if (v == s) break;
        int d = via[v];
        if (d < 4) ++orth;
        else ++diag;
        v = (uint32_t)((size_t)((int)(v / W) - DX[d]) * W + ((int)(v % W) - DY[d]));
    }
    cost->orth = orth;
    cost->diag = diag;
    if (path) path->assign(rev.rbegin(), rev.rend());
    return true;
}

// ------------------------------------------------------------------ pattern databases

std::vector<std::vector<int>> pdb_partitions(int n) {  // puzzle/CPU-solver.hpp:44-104
    if (n == 3) return {{1, 2, 3, 4, 5, 6, 7, 8}};
    if (n == 5)
        return {{3, 4, 5, 10, 15, 20}, {2, 1, 6, 11, 16, 21}, {7, 8, 9, 12, 17, 22}, {13, 14, 18, 19, 23, 24}};
    return {{1, 2, 3, 4, 5, 6, 9, 13}, {7, 8, 10, 11, 12, 14, 15}};
}

size_t pdb_table_size(int n, size_t k) {
    size_t s = 1;
    for (size_t i = 0; i < k; ++i) s *= (size_t)(n * n) - i;
    return s;
}

std::string pdb_file_name(int n, const std::vector<int> &tiles) {
    std::string f = "database_" + std::to_string(n);
    for (int t : tiles) f += "_" + std::to_string(t);
    return f + ".bin";
}

namespace {

// Rank of the placement pos[0..k) (distinct cells of an n*n board) among the ordered
// k-tuples of cells, lexicographically: digit j is pos[j] minus the earlier cells below it.
uint32_t rank_of(const int *pos, int k, int cells) {
    uint32_t code = 0, used = 0;
    for (int j = 0; j < k; ++j) {
        uint32_t digit = (uint32_t)pos[j] - (uint32_t)__builtin_popcount(used & ((1u << pos[j]) - 1u));
        code = code * (uint32_t)(cells - j) + digit;
        used |= 1u << pos[j];
    }
    return code;
}

void unrank(uint32_t code, int k, int cells, int *pos) {
    int digit[16];
    for (int j = k - 1; j >= 0; --j) {
        uint32_t radix = (uint32_t)(cells - j);
        digit[j] = (int)(code % radix);
        code /= radix;
    }
    uint32_t used = 0;
    for (int j = 0; j < k; ++j) {
        int c = -1, skip = digit[j];
        while (true) {  // the digit-th free cell
            ++c;
            if (used & (1u << c)) continue;
            if (skip-- == 0) break;
        }
        pos[j] = c;
        used |= 1u << c;
    }
}

}  // namespace

void pdb_build(int n, const std::vector<int> &tiles, uint8_t *out) {
    const int k = (int)tiles.size(), cells = n * n;
    const size_t size = pdb_table_size(n, (size_t)k);
    std::fill(out, out + size, (uint8_t)0xFF);
    int pos[16];
    for (int j = 0; j < k; ++j) pos[j] = tiles[j] - 1;  // goal: tile v on cell v - 1
    std::deque<uint32_t> q;
    uint32_t c0 = rank_of(pos, k, cells);
    out[c0] = 0;
    q.push_back(c0);
    while (!q.empty()) {
        uint32_t c = q.front();
        q.pop_front();
        unrank(c, k, cells, pos);
        uint32_t occ = 0;
        for (int j = 0; j < k; ++j) occ |= 1u << pos[j];
        for (int j = 0; j < k; ++j) {
            int x = pos[j] / n, y = pos[j] % n;
            for (int d = 0; d < 4; ++d) {
                int nx = x + DX[d], ny = y + DY[d];
                if (nx < 0 || nx >= n || ny < 0 || ny >= n || (occ & (1u << (nx * n + ny)))) continue;
                int keep = pos[j];
                pos[j] = nx * n + ny;
                uint32_t nc = rank_of(pos, k, cells);
                pos[j] = keep;
                if (out[nc] != 0xFF) continue;
                out[nc] = (uint8_t)(out[c] + 1);
                q.push_back(nc);
            }
        }
    }
}

std::vector<uint8_t> pdb_load_or_build(int n, const std::string &dir, bool write_cache, bool *from_cache) {
    std::vector<uint8_t> pdb;
    bool all_cached = true;
    for (const auto &t : pdb_partitions(n)) {
        size_t size = pdb_table_size(n, t.size()), off = pdb.size();
        pdb.resize(off + size);
        std::string path = dir + "/" + pdb_file_name(n, t);
        FILE *f = std::fopen(path.c_str(), "rb");
        size_t got = 0;
        if (f) {
            got = std::fread(pdb.data() + off, 1, size, f);
            std::fclose(f);
        }
        if (got != size) {
            all_cached = false;
            std::printf("\tGenerating pattern database %s\n", pdb_file_name(n, t).c_str());
            pdb_build(n, t, pdb.data() + off);
        }
    }
    if (from_cache) *from_cache = all_cached;
    if (!all_cached && write_cache) pdb_write_cache(n, dir, pdb);
    return pdb;
}

bool pdb_write_cache(int n, const std::string &dir, const std::vector<uint8_t> &pdb) {
    size_t off = 0;
    for (const auto &t : pdb_partitions(n)) {
        size_t size = pdb_table_size(n, t.size());
        std::string path = dir + "/" + pdb_file_name(n, t);
        FILE *f = std::fopen(path.c_str(), "wb");
        if (!f) return false;
        bool ok = std::fwrite(pdb.data() + off, 1, size, f) == size;
        ok = std::fclose(f) == 0 && ok;
        if (!ok) {
            std::remove(path.c_str());
            return false;
        }
        off += size;
    }
    return true;
}

// ------------------------------------------------------------------ puzzle A*

namespace {

struct PuzzleCtx {
    int n;
    const std::vector<uint8_t> *pdb;
    std::vector<std::vector<int>> parts;
    std::vector<size_t> offs;
    int h(const uint8_t *cell_of_tile) const {  // cell_of_tile[t] = cell holding tile t
        int s = 0;
        int pos[16];
        for (size_t p = 0; p < parts.size(); ++p) {
            for (size_t j = 0; j < parts[p].size(); ++j) pos[j] = cell_of_tile[parts[p][j]];
            s += (*pdb)[offs[p] + rank_of(pos, (int)parts[p].size(), n * n)];
        }
        return s;
    }
};

// A board as one 128-bit word, cell i at bits 5i..5i+4 (boards of up to 25 cells).
using Board = unsigned __int128;

Board pack(const std::vector<uint8_t> &b) {
    Board s = 0;
    for (size_t i = 0; i < b.size(); ++i) s |= (Board)b[i] << (5 * i);
    return s;
}

inline unsigned tile_at(Board s, int cell) { return (unsigned)(s >> (5 * cell)) & 31u; }

struct BoardHash {
    size_t operator()(Board s) const {
        uint64_t lo = (uint64_t)s, hi = (uint64_t)(s >> 64);
        lo ^= hi * 0x9E3779B97F4A7C15ull;
        lo ^= lo >> 29;
        return (size_t)(lo * 0xBF58476D1CE4E5B9ull);
    }
};

}  // namespace

bool puzzle_solvable(int n, const std::vector<uint8_t> &board) {
    // parity of the permutation (blank as n*n) equals the parity of the blank's
    // Manhattan distance to the bottom-right corner
    const int cells = n * n;
    std::vector<int> p(cells);
    int blank = 0;
    for (int i = 0; i < cells; ++i) {
        p[i] = board[i] ? board[i] : cells;
        if (!board[i]) blank = i;
    }
    int inv = 0;
    for (int i = 0; i < cells; ++i)
        for (int j = i + 1; j < cells; ++j) inv += p[i] > p[j];
    int dist = (n - 1 - blank / n) + (n - 1 - blank % n);
    return (inv & 1) == (dist & 1);
}

bool cpu_puzzle_astar(int n, const std::vector<uint8_t> &board, const std::vector<uint8_t> &pdb, int *steps,
                      std::vector<int> *moves, int64_t *expanded, int64_t *dedup) {
    int64_t n_exp = 0, n_dup = 0;
    auto counts = [&] {
        if (expanded) *expanded = n_exp;
        if (dedup) *dedup = n_dup;
    };
    if (!puzzle_solvable(n, board)) {
        counts();
        return false;
    }
    PuzzleCtx ctx{n, &pdb, pdb_partitions(n), {}};
    size_t off = 0;
    for (const auto &t : ctx.parts) {
        ctx.offs.push_back(off);
        off += pdb_table_size(n, t.size());
    }
    const int cells = n * n;
    std::vector<uint8_t> goal(cells);
    for (int i = 0; i < cells - 1; ++i) goal[i] = (uint8_t)(i + 1);
    goal[cells - 1] = 0;
    const Board goal_s = pack(goal), start = pack(board);
    auto cells_of = [&](Board s, uint8_t *cell_of_tile) {
        for (int i = 0; i < cells; ++i) cell_of_tile[tile_at(s, i)] = (uint8_t)i;
    };
    struct Rec {
        uint8_t g, dir;  // dir: move that reached the state, 15 = start
    };
    std::unordered_map<Board, Rec, BoardHash> seen;
    using Item = std::tuple<int, int, Board>;  // (f, -g, state): deeper first on ties
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> open;
    uint8_t ct[25];
    cells_of(start, ct);
    seen[start] = {0, 15};
    open.push({ctx.h(ct), 0, start});
    while (!open.empty()) {
        auto [f, ng_neg, s] = open.top();
        open.pop();
        int g = -ng_neg;
        Rec r = seen[s];
        if (r.g != g) continue;
        if (s == goal_s) {
            std::vector<int> rev;
            Board cur = s;
            while (seen[cur].dir != 15) {  // undo the moves back to the start
                int d = seen[cur].dir;
                cells_of(cur, ct);
                int bl = ct[0], bx = bl / n - DX[d], by = bl % n - DY[d];
                int from = bx * n + by;  // where the blank was before move d
                Board t = tile_at(cur, from);
                cur = cur - (t << (5 * from)) + (t << (5 * bl));
                rev.push_back(d);
            }
            *steps = g;
            if (moves) moves->assign(rev.rbegin(), rev.rend());
            counts();
            return true;
        }
        ++n_exp;
        cells_of(s, ct);
        int bl = ct[0], bx = bl / n, by = bl % n;
        for (int d = 0; d < 4; ++d) {
            if (r.dir < 4 && d == (r.dir ^ 1)) continue;  // straight back
            int nx = bx + DX[d], ny = by + DY[d];
            if (nx < 0 || nx >= n || ny < 0 || ny >= n) continue;
            int nb = nx * n + ny;
            unsigned t = tile_at(s, nb);
            Board ns = s + ((Board)t << (5 * bl)) - ((Board)t << (5 * nb));
            auto it = seen.find(ns);
            if (it != seen.end() && it->second.g <= g + 1) {
                ++n_dup;
                continue;
            }
            seen[ns] = {(uint8_t)(g + 1), (uint8_t)d};
            uint8_t nct[25];
            std::copy(ct, ct + 25, nct);
            nct[0] = (uint8_t)nb;
            nct[t] = (uint8_t)bl;
            open.push({g + 1 + ctx.h(nct), -(g + 1), ns});
        }
    }
    counts();
    return false;
}

}  // namespace cli
