// cpu_solvers.hpp -- the command line's sequential CPU A* (the --cpu side of the
// uAstar CLI, pathway/CPU-solver.cpp and puzzle/CPU-solver.hpp) and the pattern
// databases of the sliding-tile puzzle, with uAstar's cache files.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cli {

// Exact grid cost orth + diag * sqrt(2).
struct GridCost {
    int64_t orth = 0, diag = 0;
    double value() const;
};

// 8-direction grid, one byte per cell (nonzero = open; bit d = move d allowed out of
// the cell). A move u -> v needs bit d of u and v in range; returns false without a path.
// *expanded (if given) receives the number of nodes expanded.
bool cpu_grid_astar(const std::vector<uint8_t> &graph, int H, int W, int sx, int sy, int ex, int ey,
                    GridCost *cost, std::vector<int> *path, int64_t *expanded = nullptr);

// The additive disjoint partitions for an n x n board (n = 3, 4 or 5).
std::vector<std::vector<int>> pdb_partitions(int n);
size_t pdb_table_size(int n, size_t k);
// uAstar's cache file name for one partition: database_<n>_<t1>_<t2>...bin
std::string pdb_file_name(int n, const std::vector<int> &tiles);
// Breadth-first search of one partition's table (uint8 distances, 0xFF = unreached).
void pdb_build(int n, const std::vector<int> &tiles, uint8_t *out);
// All partitions concatenated, from the cache files in `dir` where present and intact,
// else built here (and then written there when write_cache).
std::vector<uint8_t> pdb_load_or_build(int n, const std::string &dir, bool write_cache, bool *from_cache);
bool pdb_write_cache(int n, const std::string &dir, const std::vector<uint8_t> &pdb);

// Optimal blank moves (0..3 = (+1,0), (-1,0), (0,+1), (0,-1)) of a board (row-major,
// 0 = blank); false when the board is not solvable. *expanded and *dedup (if given)
// receive the nodes expanded and the successors dropped as already seen at no larger g.
bool cpu_puzzle_astar(int n, const std::vector<uint8_t> &board, const std::vector<uint8_t> &pdb, int *steps,
                      std::vector<int> *moves, int64_t *expanded = nullptr, int64_t *dedup = nullptr);
bool puzzle_solvable(int n, const std::vector<uint8_t> &board);

}  // namespace cli
