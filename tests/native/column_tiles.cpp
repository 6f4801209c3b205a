// Host check of the sort-column numbering (deorbit_amd/csrc/common.h): for grids
// of many shapes and every tile width the grid builder may pick, col_index maps
// the columns one to one into [0, padded column count), tiles hold whole
// square blocks of columns, and cell_key / block_key stay inside the segment's
// cell and block ranges.  Prints a summary line; exits 1 on the first failure.
#include <cstdio>
#include <vector>

#include "common.h"

using namespace dsph;

int main() {
    int grids = 0;
    for (int dx = 1; dx <= 70; dx += 3)
        for (int dy = 1; dy <= 70; dy += 4)
            for (int sh = 0; (1 << sh) <= kColTileMax; sh++) {
                SegGrid g{};
                const int T = 1 << sh;
                g.dim[0] = dx; g.dim[1] = dy; g.dim[2] = 97;
                g.tile_sh = sh;
                g.tiles_y = (dy + T - 1) >> sh;
                const uint32_t tx = (uint32_t)((dx + T - 1) >> sh);
                const uint32_t ncol = tx * (uint32_t)g.tiles_y * (uint32_t)(T * T);
                g.cell_base = 1000;
                g.ncells = ncol * (uint32_t)g.dim[2] + 1;
                g.nblk = (g.dim[2] + kSliceBlock - 1) / kSliceBlock;
                g.blk_base = 77;
                std::vector<char> seen(ncol, 0);
                for (int cx = 0; cx < dx; cx++)
                    for (int cy = 0; cy < dy; cy++) {
                        const uint32_t c = col_index(g, cx, cy);
                        if (c >= ncol || seen[c]) {
                            std::printf("col_index not one to one: dims %d x %d tile %d (%d, %d) -> %u\n", dx, dy, T,
                                        cx, cy, c);
                            return 1;
                        }
                        seen[c] = 1;
                        // a tile is a contiguous run of T*T column numbers
                        const uint32_t tile = ((uint32_t)cx >> sh) * (uint32_t)g.tiles_y + ((uint32_t)cy >> sh);
                        if (c / (uint32_t)(T * T) != tile) {
                            std::printf("column (%d, %d) outside its tile\n", cx, cy);
                            return 1;
                        }
                        const uint32_t k0 = cell_key(g, cx, cy, 0), k1 = cell_key(g, cx, cy, g.dim[2] - 1);
                        if (k0 < g.cell_base || k1 >= g.cell_base + g.ncells - 1 || k1 - k0 != (uint32_t)g.dim[2] - 1) {
                            std::printf("cell keys of (%d, %d) out of range\n", cx, cy);
                            return 1;
                        }
                        const uint32_t b1 = block_key(g, cx, cy, g.dim[2] - 1);
                        if (block_key(g, cx, cy, 0) < g.blk_base || b1 >= g.blk_base + ncol * (uint32_t)g.nblk) {
                            std::printf("block keys of (%d, %d) out of range\n", cx, cy);
                            return 1;
                        }
                    }
                grids++;
            }
    std::printf("column tiles: %d grids one to one\n", grids);
    return 0;
}
