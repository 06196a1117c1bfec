// ifl_ring.hpp - ring-ordered cell lists of a tile and its halo, shared by the temporally blocked
// red-black sweeps (2D tiles and 3D z-march columns).
#pragma once

namespace ifl {

// Cells a K-sweep launch relaxes (those within 2K-1 of the tile), per colour, ordered by ring
// (ring 0 is the tile, ring r the cells r outside it), row-major within a ring; an entry is
// ly * 256 + lx in staged coordinates.  Every ring-r rectangle has even sides, so each colour holds half of it:
// half sweep p relaxes the first upTo(2K-1-p) cells of its colour's list.
template <int TX, int TY, int K>
struct RingOrder {
    static constexpr int H = 2 * K, R0 = 2 * K - 1;
    static constexpr int HX = TX + 2 * H, HY = TY + 2 * H;
    static_assert(HX <= 256 && HY <= 256, "cells are packed as ly * 256 + lx");
    static_assert(TX % 2 == 0 && TY % 2 == 0, "even tile sides");
    static constexpr int upTo(int r) { return (TX + 2 * r) * (TY + 2 * r) / 2; }
    static constexpr int NPC = upTo(R0);
    struct Table {
        unsigned short cell[2][NPC];
        bool ok;
    };
    static constexpr Table make() {
        Table t{};
        int n[2] = {0, 0};
        t.ok = true;
        for (int r = 0; r <= R0; ++r) {
            for (int ly = H - r; ly < H + TY + r; ++ly)
                for (int lx = H - r; lx < H + TX + r; ++lx) {
                    const int dx = lx < H ? H - lx : (lx >= H + TX ? lx - (H + TX - 1) : 0);
                    const int dy = ly < H ? H - ly : (ly >= H + TY ? ly - (H + TY - 1) : 0);
                    if ((dx > dy ? dx : dy) != r) continue;
                    const int c = (lx + ly) & 1;
                    t.cell[c][n[c]++] = (unsigned short)(ly * 256 + lx);
                }
            t.ok = t.ok && n[0] == upTo(r) && n[1] == upTo(r);
        }
        return t;
    }
    static constexpr Table table = make();
    static_assert(table.ok, "each ring holds half of its cells in each colour");
};

}  // namespace ifl
