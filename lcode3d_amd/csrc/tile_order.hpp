// Host-side plans of the deposition's tile blocks: the XCD-cluster block
// order, its spread over each XCD's CUs, and the split tiles.  Plain C++
// (no HIP), so that the CPU tests compile it on its own.
#pragma once

#include <algorithm>
#include <vector>

namespace lcode {

// Squared distance of tile t (row-major in an nt x nt grid of tiles) from
// the grid centre, in tiles.
inline double tile_centre_d2(int t, int nt) {
    const double mid = (nt - 1) / 2.;
    const double y = t / nt - mid, x = t % nt - mid;
    return x * x + y * y;
}

// The XCD-cluster tile order: order[lin] = tile for linear block id lin
// (workgroups go to the 8 XCDs round-robin, XCD lin % 8), a permutation of
// the nt * nt tiles.  The tiles come in cluster x cluster squares, each
// square on one XCD (its tiles share that XCD's L2), the squares dealt to
// the XCDs round-robin; what an XCD's slots cannot take fills the others.
inline std::vector<int> xcd_cluster_order(int nt, int cluster) {
    constexpr int kXcd = 8;
    const int kC = cluster;
    const int nc = (nt + kC - 1) / kC;
    std::vector<std::vector<int>> per(kXcd);
    for (int cy = 0, j = 0; cy < nc; ++cy)
        for (int cx = 0; cx < nc; ++cx, ++j)
            for (int dy = 0; dy < kC; ++dy)
                for (int dx = 0; dx < kC; ++dx) {
                    const int ty = cy * kC + dy, tx = cx * kC + dx;
                    if (ty < nt && tx < nt) per[j % kXcd].push_back(ty * nt + tx);
                }
    const int n = nt * nt;
    std::vector<int> order(n, -1), spare;
    std::vector<size_t> used(kXcd, 0);
    for (int lin = 0; lin < n; ++lin) {   // each XCD's own clusters first
        auto& v = per[lin % kXcd];
        size_t& u = used[lin % kXcd];
        if (u < v.size()) order[lin] = v[u++];
    }
    for (int x = 0; x < kXcd; ++x)        // what did not fit, to the free slots
        for (size_t u = used[x]; u < per[x].size(); ++u) spare.push_back(per[x][u]);
    size_t k = 0;
    for (int lin = 0; lin < n; ++lin)
        if (order[lin] < 0) order[lin] = spare[k++];
    return order;
}

// Within an XCD the k-th block of a one-round launch goes to CU
// k mod per_xcd (read from the hardware ids of such a launch), so CU c runs
// slots c, c + per_xcd, c + 2 per_xcd, ...  Reorders each XCD's tiles over
// its slots, every tile staying on its XCD: nearest the grid centre first
// (the wake's tiles, the heaviest), in snake order over the rounds, so
// that a CU's blocks mix heavy and light tiles.
inline void balance_xcd_slots(std::vector<int>& order, int nt, int per_xcd) {
    constexpr int kXcd = 8;
    const int n = (int)order.size();
    for (int x = 0; x < kXcd; ++x) {
        std::vector<int> tiles;
        for (int lin = x; lin < n; lin += kXcd) tiles.push_back(order[lin]);
        const int m = (int)tiles.size();
        std::stable_sort(tiles.begin(), tiles.end(), [&](int u, int v) {
            return tile_centre_d2(u, nt) < tile_centre_d2(v, nt);
        });
        for (int i = 0; i < m; ++i) {
            const int r = i / per_xcd, c = i % per_xcd;
            const int width = std::min(per_xcd, m - r * per_xcd);   // slots of round r
            const int k = r * per_xcd + (r % 2 == 0 ? c : width - 1 - c);
            order[x + kXcd * k] = tiles[i];
        }
    }
}

// The split tiles: the ns tiles nearest the grid centre (an axisymmetric
// beam's wake compresses and displaces the plasma there, so those tiles
// have the largest halos and the most particles).  Returns ns + nt * nt
// ints: [0, ns) the tile of each slot, then for every tile its slot or -1.
inline std::vector<int> split_tile_table(int nt, int ns) {
    std::vector<int> byd((size_t)nt * nt);
    for (int k = 0; k < nt * nt; ++k) byd[k] = k;
    std::stable_sort(byd.begin(), byd.end(), [&](int u, int v) {
        return tile_centre_d2(u, nt) < tile_centre_d2(v, nt);
    });
    std::vector<int> tab((size_t)ns + (size_t)nt * nt, -1);
    for (int s = 0; s < ns; ++s) {
        tab[s] = byd[s];
        tab[ns + byd[s]] = s;
    }
    return tab;
}

}  // namespace lcode
