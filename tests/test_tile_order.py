"""Host-side plans of the deposition's tile blocks (lcode3d_amd/csrc/tile_order.hpp):
the XCD-cluster block order, its spread over each XCD's CUs and the split tiles.

The header is plain C++; a small harness is compiled with g++ and its output
checked here, on the CPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

CSRC = os.path.join(os.path.dirname(__file__), os.pardir, "lcode3d_amd", "csrc")

HARNESS = r"""
#include <cstdio>
#include <cstdlib>
#include "tile_order.hpp"
int main(int argc, char** argv) {
    const int nt = std::atoi(argv[1]), per_xcd = std::atoi(argv[2]);
    const int cluster = std::atoi(argv[3]), ns = std::atoi(argv[4]);
    std::vector<int> order = lcode::xcd_cluster_order(nt, cluster);
    for (int v : order) std::printf("%d ", v);
    std::printf("\n");
    lcode::balance_xcd_slots(order, nt, per_xcd);
    for (int v : order) std::printf("%d ", v);
    std::printf("\n");
    for (int v : lcode::split_tile_table(nt, ns)) std::printf("%d ", v);
    std::printf("\n");
    return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("tile_order")
    src, exe = d / "h.cpp", d / "h"
    src.write_text(HARNESS)
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", CSRC, str(src), "-o", str(exe)], check=True)

    def run(nt, per_xcd, cluster=3, ns=0):
        out = subprocess.run([str(exe), str(nt), str(per_xcd), str(cluster), str(ns)],
                             check=True, capture_output=True, text=True).stdout.splitlines()
        return [np.array(line.split(), dtype=np.int64) for line in out]
    return run


def centre_d2(t, nt):
    mid = (nt - 1) / 2.
    return (t // nt - mid) ** 2 + (t % nt - mid) ** 2


@pytest.mark.parametrize("nt,per_xcd", [(27, 32), (7, 32), (54, 32), (27, 38), (5, 1)])
def test_orders_are_permutations_keeping_each_xcds_tiles(harness, nt, per_xcd):
    cl, bal, _ = harness(nt, per_xcd)
    n = nt * nt
    assert sorted(cl.tolist()) == list(range(n))
    assert sorted(bal.tolist()) == list(range(n))
    for x in range(8):
        assert sorted(cl[x::8].tolist()) == sorted(bal[x::8].tolist())


def test_clusters_share_an_xcd(harness):
    # 641^2 with 24-cell tiles: 27 x 27 tiles in 3 x 3 clusters (81 clusters over 8 XCDs)
    nt = 27
    cl, _, _ = harness(nt, 32)
    xcd = np.empty(nt * nt, dtype=np.int64)
    xcd[cl] = np.arange(nt * nt) % 8
    whole = 0
    for cy in range(nt // 3):
        for cx in range(nt // 3):
            tiles = [(cy * 3 + dy) * nt + cx * 3 + dx for dy in range(3) for dx in range(3)]
            whole += len(set(xcd[tiles].tolist())) == 1
    # an XCD's clusters that its slots cannot hold go to the other XCDs' free slots
    assert whole >= 0.9 * (nt // 3) ** 2


@pytest.mark.parametrize("nt,per_xcd", [(27, 32), (20, 32), (27, 38)])
def test_heaviest_tiles_of_an_xcd_get_distinct_cus(harness, nt, per_xcd):
    # slot k of an XCD (linear block id 8 k + x) runs on CU k mod per_xcd
    _, bal, _ = harness(nt, per_xcd)
    for x in range(8):
        tiles = bal[x::8]
        m = len(tiles)
        d2 = np.array([centre_d2(t, nt) for t in tiles])
        # the per_xcd tiles nearest the centre fill round 0, one per CU
        first = min(per_xcd, m)
        assert np.max(d2[:first]) <= np.min(d2[first:], initial=np.inf)
        cus = np.arange(m) % per_xcd
        assert len(set(cus[:first].tolist())) == first
        # and a CU's later rounds hold lighter tiles than its round-0 tile
        for c in range(min(per_xcd, m)):
            slots = np.arange(c, m, per_xcd)
            assert np.all(d2[slots[1:]] >= d2[slots[0]])


@pytest.mark.parametrize("nt,ns", [(27, 39), (7, 49), (27, 0), (6, 5)])
def test_split_table(harness, nt, ns):
    _, _, tab = harness(nt, 32, 3, ns)
    n = nt * nt
    assert len(tab) == ns + n
    tiles, slot = tab[:ns], tab[ns:]
    assert len(set(tiles.tolist())) == ns
    for s, t in enumerate(tiles):
        assert slot[t] == s
    others = [t for t in range(n) if t not in set(tiles.tolist())]
    assert np.all(slot[others] == -1)
    if 0 < ns < n:
        assert max(centre_d2(t, nt) for t in tiles) <= min(centre_d2(t, nt) for t in others)
