"""Host logic of the profile-LU ensemble groups (adhv8_kraken_amd.sw2.
profile_lu_bytes and lu_groups), CPU only.

profile_lu_bytes against a brute-force count over each mesh's node graph
(adjacency lists from every element's node pairs, the reference's serial
node partition node_order.c:64-104 written out block by block); lu_groups:
a partition of the indices, by increasing storage, each group within the
budget unless it holds a single instance."""
import numpy as np
import pytest

from adhv8_kraken_amd import mesh
from adhv8_kraken_amd.sw2 import lu_groups, profile_lu_bytes


def _brute_bytes(p, nblock):
    N = p.nnodes
    adj = [set() for _ in range(N)]
    e2 = np.asarray(p.e2d).reshape(-1, 3)
    e1 = np.asarray(p.e1d).reshape(-1, 2)
    for el in list(e2) + list(e1):
        for a in el:
            for b in el:
                if a != b:
                    adj[int(a)].add(int(b))
    nb = min(nblock, N)
    inc = max(1, N // nb)
    starts, last = [], 0
    for b in range(nb):
        starts.append(last)
        last = N if b == nb - 1 else last + inc
    ends = starts[1:] + [N]
    total = 0
    for s0, s1 in zip(starts, ends):
        d = 0
        for a in range(s0, s1):
            for c in adj[a]:
                if s0 <= c < s1:
                    d = max(d, abs(c - a))
        m, w = 3 * (s1 - s0), 3 * d + 2
        nt = -(-m // 64)
        total += 8 * (m * (2 * (w + 64) + 1) + 2 * 4096 * nt + 2 * m + 8256)
    return total


@pytest.mark.parametrize("name", ["channel", "floodplain"])
@pytest.mark.parametrize("nblock", [1, 2, 5])
def test_profile_lu_bytes_against_the_node_graph(name, nblock):
    p = mesh.channel_config1() if name == "channel" else mesh.floodplain(20, 10)
    assert profile_lu_bytes(p, nblock) == _brute_bytes(p, nblock)


def test_lu_groups_partition_and_budget():
    ps = [mesh.floodplain(n, n // 2, seed=k) for k, n in enumerate([20, 8, 14, 30, 10, 24])]
    sizes = [profile_lu_bytes(p) for p in ps]
    budget = 1.6 * max(sizes)
    groups = lu_groups(ps, 1, budget)
    flat = [i for g in groups for i in g]
    assert sorted(flat) == list(range(len(ps)))
    assert [sizes[i] for i in flat] == sorted(sizes)
    for g in groups:
        assert len(g) == 1 or sum(sizes[i] for i in g) <= budget
    # a budget below the largest instance leaves that instance alone
    small = lu_groups(ps, 1, 0.5 * max(sizes))
    assert [int(np.argmax(sizes))] in small
    assert len(lu_groups(ps, 1, 10 * sum(sizes))) == 1
