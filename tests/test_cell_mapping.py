"""CPU check of the pair pass's workgroup -> cell mapping (k_cell in
komd_amd/csrc/k_cell.hip, `cell_jobs` on the host and the mapping at the top of the
kernel): blocks go round-robin to the 8 XCDs; XCD x owns cells [x cx, (x+1) cx); with S
resident workgroups per XCD and S < n < 2S cells, the first 2S - n cells get one
workgroup and the rest two half-cell workgroups.  Every cell must be served exactly
once, whole or by both halves, for any grid and slot count."""
import pytest


def whole(n, S):
    return n if (n <= S or n >= 2 * S) else 2 * S - n


def jobs(C, S):
    cx = (C + 7) // 8
    jmax = 0
    for x in range(8):
        n = max(0, min(cx, C - x * cx))
        w = whole(n, S)
        jmax = max(jmax, w + 2 * (n - w))
    return cx, jmax


def served(C, S):
    cx, chunk = jobs(C, S)
    seen = {}
    for b in range(8 * chunk):
        x, j = b & 7, b >> 3
        n = max(0, min(cx, C - x * cx))
        w = whole(n, S)
        if j < w:
            cell, part, nparts = x * cx + j, 0, 1
        else:
            i = j - w
            cell = x * cx + w + (i >> 1) if w + (i >> 1) < n else C
            part, nparts = i & 1, 2
        if cell < C:
            seen.setdefault(cell, []).append((part, nparts))
    return seen, chunk


@pytest.mark.parametrize("C", [27, 64, 216, 1000, 1024, 1331, 2048, 2049, 2197, 3000, 4095, 4096, 4097,
                               8000, 17576, 19683])
@pytest.mark.parametrize("S", [1, 7, 224, 256, 300, 5000])
def test_every_cell_served_once(C, S):
    seen, _ = served(C, S)
    assert sorted(seen) == list(range(C))
    for v in seen.values():
        assert sorted(v) in ([(0, 1)], [(0, 2), (1, 2)])


def test_split_only_for_a_partial_second_round():
    # 32,000 atoms: 13^3 = 2,197 cells; XCDs 0-6 hold 275 (237 whole, 38 split against
    # 256 slots), XCD 7 holds 272 (240 whole, 32 split)
    seen, chunk = served(2197, 256)
    split = sum(1 for v in seen.values() if len(v) == 2)
    assert chunk == 313 and split == 7 * 38 + 32
    # 256,000 atoms: 19,683 cells, over two rounds per XCD: one workgroup each
    seen, chunk = served(19683, 256)
    assert all(len(v) == 1 for v in seen.values()) and chunk == 2461
