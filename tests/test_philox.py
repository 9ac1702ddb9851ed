"""Philox4x32-10: the draw stream shared by the oracle and the HIP kernels.

Pinned by the Random123 known-answer vectors (kat_vectors, philox4x32_10)
and, independently, by rocRAND's Philox engine (source header shipped in
/opt/rocm/include/rocrand), compiled here on the host.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import pyoracle as po

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_random123_known_answers(ctr, key, expect):
    assert po.philox(ctr, key).tolist() == expect


def test_draw_layout_is_counter_words():
    # draw d of (seed, tag, id, walker) = word d%4 of block (d/4, walker, id_lo, tag<<24|id_hi)
    seed, tag, eid, j = 0x1234567890ABCDEF, 2, (7 << 32) | 99, 311
    for d in range(12):
        blk = po.philox([d // 4, j, eid & 0xFFFFFFFF, (tag << 24) | (eid >> 32)],
                        [seed & 0xFFFFFFFF, seed >> 32])
        assert po.draw(seed, tag, eid, j, d) == int(blk[d % 4])


XCHECK = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
    unsigned long long seed = std::strtoull(argv[1], nullptr, 0);
    unsigned long long sub = std::strtoull(argv[2], nullptr, 0);
    rocrand_state_philox4x32_10 st;
    rocrand_init(seed, sub, 0, &st);
    for (int i = 0; i < 16; ++i) std::printf("%u\n", rocrand(&st));
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("hipcc") is None, reason="hipcc not available")
def test_matches_rocrand_engine():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "x.cpp")
        exe = os.path.join(d, "x")
        open(src, "w").write(XCHECK)
        subprocess.run(["hipcc", "--offload-arch=gfx950", "-O1", src, "-o", exe], check=True,
                       capture_output=True)
        for seed, sub in [(0, 0), (0x9E3779B97F4A7C15, 5), (42, (3 << 32) | 17)]:
            out = subprocess.run([exe, str(seed), str(sub)], check=True, capture_output=True,
                                 text=True).stdout.split()
            got = [int(x) for x in out]
            ours = []
            for k in range(4):
                ours += po.philox([k, 0, sub & 0xFFFFFFFF, sub >> 32],
                                  [seed & 0xFFFFFFFF, seed >> 32]).tolist()
            assert got == ours
