"""The exchange ring's row padding (ring_row in csrc/line_engine.hpp),
compiled from the header for the host: in a column workgroup of LX lines,
the lanes that the LDS serves together -- 32 for 4-byte values, 16 for
8-byte values (two banks each) -- are LX lines x B / LX consecutive
segments at element offsets line * row * S + g * S, and must fall on
distinct banks for every odd position stride S a batch can have.  The
4-byte rule applied to 8-byte values (row * S = 32 / LX mod 32) fails the
8-byte check, which is what the check is for."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "..", "3dcompactsolver_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SRC = r"""
#include <cstdio>
#include "line_engine.hpp"
using namespace cs3d;

// the 4-byte rule whatever the element size (the planted mistake)
static int row_4byte_rule(int P, int KR, int LX, int S, bool col, int)
{
    return ring_row(P, KR, LX, S, col, 4);
}

// banks of a group of lanes: 4-byte values on 32 banks, 8-byte values on
// 16 bank pairs
template <typename RowFn>
static int conflicts(RowFn row_of, int esize)
{
    const int B = esize == 8 ? 16 : 32;
    int bad = 0;
    for (int LX = 2; LX <= 64; LX *= 2)
        for (int P = 4; P <= 128; P++)
            for (int KR = 1; KR <= 48; KR++)
                for (int S = 1; S <= 17; S += 2) {
                    const int row = row_of(P, KR, LX, S, true, esize);
                    if (row < P + KR || row - (P + KR) >= B) bad++;
                    const int nseg = LX >= B ? 1 : B / LX, nl = LX >= B ? B : LX;
                    if (nseg > P) continue;
                    unsigned seen = 0;
                    for (int line = 0; line < nl; line++)
                        for (int g = 0; g < nseg; g++) {
                            const long o = (long)line * row * S + (long)g * S;
                            const unsigned bit = 1u << (unsigned)(o % B);
                            if (seen & bit) {
                                bad++;
                                goto next;
                            }
                            seen |= bit;
                        }
                next:;
                }
    return bad;
}

int main()
{
    std::printf("%d %d %d\n", conflicts(ring_row, 4), conflicts(ring_row, 8),
                conflicts(row_4byte_rule, 8));
    return 0;
}
"""


@pytest.fixture(scope="module")
def ring_counts(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    d = tmp_path_factory.mktemp("ring")
    src, exe = d / "ring.cpp", d / "ring"
    src.write_text(SRC)
    r = subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                        "-I" + INC, str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return [int(v) for v in out.split()]


def test_ring_rows_spread_4byte_groups(ring_counts):
    assert ring_counts[0] == 0


def test_ring_rows_spread_8byte_groups(ring_counts):
    assert ring_counts[1] == 0


def test_4byte_rule_conflicts_on_8byte_values(ring_counts):
    assert ring_counts[2] > 0
