"""The bit-sliced classification of the msc kernels (pyising_amd/csrc/msc_core.hpp, msc_bits),
evaluated on the host from the v_bitop3 truth tables the source compiles in: for every own spin
and four neighbours, und / a1 / pw must follow from the number of anti-aligned neighbours as
src/pyising/backends/numpy_backend.py:54-83 defines them (a0 | a1 undecided, a1 exactly one,
a2 = two or more flips). A wrong truth table or operand order fails here before any GPU run."""

import itertools
import re
from pathlib import Path

CORE = Path(__file__).resolve().parents[1] / "pyising_amd" / "csrc" / "msc_core.hpp"


def immediates() -> dict[str, int]:
    text = CORE.read_text()
    out = {}
    for name, val in re.findall(r"\b(B3_\w+) == (0x[0-9A-Fa-f]+)u", text):
        out[name] = int(val, 16)
    return out


def bitop3(a: int, b: int, c: int, imm: int) -> int:
    """v_bitop3_b32 on single bits: the result is bit (a << 2 | b << 1 | c) of the immediate."""
    return (imm >> ((a << 2) | (b << 1) | c)) & 1


def msc_bits(o, c, hz, up, dn, imm):
    """msc_bits of msc_core.hpp, operand for operand."""
    cy = bitop3(c, hz, up, imm["B3_MAJ"])
    sm = bitop3(c, hz, up, imm["B3_XOR3"])
    g = bitop3(sm, dn, o, imm["B3_BOTH_DIFFER"])
    und = bitop3(cy, o, g, imm["B3_EQ_AND_NOT"])
    a1 = bitop3(und, sm, dn, imm["B3_AND_XOR"])
    pw = 1 - (o ^ und)
    return und, a1, pw


def test_classification_truth_tables():
    imm = immediates()
    assert {"B3_MAJ", "B3_XOR3", "B3_BOTH_DIFFER", "B3_EQ_AND_NOT", "B3_AND_XOR"} <= set(imm)
    for o, n1, n2, n3, n4 in itertools.product((0, 1), repeat=5):
        k = sum(o != n for n in (n1, n2, n3, n4))
        assert msc_bits(o, n1, n2, n3, n4, imm) == (int(k <= 1), int(k == 1), o ^ int(k >= 2))


def test_a_swapped_truth_table_is_caught():
    imm = dict(immediates())
    imm["B3_EQ_AND_NOT"] ^= 0x01  # one truth-table bit wrong
    bad = 0
    for o, n1, n2, n3, n4 in itertools.product((0, 1), repeat=5):
        k = sum(o != n for n in (n1, n2, n3, n4))
        bad += msc_bits(o, n1, n2, n3, n4, imm) != (int(k <= 1), int(k == 1), o ^ int(k >= 2))
    assert bad > 0


def test_level_pair_forms_are_proved_at_compile_time(tmp_path):
    """The lazy comparison's level-pair forms (MscPairForm) are checked by static_assert in
    msc_core.hpp against the level-by-level definition on every input combination; this
    compiles the header for gfx950 (where hipcc is installed) so that a wrong form fails here,
    and checks that the assertion still covers all 16 case codes."""
    import shutil
    import subprocess

    text = CORE.read_text()
    assert all(f"msc_pair_proved<{c}>()" in text for c in range(16))
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        import pytest
        pytest.skip("hipcc not installed")
    src = tmp_path / "pair_proof.hip"
    src.write_text('#include "msc_core.hpp"\n')
    inc = CORE.parent
    r = subprocess.run([hipcc, "--offload-arch=gfx950", "-std=c++17", "--cuda-device-only",
                        "-fsyntax-only", f"-I{inc}", f"-I{inc.parents[1] / 'include'}", str(src)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]


# ------------------------------------------------ eight lanes per word (msc_resident.hip) ----
# The resident msc sweeps with small bands give a word eight lanes, lane j evaluating level pair
# j; each lane forms the pair's masks G (undecided through both levels) and T (accepted) for a
# word undecided everywhere, and the segment combines them with DPP moves inside the 8-lane
# segment (rs_lanes8_combine): quad_perm scans, quad totals, row_shr:4 on banks 1 and 3,
# row_shr:1, row_half_mirror. The lane moves are emulated here as the gfx9 DPP controls define
# them, and the combined update must equal the level pairs applied one after the other (the
# definition: the first level whose bit differs from the threshold bit decides, accepting when
# the threshold bit is 1).
_M32 = 0xFFFFFFFF


def _quad(v, sel):
    return [v[(i & ~3) + sel[i & 3]] for i in range(64)]


def _row_shr(v, d, banks, old):
    out = list(old)
    for i in range(64):
        li = i % 16
        if (banks >> (li // 4)) & 1 and li >= d:
            out[i] = v[i - d]
    return out


def _half_mirror(v):
    return [v[(i & ~7) + 7 - (i & 7)] for i in range(64)]


def _combine(G, T):
    q = [a & b for a, b in zip(G, _quad(G, [0, 0, 1, 2]))]
    q = [a & b for a, b in zip(q, _quad(q, [0, 1, 0, 1]))]
    qt = [a & b for a, b in zip(G, _quad(G, [1, 0, 3, 2]))]
    qt = [a & b for a, b in zip(qt, _quad(qt, [2, 3, 0, 1]))]
    Q = [a & b for a, b in zip(q, _row_shr(qt, 4, 0b1010, [_M32] * 64))]
    E = _row_shr(Q, 1, 0b1111, [_M32] * 64)
    E = [_M32 if i % 8 == 0 else e for i, e in enumerate(E)]
    acc = [a & b for a, b in zip(E, T)]
    a = [x | y for x, y in zip(acc, _quad(acc, [1, 0, 3, 2]))]
    a = [x | y for x, y in zip(a, _quad(a, [2, 3, 0, 1]))]
    A = [x | y for x, y in zip(a, _half_mirror(a))]
    Qall = [x & y for x, y in zip(qt, _half_mirror(qt))]
    return A, Qall


def _threshold(bit4, bit8, a1):
    """The threshold bit of every lane at one level: thr4's bit where a1, thr8's elsewhere."""
    return ((a1 if bit4 else 0) | ((~a1 & _M32) if bit8 else 0)) & _M32


def _pairs_in_order(und, pw, a1, xs, ys, codes):
    for j in range(8):
        c = codes[j]
        for x, tb in ((xs[j], _threshold(c & 1, c >> 1 & 1, a1)),
                      (ys[j], _threshold(c >> 2 & 1, c >> 3 & 1, a1))):
            decided = und & (x ^ tb)
            pw ^= decided & tb  # x = 0 < threshold bit 1: accepted
            und &= ~decided & _M32
    return und, pw


def _lane_masks(code, x, y, a1):
    """rs_pair_masks of msc_resident.hip."""
    ta = _threshold(code & 1, code >> 1 & 1, a1)
    tb = _threshold(code >> 2 & 1, code >> 3 & 1, a1)
    e = x ^ ta
    G = ~e & ~(y ^ tb) & _M32
    T = (ta & ~x | (~e & tb & ~y)) & _M32
    return G, T


def _check_segments(seed, plant=None):
    import numpy as np

    rng = np.random.default_rng(seed)
    bad = 0
    for trial in range(60):
        und = [int(v) for v in rng.integers(0, 2**32, 8)]
        if trial % 3 == 0:
            und = [_M32] * 8
        pw = [int(v) for v in rng.integers(0, 2**32, 8)]
        a1 = [int(v) for v in rng.integers(0, 2**32, 8)]
        xs = rng.integers(0, 2**32, (8, 8)).tolist()
        ys = rng.integers(0, 2**32, (8, 8)).tolist()
        if trial % 2:  # few bits differ from the thresholds: words decided late
            xs = [[x & int(m) for x, m in zip(r, rng.integers(0, 2**32, 8))] for r in xs]
        codes = rng.integers(0, 16, (8, 8)).tolist()
        G, T = [0] * 64, [0] * 64
        for lane in range(64):
            w, j = divmod(lane, 8)
            G[lane], T[lane] = _lane_masks(codes[w][j], xs[w][j], ys[w][j], a1[w])
        if plant == "shift":  # exclusive scan without the segment-start reset
            A, Qall = _combine_planted(G, T)
        else:
            A, Qall = _combine(G, T)
        for w in range(8):
            want = _pairs_in_order(und[w], pw[w], a1[w], xs[w], ys[w], codes[w])
            for lane in range(8 * w, 8 * w + 8):
                got = (und[w] & Qall[lane], pw[w] ^ (und[w] & A[lane]))
                bad += got != want
    return bad


def _combine_planted(G, T):
    q = [a & b for a, b in zip(G, _quad(G, [0, 0, 1, 2]))]
    q = [a & b for a, b in zip(q, _quad(q, [0, 1, 0, 1]))]
    qt = [a & b for a, b in zip(G, _quad(G, [1, 0, 3, 2]))]
    qt = [a & b for a, b in zip(qt, _quad(qt, [2, 3, 0, 1]))]
    Q = [a & b for a, b in zip(q, _row_shr(qt, 4, 0b1010, [_M32] * 64))]
    E = _row_shr(Q, 1, 0b1111, [_M32] * 64)  # lane 8 of a row keeps the other segment's value
    acc = [a & b for a, b in zip(E, T)]
    a = [x | y for x, y in zip(acc, _quad(acc, [1, 0, 3, 2]))]
    a = [x | y for x, y in zip(a, _quad(a, [2, 3, 0, 1]))]
    return [x | y for x, y in zip(a, _half_mirror(a))], [x & y for x, y in zip(qt, _half_mirror(qt))]


def test_eight_lane_combine_equals_pairs_in_order():
    assert _check_segments(5) == 0


def test_eight_lane_combine_catches_a_segment_leak():
    assert _check_segments(6, plant="shift") > 0
