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
