"""The march kernels mask lanes by giving a buffer access an out-of-range
lane offset (BUF_NONE = 2^31 against a 2^31 - 1 byte range), not by
branching. That is only sound if the lane offset reaches the hardware range
check unchanged: the check covers the VGPR offset plus the instruction's
immediate offset, not the scalar offset, so a compiler that moved part of a
lane offset into the scalar operand (or subtracted a constant from it) could
turn a masked lane into a real access 2 GiB past the base. This test
compiles the two kernels that use the masking to gfx950 assembly and checks
every buffer load and store: VGPR offset mode and no negative immediate; in
the two-step march a literal zero scalar offset. The 64-lane march passes
wave-uniform plane, component and psi-slot offsets in the scalar operand on
purpose (one buffer resource per array instead of one per plane); the host
keeps them below 2^31, and the lane offset, BUF_NONE for a masked lane, is
a select in the VGPR operand, so the masked lane still fails the check.
"""
import re
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "photonfdtd_amd" / "csrc"
SOURCES = ("yee3d_march64.hip", "yee3d_march2.hip")
BUF = re.compile(r"^\s*buffer_(load|store)\w*\s+(.*)$")


def _hipcc():
    sys.path.insert(0, str(ROOT))
    from photonfdtd_amd import build
    try:
        return build.hipcc(), build.COMMON
    except RuntimeError:
        return None, None


@pytest.fixture(scope="module")
def assembly(tmp_path_factory):
    cc, common = _hipcc()
    if cc is None or shutil.which(cc) is None and not Path(cc).exists():
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("isa")
    procs = {}
    for src in SOURCES:
        cmd = [cc, *common, "--cuda-device-only", "-S", "-x", "hip", str(CSRC / src),
               "-o", str(out / (src + ".s"))]
        procs[src] = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    text = {}
    for src, p in procs.items():
        _, err = p.communicate(timeout=600)
        assert p.returncode == 0, err.decode()[-2000:]
        text[src] = (out / (src + ".s")).read_text()
    return text


@pytest.mark.parametrize("src", SOURCES)
def test_buffer_accesses_keep_lane_offset_in_range_check(assembly, src):
    seen = 0
    for line in assembly[src].splitlines():
        m = BUF.match(line)
        if not m:
            continue
        seen += 1
        ops = [o.strip() for o in m.group(2).split(",")]
        # vdata, voffset, srsrc, soffset [modifiers...]
        assert len(ops) >= 4, line
        soff, mods = ops[3].split()[0], " ".join(ops[3].split()[1:] + ops[4:])
        if src == "yee3d_march64.hip":
            # a scalar register or a non-negative inline constant
            assert re.fullmatch(r"s\d+|\d+", soff), f"scalar offset operand: {line.strip()}"
        else:
            assert soff == "0", f"scalar offset in a masked access: {line.strip()}"
        assert "offen" in mods, f"no VGPR offset: {line.strip()}"
        imm = re.search(r"offset:(-?\d+)", mods)
        assert imm is None or int(imm.group(1)) >= 0, line.strip()
    assert seen > 100, f"{src}: only {seen} buffer accesses found"
