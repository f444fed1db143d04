"""The literal-edge smoothstep (sstep_c: reciprocal division, fs_divc.h) equals the plain one bit for bit.

sstep_c clamps the numerator to [-d, 2d] and divides through z = RN(1/d) with one Markstein correction.
Every edge pair the kernels use (collected from the sources, so a new call site is checked without being
listed here) is checked on every f32 numerator of its interval: the smoothstep t t (3 - 2 t) of the
clamped quotient must equal the one from IEEE division.  The quotient itself is correctly rounded except
for numerators below ~1e-30 in magnitude (the correction's remainder underflows there), where t t
underflows to 0 both ways; the test prints how many and how small.  The checker compiles the very header
the kernels include.  A planted mistake (the uncorrected product RN(x z)) must fail the same check.  The clamp itself leaves
the smoothstep unchanged for every numerator outside the interval: -d / d = -1 and 2d / d = 2 clamp to
the same t as any value beyond them, and NaN becomes -d, whose t = 0 is what the WGSL clamp makes of NaN.
"""
import os
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "procedural-fire_amd" / "csrc"
LIT = r"-?\d+\.\d*f"


def edge_pairs():
    pairs = set()
    for f in sorted(list(CSRC.glob("*.h")) + list(CSRC.glob("*.hip"))):
        for m in re.finditer(rf"\bg?sstep_c\(\s*([^,()]+?)\s*,\s*([^,()]+?)\s*,", f.read_text()):
            e0, e1 = m.group(1), m.group(2)
            if e0 in ("float e0", "e0"):
                continue                     # the definitions and the dual-path forwarder
            assert re.fullmatch(LIT, e0) and re.fullmatch(LIT, e1), f"{f.name}: sstep_c needs literal edges: {m.group(0)}"
            pairs.add((e0, e1))
    return sorted(pairs)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("divc") / "divc_check"
    flags = ["-O3", "-std=c++17", "-fopenmp", "-ffp-contract=off", f"-I{CSRC}"]
    cpu = Path("/proc/cpuinfo").read_text()
    if " fma" in cpu and " avx2" in cpu:
        flags += ["-mavx2", "-mfma"]         # vector FMA (glibc's fmaf is exact too, only slower)
    subprocess.run([gxx, *flags, str(ROOT / "tests" / "divc_check.cpp"), "-o", str(exe)], check=True)
    return exe


def run(exe, pairs, plain=False):
    args = [str(exe)] + (["--plain"] if plain else []) + [x for p in pairs for x in p]
    env = dict(os.environ, OMP_NUM_THREADS=str(min(8, os.cpu_count() or 1)))
    r = subprocess.run(args, capture_output=True, text=True, env=env, timeout=600)
    rows = [line.split() for line in r.stdout.strip().splitlines()]
    return r.returncode, {(w[0], w[1]): (float(w[2]), int(w[3]), int(w[4]), int(w[5]), float(w[6])) for w in rows}


def test_call_sites_found():
    pairs = edge_pairs()
    assert ("1000.f", "1400.f") in pairs and ("0.03f", "0.4f") in pairs
    assert len(pairs) >= 12


def test_smoothstep_equals_the_plain_division_on_every_numerator(checker):
    pairs = edge_pairs()
    rc, res = run(checker, pairs)
    for p in pairs:
        d, checked, bad_s, bad_q, max_q = res[p]
        print(f"sstep_c{p}: d = {d:.9g}, {checked} numerators, {bad_s} smoothstep mismatches; "
              f"quotient off in {bad_q}, all with |n| <= {max_q:.3g}")
        assert checked > 2 ** 30 and bad_s == 0, p
        assert max_q < 1e-29, p              # only where t t underflows
    assert rc == 0


def test_planted_mistake_is_caught(checker):
    rc, res = run(checker, [("0.f", "3.f"), ("1000.f", "1400.f")], plain=True)
    assert rc == 1 and all(v[2] > 0 for v in res.values())
