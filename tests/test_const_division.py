"""Exhaustive check of the kernels' division shortcut (div_mk in
fenix_amd/csrc/fenix_kernels.hip): q0 = n * RN(1/d) followed by one Markstein
correction, fma(fma(-q0, d, n), RN(1/d), q0), must equal the IEEE quotient n / d
that the oracle computes, for every float input that can reach it:

- smoothk(e0, e1, x): for each literal edge pair that the kernel source uses (read
  from the source, so a new call site is checked as soon as it is written), every
  numerator n = x - e0 that can reach the division: n > d clamps to d, n < 0 and NaN
  clamp to 0, and a positive float difference x - e0 is a multiple of ulp(e0), so
  with e0 >= 2^-7 it is 0 or at least 2^-30. The check covers n = 0 and every float
  from 2^-31 to d;
- height01(y, H) = (y + 0.5) / H: every y < H for every H <= 32768 (larger heights
  take the IEEE division).

The check program is C compiled with -ffp-contract=off, so products and sums round
exactly as the GPU code (built with the same flag) rounds them, and fmaf is a
fused multiply-add on both.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNELS = os.path.join(ROOT, "fenix_amd", "csrc", "fenix_kernels.hip")

CHECKER = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float f_of(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static uint32_t b_of(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }
static float div_mk(float n, float d, float r) { float q0 = n * r; return fmaf(fmaf(-q0, d, n), r, q0); }
static float poly(float t) { return t * t * (3.f - 2.f * t); }

int main(int argc, char** argv) {
  if (argc > 1 && strcmp(argv[1], "height") == 0) {
    long bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
    for (uint32_t H = 1; H <= 32768u; ++H) {
      const float d = (float)H, r = 1.f / d;
      for (uint32_t y = 0; y < H; ++y) {
        const float n = (float)y + 0.5f;
        if (b_of(div_mk(n, d, r)) != b_of(n / d)) ++bad;
      }
    }
    printf("height %ld\n", bad);
    return 0;
  }
  /* argv: pairs of edges e0 e1 as float literals (without the f suffix) */
  for (int a = 1; a + 1 < argc; a += 2) {
    const float e0 = strtof(argv[a], NULL), e1 = strtof(argv[a + 1], NULL);
    const float d = e1 - e0, r = 1.f / d;
    const uint32_t top = b_of(d);
    long bad_t = 0, bad_poly = 0;
    const uint32_t lo = 0x30000000u; /* 2^-31 */
#pragma omp parallel for schedule(static) reduction(+ : bad_t, bad_poly)
    for (uint32_t b = lo - 1u; b <= top; ++b) {
      const float n = b == lo - 1u ? 0.f : f_of(b);
      const float q = n / d, t = div_mk(n, d, r);
      bad_t += b_of(q) != b_of(t);
      bad_poly += b_of(poly(q)) != b_of(poly(t));
    }
    printf("%s %s %ld %ld\n", argv[a], argv[a + 1], bad_t, bad_poly);
  }
  return 0;
}
"""

LIT = r"(-?[0-9]*\.?[0-9]+(?:e-?[0-9]+)?)f"


def _pairs():
    src = open(KERNELS).read()
    return sorted(set(re.findall(r"smoothk\(\s*" + LIT + r"\s*,\s*" + LIT + r"\s*,", src)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    d = tmp_path_factory.mktemp("divchk")
    c = d / "divchk.c"
    c.write_text(CHECKER)
    exe = d / "divchk"
    subprocess.run(["gcc", "-O3", "-march=native", "-ffp-contract=off", "-mfma", "-fopenmp", str(c), "-o", str(exe), "-lm"], check=True)
    return str(exe)


def test_smoothk_edges_found():
    pairs = _pairs()
    assert len(pairs) >= 20, pairs
    # a literal-edge smoothf left in the source would still divide; that is allowed,
    # but each smoothk pair must be a valid (increasing) edge pair
    for e0, e1 in pairs:
        # e0 >= 2^-7 bounds the smallest positive numerator (see the module docstring)
        assert float(e1) > float(e0) >= 2.0 ** -7


def test_smoothk_division_exact(checker):
    pairs = _pairs()
    args = [x for p in pairs for x in p]
    out = subprocess.run([checker] + args, check=True, capture_output=True, text=True, timeout=600).stdout.split("\n")
    rows = [line.split() for line in out if line.strip()]
    assert len(rows) == len(pairs)
    for e0, e1, bad_t, bad_poly in rows:
        # correctly rounded quotients, hence identical smoothstep values
        assert int(bad_t) == 0, (e0, e1, bad_t)
        assert int(bad_poly) == 0, (e0, e1, bad_poly)


def test_height_division_exact(checker):
    out = subprocess.run([checker, "height"], check=True, capture_output=True, text=True, timeout=600).stdout.split()
    assert out == ["height", "0"], out


def test_checker_catches_a_wrong_reciprocal(tmp_path):
    """The check program must notice a planted error: q0 moved up by one ulp in place of
    the correction step differs from the IEEE quotient, and from its smoothstep
    value, for many inputs."""
    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    src = CHECKER.replace("return fmaf(fmaf(-q0, d, n), r, q0);", "return nextafterf(q0, 2.f);")
    c = tmp_path / "bad.c"
    c.write_text(src)
    exe = tmp_path / "bad"
    subprocess.run(["gcc", "-O3", "-ffp-contract=off", "-mfma", "-fopenmp", str(c), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe), "0.2", "0.86"], check=True, capture_output=True, text=True, timeout=600).stdout.split()
    assert int(out[2]) > 0 and int(out[3]) > 0
