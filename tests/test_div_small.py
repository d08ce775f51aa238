"""The exact small-constant quotient used by the simplex projection's theta
search (potts_math.hpp, div_small): checked against IEEE division on every
mantissa of subnormal, normal and extreme exponents and on random patterns
(tests/cpp/div_small_check.c; `div_small_check all` covers every 32-bit
pattern, which passed when the quotient was introduced).  The device code is
checked on the GPU by sct_selftest_cr_math."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def test_div_small_equals_ieee_division(tmp_path):
    exe = tmp_path / "div_small_check"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-mfma", "-o", str(exe), str(HERE / "cpp" / "div_small_check.c"),
                    "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert out == ["3", "0", "5", "0", "7", "0"], out
