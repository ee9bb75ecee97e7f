"""The flux5 / flux3 face values divide by 60 and 12 through a product by the
rounded reciprocal plus one fma correction (div_const in csrc/common.h).  This
runs the exhaustive proof that the form equals the IEEE quotient on every float
it takes (tests/float_div_const.c); subnormal and infinite inputs and
subnormal or NaN results take the true division instead."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not found")
    exe = str(tmp_path_factory.mktemp("divc") / "float_div_const")
    subprocess.run([cc, "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "float_div_const.c"),
                    "-lpthread", "-lm"], check=True)
    return exe


@pytest.mark.parametrize("d", [60, 12])
def test_product_form_equals_ieee_division_on_every_float(checker, d):
    r = subprocess.run([checker, str(d)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    fast = int(r.stdout.split("fast-path inputs")[1].split()[0])
    assert fast > 2_000_000_000         # all but the subnormal / infinite corner

