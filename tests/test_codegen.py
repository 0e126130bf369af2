"""The committed generated files are what the generator produces from the YAML."""
import filecmp
import os
import tempfile

import pytest

from pelelm_amd.mechanisms import codegen, hipgen
from pelelm_amd.mechanisms import mechanism as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", ["LiDryer", "drm19"])
def test_generated_files_are_current(name):
    mech = M.load(name)
    with tempfile.TemporaryDirectory() as tmp:
        tbl = os.path.join(tmp, f"{name}.tbl")
        codegen.write_table(mech, tbl)
        assert filecmp.cmp(tbl, os.path.join(ROOT, "pelelm_amd", "mechanisms", f"{name}.tbl"), shallow=False)
        os.makedirs(os.path.join(tmp, "csrc", "mech"))
        os.makedirs(os.path.join(tmp, "mechanisms"))
        inc = os.path.join(tmp, "csrc", "mech", f"{name}.inc")
        hipgen.write_inc(mech, inc)
        assert filecmp.cmp(inc, os.path.join(ROOT, "pelelm_amd", "csrc", "mech", f"{name}.inc"), shallow=False)


def test_op_counts_are_positive():
    for name in ("LiDryer", "drm19"):
        mech = M.load(name)
        c = hipgen.count_ops(hipgen.gen_wdot(mech))
        assert c["flops"] > 10 * mech.nreac
        assert c["by_function"]["exp"] >= mech.nreac
