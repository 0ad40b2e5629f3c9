"""The per-element flop figures bench.py divides by are the op counts of the
oracle's element routines: tests/flopcount.cpp compiles the oracle with a
counting number type (additions/subtractions, multiplications, divisions and
square roots of one hex element, VEC+MAT and VEC only).

The NS figure is the hand count of the reference's expression trees (e3LHS
181 flop per pair, e3RHS, eval_shape_fem, e3int without the zero Hessian,
tau's); the restatement drops terms multiplied by the live call's exact
zeros (uprime, the Hessian), so its count sits a few percent below."""
import os
import shutil
import subprocess

import pytest

import bench

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def counts(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("flopcount") / "flopcount")
    subprocess.run(["g++", "-O0", "-std=c++17", "-fpermissive", "-w", "-o", exe,
                    os.path.join(HERE, "flopcount.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        name, add, mul, div, sq, tot = line.split()
        res[name] = dict(add=int(add), mul=int(mul), div=int(div), sqrt=int(sq),
                         total=int(tot))
    return res


def test_quench_figures_are_the_restatement_counts(counts):
    assert counts["nsvof_vecmat"]["total"] == bench.F_NSVOF
    assert counts["nsvof_vec"]["total"] == bench.F_NSVOF_VEC
    assert counts["tem_vecmat"]["total"] == bench.F_TEM
    assert counts["tem_vec"]["total"] == bench.F_TEM_VEC


def test_ns_figure_close_to_restatement_count(counts):
    c = counts["ns_vecmat"]["total"]
    assert 0.9 * bench.F_NS <= c <= bench.F_NS
    # the matrix part dominates: VEC alone is about a tenth
    assert counts["ns_vec"]["total"] < 0.15 * c


def test_counter_sees_every_operation_kind(counts):
    for name, c in counts.items():
        assert c["add"] > 0 and c["mul"] > 0 and c["sqrt"] > 0, name
        assert c["total"] == c["add"] + c["mul"] + c["div"] + c["sqrt"]
