"""The oracle's pins (tests/oracle_pins.py), and proof that they bite: builds
of the oracle with a planted mistake (ORACLE_MUTANT=k) must each fail at
least one pin."""
from __future__ import annotations

import os
import subprocess

import pytest

import oracle_pins as P
from conftest import ROOT


@pytest.fixture(scope="module")
def lib():
    from oracle import pyoracle as po
    return po.load()


@pytest.mark.parametrize("pin", P.ALL_PINS, ids=lambda f: f.__name__)
def test_pin(lib, pin):
    pin(lib)


# planted mistake -> what it models
MUTANTS = {
    1: "triangle test: cross product operands transposed",
    2: "light attenuation: linear term dropped",
    3: "reflection: local colour not scaled by (1 - reflect)",
    4: "Beer factor: sign of the distance flipped",
    5: "dot product: summation order changed",
    6: "sphere: far root instead of near root",
    7: "ray generation: half-pixel offset added",
    8: "plane: parallel threshold compared in float",
    9: "quantisation: rounding instead of truncation",
    10: "closest hit: ties go to the later object",
    11: "plane texture coordinates: axisx and axisy swapped",
    12: "plane texture coordinates: the /5 dropped",
    13: "texel: bytes read as R, G, B instead of B, G, R",
    14: "box: the x face normal points the wrong way",
    15: "ballplane: sphere centres visited cy-major (identities reordered)",
    16: "texel: negative coordinates not wrapped into [0, 1)",
}


@pytest.mark.parametrize("k", sorted(MUTANTS), ids=lambda k: f"mutant{k}")
def test_planted_mistake_is_caught(tmp_path, k):
    from oracle import pyoracle as po
    so = os.path.join(str(tmp_path), f"liboracle_mut{k}.so")
    subprocess.run(["gcc", "-std=c11", "-O2", "-march=x86-64-v3", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                    f"-DORACLE_MUTANT={k}", "-shared", "-o", so, os.path.join(ROOT, "oracle", "oracle.c"), "-lm",
                    "-lpthread"], check=True)
    mlib = po.load(so)
    failed = []
    for pin in P.ALL_PINS:
        try:
            pin(mlib)
        except AssertionError:
            failed.append(pin.__name__)
    assert failed, f"mutant {k} ({MUTANTS[k]}) passed every pin"
