"""CPU-side checks of the C-ABI library: it loads, exports every declared entry point, and
its host-side table construction equals the oracle's bit for bit."""
import os
import re

import numpy as np
import pytest

from molsim_amd import engine, systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(msg_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = engine.lib()
    decl = _declared_functions()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(L, name), name


def test_binding_constants_match_header():
    """The Python mirror's type limit and mode codes are the header's."""
    text = open(os.path.join(ROOT, "include", "molsim_gpu.h")).read()
    assert int(re.search(r"#define MSG_MAX_TYPES (\d+)", text).group(1)) == engine.MAX_TYPES
    for name, code in (("MSG_LIST_CELL", engine.LIST_MODES["cell"]), ("MSG_EWALD_SPM", engine.EWALD_MODES["spm"])):
        m = re.search(name + r"\s*=\s*(\d+)", text) or re.search(r"#define " + name + r" (\d+)", text)
        assert m and int(m.group(1)) == code, name


@pytest.mark.parametrize("make", [systems.electrolyte_primitive, systems.polyelectrolyte])
def test_build_table_matches_oracle_bitwise(orc, make):
    # the potentials of the two workloads, built without generating positions
    if make is systems.electrolyte_primitive:
        pot = systems._pot(2, [1.0, -1.0], 0, 40.0, systems.relpermitt_from_bjerrum(7.1), 0.08,
                           rad=[2.0, 2.0])
    else:
        eps = 0.1 / systems.beta_kjmol()
        pot = systems._pot(2, [-1.0, 1.0], 1, 20.0, systems.relpermitt_from_bjerrum(7.1), 0.15,
                           sig=[4.0, 4.0], eps=[eps, eps], rad=[2.0, 2.0])
    a = engine.build_table(pot)
    b = orc.build_table(pot)
    assert np.array_equal(a["ubuf"], b["ubuf"])
    assert np.array_equal(a["iubuflow"] - 1, b["iubuflow"])
    assert np.array_equal(a["nugrid"], b["nugrid"])
    assert np.array_equal(a["r2umin"], b["r2umin"])
    assert np.array_equal(a["r2atat"], b["r2atat"])


def test_build_table_error_paths():
    pot = systems._pot(1, [1.0], 0, 10.0, 78.0, 0.3)
    with pytest.raises(engine.MolsimGpuError) as e:
        engine.build_table(pot, maxbuf=24)   # two records only
    assert e.value.code == -5


@pytest.mark.parametrize("seed", [-1, -12345, 7, -2147483646])
def test_random_generator_matches_oracle(orc, seed):
    a = engine.Random(seed)
    b = orc.Rng(seed)
    xa = [a() for _ in range(5000)]
    xb = [b() for _ in range(5000)]
    assert xa == xb
    assert a.iseed.value == b.idum.value
