"""C++ host layer (komd_amd/host, driver komd_md) on the CPU: deck parsing with the
reference's keys and checks (IO/FileInput.cpp:11-121) and the .ps checkpoint reader /
writer (IO/CheckpointIO.cpp:13-76), on the reference's own example checkpoint
(examples/IBI/ckpt_100k_2b_equil.ps, copied as a data fixture)."""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "komd_amd", "komd_md")
GOLD = os.path.join(ROOT, "tests", "golden")
IBI_PS = os.path.join(GOLD, "ref_ibi_ckpt_100k_2b_equil.ps")


@pytest.fixture(scope="module", autouse=True)
def driver():
    if not os.path.exists(EXE):
        from komd_amd import build
        build.build_host()
    return EXE


def deck(tmp_path, body, name="run.dat"):
    p = tmp_path / name
    p.write_text(body)
    return str(p)


IBI_DECK = """# reference-style deck (examples/inp.dat keys)
temperature 100
delta_t 0.0005
cutoff 5.0
cell_size 5.0
timesteps 7
checkpoint_file {ps}    0 0 0
limit 1e+9
energy_3b 10
domain_low 0 0 0
domain_high 15 15 15
enable_3b 1
COMP 0 1.0e-21 1.5 4.0 0 0 0
"""


def run(*args, check=True):
    res = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60)
    if check:
        assert res.returncode == 0, res.stderr
    return res


def test_deck_and_reference_checkpoint(tmp_path):
    d = deck(tmp_path, IBI_DECK.format(ps=IBI_PS))
    out = json.loads(run(d, "--check").stdout)
    ref = np.loadtxt(IBI_PS)
    assert out["sites"] == len(ref) == 343 and out["molecules"] == 343 and out["components"] == 1
    assert out["first_site"] == list(ref[0, 1:4])
    assert out["cutoff"] == 5.0 and out["cell_size"] == 5.0 and out["timesteps"] == 7
    assert out["domain_high"] == [15.0, 15.0, 15.0] and out["enable_3b"] == 1
    # ATM.cpp:13: nu = energy_3b * 1e9 * conv_J_Ei
    assert out["nu_internal"] == pytest.approx(10 * 1e9 * 6.022140754e22, rel=1e-15)


def test_checkpoint_round_trip(tmp_path):
    d = deck(tmp_path, IBI_DECK.format(ps=IBI_PS))
    w = str(tmp_path / "out.ps")
    run(d, "--check", "--write", w)
    a, b = np.loadtxt(IBI_PS), np.loadtxt(w)
    assert np.array_equal(a, b)
    # the written file is itself a valid checkpoint with an offset applied on load
    d2 = deck(tmp_path, IBI_DECK.format(ps=w).replace(f"{w}    0 0 0", f"{w}    1.5 0 -2"), "run2.dat")
    out = json.loads(run(d2, "--check").stdout)
    assert out["first_site"] == [a[0, 1] + 1.5, a[0, 2], a[0, 3] - 2]


def test_reference_test_ps_multisite(tmp_path):
    # examples/test/test.ps: four single-site molecules with ids 1..4
    d = deck(tmp_path, IBI_DECK.format(ps=os.path.join(GOLD, "ref_test_ps.txt")))
    out = json.loads(run(d, "--check").stdout)
    assert out["sites"] == 4 and out["molecules"] == 4


@pytest.mark.parametrize("mutation,message", [
    (lambda s: s.replace("COMP 0 1.0e-21 1.5 4.0 0 0 0\n", ""), "no components defined"),
    (lambda s: s.replace("cell_size 5.0", "cell_size 4.0"), "cell size must be larger or equal to cutoff radius"),
    (lambda s: "\n".join(l for l in s.splitlines() if not l.startswith("checkpoint_file")), "No phasespace data"),
    (lambda s: s.replace("COMP 0", "COMP -1"), "negative component id"),
    (lambda s: s + "pair_steps 0\n", "pair_steps and triple_steps must be positive"),
    (lambda s: s + "skin 0.5\n", "cell size must be at least cutoff + skin"),
    (lambda s: s + "skin -1\n", "skin must be non-negative"),
])
def test_deck_errors(tmp_path, mutation, message):
    d = deck(tmp_path, mutation(IBI_DECK.format(ps=IBI_PS)))
    res = run(d, "--check", check=False)
    assert res.returncode == 1 and message in res.stderr


def test_pair_list_keys(tmp_path):
    """The reference's pair_steps / triple_steps keys (FileInput.cpp:36-37, defaults 4 / 8 in
    util/defaults.h:39-40) are read; the device loop's Verlet-list skin is its own key."""
    d = deck(tmp_path, IBI_DECK.format(ps=IBI_PS).replace("cell_size 5.0", "cell_size 5.5")
             + "pair_steps 2\ntriple_steps 3\nskin 0.5\n")
    out = json.loads(run(d, "--check").stdout)
    assert (out["pair_steps"], out["triple_steps"], out["skin"]) == (2, 3, 0.5)
    out = json.loads(run(deck(tmp_path, IBI_DECK.format(ps=IBI_PS)), "--check").stdout)
    assert (out["pair_steps"], out["triple_steps"], out["skin"]) == (4, 8, 0.0)


@pytest.mark.parametrize("extra,pair", [("", "lj12_6"), ("enable_3b_approx 1\n", "atm2b"),
                                        ("pair_table {t}\n", "table")])
def test_pair_functor_keys(tmp_path, extra, pair):
    """The reference's enable_3b_approx key (FileInput.cpp:42) puts ATM2B in LJ's slot
    (main.cpp:56-57); pair_table puts the tabulated force there."""
    t = tmp_path / "force.txt"
    t.write_text("0.9\t10\n1.5\t2\n2.5\t0")
    d = deck(tmp_path, IBI_DECK.format(ps=IBI_PS) + extra.format(t=t))
    out = json.loads(run(d, "--check").stdout)
    assert out["pair"] == pair
