"""The C++ host layer on the GPU: the reference's per-functor loop (Simulation.cpp:49-74)
driven through komd_md against the FP64 oracle loop, the device-queued loop against the
per-functor loop, and the reference's own example checkpoint through the pair kernel."""
import json
import os
import subprocess

import numpy as np
import pytest

import komd_amd as K
from komd_amd import systems as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "komd_amd", "komd_md")
IBI_PS = os.path.join(ROOT, "tests", "golden", "ref_ibi_ckpt_100k_2b_equil.ps")
NU = 0.073
# reduced units in the reference's deck units (sigma = epsilon = m = 1, tau = 1 ps)
T_K = 0.8 / 0.8314462609869346          # T* = 0.8
E3B = NU / (1e9 * 6.022140754e22)         # nu* = 0.073


@pytest.fixture(scope="module")
def orc():
    import oracle
    oracle.build()
    return oracle


def g(x):
    return f"{float(x):.17g}"


def write_ps(path, s):
    with open(path, "w") as fh:
        for i in range(s.n):
            fh.write(f"{int(s.mol_id[i])}\t{g(s.r[i,0])} {g(s.r[i,1])} {g(s.r[i,2])}\t{g(s.v[i,0])} {g(s.v[i,1])} "
                     f"{g(s.v[i,2])}\t{g(s.mass[i])} {g(s.eps[i])} {g(s.sig[i])}\n")


def deck_for(tmp_path, s, steps, extra=""):
    ps = str(tmp_path / "start.ps")
    write_ps(ps, s)
    body = (f"temperature {g(T_K)}\ndelta_t 0.005\ncutoff 2.5\ncell_size {g(s.cell_size)}\ntimesteps {steps}\n"
            f"checkpoint_file {ps} 0 0 0\nlimit 1e+12\nenergy_3b {g(E3B)}\n"
            f"domain_low {g(s.low[0])} {g(s.low[1])} {g(s.low[2])}\n"
            f"domain_high {g(s.high[0])} {g(s.high[1])} {g(s.high[2])}\n"
            f"enable_3b 1\nCOMP 0 1.6605390688282807e-23 1 1 0 0 0\n{extra}")
    p = tmp_path / "run.dat"
    p.write_text(body)
    return str(p)


def run_md(*args):
    print("komd_md", " ".join(args), flush=True)
    res = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    return json.loads(res.stdout.strip().splitlines()[-1])


def test_reference_loop_matches_oracle(orc, tmp_path):
    s = S.fcc(6, 0.8, 0.8, 71, jitter=0.05)
    d = deck_for(tmp_path, s, 10)
    out_ps = str(tmp_path / "end.ps")
    info = run_md(d, "--write", out_ps)
    assert info["steps"] == 10 and info["sites"] == s.n
    got = np.loadtxt(out_ps)
    r0, b = orc.update_container(s)
    f0 = orc.lj_c08(s, r0, b) + orc.atm_c08(s, r0, b, NU, sign=1.0)[0]
    # the host's first thermostat call rescales the initial velocities before the loop
    T = orc.kinetic_sum(s, r0, s.v)
    s2 = s.copy()
    s2.v = s.v * np.sqrt(0.8 / (T[0] / (3 * T[1])))
    r_o, v_o, _ = orc.run(s2, 10, dt=0.005, nu=NU, atm=1, atm_sign=1.0, T0=0.8, temp_factor=1.0, nthreads=8,
                          f=f0, r=r0)
    assert np.abs(got[:, 1:4] - r_o).max() < 1e-6
    assert np.abs(got[:, 4:7] - v_o).max() < 1e-5
    assert info["temperature"] == pytest.approx(T_K, rel=1e-8)


def test_device_loop_matches_reference_loop(tmp_path):
    s = S.fcc(6, 0.8, 0.8, 72, jitter=0.05)
    d = deck_for(tmp_path, s, 12)
    a, b = str(tmp_path / "a.ps"), str(tmp_path / "b.ps")
    ia = run_md(d, "--write", a)
    ib = run_md(d, "--device-loop", "--write", b)
    ra, rb = np.loadtxt(a), np.loadtxt(b)
    assert ia["loop"] == "reference" and ib["loop"] == "device"
    # same kernels, different queueing: FP32 summation order of the pair pass only
    assert np.abs(ra[:, 1:4] - rb[:, 1:4]).max() < 1e-9
    assert np.abs(ra[:, 4:7] - rb[:, 4:7]).max() < 1e-7


def test_device_loop_skin_key(tmp_path):
    """A deck's skin key puts the device loop (NVE) on Verlet lists: the same trajectory as the
    per-step lists to the FP32 level of the relative coordinates."""
    s = S.fcc(6, 0.8, 0.8, 74, jitter=0.05, rc=2.8)   # cells of edge >= cutoff + skin
    out = []
    for extra in ("thermostat 0\n", "thermostat 0\nskin 0.3\n"):
        d = deck_for(tmp_path, s, 30, extra=extra)
        path = str(tmp_path / f"skin{len(out)}.ps")
        info = run_md(d, "--device-loop", "--write", path)
        assert info["loop"] == "device"
        out.append(np.loadtxt(path))
    dr = np.abs(out[0][:, 1:4] - out[1][:, 1:4])
    dr = np.minimum(dr, np.abs(dr - float(s.box[0])))
    print(f"device loop with and without skin: max |dr| {dr.max():.2e}")
    assert dr.max() < 1e-5


def test_surrogate_deck(tmp_path):
    s = S.fcc(6, 0.8, 0.8, 73, jitter=0.05)
    w = os.path.join(ROOT, "komd_amd", "weights", "atm_surrogate.npy")
    d = deck_for(tmp_path, s, 20, extra=f"surrogate_weights {w}\njudge_every 10\njudge_max 0.5\n")
    info = run_md(d, "--device-loop")
    assert info["steps"] == 20 and np.isfinite(info["u_3b"]) and info["u_3b"] > 0


def test_reference_checkpoint_lj_forces(orc):
    """examples/IBI/ckpt_100k_2b_equil.ps (343 sites, sigma 1.5, internal epsilon 0.6,
    mass 4) in its deck's box (15^3, cutoff 5): LJ forces against the oracle."""
    a = np.loadtxt(IBI_PS)
    s = S.System(r=a[:, 1:4].copy(), v=a[:, 4:7].copy(), mass=a[:, 7].copy(), eps=a[:, 8].copy(),
                 sig=a[:, 9].copy(), mol_id=a[:, 0].astype(np.uint64), low=np.zeros(3), high=np.full(3, 15.0),
                 cell_size=5.0, name="ibi_ckpt")
    r, b = orc.update_container(s, )
    f_or = orc.lj_c08(s, r, b, cutoff=5.0)
    with K.Context(s, cutoff=5.0) as c:
        c.update_container()
        c.forces_pair()
        f = c.forces()
    scale = np.sqrt(np.mean(f_or ** 2))
    dd = np.linalg.norm(f - f_or, axis=1)
    print(f"reference checkpoint LJ: max {dd.max() / scale:.2e} rms {np.sqrt(np.mean(dd ** 2)) / scale:.2e}")
    assert dd.max() / scale < 2e-4 and np.sqrt(np.mean(dd ** 2)) / scale < 5e-6


def test_sensor_files(orc, tmp_path):
    """RDFSensor and PressureSensor of komd_md (deck keys enable_sensor_rdf/_pres): one
    measured step, files written at the end (RDFSensor.cpp:37-57, PressureSensor.cpp:34-70),
    checked against the oracle's RDF counts and the library's stress tensor on the state
    komd_md wrote."""
    s = S.fcc(6, 0.8, 0.8, 74, jitter=0.05)
    out = tmp_path / "sensors"
    out.mkdir()
    d = deck_for(tmp_path, s, 1, extra=(f"thermostat 0\nwrite_freq 100\nenable_sensor_rdf 1\nsensor_rdf_max 4.0\n"
                                        f"sensor_rdf_dr 0.1\nenable_sensor_pres 1\nsensor_dir {out}\n"))
    end = str(tmp_path / "end.ps")
    info = run_md(d, "--write", end)
    assert info["sensors"] == 2
    a = np.loadtxt(end)
    r = a[:, 1:4].copy()
    vol = float(np.prod(s.box))
    # RDF: one sample of ordered-pair counts -> g(r) as the reference writes it
    cnt = orc.rdf(orc.molecule_coms(s, r), s.box, 4.0, 0.1, 40)
    rr, g_ref = K.rdf_normalise(cnt, 1, s.n, vol, 0.1)
    rdf = np.loadtxt(out / "rdf_1.txt")
    assert rdf.shape == (40, 8)
    assert np.allclose(rdf[:, 0], rr, rtol=1e-5)
    assert np.allclose(rdf[:, 1], g_ref, rtol=2e-5, atol=1e-12)
    assert np.allclose(rdf[:, 2], cnt, rtol=2e-5)
    # pressure: (kinetic + virial tensor) / V in Pa, lower triangle XX / XY YY / XZ YZ ZZ
    lines = (out / "pressures_1.txt").read_text().split("\n")
    vals = np.array([float(x) for ln in lines[:3] for x in ln.split()])
    assert [len(ln.split()) for ln in lines[:3]] == [1, 2, 3]
    s2 = s.copy()
    s2.r, s2.v = r, a[:, 4:7].copy()
    with K.Context(s2, nu=NU) as c:
        o = c.observables(K.ATM_REF)
    ref = (o["kinetic_tensor"] + o["virial_tensor"]) / vol * 1.66053906892e7
    assert np.allclose(vals, ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max())


def _lj_table_files(tmp_path):
    x = np.linspace(0.7, 2.5, 4001)
    f = 24.0 * (2.0 * x ** -13 - x ** -7)
    u = 4.0 * (x ** -12 - x ** -6)
    ff, uf = tmp_path / "force.txt", tmp_path / "pot.txt"
    np.savetxt(ff, np.c_[x, f], fmt="%.17g", delimiter="\t")
    np.savetxt(uf, np.c_[x, u], fmt="%.17g", delimiter="\t")
    return str(ff), str(uf)


@pytest.mark.parametrize("pair", ["atm2b", "table"])
def test_pair_functor_decks(tmp_path, pair):
    """Decks that put ATM2B (enable_3b_approx, main.cpp:56-57) or a tabulated force
    (pair_table: the IBI functor, IBI.cpp:51-55) in LJ's slot: the energy before the first
    step is LJ's times the ATM2B factor, or LJ's to interpolation error for an LJ-sampled
    table; the device loop follows the per-functor loop over 12 steps."""
    s = S.fcc(6, 0.8, 0.8, 75, jitter=0.05)
    if pair == "atm2b":
        extra = "enable_3b 0\nenable_3b_approx 1\n"
    else:
        ff, uf = _lj_table_files(tmp_path)
        extra = f"enable_3b 0\npair_table {ff}\npair_table_potential {uf}\n"
    base = run_md(deck_for(tmp_path, s, 0, extra="enable_3b 0\n"), "--steps", "0")
    d = deck_for(tmp_path, s, 12, extra=extra)
    i0 = run_md(d, "--steps", "0")
    assert i0["pair"] == pair and base["pair"] == "lj12_6"
    if pair == "atm2b":
        fac = 1.0 - 2.0 * NU * (s.n / float(np.prod(s.box))) / 3.0
        assert i0["u_lj"] == pytest.approx(fac * base["u_lj"], rel=1e-6)
    else:
        assert i0["u_lj"] == pytest.approx(base["u_lj"], rel=1e-4)
    a, b = str(tmp_path / "a.ps"), str(tmp_path / "b.ps")
    run_md(d, "--write", a)
    run_md(d, "--device-loop", "--write", b)
    ra, rb = np.loadtxt(a), np.loadtxt(b)
    assert np.abs(ra[:, 1:4] - rb[:, 1:4]).max() < 1e-9
    assert np.abs(ra[:, 4:7] - rb[:, 4:7]).max() < 1e-7
