"""The C++ module facades on the GPU: a C++ program (tests/facade_cfg1.cpp)
calls NEIGHBOR_LIST::Update, the LJ and solvent_LJ direct calls with
neighbor_list.d_nl, Long_Range_Correction, the PME excluded and reciprocal
calls, Step_Print and Particle_Mesh::Update_Box in the order of the
reference's force loop (SPONGE/main.cpp:1355-1626) on the 1,944-atom water
box, and its results are checked against the oracle: forces within 1e-5
relative RMS of the FP32 restatement and of FP64, energy and virial totals
as in the module parity tests (check_total), the d_nl rows equal to the
brute-force half list, solvent_LJ off (mdin solvent_LJ = false) and on (the
default gating of an NVT run: every atom a water)."""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

from gpu_helpers import check_total, rel, rel_rms

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORCE_TOL = 1e-5
SCALE = np.float32(1.02)
DT = np.float32(0.002)


def build_program(out_dir):
    """g++ build of the test program against the in-tree library."""
    lib = os.path.join(ROOT, "sponge_amd", "libsponge_amd.so")
    if not os.path.exists(lib):
        pytest.fail("libsponge_amd.so is not built")
    exe = os.path.join(out_dir, "facade_cfg1")
    subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", ROOT,
                    "-I", "/opt/rocm/include", os.path.join(ROOT, "tests", "facade_cfg1.cpp"),
                    lib, "-L/opt/rocm/lib", "-lamdhip64",
                    f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", exe], check=True)
    return exe


def write_inputs(d, s, mdin):
    os.makedirs(d, exist_ok=True)
    res = s.residue_starts()
    el = s.excl_list if len(s.excl_list) else np.zeros(1, np.int32)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{s.n_atoms} {s.n_types} {len(res)} {len(s.excl_list)} {s.cutoff!r} "
                f"{s.skin!r} {float(s.box[0])!r} {float(s.box[1])!r} {float(s.box[2])!r} "
                f"{float(SCALE)!r} {float(DT)!r}\n")
    with open(os.path.join(d, "mdin.txt"), "w") as f:
        f.write(mdin)
    arrays = dict(crd=s.crd, charge=s.charge, mass=s.mass, A=s.lj_A, B=s.lj_B,
                  crd2=(s.crd * SCALE).astype(np.float32))
    for k, v in arrays.items():
        np.ascontiguousarray(v, np.float32).tofile(os.path.join(d, k))
    ints = dict(type=s.lj_type, es=s.excl_start, en=s.excl_num, el=el, res=res)
    for k, v in ints.items():
        np.ascontiguousarray(v, np.int32).tofile(os.path.join(d, k))


def read(d, name, dtype=np.float32):
    return np.fromfile(os.path.join(d, name), dtype)


@pytest.fixture(scope="module")
def cfg1():
    from sponge_amd import system as S
    return S.config(1)


@pytest.fixture(scope="module")
def oracle_ref(cfg1):
    from oracle import oracle as O
    s = cfg1
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, flags = O.nl_build(s)
    assert flags == 0
    cb, sb, _ = O.nl_bruteforce(s)
    pme = O.PME(s, s.fft, beta)
    s2 = dataclasses.replace(s, box=(s.box * SCALE).astype(np.float32),
                             crd=(s.crd * SCALE).astype(np.float32))
    return dict(
        beta=beta, counts=c, serial=ser, pairs=O.pair_array(cb, sb),
        plain=O.lj_direct(s, c, ser, beta),
        solv_rest=O.lj_direct(s, c, ser, beta, solvent_n=s.n_atoms),
        solv=O.solvent_lj(s, c, ser, beta, solvent_atom=0),
        direct64=O.direct64(s, beta), excl=O.pme_excluded(s, beta),
        excl64=O.excluded64(s, beta), recip=pme.reciprocal(),
        recip64=O.pme64(s.crd, s.charge, s.box, beta, s.fft),
        long_factor=O.lj_long_range_factor(s),
        recip2=O.PME(s2, s.fft, beta).reciprocal())


MDIN = """cutoff = {cutoff}
skin = {skin}
{solvent}
[PM]
fftx = {fx}
ffty = {fy}
fftz = {fz}
Direct_Tolerance = {tol}
"""


@pytest.mark.parametrize("solvent", ["off", "default"])
def test_facade_force_loop_against_oracle(tmp_path, cfg1, oracle_ref, solvent):
    from oracle import oracle as O
    s, r = cfg1, oracle_ref
    mdin = MDIN.format(cutoff=s.cutoff, skin=s.skin, fx=s.fft[0], fy=s.fft[1], fz=s.fft[2],
                       tol=s.tolerance, solvent="solvent_LJ = false" if solvent == "off" else "")
    d = str(tmp_path / "case")
    write_inputs(d, s, mdin)
    exe = build_program(str(tmp_path))
    print(f"facade program, solvent_LJ {solvent}, N={s.n_atoms}", flush=True)
    out = subprocess.run([exe, d], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    N = s.n_atoms
    sc = read(d, "out_scalars")
    (lj_short, lj_long, ee, pm_direct, pm_recip, pm_self, pm_corr, beta, fx, fy, fz, n_solv,
     solv_on) = sc
    assert beta == np.float32(r["beta"]) and (fx, fy, fz) == tuple(float(x) for x in s.fft)
    if solvent == "off":
        assert solv_on == 0 and n_solv == 0
        direct = r["plain"]
    else:
        # NVT gating: 648 waters, all atoms in the tail
        assert solv_on == 1 and n_solv == N
        direct = {k: r["solv_rest"][k] + r["solv"][k] for k in r["plain"]}
        assert rel_rms(direct["frc"], r["plain"]["frc"]) < 1e-6

    # the list as d_nl holds it: the brute-force half list
    counts = read(d, "out_nl_counts", np.int32)
    serial = read(d, "out_nl_serial", np.int32).reshape(N, -1)
    assert np.array_equal(O.pair_array(counts, serial), r["pairs"])

    frc = read(d, "out_frc").reshape(N, 3)
    want = direct["frc"] + r["excl"]["frc"] + r["recip"]["frc"]
    want64 = r["direct64"]["frc"] + r["excl64"]["frc"] + r["recip64"]["frc"]
    assert rel_rms(frc, want) < FORCE_TOL
    assert rel_rms(frc, want64) < FORCE_TOL

    d64 = r["direct64"]
    check_total("LJ_short", lj_short, direct["lj_energy"].sum(dtype=np.float64), d64["lj"])
    volume = float(np.prod(s.box.astype(np.float64)))
    assert rel(lj_long, r["long_factor"] / volume) < 1e-6
    check_total("PM_direct", pm_direct, direct["direct_energy"].sum(dtype=np.float64),
                d64["coulomb"])
    edir = read(d, "out_direct_atom_energy")
    assert rel_rms(edir, direct["direct_energy"]) < 1e-5
    check_total("PM_reciprocal", pm_recip, r["recip"]["reciprocal"], r["recip64"]["reciprocal"])
    assert rel(pm_self, r["recip"]["self_energy"]) < 1e-6
    check_total("PM_correction", pm_corr, r["excl"]["correction_energy"].sum(dtype=np.float64),
                r["excl64"]["energy"])
    assert rel(ee, pm_direct + pm_recip + pm_self + pm_corr) < 1e-6
    # per-atom energies: LJ + direct Coulomb of the pair kernels and the
    # excluded-pair correction; the module-level sums: long-range LJ and the
    # reciprocal + self + neutralizing energy
    e_atoms = read(d, "out_energy").astype(np.float64).sum()
    check_total("atom energies", e_atoms,
                direct["atom_energy"].sum(dtype=np.float64)
                + r["excl"]["atom_energy"].sum(dtype=np.float64),
                d64["lj"] + d64["coulomb"] + r["excl64"]["energy"])
    pot = read(d, "out_potential")[0]
    assert rel(pot, r["recip"]["potential"] + lj_long) < 1e-6
    vir_atoms = read(d, "out_atom_virial").reshape(N, 6).astype(np.float64).sum(0)
    check_total("atom virials", vir_atoms,
                direct["atom_virial"].sum(0, dtype=np.float64)
                + r["excl"]["atom_virial"].sum(0, dtype=np.float64),
                d64["virial"] + r["excl64"]["virial"])
    vir = read(d, "out_virial").astype(np.float64)
    lr = np.array([2, 0, 2, 0, 0, 2]) * r["long_factor"] / volume
    check_total("module virial", vir, r["recip"]["virial"] + lr, r["recip64"]["virial"] + lr)

    # Update_Box: the next reciprocal call in the scaled box
    frc2 = read(d, "out_frc2").reshape(N, 3)
    assert rel_rms(frc2, r["recip2"]["frc"]) < FORCE_TOL
    rec2, self2, *corners, refused = read(d, "out_scalars2")
    assert rel(rec2, r["recip2"]["reciprocal"]) < 1e-5
    assert rel(read(d, "out_potential2")[0], r["recip2"]["potential"]) < 1e-5
    assert np.allclose(corners[:3], 0.0)
    assert np.allclose(corners[3:], s.box * SCALE, rtol=1e-6)
    # a list pointer no NEIGHBOR_LIST made is an error, not a silent call
    assert refused == 1
