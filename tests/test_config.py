"""The reference-configuration helpers of include/icarion_config.hpp (C++17),
compiled with g++ into a small driver (tests/cpp/config_driver.cpp) that fills
their plain structs from the reference's own example configurations and species
database (copied unchanged under tests/golden/config/ from the reference's
examples/ and data/), against closed forms:

  * n = P / (kB T) and the gas-table mass, radius and polarizability
    (EnvironmentConfig.h:150-282, src/utils/constants.h:41-65), a mixture's
    component densities x_i / sum(x) n;
  * E/N -> axial_V = E/N 1e-21 n L, and axial_V given explicitly wins
    (DomainConfigLoader.cpp:413-440);
  * HSS sigma: CCS_HSS[gas] > derived from the reference gas, pi (sqrt(sigma /
    pi) - r_ref + r_gas)^2 > the ion CCS (HSSCollisionHandler.cpp:136-172,
    263-310);
  * EHSS sigma: OAPA over the molecule's hard spheres, equal to the oracle's
    OAPA to rounding (itself pinned by the one-atom closed form), and the atom
    table centred on the centre of mass with radius LJ sigma / 2
    (geometryUtils.h:35-73);
  * the fused-step gate: the engine conditions under which a device step that
    collides every ion cannot collide any ion twice
    (SimulationEngineCollisionRuntime.cpp:26-63).
CPU only: no device is involved."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CFG = os.path.join(ROOT, "tests", "golden", "config")
KB = 1.380649e-23
AMU = 1.66053906660e-27
GAS = {"He": (4.002602 * AMU, 1.3e-10, 0.195e-30), "N2": (28.0134 * AMU, 1.82e-10, 1.74e-30),
       "O2": (31.9988 * AMU, 1.73e-10, 1.567e-30), "Ar": (39.948 * AMU, 1.70e-10, 1.596e-30)}


@pytest.fixture(scope="module")
def driver(config_driver):
    return config_driver.exe


def run(driver, config, collision=None):
    args = [driver, os.path.join(CFG, config), os.path.join(CFG, "species_database_v1.json"), CFG]
    if collision:
        args.append(collision)
    out = subprocess.run(args, check=True, capture_output=True, text=True).stdout
    return json.loads(out)


def close(a, b, rel=1e-14):
    return a == pytest.approx(b, rel=rel, abs=0.0)


def derived_sigma(sigma_ref, r_ref, r_tgt):
    r_ion = max(0.0, math.sqrt(sigma_ref / math.pi) - r_ref)
    return math.pi * (r_ion + r_tgt) ** 2


def test_ims_basic_domains_and_hss_sigma(driver):
    d = run(driver, "ims_basic.json")
    assert d["physics"] == {"integrator": 0, "collision_model": 1, "rng_seed": 123}
    drift, det = d["domains"]
    x = drift["desc"]
    n = 500.0 / (KB * 300.0)
    assert close(x["gas_density_m3"], n)
    assert close(x["dc_axial_V"], 10.0 * 1e-21 * n * 0.06)
    assert x["instrument"] == 0 and close(x["origin_z_m"], -0.01) and close(x["radius_m"], 0.01)
    m, r, a = GAS["He"]
    assert close(x["gas_mass_kg"], m) and x["gas_radius_m"] == r and x["gas_polarizability_m3"] == a
    # the detection region gives axial_V directly
    assert det["desc"]["dc_axial_V"] == 1000.0
    assert close(det["desc"]["gas_density_m3"], 200.0 / (KB * 300.0))
    # H3O+: CCS_HSS[He] = 24.9 A^2 from the database; 26DTBPH+: no map, no
    # reference gas -> its CCS_A2 (85.8 A^2)
    h3o, dtbp = drift["species"]
    assert close(h3o["sigma_hss_m2"], 24.9e-20) and close(h3o["ccs_m2"], 24.9e-20)
    assert close(dtbp["sigma_hss_m2"], 85.8e-20)
    assert h3o["K0_cm2_Vs"] == 24.1 and h3o["sigma_ehss_m2"] == 0.0
    assert drift["error"] is None and drift["atoms_xyzr"] == []


def test_quadrupole_and_orbitrap_fields(driver):
    q = run(driver, "quadrupole_basic.json")
    x = q["domains"][0]["desc"]
    assert x["instrument"] == 2 and x["rf_voltage_V"] == 100.0 and x["rf_frequency_Hz"] == 2e6
    assert x["dc_quad_V"] == 0.0 and x["dc_axial_V"] == 500.0 and x["radius_m"] == 0.005
    assert close(x["gas_mass_kg"], GAS["N2"][0]) and x["gas_radius_m"] == GAS["N2"][1]
    assert q["physics"]["collision_model"] == 0
    assert q["domains"][0]["species"][0]["sigma_hss_m2"] == 0.0   # no collisions: no sigma needed
    o = run(driver, "orbitrap_basic.json")
    x = o["domains"][0]["desc"]
    assert x["instrument"] == 3 and o["physics"]["integrator"] == 2   # RK45
    assert (x["radius_in_m"], x["radius_out_m"], x["radius_char_m"]) == (0.006, 0.015, 0.022)
    assert x["dc_radial_V"] == 3500.0 and x["dc_axial_V"] == 0.0


def test_gas_mixture_densities_and_per_component_sigma(driver):
    d = run(driver, "ims_multi_gas_air.json")["domains"][0]
    n = 101325.0 / (KB * 300.0)
    fr = {"N2": 0.78, "O2": 0.21, "Ar": 0.01}
    s = sum(fr.values())
    for comp, (g, f) in zip(d["mixture"], fr.items()):
        assert close(comp["density_m3"], f / s * n)
        assert close(comp["mass_kg"], GAS[g][0]) and comp["participates"] == 1
    # mixture means weighted by fraction
    x = d["desc"]
    assert close(x["gas_mass_kg"], sum(f / s * GAS[g][0] for g, f in fr.items()), rel=1e-14)
    assert close(x["gas_radius_m"], sum(f / s * GAS[g][1] for g, f in fr.items()), rel=1e-14)
    # H3O+ (reference gas He, CCS 24.9 A^2; CCS_HSS N2 = 104.02 A^2): N2 from the
    # database, O2 and Ar derived from the He reference
    sig = d["mixture_sigma_m2"]
    assert close(sig[0], 104.02e-20)
    assert close(sig[1], derived_sigma(24.9e-20, GAS["He"][1], GAS["O2"][1]), rel=1e-13)
    assert close(sig[2], derived_sigma(24.9e-20, GAS["He"][1], GAS["Ar"][1]), rel=1e-13)
    assert d["species"][0]["sigma_hss_m2"] == 0.0   # the mixture carries the sigma


def test_ehss_sigma_is_the_oapa_and_atoms_are_com_centred(driver, oracle):
    import ctypes as C
    d = run(driver, "ims_basic.json", collision="EHSS")["domains"][0]
    assert d["error"] is None
    atoms = np.array(d["atoms_xyzr"], dtype=np.float64).reshape(-1, 4)
    lib = oracle.load()
    r_gas = GAS["He"][1]
    for sid, name in ((0, "H3O+"), (1, "26DTBPH+")):
        sp = d["species"][sid]
        mol = json.load(open(os.path.join(CFG, "molecules", name + ".json")))["molecule"]["atoms"]
        assert sp["atom_count"] == len(mol)
        a = atoms[sp["atom_offset"]:sp["atom_offset"] + sp["atom_count"]]
        pos = np.array([at["pos"] for at in mol]) * 1e-10
        mass = np.array([at["mass_u"] for at in mol])
        com = (pos * mass[:, None]).sum(0) / mass.sum()
        # centre of mass at the origin (float32 table, to one float ulp: the
        # sums run in another order here), radius LJ sigma / 2
        a32 = a.astype(np.float32)
        np.testing.assert_allclose(a32[:, :3], (pos - com).astype(np.float32), rtol=2.4e-7, atol=1e-24)
        np.testing.assert_array_equal(
            a32[:, 3], (0.5 * np.array([at["LJ_sigma_angstrom"] for at in mol]) * 1e-10).astype(np.float32))
        # OAPA from the helper's own double geometry equals the oracle's OAPA on it
        geo = np.ascontiguousarray(
            np.concatenate([pos - com, 0.5e-10 * np.array([[at["LJ_sigma_angstrom"]] for at in mol])], 1))
        ref = lib.oracle_oapa_sigma(geo.ctypes.data_as(C.POINTER(C.c_double)), len(mol), r_gas)
        assert sp["sigma_ehss_m2"] == pytest.approx(ref, rel=1e-14, abs=0.0), name
        # bounded by the projected extent of the molecule
        rmax = np.max(np.linalg.norm(pos - com, axis=1) + a[:, 3]) + r_gas
        assert math.pi * (a[:, 3].max() + r_gas) ** 2 < ref <= math.pi * rmax ** 2 * (1 + 1e-12)


def test_one_atom_oapa_closed_form(oracle):
    """The OAPA of one atom at the centre is pi (r_atom + r_gas)^2 in every
    orientation (the oracle's OAPA, which the helper's equals to rounding)."""
    import ctypes as C
    geo = np.array([[0.0, 0.0, 0.0, 1.7e-10]])
    v = oracle.load().oracle_oapa_sigma(geo.ctypes.data_as(C.POINTER(C.c_double)), 1, 1.3e-10)
    assert v == pytest.approx(math.pi * 3.0e-10 ** 2, rel=1e-14)


def test_ehss_mixture_sigma_per_component(driver, oracle):
    """EHSS over the air mixture: each component's sigma is the OAPA of the
    geometry with that component's contact radius (the EHSS handler's mixture
    path calls compute_effective_ccs with comp.radius_m)."""
    import ctypes as C
    d = run(driver, "ims_multi_gas_air.json", collision="EHSS")["domains"][0]
    assert d["error"] is None
    assert d["species"][0]["sigma_ehss_m2"] == 0.0
    mol = json.load(open(os.path.join(CFG, "molecules", "H3O+.json")))["molecule"]["atoms"]
    pos = np.array([at["pos"] for at in mol]) * 1e-10
    mass = np.array([at["mass_u"] for at in mol])
    geo = np.ascontiguousarray(np.concatenate(
        [pos - (pos * mass[:, None]).sum(0) / mass.sum(),
         0.5e-10 * np.array([[at["LJ_sigma_angstrom"]] for at in mol])], 1))
    lib = oracle.load()
    for k, g in enumerate(("N2", "O2", "Ar")):
        ref = lib.oracle_oapa_sigma(geo.ctypes.data_as(C.POINTER(C.c_double)), len(mol), GAS[g][1])
        assert d["mixture_sigma_m2"][k] == pytest.approx(ref, rel=1e-14, abs=0.0), g
    # a larger neutral gives a larger sigma
    assert d["mixture_sigma_m2"][0] > d["mixture_sigma_m2"][1] > d["mixture_sigma_m2"][2]


def test_unsupported_configurations(driver):
    # damping over a mixture and models without a device path
    d = run(driver, "ims_multi_gas_air.json", collision="Friction")["domains"][0]
    assert d["error"].startswith("unsupported")
    out = subprocess.run([driver, os.path.join(CFG, "ims_basic.json"),
                          os.path.join(CFG, "species_database_v1.json"), CFG, "IPM"],
                         capture_output=True, text=True)
    assert out.returncode == 3 and "unsupported" in out.stdout


def test_fused_step_gate(driver):
    """Over the 64 combinations of the engine conditions:
      * the engine's own batch-collision path (SimulationEngineCollisionRuntime.cpp:46)
        needs no multi-event mode, one subcycle, no deep diagnostics, constant
        flow and a uniform dt: two rows (the maximum events per step counts only
        in multi-event mode), and each condition alone breaks it;
      * the fused device step, which runs the CPU path's collision substeps
        itself, needs no deep diagnostics, constant flow and a uniform dt: the
        eight rows with those, whatever the subcycles and multi-event mode;
      * its collision substeps are the subcycles, or max(subcycles, max events)
        in multi-event mode (:26-33)."""
    rows = json.loads(subprocess.run([driver, "--gate"], check=True, capture_output=True,
                                     text=True).stdout)
    assert len(rows) == 64
    engine = [r for r in rows if r[7] == 1]
    assert [r[:6] for r in engine] == [[0, 1, 1, 0, 1, 1], [0, 1, 3, 0, 1, 1]]
    base = dict(zip(("multi", "sub", "maxev", "deep", "flow", "udt"), (0, 1, 1, 0, 1, 1)))
    for key, bad in (("multi", 1), ("sub", 2), ("deep", 1), ("flow", 0), ("udt", 0)):
        row = dict(base, **{key: bad})
        match = [r for r in rows if r[:6] == list(row.values())]
        assert match and match[0][7] == 0, key
    for multi, sub, maxev, deep, flow, udt, ok, eng, nsub in rows:
        assert ok == int(deep == 0 and flow == 1 and udt == 1)
        assert nsub == (max(sub, maxev) if multi else sub)
        assert eng <= ok


def test_flow_models(driver):
    """to_desc's gas velocity for the flow models.  The rotated AxialUniform
    case is the reference's own engine test (test_simulation_engine.cpp:652-685,
    u = 100 m/s with local-to-global columns (0,0,-1), (0,1,0), (1,0,0)), whose
    collision handler receives (100, 0, 0) from gas_velocity_at; unrotated, the
    flow is (0, 0, u); the constant model passes gas_velocity_m_s through; the
    axial parabolic model (radius-dependent) is declined unless the caller
    applies gas_flow() through icarion_gpu_set_gas_flow, and flow_covered says
    which models the device step takes (the parabolic one on a single domain)."""
    out = json.loads(subprocess.run([driver, "--flow"], check=True, capture_output=True,
                                    text=True).stdout)
    assert out["uniform_rotated"] == pytest.approx([100.0, 0.0, 0.0], abs=1e-12)
    assert out["uniform"] == pytest.approx([0.0, 0.0, 134.0], abs=1e-12)
    assert out["constant"] == [1.0, 2.0, 3.0]
    assert out["parabolic"] == "unsupported"
    assert out["parabolic_applied"] == [0.0, 0.0, 0.0]
    assert out["parabolic_flow"] == [1, 250.0]
    assert out["covered"] == [1, 0, 1, 1]
    assert out["uniform_flow"] == [0, 0.0]
