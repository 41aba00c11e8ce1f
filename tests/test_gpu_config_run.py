"""The reference's own configuration files run on the device through the
include/icarion_config.hpp helpers (the drop-in path of simulation.enable_gpu):
the descriptors the helpers derive from examples/ims/ims_basic.json and the
species database (H3O+ and 26DTBPH+ in He at 500 Pa, 10 Td, RK4, dt 1 ns) drive
both the device and the FP64 oracle, which agree ion by ion (>= 99 % of ions to
1e-9, as the stochastic parity tests).  With the collision model set to EHSS
the helpers also build the hard-sphere atom table of the two molecules from
their geometry files, and the geometry EHSS path runs on those real
molecules."""
import math

import numpy as np
import pytest

from icarion_amd import _abi
from icarion_amd import workloads as W

pytestmark = pytest.mark.gpu


def domain_from(j):
    d = _abi.DomainDesc()
    for key in ("instrument", "use_field_array", "magnetic_enabled"):
        setattr(d, key, int(j[key]))
    for key in ("origin_m", "rot_g2l", "rot_l2g", "B_T", "gas_velocity_m_s"):
        arr = getattr(d, key)
        for k, v in enumerate(j[key]):
            arr[k] = v
    for key in ("length_m", "radius_m", "radius_in_m", "radius_out_m", "radius_char_m",
                "dc_axial_V", "dc_quad_V", "dc_radial_V", "rf_voltage_V", "rf_frequency_Hz",
                "field_array_scale", "temperature_K", "pressure_Pa", "gas_density_m3",
                "gas_mass_kg", "gas_radius_m", "gas_polarizability_m3"):
        setattr(d, key, j[key])
    d.detector_z_m = math.nan
    return d


def species_from(rows):
    return [W.species(sigma_hss=r["sigma_hss_m2"], sigma_ehss=r["sigma_ehss_m2"], ccs=r["ccs_m2"],
                      K0=r["K0_cm2_Vs"], atom_offset=r["atom_offset"], atom_count=r["atom_count"])
            for r in rows]


def cloud(rows, n_per_species, seed):
    """The config's ion block: per species a Gaussian at the origin (std 1 mm)
    with thermal velocities at 300 K."""
    sidx = np.repeat(np.arange(len(rows)), n_per_species).astype(np.uint32)
    masses = np.array([r["mass_amu"] for r in rows])[sidx]
    e = W.gaussian_cloud(sidx.size, masses, 1e-3, seed=seed)
    e.species[:] = sidx
    return e


@pytest.mark.parametrize("collision", [None, "EHSS"])
def test_ims_basic_config_on_device_equals_oracle(oracle, config_driver, collision):
    cfg = config_driver("ims_basic.json", collision)
    dom = cfg["domains"][0]
    assert dom["error"] is None
    d = domain_from(dom["desc"])
    p = cfg["physics"]
    ph = W.physics(p["integrator"], p["collision_model"], seed=p["rng_seed"])
    sp = species_from(dom["species"])
    atoms = (np.array(dom["atoms_xyzr"], dtype=np.float32).reshape(-1, 4)
             if dom["atoms_xyzr"] else None)
    ens = cloud(dom["species"], 1024, seed=5)
    g, o = ens.copy(), ens.copy()
    dt, steps = 1e-9, 2000
    from icarion_amd import gpu
    print(f"ims_basic.json drift domain, {collision or 'HSS'}: {ens.n} ions x {steps} steps",
          flush=True)
    with gpu.Handle() as h:
        h.set_domain(d, ph)
        h.set_species(sp, atoms)
        h.upload_state(g)
        sg = h.step(0.0, dt, steps, 0)
        h.download_state(g)
    so = oracle.run(d, ph, sp, o, 0.0, dt, steps, atoms=atoms)
    ok = np.ones(g.n, bool)
    for f, scale in (("x", 1e-6), ("y", 1e-6), ("z", 1e-6), ("vx", 1.0), ("vy", 1.0), ("vz", 1.0)):
        a, b = getattr(g, f), getattr(o, f)
        ok &= np.abs(a - b) <= 1e-9 * np.maximum(np.abs(b), scale)
    ok &= (g.active == o.active) & (g.n_collisions == o.n_collisions)
    print(f"{ok.mean():.4f} of ions equal the oracle; collisions {sg.collisions} vs {so.collisions}",
          flush=True)
    assert ok.mean() >= 0.99
    assert sg.ion_steps == so.ion_steps and sg.collisions > 0
    assert sg.collisions == pytest.approx(so.collisions, rel=2e-3)
