"""The C-ABI library: it loads, exports every entry point include/icarion_gpu.h
declares, and the ctypes structs match the C layout (no device needed)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from icarion_amd import _abi, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "icarion_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t)\s+(icarion_gpu_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = gpu.load()
    names = declared_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(gpu.EXPORTED)


def test_abi_version():
    assert gpu.load().icarion_gpu_abi_version() == _abi.ABI_VERSION


def test_last_error_roundtrip_without_device():
    lib = gpu.load()
    h = C.c_void_p()
    rc = lib.icarion_gpu_create(None, C.byref(h))
    if rc == _abi.OK:  # a device is visible: nothing to check here
        lib.icarion_gpu_destroy(h)
        pytest.skip("device present")
    assert rc == _abi.ERR_NO_DEVICE
    assert "device" in gpu.last_error(lib)


STRUCTS = {
    "icarion_domain_desc": (_abi.DomainDesc, ["instrument", "origin_m", "length_m", "rot_g2l",
                                              "dc_axial_V", "field_array_scale", "B_T",
                                              "temperature_K", "gas_velocity_m_s",
                                              "detector_z_m"]),
    "icarion_physics_desc": (_abi.PhysicsDesc, ["integrator", "precision", "rng_seed"]),
    "icarion_species_desc": (_abi.SpeciesDesc, ["sigma_hss_m2", "K0_cm2_Vs", "atom_offset",
                                                "atom_count"]),
    "icarion_ion_soa": (_abi.IonSoA, ["n", "x", "species", "active", "born", "n_collisions",
                                      "time_s", "dt_s"]),
    "icarion_step_stats": (_abi.StepStats, ["ion_steps", "alive", "rejected", "t_end_s",
                                            "device_ms"]),
    "icarion_gpu_desc": (_abi.GpuDesc, ["device", "block_size", "steps_per_launch"]),
    "icarion_rk45_desc": (_abi.Rk45Desc, ["atol", "absolute_min_step_s", "accept_at_dt_min"]),
    "icarion_train_desc": (_abi.TrainDesc, ["epochs", "lr", "eps", "seed"]),
    "icarion_cloud_desc": (_abi.CloudDesc, ["domain", "physics", "species", "n_species", "ions",
                                            "t0_s", "n_steps", "step_offset", "waveforms", "mixture",
                                            "n_components", "atoms_xyzr", "field_E", "field_n",
                                            "field_origin_m", "field_spacing_m", "pressure"]),
    "icarion_batch_tuning": (_abi.BatchTuning, ["block_size", "groups_per_cu"]),
    "icarion_cloud_stats": (_abi.CloudStats, ["ion_steps", "arrived", "t_end_s", "done_ms"]),
    "icarion_field_term_desc": (_abi.FieldTermDesc, ["E", "dtype", "scale_kind", "constant_scale",
                                                     "phase_rad"]),
    "icarion_gas_component": (_abi.GasComponent, ["density_m3", "mass_kg", "participates"]),
    "icarion_waveform_desc": (_abi.WaveformDesc, ["type", "clamp", "p", "times_s", "values",
                                                  "n_points"]),
}


def test_ctypes_layout_matches_c_compiler(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for cname, (_, fields) in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for cname, (cls, fields) in STRUCTS.items():
        assert int(got[cname]) == C.sizeof(cls), cname
        for f in fields:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, f"{cname}.{f}"


def test_cpp_adaptor_compiles_and_links(tmp_path):
    """include/icarion_gpu.hpp and include/icarion_config.hpp compile as C++17
    and link against the built library; without a device the never-throwing
    factory yields an empty stepper whose step_batch() reports "not handled"."""
    lib = gpu.load()
    so = lib._name
    src = tmp_path / "adaptor.cpp"
    src.write_text(
        '#include "icarion_gpu.hpp"\n'
        '#include "icarion_config.hpp"\n'
        "#include <cstdio>\n"
        "int main() {\n"
        "    icarion::BatchStepper s = icarion::BatchStepper::try_create(icarion_gpu_desc{0, 0, 0, 0});\n"
        "    if (s.handle()) {  // the config helpers feed the adaptor's configure calls\n"
        "        icarion::config::Domain d;\n"
        "        d.geometry.length_m = 0.02; d.geometry.radius_m = 0.01; d.fields.EN_Td = 10.0;\n"
        "        d.env.derive();\n"
        "        icarion::config::Physics p; p.collision_model = \"HSS\";\n"
        "        icarion::config::SpeciesRecord r; r.id = \"X\"; r.CCS_A2 = 100.0;\n"
        "        const auto t = icarion::config::species_table({r}, d, p);\n"
        "        s.configure(icarion::config::to_desc(d), icarion::config::to_physics({}, p), t.species,\n"
        "                    t.atoms_xyzr);\n"
        "    }\n"
        "    icarion_ion_soa soa{};\n"
        "    const bool handled = s.step_batch(soa, 0.0, 1e-9, 0);\n"
        "    int code = -1;\n"
        "    try { icarion::CloudBatch b(icarion_gpu_desc{0, 0, 0, 0}); code = 0; }\n"
        "    catch (const icarion::Error& e) { code = e.code(); }\n"
        '    std::printf("%d %d %d %d\\n", icarion_gpu_abi_version(), icarion_gpu_device_count(),\n'
        "                handled ? 1 : 0, code);\n"
        "    return 0;\n"
        "}\n")
    exe = tmp_path / "adaptor"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), so, f"-Wl,-rpath,{os.path.dirname(so)}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    version, devices, handled, code = map(int, out)
    assert version == _abi.ABI_VERSION
    if devices > 0:
        pytest.skip("device present")
    assert handled == 0
    assert code == _abi.ERR_NO_DEVICE
