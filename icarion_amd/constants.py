"""Physical constants and gas tables of the reference (src/utils/constants.h:13-65)."""
AMU_TO_KG = 1.66053906660e-27
ELEM_CHARGE_C = 1.602176634e-19
BOLTZMANN = 1.380649e-23
LOSCHMIDT = 2.6867811e25
EPSILON_0 = 8.854187817e-12
STP_TEMP = 273.15
STP_PRESSURE = 101325.0

# gas species -> (mass kg, hard-sphere radius m, polarizability m^3)
# (constants.h:38-62; EnvironmentConfig.h:229-270)
GASES = {
    "N2": (28.0134 * AMU_TO_KG, 1.82e-10, 1.74e-30),
    "He": (4.002602 * AMU_TO_KG, 1.3e-10, 0.195e-30),
    "Ar": (39.948 * AMU_TO_KG, 1.70e-10, 1.596e-30),
    "O2": (31.9988 * AMU_TO_KG, 1.73e-10, 1.567e-30),
    "CO2": (44.0095 * AMU_TO_KG, 1.65e-10, 2.612e-30),
    "Ne": (20.1797 * AMU_TO_KG, 1.38e-10, 0.352e-30),
    "H2O": (18.01528 * AMU_TO_KG, 1.58e-10, 1.45e-30),
}


def gas_properties(name):
    """(mass, radius, polarizability) of a single gas, EnvironmentConfig.h:229-278.

    "Air" is the N2/O2/Ar mole-fraction mix; "Default" or "" falls back to He.
    """
    if name == "Air":
        f = (0.78084, 0.20946, 0.00934)
        g = (GASES["N2"], GASES["O2"], GASES["Ar"])
        return tuple(sum(fi * gi[k] for fi, gi in zip(f, g)) for k in range(3))
    if name in ("Default", ""):
        return GASES["He"]
    if name not in GASES:
        raise ValueError(f"unknown gas species {name!r}")
    return GASES[name]
