"""Physical constants and species parameters (src/globals.cpp:13-32).

The reference keeps these as long double; the engine's FP64 path uses their
double roundings, computed here exactly as the device code expects them.
"""
PI = 3.1415926535897932384626
ELE_E = 1.602176487E-19          # elementary charge [C]
GRAVCONST = 9.80665              # [m/s^2]
C_0 = 299792458.                 # [m/s]
HBAR = 1.05457266e-34            # [J s]
MU0 = 4 * PI * 1e-7              # [V s / (A m)]
M_N = 1.674927211E-27 / ELE_E    # neutron mass [eV/c^2]
M_P = 1.672621637E-27 / ELE_E    # proton mass [eV/c^2]
M_E = 9.10938215e-31 / ELE_E     # electron mass [eV/c^2]
MU_N_SI = -0.96623641e-26        # neutron magnetic moment [J/T]
MASS_N_KG = 1.674927211E-27      # neutron mass [kg]
GAMMA_N = -1.83247185e8          # neutron gyromagnetic ratio [1/(T s)]

MAX_TRACK_DEVIATION = 0.001      # include/particle.h:20
REFLECT_TOLERANCE = 1e-8         # include/trianglemesh.h:39

NEUTRON, PROTON, ELECTRON = 0, 1, 2
SPECIES_NAMES = {NEUTRON: "neutron", PROTON: "proton", ELECTRON: "electron"}

# (q, m, mu, gamma) per species: TNeutron/TProton/TElectron constructors
# (src/neutron.cpp:20, src/proton.cpp:14, src/electron.cpp:14)
SPECIES_CONSTANTS = {
    NEUTRON: (0.0, M_N, MU_N_SI, GAMMA_N),
    PROTON: (ELE_E, M_P, 0.0, 0.0),
    ELECTRON: (-ELE_E, M_E, 0.0, 0.0),
}
