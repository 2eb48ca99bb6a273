"""Units, masses, materials and process registries.

Values follow src/PETITE/physical_constants.py:11-58, particle.py:4-13 and the
module registries of src/PETITE/shower.py:36-96 and dark_shower.py:49-63 of
the reference (same names, same codes).
"""
TeV = 1e3
GeV = 1.0
MeV = 1e-3
keV = 1e-6
eV = 1e-9
centimeter = 1.0
alpha_em = 1.0 / 137.035999
m_electron = 510.998950 * keV
m_proton = 938.272088 * MeV
m_proton_grams = 1.67262192369e-24
m_neutron = 939.56542052 * MeV
atomic_mass_unit = 931.49410241 * MeV
n_avogadro = 6.0221409e23
hbarc = 0.1973269804e-13 * GeV * centimeter
GeVsqcm2 = hbarc ** 2
cmtom = 0.01
m_muon = 105.6583755 * MeV
m_tau = 1776.93 * MeV
m_pi0 = 134.9768 * MeV
m_pi_pm = 139.57039 * MeV
m_K_pm = 493.677 * MeV
m_eta = 547.862 * MeV
m_eta_prime = 957.78 * MeV
m_omega = 782.65 * MeV

target_information = {
    "graphite": {"Z_T": 6, "A_T": 12, "mT": 11.178, "rho": 2.210},
    "lead": {"Z_T": 82, "A_T": 207, "mT": 207.2, "rho": 11.35},
    "iron": {"Z_T": 26, "A_T": 56, "mT": 55.845, "rho": 8.00},
    "hydrogen": {"Z_T": 1, "A_T": 1, "mT": 1.0, "rho": 1.0},
    "aluminum": {"Z_T": 13, "A_T": 27, "mT": 26.9815385, "rho": 2.699},
    "tungsten": {"Z_T": 74, "A_T": 183.84, "mT": 183.84, "rho": 19.3},
    "molybdenum": {"Z_T": 42, "A_T": 95.95, "mT": 95.95, "rho": 10.2},
}
for _tm in target_information:
    target_information[_tm]["dEdx"] = 2.0 * target_information[_tm]["rho"]  # MeV/cm

mass_dict = {11: m_electron, -11: m_electron, 12: 0.0, -12: 0.0, 22: 0.0, 13: m_muon, -13: m_muon,
             14: 0.0, -14: 0.0, 111: m_pi0, 211: m_pi_pm, -211: m_pi_pm, 321: m_K_pm, -321: m_K_pm,
             221: m_eta, 331: m_eta_prime, 2212: m_proton, 223: m_omega}

process_code = {"Brem": 0, "Ann": 1, "PairProd": 2, "Comp": 3, "Moller": 4, "Bhabha": 5, "MuonE": 6,
                "MuonBrem": 7}
dark_process_code = {"DarkBrem": 8, "DarkAnn": 9, "DarkComp": 10, "DarkMuonBrem": 11}
all_process_code = dict(process_code, **dark_process_code)
process_name = {v: k for k, v in all_process_code.items()}
dimensionalities = {"PairProd": 4, "Comp": 1, "Brem": 4, "MuonBrem": 4, "Ann": 1, "Moller": 1, "Bhabha": 1,
                    "MuonE": 1}
dimensionalities_dark = {"DarkComp": 1, "DarkBrem": 3, "DarkAnn": 1, "DarkMuonBrem": 3}
process_PIDS = {"PairProd": [-11, 11], "Brem": [0, 22], "MuonBrem": [0, 22], "Comp": [11, 22], "Ann": [22, 22],
                "Moller": [0, 11], "Bhabha": [0, 11], "MuonE": [0, 11]}
dark_process_codes = ["DarkBrem", "DarkAnn", "DarkComp", "TwoBody_BSMDecay", "DarkMuonBrem"]
PID_DARK_VECTOR = 4900022
# TwoBody_BSMDecay: dark process code of its output entries, and BR(M -> gamma gamma)
# of the mesons it applies to (particle.py:30-48)
TWO_BODY_CODE = 12
meson_twobody_branchingratios = {111: 0.98823, 221: 0.3936, 331: 0.02307}

# meson decays (particle.py:15-47, physical_constants.py:31-33): proper decay
# lengths c tau [m], interaction lengths of the long-lived mesons [m], and the
# decay channels as [branching ratio, product PIDs]
c_tau_pi0, c_tau_pi_pm, c_tau_K_pm = 25.3e-9, 7.8045, 3.711
int_length_dict = {211: 1.796e-1, -211: 1.796e-1, 321: 2.2875e-1, -321: 2.2875e-1}
decay_length_dict = {211: c_tau_pi_pm, -211: c_tau_pi_pm, 321: c_tau_K_pm, -321: c_tau_K_pm, 111: c_tau_pi0}
meson_decay_dict = {111: [[0.98823, [22, 22]]],
                    221: [[0.3936, [22, 22]], [0.3257, [111, 111, 111]]],
                    331: [[0.02307, [22, 22]], [0.224, [111, 111, 221]], [0.00250, [111, 111, 111]]],
                    223: [[0.0828, [22, 111]]],
                    211: [[0.9998, [-13, 14]]],
                    -211: [[0.9998, [13, -14]]],
                    321: [[0.6356, [-13, 14]]],
                    -321: [[0.6356, [13, -14]]]}
