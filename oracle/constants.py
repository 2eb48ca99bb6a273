"""Physical constants and materials (oracle restatement).

Follows src/PETITE/physical_constants.py:11-58 and the particle mass table
src/PETITE/particle.py:4-13 of the reference.
"""
import numpy as np

TeV = 1e3
GeV = 1.0
MeV = 1e-3
keV = 1e-6
alpha_em = 1.0 / 137.035999
m_electron = 510.998950 * keV
m_proton = 938.272088 * MeV
m_proton_grams = 1.67262192369e-24
hbarc = 0.1973269804e-13
GeVsqcm2 = hbarc ** 2
cmtom = 0.01
m_muon = 105.6583755 * MeV
n_avogadro = 6.0221409e23

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

mass_dict = {11: m_electron, -11: m_electron, 22: 0.0, 13: m_muon, -13: m_muon,
             12: 0.0, -12: 0.0, 14: 0.0, -14: 0.0,
             111: 134.9768e-3, 221: 547.862e-3, 331: 957.78e-3}  # physical_constants.py:31-36

# process codes (src/PETITE/shower.py:36) and dark extensions
PROCESS_CODE = {"Brem": 0, "Ann": 1, "PairProd": 2, "Comp": 3, "Moller": 4, "Bhabha": 5,
                "MuonE": 6, "MuonBrem": 7, "DarkBrem": 8, "DarkAnn": 9, "DarkComp": 10,
                "DarkMuonBrem": 11}
PROCESS_NAME = {v: k for k, v in PROCESS_CODE.items()}
DIM = {"PairProd": 4, "Comp": 1, "Brem": 4, "MuonBrem": 4, "Ann": 1, "Moller": 1, "Bhabha": 1,
       "MuonE": 1, "DarkBrem": 3, "DarkMuonBrem": 3, "DarkAnn": 1, "DarkComp": 1}
# outgoing PIDs, 0 -> parent PID (src/PETITE/shower.py:87-96)
PROCESS_PIDS = {"PairProd": (-11, 11), "Brem": (0, 22), "MuonBrem": (0, 22), "Comp": (11, 22),
                "Ann": (22, 22), "Moller": (0, 11), "Bhabha": (0, 11), "MuonE": (0, 11)}
