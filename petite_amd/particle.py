"""Particle container with the reference's interface (src/PETITE/particle.py:50-207).

Shower outputs come back from the device as a ``ParticleBatch`` (see
``petite_amd.batch``); ``ParticleBatch.to_particles()`` builds these objects
when the list-of-Particle form of ``generate_shower`` is wanted.
"""
import numpy as np

from .constants import mass_dict

default_ids = {"PID": 11, "ID": 1, "parent_PID": 22, "parent_ID": -1, "generation_number": 0,
               "generation_process": "Input", "weight": 1.0, "mass": None, "stability": "stable",
               "production_time": 0.0, "decay_time": 0.0, "interaction_time": 0.0}


class Particle:
    """Container for particle information as it is propagated through the target."""

    def __init__(self, p0, r0=np.array([0, 0, 0]), id_dictionary=None):
        if id_dictionary is None:
            id_dictionary = {}
        self.set_ids(id_dictionary)
        self.set_mass(self.get_ids()["mass"])
        if type(p0) is list:
            p0 = np.array(p0)
        elif type(p0) is int or type(p0) is float:
            if self.get_ids()["mass"] is None:
                self._mass = mass_dict[self.get_ids()["PID"]]
            p0 = np.array([p0, 0, 0, np.sqrt(p0 ** 2 - self._mass ** 2)])
        self.set_p0(p0)
        if type(r0) is list:
            r0 = np.array(r0)
        self.set_r0(r0)
        self.set_ended(False)
        self.set_pf(p0)
        self.set_rf(r0)

    def set_ids(self, value):
        self._IDs = {k: value.get(k, default_ids[k]) for k in default_ids}

    def get_ids(self):
        return self._IDs

    def update_ids(self, key, value):
        self._IDs[key] = value

    def get_pid(self):
        return self._IDs["PID"]

    def get_parent_pid(self):
        return self._IDs["parent_PID"]

    def get_weight(self):
        return self._IDs["weight"]

    def set_mass(self, value):
        self._mass = value

    def set_p0(self, value):
        self._p0 = value
        if self._mass is None:  # particle.py:125-131: infer the mass from p0
            inv = round(np.sqrt(round(value[0] ** 2 - value[1] ** 2 - value[2] ** 2 - value[3] ** 2, 12)), 6)
            self._mass = inv
            self._IDs["mass"] = inv

    def get_p0(self):
        return self._p0

    def set_pf(self, value):
        self._pf = value

    def get_pf(self):
        return self._pf

    def get_angle_to_z_0(self):
        E0, px0, py0, pz0 = self.get_p0()
        return np.arccos(pz0 / np.sqrt(px0 ** 2 + py0 ** 2 + pz0 ** 2))

    def lose_energy(self, value):
        E0, px0, py0, pz0 = self.get_pf()
        p30 = np.linalg.norm([px0, py0, pz0])
        E_updated = E0 - value
        m = self.get_ids()["mass"]
        if E_updated <= m:
            E_updated = m
        p3f = np.sqrt(E_updated ** 2 - m ** 2)
        if p3f > 0.0:
            self.set_pf([E_updated, px0 / p30 * p3f, py0 / p30 * p3f, pz0 / p30 * p3f])
        elif p3f == 0.0:
            self.set_pf([m, 0.0, 0.0, 0.0])

    def set_r0(self, value):
        self._r0 = value

    def get_r0(self):
        return self._r0

    def set_rf(self, value):
        self._rf = value

    def get_rf(self):
        return self._rf

    def set_ended(self, value):
        if value is not True and value is not False:
            raise ValueError("Ended property must be a boolean.")
        self._Ended = value

    def get_ended(self):
        return self._Ended

    def copy(self):
        return Particle(self.get_p0(), self.get_r0(), self.get_ids())

    def rotation_matrix(self):
        E0, px0, py0, pz0 = self.get_pf()
        ThZ = np.arccos(pz0 / np.sqrt(px0 ** 2 + py0 ** 2 + pz0 ** 2))
        PhiZ = np.arctan2(py0, px0)
        return [[np.cos(ThZ) * np.cos(PhiZ), -np.sin(PhiZ), np.sin(ThZ) * np.cos(PhiZ)],
                [np.cos(ThZ) * np.sin(PhiZ), np.cos(PhiZ), np.sin(ThZ) * np.sin(PhiZ)],
                [-np.sin(ThZ), 0, np.cos(ThZ)]]

    def __repr__(self):
        return f"Particle(PID={self._IDs['PID']}, p0={np.asarray(self._p0).tolist()}, gen={self._IDs['generation_number']})"
