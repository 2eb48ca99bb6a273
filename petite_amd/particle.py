"""Particle container with the reference's interface (src/PETITE/particle.py:50-423).

Shower outputs come back from the device as a ``ParticleBatch`` (see
``petite_amd.batch``); ``ParticleBatch.to_particles()`` builds these objects
when the list-of-Particle form of ``generate_shower`` is wanted.

The decay methods (two_body_decay, three_body_decay, decay_particle and their
helpers, particle.py:187-423) act on one particle with numpy's global random
generator, as the reference's do; the batched meson decays of the dark path
run on the GPU (TwoBody_BSMDecay in petite_dark_reweight).
"""
import numpy as np

from .constants import decay_length_dict, int_length_dict, mass_dict, meson_decay_dict

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

    # ------------------------------------------------------------- decays --
    def boost_matrix(self):
        """Lorentz boost from the particle's rest frame to the lab, along pf
        (particle.py:187-207); the identity for a particle at rest."""
        E0, px0, py0, pz0 = (float(v) for v in self.get_pf())
        p3 = np.array([px0, py0, pz0])
        pmag = np.linalg.norm(p3)
        if pmag == 0.0:
            return np.identity(4)
        gamma = E0 / self._mass
        beta = 1.0 if gamma == 1.0 else np.sqrt(1.0 - 1.0 / gamma ** 2)
        b = beta * p3 / pmag
        L = np.empty((4, 4))
        L[0, 0] = gamma
        L[0, 1:] = L[1:, 0] = gamma * b
        L[1:, 1:] = np.identity(3) + (gamma - 1.0) * np.outer(b, b) / beta ** 2
        return L

    def two_body_decay(self, p1_dict, p2_dict, angular_information="Isotropic"):
        """Decay into two particles of the dicts' masses (or mass_dict of their
        PIDs): back to back in the rest frame, product 2 along the direction
        (sin t sin f, sin t cos f, cos t), boosted to the lab; both start at
        rf (particle.py:209-254).  angular_information: "Isotropic", a pair of
        inverse CDFs (cos t, f) of uniforms, or one inverse CDF for cos t."""
        M = self._mass
        m1, m2 = _product_masses((p1_dict, p2_dict), "two_body_decay")
        E1 = (M * M - m2 * m2 + m1 * m1) / (2.0 * M)
        E2 = (M * M - m1 * m1 + m2 * m2) / (2.0 * M)
        q = np.sqrt(E1 * E1 - m1 * m1)
        ct, phi = _decay_angles(angular_information, 1)
        n = _direction(ct, phi)
        L = self.boost_matrix()
        k1 = L @ np.concatenate([[E1], -q * n])
        k2 = L @ np.concatenate([[E2], q * n])
        return [Particle(k1, self.get_rf(), p1_dict), Particle(k2, self.get_rf(), p2_dict)]

    def dalitz_range(self, m12sq, m1, m2, m3, M):
        """Limits of m23^2 at fixed m12^2 in the decay M -> 1 2 3, from the
        energies of 2 and 3 in the (12) rest frame (particle.py:256-263)."""
        m12 = np.sqrt(m12sq)
        E2s = (m12sq - m1 ** 2 + m2 ** 2) / (2.0 * m12)
        E3s = (M ** 2 - m12sq - m3 ** 2) / (2.0 * m12)
        q2, q3 = np.sqrt(E2s ** 2 - m2 ** 2), np.sqrt(E3s ** 2 - m3 ** 2)
        return (E2s + E3s) ** 2 - (q2 + q3) ** 2, (E2s + E3s) ** 2 - (q2 - q3) ** 2

    def three_body_decay(self, p1_dict, p2_dict, p3_dict, dalitz_information="Flat", angular_information="Isotropic"):
        """Decay into three particles (particle.py:265-357): (m12^2, m23^2)
        uniform in the Dalitz region by rejection from its bounding box, or
        accepted with probability f(m12^2, m23^2, m1, m2, m3, M) / f_max for
        dalitz_information = (f_max, f); the plane of the decay oriented by
        (cos t, f, g): product 1 along -n(t, f), product 2 at its opening angle
        to 1, rotated by g about 1's axis; boosted to the lab."""
        M = self._mass
        m1, m2, m3 = _product_masses((p1_dict, p2_dict, p3_dict), "three_body_decay")
        lo12, hi12 = (m1 + m2) ** 2, (M - m3) ** 2
        lo23, hi23 = (m2 + m3) ** 2, (M - m1) ** 2
        while True:
            s12, s23 = np.random.uniform(lo12, hi12), np.random.uniform(lo23, hi23)
            a, b = self.dalitz_range(s12, m1, m2, m3, M)
            if not (a <= s23 <= b):
                continue
            if isinstance(dalitz_information, str) and dalitz_information == "Flat":
                break
            try:
                f_max, f = dalitz_information
            except (TypeError, ValueError):
                raise ValueError("Dalitz information must be a tuple of (max value, function) for "
                                 "matrix-element-level sampling") from None
            # errors raised by the matrix element itself propagate unchanged
            if f(s12, s23, m1, m2, m3, M) > np.random.uniform(0, f_max):
                break
        E1 = (M * M - s23 + m1 * m1) / (2.0 * M)
        E3 = (M * M - s12 + m3 * m3) / (2.0 * M)
        E2 = M - E1 - E3
        q1, q2 = np.sqrt(E1 * E1 - m1 * m1), np.sqrt(E2 * E2 - m2 * m2)
        c12 = (m1 * m1 + m2 * m2 + 2.0 * E1 * E2 - s12) / (2.0 * q1 * q2)
        s12_ = np.sqrt(1.0 - c12 * c12)
        ct, phi, g = _decay_angles(angular_information, 2)
        st = np.sqrt(1.0 - ct * ct)
        n = _direction(ct, phi)
        e_t = np.array([ct * np.sin(phi), ct * np.cos(phi), -st])  # dn/dt
        e_f = np.array([np.cos(phi), -np.sin(phi), 0.0])  # (dn/df) / sin t
        k1 = -q1 * n
        k2 = q2 * (-c12 * n + s12_ * (-np.cos(g) * e_t + np.sin(g) * e_f))
        k3 = -k1 - k2
        L = self.boost_matrix()
        return [Particle(L @ np.concatenate([[E], k]), self.get_rf(), d)
                for E, k, d in ((E1, k1, p1_dict), (E2, k2, p2_dict), (E3, k3, p3_dict))]

    def _ctau_lab(self, ctau0):
        gamma = self.get_p0()[0] / self.get_ids()["mass"]
        return ctau0 * gamma * np.sqrt(1.0 - 1.0 / gamma ** 2)

    def decay_int_prob(self, x, int_length, ctau0):
        """Total rate R = 1/(c tau gamma beta) + 1/int_length [1/m] of decay
        or interaction, and its density R exp(-R x) at x (particle.py:359-366)."""
        R = 1.0 / self._ctau_lab(ctau0) + 1.0 / int_length
        return R, R * np.exp(-R * x)

    def draw_x_sample(self, int_length, ctau0):
        """Distance to the first decay or interaction: R exp(-R x) on
        [0, 4/R] by rejection against its value at 0 (particle.py:368-376)."""
        R = self.decay_int_prob(0.0, int_length, ctau0)[0]
        while True:
            x = np.random.uniform(0.0, 4.0 / R)
            if np.random.uniform(0.0, 1.0) < np.exp(-R * x):
                return x

    def prob_decay_b_int(self, int_length, ctau0):
        """(1 - exp(-R x)) / (1 + c tau_lab / int_length) at a drawn distance x
        (particle.py:378-389)."""
        x = self.draw_x_sample(int_length, ctau0)
        ctau = self._ctau_lab(ctau0)
        R = 1.0 / ctau + 1.0 / int_length
        return (1.0 - np.exp(-R * x)) / (1.0 + ctau / int_length)

    def decay_particle(self):
        """SM decay of a short- or long-lived meson into one of its two-body
        channels of meson_decay_dict, chosen with probability proportional to
        its branching ratio; products carry weight x (sum of the channels'
        branching ratios) (x the decay-before-interaction probability for a
        long-lived meson, which first moves to its decay point), IDs 2 ID and
        2 ID + 1, generation + 1, process "SMDecay" (particle.py:391-423).
        The particle is ended.  Returns the products (the reference computes
        them and returns None); three-body channels raise ValueError."""
        ids = self.get_ids()
        pid = ids["PID"]
        if pid not in meson_decay_dict:
            raise ValueError("Decay options for particle not specified. Edit dictionary in 'particle.py' to include it")
        options = meson_decay_dict[pid]
        brs = np.array([o[0] for o in options])
        br_sum = float(brs.sum())
        if len(options) == 1:
            decay = options[0][1]
        else:
            decay = options[np.random.choice(len(options), p=brs / br_sum)][1]
        if len(decay) > 2:
            raise ValueError("Three-body (and above) decays not yet implemented")
        products = []
        stability = ids["stability"]
        if stability in ("short-lived", "long-lived"):
            w = [ids["weight"] * br_sum] * 2
            if stability == "long-lived":
                L_int, ctau0 = int_length_dict[pid], decay_length_dict[pid]
                dz = self.draw_x_sample(L_int, ctau0)
                pf = np.asarray(self.get_pf(), dtype=np.float64)
                pn = np.linalg.norm(pf[1:])
                if pn > 0.0:
                    self.set_rf(np.asarray(self.get_rf(), dtype=np.float64) + pf[1:] / pn * dz)
                # one decay-before-interaction draw per product, as the reference makes
                w = [wi * self.prob_decay_b_int(L_int, ctau0) for wi in w]
            dicts = [{"PID": decay[k], "mass": mass_dict[decay[k]], "weight": w[k], "ID": 2 * ids["ID"] + k,
                      "generation_process": "SMDecay", "generation_number": ids["generation_number"] + 1,
                      "production_time": ids["decay_time"]} for k in range(2)]
            products = self.two_body_decay(dicts[0], dicts[1])
        self.set_ended(True)
        return products

    def __repr__(self):
        return f"Particle(PID={self._IDs['PID']}, p0={np.asarray(self._p0).tolist()}, gen={self._IDs['generation_number']})"


def _product_masses(dicts, who):
    """Masses of decay products: a dict's "mass", else mass_dict of its "PID"
    (written into the dict, as the reference does)."""
    out = []
    for d in dicts:
        if "mass" not in d:
            if "PID" not in d:
                raise ValueError(f"Masses must be included in the product dictionaries when calling {who}()")
            if d["PID"] not in mass_dict:
                raise ValueError("PID provided for unspecified particle.")
            d["mass"] = mass_dict[d["PID"]]
        out.append(float(d["mass"]))
    return out


def _decay_angles(angular_information, n_azimuths):
    """cos t and n_azimuths azimuths: uniform ("Isotropic"), inverse CDFs of
    uniforms for all of them (a sequence of 1 + n_azimuths callables), or an
    inverse CDF for cos t only (one callable)."""
    if isinstance(angular_information, str):
        if angular_information != "Isotropic":
            raise ValueError(f"unknown angular_information {angular_information!r}")
        ct = np.random.uniform(-1.0, 1.0)
        az = np.random.uniform(0.0, 2.0 * np.pi, size=n_azimuths)
    elif callable(angular_information):
        ct = angular_information(np.random.uniform(0.0, 1.0))
        az = np.random.uniform(0.0, 2.0 * np.pi, size=n_azimuths)
    else:
        fs = list(angular_information)
        if len(fs) != 1 + n_azimuths:
            raise ValueError(f"angular_information needs {1 + n_azimuths} inverse CDFs, got {len(fs)}")
        u = np.random.uniform(0.0, 1.0, size=1 + n_azimuths)
        ct = fs[0](u[0])
        az = [fs[k](u[k]) for k in range(1, 1 + n_azimuths)]
    return (float(np.squeeze(ct)), *(float(np.squeeze(a)) for a in az))


def _direction(ct, phi):
    st = np.sqrt(1.0 - ct * ct)
    return np.array([st * np.sin(phi), st * np.cos(phi), ct])
