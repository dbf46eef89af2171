"""Writes cpml_reference_preset.json: the CPML grading of the reference preset (P = 10, m = 3,
kappa_max = 1, alpha_max = 0.08, alpha_m = 1, sigma_max = 0.8 (m + 1) / (eta0 dx), dx = 0.01 m,
dt = 0.99 dx / (c sqrt 3)), evaluated in double precision from the published convolutional-PML
closed form (Roden & Gedney 2000; Taflove & Hagness 3rd ed. sec. 7.9):

    b = exp(-(sigma / kappa + alpha) dt / eps0),   c = sigma (b - 1) / (kappa (sigma + kappa alpha))

with E points at depths (P - 1 - q) / (P - 1), q = 0..P-1, and H points half a cell further in.
Run: python tests/golden/make_cpml_golden.py
"""
import json
import math
import os

C0 = 299792458.0
PI = 3.1415926535897
MU0 = 4e-7 * PI
EPS0 = 1.0 / (C0 * C0 * MU0)
ETA0 = math.sqrt(MU0 / EPS0)


def grading(P=10, m=3.0, kappa_max=1.0, alpha_max=0.08, alpha_m=1.0, sig_factor=0.8, dx=0.01, dt=None):
    dt = 0.99 * dx / (math.sqrt(3.0) * C0) if dt is None else dt
    smax = sig_factor * (m + 1.0) / (ETA0 * dx)
    out = {"E": [], "H": []}
    for kind, npts, shift in (("E", P, 0.0), ("H", P - 1, 0.5)):
        for q in range(npts):
            depth = (P - 1 - q - shift) / (P - 1.0)
            sig = smax * depth ** m
            kap = 1.0 + (kappa_max - 1.0) * depth ** m
            alp = alpha_max * ((q + shift) / (P - 1.0)) ** alpha_m
            b = math.exp(-(sig / kap + alp) * dt / EPS0)
            c = sig * (b - 1.0) / (kap * (sig + kap * alp)) if sig + kap * alp != 0 else 0.0
            out[kind].append({"q": q, "sigma": sig, "kappa": kap, "alpha": alp, "b": b, "c": c})
    return {"params": {"P": P, "m": m, "kappa_max": kappa_max, "alpha_max": alpha_max, "alpha_m": alpha_m,
                       "sig_factor": sig_factor, "dx": dx, "dt": dt}, **out}


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "cpml_reference_preset.json"), "w") as f:
        json.dump(grading(), f, indent=1)
