"""Helpers shared by the tests that drive gamesplat_amd/gs_headless."""
from pathlib import Path

import numpy as np

HEADLESS = Path(__file__).resolve().parents[1] / "gamesplat_amd" / "gs_headless"
FOCAL_OVER_HEIGHT = np.float32(0.8660254)  # the driver's camera


def headless_position(seed, frame, radius=15.0, orbit_deg=0.0):
    """The headless driver's camera position: splitmix64 -> Box-Muller
    direction, rotated about z by orbit_deg per frame (double, then float)."""
    import math

    mask = (1 << 64) - 1
    state = [seed & mask]

    def nxt():
        state[0] = (state[0] + 0x9E3779B97F4A7C15) & mask
        z = state[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        return ((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0)

    d = [0.0, 0.0, 0.0]
    for k in (0, 2):
        u1 = 1.0 - nxt()
        u2 = nxt()
        rr = math.sqrt(-2.0 * math.log(u1))
        d[k] = rr * math.cos(6.283185307179586 * u2)
        if k + 1 < 3:
            d[k + 1] = rr * math.sin(6.283185307179586 * u2)
    dn = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    a = orbit_deg * math.pi / 180.0 * frame
    c, s = math.cos(a), math.sin(a)
    return np.array([(c * d[0] - s * d[1]) / dn * radius, (s * d[0] + c * d[1]) / dn * radius,
                     d[2] / dn * radius], np.float32)
