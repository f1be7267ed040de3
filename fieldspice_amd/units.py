"""SI constants and unit multipliers used by the field solve.

The permittivity of free space is derived exactly as fieldspice derives it
(``fieldspice/units.py:20-26``): ``mu0 = 4e-7*pi*1.0000000000552`` and
``eps0 = 1/(mu0*c0**2)``, which gives 8.854187817131639e-12 F/m.  That value
differs from the CODATA-2018 figure by about 5e-10 relative, which is larger
than the 1e-9 parity tolerance of the capacitance path, so it is reproduced
bit for bit rather than typed in from a table.
"""

from __future__ import annotations

import math

c0 = 299_792_458.0
mu0 = 4e-7 * math.pi * 1.0000000000552
eps0 = 1.0 / (mu0 * c0 * c0)
eta0 = math.sqrt(mu0 / eps0)
q = 1.602_176_634e-19
kB = 1.380_649e-23
T_ROOM = 300.0


def thermal_voltage(T: float = T_ROOM) -> float:
    """kT/q [V]."""
    return kB * T / q


# length
m = 1.0
cm = 1e-2
mm = 1e-3
um = 1e-6
nm = 1e-9
# time / frequency
s = 1.0
ms = 1e-3
us = 1e-6
ns = 1e-9
ps = 1e-12
Hz = 1.0
kHz = 1e3
MHz = 1e6
GHz = 1e9
# electrical
V = 1.0
A = 1.0
ohm = 1.0
S = 1.0
F = 1.0
pF = 1e-12
fF = 1e-15
aF = 1e-18
W = 1.0

__all__ = ["c0", "mu0", "eps0", "eta0", "q", "kB", "T_ROOM", "thermal_voltage",
           "m", "cm", "mm", "um", "nm", "s", "ms", "us", "ns", "ps",
           "Hz", "kHz", "MHz", "GHz", "V", "A", "ohm", "S", "F", "pF", "fF",
           "aF", "W"]
