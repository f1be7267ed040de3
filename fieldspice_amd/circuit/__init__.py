"""Field-circuit coupling: a meshed quasi-static region as a circuit element
(``fieldspice/circuit/coupling.py``).  The region's stamps follow the
reference's device protocol, so its MNA solver can take the GPU region in
place of its own; the netlist and MNA machinery themselves are not part of
this package."""

from .coupling import FieldRegion, presb_fgmres, stamp_g, stamp_i

__all__ = ["FieldRegion", "stamp_g", "stamp_i", "presb_fgmres"]
