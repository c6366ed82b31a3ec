"""Error metrics for comparing two transform results.

rel_l2_slots: relative L2 per (variable, slot) with a scale floor for
derivative slots.  A derivative slot whose exact value is zero (d2/dz2 of a
field linear in z) holds only rounding noise in any FP64 implementation, so
its plain relative error is meaningless; each derivative slot is measured
against max(||y||, sigma), sigma = ||y of the next-lower derivative in the same
direction|| * 2 / L, with L the domain length in that direction (1 for lambda);
for a second derivative the lower norm is at least ||values|| * 2 / L, so a
field constant along an axis floors both derivative slots of that axis.
"""
import numpy as np


def slot_floors(ref, geometry, params):
    """ref: (npts, nslots) of one variable -> floor per slot."""
    return floors_from_norms([float(np.linalg.norm(ref[:, s])) for s in range(ref.shape[1])], geometry, params)


def floors_from_norms(n, geometry, params):
    """The same floors from the L2 norm of every slot of one variable."""
    Lr = 2.0 / (params.iMax - params.iMin)
    Lz = 2.0 / (params.kMax - params.kMin) if params.kMax != params.kMin else 0.0
    # a second derivative is floored at the larger of its first derivative's norm and the
    # values' norm, each times the inverse half-width: a field linear (or constant) along
    # an axis leaves both derivative slots of that axis at rounding level
    if geometry == "RZ":
        return [0.0, n[0] * Lr, max(n[1], n[0] * Lr) * Lr, n[0] * Lz, max(n[3], n[0] * Lz) * Lz]
    f = [0.0, n[0] * Lr, max(n[1], n[0] * Lr) * Lr, n[0], max(n[3], n[0])]
    if geometry in ("RLZ", "SLZ"):
        f += [n[0] * Lz, max(n[5], n[0] * Lz) * Lz]
    return f


def rel_l2_slot(x, ref, floor=0.0):
    den = max(float(np.linalg.norm(ref)), floor)
    return float(np.linalg.norm(x - ref) / den) if den > 0 else 0.0


def rel_l2_slots(x, ref, geometry, params):
    """x, ref: (npts, nslots) of one variable -> list of floored relative L2 per slot."""
    fl = slot_floors(ref, geometry, params)
    return [rel_l2_slot(x[:, s], ref[:, s], fl[s]) for s in range(ref.shape[1])]


def rel_l2_slots_device(x, ref, geometry, params):
    """x, ref: torch tensors (nslots, npts) of one variable on the device -> list of
    floored relative L2 per slot (the norms are reduced on the device)."""
    import torch
    n = [float(torch.linalg.vector_norm(ref[s])) for s in range(ref.shape[0])]
    fl = floors_from_norms(n, geometry, params)
    out = []
    for s in range(ref.shape[0]):
        den = max(n[s], fl[s])
        out.append(float(torch.linalg.vector_norm(x[s] - ref[s])) / den if den > 0 else 0.0)
    return out
