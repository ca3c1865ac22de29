"""double_multigrid (field-hmc.h:1011-1223): constrained fine HMC alternated
with coarse HMC on the mag-blocked lattice.

The fine chain runs with link classes (arg.mag), so its trajectories keep
every product of mag fine links along a coarse link fixed. The coarse links
are initialised as those products; every outer iteration runs
`fine_per_coarse` fine trajectories (10 in the reference), one coarse
trajectory, then update_constrain, which rewrites the first fine link of
each coarse link so the products equal the new coarse links. Both chains
stay on the device; the reference's summary-file output is returned as
rows instead.
"""
from dataclasses import replace

from . import F64, Chain, avg_plaquette, check_constrained_plaquette, coarsen_links, update_constrain


def double_multigrid(ctx, U_fine, dims, arg, arg_coarse, num_trajectory, seed=1234, precision=F64,
                     fine_per_coarse=10):
    """Runs `num_trajectory` outer iterations from the fine links U_fine
    ((V, 4, 18) host array). Returns (fine chain, coarse chain, rows) with a
    row per iteration: (iteration, constrained plaquette, coarse plaquette,
    fine results, coarse result)."""
    mag = arg.mag
    if mag < 2:
        raise ValueError("double_multigrid needs arg.mag >= 2")
    cdims = tuple(d // mag for d in dims)
    fine = Chain(ctx, dims, arg, seed=seed, precision=precision)
    fine.set_links(U_fine)
    coarse = Chain(ctx, cdims, replace(arg_coarse, mag=0), seed=seed + 1, precision=precision)
    coarsen_links(fine.links(), mag, coarse.links())
    rows = []
    for i in range(num_trajectory):
        rf = fine.run(fine_per_coarse)
        rc = coarse.run(1)
        update_constrain(fine.links(), coarse.links(), mag)
        rows.append((i, check_constrained_plaquette(fine.links(), mag), avg_plaquette(coarse.links()), rf, rc[0]))
    return fine, coarse, rows


__all__ = ["double_multigrid"]
