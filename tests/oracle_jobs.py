"""Oracle runs for the GPU workload tests, one complex per call, so that a
spawned process pool can run a sample of a batch in parallel.  Test
infrastructure: the oracle is the checker here."""

import os


def pool_workers():
    """Worker count bounded by this process's CPU share (at most 16)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def oracle_complex(c):
    """CMS of both sides, SC with per-side statistics, and every dot of one
    complex, from the NumPy oracle."""
    from oracle.cms_oracle import OracleCalculator
    o = OracleCalculator()
    o.add_binder_and_target(c["binder_xyz"], c["binder_radii"], c["target_xyz"], c["target_radii"])
    target = o.contact_ms(True)
    binder = o.contact_ms(False)
    sc = o.shape_complementarity()
    r = o.sc_results
    sides = []
    for s in r["sides"]:
        sides.append(None if s is None else {k: s[k] for k in ("d_mean", "d_median", "s_mean", "s_median")})
    dots = [dict(coor=d.coor, nml=d.nml, area=d.area, buried=d.buried, type=d.type, atom=d.atom) for d in o.dots]
    return dict(target=target, binder=binder, sc=sc, n_trimmed=list(r["n_trimmed"]),
                trimmed_area=list(r["trimmed_area"]), sides=sides, dots=dots, counts=dict(o.counts))


def oracle_many(complexes, workers=None):
    """oracle_complex over a list, in a spawned process pool."""
    import concurrent.futures
    import multiprocessing
    workers = workers or pool_workers()
    if workers <= 1 or len(complexes) < 2:
        return [oracle_complex(c) for c in complexes]
    ctx = multiprocessing.get_context("spawn")
    with concurrent.futures.ProcessPoolExecutor(max_workers=min(workers, len(complexes)), mp_context=ctx) as ex:
        return list(ex.map(oracle_complex, complexes))
