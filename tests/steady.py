"""The reference's steady cycle (integration_cyclezone.f90:80-167 for time_steady) driven on the
oracle: per iteration calc_zonetimestep with the CFL ramp (calc_zonetimestep.f90:64-73), the RK
step with the local or global time step, update_field's residual (update_field.f90:42-60) and
check_end_cycle (check_end_cycle.f90:32-33).  The checker of ty_steady_cycle."""
from __future__ import annotations

import math

import numpy as np


def oracle_cycle(o, maxit, cfl, cfl_max, maxres, local_dt=True, ref=None, residu_ref=None, wlog=0.0):
    """returns (history of cur_res, final residu_ref)"""
    ref = o.refcons() if ref is None else ref
    rref = np.finfo(float).tiny if residu_ref is None else residu_ref
    cur = rref  # integration_cyclezone.f90:50
    hist = []
    for _ in range(maxit):
        c = min(cfl * max(1.0, 1.0 - math.log10(cur / rref) - wlog), cfl_max)
        dt = o.timestep(c, np.inf)
        if dt < 0:
            raise RuntimeError("bad state (calc_ns_timestep)")
        if local_dt:
            o.rk_step_local()
        else:
            o.rk_step(dt)
        cur = o.residual_norm(*ref)
        hist.append(cur)
        rref = max(rref, cur)
        if cur / rref <= maxres:
            break
    return np.array(hist), rref
