"""Run driver with restart-on-NaN.

The reference restarts a run whose state has turned non-finite: cross1 sets
a NaN flag, the time loop ends, and initnan (trunk/mixing/src/main.f:953)
clears the work arrays and statistics, re-reads the restart file
(getfil) and calls cross1 again.  Re-reading the same file with the same
parameters repeats the same steps, so here a run returns to its last
device checkpoint (loma_snapshot / loma_restore: state, RK histories, time,
dt) and continues with a smaller CFL number, which changes the step sizes
from there on.

advance_with_restart is the host loop a driver or the request loop calls in
place of Run.advance for long runs.
"""
from __future__ import annotations

import dataclasses

from . import capi


@dataclasses.dataclass
class RestartLog:
    steps_done: int = 0
    restarts: int = 0
    cfl: float = 0.0
    events: list = dataclasses.field(default_factory=list)  # (step at failure, restored to, cfl)


def advance_with_restart(run: "capi.Run", nsteps: int, every: int = 50, cfl: float = 0.5,
                         cfl_factor: float = 0.5, max_restarts: int = 3,
                         on_chunk=None) -> RestartLog:
    """Advance `run` by `nsteps` RK3 steps, checkpointing on the device every
    `every` steps.  When a chunk ends with a non-finite state (LOMA_ERR_NAN),
    the run is restored to its last checkpoint and the CFL number is
    multiplied by `cfl_factor`; after `max_restarts` restarts the error is
    raised.  `on_chunk(run, steps_done)` is called after every good chunk
    (e.g. to write a restart file or a history line)."""
    if every < 1:
        raise ValueError("every must be >= 1")
    log = RestartLog(cfl=cfl)
    run.snapshot()
    base = 0  # steps done at the last checkpoint
    while base < nsteps:
        n = min(every, nsteps - base)
        try:
            run.advance(n)
        except capi.LomaError as e:
            if e.code != capi.LOMA_ERR_NAN or log.restarts >= max_restarts:
                raise
            run.restore()
            log.restarts += 1
            log.cfl *= cfl_factor
            run.set_cfl(log.cfl)
            log.events.append((base + n, base, log.cfl))
            continue
        base += n
        log.steps_done = base
        run.snapshot()
        if on_chunk is not None:
            on_chunk(run, base)
    return log
