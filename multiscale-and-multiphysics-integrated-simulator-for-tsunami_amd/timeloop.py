"""The time loop of VF_A1MAIN over a device context, without file output.

Source code/vf_a1main.f:919-1145 runs, per iteration: the next step size
(TIME CONST, or VF_CDTCAL from the current fields at :937-944), the end test
NNOW >= NEND or TNOW + DT/2 >= TEND (:946-947), the list-file line of the
state so far, then DTNOW = DT/LOOPS and the sub-loop of VF_V1CAL passes and
VF_F1CAL (:1029-1104).  run() does the same through the C ABI: the step size
and the end test here, one cad_step for the rest.
"""
from dataclasses import dataclass
from typing import Iterator


@dataclass
class StepRecord:
    """One list-file line (vf_a1main.f:955-961) plus the step's timers."""
    nnow: int
    tnow: float
    dtnow: float
    info: dict


def time_control(deck):
    """(IDTTYP, DTCNST, DTINIT, DTSAFE, DTMIN, DTMAX, NEND, TEND) of a deck."""
    return (deck.idttyp, deck.dt_const, deck.dtinit, deck.dtsafe, deck.dtmin, deck.dtmax,
            deck.nend, deck.tend)


def next_dt(ctx, deck, nnow):
    """The step size of step nnow + 1 (vf_a1main.f:937-944)."""
    if deck.idttyp == 0:
        return deck.dt_const
    return ctx.cdtcal(nnow, deck.dtinit, deck.dtsafe, deck.dtmin, deck.dtmax)


def run(ctx, deck, nmax=None, nnow=0) -> Iterator[StepRecord]:
    """Step ctx until the deck's TIME END (or nmax steps); yields each step.

    ctx holds the state after nnow steps (0: the initial state after
    cad_init_state) with its TNOW and, for nnow > 0, the last step's DTNOW.
    """
    tnow = ctx.get_params()["tnow"]
    done = 0
    while True:
        dt = next_dt(ctx, deck, nnow)
        if nnow >= deck.nend or tnow + 0.5 * dt >= deck.tend:
            return
        if nmax is not None and done >= nmax:
            return
        ctx.set_params(dtnow=dt)
        info = ctx.step()
        nnow += 1
        done += 1
        p = ctx.get_params()
        tnow = p["tnow"]
        yield StepRecord(nnow, tnow, p["dtnow"], info)
