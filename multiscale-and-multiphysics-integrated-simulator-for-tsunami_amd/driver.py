"""One CADMAS-SURF/3D run on the GPU from the reference's input deck.

VF_A1MAIN (Source code/vf_a1main.f) for one rank, without MPI and the model
options outside the step: the deck is read (deck.py), the grid and the set-up
arrays built (layout.py), the device context created, and the initial state
set by VF_CINIT's device part or read from a restart file <name>.res
(FILE RES step, vf_a1main.f:796-860).  The time loop (:919-1145) then runs,
per iteration:
  - the next step size (TIME CONST, or VF_CDTCAL on the device);
  - the end test (NNOW >= NEND or TNOW + DT/2 >= TEND);
  - the outputs of the state so far: the list-file lines (VF_CDIV00 for
    !VD!) and, where FILE RSL / GRP / TRN select the step, a record in
    <name>.rsl (VF_OR1TRN), <name>.grp (VF_OG1TRN) and a line of
    <name>.tran (VF_OT1TRN), all by the step / time rule of
    vf_or1trn.f:62-78;
  - the step (cad_step).
With MODEL WAVE-BC ... FUNC MATRIX the table is read from <name>.mtb in the
working directory (VF_IM1INP, vf_im1inp.f:44-46) before the initial state.
"""
from pathlib import Path

from . import capi
from . import grp as grpmod
from . import layout as layoutmod
from . import mtb as mtbmod
from . import rsl as rslmod
from . import timeloop
from . import tran as tranmod


class _RslRule:
    """Which steps VF_OR1TRN / VF_OG1TRN write (vf_or1trn.f:62-78,
    vf_og1trn.f:54-70)."""

    def __init__(self, spec, restart):
        self.spec = spec
        # RRSTRN(4): START on a fresh run (vf_a1main.f:804), 0 after a restart
        self.next_t = 0.0 if restart or spec is None or spec[0] != "TIME" else spec[1]

    def __call__(self, nnow, tnow, dt):
        if self.spec is None:
            return False
        if self.spec[0] == "STEP":
            _, a, b, c = self.spec
            return a <= nnow <= b and (nnow - a) % c == 0
        _, a, b, c = self.spec
        if tnow >= a - 0.5 * dt and tnow <= b + 0.5 * dt:
            w = (tnow + 0.5 * dt) - self.next_t
            if w >= 0.0:
                self.next_t = self.next_t + float(int(w / c) + 1) * c
                return True
        return False


def run(deck_text, name, workdir=".", nmax=None, device=0, log=None, precision=None):
    """Run the deck; writes <name>.list (step lines) and, with FILE RSL,
    <name>.rsl into workdir.  precision = capi.PRECISION_MIXED runs the
    pressure solves in the mixed form.  Returns a summary of the last step."""
    workdir = Path(workdir)
    lay = layoutmod.from_text(deck_text)
    d = lay.deck
    ni, nj, nk = lay.shape
    ctx = capi.Context(ni, nj, nk, lay.numb, device=device)
    try:
        p = lay.params()
        ctx.set_params(**{k: v for k, v in p.items() if k in ctx.params})
        if precision is not None:
            ctx.set_params(precision=precision)
        static = lay.static_arrays()
        ctx.upload_static(static)
        if d.waves:
            # the tables of FUNC MATRIX and MATRIX2: <name>.mtb and <name>.mtb2
            tables = [None, None]
            for i, (fn, ext) in enumerate((("MATRIX", "mtb"), ("MATRIX2", "mtb2"))):
                if any(wv["func"] == fn for wv in d.waves.values()):
                    tables[i] = mtbmod.parse((workdir / f"{name}.{ext}").read_text())
            ctx.set_waves(d.waves, tables[0], d.wvlvl, tables[1])
        ctx.set_opens(d.opens)
        ctx.set_drag_df(d.drag_df)
        ctx.set_damps(d.damps)
        restart = d.res_step is not None
        if restart:
            r = rslmod.restart(workdir / f"{name}.res", d.res_step)
            ctx.upload_state(r["state"])
            ctx.upload_tdoor(r["tdoor"])
            ctx.set_params(tnow=r["TNOW"], dtnow=r["DTNOW"] if r["DTNOW"] > 0.0 else p["dtnow"])
            nnow = r["NNOW"]
            info = {"fsum": 0.0, "fcut": 0.0, "cgbnrm": 0.0, "cgxnrm": 0.0, "icgitr": 0}
        else:
            ctx.upload_state(lay.raw_state())
            info = ctx.init_state(d.wvlvl, d.uini)
            info.update(cgbnrm=0.0, cgxnrm=0.0, icgitr=0)
            ctx.set_params(tnow=0.0)
            nnow = 0
        rule = _RslRule(d.rsl, restart)
        grule = _RslRule(d.grp, restart)
        trule = _RslRule(d.trn, restart)
        writer = gwriter = twriter = None
        if d.rsl is not None:
            writer = rslmod.Writer(workdir / f"{name}.rsl", lay.shape, lay.numb, static["GGV"],
                                   idrop=d.idrop)
        if d.grp is not None:
            gwriter = grpmod.Writer(workdir / f"{name}.grp", static["XX"], static["YY"], static["ZZ"],
                                    static["INDB"], static["GGV"], d.grp_area)
        if d.trn is not None and d.trn_probes:
            probes = [tranmod.parse(w, lay.shape) for w in d.trn_probes]
            twriter = tranmod.Writer(workdir / f"{name}.tran", probes, static, d.wvlvl)
        out = open(workdir / f"{name}.list", "w")
        done = 0
        try:
            while True:
                dt = timeloop.next_dt(ctx, d, nnow)
                par = ctx.get_params()
                tnow, dtnow = par["tnow"], (par["dtnow"] if nnow > 0 else 0.0)
                last = nnow >= d.nend or tnow + 0.5 * dt >= d.tend or (nmax is not None and done >= nmax)
                dv = ctx.cdiv00()
                a, b = rslmod.list_lines(nnow, tnow, dtnow, info["fsum"], info["fcut"], dv,
                                         info["cgbnrm"], info["cgxnrm"], info["icgitr"])
                out.write(a + "\n" + b + "\n")
                st = None
                if twriter is not None and trule(nnow, tnow, dt):
                    st = ctx.download_state()
                    twriter.write_step(tnow, st, {side: ctx.wave_level(side) for side in d.waves})
                if gwriter is not None and grule(nnow, tnow, dt):
                    st = st if st is not None else ctx.download_state()
                    gwriter.write_step(nnow, tnow, st)
                if writer is not None and rule(nnow, tnow, dt):
                    st = st if st is not None else ctx.download_state()
                    td = ctx.download_tdoor() if d.idrop >= 1 or d.wbub >= 1e-20 else None
                    writer.write_step(nnow, tnow, dtnow, info, st, td)
                if log:
                    log(f"step {nnow} t={tnow:.6e} dt={dtnow:.3e} itr={info['icgitr']}")
                if last:
                    break
                ctx.set_params(dtnow=dt)
                info = ctx.step()
                nnow += 1
                done += 1
        finally:
            out.close()
            for w in (writer, gwriter, twriter):
                if w is not None:
                    w.close()
        par = ctx.get_params()
        return {"nnow": nnow, "tnow": par["tnow"], "dtnow": par["dtnow"], "info": info}
    finally:
        ctx.close()
