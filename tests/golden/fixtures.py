"""Loading of the reference-generated golden fixtures (see make_golden.py)."""
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def available(case):
    return (HERE / f"{case}_static.npz").exists()


def steps(case):
    return sorted(int(p.stem.split("_s")[-1]) for p in HERE.glob(f"{case}_s[0-9]*.npz"))


def static(case):
    with np.load(HERE / f"{case}_static.npz") as z:
        return {k: z[k] for k in z.files}


def step(case, s):
    with np.load(HERE / f"{case}_s{s:05d}.npz") as z:
        out = {}
        for k in z.files:
            tag, name = k.split(":", 1)
            out.setdefault(tag, {})[name] = z[k]
        return out


def trajectory(case):
    with np.load(HERE / f"{case}_trajectory.npz") as z:
        return {k: z[k] for k in z.files}


def deck_params(deck_text):
    """Physics / solver parameters stated in a deck (defaults of vf_a2dflt.f otherwise)."""
    p = dict(rho0=1000.0, grz0=9.8, anu0=1.0e-6, cgepsr=1.0e-12, cgpara=0.95, wbub=0.2, idrop=1)
    for line in str(deck_text).splitlines():
        w = line.split("#", 1)[0].split()
        if len(w) >= 3 and w[0] == "OPTION" and w[1] == "SUB-LOOP":
            p["loops"] = int(w[2])
        if len(w) >= 4 and w[0] == "OPTION" and w[1] == "T-DOOR":
            if w[2] == "BUB":
                p["wbub"] = 0.0 if w[3] == "OFF" else float(w[3].replace("D", "E"))
            elif w[2] == "DROP":
                p["idrop"] = 0 if w[3] == "OFF" else 1
        if len(w) >= 3 and w[0] == "MATE":
            v = float(w[2].replace("D", "E"))
            if w[1] == "DENSITY":
                p["rho0"] = v
            elif w[1] == "GRAVITY":
                p["grz0"] = v
            elif w[1] == "K-VISC":
                p["anu0"] = v
        if len(w) >= 4 and w[0] == "COMP" and w[1] == "MTRX":
            if w[2] == "R-ERROR":
                p["cgepsr"] = float(w[3].replace("D", "E"))
            elif w[2] == "M-ILUBCGSTAB":
                p["cgpara"] = float(w[3].replace("D", "E"))
    return p


def time_params(deck_text):
    """TIME AUTO / LIMIT of a deck (vf_iitime.f, defaults vf_a2dflt.f:223-234)."""
    t = dict(idttyp=1, dtinit=1.0e-6, dtsafe=0.1, dtmin=1.0e-20, dtmax=1.0e20)
    for line in str(deck_text).splitlines():
        w = line.split("#", 1)[0].split()
        if len(w) >= 3 and w[0] == "TIME":
            if w[1] == "CONST":
                t["idttyp"] = 0
            elif w[1] == "AUTO":
                t.update(idttyp=1, dtinit=float(w[2]), dtsafe=float(w[3]))
            elif w[1] == "LIMIT":
                t.update(dtmin=float(w[2]), dtmax=float(w[3]))
    return t


def state_arrays(st, rec):
    """Merge static set-up arrays with one phase record into a field dict."""
    keys = ("XX", "YY", "ZZ", "ANU", "CD0", "GGV", "GGX", "GGY", "GGZ", "GLV", "GLX", "GLY",
            "GLZ", "INDX", "INDY", "INDZ", "INDB", "BCVI")
    arrays = {k: st[k] for k in keys}
    for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF", "NF", "INDC",
              "TBUB", "DROPTX", "DROPTY", "DROPTZ", "DROPUU", "DROPVV", "DROPWW"):
        if k in rec:
            arrays[k] = rec[k]
    return arrays


def solver_params(st):
    return dict(cgpara=float(st["CGPARA"][0]), cgepsa=float(st["CGEPSA"][0]),
                cgepsr=float(st["CGEPSR"][0]), cgdiv=float(st["CGDIV"][0]),
                flower=float(st["FLOWER"][0]), fupper=float(st["FUPPER"][0]),
                plower=float(st["PLOWER"][0]), scmvp=float(st["SCMVP"][0]),
                icgtyp=int(st["ICGTYP"][0]), icgmax=int(st["ICGMAX"][0]))
