"""Per-step outputs of the time loop in the reference's own formats.

- The result file (.rsl) that VF_OR1INI opens and VF_OR1TRN appends to
  (Source code/vf_or1ini.f, vf_or1trn.f:105-200), and that VF_IR1INP reads
  back as a restart (.res, Source code/vf_ir1inp.f).  It is a Fortran
  unformatted sequential file: every WRITE statement is one record, framed by
  its byte length as a 4-byte integer before and after, little endian, with
  the listed items packed without padding.
- The two list-file lines of each step (vf_a1main.f:955-961, FORMAT 9540 and
  9550).

Only the single-rank layout is written (the reference adds domain records
when NPROCS > 1), without k-epsilon, temperature or scalar fields (the step
does not carry them) and with the porosity static (IPRNT <= 1).
"""
import struct
from typing import BinaryIO, Dict, Iterator, Optional

import numpy as np

# the file version the reference writes (IVR001, IVR002 of VF_A0PRM.h)
VERSION = (0, 1)
DROP = ("DROPTX", "DROPTY", "DROPTZ", "DROPUU", "DROPVV", "DROPWW")


# ---------------------------------------------------------------- records
def _put(f: BinaryIO, payload: bytes):
    n = struct.pack("<i", len(payload))
    f.write(n)
    f.write(payload)
    f.write(n)


def _get(f: BinaryIO) -> Optional[bytes]:
    head = f.read(4)
    if not head:
        return None
    if len(head) < 4:
        raise ValueError("truncated record marker")
    (n,) = struct.unpack("<i", head)
    payload = f.read(n)
    tail = f.read(4)
    if len(payload) != n or tail != head:
        raise ValueError("corrupt record (length markers disagree)")
    return payload


def _f64(a) -> bytes:
    return np.asarray(a, dtype="<f8").ravel(order="F").tobytes()


def _i32(a) -> bytes:
    return np.asarray(a, dtype="<i4").ravel(order="F").tobytes()


# ---------------------------------------------------------------- writing
class Writer:
    """An .rsl file: the header at open, one record group per write_step."""

    def __init__(self, path, shape, numb, ggv, idrop=1):
        self.shape = tuple(int(x) for x in shape)
        self.numb = int(numb)
        self.idrop = int(idrop)
        self.f = open(path, "wb")
        _put(self.f, _i32(VERSION))
        _put(self.f, _i32(self.shape + (self.numb,)))
        # LN,LV,LP,LF,LK,LT,LS,LG,LD,L1 (vf_or1ini.f): flags, velocity,
        # pressure and F always; droplets when the model is on
        _put(self.f, _i32((1, 1, 1, 1, 0, 0, 0, 0, 1 if self.idrop >= 1 else 0, 0)))
        _put(self.f, _f64(np.asarray(ggv).reshape(self.shape, order="F")))

    def write_step(self, nnow, tnow, dtnow, info, state, tdoor=None):
        """vf_or1trn.f:105-200.  info: fsum, fcut, cgbnrm, cgxnrm, icgitr;
        state: NF, UU..FF and BCU..BCF; tdoor: TBUB and DROPTX..DROPWW."""
        f = self.f
        _put(f, struct.pack("<idd", int(nnow), float(tnow), float(dtnow)))
        _put(f, struct.pack("<ddddi", info["fsum"], info["fcut"], info["cgbnrm"], info["cgxnrm"],
                            int(info["icgitr"])))
        _put(f, _i32(state["NF"]))
        for fld, bc in (("UU", "BCU"), ("VV", "BCV"), ("WW", "BCW"), ("PP", "BCP"), ("FF", "BCF")):
            _put(f, _f64(state[fld]))
            _put(f, _f64(np.asarray(state[bc])[: self.numb]))
        numk = self.shape[2]
        if tdoor is None:
            tbub = np.zeros(numk)
        else:
            tbub = tdoor["TBUB"]
        _put(f, _f64(tbub))
        if self.idrop >= 1:
            tx = np.asarray(tdoor["DROPTX"]) if tdoor is not None else None
            if tx is None:
                _put(f, struct.pack("<i", 0))
                return
            inner = tx[1:-1, 1:-1, 1:-1]
            # droplets in the reference's loop order (K, then J, then I)
            idx = np.argwhere(np.transpose(inner >= 0.0, (2, 1, 0)))
            _put(f, struct.pack("<i", len(idx)))
            for k, j, i in idx:
                i1, j1, k1 = int(i) + 2, int(j) + 2, int(k) + 2
                vals = [float(np.asarray(tdoor[n])[i1 - 1, j1 - 1, k1 - 1]) for n in DROP]
                _put(f, struct.pack("<iii6d", i1, j1, k1, *vals))

    def close(self):
        if self.f:
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------- reading
def read(path) -> Iterator[Dict]:
    """The header (as step 'header') then every step group of an .rsl file."""
    with open(path, "rb") as f:
        iv = np.frombuffer(_get(f), "<i4")
        dims = np.frombuffer(_get(f), "<i4")
        numi, numj, numk, numb = (int(x) for x in dims)
        shape = (numi, numj, numk)
        flags = np.frombuffer(_get(f), "<i4")
        ln, lv, lp, lf, lk, lt, ls, lg, ld, l1 = (int(x) for x in flags)
        if lk or lt or ls:
            raise ValueError("k-epsilon, temperature or scalar records are not read")
        hdr = {"version": tuple(int(x) for x in iv), "shape": shape, "numb": numb, "flags": flags}
        if lg == 0:
            hdr["GGV"] = np.frombuffer(_get(f), "<f8").reshape(shape, order="F").copy()
        yield hdr
        while True:
            rec = _get(f)
            if rec is None:
                return
            nnow, tnow, dtnow = struct.unpack("<idd", rec)
            fsum, fcut, bn, xn, itr = struct.unpack("<ddddi", _get(f))
            st = {"NNOW": nnow, "TNOW": tnow, "DTNOW": dtnow, "FSUM": fsum, "FCUT": fcut,
                  "CGBNRM": bn, "CGXNRM": xn, "ICGITR": itr}
            st["NF"] = np.frombuffer(_get(f), "<i4").reshape(shape, order="F").copy()
            for fld, bc in (("UU", "BCU"), ("VV", "BCV"), ("WW", "BCW"), ("PP", "BCP"), ("FF", "BCF")):
                st[fld] = np.frombuffer(_get(f), "<f8").reshape(shape, order="F").copy()
                st[bc] = np.frombuffer(_get(f), "<f8").copy()
            if lg:
                st["GGV"] = np.frombuffer(_get(f), "<f8").reshape(shape, order="F").copy()
            st["TBUB"] = np.frombuffer(_get(f), "<f8").copy()
            if ld:
                (nd,) = struct.unpack("<i", _get(f))
                drops = [struct.unpack("<iii6d", _get(f)) for _ in range(nd)]
                st["DROPS"] = drops
            yield st


def restart(path, step):
    """VF_IR1INP (vf_ir1inp.f): the state of output step `step` as arrays for
    cad_upload_state / cad_upload_tdoor, with INDC from NF (VF_CINDC) and the
    droplet fields from the droplet records (cleared elsewhere)."""
    it = read(path)
    next(it)
    for st in it:
        if st["NNOW"] != step:
            continue
        nf = st["NF"]
        indc = np.where((nf == -1) | (nf == 8), -1, 0).astype(np.int32)
        indc[0, :, :] = indc[-1, :, :] = -1
        indc[:, 0, :] = indc[:, -1, :] = -1
        indc[:, :, 0] = indc[:, :, -1] = -1
        state = {k: st[k] for k in ("UU", "VV", "WW", "PP", "FF", "BCU", "BCV", "BCW", "BCP", "BCF")}
        state["NF"] = np.asfortranarray(nf)
        state["INDC"] = np.asfortranarray(indc)
        shape = nf.shape
        td = {"TBUB": st["TBUB"].copy()}
        for n in DROP:
            td[n] = np.full(shape, -1.0e10 if n == "DROPTX" else 0.0, order="F")
        for rec in st.get("DROPS", []):
            i, j, k = rec[:3]
            for n, v in zip(DROP, rec[3:]):
                td[n][i - 1, j - 1, k - 1] = v
        return {"NNOW": st["NNOW"], "TNOW": st["TNOW"], "DTNOW": st["DTNOW"], "state": state,
                "tdoor": td}
    raise KeyError(f"step {step} is not in {path} (VF_IR1INP: INVALID STEP.)")


# ---------------------------------------------------------------- list file
def _e12_5(x: float) -> str:
    """Fortran 1PE12.5: d.ddddd with a 2-digit exponent, or the E dropped
    for a 3-digit one."""
    s = f"{x:.5E}"
    mant, ex = s.split("E")
    e = int(ex)
    if abs(e) <= 99:
        out = f"{mant}E{e:+03d}"
    else:
        out = f"{mant}{e:+04d}"
    return out.rjust(12)


def list_lines(nnow, tnow, dtnow, fsum, fcut, dv, cgbnrm, cgxnrm, icgitr):
    """The list-file lines of one step (vf_a1main.f:956-958, FORMAT 9540/9550)."""
    a = (f" STEP={nnow:6d} : TIME= {_e12_5(tnow)} : DT  = {_e12_5(dtnow)} : FSUM= {_e12_5(fsum)}"
         f" : FCUT= {_e12_5(fcut)} : !VD!= {_e12_5(dv)}")
    b = f" {'':11s} : !B! = {_e12_5(cgbnrm)} : !R! = {_e12_5(cgxnrm)} : ITR = {icgitr:6d}"
    return a, b
