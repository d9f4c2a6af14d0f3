"""The reference's on-disk formats, written and read on the host.

Text (main.f90:455-496, every 100 steps):
    Capsule###step#######.txt   per master capsule pp: one line '(3E15.7)' per point with
                                yfp, zfp, xfp (in that order), ### = pp (i3.3),
                                ####### = istep (i7.7)
    center_velocity.txt         appended, per master capsule: '(7E16.8)' time, ap(1)%x,
                                ap(1)%y, ap(1)%z, ulmean, vlmean, wlmean (the first capsule's
                                centroid and the mean point velocity Update_Pos computed last,
                                Update_Pos.f90:136-154: the last capsule's)
    timestep.txt                appended: '(3E16.8)' 1.*istep, time, dt (main.f90:486-489)

Binary restart (loadd.f90):
    fldbg#######                loadflds (:11-137): unew, vnew, wnew, uo, vo, wo, pnew, each the
                                global interior (itot, jtot, ktot) in Fortran order (i fastest),
                                as 2DECOMP's decomp_2d_write_var writes it (real(8), native byte
                                order), then (time, 1.*nr, dt) (decomp_2d_write_scalar)
    allpartdata#######          loadpart (:147-668): one direct-access record of np particle
                                records, each x, y, z, integralv, volume, volume0 and then 22
                                arrays of nl values: xfp, yfp, zfp, dxdt, dydt, dzdt, fxl, fyl,
                                fzl, ul, vl, wl, xfpo, yfpo, zfpo, xfpold, yfpold, zfpold, xfp0,
                                yfp0, zfp0, area (type particle_restart, :151-161), real(8)

Ew.d editing follows the Fortran standard (0.ddddddd mantissa, E+ee exponent, the E dropped
for three-digit exponents; rounded to d significant digits).  A negative zero is written
without its sign, as the reference's compiler does by default (Intel Fortran's nominus0).
"""
from __future__ import annotations

import math
import os

import numpy as np

PARTICLE_SCALARS = ("x", "y", "z", "integralv", "volume", "volume0")
PARTICLE_ARRAYS = ("xfp", "yfp", "zfp", "dxdt", "dydt", "dzdt", "fxl", "fyl", "fzl", "ul", "vl",
                   "wl", "xfpo", "yfpo", "zfpo", "xfpold", "yfpold", "zfpold", "xfp0", "yfp0",
                   "zfp0", "area")
FLD_ORDER = ("unew", "vnew", "wnew", "uo", "vo", "wo", "pnew")


def fortran_e(v: float, w: int, d: int) -> str:
    """One value in Fortran Ew.d editing (w >= d + 7), right-justified in w characters."""
    v = float(v)
    if math.isnan(v):
        return "NaN".rjust(w)
    if math.isinf(v):
        s = "Infinity" if v > 0 else "-Infinity"
        return (s if len(s) <= w else ("+Inf" if v > 0 else "-Inf")).rjust(w)
    if v == 0.0:
        return ("0." + "0" * d + "E+00").rjust(w)
    # d significant digits, correctly rounded: m.mmm...E+xx -> 0.mmmm...E(xx + 1)
    s = f"{abs(v):.{d - 1}E}"
    mant, ex = s.split("E")
    digits = mant.replace(".", "")
    e = int(ex) + 1
    if abs(e) <= 99:
        es = f"E{'+' if e >= 0 else '-'}{abs(e):02d}"
    else:
        es = f"{'+' if e >= 0 else '-'}{abs(e):03d}"
    out = ("-" if v < 0 else "") + "0." + digits + es
    if len(out) > w:
        return "*" * w
    return out.rjust(w)


def format_lines(cols, w: int, d: int) -> str:
    """Rows of the given equal-length columns, each value Ew.d, one row per line."""
    cols = [np.asarray(c, dtype=np.float64).reshape(-1) for c in cols]
    n = cols[0].size
    out = []
    for i in range(n):
        out.append("".join(fortran_e(c[i], w, d) for c in cols))
    return "\n".join(out) + ("\n" if n else "")


def capsule_txt_name(pp: int, istep: int) -> str:
    """'Capsule' // i3.3 // 'step' // i7.7 // '.txt' (main.f90:459-461); pp counts from 1."""
    return f"Capsule{pp:03d}step{istep:07d}.txt"


def write_capsule_txt(path: str, x, y, z) -> None:
    """One capsule's points, '(3E15.7)' yfp, zfp, xfp per line (main.f90:462-464)."""
    with open(path, "w") as f:
        f.write(format_lines((y, z, x), 15, 7))


def read_capsule_txt(path: str):
    """(x, y, z) back from a Capsule file (7 significant digits)."""
    a = np.loadtxt(path, ndmin=2, converters=None)
    return a[:, 2].copy(), a[:, 0].copy(), a[:, 1].copy()


def append_center_velocity(path: str, time: float, centroid1, umean_last, lines: int = 1):
    """`lines` copies (one per master capsule) of '(7E16.8)' time, ap(1)%x, ap(1)%y, ap(1)%z,
    ulmean, vlmean, wlmean (main.f90:466-468)."""
    row = format_lines([[time]] + [[c] for c in centroid1] + [[u] for u in umean_last], 16, 8)
    with open(path, "a") as f:
        f.write(row * lines)


def append_timestep(path: str, istep: int, time: float, dt: float) -> None:
    """'(3E16.8)' 1.*istep, time, dt (main.f90:486-489)."""
    with open(path, "a") as f:
        f.write(format_lines(([float(istep)], [time], [dt]), 16, 8))


def fldbg_name(istep: int) -> str:
    return f"fldbg{istep:07d}"


def allpartdata_name(istep: int) -> str:
    return f"allpartdata{istep:07d}"


def write_fldbg(path: str, fields: dict, time: float, nr: int, dt: float) -> None:
    """loadflds(1, nr) (loadd.f90:84-103): fields[name] is the interior (itot, jtot, ktot)
    array of each name in FLD_ORDER (any memory order; written i fastest)."""
    shape = None
    with open(path, "wb") as f:
        for name in FLD_ORDER:
            a = np.asarray(fields[name], dtype=np.float64)
            if shape is None:
                shape = a.shape
            if a.shape != shape or a.ndim != 3:
                raise ValueError(f"{name}: expected a 3-d interior array of shape {shape}")
            f.write(np.asfortranarray(a).tobytes(order="F"))
        f.write(np.array([time, 1.0 * nr, dt], dtype=np.float64).tobytes())


def read_fldbg(path: str, itot: int, jtot: int, ktot: int):
    """loadflds(0, nr) (loadd.f90:22-45): (fields, time, nr, dt); fields[name] Fortran-ordered
    (itot, jtot, ktot) interiors."""
    n = itot * jtot * ktot
    raw = np.fromfile(path, dtype=np.float64)
    if raw.size != 7 * n + 3:
        raise ValueError(f"{path}: {raw.size} values, expected {7 * n + 3}")
    fields = {name: raw[k * n:(k + 1) * n].reshape((itot, jtot, ktot), order="F")
              for k, name in enumerate(FLD_ORDER)}
    info = raw[7 * n:]
    return fields, float(info[0]), int(info[1]), float(info[2])


def particle_record_dtype(nl: int) -> np.dtype:
    """The particle_restart record (loadd.f90:151-161): 6 + 22 nl real(8), no padding."""
    return np.dtype([(s, "<f8") for s in PARTICLE_SCALARS] +
                    [(a, "<f8", (nl,)) for a in PARTICLE_ARRAYS])


def write_allpartdata(path: str, records: np.ndarray) -> None:
    """loadpart(1, nr) (loadd.f90:588-666): records is an array of particle_record_dtype(nl)
    with one entry per particle (np of them); zeros for particles without a master."""
    with open(path, "wb") as f:
        f.write(np.ascontiguousarray(records).tobytes())


def read_allpartdata(path: str, npart: int, nl: int) -> np.ndarray:
    """loadpart(0, nr) (loadd.f90:184-188): the np records."""
    dt = particle_record_dtype(nl)
    size = os.path.getsize(path)
    if size != npart * dt.itemsize:
        raise ValueError(f"{path}: {size} bytes, expected {npart * dt.itemsize}")
    return np.fromfile(path, dtype=dt, count=npart)
