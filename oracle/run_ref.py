"""Runs a reference VAC executable built by ``oracle/build_ref.sh``.

Test infrastructure only: it is used by the golden-fixture script, by parity
tests and by the CPU-baseline leg of ``bench.py``, never by the product path.

A run happens in a throw-away directory holding
  par/DEFAULT   the default namelists VAC reads first (src/vacio.t:121-133),
  in.ini        the initial state, VAC binary format (src/vacio.t:682-755),
and VAC reads the problem namelists from stdin.  Snapshots are written to
run.out in the same binary format and read back with vac_amd.vacfile.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import tempfile
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

# The default values of every namelist entry VAC reads from par/DEFAULT.
# They restate the documented defaults (par/DEFAULT:1-86 in the reference);
# the problem file below overrides what a run changes.
DEFAULT_PAR = """ &testlist
 teststr=' '
 ixtest1=3, ixtest2=3, ixtest3=3
 iwtest=1
 idimtest=1
 /
 &filelist
 typefileini='auto'
 typefileout='auto'
 typefilelog='default'
 snapshotini=0
 snapshotout=0
 fullgridini=F
 fullgridout=F
 /
 &savelist
 /
 &stoplist
 itmax=1000000
 tmax=1.D99
 tmaxexact=F
 dtmin=1.D-99
 residmin=0.D0
 residmax=1.D99
 /
 &methodlist
 typeadvance='twostep'
 typefct='etbfct'
 typetvd='roe'
 typeaxial='slab'
 typelimited='original'
 useprimitive=F
 acmwidth=1
 acmnolim=F
 acmexpo=0.D0
 fourthorder=F
 dimsplit=T
 typedimsplit='default'
 sourcesplit=F
 sourceunsplit=T
 typesourcesplit='default'
 divbfix=T
 divbwave=T
 typepoisson='default'
 divbconstrain=F
 typeconstrain='EfieldCD'
 angmomfix=F
 muscleta=0.D0
 musclomega=0.D0
 smallfix=F
 smallpcoeff=1.D-6
 smallrhocoeff=1.D-6
 compactres=F
 implmrpc=F
 /
 &boundlist
 extraB=F
 /
 &paramlist
 courantpar=0.8D0
 dtcantgrow=F
 dtdiffpar=0.5D0
 dtpar=-1.D0
 slowsteps=0
 implmrpcpar=4
 implrelax=-0.5
 impldiffpar=1.D-12
 impldwlimit=1.D99
 impliternr=20
 impliter=100
 implrestart=20
 implrestart2=5
 typeimpliter='vac_bicg'
 impljacfast=F
 implnewton=F
 implnewmat=F
 implpred=F
 implconserv=F
 implcentered=F
 /
"""


@dataclass
class RefResult:
    snapshots: list          # vac_amd.vacfile.Snapshot per save
    stdout: str
    advance_seconds: float   # "Finish Advance" minus "Start Advance" (src/vac.t:52,115)


def ref_binary(name: str) -> str:
    return os.path.join(REF_DIR, name)


def have_ref(name: str) -> bool:
    p = ref_binary(name)
    return os.path.isfile(p) and os.access(p, os.X_OK)


def _fbool(v: bool) -> str:
    return "T" if v else "F"


def problem_par(nw: int, wnames: str, steps: int, typefull: str = "tvdlf",
                limiter: str = "minmod", typeB=None, courantpar: float = 0.8,
                divbfix: bool = True, divbwave: bool = True,
                every_step: bool = False, tmax: float | None = None,
                tmaxexact: bool = False, extra_method: str = "", extra_param: str = "") -> str:
    ndim_faces = len(typeB) if typeB is not None else 4
    if typeB is None:
        typeB = ["periodic"] * ndim_faces
    # an entry is one type for all variables of a region, or a list per variable
    tb = ",".join(f"{nw}*'{t}'" if isinstance(t, str) else ",".join(f"'{v}'" for v in t)
                  for t in typeB)
    save = " itsave(1,2)=0\n ditsave(2)=1\n" if every_step else " itsave(1,2)=0\n"
    stop = f" itmax={steps}\n"
    if tmax is not None:
        stop += f" tmax={tmax!r}\n tmaxexact={_fbool(tmaxexact)}\n"
    return (
        " &testlist\n /\n"
        " &filelist\n filenameini='in.ini'\n filename='run.log','run.out'\n"
        " typefileini='binary'\n typefileout='binary'\n /\n"
        f" &savelist\n{save} /\n"
        f" &stoplist\n{stop} /\n"
        f" &methodlist\n wnames='{wnames}'\n typefull={nw}*'{typefull}'\n"
        f" typelimiter={nw}*'{limiter}'\n divbfix={_fbool(divbfix)}\n"
        f" divbwave={_fbool(divbwave)}\n{extra_method} /\n"
        f" &boundlist\n typeB={tb}\n /\n"
        f" &paramlist\n courantpar={courantpar!r}\n{extra_param} /\n"
    )


def run_ref(binary: str, ini_bytes: bytes, par: str, timeout: float = 600.0,
            workdir: str | None = None, env_extra: dict | None = None,
            cpu: int | None = None) -> RefResult:
    """Run a reference executable on one initial state and parameter file
    (env_extra: further environment variables for the process; cpu: pin the
    process to that core)."""
    from vac_amd.vacfile import read_snapshots

    exe = binary if os.path.isabs(binary) else ref_binary(binary)
    tmp = workdir or tempfile.mkdtemp(prefix="vacrun.")
    try:
        os.makedirs(os.path.join(tmp, "par"), exist_ok=True)
        with open(os.path.join(tmp, "par", "DEFAULT"), "w") as f:
            f.write(DEFAULT_PAR)
        with open(os.path.join(tmp, "in.ini"), "wb") as f:
            f.write(ini_bytes)
        env = dict(os.environ)
        env.update(env_extra or {})
        proc = subprocess.run(
            ["bash", "-c", f"ulimit -s unlimited; exec {exe}"],
            input=par, capture_output=True, text=True, cwd=tmp,
            timeout=timeout, env=env,
            preexec_fn=(lambda: os.sched_setaffinity(0, {cpu})) if cpu is not None else None)
        out = proc.stdout + proc.stderr
        if proc.returncode != 0 or not os.path.exists(os.path.join(tmp, "run.out")):
            raise RuntimeError(f"reference run failed (rc={proc.returncode}):\n{out[-4000:]}")
        m0 = re.search(r"Start Advance\s+([-0-9.]+)", out)
        m1 = re.search(r"Finish Advance\s+([-0-9.]+)", out)
        adv = float(m1.group(1)) - float(m0.group(1)) if (m0 and m1) else float("nan")
        snaps = read_snapshots(os.path.join(tmp, "run.out"))
        return RefResult(snaps, out, adv)
    finally:
        if workdir is None:
            shutil.rmtree(tmp, ignore_errors=True)


def mhd_varnames(ndim: int) -> tuple:
    if ndim == 2:
        return "rho m1 m2 e b1 b2", "x y rho m1 m2 e b1 b2 gamma eta"
    return ("rho m1 m2 m3 e b1 b2 b3",
            "x y z rho m1 m2 m3 e b1 b2 b3 gamma eta")


def hd_varnames(ndim: int) -> tuple:
    if ndim == 2:
        return "rho m1 m2 e", "x y rho m1 m2 e gamma"
    return "rho m1 m2 m3 e", "x y z rho m1 m2 m3 e gamma"


def make_ini(w: np.ndarray, coords: list, physics: str, gamma: float,
             it: int = 0, t: float = 0.0, name: str = "state",
             special: tuple = ()) -> bytes:
    """special: (name, value) pairs of a user module's equation parameters
    (nspecialpar, e.g. ('grav1', 0.0), ('grav2', -1.0) for vacusr gravity),
    appended to eqpar and to the header's names as VAC's savefile does."""
    from vac_amd.initial import coords_array
    from vac_amd.vacfile import Snapshot, snapshot_bytes

    ndim = len(coords)
    if physics == "mhd":
        _, vn = mhd_varnames(ndim)
        eqpar = np.array([gamma, 0.0])
    else:
        _, vn = hd_varnames(ndim)
        eqpar = np.array([gamma])
    if special:
        eqpar = np.concatenate([eqpar, [float(v) for _, v in special]])
        vn = vn + " " + " ".join(n for n, _ in special)
    header = f"{name}_{physics}{ndim}{ndim}"
    snap = Snapshot(header, it, t, ndim, w.shape[0],
                    tuple(len(c) for c in coords), eqpar, vn,
                    coords_array(coords), w)
    return snapshot_bytes(snap)
