"""Writes the golden fixtures under tests/golden/ from the reference VAC.

Each fixture is one run of a reference VAC executable (built from the
reference's Fortran sources by oracle/build_ref.sh) on an initial state made
by vac_amd.initial:  the initial mesh state, the parameters, the time t after
every step (from the per-step snapshots VAC writes) and the final mesh state.

    python tests/golden/make_golden.py [name ...]  # needs /root/reference

The config-1 run (Orszag-Tang 128^2, TVDLF + minmod, 100 steps) is recorded
as SHA-256 digests of its initial and final states plus its t sequence,
which pins the full-size run without storing 0.8 MB arrays.
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from vac_amd.initial import orszag_tang, random_fourier  # noqa: E402
from run_ref import have_ref, make_ini, problem_par, run_ref  # noqa: E402

from refbuilds import BINARIES, POISSON  # noqa: E402

NAMES = {6: "rho m1 m2 e b1 b2", 4: "rho m1 m2 e", 8: "rho m1 m2 m3 e b1 b2 b3",
         5: "rho m1 m2 m3 e"}

# name: (binary, physics, builder, gamma, steps, typefull, limiter, bc, divbfix)
CASES = {
    "mhd22_rf32x24_tvdlf_minmod": ("vac_mhd22_132", "mhd", lambda g: random_fourier((32, 24), 101, g), 5 / 3, 20, "tvdlf", "minmod", "periodic", True),
    "mhd22_rf32x24_tvd_woodward": ("vac_mhd22_132", "mhd", lambda g: random_fourier((32, 24), 102, g), 5 / 3, 20, "tvd", "woodward", "periodic", True),
    "mhd22_rf24x32_tvdlf_superbee_cont": ("vac_mhd22_132", "mhd", lambda g: random_fourier((24, 32), 103, g), 1.4, 15, "tvdlf", "superbee", "cont", True),
    "mhd22_rf24x24_tvd_minmod_nodivb": ("vac_mhd22_132", "mhd", lambda g: random_fourier((24, 24), 104, g), 5 / 3, 15, "tvd", "minmod", "periodic", False),
    "mhd22_ot32_tvdlf_woodward": ("vac_mhd22_132", "mhd", lambda g: orszag_tang(32, g), 5 / 3, 30, "tvdlf", "woodward", "periodic", True),
    "hd22_rf32x24_tvdlf_minmod": ("vac_hd22_132", "hd", lambda g: random_fourier((32, 24), 105, g, physics="hd"), 1.4, 25, "tvdlf", "minmod", "periodic", True),
    "hd22_rf24x24_tvdlf_woodward_cont": ("vac_hd22_132", "hd", lambda g: random_fourier((24, 24), 106, g, physics="hd"), 5 / 3, 20, "tvdlf", "woodward", "cont", True),
    "mhd33_rf12x10x8_tvdlf_minmod": ("vac_mhd33_36", "mhd", lambda g: random_fourier((12, 10, 8), 107, g), 5 / 3, 10, "tvdlf", "minmod", "periodic", True),
    "mhd33_rf8x10x12_tvd_woodward": ("vac_mhd33_36", "mhd", lambda g: random_fourier((8, 10, 12), 108, g), 5 / 3, 10, "tvd", "woodward", "periodic", True),
    "mhd22_rf32x24_tvdmu_minmod": ("vac_mhd22_132", "mhd", lambda g: random_fourier((32, 24), 109, g), 5 / 3, 20, "tvdmu", "minmod", "periodic", True),
    "mhd22_rf24x28_tvdmu_woodward_cont": ("vac_mhd22_132", "mhd", lambda g: random_fourier((24, 28), 110, g), 1.4, 15, "tvdmu", "woodward", "cont", True),
    "mhd33_rf10x8x12_tvdmu_superbee": ("vac_mhd33_36", "mhd", lambda g: random_fourier((10, 8, 12), 111, g), 5 / 3, 8, "tvdmu", "superbee", "periodic", True),
}


# projection of B before every step (nproc=1, procpar): the reference's own
# Orszag-Tang recipe (TVD + woodward, divbfix=F, divbwave=F, procpar=0,0.01,1),
# a relative stopping rule with the thermal-pressure energy fix, cont faces
# (the Poisson boundary is then 'nul') with the total-pressure fix, a matvec
# limit that makes every solve stop early (nerror(poissonerr_)), and 3D TVD
# with the Powell/Ryu source on (its wold is the state before projection).
# name: (binary, physics, builder, gamma, steps, typefull, limiter, bc, divbfix,
#        divbwave, procpar)
PROJ_CASES = {
    "mhd22_ot32_tvd_woodward_proj": ("vacp_mhd22_132", "mhd", lambda g: orszag_tang(32, g), 5 / 3, 20, "tvd", "woodward", "periodic", False, False, (0.0, 0.01, 1.0)),
    "mhd22_rf28x32_tvdlf_minmod_projrel": ("vacp_mhd22_132", "mhd", lambda g: random_fourier((28, 32), 112, g), 5 / 3, 12, "tvdlf", "minmod", "periodic", True, True, (-0.001, -1.0, 2.0)),
    "mhd22_rf24x30_tvdlf_woodward_cont_proj3": ("vacp_mhd22_132", "mhd", lambda g: random_fourier((24, 30), 113, g), 1.4, 12, "tvdlf", "woodward", "cont", True, True, (0.01, 0.0, 3.0)),
    "mhd22_rf30x26_tvdmu_superbee_projmax3": ("vacp_mhd22_132", "mhd", lambda g: random_fourier((30, 26), 114, g), 5 / 3, 10, "tvdmu", "superbee", "periodic", True, True, (-3.0, 1e-9, 1.0)),
    "mhd33_rf10x12x8_tvd_woodward_proj": ("vacp_mhd33_36", "mhd", lambda g: random_fourier((10, 12, 8), 115, g), 5 / 3, 6, "tvd", "woodward", "periodic", True, True, (0.0, 1e-5, 1.0)),
}


# boundary types beyond cont / periodic, one entry per face (x-low, x-high,
# y-low, ...), each a type for all variables or a list per variable: walls
# (symm with the normal momentum and field asymm), fixed faces, mixes; the
# last two with the projection of B (boundscalar's symm / asymm / grad0)
WALL_X = ["symm", "asymm", "symm", "symm", "asymm", "symm"]
WALL_Y = ["symm", "symm", "asymm", "symm", "symm", "asymm"]
WALL_Y_HD = ["symm", "symm", "asymm", "symm"]
WALL_Z = ["symm", "symm", "symm", "asymm", "symm", "symm", "symm", "asymm"]
# name: (binary, physics, builder, gamma, steps, typefull, limiter, bc spec, divbfix, procpar)
BC_CASES = {
    "mhd22_rf32x24_tvdlf_minmod_bcwallx": ("vac_mhd22_132", "mhd", lambda g: random_fourier((32, 24), 401, g), 5 / 3, 12, "tvdlf", "minmod", [WALL_X, WALL_X, "periodic", "periodic"], True, None),
    "mhd22_rf28x30_tvd_woodward_bcfixedx_wally": ("vac_mhd22_132", "mhd", lambda g: random_fourier((28, 30), 402, g), 5 / 3, 10, "tvd", "woodward", ["fixed", "fixed", WALL_Y, WALL_Y], True, None),
    "hd22_rf30x26_tvdlf_woodward_bcwally": ("vac_hd22_132", "hd", lambda g: random_fourier((30, 26), 403, g, physics="hd"), 1.4, 15, "tvdlf", "woodward", ["cont", "cont", WALL_Y_HD, WALL_Y_HD], True, None),
    "mhd22_rf26x28_tvdmu_superbee_bcmixed": ("vac_mhd22_132", "mhd", lambda g: random_fourier((26, 28), 404, g), 5 / 3, 10, "tvdmu", "superbee", [WALL_X, "fixed", "periodic", "periodic"], True, None),
    "mhd33_rf10x8x12_tvdlf_minmod_bcwallz": ("vac_mhd33_36", "mhd", lambda g: random_fourier((10, 8, 12), 407, g), 5 / 3, 6, "tvdlf", "minmod", ["periodic", "periodic", "cont", "cont", WALL_Z, WALL_Z], True, None),
    "mhd22_rf30x32_tvd_woodward_bcwall_proj": ("vacp_mhd22_132", "mhd", lambda g: random_fourier((30, 32), 405, g), 5 / 3, 10, "tvd", "woodward", [WALL_X, WALL_X, "fixed", "fixed"], False, (0.0, 1e-4, 1.0)),
    "mhd22_rf32x28_tvdlf_minmod_bcsymm_proj": ("vacp_mhd22_132", "mhd", lambda g: random_fourier((32, 28), 406, g), 5 / 3, 10, "tvdlf", "minmod", [["symm"] * 6, ["symm"] * 6, "periodic", "periodic"], True, (0.0, 1e-4, 2.0)),
}


# TVD and TVD-MUSCL with the HD Roe solver (src/vacphys.hdroe.t)
HD_TVD_CASES = {
    "hd22_rf32x28_tvd_woodward": ("vac_hd22_132", "hd", lambda g: random_fourier((32, 28), 501, g, physics="hd"), 1.4, 15, "tvd", "woodward", "periodic", True, None),
    "hd22_rf28x30_tvd_minmod_cont": ("vac_hd22_132", "hd", lambda g: random_fourier((28, 30), 502, g, physics="hd"), 5 / 3, 15, "tvd", "minmod", "cont", True, None),
    "hd22_rf30x26_tvdmu_superbee": ("vac_hd22_132", "hd", lambda g: random_fourier((30, 26), 503, g, physics="hd"), 1.4, 12, "tvdmu", "superbee", "periodic", True, None),
    "hd33_rf10x12x8_tvd_woodward": ("vac_hd33_36", "hd", lambda g: random_fourier((10, 12, 8), 505, g, physics="hd"), 1.4, 8, "tvd", "woodward", "periodic", True, None),
}
BC_CASES.update(HD_TVD_CASES)


# typelimited: the corrector of TVDLF / TVD-MUSCL limits the step's start
# ('previous') or the predictor ('predictor') instead of the sweep's input
# name: (binary, physics, builder, gamma, steps, typefull, limiter, bc, typelimited)
TL_CASES = {
    "mhd22_rf32x28_tvdlf_minmod_previous": ("vac_mhd22_132", "mhd", lambda g: random_fourier((32, 28), 601, g), 5 / 3, 12, "tvdlf", "minmod", "periodic", "previous"),
    "mhd22_rf28x30_tvdlf_woodward_cont_predictor": ("vac_mhd22_132", "mhd", lambda g: random_fourier((28, 30), 602, g), 1.4, 12, "tvdlf", "woodward", "cont", "predictor"),
    "mhd22_rf30x26_tvdmu_superbee_previous": ("vac_mhd22_132", "mhd", lambda g: random_fourier((30, 26), 603, g), 5 / 3, 10, "tvdmu", "superbee", "periodic", "previous"),
    "hd22_rf26x32_tvdlf_minmod_predictor": ("vac_hd22_132", "hd", lambda g: random_fourier((26, 32), 604, g, physics="hd"), 1.4, 15, "tvdlf", "minmod", "periodic", "predictor"),
}


# entropy fixes of the Roe solvers (typeentropy for every wave; entropycoef
# keeps physini's defaults)
# name: (binary, physics, builder, gamma, steps, typefull, limiter, typeentropy)
ENT_CASES = {
    "mhd22_rf32x28_tvd_woodward_harten": ("vac_mhd22_132", "mhd", lambda g: random_fourier((32, 28), 701, g), 5 / 3, 12, "tvd", "woodward", "harten"),
    "mhd22_rf28x30_tvd_minmod_powell": ("vac_mhd22_132", "mhd", lambda g: random_fourier((28, 30), 702, g), 1.4, 12, "tvd", "minmod", "powell"),
    "mhd22_rf30x26_tvd_superbee_yee": ("vac_mhd22_132", "mhd", lambda g: random_fourier((30, 26), 703, g), 5 / 3, 12, "tvd", "superbee", "yee"),
    "mhd22_rf26x32_tvd_woodward_ratio": ("vac_mhd22_132", "mhd", lambda g: random_fourier((26, 32), 704, g), 5 / 3, 12, "tvd", "woodward", "ratio"),
    "mhd22_rf30x30_tvdmu_woodward_harten": ("vac_mhd22_132", "mhd", lambda g: random_fourier((30, 30), 705, g), 5 / 3, 10, "tvdmu", "woodward", "harten"),
    "hd22_rf32x26_tvd_woodward_harten": ("vac_hd22_132", "hd", lambda g: random_fourier((32, 26), 706, g, physics="hd"), 1.4, 15, "tvd", "woodward", "harten"),
    "hd22_rf26x28_tvdmu_superbee_powell": ("vac_hd22_132", "hd", lambda g: random_fourier((26, 28), 707, g, physics="hd"), 1.4, 12, "tvdmu", "superbee", "powell"),
    "mhd33_rf10x12x8_tvd_woodward_powell": ("vac_mhd33_36", "mhd", lambda g: random_fourier((10, 12, 8), 709, g), 5 / 3, 6, "tvd", "woodward", "powell"),
}


# useprimitive=T: upwindLR limits the primitive variables (TVDLF and
# TVD-MUSCL), the MHD Roe average takes dp = pR - pL (TVD and TVD-MUSCL)
# name: (binary, physics, builder, gamma, steps, typefull, limiter, bc)
PRIM_CASES = {
    "mhd22_rf32x28_tvdlf_minmod_prim": ("vac_mhd22_132", "mhd", lambda g: random_fourier((32, 28), 801, g), 5 / 3, 12, "tvdlf", "minmod", "periodic"),
    "mhd22_rf28x30_tvdlf_woodward_cont_prim": ("vac_mhd22_132", "mhd", lambda g: random_fourier((28, 30), 802, g), 1.4, 12, "tvdlf", "woodward", "cont"),
    "mhd22_rf30x26_tvdmu_superbee_prim": ("vac_mhd22_132", "mhd", lambda g: random_fourier((30, 26), 803, g), 5 / 3, 10, "tvdmu", "superbee", "periodic"),
    "mhd22_rf26x32_tvd_woodward_prim": ("vac_mhd22_132", "mhd", lambda g: random_fourier((26, 32), 804, g), 5 / 3, 12, "tvd", "woodward", "periodic"),
    "hd22_rf30x28_tvdlf_minmod_prim": ("vac_hd22_132", "hd", lambda g: random_fourier((30, 28), 805, g, physics="hd"), 1.4, 15, "tvdlf", "minmod", "periodic"),
    "hd22_rf28x26_tvdmu_woodward_cont_prim": ("vac_hd22_132", "hd", lambda g: random_fourier((28, 26), 806, g, physics="hd"), 1.4, 12, "tvdmu", "woodward", "cont"),
    "mhd33_rf10x12x8_tvdlf_minmod_prim": ("vac_mhd33_36", "mhd", lambda g: random_fourier((10, 12, 8), 807, g), 5 / 3, 6, "tvdlf", "minmod", "periodic"),
}


# dimsplit=F: one advect over all directions per step; VAC reads
# typelimited='original' as 'previous' then
# name: (binary, physics, builder, gamma, steps, limiter, bc, typelimited)
UNSPLIT_CASES = {
    "mhd22_rf32x28_tvdlf_minmod_unsplit": ("vac_mhd22_132", "mhd", lambda g: random_fourier((32, 28), 901, g), 5 / 3, 12, "minmod", "periodic", "original"),
    "mhd22_rf28x30_tvdlf_woodward_cont_unsplit_predictor": ("vac_mhd22_132", "mhd", lambda g: random_fourier((28, 30), 902, g), 1.4, 12, "woodward", "cont", "predictor"),
    "hd22_rf30x26_tvdlf_superbee_unsplit": ("vac_hd22_132", "hd", lambda g: random_fourier((30, 26), 903, g, physics="hd"), 1.4, 15, "superbee", "periodic", "original"),
    "mhd33_rf10x12x8_tvdlf_minmod_unsplit": ("vac_mhd33_36", "mhd", lambda g: random_fourier((10, 12, 8), 904, g), 5 / 3, 6, "minmod", "periodic", "original"),
}


# first-order extrapolating boundaries: 'cont1' every fill, 'fixed1' /
# 'grad1' on the first fill, then its values / differences to the edge
# name: (binary, physics, builder, gamma, steps, typefull, limiter, typeB per face)
EXTRAP_X = ["cont1", "fixed1", "grad1"]
EXTRAP_CASES = {
    "mhd22_rf32x28_tvdlf_minmod_cont1": ("vac_mhd22_132", "mhd", lambda g: random_fourier((32, 28), 1001, g), 5 / 3, 10, "tvdlf", "minmod", ["cont1", "cont1", "periodic", "periodic"]),
    "mhd22_rf28x30_tvd_woodward_fixed1_grad1": ("vac_mhd22_132", "mhd", lambda g: random_fourier((28, 30), 1002, g), 5 / 3, 10, "tvd", "woodward", ["fixed1", "grad1", "cont1", "fixed1"]),
    "hd22_rf30x26_tvdmu_superbee_grad1": ("vac_hd22_132", "hd", lambda g: random_fourier((30, 26), 1003, g, physics="hd"), 1.4, 12, "tvdmu", "superbee", ["periodic", "periodic", "grad1", "grad1"]),
    "mhd33_rf10x12x8_tvdlf_woodward_extrap": ("vac_mhd33_36", "mhd", lambda g: random_fourier((10, 12, 8), 1004, g), 5 / 3, 5, "tvdlf", "woodward", ["cont1", "fixed1", "grad1", "cont1", "fixed1", "grad1"]),
}


# the projection of B before every step together with a corrector that
# limits the step's start (typelimited='previous', and dimsplit=F, where
# 'original' reads as 'previous'): VAC stores wold before advance, i.e.
# before the projection, so the limiter's base is the unprojected state
# name: (binary, builder, gamma, steps, typefull, limiter, bc, typelimited, unsplit, procpar)
PROJ_TL_CASES = {
    "mhd22_rf30x28_tvdlf_minmod_previous_proj": ("vacp_mhd22_132", lambda g: random_fourier((30, 28), 1101, g), 5 / 3, 12, "tvdlf", "minmod", "periodic", "previous", False, (0.0, 1e-4, 1.0)),
    "mhd22_rf28x26_tvdlf_woodward_unsplit_proj": ("vacp_mhd22_132", lambda g: random_fourier((28, 26), 1102, g), 1.4, 12, "tvdlf", "woodward", "periodic", "original", True, (0.0, 1e-4, 2.0)),
    "mhd33_rf10x8x12_tvdlf_minmod_unsplit_proj": ("vacp_mhd33_36", lambda g: random_fourier((10, 8, 12), 1103, g), 5 / 3, 6, "tvdlf", "minmod", "periodic", "original", True, (0.0, 1e-5, 1.0)),
    "mhd22_rf26x30_tvdmu_superbee_cont_previous_proj": ("vacp_mhd22_132", lambda g: random_fourier((26, 30), 1104, g), 5 / 3, 10, "tvdmu", "superbee", "cont", "previous", False, (0.01, 0.0, 3.0)),
}


# the rest of the non-default selectors: 'symm0' faces (ghost cells as
# 'symm', zero flux through the face in every addflux), unsplit TVD-MUSCL,
# useprimitive with typelimited 'previous' / 'predictor' (VAC converts the
# limiter's base in place every sweep), and the time-step controls
# dtcantgrow and slowsteps.  Walls: normal momentum and field 'asymm', the
# rest 'symm0'.
W0X = ["symm0", "asymm", "symm0", "symm0", "asymm", "symm0"]
W0Y = ["symm0", "symm0", "asymm", "symm0", "symm0", "asymm"]
W0Y_HD = ["symm0", "symm0", "asymm", "symm0"]
W0Z = ["symm0", "symm0", "symm0", "asymm", "symm0", "symm0", "symm0", "asymm"]
# name: dict(binary, physics, build, gamma, steps, typefull, limiter, bc,
#            method: extra &methodlist lines, param: extra &paramlist lines,
#            keys: what the fixture records of them, procpar)
GEN_CASES = {
    "mhd22_rf30x28_tvdlf_minmod_symm0": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((30, 28), 1201, g), gamma=5 / 3, steps=12, typefull="tvdlf", limiter="minmod", bc=[W0X, W0X, "periodic", "periodic"]),
    "mhd22_rf28x30_tvd_woodward_symm0": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((28, 30), 1202, g), gamma=5 / 3, steps=10, typefull="tvd", limiter="woodward", bc=["periodic", "periodic", W0Y, W0Y]),
    "mhd22_rf26x28_tvdmu_superbee_symm0": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((26, 28), 1203, g), gamma=1.4, steps=10, typefull="tvdmu", limiter="superbee", bc=[W0X, "cont", W0Y, W0Y]),
    "hd22_rf30x26_tvdlf_woodward_symm0": dict(binary="vac_hd22_132", physics="hd", build=lambda g: random_fourier((30, 26), 1204, g, physics="hd"), gamma=1.4, steps=15, typefull="tvdlf", limiter="woodward", bc=["cont", "cont", W0Y_HD, W0Y_HD]),
    "hd22_rf28x28_tvd_minmod_symm0": dict(binary="vac_hd22_132", physics="hd", build=lambda g: random_fourier((28, 28), 1205, g, physics="hd"), gamma=1.4, steps=12, typefull="tvd", limiter="minmod", bc=[["symm0", "asymm", "symm0", "symm0"]] * 2 + ["periodic", "periodic"]),
    "mhd33_rf10x12x8_tvdlf_minmod_symm0": dict(binary="vac_mhd33_36", physics="mhd", build=lambda g: random_fourier((10, 12, 8), 1206, g), gamma=5 / 3, steps=6, typefull="tvdlf", limiter="minmod", bc=["periodic", "periodic", "periodic", "periodic", W0Z, W0Z]),
    "mhd22_rf28x30_tvd_woodward_symm0_proj": dict(binary="vacp_mhd22_132", physics="mhd", build=lambda g: random_fourier((28, 30), 1207, g), gamma=5 / 3, steps=10, typefull="tvd", limiter="woodward", bc=[W0X, W0X, "periodic", "periodic"], procpar=(0.0, 1e-4, 1.0)),
    "mhd22_rf30x28_tvdmu_woodward_unsplit": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((30, 28), 1211, g), gamma=5 / 3, steps=12, typefull="tvdmu", limiter="woodward", bc="periodic", method=" dimsplit=F\n typelimited='original'\n", keys=dict(unsplit=True, typelimited="original")),
    "hd22_rf28x26_tvdmu_minmod_cont_unsplit_predictor": dict(binary="vac_hd22_132", physics="hd", build=lambda g: random_fourier((28, 26), 1212, g, physics="hd"), gamma=1.4, steps=12, typefull="tvdmu", limiter="minmod", bc="cont", method=" dimsplit=F\n typelimited='predictor'\n", keys=dict(unsplit=True, typelimited="predictor")),
    "mhd33_rf10x12x8_tvdmu_superbee_unsplit": dict(binary="vac_mhd33_36", physics="mhd", build=lambda g: random_fourier((10, 12, 8), 1213, g), gamma=5 / 3, steps=5, typefull="tvdmu", limiter="superbee", bc="periodic", method=" dimsplit=F\n", keys=dict(unsplit=True)),
    "mhd22_rf30x28_tvdlf_minmod_prim_previous": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((30, 28), 1221, g), gamma=5 / 3, steps=12, typefull="tvdlf", limiter="minmod", bc="periodic", method=" useprimitive=T\n typelimited='previous'\n", keys=dict(useprimitive=True, typelimited="previous")),
    "mhd22_rf28x30_tvdlf_woodward_cont_prim_predictor": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((28, 30), 1222, g), gamma=1.4, steps=12, typefull="tvdlf", limiter="woodward", bc="cont", method=" useprimitive=T\n typelimited='predictor'\n", keys=dict(useprimitive=True, typelimited="predictor")),
    "mhd22_rf26x30_tvdmu_superbee_prim_previous": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((26, 30), 1223, g), gamma=5 / 3, steps=10, typefull="tvdmu", limiter="superbee", bc="periodic", method=" useprimitive=T\n typelimited='previous'\n", keys=dict(useprimitive=True, typelimited="previous")),
    "hd22_rf28x26_tvdlf_minmod_prim_predictor": dict(binary="vac_hd22_132", physics="hd", build=lambda g: random_fourier((28, 26), 1224, g, physics="hd"), gamma=1.4, steps=15, typefull="tvdlf", limiter="minmod", bc="periodic", method=" useprimitive=T\n typelimited='predictor'\n", keys=dict(useprimitive=True, typelimited="predictor")),
    "mhd33_rf10x12x8_tvdlf_minmod_prim_previous": dict(binary="vac_mhd33_36", physics="mhd", build=lambda g: random_fourier((10, 12, 8), 1225, g), gamma=5 / 3, steps=6, typefull="tvdlf", limiter="minmod", bc="periodic", method=" useprimitive=T\n typelimited='previous'\n", keys=dict(useprimitive=True, typelimited="previous")),
    "mhd22_rf28x26_tvdlf_woodward_prim_unsplit": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((28, 26), 1226, g), gamma=5 / 3, steps=12, typefull="tvdlf", limiter="woodward", bc="periodic", method=" useprimitive=T\n dimsplit=F\n", keys=dict(useprimitive=True, unsplit=True)),
    "mhd22_rf26x28_tvdlf_minmod_prim_unsplit_predictor": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((26, 28), 1227, g), gamma=5 / 3, steps=12, typefull="tvdlf", limiter="minmod", bc="cont", method=" useprimitive=T\n dimsplit=F\n typelimited='predictor'\n", keys=dict(useprimitive=True, unsplit=True, typelimited="predictor")),
    "mhd22_rf28x26_tvdmu_woodward_prim_unsplit": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((28, 26), 1228, g), gamma=1.4, steps=10, typefull="tvdmu", limiter="woodward", bc="periodic", method=" useprimitive=T\n dimsplit=F\n", keys=dict(useprimitive=True, unsplit=True)),
    "mhd33_rf10x12x8_tvdmu_minmod_prim_predictor": dict(binary="vac_mhd33_36", physics="mhd", build=lambda g: random_fourier((10, 12, 8), 1229, g), gamma=5 / 3, steps=5, typefull="tvdmu", limiter="minmod", bc="periodic", method=" useprimitive=T\n typelimited='predictor'\n", keys=dict(useprimitive=True, typelimited="predictor")),
    "mhd22_rf28x26_tvdlf_minmod_slowsteps5": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: random_fourier((28, 26), 1231, g), gamma=5 / 3, steps=12, typefull="tvdlf", limiter="minmod", bc="periodic", param=" slowsteps=5\n", keys=dict(slowsteps=5)),
    "mhd22_rf26x28_tvd_woodward_dtcantgrow": dict(binary="vac_mhd22_132", physics="mhd", build=lambda g: orszag_tang(26, g), gamma=5 / 3, steps=15, typefull="tvd", limiter="woodward", bc="periodic", param=" dtcantgrow=T\n", keys=dict(dtcantgrow=True)),
    "hd22_rf30x26_tvdmu_minmod_dtcantgrow_slowsteps3": dict(binary="vac_hd22_132", physics="hd", build=lambda g: random_fourier((30, 26), 1233, g, physics="hd"), gamma=1.4, steps=12, typefull="tvdmu", limiter="minmod", bc="cont", param=" dtcantgrow=T\n slowsteps=3\n", keys=dict(dtcantgrow=True, slowsteps=3)),
}


def ensure_binary(name: str) -> None:
    if have_ref(name):
        return
    if name in POISSON:
        d, phys, g = POISSON[name]
        extra = ["cpu", "poisson"]
    else:
        d, phys, g = BINARIES[name]
        extra = []
    subprocess.run([os.path.join(ROOT, "oracle", "build_ref.sh"), name, d, phys, g, *extra],
                   check=True)


def run_case(binary, physics, w, coords, gamma, steps, typefull, limiter, bc, divbfix,
             divbwave=True, procpar=None):
    nw = w.shape[0]
    ini = make_ini(w, coords, physics, gamma, name="golden")
    extra = ""
    if procpar is not None:
        extra = " nproc=1\n procpar=" + ",".join(repr(float(v)) for v in procpar) + "\n"
    typeB = [bc] * (2 * len(coords)) if isinstance(bc, str) else bc
    par = problem_par(nw, NAMES[nw], steps, typefull=typefull, limiter=limiter,
                      typeB=typeB, divbfix=divbfix, divbwave=divbwave,
                      every_step=True, extra_method=extra)
    r = run_ref(binary, ini, par)
    t = np.array([s.t for s in r.snapshots])
    return r.snapshots[-1].w, t


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def main() -> None:
    only = sys.argv[1:]  # optional: write only the fixtures whose names contain one of these
    for b in ("vac_mhd22_132", "vac_hd22_132", "vac_mhd33_36"):
        ensure_binary(b)
    for name, (binary, physics, build, gamma, steps, typefull, limiter, bc, divbfix) in CASES.items():
        if only and not any(o in name for o in only):
            continue
        w0, coords = build(gamma)
        wf, t = run_case(binary, physics, w0, coords, gamma, steps, typefull, limiter, bc, divbfix)
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=wf, t=t,
            nx=np.array([len(c) for c in coords]), gamma=gamma, steps=steps,
            physics=physics, typefull=typefull, limiter=limiter, bc=bc,
            divbfix=divbfix)
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    for name, (binary, physics, build, gamma, steps, typefull, limiter, bc, divbfix, divbwave,
               procpar) in PROJ_CASES.items():
        if only and not any(o in name for o in only):
            continue
        ensure_binary(binary)
        w0, coords = build(gamma)
        wf, t = run_case(binary, physics, w0, coords, gamma, steps, typefull, limiter, bc,
                         divbfix, divbwave, procpar)
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=wf, t=t,
            nx=np.array([len(c) for c in coords]), gamma=gamma, steps=steps,
            physics=physics, typefull=typefull, limiter=limiter, bc=bc,
            divbfix=divbfix, divbwave=divbwave, nproc1=1, procpar=np.array(procpar))
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    for name, (binary, physics, build, gamma, steps, typefull, limiter, bc, divbfix,
               procpar) in BC_CASES.items():
        if only and not any(o in name for o in only):
            continue
        ensure_binary(binary)
        w0, coords = build(gamma)
        wf, t = run_case(binary, physics, w0, coords, gamma, steps, typefull, limiter, bc,
                         divbfix, True, procpar)
        extra = {} if procpar is None else dict(nproc1=1, procpar=np.array(procpar),
                                                divbwave=True)
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=wf, t=t,
            nx=np.array([len(c) for c in coords]), gamma=gamma, steps=steps,
            physics=physics, typefull=typefull, limiter=limiter, bc=json.dumps(bc),
            divbfix=divbfix, **extra)
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    for name, (binary, physics, build, gamma, steps, typefull, limiter, bc, tl) in TL_CASES.items():
        if only and not any(o in name for o in only):
            continue
        ensure_binary(binary)
        w0, coords = build(gamma)
        nw = w0.shape[0]
        ini = make_ini(w0, coords, physics, gamma, name="golden")
        par = problem_par(nw, NAMES[nw], steps, typefull=typefull, limiter=limiter,
                          typeB=[bc] * (2 * len(coords)), every_step=True,
                          extra_method=f" typelimited='{tl}'\n")
        r = run_ref(binary, ini, par)
        t = np.array([s_.t for s_ in r.snapshots])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=r.snapshots[-1].w, t=t,
            nx=np.array([len(c) for c in coords]), gamma=gamma, steps=steps,
            physics=physics, typefull=typefull, limiter=limiter, bc=bc, divbfix=True,
            typelimited=tl)
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    for name, (binary, physics, build, gamma, steps, typefull, limiter, ent) in ENT_CASES.items():
        if only and not any(o in name for o in only):
            continue
        ensure_binary(binary)
        w0, coords = build(gamma)
        nw = w0.shape[0]
        ini = make_ini(w0, coords, physics, gamma, name="golden")
        par = problem_par(nw, NAMES[nw], steps, typefull=typefull, limiter=limiter,
                          typeB=["periodic"] * (2 * len(coords)), every_step=True,
                          extra_method=f" typeentropy={nw}*'{ent}'\n")
        r = run_ref(binary, ini, par)
        t = np.array([s_.t for s_ in r.snapshots])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=r.snapshots[-1].w, t=t,
            nx=np.array([len(c) for c in coords]), gamma=gamma, steps=steps,
            physics=physics, typefull=typefull, limiter=limiter, bc="periodic", divbfix=True,
            typeentropy=ent)
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    for name, (binary, physics, build, gamma, steps, typefull, limiter, bc) in PRIM_CASES.items():
        if only and not any(o in name for o in only):
            continue
        ensure_binary(binary)
        w0, coords = build(gamma)
        nw = w0.shape[0]
        ini = make_ini(w0, coords, physics, gamma, name="golden")
        par = problem_par(nw, NAMES[nw], steps, typefull=typefull, limiter=limiter,
                          typeB=[bc] * (2 * len(coords)), every_step=True,
                          extra_method=" useprimitive=T\n")
        r = run_ref(binary, ini, par)
        t = np.array([s_.t for s_ in r.snapshots])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=r.snapshots[-1].w, t=t,
            nx=np.array([len(c) for c in coords]), gamma=gamma, steps=steps,
            physics=physics, typefull=typefull, limiter=limiter, bc=bc, divbfix=True,
            useprimitive=True)
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    for name, (binary, physics, build, gamma, steps, limiter, bc, tl) in UNSPLIT_CASES.items():
        if only and not any(o in name for o in only):
            continue
        ensure_binary(binary)
        w0, coords = build(gamma)
        nw = w0.shape[0]
        ini = make_ini(w0, coords, physics, gamma, name="golden")
        par = problem_par(nw, NAMES[nw], steps, typefull="tvdlf", limiter=limiter,
                          typeB=[bc] * (2 * len(coords)), every_step=True,
                          extra_method=f" dimsplit=F\n typelimited='{tl}'\n")
        r = run_ref(binary, ini, par)
        t = np.array([s_.t for s_ in r.snapshots])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=r.snapshots[-1].w, t=t,
            nx=np.array([len(c) for c in coords]), gamma=gamma, steps=steps,
            physics=physics, typefull="tvdlf", limiter=limiter, bc=bc, divbfix=True,
            typelimited=tl, unsplit=True)
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    for name, (binary, physics, build, gamma, steps, typefull, limiter, bc) in EXTRAP_CASES.items():
        if only and not any(o in name for o in only):
            continue
        ensure_binary(binary)
        w0, coords = build(gamma)
        nw = w0.shape[0]
        ini = make_ini(w0, coords, physics, gamma, name="golden")
        par = problem_par(nw, NAMES[nw], steps, typefull=typefull, limiter=limiter,
                          typeB=bc, every_step=True)
        r = run_ref(binary, ini, par)
        t = np.array([s_.t for s_ in r.snapshots])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=r.snapshots[-1].w, t=t,
            nx=np.array([len(c) for c in coords]), gamma=gamma, steps=steps,
            physics=physics, typefull=typefull, limiter=limiter, bc=json.dumps(bc), divbfix=True)
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    for name, (binary, build, gamma, steps, typefull, limiter, bc, tl, unsplit,
               procpar) in PROJ_TL_CASES.items():
        if only and not any(o in name for o in only):
            continue
        ensure_binary(binary)
        w0, coords = build(gamma)
        nw = w0.shape[0]
        ini = make_ini(w0, coords, "mhd", gamma, name="golden")
        extra = (f" typelimited='{tl}'\n" + (" dimsplit=F\n" if unsplit else "")
                 + " nproc=1\n procpar=" + ",".join(repr(float(v)) for v in procpar) + "\n")
        par = problem_par(nw, NAMES[nw], steps, typefull=typefull, limiter=limiter,
                          typeB=[bc] * (2 * len(coords)), every_step=True, extra_method=extra)
        r = run_ref(binary, ini, par)
        t = np.array([s_.t for s_ in r.snapshots])
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=r.snapshots[-1].w, t=t,
            nx=np.array([len(c) for c in coords]), gamma=gamma, steps=steps,
            physics="mhd", typefull=typefull, limiter=limiter, bc=bc, divbfix=True,
            divbwave=True, typelimited=tl, unsplit=unsplit, nproc1=1,
            procpar=np.array(procpar))
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    for name, c in GEN_CASES.items():
        if only and not any(o in name for o in only):
            continue
        ensure_binary(c["binary"])
        w0, coords = c["build"](c["gamma"])
        nw = w0.shape[0]
        ini = make_ini(w0, coords, c["physics"], c["gamma"], name="golden")
        bc = c["bc"]
        typeB = [bc] * (2 * len(coords)) if isinstance(bc, str) else bc
        method = c.get("method", "")
        procpar = c.get("procpar")
        if procpar is not None:
            method += " nproc=1\n procpar=" + ",".join(repr(float(v)) for v in procpar) + "\n"
        par = problem_par(nw, NAMES[nw], c["steps"], typefull=c["typefull"], limiter=c["limiter"],
                          typeB=typeB, every_step=True, extra_method=method,
                          extra_param=c.get("param", ""))
        r = run_ref(c["binary"], ini, par)
        t = np.array([s_.t for s_ in r.snapshots])
        extra = dict(c.get("keys", {}))
        if procpar is not None:
            extra.update(nproc1=1, procpar=np.array(procpar), divbwave=True)
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"), w0=w0, wf=r.snapshots[-1].w, t=t,
            nx=np.array([len(cc) for cc in coords]), gamma=c["gamma"], steps=c["steps"],
            physics=c["physics"], typefull=c["typefull"], limiter=c["limiter"],
            bc=bc if isinstance(bc, str) else json.dumps(bc), divbfix=True, **extra)
        print("wrote", name, w0.shape, "steps", len(t) - 1)
    if only:
        return
    # config 1 at full size: digests only
    g = 5.0 / 3.0
    w0, coords = orszag_tang(128, g)
    wf, t = run_case("vac_mhd22_132", "mhd", w0, coords, g, 100, "tvdlf", "minmod", "periodic", True)
    rec = {"case": "Orszag-Tang 128x128, gamma=5/3, TVDLF+minmod, twostep, courantpar=0.8, 100 steps, periodic, divbfix=T",
           "w0_sha256": sha(w0), "wf_sha256": sha(wf), "t": [float(v) for v in t],
           "wf_sum_abs": [float(np.abs(wf[i]).sum()) for i in range(wf.shape[0])]}
    with open(os.path.join(HERE, "ot128_tvdlf_minmod_100.json"), "w") as f:
        json.dump(rec, f, indent=1)
    print("wrote ot128_tvdlf_minmod_100.json")


if __name__ == "__main__":
    main()
