"""The reference executables oracle/build_ref.sh makes (test infrastructure).

BINARIES: plain VAC 4.52 builds (the CPU oracle and baseline).
POISSON:  plain builds with the Poisson module switched on (projection of B).
HOOKED:   the same VAC host with the GPU drop-in linked in place of advect.
HOOKED_USR: drop-in builds with a user module other than nul.
TIMING:   plain builds compiled for speed (the CPU baseline of bench.py), with
          the compiler flags they take instead of the parity build's.
HOOKED_POISSON: the drop-in host with the Poisson module on (process(1)
          then projects on the device through vacgpu_process).
Each entry: name -> (setvac -d value, physics, compile-time grid ixGhi).
"""
BINARIES = {
    "vac_mhd22_132": ("22", "mhd", "132,132"),
    "vac_hd22_132": ("22", "hd", "132,132"),
    "vac_hd33_36": ("33", "hd", "36,36,36"),
    "vac_mhd33_36": ("33", "mhd", "36,36,36"),
    "vac_mhd22_2052": ("22", "mhd", "2052,2052"),
}
# VAC with the Poisson module on (setvac -on=poisson): process() projects B
POISSON = {
    "vacp_mhd22_132": ("22", "mhd", "132,132"),
    "vacp_mhd33_36": ("33", "mhd", "36,36,36"),
}
HOOKED_POISSON = {
    "vacgpup_mhd22_132": ("22", "mhd", "132,132"),
}
HOOKED = {
    "vacgpu_mhd22_132": ("22", "mhd", "132,132"),
    "vacgpu_hd22_132": ("22", "hd", "132,132"),
    "vacgpu_mhd33_36": ("33", "mhd", "36,36,36"),
    # the drop-in at the 2048^2 size (bench.py times VAC's host loop with it)
    "vacgpu_mhd22_2052": ("22", "mhd", "2052,2052"),
}
# drop-in builds with one of the reference's user modules (VAC_USR=...):
# name -> (setvac -d value, physics, compile-time grid, user module)
HOOKED_USR = {
    "vacgpu_mhd22_132_gravity": ("22", "mhd", "132,132", "gravity"),
}

# the CPU baseline: VAC compiled -O3 for a portable x86-64-v3 target (AVX2 +
# FMA; the executables are built here and run on another machine's cores, so
# -march=native of this host could emit instructions that CPU lacks); FMA
# contraction stays at the compiler's default, as a user's build would have it
TIMING_FLAGS = "-O3 -march=x86-64-v3"
TIMING = {
    "vacO3_mhd22_132": ("22", "mhd", "132,132"),
    "vacO3_mhd22_2052": ("22", "mhd", "2052,2052"),
    # config 3's 3D TVDLF, timed on a 128^3 sample of it
    "vacO3_mhd33_132": ("33", "mhd", "132,132,132"),
}
