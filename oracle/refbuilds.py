"""The reference executables oracle/build_ref.sh makes (test infrastructure).

BINARIES: plain VAC 4.52 builds (the CPU oracle and baseline).
POISSON:  plain builds with the Poisson module switched on (projection of B).
HOOKED:   the same VAC host with the GPU drop-in linked in place of advect.
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
}
