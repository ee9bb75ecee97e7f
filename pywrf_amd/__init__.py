"""MI355X-native PyWRF ARW RK3 step: HIP kernels behind the reference's
wrf_dynamics / wrf_physics / wrf_boundary function signatures."""
