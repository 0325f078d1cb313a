"""MI355X-native GAMERA MHD time step (gfx950 HIP kernels behind a C ABI)."""
