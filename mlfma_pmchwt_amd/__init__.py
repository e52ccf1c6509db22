"""MI355X-native MLFMA / PMCHWT solver engine (HIP kernels behind a C ABI)."""
