"""MI355X-native plane-current PEEC solve (HIP kernels behind a C ABI, Python mirror of the reference's solver)."""
