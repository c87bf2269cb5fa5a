"""MI355X-native DEC Gross-Pitaevskii stepper: HIP kernels behind a C ABI (libgpe.so)."""
