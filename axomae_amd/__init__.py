"""MI355X-native Nova hot path: HIP kernels for gfx950 behind a C ABI
(include/nova_hip.h), with deterministic synthetic workloads."""
