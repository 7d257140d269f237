"""MI355X-native MAS-preconditioned conjugate gradient for StiffGIPC's
Newton-step linear solve (HIP kernels for gfx950 behind the C ABI in
include/masx.h)."""

__all__ = ["native", "workloads", "build"]
