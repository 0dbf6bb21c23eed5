"""yafel_amd: MI355X (gfx950) implementation of YAFEL's assembly -> sparse ->
Krylov hot path.  The compute lives in libyafel_hip.so (hand-written HIP
kernels behind the C ABI of include/yafel_hip.h); `yafel_amd.hip` is its
ctypes binding and `yafel_amd.workloads` builds the benchmark workloads."""

__all__ = ["hip", "workloads"]
