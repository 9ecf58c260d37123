"""Python view of the MI355X vid.stab detect library (lib/libvidstab_amd.so).

The library itself is C + HIP; this package only binds its C ABI with ctypes
for tests and benchmarks (`api`), and wraps the synthetic clip generator
(`synth`).
"""
