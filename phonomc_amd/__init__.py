"""MI355X-native deviational phonon Monte Carlo step (PhonoMC drop-in).

The compute path is the HIP shared library ``_build/libpmc_device.so``
(include/pmc_capi.h) driven by the C++17 host facade ``_build/libpmc_host.so``
(include/pmc_solver.h).  ``phonomc_amd.runtime`` loads both through ctypes.
"""
