"""MI355X-native Simulon MD step: HIP kernels for the Verlet neighbor build,
LJ and EAM force fields and the fused velocity-Verlet/Langevin/Berendsen
update, behind a C ABI (include/simulon_hip.h, libsimulon_hip.so), with a
Python host layer that mirrors the reference's API (Box, AtomFileReader,
find_neighbors_gpu_pbc, LennardJonesForce, EAMForceCUDAExt, VerletIntegrator,
BerendsenBarostat, BaseModel) and a packed multi-replica engine."""
__version__ = "0.1.0"
