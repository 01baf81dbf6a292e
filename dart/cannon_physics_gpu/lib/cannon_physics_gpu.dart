/// cannon_physics on the GPU: libcannon_hip.so (gfx950) behind the reference's own classes.
///
///  * [GpuWorld]: a `World` whose `internalStep` runs on the device (state stays there; body
///    objects are refreshed after each step), with the World's contact events, sleep, ray casts,
///    `aabbQuery` and interpolated stepping on the device.
///  * [GpuBroadphase], [GpuNarrowphase], [GpuSolver]: slot drop-ins for a reference `World`
///    that keeps driving the step itself.
library cannon_physics_gpu;

export 'src/bindings.dart' show GpuContext, CannonHipException, chpPrecisionF32, chpPrecisionF64;
export 'src/gpu_world.dart';
export 'src/slots.dart';
